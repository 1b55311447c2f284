"""Host sanitizers over the CPU oracle and the seeded scene renderer: the
oracle's entry points (integration for binary32, binary16 + u8, 64-bit keys and
FP64 storage; both meshers; model splat, resolve and normals; the depth
pyramid; FP64 ICP and its sums; the binary32 term restatement; the
reference's window-search ICP; Bin; Jacobi, Horn, Kahan, scan) and the
renderer, built with -fsanitize=address,undefined (oracle/Makefile, target
`sanitize`) and run to completion: any finding aborts with a non-zero status."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


def test_oracle_and_renderer_run_clean_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE, "_san", "sanitize_driver")], env=env, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:]
    assert r.stdout.count("voxels,") == 4 and "kahan 1.000000" in r.stdout, r.stdout
