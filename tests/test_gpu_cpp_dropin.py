"""The C++ drop-in classes (merge_fusion_amd/cpp/kifi/*.h, the reference's
KinectFission/kifi/*.h API) driven by a compiled C++ program, checked against
the oracle: Integrator + Volume::Data and Mesher bit-exact, Mesher::Mesh2Obj's
format, Renderer::Bin + ICP::Align against the binary32 restatement of
ICP::Align (1 mm / 0.1 deg), Pipeline with given poses bit-exact.
"""
import os
import subprocess

import numpy as np
import pytest

import merge_fusion_amd as mf
import oracle as orc
from test_depthstream import fixture

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "merge_fusion_amd", "cpp", "tests", "test_dropin")
CAM = mf.Intrinsics(512, 424, 365.0, 365.0)


def test_dropin_headers_compile():
    """The headers compile as plain C++17 with no GPU toolchain (host code only)."""
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I", os.path.join(ROOT, "merge_fusion_amd", "cpp"), "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "merge_fusion_amd", "cpp", "tests", "test_dropin.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    if not os.path.exists(BIN):
        pytest.fail("C++ drop-in test program not built (python -m merge_fusion_amd.build)")
    out = str(tmp_path_factory.mktemp("dropin"))
    depth_file = fixture(out, "imrod_v2.depth")
    r = subprocess.run([BIN, out, depth_file], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    return r, out


def _load(out, name, dtype):
    return np.fromfile(os.path.join(out, name), dtype=dtype)


@pytest.mark.gpu
def test_program_checks_pass(run):
    r, out = run
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout and "DONE failures=0" in r.stdout


@pytest.mark.gpu
def test_integrator_and_mesher_bit_exact(run):
    r, out = run
    depth_m = _load(out, "c1_depth_m.f32", np.float32).reshape(424, 512)
    mm = np.rint(depth_m * np.float32(1000.0)).astype(np.uint16)
    assert np.array_equal(mm.astype(np.float32) * np.float32(0.001), depth_m)
    ov = orc.Volume(1024, 8.192, 0.04)
    ov.integrate(mm, CAM, np.eye(4, dtype=np.float32).reshape(16))
    ok, oval = ov.read()
    keys = _load(out, "c1_keys.u32", np.uint32)
    vals = _load(out, "c1_values.f32x2", np.float32).reshape(-1, 2)
    assert np.array_equal(keys.astype(np.uint64), ok)
    assert np.array_equal(vals.view(np.uint32), oval.view(np.uint32))
    verts = _load(out, "c1_verts.f32x3", np.float32).reshape(-1, 3)
    assert np.array_equal(verts.view(np.uint32), ov.mesh().view(np.uint32))
    ov_v, ov_i, _ = ov.mesh_indexed()
    iverts = _load(out, "c1_iverts.f32x3", np.float32).reshape(-1, 3)
    idx = _load(out, "c1_idx.u32", np.uint32)
    assert np.array_equal(iverts.view(np.uint32), ov_v.view(np.uint32))
    assert np.array_equal(idx, ov_i)
    # Mesh2Obj: "v x y z" per vertex, then 1-based "f a b c" per triangle
    lines = open(os.path.join(out, "c1_mesh.obj")).read().splitlines()
    v = [l for l in lines if l.startswith("v ")]
    f = [l for l in lines if l.startswith("f ")]
    assert len(v) == len(iverts) and len(f) * 3 == len(idx) and len(v) + len(f) == len(lines)
    assert np.allclose(np.array([list(map(float, l.split()[1:])) for l in v[:50]]), iverts[:50], atol=1e-6)
    assert np.array_equal(np.array([list(map(int, l.split()[1:])) for l in f[:50]]).reshape(-1), idx[:150] + 1)


@pytest.mark.gpu
def test_icp_align_matches_reference_restatement(run):
    r, out = run
    synth = _load(out, "icp_synth.f32x3", np.float32).reshape(768, 1024, 3)
    raw_m = _load(out, "icp_raw_m.f32", np.float32).reshape(424, 512)
    guess = _load(out, "icp_guess.f32", np.float32)
    aligned = _load(out, "icp_aligned.f32", np.float32)
    mm = np.rint(raw_m * np.float32(1000.0)).astype(np.uint16)
    opose, n = orc.icp_align_ref(mm, CAM, synth, guess, guess)
    dt, da = mf.pose_error(aligned.astype(np.float64), opose.astype(np.float64))
    assert dt <= 1e-3 and da <= 0.1, (dt, da)


@pytest.mark.gpu
def test_pipeline_given_poses_bit_exact(run):
    r, out = run
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    ov = orc.Volume(1024, 4.096, 0.02)
    for f in range(3):
        pose = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, f)
        w2e = orc.invert_transform(pose.astype(np.float32))
        ov.integrate(mf.render(sc, CAM, pose, noise_seed=10 + f), CAM, w2e)
    ok, oval = ov.read()
    keys = _load(out, "pipe_keys.u32", np.uint32)
    vals = _load(out, "pipe_values.f32x2", np.float32).reshape(-1, 2)
    assert np.array_equal(keys.astype(np.uint64), ok)
    assert np.array_equal(vals.view(np.uint32), oval.view(np.uint32))


@pytest.mark.gpu
def test_depthstream_integrator_mesher_bit_exact(run):
    """kifi::DepthStream + Integrator + Mesher on the reference's imrod_v2.depth
    (UnitTests/Integrator.cpp:21-52, UnitTests/Mesher.cpp): the frame in metres as
    the reference reads it, integrated and meshed bit for bit as the oracle."""
    r, out = run
    assert "depthstream frames read 1 of 1" in r.stdout
    depth = _load(out, "ds_depth_m.f32", np.float32).reshape(480, 640)
    view = _load(out, "ds_view.f32", np.float32)
    with mf.DepthStream(os.path.join(out, "imrod_v2.depth")) as ds:
        d, v = ds.next_frame()
    assert np.array_equal(depth, d) and np.array_equal(view, v)
    cam = mf.Intrinsics(640, 480, 571.26, 571.26)
    cam.cx, cam.cy, cam.znear, cam.zfar = 320.0, 240.0, 0.8, 4.0
    ov = orc.Volume(256, 2.0, 0.02)
    ov.integrate(depth, cam, view)
    ok, oval = ov.read()
    assert np.array_equal(_load(out, "ds_keys.u32", np.uint32), ok.astype(np.uint32))
    assert np.array_equal(_load(out, "ds_values.f32x2", np.float32).reshape(-1, 2).view(np.uint32),
                          oval.view(np.uint32))
    assert np.array_equal(_load(out, "ds_verts.f32x3", np.float32).reshape(-1, 3).view(np.uint32),
                          ov.mesh().view(np.uint32))
