"""CPU checks of exact predicates the HIP kernels use in place of divisions.

k_merge_update's projection keeps a voxel only if z/w is finite
(Integrator.cpp:305-315: (z/w)*0+0 must lie in [-1, 1)).  With w finite and
nonzero the kernel decides it without dividing: RN(z/w) is finite exactly when
|z| < |w| (2^128 - 2^103), compared in double where the product is exact.  The
test pins the constant and the boundary against IEEE binary32 division.
"""
import numpy as np

C = float.fromhex("0x1.ffffffp127")


def _pred(z, w):
    return np.abs(z.astype(np.float64)) < np.abs(w.astype(np.float64)) * C


def _truth(z, w):
    with np.errstate(all="ignore"):
        return np.isfinite(z / w)


def test_constant_is_the_overflow_boundary():
    assert C == 2.0 ** 128 - 2.0 ** 103
    fmax = np.finfo(np.float32).max
    assert float(fmax) == 2.0 ** 128 - 2.0 ** 104
    # a quotient just below the midpoint rounds to FLT_MAX, the midpoint itself to inf (ties to even)
    one = np.float32(1.0)
    with np.errstate(over="ignore"):
        assert np.isfinite(fmax / one) and not np.isfinite(np.float32(fmax) * np.float32(2.0) / np.float32(2.0 - 2 ** -22))


def test_predicate_matches_binary32_division():
    rng = np.random.default_rng(7)
    n = 400000
    # exponents spread over the whole binary32 range, subnormal w included
    with np.errstate(over="ignore"):
        z = (rng.standard_normal(n) * 2.0 ** rng.integers(-150, 128, n)).astype(np.float32)
        w = (rng.standard_normal(n) * 2.0 ** rng.integers(-150, 128, n)).astype(np.float32)
    ok = np.isfinite(w) & (w != 0)
    z, w = z[ok], w[ok]
    assert np.array_equal(_pred(z, w), _truth(z, w))


def test_predicate_at_the_rounding_boundary():
    # z/w within a few ulps of the overflow midpoint 2^128 - 2^103, both sides
    rng = np.random.default_rng(8)
    w = (rng.uniform(0.5, 1.0, 20000) * 2.0 ** rng.integers(-140, -1, 20000)).astype(np.float32)
    target = (2.0 ** 128 - 2.0 ** 103) * w.astype(np.float64)
    zs = []
    for k in range(-3, 4):
        z = target.astype(np.float32)
        z = np.nextafter(z, np.float32(np.inf)) if k > 0 else z
        for _ in range(abs(k) - 1 if k else 0):
            z = np.nextafter(z, np.float32(np.inf) if k > 0 else np.float32(0))
        zs.append(z)
    z = np.concatenate(zs)
    ww = np.tile(w, 7)
    fin = np.isfinite(z)
    z, ww = z[fin], ww[fin]
    t, p = _truth(z, ww), _pred(z, ww)
    assert t.any() and (~t).any()                  # both outcomes occur
    assert np.array_equal(p, t)


def test_special_values():
    f = np.float32
    z = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, 3e38], np.float32)
    w = np.array([1.0, -2.0, 1.0, 1.0, 1.0, 1e-45, 1e-1], np.float32)
    assert np.array_equal(_pred(z, w), _truth(z, w))
    assert not _pred(np.array([f(np.nan)]), np.array([f(1.0)]))[0]
