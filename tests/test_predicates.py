"""CPU checks of exact predicates and conservative tests the HIP kernels use in
place of divisions and of the reference-order projection.

1. k_merge_update's projection keeps a voxel only if z/w is finite
   (Integrator.cpp:305-315: (z/w)*0+0 must lie in [-1, 1)).  With w finite and
   nonzero, RN(z/w) is finite exactly when |z| < |w| (2^128 - 2^103).  The
   kernel decides it in binary32 without dividing in the common case:
   |z| < |w| 2^127 is finite, |z| >= |w| 2^128 is not, and only between the two
   does it divide.  The tests pin that rule against IEEE binary32 division,
   including quotients within a few ulps of the overflow midpoint.

2. The update's cone pre-test: the clip rows x, y, w of a voxel centre as
   three fused multiply-adds of its integer coordinates, rejected when
   |x| or |y| > 1.02 |w| + eps.  Every rejected voxel must have |x/w| or |y/w|
   above 1.01 in the reference's binary32 order (so it lands outside the
   image whatever the rounding): checked over random poses and voxels of the
   configurations' volumes, with the margins the kernel computes.
"""
import numpy as np

C = float.fromhex("0x1.ffffffp127")


def _pred(z, w):
    """The kernel's rule, in binary32 (power-of-two scalings are exact)."""
    z = np.asarray(z, np.float32)
    w = np.asarray(w, np.float32)
    za, wa = np.abs(z), np.abs(w)
    with np.errstate(all="ignore"):
        ok = za < wa * np.float32(2.0 ** 127)
        mid = ~ok & ~(za >= wa * np.float32(2.0 ** 127) * np.float32(2.0))
        ok = ok | (mid & np.isfinite(z / w))
    return ok


def _pred_double(z, w):
    """The closed form: |z| < |w| (2^128 - 2^103), exact in double."""
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
    assert np.array_equal(_pred_double(z, w), _truth(z, w))


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
    assert np.array_equal(_pred_double(z, ww), t)


def test_special_values():
    f = np.float32
    z = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, 3e38], np.float32)
    w = np.array([1.0, -2.0, 1.0, 1.0, 1.0, 1e-45, 1e-1], np.float32)
    assert np.array_equal(_pred(z, w), _truth(z, w))
    assert not _pred(np.array([f(np.nan)]), np.array([f(1.0)]))[0]


# ---------------------------------------------------------------- cone pre-test
F32 = np.float32


def _fma(a, b, c):
    """binary32 fused multiply-add: the product of two binary32 values is exact
    in double; the double sum is rounded to binary32 (its own rounding only
    matters within half an ulp of the sum, far inside the test's margins)."""
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F32)


def _eye_to_clip(W, H, f):
    """EyeToClipRH for cx = W/2, cy = H/2 (DepthSensorParams.cpp:47-63), the
    rows the update uses: x = 2f/W x_e, y = 2f/H y_e, w = -z_e."""
    P = np.zeros((4, 4), F32)
    P[0, 0] = F32(2 * f / W)
    P[1, 1] = F32(2 * f / H)
    P[2, 2], P[2, 3] = F32(-1.0), F32(-0.8)
    P[3, 2] = F32(-1.0)
    return P


def _rand_w2c(rng, P, spread):
    ax = rng.standard_normal(3)
    ax /= np.linalg.norm(ax)
    a = rng.uniform(0, np.pi)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = rng.uniform(-spread, spread, 3)
    return (P.astype(np.float64) @ T).astype(F32)


def _pre(M, vl, c0, res):
    """The kernel's clip_pre in binary32."""
    C_ = F32(F32(res) * vl + abs(c0))
    rows, S = [], F32(0)
    for r in (0, 1, 3):
        m = M[r]
        a = [F32(m[0] * vl), F32(m[1] * vl), F32(m[2] * vl)]
        a3 = F32(F32(F32(F32(m[0] * c0) + F32(m[1] * c0)) + F32(m[2] * c0)) + m[3])
        rows.append(np.array(a + [a3], F32))
        S = max(S, F32(F32(F32(abs(m[0]) + abs(m[1])) + abs(m[2])) * C_ + abs(m[3])))
    return rows, F32(F32(64.0 * 2.0 ** -24) * S)


def _ref_clip(M, X, vl, c0):
    """Reference-order binary32: centre = X vl + c0, then (m0 c0 + m1 c1) + (m2 c2 + m3)."""
    c = (X.astype(F32) * vl + c0).astype(F32)
    out = []
    for r in (0, 1, 3):
        m = M[r]
        out.append(F32(1) * (((m[0] * c[:, 0]).astype(F32) + (m[1] * c[:, 1]).astype(F32)).astype(F32) +
                              ((m[2] * c[:, 2]).astype(F32) + m[3]).astype(F32)).astype(F32))
    return out


def test_cone_pretest_rejects_only_what_the_reference_rejects():
    rng = np.random.default_rng(11)
    P = _eye_to_clip(512, 424, 365.0)
    n_rej = 0
    for res, side, spread in ((1024, 4.096, 2.5), (1024, 8.192, 4.0), (4096, 8.192, 4.0), (256, 1.0, 0.6)):
        vl = F32(F32(side) / F32(res))
        c0 = F32(vl * F32(0.5) + F32(-0.5 * side))
        for _ in range(40):
            M = _rand_w2c(rng, P, spread)
            rows, eps = _pre(M, vl, c0, res)
            X = rng.integers(0, res, (20000, 3))
            # voxels near the camera plane and near the cone's edges: camera-centred rays
            fx = X.astype(F32)
            px = _fma(rows[0][0:1], fx[:, 0], _fma(rows[0][1:2], fx[:, 1], _fma(rows[0][2:3], fx[:, 2],
                                                                                  rows[0][3:4])))
            py = _fma(rows[1][0:1], fx[:, 0], _fma(rows[1][1:2], fx[:, 1], _fma(rows[1][2:3], fx[:, 2],
                                                                                  rows[1][3:4])))
            pw = _fma(rows[2][0:1], fx[:, 0], _fma(rows[2][1:2], fx[:, 1], _fma(rows[2][2:3], fx[:, 2],
                                                                                  rows[2][3:4])))
            lim = _fma(np.full_like(pw, F32(1.02)), np.abs(pw), np.full_like(pw, eps))
            rej = (np.abs(px) > lim) | (np.abs(py) > lim)
            x, y, w = _ref_clip(M, X, vl, c0)
            ref_out = (np.abs(x) > F32(1.01) * np.abs(w)) | (np.abs(y) > F32(1.01) * np.abs(w))
            assert not (rej & ~ref_out).any()
            # the margin is tight enough to matter: most reference-rejected voxels are pre-rejected
            assert rej.sum() >= 0.95 * ref_out.sum()
            n_rej += int(rej.sum())
    assert n_rej > 1_000_000
