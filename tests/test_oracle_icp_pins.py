"""Pins of the oracle's tracking half (oracle/oracle_icp.c, oracle/oracle.c
orc_align_step) by closed forms and brute force, on the CPU.

- depth pyramid: brute-force 2x2 mean of valid depths, binary32, same order;
- model prediction: a fronto-parallel plane seen head-on predicts points on the
  plane (exactly, noise-free) with normal +z, each point inside the footprint
  of its pixel; sphere points lie within half a voxel of the analytic sphere
  with roughly radial normals;
- FP64 point-to-plane ICP: recovers known camera motions of noise-free frames
  from the previous pose to the model's own accuracy (a dropped Jacobian term
  or a sign error leaves the motion itself, centimetres, as the error);
- the reference's window-search Horn ICP (ICP.cpp:9-265): with an exact model
  and the true guess it returns the guess; from a perturbed guess it moves
  towards the truth and converges.
"""
import numpy as np
import pytest
from scipy.ndimage import binary_erosion

import merge_fusion_amd as mf
import oracle as orc

CAM = mf.Intrinsics(512, 424, 365.0, 365.0)


def _brute_level(src):
    H, W = src.shape
    out = np.zeros((H // 2, W // 2), np.float32)
    for y in range(H // 2):
        for x in range(W // 2):
            s, c = np.float32(0), np.float32(0)
            for j in range(2):
                for i in range(2):
                    d = src[2 * y + j, 2 * x + i]
                    if d > 0:
                        s = np.float32(s + d)
                        c = np.float32(c + 1)
            out[y, x] = np.float32(s / c) if c > 0 else 0
    return out


def test_depth_pyramid_brute_force():
    rng = np.random.default_rng(3)
    mm = rng.integers(400, 8000, (24, 32)).astype(np.uint16)
    mm[rng.random(mm.shape) < 0.3] = 0
    mm[0:2, 0:2] = 0                       # an all-invalid block
    lv = orc.depth_pyramid(mm, 3)
    assert lv[0].dtype == np.float32 and np.array_equal(lv[0], mm.astype(np.float32) * np.float32(0.001))
    assert np.array_equal(lv[1], _brute_level(lv[0]))
    assert np.array_equal(lv[2], _brute_level(lv[1]))
    assert lv[1][0, 0] == 0
    hand = np.array([[1000, 0], [0, 3000]], np.uint16)
    assert orc.depth_pyramid(hand, 2)[1][0, 0] == np.float32(2.0)


@pytest.fixture(scope="module")
def plane_model():
    sc = mf.make_scene(mf.SCENE_PLANE_SPHERES, 1)
    I = np.eye(4).reshape(16)
    d, truth = mf.render(sc, CAM, I, noise_seed=1, noise_k=0.0, invalid_frac=0.0, with_truth=True)
    ov = orc.Volume(1024, 8.192, 0.04)
    ov.integrate(d, CAM, np.eye(4, dtype=np.float32).reshape(16))
    pts, nrm = orc.model_predict(ov, np.eye(4, dtype=np.float32).reshape(16), CAM)
    return sc, truth, pts, nrm


def test_model_predict_plane_closed_form(plane_model):
    sc, truth, pts, nrm = plane_model
    plane = truth == np.float32(2.5)
    core = binary_erosion(plane, np.ones((11, 11), bool))
    assert core.sum() > 100000
    assert (pts[..., 3][core] == 1).all() and (nrm[..., 3][core] == 1).all()
    # every predicted point lies on the plane z = -2.5 with normal +z (towards the camera)
    assert np.abs(pts[..., 2][core] + 2.5).max() <= 1e-6
    assert (nrm[..., 2][core] == 1.0).all()
    # and projects into (at most one pixel beside) the pixel that holds it
    ys, xs = np.nonzero(core)
    p = pts[ys, xs]
    u = CAM.fx * p[:, 0] / (-p[:, 2]) + CAM.cx
    v = CAM.cy - CAM.fy * p[:, 1] / (-p[:, 2])
    assert np.abs(u - xs - 0.5).max() < 1.0 and np.abs(v - ys - 0.5).max() < 1.0


def test_model_predict_sphere_surface_and_normals(plane_model):
    sc, truth, pts, nrm = plane_model
    seen = 0
    for i in range(sc.n_spheres):
        c = np.array(sc.spheres[i][:3])
        r = sc.spheres[i][3]
        q = pts[..., :3].astype(np.float64) - c
        dist = np.linalg.norm(q, axis=-1)
        on = (pts[..., 3] > 0) & (nrm[..., 3] > 0) & (np.abs(dist - r) < 0.01)
        on = binary_erosion(on, np.ones((11, 11), bool))
        if on.sum() < 1000:
            continue                       # mostly hidden behind another sphere
        seen += 1
        assert np.abs(dist[on] - r).max() < 0.5 * 0.008      # half a voxel
        cosang = (nrm[..., :3][on] * (q[on] / dist[on][:, None])).sum(1)
        ang = np.degrees(np.arccos(np.clip(cosang, -1, 1)))
        assert np.median(ang) < 3.0 and ang.max() < 15.0
    assert seen >= 1


@pytest.fixture(scope="module")
def room_model():
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    p0 = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 0)
    w2e = mf.invert_rigid(p0).astype(np.float32)
    ov = orc.Volume(1024, 4.096, 0.02)
    ov.integrate(mf.render(sc, CAM, p0, noise_seed=1, noise_k=0.0, invalid_frac=0.0), CAM, w2e)
    pts, nrm = orc.model_predict(ov, w2e, CAM)
    return sc, p0, w2e, pts, nrm


@pytest.mark.parametrize("target", [0, 1, 3, 6])
def test_icp_f64_recovers_known_motion(room_model, target):
    sc, p0, w2e, pts, nrm = room_model
    pt = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, target)
    d = mf.render(sc, CAM, pt, noise_seed=5, noise_k=0.0, invalid_frac=0.0)
    pose, iters, rms = orc.icp_f64(d, CAM, pts, nrm, w2e.astype(np.float64), p0)
    dt, da = mf.pose_error(pose, pt)
    # the 4 mm TSDF model itself is accurate to about 2 mm / 0.05 deg
    assert dt < 2.5e-3 and da < 0.06, (dt, da)
    assert 3 <= iters <= 19 and rms < 2e-3


def _backproject_world(d, e2w):
    H, W = d.shape
    dm = d.astype(np.float64) * 0.001
    u, v = np.meshgrid(np.arange(W), np.arange(H))
    pe = np.stack([(u + 0.5 - CAM.cx) / CAM.fx * dm, (CAM.cy - v - 0.5) / CAM.fy * dm, -dm, np.ones_like(dm)], -1)
    pw = pe @ np.asarray(e2w, np.float64).reshape(4, 4)
    out = np.zeros((H, W, 4), np.float32)
    out[..., :3] = pw[..., :3]
    out[..., 3] = 1.0
    out[d == 0] = 0
    return out


def _rigid(ax, ang_deg, t):
    ax = np.asarray(ax, float) / np.linalg.norm(ax)
    a = np.radians(ang_deg)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    T = np.eye(4)
    T[:3, :3] = np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K
    T[:3, 3] = t
    return T


def _rot_err(a, b):
    A = np.asarray(a, np.float64).reshape(4, 4)
    B = np.asarray(b, np.float64).reshape(4, 4)
    return np.abs(A[:3, :3] - B[:3, :3]).max(), np.linalg.norm(A[3, :3] - B[3, :3])


def test_reference_align_exact_model_fixed_point_and_convergence():
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    p0 = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 0)
    d = mf.render(sc, CAM, p0, noise_seed=1, noise_k=0.0, invalid_frac=0.0)
    model = _backproject_world(d, p0)
    me = p0.astype(np.float32)
    out, n = orc.icp_align_ref(d, CAM, model, me, p0.astype(np.float32))
    assert n == 512 // 4 * 424                      # every 4th column of every row (ICP.cpp:66-72)
    r, t = _rot_err(out, p0)
    assert r < 1e-5 and t < 1e-5
    E = p0.reshape(4, 4).T
    g = (E @ _rigid([0.3, 1, 0.2], 0.5, [0.01, 0, 0])).T.reshape(16)
    e0 = mf.pose_error(g, p0)
    o4, _ = orc.icp_align_ref(d, CAM, model, me, g.astype(np.float32), iters=4)
    e4 = mf.pose_error(o4.astype(np.float64), p0)
    assert e4[0] < 0.7 * e0[0] and e4[1] < 0.7 * e0[1], (e0, e4)
    o12, _ = orc.icp_align_ref(d, CAM, model, me, g.astype(np.float32), iters=12)
    r, t = _rot_err(o12, p0)
    assert r < 1e-4 and t < 1e-4


# ------------------------------------------------------------------ normal equations
def _plane_maps(z=2.5):
    """Model maps of the plane z = -z seen head-on: each pixel's centre ray hits the plane; normal +z."""
    H, W = CAM.height, CAM.width
    u, v = np.meshgrid(np.arange(W) + 0.5, np.arange(H) + 0.5)
    pts = np.zeros((H, W, 4), np.float32)
    pts[..., 0] = (u - CAM.cx) / CAM.fx * z
    pts[..., 1] = (CAM.cy - v) / CAM.fy * z
    pts[..., 2] = -z
    pts[..., 3] = 1.0
    nrm = np.zeros_like(pts)
    nrm[..., 2] = 1.0
    nrm[..., 3] = 1.0
    return pts, nrm


def _tri(i, j):
    return i * (i + 1) // 2 + j


@pytest.mark.parametrize("dz", [0.0, 0.01])
def test_icp_sums_plane_closed_form(dz):
    """A flat frame at 2.5 m against the plane's own model, camera moved by dz
    along its axis: every residual is r = -dz, J = (p_y, -p_x, 0, 0, 0, 1), and
    the sums follow from the pixels' back-projections (w = 1/2.5^4)."""
    pts, nrm = _plane_maps()
    d = np.full((CAM.height, CAM.width), 2500, np.uint16)
    pose = np.eye(4)
    pose[2, 3] = dz
    e2w = pose.T.reshape(16)
    I = np.eye(4).reshape(16)
    s64 = orc.icp_sums_f64(d, CAM, pts, nrm, I, e2w)
    s32 = orc.icp_sums_f32terms(d, CAM, pts, nrm, I, e2w)
    # brute force over the pixels: back-project, keep those projecting into the image
    u, v = np.meshgrid(np.arange(CAM.width) + 0.5, np.arange(CAM.height) + 0.5)
    px = (u - CAM.cx) / CAM.fx * 2.5
    py = (CAM.cy - v) / CAM.fy * 2.5
    fu = CAM.fx * px / (2.5 - dz) + CAM.cx
    fv = CAM.cy - CAM.fy * py / (2.5 - dz)
    ok = (fu >= 0) & (fu < CAM.width) & (fv >= 0) & (fv < CAM.height)
    n, w = ok.sum(), 1.0 / 2.5 ** 4
    want = np.zeros(30)
    want[_tri(0, 0)] = w * (py[ok] ** 2).sum()
    want[_tri(1, 0)] = -w * (px[ok] * py[ok]).sum()
    want[_tri(1, 1)] = w * (px[ok] ** 2).sum()
    want[_tri(5, 0)] = w * py[ok].sum()
    want[_tri(5, 1)] = -w * px[ok].sum()
    want[_tri(5, 5)] = w * n
    want[21 + 0] = -dz * w * py[ok].sum()
    want[21 + 1] = dz * w * px[ok].sum()
    want[21 + 5] = -dz * w * n
    want[27], want[28], want[29] = n, n * dz * dz, w * n * dz * dz
    scale = np.abs(want).max()
    assert s64[27] == n == s32[27]
    assert np.abs(s64 - want).max() <= 1e-9 * scale, np.abs(s64 - want)
    # binary32 terms: the same sums to binary32 rounding of each term
    assert np.abs(s32 - want).max() <= 1e-6 * scale, np.abs(s32 - want)


def test_icp_sums_f32terms_agree_with_f64_on_a_room_frame(room_model):
    sc, p0, w2e, pts, nrm = room_model
    pt = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 3)
    d = mf.render(sc, CAM, pt, noise_seed=8)
    for level in (0, 2):
        s64 = orc.icp_sums_f64(d, CAM, pts, nrm, w2e.astype(np.float64), pt, level=level)
        s32 = orc.icp_sums_f32terms(d, CAM, pts, nrm, w2e, pt, level=level)
        assert s64[27] > 1000 and abs(s32[27] - s64[27]) <= 1e-4 * s64[27]
        diag = np.array([s64[_tri(i, i)] for i in range(6)])
        for i in range(6):
            for j in range(i + 1):
                assert abs(s32[_tri(i, j)] - s64[_tri(i, j)]) <= 1e-3 * np.sqrt(diag[i] * diag[j])
        # one iteration of the FP64 solver from the sums is the oracle's own step
        A = np.zeros((6, 6))
        for i in range(6):
            for j in range(i + 1):
                A[i, j] = A[j, i] = s64[_tri(i, j)]
        xi = np.linalg.solve(A, s64[21:27])
        if level == 0:
            pose1, it, _ = orc.icp_f64(d, CAM, pts, nrm, w2e.astype(np.float64), pt, iters=(1,))
            P = np.asarray(pt, np.float64).reshape(4, 4).T
            th = np.linalg.norm(xi[:3])
            K = np.array([[0, -xi[2], xi[1]], [xi[2], 0, -xi[0]], [-xi[1], xi[0], 0]])
            R = np.eye(3) + np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K
            D = np.eye(4)
            D[:3, :3], D[:3, 3] = R, xi[3:]
            assert it == 1 and np.abs((D @ P).T.reshape(16) - pose1).max() < 1e-9
