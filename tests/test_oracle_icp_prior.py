"""Pins of the oracle's motion prior (orc_icp_level_f64_prior, the FP64 twin of
kifi_icp_params.motion_prior), on the CPU.

The prior adds mu x inliers x (|w|^2 + |dc|^2) to the point-to-plane objective,
with w the rotation and dc the camera-centre displacement accumulated since
the pass's start.  Two independent checks:

- closed form: a flat frame at 2.5 m against the head-on plane's model, the
  camera started dz along its axis.  Each inlier weighs w = 1/2.5^4 in the
  data term and mu in the prior, so the camera's depth offset settles at
  dz mu / (w + mu) (a wrong weight scaling, a dropped inlier factor or a sign
  error of the prior gives another fraction);
- the defining objective: from a perturbed start on a room frame, the first
  and second solves equal the minimisers of
  |J xi - r|^2 + mu N |m_acc + M xi|^2, where the test derives M, the map from
  a world-frame twist to (rotation, camera-centre displacement), by finite
  differences of the exact rigid motion of the start's camera centre, and
  takes the data sums from orc_icp_sums_f64 (pinned by closed forms elsewhere).
  A wrong sign or layout of the centre's cross-product block, or an
  accumulation with the wrong sign, moves the solve by far more than the
  tolerance.
"""
import numpy as np
import pytest

import merge_fusion_amd as mf
import oracle as orc

CAM = mf.Intrinsics(512, 424, 365.0, 365.0)


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


@pytest.mark.parametrize("mu", [1e-4, 1e-3, 1e-2])
@pytest.mark.parametrize("dz", [0.01, -0.02])
def test_plane_depth_offset_closed_form(mu, dz):
    pts, nrm = _plane_maps()
    d = np.full((CAM.height, CAM.width), 2500, np.uint16)
    start = np.eye(4)
    start[2, 3] = dz
    I = np.eye(4).reshape(16)
    pose, iters, _ = orc.icp_f64(d, CAM, pts, nrm, I, start.T.reshape(16), iters=(10, 5, 4), motion_prior=mu)
    P = np.asarray(pose).reshape(4, 4).T
    w = 1.0 / 2.5 ** 4
    assert abs(P[2, 3] - dz * mu / (w + mu)) <= 1e-9, (P[2, 3], dz * mu / (w + mu))
    # nothing else moves: the frame is symmetric about the axis and the data blind in the plane
    assert np.abs(P[:3, :3] - np.eye(3)).max() <= 1e-12 and np.abs(P[:2, 3]).max() <= 1e-12
    assert iters <= 6


@pytest.fixture(scope="module")
def room_model():
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    p0 = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 0)
    w2e = mf.invert_rigid(p0).astype(np.float32)
    ov = orc.Volume(1024, 4.096, 0.02)
    ov.integrate(mf.render(sc, CAM, p0, noise_seed=1, noise_k=0.0, invalid_frac=0.0), CAM, w2e)
    pts, nrm = orc.model_predict(ov, w2e, CAM)
    return sc, p0, w2e, pts, nrm


def _exp(xi):
    """[R(w) | v] as a row-major 4x4 (the twist convention of the solver's increment)."""
    w = np.asarray(xi[:3], np.float64)
    th = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    R = np.eye(3) + (np.sin(th) / th * K + (1 - np.cos(th)) / th ** 2 * K @ K if th > 0 else K)
    T = np.eye(4)
    T[:3, :3], T[:3, 3] = R, xi[3:]
    return T


def _log(T):
    c = np.clip((np.trace(T[:3, :3]) - 1) / 2, -1, 1)
    th = np.arccos(c)
    s = np.array([T[2, 1] - T[1, 2], T[0, 2] - T[2, 0], T[1, 0] - T[0, 1]]) / 2
    f = th / np.sin(th) if th > 1e-12 else 1.0
    return np.concatenate([s * f, T[:3, 3]])


def _centre_map(start):
    """6x6 map of a world-frame twist xi to (w, dc): dc the first-order motion
    of the camera centre under exp(xi) start, by central differences."""
    c0 = np.append(start[:3, 3], 1.0)
    M = np.zeros((6, 6))
    M[:3, :3] = np.eye(3)
    h = 1e-6
    for j in range(6):
        e = np.zeros(6)
        e[j] = h
        M[3:, j] = ((_exp(e) @ c0)[:3] - (_exp(-e) @ c0)[:3]) / (2 * h)
    return M


def _sums(s):
    A = np.zeros((6, 6))
    t = 0
    for i in range(6):
        for j in range(i + 1):
            A[i, j] = A[j, i] = s[t]
            t += 1
    return A, s[21:27], s[27]


@pytest.mark.parametrize("mu", [1e-4, 3e-3])
def test_first_two_solves_minimise_the_stated_objective(room_model, mu):
    sc, p0, w2e, pts, nrm = room_model
    pt = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 4)
    d = mf.render(sc, CAM, pt, noise_seed=5, noise_k=0.0, invalid_frac=0.0)
    P = np.asarray(pt, np.float64).reshape(4, 4).T
    start = P @ _exp([0.004, -0.012, 0.006, 0.02, -0.015, 0.01])       # about 0.8 deg and 2.7 cm off
    s_col = start.T.reshape(16)
    W2E = w2e.astype(np.float64)
    M = _centre_map(start)
    # the map is the rigid motion's: rotation rows identity, centre rows [-[c]x | I]
    assert np.allclose(M[3:, 3:], np.eye(3), atol=1e-8) and np.abs(M[3:, :3]).max() > 0.1
    pose1, it1, _ = orc.icp_f64(d, CAM, pts, nrm, W2E, s_col, iters=(1,), motion_prior=mu)
    pose2, it2, _ = orc.icp_f64(d, CAM, pts, nrm, W2E, s_col, iters=(2,), motion_prior=mu)
    assert it1 == 1 and it2 == 2
    P1 = np.asarray(pose1).reshape(4, 4).T
    P2 = np.asarray(pose2).reshape(4, 4).T
    xi1, xi2 = _log(P1 @ np.linalg.inv(start)), _log(P2 @ np.linalg.inv(P1))
    A, b, n = _sums(orc.icp_sums_f64(d, CAM, pts, nrm, W2E, s_col, level=0, dist_thresh=0.1))
    want1 = np.linalg.solve(A + mu * n * M.T @ M, b)
    assert np.abs(xi1 - want1).max() <= 1e-7 * np.abs(want1).max(), (xi1, want1)
    A2, b2, n2 = _sums(orc.icp_sums_f64(d, CAM, pts, nrm, W2E, pose1, level=0, dist_thresh=0.1))
    m1 = M @ want1
    want2 = np.linalg.solve(A2 + mu * n2 * M.T @ M, b2 - mu * n2 * M.T @ m1)
    assert np.abs(xi2 - want2).max() <= 1e-6 * np.abs(want1).max(), (xi2, want2)
    # the prior is felt: without it the first solve differs
    free = np.linalg.solve(A, b)
    assert np.abs(free - want1).max() > 1e-4 * np.abs(want1).max()
