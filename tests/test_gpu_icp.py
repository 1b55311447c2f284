"""GPU parity for model prediction, ICP and Horn.

- model point/normal maps: bit-exact against the binary32 oracle restatement;
- point-to-plane ICP (binary32 compensated sums, FP64 solve): pose within
  1 mm / 0.1 deg of the FP64 oracle on identical inputs (stated tolerance);
- Horn's closed form (ICP.cpp:166-264) from exact correspondences: recovers a
  known rigid motion and agrees with the oracle's binary32 restatement.
"""
import numpy as np
import pytest

import merge_fusion_amd as mf
import oracle as orc

pytestmark = pytest.mark.gpu

CAM = mf.Intrinsics(512, 424, 365.0, 365.0)
POSE_TOL_M, POSE_TOL_DEG = 1e-3, 0.1


def _room(ctx, frames, seed=2, storage=mf.STORAGE_F32):
    sc = mf.make_scene(mf.SCENE_ROOM, seed)
    gv = mf.Volume(ctx, 1024, 4.096, 0.02, 32, storage)
    ov = orc.Volume(1024, 4.096, 0.02, 10, storage)
    for f in frames:
        pose = mf.trajectory_pose(mf.TRAJ_CIRCLE, seed, f)
        d = mf.render(sc, CAM, pose, noise_seed=10 + f)
        w2e = mf.invert_rigid(pose).astype(np.float32)
        gv.integrate(d, CAM, w2e)
        ov.integrate(d, CAM, w2e)
    return sc, gv, ov


@pytest.mark.parametrize("storage", [mf.STORAGE_F32, mf.STORAGE_F16U8])
def test_model_maps_bit_exact(ctx, storage):
    sc, gv, ov = _room(ctx, [0, 1, 2], storage=storage)
    w2e = mf.invert_rigid(mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 2)).astype(np.float32)
    m = mf.Model(ctx, CAM)
    m.predict(gv, w2e)
    gp, gn = m.download()
    op, on = orc.model_predict(ov, w2e, CAM)
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))
    assert np.array_equal(gn.view(np.uint32), on.view(np.uint32))
    assert gp[..., 3].mean() > 0.9 and gn[..., 3].mean() > 0.8


@pytest.mark.parametrize("target,eps_growth", [(3, 4.0), (6, 4.0), (9, 4.0), (6, 1.0)])
def test_icp_matches_fp64_oracle(ctx, target, eps_growth):
    """Default schedule (coarse levels converge at 4x / 16x the thresholds) and
    the same thresholds on every level."""
    sc, gv, ov = _room(ctx, [0, 1, 2])
    p2 = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 2)
    w2e = mf.invert_rigid(p2).astype(np.float32)
    m = mf.Model(ctx, CAM)
    m.predict(gv, w2e)
    pts, nrm = m.download()
    pt = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, target)
    d = mf.render(sc, CAM, pt, noise_seed=99 + target)
    prm = mf.IcpParams.default()
    assert prm.eps_level_growth == 4.0
    prm.eps_level_growth = eps_growth
    gpose, st = m.icp_align(d, p2, prm)
    opose, oit, orms = orc.icp_f64(d, CAM, pts, nrm, w2e.astype(np.float64), p2, eps_growth=eps_growth)
    dt, da = mf.pose_error(gpose, opose)
    assert dt <= POSE_TOL_M and da <= POSE_TOL_DEG, (dt, da, st, oit)
    gt_t, gt_a = mf.pose_error(gpose, pt)
    assert gt_t < 0.01 and gt_a < 0.2, (gt_t, gt_a)
    assert 1 <= st["iterations"] <= 19 and st["inliers"] > 10000 and not st["failed"]


@pytest.mark.parametrize("mu", [1e-4, 1e-3])
def test_icp_motion_prior_matches_fp64_oracle(ctx, mu):
    """The motion prior (kifi_icp_params.motion_prior) with the fast tracker's
    schedule 10/6/4 and 10/20/40 cm gates, from a start 1.2 deg / 4 cm off,
    against the FP64 oracle's prior ICP; 1 mm / 0.1 deg."""
    sc, gv, ov = _room(ctx, [0, 1, 2])
    p2 = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 2)
    w2e = mf.invert_rigid(p2).astype(np.float32)
    m = mf.Model(ctx, CAM)
    m.predict(gv, w2e)
    pts, nrm = m.download()
    pt = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 5)
    d = mf.render(sc, CAM, pt, noise_seed=77)
    P = np.asarray(pt, np.float64).reshape(4, 4).T
    ax = np.array([0.2, 1.0, -0.3]) / np.linalg.norm([0.2, 1.0, -0.3])
    a = np.radians(1.2)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    T = np.eye(4)
    T[:3, :3] = np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K
    T[:3, 3] = (0.03, -0.02, 0.015)
    start = (P @ T).T.reshape(16)
    prm = mf.IcpParams.default()
    prm.iters[0], prm.iters[1], prm.iters[2] = 10, 6, 4
    prm.level_gate_growth = 2.0
    prm.motion_prior = mu
    gpose, st = m.icp_align(d, start, prm)
    opose, oit, _ = orc.icp_f64(d, CAM, pts, nrm, w2e.astype(np.float64), start, iters=(10, 6, 4),
                                gate_growth=2.0, motion_prior=mu)
    dt, da = mf.pose_error(gpose, opose)
    assert dt <= POSE_TOL_M and da <= POSE_TOL_DEG, (dt, da, st, oit)
    assert not st["failed"] and 1 <= st["iterations"] <= 20
    # and without the prior the same start ends elsewhere (the prior is applied)
    prm.motion_prior = 0.0
    free, _ = m.icp_align(d, start, prm)
    assert max(mf.pose_error(free, gpose)[0] / POSE_TOL_M, mf.pose_error(free, gpose)[1] / POSE_TOL_DEG) > 0.05


def test_icp_horn_mode_matches_reference_align(ctx):
    """KIFI_ICP_HORN = the reference's ICP::Align (4 window-search Horn steps, ICP.cpp:9-265),
    against its binary32 restatement; tolerance 1 mm / 0.1 deg (different reduction order)."""
    sc, gv, ov = _room(ctx, [0, 1, 2])
    p2 = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 2)
    w2e = mf.invert_rigid(p2).astype(np.float32)
    m = mf.Model(ctx, CAM)
    m.predict(gv, w2e)
    pts, _ = m.download()
    pt = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 3)
    d = mf.render(sc, CAM, pt, noise_seed=7)
    prm = mf.IcpParams.default()
    prm.mode = mf.ICP_HORN
    prm.levels = 1
    prm.iters[0] = 4
    prm.eps_rot = prm.eps_trans = 0.0
    gpose, st = m.icp_align(d, p2, prm)
    assert st["iterations"] == 4 and st["inliers"] > 10000
    model_e2w = orc.invert_transform(w2e)
    opose, n = orc.icp_align_ref(d, CAM, pts, model_e2w, p2.astype(np.float32))
    dt, da = mf.pose_error(gpose, opose.astype(np.float64))
    assert dt <= POSE_TOL_M and da <= POSE_TOL_DEG, (dt, da)


def _rigid(rng, ang_deg, trans):
    ax = rng.standard_normal(3)
    ax /= np.linalg.norm(ax)
    a = np.radians(ang_deg)
    K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
    R = np.eye(3) + np.sin(a) * K + (1 - np.cos(a)) * K @ K
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = trans
    return T


@pytest.mark.parametrize("n", [100, 5000, 200000])
def test_horn_closed_form(ctx, n):
    """Known rigid motion of random points (the test ICP.cpp:134-164 sketches)."""
    rng = np.random.default_rng(n)
    src = rng.random((n, 3)).astype(np.float32) * 2.0 - 1.0
    T = _rigid(rng, 7.0, [0.1, -0.05, 0.2])
    dst = (src.astype(np.float64) @ T[:3, :3].T + T[:3, 3]).astype(np.float32)
    g = mf.horn_align(ctx, src, dst).reshape(4, 4).T
    assert np.abs(g - T).max() < 2e-5
    o = orc.horn_step(src, dst).reshape(4, 4).T
    assert np.abs(g - o).max() < 2e-5


def test_pipeline_given_poses_matches_oracle_volume_and_model(ctx):
    """Pipeline::Integrate with given poses (Pipeline.cpp:32-35) integrates at
    exactly those poses; its model, splatted inside the fused voxel pass, equals
    the oracle's stand-alone model prediction bit for bit."""
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02)
    ov = orc.Volume(1024, 4.096, 0.02)
    for f in (0, 1, 2):
        pose = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, f)
        d = mf.render(sc, CAM, pose, noise_seed=10 + f)
        w2e = mf.invert_rigid(pose).astype(np.float32)
        pipe.integrate(d, w2e)
        ov.integrate(d, CAM, w2e)
    gk, gv = pipe.volume.download()
    ok, oval = ov.read()
    assert np.array_equal(gk.astype(np.uint64), ok)
    assert np.array_equal(np.stack([gv["sum"], gv["weight"]], 1).view(np.uint32), oval.view(np.uint32))
    gp, gn = pipe.model.download()
    op, on = orc.model_predict(ov, w2e, CAM)
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))
    assert np.array_equal(gn.view(np.uint32), on.view(np.uint32))


def test_pipeline_tracks_trajectory(ctx):
    """Pipeline (Pipeline.cpp:36-55 tracking branch): ICP against the previous model."""
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02)
    p0 = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 0)
    for f in range(8):
        pt = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, f)
        d = mf.render(sc, CAM, pt, noise_seed=300 + f)
        if f == 0:
            pipe.integrate(d, mf.invert_rigid(p0).astype(np.float32))
        else:
            pipe.integrate(d)
    est = pipe.pose()
    dt, da = mf.pose_error(est, mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 7))
    st = pipe.stats()
    assert dt < 0.02 and da < 0.5, (dt, da, st)
    assert st["frames"] == 7 and 7 <= st["sum_iters"] <= 7 * 19


def test_icp_align_points_is_reference_align(ctx):
    """kifi_icp_align_points is ICP::Align (ICP.cpp:9-265) on the caller's point
    buffer, fed by the reference's own chain (Pipeline.cpp:47-52): mesh vertices
    binned into a 1024 x 768 buffer with EyeToClipRH() * worldToEye; 4 Horn steps
    against the binary32 restatement, tolerance 1 mm / 0.1 deg (reduction order)."""
    sc, gv, ov = _room(ctx, [0, 1, 2])
    p2 = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 2)
    w2e = mf.invert_rigid(p2).astype(np.float32)
    verts = gv.mesh()
    synth = ctx.bin(verts, orc.mat_mul(orc.eye_to_clip(CAM), w2e), 1024, 768)
    assert (np.abs(synth).sum(-1) > 0).mean() > 0.3
    pt = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 3)
    d = mf.render(sc, CAM, pt, noise_seed=11)
    e2w = p2.astype(np.float32)
    gpose, st = mf.icp_align_points(ctx, d, CAM, e2w, synth, e2w, iterations=4)
    assert st["iterations"] == 4 and st["inliers"] > 10000 and not st["failed"]
    opose, n = orc.icp_align_ref(d, CAM, synth, e2w, e2w)
    dt, da = mf.pose_error(gpose.astype(np.float64), opose.astype(np.float64))
    assert dt <= POSE_TOL_M and da <= POSE_TOL_DEG, (dt, da)
    # and it moves the guess towards the true pose
    assert mf.pose_error(gpose.astype(np.float64), pt)[0] < mf.pose_error(p2, pt)[0]


def _rot_y(deg):
    a = np.radians(deg)
    R = np.eye(4)
    R[0, 0], R[0, 2], R[2, 0], R[2, 2] = np.cos(a), np.sin(a), -np.sin(a), np.cos(a)
    return R


def _track(ctx, frames, w2e0, prm, guesses=None):
    """Poses after every frame and the final stats of a tracking pipeline;
    guesses: {frame: function(previous pose) -> initial pose}."""
    pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02, params=prm)
    poses = []
    for f, d in enumerate(frames):
        if f == 0:
            pipe.integrate(d, w2e0)
        else:
            if guesses and f in guesses:
                pipe.set_guess(guesses[f](poses[-1]))
            pipe.integrate(d)
        poses.append(pipe.pose())
    st = pipe.stats()
    pipe.close()
    return poses, st


def test_fallback_pass_reruns_plain_icp_from_previous_pose(ctx):
    """The fallback pass (one multi-iteration launch) re-runs ICP from the
    previous pose when the first pass fails the gate, and both passes stride the
    same workgroups over the pixels as plain tracking does:
    - a start 60 deg / 1 m off loses every correspondence, the pass fails, and
      the fallback's result is kept: every pose equals plain tracking's bit for bit;
    - a gate that always holds re-runs a pass that started at the previous pose
      anyway: twice the iterations, and the first (identical) result stands;
    - a start at the previous pose under the normal gate changes nothing."""
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    truth = [mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, f) for f in range(6)]
    frames = [mf.render(sc, CAM, p, noise_seed=500 + f) for f, p in enumerate(truth)]
    w2e0 = mf.invert_rigid(truth[0]).astype(np.float32)
    plain, st_plain = _track(ctx, frames, w2e0, mf.IcpParams.default())
    assert st_plain["fallbacks"] == 0

    def far(prev):
        P = np.asarray(prev, np.float64).reshape(4, 4).T @ _rot_y(60.0)
        P[:3, 3] += (1.0, 0.0, 0.5)
        return P.T.reshape(16)

    def same(prev):
        return prev

    for gate, guesses, fallbacks, extra_iters in ((0.005, {3: far}, 1, None),
                                                  (1e-9, {2: same, 4: same}, 2, True),
                                                  (0.005, {2: same, 4: same}, 0, False)):
        prm = mf.IcpParams.default()
        prm.fallback_rms = gate
        prm.fallback_iters = 0
        poses, st = _track(ctx, frames, w2e0, prm, guesses)
        assert st["fallbacks"] == fallbacks, (gate, st)
        if extra_iters is not None:
            assert (st["sum_iters"] > st_plain["sum_iters"]) == extra_iters, (gate, st, st_plain)
        for f in range(6):
            assert np.array_equal(poses[f], plain[f]), (gate, f)


def test_first_pass_level_caps_keep_the_fallback_schedule(ctx):
    """first_iters caps the levels of the first pass (the pass from a predicted
    pose) when a fallback is configured; the fallback pass from the previous pose
    keeps iters.  Frame 2 starts at the previous pose under a gate that always
    holds, so both passes run (the per-level counts add both): uncapped, the
    first pass is the fallback's twin; capped at one iteration per level, it
    adds exactly one per level, and the fallback still runs plain tracking's."""
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    truth = [mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, f) for f in range(3)]
    frames = [mf.render(sc, CAM, p, noise_seed=900 + f) for f, p in enumerate(truth)]
    w2e0 = mf.invert_rigid(truth[0]).astype(np.float32)
    _, st_plain = _track(ctx, frames, w2e0, mf.IcpParams.default())
    prm = mf.IcpParams.default()
    prm.fallback_rms = 1e-9
    prm.fallback_iters = 0
    _, st_full = _track(ctx, frames, w2e0, prm, {2: lambda prev: prev})
    assert st_full["fallbacks"] == 1 and st_full["iters"] == [2 * i for i in st_plain["iters"]], (st_full, st_plain)
    assert st_full["iterations"] == 2 * sum(st_plain["iters"]), (st_full, st_plain)
    for l in range(3):
        prm.first_iters[l] = 1
    _, st_cap = _track(ctx, frames, w2e0, prm, {2: lambda prev: prev})
    assert st_cap["fallbacks"] == 1 and st_cap["iters"] == [i + 1 for i in st_plain["iters"]], (st_cap, st_plain)
    assert st_cap["iterations"] == sum(st_plain["iters"]) + 3, (st_cap, st_plain)


def test_fallback_pass_wait_timeout_is_reported(ctx):
    """Every wait of the one-launch fallback pass is bounded.  A spin limit of one
    poll makes the waits give up on purpose: every workgroup leaves, the frames
    complete, and the next state read reports KIFI_ERR_STATE instead of a pose.
    A new pipeline on the same context then tracks normally."""
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    truth = [mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, f) for f in range(4)]
    frames = [mf.render(sc, CAM, p, noise_seed=700 + f) for f, p in enumerate(truth)]
    w2e0 = mf.invert_rigid(truth[0]).astype(np.float32)
    prm = mf.IcpParams.default()
    prm.fallback_rms = 1e-9          # the gate always holds, so the pass runs every frame
    prm.fallback_iters = 0
    prm.spin_limit = 1
    pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02, params=prm)
    pipe.integrate(frames[0], w2e0)
    prev = truth[0].reshape(16)
    for f in (1, 2, 3):
        pipe.set_guess(prev)
        pipe.integrate(frames[f])
    with pytest.raises(mf.KifiError, match="status %d" % mf.KIFI_ERR_STATE):
        pipe.stats()
    pipe.close()
    ctx.synchronize()
    poses, st = _track(ctx, frames, w2e0, mf.IcpParams.default())
    dt, da = mf.pose_error(poses[-1], truth[-1])
    assert st["frames"] == 3 and dt < 0.02 and da < 0.5, (dt, da, st)
