"""Pipeline state across frames and damaged volume files.

- kifi_pipeline_reset: a pipeline reset after a trajectory runs the next one
  exactly as a new pipeline does (the bench's step is a whole trajectory from
  an empty volume);
- kifi_pipeline_integrate_at: tracked as usual, fused at the caller's pose;
- the learned initial pose predicts the motion between consecutive mm frames:
  a frame tracked from the caller's guess (kifi_pipeline_set_guess) still
  becomes the network's previous frame, and a metre frame starts a new
  sequence (no prediction for the frame after it);
- kifi_volume_load checks the header before it allocates: a truncated file, a
  voxel count above res^3 or one that the file's size contradicts, and keys
  outside the volume are refused, and the context stays usable.
"""
import struct

import numpy as np
import pytest

import merge_fusion_amd as mf

pytestmark = pytest.mark.gpu

CAM = mf.Intrinsics(512, 424, 365.0, 365.0)


def _frames(n, seed=4, traj=mf.TRAJ_CIRCLE):
    sc = mf.make_scene(mf.SCENE_ROOM, seed)
    poses = [mf.trajectory_pose(traj, seed, f) for f in range(n)]
    return [mf.render(sc, CAM, p, noise_seed=40 + f) for f, p in enumerate(poses)], poses


def _run(pipe, frames, poses):
    out = []
    for f, d in enumerate(frames):
        pipe.integrate(d, mf.invert_rigid(poses[0]).astype(np.float32) if f == 0 else None)
        out.append(pipe.pose())
    k, v = pipe.volume.download()
    return np.array(out), k, np.asarray(v).view(np.uint32)


def test_reset_runs_the_next_trajectory_as_a_new_pipeline(ctx):
    frames, poses = _frames(8)
    pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02)
    first = _run(pipe, frames, poses)
    pipe.reset()
    assert pipe.volume.size() == 0
    again = _run(pipe, frames, poses)
    st = pipe.stats()
    pipe.close()
    fresh_pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02)
    fresh = _run(fresh_pipe, frames, poses)
    fresh_pipe.close()
    for a, b in ((first, again), (first, fresh)):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert st["frames"] == 2 * (len(frames) - 1)          # running totals go on across the reset
    assert sum(st["sum_level_iters"]) == st["sum_iters"]


def _relative(prev, guess):
    """prev^-1 * guess: the network's twist as a transform (eye(t) -> eye(t-1))."""
    a = np.asarray(prev, np.float64).reshape(4, 4).T
    b = np.asarray(guess, np.float64).reshape(4, 4).T
    return np.linalg.inv(a) @ b


def test_caller_guess_keeps_the_network_frame_sequence(ctx):
    """Frame 2 is tracked from the caller's guess; frame 3's predicted motion is
    still the one between frames 2 and 3 (the same as in a run without the
    caller's guess)."""
    frames, poses = _frames(4, seed=6, traj=mf.TRAJ_FAST)
    net = mf.TwistNet(ctx, 3)        # seeded initialisation: a deterministic predictor
    rel = []
    for use_guess in (False, True):
        prm = mf.IcpParams.default()
        prm.fallback_rms = 0.005
        pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02, params=prm)
        pipe.set_twistnet(net)
        for f, d in enumerate(frames):
            if f == 2 and use_guess:
                pipe.set_guess(poses[2])
            pipe.integrate(d, mf.invert_rigid(poses[0]).astype(np.float32) if f == 0 else None)
            if f == 2:
                prev = pipe.pose()
        rel.append(_relative(prev, pipe.stats()["start_pose"]))
        pipe.close()
    assert not np.allclose(rel[0], np.eye(4), atol=1e-12)          # a prediction was made
    assert np.allclose(rel[0], rel[1], rtol=0, atol=1e-9), (rel[0], rel[1])


def test_metre_frame_starts_a_new_network_sequence(ctx):
    """After a frame in metres the next mm frame has no previous mm frame: it is
    tracked from the previous pose (no prediction), as the first tracked frame
    of a sequence is."""
    frames, poses = _frames(4, seed=6, traj=mf.TRAJ_FAST)
    net = mf.TwistNet(ctx, 3)
    prm = mf.IcpParams.default()
    prm.fallback_rms = 0.005
    pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02, params=prm)
    pipe.set_twistnet(net)
    for f, d in enumerate(frames):
        depth = (d.astype(np.float32) * np.float32(0.001)).astype(np.float32) if f == 2 else d
        pipe.integrate(depth, mf.invert_rigid(poses[0]).astype(np.float32) if f == 0 else None)
        if f == 2:
            prev = pipe.pose()
    start = pipe.stats()["start_pose"]
    pipe.close()
    assert np.array_equal(np.asarray(start), np.asarray(prev))


def test_integrate_at_tracks_then_fuses_at_the_given_pose(ctx):
    """kifi_pipeline_integrate_at: the volume is the one integrated at the given
    poses (bit for bit), each frame's tracked pose is the one plain
    kifi_pipeline_integrate computes from the same state, and the next frame
    starts from the given pose."""
    frames, poses = _frames(6, seed=5)
    w2e = [mf.invert_rigid(p).astype(np.float32) for p in poses]
    pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02)
    tracked = []
    for f, d in enumerate(frames):
        pipe.integrate_at(d, w2e[f])
        tracked.append(pipe.pose())
        if f > 0:
            start = np.asarray(pipe.stats()["start_pose"])
            fused_prev = np.linalg.inv(w2e[f - 1].astype(np.float64).reshape(4, 4).T)
            assert np.allclose(start.reshape(4, 4).T, fused_prev, rtol=0, atol=1e-6), f
    k, v = pipe.volume.download()
    pipe.close()
    vol = mf.Volume(ctx, 1024, 4.096, 0.02)
    for d, w in zip(frames, w2e):
        vol.integrate(d, CAM, w)
    k2, v2 = vol.download()
    vol.close()
    assert np.array_equal(k, k2) and np.array_equal(np.asarray(v).view(np.uint32), np.asarray(v2).view(np.uint32))
    # frame 4 tracked by the plain call from the same state
    ref = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02)
    for f in range(4):
        ref.integrate_at(frames[f], w2e[f])
    ref.integrate(frames[4])
    assert np.array_equal(ref.pose(), tracked[4])
    ref.close()
    e = mf.pose_error(tracked[4], poses[4])
    assert e[0] < 0.01 and e[1] < 0.5, e
    assert not np.allclose(tracked[4], np.linalg.inv(w2e[4].astype(np.float64).reshape(4, 4).T).T.reshape(16),
                           rtol=0, atol=1e-9)


@pytest.fixture()
def saved(ctx, tmp_path):
    v = mf.Volume(ctx, 1024, 4.096, 0.02)
    frames, poses = _frames(2)
    for d, p in zip(frames, poses):
        v.integrate(d, CAM, mf.invert_rigid(p).astype(np.float32))
    path = str(tmp_path / "v.kvol")
    v.save(path)
    n = v.size()
    v.close()
    return path, n, bytearray(open(path, "rb").read())


HEADER = 40        # magic 8, version 4, res 4, side 4, trunc 4, key_bits 4, storage 4, n 8


def _refused(ctx, tmp_path, raw, match):
    bad = str(tmp_path / "bad.kvol")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(mf.KifiError, match=match):
        mf.load_volume(ctx, bad)


def test_truncated_volume_file_is_refused(ctx, tmp_path, saved):
    path, n, raw = saved
    _refused(ctx, tmp_path, raw[:-12], "file size")
    _refused(ctx, tmp_path, raw[:HEADER - 3], "header")


def test_huge_voxel_count_is_refused_before_allocating(ctx, tmp_path, saved):
    path, n, raw = saved
    for huge in (1 << 62, 1024 ** 3 + 1, (1 << 64) - 1):
        r = bytearray(raw)
        r[32:40] = struct.pack("<Q", huge)
        _refused(ctx, tmp_path, r, "res\\^3|file size")
    r = bytearray(raw)
    r[32:40] = struct.pack("<Q", n + 1)          # count one above what the file holds
    _refused(ctx, tmp_path, r, "file size")


def test_key_outside_the_volume_is_refused(ctx, tmp_path, saved):
    path, n, raw = saved
    r = bytearray(raw)
    last = HEADER + 4 * (n - 1)
    r[last:last + 4] = struct.pack("<I", 1 << 30)    # above z < 1024: bits beyond the three axes
    _refused(ctx, tmp_path, r, "outside")
    # the context stays usable and the intact file loads
    v = mf.load_volume(ctx, path)
    assert v.size() == n
    v.close()
