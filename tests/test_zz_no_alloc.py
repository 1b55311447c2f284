"""kifi_icp_align_device_noalloc: the allocation-free variant of kifi_icp_align
(its work memory is a fixed array in the library) gives kifi_icp_align's
result bit for bit, on consecutive calls of different frame sizes, and refuses
frames above its limit without disturbing the context.
"""
import numpy as np
import pytest

import merge_fusion_amd as mf

pytestmark = pytest.mark.gpu

CAM = mf.Intrinsics(512, 424, 365.0, 365.0)
CAM_SMALL = mf.Intrinsics(256, 212, 182.5, 182.5)


@pytest.fixture(scope="module")
def room(ctx):
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    vol = mf.Volume(ctx, 1024, 4.096, 0.02)
    for f in range(3):
        pose = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, f)
        vol.integrate(mf.render(sc, CAM, pose, noise_seed=10 + f), CAM, mf.invert_rigid(pose).astype(np.float32))
    yield sc, vol
    vol.close()


def _case(ctx, sc, vol, cam, target):
    """A model rendered at frame 2's pose and a frame at `target` (the guess is frame 2's pose)."""
    p2 = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 2)
    m = mf.Model(ctx, cam)
    m.predict(vol, mf.invert_rigid(p2).astype(np.float32))
    d = mf.render(sc, cam, mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, target), noise_seed=99 + target)
    return m, d, p2


def _same(a, b):
    (pa, sa), (pb, sb) = a, b
    assert np.array_equal(pa, pb), (pa, pb)
    for k in ("iterations", "iters", "rms", "rms_1m", "inliers", "failed", "fallback"):
        assert sa[k] == sb[k], (k, sa[k], sb[k])


def test_noalloc_equals_icp_align_on_consecutive_calls(ctx, room):
    sc, vol = room
    results = []
    for cam, target in ((CAM, 6), (CAM_SMALL, 5), (CAM, 6)):   # the second call is smaller
        m, d, guess = _case(ctx, sc, vol, cam, target)
        dev = ctx.device_upload(d)
        ref = m.icp_align(d, guess)
        got = m.icp_align_device_noalloc(dev.address(), guess)
        _same(got, ref)
        assert got[1]["inliers"] > 1000 and not got[1]["failed"]
        results.append(got)
        dev.free()
        m.close()
    _same(results[0], results[2])          # nothing of the smaller call stays behind


def test_noalloc_with_fallback_equals_icp_align(ctx, room):
    """The one-launch fallback pass (workgroups waiting on each other) runs in the
    array too."""
    sc, vol = room
    m, d, guess = _case(ctx, sc, vol, CAM, 8)
    prm = mf.IcpParams.default()
    prm.fallback_rms = 1e-6                # every frame falls back
    off = guess.copy()
    off[12] += 0.05                        # the first pass starts 5 cm off
    dev = ctx.device_upload(d)
    ref = m.icp_align(d, off, prm)
    got = m.icp_align_device_noalloc(dev.address(), off, prm)
    _same(got, ref)
    assert got[1]["fallback"] == 1
    dev.free()
    m.close()


def test_noalloc_refuses_frames_above_its_limit(ctx, room):
    sc, vol = room
    big = mf.Intrinsics(1024, 768, 730.0, 730.0)
    m = mf.Model(ctx, big)
    dev = ctx.device_upload(np.zeros((768, 1024), np.uint16))
    with pytest.raises(mf.KifiError, match="status %d" % mf.KIFI_ERR_CAPACITY):
        m.icp_align_device_noalloc(dev.address(), np.eye(4))
    dev.free()
    m.close()
    m, d, guess = _case(ctx, sc, vol, CAM, 6)      # the context is still usable
    dev = ctx.device_upload(d)
    _same(m.icp_align_device_noalloc(dev.address(), guess), m.icp_align(d, guess))
    dev.free()
    m.close()
