"""GPU parity: Integrator::Integrate on the MI355X vs the CPU oracle.

Keys, merge result and voxel values are compared bit for bit (the HIP path
uses the reference's binary32 operation order; -ffp-contract=off).
Reference: KinectFission/kifi/Integrator.cpp:26-71.
"""
import numpy as np
import pytest

import merge_fusion_amd as mf
import oracle as orc

pytestmark = pytest.mark.gpu

CAM = mf.Intrinsics(512, 424, 365.0, 365.0)


def _vals_bits(vals, storage):
    if storage == mf.STORAGE_F32:
        if vals.dtype.names:
            return np.stack([vals["sum"], vals["weight"]], 1).view(np.uint32)
        return np.ascontiguousarray(vals, np.float32).view(np.uint32)
    return np.ascontiguousarray(vals, np.uint32)


def assert_volume_equal(gv, ov):
    gk, gvals = gv.download()
    ok, ovals = ov.read()
    assert gk.shape == ok.shape, (gk.shape, ok.shape)
    assert np.array_equal(gk.astype(np.uint64), ok), "voxel key sets differ"
    st = gv.desc.storage
    a, b = _vals_bits(gvals, st), _vals_bits(ovals, st)
    diff = np.nonzero(np.any(a.reshape(len(gk), -1) != b.reshape(len(ok), -1), axis=1))[0]
    assert diff.size == 0, "values differ at %d voxels, first %s: gpu %s oracle %s" % (
        diff.size, diff[:3], gvals[diff[:3]], ovals[diff[:3]])


def frame(kind, seed, traj, f, cam=CAM):
    sc = mf.make_scene(kind, seed)
    pose = mf.trajectory_pose(traj, seed, f)
    d = mf.render(sc, cam, pose, noise_seed=1000 * seed + f)
    return d, mf.invert_rigid(pose).astype(np.float32)


@pytest.mark.parametrize("storage", [mf.STORAGE_F32, mf.STORAGE_F16U8])
def test_plane_spheres_one_frame(ctx, storage):
    """Config 1: one 512x424 frame into an empty 8 mm volume (res 1024, side 8.192 m)."""
    d, w2e = frame(mf.SCENE_PLANE_SPHERES, 1, mf.TRAJ_STATIC, 0)
    gv = mf.Volume(ctx, 1024, 8.192, 0.04, 32, storage)
    ov = orc.Volume(1024, 8.192, 0.04, 10, storage)
    gv.integrate(d, CAM, w2e)
    ost = ov.integrate(d, CAM, w2e)
    gst = gv.stats()
    assert gst["keys"] == ost["keys"]
    assert gst["new"] == ost["new"]
    assert gst["size"] == ost["size"]
    assert_volume_equal(gv, ov)


@pytest.mark.parametrize("storage", [mf.STORAGE_F32, mf.STORAGE_F16U8])
def test_room_trajectory_frames(ctx, storage):
    """Config 2 shape: several frames of the room along the circle, 4 mm voxels."""
    gv = mf.Volume(ctx, 1024, 4.096, 0.02, 32, storage)
    ov = orc.Volume(1024, 4.096, 0.02, 10, storage)
    for f in range(4):
        d, w2e = frame(mf.SCENE_ROOM, 2, mf.TRAJ_CIRCLE, 3 * f)
        gv.integrate(d, CAM, w2e)
        ov.integrate(d, CAM, w2e)
    assert_volume_equal(gv, ov)


@pytest.mark.parametrize("storage", [mf.STORAGE_F32, mf.STORAGE_F16U8])
def test_64bit_keys(ctx, storage):
    """64-bit keys (21 bits per axis): res 2048 at 2 mm, two frames."""
    gv = mf.Volume(ctx, 2048, 4.096, 0.02, 64, storage)
    ov = orc.Volume(2048, 4.096, 0.02, 21, storage)
    for f in range(2):
        d, w2e = frame(mf.SCENE_ROOM, 3, mf.TRAJ_CIRCLE, 5 * f)
        gv.integrate(d, CAM, w2e)
        ov.integrate(d, CAM, w2e)
    assert_volume_equal(gv, ov)


def test_empty_and_all_invalid_frames(ctx):
    gv = mf.Volume(ctx, 1024, 8.192, 0.04)
    ov = orc.Volume(1024, 8.192, 0.04)
    z = np.zeros((CAM.height, CAM.width), np.uint16)
    w2e = np.eye(4, dtype=np.float32).reshape(16)
    gv.integrate(z, CAM, w2e)
    ov.integrate(z, CAM, w2e)
    assert gv.size() == 0 == ov.size()
    # one valid pixel: 7 keys (the eighth, the maximum, is dropped by the x pass)
    z[200, 300] = 1500
    gv.integrate(z, CAM, w2e)
    ov.integrate(z, CAM, w2e)
    assert gv.size() == 7 == ov.size()
    assert_volume_equal(gv, ov)


def test_small_ragged_frames(ctx):
    """Frame widths that are multiples of 4 but not of the wave or tile sizes."""
    cam = mf.Intrinsics(36, 20, 30.0, 30.0)
    gv = mf.Volume(ctx, 256, 4.0, 0.05)
    ov = orc.Volume(256, 4.0, 0.05)
    sc = mf.make_scene(mf.SCENE_PLANE_SPHERES, 9)
    for f in range(3):
        pose = mf.trajectory_pose(mf.TRAJ_STATIC, 9, 0)
        pose[12] += 0.01 * f
        d = mf.render(sc, cam, pose, noise_seed=77 + f)
        w2e = mf.invert_rigid(pose).astype(np.float32)
        gv.integrate(d, cam, w2e)
        ov.integrate(d, cam, w2e)
    assert_volume_equal(gv, ov)


def test_upload_download_roundtrip_then_integrate(ctx):
    """flat_map content survives upload/download; integration merges into it."""
    ov = orc.Volume(1024, 4.096, 0.02)
    d, w2e = frame(mf.SCENE_ROOM, 4, mf.TRAJ_CIRCLE, 0)
    ov.integrate(d, CAM, w2e)
    k, v = ov.read()
    gv = mf.Volume(ctx, 1024, 4.096, 0.02)
    gvals = np.zeros(len(k), gv.val_dtype)
    gvals["sum"], gvals["weight"] = v[:, 0], v[:, 1]
    gv.upload(k.astype(np.uint32), gvals)
    assert_volume_equal(gv, ov)
    d2, w2e2 = frame(mf.SCENE_ROOM, 4, mf.TRAJ_CIRCLE, 7)
    gv.integrate(d2, CAM, w2e2)
    ov.integrate(d2, CAM, w2e2)
    assert_volume_equal(gv, ov)


def test_update_covers_floor_v_over_4(ctx):
    """Q2: the last V mod 4 voxels of the merged map are never updated."""
    gv = mf.Volume(ctx, 1024, 8.192, 0.04)
    d, w2e = frame(mf.SCENE_PLANE_SPHERES, 5, mf.TRAJ_STATIC, 0)
    gv.integrate(d, CAM, w2e)
    k, v = gv.download()
    n = len(k)
    assert np.all(v["weight"][n // 4 * 4:] == 0)
    assert np.all(v["weight"][: n // 4 * 4] <= 1)


@pytest.mark.parametrize("key_bits,storage", [(32, mf.STORAGE_F32), (64, mf.STORAGE_F16U8)])
def test_volume_checkpoint_resume(ctx, tmp_path, key_bits, storage):
    """kifi_volume_save / kifi_volume_load: a volume saved after 3 frames and loaded
    continues exactly as the uninterrupted one (and as the oracle)."""
    res = 1024 if key_bits == 32 else 2048
    a = mf.Volume(ctx, res, 4.096, 0.02, key_bits, storage)
    ov = orc.Volume(res, 4.096, 0.02, 10 if key_bits == 32 else 21, storage)
    frames = [frame(mf.SCENE_ROOM, 4, mf.TRAJ_CIRCLE, 2 * f) for f in range(5)]
    for d, w2e in frames[:3]:
        a.integrate(d, CAM, w2e)
        ov.integrate(d, CAM, w2e)
    path = str(tmp_path / "v.kvol")
    a.save(path)
    b = mf.load_volume(ctx, path)
    assert (b.desc.res, b.desc.key_bits, b.desc.storage) == (res, key_bits, storage)
    for d, w2e in frames[3:]:
        a.integrate(d, CAM, w2e)
        b.integrate(d, CAM, w2e)
        ov.integrate(d, CAM, w2e)
    assert_volume_equal(b, ov)
    ka, va = a.download()
    kb, vb = b.download()
    assert np.array_equal(ka, kb) and np.array_equal(np.asarray(va).view(np.uint32), np.asarray(vb).view(np.uint32))
    # a damaged file is refused
    raw = bytearray(open(path, "rb").read())
    raw[40:48] = raw[48:56]          # two keys out of order
    bad = str(tmp_path / "bad.kvol")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(mf.KifiError):
        mf.load_volume(ctx, bad)


def test_overlapped_host_frames_equal_device_frames(ctx):
    """Host frames go up on the copy stream into two staging buffers in turn while
    the previous frame computes.  Twelve tracked frames from page-locked memory,
    enqueued back to back, give the same poses and volume as the same frames
    already in HBM."""
    sc = mf.make_scene(mf.SCENE_ROOM, 5)
    poses = [mf.trajectory_pose(mf.TRAJ_CIRCLE, 5, f) for f in range(12)]
    frames = np.stack([mf.render(sc, CAM, p, noise_seed=70 + i) for i, p in enumerate(poses)])
    pinned = ctx.host_empty(frames.shape, np.uint16)
    pinned[:] = frames
    dev = ctx.device_upload(frames)
    fb = CAM.width * CAM.height * 2
    w0 = mf.invert_rigid(poses[0]).astype(np.float32)
    out = []
    for host in (True, False):
        pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02)
        for i in range(len(poses)):
            if host:
                pipe.integrate(pinned[i], w0 if i == 0 else None)
            else:
                pipe.integrate_device(dev.address(i * fb), w0 if i == 0 else None)
        out.append((pipe.pose(), pipe.volume.download()))
        pipe.close()
    (pa, (ka, va)), (pb, (kb, vb)) = out
    assert np.array_equal(pa, pb) and np.array_equal(ka, kb)
    assert np.array_equal(np.asarray(va).view(np.uint32), np.asarray(vb).view(np.uint32))
    dev.free()


def test_invalid_arguments_are_refused(ctx):
    """The reference asserts (Volume.cpp:22-26, Integrator.cpp:35-36, 89-90) become
    KIFI_ERR_ARG with a message; nothing is launched and the context stays usable."""
    for res, side, trunc, kb in ((1000, 4.096, 0.02, 32), (2048, 4.096, 0.02, 32), (1024, 0.0, 0.02, 32),
                                 (1024, 4.096, 0.0, 32), (1024, 4.096, 0.02, 16)):
        with pytest.raises(mf.KifiError, match="status %d" % mf.KIFI_ERR_ARG):
            mf.Volume(ctx, res, side, trunc, kb)
    v = mf.Volume(ctx, 1024, 4.096, 0.02)
    bad = mf.Intrinsics(510, 424, 365.0, 365.0)             # width not a multiple of 4
    with pytest.raises(mf.KifiError, match="status %d" % mf.KIFI_ERR_ARG):
        v.integrate(np.zeros((424, 510), np.uint16), bad, np.eye(4, dtype=np.float32).reshape(16))
    with pytest.raises(mf.KifiError):
        v.integrate(np.zeros((400, 512), np.uint16), CAM, np.eye(4, dtype=np.float32).reshape(16))
    pipe = mf.Pipeline(ctx, CAM, 1024, 4.096, 0.02)
    with pytest.raises(mf.KifiError):
        pipe.integrate(np.zeros((424, 256), np.uint16), np.eye(4, dtype=np.float32).reshape(16))
    pipe.close()
    sc = mf.make_scene(mf.SCENE_PLANE_SPHERES, 1)
    d = mf.render(sc, CAM, np.eye(4), noise_seed=1)
    v.integrate(d, CAM, np.eye(4, dtype=np.float32).reshape(16))
    assert v.size() > 100000
    v.close()


def test_running_totals_match_the_oracle(ctx):
    """kifi_integrate_totals over four room frames against the oracle: the number
    of integrations, the sums of V0, V1, E (new voxels) and K (valid pixels),
    and A, the voxels the update changed.  With binary32 storage an update
    always raises a voxel's weight, so the oracle's A is the count of voxels
    whose (sum, weight) differ from before the frame (a new voxel from the
    default Voxel()).  A is summed on the device over many counters."""
    gv = mf.Volume(ctx, 1024, 4.096, 0.02, 32, mf.STORAGE_F32)
    ov = orc.Volume(1024, 4.096, 0.02, 10, mf.STORAGE_F32)
    t0 = gv.totals()
    exp = dict(frames=0, v_old=0, new=0, v_new=0, keys=0, updated=0)
    for f in (0, 3, 6, 9):
        d, w2e = frame(mf.SCENE_ROOM, 2, mf.TRAJ_CIRCLE, f)
        k0, v0 = ov.read()
        gv.integrate(d, CAM, w2e)
        st = ov.integrate(d, CAM, w2e)
        k1, v1 = ov.read()
        prev = np.zeros((len(k1), 2), np.uint32)                # Voxel(): sum 0, weight 0
        prev[np.searchsorted(k1, k0)] = _vals_bits(v0, mf.STORAGE_F32).reshape(len(k0), 2)
        cur = _vals_bits(v1, mf.STORAGE_F32).reshape(len(k1), 2)
        exp["frames"] += 1
        exp["v_old"] += len(k0)
        exp["v_new"] += len(k1)
        exp["new"] += st["new"]
        exp["keys"] += st["keys"]
        exp["updated"] += int(np.any(cur != prev, axis=1).sum())
    got = {k: v - t0[k] for k, v in gv.totals().items()}
    assert got == exp
    assert exp["updated"] > 0
