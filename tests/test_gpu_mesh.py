"""GPU parity: Mesher::Mesh (both overloads) and Renderer::Bin vs the oracle.

Mesh topology (vertex order, index buffer) is compared exactly; vertex
positions bit for bit, since both sides use the reference's binary32 order.
Reference: KinectFission/kifi/Mesher.cpp:14-45, 85-229; Renderer.cpp:39-80.
"""
import numpy as np
import pytest

import merge_fusion_amd as mf
import oracle as orc

pytestmark = pytest.mark.gpu

CAM = mf.Intrinsics(512, 424, 365.0, 365.0)


def _fused(ctx, kind, seed, traj, frames, res, side, trunc, storage=mf.STORAGE_F32, key_bits=32):
    gv = mf.Volume(ctx, res, side, trunc, key_bits, storage)
    ov = orc.Volume(res, side, trunc, 10 if key_bits == 32 else 21, storage)
    sc = mf.make_scene(kind, seed)
    for f in frames:
        pose = mf.trajectory_pose(traj, seed, f)
        d = mf.render(sc, CAM, pose, noise_seed=500 + f)
        w2e = mf.invert_rigid(pose).astype(np.float32)
        gv.integrate(d, CAM, w2e)
        ov.integrate(d, CAM, w2e)
    return gv, ov


@pytest.mark.parametrize("storage", [mf.STORAGE_F32, mf.STORAGE_F16U8])
def test_mesh_vertices(ctx, storage):
    gv, ov = _fused(ctx, mf.SCENE_PLANE_SPHERES, 1, mf.TRAJ_STATIC, [0], 1024, 8.192, 0.04, storage)
    g = gv.mesh()
    o = ov.mesh()
    assert g.shape == o.shape
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))


@pytest.mark.parametrize("storage", [mf.STORAGE_F32, mf.STORAGE_F16U8])
def test_mesh_indexed(ctx, storage):
    gv, ov = _fused(ctx, mf.SCENE_ROOM, 2, mf.TRAJ_CIRCLE, [0, 4], 1024, 4.096, 0.02, storage)
    gvtx, gidx, gst = gv.mesh_indexed()
    ovtx, oidx, ost = ov.mesh_indexed()
    assert gst == ost
    assert gvtx.shape == ovtx.shape and gidx.shape == oidx.shape
    assert np.array_equal(gvtx.view(np.uint32), ovtx.view(np.uint32))
    assert np.array_equal(gidx, oidx)
    if not gst:
        assert gidx.max() < len(gvtx)


def test_mesh_64bit_keys(ctx):
    gv, ov = _fused(ctx, mf.SCENE_PLANE_SPHERES, 3, mf.TRAJ_STATIC, [0], 2048, 8.192, 0.02, key_bits=64)
    g = gv.mesh()
    o = ov.mesh()
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    gvtx, gidx, gst = gv.mesh_indexed()
    ovtx, oidx, ost = ov.mesh_indexed()
    assert gst == ost
    assert np.array_equal(gvtx.view(np.uint32), ovtx.view(np.uint32))
    assert np.array_equal(gidx, oidx)


def test_mesh_tiny_volumes(ctx):
    """Empty volume, one voxel, and hand-built 2x2x2 cubes."""
    gv = mf.Volume(ctx, 16, 1.6, 0.1)
    assert gv.mesh().shape == (0, 3)
    ov = orc.Volume(16, 1.6, 0.1)
    rng = np.random.default_rng(3)
    for trial in range(20):
        xs, ys, zs = np.meshgrid(np.arange(3), np.arange(3), np.arange(3), indexing="ij")
        keys = np.unique((zs << 20 | ys << 10 | xs).ravel().astype(np.uint64) + rng.integers(0, 4) * 1025)
        w = rng.integers(0, 3, len(keys)).astype(np.float32)
        s = (rng.standard_normal(len(keys)) * 0.05).astype(np.float32) * np.maximum(w, 1)
        vals = np.stack([s, w], 1).astype(np.float32)
        ov.write(keys, vals)
        gvals = np.zeros(len(keys), gv.val_dtype)
        gvals["sum"], gvals["weight"] = s, w
        gv.upload(keys.astype(np.uint32), gvals)
        assert np.array_equal(gv.mesh().view(np.uint32), ov.mesh().view(np.uint32))
        a, b, c = gv.mesh_indexed()
        x, y, z = ov.mesh_indexed()
        assert c == z and np.array_equal(a.view(np.uint32), x.view(np.uint32)) and np.array_equal(b, y)


def test_bin_matches_oracle(ctx):
    gv, ov = _fused(ctx, mf.SCENE_ROOM, 5, mf.TRAJ_CIRCLE, [0], 1024, 4.096, 0.02)
    verts = ov.mesh()
    pose = mf.trajectory_pose(mf.TRAJ_CIRCLE, 5, 0)
    w2c = orc.mat_mul(orc.eye_to_clip(CAM), mf.invert_rigid(pose).astype(np.float32))
    g = ctx.bin(verts, w2c, CAM.width, CAM.height)
    o = orc.bin_points(verts, w2c, CAM.width, CAM.height)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    assert (np.abs(g).sum(-1) > 0).mean() > 0.2


def test_mesh_into_page_locked_buffer(ctx):
    """kifi_host_alloc buffers: a mesh copied into page-locked memory equals the
    oracle's, a too-small buffer reports the count and writes nothing, and a
    frame integrated from page-locked memory gives the same volume."""
    gv, ov = _fused(ctx, mf.SCENE_ROOM, 2, mf.TRAJ_CIRCLE, [0, 3], 1024, 4.096, 0.02)
    o = ov.mesh()
    small = ctx.host_empty((16, 3), np.float32)
    small[:] = 7.0
    assert gv.mesh_into(small) == len(o) and (small == 7.0).all()
    buf = ctx.host_empty((len(o) + 10, 3), np.float32)
    n = gv.mesh_into(buf)
    assert n == len(o) and np.array_equal(buf[:n].view(np.uint32), o.view(np.uint32))
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    pose = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, 6)
    d = mf.render(sc, CAM, pose, noise_seed=506)
    pinned = ctx.host_empty(d.shape, np.uint16)
    pinned[:] = d
    w2e = mf.invert_rigid(pose).astype(np.float32)
    gv.integrate(pinned, CAM, w2e)
    ov.integrate(d, CAM, w2e)
    gk, gvals = gv.download()
    ok, ovals = ov.read()
    assert np.array_equal(gk, ok)
    assert np.array_equal(np.asarray(gvals).view(np.uint32).ravel(), np.asarray(ovals).view(np.uint32).ravel())


def test_mesh_async_overlaps_later_frames(ctx):
    """kifi_mesh_async / kifi_mesh_wait: each mesh is the volume as it was when
    the call was made, although frames integrated after it (and a second
    asynchronous mesh, and a synchronous one) are queued before the wait; a
    too-small buffer reports the count and queues nothing; a pageable buffer
    works too."""
    gv, ov = _fused(ctx, mf.SCENE_ROOM, 2, mf.TRAJ_CIRCLE, [0, 3], 1024, 4.096, 0.02)
    sc = mf.make_scene(mf.SCENE_ROOM, 2)
    o1 = ov.mesh()
    small = ctx.host_empty((16, 3), np.float32)
    small[:] = 7.0
    assert gv.mesh_into_async(small) == len(o1)
    ctx.mesh_wait()
    assert (small == 7.0).all()
    buf1 = ctx.host_empty((len(o1) + 10, 3), np.float32)
    assert gv.mesh_into_async(buf1) == len(o1)
    for f in (6, 9, 12):
        pose = mf.trajectory_pose(mf.TRAJ_CIRCLE, 2, f)
        d = mf.render(sc, CAM, pose, noise_seed=500 + f)
        w2e = mf.invert_rigid(pose).astype(np.float32)
        gv.integrate(d, CAM, w2e)
        ov.integrate(d, CAM, w2e)
    o2 = ov.mesh()
    buf2 = ctx.host_empty((len(o2) + 10, 3), np.float32)
    assert gv.mesh_into_async(buf2) == len(o2)
    g3 = gv.mesh()                                   # synchronous, behind the pending copies
    assert np.array_equal(g3.view(np.uint32), o2.view(np.uint32))
    pageable = np.zeros((len(o2) + 10, 3), np.float32)
    assert gv.mesh_into_async(pageable) == len(o2)
    ctx.mesh_wait()
    assert np.array_equal(buf1[:len(o1)].view(np.uint32), o1.view(np.uint32))
    assert np.array_equal(buf2[:len(o2)].view(np.uint32), o2.view(np.uint32))
    assert np.array_equal(pageable[:len(o2)].view(np.uint32), o2.view(np.uint32))
    assert len(o2) != len(o1)


@pytest.mark.parametrize("key_bits,storage", [(32, mf.STORAGE_F32), (32, mf.STORAGE_F16U8),
                                              (64, mf.STORAGE_F32), (64, mf.STORAGE_F16U8)])
def test_mesh_wide_neighbour_windows(ctx, key_bits, storage):
    """A volume whose +z key windows exceed the LDS window (2048 keys): a sparse
    z-slice under a dense 64 x 64 block one slice up, so every tile of the
    sparse slice searches its +z neighbours in global memory.  Vertices (both
    Mesh overloads, and the asynchronous copy) equal the oracle's bit for bit."""
    rng = np.random.default_rng(7 + key_bits + storage)
    kb = 10 if key_bits == 32 else 21
    res = 1024 if key_bits == 32 else 2048
    z0, x0, y0 = 300, 400, 500
    gx, gy = np.meshgrid(np.arange(64), np.arange(64))
    dense = np.stack([gx.ravel() + x0, gy.ravel() + y0, np.full(4096, z0 + 1)], 1)
    pick = rng.choice(4096, size=1500, replace=False)
    sparse = dense[pick].copy()
    sparse[:, 2] = z0
    # a few x- and y-neighbours in the sparse slice too
    extra = sparse[:300].copy()
    extra[:150, 0] += 1
    extra[150:, 1] += 1
    xyz = np.concatenate([sparse, extra, dense]).astype(np.int64)
    keys = np.unique((xyz[:, 2] << (2 * kb)) | (xyz[:, 1] << kb) | xyz[:, 0]).astype(np.uint64)
    n = keys.size
    w = rng.integers(1, 20, size=n)
    dist = rng.uniform(-0.02, 0.02, size=n).astype(np.float32)
    if storage == mf.STORAGE_F32:
        ovals = np.stack([dist * w.astype(np.float32), w.astype(np.float32)], 1).astype(np.float32)
    else:
        ovals = (dist.astype(np.float16).view(np.uint16).astype(np.uint32) | (w.astype(np.uint32) << 16))
    gv = mf.Volume(ctx, res, 8.192, 0.04, key_bits, storage)
    ov = orc.Volume(res, 8.192, 0.04, kb, storage)
    ov.write(keys, ovals)
    if storage == mf.STORAGE_F32:
        gvals = np.zeros(n, gv.val_dtype)
        gvals["sum"], gvals["weight"] = ovals[:, 0], ovals[:, 1]
    else:
        gvals = ovals.astype(np.uint32)
    gv.upload(keys.astype(gv.key_dtype), gvals)
    o = ov.mesh()
    assert len(o) > 1000
    g = gv.mesh()
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    buf = ctx.host_empty((len(o) + 1, 3), np.float32)
    assert gv.mesh_into_async(buf) == len(o)
    ctx.mesh_wait()
    assert np.array_equal(buf[:len(o)].view(np.uint32), o.view(np.uint32))
    gvtx, gidx, gst = gv.mesh_indexed()
    ovtx, oidx, ost = ov.mesh_indexed()
    assert gst == ost and np.array_equal(gvtx.view(np.uint32), ovtx.view(np.uint32))
    assert np.array_equal(gidx, oidx)
