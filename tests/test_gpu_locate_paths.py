"""GPU parity of the volume lookup's three paths (k_locate).

The sorted candidates of a frame are looked up in the volume chunk by chunk
(2048 candidates).  A chunk's ranks lie in one stretch of the volume; the
kernel reads that stretch into LDS when it holds at most 8192 keys (4096 for
64-bit keys), reads its every-16th keys into LDS in pages when it holds up to
eight pages' worth, and otherwise searches the sampled levels per candidate.
Dense frames of a grown volume almost always take the first path, so here a
large random volume is uploaded (its samples are then built once) and sparse
frames whose candidates fit one chunk are integrated, each spread so that its
stretch needs one of the three paths; the stretch is computed on the CPU from
the oracle's own candidates and checked before the frame is integrated.  Keys
and values must equal the oracle's bit for bit after every frame.

Reference: util::set_union (algorithm.h:138-198) as flat_map::insert uses it
(flat_map.h:164-197), fed by Integrator::Integrate (Integrator.cpp:26-71).
"""
import numpy as np
import pytest

import merge_fusion_amd as mf
import oracle as orc

from test_gpu_integrate import CAM, assert_volume_equal

pytestmark = pytest.mark.gpu

N_VOXELS = 4_000_000


def _random_volume(rng, res, key_bits, storage):
    """N_VOXELS distinct in-range keys (z-major packing) with random weighted values."""
    kb = 10 if key_bits == 32 else 21
    coords = rng.integers(0, res, size=(int(N_VOXELS * 1.05), 3), dtype=np.int64)
    keys = np.unique((coords[:, 2] << (2 * kb)) | (coords[:, 1] << kb) | coords[:, 0])[:N_VOXELS]
    keys = np.sort(keys)
    n = keys.size
    w = rng.integers(1, 20, size=n)
    dist = rng.uniform(-0.02, 0.02, size=n).astype(np.float32)
    if storage == mf.STORAGE_F32:
        ovals = np.stack([dist * w.astype(np.float32), w.astype(np.float32)], 1).astype(np.float32)
    else:
        ovals = (dist.astype(np.float16).view(np.uint16).astype(np.uint32) | (w.astype(np.uint32) << 16))
    return keys.astype(np.uint64), ovals


def _gpu_values(gv, ovals, storage):
    if storage == mf.STORAGE_F32:
        g = np.zeros(len(ovals), gv.val_dtype)
        g["sum"], g["weight"] = ovals[:, 0], ovals[:, 1]
        return g
    return ovals.astype(np.uint32)


def _sparse_frame(rng, region, depth_mm, count=200):
    """count valid pixels in region (v0, v1, u0, u1) at depths drawn from depth_mm (lo, hi)."""
    d = np.zeros((CAM.height, CAM.width), np.uint16)
    v0, v1, u0, u1 = region
    vs = rng.integers(v0, v1, size=count)
    us = rng.integers(u0, u1, size=count)
    d[vs, us] = rng.integers(depth_mm[0], depth_mm[1] + 1, size=count).astype(np.uint16)
    return d


def _stretch(vol_keys, d, ov):
    """Volume keys between the frame's smallest and largest candidate, and the candidate count."""
    e2w = np.eye(4, dtype=np.float32).reshape(16)
    pix = orc.keygen(d.astype(np.float32) * 0.001, CAM, e2w, ov.desc)
    cand = orc.expand(pix, ov.kb)
    lo = np.searchsorted(vol_keys, cand.min(), side="right")
    hi = np.searchsorted(vol_keys, cand.max(), side="right")
    return int(hi - max(lo - 1, 0)), len(pix)


@pytest.mark.parametrize("key_bits,storage", [(32, mf.STORAGE_F32), (32, mf.STORAGE_F16U8),
                                              (64, mf.STORAGE_F32), (64, mf.STORAGE_F16U8)])
def test_lookup_paths_match_the_oracle(ctx, key_bits, storage):
    rng = np.random.default_rng(41 + key_bits + storage)
    res, side, trunc = (1024, 8.192, 0.04) if key_bits == 32 else (2048, 8.192, 0.02)
    win = 8192 if key_bits == 32 else 4096
    gv = mf.Volume(ctx, res, side, trunc, key_bits, storage)
    ov = orc.Volume(res, side, trunc, 10 if key_bits == 32 else 21, storage)
    keys, ovals = _random_volume(rng, res, key_bits, storage)
    ov.write(keys, ovals)
    gv.upload(keys.astype(gv.key_dtype), _gpu_values(gv, ovals, storage))
    assert_volume_equal(gv, ov)
    w2e = np.eye(4, dtype=np.float32).reshape(16)
    # margins of 64 keys: the kernel's stretch rounds both ends out to its
    # every-16th samples
    frames = [
        # one depth, a small patch: at most two z-slices, the LDS window
        ("window", _sparse_frame(rng, (200, 208, 250, 258), (2000, 2000)), lambda s: s <= win - 64),
        # a 0.72 m depth band: two to eight pages of every-16th keys
        ("pages", _sparse_frame(rng, (150, 270, 200, 320), (2000, 2720)),
         lambda s: 16 * win + 64 < s <= 128 * win - 64),
        # the whole image and depth range: wider than eight pages
        ("search", _sparse_frame(rng, (0, CAM.height, 0, CAM.width), (600, 4000)),
         lambda s: s > 128 * win + 64),
    ]
    for name, d, in_range in frames:
        span, npix = _stretch(keys, d, ov)
        assert 0 < npix and 8 * npix <= 2048, (name, npix)   # one chunk of candidates
        assert in_range(span), (name, span)
        gv.integrate(d, CAM, w2e)
        ov.integrate(d, CAM, w2e)
        assert_volume_equal(gv, ov)
        keys = ov.read()[0]
