#!/usr/bin/env python3
"""Writes tests/golden/fusion_golden.json: fingerprints of the oracle's results
on small seeded inputs (the reference itself cannot be built or run here).

Each case records counts, sha256 of the raw little-endian arrays and a few
sampled entries, so both the oracle (CPU tests) and the HIP path (GPU tests)
can be checked against fixed vectors.  Regenerate only when the semantics of a
case change on purpose:

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import merge_fusion_amd as mf  # noqa: E402
import oracle as orc  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


CASES = [
    # name, scene, traj, frames, (W, H, f), (res, side, trunc), kb, storage
    ("plane_spheres_128x104_8mm", mf.SCENE_PLANE_SPHERES, mf.TRAJ_STATIC, [0], (128, 104, 91.25),
     (1024, 8.192, 0.04), 10, orc.STORE_F32),
    ("room_128x104_4mm_3frames", mf.SCENE_ROOM, mf.TRAJ_CIRCLE, [0, 5, 10], (128, 104, 91.25),
     (1024, 4.096, 0.02), 10, orc.STORE_F32),
    ("room_128x104_4mm_f16u8", mf.SCENE_ROOM, mf.TRAJ_CIRCLE, [0, 5], (128, 104, 91.25),
     (1024, 4.096, 0.02), 10, orc.STORE_F16U8),
    ("room_128x104_2mm_64bit", mf.SCENE_ROOM, mf.TRAJ_CIRCLE, [0], (128, 104, 91.25),
     (2048, 4.096, 0.02), 21, orc.STORE_F32),
]


def run_case(scene_kind, traj, frames, camspec, volspec, kb, storage, seed=11):
    W, H, f = camspec
    cam = mf.Intrinsics(W, H, f, f)
    sc = mf.make_scene(scene_kind, seed)
    v = orc.Volume(*volspec, kb=kb, storage=storage)
    for fr in frames:
        pose = mf.trajectory_pose(traj, seed, fr)
        d = mf.render(sc, cam, pose, noise_seed=100 + fr)
        v.integrate(d, cam, mf.invert_rigid(pose).astype(np.float32))
    return cam, v


def record(name, spec):
    cam, v = run_case(*spec)
    keys, vals = v.read()
    verts = v.mesh()
    iv, idx, stalled = v.mesh_indexed()
    pick = np.linspace(0, len(keys) - 1, 8).astype(int) if len(keys) else []
    return {
        "name": name,
        "voxels": int(len(keys)),
        "keys_sha256": sha(keys.astype(np.uint64)),
        "values_sha256": sha(vals),
        "samples": [[int(keys[i]), [int(x) for x in np.atleast_1d(vals[i].view(np.uint32))]] for i in pick],
        "mesh_vertices": int(len(verts)),
        "mesh_sha256": sha(verts),
        "indexed_vertices": int(len(iv)),
        "indices": int(len(idx)),
        "indexed_sha256": sha(iv),
        "indices_sha256": sha(idx),
        "stalled": bool(stalled),
    }


# Full-size cases, at the sizes the bench runs (minutes of oracle time; written
# to fullsize_golden.json by `make_golden.py --full`):
#  - the room trajectory (200 frames of 512x424, circle, seeds as the bench's
#    headline), integrated at its true poses into the 4 mm binary32 volume,
#    and the frame-199 mesh (Mesher::Mesh, generation order);
#  - the precision scene (plane + spheres) at 2 mm, 64-bit keys, binary16 + u8,
#    grown by seeded views about the identity (the bench's growth law) until it
#    holds at least 20 M voxels, and its mesh.
FULL_CASES = [
    # name, scene, scene seed, traj, traj seed, frames, noise seed base, (res, side, trunc), kb, storage, min voxels
    ("room_trajectory_512x424_4mm_200f", mf.SCENE_ROOM, 2, mf.TRAJ_CIRCLE, 2, 200, 1000, (1024, 4.096, 0.02), 10,
     orc.STORE_F32, 0),
    ("precision_512x424_2mm_64bit_f16u8_20M", mf.SCENE_PLANE_SPHERES, 1, mf.TRAJ_SWAY, 1, 400, 100,
     (4096, 8.192, 0.04), 21, orc.STORE_F16U8, 20_000_000),
]
FULL_CAM = (512, 424, 365.0)


def full_frames(spec):
    """(depth, worldToEye) of each frame of a full-size case, in order."""
    scene_kind, sseed, traj, tseed, n, nseed = spec[:6]
    W, H, f = FULL_CAM
    cam = mf.Intrinsics(W, H, f, f)
    sc = mf.make_scene(scene_kind, sseed)
    for fr in range(n):
        pose = mf.trajectory_pose(traj, tseed, fr)
        yield cam, mf.render(sc, cam, pose, noise_seed=nseed + fr), mf.invert_rigid(pose).astype(np.float32)


def full_stop(spec, frame, voxels):
    """A case ends after its frame count, or at a multiple of 8 frames once it holds min voxels."""
    n, minv = spec[4], spec[-1]
    return frame + 1 >= n or (minv and (frame + 1) % 8 == 0 and voxels >= minv)


def fingerprint(keys, vals, verts, frames):
    pick = np.linspace(0, len(keys) - 1, 8).astype(int)
    return {"frames": int(frames), "voxels": int(len(keys)), "keys_sha256": sha(keys.astype(np.uint64)),
            "values_sha256": sha(vals),
            "samples": [[int(keys[i]), [int(x) for x in np.atleast_1d(vals[i].view(np.uint32))]] for i in pick],
            "mesh_vertices": int(len(verts)), "mesh_sha256": sha(verts)}


def record_full(name, spec):
    res_side_trunc, kb, storage = spec[6], spec[7], spec[8]
    v = orc.Volume(*res_side_trunc, kb=kb, storage=storage)
    fr = -1
    for fr, (cam, d, w2e) in enumerate(full_frames(spec)):
        v.integrate(d, cam, w2e)
        if full_stop(spec, fr, v.size()):
            break
    keys, vals = v.read()
    return dict(name=name, **fingerprint(keys, vals, v.mesh(), fr + 1))


def main():
    if "--full" in sys.argv:
        out = {"generator": "tests/golden/make_golden.py --full", "cases": []}
        for name, *spec in FULL_CASES:
            out["cases"].append(record_full(name, spec))
            print(name, out["cases"][-1]["frames"], "frames", out["cases"][-1]["voxels"], "voxels", flush=True)
        with open(os.path.join(HERE, "fullsize_golden.json"), "w") as fh:
            json.dump(out, fh, indent=1)
        return
    out = {"generator": "tests/golden/make_golden.py", "cases": []}
    for name, *spec in CASES:
        out["cases"].append(record(name, spec))
        print(name, out["cases"][-1]["voxels"], "voxels")
    with open(os.path.join(HERE, "fusion_golden.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
