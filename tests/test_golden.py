"""Committed golden fixtures (tests/golden/fusion_golden.json) reproduced by the
CPU oracle and, bit for bit, by the HIP path."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as G  # noqa: E402
import merge_fusion_amd as mf  # noqa: E402

GOLD = json.load(open(os.path.join(HERE, "golden", "fusion_golden.json")))["cases"]
SPECS = {name: spec for name, *spec in G.CASES}


@pytest.mark.parametrize("case", GOLD, ids=[c["name"] for c in GOLD])
def test_oracle_reproduces_golden(case):
    assert G.record(case["name"], SPECS[case["name"]]) == case


@pytest.mark.gpu
@pytest.mark.parametrize("case", GOLD, ids=[c["name"] for c in GOLD])
def test_hip_reproduces_golden(ctx, case):
    scene_kind, traj, frames, camspec, volspec, kb, storage = SPECS[case["name"]]
    W, H, f = camspec
    cam = mf.Intrinsics(W, H, f, f)
    sc = mf.make_scene(scene_kind, 11)
    v = mf.Volume(ctx, *volspec, key_bits=32 if kb == 10 else 64, storage=storage)
    for fr in frames:
        pose = mf.trajectory_pose(traj, 11, fr)
        d = mf.render(sc, cam, pose, noise_seed=100 + fr)
        v.integrate(d, cam, mf.invert_rigid(pose).astype(np.float32))
    keys, vals = v.download()
    assert len(keys) == case["voxels"]
    assert G.sha(keys.astype(np.uint64)) == case["keys_sha256"]
    raw = np.stack([vals["sum"], vals["weight"]], 1) if vals.dtype.names else vals
    assert G.sha(raw) == case["values_sha256"]
    verts = v.mesh()
    assert len(verts) == case["mesh_vertices"] and G.sha(verts) == case["mesh_sha256"]
    iv, idx, stalled = v.mesh_indexed()
    assert G.sha(iv) == case["indexed_sha256"] and G.sha(idx) == case["indices_sha256"]
    assert stalled == case["stalled"]
