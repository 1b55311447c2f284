"""Full-size golden fingerprints (tests/golden/fullsize_golden.json), written by
the CPU oracle here with `tests/golden/make_golden.py --full`, reproduced bit
for bit by the HIP path at the sizes the bench runs:

- the 200-frame room trajectory integrated at its true poses into the 4 mm
  binary32 volume (about 20 M voxels), and the frame-199 mesh;
- the 2 mm, 64-bit-key, binary16 + u8 precision scene grown past 20 M voxels
  by the bench's seeded views, and its mesh (every voxel updated many times:
  the reciprocal-table update at every weight).

Keys, values (raw bits) and mesh vertices are compared by sha256, plus counts
and sampled entries.  The oracle side is not re-run by the CPU suite (minutes
of single-core time); the fixture's generator and this test share the frame
sequence (make_golden.full_frames)."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as G  # noqa: E402
import merge_fusion_amd as mf  # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(HERE, "golden", "fullsize_golden.json")))["cases"]
SPECS = {name: spec for name, *spec in G.FULL_CASES}


@pytest.mark.parametrize("case", GOLD, ids=[c["name"] for c in GOLD])
def test_hip_reproduces_fullsize_golden(ctx, case):
    spec = SPECS[case["name"]]
    (res, side, trunc), kb, storage = spec[6], spec[7], spec[8]
    v = mf.Volume(ctx, res, side, trunc, key_bits=32 if kb == 10 else 64, storage=storage)
    fr = -1
    for fr, (cam, d, w2e) in enumerate(G.full_frames(spec)):
        v.integrate(d, cam, w2e)
        if G.full_stop(spec, fr, v.size()):
            break
    keys, vals = v.download()
    raw = np.stack([vals["sum"], vals["weight"]], 1) if vals.dtype.names else vals
    got = G.fingerprint(keys, raw, v.mesh(), fr + 1)
    v.close()
    for k in ("frames", "voxels", "keys_sha256", "samples", "values_sha256", "mesh_vertices", "mesh_sha256"):
        assert got[k] == case[k], k
