"""The full-size golden fixture is present and describes the cases its
generator defines (the GPU test reproduces it; the oracle run that wrote it
takes minutes of single-core time and is not repeated here)."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden as G  # noqa: E402


def test_fullsize_fixture_matches_its_cases():
    gold = json.load(open(os.path.join(HERE, "golden", "fullsize_golden.json")))
    assert gold["generator"] == "tests/golden/make_golden.py --full"
    names = [c["name"] for c in gold["cases"]]
    assert names == [n for n, *_ in G.FULL_CASES]
    for c, (name, *spec) in zip(gold["cases"], G.FULL_CASES):
        assert c["voxels"] >= spec[-1] and c["frames"] <= spec[4]
        assert len(c["keys_sha256"]) == 64 and len(c["values_sha256"]) == 64 and c["mesh_vertices"] > 0
    room, prec = gold["cases"]
    assert room["frames"] == 200 and room["voxels"] > 10_000_000
    assert prec["voxels"] >= 20_000_000
