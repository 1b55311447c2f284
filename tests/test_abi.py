"""The C ABI libraries load and export every function their headers declare.

Runs without a GPU: only symbol lookup, no compute call.
"""
import ctypes
import os
import re

import pytest

import merge_fusion_amd as mf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kifi_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("kifi.h", mf.HIP_LIB_PATH), ("kifi_synth.h", mf.SYNTH_LIB_PATH)])
def test_library_exports_every_declared_symbol(header, lib):
    assert os.path.exists(lib), "build the libraries first (python -m merge_fusion_amd.build)"
    so = ctypes.CDLL(lib)
    names = declared(header)
    assert len(names) > 3
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_python_binding_covers_the_abi():
    assert sorted(mf.HIP_EXPORTS) == declared("kifi.h")


def test_no_gpu_needed_to_load():
    lib = mf.load_hip()
    assert lib.kifi_last_error(None) == b"no context"
    p = mf.IcpParams.default()
    assert list(p.iters) == [10, 5, 4] and p.levels == 3
