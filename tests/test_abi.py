"""CPU tier: the C-ABI library builds for gfx950, loads, and exports every
entry point include/orb_gpu.h declares; the product never touches the oracle."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "orb_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"\b(orb_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("orb_extractor_create", "orb_extract", "orb_descriptor_distance", "orb_match_bf",
                     "orb_stereo_match", "orb_stereo_frame", "orb_extractor_pyramid_level", "orb_batch_run"):
        assert required in names


def test_library_exports_every_declared_symbol(orb):
    out = subprocess.run(["nm", "-D", "--defined-only", orb.library_path], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (orb_[a-z0-9_]+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_python_binding_covers_header(orb):
    assert set(declared_functions()) <= set(orb.SIGNATURES), set(declared_functions()) - set(orb.SIGNATURES)
    for name in declared_functions():
        assert hasattr(orb.lib(), name)


def test_library_carries_gfx950_code(orb):
    data = open(orb.library_path, "rb").read()
    assert b"gfx950" in data


def test_descriptor_distance_needs_no_device(orb):
    import numpy as np
    a = np.arange(32, dtype=np.uint8)
    b = a ^ np.uint8(0xFF)
    assert orb.ORBmatcher.DescriptorDistance(a, a) == 0
    assert orb.ORBmatcher.DescriptorDistance(a, b) == 256


def test_bad_arguments_are_rejected_without_a_device(orb):
    import ctypes
    h = ctypes.c_void_p()
    assert orb.lib().orb_extractor_create(None, 1000, 1.2, 8, 20, 7, ctypes.byref(h)) == orb.ORB_E_ARG
    assert orb.lib().orb_descriptor_distance(None, None) == orb.ORB_E_ARG


def test_product_does_not_use_the_oracle():
    pkg = os.path.join(ROOT, "orb_slam2_detailed_comments_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", ".inc")):
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert "pyoracle" not in text and "orb_oracle" not in text and "liborb_oracle" not in text, f


def test_bench_rejects_multi_gpu():
    r = subprocess.run(["python", os.path.join(ROOT, "bench.py"), "--gpus", "8"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
