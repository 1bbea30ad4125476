"""The C ABI boundary: library loads, exports what include/phyflow_amd.h
declares, validates arguments before touching the GPU, and fails loudly
when no GPU is usable.  CPU only (no compute call reaches a device)."""
import ctypes
import os
import re

import numpy as np
import pytest

from phyflow_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(phf_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    declared = _declared()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    assert declared <= set(_native.exported_symbols()), declared - set(_native.exported_symbols())


def test_argument_validation_precedes_device_use():
    from phyflow_amd.cpp.blobs.kernels import gauss
    x = np.zeros(4)
    with pytest.raises(_native.PhfError) as e:
        gauss.vv2parGauss_gpu(x, x, x, x, x, -1.0, 128)
    assert e.value.code == -1 and "ksigmasqr" in str(e.value)
    bad = np.array([0.0, np.nan, 0.0, 0.0])
    with pytest.raises(_native.PhfError) as e:
        gauss.vv2parGauss_gpu(bad, x, x, x, x, 1.0, 128)
    assert "not finite" in str(e.value)
    with pytest.raises(ValueError):
        gauss.vv2parGauss_gpu(x, x[:3], x, x, x, 1.0, 128)
    # empty target set: nothing to do, no device needed
    vx, vy = gauss.vv2parGauss_gpu(x, x, x, np.empty(0), np.empty(0), 1.0, 128)
    assert vx.size == 0 and vy.size == 0


def test_no_gpu_fails_loudly():
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    from phyflow_amd.cpp.blobs.kernels import gauss
    x = np.linspace(0, 1, 8)
    with pytest.raises(_native.PhfError) as e:
        gauss.vv2parGauss_gpu(x, x, x, x, x, 0.1, 128)
    assert e.value.code == -2
