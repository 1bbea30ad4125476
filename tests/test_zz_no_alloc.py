"""The allocation-free direct sum (phf_vv2par_gauss_noalloc) against the
allocating entry point phf_vv2par_gauss and the oracle.

Device buffers are made by this test through the HIP runtime the library
itself links (hipMalloc here is the caller's, not the entry point's); the
entry point receives device pointers only.  Two calls in a row, the second
with a smaller input, check that the shared scratch array carries nothing
from one call into the next.
"""
import ctypes

import numpy as np
import pytest

from phyflow_amd import _native

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    _native.lib()  # loads the runtime the library is linked against
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return hip


class DeviceArrays:
    """float64 device copies of host arrays (freed on exit)."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        if a.size:
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, n):
        out = np.empty(n)
        if n:
            assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        for p in self.ptrs:
            self.hip.hipFree(p)


def _noalloc(dev, x, y, w, xt, yt, ks, prec, stream=None):
    px, py, pw = dev.put(x), dev.put(y), dev.put(w)
    pxt, pyt = dev.put(xt), dev.put(yt)
    pu, pv = dev.put(np.full(xt.size, np.nan)), dev.put(np.full(xt.size, np.nan))
    _native.call("phf_vv2par_gauss_noalloc", px, py, pw, x.size, pxt, pyt, xt.size, ks, prec,
                 pu, pv, stream)
    if stream is not None:
        assert dev.hip.hipStreamSynchronize(stream) == 0
    return dev.get(pu, xt.size), dev.get(pv, xt.size)


def _allocating(x, y, w, xt, yt, ks, prec):
    u, v = np.empty(xt.size), np.empty(xt.size)
    _native.call("phf_vv2par_gauss", _native.ptr(x), _native.ptr(y), _native.ptr(w), x.size,
                 _native.ptr(xt), _native.ptr(yt), xt.size, ks, prec, _native.ptr(u),
                 _native.ptr(v))
    return u, v


def _case(n, m, seed):
    rng = np.random.default_rng(seed)
    x, y = rng.random(n), rng.random(n)
    h = 1.0 / np.sqrt(max(n, 1))
    w = rng.standard_normal(n) * h * h
    xt, yt = rng.random(m), rng.random(m)
    return x, y, w, xt, yt, 2.0 * (h / 0.8) ** 2


def _rel(a, b):
    a, b = np.concatenate(a), np.concatenate(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


TOL = {_native.PREC_F64: 1e-12, _native.PREC_F32C: 1e-6}


@pytest.mark.parametrize("prec", [_native.PREC_F64, _native.PREC_F32C])
def test_noalloc_matches_allocating_entry_twice(O, prec):
    hip = _hip()
    with DeviceArrays(hip) as dev:
        for n, m, seed in ((3000, 2500, 11), (700, 333, 12)):  # second call smaller
            x, y, w, xt, yt, ks = _case(n, m, seed)
            got = _noalloc(dev, x, y, w, xt, yt, ks, prec)
            ref = _allocating(x, y, w, xt, yt, ks, prec)
            orc = O.vv2par_gauss(x, y, w, xt, yt, ks)
            print("noalloc n=%d m=%d prec=%d vs entry %.3e vs oracle %.3e"
                  % (n, m, prec, _rel(got, ref), _rel(got, orc)), flush=True)
            assert np.all(np.isfinite(got[0])) and np.all(np.isfinite(got[1]))
            assert _rel(got, ref) <= TOL[prec]
            assert _rel(got, orc) <= TOL[prec]


def test_noalloc_on_caller_stream_and_empty_sources(O):
    hip = _hip()
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        with DeviceArrays(hip) as dev:
            x, y, w, xt, yt, ks = _case(1500, 1024, 21)
            got = _noalloc(dev, x, y, w, xt, yt, ks, _native.PREC_F64, stream=stream)
            assert _rel(got, O.vv2par_gauss(x, y, w, xt, yt, ks)) <= 1e-12
            # no sources: the velocity is zero
            e = np.empty(0)
            u, v = _noalloc(dev, e, e, e, xt, yt, ks, _native.PREC_F32C)
            assert np.all(u == 0.0) and np.all(v == 0.0)
    finally:
        hip.hipStreamDestroy(stream)


def test_noalloc_rejects_host_memory_and_oversize():
    hip = _hip()
    x = np.linspace(0.0, 1.0, 64)
    u = np.empty(64)
    p = _native.ptr
    with pytest.raises(_native.PhfError) as e:  # host arrays are not device memory
        _native.call("phf_vv2par_gauss_noalloc", p(x), p(x), p(x), 64, p(x), p(x), 64, 0.01,
                     _native.PREC_F64, p(u), p(u), None)
    assert e.value.code == -1 and "device memory" in str(e.value)
    with DeviceArrays(hip) as dev:
        d = dev.put(x)
        with pytest.raises(_native.PhfError) as e:  # beyond the scratch limit
            _native.call("phf_vv2par_gauss_noalloc", d, d, d, 64, d, d, (1 << 21) + 1, 0.01,
                         _native.PREC_F64, d, d, None)
        assert e.value.code == -1 and "scratch" in str(e.value)
