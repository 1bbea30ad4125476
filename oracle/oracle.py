"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the
reference's own compiled CPU FMM (oracle/_ref/libref_fmm2d.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never by the shipped package.  Each wrapper names the reference
lines its C function restates (see oracle/phf_oracle.c).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_fmm2d.so")

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_ptr = ctypes.c_void_p

_lib = None
_ref = None


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_vv2par_gauss.argtypes = [_i64, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _dbl, _ptr, _ptr, ctypes.c_int]
        L.orc_vv2par_cutoff.argtypes = L.orc_vv2par_gauss.argtypes
        L.orc_ww2par_cutoff.argtypes = [_i64, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _dbl, _ptr, ctypes.c_int]
        L.orc_ww2par_gauss.argtypes = [_i64, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _dbl, _ptr, ctypes.c_int]
        L.orc_direct_oseen.argtypes = [_i64, _ptr, _ptr, _ptr, _i64, _ptr, _ptr, _dbl, _dbl, _ptr, _ptr, ctypes.c_int]
        L.orc_m4kernel_1d.argtypes = [_i64, _ptr, _dbl, _ptr]
        L.orc_regrid_m4.argtypes = [_i64, _ptr, _ptr, _ptr, _dbl, _dbl, _dbl, _dbl, _dbl,
                                    ctypes.POINTER(ctypes.POINTER(_dbl)), ctypes.POINTER(ctypes.POINTER(_dbl)),
                                    ctypes.POINTER(ctypes.POINTER(_dbl)), ctypes.POINTER(ctypes.POINTER(_i64)),
                                    ctypes.POINTER(ctypes.POINTER(_i64))]
        L.orc_regrid_m4.restype = _i64
        L.orc_regrid_tutty.argtypes = L.orc_regrid_m4.argtypes
        L.orc_regrid_tutty.restype = _i64
        L.orc_free.argtypes = [_ptr]
        L.orc_population_control.argtypes = [_i64, _ptr, _dbl, _dbl, _ptr, ctypes.POINTER(_dbl)]
        L.orc_population_control.restype = _i64
        L.orc_panel_influence.argtypes = [_i64] + [_ptr] * 9
        L.orc_panel_induced.argtypes = [_i64, _ptr, _i64] + [_ptr] * 10
        L.orc_lu_factor.argtypes = [_i64, _ptr, _ptr]
        L.orc_lu_factor.restype = ctypes.c_int
        L.orc_lu_solve.argtypes = [_i64, _ptr, _ptr, _ptr]
        L.orc_morton_keys.argtypes = [_i64, _ptr, _ptr, _dbl, _dbl, _dbl, ctypes.c_int, _ptr]
        L.orc_stable_sort_keys.argtypes = [_i64, _ptr, _ptr, _ptr]
        L.orc_stable_sort_keys.restype = ctypes.c_int
        L.orc_leaf_offsets.argtypes = [_i64, _ptr, _i64, _ptr]
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("reference FMM not built: run `make -C oracle ref`")
        R = ctypes.CDLL(REF_PATH)
        R.ref_fmm2d.argtypes = [ctypes.c_int, _ptr, _ptr, _ptr, ctypes.c_int, _ptr, _ptr,
                                ctypes.c_int, _dbl, _dbl, _dbl, _ptr, _ptr]
        _ref = R
    return _ref


# -- A.1 ---------------------------------------------------------------------
def vv2par_gauss(xb, yb, wb, xt, yt, ksigmasqr, nthreads=0):
    """vv2parGauss_C_cpu.c:99-153 restated; returns (vx, vy)."""
    xb, yb, wb, xt, yt = map(_f64, (xb, yb, wb, xt, yt))
    vx = np.empty(xt.size)
    vy = np.empty(xt.size)
    lib().orc_vv2par_gauss(xb.size, _p(xb), _p(yb), _p(wb), xt.size, _p(xt), _p(yt),
                           float(ksigmasqr), _p(vx), _p(vy), int(nthreads))
    return vx, vy


def vv2par_cutoff(xb, yb, wb, xt, yt, sigmasqr, nthreads=0):
    """vv2par_C_cpu.c:135-151 restated; returns (vx, vy)."""
    xb, yb, wb, xt, yt = map(_f64, (xb, yb, wb, xt, yt))
    vx = np.empty(xt.size)
    vy = np.empty(xt.size)
    lib().orc_vv2par_cutoff(xb.size, _p(xb), _p(yb), _p(wb), xt.size, _p(xt), _p(yt),
                            float(sigmasqr), _p(vx), _p(vy), int(nthreads))
    return vx, vy


def ww2par_cutoff(xb, yb, wb, xt, yt, sigmasqr, nthreads=0):
    """ww2par_C_cpu.c:132-144 restated; returns w."""
    xb, yb, wb, xt, yt = map(_f64, (xb, yb, wb, xt, yt))
    w = np.empty(xt.size)
    lib().orc_ww2par_cutoff(xb.size, _p(xb), _p(yb), _p(wb), xt.size, _p(xt), _p(yt),
                            float(sigmasqr), _p(w), int(nthreads))
    return w


def ww2par_gauss(xb, yb, wb, xt, yt, ksigmasqr, nthreads=0):
    """ww2parGauss_C_cpu.c:133-144 restated; returns w."""
    xb, yb, wb, xt, yt = map(_f64, (xb, yb, wb, xt, yt))
    w = np.empty(xt.size)
    lib().orc_ww2par_gauss(xb.size, _p(xb), _p(yb), _p(wb), xt.size, _p(xt), _p(yt),
                           float(ksigmasqr), _p(w), int(nthreads))
    return w


# -- A.2 ---------------------------------------------------------------------
def direct_oseen(xb, yb, wb, xt, yt, xopt, cutoff, nthreads=0):
    """direct.cpp:1409-1485 (Oseen, pot=1) restated; returns (qr, qi)."""
    xb, yb, wb, xt, yt = map(_f64, (xb, yb, wb, xt, yt))
    qr = np.empty(xt.size)
    qi = np.empty(xt.size)
    lib().orc_direct_oseen(xb.size, _p(xb), _p(yb), _p(wb), xt.size, _p(xt), _p(yt),
                           float(xopt), float(cutoff), _p(qr), _p(qi), int(nthreads))
    return qr, qi


def ref_fmm2d(xb, yb, wb, xt, yt, ndirect, xopt, cutoff, tol):
    """The reference's own compiled CPU fmm2d (tol == 0: its direct sum)."""
    xb, yb, wb, xt, yt = map(_f64, (xb, yb, wb, xt, yt))
    qr = np.empty(xt.size)
    qi = np.empty(xt.size)
    ref().ref_fmm2d(xb.size, _p(xb), _p(yb), _p(wb), xt.size, _p(xt), _p(yt),
                    int(ndirect), float(xopt), float(cutoff), float(tol), _p(qr), _p(qi))
    return qr, qi


# -- A.4 ---------------------------------------------------------------------
def m4kernel_1d(x, c=0.0):
    x = _f64(x).ravel()
    w = np.empty(x.size)
    lib().orc_m4kernel_1d(x.size, _p(x), float(c), _p(w))
    return w


def regrid_m4(x, y, w, sigma, overlap, x0, y0, c=0.0):
    """_regridding.py:118-218 restated; returns (x, y, w, i, j)."""
    return _regrid(lib().orc_regrid_m4, x, y, w, sigma, overlap, x0, y0, c)


def regrid_tutty(x, y, w, sigma, overlap, x0, y0, hnu):
    """diffusion_tutty, _evolution.py:325-445 restated (hnu = sqrt(nu dt));
    returns (x, y, w, i, j)."""
    return _regrid(lib().orc_regrid_tutty, x, y, w, sigma, overlap, x0, y0, hnu)


def _regrid(fn, x, y, w, sigma, overlap, x0, y0, par):
    x, y, w = map(_f64, (x, y, w))
    px = ctypes.POINTER(_dbl)()
    py = ctypes.POINTER(_dbl)()
    pw = ctypes.POINTER(_dbl)()
    pi = ctypes.POINTER(_i64)()
    pj = ctypes.POINTER(_i64)()
    m = fn(x.size, _p(x), _p(y), _p(w), float(sigma), float(overlap), float(x0), float(y0),
           float(par), ctypes.byref(px), ctypes.byref(py), ctypes.byref(pw), ctypes.byref(pi),
           ctypes.byref(pj))
    if m < 0:
        raise MemoryError("oracle regrid allocation failed")
    if m == 0:
        e = np.empty(0)
        return e, e.copy(), e.copy(), np.empty(0, np.int64), np.empty(0, np.int64)
    out = []
    for ptr, dt in ((px, np.float64), (py, np.float64), (pw, np.float64), (pi, np.int64), (pj, np.int64)):
        out.append(np.ctypeslib.as_array(ptr, shape=(m,)).astype(dt, copy=True))
        lib().orc_free(ctypes.cast(ptr, ctypes.c_void_p))
    return tuple(out)


def population_control(x, y, w, thr, total):
    """_regridding.py:313-330 restated; returns (x, y, w, threshold used)."""
    x, y, w = map(_f64, (x, y, w))
    keep = np.empty(max(w.size, 1), np.int64)
    t = _dbl()
    m = lib().orc_population_control(w.size, _p(w), float(thr), float(total), _p(keep),
                                     ctypes.byref(t))
    k = keep[:m]
    return x[k], y[k], w[k], t.value


# -- A.5 ---------------------------------------------------------------------
def panel_influence(xCP, yCP, xs, ys, xe, ye, cosA, sinA):
    a = list(map(_f64, (xCP, yCP, xs, ys, xe, ye, cosA, sinA)))
    M = a[0].size
    A = np.empty((M, M))
    lib().orc_panel_influence(M, *[_p(v) for v in a], _p(A))
    return A


def panel_induced(gamma, xE, yE, xs, ys, xe, ye, cosA, sinA):
    gamma = _f64(gamma)
    xE, yE = _f64(xE), _f64(yE)
    g = list(map(_f64, (xs, ys, xe, ye, cosA, sinA)))
    vx = np.empty(xE.size)
    vy = np.empty(xE.size)
    lib().orc_panel_induced(gamma.size, _p(gamma), xE.size, _p(xE), _p(yE), *[_p(v) for v in g],
                            _p(vx), _p(vy))
    return vx, vy


# -- dense solve ---------------------------------------------------------------
def lu_solve(A, b):
    """FP64 partial-pivot LU (dgetrf/dgetrs algorithm) solve of A x = b."""
    LU = np.array(A, dtype=np.float64, order="C", copy=True)
    n = LU.shape[0]
    piv = np.empty(n, np.int64)
    info = lib().orc_lu_factor(n, _p(LU), _p(piv))
    if info:
        raise np.linalg.LinAlgError("singular matrix (U[%d,%d] == 0)" % (info - 1, info - 1))
    x = np.array(b, dtype=np.float64, copy=True)
    lib().orc_lu_solve(n, _p(LU), _p(piv), _p(x))
    return x


# -- tree --------------------------------------------------------------------
def morton_keys(x, y, x0, y0, inv_s, L):
    x, y = _f64(x), _f64(y)
    keys = np.empty(x.size, np.uint32)
    lib().orc_morton_keys(x.size, _p(x), _p(y), float(x0), float(y0), float(inv_s), int(L), _p(keys))
    return keys


def stable_sort_keys(keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    sk = np.empty_like(keys)
    perm = np.empty(keys.size, np.int32)
    if lib().orc_stable_sort_keys(keys.size, _p(keys), _p(sk), _p(perm)):
        raise MemoryError("oracle sort allocation failed")
    return sk, perm


def leaf_offsets(skeys, nleaf):
    skeys = np.ascontiguousarray(skeys, dtype=np.uint32)
    start = np.empty(nleaf + 1, np.int32)
    lib().orc_leaf_offsets(skeys.size, _p(skeys), int(nleaf), _p(start))
    return start


# -- shared helpers ------------------------------------------------------------
def rel_l2(a, b):
    """||a-b||_2 / ||b||_2 over all components."""
    a = np.concatenate([np.ravel(v) for v in a]) if isinstance(a, (tuple, list)) else np.ravel(a)
    b = np.concatenate([np.ravel(v) for v in b]) if isinstance(b, (tuple, list)) else np.ravel(b)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def max_rel(a, b):
    """max|a-b| / max|b| over all components."""
    a = np.concatenate([np.ravel(v) for v in a]) if isinstance(a, (tuple, list)) else np.ravel(a)
    b = np.concatenate([np.ravel(v) for v in b]) if isinstance(b, (tuple, list)) else np.ravel(b)
    mb = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / (mb if mb > 0 else 1.0)) if a.size else 0.0
