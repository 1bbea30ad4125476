"""The reference's Python layers reach the native kernels through fixed
attribute paths after ``from pHyFlow.cpp.blobs import kernels`` (_induced.py:32),
``from pHyFlow.cpp.panels import kernels`` (panelSolver.py:26) and
``from pHyFlow.cpp.blobs.remesh import *`` (_regridding.py:32).  Each path
listed below is taken from the reference call site named beside it, with
the positional arity used there.  The CPU test resolves every path and binds
the call's positional arguments; the GPU test calls through the paths and
checks the results against the oracle."""
import inspect

import numpy as np
import pytest

from phyflow_amd.cpp import blobs as cpp_blobs
from phyflow_amd.cpp import panels as cpp_panels

# (package, dotted path below it, positional argument count, reference call site)
BLOB_PATHS = [
    ("kernels.gauss.fmm2d_cpu", 9, "_induced.py:249"),
    ("kernels.gauss.fmm2d_gpu", 9, "_induced.py:334"),
    ("kernels.gauss.vv2parGauss_C_cpu", 6, "_induced.py:401"),
    ("kernels.gauss.vv2parGauss_gpu", 7, "_induced.py:515"),
    ("kernels.cutoff.vv2par_C_cpu", 6, "_induced.py:585"),
    ("kernels.cutoff.vv2par_gpu", 7, "_induced.py:695"),
    ("kernels.gauss.ww2parGauss_C_cpu", 6, "_induced.py:862"),
    ("kernels.gauss.ww2parGauss_gpu", 7, "_induced.py:963"),
    ("kernels.cutoff.ww2par_C_cpu", 6, "_induced.py:1023"),
    ("kernels.cutoff.ww2par_gpu", 7, "_induced.py:1120"),
    ("kernels.gauss.vv2parGauss_Cython_cpu", 2, "gauss/__init__.py:26"),
    ("kernels.cutoff.vv2par_Cython_cpu", 6, "cutoff/__init__.py:25"),
    ("remesh.m4kernel_1d", 2, "_regridding.py:383"),
]
PANEL_PATHS = [
    ("kernels.vortex._vortexPanel_Cython_cpu.assemble_influenceMatrix", 8, "panelSolver.py:130"),
    ("kernels.vortex._vortexPanel_Cython_cpu.inducedVelocity", 9, "panelSolver.py:249"),
]


def _resolve(root, dotted):
    obj = root
    for part in dotted.split("."):
        obj = getattr(obj, part)
    return obj


@pytest.mark.parametrize("root,dotted,nargs,site",
                         [(cpp_blobs, *p) for p in BLOB_PATHS] + [(cpp_panels, *p) for p in PANEL_PATHS],
                         ids=[p[0] for p in BLOB_PATHS + PANEL_PATHS])
def test_reference_attribute_path_resolves(root, dotted, nargs, site):
    fn = _resolve(root, dotted)
    assert callable(fn), (dotted, site)
    inspect.signature(fn).bind(*range(nargs))


def test_remesh_star_import_exports_m4kernel_1d():
    ns = {}
    exec("from phyflow_amd.cpp.blobs.remesh import *", ns)
    assert callable(ns["m4kernel_1d"])


def test_cython_gauss_rejects_wrong_shape():
    with pytest.raises(ValueError):
        cpp_blobs.kernels.gauss.vv2parGauss_Cython_cpu(np.zeros((4, 2)), 1.0)


def _blobs(n, seed):
    rng = np.random.default_rng(seed)
    x, y = rng.random(n), rng.random(n)
    h = 1.0 / np.sqrt(n)
    return x, y, rng.standard_normal(n) * h * h, h / 0.8


@pytest.mark.gpu
def test_blob_paths_against_oracle(O):
    kernels = cpp_blobs.kernels
    x, y, w, sigma = _blobs(3000, 11)
    rng = np.random.default_rng(12)
    xt, yt = rng.random(700), rng.random(700)
    ks = 2.0 * sigma * sigma
    ref = O.vv2par_gauss(x, y, w, xt, yt, ks)
    for fn in (kernels.gauss.vv2parGauss_C_cpu,):
        assert O.rel_l2(fn(x, y, w, xt, yt, ks), ref) <= 1e-12
    assert O.rel_l2(kernels.gauss.vv2parGauss_gpu(x, y, w, xt, yt, ks, 128), ref) <= 1e-12
    # the Cython variant: blobs (N, 3), targets = blob centres, result (N, 2)
    v = kernels.gauss.vv2parGauss_Cython_cpu(np.stack((x, y, w), axis=1), ks)
    refc = O.vv2par_gauss(x, y, w, x, y, ks)
    assert v.shape == (x.size, 2)
    assert O.rel_l2((v[:, 0], v[:, 1]), refc) <= 1e-12
    # the FMM, CPU and GPU names, at tol 0 (the direct sum) and at p = 30
    xopt = np.sqrt(1.2564312086261696770 * ks)
    cutoff = 5.0 * xopt
    qref = O.direct_oseen(x, y, w, xt, yt, xopt, cutoff)
    for fn in (kernels.gauss.fmm2d_cpu, kernels.gauss.fmm2d_gpu):
        assert O.rel_l2(fn(x, y, w, xt, yt, 35, xopt, cutoff, 0.0), qref) <= 1e-12
        assert O.rel_l2(fn(x, y, w, xt, yt, 35, xopt, cutoff, 1e-6, order=30, ws=2), qref) <= 1e-12
    # vorticity and the cutoff kernel
    wref = O.ww2par_gauss(x, y, w, xt, yt, ks)
    for fn in (kernels.gauss.ww2parGauss_C_cpu,):
        assert O.rel_l2(fn(x, y, w, xt, yt, ks), wref) <= 1e-12
    assert O.rel_l2(kernels.gauss.ww2parGauss_gpu(x, y, w, xt, yt, ks, 128), wref) <= 1e-12
    s2 = sigma * sigma
    cref = O.vv2par_cutoff(x, y, w, xt, yt, s2)
    assert O.rel_l2(kernels.cutoff.vv2par_C_cpu(x, y, w, xt, yt, s2), cref) <= 1e-12
    assert O.rel_l2(kernels.cutoff.vv2par_gpu(x, y, w, xt, yt, s2, 128), cref) <= 1e-12
    cwref = O.ww2par_cutoff(x, y, w, xt, yt, s2)
    assert O.rel_l2(kernels.cutoff.ww2par_C_cpu(x, y, w, xt, yt, s2), cwref) <= 1e-12
    assert O.rel_l2(kernels.cutoff.ww2par_gpu(x, y, w, xt, yt, s2, 128), cwref) <= 1e-12
    # m4kernel_1d through the star import
    ns = {}
    exec("from phyflow_amd.cpp.blobs.remesh import *", ns)
    u = np.linspace(-2.5, 2.5, 1001)
    for c in (0.0, 0.3):
        assert np.max(np.abs(ns["m4kernel_1d"](u, c) - O.m4kernel_1d(u, c))) <= 1e-15


@pytest.mark.gpu
def test_panel_paths_against_oracle(O):
    from phyflow_amd import workloads as W
    g = W.panel_geometry(*W.cylinder_panels(128, R=0.5))
    vp = cpp_panels.kernels.vortex._vortexPanel_Cython_cpu
    args = (g["xcp"], g["ycp"], g["xs"], g["ys"], g["xe"], g["ye"], g["cosA"], g["sinA"])
    A = vp.assemble_influenceMatrix(*args)
    R = O.panel_influence(*args)
    assert np.max(np.abs(A - R)) <= 1e-13 * np.max(np.abs(R))
    rng = np.random.default_rng(5)
    r, t = rng.uniform(0.6, 2.0, 500), rng.uniform(0, 2 * np.pi, 500)
    xe, ye = r * np.cos(t), r * np.sin(t)
    s = rng.standard_normal(128)
    got = vp.inducedVelocity(s, xe, ye, g["xs"], g["ys"], g["xe"], g["ye"], g["cosA"], g["sinA"])
    ref = O.panel_induced(s, xe, ye, g["xs"], g["ys"], g["xe"], g["ye"], g["cosA"], g["sinA"])
    assert O.rel_l2(got, ref) <= 1e-13
