"""FMM (fmm2d_gpu) parity with the float64 direct Oseen sum of the oracle,
and with the reference's own compiled CPU FMM.

Tolerances (relative L2 over qr and qi):
  tol == 0 direct path:  F64 <= 1e-12, F32C <= 1e-6
  FMM F64, p=30, ws=2:   <= 1e-12
  FMM F64/F32C, p=16:    <= 1e-6  (the mixed-precision contract)
"""
import math

import numpy as np
import pytest

from phyflow_amd import workloads as W
from phyflow_amd.cpp.blobs.kernels import gauss

pytestmark = pytest.mark.gpu
F64, F32C = gauss.PREC_F64, gauss.PREC_F32C


def _sys(n, seed=0):
    s = W.uniform_system(n, seed=seed)
    xopt, cutoff = W.oseen_params(s["sigma"])
    return s["x"], s["y"], s["w"], xopt, cutoff


@pytest.mark.parametrize("prec,tol", [(F64, 1e-12), (F32C, 1e-6)])
def test_tol0_is_direct(prec, tol, O):
    x, y, w, xopt, cutoff = _sys(5000, 1)
    ref = O.direct_oseen(x, y, w, x, y, xopt, cutoff)
    got = gauss.fmm2d_gpu(x, y, w, x, y, 35, xopt, cutoff, 0.0, prec=prec)
    assert O.rel_l2(got, ref) <= tol


@pytest.mark.parametrize("prec,tol", [(F64, 1e-12), (F32C, 1e-6)])
def test_small_cutoff(prec, tol, O):
    """cutoff = xopt: the smoothing scale 1/(1 - exp(-alpha)) ~ 1.4 differs
    from 1 in float32 and pairs lie on both sides of the cutoff, so the
    kernel's cutoff branch decides (direct path and FMM near field)."""
    x, y, w, xopt, _ = _sys(6000, 11)
    cutoff = xopt
    ref = O.direct_oseen(x, y, w, x, y, xopt, cutoff)
    got = gauss.fmm2d_gpu(x, y, w, x, y, 35, xopt, cutoff, 0.0, prec=prec)
    assert O.rel_l2(got, ref) <= tol
    got = gauss.fmm2d_gpu(x, y, w, x, y, 35, xopt, cutoff, 1e-6, prec=prec, order=20, leaf=32)
    assert O.rel_l2(got, ref) <= 1e-6
    # the cutoff matters at this size: the default cutoff gives another answer
    far = O.direct_oseen(x, y, w, x, y, xopt, 5.0 * xopt)
    assert O.rel_l2(far, ref) > 1e-3


@pytest.mark.parametrize("prec,order,ws,leaf,tol", [
    (F64, 16, 1, 64, 1e-6), (F64, 30, 2, 32, 1e-12), (F32C, 16, 1, 64, 1e-6),
    (F32C, 24, 1, 16, 1e-6), (F32C, 8, 1, 64, 1e-4), (F64, 16, 2, 16, 1e-9)])
def test_fmm_vs_direct_uniform(prec, order, ws, leaf, tol, O):
    n = 32768
    x, y, w, xopt, cutoff = _sys(n, 2)
    sub = W.target_subset(n, 3000, seed=2)
    ref = O.direct_oseen(x, y, w, x[sub], y[sub], xopt, cutoff)
    qr, qi = gauss.fmm2d_gpu(x, y, w, x, y, 35, xopt, cutoff, 1e-6, prec=prec, order=order,
                             ws=ws, leaf=leaf)
    err = O.rel_l2((qr[sub], qi[sub]), ref)
    assert err <= tol, err


@pytest.mark.parametrize("prec", [F64, F32C])
def test_fmm_separate_targets_and_annulus(prec, O):
    a = W.annulus_blobs(20000, sigma=0.01, seed=3)
    xopt, cutoff = W.oseen_params(a["sigma"])
    rng = np.random.default_rng(4)
    xt, yt = rng.uniform(-2.0, 2.0, (2, 2500))  # includes targets outside the blobs' box
    ref = O.direct_oseen(a["x"], a["y"], a["w"], xt, yt, xopt, cutoff)
    got = gauss.fmm2d_gpu(a["x"], a["y"], a["w"], xt, yt, 35, xopt, cutoff, 1e-6, prec=prec,
                          order=20, leaf=32)
    assert O.rel_l2(got, ref) <= 1e-6


@pytest.mark.parametrize("prec", [F64, F32C])
def test_fmm_edge_cases(prec, O):
    x, y, w, xopt, cutoff = _sys(3000, 5)
    # a single target, targets at blob positions with duplicates
    for xt, yt in ((x[:1], y[:1]), (np.repeat(x[:5], 3), np.repeat(y[:5], 3))):
        ref = O.direct_oseen(x, y, w, xt, yt, xopt, cutoff)
        got = gauss.fmm2d_gpu(x, y, w, xt, yt, 35, xopt, cutoff, 1e-6, prec=prec, order=20)
        assert O.rel_l2(got, ref) <= 1e-6
    # no sources
    qr, qi = gauss.fmm2d_gpu(np.empty(0), np.empty(0), np.empty(0), x[:7], y[:7], 35, xopt,
                             cutoff, 1e-6, prec=prec)
    assert np.all(qr == 0) and np.all(qi == 0)
    # fewer points than a leaf: pure near field, matches direct to rounding
    ref = O.direct_oseen(x[:40], y[:40], w[:40], x[:40], y[:40], xopt, cutoff)
    got = gauss.fmm2d_gpu(x[:40], y[:40], w[:40], x[:40], y[:40], 35, xopt, cutoff, 1e-6,
                          prec=prec)
    assert O.rel_l2(got, ref) <= (1e-13 if prec == F64 else 1e-6)


def test_fmm_deterministic():
    x, y, w, xopt, cutoff = _sys(20000, 6)
    a = gauss.fmm2d_gpu(x, y, w, x, y, 35, xopt, cutoff, 1e-6, prec=F32C)
    b = gauss.fmm2d_gpu(x, y, w, x, y, 35, xopt, cutoff, 1e-6, prec=F32C)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_fmm_at_reference_tol_is_as_accurate_as_reference_fmm(O):
    # the reference's own CPU FMM (oracle/_ref) at tol=1e-6 on the same input
    if not O.ref_available():
        pytest.skip("reference CPU FMM not built")
    x, y, w, xopt, cutoff = _sys(65536, 7)
    sub = W.target_subset(x.size, 2000, seed=7)
    ref = O.direct_oseen(x, y, w, x[sub], y[sub], xopt, cutoff)
    rq = O.ref_fmm2d(x, y, w, x, y, 35, xopt, cutoff, 1e-6)
    e_ref = O.rel_l2((rq[0][sub], rq[1][sub]), ref)
    for prec in (F64, F32C):
        q = gauss.fmm2d_gpu(x, y, w, x, y, 35, xopt, cutoff, 1e-6, prec=prec)
        e = O.rel_l2((q[0][sub], q[1][sub]), ref)
        # float64: within 10x of the reference FMM's own error; mixed: the
        # float32 floor (~1e-7) is inside the 1e-6 contract
        assert e <= (10 * e_ref if prec == F64 else 1e-6), (e, e_ref)
