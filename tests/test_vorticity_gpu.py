"""Induced vorticity of Gaussian blobs (ww2parGauss) on the GPU vs the CPU
restatement of ww2parGauss_C_cpu.c.

float64: <= 1e-12 of max|w|; F32C: rel-L2 <= 1e-6.  Also the vorticity()
dispatch (targets default to the blobs, FMM and the cutoff kernel raise)
and empty inputs (the cutoff kernel is covered with the velocity tests)."""
import numpy as np
import pytest

from phyflow_amd import workloads as W
from phyflow_amd.blobs.base.induced import vorticity
from phyflow_amd.cpp.blobs.kernels.gauss import PREC_F32C, PREC_F64, ww2parGauss_gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,ne,seed", [(1, 7, 0), (1000, 3000, 1), (20000, 4096, 2)])
def test_vorticity_matches_oracle(n, ne, seed, O):
    rng = np.random.default_rng(seed)
    xb, yb = rng.uniform(-1, 1, (2, n))
    wb = rng.standard_normal(n) * 1e-3
    xt, yt = rng.uniform(-1.2, 1.2, (2, ne))
    ks = 2 * (2.0 / max(np.sqrt(n), 8)) ** 2
    ref = O.ww2par_gauss(xb, yb, wb, xt, yt, ks)
    got = ww2parGauss_gpu(xb, yb, wb, xt, yt, ks, 128)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    got32 = ww2parGauss_gpu(xb, yb, wb, xt, yt, ks, 128, prec=PREC_F32C)
    assert O.rel_l2(got32, ref) <= 1e-6


def test_vorticity_dispatch_and_edges(O):
    c = W.lamb_oseen_lattice()
    w = vorticity(c["x"], c["y"], c["w"], c["sigma"], k=c["k"], hardware=1)
    ref = O.ww2par_gauss(c["x"], c["y"], c["w"], c["x"], c["y"], c["k"] * c["sigma"] ** 2)
    assert np.max(np.abs(w - ref)) <= 1e-12 * np.max(np.abs(ref))
    w32 = vorticity(c["x"], c["y"], c["w"], c["sigma"], k=c["k"], prec=PREC_F32C)
    assert O.rel_l2(w32, ref) <= 1e-6
    e = np.empty(0)
    assert ww2parGauss_gpu(e, e, e, np.array([0.5]), np.array([0.5]), 0.01)[0] == 0.0
    assert ww2parGauss_gpu(c["x"], c["y"], c["w"], e, e, 0.01).size == 0
    with pytest.raises(Exception):
        vorticity(c["x"], c["y"], c["w"], c["sigma"], method=0)
    with pytest.raises(ValueError):
        vorticity(c["x"], c["y"], c["w"], c["sigma"], kernel=5)
