"""Resident blob systems: same results as the one-shot entry points."""
import math

import numpy as np
import pytest

from phyflow_amd import workloads as W
from phyflow_amd.cpp.blobs.kernels import gauss
from phyflow_amd.system import METHOD_DIRECT, METHOD_FMM, BlobSystem, vel_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [gauss.PREC_F64, gauss.PREC_F32C])
def test_system_matches_one_shot(prec, O):
    s = W.uniform_system(40000, seed=3)
    xopt, cutoff = W.oseen_params(s["sigma"])
    with BlobSystem(s["x"], s["y"], s["w"]) as sysm:
        vx, vy = sysm.velocity(vel_params(s["sigma"], prec=prec, order=16, leaf=64), info=True,
                              count_work=True)
        info = sysm.last_info
        assert info["p2p_pairs"] > 0 and info["L"] >= 3
        qr, qi = gauss.fmm2d_gpu(s["x"], s["y"], s["w"], s["x"], s["y"], 35, xopt, cutoff, 1e-6,
                                 prec=prec, order=16, leaf=64)
        assert np.array_equal(vx, -qi / (2 * math.pi)) and np.array_equal(vy, -qr / (2 * math.pi))
        d = sysm.velocity(vel_params(s["sigma"], method=METHOD_DIRECT, prec=prec))
        ref = O.vv2par_gauss(s["x"], s["y"], s["w"], s["x"], s["y"], 2 * s["sigma"] ** 2)
        assert O.rel_l2(d, ref) <= (1e-12 if prec == gauss.PREC_F64 else 1e-6)
        # fetch later, and resize the system
        sysm.velocity(vel_params(s["sigma"], prec=prec), fetch=False)
        a = sysm.last_velocity()
        assert a[0].size == 40000
        sysm.set(s["x"][:100], s["y"][:100], s["w"][:100])
        assert sysm.n == 100
        x, y, w = sysm.get()
        assert np.array_equal(x, s["x"][:100])


@pytest.mark.parametrize("prec", [gauss.PREC_F64, gauss.PREC_F32C])
def test_graph_replay_matches_eager(prec):
    """The recorded FMM launch sequence replays bit-identically to eager
    launches: repeated replays with eager work between them, points moved
    within the same domain side (replay reads the new corner), a wider cloud
    (new recording) and a return to the first cloud (cached recording)."""
    s = W.uniform_system(30000, seed=5)
    p = vel_params(s["sigma"], prec=prec, order=16, leaf=64)

    def same(a, b):
        return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])

    with BlobSystem(s["x"], s["y"], s["w"]) as sysm:
        ref = sysm.velocity(p, eager=True)
        for _ in range(3):
            assert same(sysm.velocity(p), ref)
            sysm.velocity(p, eager=True)
        sysm.velocity(p, info=True)
        ms = sysm.last_info["ms"]
        # a replay times the tree graph and tree + far field, not the far
        # field's stages one by one
        assert "tree_far" in ms and "tree" in ms and "p2p" in ms and "m2l" not in ms
        sysm.set(s["x"] + 0.013, s["y"] - 0.007, s["w"])
        ref2 = sysm.velocity(p, eager=True)
        assert same(sysm.velocity(p), ref2) and same(sysm.velocity(p), ref2)
        sysm.set(s["x"] * 1.7, s["y"], s["w"])
        ref3 = sysm.velocity(p, eager=True)
        assert same(sysm.velocity(p), ref3) and same(sysm.velocity(p), ref3)
        sysm.set(s["x"], s["y"], s["w"])
        assert same(sysm.velocity(p), ref)
