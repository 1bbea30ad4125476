"""Buffer lifetimes of resident systems and one-shot calls.

Resident float64 systems are created, evaluated and destroyed in turn with
growing sizes (growing tables, pinned staging and arenas), more expansion
orders per system than a lane keeps recordings for (recorded launch
sequences and their tables are evicted and freed), eager evaluations
between replays, buffers regrown by set(), and several systems alive at
once.  Every repeated evaluation must give the same float64 sums as the
first, and one per system is checked against the oracle.  One-shot calls
with growing sizes and orders exercise the shared lanes the same way."""
import math

import numpy as np
import pytest

from phyflow_amd import workloads as W
from phyflow_amd._native import PREC_F32C, PREC_F64
from phyflow_amd.cpp.blobs.kernels import gauss
from phyflow_amd.system import BlobSystem, vel_params

pytestmark = pytest.mark.gpu

ORDERS = (8, 12, 16, 20, 24, 30)  # more recordings than a lane keeps (4)


def _oracle_check(O, s, v, m=256, seed=0):
    sub = W.target_subset(s["x"].size, m, seed=seed)
    xopt, cutoff = W.oseen_params(s["sigma"])
    qr, qi = O.direct_oseen(s["x"], s["y"], s["w"], s["x"][sub], s["y"][sub], xopt, cutoff)
    return O.rel_l2((v[0][sub], v[1][sub]), (-qi / (2 * math.pi), -qr / (2 * math.pi)))


def test_resident_systems_created_evaluated_destroyed(O):
    sizes = np.geomspace(2000, 150000, 16).astype(int)
    alive = []
    try:
        for k, n in enumerate(sizes):
            s = W.uniform_system(int(n), seed=500 + k)
            sysm = BlobSystem(s["x"], s["y"], s["w"])
            alive.append(sysm)
            first = {}
            for rep in range(2):
                for p in ORDERS:
                    v = sysm.velocity(vel_params(s["sigma"], prec=PREC_F64, order=p, leaf=64))
                    if rep == 0:
                        first[p] = v
                    else:
                        assert np.array_equal(v[0], first[p][0]), (n, p)
                        assert np.array_equal(v[1], first[p][1]), (n, p)
                ve = sysm.velocity(vel_params(s["sigma"], prec=PREC_F64, order=16, leaf=64),
                                   eager=True)
                assert np.array_equal(ve[0], first[16][0]) and np.array_equal(ve[1], first[16][1])
            e = _oracle_check(O, s, first[16], seed=k)
            print("resident n=%d p=16 rel_l2=%.3e alive=%d" % (n, e, len(alive)), flush=True)
            assert e <= 1e-6
            # regrow the resident buffers (old ones freed) and evaluate again
            s2 = W.uniform_system(int(n * 1.6), seed=700 + k)
            sysm.set(s2["x"], s2["y"], s2["w"])
            v2 = sysm.velocity(vel_params(s2["sigma"], prec=PREC_F64, order=16, leaf=64))
            assert _oracle_check(O, s2, v2, seed=k) <= 1e-6
            if len(alive) > 5:
                alive.pop(0).close()
    finally:
        for a in alive:
            a.close()


def test_one_shot_calls_with_growing_tables(O):
    for k, (n, p, prec) in enumerate([(1000, 8, PREC_F64), (5000, 30, PREC_F32C),
                                      (20000, 16, PREC_F64), (80000, 24, PREC_F32C),
                                      (200000, 30, PREC_F64), (3000, 12, PREC_F32C)]):
        s = W.uniform_system(n, seed=900 + k)
        xopt, cutoff = W.oseen_params(s["sigma"])
        qr, qi = gauss.fmm2d_gpu(s["x"], s["y"], s["w"], s["x"], s["y"], 35, xopt, cutoff, 1e-6,
                                 prec=prec, order=p, leaf=64)
        v = (-qi / (2 * math.pi), -qr / (2 * math.pi))
        e = _oracle_check(O, s, v, seed=k)
        print("one-shot n=%d p=%d prec=%d rel_l2=%.3e" % (n, p, prec, e), flush=True)
        assert e <= (1e-4 if p < 12 else (1e-5 if p < 16 else 1e-6))
