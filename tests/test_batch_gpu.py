"""Batched evaluation of uneven systems (segment table + hipGraph).

Checks: each system in the batch agrees with its single-system FMM
(float64: <= 1e-12 relative; F32C: <= 1e-6) and with the float64 direct
sum (<= 1e-6); a system gives bit-identical results alone and inside the
mix; graph replays are bit-identical to eager runs, with eager work and
synchronisation between replays."""
import math

import numpy as np
import pytest

from phyflow_amd import workloads as W
from phyflow_amd.batch import PREC_F32C, PREC_F64, BatchSystems
from phyflow_amd.system import BlobSystem, vel_params

pytestmark = pytest.mark.gpu


def _mix(count, lo=1024, hi=20000, seed=3):
    sizes = W.uneven_sizes(count, lo, hi, seed=seed)
    systems = [W.uniform_system(int(n), seed=100 + i) for i, n in enumerate(sizes)]
    return sizes, systems


@pytest.mark.parametrize("prec,tol", [(PREC_F64, 1e-12), (PREC_F32C, 1e-6)])
def test_batch_matches_single(prec, tol, O):
    sizes, systems = _mix(12)
    with BatchSystems(sizes, [s["sigma"] for s in systems], prec=prec, order=16, leaf=64) as b:
        for i, s in enumerate(systems):
            b.set_system(i, s["x"], s["y"], s["w"])
        info = b.velocity(use_graph=False)
        assert info["flags"] == 0
        for i, s in enumerate(systems):
            got = b.get_velocity(i)
            with BlobSystem(s["x"], s["y"], s["w"]) as one:
                ref = one.velocity(vel_params(s["sigma"], prec=prec, order=16, leaf=64))
            assert O.rel_l2(got, ref) <= tol, i
            if i % 4 == 0:
                sub = W.target_subset(s["x"].size, 500, seed=i)
                xopt, cutoff = W.oseen_params(s["sigma"])
                qr, qi = O.direct_oseen(s["x"], s["y"], s["w"], s["x"][sub], s["y"][sub], xopt,
                                        cutoff)
                d = (-qi / (2 * math.pi), -qr / (2 * math.pi))
                assert O.rel_l2((got[0][sub], got[1][sub]), d) <= 1e-6


def test_alone_equals_in_mix_and_graph_replays():
    sizes, systems = _mix(10, seed=5)
    sizes = np.concatenate([sizes, [0, 1, 5]])  # empty and tiny systems too
    rng = np.random.default_rng(9)
    systems += [dict(x=np.empty(0), y=np.empty(0), w=np.empty(0), sigma=0.01)]
    for n in (1, 5):
        x, y = rng.random((2, n))
        systems.append(dict(x=x, y=y, w=rng.standard_normal(n) * 1e-4, sigma=0.01))
    sig = [s["sigma"] for s in systems]
    with BatchSystems(sizes, sig, prec=PREC_F32C) as b:
        for i, s in enumerate(systems):
            b.set_system(i, s["x"], s["y"], s["w"])
        b.velocity(use_graph=False)
        eager = [b.get_velocity(i) for i in range(len(systems))]
        for rep in range(3):
            b.velocity(use_graph=True)
            for i in range(len(systems)):
                g = b.get_velocity(i)
                assert np.array_equal(g[0], eager[i][0]) and np.array_equal(g[1], eager[i][1])
            b.velocity(use_graph=False)  # eager work between replays
    for k in (0, 3, 11):
        with BatchSystems([sizes[k]], [sig[k]], prec=PREC_F32C) as one:
            one.set_system(0, systems[k]["x"], systems[k]["y"], systems[k]["w"])
            one.velocity(use_graph=True)
            a = one.get_velocity(0)
        assert np.array_equal(a[0], eager[k][0]) and np.array_equal(a[1], eager[k][1])
