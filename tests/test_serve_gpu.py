"""Serving: admission by HBM footprint, in-order steps per simulation, and
each step's result identical alone and inside a concurrent mix."""
import numpy as np
import pytest

from phyflow_amd import workloads as W
from phyflow_amd._native import PhfError
from phyflow_amd.lagrangian import step_rk2
from phyflow_amd.serve import Server, poisson_run
from phyflow_amd.system import BlobSystem, vel_params

pytestmark = pytest.mark.gpu
REMESH = None


def _sim(n, seed):
    s = W.uniform_system(n, seed=seed)
    return s, vel_params(s["sigma"], prec=1, order=16, leaf=64)


def test_alone_equals_in_mix():
    sims = [_sim(n, 10 + i) for i, n in enumerate([30000, 12000, 50000, 20000, 8000, 40000])]
    dt = 1e-3
    rem = dict(sigma=sims[0][0]["sigma"], overlap=1.0)
    # alone: system 0 stepped three times by itself
    s0, p0 = sims[0]
    with BlobSystem(s0["x"], s0["y"], s0["w"]) as alone:
        for _ in range(3):
            step_rk2(alone, p0, dt, remesh=rem)
        ref = alone.get()
    systems = [BlobSystem(s["x"], s["y"], s["w"]) for s, _ in sims]
    with Server(slots=4) as srv:
        ids = [srv.add(b) for b in systems]
        rids = []
        for rep in range(3):
            for k, (s, p) in enumerate(sims):
                r = rem if k == 0 else dict(sigma=s["sigma"], overlap=1.0)
                rids.append(srv.submit(ids[k], p, dt, remesh=r))
        res = [srv.wait(r) for r in rids]
        assert all(x["status"] == 0 for x in res)
        st = srv.stats()
        assert st["completed"] == 18 and st["p50_ms"] <= st["p99_ms"]
        for i in ids:
            srv.remove(i)
    got = systems[0].get()
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
    for b in systems:
        b.close()


def test_admission_by_footprint():
    s, p = _sim(200000, 1)
    with BlobSystem(s["x"], s["y"], s["w"]) as b, Server(slots=2, hbm_budget_bytes=50 << 20) as srv:
        with pytest.raises(PhfError) as e:
            srv.add(b)
        assert e.value.code == -5


def test_poisson_run_small():
    sims = [_sim(n, 20 + i) for i, n in enumerate([20000, 10000, 30000])]
    systems = [BlobSystem(s["x"], s["y"], s["w"]) for s, _ in sims]
    with Server(slots=3) as srv:
        ids = [srv.add(b) for b in systems]
        st, lat = poisson_run(srv, ids, [p for _, p in sims], rate=200.0, duration_s=1.0,
                              dt=1e-4, seed=3)
        assert st["completed"] == st["submitted"] == lat.size and st["failed"] == 0
        for i in ids:
            srv.remove(i)
    for b in systems:
        b.close()
