"""Serving resident simulations (phf_server_*): admission by HBM footprint,
continuous batching of step requests over worker slots, per-request
latency; plus a seeded Poisson load generator.

    srv = Server(slots=8)
    sid = srv.add(system)                  # BlobSystem (ownership stays with the caller)
    rid = srv.submit(sid, vel_params(...), dt=1e-3, remesh=dict(sigma=..., overlap=...))
    res = srv.wait(rid)                    # latency_ms, queue_ms, step info
"""
import ctypes
import time

import numpy as np

from ._native import (RequestResult, ServerStats, StepParams, call, lib)

__all__ = ["Server", "poisson_run"]


def step_params(dt, vInf=(0.0, 0.0), remesh=None):
    sp = StepParams()
    sp.dt = float(dt)
    sp.vinf_x, sp.vinf_y = float(vInf[0]), float(vInf[1])
    if remesh:
        h = remesh["sigma"] * remesh.get("overlap", 1.0)
        sp.remesh = 1
        sp.remesh_sigma = float(remesh["sigma"])
        sp.remesh_overlap = float(remesh.get("overlap", 1.0))
        sp.remesh_x0 = float(remesh.get("x0", -1.5 * h))
        sp.remesh_y0 = float(remesh.get("y0", -1.5 * h))
        if "hnu" in remesh:  # Tutty's diffusion weights, hNu = sqrt(nu dt)
            sp.remesh_kernel = 1
            sp.remesh_c = float(remesh["hnu"])
        else:  # M4' with the Wee parameter c
            sp.remesh_c = float(remesh.get("c", 0.0))
        if remesh.get("population_control"):
            # Blobs defaults gThresholdLocal = gThresholdGlobal = 1e-8 (__blobsClass.py:105-106)
            local, glob = remesh["population_control"]
            sp.pc_enabled = 1
            sp.pc_w_threshold = float(local)
            sp.pc_total_threshold = float(glob)
    return sp


class Server:
    def __init__(self, slots=8, hbm_budget_bytes=0):
        h = ctypes.c_void_p()
        call("phf_server_create", int(slots), int(hbm_budget_bytes), ctypes.byref(h))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().phf_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, system):
        i = ctypes.c_int64()
        call("phf_server_add_system", self._h, system._h, ctypes.byref(i))
        return i.value

    def remove(self, sim_id):
        call("phf_server_remove_system", self._h, int(sim_id))

    def submit(self, sim_id, params, dt, vInf=(0.0, 0.0), remesh=None):
        sp = step_params(dt, vInf, remesh)
        r = ctypes.c_int64()
        call("phf_server_submit", self._h, int(sim_id), ctypes.byref(params), ctypes.byref(sp),
             ctypes.byref(r))
        return r.value

    def wait(self, req_id, timeout_s=-1.0):
        res = RequestResult()
        call("phf_server_wait", self._h, int(req_id), float(timeout_s), ctypes.byref(res))
        return {"status": res.status, "latency_ms": res.latency_ms, "queue_ms": res.queue_ms,
                "info": res.info.as_dict()}

    def drain(self):
        call("phf_server_drain", self._h)

    def stats(self, reset=False):
        st = ServerStats()
        call("phf_server_stats", self._h, 1 if reset else 0, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in st._fields_ if k != "pad"}


def poisson_run(server, sim_ids, params, rate, duration_s, dt, remesh=None, seed=0,
                progress=None):
    """Submit step requests at Poisson arrival times (rate per second) to
    uniformly chosen simulations for duration_s; waits for all of them.
    params: one velocity-parameter struct per simulation.  Returns the
    server statistics of the run and the per-request latencies."""
    rng = np.random.default_rng(seed)
    server.stats(reset=True)
    t0 = time.perf_counter()
    t_next = 0.0
    rids = []
    last_report = t0
    while True:
        t_next += rng.exponential(1.0 / rate)
        if t_next > duration_s:
            break
        k = int(rng.integers(len(sim_ids)))
        now = time.perf_counter() - t0
        if t_next > now:
            time.sleep(t_next - now)
        rids.append(server.submit(sim_ids[k], params[k], dt, remesh=remesh))
        if progress and time.perf_counter() - last_report > 10.0:
            last_report = time.perf_counter()
            progress("%.0f s: %d submitted, %s" % (last_report - t0, len(rids),
                                                  server.stats()["completed"]))
    lat = [server.wait(r)["latency_ms"] for r in rids]
    st = server.stats()
    st["submitted"] = len(rids)
    st["offered_rate"] = rate
    return st, np.array(lat)
