"""Autotuner of the FMM settings: leaf size (tree depth), leaf side against
the Oseen cutoff (the domain), near-set width ws and expansion order.

The reference fixes these by rules: the order from the tolerance with
theta = 0.5 and the depth from Ndirect (fmm.cpp:538-555).  Here they are
measured per (precision, tolerance, system size):

    table = sweep([262144, 1048576], [("f32c", 1e-6), ("f64", 1e-12)])
    save(table, path)
    p = tuned_params(sigma, n, prec=PREC_F32C, tol=1e-6)   # VelParams

For every candidate (leaf, leaf_side, ws) the smallest order whose relative
L2 error, against the float64 direct sum on a seeded subset of targets,
meets half the tolerance is found by bisection; the candidate is then timed
on a resident system (median of replayed evaluations) and the fastest one
is kept.  The measured candidates are stored with the choice.

Run on the GPU:  python -c "import __graft_entry__ as g; g.load_package();
    from phyflow_amd import autotune; autotune.main()"  [--out FILE]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import _native
from . import workloads as W
from ._native import PREC_F32C, PREC_F64
from .system import BlobSystem, vel_params

__all__ = ["sweep", "tune_system", "tuned_params", "load_table", "save", "TABLE_PATH"]

TABLE_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_table.json")
PRECS = {"f32c": PREC_F32C, "f64": PREC_F64}

LEAVES = (16, 32, 64, 128, 256)
LEAF_SIDES = (-1.0, 0.5, 1.0)  # bounding box, half-cutoff leaves, cutoff leaves
WS = (1, 2)
ORDER_MAX = 40


def _fmm_level(n, leaf):
    q = max(n, 1) / leaf
    return max(0, min(11, int(round(math.log(q) / math.log(4.0))) if q > 1.0 else 0))


def _reference_velocity(s, sub):
    """float64 direct Oseen sum (the library's own, pinned to the oracle by
    the tests) at the subset targets."""
    from .cpp.blobs.kernels import gauss
    xopt, cutoff = W.oseen_params(s["sigma"])
    qr, qi = gauss.fmm2d_gpu(s["x"], s["y"], s["w"], s["x"][sub], s["y"][sub], 35, xopt, cutoff,
                             0.0, prec=PREC_F64)
    return np.concatenate([-qi / (2 * math.pi), -qr / (2 * math.pi)])


def _error(sysm, params, sub, ref):
    sysm.velocity(params, fetch=False)
    vx, vy = sysm.last_velocity()
    got = np.concatenate([vx[sub], vy[sub]])
    return float(np.linalg.norm(got - ref) / np.linalg.norm(ref))


def _time(sysm, params, reps):
    for _ in range(3):
        sysm.velocity(params, fetch=False)
    _native.lib().phf_synchronize()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        sysm.velocity(params, fetch=False)
        ts.append(time.perf_counter() - t)
    return 1e3 * float(np.median(ts))


def tune_system(s, prec, tol, n_check=1024, reps=10, leaves=LEAVES, leaf_sides=LEAF_SIDES,
                ws_set=WS, log=None):
    """Measured choice of (leaf, leaf_side, ws, order) for system s (dict of
    x, y, w, sigma) at precision prec ('f32c' / 'f64') and tolerance tol."""
    n = s["x"].size
    sub = W.target_subset(n, n_check, seed=7)
    ref = _reference_velocity(s, sub)
    cands, seen = [], set()
    with BlobSystem(s["x"], s["y"], s["w"]) as sysm:
        for leaf in leaves:
            L = _fmm_level(n, leaf)
            if L in seen:
                continue
            seen.add(L)
            for side in leaf_sides:
                for ws in ws_set:
                    def params(p):
                        return vel_params(s["sigma"], prec=PRECS[prec], order=p, leaf=leaf, ws=ws,
                                          leaf_side=side)
                    # smallest order meeting tol / 2 (the error falls with the order)
                    lo, hi = 2, ORDER_MAX
                    if _error(sysm, params(hi), sub, ref) > 0.5 * tol:
                        cands.append({"leaf": leaf, "level": L, "leaf_side": side, "ws": ws,
                                      "order": None})
                        continue
                    while lo < hi:
                        mid = (lo + hi) // 2
                        if _error(sysm, params(mid), sub, ref) <= 0.5 * tol:
                            hi = mid
                        else:
                            lo = mid + 1
                    p = params(lo)
                    ms = _time(sysm, p, reps)
                    sysm.velocity(p, fetch=False, info=True)
                    info = sysm.last_info
                    c = {"leaf": leaf, "level": info["L"], "leaf_side": side, "ws": ws,
                         "order": lo, "ms": ms, "rel_l2": _error(sysm, p, sub, ref),
                         "domain": info["D"], "near_boxes": info["nnear"]}
                    cands.append(c)
                    if log:
                        log("  %s" % json.dumps(c))
    ok = [c for c in cands if c["order"] is not None]
    best = min(ok, key=lambda c: c["ms"]) if ok else None
    return {"prec": prec, "tol": tol, "n": n, "sigma": s["sigma"],
            "sources_in_cutoff": _sources_in_cutoff(s), "choice": best, "candidates": cands}


def _sources_in_cutoff(s):
    xopt, cutoff = W.oseen_params(s["sigma"])
    area = max((s["x"].max() - s["x"].min()) * (s["y"].max() - s["y"].min()), 1e-300)
    return float(math.pi * cutoff * cutoff * s["x"].size / area)


def sweep(sizes, cases, seed=0, log=print):
    """Tune uniform systems (positions in the unit square, N(0,1) h^2
    circulations, h = 1/sqrt(N), sigma = h/0.8) of the given sizes for the
    (precision, tolerance) cases."""
    table = {"entries": [], "device": _native.device_info()["name"]}
    for n in sizes:
        s = W.uniform_system(int(n), seed=seed)
        for prec, tol in cases:
            t = time.perf_counter()
            e = tune_system(s, prec, tol, log=log)
            e["tune_seconds"] = time.perf_counter() - t
            table["entries"].append(e)
            if log:
                log("n=%d %s tol=%g -> %s (%.1f s)" % (n, prec, tol, json.dumps(e["choice"]),
                                                        e["tune_seconds"]))
    return table


def save(table, path=TABLE_PATH):
    with open(path, "w") as f:
        json.dump(table, f, indent=1)


def load_table(path=TABLE_PATH):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def tuned_params(sigma, n, prec=PREC_F32C, tol=1e-6, k=2.0, table=None):
    """VelParams for a system of n blobs with core size sigma from the
    measured table: the entry of the same precision with the largest
    tolerance not above tol (same tolerance class), nearest in log n.  With
    no such entry, the reference's rules (order from tol, depth from
    Ndirect) apply."""
    table = table if table is not None else load_table()
    name = "f32c" if prec == PREC_F32C else "f64"
    best = None
    for e in (table or {}).get("entries", []):
        c = e.get("choice")
        if c is None or e["prec"] != name or e["tol"] > tol * (1 + 1e-9):
            continue
        key = (-e["tol"], abs(math.log(max(n, 1)) - math.log(e["n"])))
        if best is None or key < best[0]:
            best = (key, c)
    if best is None:
        return vel_params(sigma, k=k, prec=prec, tol=tol)
    c = best[1]
    return vel_params(sigma, k=k, prec=prec, tol=tol, order=c["order"], leaf=c["leaf"],
                      ws=c["ws"], leaf_side=c["leaf_side"])


def main(argv=None):
    ap = argparse.ArgumentParser(description="measure and store the FMM settings table")
    ap.add_argument("--sizes", default="65536,262144,1048576")
    ap.add_argument("--cases", default="f32c:1e-6,f64:1e-6,f64:1e-12")
    ap.add_argument("--out", default=TABLE_PATH)
    a = ap.parse_args(argv)
    sizes = [int(v) for v in a.sizes.split(",")]
    cases = [(c.split(":")[0], float(c.split(":")[1])) for c in a.cases.split(",")]
    table = sweep(sizes, cases, log=lambda m: print(m, flush=True))
    save(table, a.out)
    print("wrote %s" % a.out, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
