"""Benchmark of the vortex-blob velocity hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (headline): 1,048,576 blobs uniform in [0,1]^2 from seed 0,
circulations N(0,1) h^2 with h = 1/1024, sigma = h/0.8, k = 2; velocity at
every blob centre by the FMM (Oseen smoother, cutoff 5 xopt) at order 16,
leaf size 64, float32 pair arithmetic with compensated sums.  The blobs are
resident in HBM before timing starts; one step = one full velocity
evaluation (tree build, expansions, near field), synchronised.

Prints one JSON line: throughput in velocity evaluations/s, p50/p99 step
latency, the error against a float64 direct sum on 65,536 seeded targets,
the P2P kernel's roofline figures (HIP events on the library's stream) and
the reference's own CPU FMM timed on this host.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from __graft_entry__ import load_package  # noqa: E402

load_package()

from phyflow_amd import _native  # noqa: E402
from phyflow_amd import autotune  # noqa: E402
from phyflow_amd import workloads as W  # noqa: E402
from phyflow_amd.cpp.blobs.kernels import gauss  # noqa: E402
from phyflow_amd.system import BlobSystem, vel_params  # noqa: E402

FP32_PEAK_TFLOPS = 157.3  # MI355X fp32 (vector, packed) / dense matrix
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
FLOP_PER_PAIR = 17


def timed_steps(sysm, params, warmup, steps, counts=True):
    for _ in range(warmup):
        sysm.velocity(params, fetch=False)
    lat, p2p_ms, info = [], [], None
    _native.lib().phf_synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        sysm.velocity(params, fetch=False, p2p_timing_only=True)
        lat.append(time.perf_counter() - a)
        p2p_ms.append(sysm.last_info["ms"].get("p2p", float("nan")))
    _native.lib().phf_synchronize()
    wall = time.perf_counter() - t0
    # stage timings from one more (untimed) evaluation
    sysm.velocity(params, fetch=False, info=True)
    info = sysm.last_info
    if counts:  # work counts from one extra (untimed) evaluation
        sysm.velocity(params, fetch=False, info=True, count_work=True)
        info = dict(info, p2p_pairs=sysm.last_info["p2p_pairs"],
                    m2l_links=sysm.last_info["m2l_links"])
    return wall, np.array(lat), np.array(p2p_ms), info


class OracleReference:
    """float64 velocity of the CPU oracle (restated direct Oseen sum,
    direct.cpp:1409-1485, OpenMP on every host core) at n_check seeded
    targets; computed once and reused by every variant."""

    def __init__(self, s, n_check, seed):
        from oracle import oracle as O
        self.sub = W.target_subset(s["x"].size, n_check, seed=seed)
        xopt, cutoff = W.oseen_params(s["sigma"])
        t = time.perf_counter()
        qr, qi = O.direct_oseen(s["x"], s["y"], s["w"], s["x"][self.sub], s["y"][self.sub],
                                xopt, cutoff, nthreads=host_cores())
        self.seconds = time.perf_counter() - t
        self.ref = np.concatenate([-qi / (2 * math.pi), -qr / (2 * math.pi)])

    def errors(self, vx, vy):
        a = np.concatenate([vx[self.sub], vy[self.sub]])
        b = self.ref
        return (float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                float(np.max(np.abs(a - b)) / np.max(np.abs(b))))


def rel_errors(sysm, params, oracle_ref):
    """rel-L2 and max error of the evaluated velocity against the oracle."""
    sysm.velocity(params, fetch=False)
    vx, vy = sysm.last_velocity()
    return oracle_ref.errors(vx, vy)


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def host_cpu():
    """Model name and usable core count of this host."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "cores_usable": host_cores(), "cores_online": os.cpu_count()}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fmm_1m_p2p_hbm_traffic.json")
ISSUE_FILE = os.path.join(ROOT, "profiles", "fmm_1m_p2p_valu_issue.json")


def measured_traffic(kernel, args):
    """HBM bytes per launch of the P2P kernel at the default workload, from
    the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (each counter
    scaled by what it reads for a copy of known size); None when the
    summary is absent or was taken on another kernel or workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("kernel") != kernel or (args.n, args.order, args.leaf, args.prec) != (
            1048576, 16, 64, "f32c"):
        return None
    return {"bytes_per_launch": d["bytes_per_launch"],
            "source": "profiles/fmm_1m_p2p_hbm_traffic.json"}


def measured_issue(kernel, args):
    """Fraction of the SIMDs' issue cycles the P2P kernel keeps its VALU
    busy at the default workload, from the committed rocprofv3
    SQ_INSTS_VALU / GRBM_GUI_ACTIVE summary; None when the summary is absent
    or was taken on another kernel or workload."""
    try:
        with open(ISSUE_FILE) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("kernel") != kernel or (args.n, args.order, args.leaf, args.prec) != (
            1048576, 16, 64, "f32c"):
        return None
    return {"valu_issue_frac": d["valu_issue_frac"],
            "lane_instructions_per_pair": d["lane_instructions_per_pair"],
            "source": "profiles/fmm_1m_p2p_valu_issue.json"}


def cpu_baseline(s, budget_s=40.0):
    """The reference's compiled CPU FMM (oracle/_ref) at its default
    tol=1e-6, Ndirect=35 on the same workload; falls back to the oracle's
    direct sum on a target subset if the reference build is absent."""
    from oracle import oracle as O
    xopt, cutoff = W.oseen_params(s["sigma"])
    n = s["x"].size
    if O.ref_available():
        t = time.perf_counter()
        O.ref_fmm2d(s["x"], s["y"], s["w"], s["x"], s["y"], 35, xopt, cutoff, 1e-6)
        dt = time.perf_counter() - t
        return {"value": n / dt, "unit": "evals/s", "cores": 1, "kind": "reference",
                "sample": "full %d-blob workload, reference fmm2d CPU build (tol=1e-6, "
                          "Ndirect=35, single-threaded as shipped), %.1f s" % (n, dt),
                "host_cpu": host_cpu()}
    cores = host_cores()
    m = 4096
    sub = W.target_subset(n, m, seed=0)
    t = time.perf_counter()
    O.direct_oseen(s["x"], s["y"], s["w"], s["x"][sub], s["y"][sub], xopt, cutoff,
                   nthreads=cores)
    dt = time.perf_counter() - t
    return {"value": m / dt, "unit": "evals/s", "cores": cores, "kind": "port",
            "sample": "%d seeded targets of the %d-blob workload, float64 direct sum" % (m, n),
            "host_cpu": host_cpu()}


def bench_lamb(args):
    """Lamb-Oseen vortex on a 64x64 lattice: direct Gaussian sum (the
    vv2parGauss kernel) at all 4,096 blob centres, float64 and F32C, vs the
    CPU restatement of the reference direct sum timed on this host."""
    from oracle import oracle as O
    c = W.lamb_oseen_lattice()
    ks = c["k"] * c["sigma"] ** 2
    t = time.perf_counter()
    ref = O.vv2par_gauss(c["x"], c["y"], c["w"], c["x"], c["y"], ks, nthreads=host_cores())
    t_cpu = time.perf_counter() - t
    res = {}
    for name, prec in (("f64", gauss.PREC_F64), ("f32c", gauss.PREC_F32C)):
        p = vel_params(c["sigma"], method=1, prec=prec)
        with BlobSystem(c["x"], c["y"], c["w"]) as sysm:
            wall, lat, p2p_ms, info = timed_steps(sysm, p, args.warmup, args.steps, counts=False)
            v = sysm.last_velocity()
        res[name] = {"evals_per_s": c["x"].size / (wall / args.steps),
                     "ms_per_step": 1e3 * wall / args.steps,
                     "kernel_ms": float(np.nanmean(p2p_ms)),
                     "rel_l2": O.rel_l2(v, ref), "max_err": O.max_rel(v, ref)}
    n = c["x"].size
    pairs = n * n
    best = res["f32c"]
    ach = FLOP_PER_PAIR * pairs / (best["kernel_ms"] * 1e-3) / 1e12
    return {"metric": "blob vel. evals/s, Lamb-Oseen 64x64 direct sum", "value": best["evals_per_s"],
            "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "direct Gaussian sum, 4096 lattice blobs (Lamb-Oseen)",
                       "variants": res},
            "roofline": {"kernel": "direct_f32c", "bound": "valu",
                         "compute_pipe": "fp32 vector ALU (no matrix shape)", "achieved": ach,
                         "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP32_PEAK_TFLOPS,
                         "traffic": None},
            "cpu_baseline": {"value": n / t_cpu, "unit": "evals/s", "cores": host_cores(),
                             "kind": "port", "sample": "full 4096x4096 direct sum, float64, "
                             "restatement of vv2parGauss_C_cpu (OpenMP)",
                             "host_cpu": host_cpu()}}


def bench_cylinder(args):
    """One coupled step of the impulsively started cylinder: 512 panels,
    262,144 blobs in 0.5 < r < 1.5, Heun RK2 with FMM velocities (F32C,
    order 16) + panel velocity + vInf, panel solve by float32 LU + float64
    refinement, M4' remesh (h = 0.005) with float64 invariants.  Timed twice:
    with the body's factors kept from one step to the next (a rigid body),
    and with the influence matrix built and factored inside every step (the
    step as the configuration defines it; a moving body)."""
    from phyflow_amd.lagrangian import PanelBody, step_rk2
    xp, yp = W.cylinder_panels(512, R=0.5)
    a = W.annulus_blobs(262144, sigma=0.004, seed=0)
    t = time.perf_counter()
    body = PanelBody(xp, yp, "prescribed", 0.0)
    t_build = time.perf_counter() - t
    p_fixed = vel_params(0.004, prec=gauss.PREC_F32C, order=16, leaf=64)
    # the FMM settings measured for this blob system by the autotuner (the
    # fastest leaf size, leaf side, near-set width and order whose error
    # against the float64 direct sum on 1,024 seeded targets is <= 5e-7)
    t = time.perf_counter()
    tune = autotune.tune_system(dict(x=a["x"], y=a["y"], w=a["w"], sigma=0.004), "f32c", 1e-6)
    t_tune = time.perf_counter() - t
    ch = tune["choice"]
    p = p_fixed if ch is None else vel_params(0.004, prec=gauss.PREC_F32C, order=ch["order"],
                                              leaf=ch["leaf"], ws=ch["ws"],
                                              leaf_side=ch["leaf_side"])
    remesh = dict(sigma=0.004, overlap=1.25)

    def run(rebuild, wait=False, params=None):
        params = p if params is None else params
        times, infos = [], []
        # one resident simulation stepped repeatedly (the first steps warm
        # the system's device workspace); every step is a full coupled RK2
        # + remesh.  rebuild: the body's geometry, influence matrix and
        # factors are rebuilt before every step; wait=False enqueues that
        # rebuild, and the step's first blob velocity evaluation runs beside
        # the factorisation
        with BlobSystem(a["x"], a["y"], a["w"]) as sysm:
            for k in range(args.warmup + args.steps):
                _native.lib().phf_synchronize()
                t = time.perf_counter()
                if rebuild:
                    body.update_geometry(xp, yp, wait=wait)
                info = step_rk2(sysm, params, 0.001, vInf=(1.0, 0.0), panels=body,
                                remesh=remesh)
                dt = time.perf_counter() - t
                if k >= args.warmup:
                    times.append(dt)
                    infos.append(info)
        return np.array(times), infos

    times, infos = run(False)
    times_rb, infos_rb = run(True)
    times_rbw, _ = run(True, wait=True)
    times_fix, _ = run(True, params=p_fixed)
    # the rebuild's stage times (device events), from waited updates
    rb = [body.update_geometry(xp, yp) for _ in range(max(3, args.steps))]
    info = infos[-1]
    tot = float(np.sum(np.abs(a["w"])))
    n = 512
    lu_ms = float(np.median([r["lu_ms"] for r in rb]))
    # the 'prescribed' body's normal matrix is factored by Cholesky
    lu_flop = n ** 3 / 3.0 + n ** 2 / 2.0
    rem_ms = float(np.median([i["ms_remesh"] for i in infos]))
    rem_bytes = 3 * 8 * (info["n_before"] + info["n_after"])
    return {"metric": "cylinder step time (512 panels, 262144 blobs, RK2 + remesh)",
            "value": 1e3 * float(np.median(times_rb)), "unit": "ms", "n_gpus": 1,
            "steps": len(times_rb), "warmup": args.warmup,
            "ms_per_step": 1e3 * float(np.median(times_rb)),
            "p50_ms": 1e3 * float(np.percentile(times_rb, 50)),
            "p99_ms": 1e3 * float(np.percentile(times_rb, 99)),
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "impulsively started cylinder (512 panels, 262144 blobs, "
                                   "dt 1e-3): consecutive coupled steps of one simulation, each "
                                   "building and factoring the 512x512 influence matrix",
                       "ms_per_step_factors_kept": 1e3 * float(np.median(times)),
                       "p99_ms_factors_kept": 1e3 * float(np.percentile(times, 99)),
                       "ms_per_step_update_waited": 1e3 * float(np.median(times_rbw)),
                       "fmm_params": {"tuned": ch, "tune_seconds": t_tune,
                                      "ms_per_step_fixed_p16_leaf64":
                                          1e3 * float(np.median(times_fix))},
                       "update": "enqueued (phf_panels_update_async); the step waits for the "
                                 "factors only at its panel solve",
                       "panel_rebuild_ms": {k: float(np.median([r[k] for r in rb]))
                                            for k in rb[0]},
                       "panel_setup_first_call_ms": 1e3 * t_build, "step_info": info,
                       "step_info_rebuild": infos_rb[-1],
                       "circ_rel_change": abs(info["circ_after"] - info["circ_before"]) / tot,
                       "impulse_rel_change": max(abs(info["impulse_x_after"] - info["impulse_x_before"]),
                                                 abs(info["impulse_y_after"] - info["impulse_y_before"])) / tot},
            "roofline": {"kernel": "chol_first + chol_step + chol_to_lu (float32 Cholesky of "
                                   "the 512x512 normal matrix, in LU form)",
                         "bound": "mfma (latency-bound: the per-panel diagonal chain)",
                         "compute_pipe": "v_mfma_f32_16x16x4f32 trailing update",
                         "achieved": lu_flop / (lu_ms * 1e-3) / 1e12, "peak": FP32_PEAK_TFLOPS,
                         "unit": "TFLOP/s",
                         "frac": lu_flop / (lu_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
                         "traffic": None, "flop_rule": "n^3 / 3 + n^2 / 2, n = 512",
                         "kernel_ms": lu_ms},
            "roofline_other": {"remesh": {"bound": "hbm", "algorithmic_bytes": rem_bytes,
                                          "bytes_rule": "3 s (N + M_out), s = 8",
                                          "ms": rem_ms, "achieved": rem_bytes / (rem_ms * 1e-3) / 1e9,
                                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                          "frac": rem_bytes / (rem_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                          "note": "remesh stage of the step (events on the step's "
                                                  "stream; includes the stream's wait while the "
                                                  "host reads the output count)"}}}


def mix_leg_child(timeout_s=400):
    """mix_leg in a child process (its own device context and memory), so
    the headline line is printed whatever happens there."""
    import subprocess
    cmd = [sys.executable, os.path.abspath(__file__), "--workload", "mixleg"]
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return {"error": "timed out after %d s" % timeout_s}
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": "exit %d: %s" % (r.returncode, r.stderr[-400:])}
    return json.loads(lines[-1])


def mix_leg(args, count=256, steps=10):
    """The 256-system mix in compact form (N log-uniform in [1024, 262144],
    seed 0, uniform blobs, h = 1/sqrt(N), sigma = h/0.8): batched F32C
    velocity (order 16, leaf 64, one captured graph per evaluation) and the
    batched step (velocity + Euler + M4' remesh), against one-at-a-time
    float64 velocity of the same systems (order 16, leaf 64): one-shot
    fmm2d calls with host arrays (per-call copies, as the reference's GPU
    path) and resident systems evaluated one after another.  Eight systems
    are checked against the CPU oracle on 1,024 seeded targets each."""
    from oracle import oracle as O
    from phyflow_amd.batch import BatchSystems
    t_all = time.perf_counter()
    sizes = W.uneven_sizes(count, 1024, 262144, seed=0)
    systems = [W.uniform_system(int(n), seed=1000 + i) for i, n in enumerate(sizes)]
    total = int(np.sum(sizes))
    errs = []
    with BatchSystems(sizes, [s["sigma"] for s in systems], prec=gauss.PREC_F32C,
                      order=batch_order(systems), leaf=64) as b:
        for i, s in enumerate(systems):
            b.set_system(i, s["x"], s["y"], s["w"])
        b.set_remesh(0.8, dt=1e-3)
        for _ in range(2):
            b.velocity(use_graph=True)
        t = time.perf_counter()
        for _ in range(steps):
            b.velocity(use_graph=True)
        wall_v = time.perf_counter() - t
        for i in range(0, count, count // 8):
            s = systems[i]
            v = b.get_velocity(i)
            sub = W.target_subset(s["x"].size, 1024, seed=i)
            xopt, cutoff = W.oseen_params(s["sigma"])
            qr, qi = O.direct_oseen(s["x"], s["y"], s["w"], s["x"][sub], s["y"][sub], xopt,
                                    cutoff, nthreads=host_cores())
            errs.append(O.rel_l2((v[0][sub], v[1][sub]),
                                 (-qi / (2 * math.pi), -qr / (2 * math.pi))))
        for _ in range(2):
            b.step(use_graph=True)
        t = time.perf_counter()
        for _ in range(steps):
            b.step(use_graph=True)
        wall_s = time.perf_counter() - t
    batched_v = total / (wall_v / steps)
    # one-at-a-time float64: one-shot calls with host arrays
    p64 = dict(prec=gauss.PREC_F64, order=16, leaf=64)
    s0 = systems[0]
    x0, c0 = W.oseen_params(s0["sigma"])
    gauss.fmm2d_gpu(s0["x"], s0["y"], s0["w"], s0["x"], s0["y"], 35, x0, c0, 1e-6, **p64)
    t = time.perf_counter()
    for s in systems:
        xopt, cutoff = W.oseen_params(s["sigma"])
        gauss.fmm2d_gpu(s["x"], s["y"], s["w"], s["x"], s["y"], 35, xopt, cutoff, 1e-6, **p64)
    t_oneshot = time.perf_counter() - t
    # one-at-a-time float64: resident systems, one evaluation after another
    resident = [BlobSystem(s["x"], s["y"], s["w"]) for s in systems]
    try:
        ps = [vel_params(s["sigma"], **p64) for s in systems]
        for r, p in zip(resident, ps):
            r.velocity(p, fetch=False)
        _native.lib().phf_synchronize()
        t = time.perf_counter()
        for r, p in zip(resident, ps):
            r.velocity(p, fetch=False)
        t_res = time.perf_counter() - t
    finally:
        for r in resident:
            r.close()
    return {"workload": "%d systems, N log-uniform [1024, 262144] seed 0, total %d blobs"
                        % (count, total),
            "batched_order": batch_order(systems),
            "batched_f32c_velocity_evals_per_s": batched_v,
            "batched_f32c_step_evals_per_s": total / (wall_s / steps),
            "batched_steps_timed": steps,
            "f64_one_shot_sequential_evals_per_s": total / t_oneshot,
            "f64_resident_sequential_evals_per_s": total / t_res,
            "speedup_batched_velocity_vs_f64_one_shot": batched_v / (total / t_oneshot),
            "speedup_batched_velocity_vs_f64_resident": batched_v / (total / t_res),
            "rel_l2_vs_oracle_max": max(errs), "checked_systems": len(errs),
            "leg_seconds": time.perf_counter() - t_all}


def bench_mix(args):
    """256 independent uniform systems, N log-uniform in [1024, 262144]
    (seed 0), h = 1/sqrt(N), sigma = h/0.8.  One step of every system =
    velocity at the blob centres (F32C FMM, order 16, leaf 64), forward-Euler
    convection (dt 1e-3) and M4' remesh onto the lattice h (origin -1.5 h),
    all systems in one captured hipGraph replay.  Baselines: the same mix as
    256 one-at-a-time float64 steps (one-shot fmm2d + Regrid calls with
    per-call copies, as the reference's GPU path does) and resident float64
    velocity evaluations one after another."""
    from phyflow_amd.batch import BatchSystems
    from phyflow_amd.blobs.base.regrid import Regrid
    count = args.systems
    dt, overlap = 1e-3, 0.8
    sizes = W.uneven_sizes(count, 1024, 262144, seed=0)
    systems = [W.uniform_system(int(n), seed=1000 + i) for i, n in enumerate(sizes)]
    total = int(np.sum(sizes))
    with BatchSystems(sizes, [s["sigma"] for s in systems], prec=gauss.PREC_F32C,
                      order=batch_order(systems), leaf=64) as b:
        for i, s in enumerate(systems):
            b.set_system(i, s["x"], s["y"], s["w"])
        b.set_remesh(overlap, dt=dt)
        for _ in range(args.warmup):
            b.step(use_graph=True)
        lat = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            a = time.perf_counter()
            info = b.step(use_graph=True)
            lat.append(time.perf_counter() - a)
        wall = time.perf_counter() - t0
        dev_ms, remeshed, spilled = info["ms"], info["remeshed"], info["spilled"]
        for _ in range(args.warmup):
            b.velocity(use_graph=True)
        t = time.perf_counter()
        for _ in range(args.steps):
            vinfo = b.velocity(use_graph=True)
        wall_v = time.perf_counter() - t
        errs, remesh_exact = [], True
        check = list(range(0, count, max(1, count // 8)))
        for i in check:
            s = systems[i]
            v = b.get_velocity(i)
            sub = W.target_subset(s["x"].size, 1024, seed=i)
            xopt, cutoff = W.oseen_params(s["sigma"])
            qr, qi = gauss.fmm2d_gpu(s["x"], s["y"], s["w"], s["x"][sub], s["y"][sub], 35, xopt,
                                     cutoff, 0.0, prec=gauss.PREC_F64)
            ref = np.concatenate([-qi / (2 * math.pi), -qr / (2 * math.pi)])
            got = np.concatenate([v[0][sub], v[1][sub]])
            errs.append(float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
            h = s["sigma"] * overlap
            one = Regrid(s["x"] + dt * v[0], s["y"] + dt * v[1], s["w"], s["sigma"], overlap,
                         [-1.5 * h, 0], [-1.5 * h, 0])
            remesh_exact &= all(np.array_equal(p, q) for p, q in zip(b.get_remeshed(i), one))
    batched = total / (wall / args.steps)
    # baseline 1: one-shot float64 step per system (host buffers, per-call copies)
    t = time.perf_counter()
    for s in systems:
        xopt, cutoff = W.oseen_params(s["sigma"])
        qr, qi = gauss.fmm2d_gpu(s["x"], s["y"], s["w"], s["x"], s["y"], 35, xopt, cutoff, 1e-6,
                                 prec=gauss.PREC_F64, order=16, leaf=64)
        h = s["sigma"] * overlap
        Regrid(s["x"] + dt * (-qi / (2 * math.pi)), s["y"] + dt * (-qr / (2 * math.pi)), s["w"],
               s["sigma"], overlap, [-1.5 * h, 0], [-1.5 * h, 0])
    t_oneshot = time.perf_counter() - t
    # baseline 2: resident float64 velocity, systems one after another
    resident = [BlobSystem(s["x"], s["y"], s["w"]) for s in systems]
    ps = [vel_params(s["sigma"], prec=gauss.PREC_F64, order=16, leaf=64) for s in systems]
    for r, p in zip(resident, ps):
        r.velocity(p, fetch=False)
    _native.lib().phf_synchronize()
    t = time.perf_counter()
    for r, p in zip(resident, ps):
        r.velocity(p, fetch=False)
    t_res = time.perf_counter() - t
    for r in resident:
        r.close()
    return {"metric": "aggregate blob step evals/s (velocity + M4' remesh), 256 uneven systems, "
                      "batched F32C graph vs one-at-a-time F64",
            "value": batched, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "%d systems, N log-uniform [1024, 262144] seed 0, total %d blobs"
                                   % (count, total),
                       "order": batch_order(systems),
                       "graph_device_ms": dev_ms, "remeshed_blobs": remeshed,
                       "spilled_systems": spilled,
                       "velocity_only_evals_per_s": total / (wall_v / args.steps),
                       "velocity_only_graph_device_ms": vinfo["ms"],
                       "f64_one_shot_step_evals_per_s": total / t_oneshot,
                       "f64_resident_sequential_velocity_evals_per_s": total / t_res,
                       "speedup_vs_f64_one_shot": batched / (total / t_oneshot),
                       "speedup_vs_f64_resident_sequential_velocity": batched / (total / t_res),
                       "rel_l2_vs_fp64_max": max(errs), "checked_systems": len(check),
                       "remesh_bitexact_vs_single_system": bool(remesh_exact)}}


def batch_order(systems, tol=1e-6):
    """One expansion order for a batch (the batched FMM takes one): the
    largest the measured table gives any of its systems at this tolerance."""
    return max(autotune.tuned_params(s["sigma"], s["x"].size, gauss.PREC_F32C, tol).order
               for s in systems)


def bench_batched(args):
    """The batched half of the 256-system mix alone (no float64 one-at-a-time
    legs): the same systems as bench_mix, velocity-only graph replays, then
    step replays (velocity + Euler + M4' remesh); error of every 32nd system
    against the float64 direct sum on 1,024 seeded targets."""
    from phyflow_amd.batch import BatchSystems
    count = args.systems
    sizes = W.uneven_sizes(count, 1024, 262144, seed=0)
    systems = [W.uniform_system(int(n), seed=1000 + i) for i, n in enumerate(sizes)]
    total = int(np.sum(sizes))
    with BatchSystems(sizes, [s["sigma"] for s in systems], prec=gauss.PREC_F32C,
                      order=batch_order(systems), leaf=64) as b:
        for i, s in enumerate(systems):
            b.set_system(i, s["x"], s["y"], s["w"])
        for _ in range(args.warmup):
            b.velocity(use_graph=True)
        t = time.perf_counter()
        for _ in range(args.steps):
            vinfo = b.velocity(use_graph=True)
        wall_v = time.perf_counter() - t
        errs = []
        for i in range(0, count, max(1, count // 8)):
            s = systems[i]
            v = b.get_velocity(i)
            sub = W.target_subset(s["x"].size, 1024, seed=i)
            xopt, cutoff = W.oseen_params(s["sigma"])
            qr, qi = gauss.fmm2d_gpu(s["x"], s["y"], s["w"], s["x"][sub], s["y"][sub], 35, xopt,
                                     cutoff, 0.0, prec=gauss.PREC_F64)
            ref = np.concatenate([-qi / (2 * math.pi), -qr / (2 * math.pi)])
            got = np.concatenate([v[0][sub], v[1][sub]])
            errs.append(float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
        b.set_remesh(0.8, dt=1e-3)
        for _ in range(args.warmup):
            b.step(use_graph=True)
        lat = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            a = time.perf_counter()
            info = b.step(use_graph=True)
            lat.append(time.perf_counter() - a)
        wall = time.perf_counter() - t0
    return {"metric": "aggregate blob step evals/s (velocity + M4' remesh), 256 uneven systems, "
                      "batched F32C graph",
            "value": total / (wall / args.steps), "unit": "evals/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "%d systems, N log-uniform [1024, 262144] seed 0, total %d blobs"
                                   % (count, total),
                       "order": batch_order(systems),
                       "graph_device_ms": info["ms"], "velocity_ms": 1e3 * wall_v / args.steps,
                       "velocity_graph_device_ms": vinfo["ms"],
                       "velocity_evals_per_s": total / (wall_v / args.steps),
                       "rel_l2_vs_fp64_max": max(errs), "checked_systems": len(errs)}}


def bench_serve(args):
    """Served step requests: Poisson arrivals at --rate per second for
    --duration s over --sims resident simulations (N log-uniform in
    [1e4, 2e6], seed 0, uniform blobs, h = 1/sqrt(N), sigma = h/0.8).  Each
    request is one RK2 step (two velocity evaluations), an M4' remesh onto
    the lattice of spacing h and population control (thresholds 1e-8, the
    Blobs defaults); 80% of requests ask for 1e-6 (F32C), 20% for 1e-12
    (float64), each with the FMM settings of the measured table for its
    precision, tolerance and size (--fixed-params: order 16 / order 30 with
    ws 2).  Each class's error against the float64 direct sum is checked on
    the smallest and the largest simulation first.  Latency = submission ->
    completion."""
    from phyflow_amd.serve import Server
    rng = np.random.default_rng(0)
    sizes = np.floor(np.exp(rng.uniform(math.log(1e4), math.log(2e6), args.sims))).astype(int)
    systems, params, remesh, data = [], [], [], []
    for i, n in enumerate(sizes):
        s = W.uniform_system(int(n), seed=2000 + i)
        data.append(s)
        systems.append(BlobSystem(s["x"], s["y"], s["w"]))
        if args.fixed_params:
            params.append((vel_params(s["sigma"], prec=gauss.PREC_F32C, order=16, leaf=64),
                           vel_params(s["sigma"], prec=gauss.PREC_F64, order=30, leaf=64, ws=2)))
        else:
            params.append((autotune.tuned_params(s["sigma"], int(n), prec=gauss.PREC_F32C, tol=1e-6),
                           autotune.tuned_params(s["sigma"], int(n), prec=gauss.PREC_F64,
                                                 tol=1e-12)))
        remesh.append(dict(sigma=s["sigma"], overlap=0.8, population_control=(1e-8, 1e-8)))
    checks = []
    for i in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        s = data[i]
        sub = W.target_subset(s["x"].size, 1024, seed=i)
        xopt, cutoff = W.oseen_params(s["sigma"])
        qr, qi = gauss.fmm2d_gpu(s["x"], s["y"], s["w"], s["x"][sub], s["y"][sub], 35, xopt, cutoff,
                                 0.0, prec=gauss.PREC_F64)
        ref = np.concatenate([-qi / (2 * math.pi), -qr / (2 * math.pi)])
        for cls, p in (("f32c", params[i][0]), ("f64", params[i][1])):
            systems[i].velocity(p, fetch=False)
            vx, vy = systems[i].last_velocity()
            got = np.concatenate([vx[sub], vy[sub]])
            checks.append({"n": int(sizes[i]), "class": cls, "order": p.order, "leaf": p.leaf,
                           "ws": p.ws, "leaf_side": p.leaf_side,
                           "rel_l2": float(np.linalg.norm(got - ref) / np.linalg.norm(ref))})
    dt = 1e-4
    with Server(slots=args.slots) as srv:
        ids = [srv.add(b) for b in systems]
        for k in ids:  # warm-up: one step of each simulation, both precisions
            srv.wait(srv.submit(k, params[k - ids[0]][0], dt, remesh=remesh[k - ids[0]]))
            srv.wait(srv.submit(k, params[k - ids[0]][1], dt, remesh=remesh[k - ids[0]]))
        srv.stats(reset=True)
        t0 = time.perf_counter()
        t_next, rids, classes, evals = 0.0, [], [], 0
        last = t0
        while True:
            t_next += rng.exponential(1.0 / args.rate)
            if t_next > args.duration:
                break
            now = time.perf_counter() - t0
            if t_next > now:
                time.sleep(t_next - now)
            k = int(rng.integers(len(ids)))
            hi = rng.random() < 0.2
            rids.append(srv.submit(ids[k], params[k][1 if hi else 0], dt, remesh=remesh[k]))
            classes.append(hi)
            if time.perf_counter() - last > 10.0:
                last = time.perf_counter()
                print("serve: %.0f s, %d submitted, %d done" % (last - t0, len(rids),
                      srv.stats()["completed"]), file=sys.stderr, flush=True)
        res = [srv.wait(r) for r in rids]
        st = srv.stats()
        for i in ids:
            srv.remove(i)
    for b in systems:
        b.close()
    lat = np.array([r["latency_ms"] for r in res])
    queue = np.array([r["queue_ms"] for r in res])
    run = lat - queue
    nsys = np.array([r["info"]["n_before"] for r in res])
    cls = np.array(classes)
    evals = float(np.sum(2 * nsys))
    # what the slowest 1% are: waiting for a slot, or long to run (class, size)
    top = lat >= np.percentile(lat, 99)
    big = nsys >= np.percentile(sizes, 75)
    tail = {"p99_ms": float(np.percentile(lat, 99)),
            "top1pct_count": int(top.sum()),
            "top1pct_fp64_share": float(cls[top].mean()),
            "top1pct_mean_queue_ms": float(queue[top].mean()),
            "top1pct_mean_run_ms": float(run[top].mean()),
            "top1pct_mean_n": float(nsys[top].mean()),
            "run_ms_p50": float(np.percentile(run, 50)), "run_ms_p99": float(np.percentile(run, 99)),
            "queue_ms_p50": float(np.percentile(queue, 50)),
            "queue_ms_p99": float(np.percentile(queue, 99)),
            "run_ms_p50_fp64_largest_quartile": float(np.percentile(run[cls & big], 50))
            if (cls & big).any() else None,
            "run_ms_p50_mixed_largest_quartile": float(np.percentile(run[~cls & big], 50))
            if (~cls & big).any() else None}
    return {"metric": "served RK2 step requests: throughput and p50/p99 step latency",
            "value": len(res) / st["span_s"], "unit": "steps/s", "n_gpus": 1, "steps": len(res),
            "warmup": 2 * len(ids), "ms_per_step": float(np.mean(lat)),
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic",
            "config": {"workload": "%d resident simulations, N log-uniform [1e4, 2e6] (total %d), "
                                   "Poisson arrivals %.0f/s for %.0f s, %d slots"
                                   % (len(sizes), int(np.sum(sizes)), args.rate, args.duration,
                                      args.slots),
                       "evals_per_s": evals / st["span_s"],
                       "p50_ms_mixed": float(np.percentile(lat[~cls], 50)) if (~cls).any() else None,
                       "p99_ms_mixed": float(np.percentile(lat[~cls], 99)) if (~cls).any() else None,
                       "p50_ms_fp64": float(np.percentile(lat[cls], 50)) if cls.any() else None,
                       "p99_ms_fp64": float(np.percentile(lat[cls], 99)) if cls.any() else None,
                       "tail": tail,
                       "params": "fixed" if args.fixed_params else "autotune table",
                       "class_checks": checks,
                       "completed": st["completed"], "failed": st["failed"],
                       "hbm_used_gb": st["hbm_used"] / 1e9}}


def bench_copy(args):
    """Calibration for HBM byte counters: a plain device copy of a 1 GiB
    float32 buffer (torch's copy kernel); every launch reads and writes
    exactly 2^30 bytes."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / args.steps
    return {"metric": "device copy bandwidth (byte-counter calibration)",
            "value": 2 * 4 * n / dt / 1e9, "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * dt, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "copy of 2^28 float32", "bytes_read_per_launch": 4 * n,
                       "bytes_written_per_launch": 4 * n}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["fmm1m", "lamb", "cylinder", "mix", "batched", "serve",
                                           "copy", "mixleg"],
                    default="fmm1m")
    ap.add_argument("--sims", type=int, default=16)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--rate", type=float, default=100.0)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--fixed-params", action="store_true",
                    help="serve: order 16 / order 30 ws 2 instead of the measured table")
    ap.add_argument("--systems", type=int, default=256)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 100 for fmm1m, 30 otherwise)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps first (default 20 for fmm1m, 5 otherwise)")
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--order", type=int, default=16)
    ap.add_argument("--leaf", type=int, default=64)
    ap.add_argument("--prec", choices=["f32c", "f64"], default="f32c")
    ap.add_argument("--check", type=int, default=65536, help="targets for the error check (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mix", action="store_true",
                    help="skip the 256-system batched vs one-at-a-time float64 leg")
    ap.add_argument("--variants", action="store_true", help="also time p in {8,16,24} x {f32c,f64}")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 100 if args.workload == "fmm1m" else 30
    if args.warmup is None:
        args.warmup = 20 if args.workload == "fmm1m" else 5
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    if args.workload == "lamb":
        print(json.dumps(bench_lamb(args)), flush=True)
        return 0
    if args.workload == "cylinder":
        print(json.dumps(bench_cylinder(args)), flush=True)
        return 0
    if args.workload == "mix":
        print(json.dumps(bench_mix(args)), flush=True)
        return 0
    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return 0
    if args.workload == "serve":
        print(json.dumps(bench_serve(args)), flush=True)
        return 0
    if args.workload == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return 0
    if args.workload == "mixleg":
        print(json.dumps(mix_leg(args)), flush=True)
        return 0

    s = W.uniform_system(args.n, seed=0, h=1.0 / 1024 if args.n == 1048576 else None)
    prec = gauss.PREC_F32C if args.prec == "f32c" else gauss.PREC_F64
    params = vel_params(s["sigma"], k=2.0, prec=prec, order=args.order, leaf=args.leaf)
    dev = _native.device_info()
    with BlobSystem(s["x"], s["y"], s["w"]) as sysm:
        # untimed settle: about 0.3 s of evaluations, so that the warm-up and
        # timed steps run at the clocks a sustained stream of steps gets
        t_settle = time.perf_counter()
        while time.perf_counter() - t_settle < 0.3:
            sysm.velocity(params, fetch=False)
        _native.lib().phf_synchronize()
        wall, lat, p2p_ms, info = timed_steps(sysm, params, args.warmup, args.steps)
        ms_step = 1e3 * wall / args.steps
        # the error check after the timed steps (the CPU oracle takes seconds)
        oref = OracleReference(s, args.check, 0) if args.check else None
        rel, mx = rel_errors(sysm, params, oref) if oref else (None, None)
        # per-stage event times of one eager evaluation (every kernel
        # launched on its own; the far field still runs beside the near field)
        sysm.velocity(params, fetch=False, info=True, eager=True)
        eager_ms = sysm.last_info["ms"]
        # the same configuration in float64
        p64 = vel_params(s["sigma"], k=2.0, prec=gauss.PREC_F64, order=args.order, leaf=args.leaf)
        w64, lat64, _, _ = timed_steps(sysm, p64, 3, 10, counts=False)
        e64 = rel_errors(sysm, p64, oref) if oref else (None, None)
        f64_same = {"prec": "f64", "order": args.order, "ms_per_step": 1e2 * w64,
                    "evals_per_s": args.n / (w64 / 10), "rel_l2": e64[0], "max_err": e64[1]}
        # the measured settings table's choice for this size, precision and
        # tolerance, timed and checked the same way
        tp = autotune.tuned_params(s["sigma"], args.n, prec=prec, tol=1e-6)
        tw, tlat, _, tinfo = timed_steps(sysm, tp, 5, 50, counts=False)
        te = rel_errors(sysm, tp, oref) if oref else (None, None)
        tuned = {"table": "phyflow_amd/autotune_table.json", "tol": 1e-6, "order": tp.order,
                 "leaf": tp.leaf, "ws": tp.ws, "leaf_side": tp.leaf_side, "levels": tinfo["L"],
                 "near_boxes": tinfo["nnear"], "ms_per_step": 1e3 * tw / 50,
                 "p50_ms": float(np.percentile(tlat, 50) * 1e3),
                 "p99_ms": float(np.percentile(tlat, 99) * 1e3),
                 "evals_per_s": args.n / (tw / 50), "rel_l2": te[0], "max_err": te[1]}
        variants = []
        if args.variants:
            for pr in ("f32c", "f64"):
                for p in (8, 16, 24):
                    vp = vel_params(s["sigma"], prec=gauss.PREC_F32C if pr == "f32c" else
                                    gauss.PREC_F64, order=p, leaf=args.leaf)
                    vw, vl, _, _ = timed_steps(sysm, vp, 2, 5, counts=False)
                    e = rel_errors(sysm, vp, oref) if oref else (None, None)
                    variants.append({"prec": pr, "order": p, "ms_per_step": 1e3 * vw / 5,
                                     "evals_per_s": args.n / (vw / 5), "rel_l2": e[0], "max_err": e[1]})
    # the one-shot entry point with host arrays (copies over PCIe both ways
    # inside each call): reported beside the resident rate, never as value
    xopt, cutoff = W.oseen_params(s["sigma"])
    gauss.fmm2d_gpu(s["x"], s["y"], s["w"], s["x"], s["y"], 35, xopt, cutoff, 1e-6, prec=prec,
                    order=args.order, leaf=args.leaf)
    t = time.perf_counter()
    for _ in range(3):
        gauss.fmm2d_gpu(s["x"], s["y"], s["w"], s["x"], s["y"], 35, xopt, cutoff, 1e-6,
                        prec=prec, order=args.order, leaf=args.leaf)
    pcie_rate = 3 * args.n / (time.perf_counter() - t)
    t_p2p = float(np.nanmean(p2p_ms)) * 1e-3
    pairs = info["p2p_pairs"]
    achieved = FLOP_PER_PAIR * pairs / t_p2p / 1e12
    peak = FP32_PEAK_TFLOPS if prec == gauss.PREC_F32C else FP64_PEAK_TFLOPS
    s_bytes = 4 if prec == gauss.PREC_F32C else 8
    b_tree = args.n * (6 * s_bytes + 4)  # read x, y, w; write permuted x, y, w + int32 index
    t_tree = info["ms"].get("tree", float("nan")) * 1e-3
    m2l_flop = 8 * (args.order + 1) ** 2 * info["m2l_links"]
    t_m2l = eager_ms.get("m2l", float("nan")) * 1e-3
    out = {
        "metric": "blob vel. evals/s @1M blobs, rel err<=1e-6 vs FP64; p50/p99 step latency",
        "value": args.n / (wall / args.steps),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == gauss.PREC_F32C else "f64",
        "data": "synthetic",
        "rel_l2_vs_fp64": rel,
        "max_err_vs_fp64": mx,
        "error_reference": ("CPU oracle: float64 direct Oseen sum (restatement of "
                            "direct.cpp:1409-1485) at %d seeded targets, %d host threads, %.1f s"
                            % (args.check, host_cores(), oref.seconds)) if oref else None,
        "config": {
            "workload": "FMM velocity of %d uniform blobs at their centres (seed 0, h=1/1024, "
                        "sigma=h/0.8), order %d, leaf %d, Oseen cutoff 5*xopt, %s; error vs "
                        "float64 direct sum on %d seeded targets"
                        % (args.n, args.order, args.leaf, args.prec, args.check),
            "n_blobs": args.n, "order": args.order, "leaf": args.leaf, "levels": info["L"],
            "near_boxes": info["nnear"], "p2p_pairs": pairs, "m2l_links": info["m2l_links"],
            "stage_ms": info["ms"], "eager_stage_ms": eager_ms, "device": dev["name"],
            "pcie_inclusive_evals_per_s": pcie_rate,
            "f64_same_config": f64_same,
            "autotuned": tuned,
        },
        "roofline": {
            "kernel": "p2p_near_f32s" if prec == gauss.PREC_F32C else "p2p_near_f64s",
            "bound": "valu",
            "compute_pipe": ("fp32" if prec == gauss.PREC_F32C else "fp64") +
                            " vector ALU (the pair kernel has no matrix shape); MI355X vector "
                            "peak fp32 157.3 TFLOP/s, fp64 78.6",
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": None,
            "flop_per_pair": FLOP_PER_PAIR,
            "kernel_ms": t_p2p * 1e3,
        },
        "roofline_other": {
            "tree": {"kernels": "tree graph (bbox, keys, 2 radix passes, leaf offsets)",
                     "bound": "hbm", "algorithmic_bytes": b_tree,
                     "bytes_rule": "N*(6s+4), s=%d" % s_bytes, "ms": t_tree * 1e3,
                     "achieved": b_tree / t_tree / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": b_tree / t_tree / 1e9 / HBM_PEAK_GBS},
            "m2l": {"kernel": "m2l_box_kernel" if prec == gauss.PREC_F32C else "m2l_kernel",
                    "bound": "valu (no MFMA shape is used)", "flop": m2l_flop,
                    "flop_rule": "8(p+1)^2 per far link", "ms": t_m2l * 1e3,
                    "note": "eager event time, overlapping the near field",
                    "achieved": m2l_flop / t_m2l / 1e12, "peak": peak, "unit": "TFLOP/s",
                    "frac": m2l_flop / t_m2l / 1e12 / peak},
        },
    }
    iss = measured_issue(out["roofline"]["kernel"], args)
    if iss:
        out["roofline"]["valu_issue_frac"] = iss["valu_issue_frac"]
        out["roofline"]["lane_instructions_per_pair"] = iss["lane_instructions_per_pair"]
        out["roofline"]["issue_source"] = iss["source"]
        out["roofline"]["issue_note"] = (
            "vector-issue bound: the peak counts packed FP32 (two lanes per slot), which "
            "this build does not emit, and counts every slot as an FMA")
    tr = measured_traffic(out["roofline"]["kernel"], args)
    if tr:
        out["roofline"]["traffic"] = tr["bytes_per_launch"]
        out["roofline"]["traffic_source"] = tr["source"]
    if variants:
        out["variants"] = variants
    if not args.no_mix:
        out["mix_256_systems"] = mix_leg_child()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(s)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
