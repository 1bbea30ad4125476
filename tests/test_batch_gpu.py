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


def _single_step(s, v, dt, vinf, overlap, c=0.0):
    """Same step through the single-system entry points: Euler convection in
    float64 (the batched kernel's operation order), then Regrid."""
    from phyflow_amd.blobs.base.regrid import Regrid
    xn = s["x"] + dt * (v[0] + vinf[0])
    yn = s["y"] + dt * (v[1] + vinf[1])
    h = s["sigma"] * overlap
    return Regrid(xn, yn, s["w"], s["sigma"], overlap, [-1.5 * h, 0], [-1.5 * h, 0], c=c), (xn, yn)


@pytest.mark.parametrize("cell_factor", [2.0, 0.01])
def test_batch_step_remesh_matches_single(cell_factor, O):
    """Velocity + Euler step + M4' remesh of every system in one sequence:
    each system's remeshed blobs are bit-identical to the single-system
    Regrid of the same convected positions (cell_factor 0.01 forces every
    non-tiny system through the single-system spill path)."""
    sizes, systems = _mix(9, seed=7)
    rng = np.random.default_rng(4)
    sizes = np.concatenate([sizes, [0, 1, 3]])
    systems.append(dict(x=np.empty(0), y=np.empty(0), w=np.empty(0), sigma=0.01))
    for n in (1, 3):
        x, y = rng.random((2, n))
        systems.append(dict(x=x, y=y, w=rng.standard_normal(n) * 1e-4, sigma=0.02))
    dt, vinf, overlap = 2e-3, (0.3, -0.2), 0.8
    with BatchSystems(sizes, [s["sigma"] for s in systems], prec=PREC_F32C) as b:
        for i, s in enumerate(systems):
            b.set_system(i, s["x"], s["y"], s["w"])
        b.set_remesh(overlap, dt=dt, vInf=vinf, cell_factor=cell_factor)
        info = b.step(use_graph=False)
        assert info["flags"] == 0
        # uniform systems need ~n + 8 sqrt(n) cells; the 3-blob system spread
        # over the unit square needs more than its reserve at any factor
        if cell_factor < 1:
            assert info["spilled"] >= sum(1 for n in sizes if n >= 1024)
        else:
            assert info["spilled"] <= 1
        total = 0
        for i, s in enumerate(systems):
            v = b.get_velocity(i)
            got = b.get_remeshed(i)
            ref, (xn, yn) = _single_step(s, v, dt, vinf, overlap)
            total += got[0].size
            for a, r in zip(got, ref[:3]):
                assert np.array_equal(a, r), i
            if 0 < s["x"].size <= 20000 and i % 3 == 0:
                h = s["sigma"] * overlap
                o = O.regrid_m4(xn, yn, s["w"], s["sigma"], overlap, -1.5 * h, -1.5 * h)
                assert np.array_equal(got[0], o[0]) and np.array_equal(got[1], o[1])
                assert np.max(np.abs(got[2] - o[2])) <= 1e-15 * np.max(np.abs(o[2]))
        assert total == info["remeshed"]


@pytest.mark.parametrize("c", [0.0, 0.3])
def test_batch_step_remesh_thin_and_crowded_systems(c):
    """Lattice shapes the batched gather walks in rows of four cells: rows
    of 4..11 cells (every remainder of the row length mod 4, so a thread's
    cells wrap into the next row), a one-cell-high line of blobs, a
    one-cell-wide column, and a cell holding 40 blobs.  Each system is
    remeshed by the batch (none spills) and is bit-identical to the
    single-system Regrid of the same convected positions, with plain M4'
    weights (c = 0) and with Wee's diffusion term (c = 0.3)."""
    rng = np.random.default_rng(12)
    sigma, overlap = 0.01, 0.8
    h = sigma * overlap
    systems = []
    for bj in range(1, 9):  # touched rows of bj + 3 = 4..11 cells
        n = 60
        x = (rng.integers(0, 12, n) + rng.random(n)) * h
        y = (rng.integers(0, bj, n) + rng.random(n)) * h
        systems.append((x, y))
    n = 200  # a line along x, one cell high; then a column along y
    systems.append(((np.arange(n) + rng.random(n)) * h, np.full(n, 0.3 * h)))
    systems.append((np.full(n, 0.6 * h), (np.arange(n) + rng.random(n)) * h))
    xc = np.concatenate([np.full(40, 5.5 * h) + rng.random(40) * 0.4 * h,
                         (rng.integers(0, 12, 60) + rng.random(60)) * h])
    yc = np.concatenate([np.full(40, 3.5 * h) + rng.random(40) * 0.4 * h,
                         (rng.integers(0, 12, 60) + rng.random(60)) * h])
    systems.append((xc, yc))
    systems = [dict(x=x, y=y, w=rng.standard_normal(x.size) * 1e-4, sigma=sigma)
               for x, y in systems]
    sizes = [s["x"].size for s in systems]
    dt, vinf = 1e-3, (0.0, 0.0)
    with BatchSystems(sizes, [s["sigma"] for s in systems], prec=PREC_F32C) as b:
        for i, s in enumerate(systems):
            b.set_system(i, s["x"], s["y"], s["w"])
        b.set_remesh(overlap, dt=dt, vInf=vinf, c=c)
        info = b.step(use_graph=False)
        assert info["flags"] == 0 and info["spilled"] == 0
        for i, s in enumerate(systems):
            v = b.get_velocity(i)
            got = b.get_remeshed(i)
            ref, _ = _single_step(s, v, dt, vinf, overlap, c=c)
            for a, r in zip(got, ref[:3]):
                assert np.array_equal(a, r), i


def test_many_tiny_systems_velocity_and_remesh(O):
    """120 systems of 0..90 blobs, so most waves of the batch's point kernels
    straddle systems and take the per-lane system search, and the leaf and
    box kernels mix systems of different depths: every system's velocity
    agrees with the float64 direct sum (<= 1e-6) and its remesh is
    bit-identical to the single-system Regrid of the same convected
    positions."""
    rng = np.random.default_rng(21)
    sizes = rng.integers(0, 91, 120)
    sizes[[0, 7, 63]] = 0
    systems = []
    for i, n in enumerate(sizes):
        sigma = 0.002 * (1.0 + rng.random())
        x, y = rng.random((2, n)) * 0.05
        systems.append(dict(x=x, y=y, w=rng.standard_normal(n) * 1e-4, sigma=sigma))
    # a coarse lattice (h = 4 sigma) keeps each system's touched cells
    # within its reserve, so the batch remeshes every system itself
    dt, vinf, overlap = 1e-3, (0.1, 0.0), 4.0
    with BatchSystems(sizes, [s["sigma"] for s in systems], prec=PREC_F32C, order=16,
                      leaf=64) as b:
        for i, s in enumerate(systems):
            b.set_system(i, s["x"], s["y"], s["w"])
        b.set_remesh(overlap, dt=dt, vInf=vinf)
        info = b.step(use_graph=False)
        assert info["flags"] == 0 and info["spilled"] == 0
        for i, s in enumerate(systems):
            v = b.get_velocity(i)
            if s["x"].size:
                xopt, cutoff = W.oseen_params(s["sigma"])
                qr, qi = O.direct_oseen(s["x"], s["y"], s["w"], s["x"], s["y"], xopt, cutoff)
                d = (-qi / (2 * math.pi), -qr / (2 * math.pi))
                assert O.rel_l2(v, d) <= 1e-6, i
            got = b.get_remeshed(i)
            ref, _ = _single_step(s, v, dt, vinf, overlap)
            for a, r in zip(got, ref[:3]):
                assert np.array_equal(a, r), i


def test_batch_step_graph_replays():
    sizes, systems = _mix(6, seed=11)
    with BatchSystems(sizes, [s["sigma"] for s in systems], prec=PREC_F32C) as b:
        for i, s in enumerate(systems):
            b.set_system(i, s["x"], s["y"], s["w"])
        b.set_remesh(0.8, dt=1e-3)
        b.step(use_graph=False)
        eager = [b.get_remeshed(i) for i in range(len(systems))]
        for rep in range(3):
            info = b.step(use_graph=True)
            assert info["graph"] and info["spilled"] == 0
            for i in range(len(systems)):
                g = b.get_remeshed(i)
                assert all(np.array_equal(a, e) for a, e in zip(g, eager[i])), (rep, i)
            b.velocity(use_graph=False)  # other eager work between replays
        # parameters change -> new graph, still agrees with eager
        b.set_remesh(1.0, dt=0.0)
        g1 = b.step(use_graph=True) and [b.get_remeshed(i) for i in range(len(systems))]
        b.step(use_graph=False)
        for i in range(len(systems)):
            assert all(np.array_equal(a, e) for a, e in zip(g1[i], b.get_remeshed(i)))
