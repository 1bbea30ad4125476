"""The coupled RK2 step on resident blobs.

Pins: Heun's second order (error ratio ~4 per halving of dt on a vortex
pair whose exact motion is known), free-stream translation, one coupled
blob/panel step against a CPU restatement (oracle direct Oseen sum, float64
least squares, oracle panel velocities) to 1e-11 in float64 and to the 1e-6
mixed contract in F32C with the grouped panel series, and exact conservation of
circulation/impulse by the remesh (float64, <= 1e-13 relative)."""
import math

import numpy as np
import pytest

from phyflow_amd import workloads as W
from phyflow_amd.lagrangian import PanelBody, step_rk2
from phyflow_amd.system import BlobSystem, vel_params

pytestmark = pytest.mark.gpu


def test_heun_second_order_on_vortex_pair():
    d, G, sigma = 1.0, 2.0 * math.pi, 1e-3
    T = 0.3
    omega = G / (math.pi * d * d)  # angular velocity of a co-rotating pair
    exact = 0.5 * d * np.array([math.cos(omega * T), math.sin(omega * T)])
    errs = []
    for nsteps in (8, 16, 32):
        with BlobSystem([0.5 * d, -0.5 * d], [0.0, 0.0], [G, G]) as s:
            p = vel_params(sigma, method=1)  # direct float64
            for _ in range(nsteps):
                step_rk2(s, p, T / nsteps)
            x, y, _ = s.get()
        errs.append(math.hypot(x[0] - exact[0], y[0] - exact[1]))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.5 < r1 < 4.5 and 3.5 < r2 < 4.5, errs


def test_free_stream_translation():
    rng = np.random.default_rng(0)
    x, y = rng.random((2, 1000))
    with BlobSystem(x, y, np.zeros(1000)) as s:
        step_rk2(s, vel_params(0.01), 0.1, vInf=(1.5, -0.5))
        x1, y1, _ = s.get()
    assert np.allclose(x1, x + 0.15, rtol=0, atol=1e-15)
    assert np.allclose(y1, y - 0.05, rtol=0, atol=1e-15)


def _cpu_stage(O, x, y, w, sigma, g, s_panel, vinf):
    xopt, cutoff = W.oseen_params(sigma)
    qr, qi = O.direct_oseen(x, y, w, x, y, xopt, cutoff)
    vx, vy = -qi / (2 * math.pi), -qr / (2 * math.pi)
    px, py = O.panel_induced(s_panel, x, y, g["xs"], g["ys"], g["xe"], g["ye"], g["cosA"],
                             g["sinA"])
    return vx + px + vinf[0], vy + py + vinf[1]


def _cpu_step(O, a, g, sigma, dt, vinf):
    """CPU restatement of one coupled step: slip at the collocation points,
    least squares with the Kelvin row, Heun with oracle blob and panel
    velocities."""
    xopt, cutoff = W.oseen_params(sigma)
    A = O.panel_influence(g["xcp"], g["ycp"], g["xs"], g["ys"], g["xe"], g["ye"], g["cosA"],
                          g["sinA"])
    qr, qi = O.direct_oseen(a["x"], a["y"], a["w"], g["xcp"], g["ycp"], xopt, cutoff)
    vcx, vcy = -qi / (2 * math.pi) + vinf[0], -qr / (2 * math.pi) + vinf[1]
    rhs = -vcx * g["cosA"] - vcy * g["sinA"]
    sp = np.linalg.lstsq(np.vstack([A, g["L"][None, :]]), np.concatenate([rhs, [0.0]]),
                         rcond=None)[0]
    k1 = _cpu_stage(O, a["x"], a["y"], a["w"], sigma, g, sp, vinf)
    x1, y1 = a["x"] + dt * k1[0], a["y"] + dt * k1[1]
    k2 = _cpu_stage(O, x1, y1, a["w"], sigma, g, sp, vinf)
    return a["x"] + 0.5 * dt * (k1[0] + k2[0]), a["y"] + 0.5 * dt * (k1[1] + k2[1])


def test_coupled_step_matches_cpu_restatement(O):
    m, n, sigma, dt, vinf = 64, 2000, 0.02, 0.01, (1.0, 0.0)
    xp, yp = W.cylinder_panels(m, R=0.5)
    g = W.panel_geometry(xp, yp)
    a = W.annulus_blobs(n, gamma_std=1e-2, sigma=sigma, seed=1)
    body = PanelBody(xp, yp, "prescribed", 0.0)
    xc, yc = body.collocation_points()
    np.testing.assert_allclose(xc, g["xcp"], rtol=0, atol=1e-15)
    with BlobSystem(a["x"], a["y"], a["w"]) as s:
        info = step_rk2(s, vel_params(sigma, tol=0.0), dt, vInf=vinf, panels=body)
        xg, yg, _ = s.get()
    xr, yr = _cpu_step(O, a, g, sigma, dt, vinf)
    dx = np.max(np.abs(np.concatenate([xg - xr, yg - yr])))
    disp = np.max(np.abs(np.concatenate([xr - a["x"], yr - a["y"]])))
    assert dx <= 1e-11 * disp, (dx, disp, info)


@pytest.mark.parametrize("params,tol", [
    (dict(prec=1, order=16), 1e-6),  # F32C FMM, float32 panel series
    (dict(tol=0.0), 1e-10),          # float64 direct sums, float64 panel series
])
def test_coupled_step_grouped_panels(params, tol, O):
    """Enough blobs for the grouped panel path (512 panels in groups of 16,
    series factors cached with the body): the displacement beyond the free
    stream agrees with the float64 CPU restatement to the mixed contract
    (1e-6 relative L2) in F32C and to 1e-10 in float64."""
    m, n, sigma, dt, vinf = 512, 20000, 0.01, 0.01, (1.0, 0.0)
    xp, yp = W.cylinder_panels(m, R=0.5)
    g = W.panel_geometry(xp, yp)
    a = W.annulus_blobs(n, gamma_std=1e-2, sigma=sigma, seed=4)
    body = PanelBody(xp, yp, "prescribed", 0.0)
    with BlobSystem(a["x"], a["y"], a["w"]) as s:
        step_rk2(s, vel_params(sigma, **params), dt, vInf=vinf, panels=body)
        xg, yg, _ = s.get()
    xr, yr = _cpu_step(O, a, g, sigma, dt, vinf)
    ref = (xr - a["x"] - vinf[0] * dt, yr - a["y"] - vinf[1] * dt)
    got = (xg - a["x"] - vinf[0] * dt, yg - a["y"] - vinf[1] * dt)
    assert O.rel_l2(got, ref) <= tol


def test_remesh_conserves_invariants():
    a = W.annulus_blobs(50000, gamma_std=1e-4, sigma=0.004, seed=2)
    with BlobSystem(a["x"], a["y"], a["w"]) as s:
        info = step_rk2(s, vel_params(0.004, prec=1, order=16), 1e-3, vInf=(1.0, 0.0),
                        remesh=dict(sigma=0.004, overlap=1.25))
    tot = np.sum(np.abs(a["w"]))
    assert abs(info["circ_after"] - info["circ_before"]) <= 1e-13 * tot
    scale = tot * 1.5
    assert abs(info["impulse_x_after"] - info["impulse_x_before"]) <= 1e-13 * scale
    assert abs(info["impulse_y_after"] - info["impulse_y_before"]) <= 1e-13 * scale
    assert info["n_after"] > 0


def test_async_body_update_matches_sync():
    """phf_panels_update_async enqueues the rebuild and the step waits only
    where it needs the new body (collocation points on the device, the
    factors before the panel solve), so the blob velocity evaluation runs
    beside the factorisation: the steps must come out bit for bit as after
    the synchronous update, and solve() and the collocation points after an
    asynchronous update must see the new body."""
    xp, yp = W.cylinder_panels(128, R=0.5)
    a = W.annulus_blobs(20000, sigma=0.01, seed=5)
    p = vel_params(0.01, prec=1, order=16, leaf=64)
    rem = dict(sigma=0.01, overlap=1.25)
    out = []
    for wait in (True, False):
        body = PanelBody(xp, yp, "prescribed", 0.0)
        try:
            with BlobSystem(a["x"], a["y"], a["w"]) as s:
                for k in range(3):
                    th = 0.05 * (k + 1)
                    c, sn = math.cos(th), math.sin(th)
                    body.update_geometry(c * xp - sn * yp + 0.01 * k, sn * xp + c * yp, wait=wait)
                    info = step_rk2(s, p, 1e-3, vInf=(1.0, 0.0), panels=body, remesh=rem)
                out.append((s.get(), info["solve_iterations"], info["circ_after"]))
            # a solve and the collocation points right after an async update
            body.update_geometry(xp + 0.2, yp, wait=wait)
            xc, yc = body.collocation_points()
            out.append((xc, yc, body.solve(np.ones(body.m), np.zeros(body.m))))
        finally:
            body.close()
    (g_sync, it_sync, c_sync), pts_sync, (g_async, it_async, c_async), pts_async = out
    for u, v in zip(g_sync, g_async):
        assert np.array_equal(u, v)
    assert it_sync == it_async and c_sync == c_async
    for u, v in zip(pts_sync, pts_async):
        assert np.array_equal(u, v)
    # the collocation points are those of the moved body
    xr, yr = W.cylinder_panels(128, R=0.5)
    assert abs(float(np.mean(pts_async[0])) - (float(np.mean(0.5 * (xr[:-1] + xr[1:]))) + 0.2)) < 0.05
