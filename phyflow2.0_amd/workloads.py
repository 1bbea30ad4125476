"""Seeded synthetic workloads of the benchmark configurations.

All generators use numpy.random.default_rng(seed) (PCG64) and draw
positions first, then circulations, then any subset (from seed + 1).

  lamb_oseen_lattice   64x64 Gaussian blobs on [-1,1]^2 carrying a
                       Lamb-Oseen vortex (the pattern of the reference's
                       examples/hybrid/ex01_lambOseen_test.py:36-89)
  uniform_system       N blobs uniform in [0,1]^2, N(0,1)*h^2 circulations,
                       h = 1/sqrt(N), sigma = h/0.8
  cylinder_step        512-panel cylinder (examples/panels/ex01 recipe) and
                       blobs in the annulus 0.5 < r < 1.5
  uneven_mix           256 independent systems, N log-uniform in
                       [1024, 262144]
"""
import math

import numpy as np

ALPHA_OSEEN = 1.2564312086261696770  # _induced.py:236, fmm.cpp:1171


def oseen_params(sigma, k=2.0):
    """(xopt, cutoff) the reference passes to fmm2d (_induced.py:236-240)."""
    ks = k * sigma * sigma
    xopt = math.sqrt(ALPHA_OSEEN * ks)
    return xopt, 5.0 * xopt


def lamb_oseen_lattice(n_side=64, gamma0=1.0, c=0.5):
    """Cell-centred n_side^2 lattice on [-1,1]^2, h = 2/n_side, sigma = h/0.8,
    circulation h^2 * omega(r), omega = gamma0/(pi c^2) exp(-r^2/c^2)."""
    h = 2.0 / n_side
    g = -1.0 + (np.arange(n_side) + 0.5) * h
    X, Y = np.meshgrid(g, g, indexing="ij")
    x = X.ravel().copy()
    y = Y.ravel().copy()
    r2 = x * x + y * y
    w = h * h * gamma0 / (math.pi * c * c) * np.exp(-r2 / (c * c))
    sigma = h / 0.8
    return dict(x=x, y=y, w=w, h=h, sigma=sigma, k=2.0)


def lamb_oseen_velocity(x, y, gamma0=1.0, c=0.5):
    """Exact velocity of the continuous Lamb-Oseen vortex."""
    r2 = x * x + y * y
    with np.errstate(invalid="ignore", divide="ignore"):
        f = np.where(r2 > 0, gamma0 / (2 * math.pi * r2) * (1 - np.exp(-r2 / (c * c))), 0.0)
    return -f * y, f * x


def uniform_system(n, seed=0, h=None):
    rng = np.random.default_rng(seed)
    xy = rng.random((n, 2))
    h = (1.0 / math.sqrt(n)) if h is None else h
    w = rng.standard_normal(n) * h * h
    x = np.ascontiguousarray(xy[:, 0])
    y = np.ascontiguousarray(xy[:, 1])
    return dict(x=x, y=y, w=w, h=h, sigma=h / 0.8, k=2.0)


def target_subset(n, m, seed=0):
    """Seeded subset of m target indices out of n (stream seed + 1)."""
    rng = np.random.default_rng(seed + 1)
    if m >= n:
        return np.arange(n)
    return np.sort(rng.choice(n, m, replace=False))


def uneven_sizes(count=256, lo=1024, hi=262144, seed=0):
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(math.log(lo), math.log(hi), count))).astype(np.int64)


def cylinder_panels(n_panel=512, R=0.5, d_panel=None):
    """Panel end points of a cylinder of radius R by the reference recipe
    (examples/panels/ex01_solveNoSlip_test.py:26-38): theta from pi to -pi,
    end points at r*(cos, sin)(theta - dtheta/2), r = (R+dPanel)/cos(dtheta/2)."""
    d_panel = np.spacing(100) if d_panel is None else d_panel
    theta = np.linspace(np.pi, -np.pi, n_panel + 1)
    dtheta = theta[1] - theta[0]
    r = (R + d_panel) / np.cos(dtheta / 2.0)
    xp = r * np.cos(theta - dtheta / 2)
    yp = r * np.sin(theta - dtheta / 2)
    return xp, yp


def panel_geometry(xp, yp):
    """Start/end points, angles, normals and collocation points of closed
    panel polygons (__panelsClass.py:1036-1058): CP = mid - 0.1 L n."""
    xs, ys = xp[:-1].copy(), yp[:-1].copy()
    xe, ye = xp[1:].copy(), yp[1:].copy()
    Lx, Ly = xe - xs, ye - ys
    L = np.sqrt(Lx * Lx + Ly * Ly)
    cosA, sinA = Lx / L, Ly / L
    nx, ny = -sinA, cosA
    xcp = 0.5 * (xe + xs) - nx * L * 0.1
    ycp = 0.5 * (ye + ys) - ny * L * 0.1
    return dict(xs=xs, ys=ys, xe=xe, ye=ye, cosA=cosA, sinA=sinA, L=L, xcp=xcp, ycp=ycp)


def annulus_blobs(n=262144, r0=0.5, r1=1.5, gamma_std=1e-4, sigma=0.004, seed=0):
    rng = np.random.default_rng(seed)
    r = np.sqrt(rng.uniform(r0 * r0, r1 * r1, n))
    th = rng.uniform(0.0, 2 * math.pi, n)
    x = r * np.cos(th)
    y = r * np.sin(th)
    w = rng.normal(0.0, gamma_std, n)
    return dict(x=x, y=y, w=w, sigma=sigma, k=2.0)
