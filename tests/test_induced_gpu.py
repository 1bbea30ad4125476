"""velocity() dispatch (the Py3 restatement of _induced.py:34-166)."""
import math

import numpy as np
import pytest

from phyflow_amd import workloads as W
from phyflow_amd.blobs.base.induced import velocity

pytestmark = pytest.mark.gpu


def test_velocity_direct_and_fmm(O):
    s = W.uniform_system(20000, seed=4)
    x, y, w, sigma = s["x"], s["y"], s["w"], s["sigma"]
    ks = 2 * sigma * sigma
    ref = O.vv2par_gauss(x, y, w, x, y, ks)
    d = velocity(x, y, w, sigma, method=1, hardware=1)
    assert O.rel_l2(d, ref) <= 1e-12
    xopt, cutoff = W.oseen_params(sigma)
    qr, qi = O.direct_oseen(x, y, w, x, y, xopt, cutoff)
    f = velocity(x, y, w, sigma, method=0, hardware=1, tol=1e-10)
    assert O.rel_l2(f, (-qi / (2 * math.pi), -qr / (2 * math.pi))) <= 1e-9
    xe, ye = np.linspace(0, 1, 50), np.linspace(1, 0, 50)
    e = velocity(x, y, w, sigma, xEval=xe, yEval=ye, method=1)
    assert O.rel_l2(e, O.vv2par_gauss(x, y, w, xe, ye, ks)) <= 1e-12


def test_velocity_unsupported():
    x = np.zeros(3)
    with pytest.raises(Exception):
        velocity(x, x, x, 0.1, kernel=0, method=0)
    with pytest.raises(NotImplementedError):
        velocity(x, x, x, 0.1, kernel=0, method=1)
