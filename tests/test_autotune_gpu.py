"""The autotuner on a small system: its choice meets the tolerance against
the CPU oracle on other targets than it tuned on, and is the fastest of
the candidates it measured."""
import math

import numpy as np
import pytest

from phyflow_amd import autotune
from phyflow_amd import workloads as W
from phyflow_amd._native import PREC_F32C, PREC_F64
from phyflow_amd.system import BlobSystem

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,tol", [("f32c", 1e-6), ("f64", 1e-10)])
def test_tuned_choice_meets_tolerance(prec, tol, O):
    s = W.uniform_system(30000, seed=31)
    e = autotune.tune_system(s, prec, tol, n_check=512, reps=3, leaves=(32, 64, 128),
                             leaf_sides=(-1.0, 1.0), ws_set=(1, 2))
    c = e["choice"]
    assert c is not None
    assert c["ms"] == min(x["ms"] for x in e["candidates"] if x["order"] is not None)
    p = autotune.tuned_params(s["sigma"], s["x"].size, PREC_F32C if prec == "f32c" else PREC_F64,
                              tol, table={"entries": [e]})
    assert (p.order, p.leaf, p.ws) == (c["order"], c["leaf"], c["ws"])
    with BlobSystem(s["x"], s["y"], s["w"]) as sysm:
        vx, vy = sysm.velocity(p)
    sub = W.target_subset(s["x"].size, 800, seed=99)
    xopt, cutoff = W.oseen_params(s["sigma"])
    qr, qi = O.direct_oseen(s["x"], s["y"], s["w"], s["x"][sub], s["y"][sub], xopt, cutoff)
    err = O.rel_l2((vx[sub], vy[sub]), (-qi / (2 * math.pi), -qr / (2 * math.pi)))
    assert err <= tol, (c, err)
