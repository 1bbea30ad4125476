"""Autotuner table lookup (CPU): the entry of the same precision whose
tolerance class meets the request, nearest in system size; the reference's
rules when the table has none."""
import math

from phyflow_amd import autotune
from phyflow_amd._native import PREC_F32C, PREC_F64


def _table():
    def e(prec, tol, n, order, leaf, ws, side):
        return {"prec": prec, "tol": tol, "n": n,
                "choice": {"order": order, "leaf": leaf, "ws": ws, "leaf_side": side}}
    return {"entries": [e("f32c", 1e-6, 65536, 12, 64, 1, 1.0),
                        e("f32c", 1e-6, 1048576, 14, 64, 1, 1.0),
                        e("f64", 1e-6, 1048576, 15, 32, 1, -1.0),
                        e("f64", 1e-12, 1048576, 30, 64, 2, 1.0),
                        {"prec": "f64", "tol": 1e-9, "n": 1048576, "choice": None}]}


def test_lookup_nearest_size_and_tolerance_class():
    t = _table()
    p = autotune.tuned_params(0.001, 900000, PREC_F32C, 1e-6, table=t)
    assert (p.order, p.leaf, p.ws, p.leaf_side, p.prec) == (14, 64, 1, 1.0, PREC_F32C)
    p = autotune.tuned_params(0.001, 70000, PREC_F32C, 1e-6, table=t)
    assert p.order == 12
    # a tighter request takes the entry that meets it, never a looser one
    p = autotune.tuned_params(0.001, 1048576, PREC_F64, 1e-10, table=t)
    assert (p.order, p.ws) == (30, 2)
    p = autotune.tuned_params(0.001, 1048576, PREC_F64, 1e-6, table=t)
    assert (p.order, p.leaf, p.leaf_side) == (15, 32, -1.0)


def test_lookup_falls_back_to_reference_rules():
    t = _table()
    # no f32c entry reaches 1e-9: order from tol (fmm.cpp:538-546), depth from Ndirect
    p = autotune.tuned_params(0.002, 50000, PREC_F32C, 1e-9, table=t)
    assert (p.order, p.leaf, p.ws, p.leaf_side) == (0, 0, 0, 0.0) and math.isclose(p.tol, 1e-9)
    p = autotune.tuned_params(0.002, 50000, PREC_F64, 1e-6, table={"entries": []})
    assert p.order == 0 and p.leaf == 0


def test_committed_table_is_well_formed():
    t = autotune.load_table()
    assert t is not None, "the measured table is committed with the package"
    assert t["entries"]
    for e in t["entries"]:
        assert e["prec"] in ("f32c", "f64") and e["tol"] > 0 and e["n"] > 0
        c = e["choice"]
        if c is not None:
            assert 2 <= c["order"] <= 40 and c["ws"] in (1, 2) and c["leaf"] > 0
            assert c["rel_l2"] <= 0.5 * e["tol"]
