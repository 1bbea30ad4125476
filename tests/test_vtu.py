"""VTU point output: the written file decodes back to the same points,
cells and point data (round trip through an independent XML/base64 reader),
and the .pvd collection lists every step."""
import base64
import struct
import xml.etree.ElementTree as ET

import numpy as np

from phyflow_amd.vtu import VTUSeries, points_to_vtu


def _read(fname):
    root = ET.parse(fname).getroot()
    assert root.get("type") == "UnstructuredGrid" and root.get("byte_order") == "LittleEndian"
    piece = root.find("UnstructuredGrid/Piece")
    out = {}
    types = {"Float64": "<f8", "Int64": "<i8", "UInt8": "u1"}
    for da in piece.iter("DataArray"):
        raw = base64.b64decode(da.text)
        nbytes = struct.unpack("<I", raw[:4])[0]
        assert nbytes == len(raw) - 4
        a = np.frombuffer(raw[4:], dtype=types[da.get("type")])
        out[da.get("Name")] = a.reshape(-1, int(da.get("NumberOfComponents"))).squeeze()
    return int(piece.get("NumberOfPoints")), out


def test_points_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    n = 1000
    x, y, g = rng.standard_normal((3, n))
    vx, vy = rng.standard_normal((2, n))
    f = points_to_vtu(str(tmp_path / "b"), x, y, np.zeros(n), scalars={"g": g},
                      vectors={"v": (vx, vy, np.zeros(n))})
    m, d = _read(f)
    assert m == n
    assert np.array_equal(d["points"][:, 0], x) and np.array_equal(d["points"][:, 1], y)
    assert np.array_equal(d["points"][:, 2], np.zeros(n))
    assert np.array_equal(d["g"], g)
    assert np.array_equal(d["v"][:, 0], vx) and np.array_equal(d["v"][:, 1], vy)
    assert np.array_equal(d["connectivity"], np.arange(n))
    assert np.array_equal(d["offsets"], np.arange(1, n + 1))
    assert np.all(d["types"] == 1)


def test_empty_and_series(tmp_path):
    f = points_to_vtu(str(tmp_path / "e"), np.empty(0), np.empty(0), scalars={"g": np.empty(0)})
    assert _read(f)[0] == 0

    class FakeSystem:  # the BlobSystem surface the series uses
        def __init__(self, k):
            self.k = k

        def get(self):
            return np.arange(3.0) + self.k, np.zeros(3), np.ones(3)

    s = VTUSeries(str(tmp_path / "out" / "wake"))
    for k in range(3):
        s.write(FakeSystem(k), t=0.1 * k)
    pvd = ET.parse(str(tmp_path / "out" / "wake.pvd")).getroot()
    files = [e.get("file") for e in pvd.iter("DataSet")]
    assert files == ["wake%09d.vtu" % k for k in range(3)]
    assert _read(str(tmp_path / "out" / "wake000000002.vtu"))[1]["points"][0, 0] == 2.0
