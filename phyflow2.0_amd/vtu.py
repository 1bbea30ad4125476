"""VTU point output for blobs, the blob branch of the reference's IO class
(src/python/IO/__IO.py:216-251: ``pointsToVTK(name, x, y, 0, scalars={"g":
g}, vectors={"v": (vx, vy, 0)})`` per time step, indexed by a .pvd
collection, :132-214).

    series = VTUSeries("out/wake")          # writes out/wake.pvd
    series.write(system, t=0.0)             # out/wake000000000.vtu (+ "v")
    points_to_vtu("blobs", x, y, scalars={"g": g})

The reference writes through pyEVTK (an external dependency, not part of
its tree); this module writes the same VTK XML UnstructuredGrid content
(one VTK_VERTEX cell per blob, float64 point data) with base64-encoded
binary data arrays that ParaView and VTK read directly.  Host-side only:
the blob data come from the device once per written step.
"""
import base64
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np

__all__ = ["points_to_vtu", "VTUSeries"]

_VTK_VERTEX = 1


def _b64(a):
    """VTK inline binary: a UInt32 byte count, then the raw little-endian
    bytes, base64-encoded together."""
    raw = np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<"), copy=False).tobytes()
    return base64.b64encode(struct.pack("<I", len(raw)) + raw).decode("ascii")


def _array(parent, name, a, ncomp=1):
    types = {np.dtype(np.float64): "Float64", np.dtype(np.int64): "Int64",
             np.dtype(np.uint8): "UInt8"}
    e = ET.SubElement(parent, "DataArray", type=types[a.dtype], Name=name,
                      NumberOfComponents=str(ncomp), format="binary")
    e.text = _b64(a)
    return e


def points_to_vtu(path, x, y, z=None, scalars=None, vectors=None):
    """Write blobs as an UnstructuredGrid of vertices to ``path + ".vtu"``
    (pyEVTK's pointsToVTK naming); returns the file name."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    n = x.size
    if y.size != n:
        raise ValueError("x and y have different lengths")
    z = np.zeros(n) if z is None else np.ascontiguousarray(z, dtype=np.float64)
    root = ET.Element("VTKFile", type="UnstructuredGrid", version="0.1",
                      byte_order="LittleEndian", header_type="UInt32")
    grid = ET.SubElement(root, "UnstructuredGrid")
    piece = ET.SubElement(grid, "Piece", NumberOfPoints=str(n), NumberOfCells=str(n))
    pts = ET.SubElement(piece, "Points")
    _array(pts, "points", np.stack([x, y, z], axis=1).ravel(), 3)
    cells = ET.SubElement(piece, "Cells")
    _array(cells, "connectivity", np.arange(n, dtype=np.int64))
    _array(cells, "offsets", np.arange(1, n + 1, dtype=np.int64))
    _array(cells, "types", np.full(n, _VTK_VERTEX, dtype=np.uint8))
    pd = ET.SubElement(piece, "PointData")
    for name, v in (scalars or {}).items():
        v = np.ascontiguousarray(v, dtype=np.float64)
        if v.size != n:
            raise ValueError("scalar %r has %d values for %d points" % (name, v.size, n))
        _array(pd, name, v)
    for name, (vx, vy, vz) in (vectors or {}).items():
        comps = [np.ascontiguousarray(c, dtype=np.float64) for c in (vx, vy, vz)]
        if any(c.size != n for c in comps):
            raise ValueError("vector %r does not have %d values per component" % (name, n))
        _array(pd, name, np.stack(comps, axis=1).ravel(), 3)
    fname = path + ".vtu"
    ET.ElementTree(root).write(fname, xml_declaration=True, encoding="utf-8")
    return fname


class VTUSeries:
    """Numbered .vtu files plus a .pvd collection, one entry per write
    (the reference's IO(filename) << blobs, __IO.py:106-251, 312)."""

    def __init__(self, filename):
        self.base = filename
        d = os.path.dirname(filename)
        if d:
            os.makedirs(d, exist_ok=True)
        self.step = 0
        self.entries = []

    def write(self, system, t=None, velocity_params=None):
        """Write the blobs of a BlobSystem; with ``velocity_params`` also the
        velocity at the blobs as vector "v" (the reference's plotVelocity)."""
        x, y, g = system.get()
        vectors = None
        if velocity_params is not None:
            vx, vy = system.velocity(velocity_params)
            vectors = {"v": (vx, vy, np.zeros(g.size))}
        name = points_to_vtu(self.base + ("%09d" % self.step), x, y, np.zeros(x.size),
                             scalars={"g": g}, vectors=vectors)
        self.entries.append((float(self.step if t is None else t), os.path.basename(name)))
        self.step += 1
        self._write_pvd()
        return name

    def _write_pvd(self):
        root = ET.Element("VTKFile", type="Collection", version="0.1", byte_order="LittleEndian")
        coll = ET.SubElement(root, "Collection")
        for t, f in self.entries:
            ET.SubElement(coll, "DataSet", timestep=repr(t), group="", part="0", file=f)
        ET.ElementTree(root).write(self.base + ".pvd", xml_declaration=True, encoding="utf-8")
