"""Mesh ingestion: XDMF phantoms (geometry, topology, node and cell data,
time series) read without meshio or h5py, neither of which is installed here.

The reference reads its phantoms with ``meshio.xdmf.TimeSeriesReader`` and
falls back to ``meshio.read`` (python/feelmri/Phantom.py:76-139); the
readers here give the same objects for the files meshio, dolfin and dolfinx
write:

* ``TimeSeriesReader(path)``: ``read_points_cells()`` -> (points, cells),
  ``num_steps``, ``read_data(k)`` -> (t, point_data, cell_data), over a
  temporal grid collection (``xi:include`` of the mesh grid's geometry and
  topology resolved);
* ``read(path)`` -> ``Mesh`` (points, cells, point_data, cell_data) of the
  first uniform grid;
* ``load_phantom(path, scale_factor, dtype)`` -> ``Phantom``: the
  reference's ``_prepare_reader`` (time-series reader first, single mesh
  otherwise; nodes scaled; first cell block).

Heavy data: ``Format="XML"`` (inline text), ``"Binary"`` (raw file,
``Endian``, ``Seek``) and ``"HDF"`` (``file.h5:/path``, read by
``feelmri_amd.hdf5_lite``).  Cells are one block of one XDMF topology type.
"""
from __future__ import annotations

import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import hdf5_lite

# XDMF topology names -> meshio cell type names (and nodes per cell)
TOPOLOGY = {
    "Polyvertex": ("vertex", 1), "Polyline": ("line", 2), "Triangle": ("triangle", 3),
    "Quadrilateral": ("quad", 4), "Tetrahedron": ("tetra", 4), "Pyramid": ("pyramid", 5),
    "Wedge": ("wedge", 6), "Hexahedron": ("hexahedron", 8), "Edge_3": ("line3", 3),
    "Triangle_6": ("triangle6", 6), "Quadrilateral_8": ("quad8", 8), "Tetrahedron_10": ("tetra10", 10),
    "Hexahedron_20": ("hexahedron20", 20),
}
_NUMBER = {("Float", 4): "f4", ("Float", 8): "f8", ("Int", 4): "i4", ("Int", 8): "i8", ("Int", 2): "i2",
           ("Int", 1): "i1", ("UInt", 4): "u4", ("UInt", 8): "u8", ("UInt", 2): "u2", ("UInt", 1): "u1",
           ("Char", 1): "i1", ("UChar", 1): "u1"}
_XI = "{http://www.w3.org/2001/XInclude}include"


@dataclass
class CellBlock:
    type: str
    data: np.ndarray


@dataclass
class Mesh:
    points: np.ndarray
    cells: list
    point_data: dict = field(default_factory=dict)
    cell_data: dict = field(default_factory=dict)


def _strip(tag: str) -> str:
    return tag.rsplit("}", 1)[-1]


class _Xdmf:
    def __init__(self, path):
        self.path = Path(path)
        self.root = ET.parse(self.path).getroot()
        self._h5 = {}
        self.domain = next(e for e in self.root if _strip(e.tag) == "Domain")

    # heavy data
    def data_item(self, item) -> np.ndarray:
        if _strip(item.tag) != "DataItem":
            item = next(e for e in item if _strip(e.tag) == "DataItem")
        if item.get("ItemType") not in (None, "Uniform"):
            raise NotImplementedError(f"DataItem ItemType {item.get('ItemType')}")
        dims = tuple(int(x) for x in item.get("Dimensions").split())
        kind = item.get("DataType") or item.get("NumberType") or "Float"
        prec = int(item.get("Precision") or 4)
        fmt = item.get("Format") or "XML"
        text = (item.text or "").strip()
        if fmt == "XML":
            a = np.array(text.split(), dtype=np.dtype(_NUMBER[(kind, prec)]))
        elif fmt == "HDF":
            fname, _, inner = text.partition(":")
            f = self.path.parent / fname
            key = str(f)
            if key not in self._h5:
                self._h5[key] = hdf5_lite.File(f)
            a = self._h5[key][inner]
        elif fmt == "Binary":
            endian = {"Big": ">", "Little": "<", "Native": "="}[item.get("Endian") or "Native"]
            dt = np.dtype(endian + _NUMBER[(kind, prec)])
            a = np.fromfile(self.path.parent / text, dtype=dt, count=int(np.prod(dims)),
                            offset=int(item.get("Seek") or 0))
            a = a.astype(dt.newbyteorder("="))
        else:
            raise NotImplementedError(f"DataItem Format {fmt}")
        return np.asarray(a).reshape(dims)

    def children(self, node, tag):
        return [e for e in node if _strip(e.tag) == tag]

    # xi:include of another grid's geometry / topology
    def resolve_includes(self, grid) -> list:
        out = []
        for e in grid:
            if e.tag != _XI:
                out.append(e)
                continue
            xp = e.get("xpointer", "")
            m = re.search(r'//Grid\[@Name="([^"]+)"\](?:/Grid\[(\d+)\])?', xp)
            if not m:
                raise NotImplementedError(f"xi:include xpointer {xp}")
            src = next(g for g in self.domain.iter() if _strip(g.tag) == "Grid" and g.get("Name") == m.group(1))
            if m.group(2):
                src = self.children(src, "Grid")[int(m.group(2)) - 1]
            out += [c for c in src if _strip(c.tag) in ("Topology", "Geometry")]
        return out

    def mesh_of(self, elems) -> tuple:
        geo = next(e for e in elems if _strip(e.tag) == "Geometry")
        top = next(e for e in elems if _strip(e.tag) == "Topology")
        gtype = geo.get("GeometryType") or geo.get("Type") or "XYZ"
        if gtype == "XYZ":
            pts = self.data_item(geo)
        elif gtype == "XY":
            xy = self.data_item(geo)
            pts = np.column_stack([xy, np.zeros(len(xy), dtype=xy.dtype)])
        else:
            raise NotImplementedError(f"GeometryType {gtype}")
        ttype = top.get("TopologyType") or top.get("Type")
        if ttype not in TOPOLOGY:
            raise NotImplementedError(f"TopologyType {ttype}")
        name, npc = TOPOLOGY[ttype]
        cells = self.data_item(top).reshape(-1, npc)
        return pts, [CellBlock(name, cells)]

    def attributes(self, elems) -> tuple:
        pd, cd = {}, {}
        for a in elems:
            if _strip(a.tag) != "Attribute":
                continue
            arr = self.data_item(a)
            if (a.get("Center") or "Node") == "Node":
                pd[a.get("Name")] = arr
            else:
                cd[a.get("Name")] = [arr]
        return pd, cd


def read(path) -> Mesh:
    """The first uniform grid of an XDMF file (``meshio.read`` for XDMF)."""
    x = _Xdmf(path)
    for g in x.domain.iter():
        if _strip(g.tag) != "Grid" or (g.get("GridType") or "Uniform") != "Uniform":
            continue
        elems = x.resolve_includes(g)
        if any(_strip(e.tag) == "Geometry" for e in elems):
            pts, cells = x.mesh_of(elems)
            pd, cd = x.attributes(elems)
            return Mesh(pts, cells, pd, cd)
    raise ValueError(f"{path}: no uniform grid with a geometry")


class TimeSeriesReader:
    """XDMF time series (``meshio.xdmf.TimeSeriesReader``): the mesh of the
    first grid that has one, and per step of the temporal collection its
    time and node / cell data."""

    def __init__(self, path):
        self._x = x = _Xdmf(path)
        coll = [g for g in x.domain.iter() if _strip(g.tag) == "Grid" and g.get("GridType") == "Collection"
                and (g.get("CollectionType") or "") == "Temporal"]
        if not coll:
            raise ValueError(f"{path}: no temporal grid collection")
        self._steps = x.children(coll[0], "Grid")
        self.num_steps = len(self._steps)

    def read_points_cells(self):
        x = self._x
        for g in [*x.domain.iter()]:
            if _strip(g.tag) != "Grid" or (g.get("GridType") or "Uniform") != "Uniform":
                continue
            elems = x.resolve_includes(g)
            if any(_strip(e.tag) == "Geometry" for e in elems):
                return x.mesh_of(elems)
        raise ValueError("no grid with a geometry")

    def read_data(self, k: int):
        g = self._steps[k]
        time = next((float(e.get("Value")) for e in g if _strip(e.tag) == "Time"), float(k))
        pd, cd = self._x.attributes(self._x.resolve_includes(g))
        return time, pd, cd


@dataclass
class Phantom:
    nodes: np.ndarray
    elements: np.ndarray
    cell_type: str
    n_frames: int
    reader: object = None
    point_data: dict = field(default_factory=dict)
    cell_data: dict = field(default_factory=dict)


def load_phantom(path, scale_factor: float = 1.0, dtype=np.float64) -> Phantom:
    """The reference's ``FEMPhantom._prepare_reader`` (Phantom.py:76-139):
    a time-series reader when the file holds one (its step count as the
    number of frames), the single mesh and its data otherwise; nodes scaled
    by ``scale_factor`` and cast to ``dtype``; the first cell block."""
    try:
        reader = TimeSeriesReader(path)
        nodes, cells = reader.read_points_cells()
        n_frames, pd, cd = reader.num_steps, {}, {}
    except ValueError:
        m = read(path)
        nodes, cells, pd, cd = m.points, m.cells, m.point_data, m.cell_data
        reader, n_frames = None, 1
    nodes = (np.asarray(nodes, dtype=np.float64) * scale_factor).astype(dtype)
    return Phantom(nodes, np.ascontiguousarray(cells[0].data), cells[0].type, n_frames, reader, pd, cd)
