"""ParaView output of the mesh and its fields (pyfp3d/post/vtk_out.py:19-138,
write_vtu / read_vtu), written without meshio (absent here): a VTK XML
UnstructuredGrid with every array base64-encoded binary (UInt64 size
header), so float64 fields round-trip exactly.

read_vtu reads files written by write_vtu and returns the reference's
(nodes, elements, fields) with cell fields under "cell_<name>"."""

from __future__ import annotations

import base64
import re
from pathlib import Path
from typing import Dict, Optional, Tuple

import numpy as np

_VTK_TETRA = 10
_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.int64): "Int64",
          np.dtype(np.int32): "Int32", np.dtype(np.uint8): "UInt8"}
_DTYPES = {v: k for k, v in _TYPES.items()}


def _encode(a: np.ndarray) -> str:
    raw = np.ascontiguousarray(a).tobytes()
    head = np.array([len(raw)], dtype="<u8").tobytes()
    return base64.b64encode(head + raw).decode("ascii")


def _array(name: str, a: np.ndarray) -> str:
    a = np.asarray(a)
    if a.dtype not in _TYPES:
        a = a.astype(np.float64)
    ncomp = 1 if a.ndim == 1 else int(a.shape[1])
    nm = f' Name="{name}"' if name else ""
    return (f'<DataArray type="{_TYPES[a.dtype]}"{nm} NumberOfComponents="{ncomp}" '
            f'format="binary">{_encode(a)}</DataArray>')


def write_vtu(filepath, nodes: np.ndarray, elements: np.ndarray,
              point_data: Optional[Dict[str, np.ndarray]] = None,
              cell_data: Optional[Dict[str, np.ndarray]] = None,
              verbose: bool = False) -> None:
    """Write tets and nodal / element fields ((n,) or (n, 3)) to `filepath`."""
    filepath = Path(filepath)
    filepath.parent.mkdir(parents=True, exist_ok=True)
    nodes = np.asarray(nodes, dtype=np.float64)
    el = np.asarray(elements, dtype=np.int64)
    if nodes.ndim != 2 or nodes.shape[1] != 3:
        raise ValueError("nodes must be (n, 3)")
    if el.ndim != 2 or el.shape[1] != 4:
        raise ValueError("elements must be (m, 4) tetrahedra")
    parts = []
    for tag, data, n in (("PointData", point_data, len(nodes)),
                         ("CellData", cell_data, len(el))):
        arrs = []
        for name, a in (data or {}).items():
            a = np.asarray(a, dtype=np.float64)
            if a.shape[0] != n:
                raise ValueError(f"{tag} field {name!r}: {a.shape[0]} rows, expected {n}")
            arrs.append(_array(name, a))
        parts.append(f"<{tag}>{''.join(arrs)}</{tag}>")
    xml = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" '
        'header_type="UInt64">\n<UnstructuredGrid>\n'
        f'<Piece NumberOfPoints="{len(nodes)}" NumberOfCells="{len(el)}">\n'
        + parts[0] + "\n" + parts[1] + "\n"
        + f"<Points>{_array('', nodes)}</Points>\n"
        + "<Cells>"
        + _array("connectivity", el.reshape(-1))
        + _array("offsets", 4 * np.arange(1, len(el) + 1, dtype=np.int64))
        + _array("types", np.full(len(el), _VTK_TETRA, dtype=np.uint8))
        + "</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")
    filepath.write_text(xml)
    if verbose:
        print(f"Wrote VTU to {filepath}")


_DA = re.compile(r'<DataArray type="(\w+)"(?: Name="([^"]*)")? '
                 r'NumberOfComponents="(\d+)" format="binary">([^<]*)</DataArray>')


def _decode(typ: str, ncomp: int, text: str) -> np.ndarray:
    blob = base64.b64decode(text)
    size = int(np.frombuffer(blob[:8], dtype="<u8")[0])
    a = np.frombuffer(blob[8:8 + size], dtype=_DTYPES[typ]).copy()
    return a.reshape(-1, ncomp) if ncomp > 1 else a


def read_vtu(filepath) -> Tuple[np.ndarray, np.ndarray, Dict[str, np.ndarray]]:
    """(nodes (N, 3), elements (M, 4) int32, fields) of a write_vtu file;
    cell fields are returned as "cell_<name>"."""
    text = Path(filepath).read_text()

    def section(tag):
        m = re.search(f"<{tag}>(.*?)</{tag}>", text, re.S)
        return m.group(1) if m else ""

    fields = {}
    for typ, name, nc, data in _DA.findall(section("PointData")):
        fields[name] = _decode(typ, int(nc), data)
    for typ, name, nc, data in _DA.findall(section("CellData")):
        fields[f"cell_{name}"] = _decode(typ, int(nc), data)
    pts = _DA.findall(section("Points"))
    nodes = _decode(pts[0][0], int(pts[0][2]), pts[0][3])
    cells = {name: _decode(typ, int(nc), data)
             for typ, name, nc, data in _DA.findall(section("Cells"))}
    elements = cells["connectivity"].reshape(-1, 4).astype(np.int32)
    return nodes, elements, fields
