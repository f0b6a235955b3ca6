"""Gmsh .msh reader (formats 2.2 ASCII/binary and 4.1 ASCII).

Mirrors pyfp3d/mesh/reader.py:86-214 (read_mesh through meshio): nodes in
file order as float64 (n, 3); tetrahedra (Gmsh type 4) as int32 (n, 4) in
file order; boundary triangles (type 2) grouped by physical-surface name
(or "all_triangles" when untagged); element physical tags kept per tet.
Node ids of the file are mapped to 0-based positions in file order, as
meshio does.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass, field
from pathlib import Path
from typing import Dict

import numpy as np

_NODES_PER_TYPE = {1: 2, 2: 3, 3: 4, 4: 4, 5: 8, 6: 6, 7: 5, 8: 3, 9: 6,
                   10: 9, 11: 10, 15: 1}


@dataclass
class Mesh:
    nodes: np.ndarray
    elements: np.ndarray
    boundary_faces: Dict[str, np.ndarray] = field(default_factory=dict)
    element_tags: np.ndarray = None
    physical_names: Dict[int, str] = field(default_factory=dict)

    @property
    def n_nodes(self):
        return len(self.nodes)

    @property
    def n_tets(self):
        return len(self.elements)


def _tokens_until(lines, i, end):
    out = []
    while lines[i].strip() != end:
        out.append(lines[i])
        i += 1
    return out, i + 1


def _physical_names(text_lines):
    names = {}
    n = int(text_lines[0])
    for ln in text_lines[1:1 + n]:
        parts = ln.split(maxsplit=2)
        names[(int(parts[0]), int(parts[1]))] = parts[2].strip().strip('"')
    return names


def _finish(nodes, tets, tet_tags, tris, tri_tags, names, id_to_idx) -> Mesh:
    nodes = np.ascontiguousarray(nodes, dtype=np.float64)
    if not tets:
        raise ValueError("Mesh must contain tetrahedral elements (cell type 'tetra')")
    el = id_to_idx[np.asarray(tets, dtype=np.int64)].astype(np.int32)
    if el.min() < 0:
        raise ValueError("element references an unknown node id")
    bf = {}
    if tris:
        tri = id_to_idx[np.asarray(tris, dtype=np.int64)].astype(np.int32)
        tags = np.asarray(tri_tags, dtype=np.int64)
        surf = {tag: nm for (dim, tag), nm in names.items() if dim == 2}
        if surf and np.all(tags > 0):
            for t in np.unique(tags):
                bf[surf.get(int(t), f"tag{int(t)}")] = tri[tags == t]
        else:
            bf["all_triangles"] = tri
    phys = {tag: nm for (dim, tag), nm in names.items()}
    return Mesh(nodes=nodes, elements=np.ascontiguousarray(el), boundary_faces=bf,
                element_tags=np.asarray(tet_tags, dtype=np.int32),
                physical_names=phys)


def _read_v2(raw: bytes, binary: bool) -> Mesh:
    names = {}
    pos = 0

    def next_line():
        nonlocal pos
        j = raw.index(b"\n", pos)
        ln = raw[pos:j].decode("latin-1")
        pos = j + 1
        return ln

    nodes = []
    ids = []
    tets, tet_tags, tris, tri_tags = [], [], [], []
    while pos < len(raw):
        ln = next_line().strip()
        if ln == "$MeshFormat":
            next_line()
            if binary:
                pos += 4           # the int 1 endianness check
                next_line()
            while next_line().strip() != "$EndMeshFormat":
                pass
        elif ln == "$PhysicalNames":
            n = int(next_line())
            block = [str(n)] + [next_line() for _ in range(n)]
            names = _physical_names(block)
            next_line()
        elif ln == "$Nodes":
            n = int(next_line())
            if binary:
                rec = np.dtype([("id", "<i4"), ("xyz", "<f8", 3)])
                arr = np.frombuffer(raw, dtype=rec, count=n, offset=pos)
                pos += n * rec.itemsize
                ids = arr["id"].astype(np.int64)
                nodes = arr["xyz"].copy()
                next_line()
            else:
                vals = [next_line().split() for _ in range(n)]
                ids = np.array([int(v[0]) for v in vals], dtype=np.int64)
                nodes = np.array([[float(x) for x in v[1:4]] for v in vals])
            next_line()
        elif ln == "$Elements":
            n = int(next_line())
            if binary:
                done = 0
                while done < n:
                    etype, nfollow, ntags = struct.unpack_from("<3i", raw, pos)
                    pos += 12
                    nn = _NODES_PER_TYPE[etype]
                    width = 1 + ntags + nn
                    blk = np.frombuffer(raw, dtype="<i4", count=nfollow * width,
                                        offset=pos).reshape(nfollow, width)
                    pos += 4 * nfollow * width
                    phys = blk[:, 1] if ntags > 0 else np.zeros(nfollow, np.int32)
                    conn = blk[:, 1 + ntags:]
                    if etype == 4:
                        tets.extend(conn.tolist())
                        tet_tags.extend(phys.tolist())
                    elif etype == 2:
                        tris.extend(conn.tolist())
                        tri_tags.extend(phys.tolist())
                    done += nfollow
                next_line()
            else:
                for _ in range(n):
                    v = [int(x) for x in next_line().split()]
                    etype, ntags = v[1], v[2]
                    tags = v[3:3 + ntags]
                    conn = v[3 + ntags:]
                    if etype == 4:
                        tets.append(conn)
                        tet_tags.append(tags[0] if tags else 0)
                    elif etype == 2:
                        tris.append(conn)
                        tri_tags.append(tags[0] if tags else 0)
            next_line()
    id_to_idx = np.full(int(ids.max()) + 1, -1, dtype=np.int64)
    id_to_idx[ids] = np.arange(len(ids))
    return _finish(nodes, tets, tet_tags, tris, tri_tags, names, id_to_idx)


def _read_v41_ascii(text: str) -> Mesh:
    lines = text.splitlines()
    i = 0
    names = {}
    ent_phys = {}
    node_ids, node_xyz = [], []
    tets, tet_tags, tris, tri_tags = [], [], [], []
    while i < len(lines):
        ln = lines[i].strip()
        i += 1
        if ln == "$PhysicalNames":
            blk, i = _tokens_until(lines, i, "$EndPhysicalNames")
            names = _physical_names(blk)
        elif ln == "$Entities":
            blk, i = _tokens_until(lines, i, "$EndEntities")
            counts = [int(x) for x in blk[0].split()]
            r = 1
            for dim in range(4):
                for _ in range(counts[dim]):
                    v = blk[r].split()
                    r += 1
                    tag = int(v[0])
                    if dim == 0:
                        nph = int(v[4])
                        ph = [int(x) for x in v[5:5 + nph]]
                    else:
                        nph = int(v[7])
                        ph = [int(x) for x in v[8:8 + nph]]
                    ent_phys[(dim, tag)] = ph[0] if ph else 0
        elif ln == "$Nodes":
            hdr = [int(x) for x in lines[i].split()]
            i += 1
            for _ in range(hdr[0]):
                d, t, par, nb = (int(x) for x in lines[i].split())
                i += 1
                node_ids.extend(int(lines[i + k]) for k in range(nb))
                i += nb
                for k in range(nb):
                    node_xyz.append([float(x) for x in lines[i + k].split()[:3]])
                i += nb
            i += 1  # $EndNodes
        elif ln == "$Elements":
            hdr = [int(x) for x in lines[i].split()]
            i += 1
            for _ in range(hdr[0]):
                d, t, et, nb = (int(x) for x in lines[i].split())
                i += 1
                ph = ent_phys.get((d, t), 0)
                for k in range(nb):
                    v = [int(x) for x in lines[i + k].split()]
                    if et == 4:
                        tets.append(v[1:5])
                        tet_tags.append(ph)
                    elif et == 2:
                        tris.append(v[1:4])
                        tri_tags.append(ph)
                i += nb
            i += 1  # $EndElements
    ids = np.asarray(node_ids, dtype=np.int64)
    id_to_idx = np.full(int(ids.max()) + 1, -1, dtype=np.int64)
    id_to_idx[ids] = np.arange(len(ids))
    return _finish(np.asarray(node_xyz), tets, tet_tags, tris, tri_tags, names,
                   id_to_idx)


def read_mesh(filepath, verbose: bool = False) -> Mesh:
    """Read a Gmsh .msh file (2.2 ASCII/binary, 4.1 ASCII) into a Mesh."""
    filepath = Path(filepath)
    if not filepath.exists():
        raise FileNotFoundError(f"Mesh file not found: {filepath}")
    raw = filepath.read_bytes()
    head = raw[:64].split(b"\n")
    if head[0].strip() != b"$MeshFormat":
        raise ValueError("not a Gmsh .msh file")
    ver, ftype, _ = head[1].split()[:3]
    binary = ftype == b"1"
    if ver.startswith(b"2"):
        m = _read_v2(raw, binary)
    elif ver.startswith(b"4"):
        if binary:
            raise NotImplementedError("binary Gmsh 4.x is not supported")
        m = _read_v41_ascii(raw.decode("latin-1"))
    else:
        raise ValueError(f"unsupported Gmsh version {ver!r}")
    if verbose:
        print(f"Read mesh from {filepath}: {m.n_nodes} nodes, {m.n_tets} tets, "
              f"boundaries {list(m.boundary_faces)}")
    return m
