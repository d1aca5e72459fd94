"""Meshes: PolygonalMesh and its generators, mirroring src/mesh.jl,
src/generate_mesh.jl and src/triangle_mesh.jl.

Arrays use the reference's in-memory layouts so they pass to libhdg (and to
a Julia caller) unchanged: `nodes` (nn, 2) float64, `cells` (nc, 6) int64
with 3 node ids then 3 face ids, `faces` (nf, 4) int64 [v1 v2 cell1
cell2/0]; all ids 1-based.
"""
from __future__ import annotations

import re

import numpy as np

from ._lib import check, lib, ptr, _i64, _i8

SIDES = {1: "bottom", 2: "right", 3: "top", 4: "left"}


class PolygonalMesh:
    """PolygonalMesh{2,3,3} (src/mesh.jl:43-49)."""

    def __init__(self, nodes, cells, faces, facesets, nodesets):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64)
        self.cells = np.ascontiguousarray(cells, dtype=np.int64)
        self.faces = np.ascontiguousarray(faces, dtype=np.int64)
        self.facesets = facesets
        self.nodesets = nodesets

    @property
    def faces_colmajor(self):
        """faces as Julia stores Matrix{Int} (column-major nf x 4)."""
        return np.asfortranarray(self.faces)

    def boundary_dofs(self, nt, faceset="boundary"):
        """1-based trace dofs of the faces in `faceset`, ascending."""
        fs = np.array(sorted(self.facesets[faceset]), dtype=np.int64)
        return ((fs[:, None] - 1) * nt + np.arange(1, nt + 1)[None, :]).ravel()


# accessors of src/mesh.jl:56-77,117-124
def getncells(m): return m.cells.shape[0]
def getnfaces(m): return m.faces.shape[0]
def getnnodes(m): return m.nodes.shape[0]
def n_faces_per_cell(m): return 3
def getfaceset(m, name): return m.facesets[name]
def getnodeset(m, name): return m.nodesets[name]
def get_vertices_matrix(m): return m.nodes.copy()
def getcells_matrix(m): return m.cells[:, :3].copy()


def face_orientation(m, cell, face):
    """face_orientation(mesh, cell, face), both 1-based (src/mesh.jl:51-54)."""
    k1, k2 = ((2, 3), (3, 1), (1, 2))[face - 1]
    c = m.cells[cell - 1]
    return bool(c[k2 - 1] > c[k1 - 1])


def cell_volume(m, cell):
    x = m.nodes[m.cells[cell - 1, :3] - 1]
    a, b = x[1] - x[0], x[2] - x[0]
    return 0.5 * abs(a[0] * b[1] - a[1] * b[0])


def rectangle_mesh(nel, LL=(0.0, 0.0), UR=(1.0, 1.0), jitter=0.0, seed=0):
    """rectangle_mesh(TriangleCell, nel, LL, UR) (src/generate_mesh.jl:96-143)
    through libhdg's closed-form generator.  `jitter` > 0 moves interior
    nodes by up to jitter*h per axis after generation (seeded splitmix64)."""
    nx, ny = int(nel[0]), int(nel[1])
    L = lib()
    nf = L.hdg_rectangle_nfaces(nx, ny)
    nodes = np.zeros(((nx + 1) * (ny + 1), 2))
    cells = np.zeros((2 * nx * ny, 6), dtype=np.int64)
    faces_cm = np.zeros((4, nf), dtype=np.int64)
    side = np.zeros(nf, dtype=np.int8)
    check(L.hdg_rectangle_mesh(nx, ny, float(LL[0]), float(LL[1]), float(UR[0]), float(UR[1]),
                               float(jitter), int(seed), ptr(nodes), ptr(cells, _i64),
                               ptr(faces_cm, _i64), ptr(side, _i8)))
    faces = faces_cm.T.copy()
    facesets, nodesets = {}, {}
    for code, name in SIDES.items():
        idx = np.nonzero(side == code)[0]
        facesets[name] = set((idx + 1).tolist())
        nodesets[name] = set(np.unique(faces[idx, :2]).tolist())
    facesets["boundary"] = set().union(*(facesets[n] for n in SIDES.values()))
    nodesets["boundary"] = set().union(*(nodesets[n] for n in SIDES.values()))
    m = PolygonalMesh(nodes, cells, faces, facesets, nodesets)
    m.face_side = side
    return m


_NUM = re.compile(r"\b((\d*\.)?\d+)\b")


def _rows(path):
    out = []
    with open(path) as fh:
        for ln in fh:
            if re.match(r"^\s*(?:#|$)", ln):
                continue
            out.append([mm.group(0) for mm in _NUM.finditer(ln)])
    return out[1:]  # the first line is the header


def parse_mesh_triangle(root):
    """parse_mesh_triangle(root) for Triangle's .node/.edge/.ele files
    (src/triangle_mesh.jl:1-126): cells in file order with the vertex
    order made counter-clockwise, faces numbered by first encounter over
    local edges (2,3),(3,1),(1,2), and the "boundary" faceset from the
    edge markers."""
    nodes = np.array([[float(r[1]), float(r[2])] for r in _rows(root + ".node")])
    marker = {}
    for r in _rows(root + ".edge"):
        a, b = int(r[1]), int(r[2])
        marker.setdefault((min(a, b), max(a, b)), int(r[3]))
    cells, faces, fid = [], [], {}
    bfaces, bnodes = set(), set()
    for n_el, r in enumerate(_rows(root + ".ele"), start=1):
        v = [int(r[1]), int(r[2]), int(r[3])]
        a = nodes[v[1] - 1] - nodes[v[0] - 1]
        b = nodes[v[2] - 1] - nodes[v[0] - 1]
        if a[0] * b[1] - a[1] * b[0] < 0:
            v[1], v[2] = v[2], v[1]
        ef = []
        for i, j in ((1, 2), (2, 0), (0, 1)):
            key = (min(v[i], v[j]), max(v[i], v[j]))
            f = fid.get(key)
            if f is None:
                faces.append([v[i], v[j], n_el, 0])
                f = fid[key] = len(faces)
                if marker.get(key, -1) > 0:
                    bfaces.add(f)
                    bnodes.update((v[i], v[j]))
            else:
                faces[f - 1][3] = n_el
            ef.append(f)
        cells.append(v + ef)
    return PolygonalMesh(nodes, np.array(cells, dtype=np.int64), np.array(faces, dtype=np.int64),
                         {"boundary": bfaces}, {"boundary": bnodes})
