"""Python handle on the CPU oracle (liboracle.so built from hdg_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle restates the reference pipeline
(examples/poisson2D_HDG.jl:58-218 and the src/ files it calls) on the CPU.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product (libhdg and its Python mirror) never does.

Arrays use the reference's layouts: cells (ncells, 6) int64 1-based
(3 node ids then 3 face ids), faces (nfaces, 4) 1-based [v1 v2 cell1 cell2],
dense element matrices as numpy arrays indexed [row, col].
"""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", HERE, "liboracle.so"],
                       check=True, capture_output=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class Problem(C.Structure):
    _fields_ = [("k", C.c_int32), ("qdeg", C.c_int32), ("tau", C.c_double),
                ("src_kind", C.c_int32), ("kappa_on", C.c_int32),
                ("ka", C.c_double), ("kb", C.c_double),
                ("coef", C.c_double * 16), ("f_qp", C.c_void_p)]


def problem(k=1, qdeg=None, tau=1.0, src_kind=0, kappa=None, coef=None, f_qp=None):
    p = Problem()
    p.k = k
    p.qdeg = k + 1 if qdeg is None else qdeg
    p.tau = tau
    p.src_kind = src_kind
    if kappa is not None:
        p.kappa_on, p.ka, p.kb = 1, float(kappa[0]), float(kappa[1])
    if coef is not None:
        for i, c in enumerate(np.asarray(coef, dtype=np.float64).ravel()):
            p.coef[i] = c
    if f_qp is not None:
        p._keep = np.asfortranarray(f_qp, dtype=np.float64)
        p.f_qp = p._keep.ctypes.data
    return p


_P = C.POINTER
_d, _i64, _i8 = _P(C.c_double), _P(C.c_int64), _P(C.c_int8)


def _declare(L):
    L.orc_jacobi.restype = C.c_double
    L.orc_jacobi.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double]
    for n in ("orc_dubiner_value",):
        getattr(L, n).restype = C.c_double
        getattr(L, n).argtypes = [C.c_int, C.c_double, C.c_double]
    L.orc_dubiner_grad.argtypes = [C.c_int, C.c_double, C.c_double, _d]
    L.orc_dubiner_basis.restype = C.c_double
    L.orc_dubiner_basis.argtypes = [C.c_double, C.c_double, C.c_int]
    L.orc_dubiner_basis_grad.argtypes = [C.c_double, C.c_double, C.c_int, _d]
    L.orc_legendre_value.restype = C.c_double
    L.orc_legendre_value.argtypes = [C.c_int, C.c_double]
    L.orc_lagrange2d_p1.argtypes = [C.c_double, C.c_double, _d, _d]
    L.orc_lagrange1d_p1.argtypes = [C.c_double, _d]
    for n in ("orc_strang", "orc_gm", "orc_gauss_legendre", "orc_quad2d", "orc_quad1d"):
        getattr(L, n).argtypes = [C.c_int, _d, _d]
    L.orc_rect_nfaces.restype = C.c_int64
    L.orc_rect_nfaces.argtypes = [C.c_int64, C.c_int64]
    L.orc_rectangle_mesh.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_double,
                                     C.c_double, C.c_double, _d, _i64, _i64, _i8]
    L.orc_jitter.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double,
                             C.c_double, C.c_double, C.c_uint64, _d]
    L.orc_face_orientation.argtypes = [_i64, C.c_int64, C.c_int]
    L.orc_sizes.argtypes = [C.c_int, C.c_int] + [_P(C.c_int)] * 4
    L.orc_geometry.argtypes = [_d, _i64, C.c_int64, _d, _d, _d, _d]
    L.orc_element_blocks.argtypes = [_P(Problem), _d, _i64, C.c_int64, C.c_int64] + [_d] * 7
    L.orc_condense.argtypes = [_P(Problem), _d, _i64, C.c_int64, _d, _d, _d, _d, C.c_int]
    L.orc_assemble.argtypes = [C.c_int64, _i64, C.c_int64, C.c_int, _d, _d, _i64, _i64,
                               _d, _d, _i64]
    L.orc_dirichlet_trace.argtypes = [_P(Problem), _d, _i64, _i64, C.c_int64, _i8, C.c_int,
                                      _d, C.c_int, _i64, _d, _i64]
    L.orc_apply.argtypes = [C.c_int64, _i64, _i64, _d, _d, C.c_int64, _i64, _d]
    L.orc_banded_solve.argtypes = [C.c_int64, _i64, _i64, _d, _d, _d]
    L.orc_pcg_solve.argtypes = [C.c_int64, _i64, _i64, _d, _d, C.c_int, _i64, _d,
                                C.c_int64, C.c_double, C.c_int, _d, _P(C.c_int), _d, C.c_int]
    L.orc_recover.argtypes = [C.c_int64, _i64, C.c_int, _d, _d, _d, _d, _d, _d]
    L.orc_errornorm.restype = C.c_double
    L.orc_errornorm.argtypes = [_P(Problem), _d, _i64, C.c_int64, _d, C.c_int, _d]
    L.orc_residual.argtypes = [C.c_int64, _i64, _i64, _d, _d, _d, _d]


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _chk(code, what):
    if code != 0:
        raise RuntimeError(f"oracle {what} failed with status {code}")


# --------------------------------------------------------------------------
# reference-element numerics
# --------------------------------------------------------------------------
def dubiner_value(j, r, s):
    return lib().orc_dubiner_value(j, r, s)


def dubiner_grad(j, r, s):
    g = np.zeros(2)
    lib().orc_dubiner_grad(j, r, s, _ptr(g, _d))
    return g


def dubiner_basis(x, y, j):
    return lib().orc_dubiner_basis(x, y, j)


def dubiner_basis_grad(x, y, j):
    g = np.zeros(2)
    lib().orc_dubiner_basis_grad(x, y, j, _ptr(g, _d))
    return g


def legendre_value(k, x):
    return lib().orc_legendre_value(k, x)


def lagrange2d_p1(r, s):
    v, g = np.zeros(3), np.zeros(6)
    lib().orc_lagrange2d_p1(r, s, _ptr(v, _d), _ptr(g, _d))
    return v, g.reshape(3, 2)


def lagrange1d_p1(x):
    v = np.zeros(2)
    lib().orc_lagrange1d_p1(x, _ptr(v, _d))
    return v


def _rule(fn, arg, dim):
    p, w = np.zeros(2 * 4096), np.zeros(4096)
    n = fn(arg, _ptr(p, _d), _ptr(w, _d))
    if n < 0:
        raise ValueError(f"quadrature rule {arg} not available")
    return (p[: dim * n].reshape(n, dim) if dim == 2 else p[:n].copy()), w[:n].copy()


def strang(order):
    return _rule(lib().orc_strang, order, 2)


def grundmann_moeller(s):
    return _rule(lib().orc_gm, s, 2)


def gauss_legendre(n):
    return _rule(lib().orc_gauss_legendre, n, 1)


def quad2d(order):
    return _rule(lib().orc_quad2d, order, 2)


def quad1d(order):
    return _rule(lib().orc_quad1d, order, 1)


def sizes(k, qdeg):
    a = [C.c_int() for _ in range(4)]
    _chk(lib().orc_sizes(k, qdeg, *[C.byref(x) for x in a]), "sizes")
    return tuple(x.value for x in a)  # nu, nt, nqp, nfq


# --------------------------------------------------------------------------
# meshes
# --------------------------------------------------------------------------
class Mesh:
    """Reference PolygonalMesh arrays (src/mesh.jl:43-49)."""

    def __init__(self, xy, cells, faces, facesets=None, nodesets=None):
        self.xy = np.ascontiguousarray(xy, dtype=np.float64)
        self.cells = np.ascontiguousarray(cells, dtype=np.int64)
        self.faces = np.asarray(faces, dtype=np.int64)
        self.faces_cm = np.asfortranarray(self.faces)
        self.facesets = facesets or {}
        self.nodesets = nodesets or {}

    @property
    def ncells(self):
        return self.cells.shape[0]

    @property
    def nfaces(self):
        return self.faces.shape[0]

    @property
    def nnodes(self):
        return self.xy.shape[0]

    def boundary_mask(self, name="boundary"):
        m = np.zeros(self.nfaces, dtype=np.int8)
        for f in self.facesets[name]:
            m[f - 1] = 1
        return m


def rectangle_mesh(nx, ny, LL=(0.0, 0.0), UR=(1.0, 1.0), jitter=0.0, seed=0):
    """rectangle_mesh(TriangleCell, (nx,ny), LL, UR) (generate_mesh.jl:96-143),
    then optional interior jitter (amplitude jitter*h, splitmix64 seed)."""
    L = lib()
    nn = (nx + 1) * (ny + 1)
    nc = 2 * nx * ny
    nf = L.orc_rect_nfaces(nx, ny)
    xy = np.zeros((nn, 2))
    cells = np.zeros((nc, 6), dtype=np.int64)
    faces = np.zeros((4, nf), dtype=np.int64)  # column-major nf x 4
    side = np.zeros(nf, dtype=np.int8)
    _chk(L.orc_rectangle_mesh(nx, ny, LL[0], LL[1], UR[0], UR[1], _ptr(xy, _d), _ptr(cells, _i64),
                              _ptr(faces, _i64), _ptr(side, _i8)), "rectangle_mesh")
    if jitter:
        L.orc_jitter(nx, ny, LL[0], LL[1], UR[0], UR[1], float(jitter), int(seed), _ptr(xy, _d))
    faces = faces.T.copy()
    names = {1: "bottom", 2: "right", 3: "top", 4: "left"}
    fs = {n: set() for n in names.values()}
    ns = {n: set() for n in names.values()}
    for f in np.nonzero(side)[0]:
        nm = names[int(side[f])]
        fs[nm].add(int(f) + 1)
        ns[nm].update((int(faces[f, 0]), int(faces[f, 1])))
    fs["boundary"] = set().union(*[fs[n] for n in names.values()])
    ns["boundary"] = set().union(*[ns[n] for n in names.values()])
    m = Mesh(xy, cells, faces, fs, ns)
    m.face_side = side
    return m


_NUM = re.compile(r"\b((\d*\.)?\d+)\b")


def _read_rows(path):
    rows = []
    with open(path) as fh:
        for ln in fh:
            if re.match(r"^\s*(?:#|$)", ln):
                continue
            rows.append([m.group(0) for m in _NUM.finditer(ln)])
    return rows[0], rows[1:]


def parse_mesh_triangle(root):
    """parse_mesh_triangle (src/triangle_mesh.jl:1-126), restated."""
    _, node_rows = _read_rows(root + ".node")
    xy = np.array([[float(r[1]), float(r[2])] for r in node_rows])
    _, edge_rows = _read_rows(root + ".edge")
    faces_ref = {}
    for r in edge_rows:
        a, b = int(r[1]), int(r[2])
        key = (min(a, b), max(a, b))
        if key not in faces_ref:
            faces_ref[key] = int(r[3])
    _, ele_rows = _read_rows(root + ".ele")
    cells, faces, fdict = [], [], {}
    bfaces, bnodes = set(), set()
    for n_el, r in enumerate(ele_rows, start=1):
        n1, n2, n3 = int(r[1]), int(r[2]), int(r[3])
        a = xy[n2 - 1] - xy[n1 - 1]
        b = xy[n3 - 1] - xy[n1 - 1]
        if a[0] * b[1] - a[1] * b[0] < 0:
            n2, n3 = n3, n2
        el = (n1, n2, n3)
        ef = []
        for (i, j) in ((1, 2), (2, 0), (0, 1)):
            v1, v2 = el[i], el[j]
            key = (min(v1, v2), max(v1, v2))
            if key in fdict:
                fid = fdict[key]
                faces[fid - 1][3] = n_el
            else:
                faces.append([v1, v2, n_el, 0])
                fid = len(faces)
                fdict[key] = fid
                if faces_ref.get(key, -1) > 0:
                    bfaces.add(fid)
                    bnodes.update((v1, v2))
            ef.append(fid)
        cells.append(list(el) + ef)
    return Mesh(xy, np.array(cells, dtype=np.int64), np.array(faces, dtype=np.int64),
                {"boundary": bfaces}, {"boundary": bnodes})


def face_orientation(mesh, cell, face):
    """face_orientation(mesh, cell, face) with 1-based cell and face."""
    return bool(lib().orc_face_orientation(_ptr(mesh.cells, _i64), cell - 1, face - 1))


# --------------------------------------------------------------------------
# HDG pipeline
# --------------------------------------------------------------------------
def geometry(mesh, cell):
    detJ = np.zeros(1)
    Jinv, normals, detJf = np.zeros(4), np.zeros(6), np.zeros(3)
    _chk(lib().orc_geometry(_ptr(mesh.xy, _d), _ptr(mesh.cells, _i64), cell - 1, _ptr(detJ, _d),
                            _ptr(Jinv, _d), _ptr(normals, _d), _ptr(detJf, _d)), "geometry")
    return detJ[0], Jinv.reshape(2, 2, order="F"), normals.reshape(3, 2), detJf


def element_blocks(P, mesh, cell):
    """Local A, B, C, E, F, H, be of 1-based `cell` (poisson2D_HDG.jl:77-153)."""
    nu, nt = sizes(P.k, P.qdeg)[:2]
    nq, ntr = 2 * nu, 3 * nt
    out = dict(A=(nq, nq), B=(nq, nu), C=(nu, nu), E=(nq, ntr), F=(nu, ntr), H=(ntr, ntr), be=(nu,))
    arr = {k: np.zeros(int(np.prod(s))) for k, s in out.items()}
    _chk(lib().orc_element_blocks(C.byref(P), _ptr(mesh.xy, _d), _ptr(mesh.cells, _i64),
                                  mesh.ncells, cell - 1,
                                  *[_ptr(arr[k], _d) for k in ("A", "B", "C", "E", "F", "H", "be")]),
         "element_blocks")
    return {k: arr[k].reshape(out[k], order="F") for k in out}


def condense(P, mesh, nthreads=0):
    """(Ate, bte, Ke, be) for all cells; Ate[c] is ntr x ntr, Ke[c] m x ntr."""
    nu, nt = sizes(P.k, P.qdeg)[:2]
    m, ntr, nc = 3 * nu, 3 * nt, mesh.ncells
    Ate = np.zeros((nc, ntr, ntr))
    bte = np.zeros((nc, ntr))
    Ke = np.zeros((nc, m, ntr))
    be = np.zeros((nc, m))
    # buffers are filled column-major per cell: view them transposed
    At = np.zeros((nc, ntr * ntr))
    Kt = np.zeros((nc, m * ntr))
    _chk(lib().orc_condense(C.byref(P), _ptr(mesh.xy, _d), _ptr(mesh.cells, _i64), nc,
                            _ptr(At, _d), _ptr(bte, _d), _ptr(Kt, _d), _ptr(be, _d), nthreads),
         "condense")
    Ate[:] = At.reshape(nc, ntr, ntr).transpose(0, 2, 1)
    Ke[:] = Kt.reshape(nc, ntr, m).transpose(0, 2, 1)
    return Ate, bte, Ke, be


class CSC:
    """Julia SparseMatrixCSC{Float64,Int} arrays (1-based)."""

    def __init__(self, n, colptr, rowval, nzval):
        self.n, self.colptr, self.rowval, self.nzval = n, colptr, rowval, nzval

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csc_matrix((self.nzval.copy(), self.rowval - 1, self.colptr - 1),
                             shape=(self.n, self.n))

    def copy(self):
        return CSC(self.n, self.colptr.copy(), self.rowval.copy(), self.nzval.copy())


def assemble(mesh, nt, Ate, bte):
    """end_assemble(assembler) and the assembled rhs (poisson2D_HDG.jl:176-185)."""
    L = lib()
    nc = mesh.ncells
    ntr = 3 * nt
    At = np.ascontiguousarray(np.asarray(Ate).transpose(0, 2, 1)).reshape(nc, ntr * ntr)
    bt = np.ascontiguousarray(bte)
    nnz = C.c_int64()
    _chk(L.orc_assemble(nc, _ptr(mesh.cells, _i64), mesh.nfaces, nt, _ptr(At, _d), _ptr(bt, _d),
                        None, None, None, None, C.byref(nnz)), "assemble(nnz)")
    n = mesh.nfaces * nt
    colptr = np.zeros(n + 1, dtype=np.int64)
    rowval = np.zeros(nnz.value, dtype=np.int64)
    nzval = np.zeros(nnz.value)
    rhs = np.zeros(n)
    _chk(L.orc_assemble(nc, _ptr(mesh.cells, _i64), mesh.nfaces, nt, _ptr(At, _d), _ptr(bt, _d),
                        _ptr(colptr, _i64), _ptr(rowval, _i64), _ptr(nzval, _d), _ptr(rhs, _d),
                        None), "assemble")
    return CSC(n, colptr, rowval, nzval), rhs


def dirichlet_trace(P, mesh, faceset="boundary", g=None, quirk=False):
    """Dirichlet(u_hat, mesh, faceset, g) (boundary.jl:11-42); g None means
    g = 0, otherwise (g0, g1, g2) for g = g0 + g1 x + g2 y.  quirk False is
    the L2 projection onto the orthonormal trace basis; True is the
    reference's computation (its accumulator is not reset between modes,
    boundary.jl:27-37, and the face points are weighted with the cell rule's
    weights, :33); an int selects those two deviations as bits 1 and 2."""
    quirk = 3 if quirk is True else int(quirk)
    mask = mesh.boundary_mask(faceset)
    gk, gp = (0, np.zeros(3)) if g is None else (1, np.asarray(g, dtype=np.float64))
    nt = P.k + 1
    nmax = int(mask.sum()) * nt
    dofs = np.zeros(max(nmax, 1), dtype=np.int64)
    vals = np.zeros(max(nmax, 1))
    nd = C.c_int64()
    _chk(lib().orc_dirichlet_trace(C.byref(P), _ptr(mesh.xy, _d), _ptr(mesh.cells, _i64),
                                   _ptr(mesh.faces_cm, _i64), mesh.nfaces, _ptr(mask, _i8), gk,
                                   _ptr(gp, _d), int(quirk), _ptr(dofs, _i64), _ptr(vals, _d),
                                   C.byref(nd)), "dirichlet")
    return dofs[: nd.value], vals[: nd.value]


def apply(K, rhs, dofs, vals):
    """apply!(K, b, dbc) in place (boundary.jl:121-158)."""
    _chk(lib().orc_apply(K.n, _ptr(K.colptr, _i64), _ptr(K.rowval, _i64), _ptr(K.nzval, _d),
                         _ptr(rhs, _d), len(dofs), _ptr(dofs, _i64), _ptr(vals, _d)), "apply")


def banded_solve(K, b):
    x = np.zeros(K.n)
    _chk(lib().orc_banded_solve(K.n, _ptr(K.colptr, _i64), _ptr(K.rowval, _i64), _ptr(K.nzval, _d),
                                _ptr(np.ascontiguousarray(b), _d), _ptr(x, _d)), "banded_solve")
    return x


def pcg_solve(K, b, nt, dofs, vals, rtol=1e-12, maxit=100000, nthreads=0):
    x = np.zeros(K.n)
    it = C.c_int()
    rr = np.zeros(1)
    dofs = np.ascontiguousarray(dofs, dtype=np.int64)
    vals = np.ascontiguousarray(vals, dtype=np.float64)
    st = lib().orc_pcg_solve(K.n, _ptr(K.colptr, _i64), _ptr(K.rowval, _i64), _ptr(K.nzval, _d),
                             _ptr(np.ascontiguousarray(b), _d), nt, _ptr(dofs, _i64),
                             _ptr(vals, _d), len(dofs), rtol, maxit, _ptr(x, _d), C.byref(it),
                             _ptr(rr, _d), nthreads)
    if st < 0:
        raise RuntimeError(f"oracle pcg failed with status {st}")
    return x, it.value, rr[0]


def recover(mesh, k, Ke, be, uhat):
    nu = (k + 1) * (k + 2) // 2
    nq, nt, nc = 2 * nu, k + 1, mesh.ncells
    Kt = np.ascontiguousarray(np.asarray(Ke).transpose(0, 2, 1)).reshape(nc, -1)
    q = np.zeros((nq, nc))
    u = np.zeros((nu, nc))
    uc = np.zeros((3, nt, nc))
    _chk(lib().orc_recover(nc, _ptr(mesh.cells, _i64), k, _ptr(Kt, _d),
                           _ptr(np.ascontiguousarray(be), _d), _ptr(np.ascontiguousarray(uhat), _d),
                           _ptr(q, _d), _ptr(u, _d), _ptr(uc, _d)), "recover")
    return q.T, u.T, uc.transpose(2, 1, 0)  # (nc, nq), (nc, nu), (nc, nt, 3)


def errornorm(P, mesh, u, uex=None):
    """Squared L2 error of u (DiscreteFunctions.jl:97-120).  u: (ncells, nu).
    uex None means sin(pi x) sin(pi y); else (c0, c1, c2) for c0+c1 x+c2 y."""
    uf = np.asfortranarray(u, dtype=np.float64)
    kind, par = (0, np.zeros(3)) if uex is None else (1, np.asarray(uex, dtype=np.float64))
    return lib().orc_errornorm(C.byref(P), _ptr(mesh.xy, _d), _ptr(mesh.cells, _i64), mesh.ncells,
                               _ptr(uf, _d), kind, _ptr(par, _d))


def residual(K, x, b):
    r = np.zeros(K.n)
    lib().orc_residual(K.n, _ptr(K.colptr, _i64), _ptr(K.rowval, _i64), _ptr(K.nzval, _d),
                       _ptr(np.ascontiguousarray(x), _d), _ptr(np.ascontiguousarray(b), _d),
                       _ptr(r, _d))
    return r


def solve_pipeline(P, mesh, direct=True, rtol=1e-12, nthreads=0):
    """Full reference pipeline: condense, assemble, Dirichlet g=0 + apply!,
    K \\ b (banded direct, or PCG), recovery.  Returns a dict."""
    nt = P.k + 1
    Ate, bte, Ke, be = condense(P, mesh, nthreads)
    K, rhs = assemble(mesh, nt, Ate, bte)
    dofs, vals = dirichlet_trace(P, mesh)
    out = dict(Ate=Ate, bte=bte, Ke=Ke, be=be, K=K.copy(), rhs=rhs.copy(), dofs=dofs, vals=vals)
    if direct:
        Ka, ba = K.copy(), rhs.copy()
        apply(Ka, ba, dofs, vals)
        uhat = banded_solve(Ka, ba)
        out["iters"] = 0
    else:
        uhat, it, rr = pcg_solve(K, rhs, nt, dofs, vals, rtol=rtol, nthreads=nthreads)
        out["iters"], out["relres"] = it, rr
    q, u, uc = recover(mesh, P.k, Ke, be, uhat)
    out.update(uhat=uhat, q=q, u=u, uhat_cell=uc)
    return out
