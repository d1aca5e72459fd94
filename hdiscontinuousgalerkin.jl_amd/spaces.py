"""Function spaces, trial functions and Dirichlet data, mirroring
src/FiniteElement.jl, src/*FunctionSpaces.jl, src/DiscreteFunctions.jl and
src/boundary.jl of the reference for the HDG path.

These objects carry the discretisation choices (order, quadrature degree)
and the result arrays (`m_values`); the arithmetic runs in libhdg.
"""
from __future__ import annotations

import numpy as np

from ._lib import lib, ptr


class Dubiner:
    """Dubiner{2,RefTetrahedron,order} (src/basis.jl:50-54)."""

    def __init__(self, order):
        self.order = int(order)

    def nbasefunctions(self):
        return (self.order + 1) * (self.order + 2) // 2


class Legendre:
    """Legendre{1,RefTetrahedron,order} (src/basis.jl:341-344)."""

    def __init__(self, order):
        self.order = int(order)

    def nbasefunctions(self):
        return self.order + 1


class GenericFiniteElement:
    """GenericFiniteElement(func_basis) with P1 geometry (FiniteElement.jl:16-27)."""

    def __init__(self, basis):
        self.basis = basis
        self.order = basis.order


class ScalarFunctionSpace:
    """ScalarFunctionSpace(mesh, fe; quad_degree = order+1)
    (src/ScalarFunctionSpaces.jl:24-29)."""

    def __init__(self, mesh, fe, quad_degree=None):
        self.mesh = mesh
        self.fe = fe
        self.k = fe.order
        self.quad_degree = int(quad_degree) if quad_degree else self.k + 1
        self.nbase = fe.basis.nbasefunctions()


class VectorFunctionSpace(ScalarFunctionSpace):
    """VectorFunctionSpace: two scalar blocks, component-major
    (src/VectorFunctionSpaces.jl:10-18,28-34)."""

    def __init__(self, mesh, fe, quad_degree=None):
        super().__init__(mesh, fe, quad_degree)
        self.nbase = 2 * fe.basis.nbasefunctions()


class ScalarTraceFunctionSpace:
    """ScalarTraceFunctionSpace(Wh, fe): Legendre on faces, sharing Wh's face
    rule (src/TraceFunctionSpaces.jl:11-28)."""

    def __init__(self, Wh, fe):
        self.fs = Wh
        self.mesh = Wh.mesh
        self.fe = fe
        self.nbase = fe.basis.nbasefunctions()


def getnbasefunctions(fs):
    return fs.nbase


def getnlocaldofs(fs):
    """src/ScalarFunctionSpaces.jl:146, VectorFunctionSpaces.jl:24,
    TraceFunctionSpaces.jl:38."""
    return fs.nbase * 3 if isinstance(fs, ScalarTraceFunctionSpace) else fs.nbase


def getmesh(fs):
    return fs.mesh


class TrialFunction:
    """TrialFunction(fs): m_values ncells x nbase (ncells x nt x 3 for the
    trace), NaN-initialised (src/DiscreteFunctions.jl:38-54)."""

    def __init__(self, fs):
        self.fs = fs
        nc = fs.mesh.cells.shape[0]
        if isinstance(fs, ScalarTraceFunctionSpace):
            self.m_values = np.full((nc, fs.nbase, 3), np.nan)
            self.components = 1
        else:
            self.m_values = np.full((nc, fs.nbase), np.nan)
            self.components = 2 if isinstance(fs, VectorFunctionSpace) else 1


def quadrature(dim, order):
    """Points and weights of DefaultQuad (src/quadrature.jl:17-39) from libhdg."""
    pts = np.zeros(2 * 4096)
    w = np.zeros(4096)
    n = lib().hdg_quadrature(dim, order, 4096, ptr(pts), ptr(w))
    if n < 0:
        from ._lib import ArgumentError
        raise ArgumentError(-n, f"Quadrature rule of order {order} not available")
    return (pts[: 2 * n].reshape(n, 2) if dim == 2 else pts[:n].copy()), w[:n].copy()


def legendre01(nt, x):
    """Orthonormal Legendre basis on [0,1] at points x: (nt, len(x))."""
    x = np.asarray(x, dtype=np.float64)
    t = 2.0 * x - 1.0
    P = [np.ones_like(t), t]
    for i in range(2, nt):
        P.append(((2 * i - 1) * t * P[-1] - (i - 1) * P[-2]) / i)
    return np.array([np.sqrt(2 * i + 1) * P[i] for i in range(nt)])


class Dirichlet:
    """Dirichlet(u_hat, mesh, faceset, g) on the trace space
    (src/boundary.jl:7-42): prescribed dofs (1-based, ascending by face, mode)
    and the projection of g onto the orthonormal trace basis.

    The reference weights face point q with the cell rule's weight q
    (fs.fs.qr_weights, boundary.jl:33) and accumulates across modes without
    resetting (:27-37); both only matter for g != 0.  By default this class
    computes the L2 projection, which equals the reference whenever g = 0
    (every configuration the reference ships); reference=True reproduces the
    reference's values for any g."""

    def __init__(self, u_hat, mesh, faceset, g, reference=False):
        fs = u_hat.fs
        nt = fs.nbase
        qdeg = fs.fs.quad_degree
        x, w = quadrature(1, qdeg)
        if reference:
            _, wc = quadrature(2, qdeg)
            if len(wc) < len(w):
                raise ValueError("the cell rule has fewer points than the face rule")
            w = wc[: len(w)]
        Nt = legendre01(nt, x)  # (nt, nq)
        faces = sorted(mesh.facesets[faceset])
        dofs, vals = [], []
        for f in faces:
            if mesh.faces[f - 1, 3] != 0:
                raise AssertionError(f"Face {f} is not in boundary")
            a, b = sorted(mesh.faces[f - 1, :2])
            pa, pb = mesh.nodes[a - 1], mesh.nodes[b - 1]
            # trace parameter runs from the lower to the higher node id
            pts = pa[None, :] + x[:, None] * (pb - pa)[None, :]
            gv = np.array([float(g(p)) for p in pts]) if callable(g) else np.full(len(x), float(g))
            acc = 0.0
            for i in range(nt):
                dofs.append(f * nt - nt + i + 1)
                proj = float(np.dot(w * gv, Nt[i]))
                acc = acc + proj if reference else proj
                vals.append(acc)
        self.prescribed_dofs = np.array(dofs, dtype=np.int64)
        self.values = np.array(vals, dtype=np.float64)
