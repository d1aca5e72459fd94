"""The HDG driver API of the reference, backed by libhdg on the GPU.

Mirrors examples/poisson2D_HDG.jl:
    K, b, K_e, b_e = doassemble(Vh, Wh, Mh)        # :58-186
    apply_bc(K, b, dbc)                            # apply!, :194 / boundary.jl:121-158
    uhat = spsolve(K, b)                           # K \\ b, :195
    get_usigma(sigma_h, u_h, uhat_h, uhat, K_e, b_e, mesh)   # get_uσ!, :197-212
    errornorm(u_h, u_ex)                           # :217 / DiscreteFunctions.jl:97-120
and adds the fused device-resident path
    hdg_solve(sigma_h, u_h, uhat_h, mesh, dbc)     # assemble + condense + solve + recover
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .context import Context
from .spaces import ScalarTraceFunctionSpace, TrialFunction, quadrature

_default = {"ctx": None}


def default_context():
    if _default["ctx"] is None:
        _default["ctx"] = Context(0)
    return _default["ctx"]


def _source_args(f, Wh, mesh):
    """f: None (the reference's 2 pi^2 sin sin), a dict of libhdg source
    parameters, or a callable f(x) evaluated on the host at the cell
    quadrature points (function_value, DiscreteFunctions.jl:6-9)."""
    if f is None:
        return dict(kind=L.SRC_REFERENCE)
    if isinstance(f, dict):
        return f
    pts, _ = quadrature(2, Wh.quad_degree)
    x = mesh.nodes[mesh.cells[:, :3] - 1]  # (nc, 3, 2)
    lam = np.stack([1.0 - pts[:, 0] - pts[:, 1], pts[:, 0], pts[:, 1]], axis=1)  # (nqp, 3)
    xq = np.einsum("qv,cvd->cqd", lam, x)
    vals = np.array([[float(f(p)) for p in row] for row in xq])
    return dict(kind=L.SRC_VALUES, f_qp=vals)


def _prepare(ctx, Vh, Wh, Mh, tau, f):
    mesh = Wh.mesh
    if ctx.mesh is not mesh:
        ctx.set_mesh(mesh)
    ctx.setup(Wh.k, quad_degree=Wh.quad_degree, tau=tau)
    ctx.set_source(**_source_args(f, Wh, mesh))


def doassemble(Vh, Wh, Mh, tau=1.0, f=None, ctx=None):
    """doassemble(Vh, Wh, Mh, tau) (poisson2D_HDG.jl:58-186) on the device.
    Returns (K, rhs, K_element, b_element): K as scipy CSC (Julia's
    SparseMatrixCSC, 0-based here), K_element[c] m x ntr, b_element[c] m."""
    ctx = ctx or default_context()
    _prepare(ctx, Vh, Wh, Mh, tau, f)
    ctx.assemble_condense()
    K, rhs = ctx.get_system_scipy()
    Ke, be = ctx.get_local_ops()
    return K, rhs, Ke, be


def meandiag(K):
    """boundary.jl:169-175 (a running sum in row order, like the reference)."""
    return float(np.cumsum(np.abs(K.diagonal()))[-1] / K.shape[0])


def apply_bc(K, f, dbc):
    """apply!(K, f, dbc) with APPLY_TRANSPOSE (boundary.jl:121-158), in place
    on a scipy CSC matrix: lift the rhs, zero Dirichlet rows and columns and
    put meandiag(K) on their diagonal."""
    m = meandiag(K)
    dofs = dbc.prescribed_dofs - 1
    vals = dbc.values
    for d, v in zip(dofs, vals):
        if v != 0:
            lo, hi = K.indptr[d], K.indptr[d + 1]
            f[K.indices[lo:hi]] -= v * K.data[lo:hi]
    isd = np.zeros(K.shape[0], dtype=bool)
    isd[dofs] = True
    cols = np.repeat(np.arange(K.shape[0]), np.diff(K.indptr))
    K.data[isd[cols] | isd[K.indices]] = 0.0
    for d, v in zip(dofs, vals):
        lo, hi = K.indptr[d], K.indptr[d + 1]
        K.data[lo + np.searchsorted(K.indices[lo:hi], d)] = m
        f[d] = v * m


def get_usigma(sigma_h, u_h, uhat_h, uhat, K_e, b_e, mesh, ctx=None):
    """get_uσ! (poisson2D_HDG.jl:197-212), with the trace dofs of face f
    taken as (f-1)nt+1 .. f nt for every order.  With a context that holds
    the assembled system the recovery runs on the device; otherwise the
    given K_e, b_e are applied on the host."""
    nt = uhat_h.fs.nbase
    if ctx is not None and ctx.mesh is mesh:
        ctx.set_trace(uhat)
        q, u, uc = ctx.recover()
    else:
        idx = (mesh.cells[:, 3:][:, :, None] - 1) * nt + np.arange(nt)[None, None, :]
        ue = np.asarray(uhat)[idx.reshape(len(mesh.cells), -1)]
        dof = np.einsum("cij,cj->ci", K_e, ue) + b_e
        nq = sigma_h.fs.nbase
        q, u = dof[:, :nq], dof[:, nq:]
        uc = np.asarray(uhat)[idx].transpose(0, 2, 1)
    sigma_h.m_values[:] = q
    u_h.m_values[:] = u
    uhat_h.m_values[:] = uc


def errornorm(u_h, u_ex, ctx=None):
    """errornorm(u_h, u_ex) -- the squared L2 error (DiscreteFunctions.jl:97-120).
    u_ex: "sinsin" for sin(pi x) sin(pi y), or a callable u_ex(x)."""
    fs = u_h.fs
    mesh = fs.mesh
    if ctx is not None and u_ex == "sinsin":
        return ctx.errornorm(L.UEX_SINSIN)
    from ._lib import lib, ptr
    pts, w = quadrature(2, fs.quad_degree)
    nu = (fs.k + 1) * (fs.k + 2) // 2
    N = np.zeros((len(w), nu))
    for q, (r, s) in enumerate(pts):
        v = np.zeros(nu)
        lib().hdg_dubiner(fs.k, r, s, ptr(v), None)
        N[q] = v
    x = mesh.nodes[mesh.cells[:, :3] - 1]
    J = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]], axis=2)
    detJ = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
    lam = np.stack([1.0 - pts[:, 0] - pts[:, 1], pts[:, 0], pts[:, 1]], axis=1)
    xq = np.einsum("qv,cvd->cqd", lam, x)
    uh = u_h.m_values @ N.T  # (nc, nqp)
    if u_ex == "sinsin":
        ue = np.sin(np.pi * xq[..., 0]) * np.sin(np.pi * xq[..., 1])
    else:
        ue = np.array([[u_ex(p) for p in row] for row in xq])
    return float((((uh - ue) ** 2) * w[None, :]).sum(axis=1) @ detJ)


def hdg_solve(sigma_h, u_h, uhat_h, mesh, dbc=None, f=None, tau=1.0, precision=L.HDG_FP64,
              rtol=1e-12, maxit=0, ctx=None):
    """Fused device-resident pipeline: element build, condensation, trace
    assembly, Dirichlet elimination + PCG solve, recovery.  Fills the three
    trial functions and returns the libhdg solve record."""
    ctx = ctx or default_context()
    Wh = u_h.fs
    if ctx.mesh is not mesh:
        ctx.set_mesh(mesh)
    ctx.setup(Wh.k, quad_degree=Wh.quad_degree, tau=tau, precision=precision, rtol=rtol, maxit=maxit)
    ctx.set_source(**_source_args(f, Wh, mesh))
    if dbc is not None:
        ctx.set_dirichlet(dbc.prescribed_dofs, dbc.values)
    info = ctx.run()
    q, u, uc = ctx.recover()
    sigma_h.m_values[:] = q
    u_h.m_values[:] = u
    uhat_h.m_values[:] = uc
    return info


__all__ = ["Context", "TrialFunction", "ScalarTraceFunctionSpace", "doassemble", "apply_bc",
           "get_usigma", "errornorm", "hdg_solve", "meandiag", "default_context"]
