"""Context: one libhdg device context (hdg_ctx) with numpy in and out."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, lib, ptr, _i64


class Context:
    """Owns an hdg_ctx on HIP device `device`.  All device memory (mesh,
    tables, element operators, trace system, solver vectors) stays resident
    until close()."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._h = C.c_void_p()
        check(lib().hdg_create(C.byref(self._h), int(device), int(hbm_budget)))
        self.mesh = None
        self.sizes = None
        self.k = None

    # -- lifetime
    def close(self):
        if self._h:
            lib().hdg_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, code):
        check(code, self._h)

    def memory_used(self):
        n = C.c_size_t()
        self._chk(lib().hdg_memory_used(self._h, C.byref(n)))
        return n.value

    # -- set-up
    def set_rectangle_mesh(self, nx, ny, ll=(0.0, 0.0), ur=(1.0, 1.0), jitter=0.0, seed=0):
        """rectangle_mesh generated on the device (hdg_set_rectangle_mesh):
        the mesh rectangle_mesh((nx, ny), ...) builds, without host arrays."""
        self._chk(lib().hdg_set_rectangle_mesh(self._h, int(nx), int(ny), float(ll[0]), float(ll[1]),
                                               float(ur[0]), float(ur[1]), float(jitter), int(seed)))

    def set_dirichlet_boundary(self, value=0.0):
        """u = value on every dof of every boundary face."""
        self._chk(lib().hdg_set_dirichlet_boundary(self._h, float(value)))

    def set_mesh(self, mesh):
        faces_cm = np.asfortranarray(mesh.faces, dtype=np.int64)
        self._chk(lib().hdg_set_mesh(self._h, mesh.nodes.shape[0], ptr(mesh.nodes), mesh.cells.shape[0],
                                     ptr(mesh.cells, _i64), mesh.faces.shape[0],
                                     faces_cm.ctypes.data_as(_i64)))
        self.mesh = mesh
        self.sizes = None

    def setup(self, k, quad_degree=None, tau=1.0, precision=L.HDG_FP64, rtol=1e-12, maxit=0,
              precond=L.PC_AUTO, schedule=0):
        o = L.Options(k=int(k), quad_degree=int(quad_degree or 0), tau=float(tau),
                      precision=int(precision), rtol=float(rtol), maxit=int(maxit), schedule=int(schedule),
                      precond=int(precond))
        self._chk(lib().hdg_setup(self._h, C.byref(o)))
        s = L.Sizes()
        self._chk(lib().hdg_query_sizes(self._h, C.byref(s)))
        self.sizes = s
        self.k = int(k)

    def set_source(self, kind=L.SRC_REFERENCE, kappa=None, coef=None, f_qp=None):
        s = L.Source()
        s.kind = int(kind)
        if kappa is not None:
            s.kappa_on, s.ka, s.kb = 1, float(kappa[0]), float(kappa[1])
        if coef is not None:
            for i, c in enumerate(np.asarray(coef, dtype=np.float64).ravel()[:16]):
                s.coef[i] = c
        keep = None
        if f_qp is not None:
            keep = np.asfortranarray(f_qp, dtype=np.float64)
            s.f_qp = keep.ctypes.data
        self._chk(lib().hdg_set_source(self._h, C.byref(s)))

    def set_dirichlet_function(self, kind=L.G_LINEAR, par=(0.0, 0.0, 0.0), reference=False):
        """Dirichlet data on every boundary face computed on the device
        (hdg_set_dirichlet_function): g = par[0] + par[1] x + par[2] y or
        sin(pi x) sin(pi y); reference=True reproduces the reference's
        weights and accumulator (boundary.jl:27-37)."""
        p = np.ascontiguousarray(par, dtype=np.float64)
        self._chk(lib().hdg_set_dirichlet_function(self._h, int(kind), ptr(p),
                                                   L.DIRICHLET_REFERENCE if reference else 0))

    def get_applied_system(self):
        """apply!(K, b, dbc) on the device: (colptr, rowval, nzval, rhs,
        meandiag) of the applied system, Julia's CSC, 1-based."""
        s = self.sizes
        colptr = np.zeros(s.n_dofs + 1, dtype=np.int64)
        rowval = np.zeros(s.nnz, dtype=np.int64)
        nzval = np.zeros(s.nnz)
        rhs = np.zeros(s.n_dofs)
        md = np.zeros(1)
        self._chk(lib().hdg_get_applied_system(self._h, ptr(colptr, _i64), ptr(rowval, _i64), ptr(nzval),
                                               ptr(rhs), ptr(md)))
        return colptr, rowval, nzval, rhs, float(md[0])

    def set_dirichlet(self, dofs, vals):
        d = np.ascontiguousarray(dofs, dtype=np.int64)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        self._chk(lib().hdg_set_dirichlet(self._h, len(d), ptr(d, _i64), ptr(v)))

    # -- pipeline
    def assemble_condense(self):
        self._chk(lib().hdg_assemble_condense(self._h))

    def get_system(self):
        """(colptr, rowval, nzval, rhs): Julia's CSC of K, 1-based."""
        s = self.sizes
        colptr = np.zeros(s.n_dofs + 1, dtype=np.int64)
        rowval = np.zeros(s.nnz, dtype=np.int64)
        nzval = np.zeros(s.nnz)
        rhs = np.zeros(s.n_dofs)
        self._chk(lib().hdg_get_system(self._h, ptr(colptr, _i64), ptr(rowval, _i64), ptr(nzval),
                                       ptr(rhs)))
        return colptr, rowval, nzval, rhs

    def get_system_scipy(self):
        import scipy.sparse as sp
        colptr, rowval, nzval, rhs = self.get_system()
        n = self.sizes.n_dofs
        return sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(n, n)), rhs

    def get_local_ops(self):
        """K_element (nc, m, ntr) and b_element (nc, m)."""
        s = self.sizes
        nc = s.n_cells
        Ke = np.zeros((nc, s.ntr, s.m))
        be = np.zeros((nc, s.m))
        self._chk(lib().hdg_get_local_ops(self._h, ptr(Ke), ptr(be)))
        return Ke.transpose(0, 2, 1), be

    def get_condensed(self):
        """Ate (nc, ntr, ntr) and bte (nc, ntr)."""
        s = self.sizes
        nc = s.n_cells
        At = np.zeros((nc, s.ntr, s.ntr))
        bt = np.zeros((nc, s.ntr))
        self._chk(lib().hdg_get_condensed(self._h, ptr(At), ptr(bt)))
        return At.transpose(0, 2, 1), bt

    def element_blocks(self, first, count):
        """A, B, C, E, F, H, be of 0-based cells [first, first+count)."""
        s = self.sizes
        nq, nu, ntr = s.nq, s.nu, s.ntr
        shp = dict(A=(nq, nq), B=(nq, nu), C=(nu, nu), E=(nq, ntr), F=(nu, ntr), H=(ntr, ntr), be=(nu,))
        buf = {k: np.zeros((count,) + v[::-1]) for k, v in shp.items()}
        self._chk(lib().hdg_get_element_blocks(self._h, int(first), int(count),
                                               *[ptr(buf[k]) for k in ("A", "B", "C", "E", "F", "H", "be")]))
        return {k: (buf[k].transpose(0, 2, 1) if buf[k].ndim == 3 else buf[k]) for k in shp}

    def solve(self, return_uhat=True):
        info = L.SolveInfo()
        uh = np.zeros(self.sizes.n_dofs) if return_uhat else None
        self._chk(lib().hdg_solve(self._h, ptr(uh), C.byref(info)))
        return uh, info

    def trace(self):
        """The trace solution of the last solve / run (hdg_get_trace)."""
        u = np.zeros(self.sizes.n_dofs)
        self._chk(lib().hdg_get_trace(self._h, ptr(u)))
        return u

    def set_trace(self, uhat):
        u = np.ascontiguousarray(uhat, dtype=np.float64)
        self._chk(lib().hdg_set_trace(self._h, ptr(u)))

    def recover(self):
        """q (nc, nq), u (nc, nu), uhat_cell (nc, nt, 3), as m_values."""
        s = self.sizes
        nc = s.n_cells
        q = np.zeros((s.nq, nc))
        u = np.zeros((s.nu, nc))
        uc = np.zeros((3, s.nt, nc))
        self._chk(lib().hdg_recover(self._h, ptr(q), ptr(u), ptr(uc)))
        return q.T, u.T, uc.transpose(2, 1, 0)

    def errornorm(self, kind=L.UEX_SINSIN, par=None):
        out = np.zeros(1)
        p = None if par is None else np.ascontiguousarray(par, dtype=np.float64)
        self._chk(lib().hdg_errornorm(self._h, int(kind), ptr(p), ptr(out)))
        return float(out[0])

    def run(self):
        info = L.SolveInfo()
        self._chk(lib().hdg_run(self._h, C.byref(info)))
        return info

    def autotune(self):
        """Times the condensation kernel's schedules on the current problem
        and keeps the fastest (hdg_autotune); returns the choice as a dict.
        Assemble again before solving."""
        sc = L.Schedule()
        self._chk(lib().hdg_autotune(self._h, C.byref(sc)))
        return {f: getattr(sc, f) for f, _ in L.Schedule._fields_}

    def set_schedule(self, condense_seg=0, condense_grid=0, condense_waves=0, **_ignored):
        """Sets the condensation schedule (hdg_set_schedule); accepts the
        dict autotune returns."""
        sc = L.Schedule(condense_seg=int(condense_seg), condense_grid=int(condense_grid),
                        condense_waves=int(condense_waves))
        self._chk(lib().hdg_set_schedule(self._h, C.byref(sc)))

    def refine_probe(self, outer, inner_rtol=0.0):
        """FP64 iterative refinement whose corrections solve with the trace
        matrix rounded to float (hdg_refine_probe).  Returns (relres, inner
        iterations): ||r_j|| / ||r_0|| for j = 0..outer and the CG
        iterations of each correction.  Leaves the refined iterate as the
        trace."""
        rel = np.zeros(int(outer) + 1)
        its = np.zeros(max(int(outer), 1), dtype=np.int32)
        self._chk(lib().hdg_refine_probe(self._h, int(outer), float(inner_rtol), ptr(rel), its.ctypes.data))
        return rel, its[:int(outer)]

    def time_kernel(self, which, reps=50):
        out = np.zeros(1)
        self._chk(lib().hdg_time_kernel(self._h, int(which), int(reps), ptr(out)))
        return float(out[0])
