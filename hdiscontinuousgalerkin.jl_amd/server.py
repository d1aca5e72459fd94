"""Python handles on libhdg's serving loop (hdg_server_* in include/hdg.h)
and batches (hdg_batch_*).

    srv = Server(device=0, hbm_budget=8 << 30, slots=8)
    rid = srv.submit(nx=256, ny=256, k=2)
    res = srv.wait(rid)            # hdg_result as a dict
    srv.stats()                    # p50/p99 latency, cells/s
    srv.close()

    with Batch([make_request(...), ...]) as b:
        results = b.run()          # one dict per request, grouped launches
        uhat = b.trace(0)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import lib

QDEG_DEFAULT = {1: 2, 2: 3, 3: 6, 4: 6}


def make_request(nx, ny, k, quad_degree=None, tau=1.0, precision=L.HDG_FP64, rtol=1e-12, maxit=0,
                 precond=L.PC_AUTO, jitter=0.0, seed=0, src_kind=L.SRC_REFERENCE, kappa=None, coef=None,
                 uex_kind=-1, want_trace=False):
    """An hdg_request; quad_degree defaults to {1:2, 2:3, 3:6, 4:6}."""
    r = L.Request()
    r.nx, r.ny, r.jitter, r.seed = int(nx), int(ny), float(jitter), int(seed)
    r.opt = L.Options(k=int(k), quad_degree=int(quad_degree or QDEG_DEFAULT.get(int(k), int(k) + 1)),
                      tau=float(tau),
                      precision=int(precision), rtol=float(rtol), maxit=int(maxit), schedule=0,
                      precond=int(precond))
    r.src.kind = int(src_kind)
    if kappa is not None:
        r.src.kappa_on, r.src.ka, r.src.kb = 1, float(kappa[0]), float(kappa[1])
    if coef is not None:
        for i, c in enumerate(np.asarray(coef, dtype=np.float64).ravel()[:16]):
            r.src.coef[i] = c
    r.uex_kind = int(uex_kind)
    r.want_trace = 1 if want_trace else 0
    return r


def footprint(req):
    return int(lib().hdg_server_footprint(C.byref(req)))


class Server:
    """batch > 1: a free slot solves up to `batch` queued requests with the
    same options together in grouped launches (at most batch_cells cells,
    0 = no cap); results are bitwise those of solving them alone."""

    def __init__(self, device=0, hbm_budget=0, slots=4, batch=1, batch_cells=0):
        h = C.c_void_p()
        code = lib().hdg_server_create(C.byref(h), int(device), int(hbm_budget), int(slots))
        if code != L.HDG_OK:
            raise L.HDGError(code, "hdg_server_create failed")
        self._h = h
        self._ndofs = {}  # request id -> trace length, from its result
        if batch > 1:
            self._chk(lib().hdg_server_set_batching(self._h, int(batch), int(batch_cells)))

    def _chk(self, code):
        if code != L.HDG_OK:
            raw = lib().hdg_server_last_error(self._h)
            msg = raw.decode() if raw else ""
            if code in (1, 2, 3, 10):
                raise L.ArgumentError(code, msg)
            raise L.HDGError(code, msg)

    def submit(self, req=None, **kw):
        req = req if req is not None else make_request(**kw)
        rid = C.c_int64()
        self._chk(lib().hdg_server_submit(self._h, C.byref(req), C.byref(rid)))
        return int(rid.value)

    def wait(self, rid, timeout=-1.0):
        r = L.Result()
        self._chk(lib().hdg_server_wait(self._h, int(rid), float(timeout), C.byref(r)))
        out = {f: getattr(r, f) for f, _ in L.Result._fields_}
        self._ndofs[int(rid)] = int(out["n_dofs"])
        return out

    def trace(self, rid):
        """The kept trace solution of a completed request (want_trace),
        sized from the request's own result."""
        n = self._ndofs.get(int(rid))
        if n is None:
            raise L.HDGError(8, f"request {rid} has not been waited for")
        out = np.zeros(n)
        self._chk(lib().hdg_server_trace(self._h, int(rid), out.ctypes.data_as(C.POINTER(C.c_double)), n))
        del self._ndofs[int(rid)]
        return out

    def stats(self):
        s = L.ServerSummary()
        self._chk(lib().hdg_server_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in L.ServerSummary._fields_}

    def close(self):
        if self._h:
            lib().hdg_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Batch:
    """Many requests solved together: requests of one order and one set of
    options are packed into one mesh and share every launch (hdg_batch_*)."""

    def __init__(self, reqs, device=0, hbm_budget=0):
        reqs = list(reqs)
        arr = (L.Request * len(reqs))(*reqs)
        h = C.c_void_p()
        code = lib().hdg_batch_create(C.byref(h), int(device), int(hbm_budget), len(reqs), arr)
        self._h = h if h.value else None
        self._n = len(reqs)
        self._ndofs = [None] * len(reqs)
        if code != L.HDG_OK:
            msg = ""
            if self._h:
                raw = lib().hdg_batch_last_error(self._h)
                msg = raw.decode() if raw else ""
                self.close()
            if code in (1, 2, 3, 10):
                raise L.ArgumentError(code, msg)
            raise L.HDGError(code, msg)

    def _chk(self, code, ok=(L.HDG_OK,)):
        if code not in ok:
            raw = lib().hdg_batch_last_error(self._h)
            raise L.HDGError(code, raw.decode() if raw else "")
        return code

    def run(self, check=True):
        """Runs every instance; returns one dict per request.  With check,
        raises when an instance failed (each dict holds its own status)."""
        res = (L.Result * self._n)()
        code = lib().hdg_batch_run(self._h, res)
        if check:
            self._chk(code)
        elif code == 7:  # HDG_ERR_HIP
            self._chk(code)
        out = [{f: getattr(r, f) for f, _ in L.Result._fields_} for r in res]
        self._ndofs = [r["n_dofs"] for r in out]
        return out

    def trace(self, i):
        if self._ndofs[i] is None:
            raise L.HDGError(8, "run the batch first")
        out = np.zeros(int(self._ndofs[i]))
        self._chk(lib().hdg_batch_trace(self._h, int(i), out.ctypes.data_as(C.POINTER(C.c_double)), out.size))
        return out

    def info(self):
        ng, nb = C.c_int(), C.c_size_t()
        self._chk(lib().hdg_batch_info(self._h, C.byref(ng), C.byref(nb)))
        return {"groups": int(ng.value), "device_bytes": int(nb.value)}

    def close(self):
        if self._h:
            lib().hdg_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
