"""hdg_errornorm_noalloc (the error norm for callers that cannot allocate
device memory) against hdg_errornorm on the same solutions: two calls in a
row, the second on a smaller mesh, equal bit for bit (the same kernels and
reduction order on the same data).  The test's own device buffers come from
the HIP runtime libhdg is linked against (the test allocates; the entry
point does not)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libhdg uses."""

    def __init__(self, hdg):
        hdg.lib()  # loads libamdhip64.so.7 as libhdg's dependency
        self.rt = C.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipDeviceSynchronize.argtypes = []
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, n):
        out = np.zeros(n)
        assert self.rt.hipDeviceSynchronize() == 0
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def close(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


@pytest.mark.parametrize("order", [((64, 48, 3), (20, 17, 3)), ((40, 40, 2), (9, 13, 1)), ((33, 31, 4), (5, 4, 4))])
def test_errornorm_noalloc_equals_errornorm(hdg, order):
    lib = hdg.lib()
    hip = Hip(hdg)
    out = hip.upload(np.zeros(2))
    for call, (nx, ny, k) in enumerate(order):
        print(f"[noalloc] call {call}: {nx}x{ny} k={k}", flush=True)
        m = hdg.rectangle_mesh((nx, ny), jitter=0.2, seed=nx + 7 * ny)
        qdeg = {1: 2, 2: 3, 3: 6, 4: 6}[k]
        with hdg.Context(0) as ctx:
            ctx.set_mesh(m)
            ctx.setup(k, quad_degree=qdeg, rtol=1e-12)
            ctx.set_source()
            d = m.boundary_dofs(k + 1)
            ctx.set_dirichlet(d, np.zeros(len(d)))
            ctx.run()
            q, u, _ = ctx.recover()
            e_ctx = ctx.errornorm()
            g = np.array([0.3, -1.2, 0.7])
            e_lin = ctx.errornorm(kind=hdg._lib.UEX_LINEAR, par=g)
        xy = hip.upload(np.asarray(m.nodes, dtype=np.float64))
        cells = hip.upload(np.asarray(m.cells[:, :3] - 1, dtype=np.int32))
        qu = hip.upload(np.concatenate([q, u], axis=1).astype(np.float64))
        n = m.cells.shape[0]
        st = lib.hdg_errornorm_noalloc(k, qdeg, n, xy, cells, qu, hdg._lib.UEX_SINSIN, None, out, None)
        assert st == 0
        gp = np.ascontiguousarray(g)
        st = lib.hdg_errornorm_noalloc(k, qdeg, n, xy, cells, qu, hdg._lib.UEX_LINEAR,
                                       gp.ctypes.data_as(C.POINTER(C.c_double)), C.c_void_p(out.value + 8), None)
        assert st == 0
        got = hip.download(out, 2)
        assert got[0] == e_ctx, (got[0], e_ctx)
        assert got[1] == e_lin, (got[1], e_lin)
    hip.close()


def test_errornorm_noalloc_rejects(hdg):
    lib = hdg.lib()
    hip = Hip(hdg)
    p = hip.upload(np.zeros(8))
    assert lib.hdg_errornorm_noalloc(5, 0, 1, p, p, p, 0, None, p, None) == 9      # k outside 1..4
    assert lib.hdg_errornorm_noalloc(2, 7, 1, p, p, p, 0, None, p, None) == 3      # quad degree above 6
    assert lib.hdg_errornorm_noalloc(2, 3, 0, p, p, p, 0, None, p, None) == 1      # no cells
    assert lib.hdg_errornorm_noalloc(2, 3, 1 << 31, p, p, p, 0, None, p, None) == 1  # beyond the maximum
    assert lib.hdg_errornorm_noalloc(2, 3, 1, None, p, p, 0, None, p, None) == 1   # null input
    assert lib.hdg_errornorm_noalloc(2, 3, 1, p, p, p, 1, None, p, None) == 1      # linear without parameters
    hip.close()
