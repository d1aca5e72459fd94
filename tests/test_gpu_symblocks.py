"""At four trace modes per face (k = 3) the CG product and the multigrid
residual run on symmetric blocks: each face stores K_ff and, in each of its
cells, the block coupling it to the next local face; the other two
couplings are read transposed from the neighbours that own them.  The
solution of that operator is checked against the assembled face-ELL matrix
(hdg_get_system, the matrix the reference assembles, assembler.jl:31-49) on
the host, on a jittered rectangle (multigrid), with a kappa source (the
elimination form of the condensation) and on the unstructured Triangle
fixture (block-Jacobi, arbitrary local face numbering)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def host_relres(ctx, uhat, dofs):
    """||K uhat - rhs|| / ||rhs - K uD|| over the free rows, with K and rhs
    exported from the device in Julia's CSC (1-based)."""
    colptr, rowval, nzval, rhs = ctx.get_system()
    n = len(rhs)
    K = sp.csc_matrix((nzval, rowval - 1, colptr - 1), shape=(n, n))
    free = np.ones(n, dtype=bool)
    free[np.asarray(dofs) - 1] = False
    uD = np.where(free, 0.0, uhat)
    r = (K @ uhat - rhs)[free]
    b = (rhs - K @ uD)[free]
    return float(np.linalg.norm(r) / np.linalg.norm(b))


@pytest.mark.parametrize("case", ["rect", "kappa", "figure"])
def test_symmetric_block_product(hdg, case):
    print(f"[symmetric blocks] {case}", flush=True)
    if case == "figure":
        m = hdg.parse_mesh_triangle(os.path.join(GOLDEN, "figure.1"))
    else:
        m = hdg.rectangle_mesh((37, 23), jitter=0.25, seed=5)
    src = dict(kind=hdg._lib.SRC_KAPPA, kappa=(3.0, 5.0)) if case == "kappa" else {}
    with hdg.Context(0) as ctx:
        ctx.set_mesh(m)
        ctx.setup(3, quad_degree=6, rtol=1e-13, maxit=20000)
        ctx.set_source(**src)
        d = m.boundary_dofs(4)
        g = np.linspace(-0.5, 0.5, len(d))  # non-zero Dirichlet data: the lift reads K's columns
        ctx.set_dirichlet(d, g)
        ctx.run()
        uhat = ctx.trace()
        rr = host_relres(ctx, uhat, d)
        info_rel = ctx.solve(return_uhat=False)[1].true_relres
    print(f"{case}: host relres {rr:.2e}, device {info_rel:.2e}", flush=True)
    # recursive residual 1e-13 on the block operator; the assembled matrix
    # sees the same solution to the FP64 floor of ||K|| ||uhat|| eps
    assert rr < 1e-11
    assert info_rel < 1e-11


@pytest.mark.parametrize("nx,ny", [(1, 1), (2, 1), (1, 3), (3, 2)])
def test_symmetric_blocks_small_meshes(hdg, O, nx, ny):
    """Meshes where most faces lie on the boundary (their second block and
    neighbour are the zero placeholders): the operators, the exported K and
    the solution (block-Jacobi CG: the product is the same one) against the
    oracle's direct solve to 1e-12."""
    print(f"[symmetric blocks] {nx}x{ny}", flush=True)
    m = hdg.rectangle_mesh((nx, ny), jitter=0.1, seed=3)
    om = O.rectangle_mesh(nx, ny, jitter=0.1, seed=3)
    P = O.problem(k=3, qdeg=6)
    r = O.solve_pipeline(P, om)
    with hdg.Context(0) as ctx:
        ctx.set_mesh(m)
        ctx.setup(3, quad_degree=6, rtol=1e-15, maxit=20000, precond=hdg.PC_BLOCK_JACOBI)
        ctx.set_source()
        ctx.set_dirichlet(r["dofs"], r["vals"])
        ctx.run()
        uhat = ctx.trace()
        colptr, rowval, nzval, rhs = ctx.get_system()
    rel = lambda a, b: float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
    assert np.array_equal(colptr, r["K"].colptr) and np.array_equal(rowval, r["K"].rowval)
    assert rel(nzval, r["K"].nzval) < 1e-12 and rel(rhs, r["rhs"]) < 1e-12
    assert rel(uhat, r["uhat"]) < 1e-12
