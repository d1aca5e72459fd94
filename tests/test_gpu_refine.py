"""hdg_refine_probe on the GPU against the same refinement done with scipy's
sparse LU on the oracle's trace matrix (poisson2D_HDG.jl:176-195 assembles
it): the matrix rounded to float entry by entry, corrections solved exactly
by LU, residuals with the FP64 matrix.  The first contraction is fixed by
the rounded operator alone, so the probe (whose corrections come from CG to
1e-12) must give it to 1e-2 relative; later steps sit near the FP64 floor."""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

pytestmark = pytest.mark.gpu


def _lu_refinement(O, n, steps):
    om = O.rectangle_mesh(n, n, jitter=0.2, seed=5)
    P = O.problem(k=3, qdeg=6)
    Ate, bte, _, _ = O.condense(P, om)
    K, rhs = O.assemble(om, 4, Ate, bte)
    d, _ = O.dirichlet_trace(P, om)
    A = K.to_scipy().tocsr()
    free = np.ones(K.n, bool)
    free[d - 1] = False
    Af = A[free][:, free].tocsc()
    b = rhs[free]
    lu = sla.splu(Af.astype(np.float32).astype(np.float64))
    x = np.zeros_like(b)
    hist = []
    for _ in range(steps + 1):
        r = b - Af @ x
        hist.append(np.linalg.norm(r) / np.linalg.norm(b))
        x = x + lu.solve(r)
    return np.array(hist)


@pytest.mark.parametrize("n", [16, 32])
def test_refine_probe_vs_lu(hdg, O, n):
    print(f"[refine] n={n}", flush=True)
    ref = _lu_refinement(O, n, 3)
    m = hdg.rectangle_mesh((n, n), jitter=0.2, seed=5)
    d = m.boundary_dofs(4)
    with hdg.Context(0) as ctx:
        ctx.set_mesh(m)
        ctx.setup(3, quad_degree=6, rtol=1e-13)
        ctx.set_source()
        ctx.set_dirichlet(d, np.zeros(len(d)))
        ctx.assemble_condense()
        ctx.solve()
        u64 = ctx.trace()
        rel, its = ctx.refine_probe(3, inner_rtol=1e-12)
        u_ref = ctx.trace()
        # the assembled FP64 system is left as it was
        ctx.solve()
        assert np.array_equal(ctx.trace(), u64)
    print(f"probe {rel} inner {its}; LU {ref}", flush=True)
    assert rel[0] == 1.0
    assert abs(rel[1] - ref[1]) <= 1e-2 * ref[1]
    assert rel[2] <= 2.0 * ref[2] + 1e-12
    assert np.all(its > 0)
    assert np.abs(u_ref - u64).max() <= 1e-9 * np.abs(u64).max()


def test_refine_probe_one_step_is_the_rounded_solution(hdg, O):
    """outer = 1 leaves the solution of the rounded system: its distance to
    the FP64 solution is what the LU of the rounded matrix gives."""
    n = 16
    om = O.rectangle_mesh(n, n, jitter=0.2, seed=5)
    P = O.problem(k=3, qdeg=6)
    Ate, bte, _, _ = O.condense(P, om)
    K, rhs = O.assemble(om, 4, Ate, bte)
    d, _ = O.dirichlet_trace(P, om)
    A = K.to_scipy().tocsr()
    free = np.ones(K.n, bool)
    free[d - 1] = False
    Af = A[free][:, free].tocsc()
    x64 = sla.spsolve(Af, rhs[free])
    x32 = sla.spsolve(Af.astype(np.float32).astype(np.float64).tocsc(), rhs[free])
    m = hdg.rectangle_mesh((n, n), jitter=0.2, seed=5)
    with hdg.Context(0) as ctx:
        ctx.set_mesh(m)
        ctx.setup(3, quad_degree=6, rtol=1e-13)
        ctx.set_source()
        ctx.set_dirichlet(m.boundary_dofs(4), np.zeros(len(m.boundary_dofs(4))))
        ctx.assemble_condense()
        rel, _ = ctx.refine_probe(1, inner_rtol=1e-13)
        u = ctx.trace()
    scale = np.abs(x32).max()
    assert np.abs(u[free] - x32).max() <= 1e-9 * scale
    # and it is not the FP64 solution: the rounding shows
    assert np.abs(x32 - x64).max() > 1e-8 * scale
    assert rel[1] > 1e-7
