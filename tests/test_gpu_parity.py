"""GPU parity: libhdg (HIP, through the C ABI) against the CPU oracle and the
reference's literals.

Tolerances: element blocks, condensed operators and assembled values agree
to 1e-12 relative to the largest entry of the compared array (their
rounding differs from the oracle's by the order of operations only);
mesh connectivity, numbering and the CSR pattern are bit-exact; trace
solutions agree within the PCG stopping tolerance times the conditioning
(stated per test).
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

QD = {1: 2, 2: 3, 3: 6, 4: 6}


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def ctx(hdg):
    c = hdg.Context(0)
    yield c
    c.close()


def load(ctx, hdg, O, mesh, om, k, qdeg, src=None, tau=1.0, dirichlet=True):
    ctx.set_mesh(mesh)
    ctx.setup(k, quad_degree=qdeg, tau=tau, rtol=1e-13)
    P = O.problem(k=k, qdeg=qdeg, tau=tau, **(src or {}))
    ctx.set_source(kind=P.src_kind, kappa=(P.ka, P.kb) if P.kappa_on else None,
                   coef=list(P.coef), f_qp=getattr(P, "_keep", None))
    if dirichlet:
        d, v = O.dirichlet_trace(P, om)
        ctx.set_dirichlet(d, v)
    return P


def test_fixture_element_literals(hdg, O, ctx):
    """Element blocks on the reference fixture equal the literals of
    test/test_FunctionSpace.jl:49-72 (A = 0.5 I, B, C, E, H)."""
    from test_oracle_reference import fixture_literals
    m = hdg.parse_mesh_triangle(os.path.join(GOLDEN, "figure2.1"))
    om = O.parse_mesh_triangle(os.path.join(GOLDEN, "figure2.1"))
    P = load(ctx, hdg, O, m, om, 1, 2)
    blk = ctx.element_blocks(0, 4)
    Be, Ce, Ee, He = fixture_literals()
    for c in range(4):
        np.testing.assert_allclose(blk["A"][c], 0.5 * np.eye(6), atol=1e-14)
        np.testing.assert_allclose(blk["B"][c], Be[c], atol=1e-14)
        np.testing.assert_allclose(blk["C"][c], Ce[c], atol=1e-14)
        np.testing.assert_allclose(blk["E"][c], Ee[c], atol=1e-14)
        np.testing.assert_allclose(blk["H"][c], He[c], atol=1e-14)
        ob = O.element_blocks(P, om, c + 1)
        assert rel(blk["F"][c], ob["F"]) < 1e-13
        assert rel(blk["be"][c], ob["be"]) < 1e-13


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_element_blocks_vs_oracle(hdg, O, ctx, k):
    m = hdg.rectangle_mesh((6, 5), jitter=0.2, seed=11)
    om = O.rectangle_mesh(6, 5, jitter=0.2, seed=11)
    P = load(ctx, hdg, O, m, om, k, QD[k])
    blk = ctx.element_blocks(0, m.cells.shape[0])
    for c in range(0, m.cells.shape[0], 7):
        ob = O.element_blocks(P, om, c + 1)
        for key in ("A", "B", "C", "E", "F", "H", "be"):
            assert rel(blk[key][c], ob[key]) < 1e-12, (key, c, rel(blk[key][c], ob[key]))


@pytest.mark.parametrize("k,src,qdeg", [(1, {}, 2), (2, {}, 3), (3, {}, 6), (4, {}, 6),
                                        (2, dict(src_kind=1, kappa=(3.0, 5.0)), 3),
                                        (3, dict(src_kind=2, kappa=(2.0, 7.0), coef=np.linspace(-1, 1, 16)), 6),
                                        (2, dict(src_kind=1, kappa=(3.0, 5.0)), 4), (1, {}, 4)])
def test_condensation_vs_oracle(hdg, O, ctx, k, src, qdeg):
    """Ate, bte, K_e, b_e of every cell (poisson2D_HDG.jl:155-174).  The
    default rules at k = 1 and 3, and degree 4 at k = 1 and 2, make the q
    block SPD, and the kernel eliminates in natural order; k = 2 and 4 at
    their defaults pivot."""
    m = hdg.rectangle_mesh((7, 6), jitter=0.2, seed=5)
    om = O.rectangle_mesh(7, 6, jitter=0.2, seed=5)
    P = load(ctx, hdg, O, m, om, k, qdeg, src=src)
    ctx.assemble_condense()
    Ate, bte = ctx.get_condensed()
    Ke, be = ctx.get_local_ops()
    oA, ob, oK, obe = O.condense(P, om)
    for c in range(m.cells.shape[0]):
        assert rel(Ate[c], oA[c]) < 1e-12
        assert rel(Ke[c], oK[c]) < 1e-12
        assert rel(bte[c], ob[c]) < 1e-12
        assert rel(be[c], obe[c]) < 1e-12


def test_source_values_at_qp(hdg, O, ctx):
    """HDG_SRC_VALUES (host-evaluated closures) reproduces the built-in source."""
    m = hdg.rectangle_mesh((5, 5), jitter=0.2, seed=2)
    om = O.rectangle_mesh(5, 5, jitter=0.2, seed=2)
    load(ctx, hdg, O, m, om, 2, 3)
    ctx.assemble_condense()
    _, b0 = ctx.get_condensed()
    pts, _ = hdg.quadrature(2, 3)
    x = m.nodes[m.cells[:, :3] - 1]
    lam = np.stack([1 - pts[:, 0] - pts[:, 1], pts[:, 0], pts[:, 1]], axis=1)
    xq = np.einsum("qv,cvd->cqd", lam, x)
    fq = 2 * np.pi ** 2 * np.sin(np.pi * xq[..., 0]) * np.sin(np.pi * xq[..., 1])
    ctx.set_source(kind=3, f_qp=fq)
    ctx.assemble_condense()
    _, b1 = ctx.get_condensed()
    assert rel(b1, b0) < 1e-13


@pytest.mark.parametrize("k,n", [(1, 4), (1, 10), (2, 8), (3, 6), (4, 5)])
def test_trace_system_pattern_bit_exact(hdg, O, ctx, k, n):
    """K in Julia's CSC form: colptr and rowval bit-exact against sparse(I,J,V)
    (assembler.jl:47-49); values and rhs to 1e-12."""
    m = hdg.rectangle_mesh((n, n + 1), jitter=0.2, seed=n)
    om = O.rectangle_mesh(n, n + 1, jitter=0.2, seed=n)
    P = load(ctx, hdg, O, m, om, k, QD[k])
    ctx.assemble_condense()
    colptr, rowval, nzval, rhs = ctx.get_system()
    oA, ob, _, _ = O.condense(P, om)
    K, orhs = O.assemble(om, k + 1, oA, ob)
    assert np.array_equal(colptr, K.colptr)
    assert np.array_equal(rowval, K.rowval)
    assert rel(nzval, K.nzval) < 1e-12
    assert rel(rhs, orhs) < 1e-12
    assert ctx.sizes.nnz == len(K.nzval)


def test_fixture_solve_and_l2_bound(hdg, O, ctx):
    """The reference's HDG test (test/test_FunctionSpace.jl:214-243): the
    squared L2 error on the fixture is <= 0.12 and equals the oracle's."""
    m = hdg.parse_mesh_triangle(os.path.join(GOLDEN, "figure2.1"))
    om = O.parse_mesh_triangle(os.path.join(GOLDEN, "figure2.1"))
    P = load(ctx, hdg, O, m, om, 1, 2)
    ctx.assemble_condense()
    uhat, info = ctx.solve()
    r = O.solve_pipeline(P, om)
    assert rel(uhat, r["uhat"]) < 1e-12
    q, u, uc = ctx.recover()
    assert rel(u, r["u"]) < 1e-12 and rel(q, r["q"]) < 1e-12
    assert np.array_equal(uc, r["uhat_cell"]) or rel(uc, r["uhat_cell"]) < 1e-12
    e = ctx.errornorm()
    assert e <= 0.12
    assert e == pytest.approx(O.errornorm(P, om, r["u"]), rel=1e-12)


@pytest.mark.parametrize("k,n,tol", [(1, 10, 1e-10), (1, 32, 1e-10), (2, 16, 1e-10), (3, 12, 1e-10)])
def test_solution_vs_direct_solve(hdg, O, ctx, k, n, tol):
    """PCG to ||r|| <= 1e-13 ||b|| against the oracle's direct banded LU
    (K \\ b, poisson2D_HDG.jl:195); cond(K) <= ~1e4 at these sizes bounds
    the difference by ~1e-9 relative, the stated tolerance is 1e-10."""
    m = hdg.rectangle_mesh((n, n))
    om = O.rectangle_mesh(n, n)
    P = load(ctx, hdg, O, m, om, k, QD[k])
    ctx.assemble_condense()
    uhat, info = ctx.solve()
    r = O.solve_pipeline(P, om)
    # the recursive residual reaches 1e-13; the recomputed one stops at the
    # FP64 rounding floor of ||K|| ||uhat|| eps, about 1e-12 here
    assert info.relres <= 1e-13 and info.true_relres < 1e-11
    assert rel(uhat, r["uhat"]) < tol
    q, u, _ = ctx.recover()
    assert rel(u, r["u"]) < tol
    e = ctx.errornorm()
    eo = O.errornorm(P, om, r["u"])
    assert e == pytest.approx(eo, rel=1e-8)
    if k == 1 and n == 10:
        assert e <= 0.00006  # examples/poisson2D_HDG.jl:218


def test_nonzero_dirichlet_and_patch(hdg, O, ctx):
    """Linear g on the boundary with f = 0: the HDG solution reproduces g
    exactly (u = g, q = -grad g) through the device solve."""
    k = 2
    m = hdg.rectangle_mesh((9, 7), jitter=0.2, seed=8)
    om = O.rectangle_mesh(9, 7, jitter=0.2, seed=8)
    ctx.set_mesh(m)
    ctx.setup(k, quad_degree=3, rtol=1e-14)
    nqp = O.sizes(k, 3)[2]
    ctx.set_source(kind=3, f_qp=np.zeros((m.cells.shape[0], nqp)))
    g = (0.4, -1.1, 2.5)
    d, v = O.dirichlet_trace(O.problem(k=k, qdeg=3), om, g=g)
    ctx.set_dirichlet(d, v)
    ctx.assemble_condense()
    uhat, info = ctx.solve()
    e = ctx.errornorm(kind=1, par=np.array(g))
    assert e < 1e-22
    q, u, _ = ctx.recover()
    np.testing.assert_allclose(q[:, 0], -g[1] / np.sqrt(2), atol=1e-10)
    np.testing.assert_allclose(q[:, 6], -g[2] / np.sqrt(2), atol=1e-10)


def test_device_recovery_of_host_trace(hdg, O, ctx):
    """get_usigma with a caller-supplied uhat (hdg_set_trace) on the device."""
    m = hdg.rectangle_mesh((6, 6), jitter=0.2, seed=4)
    om = O.rectangle_mesh(6, 6, jitter=0.2, seed=4)
    P = load(ctx, hdg, O, m, om, 2, 3)
    ctx.assemble_condense()
    r = O.solve_pipeline(P, om)
    fe = hdg.GenericFiniteElement(hdg.Dubiner(2))
    Wh, Vh = hdg.ScalarFunctionSpace(m, fe), hdg.VectorFunctionSpace(m, fe)
    Mh = hdg.ScalarTraceFunctionSpace(Wh, hdg.GenericFiniteElement(hdg.Legendre(2)))
    s, u, uh = hdg.TrialFunction(Vh), hdg.TrialFunction(Wh), hdg.TrialFunction(Mh)
    hdg.get_usigma(s, u, uh, r["uhat"], None, None, m, ctx=ctx)
    assert rel(u.m_values, r["u"]) < 1e-12
    assert rel(s.m_values, r["q"]) < 1e-12


def test_reference_api_dropin(hdg, O):
    """The reference driver sequence (poisson2D_HDG.jl:189-218) through the
    Python mirror: doassemble, apply!, K \\ b, get_uσ!, errornorm."""
    import scipy.sparse.linalg as spla
    m = hdg.rectangle_mesh((10, 10))
    fe = hdg.GenericFiniteElement(hdg.Dubiner(1))
    Wh, Vh = hdg.ScalarFunctionSpace(m, fe), hdg.VectorFunctionSpace(m, fe)
    Mh = hdg.ScalarTraceFunctionSpace(Wh, hdg.GenericFiniteElement(hdg.Legendre(1)))
    uh_hat, s_h, u_h = hdg.TrialFunction(Mh), hdg.TrialFunction(Vh), hdg.TrialFunction(Wh)
    dbc = hdg.Dirichlet(uh_hat, m, "boundary", lambda x: 0.0)
    ctx = hdg.Context(0)
    K, b, K_e, b_e = hdg.doassemble(Vh, Wh, Mh, ctx=ctx)
    hdg.apply_bc(K, b, dbc)
    uhat = spla.spsolve(K.tocsc(), b)
    hdg.get_usigma(s_h, u_h, uh_hat, uhat, K_e, b_e, m, ctx=ctx)
    e = hdg.errornorm(u_h, "sinsin")
    assert e <= 0.00006
    assert e == pytest.approx(5.3645466464e-05, rel=1e-8)
    # the fused path gives the same answer
    info = hdg.hdg_solve(s_h, u_h, uh_hat, m, dbc, ctx=ctx)
    assert hdg.errornorm(u_h, "sinsin") == pytest.approx(e, rel=1e-9)
    ctx.close()


def test_deterministic(hdg, O, ctx):
    """Atomic-free assembly and fixed-order reductions: bitwise repeatable."""
    m = hdg.rectangle_mesh((40, 40), jitter=0.2, seed=1)
    ctx.set_mesh(m)
    ctx.setup(2, quad_degree=3)
    ctx.set_source()
    ctx.set_dirichlet(m.boundary_dofs(3), np.zeros(len(m.boundary_dofs(3))))
    outs = []
    for _ in range(2):
        info = ctx.run()
        uh, _ = ctx.solve()
        outs.append((uh.copy(), info.iters))
    assert outs[0][1] == outs[1][1]
    assert np.array_equal(outs[0][0].view(np.uint64), outs[1][0].view(np.uint64))


# ------------------------------------------------------------ error paths
def test_error_singular_local_matrix(hdg, ctx):
    """k=3 with the reference default quad_degree=4 makes the local (q,u)
    matrix singular; the reference silently returns garbage, libhdg raises."""
    m = hdg.rectangle_mesh((3, 3))
    ctx.set_mesh(m)
    ctx.setup(3)  # quad_degree 0 -> k+1 = 4
    ctx.set_source()
    with pytest.raises(hdg.SingularException):
        ctx.assemble_condense()


def test_error_negative_jacobian(hdg, ctx):
    m = hdg.rectangle_mesh((2, 2))
    cells = m.cells.copy()
    cells[0, [1, 2]] = cells[0, [2, 1]]  # clockwise cell
    bad = hdg.PolygonalMesh(m.nodes, cells, m.faces, m.facesets, m.nodesets)
    ctx.set_mesh(bad)
    ctx.setup(1)
    ctx.set_source()
    with pytest.raises(hdg.ArgumentError, match="NEG_JACOBIAN"):
        ctx.assemble_condense()


def test_error_quadrature_and_order(hdg, ctx):
    m = hdg.rectangle_mesh((2, 2))
    ctx.set_mesh(m)
    with pytest.raises(hdg.ArgumentError, match="QUAD"):
        ctx.setup(1, quad_degree=8)
    with pytest.raises(hdg.HDGError, match="UNSUPPORTED"):
        ctx.setup(5)


def test_error_not_converged_and_budget(hdg, O):
    m = hdg.rectangle_mesh((16, 16))
    with hdg.Context(0) as c:
        c.set_mesh(m)
        c.setup(2, quad_degree=3, maxit=3)
        c.set_source()
        c.set_dirichlet(m.boundary_dofs(3), np.zeros(len(m.boundary_dofs(3))))
        c.assemble_condense()
        with pytest.raises(hdg.HDGError, match="NOT_CONVERGED"):
            c.solve()
    with hdg.Context(0, hbm_budget=1 << 16) as c:
        with pytest.raises(hdg.HDGError, match="HBM_BUDGET"):
            c.set_mesh(hdg.rectangle_mesh((64, 64)))


def test_full_size_c2_properties(hdg, O, ctx):
    """512x512 cells (524,288 triangles), P2, jittered: the configuration the
    benchmark runs.  Size-independent checks: trace solve residual, the
    squared L2 error against the manufactured solution at the rate k+1
    relative to 256x256, and nnz = (5 n_int + 3 n_bdry) nt^2."""
    errs = []
    for n in (256, 512):
        m = hdg.rectangle_mesh((n, n), jitter=0.2, seed=20261202)
        ctx.set_mesh(m)
        ctx.setup(2, quad_degree=3, rtol=1e-12)
        ctx.set_source()
        ctx.set_dirichlet(m.boundary_dofs(3), np.zeros(len(m.boundary_dofs(3))))
        info = ctx.run()
        # recursive residual at 1e-12; the recomputed FP64 residual stops at
        # the attainable floor eps ||K|| ||uhat|| / ||b|| (cond ~ h^-2): the
        # oracle's CPU PCG reaches 1.46e-9 at 256^2, libhdg 1.47e-9
        assert info.relres <= 1e-12
        assert info.true_relres < (5e-9 if n == 256 else 2e-8), info.true_relres
        errs.append(ctx.errornorm())
        nb = 4 * n
        assert ctx.sizes.nnz == (5 * (m.faces.shape[0] - nb) + 3 * nb) * 9
    # squared L2 error: halving h divides it by 2^(2(k+1)) = 64
    assert 40.0 < errs[0] / errs[1] < 90.0, errs


def test_full_size_p3_mixed_properties(hdg, ctx):
    """1024x1024 cells (2,097,152 triangles), P3 at quadrature degree 6,
    jittered: the mixed-precision configuration of the benchmark, 12.6M trace
    dofs.  Size-independent checks: FP64 and HDG_MIXED solves reach the
    recursive tolerance, their recomputed FP64 residuals stay at the
    attainable floor, the two solutions agree within 1e-10 relative (the
    mixed form's target against FP64), the L2 error against the manufactured
    solution sits at the FP64 rounding floor, and nnz = (5 n_int + 3 n_bdry) nt^2."""
    n, k = 1024, 3
    m = hdg.rectangle_mesh((n, n), jitter=0.2, seed=20261202)
    ctx.set_mesh(m)
    d = m.boundary_dofs(k + 1)
    sols = []
    for prec in (hdg.HDG_FP64, hdg.HDG_MIXED):
        ctx.setup(k, quad_degree=6, rtol=1e-12, precision=prec, precond=hdg.PC_MULTIGRID)
        ctx.set_source()
        ctx.set_dirichlet(d, np.zeros(len(d)))
        info = ctx.run()
        print(f"precision {prec}: iters {info.iters} relres {info.relres:.3g} true {info.true_relres:.3g}", flush=True)
        assert info.relres <= 1e-12
        assert info.true_relres < 5e-9, info.true_relres
        assert 0.0 < ctx.errornorm() < 1e-20  # L2 error < 1e-10: the rounding floor at this size
        sols.append(ctx.trace())
    nb = 4 * n
    assert ctx.sizes.nnz == (5 * (m.faces.shape[0] - nb) + 3 * nb) * (k + 1) ** 2
    assert rel(sols[1], sols[0]) < 1e-10


# ------------------------------------------------------- multigrid path
@pytest.mark.parametrize("k,n", [(1, 32), (2, 32), (3, 16), (4, 16), (2, 64)])
def test_multigrid_matches_block_jacobi_and_direct(hdg, O, ctx, k, n):
    """The multigrid-preconditioned CG (structured meshes) converges in a
    mesh-independent number of iterations and gives the same trace solution
    as the block-Jacobi PCG and the oracle's direct solve."""
    m = hdg.rectangle_mesh((n, n), jitter=0.2, seed=3)
    om = O.rectangle_mesh(n, n, jitter=0.2, seed=3)
    sols = {}
    for pc in (hdg.PC_BLOCK_JACOBI, hdg.PC_MULTIGRID):
        ctx.set_mesh(m)
        ctx.setup(k, quad_degree=QD[k], rtol=1e-13, precond=pc)
        ctx.set_source()
        P = O.problem(k=k, qdeg=QD[k])
        d, v = O.dirichlet_trace(P, om)
        ctx.set_dirichlet(d, v)
        ctx.assemble_condense()
        sols[pc] = ctx.solve()
    (ub, ib), (um, im) = sols[hdg.PC_BLOCK_JACOBI], sols[hdg.PC_MULTIGRID]
    assert im.iters < 60 < ib.iters or n < 64
    assert im.relres <= 1e-13
    r = O.solve_pipeline(O.problem(k=k, qdeg=QD[k]), om) if n <= 32 else None
    if r is not None:
        assert rel(um, r["uhat"]) < 1e-10
    assert rel(um, ub) < 1e-9


@pytest.mark.parametrize("k,nx,ny,itmax", [(2, 24, 40, 60), (1, 33, 17, 60), (3, 48, 32, 60), (2, 65, 63, 60),
                                           (1, 75, 75, 60), (2, 129, 77, 60), (3, 16, 100, 120)])
def test_multigrid_rectangular_and_odd_grids(hdg, ctx, k, nx, ny, itmax):
    """Rectangular grids and odd sides (coarsened with a one-cell-wide last
    column/row and 9-point Galerkin stencils) converge to the block-Jacobi
    PCG's solution in a mesh-independent number of iterations; cells of
    aspect ratio 6 (16x100) take more (point smoothing)."""
    m = hdg.rectangle_mesh((nx, ny), jitter=0.2, seed=11)
    sols = {}
    for pc in (hdg.PC_BLOCK_JACOBI, hdg.PC_MULTIGRID):
        ctx.set_mesh(m)
        ctx.setup(k, quad_degree=QD[k], rtol=1e-13, precond=pc)
        ctx.set_source()
        ctx.set_dirichlet(m.boundary_dofs(k + 1), np.zeros(len(m.boundary_dofs(k + 1))))
        ctx.assemble_condense()
        sols[pc] = ctx.solve()
    (ub, ib), (um, im) = sols[hdg.PC_BLOCK_JACOBI], sols[hdg.PC_MULTIGRID]
    assert im.relres <= 1e-13 and ib.relres <= 1e-13
    assert im.iters < itmax and im.iters < ib.iters
    assert rel(um, ub) < 1e-9


def test_multigrid_iterations_mesh_independent(hdg, ctx):
    its = []
    for n in (64, 128, 256):
        m = hdg.rectangle_mesh((n, n), jitter=0.2, seed=20261202)
        ctx.set_mesh(m)
        ctx.setup(2, quad_degree=3, rtol=1e-12, precond=hdg.PC_MULTIGRID)
        ctx.set_source()
        ctx.set_dirichlet(m.boundary_dofs(3), np.zeros(len(m.boundary_dofs(3))))
        info = ctx.run()
        its.append(info.iters)
        assert info.true_relres < 1e-9
    assert max(its) < 50, its


def test_multigrid_rejects_unstructured(hdg, ctx):
    m = hdg.parse_mesh_triangle(os.path.join(GOLDEN, "figure.1"))
    ctx.set_mesh(m)
    with pytest.raises(hdg.HDGError, match="UNSUPPORTED"):
        ctx.setup(1, precond=hdg.PC_MULTIGRID)
    ctx.setup(1)  # auto falls back to block-Jacobi
    ctx.set_source()
    ctx.set_dirichlet(m.boundary_dofs(2), np.zeros(len(m.boundary_dofs(2))))
    info = ctx.run()
    assert info.relres <= 1e-12


# ------------------------------------------------------- mixed precision
@pytest.mark.parametrize("k,n", [(1, 32), (2, 32), (3, 16), (2, 128), (3, 64)])
def test_mixed_precision_matches_fp64(hdg, O, ctx, k, n):
    """HDG_MIXED (FP64 CG on the FP64 trace matrix, multigrid preconditioner
    on float copies of the trace matrix, Jacobi blocks and S P0) reaches
    the FP64 solve's accuracy: the same stopping test on the FP64 residual
    recursion, a recomputed residual at the same FP64 floor, the FP64
    solution to 1e-10 relative and the oracle's direct solve on small
    meshes."""
    m = hdg.rectangle_mesh((n, n), jitter=0.2, seed=5)
    sols = {}
    for prec in (hdg.HDG_FP64, hdg.HDG_MIXED):
        ctx.set_mesh(m)
        ctx.setup(k, quad_degree=QD[k], rtol=1e-12, precision=prec, precond=hdg.PC_MULTIGRID)
        ctx.set_source()
        ctx.set_dirichlet(m.boundary_dofs(k + 1), np.zeros(len(m.boundary_dofs(k + 1))))
        ctx.assemble_condense()
        sols[prec] = ctx.solve()
    (u64, i64), (u32, i32) = sols[hdg.HDG_FP64], sols[hdg.HDG_MIXED]
    assert i32.relres <= 1e-12
    assert i32.true_relres < 4 * max(i64.true_relres, 1e-12)
    assert i32.iters <= i64.iters + 3
    assert rel(u32, u64) < 1e-10
    if n <= 32:
        om = O.rectangle_mesh(n, n, jitter=0.2, seed=5)
        r = O.solve_pipeline(O.problem(k=k, qdeg=QD[k]), om)
        assert rel(u32, r["uhat"]) < 1e-10


def test_mixed_precision_needs_multigrid_path(hdg, ctx):
    m = hdg.parse_mesh_triangle(os.path.join(GOLDEN, "figure.1"))
    ctx.set_mesh(m)
    with pytest.raises(hdg.HDGError, match="UNSUPPORTED"):
        ctx.setup(1, precision=hdg.HDG_MIXED)
    m = hdg.rectangle_mesh((8, 8))
    ctx.set_mesh(m)
    with pytest.raises(hdg.HDGError, match="UNSUPPORTED"):
        ctx.setup(1, precision=hdg.HDG_MIXED, precond=hdg.PC_BLOCK_JACOBI)


# ------------------------------------------------------- HIP graph replay
@pytest.mark.parametrize("k,n,precision,precond", [(2, 64, 0, 2), (3, 33, 1, 0), (1, 24, 0, 1)])
def test_graph_replay_equals_eager(hdg, ctx, k, n, precision, precond):
    """The captured chunks of CG iterations, replayed over several solves with
    other eager work and synchronisations in between, give the eager
    launches' result bit for bit every time."""
    m = hdg.rectangle_mesh((n, n), jitter=0.2, seed=21)
    other = hdg.Context(0)
    om = hdg.rectangle_mesh((17, 19), jitter=0.1, seed=2)
    other.set_mesh(om)
    other.setup(2, quad_degree=3, schedule=hdg.SCHEDULE_EAGER)
    other.set_source()
    other.set_dirichlet(om.boundary_dofs(3), np.zeros(len(om.boundary_dofs(3))))
    out = {}
    eager, graph = hdg.SCHEDULE_EAGER, 0
    for sched in (eager, graph):
        ctx.set_mesh(m)
        ctx.setup(k, quad_degree=QD[k], rtol=1e-12, precision=precision, precond=precond, schedule=sched)
        ctx.set_source()
        ctx.set_dirichlet(m.boundary_dofs(k + 1), np.zeros(len(m.boundary_dofs(k + 1))))
        sols = []
        for rep in range(4):  # the first solve after a set-up is eager, later ones replay
            info = ctx.run()
            sols.append((ctx.trace(), info.iters))
            other.run()      # eager work of another context between replays
        out[sched] = sols
    eager_u, eager_it = out[eager][0]
    for u, it in out[eager] + out[graph]:
        assert it == eager_it
        assert np.array_equal(u, eager_u)
    other.close()


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_condense_schedules_same_bits(hdg, ctx, k):
    """Every condensation schedule (lanes per element, grid cap) gives the
    same condensed operators and recovery operators bit for bit, and the
    autotuned solve equals the default one."""
    m = hdg.rectangle_mesh((37, 29), jitter=0.2, seed=31)
    ctx.set_mesh(m)
    ctx.setup(k, quad_degree=QD[k], rtol=1e-12)
    ctx.set_source(kind=hdg._lib.SRC_KAPPA, kappa=(2.5, 6.0))
    ctx.set_dirichlet(m.boundary_dofs(k + 1), np.zeros(len(m.boundary_dofs(k + 1))))
    ref = None
    nat = {1: 16, 2: 32, 3: 64, 4: 64}[k]
    seg = nat
    while seg <= 64:
        for grid in (512, 8192):
            ctx.set_schedule(seg, grid)
            ctx.assemble_condense()
            got = ctx.get_condensed() + ctx.get_local_ops()
            if ref is None:
                ref = got
            else:
                for a, b in zip(ref, got):
                    assert np.array_equal(a, b), (seg, grid)
        seg *= 2
    with pytest.raises(hdg.HDGError):
        ctx.set_schedule(nat // 2 if nat > 16 else 8, 0)
    ctx.set_schedule(0, 0)
    ctx.run()
    u_default = ctx.trace()
    sc = ctx.autotune()
    print(f"k={k}: tuned {sc}", flush=True)
    assert sc["condense_form"] == 0
    assert sc["condense_seg"] in (16, 32, 64) and sc["condense_seg"] >= nat
    assert 0 < sc["condense_us"] <= sc["default_us"] * 1.0001
    ctx.run()
    assert np.array_equal(ctx.trace(), u_default)


@pytest.mark.parametrize("nx,ny,k,box", [(37, 29, 2, ((0.0, 0.0), (1.0, 1.0))), (16, 16, 1, ((0.0, 0.0), (1.0, 1.0))),
                                         (64, 33, 3, ((-1.0, 0.5), (2.0, 3.25))), (1, 1, 1, ((0.0, 0.0), (1.0, 1.0))),
                                         (5, 100, 4, ((0.0, 0.0), (1.0, 1.0)))])
def test_device_rectangle_mesh_same_bits(hdg, ctx, nx, ny, k, box):
    """The mesh generated on the device (hdg_set_rectangle_mesh) with u = 0
    on the boundary (hdg_set_dirichlet_boundary) gives the trace system,
    solution, recovery and error norm of the host-generated mesh with
    host-listed Dirichlet dofs, bit for bit."""
    m = hdg.rectangle_mesh((nx, ny), LL=box[0], UR=box[1], jitter=0.2, seed=77)
    out = []
    for dev in (False, True):
        if dev:
            ctx.set_rectangle_mesh(nx, ny, ll=box[0], ur=box[1], jitter=0.2, seed=77)
        else:
            ctx.set_mesh(m)
        ctx.setup(k, quad_degree=QD[k], rtol=1e-12)
        ctx.set_source(kind=hdg._lib.SRC_KAPPA, kappa=(2.5, 6.0))
        if dev:
            ctx.set_dirichlet_boundary(0.0)
        else:
            d = m.boundary_dofs(k + 1)
            ctx.set_dirichlet(d, np.zeros(len(d)))
        info = ctx.run()
        colptr, rowval, nzval, rhs = ctx.get_system()
        q, u, uc = ctx.recover()
        out.append((ctx.trace(), info.iters, ctx.errornorm(), colptr, rowval, nzval, rhs, q, u, uc))
    for a, b in zip(*out):
        assert np.array_equal(np.asarray(a), np.asarray(b))


def test_device_rectangle_mesh_boundary_error(hdg, ctx):
    """A box whose interpolated edge nodes miss its sides fails on the
    device as on the host (HDG_ERR_BOUNDARY, generate_mesh.jl:80)."""
    with pytest.raises(hdg.ArgumentError, match="BOUNDARY"):
        hdg.rectangle_mesh((64, 33), LL=(-2.0, -1.0), UR=(3.0, 1.5), jitter=0.2, seed=77)
    with pytest.raises(hdg.ArgumentError, match="BOUNDARY"):
        ctx.set_rectangle_mesh(64, 33, ll=(-2.0, -1.0), ur=(3.0, 1.5), jitter=0.2, seed=77)
