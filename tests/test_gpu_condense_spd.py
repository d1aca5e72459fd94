"""GPU parity of the Schur-complement condensation (kappa = 1 and an SPD
reference mass: A = detJ diag(Mref, Mref) is eliminated analytically and
S = C + B'A^-1 B is factored by Cholesky) against the oracle's LU of the full
local matrix (poisson2D_HDG.jl:155-174) and against the library's own
elimination (HDG_SCHEDULE_LU), to 1e-12 relative per cell; kappa sources and
rules whose mass is singular keep the elimination, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def per_cell_rel(a, b):
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    return float((np.abs(a - b).max(axis=1) / np.maximum(np.abs(b).max(axis=1), 1e-300)).max())


def _ops(hdg, ctx, m, k, qdeg, tau, src, schedule):
    ctx.set_mesh(m)
    ctx.setup(k, quad_degree=qdeg, tau=tau, rtol=1e-13, schedule=schedule)
    ctx.set_source(**src)
    d = m.boundary_dofs(k + 1)
    ctx.set_dirichlet(d, np.zeros(len(d)))
    ctx.assemble_condense()
    Ate, bte = ctx.get_condensed()
    Ke, be = ctx.get_local_ops()
    colptr, rowval, nzval, rhs = ctx.get_system()
    return Ate, bte, Ke, be, nzval, rhs


@pytest.mark.parametrize("k,qdeg,tau,kind", [(1, 2, 1.0, 0), (3, 6, 1.0, 0), (3, 6, 2.5, 2), (1, 4, 1.0, 2),
                                             (2, 4, 1.0, 0), (1, 2, 0.7, 0)])
def test_spd_condensation_vs_oracle_and_lu(hdg, O, k, qdeg, tau, kind):
    print(f"[spd] k={k} qdeg={qdeg} tau={tau} source {kind}", flush=True)
    coef = np.linspace(-1.0, 1.3, 16)
    m = hdg.rectangle_mesh((13, 11), jitter=0.2, seed=9)
    om = O.rectangle_mesh(13, 11, jitter=0.2, seed=9)
    src = dict(kind=kind, coef=coef if kind == 2 else None)
    P = O.problem(k=k, qdeg=qdeg, tau=tau, src_kind=kind, coef=coef if kind == 2 else None)
    oA, ob, oK, obe = O.condense(P, om)
    with hdg.Context(0) as ctx:
        fast = _ops(hdg, ctx, m, k, qdeg, tau, src, 0)
        lu = _ops(hdg, ctx, m, k, qdeg, tau, src, hdg.SCHEDULE_LU)
    for x, y, z in zip(fast[:4], lu[:4], (oA, ob, oK, obe)):
        assert per_cell_rel(x, z) < 1e-12
        assert per_cell_rel(x, y) < 1e-12
    assert not all(np.array_equal(x, y) for x, y in zip(fast[:4], lu[:4]))  # two different computations
    nz = np.abs(fast[4] - lu[4]).max() / np.abs(lu[4]).max()
    assert nz < 1e-12


def test_spd_solution_and_kappa_fallback(hdg, O):
    """The solve on the Schur-complement operators equals the oracle's
    direct solve to 1e-12 (32 x 32 cells, P3); with kappa the elimination
    runs whatever the schedule says (same bits with and without
    HDG_SCHEDULE_LU)."""
    n, k = 32, 3
    m = hdg.rectangle_mesh((n, n), jitter=0.2, seed=4)
    om = O.rectangle_mesh(n, n, jitter=0.2, seed=4)
    P = O.problem(k=k, qdeg=6)
    r = O.solve_pipeline(P, om)
    with hdg.Context(0) as ctx:
        ctx.set_mesh(m)
        ctx.setup(k, quad_degree=6, rtol=1e-15, maxit=20000)
        ctx.set_source()
        ctx.set_dirichlet(r["dofs"], r["vals"])
        ctx.run()
        uhat = ctx.trace()
        q, u, _ = ctx.recover()
        rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
        assert rel(uhat, r["uhat"]) < 1e-12 and rel(u, r["u"]) < 1e-12 and rel(q, r["q"]) < 1e-12
        kap = dict(kind=hdg._lib.SRC_KAPPA, kappa=(3.0, 5.0))
        a = _ops(hdg, ctx, m, k, 6, 1.0, kap, 0)
        b = _ops(hdg, ctx, m, k, 6, 1.0, kap, hdg.SCHEDULE_LU)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("k,qdeg", [(1, 2), (2, 4), (3, 6)])
def test_spd_schedules_same_bits(hdg, k, qdeg):
    """Both register budgets and every grid cap of the Schur-complement
    kernel give the same operators bit for bit; hdg_autotune reports the
    Schur form, and its choice set back through hdg_set_schedule gives the
    same solution as the default."""
    print(f"[spd schedules] k={k} qdeg={qdeg}", flush=True)
    m = hdg.rectangle_mesh((37, 29), jitter=0.2, seed=31)
    with hdg.Context(0) as ctx:
        ctx.set_mesh(m)
        ctx.setup(k, quad_degree=qdeg, rtol=1e-12)
        ctx.set_source()
        d = m.boundary_dofs(k + 1)
        ctx.set_dirichlet(d, np.zeros(len(d)))
        ref = None
        for waves in (2, 3):
            for grid in (7, 512, 16384):
                ctx.set_schedule(0, grid, waves)
                ctx.assemble_condense()
                got = ctx.get_condensed() + ctx.get_local_ops()
                if ref is None:
                    ref = got
                else:
                    for a, b in zip(ref, got):
                        assert np.array_equal(a, b), (waves, grid)
        with pytest.raises(hdg.HDGError):
            ctx.set_schedule(0, 0, 1)
        ctx.set_schedule(0, 0, 0)
        ctx.run()
        u_default = ctx.trace()
        sc = ctx.autotune()
        print(f"k={k}: tuned {sc}", flush=True)
        assert sc["condense_form"] == 1 and sc["condense_seg"] == 16 and sc["condense_waves"] in (2, 3)
        assert 0 < sc["condense_us"] <= sc["default_us"] * 1.0001
        ctx.set_schedule(**sc)
        ctx.run()
        assert np.array_equal(ctx.trace(), u_default)


@pytest.mark.parametrize("k,qdeg", [(1, 2), (1, 4), (2, 4), (3, 6)])
def test_spd_recovery_vs_oracle(hdg, O, k, qdeg):
    """Recovery on the Schur form (the kernel keeps only the u rows of K_e
    and forms q from u and the geometry) for a trace given from outside
    (hdg_set_trace), against the oracle's recovery with its own K_e and b_e
    (poisson2D_HDG.jl:197-212) to 1e-12, and against the exported K_e
    applied on the host."""
    print(f"[spd recovery] k={k} qdeg={qdeg}", flush=True)
    m = hdg.rectangle_mesh((11, 9), jitter=0.2, seed=21)
    om = O.rectangle_mesh(11, 9, jitter=0.2, seed=21)
    P = O.problem(k=k, qdeg=qdeg)
    _, _, oK, obe = O.condense(P, om)
    rng = np.random.default_rng(k * 10 + qdeg)
    nt = k + 1
    uhat = rng.standard_normal(om.nfaces * nt)
    oq, ou, _ = O.recover(om, k, oK, obe, uhat)
    with hdg.Context(0) as ctx:
        ctx.set_mesh(m)
        ctx.setup(k, quad_degree=qdeg)
        ctx.set_source()
        d = m.boundary_dofs(nt)
        ctx.set_dirichlet(d, np.zeros(len(d)))
        ctx.assemble_condense()
        ctx.set_trace(uhat)
        q, u, _ = ctx.recover()
        Ke, be = ctx.get_local_ops()
        # the export rewrote K_e in its full layout: recovery reads that now
        ctx.set_trace(uhat)
        q2, u2, _ = ctx.recover()
    rel = lambda a, b: float(np.abs(a - b).max() / np.abs(b).max())
    assert rel(u, ou) < 1e-12 and rel(q, oq) < 1e-12
    assert rel(u2, u) < 1e-13 and rel(q2, q) < 1e-13
    # the exported K_e (its q rows filled in on demand) applied on the host
    faces = np.asarray(m.cells)[:, 3:] - 1          # local faces 0..2 of every cell
    uc = uhat.reshape(-1, nt)[faces].reshape(len(faces), -1)
    qu = np.einsum("cmt,ct->cm", Ke, uc) + be
    nq = q.shape[1]
    assert rel(qu[:, :nq], q) < 1e-12 and rel(qu[:, nq:], u) < 1e-12
