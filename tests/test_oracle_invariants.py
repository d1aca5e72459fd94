"""Pins the oracle where the reference has no literal: closed forms,
invariants and textbook reductions chosen so that a dropped term, a wrong
sign, a wrong orientation flip or a transposed block fails one of them."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

QD = {1: 2, 2: 3, 3: 6, 4: 6}


def linear_trace(O, mesh, k, g):
    """L2 projection of g = g0 + g1 x + g2 y onto the orthonormal Legendre
    trace basis of every face, parametrised from the lower to the higher
    global node id (the face_orientation convention, mesh.jl:51-54)."""
    nt = k + 1
    x, w = O.quad1d(k + 1)
    uhat = np.zeros(mesh.nfaces * nt)
    for f in range(mesh.nfaces):
        a, b = sorted(mesh.faces[f, :2])
        pa, pb = mesh.xy[a - 1], mesh.xy[b - 1]
        for i in range(nt):
            s = 0.0
            for t, wt in zip(x, w):
                p = pa + t * (pb - pa)
                s += wt * (g[0] + g[1] * p[0] + g[2] * p[1]) * O.legendre_value(i + 1, t)
            uhat[f * nt + i] = s
    return uhat


@pytest.mark.parametrize("k,qdeg", [(1, 2), (3, 6)])
def test_mass_is_identity_under_exact_quadrature(O, k, qdeg):
    """Orthonormal Dubiner + affine map + exact rule => A = detJ I_nq and
    H = |e| I per face (orthonormal Legendre, exact Gauss rule)."""
    m = O.rectangle_mesh(3, 2, jitter=0.2, seed=5)
    P = O.problem(k=k, qdeg=qdeg)
    nt = k + 1
    for c in (1, 4, 9):
        b = O.element_blocks(P, m, c)
        detJ, _, _, detJf = O.geometry(m, c)
        np.testing.assert_allclose(b["A"], detJ * np.eye(b["A"].shape[0]), atol=1e-14)
        np.testing.assert_allclose(b["H"], np.kron(np.diag(detJf), np.eye(nt)), atol=1e-14)


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_constant_and_linear_patch(O, k):
    """With f = 0 the HDG local solver reproduces linear functions exactly:
    u_e = K_e uhat_e recovers g and q = -grad g, for every k, orientation
    and jittered geometry.  Pins K_e and therefore E, F, B, C and the
    orientation flip (a wrong flip breaks it on the oriented faces)."""
    m = O.rectangle_mesh(4, 3, jitter=0.2, seed=7)
    nu = (k + 1) * (k + 2) // 2
    nqp = O.sizes(k, QD[k])[2]
    P = O.problem(k=k, qdeg=QD[k], src_kind=3, f_qp=np.zeros((m.ncells, nqp)))
    Ate, bte, Ke, be = O.condense(P, m)
    np.testing.assert_allclose(be, 0.0, atol=0)
    for g in ([1.3, 0.0, 0.0], [0.4, -1.1, 2.5]):
        uhat = linear_trace(O, m, k, g)
        q, u, _ = O.recover(m, k, Ke, be, uhat)
        assert O.errornorm(P, m, u, uex=g) < 1e-26
        # q = -grad g: coefficient -g1/sqrt2 (x) and -g2/sqrt2 (y) on phi_1
        expect = np.zeros(2 * nu)
        expect[0], expect[nu] = -g[1] / np.sqrt(2), -g[2] / np.sqrt(2)
        np.testing.assert_allclose(q, np.broadcast_to(expect, q.shape), atol=1e-11)
        # Ate ue = <q.n + tau(u - uhat), mu> = -(grad g . n_f)|e_f| on mode 1
        # of each face (orthonormal Legendre: int mu_i = delta_i1), 0 else;
        # for a constant g this is the kernel statement Ate ue = 0.
        nt = k + 1
        for c in range(m.ncells):
            ue = np.concatenate([uhat[(f - 1) * nt:f * nt] for f in m.cells[c, 3:]])
            _, _, normals, detJf = O.geometry(m, c + 1)
            expect = np.zeros(3 * nt)
            for lf in range(3):
                expect[lf * nt] = -(g[1] * normals[lf, 0] + g[2] * normals[lf, 1]) * detJf[lf]
            np.testing.assert_allclose(Ate[c] @ ue, expect, atol=1e-12 * np.abs(Ate[c]).max())


@pytest.mark.parametrize("k,tau", [(1, 0.7), (2, 2.5), (3, 2.5), (3, 1.0)])
def test_constant_trace_with_tau(O, k, tau):
    """A constant trace c is reproduced by u = c, q = 0 for any tau, and
    since H carries no tau (poisson2D_HDG.jl:149) the condensed operator maps
    it to Ate c = (tau - 1) H c = (tau - 1) |e_f| c on mode 0 of each face."""
    m = O.rectangle_mesh(4, 3, jitter=0.2, seed=7)
    nt = k + 1
    P = O.problem(k=k, qdeg=QD[k], tau=tau)
    Ate, _, Ke, _ = O.condense(P, m)
    c = np.zeros(3 * nt)
    c[0::nt] = 1.0
    for e in range(m.ncells):
        _, _, _, detJf = O.geometry(m, e + 1)
        expect = np.zeros(3 * nt)
        expect[0::nt] = (tau - 1.0) * np.asarray(detJf)
        np.testing.assert_allclose(Ate[e] @ c, expect, atol=1e-12 * np.abs(Ate[e]).max())
        # K_e c: q = 0, u = 1 (coefficient 1 / sqrt 2 on phi_1 = sqrt 2)
        nu = (k + 1) * (k + 2) // 2
        want = np.zeros(3 * nu)
        want[2 * nu] = 1.0 / np.sqrt(2.0)
        np.testing.assert_allclose(Ke[e] @ c, want, atol=1e-11 * np.abs(Ke[e]).max())


@pytest.mark.parametrize("k", [1, 2, 3])
def test_trace_operator_symmetric_negative_definite(O, k):
    m = O.rectangle_mesh(5, 4, jitter=0.2, seed=3)
    P = O.problem(k=k, qdeg=QD[k])
    Ate, bte, Ke, be = O.condense(P, m)
    for c in range(0, m.ncells, 7):
        A = Ate[c]
        assert np.abs(A - A.T).max() <= 1e-13 * np.abs(A).max()
        assert np.linalg.eigvalsh(0.5 * (A + A.T)).max() < 1e-12 * np.abs(A).max()
    K, rhs = O.assemble(m, k + 1, Ate, bte)
    Kd = K.to_scipy().toarray()
    dofs, _ = O.dirichlet_trace(P, m)
    inner = np.setdiff1d(np.arange(K.n), dofs - 1)
    ev = np.linalg.eigvalsh(0.5 * (Kd + Kd.T)[np.ix_(inner, inner)])
    assert ev.max() < 0


@pytest.mark.parametrize("k", [1, 2])
def test_condensation_matches_dense_solve(O, k):
    """K_e = Me \\ EFe and Ate = Ge K_e - H against numpy's LAPACK solve."""
    m = O.rectangle_mesh(2, 2, jitter=0.2, seed=11)
    P = O.problem(k=k, qdeg=QD[k])
    Ate, bte, Ke, be = O.condense(P, m)
    for c in range(m.ncells):
        b = O.element_blocks(P, m, c + 1)
        Me = np.block([[b["A"], -b["B"]], [b["B"].T, b["C"]]])
        EF = np.vstack([-b["E"], b["F"]])
        G = np.vstack([b["E"], b["F"]]).T
        Kr = np.linalg.solve(Me, EF)
        br = np.linalg.solve(Me, np.concatenate([np.zeros(b["A"].shape[0]), b["be"]]))
        np.testing.assert_allclose(Ke[c], Kr, atol=1e-12 * np.abs(Kr).max())
        np.testing.assert_allclose(Ate[c], G @ Kr - b["H"], atol=1e-12 * np.abs(Ate[c]).max())
        np.testing.assert_allclose(be[c], br, atol=1e-12 * max(np.abs(br).max(), 1e-300))
        np.testing.assert_allclose(bte[c], -G @ br, atol=1e-12 * max(np.abs(bte[c]).max(), 1e-300))


@pytest.mark.parametrize("k", [1, 2, 3])
def test_assembly_pattern_and_sparse_semantics(O, k):
    """sparse(I,J,V) (assembler.jl:47-49): CSC with sorted rows, duplicates
    summed, nnz = (5 n_int + 3 n_bdry) nt^2 and values equal to a dense
    scatter-add of the element blocks."""
    m = O.rectangle_mesh(4, 3)
    nt = k + 1
    P = O.problem(k=k, qdeg=QD[k])
    Ate, bte, Ke, be = O.condense(P, m)
    K, rhs = O.assemble(m, nt, Ate, bte)
    nb = int((m.faces[:, 3] == 0).sum())
    ni = m.nfaces - nb
    assert K.colptr[-1] - 1 == (5 * ni + 3 * nb) * nt * nt
    for c in range(K.n):
        rows = K.rowval[K.colptr[c] - 1:K.colptr[c + 1] - 1]
        assert np.all(np.diff(rows) > 0)
    D = np.zeros((K.n, K.n))
    r = np.zeros(K.n)
    for c in range(m.ncells):
        g = np.concatenate([(f - 1) * nt + np.arange(nt) for f in m.cells[c, 3:]])
        D[np.ix_(g, g)] += Ate[c]
        r[g] += bte[c]
    np.testing.assert_array_equal(K.to_scipy().toarray(), D)
    np.testing.assert_array_equal(rhs, r)


def test_dirichlet_projection_and_quirk(O):
    """Trace Dirichlet (boundary.jl:11-42): quirk=0 is the projection onto
    the orthonormal trace basis; quirk bit 1 keeps the reference's
    accumulator that is not reset between modes, i.e. the running sum of
    quirk=0."""
    m = O.rectangle_mesh(3, 3, jitter=0.2, seed=2)
    k = 2
    P = O.problem(k=k, qdeg=3)
    g = [0.3, 1.7, -0.9]
    d0, v0 = O.dirichlet_trace(P, m, g=g)
    d1, v1 = O.dirichlet_trace(P, m, g=g, quirk=1)
    full = linear_trace(O, m, k, g)
    np.testing.assert_array_equal(d0, d1)
    np.testing.assert_allclose(v0, full[d0 - 1], atol=1e-15)
    np.testing.assert_allclose(v1, np.cumsum(v0.reshape(-1, k + 1), axis=1).ravel(), atol=1e-15)
    # face order ascending, modes 1..nt (boundary.jl:19-37)
    assert np.all(np.diff(d0) > 0)
    dz, vz = O.dirichlet_trace(P, m)
    np.testing.assert_array_equal(vz, 0.0)


def test_dirichlet_reference_weights(O):
    """The reference weights its face points with the cell rule's weights
    (fs.fs.qr_weights, boundary.jl:33) and does not reset its accumulator
    (:27-37); quirk=True reproduces both.  Pinned by closed forms:
    k = 1 (Strang 3 points of weight 1/6 against Gauss-Legendre 2 points of
    weight 1/2): every term is a third of the projection's, so the values
    are a third of its running sum; k = 2, constant g (Strang 6 points of
    weight 1/12, StrangQuad.jl:12-19, at the 3 Gauss points 1/2, 1/2 +- sqrt(15)/10):
    mode 0 gets 3 g/12, mode 1 adds 0 (P1 is odd about 1/2), mode 2 adds
    g/12 (P2(1/2) + 2 P2(1/2 + sqrt(15)/10)) = g/12 sqrt(5) (-1/2 + 4/5)."""
    m = O.rectangle_mesh(4, 3, jitter=0.2, seed=5)
    P1 = O.problem(k=1, qdeg=2)
    g = [0.3, 1.7, -0.9]
    _, v0 = O.dirichlet_trace(P1, m, g=g)
    _, vr = O.dirichlet_trace(P1, m, g=g, quirk=True)
    np.testing.assert_allclose(vr, np.cumsum(v0.reshape(-1, 2), axis=1).ravel() / 3.0, rtol=1e-14, atol=1e-15)
    # weights alone, accumulator reset: a third of the projection
    _, vw = O.dirichlet_trace(P1, m, g=g, quirk=2)
    np.testing.assert_allclose(vw, v0 / 3.0, rtol=1e-14, atol=1e-15)
    P2 = O.problem(k=2, qdeg=3)
    c = 1.37
    _, w0 = O.dirichlet_trace(P2, m, g=[c, 0.0, 0.0])
    _, wr = O.dirichlet_trace(P2, m, g=[c, 0.0, 0.0], quirk=True)
    np.testing.assert_allclose(w0.reshape(-1, 3), np.tile([c, 0.0, 0.0], (len(w0) // 3, 1)), atol=1e-15)
    want = [c / 4, c / 4, c / 4 + c * np.sqrt(5.0) * 0.3 / 12]
    np.testing.assert_allclose(wr.reshape(-1, 3), np.tile(want, (len(wr) // 3, 1)), rtol=1e-14, atol=1e-15)


def test_apply_and_direct_solve_vs_scipy(O):
    """apply! (boundary.jl:121-158) + K \\ b against scipy's sparse LU, and
    the banded solve against the same."""
    m = O.rectangle_mesh(6, 5, jitter=0.2, seed=4)
    P = O.problem(k=2, qdeg=3)
    Ate, bte, Ke, be = O.condense(P, m)
    K, rhs = O.assemble(m, 3, Ate, bte)
    g = [0.1, 0.2, -0.3]
    dofs, vals = O.dirichlet_trace(P, m, g=g)
    Ka, ba = K.copy(), rhs.copy()
    O.apply(Ka, ba, dofs, vals)
    S = Ka.to_scipy()
    # Dirichlet rows and columns are decoupled with diagonal meandiag(K)
    Kd = K.to_scipy().toarray()
    md = np.abs(np.diag(Kd)).mean()
    Sd = S.toarray()
    di = dofs - 1
    np.testing.assert_array_equal(Sd[di][:, di], md * np.eye(len(di)))
    xs = spla.spsolve(S.tocsc(), ba)
    xb = O.banded_solve(Ka, ba)
    np.testing.assert_allclose(xb, xs, rtol=1e-11, atol=1e-13)
    np.testing.assert_allclose(xb[di], vals, rtol=1e-13, atol=1e-15)
    xp, it, rr = O.pcg_solve(K, rhs, 3, dofs, vals, rtol=1e-13)
    assert rr <= 1e-13
    np.testing.assert_allclose(xp, xs, rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_convergence_rate(O, k):
    """L2 error of u decays like h^(k+1) for the reference source."""
    P = O.problem(k=k, qdeg=QD[k])
    errs = []
    for n in (4, 8, 16):
        m = O.rectangle_mesh(n, n)
        r = O.solve_pipeline(P, m)
        errs.append(np.sqrt(O.errornorm(P, m, r["u"])))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > k + 1 - 0.25, rates


def test_kappa_manufactured_rate(O):
    """Variable kappa (an extension: A = (kappa^-1 q, v)) with the matching
    manufactured source converges at rate k+1; kappa=1 reduces to kind 0."""
    P1 = O.problem(k=2, qdeg=6, src_kind=1, kappa=(3.0, 5.0))
    errs = []
    for n in (8, 16):
        m = O.rectangle_mesh(n, n)
        r = O.solve_pipeline(P1, m)
        errs.append(np.sqrt(O.errornorm(P1, m, r["u"])))
    assert np.log2(errs[0] / errs[1]) > 2.6
    m = O.rectangle_mesh(3, 3)
    a = O.condense(O.problem(k=2, qdeg=3, src_kind=1), m)
    b = O.condense(O.problem(k=2, qdeg=3, src_kind=0), m)
    for x, y in zip(a, b):
        np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-13)


def test_triangle_figure1_mesh(O):
    """The unstructured fixture figure.1 (62 triangles): every cell has a
    positive Jacobian after parsing and the pipeline converges."""
    import os
    from conftest import GOLDEN
    m = O.parse_mesh_triangle(os.path.join(GOLDEN, "figure.1"))
    assert m.ncells == 62 and m.nnodes == 41
    P = O.problem(k=2, qdeg=3)
    r = O.solve_pipeline(P, m)
    assert O.errornorm(P, m, r["u"]) < 1e-4


def _sine_series_error(O, P, m, u, coef):
    """Squared L2 error (errornorm's quadrature: the space's cell rule) of u
    against u_ex = sum c_pq sin(p pi x) sin(q pi y) / (pi^2 (p^2 + q^2))."""
    pts, w = O.quad2d(P.qdeg)
    nu = (P.k + 1) * (P.k + 2) // 2
    N = np.array([[O.dubiner_value(j + 1, r, s) for j in range(nu)] for r, s in pts])  # (nqp, nu)
    x = m.xy[m.cells[:, :3] - 1]  # (nc, 3, 2)
    J = np.stack([x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]], axis=2)
    detJ = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
    lam = np.stack([1.0 - pts[:, 0] - pts[:, 1], pts[:, 0], pts[:, 1]], axis=1)
    xq = np.einsum("qv,cvd->cqd", lam, x)
    ue = np.zeros(xq.shape[:2])
    for p in range(1, 5):
        for q in range(1, 5):
            ue += (coef[(p - 1) * 4 + (q - 1)] / (np.pi ** 2 * (p * p + q * q))
                   * np.sin(p * np.pi * xq[..., 0]) * np.sin(q * np.pi * xq[..., 1]))
    uh = u @ N.T
    return float((((uh - ue) ** 2) * w[None, :]).sum(axis=1) @ detJ)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_sine_series_source_manufactured(O, k):
    """The serving workload's source f = sum_{p,q<=4} c_pq sin(p pi x)
    sin(q pi y) (coefficients p-major) has the exact solution
    u = sum c_pq sin(p pi x) sin(q pi y) / (pi^2 (p^2 + q^2)) with u = 0 on
    the boundary of the unit square; the discrete u converges to it at rate
    k+1.  A transposed coefficient index, a wrong frequency or a sign fails
    (the error then stalls at O(1))."""
    coef = np.random.default_rng(11).standard_normal(16)
    P = O.problem(k=k, qdeg=QD[k], src_kind=2, coef=coef)
    errs = []
    sizes = (8, 16, 32) if k == 1 else (8, 16)
    for n in sizes:
        m = O.rectangle_mesh(n, n, jitter=0.1, seed=3)
        r = O.solve_pipeline(P, m)
        errs.append(np.sqrt(_sine_series_error(O, P, m, r["u"], coef)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates[-1] > k + 1 - 0.35, (errs, rates)
    scale = np.sqrt(np.sum((coef / (np.pi ** 2 * 2)) ** 2))
    assert errs[-1] < 0.05 * scale, errs
    # the coefficient index is p-major: transposing it changes the source
    Pt = O.problem(k=k, qdeg=QD[k], src_kind=2, coef=coef.reshape(4, 4).T.ravel())
    m = O.rectangle_mesh(sizes[-1], sizes[-1], jitter=0.1, seed=3)
    rt = O.solve_pipeline(Pt, m)
    assert np.sqrt(_sine_series_error(O, Pt, m, rt["u"], coef)) > 10 * errs[-1]
