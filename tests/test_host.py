"""Host-side tests of libhdg and its Python mirror (no GPU needed): the C ABI
exports, the mesh generator (bit-exact against the oracle's restatement of
rectangle_mesh), the reference-element tables, the Triangle reader,
Dirichlet data and apply!."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, REPO


def header_symbols():
    txt = open(os.path.join(REPO, "include", "hdg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hdg_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(hdg):
    L = hdg.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(hdg.EXPORTS)


@pytest.mark.parametrize("nx,ny,jit", [(1, 1, 0.0), (2, 2, 0.0), (3, 5, 0.0), (7, 4, 0.2), (32, 32, 0.0),
                                       (16, 16, 0.2), (64, 48, 0.2)])
def test_rectangle_mesh_bit_exact(hdg, O, nx, ny, jit):
    """Connectivity, face numbering, coordinates and boundary sets of the
    closed-form generator equal the oracle's first-encounter (hash map)
    restatement of rectangle_mesh bit for bit."""
    m = hdg.rectangle_mesh((nx, ny), jitter=jit, seed=20261202)
    o = O.rectangle_mesh(nx, ny, jitter=jit, seed=20261202)
    assert np.array_equal(m.cells, o.cells)
    assert np.array_equal(m.faces, o.faces)
    assert np.array_equal(m.nodes.view(np.uint64), o.xy.view(np.uint64))
    assert m.facesets == o.facesets
    assert m.nodesets == o.nodesets


@pytest.mark.parametrize("n", [(1, 1), (2, 3), (5, 3), (16, 9)])
@pytest.mark.parametrize("LL,UR", [((1.0, 0.0), (0.0, 1.0)), ((0.0, 1.0), (1.0, 0.0)), ((1.0, 1.0), (0.0, 0.0)),
                                   ((2.0, -1.0), (-1.0, 3.0)), ((0.5, 0.0), (0.5, 1.0)), ((0.0, 0.25), (1.0, 0.25))])
def test_rectangle_mesh_inverted_box(hdg, O, n, LL, UR):
    """Boxes given with a decreasing side: every triangle comes out
    clockwise and rectangle_mesh swaps its vertices 2 and 3
    (_check_node_data, generate_mesh.jl:49-57), which reorders its local
    faces, reverses every stored face traversal and renumbers the first two
    boundary faces; both sides decreasing is a rotation and swaps nothing.
    Degenerate boxes (zero width or height) give a mesh too (the reference
    fails later, at det(J) <= 0).  Bit-exact against the oracle, which
    follows the reference's hash-map sweep."""
    try:
        o = O.rectangle_mesh(n[0], n[1], LL=LL, UR=UR)
    except RuntimeError:
        with pytest.raises(hdg.ArgumentError, match="HDG_ERR_BOUNDARY"):
            hdg.rectangle_mesh(n, LL=LL, UR=UR)
        return
    m = hdg.rectangle_mesh(n, LL=LL, UR=UR)
    assert np.array_equal(m.cells, o.cells)
    assert np.array_equal(m.faces, o.faces)
    assert np.array_equal(m.nodes.view(np.uint64), o.xy.view(np.uint64))
    assert m.facesets == o.facesets and m.nodesets == o.nodesets


@pytest.mark.parametrize("n", [(5, 3), (7, 9), (16, 16)])
@pytest.mark.parametrize("LL,UR", [((-1.0, 0.25), (2.5, 3.0)), ((-1.0, -2.0), (3.0, 2.0)),
                                   ((0.0, 0.0), (3.0, 7.0)), ((-0.5, 0.25), (1.5, 4.25))])
def test_rectangle_mesh_nonunit_box(hdg, O, n, LL, UR):
    """Same nodes and faces on general boxes, and the same failure where the
    reference's interpolated edge nodes miss the box side by an ulp
    (generate_mesh.jl:80 throws; libhdg returns HDG_ERR_BOUNDARY)."""
    try:
        o = O.rectangle_mesh(n[0], n[1], LL=LL, UR=UR)
    except RuntimeError:
        with pytest.raises(hdg.ArgumentError, match="HDG_ERR_BOUNDARY"):
            hdg.rectangle_mesh(n, LL=LL, UR=UR)
        return
    m = hdg.rectangle_mesh(n, LL=LL, UR=UR)
    assert np.array_equal(m.nodes.view(np.uint64), o.xy.view(np.uint64))
    assert np.array_equal(m.faces, o.faces)
    assert m.facesets == o.facesets


def test_rectangle_mesh_rejects_bad_args(hdg):
    with pytest.raises(hdg.ArgumentError):
        hdg.rectangle_mesh((0, 3))
    with pytest.raises(hdg.ArgumentError):
        hdg.rectangle_mesh((3, 0))


def test_jittered_mesh_stays_positive(hdg):
    m = hdg.rectangle_mesh((128, 128), jitter=0.2, seed=20261202)
    x = m.nodes[m.cells[:, :3] - 1]
    a, b = x[:, 1] - x[:, 0], x[:, 2] - x[:, 0]
    det = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    h2 = (1 / 128) ** 2
    assert det.min() > 0.3 * h2


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_dubiner_tables_match_closed_forms(hdg, O, k):
    """libhdg's Dubiner basis (collapsed-coordinate recursion) against the
    reference's closed forms (basis.jl:65-123) restated in the oracle."""
    from hdg_amd._lib import ptr
    rng = np.random.default_rng(k)
    nu = (k + 1) * (k + 2) // 2
    for _ in range(40):
        r, s = rng.random(2)
        if r + s > 1:
            r, s = 1 - r, 1 - s
        v, g = np.zeros(nu), np.zeros(2 * nu)
        assert hdg.lib().hdg_dubiner(k, r, s, ptr(v), ptr(g)) == 0
        # the degree-4 closed forms cancel to ~2e-13 (see the next test)
        for j in range(nu):
            assert v[j] == pytest.approx(O.dubiner_value(j + 1, r, s), abs=4e-13)
            np.testing.assert_allclose(g[2 * j:2 * j + 2], O.dubiner_grad(j + 1, r, s), atol=4e-12)
    # edges and vertices (face quadrature points lie on edges)
    for r, s in ((0.0, 0.3), (0.4, 0.6), (0.7, 0.0), (1.0, 0.0), (0.0, 0.0)):
        v = np.zeros(nu)
        hdg.lib().hdg_dubiner(k, r, s, ptr(v), None)
        for j in range(nu):
            assert v[j] == pytest.approx(O.dubiner_value(j + 1, r, s), abs=1e-13)


def test_dubiner_high_precision(hdg):
    """Degree-4 basis functions j = 11 and 15 (basis.jl:79,83) evaluated in
    40-digit arithmetic: libhdg's recursion is accurate to ~1e-14."""
    mp = pytest.importorskip("mpmath")
    from hdg_amd._lib import ptr
    mp.mp.dps = 40
    rng = np.random.default_rng(0)
    worst = 0.0
    for _ in range(100):
        r, s = rng.random(2)
        if r + s > 1:
            r, s = 1 - r, 1 - s
        R, S = mp.mpf(r), mp.mpf(s)
        e11 = 3 * mp.sqrt(10) * (70 * R**4 + 140 * R**3 * (S - 1) + 90 * R**2 * (S - 1)**2
                                 + 20 * R * (S - 1)**3 + (S - 1)**4)
        e15 = mp.sqrt(10) * (126 * S**4 - 224 * S**3 + 126 * S**2 - 24 * S + 1)
        v = np.zeros(15)
        hdg.lib().hdg_dubiner(4, r, s, ptr(v), None)
        worst = max(worst, abs(v[10] - float(e11)), abs(v[14] - float(e15)))
    assert worst < 2e-14


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7, 9, 11])
def test_quadrature_matches_reference_rules(hdg, O, order):
    p, w = hdg.quadrature(2, order)
    po, wo = O.quad2d(order)
    key = lambda P: np.lexsort((P[:, 1], P[:, 0]))
    a, b = key(p), key(po)
    np.testing.assert_allclose(p[a], po[b], atol=1e-15)
    np.testing.assert_allclose(w[a], wo[b], rtol=1e-13)
    x, w1 = hdg.quadrature(1, order)
    xo, w1o = O.quad1d(order)
    np.testing.assert_allclose(x, xo, atol=4e-16)
    np.testing.assert_allclose(w1, w1o, rtol=1e-14)
    with pytest.raises(hdg.ArgumentError):
        hdg.quadrature(2, 8)


def test_triangle_reader_matches_oracle(hdg, O):
    for name in ("figure2.1", "figure.1"):
        m = hdg.parse_mesh_triangle(os.path.join(GOLDEN, name))
        o = O.parse_mesh_triangle(os.path.join(GOLDEN, name))
        assert np.array_equal(m.cells, o.cells)
        assert np.array_equal(m.faces, o.faces)
        assert np.array_equal(m.nodes, o.xy)
        assert m.facesets["boundary"] == o.facesets["boundary"]
    m = hdg.parse_mesh_triangle(os.path.join(GOLDEN, "figure2.1"))
    assert [hdg.face_orientation(m, 1, i) for i in (1, 2, 3)] == [True, False, True]
    assert hdg.getcells_matrix(m).tolist() == [[2, 3, 5], [4, 1, 5], [5, 3, 4], [1, 2, 5]]


def test_reference_api_objects(hdg):
    """getnlocaldofs / getnbasefunctions (test/test_FunctionSpace.jl:30-32)."""
    m = hdg.parse_mesh_triangle(os.path.join(GOLDEN, "figure2.1"))
    fe = hdg.GenericFiniteElement(hdg.Dubiner(1))
    Wh, Vh = hdg.ScalarFunctionSpace(m, fe), hdg.VectorFunctionSpace(m, fe)
    Mh = hdg.ScalarTraceFunctionSpace(Wh, hdg.GenericFiniteElement(hdg.Legendre(1)))
    assert (hdg.getnlocaldofs(Vh), hdg.getnlocaldofs(Wh), hdg.getnlocaldofs(Mh)) == (6, 3, 6)
    assert hdg.getnbasefunctions(Mh) == 2
    assert hdg.TrialFunction(Mh).m_values.shape == (4, 2, 3)
    assert hdg.TrialFunction(Vh).m_values.shape == (4, 6)


@pytest.mark.parametrize("k", [1, 2, 3])
def test_dirichlet_projection_matches_oracle(hdg, O, k):
    m = hdg.rectangle_mesh((5, 4), jitter=0.2, seed=3)
    om = O.rectangle_mesh(5, 4, jitter=0.2, seed=3)
    qd = {1: 2, 2: 3, 3: 6}[k]
    fe = hdg.GenericFiniteElement(hdg.Dubiner(k))
    Wh = hdg.ScalarFunctionSpace(m, fe, quad_degree=qd)
    Mh = hdg.ScalarTraceFunctionSpace(Wh, hdg.GenericFiniteElement(hdg.Legendre(k)))
    g = (0.3, -1.2, 0.7)
    dbc = hdg.Dirichlet(hdg.TrialFunction(Mh), m, "boundary", lambda x: g[0] + g[1] * x[0] + g[2] * x[1])
    d, v = O.dirichlet_trace(O.problem(k=k, qdeg=qd), om, g=g)
    assert np.array_equal(dbc.prescribed_dofs, d)
    np.testing.assert_allclose(dbc.values, v, atol=1e-14)
    z = hdg.Dirichlet(hdg.TrialFunction(Mh), m, "boundary", lambda x: 0.0)
    assert np.all(z.values == 0.0)
    # the reference's own computation (cell-rule weights, running accumulator)
    dr = hdg.Dirichlet(hdg.TrialFunction(Mh), m, "boundary", lambda x: g[0] + g[1] * x[0] + g[2] * x[1],
                       reference=True)
    _, vr = O.dirichlet_trace(O.problem(k=k, qdeg=qd), om, g=g, quirk=True)
    np.testing.assert_allclose(dr.values, vr, atol=1e-14)
    assert np.abs(dr.values - dbc.values).max() > 1e-3


def test_apply_bc_matches_oracle(hdg, O):
    om = O.rectangle_mesh(4, 4, jitter=0.2, seed=9)
    P = O.problem(k=2, qdeg=3)
    Ate, bte, Ke, be = O.condense(P, om)
    K, rhs = O.assemble(om, 3, Ate, bte)
    d, v = O.dirichlet_trace(P, om, g=(0.5, 1.0, -2.0))
    Ka, ba = K.copy(), rhs.copy()
    O.apply(Ka, ba, d, v)

    class DBC:
        prescribed_dofs, values = d, v
    S = K.to_scipy()
    b2 = rhs.copy()
    hdg.apply_bc(S, b2, DBC)
    np.testing.assert_array_equal(S.toarray(), Ka.to_scipy().toarray())
    np.testing.assert_allclose(b2, ba, rtol=1e-15, atol=1e-15)


def test_host_recovery_matches_oracle(hdg, O):
    """get_usigma without a device context applies K_e, b_e on the host."""
    m = hdg.rectangle_mesh((3, 3), jitter=0.2, seed=1)
    om = O.rectangle_mesh(3, 3, jitter=0.2, seed=1)
    P = O.problem(k=2, qdeg=3)
    r = O.solve_pipeline(P, om)
    fe = hdg.GenericFiniteElement(hdg.Dubiner(2))
    Wh, Vh = hdg.ScalarFunctionSpace(m, fe), hdg.VectorFunctionSpace(m, fe)
    Mh = hdg.ScalarTraceFunctionSpace(Wh, hdg.GenericFiniteElement(hdg.Legendre(2)))
    s, u, uh = hdg.TrialFunction(Vh), hdg.TrialFunction(Wh), hdg.TrialFunction(Mh)
    hdg.get_usigma(s, u, uh, r["uhat"], r["Ke"], r["be"], m)
    np.testing.assert_allclose(u.m_values, r["u"], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(s.m_values, r["q"], rtol=1e-13, atol=1e-15)
    np.testing.assert_array_equal(uh.m_values, r["uhat_cell"])
    e_host = hdg.errornorm(u, "sinsin")
    assert e_host == pytest.approx(O.errornorm(P, om, r["u"]), rel=1e-12)
