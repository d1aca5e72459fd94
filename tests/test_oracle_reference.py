"""Pins the CPU oracle to the reference's own tests and literals.

Each test cites the reference test it transcribes.  The fixture mesh
figure2.1 (5 nodes, 4 triangles) is the reference's test/mesh/figure2.1
data, copied into tests/golden/.
"""
import numpy as np
import pytest

sq2, sq3, sq6 = np.sqrt(2.0), np.sqrt(3.0), np.sqrt(6.0)


# ---------------------------------------------------------------- meshes
def test_triangle_fixture_mesh(O, fixture_mesh_root):
    """test/test_mesh.jl:7-26"""
    m = O.parse_mesh_triangle(fixture_mesh_root)
    assert m.ncells == 4 and m.nnodes == 5
    np.testing.assert_array_equal(m.cells[:, :3], [[2, 3, 5], [4, 1, 5], [5, 3, 4], [1, 2, 5]])
    np.testing.assert_array_equal(m.xy, [[0, 0], [1, 0], [1, 1], [0, 1], [0.5, 0.5]])
    assert tuple(m.cells[0, 3:]) == (1, 2, 3)
    assert [O.face_orientation(m, 1, i) for i in (1, 2, 3)] == [True, False, True]
    assert m.facesets["boundary"] == {3, 6, 7, 8}
    # get_coordinates(1, mesh) == [(1,1), (0.5,0.5)]
    np.testing.assert_array_equal(m.xy[m.faces[0, :2] - 1], [[1.0, 1.0], [0.5, 0.5]])
    # cell_diameter(mesh, 1) == 1.0 (mesh.jl:126-134)
    diam = max(np.linalg.norm(m.xy[m.faces[f - 1, 1] - 1] - m.xy[m.faces[f - 1, 0] - 1])
               for f in m.cells[0, 3:])
    assert diam == 1.0
    for c in range(4):
        x = m.xy[m.cells[c, :3] - 1]
        a, b = x[1] - x[0], x[2] - x[0]
        vol = 0.5 * abs(a[0] * b[1] - a[1] * b[0])
        assert vol == pytest.approx(0.25)


def test_rectangle_mesh_2x2(O):
    """test/test_mesh.jl:29-44"""
    m = O.rectangle_mesh(2, 2)
    assert m.ncells == 8 and m.nnodes == 9
    np.testing.assert_array_equal(m.cells[:, :3], [[1, 2, 4], [2, 5, 4], [2, 3, 5], [3, 6, 5],
                                                   [4, 5, 7], [5, 8, 7], [5, 6, 8], [6, 9, 8]])
    np.testing.assert_array_equal(m.xy, [[0, 0], [0.5, 0], [1, 0], [0, 0.5], [0.5, 0.5],
                                         [1, 0.5], [0, 1], [0.5, 1], [1, 1]])
    assert tuple(m.cells[0, 3:]) == (1, 2, 3)
    assert [O.face_orientation(m, 1, i) for i in (1, 2, 3)] == [True, False, True]
    assert m.facesets["boundary"] == {3, 7, 9, 16, 2, 11, 12, 15}
    diam = max(np.linalg.norm(m.xy[m.faces[f - 1, 1] - 1] - m.xy[m.faces[f - 1, 0] - 1])
               for f in m.cells[0, 3:])
    assert diam == np.sqrt(2) / 2


@pytest.mark.parametrize("n", [1, 3, 7, 16])
def test_rectangle_face_count_and_first_encounter(O, n):
    """nfaces = nx + ny + 3 nx ny (generate_mesh.jl:121); faces are numbered
    by first encounter, so face ids along the cell sweep never skip."""
    m = O.rectangle_mesh(n, n + 1)
    assert m.nfaces == n + (n + 1) + 3 * n * (n + 1)
    seen = 0
    for c in range(m.ncells):
        for f in m.cells[c, 3:]:
            if f > seen:
                assert f == seen + 1
                seen = f
    # every interior face has two cells, boundary faces one
    nb = (m.faces[:, 3] == 0).sum()
    assert nb == 2 * (n + n + 1) == len(m.facesets["boundary"])


# ------------------------------------------------------------ quadrature
def test_quadrature_literals(O):
    """test/test_quadrature.jl:6-25"""
    p, w = O.grundmann_moeller(0)
    np.testing.assert_allclose(w, [0.5])
    np.testing.assert_allclose(p[0], [1 / 3, 1 / 3])
    p, w = O.grundmann_moeller(1)
    np.testing.assert_allclose(w, 0.5 * np.array([0.520833333333333] * 3 + [-0.5625]), rtol=1e-12)
    np.testing.assert_allclose(p, [[1 / 5, 1 / 5], [3 / 5, 1 / 5], [1 / 5, 3 / 5], [1 / 3, 1 / 3]])
    for s in range(6):
        assert O.grundmann_moeller(s)[1].sum() == pytest.approx(0.5)
    for o in range(1, 7):
        assert O.strang(o)[1].sum() == pytest.approx(0.5)


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7, 9, 11])
def test_default_quad_exactness(O, order):
    """DefaultQuad (quadrature.jl:17-26): the rule of `order` integrates
    monomials x^a y^b, a+b <= order, exactly: a! b! / (a+b+2)!."""
    from math import factorial
    p, w = O.quad2d(order)
    for a in range(order + 1):
        for b in range(order + 1 - a):
            exact = factorial(a) * factorial(b) / factorial(a + b + 2)
            assert (w * p[:, 0] ** a * p[:, 1] ** b).sum() == pytest.approx(exact, rel=2e-12, abs=1e-15)
    with pytest.raises(ValueError):
        O.quad2d(8)


@pytest.mark.parametrize("n", range(1, 13))
def test_gauss_legendre_textbook(O, n):
    """gausslegendre (quadrature.jl:30) against numpy's Golub-Welsch."""
    x, w = O.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    np.testing.assert_allclose(x, xr, atol=4e-16)
    np.testing.assert_allclose(w, wr, rtol=1e-14)
    x01, w01 = O.quad1d(n)
    np.testing.assert_allclose(x01, (xr + 1) / 2, atol=4e-16)
    np.testing.assert_allclose(w01, wr / 2, rtol=1e-14)


# ----------------------------------------------------------------- basis
def test_dubiner_closed_forms_vs_recursion(O):
    """test/test_basis.jl:7-12"""
    p, w = O.strang(5)
    for j in range(1, 16):
        dv = [(O.dubiner_basis(x, y, j) - O.dubiner_value(j, x, y)) ** 2 for x, y in p]
        dg = [np.sum((O.dubiner_basis_grad(x, y, j) - O.dubiner_grad(j, x, y)) ** 2) for x, y in p]
        assert np.dot(dv, w) < np.finfo(float).eps
        assert np.dot(dg, w) < np.finfo(float).eps


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_dubiner_orthonormal(O, k):
    """Orthonormality of the Dubiner basis on the reference triangle (the
    normalisation of basis.jl:179-180), with an exact rule of degree 2k."""
    nu = (k + 1) * (k + 2) // 2
    p, w = O.quad2d(2 * k if 2 * k <= 6 else 2 * k + 1)
    V = np.array([[O.dubiner_value(j, x, y) for (x, y) in p] for j in range(1, nu + 1)])
    G = (V * w) @ V.T
    # the degree-9 Grundmann-Moeller rule used for k=4 has mixed-sign weights
    np.testing.assert_allclose(G, np.eye(nu), atol=2e-14 if k < 4 else 2e-13)


def test_lagrange_and_legendre(O):
    """test/test_basis.jl:32-93"""
    v, g = O.lagrange2d_p1(0.5, 0.5)
    np.testing.assert_allclose(v, [0.0, 0.5, 0.5], atol=np.finfo(float).eps)
    np.testing.assert_allclose(g, [[-1, -1], [1, 0], [0, 1]], atol=np.finfo(float).eps)
    np.testing.assert_allclose(O.lagrange1d_p1(0.5), [0.5, 0.5], atol=np.finfo(float).eps)
    ref = [lambda x: 1.0, lambda x: sq3 * (2 * x - 1), lambda x: np.sqrt(5) * (6 * x * x - 6 * x + 1),
           lambda x: np.sqrt(7) * (2 * x - 1) * (10 * x * x - 10 * x + 1)]
    x, w = O.quad1d(3)
    for i in range(4):
        d = [(ref[i](t) - O.legendre_value(i + 1, t)) ** 2 for t in x]
        assert np.dot(d, w) < np.finfo(float).eps


# ------------------------------------------------------ function spaces
def test_fixture_geometry(O, fixture_mesh_root):
    """test/test_ScalarFuncSp.jl:15-33 and test/test_FunctionSpace.jl:39-41"""
    m = O.parse_mesh_triangle(fixture_mesh_root)
    invs = [[[1.0, 1.0], [-2.0, 0.0]], [[-1.0, -1.0], [2.0, 0.0]],
            [[1.0, 1.0], [-1.0, 1.0]], [[1.0, -1.0], [0.0, 2.0]]]
    dets = [[sq2 / 2, sq2 / 2, 1], [sq2 / 2, sq2 / 2, 1], [1, sq2 / 2, sq2 / 2], [sq2 / 2, sq2 / 2, 1]]
    for c in range(4):
        detJ, Jinv, normals, detJf = O.geometry(m, c + 1)
        assert detJ == pytest.approx(0.5)
        np.testing.assert_allclose(Jinv, invs[c], atol=1e-14)
        np.testing.assert_allclose(detJf, dets[c], rtol=1e-14)
        if c == 0:
            np.testing.assert_allclose(normals, [[-sq2 / 2, sq2 / 2], [-sq2 / 2, -sq2 / 2], [1, 0]],
                                       atol=1e-15)
    # after the iteration the spaces hold cell 4: getfacedetJdS(Mh, f, 1)
    _, _, _, detJf = O.geometry(m, 4)
    w1 = O.quad1d(2)[1][0]
    np.testing.assert_allclose(detJf * w1, [sq2 / 4, sq2 / 4, 0.5], rtol=1e-14)


def fixture_literals():
    """Exact element matrices, test/test_FunctionSpace.jl:49-72."""
    Be = [np.array([[0, 0, 0], [0, 0, 0], [-3 * sq2, 0, 0], [0, 0, 0], [sq6, 0, 0], [0, 0, 0]]),
          np.array([[0, 0, 0], [0, 0, 0], [3 * sq2, 0, 0], [0, 0, 0], [-sq6, 0, 0], [0, 0, 0]]),
          np.array([[0, 0, 0], [sq6 / 2, 0, 0], [-1.5 * sq2, 0, 0], [0, 0, 0], [1.5 * sq6, 0, 0],
                    [1.5 * sq2, 0, 0]]),
          np.array([[0, 0, 0], [sq6, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0], [3 * sq2, 0, 0]])]
    c1 = np.array([[2 * sq2 + 2, 0, 2 - 2 * sq2], [0, 4 * sq2 + 4, 0], [2 - 2 * sq2, 0, 4 * sq2 + 4]])
    Ce = [c1, c1, np.array([[2 * sq2 + 2, sq6 - sq3, sq2 - 1], [sq6 - sq3, 4 * sq2 + 4, 0],
                            [sq2 - 1, 0, 4 * sq2 + 4]]), c1]
    Ee = [-np.array([[sq2 / 2, 0, sq2 / 2, 0, -sq2, 0], [sq3 / 2, -1 / 2, -sq3 / 2, 1 / 2, 0, -2],
                     [1 / 2, sq3 / 2, 1 / 2, sq3 / 2, 2, 0], [-sq2 / 2, 0, sq2 / 2, 0, 0, 0],
                     [-sq3 / 2, 1 / 2, -sq3 / 2, 1 / 2, 0, 0], [-1 / 2, -sq3 / 2, 1 / 2, sq3 / 2, 0, 0]]),
          -np.array([[-sq2 / 2, 0, -sq2 / 2, 0, sq2, 0], [-sq3 / 2, 1 / 2, sq3 / 2, -1 / 2, 0, -2],
                     [-1 / 2, -sq3 / 2, -1 / 2, -sq3 / 2, -2, 0], [sq2 / 2, 0, -sq2 / 2, 0, 0, 0],
                     [sq3 / 2, -1 / 2, sq3 / 2, -1 / 2, 0, 0], [1 / 2, sq3 / 2, -1 / 2, -sq3 / 2, 0, 0]]),
          -np.array([[0, 0, sq2 / 2, 0, -sq2 / 2, 0], [0, 0, -sq3 / 2, -1 / 2, 0, 1],
                     [0, 0, 1 / 2, -sq3 / 2, 1, 0], [-sq2, 0, sq2 / 2, 0, sq2 / 2, 0],
                     [-sq3, 1, -sq3 / 2, -1 / 2, 0, -1], [-1, -sq3, 1 / 2, -sq3 / 2, -1, 0]]),
          -np.array([[-sq2 / 2, 0, sq2 / 2, 0, 0, 0], [-sq3 / 2, 1 / 2, -sq3 / 2, 1 / 2, 0, 0],
                     [-1 / 2, -sq3 / 2, 1 / 2, sq3 / 2, 0, 0], [-sq2 / 2, 0, -sq2 / 2, 0, sq2, 0],
                     [-sq3 / 2, 1 / 2, sq3 / 2, -1 / 2, 0, 2], [-1 / 2, -sq3 / 2, -1 / 2, -sq3 / 2, -2, 0]])]
    h1 = np.diag([sq2 / 2, sq2 / 2, sq2 / 2, sq2 / 2, 1, 1])
    He = [h1, h1, np.diag([1, 1, sq2 / 2, sq2 / 2, sq2 / 2, sq2 / 2]), h1]
    return Be, Ce, Ee, He


def test_fixture_element_matrices(O, fixture_mesh_root):
    """test/test_FunctionSpace.jl:125-126,176-178 (Julia's isapprox: rtol
    sqrt(eps) in norm; checked here to 1e-14 absolute)."""
    m = O.parse_mesh_triangle(fixture_mesh_root)
    P = O.problem(k=1)
    Be, Ce, Ee, He = fixture_literals()
    for c in range(4):
        b = O.element_blocks(P, m, c + 1)
        np.testing.assert_allclose(b["A"], 0.5 * np.eye(6), atol=1e-14)
        np.testing.assert_allclose(b["B"], Be[c], atol=1e-14)
        np.testing.assert_allclose(b["C"], Ce[c], atol=1e-14)
        np.testing.assert_allclose(b["E"], Ee[c], atol=1e-14)
        np.testing.assert_allclose(b["H"], He[c], atol=1e-14)
        # F has no literal: with tau = 1 it equals the trace-test block of C
        # evaluated at the oriented points; pinned by the patch tests below.


def test_fixture_l2_bound(O, fixture_mesh_root):
    """test/test_FunctionSpace.jl:214-243: squared L2 <= 0.12."""
    m = O.parse_mesh_triangle(fixture_mesh_root)
    P = O.problem(k=1)
    r = O.solve_pipeline(P, m)
    e = O.errornorm(P, m, r["u"])
    assert e <= 0.12
    assert e == pytest.approx(0.11019700386, rel=1e-8)


def test_example_l2_bound(O):
    """examples/poisson2D_HDG.jl:189-218: 10x10 mesh, squared L2 <= 6e-5."""
    m = O.rectangle_mesh(10, 10)
    P = O.problem(k=1)
    r = O.solve_pipeline(P, m)
    e = O.errornorm(P, m, r["u"])
    assert e <= 0.00006
    assert e == pytest.approx(5.3645466464e-05, rel=1e-8)
