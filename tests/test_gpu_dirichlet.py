"""Dirichlet data and apply! on the device (src/boundary.jl:11-42,121-175)
against the oracle: the projection of g and the reference's own
computation (cell-rule weights, running accumulator), and the applied
system (rhs lift, zeroed rows and columns, meandiag on the diagonal,
b_d = v_d meandiag) exported as Julia's CSC.  Tolerances: pattern
bit-exact, values and rhs 1e-12 relative to the largest entry, meandiag
1e-14 relative (its sum runs in another order than the reference's)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

QD = {1: 2, 2: 3, 3: 6}


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _meshes(hdg, O, which):
    if which == "rect":
        return hdg.rectangle_mesh((6, 5), jitter=0.2, seed=12), O.rectangle_mesh(6, 5, jitter=0.2, seed=12)
    root = os.path.join(GOLDEN, "figure.1")
    return hdg.parse_mesh_triangle(root), O.parse_mesh_triangle(root)


@pytest.mark.parametrize("which,k", [("rect", 1), ("rect", 2), ("rect", 3), ("fig", 2)])
@pytest.mark.parametrize("reference", [False, True])
def test_dirichlet_function_and_apply_vs_oracle(hdg, O, which, k, reference):
    print(f"[dirichlet] {which} k={k} reference={reference}", flush=True)
    m, om = _meshes(hdg, O, which)
    P = O.problem(k=k, qdeg=QD[k])
    g = (0.5, 1.0, -2.0)
    Ate, bte, _, _ = O.condense(P, om)
    K, rhs = O.assemble(om, k + 1, Ate, bte)
    d, v = O.dirichlet_trace(P, om, g=g, quirk=reference)
    Ka, ba = K.copy(), rhs.copy()
    O.apply(Ka, ba, d, v)
    with hdg.Context(0) as ctx:
        ctx.set_mesh(m)
        ctx.setup(k, quad_degree=QD[k])
        ctx.set_source()
        ctx.set_dirichlet_function(hdg._lib.G_LINEAR, g, reference=reference)
        ctx.assemble_condense()
        colptr, rowval, nzval, b, md = ctx.get_applied_system()
        # the unapplied system is left as it was
        c0, r0, nz0, b0 = ctx.get_system()
    assert np.array_equal(colptr, Ka.colptr) and np.array_equal(rowval, Ka.rowval)
    assert rel(nzval, Ka.nzval) < 1e-12
    # prescribed rows and columns are zero off the diagonal, meandiag on it
    isd = np.zeros(K.n, bool)
    isd[d - 1] = True
    cols = np.repeat(np.arange(K.n), np.diff(colptr))
    rows = rowval - 1
    hit = isd[cols] | isd[rows]
    diag = hit & (rows == cols)
    assert np.all(nzval[hit & ~diag] == 0.0)
    assert np.all(nzval[diag] == md)
    assert rel(b, ba) < 1e-12
    md_ref = float(np.abs(K.to_scipy().diagonal()).sum() / K.n)
    assert md == pytest.approx(md_ref, rel=1e-14)
    # prescribed values recovered from b_d = v_d meandiag
    assert rel(b[d - 1] / md, v) < 1e-13
    assert rel(nz0, K.nzval) < 1e-12 and rel(b0, rhs) < 1e-12


@pytest.mark.parametrize("reference", [False, True])
def test_dirichlet_function_sinsin_vs_host(hdg, reference):
    """g = sin(pi x) sin(pi y) on a box that is not the unit square, against
    the host mirror's Dirichlet (the same computation in numpy)."""
    k = 2
    m = hdg.rectangle_mesh((7, 6), LL=(0.25, 0.25), UR=(1.25, 1.5), jitter=0.2, seed=3)
    fe = hdg.GenericFiniteElement(hdg.Dubiner(k))
    Wh = hdg.ScalarFunctionSpace(m, fe, quad_degree=3)
    Mh = hdg.ScalarTraceFunctionSpace(Wh, hdg.GenericFiniteElement(hdg.Legendre(k)))
    dbc = hdg.Dirichlet(hdg.TrialFunction(Mh), m, "boundary",
                        lambda x: np.sin(np.pi * x[0]) * np.sin(np.pi * x[1]), reference=reference)
    with hdg.Context(0) as ctx:
        ctx.set_mesh(m)
        ctx.setup(k, quad_degree=3)
        ctx.set_source()
        ctx.set_dirichlet_function(hdg._lib.G_SINSIN, (0.0, 0.0, 0.0), reference=reference)
        ctx.assemble_condense()
        _, _, _, b, md = ctx.get_applied_system()
    d = dbc.prescribed_dofs
    np.testing.assert_allclose(b[d - 1] / md, dbc.values, rtol=1e-13, atol=1e-15)


def test_dirichlet_function_solution(hdg, O):
    """Linear g with f = 0 (the patch test): the device projection gives the
    solution u = g exactly and the same trace as the oracle's values passed
    through hdg_set_dirichlet."""
    k = 3
    m = hdg.rectangle_mesh((9, 7), jitter=0.2, seed=8)
    om = O.rectangle_mesh(9, 7, jitter=0.2, seed=8)
    g = (0.4, -1.1, 2.5)
    nqp = O.sizes(k, 6)[2]
    with hdg.Context(0) as ctx:
        ctx.set_mesh(m)
        ctx.setup(k, quad_degree=6, rtol=1e-14)
        ctx.set_source(kind=3, f_qp=np.zeros((m.cells.shape[0], nqp)))
        ctx.set_dirichlet_function(hdg._lib.G_LINEAR, g)
        ctx.run()
        u1 = ctx.trace()
        assert ctx.errornorm(kind=1, par=np.array(g)) < 1e-24
        d, v = O.dirichlet_trace(O.problem(k=k, qdeg=6), om, g=g)
        ctx.set_dirichlet(d, v)
        ctx.run()
        u2 = ctx.trace()
    assert rel(u1, u2) < 1e-13
