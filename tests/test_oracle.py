"""CPU tests that pin the oracle (oracle/oracle.cpp) before it is trusted as the checker.

Each test names what pins it: a reference test's numbers (file:line), an exact value the
mathematics fixes (closed forms, rational brute force in tests/golden/), an invariant, or an
independent library routine (scipy).
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle_lib as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = json.load(open(os.path.join(HERE, "golden", "known_answers.json")))


def p2_stats(geom, kind):
    m = O.Mesh.structured(geom, [9, 2], [0, 0], [3, 1])
    s = O.Space(m, True, 2)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    (op.append_laplace if kind == "laplace" else op.append_product)(1.0)
    op.assemble(4)  # assemble(true) in the reference test: thread-parallel
    A = op.csr().toarray()
    return A.min(), A.max(), float((A * A).sum())


@pytest.mark.parametrize("geom,kind", [(O.QUAD, "laplace"), (O.TRI, "laplace"), (O.QUAD, "mass"), (O.TRI, "mass")])
def test_p2_known_answers(geom, kind):
    """dune/gdt/test/integrands/integrands_laplace.cc:115-135, integrands_product.cc:107-126"""
    name = f"p2_{kind}_{'cube' if geom == O.QUAD else 'simplex'}"
    ref = GOLDEN["reference"][name]
    exact = {k: float(Fraction(*v)) for k, v in GOLDEN["exact"][name].items()}
    mn, mx, ss = p2_stats(geom, kind)
    tmin, tmax, tss = ref["tol"]
    assert abs(mn - ref["min"]) <= tmin
    assert abs(mx - ref["max"]) <= tmax
    # exact rational value (tests/golden/make_golden.py): the oracle is within the reference's tolerance of it
    assert abs(mn - exact["min"]) <= tmin and abs(mx - exact["max"]) <= tmax
    assert abs(ss - exact["sumsq"]) <= tss
    if name == "p2_laplace_simplex":
        # the reference's printed 2481.524691358029 is itself 4.5e-12 above the exact 402007/162;
        # the oracle is 1.8e-12 below it, so the two differ by 6.3e-12 (> the reference's 5e-12)
        assert abs(ss - ref["sumsq"]) <= 1e-11
    else:
        assert abs(ss - ref["sumsq"]) <= tss


def test_pointwise_laplace_integrand():
    """dune/gdt/test/integrands/integrands_laplace.cc:71-113 with the bases of integrands.hh:72-141"""
    rng = np.random.default_rng(0)
    for x0, x1 in rng.random((5, 2)):
        W = np.array([[x0, x1], [1.0, 2.0]]) * x0 * x1
        # scalar bases: ansatz {x, x^2 y}, test {y, x y^3}
        ag = np.array([[[1.0, 0.0]], [[2 * x0 * x1, x0 ** 2]]])
        tg = np.array([[[0.0, 1.0]], [[x1 ** 3, 3 * x0 * x1 ** 2]]])
        res = O.laplace_evaluate(W, tg, ag)
        expected = np.array([[1, 2 * (x0 * x1 + x0 ** 2)],
                             [x0 * x1 ** 3 + 3 * x0 * x1 ** 2,
                              3 * x0 ** 2 * x1 ** 4 + 6 * (x0 ** 2 * x1 ** 3 + x0 ** 3 * x1 ** 2)]]) * x0 * x1
        np.testing.assert_allclose(res, expected, rtol=4 * np.finfo(float).eps, atol=0)
        # vector bases: ansatz {(x,y), (x^2,y^2)}, test {(1,2), (xy, x+y)}
        ag = np.array([[[1, 0], [0, 1]], [[2 * x0, 0], [0, 2 * x1]]], float)
        tg = np.array([[[0, 0], [0, 0]], [[x1, x0], [1, 1]]], float)
        res = O.laplace_evaluate(W, tg, ag)
        expected_row1 = np.array([x0 * x1 + x0 + x1 + 2,
                                  2 * (x0 ** 2 * x1 + x1 ** 2 + x0 ** 2 + 2 * x1)]) * x0 * x1
        np.testing.assert_allclose(res[0], 0.0, atol=0)
        np.testing.assert_allclose(res[1], expected_row1, rtol=4 * np.finfo(float).eps, atol=0)


def test_pointwise_product_integrand():
    """dune/gdt/test/integrands/integrands_product.cc:58-105"""
    rng = np.random.default_rng(1)
    for x0, x1 in rng.random((5, 2)):
        # scalar: weight x y, test {y, x y^3}, ansatz {x, x^2 y}
        res = O.product_evaluate(np.array([[x0 * x1]]), np.array([[x1], [x0 * x1 ** 3]]),
                                 np.array([[x0], [x0 ** 2 * x1]]))
        expected = np.array([[(x0 * x1) ** 2, (x0 * x1) ** 3], [x0 ** 3 * x1 ** 4, x0 ** 4 * x1 ** 5]])
        np.testing.assert_allclose(res, expected, rtol=4 * np.finfo(float).eps)
        # vector: weight [[x, y], [1, 2]], test {(1,2), (xy, x+y)}, ansatz {(x,y), (x^2,y^2)}
        W = np.array([[x0, x1], [1.0, 2.0]])
        tv = np.array([[1.0, 2.0], [x0 * x1, x0 + x1]])
        av = np.array([[x0, x1], [x0 ** 2, x1 ** 2]])
        res = O.product_evaluate(W, tv, av)
        expected = np.array([
            [x0 ** 2 + 2 * x0 * x1 + 5 * x1, x0 ** 3 + 2 * x0 ** 2 * x1 + 5 * x1 ** 2],
            [x0 ** 3 * x1 + x0 ** 2 * x1 + 2 * x0 * x1 ** 2 + 2 * x0 * x1 + 2 * x1 ** 2,
             x0 ** 4 * x1 + x0 ** 3 * x1 + x0 ** 2 * x1 ** 2 + x0 * x1 ** 3 + 2 * x0 * x1 ** 2 + 2 * x1 ** 3]])
        np.testing.assert_allclose(res, expected, rtol=8 * np.finfo(float).eps)


def swipdg_operator(n, sigma_inner=8.0, sigma_dirichlet=14.0, kappa=1.0, omega=1.0, source=None):
    m = O.Mesh.structured(O.QUAD, [n, n], [-1, -1], [1, 1])
    s = O.Space(m, False, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_laplace(kappa)
    op.append_inner_coupling(1.0, kappa, omega)
    op.append_inner_penalty(sigma_inner, omega)
    op.append_dirichlet_coupling(1.0, kappa)
    op.append_boundary_penalty(sigma_dirichlet, omega)
    if source is None:
        op.append_source(esv2007=True, f_order=4)
    else:
        op.append_source(source)
    op.assemble(1)
    return m, s, op


def test_esv2007_swipdg_h1_errors():
    """stationary_heat_equation__ESV2007__table_1.mini:31-33 (cube grids, Q1-DG, sigma 8 / 14,
    h_F = face volume, force of order 4, ESV2007.hh:51-61,107-113)"""
    ref = GOLDEN["reference"]["esv2007_q1dg_cube"]
    for n, h_ref, e_ref in zip((8, 16, 32), ref["h"], ref["h1_semi"]):
        m, s, op = swipdg_operator(n)
        rp, cols = op.pattern()
        x, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-12)
        assert rc == 0
        err = O.esv2007_h1_semi_error(s, x, 10)
        assert float(f"{err:.2e}") == e_ref
        assert float(f"{2 * math.sqrt(2) / n:.2e}") == h_ref


def test_swipdg_structure():
    """SWIPDG matrix: symmetric; A 1 only sees the boundary penalty (consistency terms of a constant
    vanish, inner jumps of a constant vanish); nnz = 16 E + 32 inner faces (sparsity-pattern.hh:126-140)"""
    n = 6
    m, s, op = swipdg_operator(n, source=0.0)
    A = op.csr()
    assert abs(A - A.T).max() < 1e-14 * abs(A).max()
    assert op.nnz == 16 * n * n + 32 * 2 * n * (n - 1)
    ones = np.ones(A.shape[0])
    m2, s2, op2 = swipdg_operator(n, source=0.0)
    # boundary-only operator: Dirichlet coupling + penalty alone
    mb = O.Mesh.structured(O.QUAD, [n, n], [-1, -1], [1, 1])
    sb = O.Space(mb, False, 1)
    ob = O.Operator(sb, O.STENCIL_ELEMENT_AND_INTERSECTION)
    ob.append_dirichlet_coupling(1.0, 1.0)
    ob.append_boundary_penalty(14.0, 1.0)
    ob.assemble(1)
    np.testing.assert_allclose(A @ ones, ob.csr() @ ones, atol=1e-12)


def q1_heat(n, kappa=1.0, f=1.0, dirichlet=True):
    m = O.Mesh.structured(O.QUAD, [n, n], [0, 0], [1, 1])
    s = O.Space(m, True, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    op.append_laplace(kappa)
    op.append_source(f)
    op.set_dirichlet(dirichlet)
    op.assemble(1)
    return m, s, op


def test_q1_stencil_closed_form():
    """Q1 Laplace on squares: interior row = 8/3 and eight -1/3; b_i = h^2 for f = 1"""
    n = 8
    m, s, op = q1_heat(n, dirichlet=False)
    A = op.csr()
    b = op.rhs()
    i = 4 * (n + 1) + 4
    row = A.getrow(i).toarray().ravel()
    nz = np.nonzero(row)[0]
    assert len(nz) == 9
    np.testing.assert_allclose(row[i], 8.0 / 3.0, rtol=1e-15)
    np.testing.assert_allclose(row[nz[nz != i]], -1.0 / 3.0, rtol=1e-14)
    np.testing.assert_allclose(b[i], 1.0 / n ** 2, rtol=1e-14)
    np.testing.assert_allclose(A @ np.ones(A.shape[0]), 0.0, atol=1e-14)  # row sums before constraints
    assert op.nnz == (3 * n + 1) ** 2


def test_mass_and_volume():
    """sum of all mass-matrix entries = |Omega| (Q1 2d, P1 3d Kuhn)"""
    for geom, cells, dim in ((O.QUAD, [5, 3], 2), (O.TET, [3, 2, 2], 3), (O.TRI, [4, 5], 2)):
        m = O.Mesh.structured(geom, cells, [0] * dim, [1, 2, 3][:dim])
        s = O.Space(m, True, 1)
        op = O.Operator(s, O.STENCIL_ELEMENT)
        op.append_product(1.0)
        op.assemble(1)
        np.testing.assert_allclose(op.values().sum(), np.prod([1, 2, 3][:dim]), rtol=1e-13)


def test_p1_tet_closed_form():
    """P1 tet: A = kappa |T| G G^T (G rows grad lambda_i), M = |T|/20 (1 + delta_ij)"""
    coords = np.array([[0.1, 0.2, 0.0], [1.3, 0.1, 0.2], [0.2, 0.9, 0.1], [0.3, 0.4, 1.1]])
    conn = np.array([[0, 1, 2, 3]], np.int32)
    m = O.Mesh.explicit(O.TET, coords, conn)
    s = O.Space(m, True, 1)
    J = (coords[1:] - coords[0]).T
    vol = abs(np.linalg.det(J)) / 6
    G = np.vstack([-np.ones(3), np.eye(3)]) @ np.linalg.inv(J)
    for kind, expected in (("laplace", 2.5 * vol * G @ G.T), ("mass", vol / 20 * (np.ones((4, 4)) + np.eye(4)))):
        op = O.Operator(s, O.STENCIL_ELEMENT)
        if kind == "laplace":
            op.append_laplace(2.5)
        else:
            op.append_product(1.0)
        op.assemble(1)
        np.testing.assert_allclose(op.csr().toarray(), expected, rtol=1e-13, atol=1e-15)


def kuhn_nnz(c):
    return (c + 1) ** 3 + 2 * (3 * c * (c + 1) ** 2 + 3 * c ** 2 * (c + 1) + c ** 3)


def test_pattern_counts_and_mesh():
    """P1 Kuhn stencil: nnz = N + 2 (#edges); conforming: every interior face shared by 2 tets"""
    for c in (1, 2, 4):
        m = O.Mesh.structured(O.TET, [c, c, c], [0, 0, 0], [1, 1, 1])
        s = O.Space(m, True, 1)
        op = O.Operator(s, O.STENCIL_ELEMENT)
        assert op.nnz == kuhn_nnz(c)
        nb = m.neighbors()
        boundary = int((nb < 0).sum())
        assert boundary == 2 * 6 * c * c  # 2 triangles per boundary square
        # neighbour relation symmetric
        for e, f in zip(*np.nonzero(nb >= 0)):
            assert e in nb[nb[e, f]]


def test_dirichlet_constraints():
    """tools/dirichlet-constraints.hh:153-184: unit rows and columns, b_i = 0, symmetry kept"""
    n = 16
    m, s, op = q1_heat(n)
    A = op.csr()
    b = op.rhs()
    d = op.dirichlet_dofs()
    assert len(d) == 4 * n
    for i in d[:: max(1, len(d) // 10)]:
        row = A.getrow(i).toarray().ravel()
        col = A.getcol(i).toarray().ravel()
        assert row[i] == 1.0 and col[i] == 1.0
        assert np.count_nonzero(row) == 1 and np.count_nonzero(col) == 1
        assert b[i] == 0.0
    assert abs(A - A.T).max() < 1e-15


def test_pcg_against_direct_solve():
    """Jacobi-PCG to 1e-10 vs scipy's direct solve; Q1 on 128^2 quads (16 641 DoFs): max u_h vs the
    continuous max 0.0736713 of -Laplace u = 1 on the unit square"""
    m, s, op = q1_heat(128)
    rp, cols = op.pattern()
    vals, b = op.values(), op.rhs()
    x, it, rel, rc = O.pcg_jacobi(rp, cols, vals, b, rtol=1e-10)
    assert rc == 0 and rel <= 1e-10
    A = sp.csr_matrix((vals, cols, rp))
    xd = spla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 2e-10  # true residual
    assert abs(x.max() - 0.0736713) < 2e-5
    assert s.size == 16641 and op.nnz == 148225


def test_pcg_failure_modes():
    """host-forced failures: max_iter exhaustion and a non-SPD diagonal are reported, never silent"""
    m, s, op = q1_heat(16)
    rp, cols = op.pattern()
    x, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-10, max_iter=3)
    assert rc != 0 and it == 3
    m, s, op = q1_heat(8, kappa=-1.0, dirichlet=False)
    rp, cols = op.pattern()
    x, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-10)
    assert rc == -2


def test_adjoint_gradient_central_differences():
    """dJ/dlog kappa_e = -kappa_e lambda_e^T Khat_e u_e vs central differences of J = 1/2 |u - u_obs|^2"""
    rng = np.random.default_rng(5)
    m = O.Mesh.structured(O.TET, [3, 3, 3], [0, 0, 0], [1, 1, 1])
    s = O.Space(m, True, 1)
    ne = m.ne

    def solve(logk):
        op = O.Operator(s, O.STENCIL_ELEMENT)
        op.append_laplace(np.exp(logk))
        op.append_source(1.0)
        op.set_dirichlet(True)
        op.assemble(1)
        rp, cols = op.pattern()
        x, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-15, max_iter=10000)
        return x, (rp, cols, op.values())

    g0 = rng.standard_normal(ne)
    u_obs, _ = solve(rng.standard_normal(ne))
    u, (rp, cols, vals) = solve(g0)
    lam, _, _, _ = O.pcg_jacobi(rp, cols, vals, u - u_obs, rtol=1e-15, max_iter=10000)
    grad = O.adjoint_log_kappa_gradient(s, np.exp(g0), u, lam)
    h = 1e-6
    for e in rng.choice(ne, 8, replace=False):
        gp, gm = g0.copy(), g0.copy()
        gp[e] += h
        gm[e] -= h
        up, _ = solve(gp)
        um, _ = solve(gm)
        fd = (0.5 * np.sum((up - u_obs) ** 2) - 0.5 * np.sum((um - u_obs) ** 2)) / (2 * h)
        assert abs(fd - grad[e]) <= 1e-6 * max(abs(grad).max(), 1e-30)


def test_error_norms_closed_forms():
    """BilinearForm(e, e).apply2() with product and Laplace integrands (operators/bilinear-form.hh:344-452):
    interpolants of linear functions are exact on P1 / Q1, so against the function itself both norms vanish;
    against 0 they are integrals known in closed form"""
    for geom, dim in ((O.QUAD, 2), (O.TRI, 2), (O.HEX, 3), (O.TET, 3)):
        cells = [3, 4] if dim == 2 else [2, 3, 2]
        lo, hi = [0.0] * dim, [2.0, 1.0, 3.0][:dim]
        m = O.Mesh.structured(geom, cells, lo, hi)
        s = O.Space(m, True, 1)
        X = m.coords()
        a = np.array([0.5, -1.25, 2.0][:dim])
        u = X @ a + 0.75  # interpolant of v(x) = a.x + 0.75
        order = 4
        P = m.quadrature_points(order)
        gv = P @ a + 0.75
        gg = np.broadcast_to(a, P.shape).copy()
        l2, h1 = O.error_norms(s, u, order, gv, gg)
        assert abs(l2) < 1e-26 and abs(h1) < 1e-26
        vol = float(np.prod(hi))
        # against g = 0: ||v||^2 over the box and |a|^2 vol, per-element kappa weighting the H1 part
        kappa = 1.0 + np.arange(m.ne) % 3
        l2, h1 = O.error_norms(s, u, order, kappa=kappa)
        # int (a.x + c)^2 over [0, hi]: exact by expanding (independent coordinates)
        c = 0.75
        mean = sum(a[k] * hi[k] / 2 for k in range(dim)) + c
        var = sum(a[k] ** 2 * hi[k] ** 2 / 12 for k in range(dim))
        assert abs(l2 - vol * (mean ** 2 + var)) < 1e-12 * vol * (mean ** 2 + var)
        ev = vol / m.ne  # structured grids: every element has the same volume
        assert abs(h1 - float(a @ a) * ev * kappa.sum()) < 1e-12 * h1


def test_error_norms_match_esv2007_h1_error():
    """the H1-semi part against the exact gradient equals the ESV2007 error the fixture pins"""
    n = 8
    m, s, op = swipdg_operator(n)
    rp, cols = op.pattern()
    x, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-12)
    P = m.quadrature_points(10)
    X, Y = P[..., 0], P[..., 1]
    gg = np.stack([-np.pi / 2 * np.sin(np.pi / 2 * X) * np.cos(np.pi / 2 * Y),
                   -np.pi / 2 * np.cos(np.pi / 2 * X) * np.sin(np.pi / 2 * Y)], axis=-1)
    l2, h1 = O.error_norms(s, x, 10, np.cos(np.pi / 2 * X) * np.cos(np.pi / 2 * Y), gg)
    ref = O.esv2007_h1_semi_error(s, x, 10)
    assert abs(np.sqrt(h1) - ref) <= 1e-13 * ref
    assert float(f"{np.sqrt(h1):.2e}") == GOLDEN["reference"]["esv2007_q1dg_cube"]["h1_semi"][0]
    assert 0 < l2 < h1


def _boundary_points(m, order):
    return m.boundary_quadrature_points(order)


@pytest.mark.parametrize("geom,continuous", [(O.QUAD, False), (O.TRI, True)])
def test_boundary_functional_sums(geom, continuous):
    """LocalIntersectionIntegralFunctional on the boundary (local/functionals/integrals.hh:157-185): with
    constant data the sum over all DoFs follows from the partition of unity -- product: w g |boundary|;
    penalty: sigma w g per boundary face (penalty sigma w / h_F times |F| = h_F); Dirichlet coupling: the
    gradient of the partition of unity vanishes, so the sum is 0"""
    m = O.Mesh.structured(geom, [6, 3], [0, 0], [2, 1])
    s = O.Space(m, continuous, 1)
    nb = int(np.sum(m.neighbors() < 0))
    for kind, param, coeff, g, expect in [(2, 0.0, 1.5, 0.7, 1.5 * 0.7 * 6.0), (1, 8.0, 2.0, 0.5, 8.0 * 2.0 * 0.5 * nb),
                                          (0, 1.0, 3.0, 0.9, 0.0)]:
        op = O.Operator(s, O.STENCIL_ELEMENT)
        op.append_boundary_source(kind, param, coeff, g, g_order=0)
        op.assemble(1)
        total = op.rhs().sum()
        assert abs(total - expect) <= 1e-13 * max(1.0, abs(expect)), (kind, total, expect)
    # data given at the quadrature points instead of a constant gives the same vector
    op1 = O.Operator(s, O.STENCIL_ELEMENT)
    op1.append_boundary_source(1, 8.0, 2.0, 0.5, g_order=0)
    op1.assemble(1)
    P = _boundary_points(m, 0 + 1 + 0)
    op2 = O.Operator(s, O.STENCIL_ELEMENT)
    op2.append_boundary_source(1, 8.0, 2.0, np.full(P.shape[:2], 0.5), g_order=0)
    op2.assemble(1)
    assert np.array_equal(op1.rhs(), op2.rhs())


@pytest.mark.parametrize("geom", [O.QUAD, O.TRI])
def test_swipdg_reproduces_linear_solutions(geom):
    """consistency of SIPG with weakly imposed Dirichlet data (laplace-ipdg.hh:292-320 for the coupling
    right-hand side, ipdg.hh:254-282 with the data as ansatz for the penalty): u = 0.3 + 2x - 1.5y solves
    -div(grad u) = 0 and lies in the discrete space, so the discrete solution reproduces it.  On quadrilaterals
    the reference's order for the coupling term (data + diffusion + max(test - 1, 0)) under-integrates the
    bilinear normal derivative along a face: one order more makes the face rule exact; at the reference's
    own order the interpolant does not satisfy the system (residual 5.8e-2)"""
    n, sigma = 5, 16.0
    m = O.Mesh.structured(geom, [n, n], [0, 0], [1, 1])
    s = O.Space(m, False, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_laplace(1.0)
    op.append_inner_coupling(1.0, 1.0, 1.0)
    op.append_inner_penalty(sigma, 1.0)
    op.append_dirichlet_coupling(1.0, 1.0)
    op.append_boundary_penalty(sigma, 1.0)

    def u(X):
        return 0.3 + 2.0 * X[..., 0] - 1.5 * X[..., 1]

    over = 1 if geom == O.QUAD else 0
    op.append_boundary_source(0, 1.0, 1.0, u(_boundary_points(m, 1 + 0 + 0 + over)), g_order=1, over_integrate=over)
    op.append_boundary_source(1, sigma, 1.0, u(_boundary_points(m, 0 + 1 + 1)), g_order=1)
    op.assemble(1)
    rp, cols = op.pattern()
    x, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-14, max_iter=10000)
    X = m.coords()[m.conn()]  # [ne, local vertex, 2]: DG DoF (e, l) sits at vertex l of e
    gi = s.global_indices()
    assert np.abs(x[gi] - u(X)).max() <= 1e-10


def swipdg_operator_on(mesh, sigma_inner=8.0, sigma_dirichlet=14.0):
    s = O.Space(mesh, False, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_laplace(1.0)
    op.append_inner_coupling(1.0, 1.0, 1.0)
    op.append_inner_penalty(sigma_inner, 1.0)
    op.append_dirichlet_coupling(1.0, 1.0)
    op.append_boundary_penalty(sigma_dirichlet, 1.0)
    op.append_source(esv2007=True, f_order=4)
    op.assemble(4)
    return s, op


def test_esv2007_swipdg_h1_errors_simplex():
    """stationary_heat_equation__ESV2007__table_1.mini:11-13: P1-DG on the ALU conforming simplex grid of
    ESV2007.hh:66-69 (4 x 4 cubes split into triangles, bisection-refined), sigma 8 / 14, h_F = face length,
    force of order 4.  The refined grid has criss-cross diagonals (tests/problems.py); the published errors
    are reproduced to three significant figures on it, not on a uniform Kuhn split of the same cells (whose
    errors, 3.32e-01 / 1.63e-01 / 8.07e-02, are asserted to differ, so the grid is what is pinned too)"""
    from problems import bisection_refined_triangles

    ref = GOLDEN["reference"]["esv2007_p1dg_simplex"]
    for n, h_ref, e_ref in zip((8, 16, 32), ref["h"], ref["h1_semi"]):
        X, conn = bisection_refined_triangles(n)
        s, op = swipdg_operator_on(O.Mesh.explicit(O.TRI, X, conn))
        rp, cols = op.pattern()
        x, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-12, max_iter=100000)
        assert rc == 0
        err = O.esv2007_h1_semi_error(s, x, 10)
        assert float(f"{err:.2e}") == e_ref, (n, err)
        # h: the largest element diameter (the legs' hypotenuse)
        assert float(f"{math.sqrt(2) * 2 / n:.2e}") == h_ref
        sk, opk = swipdg_operator_on(O.Mesh.structured(O.TRI, [n, n], [-1, -1], [1, 1]))
        rp, cols = opk.pattern()
        xk, *_ = O.pcg_jacobi(rp, cols, opk.values(), opk.rhs(), rtol=1e-12, max_iter=100000)
        assert float(f"{O.esv2007_h1_semi_error(sk, xk, 10):.2e}") != e_ref


def _swip_face_matrix(a, c, b, kappa, omega, s, sigma):
    """The SWIPDG face form of two Q1 elements [0, a] x [0, b] (inside, -) and [a, a + c] x [0, b] (outside,
    +) written from its textbook definition (Ern, Stephansen, Zunino 2009, with the half harmonic penalty
    weight of the reference): a_F(u, v) = -int {kappa grad u}_w . n [v] - s int {kappa grad v}_w . n [u]
    + sigma gamma / h_F int [u] [v], with n = e_x, [v] = v- - v+, {q}_w = w- q- + w+ q+,
    w- = d+ / (d- + d+), w+ = d- / (d- + d+), gamma = d- d+ / (d- + d+), d = n . (omega n), h_F = b.
    Traces and fluxes are linear along the face and integrated with the 1-d mass matrix; global DoF
    4 e + local vertex (x bit 0, y bit 1).  Row = test function, column = ansatz function."""
    M1 = np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
    km, kp = kappa
    dm, dp = omega
    wm, wp = dp / (dm + dp), dm / (dm + dp)
    gamma = dm * dp / (dm + dp)
    jump, avg = np.zeros((8, 2)), np.zeros((8, 2))
    for g in range(8):
        e, loc = divmod(g, 4)
        bx, by = loc & 1, loc >> 1
        ey = np.eye(2)[by]
        h = a if e == 0 else c
        on_face = bx == 1 if e == 0 else bx == 0
        jump[g] = (1.0 if on_face else 0.0) * ey * (1 if e == 0 else -1)
        avg[g] = (wm * km if e == 0 else wp * kp) * (1 / h if bx == 1 else -1 / h) * ey
    A = np.empty((8, 8))
    for i in range(8):
        for j in range(8):
            A[i, j] = b * (-avg[j] @ M1 @ jump[i] - s * avg[i] @ M1 @ jump[j]
                           + sigma * gamma / b * jump[j] @ M1 @ jump[i])
    return A


@pytest.mark.parametrize("kappa,omega,s", [((3.0, 0.25), (3.0, 0.25), 1.0), ((3.0, 0.25), (2.0, 5.0), 1.0),
                                           ((1.0, 1e-4), (1.0, 1e-4), -1.0), ((3.0, 0.25), (2.0, 5.0), 0.0)])
def test_swipdg_heterogeneous_face_closed_form(kappa, omega, s):
    """laplace-ipdg.hh:149-152 (weights) and ipdg.hh:149-153 (penalty) with kappa- != kappa+, weights omega
    distinct from kappa, symmetric / non-symmetric / incomplete variants, non-square cells: the oracle's
    coupling + penalty blocks equal the face form built from jumps and weighted averages; with the weights
    exchanged the matrix differs at O(1)"""
    a, c, b = 0.5, 1.25, 0.8
    X = np.array([[0, 0], [a, 0], [a + c, 0], [0, b], [a, b], [a + c, b]], float)
    m = O.Mesh.explicit(O.QUAD, X, np.array([[0, 1, 3, 4], [1, 2, 4, 5]], np.int32))
    op = O.Operator(O.Space(m, False, 1), O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_inner_coupling(s, np.array(kappa), np.array(omega))
    op.append_inner_penalty(16.0, np.array(omega))
    op.assemble(1)
    A = op.csr().toarray()
    E = _swip_face_matrix(a, c, b, kappa, omega, s, 16.0)
    assert np.abs(A - E).max() <= 1e-14 * np.abs(E).max()
    W = _swip_face_matrix(a, c, b, kappa, omega[::-1], s, 16.0)
    assert np.abs(A - W).max() >= 0.05 * np.abs(E).max()


def _channel_swipdg(n, kappa, weights, sigma=16.0):
    m = O.Mesh.structured(O.QUAD, [n, n], [0, 0], [1, 1])
    s = O.Space(m, False, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_laplace(kappa)
    op.append_inner_coupling(1.0, kappa, weights)
    op.append_inner_penalty(sigma, kappa)
    op.append_dirichlet_coupling(1.0, kappa)
    op.append_boundary_penalty(sigma, kappa)
    return m, s, op


def test_swipdg_high_contrast_is_coercive():
    """kappa in {1, 1e-4} channels (the benchmark's SWIPDG field on 8 x 8 cells), sigma = 16: the symmetric
    weighted scheme is symmetric positive definite (dense eigenvalues); the same scheme with the averages'
    weights exchanged -- omega = 1 / kappa gives w- = k- / (k- + k+) -- is indefinite, so coercivity pins
    which side each weight belongs to"""
    import hgdt.fields as F

    n = 8
    kappa = F.channel_field(3, n, n)
    assert set(np.unique(kappa)) == {1e-4, 1.0}
    lam = {}
    for name, w in (("weighted", kappa), ("exchanged", 1.0 / kappa)):
        m, s, op = _channel_swipdg(n, kappa, w)
        op.assemble(1)
        A = op.csr().toarray()
        assert np.abs(A - A.T).max() <= 1e-15 * np.abs(A).max()
        lam[name] = np.linalg.eigvalsh(A).min()
    assert lam["weighted"] > 1e-5 and lam["exchanged"] < -0.1, lam


def test_swipdg_reproduces_layered_solution_across_channels():
    """kappa constant along every cell row with channels of 1e-4 (contrast 1e4), f = 0, weak Dirichlet data
    g = u: u(y) = 0.5 + int_0^y 1 / kappa is continuous, piecewise linear (in the Q1-DG space) and has the
    continuous flux kappa u' = 1, so the consistent scheme reproduces it -- which needs kappa (not the
    weights) in the fluxes and per-element kappa in the boundary terms (laplace-ipdg.hh:158-185, 300-368)"""
    import scipy.sparse.linalg as spla

    import hgdt.fields as F

    n, sigma = 8, 16.0
    kappa = F.channel_field(3, n, n)
    krow = kappa.reshape(n, n)[:, 0]
    assert np.all(kappa.reshape(n, n) == krow[:, None]) and krow.min() == 1e-4
    h = 1.0 / n
    nodes = np.concatenate([[0.0], np.cumsum(h / krow)])

    def u(X):
        y = X[..., 1]
        r = np.minimum((y / h).astype(int), n - 1)
        return 0.5 + nodes[r] + (y - r * h) / krow[r]

    m, s, op = _channel_swipdg(n, kappa, kappa, sigma)
    over = 1  # as in test_swipdg_reproduces_linear_solutions: an exact face rule for the coupling data term
    op.append_boundary_source(0, 1.0, kappa, u(m.boundary_quadrature_points(1 + over)), g_order=1,
                              over_integrate=over)
    op.append_boundary_source(1, sigma, kappa, u(m.boundary_quadrature_points(2)), g_order=1)
    op.assemble(1)
    x = spla.spsolve(op.csr().tocsc(), op.rhs())
    ue = u(m.coords()[m.conn()])
    assert np.abs(x[s.global_indices()] - ue).max() <= 1e-12 * np.abs(ue).max()


@pytest.mark.parametrize("kind", ["laplace", "mass"])
def test_p2_tetrahedra_exact(kind):
    """P2 on Kuhn tetrahedra (2 x 1 x 1 cells of widths 1/2, 1/2, 3/4): the oracle's global matrix against the
    exact rational one of tests/golden/make_golden.py (monomials integrated exactly over the reference
    tetrahedron).  DoF numbering is the reference's own and not fixed by any fixture, so the comparison is
    permutation-invariant: DoF count, min / max / sum of squares of the entries, and the whole spectrum."""
    ex = GOLDEN["exact"][f"p2_{kind}_tet_kuhn"]
    widths = [float(Fraction(*w)) for w in ex["widths"]]
    cells = ex["cells"]
    m = O.Mesh.structured(O.TET, cells, [0, 0, 0], [c * w for c, w in zip(cells, widths)])
    s = O.Space(m, True, 2)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    (op.append_laplace if kind == "laplace" else op.append_product)(1.0)
    op.assemble(1)
    A = op.csr().toarray()
    assert A.shape[0] == ex["size"]
    scale = np.abs(A).max()
    assert abs(A.min() - float(Fraction(*ex["min"]))) <= 1e-14 * scale
    assert abs(A.max() - float(Fraction(*ex["max"]))) <= 1e-14 * scale
    ss = float(Fraction(*ex["sumsq"]))
    assert abs(float((A * A).sum()) - ss) <= 1e-13 * ss
    lam = np.linalg.eigvalsh(A)
    assert np.abs(lam - np.array(ex["eigenvalues"])).max() <= 1e-13 * scale


def test_q1_hexahedron_closed_form():
    """Q1 on one axis-parallel hexahedron of widths (hx, hy, hz): stiffness
    kappa [hy hz / hx K1 (x) M1 (x) M1 + hx hz / hy M1 (x) K1 (x) M1 + hx hy / hz M1 (x) M1 (x) K1] and mass
    hx hy hz M1 (x) M1 (x) M1, K1 = [[1, -1], [-1, 1]], M1 = [[1/3, 1/6], [1/6, 1/3]], local vertex b with
    x = bit 0 (numpy's kron puts the last factor fastest, so the factors are listed z, y, x)"""
    K1 = np.array([[1.0, -1.0], [-1.0, 1.0]])
    M1 = np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
    hx, hy, hz, kappa = 0.5, 1.0, 2.0, 1.7
    m = O.Mesh.structured(O.HEX, [1, 1, 1], [0.1, -0.2, 0.3], [0.1 + hx, -0.2 + hy, 0.3 + hz])
    s = O.Space(m, True, 1)
    stiff = kappa * (hy * hz / hx * np.kron(M1, np.kron(M1, K1)) + hx * hz / hy * np.kron(M1, np.kron(K1, M1))
                     + hx * hy / hz * np.kron(K1, np.kron(M1, M1)))
    mass = hx * hy * hz * np.kron(M1, np.kron(M1, M1))
    for kind, expected in (("laplace", stiff), ("mass", mass)):
        op = O.Operator(s, O.STENCIL_ELEMENT)
        if kind == "laplace":
            op.append_laplace(kappa)
        else:
            op.append_product(1.0)
        op.assemble(1)
        np.testing.assert_allclose(op.csr().toarray(), expected, rtol=1e-14, atol=1e-15 * np.abs(expected).max())
    # on a 3 x 2 x 2 grid (no constraints) every row sums to 0, and the pattern is the 27-point stencil
    m = O.Mesh.structured(O.HEX, [3, 2, 2], [0, 0, 0], [1.5, 1.0, 2.0])
    s = O.Space(m, True, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    op.append_laplace(kappa)
    op.assemble(1)
    np.testing.assert_allclose(op.csr() @ np.ones(s.size), 0.0, atol=1e-14)
    assert op.nnz == (3 * 3 + 1) * (3 * 2 + 1) * (3 * 2 + 1)  # tensor product of the 1-d counts 3 n + 1


def _face_midpoints(m):
    """midpoints of the boundary intersections in walk order (one-point line rule)"""
    return m.boundary_quadrature_points(1)[:, 0, :]


def test_mixed_dirichlet_neumann_q1_closed_form():
    """DirichletConstraints with a boundary info (tools/dirichlet-constraints.hh:85-110): Dirichlet on x = 0, 1,
    no constraint (homogeneous Neumann) on y = 0, 1.  -u'' = 1 reduces to the 1-d problem, whose P1/Q1 nodal
    values are exact: u_h = x (1 - x) / 2 at every vertex.  With every boundary intersection Dirichlet the
    solution vanishes on y = 0, 1 instead, so the face selection is what is checked"""
    import scipy.sparse.linalg as spla

    n = 8
    m = O.Mesh.structured(O.QUAD, [n, 6], [0, 0], [1, 0.75])
    s = O.Space(m, True, 1)
    mid = _face_midpoints(m)
    dmask = (np.abs(mid[:, 0]) < 1e-12) | (np.abs(mid[:, 0] - 1) < 1e-12)
    assert dmask.sum() == 2 * 6
    op = O.Operator(s, O.STENCIL_ELEMENT)
    op.append_laplace(1.0)
    op.append_source(1.0)
    op.set_dirichlet(True)
    op.set_dirichlet_faces(dmask)
    op.assemble(1)
    X = m.coords()
    on_x = (np.abs(X[:, 0]) < 1e-12) | (np.abs(X[:, 0] - 1) < 1e-12)
    assert np.array_equal(np.sort(op.dirichlet_dofs()), np.nonzero(on_x)[0])
    x = spla.spsolve(op.csr().tocsc(), op.rhs())
    assert np.abs(x - X[:, 0] * (1 - X[:, 0]) / 2).max() <= 1e-13
    op2 = O.Operator(s, O.STENCIL_ELEMENT)
    op2.append_laplace(1.0)
    op2.append_source(1.0)
    op2.set_dirichlet(True)
    op2.assemble(1)
    assert len(op2.dirichlet_dofs()) > len(op.dirichlet_dofs())


@pytest.mark.parametrize("geom", [O.QUAD, O.TRI])
def test_swipdg_mixed_boundary_reproduces_linear_solution(geom):
    """SWIPDG with the weak Dirichlet forms and data only on the Dirichlet intersections (x = 0, 1:
    ApplyOn::CustomBoundaryIntersections(boundary_info, DirichletBoundary()), diffusion-ipdg.hh:302-311) and
    natural homogeneous Neumann conditions on y = 0, 1: u = 0.3 + 2 x has zero normal flux there, lies in the
    space and is reproduced; applying the Dirichlet forms on every boundary intersection instead does not
    reproduce it (u != 0.3 + 2 x is imposed on y = 0, 1 only weakly, with data 0 there)"""
    import scipy.sparse.linalg as spla

    n, sigma = 5, 16.0
    m = O.Mesh.structured(geom, [n, n], [0, 0], [1, 1])
    s = O.Space(m, False, 1)
    mid = _face_midpoints(m)
    dmask = ((np.abs(mid[:, 0]) < 1e-12) | (np.abs(mid[:, 0] - 1) < 1e-12)).astype(np.uint8)

    def u(X):
        return 0.3 + 2.0 * X[..., 0]

    over = 1 if geom == O.QUAD else 0
    results = []
    for mask in (dmask, None):
        op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
        op.append_laplace(1.0)
        op.append_inner_coupling(1.0, 1.0, 1.0)
        op.append_inner_penalty(sigma, 1.0)
        op.set_boundary_filter(mask)
        op.append_dirichlet_coupling(1.0, 1.0)
        op.append_boundary_penalty(sigma, 1.0)
        g1 = u(m.boundary_quadrature_points(1 + over))
        g2 = u(m.boundary_quadrature_points(2))
        if mask is None:  # data 0 on the y faces, as a Dirichlet condition there would impose
            g1[dmask == 0] = 0.
            g2[dmask == 0] = 0.
        op.append_boundary_source(0, 1.0, 1.0, g1, g_order=1, over_integrate=over)
        op.append_boundary_source(1, sigma, 1.0, g2, g_order=1)
        op.assemble(1)
        x = spla.spsolve(op.csr().tocsc(), op.rhs())
        results.append(np.abs(x[s.global_indices()] - u(m.coords()[m.conn()])).max())
    assert results[0] <= 1e-11 and results[1] > 1e-3, results


def _face_form_from_traces(jump, avg_flux, M, area, hF, s, sigma, gamma):
    """a_F(u, v) = -int {kappa grad u}_w . n [v] - s int {kappa grad v}_w . n [u] + sigma gamma / h_F int [u][v]
    for traces expanded in a face basis with mass matrix M (of the unit face; scaled by area): jump[g] and
    avg_flux[g] are the coefficient vectors of [phi_g] and {kappa grad phi_g}_w . n"""
    G = len(jump)
    A = np.empty((G, G))
    for i in range(G):
        for j in range(G):
            A[i, j] = area * (-avg_flux[j] @ M @ jump[i] - s * avg_flux[i] @ M @ jump[j]
                              + sigma * gamma / hF * jump[j] @ M @ jump[i])
    return A


def _swip_face_matrix_hex(a, a2, b, c, kappa, omega, s, sigma):
    """two Q1 hexahedra [0, a] x [0, b] x [0, c] (-) and [a, a + a2] x [0, b] x [0, c] (+), face x = a with
    n = e_x, area b c and diameter h_F = sqrt(b^2 + c^2) (its diagonal): traces are bilinear in (y, z), so
    the face mass matrix of the unit square is M1 (x) M1; the x-derivative of a trilinear basis function on
    the face is +-1/h_x times its (y, z) factor.  Global DoF 8 e + local vertex (x bit 0, y bit 1, z bit 2)"""
    M1 = np.array([[1 / 3, 1 / 6], [1 / 6, 1 / 3]])
    M = np.kron(M1, M1)
    km, kp = kappa
    dm, dp = omega
    wm, wp = dp / (dm + dp), dm / (dm + dp)
    jump, avg = np.zeros((16, 4)), np.zeros((16, 4))
    for g in range(16):
        e, loc = divmod(g, 8)
        bx, by, bz = loc & 1, (loc >> 1) & 1, loc >> 2
        eyz = np.kron(np.eye(2)[by], np.eye(2)[bz])
        hx = a if e == 0 else a2
        on_face = bx == 1 if e == 0 else bx == 0
        jump[g] = (1.0 if on_face else 0.0) * eyz * (1 if e == 0 else -1)
        avg[g] = (wm * km if e == 0 else wp * kp) * (1 / hx if bx == 1 else -1 / hx) * eyz
    return _face_form_from_traces(jump, avg, M, b * c, np.hypot(b, c), s, sigma, dm * dp / (dm + dp))


@pytest.mark.parametrize("kappa,omega,s", [((3.0, 0.25), (3.0, 0.25), 1.0), ((3.0, 0.25), (2.0, 5.0), 1.0),
                                           ((1.0, 1e-4), (1.0, 1e-4), -1.0), ((3.0, 0.25), (2.0, 5.0), 0.0)])
def test_swipdg_face_closed_form_hexahedra(kappa, omega, s):
    """3-d intersections (integrals.hh:197-267 over the faces of hexahedra): the oracle's coupling + penalty
    blocks of two box cells of different widths equal the textbook face form with the bilinear face traces;
    exchanging the weights, or taking the face's edge instead of its diagonal as h_F, changes the matrix"""
    a, a2, b, c = 0.5, 1.25, 0.8, 0.6
    m = O.Mesh.explicit(O.HEX, np.array([[x, y, z] for z in (0, c) for y in (0, b) for x in (0, a, a + a2)], float),
                        np.array([[0, 1, 3, 4, 6, 7, 9, 10], [1, 2, 4, 5, 7, 8, 10, 11]], np.int32))
    op = O.Operator(O.Space(m, False, 1), O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_inner_coupling(s, np.array(kappa), np.array(omega))
    op.append_inner_penalty(16.0, np.array(omega))
    op.assemble(1)
    A = op.csr().toarray()
    E = _swip_face_matrix_hex(a, a2, b, c, kappa, omega, s, 16.0)
    assert np.abs(A - E).max() <= 1e-14 * np.abs(E).max()
    W = _swip_face_matrix_hex(a, a2, b, c, kappa, omega[::-1], s, 16.0)
    assert np.abs(A - W).max() >= 0.05 * np.abs(E).max()
    H = _swip_face_matrix_hex(a, a2, b, c, kappa, omega, s, 16.0 * np.hypot(b, c) / b)
    assert np.abs(A - H).max() >= 0.05 * np.abs(E).max()


def _p1_gradients(V):
    """gradients of the barycentric coordinates of the tetrahedron with vertex rows V"""
    T = np.hstack([np.ones((4, 1)), V])
    return np.linalg.inv(T)[1:].T  # row i: grad lambda_i


def test_swipdg_face_closed_form_tetrahedra():
    """3-d intersections over the faces of tetrahedra: two P1 tetrahedra sharing the face F = {(1,0,0),
    (0,1,0), (0,0,1)}; traces are linear on F (face mass matrix |F| / 12 [[2,1,1],[1,2,1],[1,1,2]]), fluxes
    constant per element (gradients from the barycentric system), n = (1,1,1)/sqrt(3) out of the first
    element, h_F = the longest edge sqrt(2)"""
    V0 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    apex = np.array([0.9, 0.7, 0.8])
    X = np.vstack([V0, apex])
    conn = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], np.int32)
    m = O.Mesh.explicit(O.TET, X, conn)
    kappa, omega, s, sigma = (2.0, 0.5), (1.5, 4.0), 1.0, 16.0
    op = O.Operator(O.Space(m, False, 1), O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_inner_coupling(s, np.array(kappa), np.array(omega))
    op.append_inner_penalty(sigma, np.array(omega))
    op.assemble(1)
    A = op.csr().toarray()
    n = np.ones(3) / np.sqrt(3)
    area = np.sqrt(3) / 2
    Mf = np.array([[2, 1, 1], [1, 2, 1], [1, 1, 2]]) / 12  # P1 mass matrix of a triangle of area 1
    dm, dp = omega
    wm, wp = dp / (dm + dp), dm / (dm + dp)
    face = [1, 2, 3]  # global vertices of F, the face basis
    jump, avg = np.zeros((8, 3)), np.zeros((8, 3))
    for e in range(2):
        G = _p1_gradients(X[conn[e]])
        for loc in range(4):
            g = 4 * e + loc
            v = conn[e][loc]
            tr = np.zeros(3)
            if v in face:
                tr[face.index(v)] = 1.0
            jump[g] = tr * (1 if e == 0 else -1)
            # a constant flux: its coefficient vector in the (partition of unity) face basis is constant
            avg[g] = (wm * kappa[0] if e == 0 else wp * kappa[1]) * (G[loc] @ n) * np.ones(3)
    E = _face_form_from_traces(jump, avg, Mf, area, np.sqrt(2), s, sigma, dm * dp / (dm + dp))
    assert np.abs(A - E).max() <= 1e-13 * np.abs(E).max()


def _swipdg_3d(m, s, kappa, sigma, g_vals, over):
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_laplace(kappa)
    op.append_inner_coupling(1.0, kappa, kappa)
    op.append_inner_penalty(sigma, kappa)
    op.append_dirichlet_coupling(1.0, kappa)
    op.append_boundary_penalty(sigma, kappa)
    op.append_boundary_source(0, 1.0, kappa, g_vals(1 + over), g_order=1, over_integrate=over)
    op.append_boundary_source(1, sigma, kappa, g_vals(2), g_order=1)
    return op


@pytest.mark.parametrize("geom", [O.HEX, O.TET])
def test_swipdg_3d_reproduces_linear_and_layered_solutions(geom):
    """SWIPDG on 3-d grids with weak Dirichlet data g = u (boundary functionals over the faces,
    local/functionals/integrals.hh:157-185): a linear u with kappa = 1, and u(z) = 0.2 + int_0^z 1 / kappa
    for kappa constant on layers of cells (contrast 1e3, continuous flux kappa u' = 1).  Both lie in the
    P1 / Q1-DG space and are reproduced; the linear u with the layered kappa is not a solution (its flux
    jumps across the layers) and is not reproduced"""
    import scipy.sparse.linalg as spla

    m = O.Mesh.structured(geom, [3, 3, 4], [0, 0, 0], [1, 1, 1])
    s = O.Space(m, False, 1)
    over = 1 if geom == O.HEX else 0
    Xe = m.coords()[m.conn()]
    layer = np.minimum((Xe[..., 2].mean(axis=1) * 4).astype(int), 3)
    kl = np.array([1.0, 1e-3, 1.0, 0.2])
    zb = np.concatenate([[0], np.cumsum(0.25 / kl)])

    def u_lin(P):
        return 0.3 + 2.0 * P[..., 0] - 1.0 * P[..., 1] + 0.5 * P[..., 2]

    def u_layer(P):
        z = P[..., 2]
        k = np.minimum((z * 4).astype(int), 3)
        return 0.2 + zb[k] + (z - 0.25 * k) / kl[k]

    # vertices nudged into their element, so that a vertex on a layer interface takes its element's layer
    Pn = Xe + 1e-12 * (Xe.mean(axis=1, keepdims=True) - Xe)
    res = {}
    for name, kappa, u in (("linear", np.ones(m.ne), u_lin), ("layered", kl[layer], u_layer),
                           ("linear, layered kappa", kl[layer], u_lin)):
        op = _swipdg_3d(m, s, kappa, 16.0, lambda order: u(m.boundary_quadrature_points(order)), over)
        op.assemble(1)
        x = spla.spsolve(op.csr().tocsc(), op.rhs())
        res[name] = np.abs(x[s.global_indices()] - u(Pn)).max()
    assert res["linear"] <= 1e-10 and res["layered"] <= 1e-9 and res["linear, layered kappa"] > 1e-3, res


@pytest.mark.parametrize("geom,cells", [(O.QUAD, [4, 3]), (O.TRI, [3, 3]), (O.TET, [2, 2, 2])])
def test_operator_between_continuous_and_discontinuous_spaces(geom, cells):
    """make_matrix_operator(source_space, range_space) (operators/matrix-based.hh:200-216): rows are the range
    space's test functions, columns the source space's ansatz functions.  A continuous function u_c is the
    same function in the discontinuous space, u_d[e, i] = u_c[cg(e, i)], so with a product and a Laplace
    form  A(CG <- DG) u_d = A(CG <- CG) u_c  and  A(DG <- CG) u_c = A(DG <- DG) u_d  (an exchanged row or column
    mapping breaks both); the rectangular shapes and the nnz of the element pattern follow the mappers"""
    rng = np.random.default_rng(5)
    m = O.Mesh.structured(geom, cells, [0] * len(cells), [1.0] * len(cells))
    cg, dg = O.Space(m, True, 1), O.Space(m, False, 1)
    kappa = np.exp(rng.standard_normal(m.ne))
    ops = {}
    for name, rng_space, src in (("cc", cg, None), ("cd", cg, dg), ("dc", dg, cg), ("dd", dg, None)):
        op = O.Operator(rng_space, O.STENCIL_ELEMENT, source=src)
        op.append_laplace(kappa)
        op.append_product(2.5)
        op.assemble(1)
        ops[name] = op.csr()
    gi = cg.global_indices()  # [ne, ls] continuous DoFs of each element's local DoFs
    u_c = rng.standard_normal(cg.size)
    u_d = u_c[gi].reshape(-1)
    assert ops["cd"].shape == (cg.size, dg.size) and ops["dc"].shape == (dg.size, cg.size)
    a = ops["cc"] @ u_c
    assert np.abs(ops["cd"] @ u_d - a).max() <= 1e-13 * np.abs(a).max()
    b = ops["dd"] @ u_d
    assert np.abs(ops["dc"] @ u_c - b).max() <= 1e-13 * np.abs(b).max()
    # element pattern of CG <- DG: every continuous DoF couples to all DoFs of its elements
    ls = gi.shape[1]
    counts = np.bincount(gi.reshape(-1), minlength=cg.size)
    assert ops["cd"].nnz == int((counts * ls).sum())
    with pytest.raises(RuntimeError):
        O.Operator(cg, O.STENCIL_ELEMENT, source=dg).set_dirichlet(True)


def test_q2_hexahedron_closed_form():
    """Q2 on a hexahedron [0, a] x [0, b] x [0, c] (27 nodes, x fastest): the stiffness matrix is
    Kx(x)My(x)Mz + Mx(x)Ky(x)Mz + Mx(x)My(x)Kz and the mass matrix Mx(x)My(x)Mz with the 1-d quadratic
    Lagrange matrices K = 1/(3h) [[7,-8,1],[-8,16,-8],[1,-8,7]], M = h/30 [[4,2,-1],[2,16,2],[-1,2,4]]
    (Kronecker products in the order z, y, x of the node index iz*9 + iy*3 + ix)"""
    a, b, c = 0.7, 1.1, 0.45

    def K1(h):
        return np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / (3 * h)

    def M1(h):
        return np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) * h / 30

    K = (np.kron(M1(c), np.kron(M1(b), K1(a))) + np.kron(M1(c), np.kron(K1(b), M1(a)))
         + np.kron(K1(c), np.kron(M1(b), M1(a))))
    M = np.kron(M1(c), np.kron(M1(b), M1(a)))
    m = O.Mesh.structured(O.HEX, [1, 1, 1], [0, 0, 0], [a, b, c])
    s = O.Space(m, False, 2)
    for form, E in (("laplace", K), ("product", M)):
        op = O.Operator(s, O.STENCIL_ELEMENT)
        (op.append_laplace if form == "laplace" else op.append_product)(1.0)
        op.assemble(1)
        A = op.csr().toarray()
        assert A.shape == (27, 27)
        assert np.abs(A - E).max() <= 1e-14 * np.abs(E).max(), form


@pytest.mark.parametrize("geom", [O.TRI, O.QUAD])
def test_swipdg_order2_reproduces_quadratic_solution(geom):
    """SWIPDG with P2 / Q2 discontinuous spaces (the intersection integrands of laplace-ipdg.hh / ipdg.hh
    with order-2 bases, quadrature of order test + ansatz): u = x^2 - xy + y^2/2 + 0.3 x, -Laplace(u) = -3,
    weak Dirichlet data g = u, lies in the space and is reproduced (every term integrated exactly: for
    quadrilaterals the Dirichlet coupling functional over-integrates by one)"""
    import scipy.sparse.linalg as spla

    n, sigma = 4, 40.0
    m = O.Mesh.structured(geom, [n, n], [0, 0], [1, 1])
    s = O.Space(m, False, 2)

    def u(P):
        x, y = P[..., 0], P[..., 1]
        return x * x - x * y + 0.5 * y * y + 0.3 * x

    over = 1 if geom == O.QUAD else 0
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_laplace(1.0)
    op.append_inner_coupling(1.0, 1.0, 1.0)
    op.append_inner_penalty(sigma, 1.0)
    op.append_dirichlet_coupling(1.0, 1.0)
    op.append_boundary_penalty(sigma, 1.0)
    op.append_source(-3.0)
    op.append_boundary_source(0, 1.0, 1.0, u(m.boundary_quadrature_points(2 + 1 + over)), g_order=2,
                              over_integrate=over)
    op.append_boundary_source(1, sigma, 1.0, u(m.boundary_quadrature_points(2 + 2)), g_order=2)
    op.assemble(1)
    x = spla.spsolve(op.csr().tocsc(), op.rhs())
    # nodal values: the Lagrange nodes of P2 (vertices, edge midpoints) / Q2 (3 x 3 lattice) per element
    Xe = m.coords()[m.conn()]
    gi = s.global_indices()
    if geom == O.TRI:
        nodes = np.concatenate([Xe, (Xe[:, [0, 0, 1]] + Xe[:, [1, 2, 2]]) / 2], axis=1)
    else:
        t = np.array([0, 0.5, 1.0])
        lo, hi = Xe[:, 0], Xe[:, 3]
        nodes = np.stack([np.stack([lo[:, 0] + t[i % 3] * (hi[:, 0] - lo[:, 0]),
                                    lo[:, 1] + t[i // 3] * (hi[:, 1] - lo[:, 1])], -1) for i in range(9)], 1)
    err = np.abs(x[gi] - u(nodes)).max()
    assert err <= 1e-10, err

@pytest.mark.parametrize("geom", [O.HEX, O.TET])
def test_swipdg_order2_reproduces_quadratic_solution_3d(geom):
    """the same on Q2 hexahedra (27 basis functions per side of a face) and P2 tetrahedra:
    u = x^2 - xy + y^2/2 + 0.3x + z^2 - 0.2yz, -Laplace(u) = -5, weak Dirichlet data g = u, reproduced at every
    Lagrange node (for hexahedra the Dirichlet coupling functional over-integrates by one)"""
    import scipy.sparse.linalg as spla

    n, sigma = 2, 60.0
    m = O.Mesh.structured(geom, [n, n, n], [0, 0, 0], [1, 1, 1])
    s = O.Space(m, False, 2)

    def u(P):
        x, y, z = P[..., 0], P[..., 1], P[..., 2]
        return x * x - x * y + 0.5 * y * y + 0.3 * x + z * z - 0.2 * y * z

    over = 1 if geom == O.HEX else 0
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_laplace(1.0)
    op.append_inner_coupling(1.0, 1.0, 1.0)
    op.append_inner_penalty(sigma, 1.0)
    op.append_dirichlet_coupling(1.0, 1.0)
    op.append_boundary_penalty(sigma, 1.0)
    op.append_source(-5.0)
    op.append_boundary_source(0, 1.0, 1.0, u(m.boundary_quadrature_points(2 + 1 + over)), g_order=2,
                              over_integrate=over)
    op.append_boundary_source(1, sigma, 1.0, u(m.boundary_quadrature_points(2 + 2)), g_order=2)
    op.assemble(1)
    x = spla.spsolve(op.csr().tocsc(), op.rhs())
    Xe = m.coords()[m.conn()]  # [ne, local vertex, 3]
    gi = s.global_indices()
    if geom == O.TET:
        pairs = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
        nodes = np.concatenate([Xe, np.stack([(Xe[:, a] + Xe[:, b]) / 2 for a, b in pairs], 1)], axis=1)
    else:
        x0 = Xe[:, 0]
        E = np.stack([Xe[:, 1] - x0, Xe[:, 2] - x0, Xe[:, 4] - x0], 1)  # edge vectors along x, y, z
        ref = np.array([[t % 3, (t // 3) % 3, t // 9] for t in range(27)]) / 2.0
        nodes = x0[:, None, :] + np.einsum("ta,eab->etb", ref, E)
    err = np.abs(x[gi] - u(nodes)).max()
    assert err <= 1e-9, err


@pytest.mark.parametrize("geom,cells", [(O.TRI, [4, 3]), (O.QUAD, [3, 3])])
def test_adjoint_gradient_order2_central_differences(geom, cells):
    """the adjoint gradient on P2 triangles and Q2 quadrilaterals: -kappa_e lambda_e^T Khat_e u_e with the
    order-2 local Laplace matrix, against central differences of J = 1/2 |u - u_obs|^2"""
    rng = np.random.default_rng(11)
    m = O.Mesh.structured(geom, cells, [0, 0], [1, 1])
    s = O.Space(m, True, 2)
    ne = m.ne

    def solve(logk):
        op = O.Operator(s, O.STENCIL_ELEMENT)
        op.append_laplace(np.exp(logk))
        op.append_source(1.0)
        op.set_dirichlet(True)
        op.assemble(1)
        rp, cols = op.pattern()
        x, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-15, max_iter=10000)
        return x, (rp, cols, op.values())

    g0 = rng.standard_normal(ne)
    u_obs, _ = solve(rng.standard_normal(ne))
    u, (rp, cols, vals) = solve(g0)
    lam, _, _, _ = O.pcg_jacobi(rp, cols, vals, u - u_obs, rtol=1e-15, max_iter=10000)
    grad = O.adjoint_log_kappa_gradient(s, np.exp(g0), u, lam)
    h = 1e-6
    gmax = np.abs(grad).max()
    for e in rng.choice(ne, 6, replace=False):
        gp, gm = g0.copy(), g0.copy()
        gp[e] += h
        gm[e] -= h
        up, _ = solve(gp)
        um, _ = solve(gm)
        fd = (0.5 * np.sum((up - u_obs) ** 2) - 0.5 * np.sum((um - u_obs) ** 2)) / (2 * h)
        assert abs(fd - grad[e]) <= 1e-6 * gmax, (e, fd, grad[e])


@pytest.mark.parametrize("only_clear,ensure_symmetry", [(False, True), (False, False), (True, True), (True, False)])
def test_dirichlet_apply_variants(only_clear, ensure_symmetry):
    """DirichletConstraints::apply(matrix, vector, only_clear, ensure_symmetry) (dirichlet-constraints.hh:152-184)
    replayed on the dense unconstrained matrix DoF by DoF as the reference does it: [clear|unit]_col (with
    ensure_symmetry), then [clear|unit]_row, vector entry 0; the oracle's constrained matrix and rhs equal it"""
    import scipy.sparse as sp

    m = O.Mesh.structured(O.QUAD, [5, 4], [0, 0], [1, 1])
    kap = 1.0 + 0.2 * np.arange(m.ne)

    def build(dirichlet):
        s = O.Space(m, True, 1)
        op = O.Operator(s, O.STENCIL_ELEMENT)
        op.append_laplace(kap)
        op.append_source(1.0)
        op.set_dirichlet(dirichlet)
        op.set_dirichlet_mode(only_clear, ensure_symmetry)
        op.assemble(1)
        rp, cols = op.pattern()
        return sp.csr_matrix((op.values(), cols, rp), shape=(len(rp) - 1,) * 2).toarray(), op.rhs(), op

    A0, b0, _ = build(False)
    A1, b1, op = build(True)
    dofs = op.dirichlet_dofs()
    assert len(dofs) == 2 * 6 + 2 * 3
    A, b = A0.copy(), b0.copy()
    pattern = A0 != 0
    for d in dofs:
        if ensure_symmetry:
            A[:, d] = 0.0
            if not only_clear:
                A[d, d] = 1.0
        A[d, :] = 0.0
        if not only_clear:
            A[d, d] = 1.0
        b[d] = 0.0
    assert np.array_equal(A1[pattern], A[pattern]) and np.array_equal(b1, b)
