"""Pins of the oracle's element forms whose coefficient varies inside the elements (LocalLaplaceIntegrand /
LocalElementProductIntegrand with a GridFunction of order c: laplace.hh:74-101, product.hh:88-128), given at
the points of the element rules of order c + 2 k + over_integrate.

* values equal at every point of an element reproduce the per-element form bit for bit (layout, rule order);
* P1 on triangles and tetrahedra with a linear diffusion: the gradients are constant, so the element matrix is
  (integral of kappa) G_i . G_j = |T| kappa(centroid) G_i . G_j, the per-element form with kappa at the
  centroids (to rounding);
* Q1 on quadrilaterals and hexahedra, P2 on triangles: every element matrix against a brute-force tensor /
  Dunavant-free evaluation with an independent 5-point Gauss rule per direction on the reference cube
  (the integrands are polynomials of degree <= 4 per direction);
* linearity in the coefficient, symmetry and positive definiteness (with Dirichlet rows).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O


def _op(m, cg=True, k=1):
    s = O.Space(m, cg, k)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    return s, op


def _dense(op):
    op.assemble(1)
    rp, cols = op.pattern()
    return sp.csr_matrix((op.values(), cols, rp), shape=(len(rp) - 1, len(rp) - 1)).toarray()


@pytest.mark.parametrize("geom,cells,k", [(O.QUAD, [4, 3], 1), (O.TRI, [3, 3], 2), (O.HEX, [2, 2, 2], 1),
                                          (O.TET, [2, 2, 1], 1)])
def test_constant_points_equal_per_element(geom, cells, k):
    m = O.Mesh.structured(geom, cells, [0] * len(cells), [1] * len(cells))
    kap = 1.0 + np.arange(m.ne) * 0.37
    _, a = _op(m, True, k)
    a.append_laplace(kap)
    a.append_product(kap[::-1].copy(), scaling=0.5)
    order = 0 + 2 * k
    off = m.quadrature_offsets(order)
    per_pt = np.repeat(kap, np.diff(off))
    per_pt_w = np.repeat(kap[::-1], np.diff(off))
    _, b = _op(m, True, k)
    b.append_laplace_points(per_pt, 0)
    b.append_product_points(per_pt_w, 0, scaling=0.5)
    assert np.array_equal(_dense(a), _dense(b))


@pytest.mark.parametrize("geom,cells", [(O.TRI, [4, 3]), (O.TET, [2, 2, 2])])
def test_linear_diffusion_on_simplices_is_centroid_diffusion(geom, cells):
    d = len(cells)
    m = O.Mesh.structured(geom, cells, [0] * d, [1] * d)
    coef = np.array([1.0, 2.0, -0.7, 0.4])[: d + 1]

    def kappa(X):
        return coef[0] + X[..., :d] @ coef[1:]

    X = m.quadrature_points(1 + 2)
    _, a = _op(m)
    a.append_laplace_points(kappa(X), 1)
    cent = m.coords()[m.conn()].mean(axis=1)
    _, b = _op(m)
    b.append_laplace(kappa(cent))
    A, B = _dense(a), _dense(b)
    assert np.abs(A - B).max() <= 1e-13 * np.abs(B).max()
    # a per-point mistake (the points of an element reversed) is seen
    Xr = X[:, ::-1, :]
    _, c = _op(m)
    c.append_laplace_points(kappa(Xr) * (1 + 0.1 * np.arange(X.shape[1])), 1)
    assert np.abs(_dense(c) - B).max() > 1e-6


def _q_basis(d, xh):
    """Q1 values and reference gradients at xh (vertex b: bits of b = directions)"""
    nv = 1 << d
    phi, dphi = np.ones(nv), np.ones((nv, d))
    for b in range(nv):
        for a in range(d):
            hi = (b >> a) & 1
            phi[b] *= xh[a] if hi else 1 - xh[a]
            for c in range(d):
                dphi[b, c] *= (1.0 if hi else -1.0) if c == a else (xh[a] if hi else 1 - xh[a])
    return phi, dphi


@pytest.mark.parametrize("geom,cells", [(O.QUAD, [3, 2]), (O.HEX, [2, 1, 2])])
def test_q1_element_matrices_against_brute_force(geom, cells):
    """kappa(x) = 1 + x^2 - 0.5 x y (order 2): the assembled matrix equals the sum of element matrices
    integrated with a 5-point Gauss rule per direction (exact for these degrees), independent of the oracle's
    rules"""
    d = len(cells)
    lo, hi = np.zeros(d), np.array([1.0, 0.8, 1.2][:d])
    m = O.Mesh.structured(geom, cells, lo.tolist(), hi.tolist())

    def kappa(X):
        return 1.0 + X[..., 0] ** 2 - 0.5 * X[..., 0] * X[..., 1]

    _, a = _op(m, True, 1)
    a.append_laplace_points(kappa(m.quadrature_points(2 + 2)), 2)
    a.append_product_points(kappa(m.quadrature_points(2 + 2)), 2, scaling=3.0)
    A = _dense(a)
    h = (hi - lo) / np.array(cells)
    g, w = np.polynomial.legendre.leggauss(5)
    g, w = (g + 1) / 2, w / 2
    ref = np.zeros_like(A)
    conn, X0 = m.conn(), m.coords()
    for e in range(m.ne):
        x0 = X0[conn[e]].min(axis=0)
        Ke = np.zeros((1 << d, 1 << d))
        for idx in np.ndindex(*([5] * d)):
            xh = np.array([g[i] for i in idx])
            wt = np.prod([w[i] for i in idx]) * np.prod(h)
            phi, dphi = _q_basis(d, xh)
            G = dphi / h
            kv = kappa(x0 + xh * h)
            Ke += wt * kv * (G @ G.T + 3.0 * np.outer(phi, phi))
        ids = conn[e]
        ref[np.ix_(ids, ids)] += Ke
    assert np.abs(A - ref).max() <= 1e-13 * np.abs(ref).max()


def test_linearity_symmetry_and_spd():
    m = O.Mesh.structured(O.TRI, [5, 4], [0, 0], [1, 1])
    X = m.quadrature_points(2 + 2)
    k1 = 1.0 + X[..., 0] ** 2
    k2 = 0.5 + X[..., 0] * X[..., 1]

    def mat(v):
        s, op = _op(m)
        op.append_laplace_points(v, 2)
        op.set_dirichlet(True)
        return _dense(op)

    A1, A2, A12 = mat(k1), mat(k2), mat(k1 + k2)
    n = A1.shape[0]
    interior = ~np.isclose(np.diag(A1), 1.0) | (np.abs(A1).sum(axis=1) != 1.0)
    D = A1 + A2 - A12
    assert np.abs(D[np.ix_(interior, interior)]).max() <= 1e-13 * np.abs(A12).max()
    assert np.abs(A12 - A12.T).max() <= 1e-15 * np.abs(A12).max()
    assert np.linalg.eigvalsh(A12).min() > 0 and n == m.nv


def test_wrong_point_count_is_refused():
    m = O.Mesh.structured(O.QUAD, [2, 2], [0, 0], [1, 1])
    _, op = _op(m)
    with pytest.raises(RuntimeError):
        op.append_laplace_points(np.ones(7), 0)
