"""Pins of the oracle's Laplace form with a matrix-valued diffusion K (LocalLaplaceIntegrand(GridFunction<E, d,
d>): (K grad phi_j) . grad psi_i, laplace.hh:94-101).

* K = kappa_e I per element reproduces the scalar per-element form to rounding (the scalar-to-matrix
  promotion; the two sums contract into fused multiply-adds differently);
* rotating the mesh by R and the tensor to R K R^T leaves the matrix unchanged (the integrand is invariant);
* Q1 element matrices with a constant anisotropic K against a brute-force 5-point Gauss rule per direction;
* with per-element SPD tensors the matrix is symmetric and, with Dirichlet rows, positive definite.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O


def _dense(m, append, k=1, dirichlet=False):
    s = O.Space(m, True, k)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    append(op)
    op.set_dirichlet(dirichlet)
    op.assemble(1)
    rp, cols = op.pattern()
    return sp.csr_matrix((op.values(), cols, rp), shape=(len(rp) - 1,) * 2).toarray()


def _spd_tensors(rng, ne, d):
    B = rng.standard_normal((ne, d, d))
    return np.einsum("eab,ecb->eac", B, B) + 0.5 * np.eye(d)


@pytest.mark.parametrize("geom,cells,k", [(O.QUAD, [4, 3], 1), (O.TRI, [3, 3], 2), (O.HEX, [2, 2, 1], 1),
                                          (O.TET, [2, 1, 2], 2)])
def test_isotropic_tensor_is_the_scalar_form(geom, cells, k):
    d = len(cells)
    m = O.Mesh.structured(geom, cells, [0] * d, [1] * d)
    kap = 1.0 + 0.3 * np.arange(m.ne)
    A = _dense(m, lambda op: op.append_laplace(kap), k)
    B = _dense(m, lambda op: op.append_laplace_tensor(kap[:, None, None] * np.eye(d)), k)
    assert np.abs(A - B).max() <= 4e-16 * np.abs(A).max()


@pytest.mark.parametrize("geom", [O.TRI, O.QUAD])
def test_rotation_invariance(geom):
    rng = np.random.default_rng(3)
    m0 = O.Mesh.structured(geom, [3, 3], [0, 0], [1, 1])
    X, conn = m0.coords(), m0.conn()
    th = 0.7
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    m1 = O.Mesh.explicit(geom, X @ R.T, conn)
    K = _spd_tensors(rng, m0.ne, 2)
    KR = np.einsum("ab,ebc,dc->ead", R, K, R)
    A0 = _dense(m0, lambda op: op.append_laplace_tensor(K))
    A1 = _dense(m1, lambda op: op.append_laplace_tensor(KR))
    assert np.abs(A0 - A1).max() <= 1e-13 * np.abs(A0).max()
    # a tensor not rotated along gives another matrix
    A2 = _dense(m1, lambda op: op.append_laplace_tensor(K))
    assert np.abs(A0 - A2).max() > 1e-3 * np.abs(A0).max()


@pytest.mark.parametrize("geom,cells", [(O.QUAD, [3, 2]), (O.HEX, [2, 1, 2])])
def test_q1_anisotropic_against_brute_force(geom, cells):
    d = len(cells)
    lo, hi = np.zeros(d), np.array([1.0, 0.8, 1.2][:d])
    K = np.array([[2.0, 0.5, -0.2], [0.5, 1.0, 0.3], [-0.2, 0.3, 1.5]])[:d, :d]
    m = O.Mesh.structured(geom, cells, lo.tolist(), hi.tolist())
    A = _dense(m, lambda op: op.append_laplace_tensor(K))
    h = (hi - lo) / np.array(cells)
    g, w = np.polynomial.legendre.leggauss(5)
    g, w = (g + 1) / 2, w / 2
    ref = np.zeros_like(A)
    conn = m.conn()
    for e in range(m.ne):
        Ke = np.zeros((1 << d, 1 << d))
        for idx in np.ndindex(*([5] * d)):
            xh = np.array([g[i] for i in idx])
            wt = np.prod([w[i] for i in idx]) * np.prod(h)
            G = np.empty((1 << d, d))
            for b in range(1 << d):
                for a in range(d):
                    v = 1.0 / h[a] * (1.0 if (b >> a) & 1 else -1.0)
                    for c in range(d):
                        if c != a:
                            v *= xh[c] if (b >> c) & 1 else 1 - xh[c]
                    G[b, a] = v
            Ke += wt * (G @ K.T @ G.T).T  # entry (i, j): (K G_j) . G_i
        ref[np.ix_(conn[e], conn[e])] += Ke
    assert np.abs(A - ref).max() <= 1e-13 * np.abs(ref).max()


def test_spd_tensors_give_symmetric_positive_definite_matrix():
    rng = np.random.default_rng(9)
    m = O.Mesh.structured(O.TET, [2, 2, 2], [0, 0, 0], [1, 1, 1])
    A = _dense(m, lambda op: op.append_laplace_tensor(_spd_tensors(rng, m.ne, 3)), dirichlet=True)
    assert np.abs(A - A.T).max() <= 1e-14 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > 0
