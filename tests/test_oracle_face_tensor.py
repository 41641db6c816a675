"""Pins of the oracle's SWIPDG intersection forms with a matrix-valued diffusion and weight (laplace-ipdg.hh:
47-61,145-186: delta = n . (W n), fluxes (K grad phi) . n; ipdg.hh:59-72,149-153).

* K = kappa_e I and W = omega_e I reproduce the scalar forms (to an ulp: the sums contract differently);
* rotating the mesh by R and both tensors to R T R^T leaves the SWIPDG matrix unchanged; unrotated tensors
  change it;
* symmetric prefactor 1 with SPD tensors and a large penalty: symmetric positive definite.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle_lib as O


def _spd(seed, ne, d):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((ne, d, d))
    return np.einsum("eab,ecb->eac", B, B) + 0.5 * np.eye(d)


def _swipdg(m, K, W, sigma=20.0, tensors=True):
    s = O.Space(m, False, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    k = np.ones(m.ne) if tensors else K
    w = np.ones(m.ne) if tensors else W
    if tensors:
        op.append_laplace_tensor(K)
    else:
        op.append_laplace(k)
    op.append_inner_coupling(1.0, k, w)
    if tensors:
        op.set_face_tensors(False, K, W)
    op.append_inner_penalty(sigma, w)
    if tensors:
        op.set_face_tensors(False, None, W)
    op.append_dirichlet_coupling(1.0, k)
    if tensors:
        op.set_face_tensors(True, K, None)
    op.append_boundary_penalty(sigma, w)
    if tensors:
        op.set_face_tensors(True, None, W)
    op.assemble(1)
    rp, cols = op.pattern()
    return sp.csr_matrix((op.values(), cols, rp), shape=(len(rp) - 1,) * 2).toarray()


@pytest.mark.parametrize("geom,cells", [(O.QUAD, [3, 3]), (O.TRI, [3, 2]), (O.HEX, [2, 2, 1]), (O.TET, [1, 2, 1])])
def test_isotropic_tensors_are_the_scalar_forms(geom, cells):
    d = len(cells)
    m = O.Mesh.structured(geom, cells, [0] * d, [1] * d)
    kap = 1.0 + 0.3 * np.arange(m.ne)
    om = 2.0 - 0.05 * np.arange(m.ne)
    A = _swipdg(m, kap, om, tensors=False)
    B = _swipdg(m, kap[:, None, None] * np.eye(d), om[:, None, None] * np.eye(d))
    assert np.abs(A - B).max() <= 4e-16 * np.abs(A).max() * 4


@pytest.mark.parametrize("geom", [O.TRI, O.QUAD])
def test_rotation_invariance(geom):
    m0 = O.Mesh.structured(geom, [3, 3], [0, 0], [1, 1])
    th = 0.6
    R = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    m1 = O.Mesh.explicit(geom, m0.coords() @ R.T, m0.conn())
    K, W = _spd(1, m0.ne, 2), _spd(2, m0.ne, 2)
    rot = lambda T: np.einsum("ab,ebc,dc->ead", R, T, R)
    A0 = _swipdg(m0, K, W)
    A1 = _swipdg(m1, rot(K), rot(W))
    assert np.abs(A0 - A1).max() <= 1e-12 * np.abs(A0).max()
    A2 = _swipdg(m1, K, W)
    assert np.abs(A0 - A2).max() > 1e-3 * np.abs(A0).max()


def test_spd_with_tensors():
    m = O.Mesh.structured(O.TRI, [3, 3], [0, 0], [1, 1])
    A = _swipdg(m, _spd(3, m.ne, 2), _spd(4, m.ne, 2), sigma=60.0)
    assert np.abs(A - A.T).max() <= 1e-13 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > 0


@pytest.mark.parametrize("geom,cells", [(O.QUAD, [3, 3]), (O.TRI, [3, 3]), (O.HEX, [2, 2, 2]), (O.TET, [2, 2, 1])])
def test_linear_solution_with_anisotropic_tensor(geom, cells):
    """SWIPDG with a constant anisotropic K as diffusion and weight, f = 0 and weak Dirichlet data g = u for a
    linear u (then -div(K grad u) = 0): the discrete solution is u at every DoF (the forms and the boundary
    functionals with tensors are consistent)"""
    import scipy.sparse.linalg as spla

    d = len(cells)
    m = O.Mesh.structured(geom, cells, [0] * d, [1] * d)
    Kc = np.array([[2.0, 0.6, -0.3], [0.6, 1.0, 0.2], [-0.3, 0.2, 1.5]])[:d, :d]
    Kt = np.ascontiguousarray(np.broadcast_to(Kc, (m.ne, d, d)))
    cvec = np.array([0.7, -1.1, 0.4])[:d]

    def u(P):
        return 0.3 + P @ cvec

    sigma, one = 30.0, np.ones(m.ne)
    s = O.Space(m, False, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_laplace_tensor(Kt)
    op.append_inner_coupling(1.0, one, one)
    op.set_face_tensors(False, Kt, Kt)
    op.append_inner_penalty(sigma, one)
    op.set_face_tensors(False, None, Kt)
    op.append_dirichlet_coupling(1.0, one)
    op.set_face_tensors(True, Kt, None)
    op.append_boundary_penalty(sigma, one)
    op.set_face_tensors(True, None, Kt)
    over = 1 if geom in (O.QUAD, O.HEX) else 0  # (K grad psi) . n g is of degree 2 along a cube's face
    op.append_boundary_source(0, 1.0, 1.0, u(m.boundary_quadrature_points(1 + over)), g_order=1,
                              over_integrate=over)
    op.set_boundary_source_tensor(Kt)
    op.append_boundary_source(1, sigma, 1.0, u(m.boundary_quadrature_points(2)), g_order=1)
    op.set_boundary_source_tensor(Kt)
    op.assemble(1)
    x = spla.spsolve(op.csr().tocsc(), op.rhs())
    nodes = m.coords()[m.conn()]
    err = np.abs(x[s.global_indices()] - u(nodes)).max()
    assert err <= 1e-10, err
