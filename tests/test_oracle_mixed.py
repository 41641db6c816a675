"""Pins of the oracle on mixed triangle/quadrilateral meshes (CPU only).

The mixed mesh path dispatches every element to its own reference element and local finite element
(spaces/mapper/continuous.hh:117-182, mapper/discontinuous.hh:102-125).  Checked against what the mathematics
fixes, independently of the oracle's own formulas:
  * topology: intersections and DoF counts counted from the element list;
  * additivity: the matrix of a mixed mesh is the sum of the matrices of its quadrilateral part and its
    triangle part, each assembled as a homogeneous mesh (pinned by test_oracle.py) over the same vertices;
  * closed forms: the mass matrix sums to the area (the Lagrange basis is a partition of unity); the
    stiffness matrix maps the interpolant of a linear function (orders 1, 2) and of a quadratic (order 2) to
    -Laplace(u) times the lumped mass in every interior row, and u^T K u = int |grad u|^2;
  * the element-ordered quadrature-point layout: the L2 error of the interpolant of a linear function
    against its own values at the points is zero, and without them it equals u^T M u;
  * intersections between a triangle and a quadrilateral: the penalty form of a two-element grid in closed
    form (face mass of the linear traces), SWIPDG with weak Dirichlet data reproducing a linear (order 1) and
    a quadratic (order 2) solution on a sheared mixed grid, and the SWIPDG matrix symmetric positive definite.
"""
import numpy as np
import pytest

import oracle_lib as O
from mixed_grids import grid_area, refined_reference_mixed_grid, strip_mixed_grid


def grids():
    return {"reference": refined_reference_mixed_grid(), "strip": strip_mixed_grid(6, 5, shear=0.3)}


def edge_counts(E):
    edges = {}
    for el in E:
        loc = [(0, 2), (1, 3), (0, 1), (2, 3)] if len(el) == 4 else [(0, 1), (0, 2), (1, 2)]
        for a, b in loc:
            k = tuple(sorted((el[a], el[b])))
            edges[k] = edges.get(k, 0) + 1
    return edges


def dof_coordinates(mesh, V, E, space, order):
    """physical position of every DoF of a continuous space (from its global indices)"""
    gi = space.global_indices()
    X = np.full((space.size, 2), np.nan)
    for e, el in enumerate(E):
        p = V[el]
        a, b = p[1] - p[0], p[2] - p[0]
        if len(el) == 4:
            ref = [(i / order, j / order) for j in range(order + 1) for i in range(order + 1)]
        else:
            ref = [(0., 0.), (1., 0.), (0., 1.)]
            if order == 2:
                ref += [(.5, 0.), (0., .5), (.5, .5)]
        for i, (xi, eta) in enumerate(ref):
            x = p[0] + xi * a + eta * b
            d = gi[e, i]
            assert np.all(np.isnan(X[d])) or np.allclose(X[d], x, atol=1e-14)
            X[d] = x
    return X


@pytest.mark.parametrize("name", ["reference", "strip"])
def test_mixed_topology_and_dof_counts(name):
    V, E = grids()[name]
    m = O.Mesh.mixed(V, E)
    assert m.is_mixed and m.ne == len(E) and m.nv == len(V)
    nb = m.neighbors()
    edges = edge_counts(E)
    assert int(np.sum(nb == -1)) == sum(1 for c in edges.values() if c == 1)
    assert int(np.sum(nb >= 0)) == 2 * sum(1 for c in edges.values() if c == 2)
    # absent faces: the fourth face of every triangle
    assert int(np.sum(nb == -2)) == sum(1 for el in E if len(el) == 3)
    for e in range(m.ne):
        for o in nb[e]:
            if o >= 0:
                assert e in nb[o]
    nq = sum(1 for el in E if len(el) == 4)
    nt = len(E) - nq
    assert O.Space(m, True, 1).size == len(V)
    assert O.Space(m, True, 2).size == len(V) + len(edges) + nq
    assert O.Space(m, False, 1).size == 4 * nq + 3 * nt
    assert O.Space(m, False, 2).size == 9 * nq + 6 * nt
    gi = O.Space(m, False, 2).global_indices()
    assert gi.shape == (m.ne, 9)
    assert all((gi[e, 6:] == -1).all() for e in range(m.ne) if len(E[e]) == 3)


@pytest.mark.parametrize("name", ["reference", "strip"])
@pytest.mark.parametrize("continuous,order", [(True, 1), (True, 2), (False, 1), (False, 2)])
def test_mixed_mass_sums_to_area(name, continuous, order):
    V, E = grids()[name]
    m = O.Mesh.mixed(V, E)
    s = O.Space(m, continuous, order)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    op.append_product(1.0)
    op.assemble()
    assert abs(op.values().sum() - grid_area(V, E)) < 1e-14 * len(E)


def test_mixed_is_sum_of_its_parts():
    V, E = strip_mixed_grid(8, 6, shear=0.2)
    rng = np.random.default_rng(7)
    kappa = np.exp(rng.normal(size=len(E)))
    quads = [e for e, el in enumerate(E) if len(el) == 4]
    tris = [e for e, el in enumerate(E) if len(el) == 3]
    m = O.Mesh.mixed(V, E)
    s = O.Space(m, True, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    op.append_laplace(kappa)
    op.append_product(2.0)
    op.assemble()
    A = op.csr().toarray()
    total = np.zeros_like(A)
    for geom, ids in ((O.QUAD, quads), (O.TRI, tris)):
        mp = O.Mesh.explicit(geom, V, np.array([E[e] for e in ids]))
        sp_ = O.Space(mp, True, 1)
        o = O.Operator(sp_, O.STENCIL_ELEMENT)
        o.append_laplace(kappa[ids])
        o.append_product(2.0)
        o.assemble()
        total += o.csr().toarray()
    assert np.abs(A - total).max() <= 1e-14 * np.abs(A).max()


@pytest.mark.parametrize("order", [1, 2])
def test_mixed_stiffness_closed_forms(order):
    V, E = strip_mixed_grid(6, 6, shear=0.25)
    m = O.Mesh.mixed(V, E)
    s = O.Space(m, True, order)
    X = dof_coordinates(m, V, E, s, order)
    K = O.Operator(s, O.STENCIL_ELEMENT)
    K.append_laplace(1.0)
    K.assemble()
    K = K.csr()
    M = O.Operator(s, O.STENCIL_ELEMENT)
    M.append_product(1.0)
    M.assemble()
    M = M.csr()
    d = O.Operator(s, O.STENCIL_ELEMENT)
    d.append_laplace(1.0)
    d.set_dirichlet(True)
    d.assemble()
    boundary = np.zeros(s.size, bool)
    boundary[d.dirichlet_dofs()] = True
    x, y = X[:, 0], X[:, 1]
    area = grid_area(V, E)
    u = 0.3 + 1.7 * x - 0.6 * y  # linear: interpolated exactly by orders 1 and 2
    Ku = K @ u
    assert np.abs(Ku[~boundary]).max() < 1e-12
    assert abs(u @ Ku - (1.7 ** 2 + 0.6 ** 2) * area) < 1e-12
    if order == 2:
        # u = x^2 + 2 y^2 - x y: -Laplace(u) = -6; the parallelogram maps are affine, so Q2 / P2 hold it exactly
        u = x ** 2 + 2 * y ** 2 - x * y
        Ku = K @ u
        lumped = M @ np.ones(s.size)
        assert np.abs(Ku[~boundary] - (-6.0) * lumped[~boundary]).max() < 1e-11


def test_mixed_error_norm_layout():
    V, E = strip_mixed_grid(5, 4, shear=0.1)
    m = O.Mesh.mixed(V, E)
    s = O.Space(m, True, 1)
    X = dof_coordinates(m, V, E, s, 1)
    u = 1.0 + 2.0 * X[:, 0] + 0.5 * X[:, 1]
    order = 3
    pts = m.quadrature_points(order)
    off = m.quadrature_offsets(order)
    assert pts.shape == (off[-1], 2)
    # triangles and quadrilaterals have rules of different sizes
    sizes = np.diff(off)
    assert len(set(sizes[[i for i, el in enumerate(E) if len(el) == 4]])) == 1
    assert len(set(sizes.tolist())) == 2
    g = 1.0 + 2.0 * pts[:, 0] + 0.5 * pts[:, 1]
    gg = np.tile([2.0, 0.5], (len(g), 1))
    err = O.error_norms(s, u, order, g, gg)
    assert abs(err[0]) < 1e-26 and abs(err[1]) < 1e-26
    M = O.Operator(s, O.STENCIL_ELEMENT)
    M.append_product(1.0)
    M.assemble()
    norms = O.error_norms(s, u, order)
    assert abs(norms[0] - u @ (M.csr() @ u)) < 1e-13
    assert abs(norms[1] - (2.0 ** 2 + 0.5 ** 2) * grid_area(V, E)) < 1e-12


def element_nodes(V, E, order):
    """Lagrange node positions of every element's local DoFs (cubes: tensor index, simplices: the vertices,
    then the edge midpoints (a, b), a < b)"""
    out = []
    for el in E:
        p = V[el]
        a, b = p[1] - p[0], p[2] - p[0]
        if len(el) == 4:
            ref = [(i / order, j / order) for j in range(order + 1) for i in range(order + 1)]
        else:
            ref = [(0., 0.), (1., 0.), (0., 1.)] + ([(.5, 0.), (0., .5), (.5, .5)] if order == 2 else [])
        out.append(np.array([p[0] + xi * a + eta * b for xi, eta in ref]))
    return out


def test_mixed_face_penalty_closed_form():
    """one quadrilateral [0,1]^2 and one triangle (1,0), (2,0.5), (1,1) sharing the edge x = 1 (length and
    diameter 1), inner penalty sigma with unit weight only (ipdg.hh:149-153: the half harmonic average of the
    weights, 1/2): the face mass of the linear traces is (1/6)[[2, 1], [1, 2]], so the quadrilateral's trace
    DoFs (local 1 at (1,0), 3 at (1,1)) get sigma/6 and sigma/12, the triangle's (local 0 at (1,0), 2 at (1,1))
    the same, and the coupling block -sigma/6 between DoFs at the same vertex, -sigma/12 across; every other
    entry 0"""
    V = np.array([(0., 0.), (1., 0.), (0., 1.), (1., 1.), (2., 0.5)])
    E = [[0, 1, 2, 3], [1, 4, 3]]
    m = O.Mesh.mixed(V, E)
    s = O.Space(m, False, 1)
    sigma = 7.0
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_inner_penalty(sigma, 1.0)
    op.assemble()
    A = op.csr().toarray()
    R = np.zeros((7, 7))
    q, t = [1, 3], [4 + 0, 4 + 2]  # trace DoFs at (1, 0) and (1, 1)
    M = 0.5 * np.array([[2., 1.], [1., 2.]]) / 6
    for a in range(2):
        for b in range(2):
            R[q[a], q[b]] += sigma * M[a, b]
            R[t[a], t[b]] += sigma * M[a, b]
            R[q[a], t[b]] -= sigma * M[a, b]
            R[t[a], q[b]] -= sigma * M[a, b]
    assert A.shape == (7, 7) and np.abs(A - R).max() <= 1e-14


def _swipdg_mixed(V, E, order, u, f, sigma):
    m = O.Mesh.mixed(V, E)
    s = O.Space(m, False, order)
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_laplace(1.0)
    op.append_inner_coupling(1.0, 1.0, 1.0)
    op.append_inner_penalty(sigma, 1.0)
    op.append_dirichlet_coupling(1.0, 1.0)
    op.append_boundary_penalty(sigma, 1.0)
    if f != 0.0:
        op.append_source(f)
    # the quadrilaterals' coupling functional needs one order more than the reference's to be exact
    op.append_boundary_source(0, 1.0, 1.0, u(m.boundary_quadrature_points(2 * order)), g_order=order,
                              over_integrate=1)
    op.append_boundary_source(1, sigma, 1.0, u(m.boundary_quadrature_points(2 * order)), g_order=order)
    op.assemble(1)
    return m, s, op


@pytest.mark.parametrize("order", [1, 2])
def test_mixed_swipdg_reproduces_polynomial_solutions(order):
    """SWIPDG with weak Dirichlet data on a sheared mixed grid reproduces a solution in the discrete space:
    u = 0.3 + 2x - 1.5y (order 1) and u = x^2 - xy + y^2/2 + 0.3x with -Laplace(u) = -3 (order 2); the faces
    between triangles and quadrilaterals evaluate each side's basis at the face points"""
    import scipy.sparse.linalg as spla

    V, E = strip_mixed_grid(5, 4, shear=0.2)
    if order == 1:
        u = lambda P: 0.3 + 2.0 * P[..., 0] - 1.5 * P[..., 1]
        f = 0.0
    else:
        u = lambda P: P[..., 0] ** 2 - P[..., 0] * P[..., 1] + 0.5 * P[..., 1] ** 2 + 0.3 * P[..., 0]
        f = -3.0
    m, s, op = _swipdg_mixed(V, E, order, u, f, 40.0)
    x = spla.spsolve(op.csr().tocsc(), op.rhs())
    gi = s.global_indices()
    nodes = element_nodes(V, E, order)
    err = max(np.abs(x[gi[e, :len(nodes[e])]] - u(nodes[e])).max() for e in range(len(E)))
    assert err <= 1e-10, err


def test_mixed_swipdg_is_spd():
    """the SWIPDG matrix (symmetric prefactor 1, sigma = 16) on the reference's mixed grid is symmetric and
    positive definite"""
    V, E = refined_reference_mixed_grid()
    m = O.Mesh.mixed(V, E)
    s = O.Space(m, False, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    op.append_laplace(1.0)
    op.append_inner_coupling(1.0, 1.0, 1.0)
    op.append_inner_penalty(16.0, 1.0)
    op.append_dirichlet_coupling(1.0, 1.0)
    op.append_boundary_penalty(16.0, 1.0)
    op.assemble()
    A = op.csr().toarray()
    assert np.abs(A - A.T).max() <= 1e-13 * np.abs(A).max()
    assert np.linalg.eigvalsh(A).min() > 0


@pytest.mark.parametrize("order", [1, 2])
def test_adjoint_gradient_on_mixed_mesh_central_differences(order):
    """the adjoint gradient on a mixed mesh (each element's own local Laplace matrix) against central
    differences of J = 1/2 |u - u_obs|^2"""
    import scipy.sparse  # noqa: F401  (the solves below are the oracle's PCG)

    rng = np.random.default_rng(13)
    V, E = strip_mixed_grid(4, 3, shear=0.3)
    m = O.Mesh.mixed(V, E)
    s = O.Space(m, True, order)
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
    tri = [e for e in range(ne) if len(E[e]) == 3]
    quad = [e for e in range(ne) if len(E[e]) == 4]
    for e in list(rng.choice(tri, 3, replace=False)) + list(rng.choice(quad, 3, replace=False)):
        gp, gm = g0.copy(), g0.copy()
        gp[e] += h
        gm[e] -= h
        up, _ = solve(gp)
        um, _ = solve(gm)
        fd = (0.5 * np.sum((up - u_obs) ** 2) - 0.5 * np.sum((um - u_obs) ** 2)) / (2 * h)
        assert abs(fd - grad[e]) <= 1e-6 * gmax, (e, fd, grad[e])
