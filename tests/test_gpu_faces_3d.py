"""Intersection integrals on 3-d grids, and of order-2 spaces, on the HIP path (integrals.hh:197-267,338-369 over the faces of
hexahedra and tetrahedra; boundary functionals, local/functionals/integrals.hh:157-185).

Bar: pattern bit-exact, matrix and right-hand side within 1e-12 (row-relative / max-relative) of the oracle,
Jacobi-PCG solution within 1e-10 of the oracle's with its iteration count (+-2); the linear solution with
weak Dirichlet data reproduced to 1e-9; the Dirichlet forms restricted to the x faces by a boundary info
(ApplyOn::CustomBoundaryIntersections) give the oracle's matrix with the same face mask.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from problems import matrix_error

pytestmark = pytest.mark.gpu

GEOMS = [("hex", O.HEX, [5, 4, 3]), ("tet", O.TET, [3, 3, 2])]


def _hgdt_swipdg(H, grid, kappa, sigma, u, over, on=None):
    space = H.DiscontinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element_and_intersection)
    kap = torch.from_numpy(kappa).cuda()
    op += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(kap))
    op.append(H.LocalCouplingIntersectionIntegralBilinearForm(
        H.LocalLaplaceIPDGInnerCoupling(1.0, kap, kap) + H.LocalIPDGInnerPenalty(sigma, kap)), {},
        H.ApplyOn.InnerIntersectionsOnce())
    where = on if on is not None else H.ApplyOn.BoundaryIntersections()
    op.append(H.LocalIntersectionIntegralBilinearForm(
        H.LocalIPDGBoundaryPenalty(sigma, kap) + H.LocalLaplaceIPDGDirichletCoupling(1.0, kap)), {}, where)
    g = H.GenericFunction(1, u)
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalLaplaceIPDGDirichletCoupling(1.0, kap, g),
                                                     over_integrate=over), {}, where)
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalIPDGBoundaryPenalty(sigma, kap).with_ansatz(g)), {},
               where)
    op.append(rhs)
    op.assemble()
    return space, op, kap


def _oracle_swipdg(m, kappa, sigma, u, over, mask=None):
    s = O.Space(m, False, 1)
    oop = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    oop.append_laplace(kappa)
    oop.append_inner_coupling(1.0, kappa, kappa)
    oop.append_inner_penalty(sigma, kappa)
    oop.set_boundary_filter(mask)
    oop.append_boundary_penalty(sigma, kappa)
    oop.append_dirichlet_coupling(1.0, kappa)
    oop.append_boundary_source(0, 1.0, kappa, u(np.moveaxis(m.boundary_quadrature_points(1 + over), -1, 0)),
                               g_order=1, over_integrate=over)
    oop.append_boundary_source(1, sigma, kappa, u(np.moveaxis(m.boundary_quadrature_points(2), -1, 0)), g_order=1)
    oop.assemble(4)
    return s, oop


def _u(X):
    return 0.3 + 2.0 * X[0] - 1.0 * X[1] + 0.5 * X[2]


@pytest.mark.parametrize("name,geom,cells", GEOMS)
def test_swipdg_3d_matches_oracle(hgdt_mod, name, geom, cells):
    H = hgdt_mod
    print(f"case SWIPDG on {name} {cells}, lognormal kappa, weak Dirichlet data", flush=True)
    grid = H.Grid.structured(name, [0, 0, 0], [1.0, 0.8, 0.6], cells)
    kappa = H.fields.lognormal(91, grid.num_elements)
    over = 1 if name == "hex" else 0
    space, op, _ = _hgdt_swipdg(H, grid, kappa, 16.0, _u, over)
    m = O.Mesh.structured(geom, cells, [0, 0, 0], [1.0, 0.8, 0.6])
    s, oop = _oracle_swipdg(m, kappa, 16.0, _u, over)
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    ob = oop.rhs()
    berr = np.abs(op.rhs.cpu().numpy() - ob).max() / np.abs(ob).max()
    x = op.apply_inverse("rhs", opts={"precision": 1e-10, "max_iter": 100000}).cpu().numpy()
    xo, ito, relo, rc = O.pcg_jacobi(orp, ocols, oop.values(), ob, rtol=1e-10, max_iter=100000)
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"  matrix {err:.2e}, rhs {berr:.2e}, iterations {op.last_solve['iterations']} (oracle {ito}), "
          f"solution rel L2 {rel:.2e}")
    assert err <= 1e-12 and berr <= 1e-12
    assert rc == 0 and rel <= 1e-10 and abs(op.last_solve["iterations"] - ito) <= 2


@pytest.mark.parametrize("name,geom,cells", GEOMS)
def test_swipdg_3d_reproduces_linear_solution(hgdt_mod, name, geom, cells):
    H = hgdt_mod
    print(f"case SWIPDG on {name} {cells}, kappa = 1, linear solution", flush=True)
    grid = H.Grid.structured(name, [0, 0, 0], [1, 1, 1], cells)
    space, op, _ = _hgdt_swipdg(H, grid, np.ones(grid.num_elements), 16.0, _u, 1 if name == "hex" else 0)
    x = op.apply_inverse("rhs", opts={"precision": 1e-14, "max_iter": 100000}).cpu().numpy()
    Xe = grid.coordinates()[grid.connectivity()]
    rep = np.abs(x[space.global_indices()] - _u(np.moveaxis(Xe, -1, 0))).max()
    print(f"  linear solution reproduced to {rep:.2e}")
    assert rep <= 1e-9


@pytest.mark.parametrize("name,geom,cells", GEOMS)
def test_swipdg_3d_dirichlet_forms_on_x_faces(hgdt_mod, name, geom, cells):
    H = hgdt_mod
    print(f"case SWIPDG on {name} {cells} with the Dirichlet forms on the x faces only", flush=True)
    grid = H.Grid.structured(name, [0, 0, 0], [1.0, 0.8, 0.6], cells)
    info = H.NormalBasedBoundaryInfo(default_boundary_type=H.NeumannBoundary())
    info.register_new_normal([1, 0, 0], H.DirichletBoundary())
    info.register_new_normal([-1, 0, 0], H.DirichletBoundary())
    mask = info.mask(grid, H.DirichletBoundary)
    _, normals = grid.boundary_intersections()
    assert np.array_equal(mask.astype(bool), np.abs(np.abs(normals[:, 0]) - 1) < 1e-12)
    kappa = H.fields.lognormal(92, grid.num_elements)
    over = 1 if name == "hex" else 0
    on = H.ApplyOn.CustomBoundaryIntersections(info, H.DirichletBoundary())
    space, op, _ = _hgdt_swipdg(H, grid, kappa, 16.0, _u, over, on)
    m = O.Mesh.structured(geom, cells, [0, 0, 0], [1.0, 0.8, 0.6])
    s, oop = _oracle_swipdg(m, kappa, 16.0, _u, over, mask)
    rp, _ = oop.pattern()
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), rp)
    berr = np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() / np.abs(oop.rhs()).max()
    print(f"  {int(mask.sum())} of {len(mask)} boundary faces Dirichlet, matrix {err:.2e}, rhs {berr:.2e}")
    assert err <= 1e-12 and berr <= 1e-12


@pytest.mark.parametrize("geometry,geom", [("tri", O.TRI), ("quad", O.QUAD)])
def test_swipdg_order2_matches_oracle(hgdt_mod, geometry, geom):
    """SWIPDG on P2 / Q2 discontinuous spaces (order-2 intersection integrals): matrix and rhs within 1e-12
    of the oracle's, and the quadratic solution u = x^2 - xy + y^2/2 + 0.3 x (-Laplace u = -3, g = u)
    reproduced to 1e-9"""
    H = hgdt_mod
    print(f"case SWIPDG order 2 on {geometry}", flush=True)
    n, sigma = 4, 40.0
    over = 1 if geometry == "quad" else 0

    def u2(X):
        return X[0] * X[0] - X[0] * X[1] + 0.5 * X[1] * X[1] + 0.3 * X[0]

    grid = H.Grid.structured(geometry, [0, 0], [1, 1], [n, n])
    space = H.DiscontinuousLagrangeSpace(grid, 2)
    op = H.make_matrix_operator(space, H.Stencil.element_and_intersection)
    op += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0))
    op.append(H.LocalCouplingIntersectionIntegralBilinearForm(
        H.LocalLaplaceIPDGInnerCoupling(1.0, 1.0, 1.0) + H.LocalIPDGInnerPenalty(sigma, 1.0)), {},
        H.ApplyOn.InnerIntersectionsOnce())
    op.append(H.LocalIntersectionIntegralBilinearForm(
        H.LocalIPDGBoundaryPenalty(sigma, 1.0) + H.LocalLaplaceIPDGDirichletCoupling(1.0, 1.0)), {},
        H.ApplyOn.BoundaryIntersections())
    g = H.GenericFunction(2, u2)
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(-3.0)))
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalLaplaceIPDGDirichletCoupling(1.0, 1.0, g),
                                                     over_integrate=over), {}, H.ApplyOn.BoundaryIntersections())
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalIPDGBoundaryPenalty(sigma, 1.0).with_ansatz(g)), {},
               H.ApplyOn.BoundaryIntersections())
    op.append(rhs)
    op.assemble()
    m = O.Mesh.structured(geom, [n, n], [0, 0], [1, 1])
    s = O.Space(m, False, 2)
    oop = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    oop.append_laplace(1.0)
    oop.append_inner_coupling(1.0, 1.0, 1.0)
    oop.append_inner_penalty(sigma, 1.0)
    oop.append_boundary_penalty(sigma, 1.0)
    oop.append_dirichlet_coupling(1.0, 1.0)
    oop.append_source(-3.0)
    oop.append_boundary_source(0, 1.0, 1.0, u2(np.moveaxis(m.boundary_quadrature_points(3 + over), -1, 0)),
                               g_order=2, over_integrate=over)
    oop.append_boundary_source(1, sigma, 1.0, u2(np.moveaxis(m.boundary_quadrature_points(4), -1, 0)), g_order=2)
    oop.assemble(1)
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    berr = np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() / np.abs(oop.rhs()).max()
    x = op.apply_inverse("rhs", opts={"precision": 1e-14, "max_iter": 100000}).cpu().numpy()
    import scipy.sparse.linalg as spla

    xo = spla.spsolve(oop.csr().tocsc(), oop.rhs())
    rep = np.abs(x - xo).max() / np.abs(xo).max()
    print(f"  matrix {err:.2e}, rhs {berr:.2e}, solution vs the oracle's exact solve {rep:.2e}")
    assert err <= 1e-12 and berr <= 1e-12 and rep <= 1e-9


@pytest.mark.parametrize("geometry,geom,n", [("tet", O.TET, 2), ("hex", O.HEX, 3)])
def test_swipdg_order2_3d_matches_oracle(hgdt_mod, geometry, geom, n):
    """SWIPDG on P2 tetrahedra and Q2 hexahedra (27 basis functions per side of a face: one lane per face
    row): matrix and rhs within 1e-12 of the oracle's, and the quadratic solution
    u = x^2 - xy + y^2/2 + 0.3x + z^2 - 0.2yz (-Laplace u = -5, g = u) reproduced to 1e-9"""
    H = hgdt_mod
    print(f"case SWIPDG order 2 on {geometry} {n}^3", flush=True)
    sigma = 60.0
    over = 1 if geometry == "hex" else 0

    def u3(X):
        return X[0] * X[0] - X[0] * X[1] + 0.5 * X[1] * X[1] + 0.3 * X[0] + X[2] * X[2] - 0.2 * X[1] * X[2]

    grid = H.Grid.structured(geometry, [0, 0, 0], [1, 1, 1], [n, n, n])
    space = H.DiscontinuousLagrangeSpace(grid, 2)
    op = H.make_matrix_operator(space, H.Stencil.element_and_intersection)
    op += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0))
    op.append(H.LocalCouplingIntersectionIntegralBilinearForm(
        H.LocalLaplaceIPDGInnerCoupling(1.0, 1.0, 1.0) + H.LocalIPDGInnerPenalty(sigma, 1.0)), {},
        H.ApplyOn.InnerIntersectionsOnce())
    op.append(H.LocalIntersectionIntegralBilinearForm(
        H.LocalIPDGBoundaryPenalty(sigma, 1.0) + H.LocalLaplaceIPDGDirichletCoupling(1.0, 1.0)), {},
        H.ApplyOn.BoundaryIntersections())
    g = H.GenericFunction(2, u3)
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(-5.0)))
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalLaplaceIPDGDirichletCoupling(1.0, 1.0, g),
                                                     over_integrate=over), {}, H.ApplyOn.BoundaryIntersections())
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalIPDGBoundaryPenalty(sigma, 1.0).with_ansatz(g)), {},
               H.ApplyOn.BoundaryIntersections())
    op.append(rhs)
    op.assemble()
    m = O.Mesh.structured(geom, [n, n, n], [0, 0, 0], [1, 1, 1])
    s = O.Space(m, False, 2)
    oop = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    oop.append_laplace(1.0)
    oop.append_inner_coupling(1.0, 1.0, 1.0)
    oop.append_inner_penalty(sigma, 1.0)
    oop.append_boundary_penalty(sigma, 1.0)
    oop.append_dirichlet_coupling(1.0, 1.0)
    oop.append_source(-5.0)
    oop.append_boundary_source(0, 1.0, 1.0, u3(np.moveaxis(m.boundary_quadrature_points(3 + over), -1, 0)),
                               g_order=2, over_integrate=over)
    oop.append_boundary_source(1, sigma, 1.0, u3(np.moveaxis(m.boundary_quadrature_points(4), -1, 0)), g_order=2)
    oop.assemble(1)
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    berr = np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() / np.abs(oop.rhs()).max()
    import scipy.sparse.linalg as spla

    xo = spla.spsolve(oop.csr().tocsc(), oop.rhs())
    x = op.apply_inverse("rhs", opts={"precision": 1e-14, "max_iter": 100000}).cpu().numpy()
    rep = np.abs(x - xo).max() / np.abs(xo).max()
    A = op.matrix.to_scipy()
    sym = abs(A - A.T).max() / abs(A).max()
    print(f"  {s.size} DoFs: matrix {err:.2e}, rhs {berr:.2e}, symmetric to {sym:.1e}, solution vs the "
          f"oracle's exact solve {rep:.2e}")
    assert err <= 1e-12 and berr <= 1e-12 and rep <= 1e-9 and sym <= 1e-14
