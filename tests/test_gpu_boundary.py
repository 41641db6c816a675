"""Right-hand sides on boundary intersections (LocalIntersectionIntegralFunctional) on the HIP path:
weakly imposed nonzero Dirichlet data for SWIPDG (Dirichlet coupling as a unary integrand, boundary
penalty with the data as ansatz) and boundary product terms.

Bar: right-hand side within 1e-12 (relative to its maximum) of the oracle with the data at the same
face quadrature points; the SWIPDG solution within 1e-10 relative L2 of the oracle's; a linear exact
solution reproduced to 1e-10 (consistency of the weak Dirichlet terms).
"""
import numpy as np
import pytest
import torch

import oracle_lib as O

pytestmark = pytest.mark.gpu


def _g(X):
    return 0.5 + np.sin(1.3 * X[0]) * np.cos(0.7 * X[1])


def _swipdg(H, grid, space, kappa, sigma, g, over_dc=0, f=0.0):
    op = H.make_matrix_operator(space, H.Stencil.element_and_intersection)
    op += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(kappa))
    op.append(H.LocalCouplingIntersectionIntegralBilinearForm(
        H.LocalLaplaceIPDGInnerCoupling(1.0, kappa, kappa) + H.LocalIPDGInnerPenalty(sigma, kappa)), {},
        H.ApplyOn.InnerIntersectionsOnce())
    boundary = H.ApplyOn.CustomBoundaryIntersections(H.AllDirichletBoundaryInfo())
    op.append(H.LocalIntersectionIntegralBilinearForm(
        H.LocalIPDGBoundaryPenalty(sigma, kappa) + H.LocalLaplaceIPDGDirichletCoupling(1.0, kappa)), {}, boundary)
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(f)))
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalLaplaceIPDGDirichletCoupling(1.0, kappa, g),
                                                     over_integrate=over_dc), {}, boundary)
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalIPDGBoundaryPenalty(sigma, kappa).with_ansatz(g)), {},
               boundary)
    op.append(rhs)
    op.assemble()
    return op


@pytest.mark.parametrize("geometry,geom_o", [("quad", O.QUAD), ("tri", O.TRI)])
def test_weak_dirichlet_rhs_and_solve_parity(hgdt_mod, geometry, geom_o):
    H = hgdt_mod
    print(f"case weak Dirichlet data {geometry}", flush=True)
    cells, sigma = [14, 10], 16.0
    grid = H.Grid.structured(geometry, [0, 0], [1.5, 1], cells)
    kappa = H.fields.lognormal(37, grid.num_elements)
    space = H.DiscontinuousLagrangeSpace(grid, 1)
    g = H.GenericFunction(2, _g)
    op = _swipdg(H, grid, space, torch.from_numpy(kappa).cuda(), sigma, g, f=1.0)
    m = O.Mesh.structured(geom_o, cells, [0, 0], [1.5, 1])
    s = O.Space(m, False, 1)
    oop = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    oop.append_laplace(kappa)
    oop.append_inner_coupling(1.0, kappa, kappa)
    oop.append_inner_penalty(sigma, kappa)
    oop.append_dirichlet_coupling(1.0, kappa)
    oop.append_boundary_penalty(sigma, kappa)
    oop.append_source(1.0)
    P1 = m.boundary_quadrature_points(2 + 0)  # coupling: data order 2 + max(1 - 1, 0)
    P2 = m.boundary_quadrature_points(1 + 2)  # penalty: test order 1 + data order 2
    assert np.abs(P1 - grid.boundary_quadrature_points(2)).max() <= 1e-15
    oop.append_boundary_source(0, 1.0, kappa, _g(P1.transpose(2, 0, 1)), g_order=2)
    oop.append_boundary_source(1, sigma, kappa, _g(P2.transpose(2, 0, 1)), g_order=2)
    oop.assemble(1)
    b, ob = op.rhs.cpu().numpy(), oop.rhs()
    berr = np.abs(b - ob).max() / np.abs(ob).max()
    print(f"  rhs max rel err {berr:.3e}")
    assert berr <= 1e-12
    rp, cols = oop.pattern()
    xo, ito, relo, rc = O.pcg_jacobi(rp, cols, oop.values(), ob, rtol=1e-10, max_iter=100000)
    x = op.apply_inverse("rhs", opts={"precision": 1e-10, "max_iter": 100000}).cpu().numpy()
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"  iters {op.last_solve['iterations']} (oracle {ito}) rel L2 {rel:.2e}")
    assert rel <= 1e-10 and abs(op.last_solve["iterations"] - ito) <= 2


@pytest.mark.parametrize("geometry", ["quad", "tri"])
def test_swipdg_reproduces_linear_solution_on_device(hgdt_mod, geometry):
    """u = 0.3 + 2x - 1.5y with its boundary values as weak Dirichlet data and f = 0: the SWIPDG solution is u
    at every DoF (on quadrilaterals with one order more for the coupling term, whose reference order
    under-integrates the bilinear normal derivative along a face)"""
    H = hgdt_mod
    print(f"case linear reproduction {geometry}", flush=True)
    grid = H.Grid.structured(geometry, [0, 0], [1, 1], [6, 6])
    space = H.DiscontinuousLagrangeSpace(grid, 1)
    u = H.GenericFunction(1, lambda X: 0.3 + 2.0 * X[0] - 1.5 * X[1])
    op = _swipdg(H, grid, space, 1.0, 16.0, u, over_dc=1 if geometry == "quad" else 0)
    x = op.apply_inverse("rhs", opts={"precision": 1e-14, "max_iter": 100000}).cpu().numpy()
    X = grid.coordinates()[grid.connectivity()]
    gi = space.global_indices()
    err = np.abs(x[gi] - (0.3 + 2.0 * X[..., 0] - 1.5 * X[..., 1])).max()
    print(f"  max nodal error {err:.2e}")
    assert err <= 1e-10


def test_boundary_product_functional(hgdt_mod):
    """LocalIntersectionProductIntegrand(weight).with_ansatz(g) on a continuous space (no constraints):
    parity with the oracle, and with g = 1 the entries sum to weight x |boundary|"""
    H = hgdt_mod
    print("case boundary product functional", flush=True)
    cells = [9, 7]
    grid = H.Grid.structured("quad", [0, 0], [2, 1], cells)
    space = H.ContinuousLagrangeSpace(grid, 1)
    w = H.fields.uniform(41, grid.num_elements) + 0.5
    for g in (1.0, H.GenericFunction(2, _g)):
        op = H.make_matrix_operator(space, H.Stencil.element)
        op += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0))
        rhs = H.make_vector_functional(space)
        rhs.append(H.LocalIntersectionIntegralFunctional(
            H.LocalIntersectionProductIntegrand(torch.from_numpy(w).cuda()).with_ansatz(g)), {},
            H.ApplyOn.CustomBoundaryIntersections(H.AllDirichletBoundaryInfo()))
        op.append(rhs)
        op.assemble()
        m = O.Mesh.structured(O.QUAD, cells, [0, 0], [2, 1])
        s = O.Space(m, True, 1)
        oop = O.Operator(s, O.STENCIL_ELEMENT)
        oop.append_laplace(1.0)
        if isinstance(g, float):
            oop.append_boundary_source(2, 0.0, w, g, g_order=0)
        else:
            P = m.boundary_quadrature_points(1 + 2)
            oop.append_boundary_source(2, 0.0, w, _g(P.transpose(2, 0, 1)), g_order=2)
        oop.assemble(1)
        b, ob = op.rhs.cpu().numpy(), oop.rhs()
        assert np.abs(b - ob).max() <= 1e-12 * np.abs(ob).max()
        if isinstance(g, float):
            # every boundary face F of element e contributes w_e |F| (faces 0, 1 have length hy, faces 2, 3 hx)
            nb = m.neighbors()
            hy, hx = 1.0 / cells[1], 2.0 / cells[0]
            lens = np.array([hy, hy, hx, hx])
            total = sum(w[e] * lens[f] for e in range(m.ne) for f in range(4) if nb[e, f] < 0)
            assert abs(b.sum() - total) <= 1e-12 * total
