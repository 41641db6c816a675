"""Mixed triangle/quadrilateral grids on the HIP path (one kernel launch per element type).

Parity with the oracle (same element list, same DoF numbering) on the reference's mixed grid
(dune/gdt/test/spaces/base.hh:355-376, globally refined once) and on a strip grid with sheared cells:
global indices, sparsity pattern, matrix and right-hand side (Laplace with a per-element coefficient,
product, constant / per-element / quadrature-point sources), Dirichlet DoFs, the Jacobi-PCG solution and
the error norms, for continuous and discontinuous Lagrange spaces of orders 1 and 2.  The device's mass
matrix also sums to the grid's area (closed form), and what needs one element type is refused.  SWIPDG
with weak Dirichlet data on the same grids (faces joining triangles and quadrilaterals), orders 1 and 2:
matrix, right-hand side and solve against the oracle.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from mixed_grids import grid_area, refined_reference_mixed_grid, strip_mixed_grid
from problems import matrix_error

pytestmark = pytest.mark.gpu

ESV_FORCE = (lambda X: 0.5 * np.pi * np.pi * np.cos(0.5 * np.pi * X[0]) * np.cos(0.5 * np.pi * X[1]))


def _grid(name):
    if name == "reference":
        return refined_reference_mixed_grid()
    if name == "strip":
        return strip_mixed_grid(24, 18, shear=0.35)
    return strip_mixed_grid(200, 160)  # 48 000 elements


def _pair(H, name, continuous, order, source):
    V, E = _grid(name)
    rng = np.random.default_rng(11)
    kappa = np.exp(rng.normal(size=len(E)))
    m = O.Mesh.mixed(V, E)
    s = O.Space(m, continuous, order)
    oop = O.Operator(s, O.STENCIL_ELEMENT)
    oop.append_laplace(kappa)
    oop.append_product(0.5)
    if source == "element":
        f = rng.normal(size=len(E))
        oop.append_source(f)
    elif source == "esv":
        oop.append_source(esv2007=True, f_order=3)
    else:
        oop.append_source(1.0)
    if continuous:
        oop.set_dirichlet(True)
    oop.assemble(1)
    grid = H.Grid.mixed(V, E)
    space = (H.ContinuousLagrangeSpace if continuous else H.DiscontinuousLagrangeSpace)(grid, order)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(torch.from_numpy(kappa).cuda())))
    op.append(H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand(0.5)))
    rhs = H.make_vector_functional(space)
    if source == "element":
        fn = torch.from_numpy(f).cuda()
    elif source == "esv":
        fn = H.GenericFunction(3, ESV_FORCE)
    else:
        fn = 1.0
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(fn)))
    op.append(rhs)
    if continuous:
        op.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    op.assemble()
    return V, E, m, s, oop, grid, space, op


@pytest.mark.parametrize("name", ["reference", "strip"])
@pytest.mark.parametrize("continuous,order", [(True, 1), (True, 2), (False, 1), (False, 2)])
@pytest.mark.parametrize("source", ["constant", "element", "esv"])
def test_mixed_assembly_parity(hgdt_mod, name, continuous, order, source):
    H = hgdt_mod
    print(f"case mixed assembly {name} continuous={continuous} order={order} source={source}", flush=True)
    V, E, m, s, oop, grid, space, op = _pair(H, name, continuous, order, source)
    assert grid.is_mixed and grid.num_elements == len(E)
    assert space.num_DoFs == s.size
    assert np.array_equal(space.global_indices(), s.global_indices())
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    b, ob = op.rhs.cpu().numpy(), oop.rhs()
    berr = np.abs(b - ob).max() / np.abs(ob).max()
    print(f"  {len(E)} elements, {s.size} DoFs: matrix max rel err {err:.2e}, rhs {berr:.2e}")
    assert err <= 1e-12 and berr <= 1e-12
    if continuous:
        assert np.array_equal(op.dirichlet_dofs(), oop.dirichlet_dofs())


@pytest.mark.parametrize("name", ["strip", "large"])
@pytest.mark.parametrize("order", [1, 2])
def test_mixed_solve_parity(hgdt_mod, name, order):
    H = hgdt_mod
    print(f"case mixed solve {name} order={order}", flush=True)
    V, E, m, s, oop, grid, space, op = _pair(H, name, True, order, "esv")
    orp, ocols = oop.pattern()
    xo, ito, relo, rc = O.pcg_jacobi(orp, ocols, oop.values(), oop.rhs(), rtol=1e-10)
    x = op.apply_inverse("rhs", opts={"precision": 1e-10}).cpu().numpy()
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    st = op.last_solve
    print(f"  {s.size} DoFs: iterations {st['iterations']} (oracle {ito}), rel L2 {rel:.2e}, true relres "
          f"{st['true_relres']:.2e}")
    assert rel <= 1e-9 and abs(st["iterations"] - ito) <= 2 and st["true_relres"] <= 2e-10
    # error norms against an analytic function at the grid's element-ordered quadrature points
    g = H.GenericFunction(3, lambda X: np.sin(X[0]) * np.cos(X[1]),
                          jacobian=lambda X: np.stack([np.cos(X[0]) * np.cos(X[1]), -np.sin(X[0]) * np.sin(X[1])]))
    uh = H.make_discrete_function(space, torch.from_numpy(x))
    f1 = H.make_bilinear_form(grid, uh - g)
    f1 += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0))
    f2 = H.make_bilinear_form(grid, uh - g)
    f2 += H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand())
    h1, l2 = f1.apply2(), f2.apply2()
    P = m.quadrature_points(6)
    assert np.abs(P - grid.quadrature_points(6)).max() <= 1e-14
    ol2, oh1 = O.error_norms(s, x, 6, g.evaluate(P.T), np.ascontiguousarray(g.jacobian(P.T).T))
    print(f"  error norms: H1 {h1:.6e} (oracle {oh1:.6e}), L2 {l2:.6e} (oracle {ol2:.6e})")
    assert abs(h1 - oh1) <= 1e-12 * oh1 and abs(l2 - ol2) <= 1e-12 * ol2


@pytest.mark.parametrize("continuous,order", [(True, 1), (True, 2), (False, 2)])
def test_mixed_mass_sums_to_area_on_device(hgdt_mod, continuous, order):
    H = hgdt_mod
    print(f"case mixed mass closed form continuous={continuous} order={order}", flush=True)
    V, E = refined_reference_mixed_grid()
    grid = H.Grid.mixed(V, E)
    space = (H.ContinuousLagrangeSpace if continuous else H.DiscontinuousLagrangeSpace)(grid, order)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand(1.0)))
    op.assemble()
    total = float(op.matrix.values.sum())
    print(f"  mass total {total!r}, area {grid_area(V, E)!r}")
    assert abs(total - grid_area(V, E)) <= 1e-14


def test_mixed_single_type_is_plain_grid(hgdt_mod):
    """a 'mixed' element list of one type gives that type's grid: the same matrix as the explicit grid"""
    H = hgdt_mod
    print("case mixed grid of one element type", flush=True)
    V, E = strip_mixed_grid(6, 4)
    tris = [el for el in E if len(el) == 3]
    g1 = H.Grid.mixed(V, tris)
    g2 = H.Grid.from_arrays("tri", V, np.array(tris))
    assert not g1.is_mixed and g1.geometry == g2.geometry
    mats = []
    for g in (g1, g2):
        sp_ = H.ContinuousLagrangeSpace(g, 1)
        op = H.make_matrix_operator(sp_, H.Stencil.element)
        op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0)))
        op.assemble()
        mats.append(op.matrix.values.cpu().numpy())
    assert np.array_equal(mats[0], mats[1])


def _g(X):
    return 0.5 + np.sin(1.3 * X[0]) * np.cos(0.7 * X[1])


@pytest.mark.parametrize("name", ["reference", "strip"])
@pytest.mark.parametrize("order", [1, 2])
def test_mixed_swipdg_parity(hgdt_mod, name, order):
    """SWIPDG (heterogeneous kappa, penalty weighted by kappa) with weak Dirichlet data on a mixed grid, whose
    intersections join triangles and quadrilaterals: matrix and right-hand side against the oracle (1e-12),
    the solve against the oracle's (1e-10, its iteration count)"""
    H = hgdt_mod
    print(f"case mixed SWIPDG {name} order={order}", flush=True)
    V, E = _grid(name)
    sigma = 16.0 if order == 1 else 40.0
    kappa = np.exp(np.random.default_rng(5).normal(size=len(E)))
    grid = H.Grid.mixed(V, E)
    space = H.DiscontinuousLagrangeSpace(grid, order)
    kd = torch.from_numpy(kappa).cuda()
    g = H.GenericFunction(2, _g)
    op = H.make_matrix_operator(space, H.Stencil.element_and_intersection)
    op += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(kd))
    op.append(H.LocalCouplingIntersectionIntegralBilinearForm(
        H.LocalLaplaceIPDGInnerCoupling(1.0, kd, kd) + H.LocalIPDGInnerPenalty(sigma, kd)), {},
        H.ApplyOn.InnerIntersectionsOnce())
    boundary = H.ApplyOn.CustomBoundaryIntersections(H.AllDirichletBoundaryInfo())
    op.append(H.LocalIntersectionIntegralBilinearForm(
        H.LocalIPDGBoundaryPenalty(sigma, kd) + H.LocalLaplaceIPDGDirichletCoupling(1.0, kd)), {}, boundary)
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalLaplaceIPDGDirichletCoupling(1.0, kd, g)), {}, boundary)
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalIPDGBoundaryPenalty(sigma, kd).with_ansatz(g)), {},
               boundary)
    op.append(rhs)
    op.assemble()
    m = O.Mesh.mixed(V, E)
    s = O.Space(m, False, order)
    oop = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    oop.append_laplace(kappa)
    oop.append_inner_coupling(1.0, kappa, kappa)
    oop.append_inner_penalty(sigma, kappa)
    oop.append_dirichlet_coupling(1.0, kappa)
    oop.append_boundary_penalty(sigma, kappa)
    oop.append_source(1.0)
    P1 = m.boundary_quadrature_points(2 + max(order - 1, 0))  # coupling: data order 2 + max(test - 1, 0)
    P2 = m.boundary_quadrature_points(order + 2)  # penalty: test order + data order 2
    assert np.abs(P1 - grid.boundary_quadrature_points(2 + max(order - 1, 0))).max() <= 1e-15
    oop.append_boundary_source(0, 1.0, kappa, _g(P1.transpose(2, 0, 1)), g_order=2)
    oop.append_boundary_source(1, sigma, kappa, _g(P2.transpose(2, 0, 1)), g_order=2)
    oop.assemble(1)
    assert np.array_equal(space.global_indices(), s.global_indices())
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    b, ob = op.rhs.cpu().numpy(), oop.rhs()
    berr = np.abs(b - ob).max() / np.abs(ob).max()
    xo, ito, relo, rc = O.pcg_jacobi(orp, ocols, oop.values(), ob, rtol=1e-10, max_iter=100000)
    x = op.apply_inverse("rhs", opts={"precision": 1e-10, "max_iter": 100000}).cpu().numpy()
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"  {len(E)} elements, {s.size} DoFs: matrix {err:.2e}, rhs {berr:.2e}; iterations "
          f"{op.last_solve['iterations']} (oracle {ito}), rel L2 {rel:.2e}")
    assert err <= 1e-12 and berr <= 1e-12
    assert rel <= 1e-10 and abs(op.last_solve["iterations"] - ito) <= 2


def test_mixed_refusals(hgdt_mod):
    H = hgdt_mod
    print("case mixed refusals", flush=True)
    V, E = refined_reference_mixed_grid()
    grid = H.Grid.mixed(V, E)
    with pytest.raises(H.not_supported):
        H.Grid.union([grid, grid])
    with pytest.raises(H.wrong_input_given):
        H.Grid.mixed(V, [[0, 1]])
    space = H.ContinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0)))
    op.set_format("matrix_free")
    op.assemble()
    with pytest.raises(H.not_supported):
        op.apply(torch.zeros(space.num_DoFs, dtype=torch.float64, device="cuda"))


@pytest.mark.parametrize("order", [1, 2])
def test_adjoint_gradient_on_mixed_grid(hgdt_mod, order):
    """the adjoint element kernel on a mixed grid (one launch per element type, each element's gradient at its
    grid index): for the solve u and the adjoint lambda (A lambda = u), within 1e-12 of the oracle's gradient
    (pinned by central differences)"""
    H = hgdt_mod
    print(f"case adjoint gradient on a mixed grid, order {order}", flush=True)
    V, E = strip_mixed_grid(10, 8, shear=0.25)
    grid = H.Grid.mixed(V, E)
    space = H.ContinuousLagrangeSpace(grid, order)
    kap = np.exp(H.fields.normal(43, grid.num_elements))
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(torch.from_numpy(kap).cuda())))
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    op.assemble()
    u = op.apply_inverse("rhs", opts={"precision": 1e-13})
    lam = op.apply_inverse(u.clone(), opts={"precision": 1e-13})
    g = op.adjoint_log_diffusion_gradient(u, lam).cpu().numpy()
    s = O.Space(O.Mesh.mixed(V, E), True, order)
    go = O.adjoint_log_kappa_gradient(s, kap, u.cpu().numpy(), lam.cpu().numpy())
    rel = np.abs(g - go).max() / np.abs(go).max()
    print(f"  {grid.num_elements} elements: gradient vs oracle {rel:.2e}")
    assert rel <= 1e-12
