"""SWIPDG with a matrix-valued diffusion and weight on the HIP path (laplace-ipdg.hh:47-61,145-186,
ipdg.hh:59-72,149-153): per-element symmetric positive definite tensors in the element Laplace form, the inner
coupling and penalty, the Dirichlet coupling and the boundary penalty, on quadrilaterals, triangles,
hexahedra (order 2: Q2 row lanes) and tetrahedra, and on a mixed grid.  Bar: pattern bit-exact, matrix within
1e-12 of the oracle's (row-relative), symmetric to 1e-14; Jacobi-PCG solution within 1e-10 of the oracle's with
its iteration count (+-2).
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from mixed_grids import strip_mixed_grid
from problems import matrix_error

pytestmark = pytest.mark.gpu


def _spd(seed, ne, d):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((ne, d, d))
    return np.einsum("eab,ecb->eac", B, B) + 0.5 * np.eye(d)


def _run(H, grid, m, k, sigma, label, solve=True):
    d = grid.dimension
    # the weight is the diffusion (SWIPDG's choice, which the penalty then controls) for the solves; an
    # independent weight for the matrix alone
    K = _spd(11, grid.num_elements, d)
    W = K if solve else _spd(12, grid.num_elements, d)
    Kt, Wt = torch.from_numpy(K).cuda(), torch.from_numpy(W).cuda()
    space = H.DiscontinuousLagrangeSpace(grid, k)
    op = H.make_matrix_operator(space, H.Stencil.element_and_intersection)
    op += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(Kt))
    op.append(H.LocalCouplingIntersectionIntegralBilinearForm(H.LocalLaplaceIPDGInnerCoupling(1.0, Kt, Wt)), {},
              H.ApplyOn.InnerIntersectionsOnce())
    op.append(H.LocalCouplingIntersectionIntegralBilinearForm(H.LocalIPDGInnerPenalty(sigma, Wt)), {},
              H.ApplyOn.InnerIntersectionsOnce())
    op.append(H.LocalIntersectionIntegralBilinearForm(H.LocalLaplaceIPDGDirichletCoupling(1.0, Kt)), {},
              H.ApplyOn.BoundaryIntersections())
    op.append(H.LocalIntersectionIntegralBilinearForm(H.LocalIPDGBoundaryPenalty(sigma, Wt)), {},
              H.ApplyOn.BoundaryIntersections())
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.assemble()
    s = O.Space(m, False, k)
    oop = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    one = np.ones(m.ne)
    oop.append_laplace_tensor(K)
    oop.append_inner_coupling(1.0, one, one)
    oop.set_face_tensors(False, K, W)
    oop.append_inner_penalty(sigma, one)
    oop.set_face_tensors(False, None, W)
    oop.append_dirichlet_coupling(1.0, one)
    oop.set_face_tensors(True, K, None)
    oop.append_boundary_penalty(sigma, one)
    oop.set_face_tensors(True, None, W)
    oop.append_source(1.0)
    oop.assemble(1)
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    A = op.matrix.to_scipy()
    if not solve:
        print(f"  {label}: matrix {err:.2e}")
        assert err <= 1e-12 and abs(A - A.T).max() <= 1e-14 * abs(A).max()
        return
    x = op.apply_inverse("rhs").cpu().numpy()
    xo, ito, relo, rc = O.pcg_jacobi(orp, ocols, oop.values(), oop.rhs(), rtol=1e-10)
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    its = op.last_solve["iterations"]
    print(f"  {label}: matrix {err:.2e}, solution {rel:.2e}, iterations {its} (oracle {ito})")
    assert err <= 1e-12 and abs(A - A.T).max() <= 1e-14 * abs(A).max() and rel <= 1e-10 and abs(its - ito) <= 2


@pytest.mark.parametrize("geometry,geom,cells,k", [("quad", O.QUAD, [6, 5], 1), ("tri", O.TRI, [4, 4], 2),
                                                   ("hex", O.HEX, [3, 2, 2], 1), ("hex", O.HEX, [2, 2, 2], 2),
                                                   ("tet", O.TET, [2, 2, 2], 1), ("tet", O.TET, [2, 1, 2], 2)])
def test_swipdg_tensors_match_oracle(hgdt_mod, geometry, geom, cells, k):
    H = hgdt_mod
    print(f"case SWIPDG tensors {geometry} cells={cells} order {k}", flush=True)
    d = len(cells)
    grid = H.Grid.structured(geometry, [0] * d, [1] * d, cells)
    m = O.Mesh.structured(geom, cells, [0] * d, [1] * d)
    _run(H, grid, m, k, 40.0 if k == 1 else 200.0, f"{grid.num_elements} elements")
    _run(H, grid, m, k, 40.0, "independent weight", solve=False)


@pytest.mark.parametrize("k", [1, 2])
def test_swipdg_tensors_on_mixed_grid(hgdt_mod, k):
    H = hgdt_mod
    print(f"case SWIPDG tensors mixed grid order {k}", flush=True)
    V, E = strip_mixed_grid(6, 4, shear=0.2)
    _run(H, H.Grid.mixed(V, E), O.Mesh.mixed(V, E), k, 40.0 if k == 1 else 200.0, "mixed")


def test_face_tensor_refusals(hgdt_mod):
    H = hgdt_mod
    print("case SWIPDG tensor refusals", flush=True)
    grid = H.Grid.structured("quad", [0, 0], [1, 1], [4, 4])
    space = H.DiscontinuousLagrangeSpace(grid, 1)
    K = _spd(5, grid.num_elements, 2)
    K[3, 1, 0] += 1e-3
    op = H.make_matrix_operator(space, H.Stencil.element_and_intersection)
    with pytest.raises(H._lib.HgdtError):  # not symmetric
        op.append(H.LocalCouplingIntersectionIntegralBilinearForm(
            H.LocalLaplaceIPDGInnerCoupling(1.0, torch.from_numpy(K).cuda(), 1.0)), {},
            H.ApplyOn.InnerIntersectionsOnce())
    with pytest.raises(H._lib.HgdtError):  # the boundary product term takes a scalar weight
        H.LocalIntersectionProductIntegrand([[1.0, 0.0], [0.0, 2.0]])


@pytest.mark.parametrize("geometry,geom,cells", [("quad", O.QUAD, [4, 3]), ("tri", O.TRI, [4, 3]),
                                                 ("hex", O.HEX, [2, 2, 2]), ("tet", O.TET, [2, 2, 1])])
def test_swipdg_tensor_boundary_functionals_reproduce_linear_solution(hgdt_mod, geometry, geom, cells):
    """a constant anisotropic K as diffusion and weight of every form and of the boundary functionals (weak
    Dirichlet data g = u for a linear u, f = 0): rhs within 1e-12 of the oracle's and the solution equal to u at
    every DoF to 1e-9"""
    H = hgdt_mod
    print(f"case SWIPDG tensor boundary functionals {geometry}", flush=True)
    d = len(cells)
    Kc = np.array([[2.0, 0.6, -0.3], [0.6, 1.0, 0.2], [-0.3, 0.2, 1.5]])[:d, :d]
    cvec = np.array([0.7, -1.1, 0.4])[:d]

    def u(X):
        return 0.3 + sum(cvec[a] * X[a] for a in range(d))

    sigma = 30.0
    over = 1 if geometry in ("quad", "hex") else 0
    grid = H.Grid.structured(geometry, [0] * d, [1] * d, cells)
    space = H.DiscontinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element_and_intersection)
    op += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(Kc))
    op.append(H.LocalCouplingIntersectionIntegralBilinearForm(
        H.LocalLaplaceIPDGInnerCoupling(1.0, Kc, Kc) + H.LocalIPDGInnerPenalty(sigma, Kc)), {},
        H.ApplyOn.InnerIntersectionsOnce())
    op.append(H.LocalIntersectionIntegralBilinearForm(
        H.LocalIPDGBoundaryPenalty(sigma, Kc) + H.LocalLaplaceIPDGDirichletCoupling(1.0, Kc)), {},
        H.ApplyOn.BoundaryIntersections())
    g = H.GenericFunction(1, u)
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalLaplaceIPDGDirichletCoupling(1.0, Kc, g),
                                                     over_integrate=over), {}, H.ApplyOn.BoundaryIntersections())
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalIPDGBoundaryPenalty(sigma, Kc).with_ansatz(g)), {},
               H.ApplyOn.BoundaryIntersections())
    op.append(rhs)
    op.assemble()
    m = O.Mesh.structured(geom, cells, [0] * d, [1] * d)
    Kt = np.ascontiguousarray(np.broadcast_to(Kc, (m.ne, d, d)))
    s = O.Space(m, False, 1)
    oop = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    oop.append_boundary_source(0, 1.0, 1.0, u(np.moveaxis(m.boundary_quadrature_points(1 + over), -1, 0)),
                               g_order=1, over_integrate=over)
    oop.set_boundary_source_tensor(Kt)
    oop.append_boundary_source(1, sigma, 1.0, u(np.moveaxis(m.boundary_quadrature_points(2), -1, 0)), g_order=1)
    oop.set_boundary_source_tensor(Kt)
    oop.assemble(1)
    berr = np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() / np.abs(oop.rhs()).max()
    x = op.apply_inverse("rhs", opts={"precision": 1e-13}).cpu().numpy()
    nodes = m.coords()[m.conn()]
    err = np.abs(x[space.global_indices()] - u(np.moveaxis(nodes, -1, 0))).max()
    print(f"  rhs vs oracle {berr:.2e}, solution vs u {err:.2e}")
    assert berr <= 1e-12 and err <= 1e-9
