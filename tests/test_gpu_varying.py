"""Element forms with a coefficient varying inside the elements on the HIP path: LocalLaplaceIntegrand and
LocalElementProductIntegrand with a GridFunction of order c (laplace.hh:74-101, product.hh:88-128), given
as a GenericFunction evaluated at the points of the element rules of order c + 2 k + over_integrate.

Bar: pattern bit-exact, matrix and right-hand side within 1e-12 of the oracle given the same values at the
same points, Jacobi-PCG solution within 1e-10 of the oracle's with its iteration count (+-2); on quadrilaterals
and triangles, on hexahedra (Q2: row lanes) and tetrahedra (P1: the quadrature loop, not the one-point
shortcut of a constant coefficient), and on a mixed triangle / quadrilateral grid.  Refused: the matrix-free
format, the adjoint gradient and intersection integrands (their coefficients are per element).
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from mixed_grids import strip_mixed_grid
from problems import matrix_error

pytestmark = pytest.mark.gpu


def kappa_fn(X):
    v = 1.0 + X[0] * X[0] - 0.5 * X[0] * X[1]
    return v + 0.3 * X[2] if len(X) > 2 else v


CASES = [("quad", O.QUAD, [6, 5], 1), ("tri", O.TRI, [5, 4], 2), ("hex", O.HEX, [3, 3, 2], 1),
         ("hex", O.HEX, [2, 2, 2], 2), ("tet", O.TET, [3, 2, 2], 1), ("tet", O.TET, [2, 2, 2], 2)]


def _hgdt(H, grid, k, over=0):
    space = H.ContinuousLagrangeSpace(grid, k)
    op = H.make_matrix_operator(space, H.Stencil.element)
    kap = H.GenericFunction(2, kappa_fn)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(kap), over_integrate=over))
    op.append(H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand(kap)), {"matrixoperator.scaling": 0.5})
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    return space, op


def _oracle(m, grid, k, over=0):
    d = grid.dimension
    s = O.Space(m, True, k)
    oop = O.Operator(s, O.STENCIL_ELEMENT)
    # the values the library evaluated: the GenericFunction at the library's points
    v1 = kappa_fn(grid.quadrature_points(2 + 2 * k + over).reshape(-1, d).T)
    v2 = kappa_fn(grid.quadrature_points(2 + 2 * k).reshape(-1, d).T)
    oop.append_laplace_points(v1, 2, over_integrate=over)
    oop.append_product_points(v2, 2, scaling=0.5)
    oop.append_source(1.0)
    oop.set_dirichlet(True)
    oop.assemble(1)
    return oop


def _check(op, oop, label):
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    berr = np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() / np.abs(oop.rhs()).max()
    x = op.apply_inverse("rhs").cpu().numpy()
    xo, ito, relo, rc = O.pcg_jacobi(rp, cols, oop.values(), oop.rhs(), rtol=1e-10)
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    its = op.last_solve["iterations"]
    print(f"  {label}: matrix {err:.2e}, rhs {berr:.2e}, solution {rel:.2e}, iterations {its} (oracle {ito})")
    assert err <= 1e-12 and berr <= 1e-12 and rel <= 1e-10 and abs(its - ito) <= 2


@pytest.mark.parametrize("geometry,geom,cells,k", CASES)
def test_varying_coefficient_matches_oracle(hgdt_mod, geometry, geom, cells, k):
    H = hgdt_mod
    print(f"case varying coefficient {geometry} cells={cells} order {k}", flush=True)
    d = len(cells)
    lo, hi = [0.0] * d, [1.0, 0.8, 1.2][:d]
    grid = H.Grid.structured(geometry, lo, hi, cells)
    over = 1 if k == 1 else 0
    space, op = _hgdt(H, grid, k, over)
    op.assemble()
    assert not op.assembly_info()["structured"]
    m = O.Mesh.structured(geom, cells, lo, hi)
    _check(op, _oracle(m, grid, k, over), f"{space.num_DoFs} DoFs")


@pytest.mark.parametrize("k", [1, 2])
def test_varying_coefficient_on_mixed_grid(hgdt_mod, k):
    H = hgdt_mod
    print(f"case varying coefficient mixed grid order {k}", flush=True)
    V, E = strip_mixed_grid(6, 5, shear=0.2)
    grid = H.Grid.mixed(V, E)
    space, op = _hgdt(H, grid, k)
    op.assemble()
    _check(op, _oracle(O.Mesh.mixed(V, E), grid, k), f"mixed, {space.num_DoFs} DoFs")


@pytest.mark.parametrize("geometry", ["quad", "tri"])
def test_constant_points_equal_per_element_field(hgdt_mod, geometry):
    """values equal at all points of an element give the per-element form: the same bits on quadrilaterals
    (the same quadrature sum); to rounding on triangles, where a per-element P1 Laplace form is one evaluation
    (constant gradients) and the values at points are summed over the rule"""
    H = hgdt_mod
    print(f"case varying coefficient: constant per element on {geometry}", flush=True)
    grid = H.Grid.structured(geometry, [0, 0], [1, 1], [7, 6])
    kap = 1.0 + 0.1 * np.arange(grid.num_elements)
    space = H.ContinuousLagrangeSpace(grid, 1)
    a = H.make_matrix_operator(space, H.Stencil.element)
    a.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(torch.from_numpy(kap).cuda())))
    a.assemble()
    b = H.make_matrix_operator(space, H.Stencil.element)
    nq = grid.quadrature_points(2).shape[1]
    qp = H.QuadraturePointFunction(grid, np.repeat(kap, nq), 2)
    b.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(qp)))
    b.assemble()
    va, vb = a.matrix.values.cpu().numpy(), b.matrix.values.cpu().numpy()
    if geometry == "quad":
        assert np.array_equal(va, vb)
    else:
        assert np.abs(va - vb).max() <= 1e-14 * np.abs(va).max()


def test_varying_coefficient_refusals(hgdt_mod):
    H = hgdt_mod
    L = H._lib
    print("case varying coefficient refusals", flush=True)
    kap = H.GenericFunction(1, lambda X: 1.0 + X[0])
    grid = H.Grid.structured("quad", [0, 0], [1, 1], [4, 4])
    space = H.ContinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(kap)))
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    op.set_format("matrix_free")
    op.assemble()
    with pytest.raises(L.HgdtError):
        op.apply(torch.ones(space.num_DoFs, dtype=torch.float64, device="cuda"))
    op.set_format("auto")
    op.assemble()
    u = op.apply_inverse("rhs")
    with pytest.raises(L.HgdtError):
        op.adjoint_log_diffusion_gradient(u, u)
    dspace = H.DiscontinuousLagrangeSpace(grid, 1)
    dop = H.make_matrix_operator(dspace, H.Stencil.element_and_intersection)
    with pytest.raises(L.HgdtError):
        dop.append(H.LocalCouplingIntersectionIntegralBilinearForm(H.LocalLaplaceIPDGInnerCoupling(1.0, kap, 1.0)), {},
                   H.ApplyOn.InnerIntersectionsOnce())


def test_wrong_number_of_values_is_refused(hgdt_mod):
    H = hgdt_mod
    print("case varying coefficient: wrong number of values", flush=True)
    grid = H.Grid.structured("quad", [0, 0], [1, 1], [3, 3])
    space = H.ContinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    nq = grid.quadrature_points(2).shape[1]
    qp = H.QuadraturePointFunction(grid, np.ones(grid.num_elements * nq - 1), 2)
    with pytest.raises(H._lib.HgdtError):
        op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(qp)))
