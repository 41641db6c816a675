"""Boundary infos on the HIP path: Dirichlet constraints and boundary forms restricted to the intersections of one
boundary type (tools/dirichlet-constraints.hh:85-110 with boundary_info.type(intersection); ApplyOn::
CustomBoundaryIntersections(boundary_info, type), test/stationary-eocstudies/diffusion-ipdg.hh:302-311), with a
NormalBasedBoundaryInfo (Dirichlet on x = 0, 1, default Neumann on the rest).

Bar: Dirichlet DoFs bit-exact and matrix / right-hand side within 1e-12 of the oracle given the same face mask;
solutions equal to the discrete closed forms (x (1 - x) / 2 at the vertices for -u'' = 1 with natural conditions
on the other faces, continuous Q1 and P1 on the structured tetrahedral path; u = 0.3 + 2 x for SWIPDG).
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from problems import matrix_error

pytestmark = pytest.mark.gpu


def _x_faces(H):
    info = H.NormalBasedBoundaryInfo(default_boundary_type=H.NeumannBoundary())
    info.register_new_normal([1, 0, 0], H.DirichletBoundary())
    info.register_new_normal([-1, 0, 0], H.DirichletBoundary())
    return info


@pytest.mark.parametrize("geometry,cells,hi", [("quad", [8, 6], [1, 0.75]), ("tet", [6, 4, 3], [1, 0.8, 0.6])])
def test_mixed_dirichlet_neumann_continuous(hgdt_mod, geometry, cells, hi):
    H = hgdt_mod
    print(f"case mixed Dirichlet / Neumann, continuous {geometry} {cells}", flush=True)
    dim = len(cells)
    grid = H.Grid.structured(geometry, [0] * dim, hi, cells)
    info = _x_faces(H)
    mask = info.mask(grid, H.DirichletBoundary)
    centres, normals = grid.boundary_intersections()
    assert np.array_equal(mask.astype(bool), np.abs(np.abs(normals[:, 0]) - 1) < 1e-12)
    assert np.all(np.isclose(centres[mask == 1, 0], 0) | np.isclose(centres[mask == 1, 0], 1))
    space = H.ContinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0)))
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.append(H.make_dirichlet_constraints(space, info))
    op.assemble()
    if geometry == "tet":
        assert op.assembly_info()["structured"]
    m = O.Mesh.structured(O.QUAD if geometry == "quad" else O.TET, cells, [0] * dim, hi)
    s = O.Space(m, True, 1)
    oop = O.Operator(s, O.STENCIL_ELEMENT)
    oop.append_laplace(1.0)
    oop.append_source(1.0)
    oop.set_dirichlet(True)
    oop.set_dirichlet_faces(mask)
    oop.assemble(1)
    assert np.array_equal(op.dirichlet_dofs(), oop.dirichlet_dofs())
    rp, _ = oop.pattern()
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), rp)
    berr = np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() / np.abs(oop.rhs()).max()
    x = op.apply_inverse("rhs", opts={"precision": 1e-13}).cpu().numpy()
    X = grid.coordinates()
    cf = np.abs(x - X[:, 0] * (1 - X[:, 0]) / 2).max()
    print(f"  Dirichlet DoFs {len(op.dirichlet_dofs())}, matrix {err:.2e}, rhs {berr:.2e}, closed form {cf:.2e}")
    assert err <= 1e-12 and berr <= 1e-12 and cf <= 1e-11


@pytest.mark.parametrize("geometry", ["quad", "tri"])
def test_swipdg_boundary_forms_on_dirichlet_faces(hgdt_mod, geometry):
    H = hgdt_mod
    print(f"case SWIPDG with boundary forms and data on the Dirichlet faces only, {geometry}", flush=True)
    n, sigma = 5, 16.0
    grid = H.Grid.structured(geometry, [0, 0], [1, 1], [n, n])
    info = _x_faces(H)
    on_dirichlet = H.ApplyOn.CustomBoundaryIntersections(info, H.DirichletBoundary())
    space = H.DiscontinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element_and_intersection)
    op += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0))
    op.append(H.LocalCouplingIntersectionIntegralBilinearForm(
        H.LocalLaplaceIPDGInnerCoupling(1.0, 1.0, 1.0) + H.LocalIPDGInnerPenalty(sigma, 1.0)), {},
        H.ApplyOn.InnerIntersectionsOnce())
    op.append(H.LocalIntersectionIntegralBilinearForm(
        H.LocalIPDGBoundaryPenalty(sigma, 1.0) + H.LocalLaplaceIPDGDirichletCoupling(1.0, 1.0)), {}, on_dirichlet)

    def u(X):
        return 0.3 + 2.0 * X[0]

    g = H.GenericFunction(1, u)
    over = 1 if geometry == "quad" else 0
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalLaplaceIPDGDirichletCoupling(1.0, 1.0, g),
                                                     over_integrate=over), {}, on_dirichlet)
    rhs.append(H.LocalIntersectionIntegralFunctional(H.LocalIPDGBoundaryPenalty(sigma, 1.0).with_ansatz(g)), {},
               on_dirichlet)
    op.append(rhs)
    op.assemble()
    mask = info.mask(grid, H.DirichletBoundary)
    m = O.Mesh.structured(O.QUAD if geometry == "quad" else O.TRI, [n, n], [0, 0], [1, 1])
    s = O.Space(m, False, 1)
    oop = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    oop.append_laplace(1.0)
    oop.append_inner_coupling(1.0, 1.0, 1.0)
    oop.append_inner_penalty(sigma, 1.0)
    oop.set_boundary_filter(mask)
    oop.append_boundary_penalty(sigma, 1.0)
    oop.append_dirichlet_coupling(1.0, 1.0)
    ug = lambda P: 0.3 + 2.0 * P[..., 0]
    oop.append_boundary_source(0, 1.0, 1.0, ug(m.boundary_quadrature_points(1 + over)), g_order=1,
                               over_integrate=over)
    oop.append_boundary_source(1, sigma, 1.0, ug(m.boundary_quadrature_points(2)), g_order=1)
    oop.assemble(1)
    rp, _ = oop.pattern()
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), rp)
    berr = np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() / np.abs(oop.rhs()).max()
    x = op.apply_inverse("rhs", opts={"precision": 1e-14, "max_iter": 100000}).cpu().numpy()
    Xe = grid.coordinates()[grid.connectivity()]
    rep = np.abs(x[space.global_indices()] - u(np.moveaxis(Xe, -1, 0))).max()
    print(f"  matrix {err:.2e}, rhs {berr:.2e}, linear solution reproduced to {rep:.2e}")
    assert err <= 1e-12 and berr <= 1e-12 and rep <= 1e-9


@pytest.mark.parametrize("only_clear,ensure_symmetry", [(False, False), (True, True), (True, False), (False, True)])
def test_dirichlet_apply_variants_match_oracle(hgdt_mod, only_clear, ensure_symmetry):
    """DirichletConstraints::apply(matrix, vector, only_clear, ensure_symmetry) (dirichlet-constraints.hh:
    152-184) through constraints.apply(...) after assembly (the operator is assembled again with the variant):
    matrix and rhs bit-exact against the oracle's on the constrained slots, 1e-12 elsewhere; apply() of the
    variant equals the oracle matrix's product; Jacobi-PCG refuses the non-symmetric and the cleared variants"""
    H = hgdt_mod
    print(f"case Dirichlet apply only_clear={only_clear} ensure_symmetry={ensure_symmetry}", flush=True)
    cells = [9, 7]
    kappa = H.fields.lognormal(5, cells[0] * cells[1])
    grid = H.Grid.structured("quad", [0, 0], [1, 1], cells)
    space = H.ContinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(torch.from_numpy(kappa).cuda())))
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    constraints = H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo())
    op.append(constraints)
    op.assemble()
    constraints.apply(op.matrix, op.rhs, only_clear, ensure_symmetry)
    m = O.Mesh.structured(O.QUAD, cells, [0, 0], [1, 1])
    s = O.Space(m, True, 1)
    oop = O.Operator(s, O.STENCIL_ELEMENT)
    oop.append_laplace(kappa)
    oop.append_source(1.0)
    oop.set_dirichlet(True)
    oop.set_dirichlet_mode(only_clear, ensure_symmetry)
    oop.assemble(1)
    rp, cols = oop.pattern()
    assert np.array_equal(op.pattern()[0], rp) and np.array_equal(op.pattern()[1], cols)
    v, ov = op.matrix.values.cpu().numpy(), oop.values()
    dof = np.zeros(len(rp) - 1, bool)
    dof[oop.dirichlet_dofs()] = True
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    constrained = dof[rows] | (ensure_symmetry & dof[cols])
    assert np.array_equal(v[constrained], ov[constrained])
    assert matrix_error(v, ov, rp) <= 1e-12
    assert np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() <= 1e-12 * np.abs(oop.rhs()).max()
    x = torch.from_numpy(np.linspace(-1, 1, len(rp) - 1)).cuda()
    y = op.apply(x).cpu().numpy()
    import scipy.sparse as sp

    yo = sp.csr_matrix((ov, cols, rp), shape=(len(rp) - 1,) * 2) @ x.cpu().numpy()
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()
    if (only_clear, ensure_symmetry) == (False, True):
        op.apply_inverse("rhs")
    else:
        with pytest.raises(H._lib.HgdtError):
            op.apply_inverse("rhs")
