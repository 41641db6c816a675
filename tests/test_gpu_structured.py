"""Structured assembly straight into the solver's storage, and the solver's compact symmetric DIA copy.

* P1 Laplace on the Kuhn tetrahedra of axis-parallel lattices (single grids and unions) is assembled by one
  pass over the DoFs into symmetric DIA storage; the CSR values (computed from that storage when asked for)
  must match the oracle's to 1e-12 (row-relative), the right-hand side to 1e-12, and the solve must match
  the oracle's Jacobi-PCG to 1e-10 with its iteration count (+-2).  Several Laplace forms with scalings and
  per-element / constant coefficients, and several constant / per-element sources, are summed.
* The same operator assembled through the generic element-walk + fixed-order scatter (forced by requesting
  non-symmetric storage) gives the same matrix to rounding (1e-14 row-relative).
* The solver's scaled copy keeps only strict upper diagonals with a nonzero entry: 3 of 7 for P1 Laplace on
  Kuhn tetrahedra (the face and body diagonals are exactly 0, in the oracle's assembly too), all 7 once a
  product (mass) form couples every edge; the solution is bitwise equal to the full non-symmetric DIA solve.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from problems import matrix_error

pytestmark = pytest.mark.gpu


def _operator(H, grid, forms, sources, dirichlet=True):
    space = H.ContinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    for kind, coeff, scaling in forms:
        integrand = H.LocalLaplaceIntegrand(coeff) if kind == "laplace" else H.LocalElementProductIntegrand(coeff)
        op.append(H.LocalElementIntegralBilinearForm(integrand), {"matrixoperator.scaling": scaling})
    rhs = H.make_vector_functional(space)
    for f in sources:
        rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(f)))
    op.append(rhs)
    if dirichlet:
        op.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    return space, op


def _oracle(mesh, forms, sources, dirichlet=True):
    s = O.Space(mesh, True, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    for kind, coeff, scaling in forms:
        (op.append_laplace if kind == "laplace" else op.append_product)(coeff, scaling=scaling)
    for f in sources:
        op.append_source(f)
    op.set_dirichlet(dirichlet)
    op.assemble(4)
    return s, op


CELLS, LO, HI = [7, 5, 6], [0.1, -0.3, 0.2], [1.5, 0.2, 2.0]


@pytest.mark.parametrize("dirichlet", [True, False])
def test_structured_assembly_matches_oracle(hgdt_mod, dirichlet):
    H = hgdt_mod
    print(f"case structured P1 tets {CELLS} on an anisotropic box, two Laplace forms, two sources, "
          f"dirichlet={dirichlet}", flush=True)
    grid = H.Grid.structured("tet", LO, HI, CELLS)
    ne = grid.num_elements
    k1 = H.fields.lognormal(71, ne)
    f2 = H.fields.uniform(72, ne)
    forms = [("laplace", k1, 0.7), ("laplace", 2.5, 1.3)]
    sources = [1.5, f2]
    gpu_forms = [(k, c if np.isscalar(c) else torch.from_numpy(c).cuda(), s) for k, c, s in forms]
    space, op = _operator(H, grid, gpu_forms, [1.5, torch.from_numpy(f2).cuda()], dirichlet)
    op.assemble()
    assert op.assembly_info()["structured"]
    m = O.Mesh.structured(O.TET, CELLS, LO, HI)
    s, oop = _oracle(m, forms, sources, dirichlet)
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    ob = oop.rhs()
    berr = np.abs(op.rhs.cpu().numpy() - ob).max() / np.abs(ob).max()
    print(f"  matrix max rel err {err:.2e}, rhs max rel err {berr:.2e}")
    assert err <= 1e-12 and berr <= 1e-12
    if not dirichlet:
        return  # the pure Neumann Laplacian is singular: no solve
    x = op.apply_inverse("rhs").cpu().numpy()
    info = op.assembly_info()
    xo, ito, relo, rc = O.pcg_jacobi(orp, ocols, oop.values(), ob, rtol=1e-10)
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"  solver upper diagonals {info['solver_upper_diagonals']}, iterations {op.last_solve['iterations']} "
          f"(oracle {ito}), solution rel L2 {rel:.2e}, true relres {op.last_solve['true_relres']:.2e}")
    assert info["solver_upper_diagonals"] == 3
    assert rel <= 1e-10 and abs(op.last_solve["iterations"] - ito) <= 2
    assert op.last_solve["true_relres"] <= 2e-10


def test_structured_union_and_generic_path_agree(hgdt_mod):
    """a union of three lattices of different sizes and widths: structured vs generic assembly of the same
    operator (the generic one forced by non-symmetric storage) agree to rounding, and each instance's solve
    matches the oracle"""
    H = hgdt_mod
    print("case structured union of 3 tetrahedral lattices vs the generic assembly", flush=True)
    parts = [H.Grid.structured("tet", [0, 0, 0], [1.0 + 0.3 * i, 1.0, 0.8], c)
             for i, c in enumerate([[3, 4, 2], [6, 5, 7], [2, 2, 2]])]
    union = H.Grid.union(parts)
    kappa = H.fields.lognormal(73, union.num_elements)
    forms = [("laplace", torch.from_numpy(kappa).cuda(), 1.0)]
    space, op = _operator(H, union, forms, [1.0])
    op.assemble()
    assert op.assembly_info()["structured"]
    vals = op.matrix.values.cpu().numpy()
    x = op.apply_inverse("rhs").cpu().numpy()
    assert op.assembly_info()["solver_upper_diagonals"] == 3
    its = op.instance_stats()[0]
    space2, op2 = _operator(H, union, forms, [1.0])
    op2.set_format(H._lib.FORMAT_DIA_NONSYMMETRIC)
    op2.assemble()
    assert not op2.assembly_info()["structured"]
    rp, _ = op2.pattern()
    err = matrix_error(vals, op2.matrix.values.cpu().numpy(), rp)
    print(f"  structured vs generic matrix max rel diff {err:.2e}")
    assert err <= 1e-14
    eoff, doff = union.instance_elements(), space.instance_dofs()
    for i, p in enumerate(parts):
        lo, hi = [0, 0, 0], [1.0 + 0.3 * i, 1.0, 0.8]
        m = O.Mesh.structured(O.TET, [[3, 4, 2], [6, 5, 7], [2, 2, 2]][i], lo, hi)
        s, oop = _oracle(m, [("laplace", kappa[eoff[i]:eoff[i + 1]], 1.0)], [1.0])
        orp, ocols = oop.pattern()
        xo, ito, relo, rc = O.pcg_jacobi(orp, ocols, oop.values(), oop.rhs(), rtol=1e-10)
        rel = np.linalg.norm(x[doff[i]:doff[i + 1]] - xo) / np.linalg.norm(xo)
        print(f"  instance {i}: iterations {its[i]} (oracle {ito}), rel L2 {rel:.2e}")
        assert rel <= 1e-10 and abs(its[i] - ito) <= 2


def test_solver_copy_keeps_nonzero_diagonals(hgdt_mod):
    """P1 Laplace + mass on Kuhn tetrahedra (generic assembly: a product form couples every edge): all 7 strict
    upper diagonals are kept; the compact Laplace-only copy (3 kept) solves to the bits of the full
    non-symmetric DIA solve of the same matrix"""
    H = hgdt_mod
    print("case solver copy: kept diagonals with and without a mass form", flush=True)
    grid = H.Grid.structured("tet", [0, 0, 0], [1, 1, 1], [6, 7, 5])
    kappa = torch.from_numpy(H.fields.lognormal(74, grid.num_elements)).cuda()
    _, opm = _operator(H, grid, [("laplace", kappa, 1.0), ("product", 3.0, 1.0)], [1.0])
    opm.assemble()
    assert not opm.assembly_info()["structured"]
    opm.apply_inverse("rhs")
    assert opm.assembly_info()["solver_upper_diagonals"] == 7
    _, op = _operator(H, grid, [("laplace", kappa, 1.0)], [1.0])
    op.assemble()
    x = op.apply_inverse("rhs").cpu().numpy()
    it = op.last_solve["iterations"]
    assert op.assembly_info()["solver_upper_diagonals"] == 3
    op.set_format(H._lib.FORMAT_DIA_NONSYMMETRIC)  # the same CSR values, every diagonal read
    xd = op.apply_inverse("rhs").cpu().numpy()
    assert op.format_info()["format"] == "dia" and op.last_solve["iterations"] == it
    assert np.array_equal(x, xd)


def test_reported_algorithmic_bytes(hgdt_mod):
    """the algorithmic bytes the library reports for an assembly and a PCG iteration: the structured pass
    reads the coefficient field (8 B per element) and the Dirichlet mask (1 B per DoF) and writes 88 B per
    DoF; the generic path's compulsory bytes are 8 per element and field, 8 per nonzero and 8 per DoF; an
    iteration moves the application's bytes plus 36 B per DoF (direction update) plus 32 B per DoF (update),
    and the matrix-free SWIPDG application, which does the direction update itself, plus 32 B per DoF only"""
    H = hgdt_mod
    print("case reported algorithmic bytes", flush=True)
    cells = [10, 9, 8]
    grid = H.Grid.structured("tet", [0, 0, 0], [1, 1, 1], cells)
    kappa = torch.from_numpy(H.fields.lognormal(3, grid.num_elements)).cuda()
    space, op = _operator(H, grid, [("laplace", kappa, 1.0)], [1.0])
    op.assemble()
    n, ne = space.num_DoFs, grid.num_elements
    assert op.assembly_info()["structured"]
    assert op.assembly_bytes() == 8 * ne + 89 * n
    op.apply_inverse("rhs")
    st = op.last_solve
    assert st["step_bytes"] == (3 * 8 + 16) * n
    assert st["iteration_bytes"] == st["step_bytes"] + 68 * n
    # the generic element walk (non-symmetric storage requested)
    space2, op2 = _operator(H, grid, [("laplace", kappa, 1.0)], [1.0])
    op2.set_format(H._lib.FORMAT_DIA_NONSYMMETRIC)
    op2.assemble()
    assert not op2.assembly_info()["structured"]
    assert op2.assembly_bytes() == 8 * ne + 8 * op2.nnz + 8 * n
    # matrix-free SWIPDG on quadrilaterals
    from problems import hgdt_problem

    cells2 = [12, 10]
    k2 = torch.from_numpy(H.fields.lognormal(5, cells2[0] * cells2[1])).cuda()
    _, s3, op3 = hgdt_problem(H, "swipdg", cells2, k2)
    op3.set_format("matrix_free")
    op3.assemble()
    assert op3.assembly_bytes() == 8 * cells2[0] * cells2[1] + 8 * s3.num_DoFs
    op3.apply_inverse("rhs")
    st3 = op3.last_solve
    assert st3["iteration_bytes"] == st3["step_bytes"] + 32 * s3.num_DoFs
