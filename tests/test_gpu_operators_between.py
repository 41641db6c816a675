"""MatrixOperator between two spaces and on a caller's matrix (operators/matrix-based.hh:200-216,279-311).

* Source != range: rows from the range space (test functions), columns from the source space (ansatz);
  pattern bit-exact, values within 1e-12 of the oracle's, y = A x within 1e-12; apply_inverse and Dirichlet
  constraints refuse such an operator (not_supported).
* On a caller's matrix: its pattern (here the element-and-intersection pattern, larger than the operator's
  element stencil) is the operator's; assemble() adds the contributions to the values held, in walk order
  (add_to_entry): once -> V0 + A, twice -> V0 + 2A, within 1e-12 of the oracle's A added to V0; apply and
  apply_inverse use the whole matrix (the solve within 1e-10 of scipy's); constrained rows and columns are
  set; entries outside the stencil keep V0.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from problems import matrix_error

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("geometry,cells", [("quad", [6, 5]), ("tet", [3, 3, 2])])
def test_operator_between_spaces_matches_oracle(hgdt_mod, geometry, cells):
    H = hgdt_mod
    print(f"case operators between CG and DG spaces, {geometry} {cells}", flush=True)
    dim = len(cells)
    grid = H.Grid.structured(geometry, [0] * dim, [1.0] * dim, cells)
    cg, dg = H.ContinuousLagrangeSpace(grid, 1), H.DiscontinuousLagrangeSpace(grid, 1)
    kappa = H.fields.lognormal(81, grid.num_elements)
    kd = torch.from_numpy(kappa).cuda()
    m = O.Mesh.structured(O.QUAD if geometry == "quad" else O.TET, cells, [0] * dim, [1.0] * dim)
    ocg, odg = O.Space(m, True, 1), O.Space(m, False, 1)
    rng = np.random.default_rng(6)
    for (rs, ss), (ors, oss) in (((cg, dg), (ocg, odg)), ((dg, cg), (odg, ocg))):
        op = H.make_matrix_operator(ss, H.Stencil.element, range_space=rs)
        op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(kd)))
        op.append(H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand(2.5)))
        op.assemble()
        oop = O.Operator(ors, O.STENCIL_ELEMENT, source=oss)
        oop.append_laplace(kappa)
        oop.append_product(2.5)
        oop.assemble(1)
        assert op.shape == (ors.size, oss.size)
        rp, cols = op.pattern()
        orp, ocols = oop.pattern()
        assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
        err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
        x = rng.standard_normal(oss.size)
        y = op.apply(torch.from_numpy(x).cuda()).cpu().numpy()
        yo = oop.csr() @ x
        yerr = np.abs(y - yo).max() / np.abs(yo).max()
        print(f"  {ors.size} x {oss.size}: matrix {err:.2e}, apply {yerr:.2e}")
        assert err <= 1e-12 and yerr <= 1e-12
        with pytest.raises(H.not_supported):
            op.apply_inverse(torch.zeros(ors.size, dtype=torch.float64, device="cuda"))
        with pytest.raises(H.not_supported):
            op.append(H.make_dirichlet_constraints(rs, H.AllDirichletBoundaryInfo()))
            op.assemble()


def test_operator_on_callers_matrix(hgdt_mod):
    H = hgdt_mod
    print("case MatrixOperator on a caller's matrix (Q1 on 8 x 7 quads, element stencil in the "
          "element-and-intersection pattern)", flush=True)
    cells = [8, 7]
    grid = H.Grid.structured("quad", [0, 0], [1, 1], cells)
    space = H.ContinuousLagrangeSpace(grid, 1)
    big = H.make_matrix_operator(space, H.Stencil.element_and_intersection)
    rp, cols = big.pattern()
    n = space.num_DoFs
    m = O.Mesh.structured(O.QUAD, cells, [0, 0], [1, 1])
    s = O.Space(m, True, 1)
    # V0: the oracle's mass matrix scattered into the larger pattern (SPD), plus a symmetric entry outside
    # the element stencil at every such slot
    mass = O.Operator(s, O.STENCIL_ELEMENT)
    mass.append_product(1.0)
    mass.assemble(1)
    Mm = mass.csr()
    rows = np.repeat(np.arange(n), np.diff(rp))
    v0 = np.asarray(Mm[rows, cols]).reshape(-1)
    outside = v0 == 0.
    v0[outside] = 1e-3
    vals = torch.from_numpy(v0.copy()).cuda()
    mat = H.CsrMatrix(rp, cols, vals, (n, n))
    kappa = H.fields.lognormal(82, grid.num_elements)
    op = H.MatrixOperator(grid, space, space, stencil=H.Stencil.element, matrix=mat)
    assert op.matrix is mat
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(torch.from_numpy(kappa).cuda())))
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.assemble()
    ok = O.Operator(s, O.STENCIL_ELEMENT)
    ok.append_laplace(kappa)
    ok.append_source(1.0)
    ok.assemble(1)
    K = ok.csr()
    a = np.asarray(K[rows, cols]).reshape(-1)
    got = vals.cpu().numpy()
    e1 = np.abs(got - (v0 + a)).max() / np.abs(v0 + a).max()
    assert np.array_equal(got[outside], v0[outside])
    x = np.random.default_rng(7).standard_normal(n)
    y = op.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    import scipy.sparse as sp

    full = sp.csr_matrix((v0 + a, cols, rp), shape=(n, n))
    yerr = np.abs(y - full @ x).max() / np.abs(full @ x).max()
    op.assemble()  # the reference adds again
    e2 = np.abs(vals.cpu().numpy() - (v0 + 2 * a)).max() / np.abs(v0 + 2 * a).max()
    print(f"  once {e1:.2e}, apply {yerr:.2e}, twice {e2:.2e}")
    assert e1 <= 1e-12 and yerr <= 1e-12 and e2 <= 1e-12
    # a solve with the whole matrix, and Dirichlet constraints setting rows and columns
    vals.copy_(torch.from_numpy(v0))
    op2 = H.MatrixOperator(grid, space, space, stencil=H.Stencil.element, matrix=mat)
    op2.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(torch.from_numpy(kappa).cuda())))
    rhs2 = H.make_vector_functional(space)
    rhs2.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op2.append(rhs2)
    op2.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    op2.assemble()
    A = sp.csr_matrix((vals.cpu().numpy(), cols, rp), shape=(n, n))
    dd = op2.dirichlet_dofs()
    Ad = A[dd].toarray()
    expect = np.zeros_like(Ad)
    expect[np.arange(len(dd)), dd] = 1.
    assert np.array_equal(Ad, expect) and np.array_equal(A[:, dd].toarray(), expect.T)
    u = op2.apply_inverse("rhs", opts={"precision": 1e-12}).cpu().numpy()
    import scipy.sparse.linalg as spla

    ue = spla.spsolve(A.tocsc(), op2.rhs.cpu().numpy())
    rel = np.linalg.norm(u - ue) / np.linalg.norm(ue)
    print(f"  Dirichlet rows/columns set, solve vs scipy {rel:.2e}")
    assert rel <= 1e-10
    with pytest.raises(H.shapes_do_not_match):
        small = H.make_matrix_operator(space, H.Stencil.element)
        srp, scols = small.pattern()
        H.MatrixOperator(grid, space, space, stencil=H.Stencil.element_and_intersection,
                         matrix=H.CsrMatrix(srp, scols, torch.zeros(len(scols), dtype=torch.float64,
                                                                      device="cuda"), (n, n)))
