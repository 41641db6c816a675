"""Matrix-free operator application (format "matrix_free": the operator is recomputed from the coefficient
fields at every application, element by element and intersection by intersection, as
LocalizableOperator::apply does, operators/localizable-operator.hh:352-379) on the HIP path.

Bar: y = A x equal to the oracle's assembled matrix times x up to round-off, row by row
|y_i - (A x)_i| <= 1e-12 sum_j |a_ij x_j| (the sums are associated differently from the assembled
matrix's); apply_inverse within 1e-10 relative L2 of the oracle's Jacobi-PCG solution with the iteration
count within 2; combinations outside the matrix-free set raise not_supported; the CSR values are refused
after a matrix-free assembly and come back after an assembly in another format.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

import oracle_lib as O
from problems import hgdt_problem, oracle_problem

pytestmark = pytest.mark.gpu


def _csr(oop):
    rp, cols = oop.pattern()
    n = len(rp) - 1
    return sp.csr_matrix((oop.values(), cols, rp), shape=(n, n))


def _apply_error(H, op, A, seed):
    x = np.random.default_rng(seed).standard_normal(A.shape[0])
    y = op.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    scale = abs(A) @ np.abs(x)
    err = np.abs(y - A @ x) / np.maximum(scale, 1e-300)
    return float(err.max())


def _solve_check(op, oop, label, precision=1e-10):
    rp, cols = oop.pattern()
    xo, ito, relo, rc = O.pcg_jacobi(rp, cols, oop.values(), oop.rhs(), rtol=precision, max_iter=200000)
    x = op.apply_inverse("rhs", opts={"precision": precision, "max_iter": 200000}).cpu().numpy()
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"  {label}: iters {op.last_solve['iterations']} (oracle {ito}) rel L2 {rel:.2e}")
    assert rel <= 1e-10 and abs(op.last_solve["iterations"] - ito) <= 2, label


def _kappa(H, n, seed, channels=False, nx=None):
    if channels:
        # rows of cells with kappa 1e-4 in a background of 1 (the high-contrast SWIPDG workload, small)
        k = np.ones(n)
        rows = np.random.default_rng(seed).choice(n // nx, size=max(1, n // nx // 5), replace=False)
        for r in rows:
            k[r * nx:(r + 1) * nx] = 1e-4
        return k
    return H.fields.lognormal(seed, n)


@pytest.mark.parametrize("cells,channels", [([13, 9], False), ([16, 16], True)])
def test_swipdg_matrix_free_parity(hgdt_mod, cells, channels):
    H = hgdt_mod
    print(f"case matrix-free SWIPDG {cells} channels={channels}", flush=True)
    lo, hi = [0.0, 0.0], [1.5, 1.0]
    kappa = _kappa(H, cells[0] * cells[1], 3, channels, cells[0])
    grid, space, op = hgdt_problem(H, "swipdg", cells, torch.from_numpy(kappa).cuda(), lo=lo, hi=hi)
    op.set_format("matrix_free")
    op.assemble()
    m, s, oop = oracle_problem("swipdg", cells, kappa, lo=lo, hi=hi)
    oop.assemble(1)
    assert np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() <= 1e-12 * np.abs(oop.rhs()).max()
    err = _apply_error(H, op, _csr(oop), 5)
    print(f"  apply row-relative error {err:.2e}")
    assert err <= 1e-12
    assert op.format_info()["format"] == "matrix_free"
    _solve_check(op, oop, "swipdg")


@pytest.mark.parametrize("symmetry", [0.0, -1.0])
def test_swipdg_variants_apply(hgdt_mod, symmetry):
    """IIPG / NIPG symmetry prefactors (the transposed face terms weighted differently)"""
    H = hgdt_mod
    print(f"case matrix-free IPDG symmetry {symmetry}", flush=True)
    cells = [7, 11]
    kappa = H.fields.lognormal(19, cells[0] * cells[1])
    grid, space, op = hgdt_problem(H, "swipdg", cells, torch.from_numpy(kappa).cuda(), symmetry=symmetry,
                                   hi=[1.0, 2.0])
    op.set_format("matrix_free")
    op.assemble()
    m, s, oop = oracle_problem("swipdg", cells, kappa, symmetry=symmetry, hi=[1.0, 2.0])
    oop.assemble(1)
    err = _apply_error(H, op, _csr(oop), 6)
    print(f"  apply row-relative error {err:.2e}")
    assert err <= 1e-12


def test_dg_generic_forms_matrix_free(hgdt_mod):
    """the general form set (runtime form counts): a scaled product element form, omega different from
    kappa, scaled and over-integrated intersection forms; apply against the oracle's assembled operator"""
    H = hgdt_mod
    print("case matrix-free DG generic forms", flush=True)
    cells, hi = [19, 13], [1.3, 0.7]
    ne = cells[0] * cells[1]
    kappa = H.fields.lognormal(43, ne)
    omega = H.fields.uniform(44, ne) + 0.25
    w = H.fields.uniform(45, ne) + 0.5
    tk, to, tw = (torch.from_numpy(v).cuda() for v in (kappa, omega, w))
    grid = H.Grid.structured("quad", [0, 0], hi, cells)
    space = H.DiscontinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element_and_intersection)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(tk)))
    op.append(H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand(tw)), {"matrixoperator.scaling": 0.3})
    op.append(H.LocalCouplingIntersectionIntegralBilinearForm(H.LocalLaplaceIPDGInnerCoupling(1.0, tk, to),
                                                              over_integrate=1),
              {"matrixoperator.scaling": 0.5}, H.ApplyOn.InnerIntersectionsOnce())
    op.append(H.LocalCouplingIntersectionIntegralBilinearForm(H.LocalIPDGInnerPenalty(12.0, to)),
              {"matrixoperator.scaling": 2.0}, H.ApplyOn.InnerIntersectionsOnce())
    boundary = H.ApplyOn.CustomBoundaryIntersections(H.AllDirichletBoundaryInfo())
    op.append(H.LocalIntersectionIntegralBilinearForm(H.LocalLaplaceIPDGDirichletCoupling(1.0, tk), over_integrate=2),
              {}, boundary)
    op.append(H.LocalIntersectionIntegralBilinearForm(H.LocalIPDGBoundaryPenalty(20.0, to)), {}, boundary)
    op.set_format("matrix_free")
    op.assemble()
    m = O.Mesh.structured(O.QUAD, cells, [0, 0], hi)
    s = O.Space(m, False, 1)
    oop = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    oop.append_laplace(kappa)
    oop.append_product(w, scaling=0.3)
    oop.append_inner_coupling(1.0, kappa, omega, over_integrate=1, scaling=0.5)
    oop.append_inner_penalty(12.0, omega, scaling=2.0)
    oop.append_dirichlet_coupling(1.0, kappa, over_integrate=2)
    oop.append_boundary_penalty(20.0, omega)
    oop.assemble(1)
    err = _apply_error(H, op, _csr(oop), 46)
    print(f"  apply row-relative error {err:.2e}")
    assert err <= 1e-12


def test_matrix_free_unaligned_vectors(hgdt_mod):
    """source and range at an 8-byte offset (no 16-byte accesses): the same bits as aligned vectors"""
    H = hgdt_mod
    print("case matrix-free unaligned vectors", flush=True)
    cells = [70, 9]  # more than one tile of 64 elements per row
    kappa = torch.from_numpy(H.fields.lognormal(23, cells[0] * cells[1])).cuda()
    grid, space, op = hgdt_problem(H, "swipdg", cells, kappa)
    op.set_format("matrix_free")
    op.assemble()
    n = space.num_DoFs
    x = torch.randn(n, dtype=torch.float64, device="cuda")
    y1 = op.apply(x)
    xb = torch.empty(n + 1, dtype=torch.float64, device="cuda")
    yb = torch.empty(n + 1, dtype=torch.float64, device="cuda")
    xb[1:] = x
    op.apply(xb[1:], yb[1:])
    assert torch.equal(y1, yb[1:])


def test_heat_q1_matrix_free_parity(hgdt_mod):
    H = hgdt_mod
    print("case matrix-free heat Q1", flush=True)
    cells, lo, hi = [17, 11], [0.0, 0.0], [2.0, 1.0]
    kappa = H.fields.lognormal(7, cells[0] * cells[1])
    grid, space, op = hgdt_problem(H, "heat_q1", cells, torch.from_numpy(kappa).cuda(), lo=lo, hi=hi)
    op.set_format("matrix_free")
    op.assemble()
    m, s, oop = oracle_problem("heat_q1", cells, kappa, lo=lo, hi=hi)
    oop.assemble(1)
    assert np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() <= 1e-12 * np.abs(oop.rhs()).max()
    err = _apply_error(H, op, _csr(oop), 8)
    print(f"  apply row-relative error {err:.2e}")
    assert err <= 1e-12
    _solve_check(op, oop, "heat q1")


def test_hex_q1_laplace_plus_product(hgdt_mod):
    """Q1 on hexahedra: Laplace with a per-element field plus a scaled product form, zero Dirichlet"""
    H = hgdt_mod
    print("case matrix-free hex Q1", flush=True)
    cells, hi = [6, 5, 4], [1.0, 1.2, 0.8]
    ne = int(np.prod(cells))
    kappa = H.fields.lognormal(9, ne)
    w = H.fields.uniform(10, ne) + 0.5
    grid = H.Grid.structured("hex", [0, 0, 0], hi, cells)
    space = H.ContinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(torch.from_numpy(kappa).cuda())))
    op.append(H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand(torch.from_numpy(w).cuda())),
              {"matrixoperator.scaling": 0.75})
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    op.set_format("matrix_free")
    op.assemble()
    m = O.Mesh.structured(O.HEX, cells, [0, 0, 0], hi)
    s = O.Space(m, True, 1)
    oop = O.Operator(s, O.STENCIL_ELEMENT)
    oop.append_laplace(kappa)
    oop.append_product(w, scaling=0.75)
    oop.append_source(1.0)
    oop.set_dirichlet(True)
    oop.assemble(1)
    err = _apply_error(H, op, _csr(oop), 11)
    print(f"  apply row-relative error {err:.2e}")
    assert err <= 1e-12
    _solve_check(op, oop, "hex q1")


@pytest.mark.parametrize("kind", ["heat_q1", "swipdg"])
def test_ragged_union_matrix_free(hgdt_mod, kind):
    """a union of structured grids of different sizes (different cell widths): every instance solved with
    its own scalars; the matrix-free solution of each instance within 1e-10 of its assembled solve"""
    H = hgdt_mod
    print(f"case matrix-free union {kind}", flush=True)
    sizes = [3, 12, 1, 7] if kind == "heat_q1" else [4, 9, 2]
    parts = [H.Grid.structured("quad", [0, 0], [1, 1 + 0.25 * i], [n, n + 1]) for i, n in enumerate(sizes)]
    union = H.Grid.union(parts)
    kappa = torch.from_numpy(H.fields.lognormal(13, union.num_elements)).cuda()
    _, space, op_mf = hgdt_problem(H, kind, None, kappa, grid=union)
    _, _, op_as = hgdt_problem(H, kind, None, kappa, grid=union)
    op_mf.set_format("matrix_free")
    op_mf.assemble()
    op_as.assemble()
    A = op_as.matrix.to_scipy()
    err = _apply_error(H, op_mf, A, 14)
    print(f"  apply row-relative error vs assembled {err:.2e}")
    assert err <= 1e-12
    # frequent checks, so that the launches are narrowed to the unfinished instances while they iterate
    x_mf = op_mf.apply_inverse("rhs", check_every=2).cpu().numpy()
    assert op_mf.last_solve["compactions"] >= 1
    x_as = op_as.apply_inverse("rhs").cpu().numpy()
    its_mf, _ = op_mf.instance_stats()
    its_as, _ = op_as.instance_stats()
    doff = space.instance_dofs()
    for i in range(len(sizes)):
        a, b = x_mf[doff[i]:doff[i + 1]], x_as[doff[i]:doff[i + 1]]
        nrm = np.linalg.norm(b)
        assert (np.linalg.norm(a - b) / nrm if nrm > 0 else np.linalg.norm(a)) <= 1e-10
        assert abs(int(its_mf[i]) - int(its_as[i])) <= 2


def test_matrix_free_graph_replay_matches_eager(hgdt_mod):
    H = hgdt_mod
    print("case matrix-free graph replay", flush=True)
    cells = [24, 20]
    kappa = torch.from_numpy(H.fields.lognormal(21, cells[0] * cells[1])).cuda()
    grid, space, op = hgdt_problem(H, "swipdg", cells, kappa)
    op.set_format("matrix_free")
    op.assemble()
    x1 = op.apply_inverse("rhs").cpu().numpy()
    it1 = op.last_solve["iterations"]
    for _ in range(3):
        x2 = op.apply_inverse("rhs", use_graph=True, check_every=8).cpu().numpy()
        assert op.last_solve["iterations"] == it1
        assert np.array_equal(x1, x2)


def test_matrix_free_mode_rules(hgdt_mod):
    H = hgdt_mod
    print("case matrix-free mode rules", flush=True)
    cells = [5, 4]
    grid, space, op = hgdt_problem(H, "heat_q1", cells, 1.0)
    op.set_format("matrix_free")
    op.assemble()
    with pytest.raises(H.operator_error):
        op.matrix
    # another format needs the matrix: refused until assembled again, then the values are there
    op.set_format("auto")
    with pytest.raises(H.operator_error):
        op.apply(torch.ones(space.num_DoFs, dtype=torch.float64).cuda())
    op.assemble()
    m, s, oop = oracle_problem("heat_q1", cells, 1.0)
    oop.assemble(1)
    assert np.abs(op.matrix.values.cpu().numpy() - oop.values()).max() <= 1e-12 * np.abs(oop.values()).max()
    # combinations outside the matrix-free set
    unsupported = []
    g_tet = H.Grid.structured("tet", [0, 0, 0], [1, 1, 1], [2, 2, 2])
    unsupported.append(hgdt_problem(H, "p1_tets", None, 1.0, grid=g_tet)[2])
    X = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], dtype=np.float64)
    g_expl = H.Grid.from_arrays("quad", X, np.array([[0, 1, 2, 3]], dtype=np.int32))
    unsupported.append(hgdt_problem(H, "heat_q1", None, 1.0, grid=g_expl)[2])
    g2 = H.Grid.structured("quad", [0, 0], [1, 1], [3, 3])
    sp2 = H.ContinuousLagrangeSpace(g2, 2)
    op2 = H.make_matrix_operator(sp2, H.Stencil.element)
    op2.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0)))
    unsupported.append(op2)
    for o in unsupported:
        o.set_format("matrix_free")
        o.assemble()
        with pytest.raises(H.not_supported):
            o.apply(torch.ones(o.shape[0], dtype=torch.float64).cuda())


def test_matrix_free_max_iter_stop_matches_assembled(hgdt_mod):
    """a solve stopped by max_iter: the matrix-free solver (direction update inside the application kernel,
    x += alpha p deferred to the stopping step) leaves the same iterate as the assembled solver after the
    same number of steps, up to the round-off of the differently associated applications"""
    H = hgdt_mod
    print("case matrix-free max_iter stop", flush=True)
    cells = [20, 17]
    kappa = torch.from_numpy(H.fields.lognormal(23, cells[0] * cells[1])).cuda()
    xs = []
    for fmt in ("matrix_free", "auto"):
        grid, space, op = hgdt_problem(H, "swipdg", cells, kappa)
        op.set_format(fmt)
        op.assemble()
        x = torch.zeros(space.num_DoFs, dtype=torch.float64, device="cuda")
        with pytest.raises(H.linear_solver_failed):
            op.apply_inverse("rhs", x, opts={"max_iter": 40}, check_every=6)
        assert op.last_solve["iterations"] == 40 and not op.last_solve["converged"]
        xs.append(x.cpu().numpy())
    rel = np.linalg.norm(xs[0] - xs[1]) / np.linalg.norm(xs[1])
    print(f"  iterate after 40 steps, matrix-free vs assembled: rel L2 {rel:.2e}")
    assert rel <= 1e-9


def test_matrix_free_zero_rhs_instance(hgdt_mod):
    """an instance of a union with a zero right-hand side stops at step 0 with x = 0; the other instances'
    iterates do not depend on it (bitwise equal to the solve with the full right-hand side)"""
    H = hgdt_mod
    print("case matrix-free zero rhs instance", flush=True)
    sizes = [4, 9, 2]
    parts = [H.Grid.structured("quad", [0, 0], [1, 1 + 0.25 * i], [n, n + 1]) for i, n in enumerate(sizes)]
    union = H.Grid.union(parts)
    kappa = torch.from_numpy(H.fields.lognormal(29, union.num_elements)).cuda()
    _, space, op = hgdt_problem(H, "swipdg", None, kappa, grid=union)
    op.set_format("matrix_free")
    op.assemble()
    doff = space.instance_dofs()
    x_full = op.apply_inverse("rhs", check_every=2).cpu().numpy()
    its_full, _ = op.instance_stats()
    b = op.rhs.clone()
    b[doff[1]:doff[2]] = 0.
    x = op.apply_inverse(b, check_every=2).cpu().numpy()
    its, _ = op.instance_stats()
    assert its[1] == 0 and np.all(x[doff[1]:doff[2]] == 0.)
    for i in (0, 2):
        assert its[i] == its_full[i]
        assert np.array_equal(x[doff[i]:doff[i + 1]], x_full[doff[i]:doff[i + 1]])
