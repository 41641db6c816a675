"""Parity of the HIP path (libhgdt.so through the hgdt API) with the CPU oracle.

Bar: sparsity pattern and Dirichlet DoFs bit-exact; matrix entries within 1e-12 relative to the
row's largest entry; right-hand side within 1e-12 relative; discrete solutions of Jacobi-PCG to
1e-10 within 1e-10 relative L2 of the oracle's PCG solution.  Each test prints its case first.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from problems import hgdt_problem, matrix_error, oracle_problem

pytestmark = pytest.mark.gpu

MATRIX_TOL = 1e-12
SOLUTION_TOL = 1e-10


def fields(hgdt, kind, cells, seed):
    ne = int(np.prod(cells)) * (6 if kind == "p1_tets" else 1)
    if kind == "swipdg":
        return hgdt.fields.channel_field(seed, cells[0], cells[1])
    return hgdt.fields.lognormal(seed, ne)


CASES = [
    ("heat_q1", [16, 16], None),      # constant kappa
    ("heat_q1", [33, 20], 11),        # lognormal kappa, non-square
    ("heat_q1", [128, 128], None),    # the CPU-runnable reference configuration, full size
    ("p1_tets", [6, 6, 6], 3),
    ("p1_tets", [12, 10, 9], 7),
    ("swipdg", [12, 12], 5),          # high-contrast channels, sigma 16
    ("swipdg", [32, 24], 9),
]


def _build(hgdt, kind, cells, seed):
    kappa = 1.0 if seed is None else fields(hgdt, kind, cells, seed)
    m, s, oop = oracle_problem(kind, cells, kappa)
    oop.assemble(4)
    kap = kappa if np.isscalar(kappa) else torch.from_numpy(kappa).cuda()
    grid, space, op = hgdt_problem(hgdt, kind, cells, kap)
    op.assemble()
    torch.cuda.synchronize()
    return oop, op


@pytest.mark.parametrize("kind,cells,seed", CASES)
def test_assembly_parity(hgdt_mod, kind, cells, seed):
    print(f"case assembly {kind} cells={cells} seed={seed}", flush=True)
    oop, op = _build(hgdt_mod, kind, cells, seed)
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    vals = op.matrix.values.cpu().numpy()
    err = matrix_error(vals, oop.values(), orp)
    print(f"  matrix max rel err {err:.3e}")
    assert err <= MATRIX_TOL
    b, ob = op.rhs.cpu().numpy(), oop.rhs()
    berr = np.abs(b - ob).max() / np.abs(ob).max()
    print(f"  rhs max rel err {berr:.3e}")
    assert berr <= MATRIX_TOL
    if kind != "swipdg":
        assert np.array_equal(op.dirichlet_dofs(), oop.dirichlet_dofs())


@pytest.mark.parametrize("kind,cells,seed", CASES)
def test_solve_parity(hgdt_mod, kind, cells, seed):
    print(f"case solve {kind} cells={cells} seed={seed}", flush=True)
    oop, op = _build(hgdt_mod, kind, cells, seed)
    orp, ocols = oop.pattern()
    xo, ito, relo, rc = O.pcg_jacobi(orp, ocols, oop.values(), oop.rhs(), rtol=1e-10, nthreads=4)
    assert rc == 0
    for fmt in (hgdt_mod._lib.FORMAT_AUTO, hgdt_mod._lib.FORMAT_SELL):
        op.set_format(fmt)
        x = op.apply_inverse(op.rhs, opts={"precision": 1e-10})
        st = op.last_solve
        xg = x.cpu().numpy()
        rel = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
        info = op.format_info()
        print(f"  {info['format']}: iters {st['iterations']} (oracle {ito}) relres {st['relres']:.2e} "
              f"rel L2 diff {rel:.3e}")
        assert st["converged"] == 1 and st["relres"] <= 1e-10
        assert abs(st["iterations"] - ito) <= 2
        assert rel <= SOLUTION_TOL
        # true residual of the GPU solution with the oracle's matrix, and the one the solver reports at exit
        A = oop.csr()
        b = oop.rhs()
        true_host = np.linalg.norm(b - A @ xg) / np.linalg.norm(b)
        print(f"  true relres: device {st['true_relres']:.3e}, host {true_host:.3e}")
        assert true_host <= 2e-10
        assert abs(st["true_relres"] - true_host) <= 0.05 * true_host + 1e-14
        tr = np.empty(1)
        hgdt_mod._lib.call("hgdt_operator_instance_true_residuals", op._h, tr.ctypes.data_as(hgdt_mod._lib._pdbl))
        assert tr[0] == st["true_relres"]


@pytest.mark.parametrize("kind,cells,seed", [CASES[1], CASES[3], CASES[5]])
def test_apply_parity(hgdt_mod, kind, cells, seed):
    print(f"case apply {kind} cells={cells} seed={seed}", flush=True)
    oop, op = _build(hgdt_mod, kind, cells, seed)
    orp, ocols = oop.pattern()
    rng = np.random.default_rng(3)
    x = rng.standard_normal(op.shape[0])
    yo = O.spmv(orp, ocols, oop.values(), x)
    L = hgdt_mod._lib
    ys = {}
    fmts = [L.FORMAT_AUTO, L.FORMAT_DIA, L.FORMAT_DIA_NONSYMMETRIC, L.FORMAT_SELL]
    if kind == "swipdg":
        fmts.append(L.FORMAT_BLOCK)
    for fmt in fmts:
        op.set_format(fmt)
        y = op.apply(torch.from_numpy(x).cuda()).cpu().numpy()
        err = np.abs(y - yo).max() / np.abs(yo).max()
        name = op.format_info()["format"]
        ys[name] = y
        print(f"  {name}: max rel err {err:.3e}")
        assert err <= 1e-13
    # the matrices (continuous Lagrange, and symmetric interior penalty) are assembled exactly symmetric, so
    # the symmetric DIA (and, for the discontinuous space, element-block) paths are taken and, summing every
    # row in column order, agree bit for bit with full DIA and SELL
    assert "dia_symmetric" in ys and ("block4" in ys) == (kind == "swipdg")
    for name, y in ys.items():
        assert np.array_equal(y, ys["sell64"]), name


@pytest.mark.parametrize("kind,cells,seed", [CASES[1], CASES[4], CASES[6]])
def test_assembled_matrix_exactly_symmetric(hgdt_mod, kind, cells, seed):
    print(f"case symmetry {kind} cells={cells} seed={seed}", flush=True)
    oop, op = _build(hgdt_mod, kind, cells, seed)
    A = op.matrix.to_scipy()
    assert (A != A.T).nnz == 0


def test_nonsymmetric_interior_penalty(hgdt_mod):
    """NIPG (symmetry prefactor -1): blocks stored as computed, no symmetric storage, parity with the oracle"""
    print("case NIPG swipdg 10x8", flush=True)
    cells, seed = [10, 8], 4
    kappa = fields(hgdt_mod, "swipdg", cells, seed)
    m, s, oop = oracle_problem("swipdg", cells, kappa, symmetry=-1.0)
    oop.assemble(1)
    grid, space, op = hgdt_problem(hgdt_mod, "swipdg", cells, torch.from_numpy(kappa).cuda(), symmetry=-1.0)
    op.assemble()
    orp, ocols = oop.pattern()
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    print(f"  matrix max rel err {err:.3e}")
    assert err <= MATRIX_TOL
    A = op.matrix.to_scipy()
    assert (A != A.T).nnz > 0
    x = np.random.default_rng(5).standard_normal(op.shape[0])
    yo = O.spmv(orp, ocols, oop.values(), x)
    y = op.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    # element blocks read in full (the transposed blocks differ), the same bits as DIA
    assert op.format_info()["format"] == "block4"
    assert np.abs(y - yo).max() <= 1e-13 * np.abs(yo).max()
    op.set_format(hgdt_mod._lib.FORMAT_DIA)
    y2 = op.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    assert op.format_info()["format"] == "dia"
    assert np.array_equal(y, y2)


def test_bitwise_reproducible(hgdt_mod):
    """no floating-point atomics: two assemblies and two solves agree bit for bit"""
    print("case reproducibility p1_tets 10^3", flush=True)
    kappa = torch.from_numpy(hgdt_mod.fields.lognormal(2, 6 * 1000)).cuda()
    grid, space, op = hgdt_problem(hgdt_mod, "p1_tets", [10, 10, 10], kappa)
    op.assemble()
    v1, b1 = op.matrix.values.clone(), op.rhs.clone()
    x1 = op.apply_inverse(b1).clone()
    op.assemble()
    v2, b2 = op.matrix.values.clone(), op.rhs.clone()
    x2 = op.apply_inverse(b2)
    assert torch.equal(v1, v2) and torch.equal(b1, b2) and torch.equal(x1, x2)


def test_graph_and_eager_agree(hgdt_mod):
    print("case graph vs eager heat_q1 64^2", flush=True)
    kappa = torch.from_numpy(hgdt_mod.fields.lognormal(8, 64 * 64)).cuda()
    grid, space, op = hgdt_problem(hgdt_mod, "heat_q1", [64, 64], kappa)
    op.assemble()
    b = op.rhs
    xe = op.apply_inverse(b).clone()
    ite = op.last_solve["iterations"]
    for rep in range(3):  # graph replays checked against the eager result every time
        xg = op.apply_inverse(b, use_graph=True, check_every=8)
        assert op.last_solve["iterations"] == ite
        assert torch.equal(xe, xg)


def test_solver_failure_is_reported(hgdt_mod):
    print("case solver failure", flush=True)
    grid, space, op = hgdt_problem(hgdt_mod, "heat_q1", [16, 16], 1.0)
    op.assemble()
    with pytest.raises(hgdt_mod.linear_solver_failed):
        op.apply_inverse(op.rhs, opts={"max_iter": 3})
    assert op.last_solve["iterations"] == 3 and op.last_solve["converged"] == 0
    grid, space, op = hgdt_problem(hgdt_mod, "heat_q1", [8, 8], -1.0)
    op.assemble()
    with pytest.raises(hgdt_mod.linear_solver_failed):
        op.apply_inverse(op.rhs)


def test_launch_shapes_and_autotuner(hgdt_mod, tmp_path):
    """every launch shape gives the oracle's solution (1e-10 rel. L2); the tuner picks one of its candidates"""
    print("case launch shapes p1_tets 8^3", flush=True)
    from hgdt import autotune

    kind, cells, seed = "p1_tets", [8, 8, 8], 4
    oop, op = _build(hgdt_mod, kind, cells, seed)
    orp, ocols = oop.pattern()
    xo, ito, relo, rc = O.pcg_jacobi(orp, ocols, oop.values(), oop.rhs(), rtol=1e-10)
    L = hgdt_mod._lib
    yo = op.apply(torch.from_numpy(xo).cuda()).cpu().numpy()  # default launch shapes
    assert np.abs(yo - O.spmv(orp, ocols, oop.values(), xo)).max() <= 1e-13 * np.abs(yo).max()
    # (256, 8) and (128, 16): operator-application blocks that process their rows in chunks
    for a, v in [((64, 1), (256, 8)), ((512, 2), (1024, 4)), ((128, 1), (128, 1)), ((1024, 1), (64, 2)),
                 ((256, 8), (256, 4)), ((128, 16), (512, 1))]:
        op.set_launch(L.KERNEL_APPLY, *a)
        op.set_launch(L.KERNEL_VECTOR, *v)
        x = op.apply_inverse("rhs").cpu().numpy()
        rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        print(f"  apply {a} vector {v}: iters {op.last_solve['iterations']} rel {rel:.2e}")
        assert rel <= SOLUTION_TOL and abs(op.last_solve["iterations"] - ito) <= 2
        # the operator application is the same sum in every launch shape and format
        for fmt in (L.FORMAT_AUTO, L.FORMAT_SELL):
            op.set_format(fmt)
            y = op.apply(torch.from_numpy(xo).cuda()).cpu().numpy()
            assert np.array_equal(y, yo), f"apply differs for launch {a} format {fmt}"
        op.set_format(L.FORMAT_AUTO)
    choice = autotune.tune_solver(op, iterations=10, repeats=1)
    assert tuple(choice["apply"]) in autotune.APPLY_CANDIDATES
    assert tuple(choice["vector"]) in autotune.VECTOR_CANDIDATES
    # symmetric DIA: column staging is timed off and on where the staged kernel runs at the chosen shapes (one
    # row per lane), else the automatic setting (-1) is recorded
    staging = choice["staging"]
    op.set_column_staging(1)
    op.apply_inverse("rhs")
    applies = bool(op.last_solve["column_staged"])
    print(f"  tuned apply {choice['apply']} vector {choice['vector']} staging {staging} (staged kernel "
          f"{'applies' if applies else 'does not apply'})")
    assert applies == (choice["apply"][1] == 1)
    assert (staging in (0, 1)) if applies else (staging == -1)
    op.set_column_staging(staging)
    x = op.apply_inverse("rhs").cpu().numpy()
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= SOLUTION_TOL
    # a recorded choice is applied again from a launch table, staging included
    table = str(tmp_path / "table.json")
    autotune.save_table(op, choice, table)
    entry = autotune.apply_table(op, table)
    assert entry["apply"] == list(choice["apply"]) and entry["staging"] == choice["staging"]
    assert np.array_equal(op.apply_inverse("rhs").cpu().numpy(), x)


def test_element_blocks_on_a_union(hgdt_mod):
    """element-block storage over a ragged union of SWIPDG problems (elements numbered across the
    instances): y = A x bit for bit equal to SELL and symmetric DIA, and the solve within 1e-12 of the
    symmetric-DIA solve with the same iteration counts per instance"""
    H = hgdt_mod
    L = H._lib
    print("case element blocks union swipdg", flush=True)
    sizes = [5, 11, 3]
    parts = [H.Grid.structured("quad", [0, 0], [1, 1 + 0.3 * i], [n, n + 2]) for i, n in enumerate(sizes)]
    union = H.Grid.union(parts)
    kappa = torch.from_numpy(H.fields.lognormal(17, union.num_elements)).cuda()
    _, space, op = hgdt_problem(H, "swipdg", None, kappa, grid=union)
    op.assemble()
    x = torch.from_numpy(np.random.default_rng(8).standard_normal(space.num_DoFs)).cuda()
    ys, xs, its = {}, {}, {}
    for fmt in (L.FORMAT_BLOCK, L.FORMAT_DIA, L.FORMAT_SELL):
        op.set_format(fmt)
        ys[fmt] = op.apply(x).cpu().numpy()
        xs[fmt] = op.apply_inverse("rhs").cpu().numpy()
        its[fmt] = op.instance_stats()[0].copy()
        print(f"  {op.format_info()['format']}: iterations {list(its[fmt])}")
    assert np.array_equal(ys[L.FORMAT_BLOCK], ys[L.FORMAT_SELL])
    assert np.array_equal(ys[L.FORMAT_BLOCK], ys[L.FORMAT_DIA])
    # the element-block solve replayed as captured iteration chunks gives the eager bits, replay after replay
    op.set_format(L.FORMAT_BLOCK)
    for _ in range(3):
        xg = op.apply_inverse("rhs", use_graph=True, check_every=8).cpu().numpy()
        assert np.array_equal(xg, xs[L.FORMAT_BLOCK])
    assert np.array_equal(its[L.FORMAT_BLOCK], its[L.FORMAT_DIA])
    rel = np.linalg.norm(xs[L.FORMAT_BLOCK] - xs[L.FORMAT_DIA]) / np.linalg.norm(xs[L.FORMAT_DIA])
    print(f"  block vs symmetric DIA solution: rel L2 {rel:.2e}")
    assert rel <= 1e-12


def test_element_blocks_without_neighbours(hgdt_mod):
    """a discontinuous operator with element forms only (block diagonal, no neighbour slots): element-block
    storage gives SELL's bits and the oracle's solution"""
    H = hgdt_mod
    L = H._lib
    print("case element blocks, block-diagonal DG mass + Laplace", flush=True)
    cells = [9, 7]
    kappa = H.fields.lognormal(23, cells[0] * cells[1])
    grid = H.Grid.structured("quad", [0, 0], [1, 1], cells)
    space = H.DiscontinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(torch.from_numpy(kappa).cuda())))
    op.append(H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand(1.0)))
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.assemble()
    m = O.Mesh.structured(O.QUAD, cells, [0, 0], [1, 1])
    s = O.Space(m, False, 1)
    oop = O.Operator(s, O.STENCIL_ELEMENT)
    oop.append_laplace(kappa)
    oop.append_product(1.0)
    oop.append_source(1.0)
    oop.assemble(1)
    x = torch.from_numpy(np.random.default_rng(2).standard_normal(space.num_DoFs)).cuda()
    y = op.apply(x).cpu().numpy()
    assert op.format_info()["format"] == "block4" and op.format_info()["num_diagonals"] == 0
    op.set_format(L.FORMAT_SELL)
    assert np.array_equal(y, op.apply(x).cpu().numpy())
    op.set_format(L.FORMAT_AUTO)
    rp, cols = oop.pattern()
    xo, ito, relo, rc = O.pcg_jacobi(rp, cols, oop.values(), oop.rhs(), rtol=1e-10)
    xg = op.apply_inverse("rhs").cpu().numpy()
    rel = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    print(f"  iterations {op.last_solve['iterations']} (oracle {ito}), rel L2 {rel:.2e}")
    assert rel <= 1e-10 and abs(op.last_solve["iterations"] - ito) <= 2


@pytest.mark.parametrize("case", ["p1_tets_three_windows", "heat_q1_union"])
def test_symmetric_dia_solve_bits(hgdt_mod, case):
    """the symmetric DIA solve with its application's p entries staged in LDS (forced on: these problems are
    below the automatic threshold) and without gives the bits and iteration counts of the full DIA solve
    (every diagonal read, the same sums in the same order) in launch shapes of one row per lane, eagerly
    and replayed as captured chunks"""
    H = hgdt_mod
    L = H._lib
    print(f"case symmetric DIA solve bits {case}", flush=True)
    if case == "p1_tets_three_windows":
        # 25 x 25 vertices per plane: the column offsets +-1, +-25, +-26 and +-(625 .. 651) form three windows
        cells = [24, 24, 6]
        kappa = torch.from_numpy(fields(H, "p1_tets", cells, 3)).cuda()
        _, space, op = hgdt_problem(H, "p1_tets", cells, kappa)
    else:
        parts = [H.Grid.structured("quad", [0, 0], [1, 1 + 0.2 * i], [n, n + 2]) for i, n in enumerate([7, 30, 4])]
        union = H.Grid.union(parts)
        kappa = torch.from_numpy(H.fields.lognormal(29, union.num_elements)).cuda()
        _, space, op = hgdt_problem(H, "heat_q1", None, kappa, grid=union)
    op.assemble()
    with pytest.raises(H._lib.HgdtError):
        op.set_column_staging(2)
    for block in (64, 256, 1024):
        op.set_launch(L.KERNEL_APPLY, block, 1)
        op.set_format(L.FORMAT_DIA_NONSYMMETRIC)
        xd = op.apply_inverse("rhs").cpu().numpy()
        itd = op.instance_stats()[0].copy()
        assert op.format_info()["format"] == "dia"
        op.set_format(L.FORMAT_AUTO)
        for staging in (1, 0):
            op.set_column_staging(staging)
            x = op.apply_inverse("rhs").cpu().numpy()
            its = op.instance_stats()[0].copy()
            assert op.format_info()["format"] == "dia_symmetric"
            print(f"  apply block {block}, staging {staging}: iterations {list(its)}, "
                  f"staged {op.last_solve['column_staged']}")
            # the staged kernel really ran (staging 1) or did not (staging 0)
            assert op.last_solve["column_staged"] == staging
            assert np.array_equal(its, itd)
            assert np.array_equal(x, xd)
            for _ in range(2):
                xg = op.apply_inverse("rhs", use_graph=True, check_every=8).cpu().numpy()
                assert np.array_equal(xg, xd)
        op.set_column_staging(-1)
