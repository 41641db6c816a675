"""Edge cases of the HIP path: degenerate sizes, general (non axis-aligned, unstructured) meshes,
ragged batches and invalid input.

Bar as in test_gpu_parity: pattern bit-exact, matrix within 1e-12 (row-relative), solution within
1e-10 relative L2 of the oracle's Jacobi-PCG; invalid input raises the exception family the reference
raises, never a crash.
"""
import numpy as np
import pytest
import torch

import oracle_lib as O
from problems import matrix_error

pytestmark = pytest.mark.gpu

GEO = {"quad": O.QUAD, "tri": O.TRI, "hex": O.HEX, "tet": O.TET}


def _heat(H, grid, kappa=1.0, f=1.0, order=1, dirichlet=True):
    space = H.ContinuousLagrangeSpace(grid, order)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(kappa)))
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(f)))
    op.append(rhs)
    if dirichlet:
        op.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    op.assemble()
    return space, op


def _oracle_heat(mesh, kappa=1.0, f=1.0, order=1, dirichlet=True):
    s = O.Space(mesh, True, order)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    op.append_laplace(kappa)
    op.append_source(f)
    op.set_dirichlet(dirichlet)
    op.assemble(1)
    return s, op


def _compare(op, oop, label):
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols), label
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    assert err <= 1e-12, (label, err)
    ob = oop.rhs()
    if np.abs(ob).max() > 0:
        assert np.abs(op.rhs.cpu().numpy() - ob).max() <= 1e-12 * np.abs(ob).max(), label
    xo, ito, relo, rc = O.pcg_jacobi(orp, ocols, oop.values(), ob, rtol=1e-10)
    x = op.apply_inverse("rhs").cpu().numpy()
    nrm = np.linalg.norm(xo)
    rel = np.linalg.norm(x - xo) / nrm if nrm > 0 else np.linalg.norm(x)
    print(f"  {label}: iters {op.last_solve['iterations']} (oracle {ito}) rel {rel:.2e}")
    assert rel <= 1e-10 and abs(op.last_solve["iterations"] - ito) <= 2, label


@pytest.mark.parametrize("geometry", ["quad", "tri", "hex", "tet"])
def test_single_cell_all_dirichlet(hgdt_mod, geometry):
    """one cube (split into simplices): every DoF lies on the boundary, the matrix is the identity, x = 0"""
    H = hgdt_mod
    print(f"case single cell {geometry}", flush=True)
    dim = 2 if geometry in ("quad", "tri") else 3
    grid = H.Grid.structured(geometry, [0] * dim, [1] * dim, [1] * dim)
    space, op = _heat(H, grid)
    A = op.matrix.to_scipy().toarray()
    assert np.array_equal(A, np.eye(space.num_DoFs))
    assert np.array_equal(op.dirichlet_dofs(), np.arange(space.num_DoFs))
    x = op.apply_inverse("rhs")
    assert op.last_solve["iterations"] == 0 and torch.count_nonzero(x) == 0


def test_zero_right_hand_side(hgdt_mod):
    """b = 0: converged before the first iteration, x = 0 (relative residual reported 0)"""
    H = hgdt_mod
    print("case zero rhs", flush=True)
    grid = H.Grid.structured("quad", [0, 0], [1, 1], [8, 8])
    space, op = _heat(H, grid, f=0.0)
    x = op.apply_inverse("rhs")
    assert op.last_solve["iterations"] == 0 and op.last_solve["converged"] == 1
    assert torch.count_nonzero(x) == 0


def _affine_mesh(geometry, cells, A, shift):
    """structured mesh mapped by x -> A x + shift (sheared, rotated: no axis-aligned cells)"""
    m = O.Mesh.structured(GEO[geometry], cells, [0] * len(cells), [1] * len(cells))
    X = m.coords() @ np.asarray(A).T + np.asarray(shift)
    return X, m.conn()


@pytest.mark.parametrize("geometry,cells", [("quad", [7, 5]), ("tri", [6, 6]), ("hex", [4, 3, 5]), ("tet", [4, 4, 3])])
def test_sheared_explicit_mesh(hgdt_mod, geometry, cells):
    """explicit grids from coordinates and connectivity, with parallelogram / parallelepiped cells"""
    H = hgdt_mod
    print(f"case sheared {geometry} {cells}", flush=True)
    A = [[1.0, 0.35], [-0.2, 0.9]] if len(cells) == 2 else [[1.0, 0.3, -0.1], [0.1, 0.8, 0.25], [-0.2, 0.15, 1.1]]
    X, conn = _affine_mesh(geometry, cells, A, [0.5] * len(cells))
    grid = H.Grid.from_arrays(geometry, X, conn)
    kappa = H.fields.lognormal(17, grid.num_elements)
    space, op = _heat(H, grid, kappa=torch.from_numpy(kappa).cuda())
    mesh = O.Mesh.explicit(GEO[geometry], X, conn)
    s, oop = _oracle_heat(mesh, kappa=kappa)
    _compare(op, oop, f"sheared {geometry}")


@pytest.mark.parametrize("geometry,cells,order", [("tri", [9, 7], 1), ("tet", [5, 4, 4], 1), ("tri", [6, 5], 2)])
def test_unstructured_simplices(hgdt_mod, geometry, cells, order):
    """interior vertices moved at random (a general simplex mesh, no constant stencil offsets): the
    operator falls back to SELL storage"""
    H = hgdt_mod
    print(f"case perturbed {geometry} {cells} order {order}", flush=True)
    m = O.Mesh.structured(GEO[geometry], cells, [0] * len(cells), [1] * len(cells))
    X = m.coords().copy()
    h = 1.0 / max(cells)
    interior = np.all((X > 1e-12) & (X < 1 - 1e-12), axis=1)
    rng = np.random.default_rng(23)
    X[interior] += rng.uniform(-0.15 * h, 0.15 * h, size=(interior.sum(), X.shape[1]))
    # randomise the vertex numbering too: the pattern is no longer banded
    perm = rng.permutation(X.shape[0])
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    Xp, conn = X[perm], inv[m.conn()]
    grid = H.Grid.from_arrays(geometry, Xp, conn)
    kappa = H.fields.lognormal(29, grid.num_elements)
    space, op = _heat(H, grid, kappa=torch.from_numpy(kappa).cuda(), order=order)
    mesh = O.Mesh.explicit(GEO[geometry], Xp, conn)
    s, oop = _oracle_heat(mesh, kappa=kappa, order=order)
    _compare(op, oop, f"perturbed {geometry} order {order}")
    assert op.format_info()["format"] == "sell64"


def test_ragged_batch_with_tiny_instances(hgdt_mod):
    """instances of 1, 2, 3 and 40 cells per direction in one union: each solved to its own tolerance,
    equal bit for bit to its plain solve"""
    H = hgdt_mod
    print("case ragged batch", flush=True)
    sizes = [1, 40, 2, 3]
    parts = [H.Grid.structured("quad", [0, 0], [1, 1], [n, n]) for n in sizes]
    union = H.Grid.union(parts)
    kappa = H.fields.lognormal(31, union.num_elements)
    space, op = _heat(H, union, kappa=torch.from_numpy(kappa).cuda())
    xb = op.apply_inverse("rhs").cpu().numpy()
    its, rels = op.instance_stats()
    eoff, doff = union.instance_elements(), space.instance_dofs()
    assert its[0] == 0  # the single cell is all Dirichlet
    assert op.last_solve["compactions"] >= 1
    for i, n in enumerate(sizes):
        s1, op1 = _heat(H, parts[i], kappa=torch.from_numpy(kappa[eoff[i]:eoff[i + 1]]).cuda())
        for k in (3, 4):
            op1.set_launch(k, *op.get_launch(k))
        x1 = op1.apply_inverse("rhs").cpu().numpy()
        assert op1.last_solve["iterations"] == its[i]
        assert np.array_equal(x1, xb[doff[i]:doff[i + 1]])


def test_invalid_input_raises(hgdt_mod):
    H = hgdt_mod
    print("case invalid input", flush=True)
    grid = H.Grid.structured("quad", [0, 0], [1, 1], [4, 4])
    space = H.ContinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    # a coefficient field of the wrong length
    with pytest.raises(H.shapes_do_not_match):
        op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(torch.ones(7, dtype=torch.float64))))
    # non-affine quadrilateral
    X = np.array([[0, 0], [1, 0], [0, 1], [1.3, 1.2]], dtype=np.float64)
    with pytest.raises(H.not_supported):
        H.Grid.from_arrays("quad", X, np.array([[0, 1, 2, 3]], dtype=np.int32))
    # connectivity out of range
    with pytest.raises(H.wrong_input_given):
        H.Grid.from_arrays("tri", X[:3], np.array([[0, 1, 5]], dtype=np.int32))
    # order 3 spaces are not provided
    with pytest.raises(H.not_supported):
        H.ContinuousLagrangeSpace(grid, 3)
    # the adjoint kernel for a form that is not a Laplace form
    dg2 = H.ContinuousLagrangeSpace(grid, 2)
    op2 = H.make_matrix_operator(dg2, H.Stencil.element)
    op2.append(H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand(torch.ones(16, dtype=torch.float64,
                                                                                             device="cuda"))))
    op2.assemble()
    u2 = torch.zeros(dg2.num_DoFs, dtype=torch.float64, device="cuda")
    with pytest.raises(H.wrong_input_given):
        op2.adjoint_log_diffusion_gradient(u2, u2)
    # an operator between spaces of different orders
    with pytest.raises(H.not_supported):
        H.make_matrix_operator(space, H.Stencil.element, range_space=H.ContinuousLagrangeSpace(grid, 2))
    # launch shapes outside the kernels' bounds
    with pytest.raises(H.wrong_input_given):
        op.set_launch(H._lib.KERNEL_APPLY, 96, 1)
    with pytest.raises(H.wrong_input_given):
        op.set_launch(H._lib.KERNEL_ELEMENT, 512, 1)


def _first_nonpositive_curvature(A, b, rtol=1e-10, max_iter=2000):
    """step j of textbook Jacobi-PCG (x0 = 0) at which p_j . A p_j <= 0 first occurs (None if it converges
    first), found on the host to choose the oracle's comparison step; the margin is |cos(p, Ap)|"""
    import scipy.sparse  # noqa: F401  (A is a scipy matrix)
    d = A.diagonal()
    x, r = np.zeros_like(b), b.copy()
    z = r / d
    p, rz, tol2 = z.copy(), r @ z, rtol * rtol * (b @ b)
    for k in range(max_iter):
        q = A @ p
        pq = p @ q
        if not pq > 0:
            return k, abs(pq) / (np.linalg.norm(p) * np.linalg.norm(q))
        alpha = rz / pq
        x += alpha * p
        r -= alpha * q
        if r @ r <= tol2:
            return None, 1.
        z = r / d
        rzn = r @ z
        p = z + (rzn / rz) * p
        rz = rzn
    return None, 1.


def test_breakdown_leaves_the_iterate_of_the_last_full_step(hgdt_mod):
    """indefinite operators with a positive diagonal (Laplace minus c times the mass matrix): the solve
    stops each instance at its first non-positive curvature p_j . A p_j <= 0 (reported as breakdown) and
    leaves the iterate after j updates, x_j, equal to the oracle's Jacobi-PCG run for j iterations.  The
    breakdown steps (3, 2, 1, 0) cover both parities of the iterate's every-second-step update; the
    instance with c = 0 converges and equals its plain solve"""
    import scipy.sparse as sp
    H = hgdt_mod
    print("case breakdown iterates", flush=True)
    n, cs = 16, [0., 25., 30., 40., 80.]
    parts = [H.Grid.structured("quad", [0, 0], [1, 1], [n, n]) for _ in cs]
    union = H.Grid.union(parts)
    w = np.repeat(np.array(cs), n * n)
    space = H.ContinuousLagrangeSpace(union, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0)))
    op.append(H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand(torch.from_numpy(w).cuda())),
              {"matrixoperator.scaling": -1.0})
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    op.assemble()
    x = torch.zeros(space.num_DoFs, dtype=torch.float64, device="cuda")
    with pytest.raises(H.linear_solver_failed):
        op.apply_inverse("rhs", x, check_every=2)
    assert op.last_solve["breakdown"] == 1
    x = x.cpu().numpy()
    its, _ = op.instance_stats()
    doff = space.instance_dofs()
    seen = set()
    for i, c in enumerate(cs):
        m = O.Mesh.structured(O.QUAD, [n, n], [0, 0], [1, 1])
        s = O.Space(m, True, 1)
        oop = O.Operator(s, O.STENCIL_ELEMENT)
        oop.append_laplace(1.0)
        oop.append_product(c, scaling=-1.0)
        oop.append_source(1.0)
        oop.set_dirichlet(True)
        oop.assemble(1)
        rp, cols = oop.pattern()
        A = sp.csr_matrix((oop.values(), cols, rp), shape=(len(rp) - 1, len(rp) - 1))
        j, margin = _first_nonpositive_curvature(A, oop.rhs())
        xi = x[doff[i]:doff[i + 1]]
        if j is None:
            xo, ito, relo, rc = O.pcg_jacobi(rp, cols, oop.values(), oop.rhs(), rtol=1e-10)
            print(f"  c={c}: converged, iters {its[i]} (oracle {ito})")
            assert rc == 0 and its[i] == ito
        else:
            assert margin > 1e-3  # the sign of p.Ap is not a rounding matter
            seen.add(j)
            xo = np.zeros_like(xi) if j == 0 else O.pcg_jacobi(rp, cols, oop.values(), oop.rhs(), rtol=1e-10,
                                                                 max_iters_run=j)[0]
            print(f"  c={c}: breakdown at step {j} (steps begun {its[i]})")
            assert its[i] == j + 1
        rel = np.linalg.norm(xi - xo) / max(np.linalg.norm(xo), 1e-300)
        print(f"    iterate vs oracle: rel L2 {rel:.2e}")
        assert (rel <= 1e-10) if np.linalg.norm(xo) > 0 else np.all(xi == 0.)
    assert {0, 1, 2, 3} <= seen


def test_breakdown_with_more_blocks_than_resident(hgdt_mod):
    """the breakdown flag of the update kernel at full size: four indefinite instances (Laplace minus c
    times the mass matrix on 600 x 600 cells, 1.44 M DoFs) with vector blocks of 64 rows, about 22 600
    blocks, far more than the GPU holds at once, so most blocks of an instance start after the block that
    flags the breakdown has finished.  Every instance must stop at its first non-positive curvature with the
    oracle's iterate after j steps, including the pending half update at odd j"""
    import scipy.sparse as sp
    H = hgdt_mod
    n, cs = 600, [2910.4, 1862.6, 953.7, 762.9]  # breakdown at steps 0, 1, 2, 3 (found with the oracle below)
    print(f"case breakdown at full size: 4 x Q1 {n}^2, vector blocks 64 x 1", flush=True)
    parts = [H.Grid.structured("quad", [0, 0], [1, 1], [n, n]) for _ in cs]
    union = H.Grid.union(parts)
    w = np.repeat(np.array(cs), n * n)
    space = H.ContinuousLagrangeSpace(union, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0)))
    op.append(H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand(torch.from_numpy(w).cuda())),
              {"matrixoperator.scaling": -1.0})
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    op.set_launch(H._lib.KERNEL_VECTOR, 64, 1)
    op.assemble()
    x = torch.zeros(space.num_DoFs, dtype=torch.float64, device="cuda")
    with pytest.raises(H.linear_solver_failed):
        op.apply_inverse("rhs", x, check_every=2)
    assert op.last_solve["breakdown"] == 1 and op.last_solve["converged"] == 0
    x = x.cpu().numpy()
    its, _ = op.instance_stats()
    doff = space.instance_dofs()
    m = O.Mesh.structured(O.QUAD, [n, n], [0, 0], [1, 1])
    s = O.Space(m, True, 1)
    seen = set()
    for i, c in enumerate(cs):
        oop = O.Operator(s, O.STENCIL_ELEMENT)
        oop.append_laplace(1.0)
        oop.append_product(c, scaling=-1.0)
        oop.append_source(1.0)
        oop.set_dirichlet(True)
        oop.assemble(4)
        rp, cols = oop.pattern()
        A = sp.csr_matrix((oop.values(), cols, rp), shape=(len(rp) - 1, len(rp) - 1))
        j, margin = _first_nonpositive_curvature(A, oop.rhs(), max_iter=16)
        assert j is not None and margin > 1e-3, (c, j, margin)
        seen.add(j)
        xi = x[doff[i]:doff[i + 1]]
        xo = np.zeros_like(xi) if j == 0 else O.pcg_jacobi(rp, cols, oop.values(), oop.rhs(), rtol=1e-10,
                                                             max_iters_run=j)[0]
        rel = np.linalg.norm(xi - xo) / max(np.linalg.norm(xo), 1e-300)
        print(f"  c={c}: breakdown at step {j} (steps begun {its[i]}), iterate vs oracle rel L2 {rel:.2e}")
        assert its[i] == j + 1
        assert (rel <= 1e-10) if j > 0 else np.all(xi == 0.)
    assert {1, 3} <= seen


def test_coefficient_deleted_before_assemble(hgdt_mod):
    """python/test/segfault_laplace_integrand.py:1-16: the coefficient function (here a device field), its
    integrand and its form are dropped before assemble(); the operator keeps what the appended forms reference,
    so the assembled matrix is the oracle's even after the freed memory of the caller's objects was reused"""
    import gc
    H = hgdt_mod
    print("case coefficient deleted before assemble", flush=True)
    grid = H.Grid.structured("quad", [0, 0], [1, 1], [12, 10])
    space = H.ContinuousLagrangeSpace(grid, 1)
    op = H.MatrixOperator(grid, space, space)
    kappa = H.fields.lognormal(81, grid.num_elements)
    func = torch.from_numpy(kappa).cuda()
    form = H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(H.GridFunction(func)))
    op += form
    del func, form
    gc.collect()
    # reuse what the caching allocator freed: same-size blocks filled with garbage
    junk = [torch.full((grid.num_elements,), float("nan"), dtype=torch.float64, device="cuda") for _ in range(8)]
    torch.cuda.synchronize()
    op.assemble()
    m = O.Mesh.structured(O.QUAD, [12, 10], [0, 0], [1, 1])
    s = O.Space(m, True, 1)
    oop = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
    oop.append_laplace(kappa)
    oop.assemble(1)
    rp, _ = oop.pattern()
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), rp)
    print(f"  matrix max rel err {err:.2e}")
    assert err <= 1e-12
    del junk
