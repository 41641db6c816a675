"""Order-2 Lagrange spaces (P2 triangles and tetrahedra, Q2 quadrilaterals and hexahedra) on the HIP path.

The reference's own known-answer tests run on the device: P2 continuous Lagrange Laplace and mass
matrices on the grid of dune/gdt/test/integrands/integrands.hh:64-70 (9 x 2 cells on [0, 3] x [0, 1]),
statistics min / max / sum of squares against integrands_laplace.cc:115-135 and
integrands_product.cc:107-126 with the reference's tolerances (and against the exact rational values of
tests/golden/make_golden.py).  Entry-wise parity with the oracle (same DoF numbering) for assembly,
Dirichlet DoFs, solve and error norms; Q2 hexahedra (27 local DoFs, assembled by one lane per element row)
also against the Kronecker closed form of a single element.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest
import torch

import oracle_lib as O
from problems import matrix_error

pytestmark = pytest.mark.gpu

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


@pytest.mark.parametrize("geometry,kind", [("quad", "laplace"), ("tri", "laplace"), ("quad", "mass"), ("tri", "mass")])
def test_p2_known_answers_on_device(hgdt_mod, geometry, kind):
    H = hgdt_mod
    print(f"case P2 known answers {geometry} {kind}", flush=True)
    grid = H.Grid.structured(geometry, [0, 0], [3, 1], [9, 2])
    space = H.ContinuousLagrangeSpace(grid, 2)
    op = H.make_matrix_operator(space, H.Stencil.element)
    integrand = H.LocalLaplaceIntegrand(1.0) if kind == "laplace" else H.LocalElementProductIntegrand(1.0)
    op.append(H.LocalElementIntegralBilinearForm(integrand))
    op.assemble()
    A = op.matrix.to_scipy().toarray()
    mn, mx, ss = A.min(), A.max(), float((A * A).sum())
    name = f"p2_{kind}_{'cube' if geometry == 'quad' else 'simplex'}"
    ref = GOLDEN["reference"][name]
    exact = {k: float(Fraction(*v)) for k, v in GOLDEN["exact"][name].items()}
    tmin, tmax, tss = ref["tol"]
    print(f"  min {mn:.15g} max {mx:.15g} sumsq {ss:.15g} (reference {ref['min']}, {ref['max']}, {ref['sumsq']})")
    assert abs(mn - ref["min"]) <= tmin and abs(mx - ref["max"]) <= tmax
    assert abs(mn - exact["min"]) <= tmin and abs(mx - exact["max"]) <= tmax
    assert abs(ss - exact["sumsq"]) <= tss
    # the reference's printed simplex Laplace sum is itself 4.5e-12 above the exact 402007/162
    assert abs(ss - ref["sumsq"]) <= (1e-11 if name == "p2_laplace_simplex" else tss)


CASES = [("tri", O.TRI, [7, 5]), ("quad", O.QUAD, [6, 9]), ("tet", O.TET, [3, 4, 3]), ("hex", O.HEX, [3, 4, 3])]


def _pair(H, geometry, geom_o, cells, seed):
    dim = len(cells)
    lo, hi = [0.0] * dim, [1.0, 2.0, 1.5][:dim]
    kappa = H.fields.lognormal(seed, int(np.prod(cells)) * (1 if geometry in ("quad", "hex") else (2 if dim == 2 else 6)))
    m = O.Mesh.structured(geom_o, cells, lo, hi)
    s = O.Space(m, True, 2)
    oop = O.Operator(s, O.STENCIL_ELEMENT)
    oop.append_laplace(kappa)
    oop.append_source(1.0)
    oop.set_dirichlet(True)
    oop.assemble(1)
    grid = H.Grid.structured(geometry, lo, hi, cells)
    space = H.ContinuousLagrangeSpace(grid, 2)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(torch.from_numpy(kappa).cuda())))
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    op.assemble()
    return m, s, oop, grid, space, op


@pytest.mark.parametrize("geometry,geom_o,cells", CASES)
def test_order2_parity(hgdt_mod, geometry, geom_o, cells):
    H = hgdt_mod
    print(f"case order 2 parity {geometry} {cells}", flush=True)
    m, s, oop, grid, space, op = _pair(H, geometry, geom_o, cells, 13)
    assert space.num_DoFs == s.size
    assert np.array_equal(space.global_indices(), s.global_indices())
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    berr = np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() / np.abs(oop.rhs()).max()
    print(f"  matrix max rel err {err:.3e}, rhs {berr:.3e}")
    assert err <= 1e-12 and berr <= 1e-12
    assert np.array_equal(op.dirichlet_dofs(), oop.dirichlet_dofs())
    xo, ito, relo, rc = O.pcg_jacobi(orp, ocols, oop.values(), oop.rhs(), rtol=1e-10)
    x = op.apply_inverse("rhs", opts={"precision": 1e-10}).cpu().numpy()
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"  iters {op.last_solve['iterations']} (oracle {ito}) rel L2 {rel:.2e}")
    assert rel <= 1e-10 and abs(op.last_solve["iterations"] - ito) <= 2
    # error norms of the P2 solution against an analytic function, device vs oracle at the same points
    g = H.GenericFunction(3, lambda X: np.sin(X[0]) * np.cos(X[1]),
                          jacobian=lambda X: np.stack([np.cos(X[0]) * np.cos(X[1]), -np.sin(X[0]) * np.sin(X[1])]
                                                      + ([0 * X[0]] if len(cells) == 3 else [])))
    uh = H.make_discrete_function(space, torch.from_numpy(x))
    f1 = H.make_bilinear_form(grid, uh - g)
    f1 += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0))
    f2 = H.make_bilinear_form(grid, uh - g)
    f2 += H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand())
    h1, l2 = f1.apply2(), f2.apply2()
    P = m.quadrature_points(6)
    Xp = P.reshape(-1, len(cells)).T
    ol2, oh1 = O.error_norms(s, x, 6, g.evaluate(Xp), np.ascontiguousarray(g.jacobian(Xp).T))
    assert abs(h1 - oh1) <= 1e-12 * oh1 and abs(l2 - ol2) <= 1e-12 * ol2


def test_order2_interpolates_quadratics_exactly(hgdt_mod):
    """the P2 space contains the quadratics: the nodal interpolant of q(x, y) = x^2 - 3xy + 2y has zero
    error in L2 and H1 (to round-off) on the device"""
    H = hgdt_mod
    print("case P2 quadratic interpolation", flush=True)
    cells = [5, 4]
    grid = H.Grid.structured("tri", [0, 0], [1, 1], cells)
    space = H.ContinuousLagrangeSpace(grid, 2)
    m = O.Mesh.structured(O.TRI, cells, [0, 0], [1, 1])
    s = O.Space(m, True, 2)
    # node coordinates from the reference node positions (vertices, then edge midpoints)
    X = m.coords()
    conn = m.conn()
    gi = s.global_indices()
    u = np.zeros(s.size)
    for e in range(m.ne):
        v = X[conn[e]]
        nodes = [v[0], v[1], v[2], (v[0] + v[1]) / 2, (v[0] + v[2]) / 2, (v[1] + v[2]) / 2]
        for i, p in enumerate(nodes):
            u[gi[e, i]] = p[0] ** 2 - 3 * p[0] * p[1] + 2 * p[1]
    q = H.GenericFunction(2, lambda Z: Z[0] ** 2 - 3 * Z[0] * Z[1] + 2 * Z[1],
                          jacobian=lambda Z: np.stack([2 * Z[0] - 3 * Z[1], -3 * Z[0] + 2]))
    uh = H.make_discrete_function(space, torch.from_numpy(u))
    f = H.make_bilinear_form(grid, uh - q)
    f += H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(1.0))
    f += H.LocalElementIntegralBilinearForm(H.LocalElementProductIntegrand())
    assert f.apply2() < 1e-26


def test_q2_hexahedron_element_matrices(hgdt_mod):
    """one Q2 hexahedron [0, a] x [0, b] x [0, c] in a discontinuous space (global = local numbering): the
    device's stiffness and mass matrices equal the Kronecker closed forms of the 1-d quadratic matrices"""
    H = hgdt_mod
    print("case Q2 hexahedron element matrices against the closed form", flush=True)
    a, b, c = 0.7, 1.1, 0.45

    def K1(h):
        return np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / (3 * h)

    def M1(h):
        return np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) * h / 30

    K = (np.kron(M1(c), np.kron(M1(b), K1(a))) + np.kron(M1(c), np.kron(K1(b), M1(a)))
         + np.kron(K1(c), np.kron(M1(b), M1(a))))
    M = np.kron(M1(c), np.kron(M1(b), M1(a)))
    grid = H.Grid.structured("hex", [0, 0, 0], [a, b, c], [1, 1, 1])
    space = H.DiscontinuousLagrangeSpace(grid, 2)
    for integrand, E in ((H.LocalLaplaceIntegrand(1.0), K), (H.LocalElementProductIntegrand(1.0), M)):
        op = H.make_matrix_operator(space, H.Stencil.element)
        op.append(H.LocalElementIntegralBilinearForm(integrand))
        op.assemble()
        A = op.matrix.to_scipy().toarray()
        err = np.abs(A - E).max() / np.abs(E).max()
        print(f"  max rel err {err:.2e}")
        assert A.shape == (27, 27) and err <= 1e-14


@pytest.mark.parametrize("geometry,geom,cells", [("quad", O.QUAD, [5, 4]), ("tri", O.TRI, [4, 4]),
                                                 ("hex", O.HEX, [2, 2, 2]), ("tet", O.TET, [2, 2, 2])])
def test_adjoint_gradient_order2_matches_oracle(hgdt_mod, geometry, geom, cells):
    """the adjoint element kernel on order-2 spaces (P2, Q2; Q2 hexahedra with 27 local DoFs): for the solve u
    and the adjoint lambda (A lambda = u) of a lognormal diffusion problem, -kappa_e lambda_e^T Khat_e u_e
    within 1e-12 (max-relative) of the oracle's, whose gradient is pinned by central differences"""
    H = hgdt_mod
    print(f"case adjoint gradient order 2 on {geometry} cells={cells}", flush=True)
    d = len(cells)
    grid = H.Grid.structured(geometry, [0] * d, [1] * d, cells)
    space = H.ContinuousLagrangeSpace(grid, 2)
    kap = np.exp(H.fields.normal(41, grid.num_elements))
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
    m = O.Mesh.structured(geom, cells, [0] * d, [1] * d)
    s = O.Space(m, True, 2)
    go = O.adjoint_log_kappa_gradient(s, kap, u.cpu().numpy(), lam.cpu().numpy())
    rel = np.abs(g - go).max() / np.abs(go).max()
    print(f"  {grid.num_elements} elements: gradient vs oracle {rel:.2e}")
    assert rel <= 1e-12
