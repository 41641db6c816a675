"""The Laplace form with a matrix-valued diffusion on the HIP path (LocalLaplaceIntegrand(GridFunction<E, d, d>),
laplace.hh:40-48,94-101): per-element symmetric positive definite tensors and a constant anisotropic tensor on
quadrilaterals, triangles, hexahedra (Q2: row lanes) and tetrahedra (P1: the one-point path), orders 1 and 2,
and on a mixed grid.  Bar: pattern bit-exact, matrix and rhs within 1e-12 of the oracle's, Jacobi-PCG solution
within 1e-10 with the oracle's iteration count (+-2).  Refused: non-symmetric tensors, the adjoint gradient,
the matrix-free format.
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


def _hgdt(H, grid, k, K):
    space = H.ContinuousLagrangeSpace(grid, k)
    op = H.make_matrix_operator(space, H.Stencil.element)
    op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(K)))
    rhs = H.make_vector_functional(space)
    rhs.append(H.LocalElementIntegralFunctional(H.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    op.append(H.make_dirichlet_constraints(space, H.AllDirichletBoundaryInfo()))
    return space, op


def _check(op, m, k, K, label):
    s = O.Space(m, True, k)
    oop = O.Operator(s, O.STENCIL_ELEMENT)
    oop.append_laplace_tensor(K)
    oop.append_source(1.0)
    oop.set_dirichlet(True)
    oop.assemble(1)
    rp, cols = op.pattern()
    orp, ocols = oop.pattern()
    assert np.array_equal(rp, orp) and np.array_equal(cols, ocols)
    err = matrix_error(op.matrix.values.cpu().numpy(), oop.values(), orp)
    berr = np.abs(op.rhs.cpu().numpy() - oop.rhs()).max() / np.abs(oop.rhs()).max()
    x = op.apply_inverse("rhs").cpu().numpy()
    xo, ito, relo, rc = O.pcg_jacobi(orp, ocols, oop.values(), oop.rhs(), rtol=1e-10)
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    its = op.last_solve["iterations"]
    A = op.matrix.to_scipy()
    print(f"  {label}: matrix {err:.2e}, rhs {berr:.2e}, solution {rel:.2e}, iterations {its} (oracle {ito})")
    assert err <= 1e-12 and berr <= 1e-12 and rel <= 1e-10 and abs(its - ito) <= 2
    assert abs(A - A.T).max() == 0.0


@pytest.mark.parametrize("geometry,geom,cells,k", [("quad", O.QUAD, [6, 5], 1), ("tri", O.TRI, [5, 4], 2),
                                                   ("hex", O.HEX, [3, 2, 2], 1), ("hex", O.HEX, [2, 2, 2], 2),
                                                   ("tet", O.TET, [3, 2, 2], 1), ("tet", O.TET, [2, 2, 2], 2)])
def test_tensor_diffusion_matches_oracle(hgdt_mod, geometry, geom, cells, k):
    H = hgdt_mod
    print(f"case tensor diffusion {geometry} cells={cells} order {k}", flush=True)
    d = len(cells)
    lo, hi = [0.0] * d, [1.0, 0.8, 1.2][:d]
    grid = H.Grid.structured(geometry, lo, hi, cells)
    K = _spd(7, grid.num_elements, d)
    space, op = _hgdt(H, grid, k, torch.from_numpy(K).cuda())
    op.assemble()
    assert not op.assembly_info()["structured"]
    _check(op, O.Mesh.structured(geom, cells, lo, hi), k, K, f"per element, {space.num_DoFs} DoFs")
    Kc = np.array([[2.0, 0.5, -0.2], [0.5, 1.0, 0.3], [-0.2, 0.3, 1.5]])[:d, :d]
    space, op = _hgdt(H, grid, k, Kc)
    op.assemble()
    _check(op, O.Mesh.structured(geom, cells, lo, hi), k, Kc, "constant")


@pytest.mark.parametrize("k", [1, 2])
def test_tensor_diffusion_on_mixed_grid(hgdt_mod, k):
    H = hgdt_mod
    print(f"case tensor diffusion mixed grid order {k}", flush=True)
    V, E = strip_mixed_grid(6, 5, shear=0.2)
    grid = H.Grid.mixed(V, E)
    K = _spd(8, grid.num_elements, 2)
    space, op = _hgdt(H, grid, k, torch.from_numpy(K).cuda())
    op.assemble()
    _check(op, O.Mesh.mixed(V, E), k, K, f"mixed, {space.num_DoFs} DoFs")


def test_tensor_diffusion_refusals(hgdt_mod):
    H = hgdt_mod
    L = H._lib
    print("case tensor diffusion refusals", flush=True)
    grid = H.Grid.structured("quad", [0, 0], [1, 1], [4, 4])
    space = H.ContinuousLagrangeSpace(grid, 1)
    op = H.make_matrix_operator(space, H.Stencil.element)
    K = _spd(9, grid.num_elements, 2)
    K[5, 0, 1] += 1e-3  # one element's tensor not symmetric
    with pytest.raises(L.HgdtError):
        op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand(torch.from_numpy(K).cuda())))
    with pytest.raises(L.HgdtError):
        op.append(H.LocalElementIntegralBilinearForm(H.LocalLaplaceIntegrand([[1.0, 0.2], [0.0, 1.0]])))
    space, op = _hgdt(H, grid, 1, [[2.0, 0.3], [0.3, 1.0]])
    op.assemble()
    u = op.apply_inverse("rhs")
    with pytest.raises(L.HgdtError):
        op.adjoint_log_diffusion_gradient(u, u)
    op.set_format("matrix_free")
    op.assemble()
    with pytest.raises(L.HgdtError):
        op.apply(u)
