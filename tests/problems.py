"""Identical problem set-ups on the CPU oracle and on the HIP path (hgdt).

Configurations follow the workload list of the benchmark (reduced sizes for parity):
  heat_q1   -- Q1 continuous Lagrange, -div(kappa grad u) = f on the unit square, zero Dirichlet
  p1_tets   -- P1 continuous Lagrange on the unit cube split into Kuhn tetrahedra, zero Dirichlet
  swipdg    -- Q1 discontinuous Lagrange, SWIPDG (element + inner coupling + penalty + weak Dirichlet)
"""
import numpy as np

import oracle_lib as O


def oracle_problem(kind, cells, kappa, f=1.0, sigma=16.0, lo=None, hi=None, symmetry=1.0):
    if kind == "heat_q1":
        m = O.Mesh.structured(O.QUAD, cells, lo or [0, 0], hi or [1, 1])
        s = O.Space(m, True, 1)
        op = O.Operator(s, O.STENCIL_ELEMENT)
        op.append_laplace(kappa)
        op.append_source(f)
        op.set_dirichlet(True)
    elif kind == "p1_tets":
        m = O.Mesh.structured(O.TET, cells, lo or [0, 0, 0], hi or [1, 1, 1])
        s = O.Space(m, True, 1)
        op = O.Operator(s, O.STENCIL_ELEMENT)
        op.append_laplace(kappa)
        op.append_source(f)
        op.set_dirichlet(True)
    elif kind == "swipdg":
        m = O.Mesh.structured(O.QUAD, cells, lo or [0, 0], hi or [1, 1])
        s = O.Space(m, False, 1)
        op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
        op.append_laplace(kappa)
        op.append_inner_coupling(symmetry, kappa, kappa)
        op.append_inner_penalty(sigma, kappa)
        op.append_dirichlet_coupling(symmetry, kappa)
        op.append_boundary_penalty(sigma, kappa)
        op.append_source(f)
    else:
        raise ValueError(kind)
    return m, s, op


def hgdt_problem(hgdt, kind, cells, kappa, f=1.0, sigma=16.0, lo=None, hi=None, grid=None, symmetry=1.0):
    """Same problem through the reference-named API of hgdt (the HIP path)."""
    if kind == "heat_q1":
        grid = grid or hgdt.Grid.structured("quad", lo or [0, 0], hi or [1, 1], cells)
        space = hgdt.ContinuousLagrangeSpace(grid, 1)
        op = hgdt.make_matrix_operator(space, hgdt.Stencil.element)
        op.append(hgdt.LocalElementIntegralBilinearForm(hgdt.LocalLaplaceIntegrand(kappa)))
    elif kind == "p1_tets":
        grid = grid or hgdt.Grid.structured("tet", lo or [0, 0, 0], hi or [1, 1, 1], cells)
        space = hgdt.ContinuousLagrangeSpace(grid, 1)
        op = hgdt.make_matrix_operator(space, hgdt.Stencil.element)
        op.append(hgdt.LocalElementIntegralBilinearForm(hgdt.LocalLaplaceIntegrand(kappa)))
    elif kind == "swipdg":
        grid = grid or hgdt.Grid.structured("quad", lo or [0, 0], hi or [1, 1], cells)
        space = hgdt.DiscontinuousLagrangeSpace(grid, 1)
        op = hgdt.make_matrix_operator(space, hgdt.Stencil.element_and_intersection)
        op += hgdt.LocalElementIntegralBilinearForm(hgdt.LocalLaplaceIntegrand(kappa))
        op.append(hgdt.LocalCouplingIntersectionIntegralBilinearForm(
            hgdt.LocalLaplaceIPDGInnerCoupling(symmetry, kappa, kappa)), {}, hgdt.ApplyOn.InnerIntersectionsOnce())
        op.append(hgdt.LocalCouplingIntersectionIntegralBilinearForm(hgdt.LocalIPDGInnerPenalty(sigma, kappa)), {},
                  hgdt.ApplyOn.InnerIntersectionsOnce())
        op.append(hgdt.LocalIntersectionIntegralBilinearForm(hgdt.LocalLaplaceIPDGDirichletCoupling(symmetry, kappa)),
                  {}, hgdt.ApplyOn.CustomBoundaryIntersections(hgdt.AllDirichletBoundaryInfo()))
        op.append(hgdt.LocalIntersectionIntegralBilinearForm(hgdt.LocalIPDGBoundaryPenalty(sigma, kappa)), {},
                  hgdt.ApplyOn.CustomBoundaryIntersections(hgdt.AllDirichletBoundaryInfo()))
    else:
        raise ValueError(kind)
    rhs = hgdt.make_vector_functional(space)
    rhs.append(hgdt.LocalElementIntegralFunctional(hgdt.LocalProductIntegrand().with_ansatz(f)))
    op.append(rhs)
    if kind != "swipdg":
        op.append(hgdt.make_dirichlet_constraints(space, hgdt.AllDirichletBoundaryInfo()))
    return grid, space, op


def matrix_error(vals, ref_vals, row_ptr):
    """max over rows of max_k |a_k - r_k| / max_k |r_k| (entries compared relative to their row's scale)"""
    err = np.abs(vals - ref_vals)
    scale = np.abs(ref_vals)
    n = len(row_ptr) - 1
    row_of = np.repeat(np.arange(n), np.diff(row_ptr))
    row_err = np.zeros(n)
    row_scale = np.zeros(n)
    np.maximum.at(row_err, row_of, err)
    np.maximum.at(row_scale, row_of, scale)
    return float((row_err / np.maximum(row_scale, 1e-300)).max())
