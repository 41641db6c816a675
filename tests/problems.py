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


def bisection_refined_triangles(n, lo=-1.0, hi=1.0):
    """The simplex grid of the reference's ESV2007 test (ESV2007.hh:66-69): a 4 x 4 cube grid split into
    Kuhn triangles, then refined by newest-vertex bisection (ALU conforming, two bisections per halving of
    h).  Each bisection pair splits a triangle's hypotenuse and then its legs, so every 2 x 2 block of
    cells of width (hi - lo) / n ends up with its diagonals meeting at the block's centre ("union jack");
    n = 8, 16, 32 are the grids after 2, 4 and 6 bisections.  Returns (coords [N, 2], connectivity [E, 3])."""
    assert n % 2 == 0
    xs = np.linspace(lo, hi, n + 1)
    X = np.array([[x, y] for y in xs for x in xs])
    conn = []
    for j in range(n):
        for i in range(n):
            v00, v10 = j * (n + 1) + i, j * (n + 1) + i + 1
            v01, v11 = v00 + n + 1, v10 + n + 1
            if i % 2 == j % 2:  # the block centre is this cell's corner v11 or v00
                conn += [[v00, v10, v11], [v00, v01, v11]]
            else:  # the block centre is at v10 or v01
                conn += [[v10, v11, v01], [v10, v00, v01]]
    return X, np.array(conn, np.int32)


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
