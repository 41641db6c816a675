"""Writes tests/golden/known_answers.json.

Two kinds of vectors:
  * values transcribed from the reference's own tests (file:line given per entry);
  * exact rational values of the same quantities, computed here by brute force with
    fractions.Fraction (exact polynomial integration of the P2/Q2 Lagrange bases on the
    9x2 grid of dune/gdt/test/integrands/integrands.hh:64-70), which pin the oracle to the
    mathematics independently of the reference's floating-point rounding.
Run:  python tests/golden/make_golden.py
"""
import json
import os
from fractions import Fraction as F

HERE = os.path.dirname(os.path.abspath(__file__))
NX, NY, LX, LY = 9, 2, F(3), F(1)


def pmul(a, b):
    r = [F(0)] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            r[i + j] += x * y
    return r


def pint(a):
    return sum(c / (k + 1) for k, c in enumerate(a))


def pder(a):
    return [k * a[k] for k in range(1, len(a))] or [F(0)]


def lagrange_1d(order):
    nodes = [F(i, order) for i in range(order + 1)]
    out = []
    for m in range(order + 1):
        p = [F(1)]
        for l in range(order + 1):
            if l != m:
                p = pmul(p, [-nodes[l] / (nodes[m] - nodes[l]), 1 / (nodes[m] - nodes[l])])
        out.append(p)
    return out


def stats(A, n):
    vals = list(A.values()) + [F(0)] * (n * n - len(A))
    return min(vals), max(vals), sum(v * v for v in vals)


def q2_matrices():
    """Q2 on the 9x2 rectangle grid: DoFs identified by lattice coordinates of the nodes."""
    hx, hy = LX / NX, LY / NY
    L = lagrange_1d(2)
    dL = [pder(p) for p in L]
    ii = lambda p, q: pint(pmul(p, q))
    K, M = {}, {}
    for ey in range(NY):
        for ex in range(NX):
            nodes = [(2 * ex + t % 3, 2 * ey + t // 3) for t in range(9)]
            for t in range(9):
                for u in range(9):
                    i, j, k, l = t % 3, t // 3, u % 3, u // 3
                    kv = ii(dL[i], dL[k]) * ii(L[j], L[l]) * hy / hx + ii(L[i], L[k]) * ii(dL[j], dL[l]) * hx / hy
                    mv = ii(L[i], L[k]) * ii(L[j], L[l]) * hx * hy
                    key = (nodes[t], nodes[u])
                    K[key] = K.get(key, F(0)) + kv
                    M[key] = M.get(key, F(0)) + mv
    n = (2 * NX + 1) * (2 * NY + 1)
    return K, M, n


def p2_triangle_matrices():
    """P2 on the Kuhn triangulation (diagonal from (0,0) to (1,1) of each cell), exact monomial
    integration over the reference triangle."""
    hx, hy = LX / NX, LY / NY

    def mono_int(a, b):  # int_T x^a y^b over the reference triangle = a! b! / (a+b+2)!
        from math import factorial

        return F(factorial(a) * factorial(b), factorial(a + b + 2))

    # P2 basis as polynomials in (x, y): dict {(a,b): coeff}
    def poly_mul(p, q):
        r = {}
        for (a, b), c in p.items():
            for (d, e), f in q.items():
                r[(a + d, b + e)] = r.get((a + d, b + e), F(0)) + c * f
        return r

    lam = [{(0, 0): F(1), (1, 0): F(-1), (0, 1): F(-1)}, {(1, 0): F(1)}, {(0, 1): F(1)}]
    basis = []
    for a in range(3):
        basis.append(poly_mul(lam[a], {**{k: 2 * v for k, v in lam[a].items()},
                                       (0, 0): 2 * lam[a].get((0, 0), F(0)) - 1}))
    for a, b in ((0, 1), (0, 2), (1, 2)):
        basis.append({k: 4 * v for k, v in poly_mul(lam[a], lam[b]).items()})

    def deriv(p, axis):
        r = {}
        for (a, b), c in p.items():
            if axis == 0 and a > 0:
                r[(a - 1, b)] = r.get((a - 1, b), F(0)) + a * c
            if axis == 1 and b > 0:
                r[(a, b - 1)] = r.get((a, b - 1), F(0)) + b * c
        return r

    def integrate(p):
        return sum(c * mono_int(a, b) for (a, b), c in p.items())

    # node positions in reference coordinates (x2 lattice): vertices then edge midpoints
    ref_nodes = [(0, 0), (2, 0), (0, 2), (1, 0), (0, 1), (1, 1)]
    K, M = {}, {}
    for ey in range(NY):
        for ex in range(NX):
            # two triangles: (v00, v10, v11) and (v00, v01, v11): affine maps x = x0 + J xh
            for J in (((hx, hx), (F(0), hy)), ((F(0), hx), (hy, hy))):
                # columns of J: v1 - v0, v2 - v0
                det = J[0][0] * J[1][1] - J[0][1] * J[1][0]
                Ji = ((J[1][1] / det, -J[0][1] / det), (-J[1][0] / det, J[0][0] / det))
                nodes = []
                for (rx, ry) in ref_nodes:
                    px = 2 * ex * 1 + (J[0][0] * rx + J[0][1] * ry) / hx
                    py = 2 * ey * 1 + (J[1][0] * rx + J[1][1] * ry) / hy
                    nodes.append((px, py))
                for i in range(6):
                    for j in range(6):
                        # physical gradient components J^{-T} grad_ref as polynomials
                        def phys(idx):
                            gx, gy = deriv(basis[idx], 0), deriv(basis[idx], 1)
                            cx, cy = {}, {}
                            for k, v in gx.items():
                                cx[k] = cx.get(k, F(0)) + Ji[0][0] * v
                                cy[k] = cy.get(k, F(0)) + Ji[0][1] * v
                            for k, v in gy.items():
                                cx[k] = cx.get(k, F(0)) + Ji[1][0] * v
                                cy[k] = cy.get(k, F(0)) + Ji[1][1] * v
                            return cx, cy
                        ax, ay = phys(i)
                        bx, by = phys(j)
                        kv = abs(det) * (integrate(poly_mul(ax, bx)) + integrate(poly_mul(ay, by)))
                        mv = abs(det) * integrate(poly_mul(basis[i], basis[j]))
                        key = (nodes[i], nodes[j])
                        K[key] = K.get(key, F(0)) + kv
                        M[key] = M.get(key, F(0)) + mv
    n = (2 * NX + 1) * (2 * NY + 1)
    return K, M, n


# P2 on Kuhn tetrahedra: grid of TNX x TNY x TNZ cells of widths TH (anisotropic), exact rationals
TCELLS, TH = (2, 1, 1), (F(1, 2), F(1, 2), F(3, 4))


def p2_tet_matrices():
    """P2 Laplace and mass matrices on the Kuhn tetrahedra of TCELLS cells (every cube split into the six
    tetrahedra (v0, v0 + e_p0, v0 + e_p0 + e_p1, v7) for the axis permutations p), by exact integration of
    the monomials over the reference tetrahedron (int x^a y^b z^c = a! b! c! / (a + b + c + 3)!).  The
    basis is lambda_a (2 lambda_a - 1) at the vertices and 4 lambda_a lambda_b at the edge midpoints; DoFs
    are identified by their node's coordinates on the half-cell lattice."""
    from itertools import permutations
    from math import factorial

    def mul(p, q):
        r = {}
        for a, c in p.items():
            for b, d in q.items():
                k = (a[0] + b[0], a[1] + b[1], a[2] + b[2])
                r[k] = r.get(k, F(0)) + c * d
        return r

    def der(p, axis):
        r = {}
        for a, c in p.items():
            if a[axis] > 0:
                k = list(a)
                k[axis] -= 1
                r[tuple(k)] = r.get(tuple(k), F(0)) + a[axis] * c
        return r

    def integrate(p):
        return sum(c * F(factorial(a[0]) * factorial(a[1]) * factorial(a[2]), factorial(sum(a) + 3))
                   for a, c in p.items())

    lam = [{(0, 0, 0): F(1), (1, 0, 0): F(-1), (0, 1, 0): F(-1), (0, 0, 1): F(-1)},
           {(1, 0, 0): F(1)}, {(0, 1, 0): F(1)}, {(0, 0, 1): F(1)}]
    basis = [mul(lam[a], {k: 2 * v for k, v in lam[a].items()} | {(0, 0, 0): 2 * lam[a].get((0, 0, 0), F(0)) - 1})
             for a in range(4)]
    edges = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    basis += [{k: 4 * v for k, v in mul(lam[a], lam[b]).items()} for a, b in edges]
    ref_vert = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    grads = [[der(p, ax) for ax in range(3)] for p in basis]
    # integrals of products of reference derivatives / values (the same for every tetrahedron)
    DD = [[[[integrate(mul(grads[i][a], grads[j][b])) for b in range(3)] for a in range(3)] for j in range(10)]
          for i in range(10)]
    VV = [[integrate(mul(basis[i], basis[j])) for j in range(10)] for i in range(10)]
    K, Mm = {}, {}
    for cz in range(TCELLS[2]):
        for cy in range(TCELLS[1]):
            for cx in range(TCELLS[0]):
                v0 = (cx, cy, cz)
                for perm in permutations(range(3)):
                    pts = [v0]
                    cur = list(v0)
                    for ax in perm:
                        cur[ax] += 1
                        pts.append(tuple(cur))
                    # columns of J: physical edge vectors from vertex 0 (lattice units times TH)
                    J = [[(pts[c + 1][r] - pts[0][r]) * TH[r] for c in range(3)] for r in range(3)]
                    det = (J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0])
                           + J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]))
                    cof = [[(J[(r + 1) % 3][(c + 1) % 3] * J[(r + 2) % 3][(c + 2) % 3]
                             - J[(r + 1) % 3][(c + 2) % 3] * J[(r + 2) % 3][(c + 1) % 3]) for c in range(3)] for r in range(3)]
                    Ji = [[cof[c][r] / det for c in range(3)] for r in range(3)]  # J^{-1}
                    G = [[sum(Ji[a][k] * Ji[b][k] for k in range(3)) for b in range(3)] for a in range(3)]
                    # node keys: doubled lattice coordinates of vertices and edge midpoints
                    lat = [tuple(2 * p[r] for r in range(3)) for p in pts]
                    nodes = lat + [tuple((lat[a][r] + lat[b][r]) // 2 for r in range(3)) for a, b in edges]
                    for i in range(10):
                        for j in range(10):
                            kv = abs(det) * sum(G[a][b] * DD[i][j][a][b] for a in range(3) for b in range(3))
                            key = (nodes[i], nodes[j])
                            K[key] = K.get(key, F(0)) + kv
                            Mm[key] = Mm.get(key, F(0)) + abs(det) * VV[i][j]
    keys = sorted({k[0] for k in K})
    return K, Mm, keys


def dense_eigenvalues(A, keys):
    import numpy as np

    idx = {k: i for i, k in enumerate(keys)}
    D = np.zeros((len(keys), len(keys)))
    for (a, b), v in A.items():
        D[idx[a], idx[b]] = float(v)
    return [float(x) for x in np.linalg.eigvalsh(D)]


def main():
    out = {"reference": {
        # dune/gdt/test/integrands/integrands_laplace.cc:131-133 (tol 1e-13, 1e-13, 5e-12)
        "p2_laplace_cube": {"min": -1.896296296296300, "max": 6.162962962962970, "sumsq": 1704.099039780521,
                            "tol": [1e-13, 1e-13, 5e-12]},
        "p2_laplace_simplex": {"min": -2.0, "max": 5.777777777777780, "sumsq": 2481.524691358029,
                               "tol": [1e-13, 1e-13, 5e-12]},
        # dune/gdt/test/integrands/integrands_product.cc:122-124 (tol 1e-13)
        "p2_mass_cube": {"min": -0.002962962962963, "max": 0.047407407407407, "sumsq": 0.066475994513031,
                         "tol": [1e-13, 1e-13, 1e-13]},
        "p2_mass_simplex": {"min": -0.001851851851852, "max": 0.029629629629630, "sumsq": 0.058804012345679,
                            "tol": [1e-13, 1e-13, 1e-13]},
        # dune/gdt/test/stationary-heat-equation/stationary_heat_equation__ESV2007__table_1.mini:31-33
        "esv2007_q1dg_cube": {"h": [3.54e-01, 1.77e-01, 8.84e-02], "h1_semi": [2.52e-01, 1.26e-01, 6.30e-02]},
        # same file, :11-13: P1-DG on the ALU conforming simplex grid (4 x 4 cubes, Kuhn split, bisection-refined)
        "esv2007_p1dg_simplex": {"h": [3.54e-01, 1.77e-01, 8.84e-02], "h1_semi": [3.28e-01, 1.62e-01, 8.04e-02]},
    }, "exact": {}}
    K, M, n = q2_matrices()
    for name, A in (("p2_laplace_cube", K), ("p2_mass_cube", M)):
        mn, mx, ss = stats(A, n)
        out["exact"][name] = {"min": [mn.numerator, mn.denominator], "max": [mx.numerator, mx.denominator],
                              "sumsq": [ss.numerator, ss.denominator]}
    K, M, n = p2_triangle_matrices()
    for name, A in (("p2_laplace_simplex", K), ("p2_mass_simplex", M)):
        mn, mx, ss = stats(A, n)
        out["exact"][name] = {"min": [mn.numerator, mn.denominator], "max": [mx.numerator, mx.denominator],
                              "sumsq": [ss.numerator, ss.denominator]}
    K, M, keys = p2_tet_matrices()
    for name, A in (("p2_laplace_tet_kuhn", K), ("p2_mass_tet_kuhn", M)):
        mn, mx, ss = stats(A, len(keys))
        out["exact"][name] = {"min": [mn.numerator, mn.denominator], "max": [mx.numerator, mx.denominator],
                              "sumsq": [ss.numerator, ss.denominator], "size": len(keys),
                              "cells": list(TCELLS), "widths": [[w.numerator, w.denominator] for w in TH],
                              "eigenvalues": dense_eigenvalues(A, keys)}
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(out, f, indent=1)
    for k, v in out["exact"].items():
        print(k, {kk: float(F(*vv)) for kk, vv in v.items() if kk in ("min", "max", "sumsq")})


if __name__ == "__main__":
    main()
