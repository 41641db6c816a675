"""Mixed triangle/quadrilateral grids for the tests: the reference's mixed grid and a larger strip grid.

The reference grid is the one of dune/gdt/test/spaces/base.hh:355-376 (make_mixed_grid, 2d): two
quadrilaterals and two triangles, globally refined once.  UGGrid's red refinement splits a quadrilateral at
its edge midpoints and centre into four and a triangle at its edge midpoints into four; the children are built
here in the same way (the element order of UGGrid's leaf view is not reproduced: the tests compare a library
with the oracle on the same element list, and the closed forms do not depend on the order).
"""
import numpy as np


def reference_mixed_grid():
    """vertices and elements (reference vertex order) of make_mixed_grid<UG_2D>() before refinement"""
    V = np.array([(-1., -1.5), (-1., -1.25), (-1., -1.), (-1.5, -1.5), (-1.5, -1.25), (-1.5, -1.), (-1.75, -1.25)])
    E = [[3, 0, 4, 1], [4, 1, 5, 2], [4, 6, 3], [4, 5, 6]]
    return V, E


def refine(V, E):
    """one red refinement of every element (shared edge midpoints)"""
    V = [np.asarray(v, np.float64) for v in V]
    mid = {}

    def m(a, b):
        k = (min(a, b), max(a, b))
        if k not in mid:
            V.append((V[a] + V[b]) / 2)
            mid[k] = len(V) - 1
        return mid[k]

    out = []
    for el in E:
        if len(el) == 4:  # cube vertex b: bits (x, y)
            v0, v1, v2, v3 = el
            a, b, l, r = m(v0, v1), m(v2, v3), m(v0, v2), m(v1, v3)
            V.append((V[v0] + V[v3]) / 2)
            c = len(V) - 1
            out += [[v0, a, l, c], [a, v1, c, r], [l, c, v2, b], [c, r, b, v3]]
        else:
            v0, v1, v2 = el
            a, b, c = m(v0, v1), m(v0, v2), m(v1, v2)
            out += [[v0, a, b], [a, v1, c], [b, c, v2], [c, b, a]]
    return np.array(V), out


def refined_reference_mixed_grid():
    """make_mixed_grid<UG_2D>() after global_refine(1): 8 quadrilaterals, 8 triangles, area 0.3125"""
    return refine(*reference_mixed_grid())


def strip_mixed_grid(nx, ny, shear=0.0):
    """the unit square in nx x ny cells: columns left of nx // 2 quadrilaterals, the others split into two
    triangles (v0, v0+e_x, v0+e_x+e_y) and (v0, v0+e_y, v0+e_x+e_y); the triangles' and quadrilaterals' shared
    edges are whole cell edges, so the grid is conforming.  shear != 0 maps (x, y) -> (x + shear y, y)
    (parallelograms)."""
    xs, ys = np.meshgrid(np.linspace(0., 1., nx + 1), np.linspace(0., 1., ny + 1))
    V = np.stack([xs.ravel() + shear * ys.ravel(), ys.ravel()], axis=1)
    vid = lambda i, j: j * (nx + 1) + i
    E = []
    for j in range(ny):
        for i in range(nx):
            c = [vid(i, j), vid(i + 1, j), vid(i, j + 1), vid(i + 1, j + 1)]
            if i < nx // 2:
                E.append(c)
            else:
                E.append([c[0], c[1], c[3]])
                E.append([c[0], c[2], c[3]])
    return V, E


def grid_area(V, E):
    total = 0.
    for el in E:
        p = V[el]
        if len(el) == 4:
            a, b = p[1] - p[0], p[2] - p[0]
        else:
            a, b = p[1] - p[0], p[2] - p[0]
        det = abs(a[0] * b[1] - a[1] * b[0])
        total += det if len(el) == 4 else det / 2
    return total
