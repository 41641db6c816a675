"""The header-only C++ classes (include/hgdt.hh) and the example written with them
(dune-gdt_amd/examples/stationary_heat_equation.cc, the counterpart of dune-gdt's
examples/stationary-heat-equation.cc).

CPU: the header compiles on its own with warnings as errors; the built example refuses to run without a GPU
(status 8, no CPU fallback).  GPU: the example's solution equals the oracle's Jacobi-PCG solution of the same
problem to 1e-10 relative L2, and its H1-semi error equals the oracle's error of that solution to 1e-12.
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_lib as O
from test_capi import gpu_available

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXAMPLE = os.path.join(ROOT, "dune-gdt_amd", "bin", "stationary_heat_equation")


def test_header_compiles_standalone():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.cc")
        with open(src, "w") as f:
            f.write('#include "hgdt.hh"\nint main() { return hgdt::Stencil::element == hgdt::Stencil::element ? 0 : 1; }\n')
        r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                            "-I", os.path.join(ROOT, "include"), src], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU refusal path")
def test_example_refuses_without_gpu():
    assert os.path.exists(EXAMPLE), "build the example with __graft_entry__.build()"
    r = subprocess.run([EXAMPLE, "4"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "status 8" in r.stderr


@pytest.mark.gpu
def test_example_matches_oracle():
    n = 64
    print(f"case C++ example stationary_heat_equation n={n}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "u.bin")
        r = subprocess.run([EXAMPLE, str(n), out], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        res = json.loads(r.stdout.strip().splitlines()[-1])
        u = np.fromfile(out, dtype=np.float64)
    m = O.Mesh.structured(O.QUAD, [n, n], [-1, -1], [1, 1])
    s = O.Space(m, True, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    op.append_laplace(1.0)
    op.append_source(esv2007=True, f_order=3)
    op.set_dirichlet(True)
    op.assemble(4)
    rp, cols = op.pattern()
    xo, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-10, nthreads=4)
    assert rc == 0 and u.shape == xo.shape
    diff = np.linalg.norm(u - xo) / np.linalg.norm(xo)
    err_o = O.esv2007_h1_semi_error(s, u, 6)
    print(f"  iterations {res['iterations']} (oracle {it}), rel L2 {diff:.2e}, h1 error {res['h1_semi_error']:.6e} "
          f"(oracle on the same solution {err_o:.6e}), l2 error {res['l2_error']:.3e}")
    assert diff <= 1e-10 and abs(res["iterations"] - it) <= 2
    assert abs(res["h1_semi_error"] - err_o) <= 1e-12 * err_o
    assert 0 < res["l2_error"] < res["h1_semi_error"]


@pytest.mark.gpu
def test_example_on_a_mixed_grid_matches_oracle():
    """the same example on a grid of quadrilaterals (left half) and triangles (right half) built with
    Grid::mixed: solution parity with the oracle on the same element list (1e-10) and the H1-semi and L2
    errors equal the oracle's errors of that solution at its quadrature points (1e-12)"""
    from mixed_grids import strip_mixed_grid

    n = 32
    print(f"case C++ example stationary_heat_equation n={n} on a mixed grid", flush=True)
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "u.bin")
        r = subprocess.run([EXAMPLE, str(n), out, "mixed"], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        res = json.loads(r.stdout.strip().splitlines()[-1])
        u = np.fromfile(out, dtype=np.float64)
    V, E = strip_mixed_grid(n, n)
    m = O.Mesh.mixed(2.0 * V - 1.0, E)
    s = O.Space(m, True, 1)
    op = O.Operator(s, O.STENCIL_ELEMENT)
    op.append_laplace(1.0)
    op.append_source(esv2007=True, f_order=3)
    op.set_dirichlet(True)
    op.assemble(4)
    rp, cols = op.pattern()
    xo, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-10, nthreads=4)
    assert rc == 0 and u.shape == xo.shape
    diff = np.linalg.norm(u - xo) / np.linalg.norm(xo)
    P = m.quadrature_points(6)
    c = np.pi / 2
    g = np.cos(c * P[:, 0]) * np.cos(c * P[:, 1])
    gg = np.stack([-c * np.sin(c * P[:, 0]) * np.cos(c * P[:, 1]), -c * np.cos(c * P[:, 0]) * np.sin(c * P[:, 1])], 1)
    l2o, h1o = np.sqrt(O.error_norms(s, u, 6, g, np.ascontiguousarray(gg)))
    print(f"  iterations {res['iterations']} (oracle {it}), rel L2 {diff:.2e}, h1 error {res['h1_semi_error']:.6e} "
          f"(oracle {h1o:.6e}), l2 error {res['l2_error']:.6e} (oracle {l2o:.6e})")
    assert diff <= 1e-10 and abs(res["iterations"] - it) <= 2
    assert abs(res["h1_semi_error"] - h1o) <= 1e-12 * h1o and abs(res["l2_error"] - l2o) <= 1e-12 * l2o


SWIPDG = os.path.join(ROOT, "dune-gdt_amd", "bin", "elliptic_swipdg")


def _exact(X):
    return np.cos(np.pi / 2 * X[..., 0]) * np.cos(np.pi / 2 * X[..., 1]) + 0.25 * X[..., 0] - 0.5 * X[..., 1]


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["assembled", "matrix-free"])
def test_swipdg_example_with_dirichlet_data_matches_oracle(mode):
    """dune-gdt_amd/examples/elliptic_swipdg.cc: SWIPDG with nonzero Dirichlet data through boundary
    functionals, with the assembled matrix or matrix-free; solution parity with the oracle (1e-10), H1-semi
    error parity (1e-12) and first-order convergence under refinement"""
    errs = []
    for n in (16, 32):
        print(f"case C++ example elliptic_swipdg n={n} {mode}", flush=True)
        with tempfile.TemporaryDirectory() as d:
            out = os.path.join(d, "u.bin")
            r = subprocess.run([SWIPDG, str(n), out, mode], capture_output=True, text=True, timeout=120)
            assert r.returncode == 0, r.stderr
            res = json.loads(r.stdout.strip().splitlines()[-1])
            u = np.fromfile(out, dtype=np.float64)
        m = O.Mesh.structured(O.QUAD, [n, n], [-1, -1], [1, 1])
        s = O.Space(m, False, 1)
        op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
        op.append_laplace(1.0)
        op.append_inner_coupling(1.0, 1.0, 1.0)
        op.append_inner_penalty(16.0, 1.0)
        op.append_dirichlet_coupling(1.0, 1.0)
        op.append_boundary_penalty(16.0, 1.0)
        op.append_source(esv2007=True, f_order=3)
        op.append_boundary_source(0, 1.0, 1.0, _exact(m.boundary_quadrature_points(3)), g_order=3)
        op.append_boundary_source(1, 16.0, 1.0, _exact(m.boundary_quadrature_points(1 + 3)), g_order=3)
        op.assemble(4)
        rp, cols = op.pattern()
        xo, it, rel, rc = O.pcg_jacobi(rp, cols, op.values(), op.rhs(), rtol=1e-10, nthreads=4)
        diff = np.linalg.norm(u - xo) / np.linalg.norm(xo)
        P = m.quadrature_points(10)
        X, Y = P[..., 0], P[..., 1]
        gg = np.stack([-np.pi / 2 * np.sin(np.pi / 2 * X) * np.cos(np.pi / 2 * Y) + 0.25,
                       -np.pi / 2 * np.cos(np.pi / 2 * X) * np.sin(np.pi / 2 * Y) - 0.5], axis=-1)
        l2o, h1o = O.error_norms(s, u, 10, _exact(P), gg)
        print(f"  iterations {res['iterations']} (oracle {it}), rel L2 {diff:.2e}, h1 error {res['h1_semi_error']:.6e} "
              f"(oracle on the same solution {np.sqrt(h1o):.6e})")
        assert diff <= 1e-10 and abs(res["iterations"] - it) <= 2
        assert abs(res["h1_semi_error"] - np.sqrt(h1o)) <= 1e-12 * np.sqrt(h1o)
        errs.append(res["h1_semi_error"])
    assert 1.8 < errs[0] / errs[1] < 2.2  # O(h) in the broken H1 seminorm
