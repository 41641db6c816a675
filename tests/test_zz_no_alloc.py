"""hgdt_space_error_norms_noalloc (fixed device array, no allocation) against hgdt_space_error_norms.

Both are called through the C ABI on the same inputs; the no-allocation variant runs the same kernels over
the same elements and must give the same bits.  Two calls in a row: a Q1 grid of 1100 x 1000 cells (1.1 M
elements: several chunks and two windows of the fixed array), then a smaller union of two P1 tetrahedral
grids (the array keeps the first call's contents).  The device's free memory does not change across the
no-allocation calls, and a grid needing more chunk-table entries than the array holds is refused.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _norms(H, fn, space, u, order, gv, gg, kappa):
    L = H._lib
    out = np.empty(2 * space.grid.num_instances)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
    L.call(fn, space._h, ptr(u), order, ptr(gv), ptr(gg), ptr(kappa), 1.0, out.ctypes.data_as(L._pdbl),
           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return out


def _inputs(H, grid, order, seed):
    space = H.ContinuousLagrangeSpace(grid, 1)
    rng = np.random.default_rng(seed)
    u = torch.from_numpy(rng.standard_normal(space.num_DoFs)).cuda()
    P = grid.quadrature_points(order)
    dim = P.shape[-1]
    X = P.reshape(-1, dim)
    gv = torch.from_numpy(np.sin(X[:, 0]) * np.cos(0.5 * X[:, 1])).cuda()
    gg = torch.from_numpy(np.ascontiguousarray(np.stack([np.cos(X[:, 0]), X[:, 1]] + ([X[:, 2]] if dim == 3 else []),
                                                        axis=1))).cuda()
    kappa = torch.from_numpy(H.fields.lognormal(seed, grid.num_elements)).cuda()
    return space, u, gv, gg, kappa


def test_noalloc_error_norms_equal_the_allocating_entry_point(hgdt_mod):
    H = hgdt_mod
    L = H._lib
    nbytes = ctypes.c_int64()
    L.call("hgdt_noalloc_scratch_bytes", ctypes.byref(nbytes))
    print(f"case no-allocation error norms: fixed array {nbytes.value} bytes", flush=True)
    assert nbytes.value <= 32 << 20
    big = H.Grid.structured("quad", [0, 0], [2.0, 1.0], [1100, 1000])
    small_part = H.Grid.structured("tet", [0, 0, 0], [1, 1, 1], [6, 5, 4])
    small = H.Grid.union([small_part, small_part])
    cases = [("quad 1100 x 1000", big, 4, 3), ("2 x tet 6 x 5 x 4", small, 3, 4)]
    inputs = [_inputs(H, g, order, seed) for _, g, order, seed in cases]
    # reference results from the allocating entry point, with and without g and a diffusion field
    expected = [(_norms(H, "hgdt_space_error_norms", s, u, order, gv, gg, k),
                 _norms(H, "hgdt_space_error_norms", s, u, order, None, None, None))
                for (label, g, order, _), (s, u, gv, gg, k) in zip(cases, inputs)]
    torch.cuda.synchronize()
    # the first call on the device looks the array up; from then on only no-allocation calls run, and the
    # device's free memory must not change
    s, u, gv, gg, k = inputs[1]
    _norms(H, "hgdt_space_error_norms_noalloc", s, u, cases[1][2], gv, gg, k)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for (label, g, order, _), (s, u, gv, gg, k), (ref, ref0) in zip(cases, inputs, expected):
        print(f"  {label}: {g.num_elements} elements, order {order}", flush=True)
        got = _norms(H, "hgdt_space_error_norms_noalloc", s, u, order, gv, gg, k)
        got0 = _norms(H, "hgdt_space_error_norms_noalloc", s, u, order, None, None, None)
        print(f"    allocating {ref} / {ref0}, no-allocation {got} / {got0}", flush=True)
        assert np.array_equal(got, ref) and np.array_equal(got0, ref0), label
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0


def test_noalloc_error_norms_refuses_grids_beyond_its_array(hgdt_mod):
    """8193 single-cell instances need 8193 chunks (a chunk never straddles two instances) plus one window
    mark: more than the array's 8192 chunk-table entries, refused before any launch"""
    H = hgdt_mod
    print("case no-allocation error norms: 8193 one-cell instances (refused)", flush=True)
    cell = H.Grid.structured("quad", [0, 0], [1, 1], [1, 1])
    grid = H.Grid.union([cell] * 8193)
    space = H.ContinuousLagrangeSpace(grid, 1)
    u = torch.zeros(space.num_DoFs, dtype=torch.float64, device="cuda")
    with pytest.raises(H.not_supported, match="chunk-table entries"):
        _norms(H, "hgdt_space_error_norms_noalloc", space, u, 2, None, None, None)
