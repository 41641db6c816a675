"""Launch-shape autotuner for libhgdt's assembly and Jacobi-PCG kernels.

A candidate launch shape is timed on the operator itself: for the solver, a fixed number of PCG
iterations (relative tolerance 0, so exactly `iterations` iterations run) measured with HIP events
inside libhgdt; for assembly, MatrixOperator.assemble() bracketed by torch events.  The operator
application (kernel 3) and the vector kernels (kernel 4) have separate row partitions and are tuned
one after the other; with symmetric DIA storage the tuner then times the application with and without
column staging (p entries of a row block read from LDS) at the chosen shapes and keeps the faster.
Results are cached per problem signature in memory and optionally in a JSON file.
"""
import json
import os

import torch

from . import _lib as L

APPLY_CANDIDATES = [(64, 1), (128, 1), (256, 1), (512, 1), (1024, 1), (256, 2), (512, 2), (128, 2), (256, 4),
                    (512, 4), (256, 8), (512, 8), (1024, 4), (256, 16)]
VECTOR_CANDIDATES = [(256, 1), (256, 2), (256, 4), (256, 8), (512, 1), (512, 2), (512, 4), (512, 8),
                     (1024, 1), (1024, 2), (1024, 4), (128, 4), (128, 8)]
ASSEMBLY_CANDIDATES = [64, 128, 256]

_memory = {}

# launch shapes measured on MI355X for the benchmark workloads, keyed by signature()
TABLE_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "launch_table.json")


def signature(op):
    fi = op.format_info()
    return f"n{op.shape[0]}_nnz{op.nnz}_inst{op.grid.num_instances}_{fi['format']}_d{fi['num_diagonals']}"


def _solve_ms(op, iterations):
    try:
        op.apply_inverse("rhs", opts={"precision": 0.0, "max_iter": iterations}, check_every=16)
    except L.linear_solver_failed:
        pass
    return op.last_solve["solve_ms"], op.last_solve["iterations"]


def _solver_us_per_iteration(op, iterations):
    """time per PCG iteration from two run lengths (fixed set-up and drain costs cancel)"""
    t1, i1 = _solve_ms(op, iterations)
    t2, i2 = _solve_ms(op, 2 * iterations)
    if i2 <= i1:  # converged early: fall back to the plain ratio
        return t2 * 1e3 / max(1, i2)
    return (t2 - t1) * 1e3 / (i2 - i1)


def _staging_applies(op):
    """whether the column-staged application runs for this operator at its current launch shapes"""
    op.set_column_staging(1)
    _solve_ms(op, 2)
    return bool(op.last_solve["column_staged"])


def tune_solver(op, iterations=40, repeats=2, cache_file=None, verbose=False):
    """Pick (block, items) for operator application and vector kernels; returns the choice and timings."""
    op.apply_inverse("rhs")  # prepares the storage format, warms the kernels
    key = signature(op)
    cached = _memory.get(key)
    if cached is None and cache_file and os.path.exists(cache_file):
        with open(cache_file) as f:
            cached = json.load(f).get(key)
    if cached is not None:
        op.set_launch(L.KERNEL_APPLY, *cached["apply"])
        op.set_launch(L.KERNEL_VECTOR, *cached["vector"])
        op.set_column_staging(cached.get("staging", -1))
        return dict(cached, cached=True)
    timings = {}

    def best_of(kernel, candidates, fixed):
        results = {}
        for cand in candidates:
            op.set_launch(kernel, *cand)
            op.set_launch(*fixed)
            t = min(_solver_us_per_iteration(op, iterations) for _ in range(repeats))
            results[cand] = t
            if verbose:
                print(f"  kernel {kernel} block {cand[0]:4d} items {cand[1]}: {t:8.2f} us/iteration", flush=True)
        return min(results, key=results.get), results

    vec0 = op.get_launch(L.KERNEL_VECTOR)
    best_a, ra = best_of(L.KERNEL_APPLY, APPLY_CANDIDATES, (L.KERNEL_VECTOR, *vec0))
    best_v, rv = best_of(L.KERNEL_VECTOR, VECTOR_CANDIDATES, (L.KERNEL_APPLY, *best_a))
    op.set_launch(L.KERNEL_APPLY, *best_a)
    op.set_launch(L.KERNEL_VECTOR, *best_v)
    timings = {"apply": {f"{b}x{i}": t for (b, i), t in ra.items()},
               "vector": {f"{b}x{i}": t for (b, i), t in rv.items()}}
    choice = {"apply": list(best_a), "vector": list(best_v), "us_per_iteration": rv[best_v]}
    if op.format_info()["format"] == "dia_symmetric" and best_a[1] == 1 and _staging_applies(op):
        # column staging off / on at the chosen shapes (only where the staged kernel applies: one row per
        # lane and a stencil whose windows fit; elsewhere both settings would time the same kernel and the
        # automatic choice, -1, is kept)
        rs = {}
        for mode in (0, 1):
            op.set_column_staging(mode)
            rs[mode] = min(_solver_us_per_iteration(op, iterations) for _ in range(repeats))
            if verbose:
                print(f"  column staging {mode}: {rs[mode]:8.2f} us/iteration", flush=True)
        best_s = min(rs, key=rs.get)
        op.set_column_staging(best_s)
        timings["staging"] = {str(m): t for m, t in rs.items()}
        choice["staging"], choice["us_per_iteration"] = best_s, rs[best_s]
    else:
        op.set_column_staging(-1)
        choice["staging"] = -1
    _memory[key] = choice
    if cache_file:
        data = {}
        if os.path.exists(cache_file):
            with open(cache_file) as f:
                data = json.load(f)
        data[key] = choice
        os.makedirs(os.path.dirname(os.path.abspath(cache_file)), exist_ok=True)
        with open(cache_file, "w") as f:
            json.dump(data, f, indent=1)
    return dict(choice, timings=timings, cached=False)


def apply_table(op, table_file=TABLE_FILE):
    """Set the solver launch shapes recorded for this operator's signature; returns the entry or None."""
    if not os.path.exists(table_file):
        return None
    with open(table_file) as f:
        entry = json.load(f).get(signature(op))
    if entry is None:
        return None
    op.set_launch(L.KERNEL_APPLY, *entry["apply"])
    op.set_launch(L.KERNEL_VECTOR, *entry["vector"])
    op.set_column_staging(entry.get("staging", -1))
    return dict(entry, cached=True)


def save_table(op, choice, table_file):
    """Record a tune_solver() choice for this operator's signature in a JSON launch table."""
    data = {}
    if os.path.exists(table_file):
        with open(table_file) as f:
            data = json.load(f)
    data[signature(op)] = {"apply": list(choice["apply"]), "vector": list(choice["vector"]),
                           "us_per_iteration": choice["us_per_iteration"]}
    if "staging" in choice:
        data[signature(op)]["staging"] = choice["staging"]
    os.makedirs(os.path.dirname(os.path.abspath(table_file)), exist_ok=True)
    with open(table_file, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def tune_assembly(op, repeats=3, verbose=False):
    """Pick the block size of the element / intersection / scatter kernels (one launch shape each)."""
    best = {}
    for kernel in (L.KERNEL_ELEMENT, L.KERNEL_INTERSECTION, L.KERNEL_GATHER):
        results = {}
        for block in ASSEMBLY_CANDIDATES:
            op.set_launch(kernel, block, 1)
            op.assemble()
            torch.cuda.synchronize()
            ts = []
            for _ in range(repeats):
                s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                s.record()
                op.assemble()
                e.record()
                torch.cuda.synchronize()
                ts.append(s.elapsed_time(e))
            results[block] = min(ts)
            if verbose:
                print(f"  kernel {kernel} block {block}: {results[block]:.3f} ms", flush=True)
        b = min(results, key=results.get)
        op.set_launch(kernel, b, 1)
        best[kernel] = b
    return best
