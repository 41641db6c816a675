"""Benchmark: global assembly + Jacobi-PCG (1e-10) on one MI355X, FP64.

Headline workload (the configuration the metric is quoted on): P1 continuous Lagrange on the unit cube
split into 128^3 x 6 Kuhn tetrahedra (2 146 689 DoFs), kappa_e = exp(g_e), g ~ N(0,1) from seed, f = 1,
zero Dirichlet.  A step = MatrixOperator.assemble() (values + rhs + Dirichlet) followed by apply_inverse
from x0 = 0 to ||r|| <= 1e-10 ||b|| (with the true residual recomputed at exit); inputs are resident in
HBM before timing.

Prints one JSON line: the headline fields (value = DoFs / step time), the FP64 errors of the step's
results against the oracle (CPU restatement), the roofline of the dominant kernel, the CPU baseline (the
oracle on this host, all threads and one thread), and -- unless --no-extras -- the other workloads:
  c3        SWIPDG Q1-DG 512^2 with high-contrast channels, element-block storage and matrix-free;
  batched   256 uneven Q1 problems in one batch against 256 sequential plain GPU solves;
  adjoint   dJ/dlog(kappa) for 32 x P1 64^3 x 6 through torch.autograd against the per-instance path,
            checked by central differences on 64 elements and by the oracle's gradient of one instance.
--config c1|c3|c4|c5 prints a line for that workload alone.
"""
import argparse
import ctypes
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dune-gdt_amd"))

METRIC = "DoFs/s assembly+CG @1e-10 rel. residual FP64; batched/adjoint vs plain; % HBM peak"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    return O


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def host_cpu():
    """model name and physical core count of this host (/proc/cpuinfo)"""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    if phys is not None:
        cores.add((phys, core))
    return {"cpu_model": model, "physical_cores": len(cores) or None, "logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def row_relative_error(vals, ref, row_ptr):
    """max over rows of max_k |a_k - r_k| / max_k |r_k|"""
    import numpy as np

    starts = row_ptr[:-1]
    nonempty = np.diff(row_ptr) > 0
    err = np.maximum.reduceat(np.abs(vals - ref), starts[nonempty])
    scale = np.maximum.reduceat(np.abs(ref), starts[nonempty])
    return float((err / np.maximum(scale, 1e-300)).max())


def free_gpu(torch):
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


# ---------------------------------------------------------------------------------------------------------
# problem set-ups
# ---------------------------------------------------------------------------------------------------------
def build_problem(hgdt, torch, config):
    F = hgdt.fields
    if config == "c2":
        c = 128
        grid = hgdt.Grid.structured("tet", [0, 0, 0], [1, 1, 1], [c, c, c])
        kappa = torch.from_numpy(F.lognormal(1, grid.num_elements)).cuda()
        space = hgdt.ContinuousLagrangeSpace(grid, 1)
        op = hgdt.make_matrix_operator(space, hgdt.Stencil.element)
        op.append(hgdt.LocalElementIntegralBilinearForm(hgdt.LocalLaplaceIntegrand(kappa)))
        workload = "p1_kuhn_tets_128x128x128x6_lognormal_kappa_f1_dirichlet0"
    elif config == "c1":
        grid = hgdt.Grid.structured("quad", [0, 0], [1, 1], [128, 128])
        kappa = 1.0
        space = hgdt.ContinuousLagrangeSpace(grid, 1)
        op = hgdt.make_matrix_operator(space, hgdt.Stencil.element)
        op.append(hgdt.LocalElementIntegralBilinearForm(hgdt.LocalLaplaceIntegrand(kappa)))
        workload = "q1_quads_128x128_kappa1_f1_dirichlet0"
    elif config == "c3":
        n = 512
        grid = hgdt.Grid.structured("quad", [0, 0], [1, 1], [n, n])
        kappa = torch.from_numpy(F.channel_field(3, n, n)).cuda()
        space = hgdt.DiscontinuousLagrangeSpace(grid, 1)
        op = hgdt.make_matrix_operator(space, hgdt.Stencil.element_and_intersection)
        op += hgdt.LocalElementIntegralBilinearForm(hgdt.LocalLaplaceIntegrand(kappa))
        op.append(hgdt.LocalCouplingIntersectionIntegralBilinearForm(
            hgdt.LocalLaplaceIPDGInnerCoupling(1.0, kappa, kappa) + hgdt.LocalIPDGInnerPenalty(16.0, kappa)), {},
            hgdt.ApplyOn.InnerIntersectionsOnce())
        op.append(hgdt.LocalIntersectionIntegralBilinearForm(
            hgdt.LocalIPDGBoundaryPenalty(16.0, kappa) + hgdt.LocalLaplaceIPDGDirichletCoupling(1.0, kappa)), {},
            hgdt.ApplyOn.CustomBoundaryIntersections(hgdt.AllDirichletBoundaryInfo()))
        workload = "q1dg_swipdg_quads_512x512_channels_kappa_1e-4_sigma16_f1"
    else:
        raise SystemExit(f"unknown config {config}")
    rhs = hgdt.make_vector_functional(space)
    rhs.append(hgdt.LocalElementIntegralFunctional(hgdt.LocalProductIntegrand().with_ansatz(1.0)))
    op.append(rhs)
    if config != "c3":
        op.append(hgdt.make_dirichlet_constraints(space, hgdt.AllDirichletBoundaryInfo()))
    return grid, space, op, kappa, workload


def oracle_operator(O, config, kappa_np):
    if config == "c2":
        m = O.Mesh.structured(O.TET, [128, 128, 128], [0, 0, 0], [1, 1, 1])
        s = O.Space(m, True, 1)
        op = O.Operator(s, O.STENCIL_ELEMENT)
        op.append_laplace(kappa_np)
        op.append_source(1.0)
        op.set_dirichlet(True)
    elif config == "c1":
        m = O.Mesh.structured(O.QUAD, [128, 128], [0, 0], [1, 1])
        s = O.Space(m, True, 1)
        op = O.Operator(s, O.STENCIL_ELEMENT)
        op.append_laplace(1.0)
        op.append_source(1.0)
        op.set_dirichlet(True)
    else:
        m = O.Mesh.structured(O.QUAD, [512, 512], [0, 0], [1, 1])
        s = O.Space(m, False, 1)
        op = O.Operator(s, O.STENCIL_ELEMENT_AND_INTERSECTION)
        op.append_laplace(kappa_np)
        op.append_inner_coupling(1.0, kappa_np, kappa_np)
        op.append_inner_penalty(16.0, kappa_np)
        op.append_dirichlet_coupling(1.0, kappa_np)
        op.append_boundary_penalty(16.0, kappa_np)
        op.append_source(1.0)
    return m, s, op


def batch_problem(hgdt, grid, kappa, f):
    space = hgdt.ContinuousLagrangeSpace(grid, 1)
    op = hgdt.make_matrix_operator(space, hgdt.Stencil.element)
    op.append(hgdt.LocalElementIntegralBilinearForm(hgdt.LocalLaplaceIntegrand(kappa)))
    rhs = hgdt.make_vector_functional(space)
    rhs.append(hgdt.LocalElementIntegralFunctional(hgdt.LocalProductIntegrand().with_ansatz(f)))
    op.append(rhs)
    op.append(hgdt.make_dirichlet_constraints(space, hgdt.AllDirichletBoundaryInfo()))
    return space, op


def choose_launch(op, args):
    """Launch shapes of the solver kernels: the recorded table (tuning only for an unknown signature),
    a fresh tuning run, or the library defaults."""
    if args.tune == "off":
        return None
    from hgdt import autotune

    if args.tune == "table":
        entry = autotune.apply_table(op)
        if entry is not None:
            return entry
    choice = autotune.tune_solver(op)
    choice.pop("timings", None)
    if args.tune == "run" and args.tune_out:
        autotune.save_table(op, choice, args.tune_out)
    return choice


def kernel_roofline(op, kernel, rhs="rhs", max_iter=None):
    """live roofline of the solver's operator application: HIP events around two consecutive applications in
    every 8 iterations of a solve (max_iter bounds the solve to the first iterations of a batch)"""
    hgdt = sys.modules["hgdt"]
    try:
        op.apply_inverse(rhs, opts={"max_iter": max_iter} if max_iter else None, time_spmv=8)
    except hgdt.linear_solver_failed:
        if not max_iter:
            raise
    st = op.last_solve
    us = st["spmv_avg_ms"] * 1e3
    achieved = st["step_bytes"] / (us * 1e-6) / 1e9 if us > 0 else None
    return {"bound": "hbm", "kernel": kernel, "bytes_per_launch": int(st["step_bytes"]), "avg_launch_us": us,
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "launches_timed": int(st["spmv_count"])}


def spmv_kernel_name(op, config):
    fi = op.format_info()
    if fi["format"] == "matrix_free":
        return {"c3": "k_mf_dg<3, 1, 2, 2, true>"}.get(config, "k_mf_cg<2, 1, 1>")
    if fi["format"] == "block4":
        return "k_cg_spmv_blk<symmetric>"
    if fi["format"] == "dia_symmetric" and op.last_solve and op.last_solve.get("column_staged"):
        return "k_cg_spmv_win"
    return f"k_cg_spmv<{fi['format']}>"


def traffic_record(config, mf):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3 counter summary"""
    tfile = os.path.join(ROOT, "profiles", f"pmc_{'mf' if mf else 'cg_spmv'}_{config}.json")
    if not os.path.exists(tfile):
        return None, None
    with open(tfile) as f:
        d = json.load(f)
    src = f"profiles/{os.path.basename(tfile)} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; {d.get('date', 'undated')})"
    return d.get("hbm_bytes_per_launch"), src


# ---------------------------------------------------------------------------------------------------------
# CPU baseline and the FP64 errors of the headline step
# ---------------------------------------------------------------------------------------------------------
def cpu_leg(O, config, kappa_np, threads, sample_iters):
    m, s, op = oracle_operator(O, config, kappa_np)
    t0 = time.perf_counter()
    op.assemble(threads)
    t_asm = time.perf_counter() - t0
    rp, cols = op.pattern()
    vals, b = op.values(), op.rhs()
    t0 = time.perf_counter()
    x, it, rel, rc = O.pcg_jacobi(rp, cols, vals, b, rtol=1e-10, nthreads=threads, max_iters_run=sample_iters)
    t_pcg = time.perf_counter() - t0
    return {"t_assembly_s": t_asm, "t_pcg_s": t_pcg, "iterations": it, "relres": rel, "converged": rc == 0,
            "x": x, "op": op, "rp": rp, "cols": cols, "vals": vals, "b": b, "space": s}


def cpu_baseline(O, config, kappa_np, gpu_iterations, args):
    """The oracle on this host: one full step of the same workload with all threads (the baseline value),
    and one thread on a bounded sample (assembly + the first iterations of the PCG, projected to the full
    iteration count with the measured time per iteration)."""
    threads = host_threads()
    full = cpu_leg(O, config, kappa_np, threads, args.cpu_sample_iters)
    it_all = full["iterations"]
    t_it = full["t_pcg_s"] / max(it_all, 1)
    if args.cpu_sample_iters == 0:
        t_total = full["t_assembly_s"] + full["t_pcg_s"]
        sample = (f"one full step of the workload: oracle assembly ({full['t_assembly_s']:.2f} s) + Jacobi-PCG to "
                  f"1e-10 ({it_all} iterations, {full['t_pcg_s']:.2f} s, relres {full['relres']:.1e}), "
                  f"{threads} threads")
    else:
        t_total = full["t_assembly_s"] + t_it * gpu_iterations
        sample = (f"oracle assembly of the full workload ({full['t_assembly_s']:.2f} s, {threads} threads) + "
                  f"{it_all} PCG iterations ({t_it * 1e3:.2f} ms each) projected to {gpu_iterations} iterations")
    n = full["space"].size
    res = {"value": n / t_total, "unit": "DoFs/s", "cores": threads, "kind": "port", "sample": sample,
           "t_assembly_s": full["t_assembly_s"], "t_iteration_s": t_it, "iterations": it_all, **host_cpu()}
    if args.cpu_single_thread_iters > 0:
        one = cpu_leg(O, config, kappa_np, 1, args.cpu_single_thread_iters)
        t1 = one["t_pcg_s"] / max(one["iterations"], 1)
        total1 = one["t_assembly_s"] + t1 * gpu_iterations
        res["single_thread"] = {
            "value": n / total1, "unit": "DoFs/s", "cores": 1, "t_assembly_s": one["t_assembly_s"],
            "t_iteration_s": t1,
            "sample": (f"oracle assembly of the full workload ({one['t_assembly_s']:.2f} s) + {one['iterations']} PCG "
                       f"iterations ({t1 * 1e3:.2f} ms each) projected to {gpu_iterations} iterations, 1 thread")}
        res["speedup_all_threads_vs_one"] = res["value"] / res["single_thread"]["value"]
    return res, full


def step_errors(np, op, x_gpu, full, gpu_iters):
    """FP64 errors of the GPU step against the oracle's full step on the same inputs"""
    rp = full["rp"]
    vals = op.matrix.values.cpu().numpy()
    b = op.rhs.cpu().numpy()
    xo = full["x"]
    return {"matrix_max_rel": row_relative_error(vals, full["vals"], rp),
            "rhs_max_rel": float(np.abs(b - full["b"]).max() / np.abs(full["b"]).max()),
            "solution_rel_l2": float(np.linalg.norm(x_gpu - xo) / np.linalg.norm(xo)),
            "true_relres": float(op.last_solve["true_relres"]),
            "oracle_iterations": int(full["iterations"]), "iterations": int(gpu_iters),
            "bars": {"matrix": 1e-12, "solution": 1e-10}}


# ---------------------------------------------------------------------------------------------------------
# single workloads
# ---------------------------------------------------------------------------------------------------------
def run_single(hgdt, torch, args, config, fmt, steps, warmup, spmv_sample):
    import numpy as np

    t_setup = time.perf_counter()
    grid, space, op, kappa, workload = build_problem(hgdt, torch, config)
    op.set_format(fmt)
    op.assemble()
    op.apply_inverse("rhs")
    torch.cuda.synchronize()
    tuned = choose_launch(op, args)
    op.assemble()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    log(f"[{config}/{fmt}] setup {t_setup:.1f} s: {space.num_DoFs} DoFs, {grid.num_elements} elements, nnz {op.nnz}, "
        f"{op.format_info()}, launch {tuned}")

    def step(time_spmv):
        op.assemble()
        return op.apply_inverse("rhs", use_graph=args.graph and not time_spmv, time_spmv=time_spmv,
                                check_every=args.check_every)

    for w in range(warmup):
        step(False)
        log(f"[{config}] warmup {w + 1}/{warmup}: {op.last_solve['iterations']} iterations, "
            f"{op.last_solve['solve_ms']:.2f} ms solve")
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(steps):
        x = step(time_spmv=0)
        stats.append(dict(op.last_solve))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # the application's roofline comes from one more step with events bound to the sampled dispatches, outside
    # the timed region (the events' own cost stays out of ms_per_step)
    step(time_spmv=spmv_sample)
    timed = dict(op.last_solve)
    for k, st in enumerate(stats):
        log(f"[{config}] step {k + 1}: {st['iterations']} iterations, relres {st['relres']:.2e}, true relres "
            f"{st['true_relres']:.2e}, solve {st['solve_ms']:.2f} ms")
    log(f"[{config}] sampled step: {timed['spmv_count']} applications timed, {timed['spmv_avg_ms'] * 1e3:.2f} us")
    ms_per_step = elapsed / steps * 1e3
    n = space.num_DoFs
    fi = op.format_info()
    spmv_ms = float(timed["spmv_avg_ms"])
    step_bytes = int(timed["step_bytes"])
    achieved = step_bytes / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else None
    solve_ms = float(np.mean([s["solve_ms"] for s in stats]))
    mf = fi["format"] == "matrix_free"
    traffic, tsrc = traffic_record(config, mf)
    # the whole step at its algorithmic bytes: the assembly's plus the iterations' (application, direction
    # update, update), over the timed step
    asm_bytes = op.assembly_bytes()
    it_bytes = int(stats[-1]["iteration_bytes"])
    mean_iters = float(np.mean([s["iterations"] for s in stats]))
    total = asm_bytes + mean_iters * it_bytes
    step_rate = total / (ms_per_step * 1e-3) / 1e9
    res = {
        "value": n / (ms_per_step * 1e-3), "ms_per_step": ms_per_step,
        "config": {"workload": workload, "dofs": n, "elements": grid.num_elements, "nnz": op.nnz,
                   "iterations": stats[-1]["iterations"], "relres": stats[-1]["relres"],
                   "true_relres": stats[-1]["true_relres"], "format": fi["format"],
                   "t_cg_ms": solve_ms, "t_assembly_ms": ms_per_step - solve_ms, "setup_s": t_setup,
                   "launch": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "traffic_source": tsrc, "kernel": spmv_kernel_name(op, config),
                     "bytes_per_launch": step_bytes, "avg_launch_us": spmv_ms * 1e3,
                     "timing": f"one step after the timed ones, HIP events bound to the dispatches of two consecutive "
                               f"iterations (both parities of the deferred iterate update) in every {spmv_sample}: "
                               "from the end of "
                               "the kernel before the application (the direction update; for a matrix-free application "
                               "fused with it, the previous iteration's update) to the application's end"},
        "step_hbm": {"achieved": step_rate, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_rate / HBM_PEAK_GBS,
                     "bytes_per_step": total, "assembly_bytes": asm_bytes, "iteration_bytes": it_bytes,
                     "basis": "algorithmic bytes of the assembly and of the step's iterations (operator application, "
                              "direction update, update; the formats' compulsory bytes, as the roofline's) over "
                              "ms_per_step; the solve's set-up and exit passes (initial residual, the solver's scaled "
                              "copy, true residual) are not counted"},
    }
    return res, (grid, space, op, kappa, x)


def run_c3(hgdt, torch, args, O):
    """SWIPDG 512^2 in element-block storage and matrix-free: step time, roofline of each format's operator
    application, the assembled matrix against the oracle's (full size), the true residual of both solutions
    and their difference"""
    import numpy as np

    out = {}
    xs = {}
    for fmt in ("block4", "matrix_free"):
        res, (grid, space, op, kappa, x) = run_single(hgdt, torch, args, "c3", fmt, args.extra_steps, 1,
                                                      args.spmv_sample)
        xs[fmt] = x.cpu().numpy()
        if fmt == "block4":
            kappa_np = kappa.cpu().numpy()
            m, s, oop = oracle_operator(O, "c3", kappa_np)
            oop.assemble(host_threads())
            rp, _ = oop.pattern()
            res["errors"] = {"matrix_max_rel": row_relative_error(op.matrix.values.cpu().numpy(), oop.values(), rp),
                             "rhs_max_rel": float(np.abs(op.rhs.cpu().numpy() - oop.rhs()).max()
                                                  / np.abs(oop.rhs()).max())}
            del m, s, oop
        res.setdefault("errors", {})["true_relres"] = float(op.last_solve["true_relres"])
        if fmt == "matrix_free":
            # the same solve with restarts from the true residual (solver option, off in the timed steps: the
            # reference stops on the recursive residual), outside the timed region
            t0 = time.perf_counter()
            op.apply_inverse("rhs", opts={"restarts": 3})
            torch.cuda.synchronize()
            st = op.last_solve
            res["with_restarts"] = {"iterations": int(st["iterations"]), "restarts": int(st["restarts"]),
                                    "true_relres": float(st["true_relres"]), "solve_ms": float(st["solve_ms"]),
                                    "wall_ms": (time.perf_counter() - t0) * 1e3}
        out[fmt] = res
        del grid, space, op, kappa, x
        free_gpu(torch)
    d = float(np.linalg.norm(xs["block4"] - xs["matrix_free"]) / np.linalg.norm(xs["block4"]))
    out["matrix_free"]["errors"]["solution_rel_l2_vs_block4"] = d
    out["matrix_free_speedup_vs_block4"] = out["block4"]["ms_per_step"] / out["matrix_free"]["ms_per_step"]
    return out


# ---------------------------------------------------------------------------------------------------------
# batched (C4) and adjoint (C5)
# ---------------------------------------------------------------------------------------------------------
def tune_plain_buckets(ops_by_n, args):
    """Launch shapes of the plain per-instance operators: tuned (or taken from the launch table) on the
    instance nearest to each of n = 32, 64, 128, 256, 512 and applied to every instance of its bucket"""
    import numpy as np

    if args.tune == "off":
        return "library defaults"
    from hgdt import _lib as L

    ns = np.array(sorted(ops_by_n))
    reps = sorted({int(ns[np.abs(np.log(ns) - np.log(t)).argmin()]) for t in (32, 64, 128, 256, 512)})
    shapes = {}
    for n in reps:
        op = ops_by_n[n][0]
        choice = choose_launch(op, args)
        shapes[n] = (op.get_launch(L.KERNEL_APPLY), op.get_launch(L.KERNEL_VECTOR),
                     (choice or {}).get("staging", -1))
    for n, ops in ops_by_n.items():
        rep = reps[int(np.abs(np.log(np.array(reps)) - np.log(n)).argmin())]
        a, v, s = shapes[rep]
        for op in ops:
            op.set_launch(L.KERNEL_APPLY, *a)
            op.set_launch(L.KERNEL_VECTOR, *v)
            op.set_column_staging(s)
    return {str(n): {"apply": list(a), "vector": list(v), "staging": s} for n, (a, v, s) in shapes.items()}


def run_batched(hgdt, torch, args, O):
    """256 independent Q1 problems on n x n grids, n ~ U{32..512}: one assembly launch set and one
    graph-replayed PCG for the whole batch vs 256 sequential plain GPU solves"""
    import numpy as np

    F = hgdt.fields
    t_setup = time.perf_counter()
    sizes = F.batch_sizes(4, 256)
    parts = [hgdt.Grid.structured("quad", [0, 0], [1, 1], [int(n), int(n)]) for n in sizes]
    union = hgdt.Grid.union(parts)
    kappa = F.lognormal(41, union.num_elements)
    f = F.uniform(42, union.num_elements)
    space, op = batch_problem(hgdt, union, torch.from_numpy(kappa).cuda(), torch.from_numpy(f).cuda())
    op.set_format(args.format)
    op.assemble()
    op.apply_inverse("rhs", use_graph=True)
    tuned = choose_launch(op, args)
    eoff = union.instance_elements()
    plain, by_n = [], {}
    for i, n in enumerate(sizes):
        s1, op1 = batch_problem(hgdt, parts[i], torch.from_numpy(kappa[eoff[i]:eoff[i + 1]]).cuda(),
                                torch.from_numpy(f[eoff[i]:eoff[i + 1]]).cuda())
        op1.set_format(args.format)
        op1.assemble()
        op1.apply_inverse("rhs", use_graph=True)
        plain.append((s1, op1))
        by_n.setdefault(int(n), []).append(op1)
    plain_launch = tune_plain_buckets(by_n, args)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    log(f"[batched] set-up {t_setup:.1f} s: {len(sizes)} instances, {space.num_DoFs} DoFs, nnz {op.nnz}, "
        f"launch {tuned}")

    def step():
        op.assemble()
        return op.apply_inverse("rhs", use_graph=True, check_every=16)

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.extra_steps):
        xb = step()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / args.extra_steps
    st = dict(op.last_solve)
    its, rels = op.instance_stats()
    true_rel = np.empty(len(sizes))
    hgdt._lib.call("hgdt_operator_instance_true_residuals", op._h, true_rel.ctypes.data_as(hgdt._lib._pdbl))
    xb = xb.cpu().numpy()
    # plain: every instance assembled and solved on its own, one after the other
    for s1, op1 in plain:
        op1.assemble()
        op1.apply_inverse("rhs", use_graph=True, check_every=16)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xp = []
    for s1, op1 in plain:
        op1.assemble()
        xp.append(op1.apply_inverse("rhs", use_graph=True, check_every=16))
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    doff = space.instance_dofs()
    # plain solves use their own tuned launch shapes (other block partitions, so other summation orders of the
    # dot products): equal to the batched solutions to the solver's accuracy, not bit for bit
    vs_plain = max(float(np.linalg.norm(xp[i].cpu().numpy() - xb[doff[i]:doff[i + 1]]) / np.linalg.norm(xb[doff[i]:doff[i + 1]]))
                   for i in range(len(sizes)))
    # the smallest and the largest instance against the oracle's Jacobi-PCG
    vs_oracle = {}
    for i in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        n = int(sizes[i])
        m = O.Mesh.structured(O.QUAD, [n, n], [0, 0], [1, 1])
        so = O.Space(m, True, 1)
        oop = O.Operator(so, O.STENCIL_ELEMENT)
        oop.append_laplace(kappa[eoff[i]:eoff[i + 1]])
        oop.append_source(f[eoff[i]:eoff[i + 1]])
        oop.set_dirichlet(True)
        oop.assemble(host_threads())
        rp, cols = oop.pattern()
        xo, ito, relo, rc = O.pcg_jacobi(rp, cols, oop.values(), oop.rhs(), rtol=1e-10, nthreads=host_threads())
        vs_oracle[f"n{n}"] = {"solution_rel_l2": float(np.linalg.norm(xb[doff[i]:doff[i + 1]] - xo) / np.linalg.norm(xo)),
                              "iterations": int(its[i]), "oracle_iterations": int(ito)}
    roof = kernel_roofline(op, "batched operator application", max_iter=96)
    roof["kernel"] = spmv_kernel_name(op, "c4")
    n = space.num_DoFs
    return {"value": n / t_batch, "unit": "DoFs/s", "plain_dofs_per_s": n / t_plain,
            "speedup_vs_plain": t_plain / t_batch, "ms_per_batch": t_batch * 1e3, "ms_plain_sequential": t_plain * 1e3,
            "instances": len(sizes), "dofs": n, "nnz": op.nnz, "iterations_min": int(its.min()),
            "iterations_max": int(its.max()), "relres_max": float(rels.max()), "compactions": int(st["compactions"]),
            "errors": {"true_relres_max": float(true_rel.max()), "batched_vs_plain_max_rel_l2": vs_plain,
                       "vs_oracle": vs_oracle},
            "roofline": roof, "launch": tuned, "plain_launch": plain_launch, "setup_s": t_setup,
            "format": op.format_info()["format"], "steps": args.extra_steps,
            "workload": "256_q1_quads_n_U32-512_lognormal_kappa_uniform_f_dirichlet0"}


def run_adjoint(hgdt, torch, args, O):
    """32 P1 problems on 64^3 x 6 tets: J = 1/2 |u(log kappa) - u_obs|^2, forward CG + adjoint CG +
    adjoint element kernel through torch.autograd, batched vs the plain per-instance path; central
    differences (h = 1e-6) on 64 random elements, two per instance, each pair perturbed in one batched
    solve; the oracle's gradient of one instance"""
    import numpy as np
    from hgdt import _lib as L
    from hgdt.autograd import DiffusionProblem, solve_log_diffusion

    F = hgdt.fields
    ninst, c = 32, 64
    t_setup = time.perf_counter()
    part = hgdt.Grid.structured("tet", [0, 0, 0], [1, 1, 1], [c, c, c])
    union = hgdt.Grid.union([part] * ninst)
    E1 = part.num_elements
    logk_np = F.normal(51, E1 * ninst)
    logk = torch.from_numpy(logk_np).cuda()
    logk_obs = torch.from_numpy(F.normal(52, E1 * ninst)).cuda()
    prob = DiffusionProblem(union, source=1.0, rtol=1e-10)
    with torch.no_grad():
        u_obs = solve_log_diffusion(logk_obs, prob).clone()
    tuned = choose_launch(prob.op, args)  # the batched operator's solver launch shapes
    singles = [DiffusionProblem(part, source=1.0, rtol=1e-10) for _ in range(ninst)]
    singles[0].solve(torch.exp(logk[:E1]))
    single_tuned = choose_launch(singles[0].op, args)
    for sp in singles[1:]:
        sp.op.set_launch(L.KERNEL_APPLY, *singles[0].op.get_launch(L.KERNEL_APPLY))
        sp.op.set_launch(L.KERNEL_VECTOR, *singles[0].op.get_launch(L.KERNEL_VECTOR))
        sp.op.set_column_staging((single_tuned or {}).get("staging", -1))
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    n1 = singles[0].space.num_DoFs
    log(f"[adjoint] set-up {t_setup:.1f} s: {ninst} x {n1} DoFs")

    def batched():
        theta = logk.clone().requires_grad_(True)
        u = solve_log_diffusion(theta, prob)
        J = 0.5 * torch.sum((u - u_obs) ** 2)
        J.backward()
        return theta.grad

    def plain():
        grads = []
        for i, sp in enumerate(singles):
            theta = logk[i * E1:(i + 1) * E1].clone().requires_grad_(True)
            u = solve_log_diffusion(theta, sp)
            J = 0.5 * torch.sum((u - u_obs[i * n1:(i + 1) * n1]) ** 2)
            J.backward()
            grads.append(theta.grad)
        return torch.cat(grads)

    batched()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.extra_steps):
        gb = batched()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / args.extra_steps
    fwd, adj = dict(prob.last_forward), dict(prob.last_adjoint)
    plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gp = plain()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    agree = float((gb - gp).abs().max() / gp.abs().max())
    g = gb.cpu().numpy()
    # central differences: 64 elements, two per instance; one batched solve per sign perturbs one element of
    # every instance (instances are independent, so each instance's J changes only through its own element)
    rng = np.random.default_rng(53)
    picks = np.stack([rng.choice(E1, 2, replace=False) + i * E1 for i in range(ninst)])  # [ninst, 2]
    h = 1e-6
    prob.rtol = 1e-13
    fd = np.empty(picks.shape)
    doff = prob.space.instance_dofs()
    for k in range(2):
        Js = []
        for sgn in (1, -1):
            lk = logk_np.copy()
            lk[picks[:, k]] += sgn * h
            with torch.no_grad():
                up = solve_log_diffusion(torch.from_numpy(lk).cuda(), prob)
                d2 = ((up - u_obs) ** 2).cpu().numpy()
            Js.append(np.array([0.5 * d2[doff[i]:doff[i + 1]].sum() for i in range(ninst)]))
        fd[:, k] = (Js[0] - Js[1]) / (2 * h)
    # the gradient the checks compare with: forward and adjoint solved to 1e-12 (the timed batches use 1e-10)
    prob.rtol = 1e-12
    g = batched().cpu().numpy()
    prob.rtol = 1e-10
    gsel = g[picks]
    gmax_inst = np.array([np.abs(g[i * E1:(i + 1) * E1]).max() for i in range(ninst)])
    fd_err = float((np.abs(fd - gsel) / gmax_inst[:, None]).max())
    # the oracle's gradient of instance 0 (forward and adjoint PCG to 1e-12 on the host)
    m = O.Mesh.structured(O.TET, [c, c, c], [0, 0, 0], [1, 1, 1])
    so = O.Space(m, True, 1)
    kap0 = np.exp(logk_np[:E1])
    oop = O.Operator(so, O.STENCIL_ELEMENT)
    oop.append_laplace(kap0)
    oop.append_source(1.0)
    oop.set_dirichlet(True)
    oop.assemble(host_threads())
    rp, cols = oop.pattern()
    u0, *_ = O.pcg_jacobi(rp, cols, oop.values(), oop.rhs(), rtol=1e-12, nthreads=host_threads())
    lam0, *_ = O.pcg_jacobi(rp, cols, oop.values(), u0 - u_obs[:n1].cpu().numpy(), rtol=1e-12,
                            nthreads=host_threads())
    go = O.adjoint_log_kappa_gradient(so, kap0, u0, lam0)
    cpu_err = float(np.abs(g[:E1] - go).max() / np.abs(go).max())
    roof = kernel_roofline(prob.op, "adjoint batch operator application", rhs=prob.op.rhs, max_iter=96)
    roof["kernel"] = spmv_kernel_name(prob.op, "c5")
    n = ninst * n1
    return {"value": n / t_batch, "unit": "DoFs/s", "plain_dofs_per_s": n / t_plain,
            "speedup_vs_plain": t_plain / t_batch, "ms_per_gradient_batch": t_batch * 1e3,
            "ms_plain_sequential": t_plain * 1e3, "instances": ninst, "dofs": n,
            "forward_iterations": fwd["iterations"], "adjoint_iterations": adj["iterations"],
            "errors": {"batched_vs_plain_max_rel_diff": agree, "fd_elements": int(picks.size), "fd_h": h,
                       "fd_max_abs_err_over_max_grad": fd_err, "cpu_gradient_instance0_max_rel": cpu_err,
                       "forward_true_relres": fwd["true_relres"], "adjoint_true_relres": adj["true_relres"],
                       "check_rtol": 1e-12, "bars": {"cpu_gradient": 1e-8, "fd": 1e-5}},
            "roofline": roof, "setup_s": t_setup, "launch": tuned, "plain_launch": single_tuned,
            "steps": args.extra_steps, "workload": "32_p1_kuhn_tets_64x64x64x6_logkappa_N01_J_half_l2_misfit"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--no-extras", action="store_true", help="the headline workload only (no c3 / batched / adjoint)")
    ap.add_argument("--extra-steps", type=int, default=2, help="timed steps of the c3, batched and adjoint workloads")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-iters", type=int, default=None,
                    help="oracle PCG iterations timed for the all-threads CPU baseline (0: to convergence; default 0)")
    ap.add_argument("--cpu-single-thread-iters", type=int, default=60,
                    help="oracle PCG iterations timed with one thread (0: no single-thread leg)")
    ap.add_argument("--graph", action="store_true", help="replay PCG iteration blocks as a hipGraph")
    ap.add_argument("--tune", default="table", choices=["table", "run", "off"],
                    help="solver launch shapes: recorded table (tuning an unknown workload), fresh tuning run, "
                         "or library defaults")
    ap.add_argument("--tune-out", default=None, help="with --tune run: record the choice in this JSON table")
    ap.add_argument("--no-tune", action="store_true", help="same as --tune off")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="issue one device-to-device copy of this size first (reference dispatch for HBM counters)")
    ap.add_argument("--format", default="auto", choices=["auto", "matrix_free", "block4", "dia_symmetric", "dia", "sell64"],
                    help="operator storage of apply_inverse for --config runs")
    ap.add_argument("--check-every", type=int, default=16, help="PCG iterations between host-side convergence checks")
    ap.add_argument("--spmv-sample", type=int, default=8,
                    help="HIP events around the operator applications of two consecutive PCG iterations in every k, "
                         "in one step after the timed ones")
    args = ap.parse_args()
    if args.no_tune:
        args.tune = "off"
    if args.cpu_sample_iters is None:
        args.cpu_sample_iters = 400 if args.config == "c3" else 0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported by this build "
              f"(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import hgdt

    torch.cuda.set_device(0)
    t_start = time.perf_counter()
    if args.calibrate_copy_mib:
        nb = args.calibrate_copy_mib << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        hgdt._lib.call("hgdt_memcpy", ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), nb,
                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        del src, dst
    O = oracle() if not args.no_cpu_baseline or not args.no_extras else None
    common = {"metric": METRIC, "unit": "DoFs/s", "n_gpus": 1, "higher_is_better": True, "scaling": "strong",
              "vs_baseline": None, "dtype": "f64"}
    if args.config in ("c4", "c5"):
        O = O or oracle()
        res = run_batched(hgdt, torch, args, O) if args.config == "c4" else run_adjoint(hgdt, torch, args, O)
        line = dict(common, value=res["value"], steps=args.extra_steps, warmup=1,
                    ms_per_step=res.get("ms_per_batch", res.get("ms_per_gradient_batch")),
                    data="synthetic (seeded fields)", config=res, roofline=res["roofline"])
        print(json.dumps(line), flush=True)
        return
    res, (grid, space, op, kappa, x) = run_single(hgdt, torch, args, args.config, args.format, args.steps,
                                                  args.warmup, args.spmv_sample)
    line = dict(common, value=res["value"], steps=args.steps, warmup=args.warmup, ms_per_step=res["ms_per_step"],
                data="synthetic (seeded splitmix64 / Box-Muller fields, structured grid)", config=res["config"],
                roofline=res["roofline"], step_hbm=res["step_hbm"])
    iters = res["config"]["iterations"]
    x_np = x.cpu().numpy()
    kappa_np = kappa.cpu().numpy() if hasattr(kappa, "cpu") else kappa
    if not args.no_cpu_baseline:
        cb, full = cpu_baseline(O, args.config, kappa_np, iters, args)
        line["cpu_baseline"] = cb
        line["config"]["gpu_vs_cpu"] = line["value"] / cb["value"]
        if args.cpu_sample_iters == 0:
            line["config"]["errors"] = step_errors(np, op, x_np, full, iters)
        del full
    del grid, space, op, kappa, x
    free_gpu(torch)
    if not args.no_extras and args.config == "c2":
        t0 = time.perf_counter()
        line["c3"] = run_c3(hgdt, torch, args, O)
        free_gpu(torch)
        line["batched"] = run_batched(hgdt, torch, args, O)
        free_gpu(torch)
        line["adjoint"] = run_adjoint(hgdt, torch, args, O)
        line["extras_wall_s"] = time.perf_counter() - t0
    line["wall_s"] = time.perf_counter() - t_start
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
