"""Benchmark: global assembly + Jacobi-PCG (1e-10) on one MI355X, FP64.

Default workload (the configuration the metric is quoted on): P1 continuous Lagrange on the unit
cube split into 128^3 x 6 Kuhn tetrahedra (2 146 689 DoFs), kappa_e = exp(g_e), g ~ N(0,1) from seed,
f = 1, zero Dirichlet.  A step = MatrixOperator.assemble() (values + rhs + Dirichlet) followed by
apply_inverse from x0 = 0 to ||r|| <= 1e-10 ||b||; inputs are resident in HBM before timing.

Prints one JSON line.  Other workloads: --config c1 (Q1 128^2), c3 (Q1-DG SWIPDG 512^2).
"""
import argparse
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


def batch_problem(hgdt, grid, kappa, f):
    space = hgdt.ContinuousLagrangeSpace(grid, 1)
    op = hgdt.make_matrix_operator(space, hgdt.Stencil.element)
    op.append(hgdt.LocalElementIntegralBilinearForm(hgdt.LocalLaplaceIntegrand(kappa)))
    rhs = hgdt.make_vector_functional(space)
    rhs.append(hgdt.LocalElementIntegralFunctional(hgdt.LocalProductIntegrand().with_ansatz(f)))
    op.append(rhs)
    op.append(hgdt.make_dirichlet_constraints(space, hgdt.AllDirichletBoundaryInfo()))
    return space, op


def run_batched(hgdt, torch, args):
    """256 independent Q1 problems on n x n grids, n ~ U{32..512}: one assembly launch set and one
    graph-replayed PCG for the whole batch vs 256 sequential plain GPU solves."""
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
    plain = []
    for i, n in enumerate(sizes):
        s1, op1 = batch_problem(hgdt, parts[i], torch.from_numpy(kappa[eoff[i]:eoff[i + 1]]).cuda(),
                                torch.from_numpy(f[eoff[i]:eoff[i + 1]]).cuda())
        op1.set_format(args.format)
        op1.assemble()
        op1.apply_inverse("rhs", use_graph=True)
        plain.append((s1, op1))
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    log(f"batched set-up {t_setup:.1f} s: {len(sizes)} instances, {space.num_DoFs} DoFs, nnz {op.nnz}, "
        f"launch {tuned}")

    def step():
        op.assemble()
        return op.apply_inverse("rhs", use_graph=True, check_every=16)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / args.steps
    its, rels = op.instance_stats()
    # plain: every instance assembled and solved on its own, one after the other
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s1, op1 in plain:
        op1.assemble()
        op1.apply_inverse("rhs", use_graph=True, check_every=16)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    n = space.num_DoFs
    return {"batched_dofs_per_s": n / t_batch, "plain_dofs_per_s": n / t_plain, "speedup_vs_plain": t_plain / t_batch,
            "ms_per_batch": t_batch * 1e3, "ms_plain_sequential": t_plain * 1e3, "instances": len(sizes),
            "dofs": n, "nnz": op.nnz, "iterations_min": int(its.min()), "iterations_max": int(its.max()),
            "relres_max": float(rels.max()), "launch": tuned, "setup_s": t_setup, "format": op.format_info()["format"],
            "workload": "256_q1_quads_n_U32-512_lognormal_kappa_uniform_f_dirichlet0"}


def run_adjoint(hgdt, torch, args):
    """32 P1 problems on 64^3 x 6 tets: J = 1/2 |u(log kappa) - u_obs|^2, forward CG + adjoint CG +
    adjoint element kernel through torch.autograd, batched vs the plain per-instance path."""
    from hgdt.autograd import DiffusionProblem, solve_log_diffusion

    F = hgdt.fields
    ninst, c = 32, 64
    t_setup = time.perf_counter()
    part = hgdt.Grid.structured("tet", [0, 0, 0], [1, 1, 1], [c, c, c])
    union = hgdt.Grid.union([part] * ninst)
    E1 = part.num_elements
    logk = torch.from_numpy(F.normal(51, E1 * ninst)).cuda()
    logk_obs = torch.from_numpy(F.normal(52, E1 * ninst)).cuda()
    prob = DiffusionProblem(union, source=1.0, rtol=1e-10)
    with torch.no_grad():
        u_obs = solve_log_diffusion(logk_obs, prob).clone()
    tuned = choose_launch(prob.op, args)  # the batched operator's solver launch shapes
    singles = [DiffusionProblem(part, source=1.0, rtol=1e-10) for _ in range(ninst)]
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    n1 = singles[0].space.num_DoFs
    log(f"adjoint set-up {t_setup:.1f} s: {ninst} x {n1} DoFs")

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

    for _ in range(args.warmup):
        batched()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gb = batched()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / args.steps
    plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gp = plain()
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    agree = float((gb - gp).abs().max() / gp.abs().max())
    n = ninst * n1
    return {"batched_dofs_per_s": n / t_batch, "plain_dofs_per_s": n / t_plain, "speedup_vs_plain": t_plain / t_batch,
            "ms_per_gradient_batch": t_batch * 1e3, "ms_plain_sequential": t_plain * 1e3, "instances": ninst,
            "dofs": n, "forward_iterations": prob.last_forward["iterations"],
            "adjoint_iterations": prob.last_adjoint["iterations"], "batched_vs_plain_max_rel_diff": agree,
            "setup_s": t_setup, "launch": tuned, "workload": "32_p1_kuhn_tets_64x64x64x6_logkappa_N01_J_half_l2_misfit"}


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


def cpu_baseline(config, kappa_np, iterations, sample_iters):
    """The oracle (CPU restatement) on this host: assembly of the full workload + Jacobi-PCG, run to
    convergence (sample_iters 0) or for a bounded number of iterations whose time is projected to the
    GPU's iteration count (same recursion and stopping rule)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
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
    t0 = time.perf_counter()
    op.assemble(cores)
    t_asm = time.perf_counter() - t0
    rp, cols = op.pattern()
    vals, b = op.values(), op.rhs()
    t0 = time.perf_counter()
    x, it, rel, rc = O.pcg_jacobi(rp, cols, vals, b, rtol=1e-10, nthreads=cores, max_iters_run=sample_iters)
    t_pcg = time.perf_counter() - t0
    t_it = t_pcg / max(it, 1)
    if sample_iters == 0:
        t_total = t_asm + t_pcg
        sample = (f"one full step of the workload: oracle assembly ({t_asm:.2f} s) + Jacobi-PCG to 1e-10 "
                  f"({it} iterations, {t_pcg:.2f} s, relres {rel:.1e}), {cores} threads")
    else:
        t_total = t_asm + t_it * iterations
        sample = (f"oracle assembly of the full workload ({t_asm:.2f} s, {cores} threads) + {it} PCG "
                  f"iterations ({t_it * 1e3:.2f} ms each) projected to {iterations} iterations")
    return {"value": s.size / t_total, "unit": "DoFs/s", "cores": cores, "kind": "port", "sample": sample,
            "t_assembly_s": t_asm, "t_iteration_s": t_it, "iterations": it}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--extras", action="store_true",
                    help="add the batched (256 uneven Q1) and adjoint (32 x P1 64^3) comparisons to the line")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-iters", type=int, default=None,
                    help="oracle PCG iterations timed for the CPU baseline (0: to convergence; default 0, "
                         "400 for c3)")
    ap.add_argument("--graph", action="store_true", help="replay PCG iteration blocks as a hipGraph")
    ap.add_argument("--tune", default="table", choices=["table", "run", "off"],
                    help="solver launch shapes: recorded table (tuning an unknown workload), fresh tuning run, "
                         "or library defaults")
    ap.add_argument("--tune-out", default=None, help="with --tune run: record the choice in this JSON table")
    ap.add_argument("--no-tune", action="store_true", help="same as --tune off")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="issue one device-to-device copy of this size first (reference dispatch for HBM counters)")
    ap.add_argument("--format", default="auto", choices=["auto", "matrix_free", "block4", "dia_symmetric", "dia", "sell64"],
                    help="operator storage of apply_inverse (matrix_free: recomputed from the coefficients at every "
                         "application; the assembly then builds the right-hand side only)")
    ap.add_argument("--check-every", type=int, default=16,
                    help="PCG iterations between host-side convergence checks")
    ap.add_argument("--spmv-sample", type=int, default=32,
                    help="HIP events around the operator application of every k-th PCG iteration of the timed steps")
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
    if args.calibrate_copy_mib:
        import ctypes

        nb = args.calibrate_copy_mib << 20
        src = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        hgdt._lib.call("hgdt_memcpy", ctypes.c_void_p(dst.data_ptr()), ctypes.c_void_p(src.data_ptr()), nb,
                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        del src, dst
    if args.config in ("c4", "c5"):
        res = run_batched(hgdt, torch, args) if args.config == "c4" else run_adjoint(hgdt, torch, args)
        line = {"metric": METRIC, "value": res["batched_dofs_per_s"], "unit": "DoFs/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup,
                "ms_per_step": res.get("ms_per_batch", res.get("ms_per_gradient_batch")), "higher_is_better": True,
                "scaling": "strong", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded fields)",
                "config": res}
        print(json.dumps(line), flush=True)
        return
    t_setup = time.perf_counter()
    grid, space, op, kappa, workload = build_problem(hgdt, torch, args.config)
    op.set_format(args.format)
    op.assemble()
    x = op.apply_inverse("rhs")
    torch.cuda.synchronize()
    tuned = choose_launch(op, args)
    op.assemble()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    log(f"setup {t_setup:.1f} s: {space.num_DoFs} DoFs, {grid.num_elements} elements, nnz {op.nnz}, "
        f"{op.format_info()}, launch {tuned}")

    def step(time_spmv):
        op.assemble()
        return op.apply_inverse("rhs", use_graph=args.graph and not time_spmv, time_spmv=time_spmv,
                                check_every=args.check_every)

    for w in range(args.warmup):
        step(False)
        log(f"warmup {w + 1}/{args.warmup}: {op.last_solve['iterations']} iterations, "
            f"{op.last_solve['solve_ms']:.2f} ms solve")
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        x = step(time_spmv=args.spmv_sample)
        stats.append(dict(op.last_solve))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    for k, st in enumerate(stats):
        log(f"step {k + 1}: {st['iterations']} iterations, relres {st['relres']:.2e}, solve {st['solve_ms']:.2f} ms, "
            f"spmv {st['spmv_avg_ms'] * 1e3:.1f} us")
    ms_per_step = elapsed / args.steps * 1e3
    n = space.num_DoFs
    iters = stats[-1]["iterations"]
    fi = op.format_info()
    spmv_ms = float(np.mean([s["spmv_avg_ms"] for s in stats]))
    step_bytes = int(stats[-1]["step_bytes"])
    # --spmv-sample 0 (or graph replay) records no events: no live roofline figure then
    achieved = step_bytes / (spmv_ms * 1e-3) / 1e9 if spmv_ms > 0 else None
    solve_ms = float(np.mean([s["solve_ms"] for s in stats]))
    # the solver's operator application: the storage format's SpMV, or the matrix-free kernel
    mf = fi["format"] == "matrix_free"
    # (discontinuous: the application with the solver's direction update fused in, mode 3)
    kernel = ({"c3": "k_mf_dg<3, 1, 2, 2, true>"}.get(args.config, "k_mf_cg<2, 1, 1>") if mf
              else "k_cg_spmv_blk<symmetric>" if fi["format"] == "block4" else f"k_cg_spmv<{fi['format']}>")
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"pmc_{'mf' if mf else 'cg_spmv'}_{args.config}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    line = {
        "metric": METRIC,
        "value": n / (ms_per_step * 1e-3),
        "unit": "DoFs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded splitmix64 / Box-Muller fields, structured grid)",
        "config": {"workload": workload, "dofs": n, "elements": grid.num_elements, "nnz": op.nnz,
                   "iterations": iters, "relres": stats[-1]["relres"], "format": fi["format"],
                   "t_cg_ms": solve_ms, "t_assembly_ms": ms_per_step - solve_ms, "setup_s": t_setup,
                   "launch": tuned},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": traffic,
                     "kernel": kernel, "bytes_per_launch": step_bytes,
                     "avg_launch_us": spmv_ms * 1e3},
    }
    if args.extras:
        line["batched"] = run_batched(hgdt, torch, args)
        line["adjoint"] = run_adjoint(hgdt, torch, args)
    if not args.no_cpu_baseline:
        kappa_np = kappa.cpu().numpy() if hasattr(kappa, "cpu") else kappa
        cb = cpu_baseline(args.config, kappa_np, iters, args.cpu_sample_iters)
        line["cpu_baseline"] = cb
        line["config"]["gpu_vs_cpu"] = line["value"] / cb["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
