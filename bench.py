"""Benchmark of the cell-assembly hot path on one MI355X.

Headline workload: P1 variable-coefficient Poisson stiffness matrix + load vector on
UnitCubeMesh 128^3 (12,582,912 tets, 2,146,689 dofs, nnz 31,802,497), interior vertices
jittered +-0.2h, kappa ~ U[0.5,2], f ~ U[-1,1] (seeded). One step = one fused A+b assembly
pass (SystemAssembler semantics, values only: the CSR structure is built once before timing,
as in DOLFIN's re-assembly). Inputs are resident in HBM when timing starts.

Also reported (field "secondary"): vector-P2 elasticity 64^3 (matrix + body force), the
adjoint gradient of J = u^T K u - F^T u on P1 128^3 and vector P2 64^3 and its ratio to the
forward assembly,
and y = A x on the assembled P1 128^3 and vector-P2 64^3 matrices.

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 (vector and matrix)
METRIC = "cells assembled/s + % HBM roofline, 1 GPU (P1 128^3, P2 elast.); adjoint/forward time"
HERE = os.path.dirname(os.path.abspath(__file__))
SECTIONS = "matrix,adjoint,spmv,strategies,rcm,vp2,batch"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=128, help="P1 mesh size (cubes per side)")
    ap.add_argument("--p2-n", type=int, default=64, help="vector P2 mesh size (0 = skip)")
    ap.add_argument("--scatter", default="auto", choices=["auto", "atomic", "blocked", "colored"])
    ap.add_argument("--settle", type=float, default=0.3,
                    help="seconds of untimed steps after the warmup steps, before timing (clock ramp)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--sections", default=SECTIONS,
                    help="comma-separated secondary workloads to run (default: all of " + SECTIONS + ")")
    ap.add_argument("--batch", type=int, default=256,
                    help="members of the batched launch (0 = skip): n ~ U{8..48}, half P1, half P2")
    return ap.parse_args()


def timed_steps(fn, steps, warmup, settle_s=0.0):
    """Run warmup steps, then untimed steps for settle_s more seconds (the GPU raises its
    clocks over the first tens of milliseconds of sustained work: at 128^3 the first 50 A+b
    passes average 0.227 ms, passes after 40 ms of work 0.197 ms), then `steps` timed steps
    back to back between two HIP events on the launch stream (no event between steps, so
    none of their overhead is counted). Returns (wall seconds, GPU ms per step repeated per
    step, settle steps run); the host queues launches faster than the kernels run, so the GPU
    time per step is the kernels' time plus launch gaps."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    settle = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < settle_s:
        for _ in range(10):
            fn()
        settle += 10
        torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record(s)
    for _ in range(steps):
        fn()
    b.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, [a.elapsed_time(b) / steps] * steps, settle


def load_traffic(name):
    """HBM bytes per launch of a kernel and workload, as measured by the rocprofv3 counter
    runs recorded in profiles/traffic.json (not measured in this run), or None."""
    path = os.path.join(HERE, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(name)
    except Exception:
        return None


def traffic_source(name):
    try:
        src = json.load(open(os.path.join(HERE, "profiles", "traffic.json")))["_sources"][name]
    except Exception:
        return None
    return f"profiles/traffic.json[{name}]: rocprofv3 --pmc counters, {src}"


def bench_batch(nmembers, log, seed=5, settle_s=0.0):
    """Batched launch: nmembers UnitCubeMesh(n) instances, n ~ U{8..48}, half P1 kappa-Poisson
    and half vector-P2 elasticity with random coefficients, assembled into separate CSR
    matrices by one graph-captured launch of two kernels (vector-P2, P1). Autotuned: the cells
    per workgroup of each kernel (P1 rows per block, vector-P2 LDS values per block) on a
    sample of the members, then the launch mode (one workgroup per row block, or persistent
    workgroups per CU) on the whole batch. Setup is reported by phase."""
    import torch
    import dolfin_amd as dfa
    from dolfin_amd import autotune, workloads
    from dolfin_amd.batch import BatchAssembler
    t = time.perf_counter()
    specs = workloads.batch_specs(nmembers, seed)
    forms = workloads.batch_members(nmembers, seed, log=log)
    t_members = time.perf_counter() - t
    t = time.perf_counter()
    sample = workloads.batch_sample(specs)
    (p1_rows, p2_vals), plan_times = autotune.tune_batch_plans(sample, workloads.batch_form)
    t_tune_plans = time.perf_counter() - t
    log(f"batch: plan sizes tuned ({t_tune_plans:.1f} s): P1 {p1_rows} rows, P2 {p2_vals} values")
    t = time.perf_counter()
    # builds every member's structure and plan (8 members at a time), then the launch
    batch = BatchAssembler(forms, grid_per_cu=0, plan_defaults=(p1_rows, 0, p2_vals))
    torch.cuda.synchronize()
    t_plans = time.perf_counter() - t
    log(f"batch: structures and plans built ({t_plans:.1f} s)")
    t = time.perf_counter()
    best, gtimes = autotune.tune_batch_grid(
        lambda g: batch if g == 0 else BatchAssembler(forms, grid_per_cu=g), grids=(1, 2, 0), reps=3)
    if best != 0:
        batch = BatchAssembler(forms, grid_per_cu=best)
    t_tune_grid = time.perf_counter() - t
    alg = 0
    cells = sum(a.V.mesh().num_cells() for a in forms)
    for a, A in zip(forms, batch.tensors):
        e = dfa.engine_for(a.V)
        m = a.V.mesh()
        if a.kind == dfa._lib.FORM_P1_POISSON:
            alg += workloads.p1_algorithmic_bytes(m.num_cells(), 4, m.num_vertices(), 3, e.nnz, a.V.dim())
        else:
            alg += workloads.vp2_algorithmic_bytes(m.num_cells(), m.num_vertices(), e.nnz, a.V.dim())
    w, ev, _ = timed_steps(batch.assemble, 5, 2, settle_s=settle_s)
    ms = float(np.mean(ev))
    return {"workload": f"{nmembers} UnitCubeMesh instances n~U{{8..48}} (seed {seed}), half P1 "
                        f"kappa-Poisson, half vector-P2 elasticity, {cells} cells in total, one "
                        f"hipGraph-captured launch (vector-P2 and P1 kernels)",
            "ms": ms, "cells_per_s": cells / (ms * 1e-3),
            "hbm_achieved_GBs": alg / (ms * 1e-3) / 1e9, "hbm_frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "algorithmic_bytes": alg, "info": batch.info, "grid_per_cu": best,
            "grid_sweep_ms": {str(k): v for k, v in gtimes.items()},
            "plan_sizes": {"p1_rows_per_block": p1_rows, "vp2_values_per_block": p2_vals,
                           "sample_members": len(sample),
                           "sweep_ms": {k: {str(c): v for c, v in d.items()}
                                        for k, d in plan_times.items()}},
            "setup_s": {"members": t_members, "tune_plan_sizes": t_tune_plans,
                        "structures_and_plans": t_plans, "tune_launch_mode": t_tune_grid}}


def device_info(dev):
    """Name, CUs, clocks, memory bus and sizes from hipDeviceGetAttribute (through the
    library), beside the HBM peak the roofline uses: MI355X's published 8 TB/s, i.e. the
    8192-bit bus at 8 Gb/s per pin."""
    import ctypes
    from dolfin_amd import _lib
    v = (ctypes.c_int64 * 6)()
    _lib.check(_lib.load().dfa_device_info(0, ctypes.cast(v, ctypes.c_void_p)))
    return {"name": dev.name, "arch": getattr(dev, "gcnArchName", ""), "compute_units": v[0],
            "engine_clock_mhz": v[1] / 1e3, "memory_clock_mhz": v[2] / 1e3,
            "memory_bus_bits": v[3], "l2_bytes": v[4], "memory_bytes": v[5],
            "hbm_peak_used_GBs": HBM_PEAK_GBS}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the assembly path is "
              "single-GPU in this build.", file=sys.stderr)
        sys.exit(2)
    import torch
    sys.path.insert(0, HERE)
    import dolfin_amd as dfa
    from dolfin_amd import adjoint, workloads

    scatter = {"auto": dfa.SCATTER_AUTO, "atomic": dfa.SCATTER_ATOMIC,
               "blocked": dfa.SCATTER_BLOCKED, "colored": dfa.SCATTER_COLORED}[args.scatter]
    dev = torch.cuda.get_device_properties(0)

    # ---- headline: P1 128^3 fused A + b --------------------------------------------------
    t = time.perf_counter()
    mesh, V, kappa, f, u = workloads.p1_poisson_3d(args.n, with_u=True)
    K = workloads.to_function(V, kappa)
    F = workloads.to_function(V, f)
    a, L = dfa.poisson_form(V, K), dfa.load_form(V, F)
    asm = dfa.SystemAssembler(a, L, scatter=scatter)
    A, b = dfa.Matrix(), dfa.Vector()
    t_inputs = time.perf_counter() - t
    torch.cuda.synchronize()
    t = time.perf_counter()
    log(f"inputs ready ({t_inputs:.1f} s); first assembly")
    asm.assemble(A, b)  # first assembly: structure + plan + values
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t
    log(f"first assembly {t_first:.1f} s; timing {args.steps} steps")
    eng = dfa.engine_for(V)
    info = eng.plan_info(scatter)
    _, ev_cold, _ = timed_steps(lambda: asm.assemble(A, b), args.steps, args.warmup)
    wall, ev, n_settle = timed_steps(lambda: asm.assemble(A, b), args.steps, args.warmup,
                                     settle_s=args.settle)
    nc = mesh.num_cells()
    ms = 1e3 * wall / args.steps
    ev_ms = float(np.mean(ev))
    alg_bytes = workloads.p1_algorithmic_bytes(nc, 4, mesh.num_vertices(), 3, eng.nnz, V.dim(),
                                               n_coeff=1, with_vector=True)
    achieved = alg_bytes / (ev_ms * 1e-3) / 1e9
    used_scatter = ("colored" if scatter == dfa.SCATTER_COLORED else
                    "blocked" if info["num_blocks"] > 0 else "atomic")
    traffic = load_traffic(f"p1_{used_scatter}_{args.n}")
    kernel = eng.pass_kernel(dfa._lib.FORM_P1_POISSON, dfa._lib.FORM_P1_LOAD, scatter=scatter)
    result = {
        "metric": METRIC,
        "value": nc / (ms * 1e-3),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded UnitCubeMesh, jitter and P1 coefficients)",
        "config": {"workload": f"P1 kappa-Poisson stiffness + load vector, UnitCubeMesh "
                               f"{args.n}^3 tets ({nc} cells, {V.dim()} dofs, nnz {eng.nnz}), "
                               f"fused A+b re-assembly, {used_scatter} scatter",
                   "cells": nc, "dofs": V.dim(), "nnz": eng.nnz, "scatter": used_scatter,
                   "kernel": kernel,
                   "plan": info, "device": device_info(dev)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_source(f"p1_{used_scatter}_{args.n}"),
                     "algorithmic_bytes": alg_bytes, "kernel_ms": ev_ms},
        "first_assembly_s": t_first,
        "settle": {"seconds": args.settle, "steps": n_settle,
                   "cold_kernel_ms": float(np.mean(ev_cold)),
                   "note": "untimed steps after the warmup steps while the GPU raises its clocks; "
                           "cold_kernel_ms is the same K steps timed right after the warmup steps"},
    }

    # ---- CPU baseline: the oracle's SystemAssembler cell loop on the host ----------------------
    if not args.no_cpu_baseline:
        import oracle as O
        log("cpu baseline: building the oracle with -O3 -march=native for this host")
        cflags = O.use_native_build()
        rp, col = A.structure()
        rp, col = rp.cpu().numpy(), col.cpu().numpy()
        x, cells, cd = mesh.coordinates(), mesh.cells(), V.cell_dofs()
        mask = np.zeros(V.dim(), np.uint8)
        zero = np.zeros(V.dim())
        nthreads = max(1, min(16, os.cpu_count() or 1))

        def cpu_pass(threads):
            t0 = time.perf_counter()
            O.assemble_system(O.P1_POISSON, O.P1_LOAD, x, cells, cd, rp, col, mask, zero,
                              a_c0=kappa, L_c0=f, threads=threads)
            return time.perf_counter() - t0
        log(f"cpu baseline: structure + values, 1 thread")
        t0 = time.perf_counter()
        O.sparsity(cd, V.dim())
        cpu_struct = time.perf_counter() - t0
        t1 = [cpu_pass(0) for _ in range(5)]
        log(f"cpu baseline: values, {nthreads} threads")
        cpu_pass(nthreads)  # warm-up
        tn = [cpu_pass(nthreads) for _ in range(5)]
        s1, sn = float(np.median(t1)), float(np.median(tn))
        result["cpu_baseline"] = {
            "value": nc / sn, "unit": "cells/s", "cores": nthreads, "kind": "port",
            "sample": f"oracle SystemAssembler cell loop (Eigen coeffRef semantics), the full "
                      f"{args.n}^3 workload (A+b values, structure prebuilt), {nthreads} host "
                      f"threads over contiguous cell ranges with atomic scatter, median of 5, "
                      f"g++ {cflags}"}
        result["cpu_baseline_1thread"] = {
            "value": nc / s1, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": f"the same loop serial (DOLFIN's per-process semantics), median of 5, "
                      f"g++ {cflags}",
            "first_assembly_s": cpu_struct + t1[0], "structure_s": cpu_struct}
        result["cpu_baseline_host"] = {"nproc": os.cpu_count(), "cpu_model": O.cpu_model(),
                                       "compiler_flags": cflags}

    # ---- secondary: adjoint on P1 and vector P2 elasticity -----------------------------------
    sections = set(args.sections.split(","))
    unknown = sections - set(SECTIONS.split(","))
    if unknown:
        print(f"unknown --sections {sorted(unknown)}; choose from {SECTIONS}", file=sys.stderr)
        sys.exit(2)

    def want(name):
        return name in sections

    if not args.no_secondary:
        sec = {}
        kk = torch.from_numpy(kappa).cuda()
        ff = torch.from_numpy(f).cuda()
        uu = torch.from_numpy(u).cuda()
        xx = mesh.coordinates_device()

        if want("matrix"):
            # the bilinear form alone (Assembler::assemble(A, a)): the P1 Poisson assembly that
            # the >= 50 % HBM target names, 660,237,364 algorithmic bytes
            Am = dfa.Matrix()
            masm = dfa.Assembler(scatter)
            masm.assemble(Am, a)
            w_m, ev_m, _ = timed_steps(lambda: masm.assemble(Am, a), args.steps, args.warmup,
                                       settle_s=args.settle)
            m_ms = float(np.mean(ev_m))
            m_bytes = workloads.p1_algorithmic_bytes(nc, 4, mesh.num_vertices(), 3, eng.nnz, V.dim())
            sec["p1_matrix"] = {"ms": m_ms, "cells_per_s": nc / (m_ms * 1e-3),
                                "hbm_achieved_GBs": m_bytes / (m_ms * 1e-3) / 1e9,
                                "hbm_frac": m_bytes / (m_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                "algorithmic_bytes": m_bytes,
                                "kernel": eng.pass_kernel(dfa._lib.FORM_P1_POISSON, -1, scatter=scatter),
                                "workload": f"P1 kappa-Poisson stiffness matrix alone, {args.n}^3"}
            del Am
            log("matrix-only pass timed")
            # SystemAssembler with a Dirichlet condition on the whole boundary (apply_bc per cell,
            # SystemAssembler.cpp:1036-1112): the same forms, the same bytes plus the mask
            bc = dfa.DirichletBC(V, 0.5)
            asm_bc = dfa.SystemAssembler(a, L, bc, scatter=scatter)
            Abc, bbc = dfa.Matrix(), dfa.Vector()
            asm_bc.assemble(Abc, bbc)
            w_bc, ev_bc, _ = timed_steps(lambda: asm_bc.assemble(Abc, bbc), args.steps, args.warmup,
                                         settle_s=args.settle)
            bc_ms = float(np.mean(ev_bc))
            sec["p1_system_bc"] = {"ms": bc_ms, "cells_per_s": nc / (bc_ms * 1e-3),
                                   "hbm_frac": alg_bytes / (bc_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                                   "kernel": eng.pass_kernel(dfa._lib.FORM_P1_POISSON, dfa._lib.FORM_P1_LOAD,
                                                             has_bc=True, scatter=scatter),
                                   "workload": f"P1 kappa-Poisson + load with a Dirichlet condition on the "
                                               f"boundary ({len(bc.dofs)} dofs), {args.n}^3"}
            del Abc, bbc
            log("Dirichlet system timed")
        if want("adjoint"):
            def adj():
                adjoint._run(V, dfa._lib.FORM_P1_POISSON, xx, kk, None, ff, None, uu)
            w_adj, ev_adj, _ = timed_steps(adj, max(10, args.steps // 2), 3, settle_s=args.settle)
            sec["p1_adjoint"] = {"ms": float(np.mean(ev_adj)),
                                 "ratio_to_forward": float(np.mean(ev_adj)) / ev_ms,
                                 "workload": f"dJ/dkappa, dJ/dx of u^T K u - F^T u, P1 {args.n}^3"}
            log("adjoint timed")
        if want("spmv"):
            xs = torch.rand(V.dim(), dtype=torch.float64, device="cuda")
            ys = torch.empty_like(xs)
            w_mv, ev_mv, _ = timed_steps(lambda: A.mult(xs, ys), max(20, args.steps), 5, settle_s=args.settle)
            mv_ms = float(np.mean(ev_mv))
            mv_bytes = 12 * eng.nnz + 4 * (V.dim() + 1) + 16 * V.dim()
            sec["p1_spmv"] = {"ms": mv_ms, "hbm_achieved_GBs": mv_bytes / (mv_ms * 1e-3) / 1e9,
                              "hbm_frac": mv_bytes / (mv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                              "algorithmic_bytes": mv_bytes,
                              "kernel": "spmv_stream_kernel, 16-bit column offsets per row block",
                              "format_bytes": 10 * eng.nnz + 4 * (V.dim() + 1) + 16 * V.dim(),
                              "traffic": load_traffic(f"p1_spmv_{args.n}"),
                              "workload": f"y = A x on the assembled P1 {args.n}^3 matrix"}
            # Jacobi-preconditioned CG on the same matrix (dfa_cg_solve: the SpMV and three fused
            # vector kernels per iteration): 50 iterations per solve, tolerance 0 so that all run
            bcg = torch.rand(V.dim(), dtype=torch.float64, device="cuda")
            xcg = torch.zeros_like(bcg)
            cg_err = []

            def cg50():
                xcg.zero_()
                try:
                    dfa.solve(A, xcg, bcg, "cg", "jacobi", rtol=0.0, atol=0.0, maxiter=50, check_every=50)
                except RuntimeError as e:
                    if "did not converge" not in str(e):
                        cg_err.append(str(e))
            w_cg, ev_cg, _ = timed_steps(cg50, 4, 1, settle_s=0.0)
            cg_it = float(np.mean(ev_cg)) / 50
            cg_bytes = mv_bytes + 8 * 13 * V.dim()  # + p.Ap (2n), x, r update (7n), p update (4n)
            sec["p1_cg"] = {"ms_per_iteration": cg_it, "spmv_ms": mv_ms,
                            "hbm_frac": cg_bytes / (cg_it * 1e-3) / 1e9 / HBM_PEAK_GBS,
                            "algorithmic_bytes_per_iteration": cg_bytes,
                            "errors": cg_err,
                            "workload": f"Jacobi-preconditioned CG iterations on the P1 {args.n}^3 matrix "
                                        "(50 per solve, 4 solves)"}
            log("spmv timed")
        if want("strategies"):
            # the other two scatter strategies on the same A+b workload (the choice between
            # owner-computes, graph-coloured read-add-write and global atomics)
            strat = {}
            for name, sc in (("atomic", dfa.SCATTER_ATOMIC), ("colored", dfa.SCATTER_COLORED)):
                if sc == scatter:
                    continue
                asm_s = dfa.SystemAssembler(a, L, scatter=sc)
                t = time.perf_counter()
                asm_s.assemble(A, b)
                torch.cuda.synchronize()
                t_s = time.perf_counter() - t
                w_s, ev_s, _ = timed_steps(lambda: asm_s.assemble(A, b), 10, 2, settle_s=args.settle)
                strat[name] = {"ms": float(np.mean(ev_s)), "first_call_s": t_s,
                               "hbm_frac": alg_bytes / (float(np.mean(ev_s)) * 1e-3) / 1e9 / HBM_PEAK_GBS}
                if sc == dfa.SCATTER_COLORED:
                    strat[name]["colors"] = len(eng.color_ptr()) - 1
                    strat[name]["traffic"] = load_traffic(f"p1_colored_{args.n}")
                else:
                    strat[name]["traffic"] = load_traffic(f"p1_atomic_{args.n}")
                log(f"{name} scatter timed")
            asm.assemble(A, b)
            from dolfin_amd import autotune
            pick, ttimes = autotune.tune_scatter([a], reps=3)
            names = {dfa.SCATTER_BLOCKED: "blocked", dfa.SCATTER_COLORED: "colored",
                     dfa.SCATTER_ATOMIC: "atomic"}
            sec["p1_scatter_strategies"] = dict(
                strat, workload=f"P1 {args.n}^3 A+b, same inputs as the headline, other scatters",
                autotuned={"choice": names[pick],
                           "matrix_ms": {names[k]: v for k, v in ttimes.items()}})
        if want("rcm"):
            # the same workload with DOLFIN's default numbering (reverse Cuthill-McKee nodes)
            t = time.perf_counter()
            Vr = dfa.FunctionSpace(mesh, "P", 1, reorder=True)
            t_rcm = time.perf_counter() - t
            new_of = np.empty(V.dim(), np.int64)  # new dof of each vertex
            new_of[mesh.cells().ravel()] = Vr.cell_dofs().ravel()
            kr = np.empty_like(kappa)
            fr = np.empty_like(f)
            kr[new_of] = kappa
            fr[new_of] = f
            asm_r = dfa.SystemAssembler(dfa.poisson_form(Vr, workloads.to_function(Vr, kr)),
                                        dfa.load_form(Vr, workloads.to_function(Vr, fr)), scatter=scatter)
            Ar, br = dfa.Matrix(), dfa.Vector()
            t = time.perf_counter()
            asm_r.assemble(Ar, br)
            torch.cuda.synchronize()
            t_rfirst = time.perf_counter() - t
            w_r, ev_r, _ = timed_steps(lambda: asm_r.assemble(Ar, br), max(20, args.steps), 5, settle_s=args.settle)
            sec["p1_rcm"] = {"ms": float(np.mean(ev_r)), "first_assembly_s": t_rfirst,
                             "reordering_s": t_rcm,
                             "workload": f"P1 {args.n}^3 A+b with reverse Cuthill-McKee dof numbering "
                                         f"(reorder_dofs_serial)"}
            del Ar, br, asm_r
            log("rcm-numbered assembly timed")
        if args.p2_n > 0 and want("vp2"):
            log(f"vector P2 {args.p2_n}^3 setup")
            del A, b
            torch.cuda.empty_cache()
            m2, V2, mu, lmbda, f3, u2 = workloads.vp2_elasticity_3d(args.p2_n, with_u=True)
            Q = V2.coefficient_space()
            a2 = dfa.elasticity_form(V2, workloads.to_function(Q, mu), workloads.to_function(Q, lmbda))
            L2 = dfa.body_force_form(V2, f3)
            asm2 = dfa.SystemAssembler(a2, L2)
            A2, b2 = dfa.Matrix(), dfa.Vector()
            t = time.perf_counter()
            asm2.assemble(A2, b2)
            torch.cuda.synchronize()
            t2_first = time.perf_counter() - t
            log(f"vector P2 first assembly {t2_first:.1f} s")
            e2 = dfa.engine_for(V2)
            info2 = e2.plan_info()
            sc2 = "blocked" if info2["num_blocks"] > 0 else "atomic"
            w2, ev2, _ = timed_steps(lambda: asm2.assemble(A2, b2), max(5, args.steps // 5), 2, settle_s=args.settle)
            nc2 = m2.num_cells()
            ms2 = float(np.mean(ev2))
            alg2 = workloads.vp2_algorithmic_bytes(nc2, m2.num_vertices(), e2.nnz, V2.dim())
            sec["vp2_elasticity"] = {
                "workload": f"vector P2 elasticity + body force, UnitCubeMesh {args.p2_n}^3 "
                            f"({nc2} cells, {V2.dim()} dofs, nnz {e2.nnz}), fused, {sc2} scatter",
                "plan": info2,
                "ms": ms2, "cells_per_s": nc2 / (ms2 * 1e-3),
                "hbm_achieved_GBs": alg2 / (ms2 * 1e-3) / 1e9,
                "hbm_frac": alg2 / (ms2 * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "fp64_TFLOPs_at_39900_per_cell": nc2 * 39900 / (ms2 * 1e-3) / 1e12,
                "algorithmic_bytes": alg2, "traffic": load_traffic(f"vp2_{sc2}_{args.p2_n}"),
                "traffic_source": traffic_source(f"vp2_{sc2}_{args.p2_n}"),
                "first_assembly_s": t2_first}
            # SystemAssembler with a Dirichlet condition on the whole boundary: the blocked pass
            # without it, then the node-graph passes that apply it to the assembled rows
            bc2 = dfa.DirichletBC(V2, 0.0)
            asm2_bc = dfa.SystemAssembler(a2, L2, bc2)
            asm2_bc.assemble(A2, b2)
            w_bc2, ev_bc2, _ = timed_steps(lambda: asm2_bc.assemble(A2, b2), max(5, args.steps // 5), 2,
                                           settle_s=args.settle)
            bc2_ms = float(np.mean(ev_bc2))
            sec["vp2_system_bc"] = {"ms": bc2_ms, "cells_per_s": nc2 / (bc2_ms * 1e-3),
                                    "ratio_to_no_bc": bc2_ms / ms2,
                                    "kernel": e2.pass_kernel(dfa._lib.FORM_VP2_ELASTICITY,
                                                             dfa._lib.FORM_VP2_BODY_FORCE, has_bc=True),
                                    "workload": f"vector P2 elasticity + body force with a Dirichlet condition "
                                                f"on the boundary ({len(bc2.dofs)} dofs), {args.p2_n}^3"}
            del asm2_bc
            log("vector P2 Dirichlet system timed")
            x2 = m2.coordinates_device()
            mu_t, la_t, u_t = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (mu, lmbda, u2))
            f3a = np.ascontiguousarray(f3, dtype=np.float64)

            def adj2():
                adjoint._run(V2, dfa._lib.FORM_VP2_ELASTICITY, x2, mu_t, la_t, None, f3a, u_t)
            w_a2, ev_a2, _ = timed_steps(adj2, max(5, args.steps // 5), 2, settle_s=args.settle)
            sec["vp2_adjoint"] = {"ms": float(np.mean(ev_a2)),
                                  "ratio_to_forward": float(np.mean(ev_a2)) / ms2,
                                  "workload": f"dJ/dmu, dJ/dlmbda, dJ/dx of u^T K u - F^T u, "
                                              f"vector P2 {args.p2_n}^3"}
            xs2 = torch.rand(V2.dim(), dtype=torch.float64, device="cuda")
            ys2 = torch.empty_like(xs2)
            w_mv2, ev_mv2, _ = timed_steps(lambda: A2.mult(xs2, ys2), max(10, args.steps // 5), 2, settle_s=args.settle)
            mv2_ms = float(np.mean(ev_mv2))
            mv2_bytes = 12 * e2.nnz + 4 * (V2.dim() + 1) + 16 * V2.dim()
            sec["vp2_spmv"] = {"ms": mv2_ms,
                               "hbm_achieved_GBs": mv2_bytes / (mv2_ms * 1e-3) / 1e9,
                               "hbm_frac": mv2_bytes / (mv2_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                               "algorithmic_bytes": mv2_bytes,
                               "kernel": "spmv_node_kernel, one node column per 3 x 3 block",
                               "format_bytes": 8 * e2.nnz + 4 * (e2.nnz // 9) + 4 * (V2.dim() // 3 + 1)
                                               + 16 * V2.dim(),
                               "traffic": load_traffic(f"vp2_spmv_{args.p2_n}"),
                               "workload": f"y = A x on the assembled vector-P2 {args.p2_n}^3 matrix"}
            # Jacobi-preconditioned CG on the elasticity matrix (pure Neumann, so the iterations
            # only measure speed): 20 iterations per solve, tolerance 0 so that all run
            bcg2 = torch.rand(V2.dim(), dtype=torch.float64, device="cuda")
            xcg2 = torch.zeros_like(bcg2)
            cg2_err = []

            def cg20():
                xcg2.zero_()
                try:
                    dfa.solve(A2, xcg2, bcg2, "cg", "jacobi", rtol=0.0, atol=0.0, maxiter=20, check_every=20)
                except RuntimeError as e:
                    if "did not converge" not in str(e):
                        cg2_err.append(str(e))
            w_cg2, ev_cg2, _ = timed_steps(cg20, 3, 1, settle_s=0.0)
            cg2_it = float(np.mean(ev_cg2)) / 20
            cg2_bytes = mv2_bytes + 8 * 13 * V2.dim()
            sec["vp2_cg"] = {"ms_per_iteration": cg2_it, "spmv_ms": mv2_ms,
                             "hbm_frac": cg2_bytes / (cg2_it * 1e-3) / 1e9 / HBM_PEAK_GBS,
                             "algorithmic_bytes_per_iteration": cg2_bytes, "errors": cg2_err,
                             "workload": f"Jacobi-preconditioned CG iterations on the vector-P2 {args.p2_n}^3 "
                                         "matrix (20 per solve, 3 solves)"}
        if args.batch > 0 and want("batch"):
            sec["batch"] = bench_batch(args.batch, log, settle_s=args.settle)
        result["secondary"] = sec
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
