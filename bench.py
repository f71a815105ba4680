#!/usr/bin/env python3
"""Benchmark of the MI355X residual-distribution cell loop.

One step = one explicit pseudo-time iteration of PGRDMethodExplicit over the
whole mesh on the GPU: the fused cell kernel (gather, metrics, P+1 flux,
K+, scheme distribution, scatter), the strong Dirichlet values and the
nodal pass (residual / V, dt = CFL V / sum lambda, u -= dt res). The default
workload is the 2D Euler channel over a bump with P2 triangles and Q2
quadrilaterals (786 432 elements, 2 100 225 nodes, LDA scheme;
pdekit_amd.workloads.c3p2); --workload c2 gives 2D Burgers on 1 048 576
jittered P2 triangles with the B scheme. Inputs are resident in HBM before
the timed region.

Prints one JSON line: throughput in element residual evaluations per second,
the roofline of the cell kernel (HIP events around every cell-loop launch in
the timed region) and the CPU oracle timed on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, GB/s
SCATTER_NAMES = {0: "atomic", 1: "colour", 2: "patch"}
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 vector, TFLOP/s


def algorithmic_bytes(wl, scatter: int):
    """Compulsory HBM bytes of the dominant kernel and of one whole iteration.

    Every iteration must read the solution connectivity (u32 per cell node),
    the vertex coordinates, u and the dual volumes, and write u. With the
    patch assembly (scatter 2) the cell loop and the nodal step are one
    kernel, so both figures are the iteration's. With a global scatter the
    cell loop alone reads connectivity, coordinates and u and writes the
    residual and wave-speed accumulators."""
    m = wl.mesh
    neq = wl.state.shape[1]
    conn = sum(4 * b.sol_conn.size for b in m.blocks)
    coords = 8 * m.coords.size
    N = m.nb_nodes
    sweep = conn + coords + 8 * neq * N + 8 * N + 8 * neq * N
    if scatter == 2:
        return sweep, sweep
    cell = conn + coords + 8 * neq * N + 8 * neq * N + 8 * N
    return cell, sweep


def linear_bytes(wl):
    """Compulsory HBM bytes of one linear sweep of the mixed-precision form:
    the FP32 element Jacobians ((nd m)^2 per cell), the connectivity (u32 per
    cell node), and per node u_ref, R(u_ref) and u read, dt(u_ref) and V read,
    u written."""
    m = wl.mesh
    neq = wl.state.shape[1]
    jac = sum(4 * b.sol_conn.shape[0] * (b.sol_conn.shape[1] * neq) ** 2 for b in m.blocks)
    conn = sum(4 * b.sol_conn.size for b in m.blocks)
    N = m.nb_nodes
    return jac + conn + 8 * N * (4 * neq + 2)


def profile_entry(workload: str) -> dict:
    """Figures of the patch kernel measured with rocprofv3 --pmc, if a summary
    for this workload is committed under profiles/: HBM bytes per launch
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor) and FP64 flops
    per launch (SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            data = json.load(f)
        return data.get(workload, {})
    except (OSError, ValueError):
        return {}


def cpu_baseline(wl, seconds: float):
    """The CPU oracle (FP64 restatement of the reference cell loop) on the same
    mesh and state, timed per sweep with std::chrono::steady_clock around the
    cell loop, boundary conditions, dt and update: split over the host threads
    as PGRDMethodExplicit splits its cell loop (nb_threads std::threads over
    cell ranges; at most 16, the box's CPU share), 3 warm-up sweeps then the
    median of at least 20; and single-threaded (the reference's default
    nb_threads = 1), the median of a bounded number of sweeps."""
    from oracle.pyoracle import Problem
    from pdekit_amd.fe import element_tables

    m = wl.mesh
    tables = {b.shape: element_tables(b.shape, m.order, wl.quad_order[b.shape]) for b in m.blocks}
    vol = {b.shape: element_tables(b.shape, m.order, m.order) for b in m.blocks}
    P = Problem(m, tables, vol, wl.physics, wl.scheme, wl.params)
    threads = max(1, min(16, os.cpu_count() or 1))

    def sweeps(th, warm, timed):
        t = np.zeros(warm + timed)
        P.iterate_threads(wl.state, wl.cfl, warm + timed, th, wl.dirichlet_nodes, wl.dirichlet_values,
                          sweep_times=t)
        return t[warm:]

    warm, timed = 3, 20
    tm = sweeps(threads, warm, timed)
    med = float(np.median(tm))
    # single thread: as many sweeps as fit in the remaining budget (at least 3)
    probe = sweeps(1, 0, 1)[0]
    n1 = max(3, min(20, int(seconds / max(probe, 1e-3))))
    t1 = sweeps(1, 0, n1)
    med1 = float(np.median(t1))
    return {"value": wl.nb_elements / med, "unit": "element residual evals/s", "cores": threads,
            "kind": "port",
            "sample": f"{wl.nb_elements} elements of the same mesh and state, oracle/rd_oracle.cpp, {threads} "
                      f"threads over cell ranges: {warm} warm-up sweeps then the median of {timed} "
                      f"({med:.3f} s per sweep); single thread: median of {n1} sweeps after one probe "
                      f"({med1:.3f} s per sweep)",
            "single_thread_value": wl.nb_elements / med1}


SCHEMES = {"lda": 0, "n": 1, "b": 2, "psi": 3, "bpsi": 4}


def serving_bench(args):
    """The serving stream: seeded Poisson arrivals (20 requests/s, seed 5) of
    steady advection P1-P3, Burgers P2 and Euler bump (P1 / Q1, Mach uniform
    in [0.3, 0.6]) solves to a residual norm of 1e-10, served with continuous
    batching and, as the baseline, one at a time in arrival order; p50 / p99
    time-to-converge and throughput of both."""
    from pdekit_amd import serving as S

    sc = args.size_scale
    e0, e1 = (int(x) for x in args.euler_cells.split(","))
    reqs = S.request_stream(args.requests, rate=20.0, seed=5,
                            adv_cells=(int(10_000 * sc), int(1_000_000 * sc)),
                            burgers_cells=(int(50_000 * sc), int(500_000 * sc)), tol=1e-10,
                            euler_cells=(e0, e1))
    t0 = time.perf_counter()
    probs = {r.rid: S.build_problem(r) for r in reqs}
    print(f"[bench] c5: {len(reqs)} requests, {sum(p.mesh.nb_cells for p in probs.values())} cells in all, "
          f"meshes built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    # warm-up: load the code objects of every kernel the stream uses
    warm = S.Server(S.GpuExecutor(), chunk=2, max_iters=2)
    kinds = {}
    for r in reqs:
        kinds.setdefault((r.family, r.order), r)
    small = [S.Request(i, 0.0, k[0], k[1], 2_000, 1e-10, 0.5) for i, k in enumerate(kinds)]
    warm.run(small, {q.rid: S.build_problem(q) for q in small}, mode="batched")
    out = {}
    modes = args.modes.split(",")
    for mode in modes:
        srv = S.Server(S.GpuExecutor(), chunk=50, max_iters=args.max_iters)
        done = [0, time.perf_counter()]

        def progress(res, mode=mode, done=done):
            done[0] += 1
            if done[0] % 64 == 0 or time.perf_counter() - done[1] > 30.0:
                done[1] = time.perf_counter()
                print(f"[bench] c5 {mode}: {done[0]} of {len(reqs)} requests done (request {res.rid}, "
                      f"{res.family}, {res.iterations} iterations)", file=sys.stderr, flush=True)

        res, peak = srv.run(reqs, probs, mode=mode, on_finish=progress)
        out[mode] = S.summary(res)
        out[mode]["peak_footprint_bytes"] = int(peak)
        print(f"[bench] c5 {mode}: {out[mode]}", file=sys.stderr, flush=True)
    if len(modes) < 2:
        print(json.dumps({"metric": "p99 time-to-converge (serving stream)", "modes": out,
                          "config": {"workload": "serving_c5", "requests": len(reqs), "size_scale": sc,
                                     "euler_cells": [e0, e1]}}), flush=True)
        return
    b, q = out["batched"], out["sequential"]
    line = {
        "metric": "p99 time-to-converge (serving stream)",
        "value": round(b["p99_s"], 4),
        "unit": "s",
        "n_gpus": 1,
        "higher_is_better": False,
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "serving_c5", "requests": len(reqs), "rate_per_s": 20.0, "seed": 5,
                   "size_scale": sc, "euler_cells": [e0, e1], "tol": 1e-10, "chunk": 50,
                   "max_iters": args.max_iters, "parallelism": "single GPU",
                   "families": {f: sum(r.family == f for r in reqs) for f in S.FAMILIES}},
        "batched": b,
        "sequential": q,
        "p99_speedup_vs_sequential": round(q["p99_s"] / b["p99_s"], 3),
        "p50_speedup_vs_sequential": round(q["p50_s"] / b["p50_s"], 3),
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c3p2", choices=["c1", "c2", "c3", "c4", "c3p2", "c4p2", "c5"],
                    help="c1-c4, c3p2, c4p2: per-sweep throughput of one mesh (default c3p2, the P2 Euler "
                         "channel); c5: the serving stream")
    ap.add_argument("--requests", type=int, default=64, help="c5: requests in the stream")
    ap.add_argument("--size-scale", type=float, default=0.1,
                    help="c5: scale of the request sizes (1.0: advection 10k-1M, Burgers 50k-500k triangles)")
    ap.add_argument("--euler-cells", default="1000,4000",
                    help="c5: size range of the Euler bump requests (elements; the explicit LDA iteration "
                         "of the bump needs ~3e4 iterations to 1e-10 at 1.7k elements)")
    ap.add_argument("--max-iters", type=int, default=150_000, help="c5: iteration cap of one request")
    ap.add_argument("--modes", default="batched,sequential",
                    help="c5: serving modes to run (the JSON line compares them only when both run)")
    ap.add_argument("--scatter", default="auto", choices=["auto", "atomic", "colour", "patch"])
    ap.add_argument("--scheme", default="default", choices=["default"] + list(SCHEMES),
                    help="element scheme in place of the workload's own (the workload name then carries it)")
    ap.add_argument("--precision", default="plain", choices=["plain", "mixed"],
                    help="plain FP64 sweeps, or the mixed-precision form (linear sweeps with FP32 element "
                         "Jacobians between FP64 sweeps); mixed also times the plain path for the speed-up")
    ap.add_argument("--refresh", type=int, default=50,
                    help="mixed: iterations between FP64 sweeps with fresh Jacobians (PDRD_OPT_REFRESH)")
    ap.add_argument("--patch-cells", default="default", choices=["default", "auto", "64", "128", "256"],
                    help="cells per patch workgroup; auto: pdekit_amd.autotune times 64/128/256 first")
    ap.add_argument("--chain", default="default", choices=["default", "0", "1"],
                    help="patch assembly: shared-node sums of a sweep in the next sweep's gather (1) or a "
                         "fix-up pass after every sweep (0); default: the library's choice")
    ap.add_argument("--jacobians", action="store_true",
                    help="time the element Jacobians of the implicit forms (pdrd_element_jacobians) on the "
                         "workload's mesh and state instead of the explicit iterations")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only run 5 device-to-device copies of this many MiB (reference kernel for the "
                         "FETCH_SIZE / WRITE_SIZE scale of rocprofv3 --pmc) and exit")
    args = ap.parse_args()

    if args.copy_calibration_mb:
        import torch

        n = args.copy_calibration_mb * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "copies": 5}))
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU: the cell loop runs on a single GPU and has no multi-GPU path (--gpus {args.gpus} requested). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "c5":
        serving_bench(args)
        return

    from pdekit_amd import _lib as L
    from pdekit_amd.rdm import PGRDMethodExplicit
    from pdekit_amd.workloads import BUILDERS

    t_setup = time.perf_counter()
    wl = BUILDERS[args.workload]() if args.scheme == "default" else \
        BUILDERS[args.workload](scheme=SCHEMES[args.scheme])
    E = wl.nb_elements
    cells = None
    if args.patch_cells == "auto":
        from pdekit_amd import autotune

        cells = autotune.tune(wl.mesh, wl.physics, wl.scheme, wl.params, wl.quad_order, wl.state, wl.cfl,
                              precision=1 if args.precision == "mixed" else 0)
        print(f"[bench] autotuned cells per patch: {cells}", file=sys.stderr, flush=True)
    elif args.patch_cells != "default":
        cells = int(args.patch_cells)
    meth = PGRDMethodExplicit(wl.physics, wl.scheme, params=wl.params if wl.params else None)
    meth.configure_mesh_data(wl.mesh).initialize_work_data(quad_order=wl.quad_order, patch_cells=cells)
    if args.precision == "mixed":
        meth.set_option(L.OPT_PRECISION, 1)
        meth.set_option(L.OPT_REFRESH, args.refresh)
    if args.chain != "default":
        meth.set_option(L.OPT_CHAIN, int(args.chain))
    meth.set_vec_function("solution", wl.state)
    if wl.dirichlet_nodes.size:
        meth.add_boundary_condition("StrongDirichlet", "inflow", wl.dirichlet_nodes, wl.dirichlet_values)
    if args.jacobians:
        jacobian_bench(args, wl, meth)
        return
    setup_s = time.perf_counter() - t_setup
    print(f"[bench] {wl.name}: {E} elements, {wl.mesh.nb_nodes} nodes, setup {setup_s:.1f} s", file=sys.stderr,
          flush=True)

    # scatter: measured choice (cell loop + nodal pass) between the patch
    # assembly, colour-by-colour launches and FP64 atomics
    modes = {"atomic": [0], "colour": [1], "patch": [2], "auto": [2, 1, 0]}[args.scatter]
    if args.precision == "mixed":
        modes = [2]  # the mixed cells run with the patch assembly only
    best = None
    for mode in modes:
        meth.set_option(L.OPT_SCATTER, mode)
        meth.set_option(L.OPT_TIMING, 0)
        meth.set_vec_function("solution", wl.state)
        meth.iterate(wl.cfl, max(2, args.warmup))  # first launches load the code object
        meth.set_option(L.OPT_TIMING, 1)
        meth.iterate(wl.cfl, max(2, args.warmup))
        cell_ms, node_ms = meth.last_timing()
        print(f"[bench] scatter={SCATTER_NAMES[mode]}: cell loop {cell_ms * 1e3:.1f} us, "
              f"nodal pass {node_ms * 1e3:.1f} us", file=sys.stderr, flush=True)
        if best is None or cell_ms + node_ms < best[1]:
            best = (mode, cell_ms + node_ms)
    meth.set_option(L.OPT_SCATTER, best[0])

    # warm-up, then the timed region
    meth.set_option(L.OPT_TIMING, 0)
    meth.set_vec_function("solution", wl.state)
    meth.iterate(wl.cfl, max(1, args.warmup))
    meth.synchronize()
    t0 = time.perf_counter()
    meth.iterate(wl.cfl, args.steps)
    meth.synchronize()
    t = time.perf_counter() - t0
    meth.check_status()
    value = E * args.steps / t

    # per-launch duration of the cell loop in a second timed pass of the same
    # length, HIP events on the method's stream around every launch (mixed:
    # the linear sweeps' kernels only)
    mixed = args.precision == "mixed"
    meth.set_option(L.OPT_TIMING, 2 if mixed else 1)
    meth.iterate(wl.cfl, args.steps)
    cell_ms, node_ms = meth.last_timing()
    info = meth.info()
    plain = None
    if mixed:
        # the plain FP64 path on the same method and state, same steps
        meth.set_option(L.OPT_TIMING, 0)
        meth.set_option(L.OPT_PRECISION, 0)
        meth.set_vec_function("solution", wl.state)
        meth.iterate(wl.cfl, max(1, args.warmup))
        meth.synchronize()
        t0 = time.perf_counter()
        meth.iterate(wl.cfl, args.steps)
        meth.synchronize()
        tp = time.perf_counter() - t0
        plain = {"value": round(E * args.steps / tp, 1), "ms_per_step": round(tp / args.steps * 1e3, 4),
                 "speedup": round((E * args.steps / t) / (E * args.steps / tp), 3)}
        print(f"[bench] mixed {t / args.steps * 1e3:.4f} ms/step, plain {tp / args.steps * 1e3:.4f} ms/step, "
              f"linear sweep kernels {cell_ms * 1e3:.1f} + {node_ms * 1e3:.1f} us", file=sys.stderr, flush=True)

    cell_bytes, sweep_bytes = algorithmic_bytes(wl, best[0])
    if mixed:
        cell_bytes = linear_bytes(wl)
    achieved = cell_bytes / (cell_ms * 1e-3) / 1e9
    prof = profile_entry(wl.name) if (best[0] == 2 and args.precision == "plain") else {}
    traffic = prof.get("bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": ("rd_patch_kernel, linear sweep (J delta + nodal step)" if mixed else
                           "rd_patch_kernel (cell loop + nodal step)" if best[0] == 2 else "rd_cell_kernel (cell loop)"),
                "algorithmic_bytes": int(cell_bytes),
                "avg_launch_us": round(cell_ms * 1e3, 2)}
    out = {
        "metric": "element residual evals/s",
        "value": round(value, 1),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "plain" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": wl.name, "precision": args.precision, "elements": E, "nodes": wl.mesh.nb_nodes,
                   "order": wl.mesh.order, "scheme": ["PGLDA", "PGN", "PGB", "PSI", "PSI/LDA blend"][wl.scheme],
                   "scatter": SCATTER_NAMES[best[0]], "colours": info["nb_colours"],
                   "patch_cells": meth.get_option(L.OPT_PATCH_CELLS) if best[0] == 2 else None,
                   "chain": meth.get_option(L.OPT_CHAIN) if best[0] == 2 else None,
                   "parallelism": "single GPU"},
        "roofline": roofline,
        # the FP64 vector pipe is the binding roof of this kernel: flops per
        # launch from the committed rocprofv3 counters over the live launch time
        "fp64": ({"achieved": round(prof["fp64_flops_per_launch"] / (cell_ms * 1e-3) / 1e12, 2),
                  "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                  "frac": round(prof["fp64_flops_per_launch"] / (cell_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
                  "flops_per_element": round(prof["fp64_flops_per_launch"] / E, 1)}
                 if "fp64_flops_per_launch" in prof else None),
        "sweep_hbm_frac": None if mixed else round(sweep_bytes / (t / args.steps) / 1e9 / HBM_PEAK_GBS, 4),
        "nodal_pass_us": round(node_ms * 1e3, 2),  # shared-node fix-up with the patch assembly
    }
    if mixed:
        out["config"]["refresh"] = args.refresh
        out["plain"] = plain
    # a caller that hands over host buffers every step (set the state, one
    # sweep, read the state back through the C ABI): the PCIe-inclusive rate,
    # reported beside the resident one, never as `value`
    meth.set_option(L.OPT_TIMING, 0)
    u_host = meth.vec_function("solution")
    n_rt = 10
    meth.synchronize()
    t0 = time.perf_counter()
    for _ in range(n_rt):
        meth.set_vec_function("solution", u_host)
        meth.iterate(wl.cfl, 1)
        u_host = meth.vec_function("solution")
    t_rt = time.perf_counter() - t0
    out["host_roundtrip"] = {"value": round(E * n_rt / t_rt, 1), "unit": "elements/s",
                             "ms_per_step": round(t_rt / n_rt * 1e3, 4), "steps": n_rt,
                             "bytes_per_step": int(2 * 8 * u_host.size)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def jacobian_bench(args, wl, meth):
    """Element Jacobians (pdrd_element_jacobians) of every block: the kernel's
    duration from HIP events around its launch (pdrd_last_timing), mean of
    `steps` calls per block after `warmup`; compulsory bytes = the kernel's
    outputs (residual [nd m], Jacobian [(nd m)^2], wave speed per cell) and
    inputs (connectivity, coordinates, u)."""
    from pdekit_amd import _lib as L

    m = wl.mesh
    neq = wl.state.shape[1]
    need = sum(b.sol_conn.shape[0] * (b.sol_conn.shape[1] * neq) ** 2 * 8 for b in m.blocks)
    if need > 4 * 2**30:
        print(f"bench.py --jacobians: the Jacobians of {wl.name} fill {need / 2**30:.1f} GiB of host memory "
              "per call; use c1, c2 or c3", file=sys.stderr)
        sys.exit(2)
    total_ms, out_bytes = 0.0, 0
    for blk, b in enumerate(m.blocks):
        ne, nd = b.sol_conn.shape
        for _ in range(max(1, args.warmup)):
            meth.element_jacobians(blk)
        ms = 0.0
        for _ in range(args.steps):
            meth.element_jacobians(blk)
            ms += meth.last_timing()[0]
        ms /= args.steps
        total_ms += ms
        ndm = nd * neq
        out_bytes += ne * 8 * (ndm * ndm + ndm + 1) + 4 * b.sol_conn.size + 4 * b.geo_conn.size
        print(f"[bench] jacobians block {blk}: {ne} cells, {ms * 1e3:.1f} us", file=sys.stderr, flush=True)
    in_bytes = 8 * m.coords.size + 8 * neq * m.nb_nodes
    alg = out_bytes + in_bytes
    achieved = alg / (total_ms * 1e-3) / 1e9
    print(json.dumps({
        "metric": "element Jacobian evals/s", "value": round(m.nb_cells / (total_ms * 1e-3), 1),
        "unit": "elements/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total_ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": wl.name, "scheme": ["PGLDA", "PGN", "PGB", "PSI", "PSI/LDA blend"][wl.scheme],
                   "elements": m.nb_cells, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "rd_jacobian_qp_kernel" if neq > 1 else "rd_jacobian_kernel",
                     "algorithmic_bytes": int(alg), "avg_launch_us": round(total_ms * 1e3, 2)},
    }), flush=True)


if __name__ == "__main__":
    main()
