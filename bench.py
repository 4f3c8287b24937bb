"""Benchmark: particle updates per second of the cutting step on one MI355X.

Workload (default): the 4M-particle Ti6Al4V case -- 4x1 mm workpiece on a 4000x1000 jittered
lattice (dx = 1 um), a ~50k-triangle rigid FEM tool with transient heat conduction and Usui
nodal-shift wear, 70 m/min.  The step is the full leap-frog step (cell list, pair sums,
corrector, friction heat, plasticity, FEM tool heat, wear, tool motion, contact, boundary
conditions), inputs resident in HBM.

Prints one JSON line with the metric, a roofline object for the interaction kernel (HIP
events on the library's stream inside the timed region) and a CPU baseline (the FP64 host
oracle on a bounded sample of the same workload, on this host's cores).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
INTERACTION_KERNEL = {"fp64": "pairs_fp64", "fp32c": "pairs_fp32c"}
FUSED_KERNEL = "pairs_update_fp32c"      # the FP32C pair kernel that also runs the update
# Algorithmic bytes per particle of the FP32C update: reads idx 4, blanked, tool flag, fixed 3x8,
# pos 16, vel 16, S 32, rho 8, T 8, eps_pl 8, eps_pl_dot 8, on_seg 4, h 8, the FP32 rate carry
# (vel_t 8, S_t 16, rho_t 4, T_t 4) = 176; writes the sorted state (pos 16, vel 16, S 32, rho, T,
# eps_pl, eps_pl_dot, blanked 5x8, fc, ft, n 3x16, on_seg 4, h, fixed, tool flag, p 4x8) = 192 and
# the new carry (pos_t 8, vel_t 8, S_t 16, rho_t 4, T_t 4) = 40.  The contact fields of particles
# far from the tool are known to be zero and not read.
UPDATE_BYTES_FP32C = 176 + 192 + 40


def interaction_bytes(N, n_occ_cells, w):
    """Algorithmic bytes of the fused pair-sum stage: 14 input and 10 output scalars of w bytes,
    2 flag bytes per particle, 8 bytes per occupied cell."""
    return N * (24 * w + 2) + 8 * n_occ_cells


def pairs_update_bytes(N, n_occ_cells):
    """Algorithmic bytes of the FP32C pair kernel that runs each particle's update after its sums:
    the pair stage's inputs (14 FP32 scalars, 2 flag bytes per particle, 8 bytes per occupied cell;
    its sums stay in registers) plus the update's reads and writes."""
    return N * (14 * 4 + 2 + UPDATE_BYTES_FP32C) + 8 * n_occ_cells


def measured_traffic(kernel, workload_key):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of
    this workload (profiles/traffic_<workload>.json), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload_key}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    # (name prefix, template-argument tail) of the kernel's instantiations
    sel = {"pairs_fp32c": ("sphfem::k_pairs_fp32c<", ", false, false>"),
           "pairs_update_fp32c": ("sphfem::k_pairs_fp32c<", ", true, false>"),
           "pairs_fp64": ("sphfem::k_pairs_fp64<", ">"),
           "fused_update": ("sphfem::k_fused_update<false,", ", false>"),
           "update_fp32c": ("sphfem::k_fused_update<true,", ", false>")}.get(kernel)
    if sel is None:
        return None
    prefix, tail = sel
    # the launch shapes of one kernel move the same bytes: take the most profiled one
    hits = [v for k, v in t.get("per_launch", {}).items() if k.startswith(prefix) and k.endswith(tail)]
    if not hits:
        return None
    v = max(hits, key=lambda x: x["dispatches"])
    return v["read_bytes"] + v["write_bytes"]


def make_case(name):
    from sph_fem_2d_cutting_gpu_amd import cases
    if name == "ti64_4m":
        return cases.config3(), "ti64_4M_cutting: 4000x1000 lattice dx=1um, ~50k-triangle FEM tool, FEM heat + Usui wear"
    if name == "aisi1045_160k":
        return cases.config2(), "aisi1045_160k_cutting: 800x200 lattice dx=1.25um, ~8k-triangle FEM tool"
    if name == "ti64_6400":
        return cases.config1(), "ti64_6400_cutting: 160x40 lattice dx=2.5um, ~2k-triangle FEM tool"
    if name == "chip16m":
        c = cases.chip_case(n_target=16_000_000)
        return c, (f"chip16m: chip-shaped domain with a 10:1 density range (slab, chip root, chip band), "
                   f"{c.N} particles, dx=1um, no tool; kernel-only pass")
    if name.startswith("ref_cut_sph_ny"):
        ny = int(name[len("ref_cut_sph_ny"):])
        c = cases.ref_cut_sph(ny=ny)
        return c, (f"ref_cut_sph_ny{ny}: setup_ref_cut rigid tool (rake 0, clearance 4.5 deg, fillet 3.5 um, "
                   f"feed 100 um, 750 m/min), {c.extra['n_workpiece']} workpiece + "
                   f"{c.N - c.extra['n_workpiece']} tool particles, dx={c.dx * 1e4:.3f}um")
    if name.startswith("rings_nbox"):
        nb = int(name[len("rings_nbox"):])
        c = cases.rings(nb)
        return c, f"rings_nbox{nb}: setup_rings rubber rings at 200 m/s, {c.N} particles, dx={c.dx * 1e3:.4f}mm"
    if name.startswith("impact_nbox"):
        nb = int(name[len("impact_nbox"):])
        c = cases.impact(nb)
        return c, (f"impact_nbox{nb}: setup_impact steel disc at 1000 m/s on a wall, {c.N} particles, "
                   f"dx={c.dx * 1e3:.3f}mm, unlocked grid (eager steps)")
    raise SystemExit(f"unknown workload {name}")


KERNEL_PASS = {"fp64": "interactions_ref", "fp32c": "pairs_fp32c"}


PAIR_KERNELS = {"list": 0, "staged": 1, "list_fused": 2, "staged_fused": 3}   # the library's pair_kernel values


def run_kernel_pass(case, precision, steps, warmup, tune=True, pair_kernel="auto"):
    """Kernel-only step (the 16M chip domain): hash, stable sort with the gather of every particle
    field (grid_gpu_green::sort), cell table, then one interaction pass -- the FP32C pass of the
    step's pair kernel (records, pair sums, expansion into the reference's fields) or the FP64
    interactions_monaghan of interactions_gpu.cu:143-379 -- on particles that do not move (the
    domain's velocities would carry them out of its tight grid).  Eager launches.  Returns
    (seconds for `steps` passes, interaction-kernel seconds per launch, occupied cells, tuning)."""
    import ctypes as C
    from sph_fem_2d_cutting_gpu_amd import _lib, autotune
    from sph_fem_2d_cutting_gpu_amd import solver as S
    prec = {"fp64": _lib.FP64, "fp32c": _lib.FP32C}[precision]
    sim = S.Simulation(case, precision=prec)
    p, g, ctx = sim.p, sim.grid, sim.ctx
    nc = g.num_cell
    cs, ce = C.c_void_p(), C.c_void_p()
    ctx.call("sphfem_malloc", C.byref(cs), 4 * nc)
    ctx.call("sphfem_malloc", C.byref(ce), 4 * nc)
    S.interactions_setup_geometry_constants(g)
    S.material_eos(p)
    S.corrector_artificial_stress(p)

    def passes(n):
        for _ in range(n):
            g.assign_hashes(p)
            g.sort(p)
            g.get_cells(p, cs.value, ce.value)
            if prec == _lib.FP32C:
                S.interactions_monaghan_fp32c(sim.stepper, p, cs.value, ce.value, nc)
            else:
                S.interactions_monaghan(p, cs.value, ce.value, nc)

    passes(max(warmup, 1))
    ctx.synchronize()
    ctx.check_error()
    tuning = None
    if tune and prec == _lib.FP32C:
        rows = []
        for cand in autotune.candidates(prec):
            # the pass has no update: the fused forms would run the same pair kernels
            if cand["pair_kernel"] in (_lib.PAIRS_LIST_FUSED, _lib.PAIRS_STAGED_FUSED):
                continue
            if pair_kernel != "auto" and cand["pair_kernel"] != PAIR_KERNELS[pair_kernel.replace("_fused", "")]:
                continue
            sim.set_tuning(**cand)
            passes(1)
            ctx.synchronize()
            t0 = time.perf_counter()
            passes(3)
            ctx.synchronize()
            rows.append((time.perf_counter() - t0, cand))
        tuning = min(rows, key=lambda r: r[0])[1]
        sim.set_tuning(**tuning)
    kname = KERNEL_PASS[precision]
    ctx.timing_select(kname)
    ctx.timing(True)
    ctx.timing_reset()
    passes(min(steps, 10))
    ctx.synchronize()
    ctx.timing(False)
    k_ms, k_n = ctx.timing_get(kname)
    ctx.synchronize()
    t0 = time.perf_counter()
    passes(steps)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.check_error()
    cs_h = np.empty(nc, np.int32)
    ctx.call("sphfem_memcpy_dtoh", cs_h.ctypes.data_as(C.c_void_p), cs, 4 * nc)
    ctx.call("sphfem_free", cs)
    ctx.call("sphfem_free", ce)
    sim.close()
    return elapsed, k_ms / max(k_n, 1) / 1e3, int(np.count_nonzero(cs_h != -1)), tuning


def cpu_kernel_pass(case, budget_s):
    """The FP64 host oracle's kernel-only pass (hash, stable sort and cells, interactions) on this
    host's cores, a bounded number of passes; the one-thread figure times the interaction pass
    over the first 500,000 sorted particles and scales it to N."""
    from oracle.oracle import Oracle, lib
    o = Oracle(case)
    threads = lib().o_num_threads()
    o.call("o_material_eos", "cp")
    o.call("o_artificial_stress", "cp")

    def one():
        o.call("o_compute_hashes", "cp")
        o.call("o_sort_and_cells", "cps")
        o.call("o_interactions_monaghan", "cps")

    t0 = time.perf_counter()
    one()
    t1 = time.perf_counter() - t0
    n = int(max(1, min(10, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    for _ in range(n):
        one()
    dt = time.perf_counter() - t0
    w = min(500_000, case.N)
    lib().o_set_num_threads(1)
    t0 = time.perf_counter()
    o.call("o_interactions_monaghan_range", "cps", 0, w)
    d1 = time.perf_counter() - t0
    lib().o_set_num_threads(threads)
    return {"value": case.N * n / dt, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} kernel-only passes (hash, stable sort and cells, FP64 interactions; after one "
                      f"warm-up pass) on the FP64 host oracle, {threads} OpenMP threads; one-thread figure: "
                      f"the interaction pass over {w} particles, scaled",
            "single_thread": {"value": w / d1, "interactions_only": True},
            "cpu_model": cpu_model(), "nproc": os.cpu_count()}


def main_kernel_pass(args, case, workload):
    elapsed, k_s, n_occ, tuning = run_kernel_pass(case, args.precision, args.steps, args.warmup,
                                                  not args.no_tune, args.pair_kernel)
    w = 8 if args.precision == "fp64" else 4
    b_int = interaction_bytes(case.N, n_occ, w)
    achieved = b_int / k_s / 1e9
    ups = case.N * args.steps / elapsed
    out = {
        "metric": "particle-updates/s of the kernel-only pass (sort + interactions), FP32-comp vs FP64",
        "value": ups, "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 pair sums (FP64 block accumulation) + f64 state",
        "data": "synthetic (seeded chip-shaped point set)",
        "config": {"workload": workload, "particles": case.N, "precision": args.precision,
                   "parallelism": "single GPU", "step": "hash + sort + cells + one interaction pass, eager",
                   "tuning": tuning},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": KERNEL_PASS[args.precision],
                     "kernel_ms": k_s * 1e3, "algorithmic_bytes": b_int, "occupied_cells": n_occ},
    }
    if args.precision == "fp32c" and not args.no_compare:
        e64, k64, _, _ = run_kernel_pass(case, "fp64", args.steps, args.warmup, False)
        ups64 = case.N * args.steps / e64
        out["fp64_form"] = {"value": ups64, "ms_per_step": e64 / args.steps * 1e3,
                            "kernel": KERNEL_PASS["fp64"], "kernel_ms": k64 * 1e3,
                            "roofline_frac": interaction_bytes(case.N, n_occ, 8) / k64 / 1e9 / HBM_PEAK_GBS}
        out["speedup_vs_fp64"] = ups / ups64
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_kernel_pass(case, args.cpu_budget)
    print(json.dumps(out))


def cpu_model():
    """The host CPU's model name (what lscpu reports), from /proc/cpuinfo."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, budget_s):
    """The FP64 host oracle on the host's cores: a bounded sample of steps with the OpenMP threads the
    environment gives (the reported value), then one step on one thread."""
    from oracle.oracle import Oracle, lib
    o = Oracle(case)
    threads = lib().o_num_threads()
    t0 = time.perf_counter()
    o.step(1)
    t1 = time.perf_counter() - t0
    n = int(max(1, min(20, budget_s / max(t1, 1e-6))))
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    lib().o_set_num_threads(1)
    t0 = time.perf_counter()
    o.step(1)
    d1 = time.perf_counter() - t0
    lib().o_set_num_threads(threads)
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"value": case.N * n / dt, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": f"{n} steps (after 1 warm-up step) of the same workload on the FP64 host oracle, "
                      f"{threads} OpenMP threads; then 1 step on 1 thread",
            "single_thread": {"value": case.N / d1, "steps": 1},
            "cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity": affinity}


def run_form(case, precision, steps, warmup, breakdown=False, graph=True, tune=True, pair_kernel="auto"):
    """Time `steps` full steps of one arithmetic form; returns (seconds, interaction-kernel seconds, sim).

    The timed loop replays a captured two-step HIP graph (an odd step count ends with one eager
    step).  The interaction kernel's time comes from a separate eager pass with HIP events around
    that kernel only, on the library's stream."""
    from sph_fem_2d_cutting_gpu_amd import _lib
    from sph_fem_2d_cutting_gpu_amd import solver as S
    prec = {"fp64": _lib.FP64, "fp32c": _lib.FP32C}[precision]
    sim = S.Simulation(case, precision=prec)
    ctx = sim.ctx
    for _ in range(max(warmup, 1)):
        sim.step(1)
    ctx.synchronize()
    ctx.check_error()
    tuning = None
    if tune:                                           # untimed: advances the warm-up
        from sph_fem_2d_cutting_gpu_amd import autotune
        kernels = None if pair_kernel == "auto" else PAIR_KERNELS[pair_kernel]
        tuning = autotune.tune_simulation(sim, prec, steps=4, repeats=2, pair_kernel=kernels)["best"]
    elif pair_kernel != "auto":
        sim.set_tuning(pair_kernel=PAIR_KERNELS[pair_kernel])
    # interaction-kernel time (eager steps, events around that kernel only): the FP32C pair kernel,
    # or the pair kernel that also runs the update
    kname = INTERACTION_KERNEL[precision]
    if prec == _lib.FP32C and sim.get_tuning()["pair_kernel"] in (_lib.PAIRS_LIST_FUSED, _lib.PAIRS_STAGED_FUSED):
        kname = FUSED_KERNEL
    ctx.timing_select(kname)
    ctx.timing(True)
    ctx.timing_reset()
    sim.step(min(steps, 20))
    ctx.synchronize()
    ctx.timing(False)
    k_ms, k_n = ctx.timing_get(kname)
    g = sim.capture(2) if graph else None
    if g is not None:
        g.launch(1)                                    # graph warm-up
    ctx.synchronize()
    t0 = time.perf_counter()
    if g is not None:
        g.launch(steps // 2)
        sim.step(steps % 2)
    else:
        sim.step(steps)
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    ctx.check_error()
    if g is not None:
        g.destroy()
    if breakdown:
        # separate, untimed eager pass with an event pair around every kernel
        ctx.timing_select("")
        ctx.timing(True)
        ctx.timing_reset()
        sim.step(steps)
        ctx.synchronize()
        ctx.timing(False)
        rows = []
        for name in ctx.timing_names():
            ms, n = ctx.timing_get(name)
            if n:
                rows.append((ms / steps, n / steps, name))
        rows.sort(reverse=True)
        print(f"[{precision}] per-step kernel time (ms), {steps} eager steps:", file=sys.stderr)
        for ms, n, name in rows:
            print(f"  {name:28s} {ms:9.4f}  x{n:.0f}", file=sys.stderr)
    sim.bench_tuning = tuning
    sim.bench_kernel = kname
    return elapsed, (k_ms / 1e3) / max(k_n, 1), sim


def run_batch(args):
    """Batched sweep: cases advance together (one stream and captured graph per case) vs one after
    another; aggregate particle-updates/s of both."""
    from sph_fem_2d_cutting_gpu_amd import _lib, cases
    from sph_fem_2d_cutting_gpu_amd import solver as S
    cs = cases.batch_cases(n_cases=args.batch_cases)
    prec = {"fp64": _lib.FP64, "fp32c": _lib.FP32C}[args.precision]
    batch = S.BatchedSimulation(cs, precision=prec)
    batch.step(max(args.warmup, 1))
    batch.synchronize()
    batch.check_error()
    tuned = None
    if not args.no_tune:
        from sph_fem_2d_cutting_gpu_amd import autotune
        tuned = autotune.tune_batch(batch, prec)
    else:
        batch.capture()
    batch.launch(1)
    batch.synchronize()
    reps = max(1, args.steps // batch.graph_steps)
    t0 = time.perf_counter()
    batch.launch(reps)
    batch.synchronize()
    t_batch = time.perf_counter() - t0
    batch.launch_sequential(1)          # per-case graphs, captured on first use
    batch.synchronize()
    t0 = time.perf_counter()
    batch.launch_sequential(reps)
    t_seq = time.perf_counter() - t0
    batch.check_error()
    n_tot = batch.total_particles
    steps = batch.graph_steps * reps
    ns = [c.N for c in cs]
    out = {
        "metric": f"aggregate particle-updates/s, {len(cs)}-case batched cutting sweep",
        "value": n_tot * steps / t_batch, "unit": "particle-updates/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": t_batch / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 pair sums (FP64 block accumulation) + f64 state",
        "data": "synthetic (seeded parameter sweep, generated tool meshes)",
        "config": {"workload": f"batch{len(cs)}_sweep: dx=1.25um, length U(0.5,3) mm, speed U(30,300) m/min, "
                               "rake U(-10,15) deg, doc U(20,80) um, Ti6Al4V/AISI1045",
                   "cases": len(cs), "particles_total": n_tot, "particles_min": min(ns), "particles_max": max(ns),
                   "precision": args.precision,
                   "parallelism": ("single GPU, one HIP graph: each case a branch on its own stream"
                                   if batch.graph is not None else "single GPU, one step graph per case, on its own stream"),
                   "graph_steps": batch.graph_steps,
                   "tuning_ms_per_step": tuned["ms_per_step"] if tuned else None},
        "sequential": {"value": n_tot * steps / t_seq, "ms_per_step": t_seq / steps * 1e3},
        "speedup_vs_sequential": t_seq / t_batch,
    }
    batch.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="ti64_4m")
    ap.add_argument("--precision", default="fp32c", choices=["fp32c", "fp64"])
    ap.add_argument("--no-compare", action="store_true", help="skip timing the FP64 form beside FP32C")
    ap.add_argument("--breakdown", action="store_true", help="per-kernel times on stderr")
    ap.add_argument("--eager", action="store_true", help="time eager steps instead of the captured step graph")
    ap.add_argument("--batch-cases", type=int, default=64, help="cases of the batch64 workload")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default launch parameters)")
    ap.add_argument("--pair-kernel", default="auto", choices=["auto"] + list(PAIR_KERNELS),
                    help="FP32C pair kernel the tuner may choose: any (auto) or only one; the _fused forms "
                         "also run the update")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 256 MiB elementwise copy (y = x * 1, fp32) 5 times: the reference "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        x = torch.ones(64 * 1024 * 1024, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            torch.mul(x, 1.0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": x.numel() * 4, "bytes_written": y.numel() * 4}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (the solver has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch64":
        run_batch(args)
        return
    case, workload = make_case(args.workload)
    if args.workload in ("chip16m",):
        main_kernel_pass(args, case, workload)
        return
    graph = not args.eager and case.max_cell is None      # an unlocked grid reads its geometry back
    elapsed, k_avg_s, sim = run_form(case, args.precision, args.steps, args.warmup, args.breakdown, graph,
                                     not args.no_tune, args.pair_kernel)
    tuning = sim.bench_tuning
    kname = sim.bench_kernel
    cs, _ = sim.grid.download_cells()
    n_occ = int(np.count_nonzero(cs != -1))
    w = 8 if args.precision == "fp64" else 4
    b_int = pairs_update_bytes(case.N, n_occ) if kname == FUSED_KERNEL else interaction_bytes(case.N, n_occ, w)
    achieved = b_int / k_avg_s / 1e9
    ups = case.N * args.steps / elapsed
    sim.close()
    del sim

    out = {
        "metric": "particle-updates/s at 4M particles, FP32-comp vs FP64; % HBM roofline",
        "value": ups,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 pair sums (FP64 block accumulation) + f64 state",
        "data": "synthetic (seeded jittered lattice and generated tool mesh)",
        "config": {"workload": workload, "particles": case.N,
                   "tool_triangles": int(len(case.tool.elems)) if case.tool is not None else 0,
                   "cutting_segments": int(len(case.tool.seg_left)) if case.tool is not None else 0,
                   "rigid_tool": case.rigid is not None, "wear": bool(case.wear),
                   "precision": args.precision, "parallelism": "single GPU", "step_graph": graph,
                   "tuning": tuning},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kname, args.workload),
                     "kernel": kname, "kernel_ms": k_avg_s * 1e3,
                     "algorithmic_bytes": b_int, "occupied_cells": n_occ},
    }
    if args.precision == "fp32c" and not args.no_compare:
        e64, k64, sim64 = run_form(case, "fp64", args.steps, args.warmup, args.breakdown, graph, not args.no_tune)
        sim64.close()
        ups64 = case.N * args.steps / e64
        out["fp64_form"] = {"value": ups64, "ms_per_step": e64 / args.steps * 1e3,
                            "kernel": INTERACTION_KERNEL["fp64"], "kernel_ms": k64 * 1e3,
                            "roofline_frac": interaction_bytes(case.N, n_occ, 8) / k64 / 1e9 / HBM_PEAK_GBS}
        out["speedup_vs_fp64"] = ups / ups64
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, args.cpu_budget)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
