"""Autotuner for the step's launch parameters and the batch's graph depth.

Tunables (all result-neutral except the FP32 summation blocks of the FP32C form):
  * pair_block   -- pair kernel workgroup size (128, 256), both forms
  * pair_chunk   -- FP32C neighbour-list entries per lane in LDS (16, 32, 48): larger lists run
                    fewer list flushes in dense regions, smaller ones leave LDS for more waves
  * update_block -- update kernel workgroup size (128, 256)
  * pair_kernel  -- FP32C pair kernel: the list kernel (every neighbour record gathered through L1)
                    or the staged kernel (each wavefront's window positions copied into LDS first;
                    one wavefront per workgroup, so pair_block does not apply); same sums.  Each
                    also fused with the update (the particle's update right after its sums: no
                    pair-sum round trip, update_block does not apply); same step result
  * resort_one_block -- the step's full-sort fallback by one workgroup in one launch instead of the
                    multi-block radix sort's eight per-step launches (N <= 2^19; tried with the best
                    of the others: it pays when launches dominate and full sorts are rare)
  * graph_steps  -- steps per captured graph in a batch (2 ... 32): fewer graph launches per step
The tuner times each candidate on the live workload (it advances the simulation: run it in the
warm-up), keeps the fastest and can store the choice in a JSON cache keyed by workload size,
precision and device.
"""
from __future__ import annotations

import itertools
import json
import os
import time

from . import _lib

PAIR_BLOCKS = (128, 256)
PAIR_CHUNKS = (16, 32, 48)
UPDATE_BLOCKS = (128, 256)
GRAPH_STEPS = (2, 4, 8, 16, 32)


def candidates(precision):
    if precision == _lib.FP32C:
        listed = [dict(pair_block=b, pair_chunk=c, update_block=u, pair_kernel=_lib.PAIRS_LIST, resort_one_block=0)
                  for b, c, u in itertools.product(PAIR_BLOCKS, PAIR_CHUNKS, UPDATE_BLOCKS)]
        staged = [dict(pair_block=256, pair_chunk=c, update_block=u, pair_kernel=_lib.PAIRS_STAGED, resort_one_block=0)
                  for c, u in itertools.product(PAIR_CHUNKS, UPDATE_BLOCKS)]
        fused = [dict(pair_block=b, pair_chunk=c, update_block=256, pair_kernel=_lib.PAIRS_LIST_FUSED,
                      resort_one_block=0) for b, c in itertools.product(PAIR_BLOCKS, PAIR_CHUNKS)]
        fused += [dict(pair_block=256, pair_chunk=c, update_block=256, pair_kernel=_lib.PAIRS_STAGED_FUSED,
                       resort_one_block=0) for c in PAIR_CHUNKS]
        return listed + staged + fused
    return [dict(pair_block=b, pair_chunk=32, update_block=u, pair_kernel=_lib.PAIRS_LIST, resort_one_block=0)
            for b, u in itertools.product(PAIR_BLOCKS, UPDATE_BLOCKS)]


def _time_steps(sim, steps):
    sim.ctx.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.ctx.synchronize()
    return (time.perf_counter() - t0) / steps


def tune_simulation(sim, precision, steps=6, repeats=2, pair_kernel=None) -> dict:
    """Time every candidate for `steps` steps (best of `repeats`), apply the fastest; returns the
    choice and all timings (ms per step).  pair_kernel limits the candidates to one FP32C pair
    kernel.  Advances the simulation by len(candidates) * steps * repeats steps."""
    rows = []
    for cand in candidates(precision):
        if pair_kernel is not None and cand["pair_kernel"] != pair_kernel:
            continue
        sim.set_tuning(**cand)
        t = min(_time_steps(sim, steps) for _ in range(repeats))
        rows.append((t, cand))
    sim.ctx.check_error()
    t_best, best = min(rows, key=lambda r: r[0])
    if sim.case.N <= _lib.SORT_ONE_BLOCK_MAX:
        cand = dict(best, resort_one_block=1)
        sim.set_tuning(**cand)
        t = min(_time_steps(sim, steps) for _ in range(repeats))
        rows.append((t, cand))
        sim.ctx.check_error()
        if t < t_best:
            t_best, best = t, cand
    sim.set_tuning(**best)
    return {"best": best, "ms_per_step": t_best * 1e3,
            "timings": [{**c, "ms_per_step": t * 1e3} for t, c in rows]}


def tune_batch(batch, precision, steps=4, launches=2) -> dict:
    """Per-case kernel tuning, then how the batch is replayed: one graph holding every case as a
    branch, or one graph per case queued on the cases' own streams, each at graph depths
    GRAPH_STEPS (recaptures the graphs; the fastest is left captured)."""
    per_case = [tune_simulation(s, precision, steps=steps, repeats=1)["best"] for s in batch.sims]
    timings = {}
    for one in (True, False):
        for gs in GRAPH_STEPS:
            batch.release_graphs()
            batch.graph_steps = gs
            batch.capture(one_graph=one)
            batch.launch(1)
            batch.synchronize()
            reps = max(1, (launches * max(GRAPH_STEPS)) // gs)
            t0 = time.perf_counter()
            batch.launch(reps)
            batch.synchronize()
            timings[(one, gs)] = (time.perf_counter() - t0) / (reps * gs)
    one, best = min(timings, key=timings.get)
    batch.release_graphs()
    batch.graph_steps = best
    batch.capture(one_graph=one)
    batch.check_error()
    return {"graph_steps": best, "one_graph": one,
            "ms_per_step": {f"{'one graph' if o else 'graph per case'}, {g} steps": v * 1e3
                            for (o, g), v in timings.items()},
            "per_case": per_case}


def cache_key(n_particles, precision, device_name=""):
    bucket = max(0, int(n_particles).bit_length() - 1)
    return f"{device_name}|{'fp32c' if precision == _lib.FP32C else 'fp64'}|2^{bucket}"


def load_cache(path) -> dict:
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_cache(path, cache: dict):
    if path:
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
