"""Every launch-parameter candidate of the autotuner gives the default's results: bit-identical
sorted order; FP64 state identical (to the segment-atomics rounding); FP32C derivatives within
1e-6 norm-wise (the FP32 summation blocks follow the list chunks), on the cutting case and on the
10:1-density chip domain where the neighbour lists overflow their chunks."""
import numpy as np
import pytest

from helpers import assert_equal
from sph_fem_2d_cutting_gpu_amd import _lib, autotune, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu

DERIVS = ("pos_t", "vel_t", "rho_t", "S_t", "T_t")


def _run(case, prec, tuning, steps):
    sim = S.Simulation(case, precision=prec)
    sim.set_tuning(**tuning)
    sim.step(steps)
    sim.ctx.check_error()
    out = sim.state(("hash", "idx") + DERIVS + ("pos", "vel", "S", "T"))
    sim.close()
    return out


@pytest.mark.parametrize("prec", [_lib.FP64, _lib.FP32C])
@pytest.mark.parametrize("which", ["cutting", "chip"])
def test_all_candidates_agree(prec, which):
    if which == "cutting":
        case, steps = cases.perturbed(cases.small_cutting_case(nx=48, ny=16, wear=True)), 3
    else:
        case, steps = cases.chip_case(n_target=50_000), 1      # 51,570 particles: a partial last wavefront
    ref = _run(case, prec, {"pair_block": 256, "pair_chunk": 32, "update_block": 256}, steps)
    tol = 1e-11 if prec == _lib.FP64 else 1e-6
    for cand in autotune.candidates(prec) + [dict(autotune.candidates(prec)[0], resort_one_block=1)]:
        got = _run(case, prec, cand, steps)
        assert_equal(f"{cand} hash", got["hash"], ref["hash"])
        assert_equal(f"{cand} idx", got["idx"], ref["idx"])
        for f in DERIVS + ("pos", "vel", "S", "T"):
            scale = max(float(np.max(np.abs(ref[f]))), 1e-300)
            err = float(np.max(np.abs(got[f] - ref[f]))) / scale
            assert err <= tol, f"{cand} {f}: {err:.3e}"


@pytest.mark.parametrize("which", ["cutting", "cutting_rothlin", "chip50k", "chip1m"])
def test_staged_pair_kernel_equals_list_kernel(which):
    """The LDS-staged FP32C pair kernel against the list kernel at each list chunk: the same sums bit
    for bit, on cases where most wavefronts stage their windows (the chip domain, 92 % at 1M) and on
    a cutting case where most run the list path (64 particles span several short columns), on the
    Green and the Rothlin (dense) cell table, with partial last wavefronts (N mod 64 = 0, 50, 16).
    Both fused forms (each particle's update run by the pair kernel) give the same step bit for bit."""
    grid = "green"
    if which.startswith("cutting"):
        case, steps = cases.perturbed(cases.small_cutting_case(nx=160, ny=40, wear=True)), 3
        if which == "cutting_rothlin":           # dense cell table: empty cells hold offsets, not -1
            grid = "rothlin"
    elif which == "chip50k":
        case, steps = cases.chip_case(n_target=50_000), 2
    else:
        case, steps = cases.chip_case(n_target=1_000_000), 1
    print(f"staged vs list: {which}, N = {case.N}, N mod 64 = {case.N % 64}", flush=True)
    fields = ("hash", "idx", "blanked") + DERIVS + ("pos", "vel", "S", "rho", "T")
    for chunk in autotune.PAIR_CHUNKS:
        out = {}
        for kernel in (_lib.PAIRS_LIST, _lib.PAIRS_STAGED, _lib.PAIRS_LIST_FUSED, _lib.PAIRS_STAGED_FUSED):
            sim = S.Simulation(case, precision=_lib.FP32C, grid=grid)
            sim.set_tuning(pair_block=256, pair_chunk=chunk, update_block=256, pair_kernel=kernel)
            sim.step(steps)
            sim.ctx.check_error()          # also the staged kernel's containment check
            out[kernel] = sim.state(fields)
            sim.close()
        for kernel in (_lib.PAIRS_STAGED, _lib.PAIRS_LIST_FUSED, _lib.PAIRS_STAGED_FUSED):
            for f in fields:
                assert_equal(f"chunk {chunk} kernel {kernel} {f}", out[kernel][f], out[_lib.PAIRS_LIST][f])


def test_tune_simulation_applies_a_candidate():
    sim = S.Simulation(cases.small_cutting_case(nx=32, ny=12), precision=_lib.FP32C)
    sim.step(1)
    res = autotune.tune_simulation(sim, _lib.FP32C, steps=2, repeats=1)
    assert dict(res["best"], resort_one_block=0) in autotune.candidates(_lib.FP32C)
    assert sim.get_tuning() == res["best"]
    assert len(res["timings"]) == len(autotune.candidates(_lib.FP32C)) + 1     # + the one-block fallback
    sim.close()


def test_invalid_tuning_is_rejected():
    sim = S.Simulation(cases.small_cutting_case(nx=16, ny=8))
    with pytest.raises(_lib.SphfemError):
        sim.set_tuning(pair_block=96)
    sim.close()
