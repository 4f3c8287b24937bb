"""Bitwise reproducibility of the step: two simulations of the same case in one process, 100 steps
each, agree in every particle field and every tool field bit for bit (FP64 and FP32C forms).

What used to vary between runs were the tool's segment sums (friction heat, Usui inputs, sliding
force, contact heat exchange), accumulated with atomics in whatever order the lanes arrived.  They
are now summed per segment in increasing particle index (the reference's loop order,
actions_gpu.cu:398-418, 607-608), so the order in which particles are listed no longer matters."""
import numpy as np
import pytest

from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        return np.array_equal(a.view(np.uint64 if a.dtype == np.float64 else np.uint32),
                              b.view(np.uint64 if b.dtype == np.float64 else np.uint32))
    return np.array_equal(a, b)


@pytest.mark.parametrize("prec", [_lib.FP64, _lib.FP32C])
def test_two_runs_are_bitwise_equal(prec):
    case = cases.perturbed(cases.small_cutting_case(nx=64, ny=20, wear=True), vel_sd=2e-4)
    runs = []
    for r in range(2):
        sim = S.Simulation(case, precision=prec)
        for k in range(10):
            sim.step(10)
        sim.ctx.check_error()
        runs.append((sim.state(), sim.tool_state()))
        sim.close()
    (pa, ta), (pb, tb) = runs
    assert (pa["on_seg"] >= 0).sum() > 0, "no contact: the segment sums would not be exercised"
    # accumulated during the last step (friction heat is consumed and cleared within the step)
    assert ta["heat_exchange"][:, 1].max() > 0 and ta["physical_para"][:, 0].max() > 0
    bad = [f for f in pa if not _same(pa[f], pb[f])]
    assert not bad, f"particle fields differ between runs: {bad}"
    bad = [f for f in ta if isinstance(ta[f], np.ndarray) and not _same(ta[f], tb[f])]
    assert not bad, f"tool fields differ between runs: {bad}"
