"""Incremental re-sort of the step against the full stable radix sort and the oracle.

After a step the particles are stored in the previous sorted order; the step re-sorts the few
particles that changed cells and merges them in (k_resort_*), falling back to the full radix
sort on the first step or when more than the limit changed.  Both paths must give the stable
sort exactly: hash and sorted index are compared bit for bit.
"""
import numpy as np
import pytest

from helpers import assert_equal
from oracle.oracle import Oracle
from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu


def fast_case():
    """Perturbed 768-particle cutting case with fast particles: ~1-2 % change cells per step."""
    return cases.perturbed(cases.small_cutting_case(nx=48, ny=16, wear=True), vel_sd=0.05)


@pytest.mark.parametrize("one_block", [0, 1])
@pytest.mark.parametrize("limit,expect_mode", [(4096, 2), (0, 1), (2, 1)])
def test_resort_paths_match_oracle(limit, expect_mode, one_block):
    """one_block: the full-sort fallback by one workgroup instead of the multi-block radix sort."""
    case = fast_case()
    o = Oracle(case)
    sim = S.Simulation(case)
    sim.set_tuning(resort_one_block=one_block)
    sim.grid.set_resort_limit(limit)
    modes, changed = [], []
    for k in range(10):
        o.step(1)
        sim.step(1)
        m, c = sim.grid.resort_stats()
        modes.append(m)
        changed.append(c)
        a = sim.state(("hash", "idx"))
        assert_equal(f"hash after step {k + 1}", a["hash"], o.P["hash"])
        assert_equal(f"idx after step {k + 1}", a["idx"], o.P["idx"])
    sim.ctx.check_error()
    print("limit", limit, "modes", modes, "changed", changed)
    assert modes[0] == 1                                   # first step: full sort
    assert all(m == expect_mode for m in modes[2:])
    assert sum(changed[1:]) > 10                          # the case does move particles across cells


def test_incremental_equals_full_on_200k_with_thousands_of_changes():
    """Chip domain (10:1 density), velocities scaled so that ~1 % of 200k particles change cells
    per step: the incremental path (m in the thousands, several rank iterations) against the
    full sort; no tool, so the whole state is bit-identical."""
    base = cases.chip_case(n_target=200_000)
    vscale = 0.004 * base.h / base.dt / max(float(np.sqrt(np.mean(base.vel ** 2))), 1e-30)
    import dataclasses
    case = dataclasses.replace(base, vel=base.vel * vscale)
    inc = S.Simulation(case, precision=_lib.FP32C)
    full = S.Simulation(case, precision=_lib.FP32C)
    full.grid.set_resort_limit(0)
    one = S.Simulation(case, precision=_lib.FP32C)      # the full sort by one workgroup every step
    one.set_tuning(resort_one_block=1)
    one.grid.set_resort_limit(0)
    seen = []
    for k in range(6):
        inc.step(1)
        full.step(1)
        one.step(1)
        seen.append(inc.grid.resort_stats())
        assert one.grid.resort_stats()[0] == 1
    for s_ in (inc, full, one):
        s_.ctx.check_error()
    print("incremental (mode, changed):", seen)
    assert all(m == 2 for m, _ in seen[1:]), seen
    assert max(c for _, c in seen[1:]) > 500, seen
    fields = ("hash", "idx", "pos", "vel", "S", "rho", "T")
    a, b, c = inc.state(fields), full.state(fields), one.state(fields)
    for f in fields:
        assert_equal(f, a[f], b[f])
        assert_equal(f"{f} (one-workgroup sort)", c[f], b[f])
