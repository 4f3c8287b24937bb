"""Batched stepping of independent cases (solver.BatchedSimulation): every case of a concurrent
batch (one context, stream and captured graph per case) must equal the same case stepped alone,
and a batch case must follow the oracle."""
import numpy as np
import pytest

from helpers import assert_equal
from oracle.oracle import Oracle
from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu

FIELDS = ("hash", "idx", "pos", "vel", "S", "rho", "T", "eps_pl", "fc", "ft")


def small_batch(n=6):
    return cases.batch_cases(n_cases=n, rows=(16, 24), tool_tris=400)


@pytest.mark.parametrize("prec", [_lib.FP64, _lib.FP32C])
def test_batch_matches_cases_stepped_alone(prec):
    cs = small_batch()
    batch = S.BatchedSimulation(cs, precision=prec)
    alone = [S.Simulation(c, precision=prec) for c in cs]
    batch.step(1)
    batch.capture()
    batch.launch(3)                     # 1 + 6 steps, all cases in flight together
    batch.synchronize()
    batch.check_error()
    for sim in alone:
        sim.step(7)
        sim.ctx.check_error()
    for c, bs, al in zip(cs, batch.sims, alone):
        a, b = bs.state(FIELDS), al.state(FIELDS)
        assert_equal(c.name + " hash", a["hash"], b["hash"])
        assert_equal(c.name + " idx", a["idx"], b["idx"])
        for f in FIELDS[2:]:
            scale = max(float(np.max(np.abs(b[f]))), 1e-300)
            err = float(np.max(np.abs(a[f] - b[f]))) / scale
            assert err <= 1e-11, f"{c.name} {f}: {err:.3e}"
    batch.close()


def test_batch_case_follows_oracle():
    cs = small_batch(3)
    c = min(cs, key=lambda x: x.N)
    batch = S.BatchedSimulation(cs)
    o = Oracle(c)
    batch.step(1)
    batch.capture()
    batch.launch(2)
    batch.synchronize()
    o.step(5)
    sim = batch.sims[cs.index(c)]
    a = sim.state(("hash", "idx", "pos", "vel", "T"))
    assert_equal("hash", a["hash"], o.P["hash"])
    assert_equal("idx", a["idx"], o.P["idx"])
    live = o.P["blanked"] != 1.0
    for f in ("pos", "vel", "T"):
        scale = max(float(np.max(np.abs(o.P[f][live]))), 1e-300)
        err = float(np.max(np.abs(a[f][live] - o.P[f][live]))) / scale
        assert err <= 1e-6, f"{f}: {err:.3e}"
    batch.close()
