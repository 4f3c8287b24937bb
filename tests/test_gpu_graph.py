"""Captured step graphs (sphfem_leap_frog_graph_*) against eager stepping.

The graph records two steps (the array-pointer swaps of a step return to the start after two).
Every replay is checked against the same number of eager steps, with eager work (force
reduction, downloads, eager steps on the graph's own state) and synchronisations between
replays.  Sorted order must agree exactly; the state to the rounding of the order-dependent
segment atomics.
"""
import numpy as np
import pytest

from helpers import assert_equal
from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu


def _compare(a_sim, b_sim, label):
    a = a_sim.state(("hash", "idx", "pos", "vel", "S", "rho", "T", "eps_pl", "fc", "ft"))
    b = b_sim.state(("hash", "idx", "pos", "vel", "S", "rho", "T", "eps_pl", "fc", "ft"))
    assert_equal(label + " hash", a["hash"], b["hash"])
    assert_equal(label + " idx", a["idx"], b["idx"])
    for f in ("pos", "vel", "S", "rho", "T", "eps_pl", "fc", "ft"):
        scale = max(float(np.max(np.abs(b[f]))), 1e-300)
        err = float(np.max(np.abs(a[f] - b[f]))) / scale
        assert err <= 1e-11, f"{label} {f}: {err:.3e}"


@pytest.mark.parametrize("prec", [_lib.FP64, _lib.FP32C])
def test_graph_replays_match_eager(prec):
    case = cases.perturbed(cases.small_cutting_case(nx=48, ny=16, wear=True))
    eager = S.Simulation(case, precision=prec)
    gsim = S.Simulation(case, precision=prec)
    eager.step(1)
    gsim.step(1)
    graph = gsim.capture(steps=2)
    for r in range(4):
        graph.launch(1)
        eager.step(2)
        fg = S.report_force(gsim.p)                    # eager work + synchronisation between replays
        fe = S.report_force(eager.p)
        assert np.allclose(fg, fe, rtol=1e-11, atol=0.0)
        _compare(gsim, eager, f"replay {r}")
        if r == 1:                                     # eager steps on the graph's own state
            gsim.step(2)
            eager.step(2)
    graph.launch(3)                                    # back-to-back replays
    eager.step(6)
    _compare(gsim, eager, "3 replays")
    gsim.ctx.check_error()
    graph.destroy()


def test_graph_needs_an_eager_step_and_even_steps():
    case = cases.small_cutting_case(nx=24, ny=8)
    sim = S.Simulation(case)
    with pytest.raises(_lib.SphfemError):
        sim.capture(steps=2)                           # no eager step yet (first sort)
    sim.step(1)
    with pytest.raises(_lib.SphfemError):
        sim.capture(steps=3)
    g = sim.capture(steps=4)
    g.launch(1)
    sim.ctx.check_error()
    g.destroy()


def test_graph_refused_after_settings_change():
    """A graph bakes in the constants, dt, wear switch, precision and tuning: after any of them changes
    its launch is refused (the reference driver turns wear on mid-run, mfree_gpu.cpp:306); setting the
    captured values back makes it valid again, and a new capture follows the new settings."""
    case = cases.small_cutting_case(nx=24, ny=8, wear=True)
    sim = S.Simulation(case, wear=False)
    sim.step(1)
    g = sim.capture(steps=2)
    g.launch(1)
    sim.ctx.call("sphfem_set_wear", 1)
    with pytest.raises(_lib.SphfemError, match="captured"):
        g.launch(1)
    sim.ctx.call("sphfem_set_wear", 0)
    g.launch(1)
    sim.ctx.call("sphfem_set_dt", case.dt * 0.5)
    with pytest.raises(_lib.SphfemError, match="captured"):
        g.launch(1)
    sim.ctx.call("sphfem_set_dt", case.dt)
    sim.set_tuning(update_block=128 if sim.get_tuning()["update_block"] == 256 else 256)
    with pytest.raises(_lib.SphfemError, match="captured"):
        g.launch(1)
    g.destroy()
    sim.ctx.call("sphfem_set_wear", 1)
    g2 = sim.capture(steps=2)                          # wear on from here
    g2.launch(2)
    sim.ctx.check_error()
    g2.destroy()
