"""Parity of the fused FP64 step (sphfem_leap_frog_step) with the reference step.

Three comparisons:
* fused step vs the same step assembled from the drop-in entry points in the order of
  leap_frog.cu:162-248 (both on the GPU): identical arithmetic, so the only admissible
  differences come from the order of the atomic segment sums;
* fused step vs the FP64 oracle over several steps (libm-level differences compound);
* the cutting-force and peak-temperature observables over 2000 steps of the 6,400-particle
  case, within 0.5 % of the oracle (force relative to its running maximum, temperature as
  the rise above T0).
"""
import ctypes as C

import numpy as np
import pytest

from helpers import assert_close, assert_equal, max_rel_err
from oracle.oracle import Oracle
from sph_fem_2d_cutting_gpu_amd import cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "S", "rho", "T", "eps_pl", "eps_pl_dot", "fc", "ft", "n", "p", "R")
DERIVS = ("pos_t", "vel_t", "rho_t", "S_t", "T_t")


class Unfused:
    """leap_frog::step(particles, grid, tool_FEM) written with the drop-in entry points."""

    def __init__(self, sim):
        self.sim = sim
        nc = sim.grid.num_cell
        self.cs, self.ce = C.c_void_p(), C.c_void_p()
        sim.ctx.call("sphfem_malloc", C.byref(self.cs), 4 * nc)
        sim.ctx.call("sphfem_malloc", C.byref(self.ce), 4 * nc)

    def step(self):
        sim = self.sim
        p, g, t, lf = sim.p, sim.grid, sim.tool, sim.stepper
        dt = sim.case.dt
        g.assign_hashes(p)
        g.sort(p)
        g.get_cells(p, self.cs.value, self.ce.value)
        lf.init(p)
        t.reset_FEM_server()
        t.segments_wear_rate_reset()
        lf.predict(p, dt)
        t.calculate_temp(0.5 * dt)
        S.material_eos(p)
        S.corrector_artificial_stress(p)
        S.interactions_setup_geometry_constants(g)
        S.interactions_monaghan(p, self.cs.value, self.ce.value, g.num_cell)
        S.material_stress_rate_jaumann(p)
        S.contmech_continuity(p)
        S.contmech_momentum(p)
        S.contmech_advection(p)
        lf.correct(p, dt)
        S.material_fric_heat_gen_FEM_v2(p, t)
        t.apply_heat_boundary_condition()
        t.calculate_temp_rate()
        t.calculate_temp(dt)
        S.plasticity_johnson_cook(p)
        if sim.case.wear:
            S.wear_adjustment(t)
        t.segments_update(dt)
        t.tool_update_FEM_gpu(dt)
        S.compute_contact_forces_FEM_tool_v3(p, t)
        S.perform_boundary_conditions(p)
        S.perform_boundary_conditions_thermal(p)
        S.debug_invalidate(p)


def _compare_states(a, b, live, rtol, fields, label):
    errs = {}
    for f in fields:
        errs[f] = max_rel_err(a[f], b[f], live)
    print(label, {k: f"{v:.2e}" for k, v in errs.items()})
    bad = {k: v for k, v in errs.items() if v > rtol}
    assert not bad, f"{label}: fields above {rtol:.0e}: {bad}"


@pytest.mark.parametrize("wear", [False, True])
def test_fused_matches_unfused(wear):
    case = cases.small_cutting_case(nx=48, ny=16, wear=wear)
    fused = S.Simulation(case)
    ref = S.Simulation(case)
    unf = Unfused(ref)
    for k in range(40):
        fused.step(1)
        unf.step()
    fused.ctx.check_error()
    ref.ctx.check_error()
    a, b = fused.state(), ref.state()
    for f in ("hash", "idx", "on_seg", "blanked"):
        assert_equal(f, a[f], b[f])
    live = b["blanked"] != 1.0
    _compare_states(a, b, live, 1e-11, STATE + DERIVS, f"fused vs unfused (wear={wear})")
    ta, tb = fused.tool_state(), ref.tool_state()
    # heat_exchange: the contact heat-exchange sums of the last step (contact s2, which the fused
    # step runs in the near-tool pass and the unfused step over all particles)
    for f in ("T", "node_pos", "left", "right", "K", "M", "heat_exchange"):
        assert_close("tool " + f, ta[f], tb[f], 1e-11)


@pytest.mark.parametrize("nsteps", [1, 5, 25])
def test_fused_matches_oracle(nsteps):
    case = cases.small_cutting_case(nx=48, ny=16, wear=True)
    o = Oracle(case)
    sim = S.Simulation(case)
    o.step(nsteps)
    sim.step(nsteps)
    sim.ctx.check_error()
    a = sim.state()
    b = o.P
    for f in ("hash", "idx", "blanked"):
        assert_equal(f, a[f], b[f])
    live = b["blanked"] != 1.0
    # step 1 starts from S = 0, so the artificial stress is zero and everything is exact
    tol = 1e-13 if nsteps == 1 else 1e-6
    _compare_states(a, b, live, tol, STATE + DERIVS, f"fused vs oracle ({nsteps} steps)")
    frac = np.mean(a["on_seg"][live] == b["on_seg"][live])
    assert frac >= (1.0 if nsteps == 1 else 0.999), f"on_seg agreement {frac}"
    tg = sim.tool_state()
    for f in ("T", "node_pos", "left", "right", "n", "K", "M"):
        assert_close("tool " + f, tg[f], o.t[f], tol)
    if nsteps == 1:     # the contact set is the oracle's exactly after one step
        assert_close("tool heat_exchange", tg["heat_exchange"], o.t["heat_exchange"], tol)


def test_observables_2000_steps():
    """6,400-particle Ti6Al4V case: cutting force and peak temperature after 2000 steps."""
    case = cases.config1()
    o = Oracle(case)
    sim = S.Simulation(case)
    F_o, F_g = [], []
    for k in range(100):
        o.step(20)
        sim.step(20)
        F_o.append(o.report_force())
        F_g.append(S.report_force(sim.p))
    sim.ctx.check_error()
    F_o, F_g = np.array(F_o), np.array(F_g)
    scale = np.max(np.abs(F_o), axis=0)
    ferr = np.max(np.abs(F_g - F_o), axis=0) / scale
    a = sim.state(("T", "eps_pl", "blanked"))
    T0 = case.const.joco["Tref"]
    dT_o = o.P["T"].max() - T0
    dT_g = a["T"].max() - T0
    eps_o, eps_g = o.P["eps_pl"].max(), a["eps_pl"].max()
    print(f"force error {ferr}, dT {dT_g:.6g} vs {dT_o:.6g}, eps {eps_g:.6g} vs {eps_o:.6g}")
    assert np.all(ferr <= 5e-3), ferr
    assert abs(dT_g - dT_o) <= 5e-3 * dT_o
    assert abs(eps_g - eps_o) <= 5e-3 * max(eps_o, 1e-12)
