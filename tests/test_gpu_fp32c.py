"""FP32-compensated step (SPHFEM_FP32C) against the FP64 oracle and the FP64 GPU step.

The FP32C form changes only the arithmetic of the pair sums (FP32 terms on cell-relative
positions, block sums accumulated in FP64); the cell list, the corrector, plasticity, contact
and the tool are the FP64 code.  So:
* integer outputs (hash, sorted index, blanked) are identical to the oracle's;
* the derivatives of one step from a perturbed state (every pair term non-zero) match the
  oracle within 1e-5 of each field's largest magnitude (norm-wise relative error);
* on a chip-shaped domain with a 10:1 density range (several neighbour chunks per particle)
  the same bound holds against the FP64 GPU step;
* after 2000 steps of the 6,400-particle case the cutting force is within 0.5 % of the
  oracle's running maximum and the peak temperature rise within 0.5 %.
"""
import numpy as np
import pytest

from helpers import assert_equal
from oracle.oracle import Oracle
from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu

DERIVS = ("pos_t", "vel_t", "rho_t", "S_t", "T_t")
STATE = ("pos", "vel", "S", "rho", "T", "eps_pl", "fc", "ft")
TOL_DERIV = 1e-5     # FP32C pair sums: norm-wise relative error of one step's derivatives
TOL_STATE = 1e-6     # the state moves by dt * derivative from an FP64 start


def normwise(a, b, live):
    a = np.asarray(a)[live].astype(np.float64)
    b = np.asarray(b)[live].astype(np.float64)
    scale = np.max(np.abs(b))
    if scale == 0.0:
        return float(np.max(np.abs(a)))
    return float(np.max(np.abs(a - b)) / scale)


def _check(a, b, fields, tol, label):
    live = b["blanked"] != 1.0
    errs = {f: normwise(a[f], b[f], live) for f in fields}
    print(label, {k: f"{v:.2e}" for k, v in errs.items()})
    bad = {k: v for k, v in errs.items() if v > tol}
    assert not bad, f"{label}: above {tol:.0e}: {bad}"


def test_fp32c_first_step_vs_oracle():
    case = cases.perturbed(cases.small_cutting_case(nx=48, ny=16, wear=True))
    o = Oracle(case)
    sim = S.Simulation(case, precision=_lib.FP32C)
    o.step(1)
    sim.step(1)
    sim.ctx.check_error()
    a, b = sim.state(), o.P
    for f in ("hash", "idx", "blanked"):
        assert_equal(f, a[f], b[f])
    for f in DERIVS:
        assert np.max(np.abs(b[f])) > 0.0, f"{f} is identically zero: the test state is too trivial"
    _check(a, b, DERIVS, TOL_DERIV, "fp32c vs oracle, derivatives")
    _check(a, b, STATE, TOL_STATE, "fp32c vs oracle, state")


def test_fp32c_several_steps_vs_oracle():
    case = cases.perturbed(cases.small_cutting_case(nx=48, ny=16, wear=True))
    o = Oracle(case)
    sim = S.Simulation(case, precision=_lib.FP32C)
    o.step(10)
    sim.step(10)
    sim.ctx.check_error()
    a, b = sim.state(), o.P
    assert_equal("hash", a["hash"], b["hash"])
    _check(a, b, DERIVS, 1e-4, "fp32c vs oracle, 10 steps, derivatives")
    _check(a, b, ("pos", "vel", "S", "rho", "T"), 1e-6, "fp32c vs oracle, 10 steps, state")


def test_fp32c_matches_fp64_on_chip_density():
    """Chip-shaped domain, 10:1 number density: dense particles have several neighbour chunks."""
    case = cases.chip_case(n_target=200_000)
    a_sim = S.Simulation(case, precision=_lib.FP32C)
    b_sim = S.Simulation(case, precision=_lib.FP64)
    a_sim.step(1)
    b_sim.step(1)
    a_sim.ctx.check_error()
    b_sim.ctx.check_error()
    a, b = a_sim.state(), b_sim.state()
    for f in ("hash", "idx", "blanked"):
        assert_equal(f, a[f], b[f])
    _check(a, b, DERIVS, TOL_DERIV, "fp32c vs fp64, chip domain")


def test_fp32c_observables_2000_steps():
    """6,400-particle Ti6Al4V case: cutting force and peak temperature after 2000 steps."""
    case = cases.config1()
    o = Oracle(case)
    sim = S.Simulation(case, precision=_lib.FP32C)
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
    print(f"fp32c force error {ferr}, dT {dT_g:.6g} vs {dT_o:.6g}")
    assert np.all(ferr <= 5e-3), ferr
    assert abs(dT_g - dT_o) <= 5e-3 * dT_o


@pytest.mark.parametrize("prec", [_lib.FP64, _lib.FP32C])
def test_plastic_observables_2000_steps(prec):
    """6,400-particle case started from a random deviatoric stress near yield (600 MPa per component):
    plastic flow and plastic-work heating act over the whole 2000-step horizon (peak dT ~80 K, max
    eps_pl ~0.04).  Cutting force within 0.5 % of the oracle's running maximum; peak temperature rise,
    max and mean plastic strain within 0.5 %."""
    case = cases.perturbed(cases.config1(), vel_sd=1e-5, S_sd=6e-3, rho_rel_sd=1e-4, T_sd=20.0)
    o = Oracle(case)
    sim = S.Simulation(case, precision=prec)
    F_o, F_g = [], []
    for k in range(100):
        o.step(20)
        sim.step(20)
        F_o.append(o.report_force())
        F_g.append(S.report_force(sim.p))
    sim.ctx.check_error()
    F_o, F_g = np.array(F_o), np.array(F_g)
    ferr = np.max(np.abs(F_g - F_o), axis=0) / np.max(np.abs(F_o), axis=0)
    a = sim.state(("T", "eps_pl", "blanked"))
    T0 = case.const.joco["Tref"]
    dT_o, dT_g = o.P["T"].max() - T0, a["T"].max() - T0
    print(f"force error {ferr}, dT {dT_g:.8g} vs {dT_o:.8g}, eps max {a['eps_pl'].max():.8g} vs "
          f"{o.P['eps_pl'].max():.8g}")
    assert dT_o > 10.0 and o.P["eps_pl"].max() > 0.01          # the horizon is plastic
    assert np.all(ferr <= 5e-3), ferr
    assert abs(dT_g - dT_o) <= 5e-3 * dT_o
    assert abs(a["eps_pl"].max() - o.P["eps_pl"].max()) <= 5e-3 * o.P["eps_pl"].max()
    assert abs(a["eps_pl"].mean() - o.P["eps_pl"].mean()) <= 5e-3 * o.P["eps_pl"].mean()
