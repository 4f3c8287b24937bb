"""Failure and blanking branches on the GPU against the oracle, from forced states.

The branches (actions_gpu.cu:375-380 NaN relative velocity, :479-491 deep penetration, :750-753
melt, :787-824 non-finite rates, interactions_gpu.cu:152/232 blanked particles left out of the
neighbour sums) fire only in states a healthy run rarely reaches, so each test builds such a state
on the host and runs the drop-in entry point or the whole step from it:
* per-stage: the GPU entry point must reproduce the oracle's stage bit for bit (blanked flags,
  clamped temperatures, zeroed contact fields) -- the stage arithmetic is the same IEEE operations;
* whole step, FP64 and FP32C: a particle with a NaN stress (its neighbours' sums turn NaN and they
  are blanked with it), a particle above the melting temperature with a small stress, and a particle
  3 um behind the rake face; after one and three steps the blanked set equals the oracle's and the
  live particles agree to the tolerances of the healthy-state tests.
Positions stay finite in every case, so every hash stays inside the grid.
"""
import dataclasses

import numpy as np
import pytest

from helpers import assert_close, assert_equal
from oracle.oracle import Oracle
from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S
from test_gpu_stages import _run_unfused_from

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    case = cases.small_cutting_case(nx=48, ny=16, wear=True)
    o = Oracle(case)
    o.step(30)
    sim = S.Simulation(case)
    return case, o, sim


def _stage(world, stage, force):
    case, o0, sim = world
    o = o0.clone()
    force(o)
    print(f"stage {stage} from a forced state", flush=True)
    _run_unfused_from(sim, o, stage)
    return sim, o


def test_invalidate_each_non_finite_rate(world):
    """NaN, +Inf and -Inf in each of the ten rate components of its own particle."""
    comps = [("pos_t", 0), ("pos_t", 1), ("vel_t", 0), ("vel_t", 1), ("S_t", 0), ("S_t", 1), ("S_t", 2),
             ("S_t", 3), ("rho_t", None), ("T_t", None)]

    def force(o):
        i = 0
        for f, c in comps:
            for v in (np.nan, np.inf, -np.inf):
                k = 40 + 17 * i
                if c is None:
                    o.P[f][k] = v
                else:
                    o.P[f][k, c] = v
                i += 1
        o.P["vel_t"][5] = 1e300              # large but finite

    sim, o = _stage(world, "invalidate", force)
    assert o.P["blanked"].sum() >= 30
    assert_equal("blanked", sim.p.download("blanked"), o.P["blanked"])


def test_plasticity_melt_branch(world):
    """Particles above Tmelt with a stress: the return runs with sigma_Y = 0, T is set to Tmelt and the
    particle is blanked; particles just below stay live."""
    case, _, _ = world
    Tm = case.const.joco["Tmelt"]

    def force(o):
        o.P["T"][::5] = Tm + 10.0
        o.P["T"][1::5] = Tm - 1e-3
        o.P["S"][::5, 0] += 1e-4
        o.P["S"][::5, 2] -= 1e-4

    sim, o = _stage(world, "plasticity", force)
    melted = o.P["blanked"] == 1.0
    assert melted.sum() >= 100 and np.all(o.P["T"][melted] == Tm)
    assert_equal("blanked", sim.p.download("blanked"), o.P["blanked"])
    assert_equal("T (melted)", sim.p.download("T")[melted], o.P["T"][melted])
    live = ~melted
    for f in ("S", "eps_pl", "eps_pl_dot", "T"):
        assert_close(f, sim.p.download(f), o.P[f], 1e-10, mask=live)


def test_contact_deep_penetration_branch(world):
    """Particles moved 2.6-4 um behind the rake face are blanked with zero forces, normal and no segment;
    one at 2.4 um gets the penalty force."""
    case, o0, _ = world

    def force(o):
        t = o.t
        rake = [s for s in range(len(t["n"])) if t["n"][s][0] < -0.99][2:8]
        for j, (s, d) in enumerate(zip(rake, (2.4e-4, 2.6e-4, 3e-4, 3.5e-4, 4e-4, 2.45e-4))):
            k = 100 + 9 * j
            xr = t["left"][s][0]
            o.P["pos"][k] = [xr - d, 0.5 * (t["left"][s][1] + t["right"][s][1])]

    sim, o = _stage(world, "contact", force)
    assert o.P["blanked"].sum() >= 4
    for f in ("fc", "ft", "n", "on_seg", "blanked", "T"):
        assert_equal(f, sim.p.download(f), o.P[f])


def test_fric_heat_nan_velocity_branch(world):
    """Sixteen particles put in contact with rake segments, every other one with a NaN velocity: those
    are blanked and add nothing to the segment sums; the others heat the particle and the segment."""
    def force(o):
        t = o.t
        o.P["on_seg"][:] = -1
        rake = [s for s in range(len(t["n"])) if t["n"][s][0] < -0.99][:8]
        cont = 200 + 7 * np.arange(16)
        o.P["on_seg"][cont] = np.repeat(rake, 2)
        o.P["n"][cont] = [1.0, 0.0]
        o.P["ft"][cont] = [0.0, -2e-9]
        o.P["fc"][cont] = [5e-9, 0.0]
        o.P["vel"][cont] = [1e-4, 2e-3]
        o.P["vel"][cont[::2], 1] = np.nan

    sim, o = _stage(world, "fric_heat", force)
    assert o.P["blanked"].sum() >= 2
    assert_equal("blanked", sim.p.download("blanked"), o.P["blanked"])
    assert_equal("T", sim.p.download("T"), o.P["T"])
    tg = sim.tool.download()
    for f in ("fric_heat", "physical_para", "sliding_force"):
        assert_equal(f, tg[f], o.t[f])


def test_interactions_skip_blanked_neighbours(world):
    """Every fifth particle blanked (and given a NaN stress, which must not leak into any sum)."""
    def force(o):
        o.P["blanked"][3::5] = 1.0
        o.P["S"][3::5] = np.nan

    sim, o = _stage(world, "interactions", force)
    live = o.P["blanked"] != 1.0
    for f in ("S_der", "v_der", "pos_t", "vel_t", "T_t"):
        g = sim.p.download(f)
        assert np.all(np.isfinite(g[live])), f
        assert_close(f, g, o.P[f], 1e-15, mask=live)


def _forced_case():
    """The 768-particle cutting case with three forced particles: a NaN stress far from the tool, one
    above Tmelt with a small stress, one 3 um behind the rake face.  Tmelt is lowered to 1000 K: in the
    first step every particle still has the constructor's on_seg = 0, so the friction-heat pass caps
    every temperature at 1550 K (actions_gpu.cu:397-400) before plasticity could see a melt."""
    case = cases.small_cutting_case(nx=48, ny=16, wear=True)
    joco = dict(case.const.joco, Tmelt=1000.0)
    case = dataclasses.replace(case, const=dataclasses.replace(case.const, joco=joco))
    N = case.N
    pos, T = case.pos.copy(), case.T.copy()
    Sd = np.zeros((N, 4))
    free = case.fixed == 0.0

    def nearest(x, y):
        d = np.hypot(pos[:, 0] - x, pos[:, 1] - y)
        d[~free] = np.inf
        return int(np.argmin(d))

    k_nan = nearest(0.0140, 0.0590)
    Sd[k_nan] = [np.nan, 0.0, 0.0, 0.0]
    k_melt = nearest(0.0110, 0.0592)
    T[k_melt] = case.const.joco["Tmelt"] + 10.0
    Sd[k_melt] = [1e-4, 2e-5, -1e-4, 0.0]
    tm = case.tool
    rake_x = tm.seg_left[0][0]
    k_pen = nearest(0.0120, 0.0580)
    pos[k_pen] = [rake_x - 3e-4, tm.br[1] + 0.35 * (tm.tl[1] - tm.br[1])]
    forced = dataclasses.replace(case, pos=pos, T=T, S=Sd)
    return forced, (k_nan, k_melt, k_pen)


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_step_from_forced_state(precision):
    case, ks = _forced_case()
    o = Oracle(case)
    prec = _lib.FP32C if precision == "fp32c" else _lib.FP64
    sim = S.Simulation(case, precision=prec)
    done = 0
    for n in (1, 2):
        print(f"forced-state step, {precision}, steps {done + 1}..{done + n}", flush=True)
        o.step(n)
        sim.step(n)
        done += n
        sim.ctx.check_error()
        a, b = sim.state(), o.P
        for f in ("hash", "idx", "blanked"):
            assert_equal(f"{f} after {done} steps", a[f], b[f])
        live = b["blanked"] != 1.0
        tol = 1e-6
        for f in ("pos", "vel", "S", "rho", "T", "eps_pl", "fc", "ft"):
            assert_close(f"{f} after {done} steps", a[f], b[f], tol, mask=live)
        for f in ("pos_t", "vel_t", "rho_t", "S_t", "T_t"):
            assert_close(f"{f} after {done} steps", a[f], b[f], 1e-5 if precision == "fp32c" else tol, mask=live)
        if done == 1:
            _assert_forced_branches_fired(case, b)


def test_fused_pair_kernels_equal_unfused_from_forced_state():
    """The fused FP32C pair kernels (each particle's update run right after its sums) against the
    separate pair and update kernels, from the forced state: the first step runs the window form
    (a live particle with a NaN stress), the later ones the in-support form again once the step
    start has cleared the switch; every field equal bit for bit after each step."""
    case, _ = _forced_case()
    fields = ("hash", "idx", "blanked", "pos", "vel", "S", "rho", "T", "eps_pl", "fc", "ft",
              "pos_t", "vel_t", "rho_t", "S_t", "T_t")
    sims = {}
    for kernel in (_lib.PAIRS_LIST, _lib.PAIRS_LIST_FUSED, _lib.PAIRS_STAGED_FUSED):
        sims[kernel] = S.Simulation(case, precision=_lib.FP32C)
        sims[kernel].set_tuning(pair_block=256, pair_chunk=32, update_block=256, pair_kernel=kernel)
    for step in range(1, 4):
        print(f"fused vs separate pair kernels, forced state, step {step}", flush=True)
        out = {}
        for kernel, sim in sims.items():
            sim.step(1)
            sim.ctx.check_error()
            out[kernel] = sim.state(fields)
        for kernel in (_lib.PAIRS_LIST_FUSED, _lib.PAIRS_STAGED_FUSED):
            for f in fields:
                np.testing.assert_array_equal(out[kernel][f], out[_lib.PAIRS_LIST][f],
                                              err_msg=f"kernel {kernel} {f} after step {step}")
    for sim in sims.values():
        sim.close()


def _assert_forced_branches_fired(case, P):
    """The oracle's state after the first step holds the three forced failures."""
    bl = P["blanked"] == 1.0
    assert np.any(bl & (P["T"] == case.const.joco["Tmelt"])), "melt branch"
    assert np.any(bl & np.isnan(P["S"][:, 0])), "NaN stress"
    inside = P["pos"][:, 0] < case.tool.seg_left[0][0] - 2.5e-4 + case.tool_vel[0] * case.dt
    assert np.any(bl & inside), "deep penetration"
    assert bl.sum() >= 10, "the NaN particle's neighbours"


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_fixed_particles_near_the_tool(precision):
    """Fixed particles among the near-tool ones (the near-contact pass applies their boundary condition
    and the invalidation check): after every step their velocity, contact force and position/velocity
    rates are zero and their temperature is T_init, and the run follows the oracle."""
    base = cases.small_cutting_case(nx=48, ny=16, wear=True)
    fixed = base.fixed.copy()
    tip = base.tool.seg_left[0][0]
    near = np.nonzero(base.pos[:, 0] < tip + 12 * base.dx)[0]
    fixed[near[::3]] = 1.0
    case = dataclasses.replace(base, fixed=fixed)
    o = Oracle(case)
    prec = _lib.FP32C if precision == "fp32c" else _lib.FP64
    sim = S.Simulation(case, precision=prec)
    T0 = case.const.trml_wp["T_init"]
    for k in range(4):
        print(f"near-tool fixed particles, {precision}, steps {3 * k + 1}..{3 * k + 3}", flush=True)
        o.step(3)
        sim.step(3)
        sim.ctx.check_error()
        a = sim.state()
        fx = a["fixed"] == 1.0
        assert fx.sum() == fixed.sum()
        for f in ("vel", "fc", "pos_t", "vel_t"):
            assert np.all(a[f][fx] == 0.0), f
        assert np.all(a["T"][fx] == T0)
        for f in ("hash", "idx", "blanked"):
            assert_equal(f, a[f], o.P[f])
        live = o.P["blanked"] != 1.0
        for f in ("pos", "vel", "S", "rho", "T"):
            assert_close(f"{f} after {3 * k + 3} steps", a[f], o.P[f], 1e-6, mask=live)
    assert (o.P["on_seg"] >= 0).sum() > 0
