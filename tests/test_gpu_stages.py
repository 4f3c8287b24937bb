"""Per-function parity of the drop-in entry points against the FP64 oracle.

Every stage of the reference step (leap_frog.cu:162-248) is run on the GPU from exactly
the oracle's state before that stage, and compared with the oracle's state after it.
Stages whose arithmetic is the same IEEE operations in the same order must match
bit for bit -- the tool's segment sums included, which both sides add in increasing particle
index; the tolerances below name the few places where results depend on the device's libm
(atan2f/cosf/sinf in the artificial stress, pow/log/exp in Johnson-Cook and Usui).
"""
import copy

import numpy as np
import pytest

from helpers import assert_close, assert_equal
from oracle.oracle import Oracle
from sph_fem_2d_cutting_gpu_amd import cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu

LIBM_RTOL = 1e-6       # float trig of the artificial stress (ulp-level differences, relative to max|R|)
POW_RTOL = 1e-10       # double pow/log/exp of Johnson-Cook / Usui
PSE_RTOL = 1e-12       # double exp of the PSE kernel, summed in the same order


@pytest.fixture(scope="module")
def world():
    case = cases.small_cutting_case(nx=48, ny=16, wear=True)
    o = Oracle(case)
    o.step(30)   # stresses, plasticity, heat and contact are all active by now
    sim = S.Simulation(case)
    return case, o, sim


def upload_all(sim, o):
    """Copy the oracle's particle, stepper and tool state to the device."""
    p = sim.p
    P = o.P
    # stepper init arrays through leap_frog_init (which skips blanked particles)
    for f, src in (("pos", "pos_init"), ("vel", "vel_init"), ("S", "S_init"), ("rho", "rho_init"), ("T", "T_init")):
        p.upload(f, o.st[src])
    p.upload("blanked", np.zeros(case_n(o)))
    sim.stepper.init(p)
    for f in P:
        p.upload(f, P[f])
    if o.t is not None:
        st = dict(o.t)
        st["scalars"] = o.tool_scalars
        sim.tool.upload(st)


def case_n(o):
    return o.case.N


def cells_to_device(sim, o):
    cs = o.st["cell_start"]
    ce = o.st["cell_end"]
    return cs, ce


def snapshot(o):
    return copy.deepcopy(o.P), copy.deepcopy(o.t), dict(o.tool_scalars)


def live(o):
    return o.P["blanked"] != 1.0


def test_hash_sort_cells_bitexact(world):
    case, o, sim = world
    upload_all(sim, o)
    o.call("o_compute_hashes", "cp")
    sim.grid.assign_hashes(sim.p)
    assert_equal("hash", sim.p.download("hash"), o.P["hash"])
    o.call("o_sort_and_cells", "cps")
    sim.grid.sort(sim.p)
    for f in ("hash", "idx"):
        assert_equal(f, sim.p.download(f), o.P[f])
    for f in ("pos", "vel", "fc", "ft", "h", "n", "rho", "S", "eps_pl", "eps_pl_dot", "T", "fixed", "blanked",
              "tool_particle", "pos_t", "vel_t", "rho_t", "S_t", "T_t", "on_seg"):
        assert_equal(f, sim.p.download(f), o.P[f])
    cs, ce = sim.grid.download_cells()
    assert_equal("cell_start", cs, o.st["cell_start"])
    occ = o.st["cell_start"] != -1
    assert_equal("cell_end", ce[occ], o.st["cell_end"][occ])


def _run_unfused_from(sim, o, stage):
    """Upload oracle state, run one GPU stage and the same oracle stage."""
    upload_all(sim, o)
    p, t = sim.p, sim.tool
    ctx = sim.ctx
    dt = o.c.dt
    if stage == "predict":
        o.call("o_predict", "ps", dt)
        sim.stepper.predict(p, dt)
    elif stage == "correct":
        o.call("o_correct", "ps", dt)
        sim.stepper.correct(p, dt)
    elif stage == "eos":
        o.call("o_material_eos", "cp")
        S.material_eos(p)
    elif stage == "artificial_stress":
        o.call("o_artificial_stress", "cp")
        S.corrector_artificial_stress(p)
    elif stage in ("interactions", "heat_pse"):
        # cell arrays of the oracle's last sort
        import ctypes as C
        nc = len(o.st["cell_start"])
        a, b = C.c_void_p(), C.c_void_p()
        ctx.call("sphfem_malloc", C.byref(a), 4 * nc)
        ctx.call("sphfem_malloc", C.byref(b), 4 * nc)
        ctx.htod(a.value, o.st["cell_start"])
        ctx.htod(b.value, o.st["cell_end"])
        if stage == "heat_pse":
            o.call("o_interactions_heat_pse", "cps")
            S.interactions_heat_pse(p, a.value, b.value, nc)
        else:
            o.call("o_interactions_monaghan", "cps")
            S.interactions_monaghan(p, a.value, b.value, nc)
        ctx.synchronize()
        ctx.call("sphfem_free", a)
        ctx.call("sphfem_free", b)
    elif stage == "jaumann":
        o.call("o_stress_rate_jaumann", "cp")
        S.material_stress_rate_jaumann(p)
    elif stage == "continuity":
        o.call("o_continuity", "p")
        S.contmech_continuity(p)
    elif stage == "momentum":
        o.call("o_momentum", "cp")
        S.contmech_momentum(p)
    elif stage == "advection":
        o.call("o_advection", "p")
        S.contmech_advection(p)
    elif stage == "fric_heat":
        o.call("o_fric_heat_v21", "cpt")
        S.material_fric_heat_gen_FEM_v2(p, t)
    elif stage == "plasticity":
        o.call("o_plasticity", "cp")
        S.plasticity_johnson_cook(p)
    elif stage == "contact":
        o.call("o_contact_s1", "cpt")
        o.call("o_contact_s2", "cpt")
        S.compute_contact_forces_FEM_tool_v3(p, t)
    elif stage == "bc":
        o.call("o_boundary_conditions", "p")
        o.call("o_boundary_conditions_thermal", "cp")
        S.perform_boundary_conditions(p)
        S.perform_boundary_conditions_thermal(p)
    elif stage == "invalidate":
        o.call("o_invalidate", "p")
        S.debug_invalidate(p)
    else:
        raise ValueError(stage)
    ctx.synchronize()


EXACT = {
    "predict": ("pos", "vel", "S", "rho", "T"),
    "correct": ("pos", "vel", "S", "rho", "T"),
    "eos": ("p",),
    "interactions": ("S_der", "v_der", "pos_t", "vel_t", "T_t"),
    "jaumann": ("S_t",),
    "continuity": ("rho_t",),
    "momentum": ("vel_t",),
    "advection": ("pos_t",),
    "bc": ("vel", "fc", "pos_t", "vel_t", "T"),
    "invalidate": ("blanked",),
    "contact": ("fc", "ft", "n", "on_seg", "blanked", "T"),
    "fric_heat": ("T", "blanked"),
}


@pytest.mark.parametrize("stage", ["predict", "correct", "eos", "interactions", "jaumann", "continuity",
                                   "momentum", "advection", "bc", "invalidate", "contact", "fric_heat"])
def test_stage_bitexact(world, stage):
    case, o0, sim = world
    o = o0.clone()
    _run_unfused_from(sim, o, stage)
    mask = None
    if stage in ("interactions",):
        mask = o.P["blanked"] != 1.0
    for f in EXACT[stage]:
        g = sim.p.download(f)
        r = o.P[f]
        if f in ("S_der", "v_der", "pos_t", "vel_t", "T_t") and stage == "interactions":
            # exact up to the sign of zero sums
            assert_close(f, g, r, 1e-15, mask=mask)
        else:
            assert_equal(f, g, r, mask=mask)
    if stage in ("contact", "fric_heat"):
        tg = sim.tool.download()
        for f in ("heat_exchange", "fric_heat", "physical_para", "sliding_force"):
            assert_equal(f, tg[f], o.t[f])


def test_stage_heat_pse(world):
    """PSE conduction (interactions_gpu.cu:84-140): same neighbour order as the oracle, the only
    difference is the device's double exp (ulp-level)."""
    case, o0, sim = world
    o = o0.clone()
    rng = np.random.default_rng(3)
    o.P["T"][:] += rng.normal(0.0, 20.0, case.N)         # a rough field so that every pair contributes
    _run_unfused_from(sim, o, "heat_pse")
    live_i = live(o)
    assert np.abs(o.P["T_t"][live_i]).max() > 0
    assert_close("T_t", sim.p.download("T_t"), o.P["T_t"], PSE_RTOL, mask=live_i)


def test_stage_artificial_stress(world):
    case, o0, sim = world
    o = o0.clone()
    _run_unfused_from(sim, o, "artificial_stress")
    assert_close("R", sim.p.download("R"), o.P["R"], LIBM_RTOL)


def test_stage_plasticity(world):
    case, o0, sim = world
    o = o0.clone()
    # load every other particle beyond yield (pure shear of 0.02 = 2 GPa) so the radial return runs
    o.P["S"][::2, 0] += 0.02
    o.P["S"][::2, 2] -= 0.02
    _run_unfused_from(sim, o, "plasticity")
    plastic = o.P["eps_pl_dot"] > 0
    assert plastic.sum() > 0, "the fixture should have plastic particles"
    for f in ("S", "eps_pl", "eps_pl_dot", "T"):
        assert_close(f, sim.p.download(f), o.P[f], POW_RTOL)
    assert_equal("blanked", sim.p.download("blanked"), o.P["blanked"])


def test_fem_and_tool_stages(world):
    """reset_FEM_server, calculate_temp, apply_robin_bdc, calculate_temp_rate, segments / tool update."""
    case, o0, sim = world
    o = o0.clone()
    upload_all(sim, o)
    t = sim.tool
    dt = o.c.dt
    o.call("o_fem_reset_server", "t")
    t.reset_FEM_server()
    o.call("o_fem_calculate_temp", "t", 0.5 * dt)
    t.calculate_temp(0.5 * dt)
    o.call("o_fem_apply_robin", "ct")
    t.apply_heat_boundary_condition()
    o.call("o_fem_calculate_temp_rate", "t")
    t.calculate_temp_rate()
    o.call("o_fem_calculate_temp", "t", dt)
    t.calculate_temp(dt)
    g = t.download()
    assert_equal("k", g["k"], o.t["k"])
    assert_equal("T", g["T"], o.t["T"])
    o.call("o_segments_update", "t", dt)
    t.segments_update(dt)
    o.call("o_tool_update", "t", dt)
    t.tool_update_FEM_gpu(dt)
    g = t.download()
    for f in ("node_pos", "left", "right", "heat_exchange", "fric_heat"):
        assert_equal(f, g[f], o.t[f])
    for k, v in o.tool_scalars.items():
        assert g["scalars"][k] == v, k


def test_wear_stage(world):
    """wear_adjustment: boundary-layer re-assembly, Usui nodal shift, segment update, flank retreat."""
    case, o0, sim = world
    o = o0.clone()
    # a hot, loaded edge so that the Usui rates are not negligible
    o.t["T"][:] = 1200.0
    o.t["physical_para"][:, 0] = 2.0
    o.t["physical_para"][:, 1] = 2 * 0.02
    o.t["physical_para"][:, 2] = 2 * 0.004
    upload_all(sim, o)
    o.call("o_wear_adjustment", "ct")
    S.wear_adjustment(sim.tool)
    g = sim.tool.download()
    assert np.any(g["wear_nodes"] > 0)
    for f in ("node_pos", "left", "right", "n", "K", "M", "wear_nodes"):
        assert_close(f, g[f], o.t[f], POW_RTOL)
    assert abs(g["scalars"]["flank_retreat"] - o.tool_scalars["flank_retreat"]) <= 1e-12 * max(
        1e-12, abs(o.tool_scalars["flank_contact"]))


@pytest.mark.parametrize("tool_block", [False, True])
def test_heat_pse_tool_particles_and_blanked(tool_block):
    """Tool particles take the tool diffusivity of interactions_setup_thermal_constants_tool (zero
    until it is called), blanked particles keep their rate; lattice whose grid cells are 2.5 h wide
    (the window then covers the Gaussian's support)."""
    from test_oracle_pins import _pse_lattice
    o, hp, inner = _pse_lattice(n=40)
    N = o.case.N
    sim = S.Simulation(o.case)
    if tool_block:
        trml = o.case.const.trml_tool
        S.interactions_setup_thermal_constants_tool(sim.ctx, trml)
        o.c.itool_alpha_pse = trml["alpha"]
    rng = np.random.default_rng(8)
    x = o.P["pos"] - o.P["pos"].mean(axis=0)
    o.P["T"][:] = 300.0 + (x[:, 0] ** 2 + x[:, 1] ** 2) / hp ** 2 + rng.normal(0, 1.0, N)
    o.P["tool_particle"][::3] = 1.0
    o.P["blanked"][1::7] = 1.0
    o.P["T_t"][:] = -7.0
    _run_unfused_from(sim, o, "heat_pse")
    g = sim.p.download("T_t")
    blank = o.P["blanked"] == 1.0
    tool = o.P["tool_particle"] == 1.0
    assert np.all(g[blank] == -7.0)
    assert np.any(g[tool & ~blank] != 0.0) == tool_block
    assert_close("T_t", g, o.P["T_t"], PSE_RTOL, mask=~blank)
    sim.close()
