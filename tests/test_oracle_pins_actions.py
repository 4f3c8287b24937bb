"""Pins of the oracle's per-particle actions, boundary conditions, failure branches and tool motion.

Each check compares an oracle function with a closed form, an invariant of the step or an
independent computation (numpy, a removed particle, a hand-worked formula at chosen points), so a
plausible slip -- a wrong derivative component, a wrong axis, a dropped factor of 1/2, a missing
division by the mass, a branch taken on the wrong side -- fails at least one of them.
"""
import dataclasses
import math

import numpy as np
import pytest

from oracle import oracle as O
from sph_fem_2d_cutting_gpu_amd import cases
from test_oracle_pins import _f32sqrt, _interior, _lattice_oracle, _sigma_y


def _rates(o):
    o.call("o_material_eos", "cp")
    o.call("o_artificial_stress", "cp")
    o.call("o_interactions_monaghan", "cps")


# ---------------------------------------------------------------- actions_gpu.cu:626-687
def test_continuity_of_linear_velocity_field():
    """rho_t = -rho div v: v = A x with all four gradient entries distinct gives -rho (A00 + A11) in the
    interior; a slip to v_der.z (dvy/dx) or v_der.y (dvx/dy) would give -rho (A00 + A10) or -rho (A01 + A11)."""
    o = _lattice_oracle()
    A = np.array([[0.3, -0.7], [1.1, 0.2]])
    o.P["vel"][:] = o.P["pos"] @ A.T
    rho = 4.43 * (1.0 + 0.01 * np.sin(np.arange(o.case.N)))
    o.P["rho"][:] = rho
    _rates(o)
    o.P["tool_particle"][::11] = 1.0
    o.P["rho_t"][:] = -5.0
    o.call("o_continuity", "p")
    inn = _interior(o) & (o.P["tool_particle"] == 0.0)
    want = -o.P["rho"][inn] * (A[0, 0] + A[1, 1])
    assert np.allclose(o.P["rho_t"][inn], want, rtol=0.02 * np.abs(A).max() / abs(A[0, 0] + A[1, 1]))
    # and exactly the product of the interaction sums' diagonal, not an approximation of it
    d = o.P["v_der"]
    live = o.P["tool_particle"] == 0.0
    assert np.array_equal(o.P["rho_t"][live], -o.P["rho"][live] * (d[live, 0] + d[live, 3]))
    assert np.all(o.P["rho_t"][~live] == -5.0)            # tool particles keep their rate


def test_momentum_of_linear_stress_field_and_contact_forces():
    """vel_t += (S_der.x + S_der.y, S_der.z + S_der.w) + (fc + ft) / m.  At rest (no viscosity, no XSPH)
    with p = 0, R = 0 and a linear stress field Sxx = a x + b y, Sxy = c x + d y, Syy = e x + f y the
    interior acceleration is ((a + d) / rho, (c + f) / rho); every coefficient is distinct, so a
    mispaired derivative fails.  The contact part is exactly (fc + ft) / m per axis."""
    o = _lattice_oracle()
    N = o.case.N
    rho = o.case.const.phys["rho0"]
    a, b, c, d, e, f = 2.0, -1.3, -3.0, 5.0, 0.7, 7.0
    x = o.P["pos"] - o.P["pos"].mean(axis=0)
    o.P["S"][:] = 0.0
    o.P["S"][:, 0] = a * x[:, 0] + b * x[:, 1]
    o.P["S"][:, 1] = c * x[:, 0] + d * x[:, 1]
    o.P["S"][:, 2] = e * x[:, 0] + f * x[:, 1]
    o.call("o_material_eos", "cp")
    assert np.all(o.P["p"] == 0.0)
    o.P["R"][:] = 0.0
    o.call("o_interactions_monaghan", "cps")
    base = o.P["vel_t"].copy()
    assert np.all(base == 0.0)                          # nothing moves: no viscosity term
    o.P["tool_particle"][::13] = 1.0
    rng = np.random.default_rng(21)
    m = o.case.const.phys["mass"]
    o.P["fc"][:] = rng.normal(0, 1e-9, (N, 2))
    o.P["ft"][:] = rng.normal(0, 1e-9, (N, 2))
    o.call("o_momentum", "cp")
    tool = o.P["tool_particle"] == 1.0
    assert np.all(o.P["vel_t"][tool] == 0.0)
    inn = _interior(o) & ~tool
    acc = o.P["vel_t"][inn] - (o.P["fc"][inn] + o.P["ft"][inn]) / m
    assert np.allclose(acc[:, 0], (a + d) / rho, rtol=0.03)
    assert np.allclose(acc[:, 1], (c + f) / rho, rtol=0.03)
    # the contact term alone (no stress, zero previous rate): (fc + ft) / m on each axis
    o.P["S_der"][:] = 0.0
    o.P["vel_t"][:] = 0.0
    o.call("o_momentum", "cp")
    want = (o.P["fc"][~tool] + o.P["ft"][~tool]) / m
    mag = (np.abs(o.P["fc"][~tool]) + np.abs(o.P["ft"][~tool])) / m
    assert np.all(np.abs(o.P["vel_t"][~tool] - want) <= 1e-15 * mag)


def test_advection_of_uniform_velocity():
    """pos_t = XSPH + v: a uniform velocity (vx != vy) has zero XSPH correction, so pos_t = v exactly for
    material particles; tool particles keep the interaction result (zero)."""
    o = _lattice_oracle(jitter=0.2)
    o.P["vel"][:] = [3e-4, -7e-4]
    _rates(o)
    assert np.all(o.P["pos_t"] == 0.0)
    o.P["tool_particle"][::7] = 1.0
    o.call("o_advection", "p")
    tool = o.P["tool_particle"] == 1.0
    assert np.all(o.P["pos_t"][~tool] == [3e-4, -7e-4])
    assert np.all(o.P["pos_t"][tool] == 0.0)


def test_advection_of_linear_velocity_field_in_the_interior():
    """v = A x: the XSPH sum of (v_j - v_i) W_ij vanishes on the lattice's interior by symmetry, so
    pos_t = v there (to the lattice's rounding)."""
    o = _lattice_oracle()
    A = np.array([[0.3, -0.7], [1.1, 0.2]])
    o.P["vel"][:] = (o.P["pos"] - o.P["pos"].mean(axis=0)) @ A.T
    _rates(o)
    o.call("o_advection", "p")
    inn = _interior(o)
    scale = np.abs(o.P["vel"][inn]).max()
    assert np.allclose(o.P["pos_t"][inn], o.P["vel"][inn], rtol=0, atol=1e-9 * scale)


# ---------------------------------------------------------------- actions_gpu.cu:756-781
def test_boundary_conditions_hold_after_every_step():
    """Whole-step invariant of leap_frog.cu:240-244: after each step every fixed particle has zero
    velocity, contact force and position/velocity rates, and the initial temperature, while the free
    particles near the tool move and heat."""
    case = cases.small_cutting_case(nx=32, ny=12)
    o = O.Oracle(case)
    fixed = case.fixed == 1.0
    assert fixed.sum() > 0 and (~fixed).sum() > 0
    T0 = case.const.trml_wp["T_init"]
    for _ in range(8):
        o.step(5)
        fx = o.P["fixed"] == 1.0          # sorted order: follow the flags
        assert fx.sum() == fixed.sum()
        for k in ("vel", "fc", "pos_t", "vel_t"):
            assert np.all(o.P[k][fx] == 0.0), k
        assert np.all(o.P["T"][fx] == T0)
    free = o.P["fixed"] == 0.0
    assert np.abs(o.P["vel"][free]).max() > 0 and o.P["T"][free].max() > T0


def test_boundary_conditions_flag_combinations():
    """Fixed material particles lose vel, fc, pos_t and vel_t; fixed tool particles keep them (the
    mechanical BC skips tool particles) but still take T_init (the thermal BC does not); free particles
    are untouched; ft, positions and the stress are never touched."""
    o = _lattice_oracle(nx=8, ny=8)
    N = o.case.N
    rng = np.random.default_rng(2)
    for k in ("vel", "fc", "ft", "pos_t", "vel_t", "pos"):
        o.P[k][:] = rng.normal(0, 1, (N, 2))
    o.P["T"][:] = rng.uniform(400, 900, N)
    o.P["fixed"][:] = (np.arange(N) % 2 == 0)
    o.P["tool_particle"][:] = (np.arange(N) % 3 == 0)
    before = {k: v.copy() for k, v in o.P.items()}
    o.call("o_boundary_conditions", "p")
    o.call("o_boundary_conditions_thermal", "cp")
    fx = before["fixed"] == 1.0
    tl = before["tool_particle"] == 1.0
    mech = fx & ~tl
    for k in ("vel", "fc", "pos_t", "vel_t"):
        assert np.all(o.P[k][mech] == 0.0), k
        assert np.array_equal(o.P[k][~mech], before[k][~mech]), k
    for k in ("ft", "pos", "S", "rho"):
        assert np.array_equal(o.P[k], before[k]), k
    assert np.all(o.P["T"][fx] == o.case.const.trml_wp["T_init"])
    assert np.array_equal(o.P["T"][~fx], before["T"][~fx])


# ---------------------------------------------------------------- actions_gpu.cu:787-824
RATE_COMPONENTS = [("pos_t", 0), ("pos_t", 1), ("vel_t", 0), ("vel_t", 1), ("S_t", 0), ("S_t", 1),
                   ("S_t", 2), ("S_t", 3), ("rho_t", None), ("T_t", None)]


def test_invalidate_blanks_each_non_finite_rate():
    """Every one of the ten rate components, set to NaN, +Inf or -Inf on its own particle, blanks that
    particle; finite rates (however large) and the other fields (a NaN stress) do not; blanked
    particles stay blanked.  Checked against numpy's isfinite over the same ten components."""
    o = _lattice_oracle(nx=10, ny=10)
    N = o.case.N
    rng = np.random.default_rng(4)
    for k in ("pos_t", "vel_t"):
        o.P[k][:] = rng.normal(0, 1e300, (N, 2))
    o.P["S_t"][:] = rng.normal(0, 1, (N, 4))
    o.P["rho_t"][:] = rng.normal(0, 1, N)
    o.P["T_t"][:] = rng.normal(0, 1, N)
    bad_values = (np.nan, np.inf, -np.inf)
    i = 0
    for f, comp in RATE_COMPONENTS:
        for v in bad_values:
            if comp is None:
                o.P[f][i] = v
            else:
                o.P[f][i, comp] = v
            i += 1
    o.P["S"][i] = np.nan                 # not a rate: ignored by the check
    o.P["blanked"][i + 1] = 1.0          # stays blanked
    o.P["T_t"][i + 1] = 1.0
    o.call("o_invalidate", "p")
    rates = np.concatenate([o.P["pos_t"], o.P["vel_t"], o.P["S_t"], o.P["rho_t"][:, None], o.P["T_t"][:, None]],
                           axis=1)
    want = (~np.isfinite(rates).all(axis=1)).astype(float)
    want[i + 1] = 1.0
    assert np.array_equal(o.P["blanked"], want)
    assert o.P["blanked"][:30].sum() == 30 and o.P["blanked"].sum() == 31


def test_blanked_neighbour_is_left_out_of_the_sums():
    """interactions_gpu.cu:152, 232: the pair sums of the other particles with particle k blanked equal,
    bit for bit, those of the same lattice with k removed; k's own derivatives are not written."""
    o = _lattice_oracle(nx=14, ny=14, jitter=0.25)
    N = o.case.N
    rng = np.random.default_rng(6)
    vel = rng.normal(0, 1e-3, (N, 2))
    S = np.zeros((N, 4))
    S[:, :3] = rng.normal(0, 3e-3, (N, 3))
    T = 300.0 + rng.normal(0, 50.0, N)
    rho = 4.43 * (1.0 + rng.normal(0, 1e-3, N))
    for key, v in (("vel", vel), ("S", S), ("T", T), ("rho", rho)):
        o.P[key][:] = v
    k = int(np.argmin(np.sum((o.P["pos"] - o.P["pos"].mean(axis=0)) ** 2, axis=1)))   # a central particle
    o.call("o_compute_hashes", "cp")
    o.P["blanked"][k] = 1.0
    o.call("o_sort_and_cells", "cps")
    kk = int(np.nonzero(o.P["blanked"] == 1.0)[0][0])
    for f in ("v_der", "S_der", "pos_t", "vel_t", "T_t"):
        o.P[f][kk] = 99.0
    _rates(o)
    for f in ("v_der", "S_der", "pos_t", "vel_t"):
        assert np.all(o.P[f][kk] == 99.0), f
    # the same workload without particle k, built from the sorted state (which a re-sort keeps)
    keep = np.ones(N, bool)
    keep[kk] = False
    base = o.case
    case2 = dataclasses.replace(base, N=N - 1, pos=o.P["pos"][keep].copy(), vel=o.P["vel"][keep].copy(),
                                rho=o.P["rho"][keep].copy(), T=o.P["T"][keep].copy(), hh=o.P["h"][keep].copy(),
                                fixed=o.P["fixed"][keep].copy(), S=o.P["S"][keep].copy())
    o2 = O.Oracle(case2)
    o2.call("o_compute_hashes", "cp")
    o2.call("o_sort_and_cells", "cps")
    assert np.array_equal(o2.P["idx"], np.arange(N - 1))
    _rates(o2)
    for f in ("v_der", "S_der", "pos_t", "vel_t", "T_t"):
        assert np.array_equal(o.P[f][keep], o2.P[f]), f
    # k was a neighbour of many: without the exclusion the sums would differ
    assert np.count_nonzero(np.abs(o.P["pos"] - o.P["pos"][kk]).max(axis=1) < 2 * o.case.h) > 10


def test_blanked_particle_keeps_its_hash():
    """grid.cu:30-32: a blanked particle keeps the hash it had, wherever it is now."""
    o = _lattice_oracle(nx=6, ny=6)
    o.call("o_compute_hashes", "cp")
    h0 = o.P["hash"].copy()
    o.P["pos"][3] += [5 * o.case.dx, 0.0]
    o.P["pos"][4] += [5 * o.case.dx, 0.0]
    o.P["blanked"][4] = 1.0
    o.call("o_compute_hashes", "cp")
    assert o.P["hash"][3] != h0[3] and o.P["hash"][4] == h0[4]


# ---------------------------------------------------------------- actions_gpu.cu:689-754 melt branch
def _plastic_particle_case(tq_scale):
    case = cases.small_cutting_case(nx=16, ny=8)
    case = dataclasses.replace(case, const=dataclasses.replace(case.const, trml_wp=dict(case.const.trml_wp)))
    case.const.trml_wp["tq"] *= tq_scale
    return case


@pytest.mark.parametrize("T0,tq_scale,melts", [(1877.0, 1e8, True), (600.0, 1.0, False)])
def test_melt_clamps_and_blanks(T0, tq_scale, melts):
    """After the return, plastic-work heating above Tmelt sets T = Tmelt and blanks the particle
    (actions_gpu.cu:750-753); below Tmelt T = T0 + tq / (c_p rho) d(eps) sigma_Y(new) and the particle
    stays.  Elastic, tool and blanked particles never reach the check."""
    case = _plastic_particle_case(tq_scale)
    o = O.Oracle(case)
    jc, G, dt = case.const.joco, case.const.phys["G"], case.dt
    o.P["S"][:] = 0.0
    sd = np.array([0.012, 0.004, -0.009])
    for i in range(3):
        o.P["S"][i] = [sd[0], sd[1], sd[2], -(sd[0] + sd[2])]
    o.P["p"][:] = 0.0
    o.P["eps_pl"][:3] = 0.1
    o.P["eps_pl_dot"][:3] = 5e-5
    o.P["T"][:3] = T0
    o.P["rho"][:3] = 4.43
    o.P["tool_particle"][1] = 1.0
    o.P["blanked"][2] = 1.0
    o.P["T"][3] = jc["Tmelt"] - 1e-3            # elastic (zero stress): not clamped, not blanked
    o.call("o_plasticity", "cp")
    Sx, Sxy, Sy, Sz = o.P["S"][1]
    normS = _f32sqrt(Sx * Sx + Sy * Sy + Sz * Sz + 2 * Sxy * Sxy)
    init = max(5e-5 * dt * math.sqrt(2.0 / 3.0), 1e-8)
    dl, _ = O.solve_secant(jc, init, 1e-6, normS, 0.1, T0, dt, G)
    s23 = float(np.sqrt(np.float32(2.0 / 3.0)))
    eps_new = 0.1 + s23 * float(np.float32(dl))
    sy = _sigma_y(jc, eps_new, s23 * float(np.float32(dl)) / dt, T0)
    trml = case.const.trml_wp
    T_heated = T0 + trml["tq"] / (trml["cp"] * 4.43) * (eps_new - 0.1) * sy
    assert (T_heated > jc["Tmelt"]) == melts
    if melts:
        assert o.P["T"][0] == jc["Tmelt"] and o.P["blanked"][0] == 1.0
    else:
        assert o.P["T"][0] == pytest.approx(T_heated, rel=1e-12) and o.P["blanked"][0] == 0.0
    assert o.P["T"][1] == T0 and o.P["blanked"][1] == 0.0 and o.P["eps_pl"][1] == 0.1
    assert o.P["T"][2] == T0 and o.P["eps_pl"][2] == 0.1
    assert o.P["T"][3] == jc["Tmelt"] - 1e-3 and o.P["blanked"][3] == 0.0


# ---------------------------------------------------------------- actions_gpu.cu:423-583 deep penetration
@pytest.mark.parametrize("depth,blanks", [(2.4e-4, False), (2.6e-4, True), (4e-4, True)])
def test_contact_deep_penetration_blanks(depth, blanks):
    """A particle more than 2.5e-4 cm behind the rake face is blanked with zero forces, normal and no
    segment (actions_gpu.cu:479-491); one less deep gets the penalty force m d alpha / dt^2."""
    case = cases.small_cutting_case(nx=16, ny=8)
    o = O.Oracle(case)
    t = o.t
    seg = 7
    assert np.allclose(t["n"][seg], [-1.0, 0.0])
    xr = t["left"][seg][0]
    ymid = 0.5 * (t["left"][seg][1] + t["right"][seg][1])
    o.P["pos"][:] = [xr + 1.0, ymid]
    o.P["pos"][0] = [xr - depth, ymid]
    o.P["vel"][0] = [0.0, 1e-3]
    o.P["ft"][:] = 1.0                       # stale values are cleared on every branch
    o.P["fc"][:] = 1.0
    o.call("o_contact_s1", "cpt")
    if blanks:
        assert o.P["blanked"][0] == 1.0 and o.P["on_seg"][0] == -1
        assert np.all(o.P["fc"][0] == 0.0) and np.all(o.P["ft"][0] == 0.0) and np.all(o.P["n"][0] == 0.0)
    else:
        c = case.const.tool
        assert o.P["blanked"][0] == 0.0 and o.P["on_seg"][0] == seg
        assert o.P["fc"][0][0] == pytest.approx(c["slave_mass"] * depth * c["contact_alpha"] / case.dt ** 2,
                                                rel=1e-9)
    assert np.all(o.P["blanked"][1:] == 0.0) and np.all(o.P["fc"][1:] == 0.0)


def test_fric_heat_nan_velocity_blanks():
    """A contact particle whose relative tangential speed is NaN is blanked and adds nothing to its
    segment (actions_gpu.cu:375-380); a neighbour on the same segment still contributes."""
    case = cases.small_cutting_case(nx=16, ny=8)
    o = O.Oracle(case)
    seg = 3
    o.P["on_seg"][:] = -1
    o.P["on_seg"][:2] = seg
    for k in ("fric_heat", "sliding_force", "physical_para", "heat_exchange"):
        o.t[k][:] = 0.0
    o.P["n"][:2] = [1.0, 0.0]
    o.P["vel"][0] = [0.0, np.nan]
    o.P["vel"][1] = [0.0, 2e-3]
    o.P["ft"][:2] = [0.0, -2e-9]
    o.P["fc"][:2] = [5e-9, 0.0]
    o.P["T"][:2] = 400.0
    o.call("o_fric_heat_v21", "cpt")
    assert o.P["blanked"][0] == 1.0 and o.P["T"][0] == 400.0
    assert o.P["blanked"][1] == 0.0 and o.P["T"][1] > 400.0
    assert o.t["physical_para"][seg][0] == 1.0
    value = case.const.tool["eta"] * _f32sqrt(4e-18) * _f32sqrt(4e-6)
    assert o.t["fric_heat"][seg] == pytest.approx(value * case.const.tool["beta"], rel=1e-15)


# ---------------------------------------------------------------- tool_FEM.cu:109-139, 1552-1582
def test_tool_and_segment_motion_is_affine_with_flank_retreat():
    """k steps of segments_update + tool_update_FEM_gpu translate nodes and segment ends by k dt v plus
    k flank_retreat on y only; the reference points move by k dt v without the retreat; flank_contact
    follows the lowest point; the heat exchange and friction heat accumulators are cleared."""
    case = cases.small_cutting_case(nx=16, ny=8, wear=True)
    o = O.Oracle(case)
    vx, vy, fr = 1.1667e-4, -2.5e-5, -3.0e-9
    o.Tc.vel_x, o.Tc.vel_y, o.Tc.flank_retreat = vx, vy, fr
    t = o.t
    nodes0, l0, r0 = t["node_pos"].copy(), t["left"].copy(), t["right"].copy()
    sc0 = o.tool_scalars
    dt, k = case.dt, 7
    for _ in range(k):
        t["heat_exchange"][:] = 3.0
        t["fric_heat"][:] = 4.0
        o.call("o_segments_update", "t", dt)
        o.call("o_tool_update", "t", dt)
    shift = np.array([k * dt * vx, k * (dt * vy + fr)])
    for a, a0 in ((t["node_pos"], nodes0), (t["left"], l0), (t["right"], r0)):
        assert np.allclose(a - a0, shift, rtol=1e-9, atol=1e-17)
    sc = o.tool_scalars
    assert sc["tl_x"] - sc0["tl_x"] == pytest.approx(k * dt * vx, rel=1e-9)
    assert sc["tl_y"] - sc0["tl_y"] == pytest.approx(k * dt * vy, rel=1e-9)
    assert sc["br_y"] - sc0["br_y"] == pytest.approx(k * dt * vy, rel=1e-9)
    assert sc["lowest_point"] - sc0["lowest_point"] == pytest.approx(k * dt * vy, rel=1e-9)
    assert sc["right_point"] - sc0["right_point"] == pytest.approx(k * dt * vx, rel=1e-9)
    assert sc["flank_contact"] == sc["lowest_point"] and sc["flank_retreat"] == fr
    assert np.all(t["heat_exchange"] == 0.0) and np.all(t["fric_heat"] == 0.0)


# ---------------------------------------------------------------- FEM_thermal_gpu.cu:342-384
def test_robin_increments_of_one_contact_segment():
    """One segment with contact: its two end nodes (marks z, w, at the segment's end points) get
    f += h L T_avg / 2 + q / 2 and K_ii += h L / 2 with L the (single-precision) segment length and
    T_avg = sum(T l) / sum(l); nothing else changes; a second segment sharing a node adds its own half."""
    case = cases.config1()
    o = O.Oracle(case)
    t = o.t
    h = o.c.tc.h
    for s_list in ([5], [5, 6]):
        t["heat_exchange"][:] = 0.0
        t["fric_heat"][:] = 0.0
        for s in s_list:
            t["heat_exchange"][s] = [900.0 * 3e-4 + 700.0 * 1e-4, 4e-4]      # T_avg = 850 K
            t["fric_heat"][s] = 2.5e-9
        o.call("o_fem_reset_server", "t")
        o.call("o_fem_apply_robin", "ct")
        df = t["f_temp"] - t["f"]
        dK = t["K_temp"] - t["K"]
        want_f = np.zeros_like(df)
        want_K = np.zeros_like(dK)
        rp, ci = t["row_ptr"], t["col_ind"]
        for s in s_list:
            ln, rn = t["marks"][s][2], t["marks"][s][3]
            assert np.array_equal(t["node_pos"][ln], t["left"][s]) and np.array_equal(t["node_pos"][rn], t["right"][s])
            L = _f32sqrt(np.sum((t["left"][s] - t["right"][s]) ** 2))
            for node in (ln, rn):
                want_f[node] += 0.5 * h * L * 850.0 + 0.5 * 2.5e-9
                j = rp[node] + int(np.nonzero(ci[rp[node]:rp[node + 1]] == node)[0][0])
                want_K[j] += 0.5 * h * L
        assert np.allclose(df, want_f, rtol=1e-12, atol=0)
        assert np.allclose(dK, want_K, rtol=1e-12, atol=0)
        assert np.count_nonzero(df) == 2 * len(s_list) - (len(s_list) - 1)


def _strip_tool(nx, ny, Lx, Ly, tc):
    """A rectangle [0, Lx] x [0, Ly] of right triangles: Dirichlet edge x = 0, cutting (Robin) edge
    x = Lx as a chain of ny segments from the top down, the other edges insulated."""
    xs = np.linspace(0.0, Lx, nx + 1)
    ys = np.linspace(0.0, Ly, ny + 1)
    nid = lambda i, j: i * (ny + 1) + j
    nodes = np.array([[x, y] for x in xs for y in ys])
    tris = []
    for i in range(nx):
        for j in range(ny):
            a, b, c, d = nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1)
            tris += [(a, b, c), (a, c, d)]
    tris = np.array(tris)
    n_nodes, n_elem = len(nodes), len(tris)
    node_elems = [[] for _ in range(n_nodes)]
    for e, tr in enumerate(tris):
        for v in tr:
            node_elems[v].append(e)
    cut_nodes = [nid(nx, j) for j in range(ny, -1, -1)]        # top -> bottom
    cut_edges = [(cut_nodes[k], cut_nodes[k + 1]) for k in range(ny)]
    owner = {}
    for e, (a, b, c) in enumerate(tris):
        for u, v in ((a, b), (b, c), (c, a)):
            owner[(min(u, v), max(u, v))] = e
    cut_elems = sorted({owner[(min(u, v), max(u, v))] for u, v in cut_edges})
    colour = np.zeros(n_elem, np.int32)

    def paint(es, allowed):
        for e in es:
            used = {colour[o] for v in tris[e] for o in node_elems[v] if o != e and o in allowed}
            c = 1
            while c in used:
                c += 1
            colour[e] = c
    paint(cut_elems, set(cut_elems))
    n_cut_colors = int(colour.max())
    paint([e for e in range(n_elem) if colour[e] == 0], set(range(n_elem)))
    rows = [set() for _ in range(n_nodes)]
    for tr in tris:
        for u in tr:
            rows[u].update(tr.tolist())
    row_ptr = np.zeros(n_nodes + 1, np.int32)
    row_ptr[1:] = np.cumsum([len(s) for s in rows])
    col_ind = np.concatenate([np.array(sorted(s), np.int32) for s in rows])
    elems = np.zeros((n_elem, 4), np.int32)
    elems[:, :3] = tris
    elems[:, 3] = colour
    cut_edge_node = np.zeros(n_nodes, np.int32)
    cut_edge_node[cut_nodes] = 1
    fixed_edge_node = np.zeros(n_nodes, np.int32)
    fixed_edge_node[[nid(0, j) for j in range(ny + 1)]] = 1
    K, M = cases.assemble_p1(nodes, elems, row_ptr, col_ind, tc, colour_order=True, n_colors=int(colour.max()))
    left = nodes[[a for a, _ in cut_edges]]
    right = nodes[[b for _, b in cut_edges]]
    d = right - left
    nvec = d[:, ::-1] * [1.0, -1.0] / np.linalg.norm(d, axis=1)[:, None]
    marks = np.array([(owner[(min(u, v), max(u, v))], colour[owner[(min(u, v), max(u, v))]], u, v)
                      for u, v in cut_edges], np.int32)
    tm = cases.ToolMesh(nodes=nodes, elems=elems, layers=np.zeros((n_elem, 2), np.int32), cut_edge_node=cut_edge_node,
                        fixed_edge_node=fixed_edge_node, row_ptr=row_ptr, col_ind=col_ind, K=K, M=M,
                        f=np.zeros(n_nodes), T=np.full(n_nodes, tc["T0"]), seg_left=left, seg_right=right,
                        seg_n=nvec, seg_marks=marks, n_colors=int(colour.max()), n_cut_colors=n_cut_colors,
                        tl=np.array([0.0, Ly]), br=np.array([Lx, 0.0]), ref_dist=float(np.hypot(Lx, Ly)),
                        lowest_point=0.0, right_point=Lx)
    fe, ff = cases.fixed_boundary(tm)
    flags = np.zeros((n_elem, 3), np.int32)
    flags[fe] = ff
    cases.dirichlet_rows(elems, flags, fe, row_ptr, col_ind, tm.K, tm.M, tm.f, tm.T, tc["T0"], tm.n_colors)
    return tm


def _strip_oracle(tm, tc_over):
    base = cases.small_cutting_case(nx=4, ny=2)
    const = dataclasses.replace(base.const, tool=dict(base.const.tool, **tc_over))
    case = dataclasses.replace(base, tool=tm, const=const, pos=base.pos + 1.0)   # particles far away
    return O.Oracle(case)


@pytest.mark.parametrize("q_per_length", [0.0, 3e-6])
def test_robin_dirichlet_steady_profile(q_per_length):
    """1-D conduction through the strip between T_D at x = 0 and a Robin edge at x = L (exchange h with
    contact particles at T_inf, friction heat q' per unit length): the steady profile
    T = T_D + x (h (T_inf - T_D) + q') / (k + h L) is linear, which P1 elements hold exactly, so it is
    a fixed point of the oracle's rate (k = M^-1 (f - K T) = 0 at every node).  The heat exchange
    lengths differ per segment: only the ratio sum(T l) / sum(l) may enter."""
    Lx, Ly = 0.02, 0.006
    tc0 = cases.make_constants("Ti6Al4V", 2.5e-4, 0.35).tool
    h = 2.0 * tc0["k"] / Lx                        # Biot number 2: the edge sits between the two limits
    tm = _strip_tool(10, 4, Lx, Ly, dict(tc0))
    o = _strip_oracle(tm, dict(h=h))
    t = o.t
    TD, Tinf = tc0["T0"], 800.0
    a = (h * (Tinf - TD) + q_per_length) / (tc0["k"] + h * Lx)
    T_exact = TD + a * t["node_pos"][:, 0]
    rng = np.random.default_rng(1)
    ns = len(t["left"])
    lengths = rng.uniform(1e-4, 5e-4, ns)
    seg_len = np.array([_f32sqrt(np.sum((t["left"][s] - t["right"][s]) ** 2)) for s in range(ns)])
    t["T"][:] = T_exact
    t["heat_exchange"][:, 0] = Tinf * lengths
    t["heat_exchange"][:, 1] = lengths
    t["fric_heat"][:] = q_per_length * seg_len
    o.call("o_fem_reset_server", "t")
    o.call("o_fem_apply_robin", "ct")
    o.call("o_fem_calculate_temp_rate", "t")
    # scale: the rate one Robin node would see with the wall at T_D
    scale = np.abs(t["Minv"] * h * Ly / ns * (Tinf - TD)).max()
    assert np.abs(t["k"]).max() < 1e-6 * scale
    # a profile with the wrong slope is not steady
    t["T"][:] = TD + 0.9 * a * t["node_pos"][:, 0]
    o.call("o_fem_reset_server", "t")
    o.call("o_fem_apply_robin", "ct")
    o.call("o_fem_calculate_temp_rate", "t")
    assert np.abs(t["k"]).max() > 1e-2 * scale


def test_robin_dirichlet_profile_is_reached_by_the_explicit_update():
    """The explicit midpoint update of the step (calculate_temp at dt/2, rate, calculate_temp at dt)
    driven from a uniform T_D converges to the same linear profile."""
    Lx, Ly = 0.02, 0.006
    tc0 = cases.make_constants("Ti6Al4V", 2.5e-4, 0.35).tool
    h = 2.0 * tc0["k"] / Lx
    tm = _strip_tool(8, 3, Lx, Ly, dict(tc0))
    o = _strip_oracle(tm, dict(h=h))
    t = o.t
    TD, Tinf = tc0["T0"], 800.0
    a = h * (Tinf - TD) / (tc0["k"] + h * Lx)
    T_exact = TD + a * t["node_pos"][:, 0]
    rows = np.repeat(np.arange(len(t["M"])), np.diff(t["row_ptr"]))
    gersh = t["Minv"] * np.bincount(rows, weights=np.abs(t["K"]), minlength=len(t["M"]))
    dt = 1.0 / gersh[t["M"] != 1.0].max()          # within the explicit update's stability limit
    for _ in range(20000):
        t["heat_exchange"][:, 0] = Tinf * 1e-4
        t["heat_exchange"][:, 1] = 1e-4
        o.call("o_fem_reset_server", "t")
        o.call("o_fem_calculate_temp", "t", 0.5 * dt)
        o.call("o_fem_apply_robin", "ct")
        o.call("o_fem_calculate_temp_rate", "t")
        o.call("o_fem_calculate_temp", "t", dt)
    assert np.abs(t["T"] - T_exact).max() < 1e-3 * (Tinf - TD)


# ---------------------------------------------------------------- tool_FEM.cu:589-616 mesh shift
def test_mesh_shift_profile_hand_worked():
    """remesh_low_level's node shift at chosen points of the strip, with s = min(flank_retreat * freq,
    -6.5e-6), wy = (x - tl.x) / (br.x - tl.x), wx = min(1, (y - tl.y) / (br.y - tl.y)), reduced by hand:
      wy = 0                      -> no shift
      wx = 0                      -> dx = 1.45 s (0.94^0.7 - 0.94^0.55) wy^3,            dy = 0
      wx = 1/2, wy = 1/2          -> dx = 1.45 s (0.94^0.7 - 0.44^0.55) / 8,             dy = -0.9 s / 2^12
      wx = 1 (clamped)            -> dx = 1.45 s (0.94^0.7 - 0.06^0.55) wy^2,            dy = -0.9 s wy^6 (wy < 1.01)
      wy > 1.01                   -> dy = -0.9 s wx^4 (the sixth power of wy capped at 1)
    cutting-edge and Dirichlet nodes stay; a retreat smaller than 6.5e-6 / freq is raised to it."""
    Lx, Ly = 0.02, 0.008
    tc0 = cases.make_constants("Ti6Al4V", 2.5e-4, 0.35).tool
    tm = _strip_tool(8, 8, Lx, Ly, dict(tc0))
    for fr, freq, s in ((-1e-8, 20, -6.5e-6), (-1e-6, 20, -2e-5)):
        o = _strip_oracle(tm, {})
        # tl = (0, 0), br = (Lx/2, Ly/2): wy = 2x / Lx, wx = min(1, 2y / Ly)
        o.Tc.tl_x, o.Tc.tl_y, o.Tc.br_x, o.Tc.br_y = 0.0, 0.0, Lx / 2, Ly / 2
        o.Tc.flank_retreat = fr
        p0 = o.t["node_pos"].copy()
        o.remesh_low_level(freq)
        d = o.t["node_pos"] - p0
        flags = o.t["node_flags"]
        c7, c55 = 0.94 ** 0.7, 0.94 ** 0.55
        checked = 0
        for i, (x, y) in enumerate(p0):
            wy, wx_raw = 2 * x / Lx, 2 * y / Ly
            if flags[i] & 3:
                assert np.all(d[i] == 0.0)
                continue
            if wy == 0.0:
                want = (0.0, 0.0)
            elif wx_raw == 0.0:
                want = (1.45 * s * (c7 - c55) * wy ** 3, 0.0)
            elif wx_raw == 0.5 and wy == 0.5:
                want = (1.45 * s * (c7 - 0.44 ** 0.55) / 8, -0.9 * s / 4096)
            elif wx_raw >= 1.0 and wy < 1.01:
                want = (1.45 * s * (c7 - 0.06 ** 0.55) * wy ** 2, -0.9 * s * wy ** 6)
            elif wy > 1.01 and wx_raw <= 0.94:
                want = (None, -0.9 * s * wx_raw ** 4)
            elif wy > 1.01:
                want = (1.45 * s * (c7 - 0.06 ** 0.55) * wy ** 2, -0.9 * s)
            else:
                continue
            ulp = 4e-18                  # rounding of the shifted coordinate (|x|, |y| <= 0.02)
            if want[0] is not None:
                assert d[i, 0] == pytest.approx(want[0], rel=1e-12, abs=ulp), (i, x, y)
            assert d[i, 1] == pytest.approx(want[1], rel=1e-12, abs=ulp), (i, x, y)
            checked += 1
        assert checked >= 30
