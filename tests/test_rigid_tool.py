"""The rigid (SPH) tool on the CPU: the oracle's restatement of tool.cpp / tool_gpu.cu /
actions_gpu.cu:316-343 pinned against geometry and closed forms, and the library's host geometry
(the constructors of class tool) checked bit for bit against the oracle's.

Pins that do not retype the oracle's formulas:
  * constructors: right angles give the rectangle; the fillet centre is at distance r from both
    faces it is tangent to, its tangent points at distance r from the centre, t1 = pi for a vertical
    rake face and t2 = 90 deg - clearance (the direction of the clearance face's inner normal);
  * inside test and sampling: an independent point-in-convex-polygon test by cross products;
  * contact: the contact point lies on the boundary, the normal is the outward unit normal of the
    chosen face, no face line is closer than the chosen one (distance by |cross| / length), and the
    penalty force is slave_mass * depth * alpha / dt^2 along the normal;
  * friction: the stick branch returns ft_old - alpha m v_r / dt, the slip branch has magnitude
    mu |fN| along the trial force, and v_r keeps the reference's component-wise v - v*n;
  * friction heat: eta dt |ft| |v_t| / (cp m) with the tangential velocity by projection.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import oracle as O
from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S


def _otool_angles(tl, L, H, rake, clear, r, mu=0.35):
    t = O.ORTool()
    O.lib().o_rtool_from_angles(C.byref(t), O.OD2(*tl), L, H, rake, clear, r, mu)
    return t


def _segs(t):
    return [((s.left.x, s.left.y), (s.right.x, s.right.y), s.l.vertical, (s.n.x, s.n.y))
            for s in t.seg[: t.num_seg]]


def _line_dist(p1, p2, q):
    """distance from q to the line through p1, p2 (|cross| / length)"""
    d = np.subtract(p2, p1)
    w = np.subtract(q, p1)
    return abs(d[0] * w[1] - d[1] * w[0]) / math.hypot(*d)


def _inside_convex(t, q):
    """independent inside test: q strictly right of every directed edge (clockwise polygon), or
    within the fillet circle"""
    ok = True
    for (a, b, _, _) in _segs(t):
        cross = (b[0] - a[0]) * (q[1] - a[1]) - (b[1] - a[1]) * (q[0] - a[0])
        ok = ok and cross < 0
    if t.has_fillet:
        ok = ok or math.hypot(q[0] - t.fillet.p.x, q[1] - t.fillet.p.y) < t.fillet.r
    return ok


# ---------------------------------------------------------------- tool.cpp:335-406
def test_right_angles_give_the_rectangle():
    tl = (0.1, 0.2)
    t = _otool_angles(tl, 0.04, 0.03, 0.0, 0.0, 0.0)
    assert t.num_seg == 4 and not t.has_fillet
    corners = [s[0] for s in _segs(t)]
    expect = [(0.1, 0.2), (0.14, 0.2), (0.14, 0.17), (0.1, 0.17)]
    assert np.allclose(corners, expect, atol=1e-15, rtol=0)
    normals = [s[3] for s in _segs(t)]
    assert np.allclose(normals, [(0, -1), (-1, 0), (0, 1), (1, 0)], atol=1e-12)   # into the tool
    assert [s[2] for s in _segs(t)] == [0, 1, 0, 1]
    assert t.front == pytest.approx(0.14, abs=1e-15)


@pytest.mark.parametrize("rake,clear,r", [(0.0, 4.5, 0.00035), (6.0, 7.0, 0.0005), (-5.0, 11.0, 0.001)])
def test_fillet_is_tangent_to_both_faces(rake, clear, r):
    tl = (-0.046, 0.0986074)
    sharp = _otool_angles(tl, 0.0413176, 0.0431, rake, clear, 0.0)
    t = _otool_angles(tl, 0.0413176, 0.0431, rake, clear, r)
    assert t.num_seg == 5 and t.has_fillet
    tr, br, bl = sharp.seg[1].left, sharp.seg[2].left, sharp.seg[3].left
    p = (t.fillet.p.x, t.fillet.p.y)
    assert _line_dist((tr.x, tr.y), (br.x, br.y), p) == pytest.approx(r, rel=1e-9)
    assert _line_dist((bl.x, bl.y), (br.x, br.y), p) == pytest.approx(r, rel=1e-9)
    trc, blc = t.seg[2].left, t.seg[3].left
    assert math.hypot(trc.x - p[0], trc.y - p[1]) == pytest.approx(r, rel=1e-9)
    assert math.hypot(blc.x - p[0], blc.y - p[1]) == pytest.approx(r, rel=1e-9)
    # the tangent points lie on the faces, between the sharp corner and the face ends
    assert _line_dist((tr.x, tr.y), (br.x, br.y), (trc.x, trc.y)) < 1e-12
    assert _line_dist((bl.x, bl.y), (br.x, br.y), (blc.x, blc.y)) < 1e-12
    # angles of centre - tangent point: the inner normals of the rake and clearance faces
    assert t.fillet.t1 == pytest.approx(math.pi + math.radians(rake), abs=1e-9)
    assert t.fillet.t2 == pytest.approx(math.radians(90.0 - clear), abs=1e-9)
    assert t.front == sharp.front                       # m_front is the sharp corner's x
    assert O.lib().o_rtool_low(C.byref(t)) == pytest.approx(p[1] - r, abs=0)


def test_points_constructor_matches_angles_constructor():
    a = _otool_angles((0.0, 0.05), 0.04, 0.045, 3.0, 6.0, 0.0005)
    sharp = _otool_angles((0.0, 0.05), 0.04, 0.045, 3.0, 6.0, 0.0)
    pts = [O.OD2(s.left.x, s.left.y) for s in sharp.seg[:4]]
    b = O.ORTool()
    O.lib().o_rtool_from_points(C.byref(b), *pts, 0.0005, 0.35)
    assert bytes(a) == bytes(b)


def test_move_translates_and_keeps_normals():
    t = _otool_angles((0.0, 0.05), 0.04, 0.045, 0.0, 4.5, 0.0005)
    t0 = O.ORTool()
    C.pointer(t0)[0] = t
    t.vel = O.OD2(2e-3, -1e-3)
    O.lib().o_rtool_move(C.byref(t), 7.0)
    for s0, s1 in zip(_segs(t0), _segs(t)):
        assert np.allclose(np.subtract(s1[0], s0[0]), (0.014, -0.007), atol=1e-15)
        assert np.allclose(np.subtract(s1[1], s0[1]), (0.014, -0.007), atol=1e-15)
        assert s1[3] == s0[3] and s1[2] == s0[2]
    assert (t.fillet.p.x - t0.fillet.p.x, t.fillet.p.y - t0.fillet.p.y) == pytest.approx((0.014, -0.007), abs=1e-15)
    for s in t.seg[:5]:
        if not s.l.vertical:                            # rebuilt lines pass through the moved ends
            assert s.l.a * s.left.x + s.l.b == pytest.approx(s.left.y, abs=1e-14)
            assert s.l.a * s.right.x + s.l.b == pytest.approx(s.right.y, abs=1e-14)


@pytest.mark.parametrize("r", [0.0, 0.0015])
def test_inside_and_sample_match_convex_polygon_test(r):
    t = _otool_angles((0.0, 0.05), 0.02, 0.018, 5.0, 6.0, r)
    rng = np.random.default_rng(3)
    bb = (C.c_double * 4)()
    O.lib().o_rtool_safe_bb(C.byref(t), 0.0, bb)
    q = rng.uniform((bb[0] - 0.002, bb[2] - 0.002), (bb[1] + 0.002, bb[3] + 0.002), (4000, 2))
    got = [O.lib().o_rtool_inside(C.byref(t), x, y) for x, y in q]
    exp = [_inside_convex(t, p) for p in q]
    assert got == [int(e) for e in exp]
    dx = 5e-4
    n = O.lib().o_rtool_sample(C.byref(t), dx, None, 0)
    pts = np.zeros((n, 2))
    O.lib().o_rtool_sample(C.byref(t), dx, pts.ctypes.data, n)
    assert n > 100 and all(_inside_convex(t, p) for p in pts)
    # every lattice point of the box that is inside was sampled
    O.lib().o_rtool_safe_bb(C.byref(t), 0.1 * dx, bb)
    nx, ny = int((bb[1] - bb[0]) / dx + 1), int((bb[3] - bb[2]) / dx + 1)
    lat = [(i * dx + bb[0], j * dx + bb[2]) for i in range(nx) for j in range(ny)]
    assert n == sum(_inside_convex(t, p) for p in lat)


# ---------------------------------------------------------------- tool_gpu.cu:82-202
def _contact_setup(r):
    t = _otool_angles((0.0, 0.05), 0.02, 0.018, 5.0, 6.0, r)
    t.contact_alpha = 0.075
    t.slave_mass = 4.43e-6
    t.vel = O.OD2(1.2e-3, 0.0)
    return t


@pytest.mark.parametrize("r", [0.0, 0.0015])
def test_contact_point_normal_and_depth(r):
    t = _contact_setup(r)
    rng = np.random.default_rng(5)
    bb = (C.c_double * 4)()
    O.lib().o_rtool_safe_bb(C.byref(t), 0.0, bb)
    checked = 0
    for x, y in rng.uniform((bb[0], bb[2]), (bb[1], bb[3]), (3000, 2)):
        cp, n = O.OD2(), O.OD2()
        inside = O.lib().o_rtool_establish_contact(C.byref(t), x, y, C.byref(cp), C.byref(n))
        assert inside == _inside_convex(t, (x, y))
        if not inside:
            continue
        checked += 1
        depth = math.hypot(cp.x - x, cp.y - y)
        assert math.hypot(n.x, n.y) == pytest.approx(1.0, abs=1e-14)
        # the normal points from the particle to the contact point (outward)
        assert (cp.x - x) * n.x + (cp.y - y) * n.y == pytest.approx(depth, rel=1e-9, abs=1e-18)
        on_fillet = r > 0 and abs(math.hypot(cp.x - t.fillet.p.x, cp.y - t.fillet.p.y) - r) < 1e-12 and depth > 0
        # no face line nearer than the chosen primitive (the chord of the fillet excluded)
        dists = [_line_dist(a, b, (x, y)) for k, (a, b, _, _) in enumerate(_segs(t)) if not (r > 0 and k == 2)]
        if on_fillet:
            assert depth <= min(dists) * (1 + 1e-9)
            assert depth == pytest.approx(r - math.hypot(x - t.fillet.p.x, y - t.fillet.p.y), rel=1e-9, abs=1e-17)
        else:
            assert depth == pytest.approx(min(dists), rel=1e-9, abs=1e-17)
    assert checked > 500


def _contact_one(t, pos, vel, ft_old, dt):
    P = {k: np.zeros((1,) + shp, dtype=dt_) for k, shp, dt_ in cases.PARTICLE_FIELDS}
    P["pos"][0], P["vel"][0], P["ft"][0] = pos, vel, ft_old
    Pc = O.OParticles()
    Pc.N = 1
    for k in P:
        setattr(Pc, k, P[k].ctypes.data_as(C.c_void_p))
    c = O.OConsts()
    c.dt = dt
    O.lib().o_rtool_contact(C.byref(c), C.byref(Pc), C.byref(t))
    return P


def test_penalty_force_and_friction_branches():
    t = _contact_setup(0.0)
    t.mu = 0.35
    dt = 8.86e-4
    # a point at depth d below the top face, far from the other faces: outward normal (0, 1)
    top = t.seg[0].left.y
    x = t.seg[0].left.x + 0.004
    d = 2e-5
    ms, a = t.slave_mass, t.contact_alpha
    # stick: the trial force stays below mu |fN|
    v = (t.vel.x + 1e-9, 0.0)
    P = _contact_one(t, (x, top - d), v, (0.0, 0.0), dt)
    fN = ms * d * a / dt ** 2
    assert P["n"][0] == pytest.approx((0.0, 1.0), abs=1e-15)
    assert P["fc"][0] == pytest.approx((0.0, fN), rel=1e-9)
    kd = a * ms * 1e-9 / dt
    assert P["ft"][0] == pytest.approx((-kd, 0.0), rel=1e-6)
    # slip: magnitude mu |fN| along the trial force
    v = (t.vel.x + 0.5, 0.0)
    P = _contact_one(t, (x, top - d), v, (1e-9, 0.0), dt)
    assert math.hypot(*P["ft"][0]) == pytest.approx(0.35 * fN, rel=1e-12)
    assert P["ft"][0][0] < 0 and P["ft"][0][1] == 0.0
    # v_r = v - v*n component-wise: with n = (0, 1), v_r.y = v.y - v.y = 0, so a purely normal
    # relative velocity gives no friction ...
    P = _contact_one(t, (x, top - d), (t.vel.x, 1e-3), (0.0, 0.0), dt)
    assert P["ft"][0] == pytest.approx((0.0, 0.0), abs=0)
    # ... but with n = (0, -1) (bottom of a rectangle) v_r.y = v.y + v.y = 2 v.y, as written in the
    # reference (a projection v - (v.n) n would give 0)
    rect = _otool_angles((0.0, 0.05), 0.02, 0.018, 0.0, 0.0, 0.0)
    rect.contact_alpha, rect.slave_mass, rect.mu = a, ms, 0.35
    rect.vel = O.OD2(0.0, 0.0)
    yb = rect.seg[2].left.y
    P = _contact_one(rect, (0.01, yb + d), (0.0, 1e-9), (0.0, 0.0), dt)
    assert P["n"][0] == pytest.approx((0.0, -1.0), abs=1e-15)
    assert P["ft"][0][1] == pytest.approx(-a * ms * 2e-9 / dt, rel=1e-9)
    # outside: everything zero
    P = _contact_one(t, (x, top + d), v, (1.0, 1.0), dt)
    assert not P["fc"].any() and not P["ft"].any() and not P["n"].any()


def test_friction_heat_closed_form():
    rng = np.random.default_rng(9)
    n = 64
    P = {k: np.zeros((n,) + shp, dtype=dt_) for k, shp, dt_ in cases.PARTICLE_FIELDS}
    P["vel"][:] = rng.normal(0, 1e-3, (n, 2))
    P["ft"][:] = rng.normal(0, 1e-6, (n, 2))
    P["ft"][:4] = 0.0
    ang = rng.uniform(0, 2 * np.pi, n)
    P["n"][:] = np.stack([np.cos(ang), np.sin(ang)], 1)
    P["T"][:] = 300.0
    P["tool_particle"][n - 4:] = 1.0
    Pc = O.OParticles()
    Pc.N = n
    for k in P:
        setattr(Pc, k, P[k].ctypes.data_as(C.c_void_p))
    c = O.OConsts()
    c.dt, c.trml_wp.eta, c.trml_wp.cp, c.phys.mass = 8e-4, 0.9, 580e-8, 4.4e-6
    t = O.ORTool()
    t.vel = O.OD2(1.25e-3, 0.0)
    O.lib().o_rtool_fric_heat(C.byref(c), C.byref(Pc), C.byref(t))
    vd = P["vel"] - np.array([1.25e-3, 0.0])
    vt = vd - np.sum(vd * P["n"], 1, keepdims=True) * P["n"]
    q = 0.9 * 8e-4 * np.hypot(*P["ft"].T) * np.hypot(*vt.T) / (580e-8 * 4.4e-6)
    exp = 300.0 + np.where(P["tool_particle"] == 0, q, 0.0)
    assert q.max() > 1e-3
    assert np.allclose(P["T"], exp, rtol=0, atol=1e-12 + 1e-6 * q.max())    # |.| rounded to FP32
    assert np.all(P["T"][:4] == 300.0) and np.all(P["T"][n - 4:] == 300.0)


def test_tool_particles_move_with_the_tool():
    n = 10
    P = {k: np.zeros((n,) + shp, dtype=dt_) for k, shp, dt_ in cases.PARTICLE_FIELDS}
    P["pos"][:] = np.arange(2 * n).reshape(n, 2) * 1e-3
    P["tool_particle"][::2] = 1.0
    P["blanked"][::4] = 1.0
    before = P["pos"].copy()
    Pc = O.OParticles()
    Pc.N = n
    for k in P:
        setattr(Pc, k, P[k].ctypes.data_as(C.c_void_p))
    c = O.OConsts()
    c.dt = 0.5
    t = O.ORTool()
    t.vel = O.OD2(2e-3, 1e-3)
    O.lib().o_rtool_tool_particles(C.byref(c), C.byref(Pc), C.byref(t))
    d = P["pos"] - before
    assert np.allclose(d[::2], (1e-3, 5e-4), atol=1e-18)
    assert np.all(d[1::2] == 0)


# ---------------------------------------------------------------- library host geometry
def _lib_geom_to_otool(g):
    return O.rtool_from_geometry(g, 0.0, 0.0)


@pytest.mark.parametrize("rake,clear,r", [(0.0, 4.5, 0.00035), (6.0, 7.0, 0.0), (-5.0, 11.0, 0.001), (0.0, 0.0, 0.0)])
def test_library_geometry_equals_oracle(rake, clear, r):
    tl = (-0.046, 0.0986074)
    lt = S.tool.from_angles(tl, 0.0413176, 0.0431, rake, clear, r, 0.35)
    ot = _otool_angles(tl, 0.0413176, 0.0431, rake, clear, r)
    assert bytes(_lib_geom_to_otool(lt.g)) == bytes(ot)
    lt.set_vel((1.25e-3, -2e-4))
    ot.vel = O.OD2(1.25e-3, -2e-4)
    for dt in (-3.5, 0.7, 1e-3):
        lt.update_tool(dt)
        O.lib().o_rtool_move(C.byref(ot), dt)
    assert bytes(_lib_geom_to_otool(lt.g)) == bytes(ot)
    assert lt.low() == O.lib().o_rtool_low(C.byref(ot))
    bb = (C.c_double * 4)()
    O.lib().o_rtool_safe_bb(C.byref(ot), 1e-3, bb)
    assert list(lt.safe_bb(1e-3).values()) == list(bb)
    dx = 1.0345e-3
    n = O.lib().o_rtool_sample(C.byref(ot), dx, None, 0)
    pts = np.zeros((n, 2))
    O.lib().o_rtool_sample(C.byref(ot), dx, pts.ctypes.data, n)
    assert np.array_equal(lt.sample_tool(dx), pts)


def test_ref_cut_case_matches_the_setup():
    """setup_ref_cut at the driver's defaults: lattice, fixed flags, tool placed at the workpiece's
    left edge with its lowest point at hi_y - feed, tool particles, grid box and time step."""
    case = cases.ref_cut_sph()
    n = case.extra["n_workpiece"]
    dx = case.dx
    assert dx == pytest.approx(0.03 / 29, rel=1e-15) and n == 188 * 30
    g = S.tool(case.rigid.geometry)
    assert g.low() == pytest.approx(0.06 - 0.01, abs=1e-12)
    rake_x = g.get_segments()[1]["left"][0]
    assert rake_x == pytest.approx(0.005, abs=1e-12)       # sharp front moved to lo_x
    tp = case.tool_p
    assert np.all(tp[:n] == 0) and np.all(tp[n:] == 1) and case.N - n > 1000
    assert all(g.inside(p) for p in case.pos[n:])
    v = 83.333328 * 1e-5 / 200 * 300                  # mfree_gpu.cpp:197 at speed_in_ui = 300
    assert case.tool_vel[0] == v and case.rigid.geometry.vel.x == v
    assert case.dt == pytest.approx(0.32 * 1.5 * dx / (math.sqrt(case.const.phys["K"] / 4.429998) + math.sqrt(v)),
                                    rel=1e-15)
    assert np.all(case.pos > case.grid_bbmin) and np.all(case.pos < case.grid_bbmax)
    assert case.grid_bbmax[1] >= 0.06 + 0.195 - 1e-12
    # fixed: bottom row, lower half of the left column, right column; tool: left and top edges
    wp = case.pos[:n]
    f = case.fixed[:n] == 1
    assert np.all(f[wp[:, 1] < wp[:, 1].min() + 0.5 * dx])
    assert np.all(f[wp[:, 0] > wp[:, 0].max() - dx])
    assert not f[(wp[:, 0] > 0.01) & (wp[:, 0] < 0.19) & (wp[:, 1] > 0.035)].any()
    tpos = case.pos[n:]
    tf = case.fixed[n:] == 1
    assert np.all(tf == ((tpos[:, 0] < tpos[:, 0].min() + dx / 2) | (tpos[:, 1] > tpos[:, 1].max() - dx / 2)))


def test_library_exports_rigid_symbols():
    L = _lib.lib()
    for name in ("sphfem_tool_gpu_set_up_tool", "sphfem_tool_gpu_update_tool", "sphfem_compute_contact_forces",
                 "sphfem_material_fric_heat_gen", "sphfem_leap_frog_step_rigid",
                 "sphfem_interactions_setup_thermal_constants_tool"):
        assert hasattr(L, name)
