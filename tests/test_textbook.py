"""Textbook validation cases of the reference (benchmarks.cpp:16-246) on the oracle: the whole
step pinned by physics the solution must show.

* rubber rings (Gray, Monaghan & Swift 2001): total momentum stays zero to roundoff, the motion
  stays mirror-symmetric about x = 0 and about y = 0, and the rings rebound (each ring's mean
  velocity reverses) without any particle lost;
* plastic impact of a steel disc on a wall: the motion stays mirror-symmetric about y = 0, the
  disc decelerates, plastic strain develops where it strikes, and the wall's fixed rows stay put.
"""
import dataclasses

import numpy as np

from oracle.oracle import Oracle
from sph_fem_2d_cutting_gpu_amd import cases


def _mirror_map(pos, axis):
    """index of each particle's mirror image across axis (0: x -> -x, 1: y -> -y) on a lattice that
    is mirror-symmetric up to rounding"""
    m = pos.copy()
    m[:, axis] = -m[:, axis]
    key = lambda p: np.lexsort((np.round(p[:, 1], 9), np.round(p[:, 0], 9)))
    ia, ib = key(pos), key(m)
    j = np.empty(len(pos), dtype=np.int64)
    j[ia] = ib
    assert np.allclose(pos[j, axis], -pos[:, axis], atol=1e-12, rtol=0)
    return j


def _by_label(P, labels):
    """state arrays reordered by particle label (labels[s] = label of sorted slot s)"""
    slot = np.empty(len(labels), np.int64)
    slot[labels] = np.arange(len(labels))
    return {k: v[slot] for k, v in P.items()}


def test_rings_construction():
    c = cases.rings(80)
    assert c.N == 4232 and c.N % 2 == 0
    r = np.hypot(np.abs(c.pos[:, 0]) - 0.041, c.pos[:, 1])
    assert np.all((r >= 0.03 - 1e-12) & (r < 0.04))
    assert np.all(c.vel[c.pos[:, 0] < 0, 0] == 200) and np.all(c.vel[c.pos[:, 0] > 0, 0] == -200)
    assert c.const.joco is None and c.const.trml_wp["alpha"] == 0.0


def test_rings_momentum_symmetry_and_rebound():
    """T is a passive scalar here (no heat, no plasticity), so it carries each particle's label
    through the sorts."""
    case = dataclasses.replace(cases.rings(80), T=np.arange(cases.rings(80).N, dtype=np.float64))
    jx, jy = _mirror_map(case.pos, 0), _mirror_map(case.pos, 1)
    ring_left = case.pos[:, 0] < 0
    o = Oracle(case)
    n = int(round(case.extra["t_final"] / case.dt))
    m = case.const.phys["mass"]
    # a quarter into the run (the rings in contact since step ~25): symmetric to roundoff
    o.step(n // 4)
    Q = _by_label(o.P, o.P["T"].astype(np.int64))
    assert np.max(np.abs(Q["pos"][jx, 0] + Q["pos"][:, 0])) < 1e-9
    assert np.max(np.abs(Q["pos"][jx, 1] - Q["pos"][:, 1])) < 1e-9
    assert np.max(np.abs(Q["vel"][jx, 0] + Q["vel"][:, 0])) < 1e-6 * 200
    assert np.max(np.abs(Q["pos"][jy, 1] + Q["pos"][:, 1])) < 1e-9
    assert np.max(np.abs(Q["vel"][jy, 1] + Q["vel"][:, 1])) < 1e-6 * 200
    o.step(n - n // 4)
    Q = _by_label(o.P, o.P["T"].astype(np.int64))
    p_tot = (m * Q["vel"]).sum(axis=0)
    assert np.all(np.abs(p_tot) <= 1e-12 * m * 200 * case.N), p_tot
    assert np.all(Q["blanked"] == 0) and np.all(np.isfinite(Q["pos"]))
    # rebound: each ring moves back out with most of its speed, and the rings stay apart
    vl, vr = Q["vel"][ring_left, 0].mean(), Q["vel"][~ring_left, 0].mean()
    assert vl < -100.0 and vr > 100.0 and abs(vl + vr) < 1e-9 * 200
    assert Q["pos"][ring_left, 0].max() < 0 < Q["pos"][~ring_left, 0].min()


def test_impact_symmetry_deceleration_and_plastic_zone():
    """On the Rothlin grid (idx carries labels; tool-free steps equal the Green grid's)."""
    case = cases.impact(20)
    jy = _mirror_map(case.pos, 1)
    o = Oracle(case, grid="rothlin")
    nf = case.extra["n_flyer"]
    n = int(round(case.extra["t_final"] / case.dt))
    fixed0 = case.pos[case.fixed == 1].copy()
    o.step(n // 4)
    Q = _by_label(o.P, o.P["idx"])
    assert np.max(np.abs(Q["pos"][jy, 1] + Q["pos"][:, 1])) < 1e-9
    assert np.max(np.abs(Q["pos"][jy, 0] - Q["pos"][:, 0])) < 1e-9
    assert np.max(np.abs(Q["vel"][jy, 1] + Q["vel"][:, 1])) < 1e-6 * 1e3
    assert np.max(np.abs(Q["eps_pl"][jy] - Q["eps_pl"])) < 1e-9
    o.step(n - n // 4)
    P = o.P
    assert np.all(P["blanked"] == 0)
    pos, vel = P["pos"], P["vel"]
    m = case.const.phys["mass"]
    assert abs((m * vel[:, 1]).sum()) < 1e-9 * m * 1e3 * case.N
    # the fixed rows: the boundary condition zeroes their velocity after every step; they drift only
    # by the XSPH term of the corrector, which the condition zeroes after it was applied
    fx = P["fixed"] == 1
    assert fx.sum() == len(fixed0) and np.all(vel[fx] == 0.0)
    d = np.abs(np.sort(pos[fx], axis=0) - np.sort(fixed0, axis=0))
    assert 0 < d.max() < 0.01 * case.dx
    # the disc (particles left of the wall's original face or moving fast) has slowed down
    ke = 0.5 * m * (vel ** 2).sum()
    assert ke < 0.5 * m * nf * 1e6 * 0.8
    assert P["eps_pl"].max() > 0.5
