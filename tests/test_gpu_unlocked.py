"""Unlocked grids: grid_gpu_green(max_cell, N) / grid_gpu_rothlin(max_cell, N) with the geometry
rebuilt from the particles' bounding box by update_geometry (grid.cu:88-127), which the step calls
first (leap_frog.cu:175).  The device reduction must give the geometry of the oracle's sequential
scan exactly (min and max are exact in any order), and steps on an unlocked grid follow the
oracle's unlocked step."""
import dataclasses

import numpy as np
import pytest

from helpers import assert_equal, max_rel_err
from oracle.oracle import Oracle
from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "S", "rho", "T", "eps_pl", "eps_pl_dot", "fc", "ft", "n", "p", "R")
DERIVS = ("pos_t", "vel_t", "rho_t", "S_t", "T_t")


def _geom_tuple(g):
    return (g.nx, g.ny, g.bbmin_x, g.bbmin_y, g.dx)


@pytest.mark.parametrize("N", [1, 777, 1_500_000])
def test_update_geometry_vs_rule(N):
    ctx = S.Context()
    rng = np.random.default_rng(N)
    pos = rng.uniform(-0.3, 0.7, (N, 2)) * np.array([1.0, 0.4])
    h = rng.uniform(1e-4, 2e-4, N) if N > 1 else np.array([1.5e-3])
    blanked = (rng.uniform(size=N) < 0.05).astype(np.float64)   # blanked particles count too
    p = S.particle_gpu(ctx, N, pos=pos, h=h, blanked=blanked)
    g = S.grid_gpu_green.unlocked(ctx, 2 ** 31 - 1, N)
    assert not g.is_locked() and g.max_cell == 2 ** 31 - 1
    for kw in (2.0, 3.0):
        g.update_geometry(p, kw)
        want = cases.grid_geometry(pos.min(axis=0), pos.max(axis=0), h.max() * kw / 2)
        assert _geom_tuple(g.geom) == (want["nx"], want["ny"], want["bbmin_x"], want["bbmin_y"], want["dx"])
        assert g.num_cell == want["num_cell"]
        assert (g.bbmax_x(), g.bbmax_y()) == (pos[:, 0].max() + 1e-6, pos[:, 1].max() + 1e-6)
    g.assign_hashes(p)
    hs = p.download("hash")
    live = blanked != 1.0
    assert hs[live].min() >= 0 and hs[live].max() < g.num_cell
    p.free()


def test_unlocked_capacity_assign_order_and_graph_refused():
    case = cases.small_cutting_case(nx=24, ny=8)
    ctx = S.Context()
    p = S.particle_gpu(ctx, case.N, pos=case.pos, h=case.hh)
    g = S.grid_gpu_green.unlocked(ctx, 10, case.N)
    with pytest.raises(_lib.SphfemError):
        g.assign_hashes(p)                 # no geometry yet
    with pytest.raises(_lib.SphfemError):
        g.update_geometry(p)               # needs more than 10 cells
    sim = S.Simulation(case, max_cell=10 ** 6)
    sim.step(1)
    with pytest.raises(_lib.SphfemError):
        sim.capture(steps=2)
    sim.ctx.check_error()
    p.free()


def _moving_case():
    case = cases.perturbed(cases.small_cutting_case(nx=48, ny=16, wear=True), vel_sd=0.02)
    return case


@pytest.mark.parametrize("grid,nsteps", [("green", 1), ("green", 6), ("rothlin", 6)])
def test_unlocked_step_matches_oracle(grid, nsteps):
    case = _moving_case()
    max_cell = 200_000
    o = Oracle(case, grid=grid, max_cell=max_cell)
    sim = S.Simulation(case, grid=grid, max_cell=max_cell)
    geoms = []
    for k in range(nsteps):
        o.step(1)
        sim.step(1)
        gg, go = _geom_tuple(sim.grid.geom), _geom_tuple(o.c.geom)
        if k == 0:
            assert gg == go                      # same starting state: the same geometry exactly
        else:
            # later boxes come from states that differ at the libm rounding level
            assert gg[0] == go[0] and gg[1] == go[1] and gg[4] == go[4], k
            assert abs(gg[2] - go[2]) <= 1e-13 and abs(gg[3] - go[3]) <= 1e-13, k
        geoms.append(go)
    sim.ctx.check_error()
    assert nsteps == 1 or len(set(geoms)) > 1           # the box moved
    a = sim.state()
    b = o.P
    for f in ("hash", "idx", "blanked"):
        assert_equal(f, a[f], b[f])
    live = b["blanked"] != 1.0
    # the perturbed state has stress from the start: the artificial stress (float atan2/cos/sin) and
    # what follows from it differ at the 1e-7 level already after one step
    tol = 1e-6
    errs = {f: max_rel_err(a[f], b[f], live) for f in STATE + DERIVS}
    print(grid, nsteps, {k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) <= tol, errs


@pytest.mark.parametrize("prec", [_lib.FP64, _lib.FP32C])
def test_unlocked_equals_locked_when_the_box_is_the_same(prec):
    """A body at rest (zero velocity and stress, uniform density and temperature: every rate is 0):
    the unlocked geometry stays that of the locked grid built on the particles' exact box, so every
    step agrees with the locked grid bit for bit, and the unchanged geometry keeps the incremental
    re-sort path."""
    base = cases.chip_case(n_target=40_000)
    case = dataclasses.replace(base, vel=np.zeros_like(base.vel), S=None, grid_bbmin=base.pos.min(axis=0),
                               grid_bbmax=base.pos.max(axis=0))
    lk = S.Simulation(case, precision=prec)
    un = S.Simulation(case, precision=prec, max_cell=10 ** 7)
    for k in range(3):
        lk.step(1)
        un.step(1)
    lk.ctx.check_error()
    un.ctx.check_error()
    assert un.grid.resort_stats()[0] == 2
    a, b = un.state(), lk.state()
    for f in a:
        assert_equal(f, a[f], b[f])
    lk.close()
    un.close()
