"""grid_gpu_rothlin (grid_gpu_rothlin.cu:19-249) on the device against numpy, the oracle and the Green grid.

The Rothlin sort is the same stable sort as the Green one, with in-place gathers: idx then carries
particle labels, ft and n stay in place, the array addresses do not change.  Its cell table holds
dense offsets (cell_start[c] = number of sorted keys below c, cell_end[c] = below c + 1), which the
neighbour loops read exactly as the Green table, so interactions agree bit for bit.  A step with a
Rothlin grid follows the oracle's step with the Rothlin sort.
"""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from helpers import assert_equal, max_rel_err
from oracle.oracle import Oracle
from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu

GATHERED = ("pos", "vel", "fc", "h", "rho", "S", "eps_pl", "eps_pl_dot", "fixed", "blanked", "tool_particle", "T",
            "pos_t", "vel_t", "rho_t", "S_t", "T_t", "on_seg")
STATE = ("pos", "vel", "S", "rho", "T", "eps_pl", "eps_pl_dot", "fc", "ft", "n", "p", "R")
DERIVS = ("pos_t", "vel_t", "rho_t", "S_t", "T_t")


def _random_fields(rng, N):
    f = {"pos": None}
    for k, shp in (("vel", (N, 2)), ("fc", (N, 2)), ("ft", (N, 2)), ("n", (N, 2)), ("S", (N, 4)),
                   ("pos_t", (N, 2)), ("vel_t", (N, 2)), ("S_t", (N, 4))):
        f[k] = rng.normal(0, 1, shp)
    for k in ("h", "rho", "eps_pl", "eps_pl_dot", "fixed", "tool_particle", "T", "rho_t", "T_t"):
        f[k] = rng.normal(0, 1, N)
    f["on_seg"] = rng.integers(-1, 300, N).astype(np.int32)
    f["idx"] = (rng.permutation(N) + 17).astype(np.int32)
    del f["pos"]
    return f


@pytest.mark.parametrize("N,cells_side", [(300_000, 150), (1_000_000, 3000)])
def test_rothlin_sort_and_cells_vs_numpy(N, cells_side):
    """150^2 cells: two 8-bit radix passes; 3000^2 cells: three (odd, the keys start in the scratch
    buffer).  A tenth of the particles crowd a few cells; blanked particles keep seeded hashes."""
    ctx = S.Context()
    rng = np.random.default_rng(N)
    h = 1.5e-4
    L = cells_side * h
    pos = rng.uniform(0, L, (N, 2))
    pos[: N // 10] = rng.uniform(0, 3 * h, (N // 10, 2))
    blanked = (rng.uniform(size=N) < 0.02).astype(np.float64)
    fields = _random_fields(rng, N)
    p = S.particle_gpu(ctx, N, pos=pos, blanked=blanked)
    g = S.grid_gpu_rothlin(ctx, N, pos.min(axis=0), pos.max(axis=0), h)
    for k, v in fields.items():
        p.upload(k, v)
    old = (np.arange(N, dtype=np.int64) * 7919 % g.num_cell).astype(np.int32)
    p.upload("hash", old)
    g.assign_hashes(p)
    keys = p.download("hash")
    geo = g.geom
    ix = ((pos[:, 0] - geo.bbmin_x) / geo.dx).astype(np.int32)
    iy = ((pos[:, 1] - geo.bbmin_y) / geo.dx).astype(np.int32)
    assert_equal("hash", keys, np.where(blanked == 1.0, old, ix * geo.ny + iy).astype(np.int32))
    before = {k: p.download(k) for k in GATHERED + ("ft", "n", "idx")}
    addr = {k: p.ptr(k) for k in GATHERED + ("ft", "n", "idx", "hash")}
    g.sort(p)
    ctx.check_error()
    perm = np.argsort(keys, kind="stable")
    sk = keys[perm]
    assert_equal("sorted hash", p.download("hash"), sk)
    assert_equal("idx labels", p.download("idx"), before["idx"][perm])
    for k in GATHERED:
        assert_equal(k, p.download(k), before[k][perm])
    for k in ("ft", "n"):
        assert_equal(k + " (not gathered)", p.download(k), before[k])
    assert {k: p.ptr(k) for k in addr} == addr                      # in place
    cs, ce = g.download_cells(p)
    lb = np.searchsorted(sk, np.arange(g.num_cell + 1), side="left").astype(np.int32)
    assert_equal("cell_start", cs, lb[:-1])
    assert_equal("cell_end", ce, lb[1:])
    p.free()
    ctx.call("sphfem_grid_destroy", g.h)


def test_rothlin_get_cells_needs_particles():
    ctx = S.Context()
    g = S.grid_gpu_rothlin(ctx, 100, (0.0, 0.0), (1e-2, 1e-2), 1e-3)
    with pytest.raises(_lib.SphfemError):
        g.download_cells(None)
    kind = C.c_int(-1)
    assert _lib.lib().sphfem_grid_kind(g.h, C.byref(kind)) == 0 and kind.value == 1
    ctx.call("sphfem_grid_destroy", g.h)


def _upload_oracle(sim, o):
    for f in o.P:
        sim.p.upload(f, o.P[f])


def test_rothlin_sort_matches_oracle_and_interactions_match_green():
    case = cases.small_cutting_case(nx=48, ny=16, wear=True)
    o = Oracle(case, grid="rothlin")
    o.step(20)
    rng = np.random.default_rng(2)
    o.P["idx"][:] = rng.permutation(case.N).astype(np.int32)
    o.call("o_compute_hashes", "cp")
    sim = S.Simulation(case, grid="rothlin")
    _upload_oracle(sim, o)
    o.call("o_sort_rothlin", "cps")
    sim.grid.sort(sim.p)
    sim.ctx.check_error()
    for f in ("hash", "idx") + GATHERED + ("ft", "n"):
        assert_equal(f, sim.p.download(f), o.P[f])
    cs, ce = sim.grid.download_cells(sim.p)
    assert_equal("cell_start", cs, o.st["cell_start"])
    assert_equal("cell_end", ce, o.st["cell_end"])
    # the Green table of the same sorted keys: -1 / stale for empty cells, the same ranges elsewhere
    sk = o.P["hash"]
    occ = np.unique(sk)
    assert np.all(cs[occ] < ce[occ])
    empty = np.setdiff1d(np.arange(len(cs)), occ)
    assert np.all(cs[empty] == ce[empty])
    green_cs = np.full(len(cs), -1, np.int32)
    green_ce = np.zeros(len(cs), np.int32)
    green_cs[occ], green_ce[occ] = cs[occ], ce[occ]
    ctx = sim.ctx
    nc = len(cs)
    bufs = [C.c_void_p() for _ in range(4)]
    for b in bufs:
        ctx.call("sphfem_malloc", C.byref(b), 4 * nc)
    ctx.htod(bufs[0].value, cs)
    ctx.htod(bufs[1].value, ce)
    ctx.htod(bufs[2].value, green_cs)
    ctx.htod(bufs[3].value, green_ce)
    out = []
    for a, b in ((bufs[0], bufs[1]), (bufs[2], bufs[3])):
        S.material_eos(sim.p)
        S.corrector_artificial_stress(sim.p)
        S.interactions_heat_pse(sim.p, a.value, b.value, nc)
        ctx.synchronize()
        r = {"T_t (PSE)": sim.p.download("T_t")}
        S.interactions_monaghan(sim.p, a.value, b.value, nc)
        ctx.synchronize()
        r.update({f: sim.p.download(f) for f in ("S_der", "v_der", "pos_t", "vel_t", "T_t")})
        out.append(r)
    for b in bufs:
        ctx.call("sphfem_free", b)
    for f in out[0]:
        assert_equal(f + " (dense vs Green table)", out[0][f], out[1][f])


@pytest.mark.parametrize("nsteps", [1, 6])
def test_rothlin_step_matches_oracle(nsteps):
    case = cases.small_cutting_case(nx=48, ny=16, wear=True)
    o = Oracle(case, grid="rothlin")
    sim = S.Simulation(case, grid="rothlin")
    o.step(nsteps)
    sim.step(nsteps)
    sim.ctx.check_error()
    a = sim.state()
    b = o.P
    for f in ("hash", "idx", "blanked"):
        assert_equal(f, a[f], b[f])
    live = b["blanked"] != 1.0
    tol = 1e-13 if nsteps == 1 else 1e-6
    errs = {f: max_rel_err(a[f], b[f], live) for f in STATE + DERIVS}
    print({k: f"{v:.2e}" for k, v in errs.items()})
    assert max(errs.values()) <= tol, errs
    cs, ce = sim.grid.download_cells(sim.p)
    assert_equal("cell_start", cs, o.st["cell_start"])
    assert_equal("cell_end", ce, o.st["cell_end"])


def _padded_chip(n):
    base = cases.chip_case(n_target=n)
    vscale = 0.004 * base.h / base.dt / max(float(np.sqrt(np.mean(base.vel ** 2))), 1e-30)
    return dataclasses.replace(base, vel=base.vel * vscale, grid_bbmin=base.grid_bbmin - 4 * base.h,
                               grid_bbmax=base.grid_bbmax + 4 * base.h)


@pytest.mark.parametrize("prec", [_lib.FP64, _lib.FP32C])
def test_rothlin_step_equals_green_step_without_tool(prec):
    """No tool, so ft and n play no part: both grids give the same sorted order and the same state
    bit for bit; idx holds the labels (Rothlin) or the last permutation (Green)."""
    case = _padded_chip(60_000)
    gr = S.Simulation(case, precision=prec)
    ro = S.Simulation(case, precision=prec, grid="rothlin")
    gr.step(5)
    ro.step(5)
    labels = ro.p.download("idx")
    gr.step(1)
    ro.step(1)
    gr.ctx.check_error()
    ro.ctx.check_error()
    perm = gr.p.download("idx")
    assert (perm != np.arange(case.N)).sum() > 0
    assert_equal("labels", ro.p.download("idx"), labels[perm])
    fields = ("hash", "pos", "vel", "S", "rho", "T", "eps_pl", "p") + DERIVS
    a, b = ro.state(fields), gr.state(fields)
    for f in fields:
        assert_equal(f, a[f], b[f])
    gr.close()
    ro.close()


def test_rothlin_graph_replays_match_eager():
    case = cases.perturbed(cases.small_cutting_case(nx=48, ny=16, wear=True))
    eager = S.Simulation(case, grid="rothlin")
    gsim = S.Simulation(case, grid="rothlin")
    graph = gsim.capture(steps=2)          # every Rothlin sort is a full sort: no eager step needed
    for r in range(3):
        graph.launch(1)
        eager.step(2)
        a = gsim.state(("hash", "idx", "pos", "vel", "T", "ft", "n"))
        b = eager.state(("hash", "idx", "pos", "vel", "T", "ft", "n"))
        for f in ("hash", "idx"):
            assert_equal(f"replay {r} {f}", a[f], b[f])
        for f in ("pos", "vel", "T", "ft", "n"):
            assert max_rel_err(a[f], b[f]) <= 1e-11, (r, f)
    gsim.ctx.check_error()
    graph.destroy()
