"""Cell list at scale: FP64 hash, stable radix sort and cell boundaries vs numpy.

Keys up to 2^24 cells (three 8-bit passes) and a 10:1 density variation; the
permutation must equal numpy's stable argsort, i.e. within a cell the previous order
is kept (the CUB radix sort behind thrust::sort_by_key, grid_gpu_green.cu:193, is stable).
"""
import numpy as np
import pytest

from helpers import assert_equal
from sph_fem_2d_cutting_gpu_amd import cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu


def _check(ctx, pos, blanked, bbmin, bbmax, h):
    N = len(pos)
    p = S.particle_gpu(ctx, N, pos=pos, blanked=blanked)
    g = S.grid_gpu_green(ctx, N, bbmin, bbmax, h)
    # the reference hash of blanked particles is whatever the array held before: seed it
    old = (np.arange(N, dtype=np.int64) * 7919 % g.num_cell).astype(np.int32)
    p.upload("hash", old)
    g.assign_hashes(p)
    geo = g.geom
    ix = ((pos[:, 0] - geo.bbmin_x) / geo.dx).astype(np.int32)
    iy = ((pos[:, 1] - geo.bbmin_y) / geo.dx).astype(np.int32)
    want = np.where(blanked == 1.0, old, ix * geo.ny + iy).astype(np.int32)
    got = p.download("hash")
    assert_equal("hash", got, want)
    g.sort(p)
    ctx.check_error()
    perm = np.argsort(want, kind="stable").astype(np.int32)
    assert_equal("idx", p.download("idx"), perm)
    skeys = want[perm]
    assert_equal("sorted hash", p.download("hash"), skeys)
    assert_equal("gathered pos", p.download("pos"), pos[perm])
    cs, ce = g.download_cells()
    first = np.ones(N, dtype=bool)
    first[1:] = skeys[1:] != skeys[:-1]
    starts = np.nonzero(first)[0]
    ends = np.append(starts[1:], N)
    ref_cs = np.full(g.num_cell, -1, dtype=np.int32)
    ref_cs[skeys[starts]] = starts
    assert_equal("cell_start", cs, ref_cs)
    assert_equal("cell_end", ce[skeys[starts]], ends.astype(np.int32))
    p.free()


def test_sort_4m_uniform_random():
    ctx = S.Context()
    N = 4_000_000
    rng = np.random.default_rng(7)
    h = 1.5e-4
    L = 4096 * h   # ~2^24 cells
    pos = rng.uniform(0, L, size=(N, 2))
    # many particles share cells in one corner (stability under heavy collisions)
    pos[: N // 8] = rng.uniform(0, 8 * h, size=(N // 8, 2))
    blanked = (rng.uniform(size=N) < 0.01).astype(np.float64)
    _check(ctx, pos, blanked, pos.min(axis=0), pos.max(axis=0), h)


def test_sort_chip_density():
    ctx = S.Context()
    case = cases.chip_case(n_target=1_000_000)
    _check(ctx, case.pos, np.zeros(case.N), case.grid_bbmin, case.grid_bbmax, case.h)
