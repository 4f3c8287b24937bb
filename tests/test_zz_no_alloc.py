"""sphfem_report_force_noalloc (no counterpart in the reference; for callers that cannot allocate)
against sphfem_report_force through the C ABI: bit-identical sums on two calls in a row, the second
with fewer particles, plus an exact-sum check and the size refusal."""
import copy
import ctypes as C
import math

import numpy as np
import pytest

from sph_fem_2d_cutting_gpu_amd import _lib
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu


def _view(p, n):
    """The same device arrays seen as the first n particles."""
    s = _lib.Particles()
    C.memmove(C.addressof(s), C.addressof(p.s), C.sizeof(s))
    s.N = n
    return s


def test_report_force_noalloc_equals_report_force():
    ctx = S.Context()
    rng = np.random.default_rng(31)
    N = 1_000_003
    fc = rng.normal(0, 1e-9, (N, 2))
    ft = rng.normal(0, 1e-9, (N, 2))
    p = S.particle_gpu(ctx, N, fc=fc, ft=ft)
    out_d = C.c_void_p()
    ctx.call("sphfem_malloc", C.byref(out_d), 16)
    for n in (N, 4099):
        print(f"report_force_noalloc, N = {n}", flush=True)
        s = _view(p, n)
        ref = _lib.Double2()
        ctx.call("sphfem_report_force", C.byref(s), C.byref(ref))
        ctx.call("sphfem_report_force_noalloc", C.byref(s), out_d)
        got = np.zeros(2)
        ctx.call("sphfem_memcpy_dtoh", got.ctypes.data_as(C.c_void_p), out_d, 16)
        assert got[0] == ref.x and got[1] == ref.y, (got, ref.x, ref.y)
        for k, g in enumerate(got):
            exact = math.fsum(list(fc[:n, k]) + list(ft[:n, k]))
            mag = float(np.abs(fc[:n, k]).sum() + np.abs(ft[:n, k]).sum())
            assert abs(g - exact) <= 1e-13 * mag
    big = _view(p, (1 << 30) + 1)          # refused before any device work
    with pytest.raises(_lib.SphfemError):
        ctx.call("sphfem_report_force_noalloc", C.byref(big), out_d)
    ctx.call("sphfem_free", out_d)
    p.free()
    ctx.close()
