"""Comparison helpers shared by the parity tests."""
import numpy as np

FLOAT_FIELDS = ("pos", "vel", "rho", "S", "T", "eps_pl", "eps_pl_dot", "fc", "ft", "n", "p", "R", "h",
                "pos_t", "vel_t", "rho_t", "S_t", "T_t")


def max_rel_err(gpu, ref, mask=None):
    """max |gpu - ref| over the field, relative to the field's max |ref| (1 if the field is 0)."""
    g = np.asarray(gpu, dtype=np.float64)
    r = np.asarray(ref, dtype=np.float64)
    if mask is not None:
        g, r = g[mask], r[mask]
    if g.size == 0:
        return 0.0
    scale = np.max(np.abs(r))
    if not np.isfinite(scale) or scale == 0:
        scale = 1.0
    d = np.abs(g - r)
    d[np.isnan(g) & np.isnan(r)] = 0.0
    return float(np.max(d) / scale)


def assert_close(name, gpu, ref, rtol, mask=None):
    e = max_rel_err(gpu, ref, mask)
    assert e <= rtol, f"{name}: max relative error {e:.3e} > {rtol:.1e}"
    return e


def assert_equal(name, gpu, ref, mask=None):
    g, r = np.asarray(gpu), np.asarray(ref)
    if mask is not None:
        g, r = g[mask], r[mask]
    bad = np.nonzero(g != r)[0] if g.ndim == 1 else np.nonzero(np.any(g != r, axis=tuple(range(1, g.ndim))))[0]
    assert bad.size == 0, f"{name}: {bad.size} mismatches, first at {bad[:5]}: gpu {g[bad[:3]]} ref {r[bad[:3]]}"
