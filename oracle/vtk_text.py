"""The reference's VTK writers restated in Python -- TEST INFRASTRUCTURE ONLY.

Each function returns the text that the cited vtk_writer.cu function writes for the given host
arrays, with Python's "%f" / "%d" formatting (correctly rounded, ties to even, as glibc's printf).
Only tests/ import this module, as the checker of the library's writers.
"""
from __future__ import annotations

import math

import numpy as np

HEADER = "# vtk DataFile Version 2.0\nmfree iwf\nASCII\n\nDATASET UNSTRUCTURED_GRID\n"


def particles(P: dict, skip_blanked: bool = False) -> str:
    """vtk_writer.cu:11-219 (every particle: the reference's blanked buffer is never downloaded)."""
    N = len(P["pos"])
    sel = [i for i in range(N) if not (skip_blanked and P["blanked"][i] == 1.0)]
    n = len(sel)
    out = [HEADER, "POINTS %d float\n" % n]
    out += ["%f %f %f\n" % (P["pos"][i][0], P["pos"][i][1], 0.0) for i in sel]
    out.append("\n")
    out.append("CELLS %d %d\n" % (n, 2 * n))
    out += ["%d %d\n" % (1, k) for k in range(n)]
    out.append("\n")
    out.append("CELL_TYPES %d\n" % n)
    out += ["%d\n" % 1 for _ in range(n)]
    out.append("\n")
    out.append("POINT_DATA %d\n" % n)
    out.append("SCALARS Temperature float 1\nLOOKUP_TABLE default\n")
    out += ["%f\n" % P["T"][i] for i in sel]
    out.append("\n")
    out.append("SCALARS on_seg int 1\nLOOKUP_TABLE default\n")
    out += ["%d\n" % P["on_seg"][i] for i in sel]
    out.append("\n")
    out.append("SCALARS EquivAccumPlasticStrain float 1\nLOOKUP_TABLE default\n")
    out += ["%f\n" % P["eps_pl"][i] for i in sel]
    out.append("\n")
    out.append("VECTORS Velocity float\n")
    out += ["%f %f %f\n" % (P["vel"][i][0], P["vel"][i][1], 0.0) for i in sel]
    out.append("\n")
    out.append("VECTORS fc float\n")
    out += ["%2f %2f %2f\n" % (100000 * P["fc"][i][0], 100000 * P["fc"][i][1], 0.0) for i in sel]
    out.append("\n")
    out.append("VECTORS ft float\n")
    out += ["%f %f %f\n" % (100000 * P["ft"][i][0], 100000 * P["ft"][i][1], 0.0) for i in sel]
    out.append("\n")
    out.append("VECTORS normal float\n")
    out += ["%f %f %f\n" % (P["n"][i][0], P["n"][i][1], 0.0) for i in sel]
    out.append("\n")
    out.append("SCALARS SvM float 1\nLOOKUP_TABLE default\n")
    for i in sel:
        S, p = P["S"][i], P["p"][i]
        cxx, cyy, czz, cxy = S[0] - p, S[2] - p, S[3] - p, S[1]
        svm2 = (cxx * cxx + cyy * cyy + czz * czz) - cxx * cyy - cxx * czz - cyy * czz + 3.0 * cxy * cxy
        out.append("%f\n" % (math.sqrt(svm2) if svm2 > 0 else 0.0))
    out.append("\n")
    return "".join(out)


def rigid_tool(segs: list, fillet: dict | None) -> str:
    """vtk_writer.cu:222-302; segs: list of (left, right) point pairs."""
    if len(segs) == 0:
        return ""
    tris = []
    if len(segs) == 4:
        tris.append((segs[0][0], segs[0][1], segs[1][1]))
        tris.append((segs[2][0], segs[2][1], segs[3][1]))
    else:
        tris.append((segs[0][0], segs[0][1], segs[2][1]))
        tris.append((segs[1][0], segs[1][1], segs[2][1]))
        tris.append((segs[3][0], segs[3][1], segs[4][1]))
    if fillet is not None:
        t1, t2 = min(fillet["t1"], fillet["t2"]), max(fillet["t1"], fillet["t2"])
        lo = t1 - 0.1 * t1
        d = (t2 - t1) / (20 - 1)
        r = fillet["r"]
        c = fillet["p"]
        for i in range(19):
            a1, a2 = lo + (i + 0) * d, lo + (i + 1) * d
            tris.append((c, (c[0] + r * math.sin(a1), c[1] + r * math.cos(a1)),
                         (c[0] + r * math.sin(a2), c[1] + r * math.cos(a2))))
    n = len(tris)
    out = [HEADER, "POINTS %d float\n" % (3 * n)]
    for t in tris:
        for q in t:
            out.append("%f %f %f\n" % (q[0], q[1], 0.0))
    out.append("\n")
    out.append("CELLS %d %d\n" % (n, 4 * n))
    out += ["3 %d %d %d\n" % (3 * i, 3 * i + 1, 3 * i + 2) for i in range(n)]
    out.append("\n")
    out.append("CELL_TYPES %d\n" % n)
    out += ["5\n"] * n
    return "".join(out)


def fem_tool(pos, elem, T, shift=(0.0, 0.0)) -> str:
    """vtk_writer.cu:304-366 (shift 0) and :518-587 (nodes shifted back by tl - tl_ref)."""
    nn, ne = len(pos), len(elem)
    out = [HEADER, "POINTS %d float\n" % nn]
    out += ["%f %f %f\n" % (p[0] - shift[0], p[1] - shift[1], 0.0) for p in pos]
    out.append("\n")
    out.append("CELLS %d %d\n" % (ne, 4 * ne))
    out += ["3 %d %d %d\n" % (e[0], e[1], e[2]) for e in elem]
    out.append("\n")
    out.append("CELL_TYPES %d\n" % ne)
    out += ["5\n"] * ne
    out.append("POINT_DATA %d\n" % nn)
    out.append("SCALARS temperature float 1\nLOOKUP_TABLE default\n")
    out += ["%f\n" % t for t in T]
    out.append("\n")
    return "".join(out)


def _seg_scalar(name, pp, zeros, value):
    out = ["SCALARS %s float 1\nLOOKUP_TABLE default\n" % name]
    for i in range(len(pp)):
        out += ["%f\n" % 0.0] * zeros
        out.append("%f\n" % (value(i) if pp[i][0] != 0.0 else 0.0))
    out.append("\n")
    return out


def fem_segments(left, right, n, pp, sliding, wear) -> str:
    """vtk_writer.cu:368-516"""
    ns = len(left)
    out = [HEADER, "POINTS %d float\n" % (3 * ns)]
    for i in range(ns):
        l, r = left[i], right[i]
        out.append("%f %f %f\n" % (l[0], l[1], 0.0))
        out.append("%f %f %f\n" % (r[0], r[1], 0.0))
        dis = math.sqrt((l[0] - r[0]) * (l[0] - r[0]) + (l[1] - r[1]) * (l[1] - r[1]))
        out.append("%f %f %f\n" % (r[0] + dis * n[i][0], r[1] + dis * n[i][1], 0.0))
    out.append("\n")
    out.append("CELLS %d %d\n" % (ns, 4 * ns))
    out += ["3 %d %d %d \n" % (3 * i, 3 * i + 1, 3 * i + 2) for i in range(ns)]
    out.append("\n")
    out.append("CELL_TYPES %d\n" % ns)
    out += ["5\n"] * ns
    out.append("POINT_DATA %d\n" % (3 * ns))
    out += _seg_scalar("PressureGPa", pp, 2, lambda i: pp[i][1] / pp[i][0] * 100.)
    out += _seg_scalar("Velocity", pp, 2, lambda i: pp[i][2] / pp[i][0] * 10000.)
    out += _seg_scalar("Temperature", pp, 2, lambda i: pp[i][3] / pp[i][0])
    out += _seg_scalar("Frictional_stress", pp, 2, lambda i: sliding[i] / pp[i][0] * 100.)
    out += _seg_scalar("Wear_rate", pp, 2, lambda i: wear[i] * 1e10)
    return "".join(out)


def contact_variables(left, right, pp, sliding, wear) -> str:
    """vtk_writer.cu:589-714"""
    ns = len(left)
    out = [HEADER, "POINTS %d float\n" % ns]
    out += ["%f %f %f\n" % (0.5 * (left[i][0] + right[i][0]), 0.5 * (left[i][1] + right[i][1]), 0.0)
            for i in range(ns)]
    out.append("\n")
    out.append("CELLS %d %d\n" % (ns, 2 * ns))
    out += ["%d %d\n" % (1, i) for i in range(ns)]
    out.append("\n")
    out.append("CELL_TYPES %d\n" % ns)
    out += ["%d\n" % 1] * ns
    out.append("\n")
    out.append("POINT_DATA %d\n" % ns)
    out += _seg_scalar("PressureGPa", pp, 0, lambda i: 100. * pp[i][1] / pp[i][0])
    out += _seg_scalar("Velocity", pp, 0, lambda i: pp[i][2] / pp[i][0] * 10000.)
    out += _seg_scalar("Temperature", pp, 0, lambda i: pp[i][3] / pp[i][0])
    out += _seg_scalar("Frictional_stress", pp, 0, lambda i: sliding[i] / pp[i][0] * 100.)
    out += _seg_scalar("Wear_rate", pp, 2, lambda i: wear[i] * 1e10)
    return "".join(out)


def usui(P: dict, K: float, alpha: float, tool_vel) -> tuple:
    """tool_wear.cu:14-57: (min, max, avg, count) over the particles in contact, summed in index
    order (the fourth contact test repeats fc.y, as there)."""
    mn, mx, tot, cnt = float(np.finfo(np.float32).max), 0.0, 0.0, 0
    for i in range(len(P["pos"])):
        fc, ft = P["fc"][i], P["ft"][i]
        if not (abs(fc[0]) > 0 or abs(fc[1]) > 0 or abs(ft[0]) > 0 or abs(fc[1]) > 0):
            continue
        cnt += 1
        vx, vy = P["vel"][i][0] - tool_vel[0], P["vel"][i][1] - tool_vel[1]
        nx, ny = P["n"][i]
        vrx, vry = vx - vx * nx, vy - vy * ny
        vmag = math.sqrt(vrx * vrx + vry * vry)
        fmag = math.sqrt(fc[0] * fc[0] + fc[1] * fc[1])
        A = 2 * P["h"][i] * math.pi
        w = K * math.exp(alpha / P["T"][i]) * (fmag / A) * vmag
        tot += w
        mn = min(mn, w)
        mx = max(mx, w)
    return (0.0 if cnt == 0 else mn), mx, (0.0 if cnt == 0 else tot / cnt), cnt
