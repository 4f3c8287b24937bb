"""The reference's tool-mesh set-up restated in Python -- TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, as the checker of the library's host-side mesh pipeline:

* read_abaqus / read_gmsh: mesh_read(file) (mesh_structure_GPU.cu:784-1257) and
  mesh_read(file, true) (:213-594), the two .inp dialects;
* colour_and_layers: boundary layers and element colours (:596-737, :1064-1190), with the
  first free colour counted from 1 (the reference starts at rand() % palette + 1);
* csr: the CSR pattern (:739-780);
* segments: tool_FEM::set_up_segments_from_mesh (tool_FEM.cu:771-1039);
* fem_system: the P1 stiffness / lumped mass assembly and the Dirichlet rows of
  FEM_thermal_gpu.cu:71-152, 126-225, 234-324, in colour order;
* mesh_vtk: write_mesh_file (mesh_structure_GPU.cu:1395-1462);
* remesh_geo: remesh_read_segments (tool_FEM.cu:1708-1810) as a .geo script.

Pure-Python loops, for meshes of a few thousand elements.  Pins (tests/test_mesh_inp.py,
tests/test_remesh.py): round trips of generated meshes in both dialects, the closed-form second
renumbering of Gmsh boundary edges, a hand-worked correction example, colouring / layer / CSR
invariants by brute force, P1 patch tests and lumped-mass areas, the chain and normals of the
generated mesh, and the chain-point properties of remesh_geo.  What Gmsh makes of the remesh_geo
script is parity unpinned (Gmsh is not available).
"""
from __future__ import annotations

import math

import numpy as np

FLT_MAX = 3.4028234663852886e38
DBL_MAX = 1.7976931348623157e308


def _lead_int(tok: str):
    """the integer at the start of a token, as `stringstream >> int` reads it (None if none)"""
    s = tok.lstrip()
    k = 0
    if k < len(s) and s[k] in "+-":
        k += 1
    j = k
    while j < len(s) and s[j].isdigit():
        j += 1
    return int(s[:j]) if j > k else None


def _lead_float(tok: str):
    """the double at the start of a token (std::stod / strtod on decimal input)"""
    s = tok.strip()
    for end in range(len(s), 0, -1):
        try:
            return float(s[:end])
        except ValueError:
            continue
    return None


def _ids(line):
    return [a for a in (_lead_int(t) for t in line.split(",")) if a is not None]


def _split(line):
    """std::getline(ss, tok, ',') tokens: no empty token after a final comma"""
    parts = line.split(",")
    if parts and parts[-1] == "":
        parts = parts[:-1]
    return parts


def _empty_mesh():
    return dict(x=[], y=[], cut_node=[], fixed_node=[], elem=[], cut_flags=[], fixed_flags=[], cut_elems=[],
                fixed_elems=[])


def _correct(m, flags, corr, lst):
    """mesh_structure_GPU.cu:857-876: an element with all three edges flagged keeps an edge only if
    the node opposite to it is shared with another element of the set"""
    for a in corr:
        na = m["elem"][a]
        for k, clear in ((0, 1), (1, 2), (2, 0)):
            shared = any(b != a and na[k] in m["elem"][b] for b in lst)
            if not shared:
                flags[a][clear] = 0


def read_abaqus(text: str) -> dict:
    """mesh_structure_GPU.cu:784-1050"""
    m = _empty_mesh()
    node_start = elem_start = cut_n = fixed_n = cut_e = fixed_e = False
    xs, ys = [], []
    cut_corr, fixed_corr = [], []
    for line in text.split("\n"):
        if "#" in line:
            continue
        if "*Node" in line:
            node_start = True
            continue
        if "Element" in line:
            elem_start, node_start = True, False
            m["x"] = [x / 100 for x in xs]
            m["y"] = [y / 100 for y in ys]
            m["cut_node"] = [0] * len(xs)
            m["fixed_node"] = [0] * len(xs)
            continue
        if "CuttingEdge" in line and "Nset" in line:
            cut_n, elem_start = True, False
            continue
        if "CuttingEdge" in line and "Elset" in line:
            cut_e, cut_n = True, False
            continue
        if "FixedBoundary" in line and "Nset" in line:
            fixed_n, cut_e = True, False
            _correct(m, m["cut_flags"], cut_corr, m["cut_elems"])
            continue
        if "FixedBoundary" in line and "Elset" in line:
            fixed_e, fixed_n = True, False
            continue
        if "End" in line:
            _correct(m, m["fixed_flags"], fixed_corr, m["fixed_elems"])
            break
        if node_start:
            t = line.split(",", 2)
            xs.append(_lead_float(t[1]))
            ys.append(_lead_float(t[2]))
        if elem_start:
            t = _split(line)
            m["elem"].append([_lead_int(t[1]) - 1, _lead_int(t[2]) - 1, _lead_int(t[3]) - 1])
            m["cut_flags"].append([0, 0, 0])
            m["fixed_flags"].append([0, 0, 0])
        if cut_n or fixed_n:
            for a in _ids(line):
                (m["cut_node"] if cut_n else m["fixed_node"])[a - 1] = 1
        if cut_e or fixed_e:
            flags = m["cut_flags"] if cut_e else m["fixed_flags"]
            lst = m["cut_elems"] if cut_e else m["fixed_elems"]
            corr = cut_corr if cut_e else fixed_corr
            on = m["cut_node"] if cut_e else m["fixed_node"]
            for a in _ids(line):
                e = a - 1
                lst.append(e)
                i1, i2, i3 = (on[v] != 0 for v in m["elem"][e])
                if i1 and i2:
                    flags[e][0] = 1
                if i2 and i3:
                    flags[e][1] = 1
                if i3 and i1:
                    flags[e][2] = 1
                if flags[e] == [1, 1, 1]:
                    corr.append(e)
    return m


def number_to_jump(a: int, exclude) -> int:
    """mesh_structure_GPU.cu:12-19"""
    return sum(1 for e in exclude if e < a)


def read_gmsh(text: str) -> dict:
    """mesh_structure_GPU.cu:213-594 (three passes over the file)"""
    lines = text.split("\n")
    used = set()
    on = False
    for line in lines:
        if "type=CPS3" in line:
            on = True
            continue
        if "FixedBoundary" in line:
            break
        if on:
            t = _split(line)
            for k in (1, 2, 3):
                used.add(_lead_int(t[k]) - 1)
    m = _empty_mesh()
    exclude, seq, xs, ys, bedge, cut_edges, fixed_edges = [], [], [], [], [], [], []
    node = edge = cut = fixed = False
    for line in lines:
        if "#" in line:
            continue
        if "*NODE" in line:
            node = True
            continue
        if "T3D2" in line:
            node = cut = fixed = False
            edge = True
            continue
        if "type=CPS3" in line:
            node = edge = cut = fixed = False
            continue
        if "*******" in line:
            node = edge = cut = fixed = False
            n = len(xs)
            m["x"], m["y"] = [0.0] * n, [0.0] * n
            m["cut_node"], m["fixed_node"] = [0] * n, [0] * n
            for i in range(n):
                m["x"][seq[i]] = xs[i]
                m["y"][seq[i]] = ys[i]
            continue
        if "CuttingEdge" in line:
            node = edge = fixed = False
            cut = True
            continue
        if "FixedBoundary" in line:
            node = edge = cut = False
            fixed = True
            continue
        if node:
            t = _split(line)
            a = _lead_int(t[0])
            if a - 1 not in used:
                exclude.append(a)
                continue
            seq.append(a - 1 - number_to_jump(a, exclude))
            xs.append(_lead_float(t[1]))
            ys.append(_lead_float(t[2]))
        if edge:
            t = _split(line)
            a, b = _lead_int(t[1]), _lead_int(t[2])
            bedge.append((a - 1 - number_to_jump(a, exclude), b - 1 - number_to_jump(b, exclude)))
        if cut or fixed:
            for a in _ids(line):
                u, v = bedge[a - 1]
                # renumbered again, as mesh_structure_GPU.cu:424-427 do
                e = (u - number_to_jump(u, exclude), v - number_to_jump(v, exclude))
                flag = m["cut_node"] if cut else m["fixed_node"]
                flag[e[0]] = flag[e[1]] = 1
                (cut_edges if cut else fixed_edges).append(e)
    on = False

    def has(edges, u, v):
        return any((e[0] == u and e[1] == v) or (e[0] == v and e[1] == u) for e in edges)

    for line in lines:
        if "type=CPS3" in line:
            on = True
            continue
        if "FixedBoundary" in line:
            break
        if not on:
            continue
        t = _split(line)
        a = [_lead_int(t[k]) for k in (1, 2, 3)]
        a = [x - 1 - number_to_jump(x, exclude) for x in a]
        e = len(m["elem"])
        m["elem"].append(a)
        fc = [int(has(cut_edges, a[0], a[1])), int(has(cut_edges, a[1], a[2])), int(has(cut_edges, a[2], a[0]))]
        ff = [int(has(fixed_edges, a[0], a[1])), int(has(fixed_edges, a[1], a[2])),
              int(has(fixed_edges, a[2], a[0]))]
        m["cut_flags"].append(fc)
        m["fixed_flags"].append(ff)
        if any(fc):
            m["cut_elems"].append(e)
        if any(ff):
            m["fixed_elems"].append(e)
    return m


def colour_and_layers(m: dict, color_total: int = 18, color_cutting_edge: int = 6):
    """mesh_structure_GPU.cu:596-737: cutting-edge elements first (palette color_cutting_edge,
    neighbours among the cutting-edge elements only), then boundary layer 2 (elements touching a
    cutting-edge node or a node of a cutting-edge element), then the rest, each taking the first
    colour not used by an element sharing a node, searched cyclically from colour 1."""
    elem = m["elem"]
    N = len(elem)
    col = [0] * N
    layers = [[0, 0] for _ in range(N)]

    def colour(e, palette, cands):
        c, loops = 1, 0
        while any(o != e and col[o] == c and set(elem[o]) & set(elem[e]) for o in cands):
            c = 1 if c + 1 > palette else c + 1
            loops += 1
            if loops >= palette:
                raise ValueError("palette too small")
        col[e] = c

    cut = m["cut_elems"]
    for e in cut:
        layers[e][0] = 1
        colour(e, color_cutting_edge, cut)
    cut_nodes = {i for i, f in enumerate(m["cut_node"]) if f}
    l1_nodes = {v for e in cut for v in elem[e]}
    for e in range(N):
        if layers[e][0] == 1:
            continue
        if not (set(elem[e]) & (cut_nodes | l1_nodes)):
            continue
        layers[e][1] = 1
        colour(e, color_total, range(N))
    for e in range(N):
        if col[e] == 0:
            colour(e, color_total, range(N))
    return col, layers


def csr(m: dict):
    """mesh_structure_GPU.cu:739-780: per node, the sorted set of nodes sharing an element"""
    nn = len(m["x"])
    rows = [set() for _ in range(nn)]
    for t in m["elem"]:
        for u in t:
            rows[u].update(t)
    row_ptr = [0]
    col_ind = []
    for r in rows:
        col_ind += sorted(r)
        row_ptr.append(len(col_ind))
    return row_ptr, col_ind


def _segment(l, r, tl):
    """segment(left, right, tl) (geometric_primitives.cpp:115-137)"""
    dx, dy = r[0] - l[0], r[1] - l[1]
    s = 1.0 / math.sqrt(dy * dy + (-dx) * (-dx))
    n = (dy * s, -dx * s)
    if (tl[0] - l[0]) * n[0] + (tl[1] - l[1]) * n[1] < 0:
        return r, l, (-n[0], -n[1])
    return l, r, n


def segments(m: dict, col, tl_ref):
    """tool_FEM.cu:771-1039: one segment per flagged cutting edge, oriented by tl_ref, chained from
    the highest point; tl / br node search; reference distance"""
    d, mark = [], []
    for e in m["cut_elems"]:
        t = m["elem"][e]
        for k, (u, v) in enumerate(((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))):
            if not m["cut_flags"][e][k]:
                continue
            lp, rp = (m["x"][u], m["y"][u]), (m["x"][v], m["y"][v])
            L, R, n = _segment(lp, rp, tl_ref)
            mk = [e, col[e], 0, 0]
            if L == lp:
                mk[2:] = [u, v]
            elif L == rp:
                mk[2:] = [v, u]
            d.append([L, R, n])
            mark.append(mk)
    ns = len(d)
    pts = [p for s in d for p in (s[0], s[1])]
    y_min = min(p[1] for p in pts)
    y_max = max(p[1] for p in pts)
    x_min = min(p[0] for p in pts)
    x_max = max(p[0] for p in pts)
    dist = lambda a, b: math.sqrt((a[0] - b[0]) ** 2 + (a[1] - b[1]) ** 2)
    dmax = max(max(dist(tl_ref, s[0]), dist(tl_ref, s[1])) for s in d)
    xy = (x_min, y_max)
    dmax = max(dmax, max(dist(xy, s[0]) for s in d))
    is_left, start = False, 0
    for i, s in enumerate(d):
        if s[0][1] == y_max:
            is_left, start = True, i
            break
        if s[1][1] == y_max:
            is_left, start = False, i
            break
    chain, cmark = [], []
    a_end, b_end = (0, 1) if is_left else (1, 0)      # follow left -> right, or right -> left
    b = (d[start][a_end][0], y_max)
    rep = 0
    for _ in range(ns):
        done = False
        for j in range(ns):
            if d[j][a_end] == b:
                b = d[j][b_end]
                if is_left:
                    rep = j
                chain.append(list(d[j]))
                cmark.append(list(mark[j]))
                done = True
                break
        if done:
            continue
        for j in range(ns):
            if d[j][b_end] == b and j != rep:
                s = d[j]
                s[b_end] = s[a_end]
                s[a_end] = b
                s[2] = (-s[2][0], -s[2][1])
                b = s[b_end]
                mark[j][2], mark[j][3] = mark[j][3], mark[j][2]
                rep = j
                chain.append(list(s))
                cmark.append(list(mark[j]))
                break
    tl = [FLT_MAX, -FLT_MAX]
    br = [-FLT_MAX, FLT_MAX]
    for x, y in zip(m["x"], m["y"]):
        if x <= tl[0] and y >= tl[1]:
            tl = [x, y]
        if x >= br[0] and y <= br[1]:
            br = [x, y]
    return dict(left=np.array([s[0] for s in chain]), right=np.array([s[1] for s in chain]),
                n=np.array([s[2] for s in chain]), marks=np.array(cmark, dtype=np.int32), tl=np.array(tl),
                br=np.array(br), ref_dist=dmax, lowest_point=y_min, right_point=x_max)


def _element(m, t, tc):
    """FEM_thermal_gpu.cu:71-152: lumped mass and P1 stiffness of one triangle"""
    x0, y0 = m["x"][t[0]], m["y"][t[0]]
    j00, j01 = m["x"][t[1]] - x0, m["y"][t[1]] - y0
    j10, j11 = m["x"][t[2]] - x0, m["y"][t[2]] - y0
    det = abs(j00 * j11 - j10 * j01)
    w = 1.0 / 3.0 * 0.5
    ms = w * tc["cp"] * tc["rho"] * det
    dd = j00 * j11 - j01 * j10
    ood = 1.0 / dd
    i00, i01, i10, i11 = j11 * ood, -j10 * ood, -j01 * ood, j00 * ood
    G = [(i00 * gx + i10 * gy, i01 * gx + i11 * gy) for gx, gy in ((-1.0, -1.0), (1.0, 0.0), (0.0, 1.0))]
    s = (w * 3.0) * (tc["k"] * det)
    return ms, [[(s * G[a][0]) * G[b][0] + (s * G[a][1]) * G[b][1] for b in range(3)] for a in range(3)]


def fem_system(m: dict, col, row_ptr, col_ind, tc, T_init=None):
    """assembly colour by colour (FEM_thermal_gpu.cu:126-225), Dirichlet rows of the fixed edges
    colour by colour (:234-324), temperatures T0 or the hot-tool values"""
    nn = len(m["x"])
    K = [0.0] * len(col_ind)
    M = [0.0] * nn
    f = [0.0] * nn
    T = [tc["T0"]] * nn if T_init is None else [float(v) for v in T_init]

    def find(r, c):
        for i in range(row_ptr[r], row_ptr[r + 1]):
            if col_ind[i] == c:
                return i
        return row_ptr[r + 1]

    ncol = max(col)
    for c in range(1, ncol + 1):
        for e, t in enumerate(m["elem"]):
            if col[e] != c:
                continue
            ms, Ke = _element(m, t, tc)
            for a in range(3):
                M[t[a]] += ms
                for b in range(3):
                    K[find(t[a], t[b])] += Ke[a][b]
    T0 = tc["T0"]
    for c in range(1, ncol + 1):
        for e in m["fixed_elems"]:
            if col[e] != c:
                continue
            fl = m["fixed_flags"][e]
            nf = [0, 0, 0]
            if fl[0]:
                nf[0] = nf[1] = 1
            if fl[1]:
                nf[1] = nf[2] = 1
            if fl[2]:
                nf[2] = nf[0] = 1
            for k in range(3):
                if not nf[k]:
                    continue
                node = m["elem"][e][k]
                for i in range(row_ptr[node], row_ptr[node + 1]):
                    loc = col_ind[i]
                    f[loc] += -K[i] * T0
                    K[i] = 0.0
                    K[find(loc, node)] = 0.0
                    if loc == node:
                        K[i] = 1.0
                M[node] = 1.0
                f[node] = T0
    return np.array(K), np.array(M), np.array(f), np.array(T)


def mesh_vtk(m: dict, col, layers) -> str:
    """write_mesh_file (mesh_structure_GPU.cu:1395-1462)"""
    nn, nt = len(m["x"]), len(m["elem"])
    out = ["# vtk DataFile Version 2.0\nmfree iwf\nASCII\n\nDATASET UNSTRUCTURED_GRID\n", "POINTS %d float\n" % nn]
    out += ["%f %f %f\n" % (m["x"][i], m["y"][i], 0.0) for i in range(nn)]
    out.append("\n")
    out.append("CELLS %d %d\n" % (nt, 4 * nt))
    out += ["3 %d %d %d\n" % tuple(t) for t in m["elem"]]
    out.append("\n")
    out.append("CELL_TYPES %d\n" % nt)
    out += ["5\n"] * nt
    out.append("CELL_DATA %d\n" % nt)
    out.append("SCALARS color int 1\nLOOKUP_TABLE default\n")
    out += ["%d\n" % c for c in col]
    out.append("\n")
    out.append("SCALARS layerf int 1\nLOOKUP_TABLE default\n")
    out += ["%d\n" % l[0] for l in layers]
    out.append("\n")
    out.append("SCALARS layers int 1\nLOOKUP_TABLE default\n")
    out += ["%d\n" % l[1] for l in layers]
    out.append("\n")
    return "".join(out)


def _fdist(a, b):
    """sqrtf of the double expression, as the distance helpers of tool_FEM.cu do"""
    return float(np.sqrt(np.float32((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]))))


def remesh_geo(left, right, tl) -> str:
    """remesh_read_segments (tool_FEM.cu:1708-1810): the gmsh::model::geo calls as a .geo script"""
    ns = len(left)
    down = left[0][1] > right[0][1]
    pts, cut = [], []
    length, prev = 0.004, 0.0
    i = 0
    while i < ns:
        p = left[i] if down else right[i]
        if i + 1 < ns:
            judge = _fdist(left[i], right[i + 1]) if down else _fdist(left[i + 1], right[i])
        else:
            judge = math.inf
        buf = _fdist(left[i], right[i])
        length = 6.0 * buf
        pts.append((p[0], p[1], length))
        cut.append(1 + len(pts))
        if (judge < 1.25 * buf and i != ns - 1 and prev > 0.7 * buf) or judge < 0.00013:
            i += 1
        prev = buf
        i += 1
    last = right[ns - 1] if down else left[ns - 1]
    pts.append((last[0], last[1], length))
    g = lambda v: "%.17g" % v
    np_ = len(pts) + 1
    out = ["// tool remesh geometry (remesh_read_segments): mesh with\n",
           "//   gmsh -2 -format inp -o tool_gmsh_remeshed.inp <this file>\n",
           "Point(1) = {%s, %s, 0, %s};\n" % (g(tl[0]), g(tl[1]), g(0.075))]
    out += ["Point(%d) = {%s, %s, 0, %s};\n" % (q + 2, g(x), g(y), g(s)) for q, (x, y, s) in enumerate(pts)]
    out += ["Line(%d) = {%d, %d};\n" % (q, q, q + 1) for q in range(1, np_)]
    out.append("Line(%d) = {%d, 1};\n" % (np_, np_))
    out.append("Curve Loop(1) = {%s};\n" % ", ".join(str(q) for q in range(1, np_ + 1)))
    out.append("Plane Surface(1) = {1};\n")
    out.append('Physical Curve("CuttingEdge", 67) = {%s};\n' % ", ".join(str(c) for c in cut))
    out.append('Physical Curve("FixedBoundary", 2) = {1, %d};\n' % np_)
    out.append("Mesh.MshFileVersion = 4;\nMesh.SaveAll = 1;\n")
    return "".join(out)
