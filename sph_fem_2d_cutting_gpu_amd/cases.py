"""Deterministic cutting-case generator (scenario setup, not the hot path).

Builds everything the per-step hot path consumes, in the reference's units
(cm, g, us, K; stress in units of 100 GPa):

* the jittered particle lattice and its initial state
  (benchmarks_NZ.cpp:138-185, particle_gpu.cu:11-134);
* the constant blocks of constants_structs.h:18-123 with the values of
  benchmarks_NZ.cpp:80-136 and 194-205, overridden by the workload's material;
* a rigid FEM tool: polygon as in tool_FEM.cu:623-660 / :705-735, a graded
  triangle mesh (replaces the Gmsh .inp file read by mesh_structure_GPU.cu:213),
  cutting-edge segments in chain order with inward normals
  (tool_FEM.cu:771-1039), element colouring, boundary layers, CSR pattern
  (mesh_structure_GPU.cu:596-780), lumped mass / P1 stiffness assembly and the
  Dirichlet rows of the fixed boundary (FEM_thermal_gpu.cu:126-324);
* the locked cell-list geometry (benchmarks_NZ.cpp:272-296, grid.cu:130-149)
  and the CFL time step (benchmarks_NZ.cpp:360).

Everything is a pure function of the workload parameters and the seed, so the
HIP path and the oracle are always fed identical inputs.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

M_1_PI = 0.318309886183790671538
FLT_MAX = 3.4028234663852886e38
DEFAULT_SEED = 20261202

# ----------------------------------------------------------------------------
# SplitMix64 (Steele, Lea, Flood 2014), vectorised.
# ----------------------------------------------------------------------------
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n successive SplitMix64 outputs of state seed + stream * 2**32."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed + (stream << 32)) & 0xFFFFFFFFFFFFFFFF)
        z = base + _GOLD * (np.arange(1, n + 1, dtype=np.uint64))
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits of SplitMix64."""
    return (splitmix64(seed, n, stream) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal01(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """Standard normals by Box-Muller over two SplitMix64 streams."""
    u1 = uniform01(seed, n, stream)
    u2 = uniform01(seed, n, stream + 1)
    u1 = np.maximum(u1, 2.0 ** -53)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


# ----------------------------------------------------------------------------
# materials (bomb units).  Ti6Al4V elastic/thermal values: benchmarks_NZ.cpp:81-115;
# Johnson-Cook values are the workload's.
# ----------------------------------------------------------------------------
MATERIALS = {
    "Ti6Al4V": dict(E=1.1, nu=0.35, rho0=4.43, A=0.01098, B=0.01092, n=0.93, C=0.014,
                    m=1.1, Tmelt=1878.0, cp=526e-8, k=6.8e-13, tq=0.9, eta=0.9),
    # AISI 1045: E, nu, rho0, Tmelt, cp and k are not given by the workload; stated here.
    "AISI1045": dict(E=2.0, nu=0.3, rho0=7.85, A=0.00553, B=0.00601, n=0.234, C=0.0134,
                     m=1.0, Tmelt=1733.0, cp=486e-8, k=49.8e-13, tq=0.9, eta=0.9),
}

# Usui constants of the populated reference setup (benchmarks_NZ.cpp:310-312), used for
# both the rake (C, D) and the flank (A, B) branch of tool_FEM.cu:400-408.
USUI_K = 86.0
USUI_ALPHA = 18769.9

T_REF = 300.0
HDX = 1.5


def m_per_min_to_cm_per_us(v: float) -> float:
    return v / 60.0 * 100.0 * 1e-6


@dataclass
class Constants:
    """The seven constant blocks of constants_structs.h:18-123 (plain dicts of floats)."""
    phys: dict
    trml_wp: dict
    trml_tool: dict
    corr: dict
    joco: dict
    tool: dict
    wear: dict


def make_constants(material: str, dx: float, mu: float) -> Constants:
    mat = MATERIALS[material]
    phys = dict(E=mat["E"], nu=mat["nu"], rho0=mat["rho0"])
    phys["G"] = phys["E"] / (2.0 * (1.0 + phys["nu"]))
    phys["K"] = 2.0 * phys["G"] * (1 + phys["nu"]) / (3 * (1 - 2 * phys["nu"]))
    phys["mass"] = dx * dx * phys["rho0"]
    rho_tool = 15.25
    phys["mass_tool"] = dx * dx * rho_tool

    joco = dict(A=mat["A"], B=mat["B"], C=mat["C"], n=mat["n"], m=mat["m"],
                Tmelt=mat["Tmelt"], Tref=T_REF, eps_dot_ref=1e-6,
                tanh_a=5.0, tanh_b=0.6, tanh_c=600.0, tanh_d=0.7)

    trml_wp = dict(cp=mat["cp"], tq=mat["tq"], k=mat["k"], alpha=0.0, T_init=T_REF, eta=mat["eta"])
    trml_wp["alpha"] = trml_wp["k"] / (phys["rho0"] * trml_wp["cp"])
    trml_tool = dict(cp=15e-8, tq=0.0, k=59e-13, alpha=0.0, T_init=T_REF, eta=0.0)
    trml_tool["alpha"] = trml_tool["k"] / (rho_tool * trml_tool["cp"])

    corr = dict(wdeltap=1.0, stresseps=0.3, xspheps=0.5, alpha=1.0, beta=1.0, eta=0.1)
    h1 = 1.0 / (HDX * dx)
    q = dx * h1
    fac = 10 * M_1_PI / 7.0 * h1 * h1
    corr["wdeltap"] = fac * (1 - 1.5 * q * q * (1 - 0.5 * q))

    tool = dict(cp=292e-8, tq=0.9, k=88e-13, eta=1.0, rho=15.25, T0=T_REF, h=2e-10,
                beta=0.5, mu=mu, contact_alpha=0.02, slave_mass=phys["mass"], hc=0.0)
    wear = dict(A=USUI_K, B=USUI_ALPHA, C=USUI_K, D=USUI_ALPHA)
    return Constants(phys, trml_wp, trml_tool, corr, joco, tool, wear)


# ----------------------------------------------------------------------------
# tool geometry (tool_FEM.cu:623-660, :705-758, geometric_primitives.cpp:11-76)
# ----------------------------------------------------------------------------
def _line(p1, p2):
    """(a, b, vertical) with y = a x + b, as geometric_primitives.cpp:61-76."""
    det = p1[0] - p2[0]
    if abs(det) < 1e-12:
        return (FLT_MAX, p1[0], True)
    a = (p1[1] - p2[1]) / det
    b = (p2[1] * p1[0] - p2[0] * p1[1]) / det
    return (a, b, False)


def _intersect(l1, l2):
    if l1[2] and l2[2]:
        raise ValueError("parallel vertical lines")
    if l1[2]:
        return np.array([l1[1], l2[0] * l1[1] + l2[1]])
    if l2[2]:
        return np.array([l2[1], l1[0] * l2[1] + l1[1]])
    x = (l2[1] - l1[1]) / (l1[0] - l2[0])
    return np.array([x, l1[0] * x + l1[1]])


def _closest_point(l, q):
    if l[2]:
        return np.array([l[1], q[1]])
    a, b = l[0], l[1]
    bb, cc = -1.0, b
    px = (bb * (bb * q[0] - a * q[1]) - a * cc) / (a * a + bb * bb)
    py = (a * (-bb * q[0] + a * q[1]) - bb * cc) / (a * a + bb * bb)
    return np.array([px, py])


def tool_outline(length: float, height: float, rake_deg: float, clear_deg: float, r: float):
    """Corner points tl, tr, trc, blc, bl, the fillet centre and its two angles.

    The rake face leaves tr along (sin a, -cos a), the flank leaves bl along
    (cos c, -sin c) (tool_FEM.cu:630-644); the fillet of radius r is tangent to both.
    """
    tl = np.array([0.0, 0.0])
    tr = np.array([length, 0.0])
    bl = np.array([0.0, -height])
    a = math.radians(rake_deg)
    c = math.radians(clear_deg)
    d_rake = np.array([math.sin(a), -math.cos(a)])
    d_flank = np.array([math.cos(c), -math.sin(c)])
    l1 = _line(tr, tr + d_rake)
    l2 = _line(bl, bl + d_flank)
    br = _intersect(l1, l2)
    if r <= 0.0:
        return dict(tl=tl, tr=tr, bl=bl, br=br, trc=br, blc=br, centre=None)
    # fillet centre: on the bisector of the corner at distance r from both faces
    nt = (tr - br) / np.linalg.norm(tr - br)
    nl = (bl - br) / np.linalg.norm(bl - br)
    nm = 0.5 * (nt + nl)
    nm /= np.linalg.norm(nm)
    half = 0.5 * math.acos(float(np.clip(np.dot(nt, nl), -1.0, 1.0)))
    centre = br + nm * (r / math.sin(half))
    trc = _closest_point(l1, centre)
    blc = _closest_point(l2, centre)
    return dict(tl=tl, tr=tr, bl=bl, br=br, trc=trc, blc=blc, centre=centre)


def _sample_polyline(points, spacing_fn):
    """Resample a polyline (list of (n,2)) with local spacing spacing_fn(p)."""
    out = [points[0]]
    for k in range(len(points) - 1):
        a, b = points[k], points[k + 1]
        L = float(np.linalg.norm(b - a))
        if L == 0:
            continue
        t = 0.0
        while True:
            p = a + (b - a) * t
            s = spacing_fn(p)
            t_next = t + s / L
            if t_next >= 1.0 - 0.35 * s / L:
                break
            t = t_next
            out.append(a + (b - a) * t)
        out.append(b)
    return np.array(out)


def _arc_points(centre, r, p_start, p_end, spacing):
    t1 = math.atan2(p_start[1] - centre[1], p_start[0] - centre[0])
    t2 = math.atan2(p_end[1] - centre[1], p_end[0] - centre[0])
    # clockwise from t1 to t2
    while t2 > t1:
        t2 -= 2 * math.pi
    n = max(2, int(math.ceil((t1 - t2) * r / spacing)))
    ts = t1 + (t2 - t1) * np.arange(n + 1) / n
    pts = np.stack([centre[0] + r * np.cos(ts), centre[1] + r * np.sin(ts)], axis=1)
    pts[0] = p_start
    pts[-1] = p_end
    return pts


@dataclass
class ToolMesh:
    nodes: np.ndarray           # (n,2) float64
    elems: np.ndarray           # (e,4) int32: 3 node ids + colour (mesh_structure_GPU.h:37)
    layers: np.ndarray          # (e,2) int32: boundary layer 1 / 2 flags
    cut_edge_node: np.ndarray   # (n,) int32
    fixed_edge_node: np.ndarray  # (n,) int32
    row_ptr: np.ndarray
    col_ind: np.ndarray
    K: np.ndarray               # CSR values after Dirichlet rows
    M: np.ndarray               # lumped mass after Dirichlet rows
    f: np.ndarray               # rhs after Dirichlet rows
    T: np.ndarray               # nodal temperature
    seg_left: np.ndarray        # (s,2)
    seg_right: np.ndarray
    seg_n: np.ndarray
    seg_marks: np.ndarray       # (s,4) int32: element, colour, left node, right node
    n_colors: int
    n_cut_colors: int
    tl: np.ndarray
    br: np.ndarray
    ref_dist: float
    lowest_point: float
    right_point: float
    fixed_elems: np.ndarray | None = None   # fixed-boundary elements and their edge flags (e,3), as read
    fixed_flags: np.ndarray | None = None   # from a mesh file; None: derived from the boundary edges
    outline: tuple | None = None    # (length, height, rake, clearance, edge radius) of a generated tool


def _normalize_glm(v):
    """glm::normalize: v * inversesqrt(dot(v, v))."""
    return v * (1.0 / np.sqrt(np.sum(v * v, axis=-1, keepdims=True)))


def mesh_tool(length, height, rake_deg, clear_deg, r, target_tris, tc) -> ToolMesh:
    """Graded triangle mesh of the tool with about target_tris triangles (FEM set-up of
    FEM_thermal_gpu.cu:584-603).  The base spacing is adjusted until the count is within 10 %."""
    area = length * height
    s_max = math.sqrt(2.0 * area / (target_tris * 0.62))
    # the quadtree quantises cell sizes to root / 2^k: search the root scale as well
    best = None
    for u in (0.0, 0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875):
        ntri = 2 * len(_tool_points(length, height, rake_deg, clear_deg, r, s_max, 2.0 ** u)[0])
        if best is None or abs(ntri - target_tris) < abs(best[1] - target_tris):
            best = (u, ntri)
    tm = _mesh_tool(length, height, rake_deg, clear_deg, r, s_max, tc, 2.0 ** best[0])
    tm.outline = (float(length), float(height), float(rake_deg), float(clear_deg), float(r))
    return tm


def _tool_points(length, height, rake_deg, clear_deg, r, s_max, root_scale=1.0):
    g = tool_outline(length, height, rake_deg, clear_deg, r)
    tl, tr, bl, trc, blc, centre = g["tl"], g["tr"], g["bl"], g["trc"], g["blc"], g["centre"]
    edge = g["br"] if centre is None else centre
    s_min = min(s_max, (r if r > 0 else s_max) / 5.0)
    grad = 0.22

    def spacing(p):
        d = float(np.linalg.norm(np.asarray(p) - edge)) - r
        return float(min(s_max, s_min + grad * max(d, 0.0)))

    # boundary, clockwise: tl -> tr -> trc -> arc -> blc -> bl -> tl
    cut = [_sample_polyline([tr, trc], spacing)]
    if centre is not None:
        cut.append(_arc_points(centre, r, trc, blc, s_min)[1:])
    cut.append(_sample_polyline([blc, bl], spacing)[1:])
    cut = np.concatenate(cut)
    fixed = _sample_polyline([bl, tl, tr], spacing)
    boundary = np.concatenate([cut, fixed[1:-1]])
    nb_cut = len(cut)

    # interior points from a graded quadtree over the bounding square
    lo = np.array([min(tl[0], bl[0]), min(bl[1], g["br"][1])])
    size = max(length + abs(g["br"][0] - tr[0]), height) * 1.0001 * root_scale
    interior = []
    stack = [(lo[0], lo[1], size)]
    while stack:
        x0, y0, s = stack.pop()
        c = np.array([x0 + 0.5 * s, y0 + 0.5 * s])
        if s > spacing(c) * 1.0:
            h = 0.5 * s
            stack.extend([(x0, y0, h), (x0 + h, y0, h), (x0, y0 + h, h), (x0 + h, y0 + h, h)])
        else:
            interior.append((c[0], c[1], s))
    interior = np.array(interior)
    pts = interior[:, :2]
    loc = interior[:, 2]
    # keep points inside the convex tool with a margin of half the local spacing
    poly = np.concatenate([cut, fixed[1:]])
    keep = np.ones(len(pts), dtype=bool)
    for k in range(len(poly) - 1):
        a, b = poly[k], poly[k + 1]
        d = b - a
        nrm = np.array([d[1], -d[0]]) / np.linalg.norm(d)  # inward for clockwise order
        dist = (pts - a) @ nrm
        keep &= dist > 0.55 * loc
    pts = pts[keep]
    # drop interior points too close to boundary points
    from scipy.spatial import cKDTree
    tree = cKDTree(boundary)
    dmin, _ = tree.query(pts)
    pts = pts[dmin > 0.55 * loc[keep]]
    order = np.lexsort((pts[:, 1], pts[:, 0]))
    pts = pts[order]
    nodes = np.concatenate([boundary, pts])
    return nodes, g, cut, fixed, boundary, nb_cut, s_min


def _mesh_tool(length, height, rake_deg, clear_deg, r, s_max, tc, root_scale=1.0) -> ToolMesh:
    from scipy.spatial import Delaunay

    nodes, g, cut, fixed, boundary, nb_cut, s_min = _tool_points(length, height, rake_deg, clear_deg, r, s_max,
                                                                 root_scale)
    tl = g["tl"]
    # Laplacian smoothing of the interior nodes (boundary nodes stay put); the tool is convex, so a
    # neighbour average stays inside it.  Removes the slivers of the quadtree/boundary transition.
    nbd0 = len(boundary)
    for _ in range(4):
        tri = Delaunay(nodes).simplices
        acc = np.zeros_like(nodes)
        cnt = np.zeros(len(nodes))
        for a, b in ((0, 1), (1, 2), (2, 0)):
            np.add.at(acc, tri[:, a], nodes[tri[:, b]])
            np.add.at(acc, tri[:, b], nodes[tri[:, a]])
            np.add.at(cnt, tri[:, a], 1.0)
            np.add.at(cnt, tri[:, b], 1.0)
        nodes = nodes.copy()
        nodes[nbd0:] = acc[nbd0:] / cnt[nbd0:, None]
    tri = Delaunay(nodes).simplices.astype(np.int64)
    # orient counter-clockwise
    p0, p1, p2 = nodes[tri[:, 0]], nodes[tri[:, 1]], nodes[tri[:, 2]]
    cr = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1])
    flip = cr < 0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    cr = np.abs(cr)
    tri = tri[cr > 1e-6 * s_min * s_min]
    tri = tri[np.lexsort((tri[:, 0], (nodes[tri].mean(axis=1)[:, 1]), (nodes[tri].mean(axis=1)[:, 0])))]
    n_nodes, n_elem = len(nodes), len(tri)

    # boundary edge sets
    cut_edges = [(k, k + 1) for k in range(nb_cut - 1)]
    nbd = len(boundary)
    fixed_ids = [nb_cut - 1] + list(range(nb_cut, nbd)) + [0]
    fixed_edges = [(fixed_ids[k], fixed_ids[k + 1]) for k in range(len(fixed_ids) - 1)]
    cut_edge_node = np.zeros(n_nodes, dtype=np.int32)
    cut_edge_node[:nb_cut] = 1
    fixed_edge_node = np.zeros(n_nodes, dtype=np.int32)
    fixed_edge_node[fixed_ids] = 1

    edge_to_elem = {}
    for e, (a, b, c) in enumerate(tri):
        for k, (u, v) in enumerate(((a, b), (b, c), (c, a))):
            edge_to_elem[(min(u, v), max(u, v))] = (e, k)

    def flags_for(edges):
        fl = np.zeros((n_elem, 3), dtype=np.int32)
        owner = []
        for (u, v) in edges:
            e, k = edge_to_elem[(min(u, v), max(u, v))]
            fl[e, k] = 1
            owner.append(e)
        return fl, owner

    cut_flags, cut_owner = flags_for(cut_edges)
    fixed_flags, _ = flags_for(fixed_edges)
    cut_elems = np.nonzero(cut_flags.any(axis=1))[0]
    fixed_elems = np.nonzero(fixed_flags.any(axis=1))[0]

    # node -> elements
    node_elems = [[] for _ in range(n_nodes)]
    for e, t in enumerate(tri):
        for v in t:
            node_elems[v].append(e)

    def neighbours(e):
        s = set()
        for v in tri[e]:
            s.update(node_elems[v])
        s.discard(e)
        return s

    # boundary layers (mesh_structure_GPU.cu:600-668)
    layers = np.zeros((n_elem, 2), dtype=np.int32)
    layers[cut_elems, 0] = 1
    cut_nodes = set(np.nonzero(cut_edge_node)[0].tolist())
    layer1_nodes = set(tri[cut_elems].ravel().tolist())
    for e in range(n_elem):
        if layers[e, 0]:
            continue
        if cut_nodes.intersection(tri[e].tolist()) or layer1_nodes.intersection(tri[e].tolist()):
            layers[e, 1] = 1

    # deterministic greedy colouring: cutting elements first (own palette), then the rest
    colour = np.zeros(n_elem, dtype=np.int32)
    cut_set = set(cut_elems.tolist())
    for e in cut_elems:
        used = {colour[o] for o in neighbours(e) if o in cut_set}
        c = 1
        while c in used:
            c += 1
        colour[e] = c
    n_cut_colors = int(colour.max())
    for e in range(n_elem):
        if colour[e]:
            continue
        used = {colour[o] for o in neighbours(e)}
        c = 1
        while c in used:
            c += 1
        colour[e] = c
    n_colors = int(colour.max())

    # CSR pattern (sorted columns per row, mesh_structure_GPU.cu:741-780)
    rows = [set() for _ in range(n_nodes)]
    for t in tri:
        for u in t:
            rows[u].update(t.tolist())
    row_ptr = np.zeros(n_nodes + 1, dtype=np.int32)
    row_ptr[1:] = np.cumsum([len(s) for s in rows])
    col_ind = np.concatenate([np.array(sorted(s), dtype=np.int32) for s in rows])

    elems = np.zeros((n_elem, 4), dtype=np.int32)
    elems[:, :3] = tri
    elems[:, 3] = colour

    K, M = assemble_p1(nodes, elems, row_ptr, col_ind, tc, colour_order=True, n_colors=n_colors)
    f = np.zeros(n_nodes)
    T = np.full(n_nodes, tc["T0"])
    dirichlet_rows(elems, fixed_flags, fixed_elems, row_ptr, col_ind, K, M, f, T, tc["T0"], n_colors)

    # cutting segments in chain order tr -> bl, inward normals (tool_FEM.cu:771-1039)
    left = nodes[np.arange(nb_cut - 1)]
    right = nodes[np.arange(1, nb_cut)]
    d = right - left
    nvec = _normalize_glm(np.stack([d[:, 1], -d[:, 0]], axis=1))
    marks = np.zeros((nb_cut - 1, 4), dtype=np.int32)
    for s_id, (u, v) in enumerate(cut_edges):
        e, _ = edge_to_elem[(min(u, v), max(u, v))]
        marks[s_id] = (e, colour[e], u, v)

    # tl / br node search (tool_FEM.cu:1021-1037) and ref_dist (:847-873)
    tl_x, tl_y = FLT_MAX, -FLT_MAX
    br_x, br_y = -FLT_MAX, FLT_MAX
    for x, y in nodes:
        if x <= tl_x and y >= tl_y:
            tl_x, tl_y = x, y
        if x >= br_x and y <= br_y:
            br_x, br_y = x, y
    ends = np.concatenate([left, right])
    ref_dist = float(np.max(np.sqrt(np.sum((ends - tl) ** 2, axis=1))))
    xy_ref = np.array([ends[:, 0].min(), ends[:, 1].max()])
    ref_dist = max(ref_dist, float(np.max(np.sqrt(np.sum((left - xy_ref) ** 2, axis=1)))))
    return ToolMesh(nodes=nodes, elems=elems, layers=layers, cut_edge_node=cut_edge_node,
                    fixed_edge_node=fixed_edge_node, row_ptr=row_ptr, col_ind=col_ind, K=K, M=M,
                    f=f, T=T, seg_left=left, seg_right=right, seg_n=nvec, seg_marks=marks,
                    n_colors=n_colors, n_cut_colors=n_cut_colors,
                    tl=np.array([tl_x, tl_y]), br=np.array([br_x, br_y]), ref_dist=ref_dist,
                    lowest_point=float(ends[:, 1].min()), right_point=float(ends[:, 0].max()))


def tool_mesh_edges(tm: ToolMesh):
    """(cutting edges, fixed edges) of a tool mesh as node pairs: the cutting edges are the segment
    ends, the fixed edges the boundary edges (edges of one triangle) with both nodes on the fixed
    boundary that are not cutting edges."""
    cut = [(int(a), int(b)) for a, b in tm.seg_marks[:, 2:4]]
    cut_set = {(min(a, b), max(a, b)) for a, b in cut}
    count = {}
    for t in tm.elems[:, :3]:
        for u, v in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            k = (int(min(u, v)), int(max(u, v)))
            count[k] = count.get(k, 0) + 1
    fx = tm.fixed_edge_node
    fixed = [k for k, c in count.items() if c == 1 and fx[k[0]] and fx[k[1]] and k not in cut_set]
    return cut, sorted(fixed)


def fixed_boundary(tm: ToolMesh):
    """(fixed-boundary elements, their edge flags x: 0-1, y: 1-2, z: 2-0) as mesh_read sets them
    (mesh_structure_GPU.cu:546-556): the mesh file's when the tool was read from one, else from the
    boundary edges."""
    if tm.fixed_elems is not None:
        return np.asarray(tm.fixed_elems, np.int32), np.asarray(tm.fixed_flags, np.int32)
    _, fixed = tool_mesh_edges(tm)
    want = set(fixed)
    tri = tm.elems[:, :3]
    fl = np.zeros((len(tri), 3), np.int32)
    for e, t in enumerate(tri):
        for k, (u, v) in enumerate(((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))):
            fl[e, k] = (int(min(u, v)), int(max(u, v))) in want
    sel = np.nonzero(fl.any(axis=1))[0].astype(np.int32)
    return sel, fl[sel]


def write_inp(tm: ToolMesh, path: str, gmsh: bool = False, unused_nodes=()):
    """Write a tool mesh in one of the two .inp dialects the reference reads
    (mesh_structure_GPU.cu:784-1257 and :213-594), so the same mesh can be fed to both codes.

    gmsh=False: Abaqus style, *Node (coordinates times 100, the reader divides by 100), *Element,
    node and element sets CuttingEdge / FixedBoundary, *End Part.
    gmsh=True: Gmsh export, *NODE, T3D2 boundary edges, CPS3 triangles, element sets of edge ids
    (FixedBoundary before CuttingEdge, as Gmsh orders the physical groups 2 and 67).  unused_nodes
    lists node positions (0-based, into the written numbering) where a node that no triangle uses is
    inserted, as Gmsh writes geometry points outside the surface mesh."""
    cut, fixed = tool_mesh_edges(tm)
    nn = len(tm.nodes)
    tri = tm.elems[:, :3]
    lines = []
    if not gmsh:
        lines += ["*Heading", "** tool mesh", "*Part, name=Tool", "*Node"]
        lines += ["%d, %r, %r" % (i + 1, float(x) * 100, float(y) * 100) for i, (x, y) in enumerate(tm.nodes)]
        lines.append("*Element, type=CPS3")
        lines += ["%d, %d, %d, %d" % (e + 1, t[0] + 1, t[1] + 1, t[2] + 1) for e, t in enumerate(tri)]

        def owners(edges):
            want = {(min(a, b), max(a, b)) for a, b in edges}
            out = []
            for e, t in enumerate(tri):
                for u, v in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
                    if (min(u, v), max(u, v)) in want:
                        out.append(e)
                        break
            return out

        def ids(values):
            values = list(values)
            return [", ".join(str(v) for v in values[k:k + 16]) for k in range(0, len(values), 16)]

        cut_nodes = sorted({a for e in cut for a in e})
        fixed_nodes = sorted({a for e in fixed for a in e})
        lines.append("*Nset, nset=CuttingEdge")
        lines += ids(a + 1 for a in cut_nodes)
        lines.append("*Elset, elset=CuttingEdge")
        lines += ids(e + 1 for e in owners(cut))
        lines.append("*Nset, nset=FixedBoundary")
        lines += ids(a + 1 for a in fixed_nodes)
        lines.append("*Elset, elset=FixedBoundary")
        lines += ids(e + 1 for e in owners(fixed))
        lines.append("*End Part")
    else:
        unused = sorted(int(u) for u in unused_nodes)
        file_id = np.arange(nn) + 1
        for u in unused:
            file_id[file_id >= u + 1] += 1
        lines += ["*Heading", " tool.inp", "*NODE"]
        total = nn + len(unused)
        inv = {int(f): i for i, f in enumerate(file_id)}
        lo, hi = tm.nodes.min(axis=0), tm.nodes.max(axis=0)
        for f in range(1, total + 1):
            if f in inv:
                x, y = tm.nodes[inv[f]]
            else:           # a point outside the tool, as the centre of a fillet arc
                x, y = hi[0] + (f % 7) * 1e-3, lo[1] - (f % 5) * 1e-3
            lines.append("%d, %r, %r, 0" % (f, float(x), float(y)))
        lines.append("******* E L E M E N T S *************")
        lines.append("*ELEMENT, type=T3D2, ELSET=Line1")
        edges = cut + fixed
        lines += ["%d, %d, %d" % (k + 1, file_id[a], file_id[b]) for k, (a, b) in enumerate(edges)]
        lines.append("*ELEMENT, type=CPS3, ELSET=Surface1")
        ne = len(edges)
        lines += ["%d, %d, %d, %d" % (ne + e + 1, file_id[t[0]], file_id[t[1]], file_id[t[2]])
                  for e, t in enumerate(tri)]

        def ids(values):
            values = list(values)
            return [", ".join(str(v) for v in values[k:k + 10]) + "," for k in range(0, len(values), 10)]

        lines.append("*ELSET,ELSET=FixedBoundary")
        lines += ids(range(len(cut) + 1, ne + 1))
        lines.append("*ELSET,ELSET=CuttingEdge")
        lines += ids(range(1, len(cut) + 1))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def element_matrices(p0, p1, p2, tc):
    """Lumped mass (per node) and P1 stiffness of one triangle.

    FEM_thermal_gpu.cu:71-109 and :126-152: det = |J00 J11 - J10 J01| with the
    edge vectors as Jacobian columns, mass = ((w cp) rho) det with w = 1/6,
    stiffness = ((3w) (k det)) G^T G with G = J^-T [dN/dxi].
    """
    j00, j01 = p1[0] - p0[0], p1[1] - p0[1]     # column 0
    j10, j11 = p2[0] - p0[0], p2[1] - p0[1]     # column 1
    det = abs(j00 * j11 - j10 * j01)
    w = 1.0 / 3.0 * 0.5
    mass = w * tc["cp"] * tc["rho"] * det
    # inverse of the transpose (columns (j00, j10), (j01, j11))
    dd = j00 * j11 - j01 * j10
    ood = 1.0 / dd
    i00, i01, i10, i11 = j11 * ood, -j10 * ood, -j01 * ood, j00 * ood
    # reference shape gradients of N1, N2, N3: (-1,-1), (1,0), (0,1)
    gref = ((-1.0, -1.0), (1.0, 0.0), (0.0, 1.0))
    G = [(i00 * gx + i10 * gy, i01 * gx + i11 * gy) for gx, gy in gref]
    s = (w * 3.0) * (tc["k"] * det)
    K = [[(s * G[a][0]) * G[b][0] + (s * G[a][1]) * G[b][1] for b in range(3)] for a in range(3)]
    return mass, K


def assemble_p1(nodes, elems, row_ptr, col_ind, tc, colour_order=True, n_colors=None):
    n_nodes = len(nodes)
    K = np.zeros(len(col_ind))
    M = np.zeros(n_nodes)
    order = range(len(elems))
    if colour_order:
        order = sorted(order, key=lambda e: (elems[e, 3], e))
    for e in order:
        a, b, c = elems[e, :3]
        mass, Ke = element_matrices(nodes[a], nodes[b], nodes[c], tc)
        tri = (a, b, c)
        for ia, na in enumerate(tri):
            M[na] += mass
            lo, hi = row_ptr[na], row_ptr[na + 1]
            cols = col_ind[lo:hi]
            for ib, nb in enumerate(tri):
                j = lo + int(np.searchsorted(cols, nb))
                K[j] += Ke[ia][ib]
    return K, M


def dirichlet_rows(elems, flags, fixed_elems, row_ptr, col_ind, K, M, f, T, T0, n_colors):
    """FEM_thermal_gpu.cu:234-324 applied colour by colour, element by element."""
    def find(row, col):
        lo, hi = row_ptr[row], row_ptr[row + 1]
        return lo + int(np.searchsorted(col_ind[lo:hi], col))

    for colour in range(1, n_colors + 1):
        for e in fixed_elems:
            if elems[e, 3] != colour:
                continue
            fl = flags[e]
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
                node = elems[e, k]
                for i in range(row_ptr[node], row_ptr[node + 1]):
                    loc = col_ind[i]
                    f[loc] += -K[i] * T0
                    K[i] = 0.0
                    K[find(loc, node)] = 0.0
                    if loc == node:
                        K[i] = 1.0
                M[node] = 1.0
                f[node] = T0
                T[node] = T0


# ----------------------------------------------------------------------------
# workloads
# ----------------------------------------------------------------------------
@dataclass
class Case:
    name: str
    N: int
    dx: float
    h: float
    dt: float
    pos: np.ndarray
    vel: np.ndarray
    rho: np.ndarray
    T: np.ndarray
    hh: np.ndarray
    fixed: np.ndarray
    const: Constants
    grid_bbmin: np.ndarray
    grid_bbmax: np.ndarray
    tool: ToolMesh | None
    tool_vel: np.ndarray
    wear: bool
    S: np.ndarray | None = None
    extra: dict = field(default_factory=dict)
    tool_p: np.ndarray | None = None      # tool_particle flags (sampled rigid tools)
    rigid: "RigidTool | None" = None      # rigid (SPH) tool instead of the FEM tool
    max_cell: int | None = None           # unlocked grid of this capacity instead of the locked box


@dataclass
class RigidTool:
    """A rigid tool as tool_gpu_set_up_tool receives it (tool_gpu.cu:219-266): the host geometry of
    class tool (a ctypes ToolGeometry, velocity included), the contact alpha, and the thermal block
    that interactions_setup_thermal_constants_tool gives the pair sums (interactions_gpu.cu:494-502)."""
    geometry: object
    contact_alpha: float
    trml_itool: dict


def grid_geometry(bbmin, bbmax, h):
    """grid.cu:130-149 (locked grid): margins, cell size h, ceil counts."""
    bx0, by0 = bbmin[0] - 1e-6, bbmin[1] - 1e-6
    bx1, by1 = bbmax[0] + 1e-6, bbmax[1] + 1e-6
    dxc = h * 2.0 / 2.0
    nx = int(math.ceil((bx1 - bx0) / dxc))
    ny = int(math.ceil((by1 - by0) / dxc))
    return dict(nx=nx, ny=ny, bbmin_x=bx0, bbmin_y=by0, dx=dxc, num_cell=nx * ny)


def cutting_case(name, nx, ny, dx, material, speed_m_min, doc, mu, rake, clear, r_edge,
                 tool_tris, wear=False, seed=DEFAULT_SEED, engage=0.0, tool_scale=2.5,
                 jitter=0.01) -> Case:
    """Orthogonal-cutting workload on a jittered nx x ny lattice (benchmarks_NZ.cpp:51-367)."""
    const = make_constants(material, dx, mu)
    # workpiece placed as benchmarks_NZ.cpp:67-68 (lo_x = 0.005, hi_y = 0.060): contact tests such as
    # the projection of find_shortest_distance_v1 (actions_gpu.cu:156-166) rely on coordinates away
    # from the origin, where qp.x - left.x is exact for points next to a vertical rake face
    lo_x, hi_y = 0.005, 0.060
    lo_y = hi_y - (ny - 1) * dx
    hi_x = lo_x + nx * dx
    n = nx * ny
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    pos = np.stack([ii.ravel() * dx + lo_x, jj.ravel() * dx + lo_y], axis=1)
    ju = uniform01(seed, 2 * n).reshape(n, 2)
    pos = pos + (2.0 * ju - 1.0) * (jitter * dx)
    min_x, max_x = pos[:, 0].min(), pos[:, 0].max()
    min_y = pos[:, 1].min()
    fixed = (pos[:, 1] < min_y + 0.5 * dx)
    fixed |= (pos[:, 0] < min_x + dx / 2) & (pos[:, 1] < min_y + 0.5 * (hi_y - lo_y))
    fixed |= pos[:, 0] > max_x - dx
    h = HDX * dx
    vel_tool = m_per_min_to_cm_per_us(speed_m_min)

    L = H = tool_scale * doc
    tm = mesh_tool(L, H, rake, clear, r_edge, tool_tris, const.tool)
    # place the tool: cutting edge at depth doc, rake face `engage` dx into the workpiece
    shift = np.array([lo_x + engage * dx - tm.right_point, (hi_y - doc) - tm.lowest_point])
    for arr in (tm.nodes, tm.seg_left, tm.seg_right, tm.tl, tm.br):
        arr += shift
    tm.lowest_point += shift[1]
    tm.right_point += shift[0]

    bbmin = pos.min(axis=0) - 1e-8
    bbmax = pos.max(axis=0) + 1e-8 + np.array([300 * dx, 500 * dx])
    bbmax[1] = max(bbmax[1], hi_y + hi_x - lo_x)
    dt = 0.1 * HDX * dx / (math.sqrt(const.phys["K"] / const.phys["rho0"]) + math.sqrt(vel_tool))
    return Case(name=name, N=n, dx=dx, h=h, dt=dt, pos=pos, vel=np.zeros((n, 2)),
                rho=np.full(n, const.phys["rho0"]), T=np.full(n, T_REF), hh=np.full(n, h),
                fixed=fixed.astype(np.float64), const=const, grid_bbmin=bbmin, grid_bbmax=bbmax,
                tool=tm, tool_vel=np.array([vel_tool, 0.0]), wear=wear)


def small_cutting_case(nx=48, ny=16, seed=DEFAULT_SEED, wear=False, tool_tris=600, material="Ti6Al4V",
                       engage=0.0):
    """Reduced cutting workload for fast parity tests (same physics as the 6,400-particle case)."""
    dx = 2.5e-4
    doc = 4e-3 * ny / 40.0
    return cutting_case(f"small_{nx}x{ny}", nx, ny, dx, material, 70.0, doc, 0.35, 0.0, 7.0,
                        5e-4 * ny / 40.0 * 2, tool_tris, wear=wear, seed=seed, engage=engage)


def config1(seed=DEFAULT_SEED):
    """6,400-particle Ti6Al4V case: 400x100 um, dx 2.5 um, ~2k-triangle tool, 70 m/min."""
    return cutting_case("ti64_6400", 160, 40, 2.5e-4, "Ti6Al4V", 70.0, 4e-3, 0.35, 0.0, 7.0,
                        5e-4, 2000, wear=False, seed=seed)


def config2(seed=DEFAULT_SEED):
    """160k-particle AISI 1045 case: 1.0x0.25 mm, dx 1.25 um, ~8k-triangle tool, 200 m/min."""
    return cutting_case("aisi1045_160k", 800, 200, 1.25e-4, "AISI1045", 200.0, 5e-3, 0.35, 0.0, 7.0,
                        5e-4, 8000, wear=False, seed=seed)


def config3(seed=DEFAULT_SEED, nx=4000, ny=1000, tool_tris=50000, wear=True, speed_m_min=70.0):
    """4M-particle Ti6Al4V case: 4x1 mm, dx 1 um, ~50k-triangle tool, FEM heat + Usui wear, cutting
    speed 70 m/min unless given."""
    return cutting_case(f"ti64_{nx * ny}", nx, ny, 1e-4, "Ti6Al4V", speed_m_min, 5e-3, 0.35, 0.0, 7.0,
                        5e-4, tool_tris, wear=wear, seed=seed)


def batch_cases(n_cases=64, seed=DEFAULT_SEED, dx=1.25e-4, tool_tris=1000, rows=(50, 125)):
    """Parameter sweep of independent cutting cases (dx = 1.25 um): workpiece length U(0.5, 3) mm,
    height U(rows) lattice rows (about 20k-300k particles per case), cutting speed U(30, 300) m/min,
    rake angle U(-10, 15) deg, depth of cut U(20, 80) um, Ti6Al4V or AISI 1045 at random."""
    u = uniform01(seed, 8 * n_cases, stream=20).reshape(n_cases, 8)
    out = []
    for k in range(n_cases):
        length = 0.05 + 0.25 * u[k, 0]
        nx = int(round(length / dx))
        ny = int(round(rows[0] + (rows[1] - rows[0]) * u[k, 1]))
        speed = 30.0 + 270.0 * u[k, 2]
        rake = -10.0 + 25.0 * u[k, 3]
        doc = 2e-3 + 6e-3 * u[k, 4]
        material = "Ti6Al4V" if u[k, 5] < 0.5 else "AISI1045"
        doc = min(doc, 0.8 * ny * dx)
        out.append(cutting_case(f"batch{k:02d}_{material}_{nx}x{ny}", nx, ny, dx, material, speed, doc, 0.35,
                                rake, 7.0, 5e-4, tool_tris, wear=False, seed=seed + 1000 + k))
    return out


def chip_case(n_target=16_000_000, seed=DEFAULT_SEED, dx=1e-4, material="Ti6Al4V"):
    """Kernel-only workload: slab plus inclined chip band with a 10:1 density ramp.

    Positions come from a jittered lattice whose spacing is compressed by up to
    sqrt(10) per axis towards the chip root; velocity ~ N(0, 5 m/s) per axis,
    deviatoric stress ~ N(0, 300 MPa) with Szz = -(Sxx + Syy), rho = rho0, T = 300.
    """
    const = make_constants(material, dx, 0.35)
    # slab of aspect 4:1 (4 base^2 lattice points) with a chip band of about 0.35 base^2 points on
    # top, so base = sqrt(n_target / 4.35) gives N just above n_target.
    # region A: slab [0, Lx] x [0, Ly]; region B: band along the rake at 45 degrees.
    ratio = 4.0
    base = int(math.ceil(math.sqrt(n_target / 4.35)))
    nxs, nys = int(ratio * base), base
    Lx, Ly = nxs * dx, nys * dx
    root = np.array([0.6 * Lx, Ly])
    pts = []
    # slab with local compression: map a uniform lattice through a density-preserving warp
    u = (np.arange(nxs) + 0.5) / nxs
    v = (np.arange(nys) + 0.5) / nys
    U, V = np.meshgrid(u, v, indexing="ij")
    X = U * Lx
    Y = V * Ly
    pts.append(np.stack([X.ravel(), Y.ravel()], axis=1))
    # chip band: parallelogram from the root, inclined at 45 degrees, thickness 0.25 Ly
    nb_t = max(4, int(0.25 * nys))
    nb_l = int(1.2 * nys / math.sqrt(2)) * 2
    s = (np.arange(nb_l) + 0.5) / nb_l
    t = (np.arange(nb_t) + 0.5) / nb_t
    Sg, Tg = np.meshgrid(s, t, indexing="ij")
    d_along = np.array([-1.0, 1.0]) / math.sqrt(2)
    d_across = np.array([1.0, 1.0]) / math.sqrt(2)
    band = (root[None, :] + (Sg.ravel()[:, None] * (1.2 * Ly)) * d_along[None, :]
            + (Tg.ravel()[:, None] * (0.25 * Ly)) * d_across[None, :] + np.array([0, 0.5 * dx]))
    pts.append(band[band[:, 1] > Ly + 0.5 * dx])
    P = np.concatenate(pts)
    # compress towards the chip root: r -> r * f(r), 10:1 number density at the root
    dvec = P - root
    dist = np.sqrt(np.sum(dvec ** 2, axis=1))
    L0 = 0.3 * Ly
    scale = 1.0 - (1.0 - 1.0 / math.sqrt(10.0)) * np.exp(-(dist / L0) ** 2)
    P = root + dvec * scale[:, None]
    n = len(P)
    jit = (2.0 * uniform01(seed, 2 * n).reshape(n, 2) - 1.0) * (0.01 * dx) * scale[:, None]
    P = P + jit
    vel = normal01(seed, 2 * n, stream=2).reshape(n, 2) * m_per_min_to_cm_per_us(5.0 * 60.0)
    sd = normal01(seed, 3 * n, stream=4).reshape(n, 3) * 3e-3
    S = np.zeros((n, 4))
    S[:, 0], S[:, 1], S[:, 2] = sd[:, 0], sd[:, 1], sd[:, 2]
    S[:, 3] = -(S[:, 0] + S[:, 2])
    bbmin = P.min(axis=0) - 1e-8
    bbmax = P.max(axis=0) + 1e-8
    h = HDX * dx
    dt = 0.1 * HDX * dx / math.sqrt(const.phys["K"] / const.phys["rho0"])
    return Case(name=f"chip_{n}", N=n, dx=dx, h=h, dt=dt, pos=P, vel=vel,
                rho=np.full(n, const.phys["rho0"]), T=np.full(n, T_REF), hh=np.full(n, h),
                fixed=np.zeros(n), const=const, grid_bbmin=bbmin, grid_bbmax=bbmax, tool=None,
                tool_vel=np.zeros(2), wear=False, S=S)


# ----------------------------------------------------------------------------
# initial device-state images in the reference layout
# ----------------------------------------------------------------------------
PARTICLE_FIELDS = (
    # name, trailing shape, dtype  (particle_gpu.h:23-83 order)
    ("pos", (2,), np.float64), ("vel", (2,), np.float64), ("h", (), np.float64),
    ("rho", (), np.float64), ("p", (), np.float64), ("S", (4,), np.float64), ("R", (4,), np.float64),
    ("fc", (2,), np.float64), ("ft", (2,), np.float64), ("n", (2,), np.float64),
    ("fixed", (), np.float64), ("blanked", (), np.float64), ("tool_particle", (), np.float64),
    ("eps_pl", (), np.float64), ("eps_pl_dot", (), np.float64), ("T", (), np.float64),
    ("pos_t", (2,), np.float64), ("vel_t", (2,), np.float64), ("rho_t", (), np.float64),
    ("S_t", (4,), np.float64), ("T_t", (), np.float64), ("S_der", (4,), np.float64),
    ("v_der", (4,), np.float64), ("idx", (), np.int32), ("hash", (), np.int32), ("on_seg", (), np.int32),
)


def perturbed(case: Case, seed: int = DEFAULT_SEED + 1, vel_sd: float = 5e-4, S_sd: float = 3e-3,
              rho_rel_sd: float = 1e-3, T_sd: float = 50.0) -> Case:
    """The same case with seeded random fields on the non-fixed particles: velocity N(0, vel_sd)
    per axis (cm/us), deviatoric stress N(0, S_sd) per component with S_zz = -(S_xx + S_yy),
    density rho0 * (1 + N(0, rho_rel_sd)) and temperature T + |N(0, T_sd)|.  Every term of the
    pair sums is then non-zero on the first step (from rest they all vanish)."""
    import dataclasses
    n = case.N
    free = (case.fixed == 0.0)[:, None]
    vel = case.vel + free * normal01(seed, 2 * n, stream=10).reshape(n, 2) * vel_sd
    sd = normal01(seed, 3 * n, stream=12).reshape(n, 3) * S_sd
    S = np.zeros((n, 4))
    S[:, 0], S[:, 1], S[:, 2] = sd[:, 0], sd[:, 1], sd[:, 2]
    S[:, 3] = -(S[:, 0] + S[:, 2])
    S0 = case.S if case.S is not None else np.zeros((n, 4))
    S = S0 + free * S
    rho = case.rho * (1.0 + free[:, 0] * normal01(seed, n, stream=14) * rho_rel_sd)
    T = case.T + free[:, 0] * np.abs(normal01(seed, n, stream=15)) * T_sd
    return dataclasses.replace(case, name=case.name + "_perturbed", vel=vel, S=S, rho=rho, T=T)


def ref_cut_sph(ny: int = 30, rake: float = 0.0, clear: float = 4.5, chamfer: float = 0.00035,
                speed: float = 83.333328e-5 / 200 * 300, feed: float = 0.01) -> Case:
    """setup_ref_cut (benchmarks.cpp:248-530): Ti6Al4V (Johnson 85) cut by a rigid tool sampled with
    tool particles, at the reference driver's defaults (mfree_gpu.cpp:155-162, 197: rake 0, clearance
    4.5 deg, fillet 3.5 um, ny 30, feed 100 um, 750 m/min).  The tool geometry comes from the library's
    host constructors of class tool (tool.cpp:335-406)."""
    from . import _lib
    import ctypes as C
    L = _lib.lib()

    hi_x, hi_y = 0.200, 0.060
    lo_x, lo_y = 0.005, hi_y - 3 * feed
    dy = (hi_y - lo_y) / (ny - 1)
    dx = dy
    nx = int((hi_x - lo_x) / dx)
    hdx = 1.5
    phys = dict(E=1.1, nu=0.35, rho0=4.429998)
    phys["G"] = phys["E"] / (2. * (1. + phys["nu"]))
    phys["K"] = 2.0 * phys["G"] * (1 + phys["nu"]) / (3 * (1 - 2 * phys["nu"]))
    phys["mass"] = dx * dx * phys["rho0"]
    rho_tool = 15.25
    phys["mass_tool"] = dx * dx * rho_tool
    joco = dict(A=0.0085200, B=0.0033890, C=0.02754, m=0.5961, n=0.148, Tref=300., Tmelt=1836.0,
                eps_dot_ref=1e-6, tanh_a=1.1, tanh_b=0.4, tanh_c=12., tanh_d=1.)
    trml_wp = dict(cp=580 * 1e-8, tq=0.9, k=7.3 * 1e-13, alpha=0.0, T_init=joco["Tref"], eta=0.9)
    trml_wp["alpha"] = trml_wp["k"] / (phys["rho0"] * trml_wp["cp"])
    trml_tool = dict(cp=292 * 1e-08, tq=.0, k=88 * 1e-13, alpha=0.0, T_init=joco["Tref"], eta=0.0)
    trml_tool["alpha"] = trml_tool["k"] / (rho_tool * trml_tool["cp"])
    corr = dict(wdeltap=1.0, stresseps=0.3, xspheps=0.5, alpha=1., beta=1., eta=0.1)
    h1 = 1. / (hdx * dx)
    q = dx * h1
    fac = 10 * M_1_PI / 7.0 * h1 * h1
    corr["wdeltap"] = fac * (1 - 1.5 * q * q * (1 - 0.5 * q))
    zeros_tool = dict(cp=0., tq=0., k=0., eta=0., rho=0., T0=0., h=0., beta=0., mu=0., contact_alpha=0.,
                      slave_mass=0., hc=0.)
    const = Constants(phys, trml_wp, trml_tool, corr, joco, zeros_tool, dict(A=0., B=0., C=0., D=0.))

    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    pos = np.stack([ii.ravel() * dx + lo_x, jj.ravel() * dx + lo_y], axis=1)
    n = nx * ny
    min_x, max_x = pos[:, 0].min(), pos[:, 0].max()
    min_y = pos[:, 1].min()
    fixed = pos[:, 1] < min_y + 0.5 * dx
    fixed |= (pos[:, 0] < min_x + dx / 2) & (pos[:, 1] < min_y + 0.5 * (hi_y - lo_y))
    fixed |= pos[:, 0] > max_x - dx

    vel_tool = 83.333328 * 1e-5 if speed == 0.0 else speed
    mu_fric = 0.35
    nudge = 0.004
    tl = _lib.Double2(-0.05 + nudge, 0.0986074)
    l_tool = -0.0086824 - -0.05
    h_tool = 0.0986074 - 0.0555074
    g = _lib.ToolGeometry()
    _lib.check(L.sphfem_tool_geometry_from_angles(tl, l_tool, h_tool, rake, clear, chamfer, mu_fric, C.byref(g)))
    # move the tool to the workpiece's left edge and to the target feed (benchmarks.cpp:393-408)
    distance = g.front - lo_x
    g.vel = _lib.Double2(vel_tool, 0.)
    _lib.check(L.sphfem_tool_geometry_move(C.byref(g), -(distance / vel_tool)))
    low = C.c_double()
    _lib.check(L.sphfem_tool_geometry_low(C.byref(g), C.byref(low)))
    current_feed = hi_y - low.value
    correction_time = abs(current_feed - feed) / vel_tool
    sign = 1. if current_feed > feed else -1.
    g.vel = _lib.Double2(0., vel_tool)
    _lib.check(L.sphfem_tool_geometry_move(C.byref(g), correction_time * sign))
    g.vel = _lib.Double2(vel_tool, 0.)

    cnt = C.c_long()
    _lib.check(L.sphfem_tool_geometry_sample(C.byref(g), dx, None, 0, C.byref(cnt)))
    samples = np.zeros((cnt.value, 2))
    _lib.check(L.sphfem_tool_geometry_sample(C.byref(g), dx, samples.ctypes.data, cnt.value, C.byref(cnt)))
    tfixed = (samples[:, 0] < samples[:, 0].min() + dx / 2) | (samples[:, 1] > samples[:, 1].max() - dx / 2)

    allpos = np.concatenate([pos, samples])
    N = len(allpos)
    bbmin = allpos.min(axis=0) - 1e-8
    bbmax = allpos.max(axis=0) + 1e-8
    bb = _lib.BBox()
    _lib.check(L.sphfem_tool_geometry_safe_bb(C.byref(g), 1e-3, C.byref(bb)))
    bbmin = np.array([min(bbmin[0], bb.bbmin_x), min(bbmin[1], bb.bbmin_y)])
    bbmax = np.array([max(bbmax[0], bb.bbmax_x) + 10 * dx, max(bbmax[1], bb.bbmax_y)])
    bbmax[1] = max(bbmax[1], hi_y + hi_x - lo_x)

    dt = 0.32 * hdx * dx / (math.sqrt(phys["K"] / phys["rho0"]) + math.sqrt(vel_tool))
    h = hdx * dx
    rho = np.concatenate([np.full(n, phys["rho0"]), np.full(len(samples), rho_tool)])
    return Case(name=f"ref_cut_sph_ny{ny}", N=N, dx=dx, h=h, dt=dt, pos=allpos, vel=np.zeros((N, 2)), rho=rho,
                T=np.full(N, joco["Tref"]), hh=np.full(N, h),
                fixed=np.concatenate([fixed, tfixed]).astype(np.float64), const=const, grid_bbmin=bbmin,
                grid_bbmax=bbmax, tool=None, tool_vel=np.array([vel_tool, 0.]), wear=False,
                tool_p=np.concatenate([np.zeros(n), np.ones(len(samples))]),
                rigid=RigidTool(geometry=g, contact_alpha=0.075, trml_itool=dict(trml_tool)),
                extra=dict(t_final=0.1 / vel_tool, n_workpiece=n))


# ----------------------------------------------------------------------------
# textbook validation cases (benchmarks.cpp:16-246), SI units as there
# ----------------------------------------------------------------------------
def _zero_trml():
    return dict(cp=0.0, tq=0.0, k=0.0, alpha=0.0, T_init=0.0, eta=0.0)


def _corr_for(hdx, dx):
    corr = dict(wdeltap=1.0, stresseps=0.3, xspheps=0.5, alpha=1., beta=1., eta=0.1)
    h1 = 1. / (hdx * dx)
    q = dx * h1
    fac = 10 * M_1_PI / 7.0 * h1 * h1
    corr["wdeltap"] = fac * (1 - 1.5 * q * q * (1 - 0.5 * q))
    return corr


_ZERO_TOOLC = dict(cp=0., tq=0., k=0., eta=0., rho=0., T0=0., h=0., beta=0., mu=0., contact_alpha=0., slave_mass=0.,
                   hc=0.)
_ZERO_WEAR = dict(A=0., B=0., C=0., D=0.)


def rings(nbox: int = 80) -> Case:
    """setup_rings (benchmarks.cpp:148-246): two rubber-like rings (inner radius 3 cm, outer 4 cm,
    E = 10 MPa, nu = 0.4, rho = 1) approaching each other at 200 m/s each on a locked grid; no
    plasticity (the Johnson-Cook block is never set up), no heat.  particle_gpu(pos, vel, rho, h,
    n): T and fixed are zero.  The reference driver calls it with nbox = 80 (mfree_gpu.cpp:224)."""
    ri, ro = 0.03, 0.04
    spacing = ro + 0.001
    dx = 2 * ro / (nbox - 1)
    hdx = 1.7
    phys = dict(E=1e7, nu=0.4, rho0=1.)
    phys["G"] = phys["E"] / (2. * (1. + phys["nu"]))
    phys["K"] = 2.0 * phys["G"] * (1 + phys["nu"]) / (3 * (1 - 2 * phys["nu"]))
    phys["mass"] = dx * dx * phys["rho0"]
    phys["mass_tool"] = 0.0
    pos = []
    for i in range(nbox):
        for j in range(nbox):
            px, py = -ro + i * dx, -ro + j * dx
            dist = math.sqrt(px * px + py * py)
            if ro > dist >= ri:
                pos.append((px - spacing, py))
                pos.append((px + spacing, py))
    pos = np.array(pos)
    n = len(pos)
    vel_ring = 170.
    dt = 0.5 * hdx * dx / (math.sqrt(phys["K"] / phys["rho0"]) + math.sqrt(vel_ring))
    vel = np.zeros((n, 2))
    vel[:, 0] = np.where(pos[:, 0] < 0, 200., -200.)
    const = Constants(phys, _zero_trml(), _zero_trml(), _corr_for(hdx, dx), None, dict(_ZERO_TOOLC),
                      dict(_ZERO_WEAR))
    h = hdx * dx
    return Case(name=f"rings_{nbox}", N=n, dx=dx, h=h, dt=dt, pos=pos, vel=vel, rho=np.full(n, phys["rho0"]),
                T=np.zeros(n), hh=np.full(n, h), fixed=np.zeros(n), const=const,
                grid_bbmin=np.array([-5 * ro, -2 * ro]), grid_bbmax=np.array([5 * ro, 2 * ro]), tool=None,
                tool_vel=np.zeros(2), wear=False, extra=dict(t_final=2.5 * ro / vel_ring, ri=ri, ro=ro))


def impact(nbox: int = 20) -> Case:
    """setup_impact (benchmarks.cpp:16-146): a steel 4340 disc (radius 0.2 m) at 1000 m/s against a
    wall of nbox/2 columns fixed at top and bottom, Johnson-Cook plasticity (no heat: the thermal
    blocks are never set up), on an unlocked grid grid_gpu_green(10 n, n).  The reference driver
    calls it with nbox = 20 (mfree_gpu.cpp:225)."""
    ro = 0.2
    dx = 2 * ro / (nbox - 1)
    hdx = 1.7
    phys = dict(E=200e9, nu=0.29, rho0=7830.0)
    phys["G"] = phys["E"] / (2. * (1. + phys["nu"]))
    phys["K"] = 2.0 * phys["G"] * (1 + phys["nu"]) / (3 * (1 - 2 * phys["nu"]))
    phys["mass"] = dx * dx * phys["rho0"]
    phys["mass_tool"] = 0.0
    joco = dict(A=792.0e6, B=510.0e6, C=0.014, m=1.03, n=0.26, Tref=273.0, Tmelt=1793.0, eps_dot_ref=1.,
                tanh_a=0., tanh_b=0., tanh_c=0., tanh_d=0.)
    spacing = ro + 0.05
    pos = []
    for i in range(nbox):
        for j in range(nbox):
            px, py = -ro + i * dx, -ro + j * dx
            if math.sqrt(px * px + py * py) < ro:
                pos.append((px - spacing, py))
    n_flyer = len(pos)
    for i in range(nbox // 2):
        for j in range(10 * nbox):
            if j * dx - 1. > 1.:
                continue
            pos.append((i * dx, j * dx - 1.))
    pos = np.array(pos)
    n = len(pos)
    vel_flyer = 1e3
    dt = 0.3 * hdx * dx / (math.sqrt(phys["K"] / phys["rho0"]) + math.sqrt(vel_flyer))
    vel = np.zeros((n, 2))
    vel[:n_flyer, 0] = 1e3
    fixed = ((pos[:, 1] < -1 + dx / 2) | (pos[:, 1] > 1 - dx / 2)).astype(np.float64)
    const = Constants(phys, _zero_trml(), _zero_trml(), _corr_for(hdx, dx), joco, dict(_ZERO_TOOLC),
                      dict(_ZERO_WEAR))
    h = hdx * dx
    return Case(name=f"impact_{nbox}", N=n, dx=dx, h=h, dt=dt, pos=pos, vel=vel, rho=np.full(n, phys["rho0"]),
                T=np.full(n, joco["Tref"]), hh=np.full(n, h), fixed=fixed, const=const,
                grid_bbmin=pos.min(axis=0), grid_bbmax=pos.max(axis=0), tool=None, tool_vel=np.zeros(2),
                wear=False, max_cell=10 * n, extra=dict(t_final=2 * ro / vel_flyer, n_flyer=n_flyer))


def initial_particles(case: Case) -> dict:
    """State after particle_gpu(pos, vel, rho, T, h, fixed, N) (particle_gpu.cu:11-104, 203-206):
    every array zero, idx = 0..N-1, on_seg = 0, then the given fields copied in."""
    N = case.N
    out = {name: np.zeros((N,) + shp, dtype=dt) for name, shp, dt in PARTICLE_FIELDS}
    out["idx"][:] = np.arange(N, dtype=np.int32)
    out["pos"][:] = case.pos
    out["vel"][:] = case.vel
    out["rho"][:] = case.rho
    out["T"][:] = case.T
    out["h"][:] = case.hh
    out["fixed"][:] = case.fixed
    if case.tool_p is not None:
        out["tool_particle"][:] = case.tool_p
    if case.S is not None:
        out["S"][:] = case.S
    return out


def initial_tool(case: Case) -> dict | None:
    """Per-step tool state after set_up_segments_from_mesh / set_up_FEM_tool_gpu (tool_FEM.cu:771-1039,
    1539-1544): nodal fields, CSR system, segment arrays with zeroed accumulators, kinematics."""
    tm = case.tool
    if tm is None:
        return None
    n, s = len(tm.nodes), len(tm.seg_left)
    return dict(
        node_pos=tm.nodes.copy(), elem=tm.elems.copy(),
        layer=(tm.layers[:, 0] | tm.layers[:, 1]).astype(np.int32),
        row_ptr=tm.row_ptr.copy(), col_ind=tm.col_ind.copy(),
        K=tm.K.copy(), K_temp=np.zeros_like(tm.K), M=tm.M.copy(), Minv=1.0 / tm.M,
        f=tm.f.copy(), f_temp=np.zeros(n), T=tm.T.copy(), T_old=tm.T.copy(), k=np.zeros(n),
        left=tm.seg_left.copy(), right=tm.seg_right.copy(), n=tm.seg_n.copy(),
        heat_exchange=np.zeros((s, 2)), fric_heat=np.zeros(s), wear_rate=np.zeros((s, 2)),
        wear_nodes=np.zeros(s), physical_para=np.zeros((s, 4)), sliding_force=np.zeros(s),
        marks=tm.seg_marks.copy(),
        node_flags=(tm.cut_edge_node.astype(np.int32) | (tm.fixed_edge_node.astype(np.int32) << 1)),
        fixed_elems=fixed_boundary(tm)[0], fixed_flags=fixed_boundary(tm)[1],
        scalars=dict(vel_x=float(case.tool_vel[0]), vel_y=float(case.tool_vel[1]),
                     tl_x=float(tm.tl[0]), tl_y=float(tm.tl[1]), br_x=float(tm.br[0]), br_y=float(tm.br[1]),
                     ref_dist=float(tm.ref_dist), lowest_point=float(tm.lowest_point),
                     right_point=float(tm.right_point), flank_retreat=0.0,
                     flank_contact=float(tm.lowest_point)),
        n_colors=tm.n_colors, n_cut_colors=tm.n_cut_colors,
    )


def default_rr2_min(dx: float) -> float:
    """Lower pair cut-off of kernels.cuh:38: the reference's 1e-8 cm^2 while the lattice spacing is
    at least 1.25 um; below that 1e-8 reaches the squared spacing, so the cut-off is scaled with dx^2
    to keep only coincident particles out (1e-10 cm^2 at dx = 1 um)."""
    return 1e-8 if dx >= 1.25e-4 - 1e-12 else 1e-2 * dx * dx
