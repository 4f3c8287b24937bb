"""Tool mesh read from .inp files (mesh_structure_GPU.cu:213-1462, tool_FEM.cu:771-1039,
FEM_thermal_gpu.cu:126-330, 584-659): the library's host-side reader, colouring, CSR, segments and
FEM system, on the CPU.

Pinned three ways:
* round trip: a generated tool mesh written in either dialect reads back as the same nodes,
  triangles, boundary flags and cutting segments (chain order, inward normals), also with unused
  Gmsh nodes that renumber the rest;
* invariants and closed forms: a proper colouring within the palettes, layer rules, CSR equal to
  the brute-force node adjacency, stiffness rows that annihilate linear fields at interior nodes,
  lumped masses equal to a third of the adjacent triangle areas times cp rho, Dirichlet rows;
  a hand-worked two-triangle file for the all-edges-flagged correction;
* the Python restatement of the reference (oracle/mesh_inp.py) gives the same arrays, bit for bit.
"""
import numpy as np
import pytest

from oracle import mesh_inp as R
from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S


@pytest.fixture(scope="module")
def case():
    return cases.small_cutting_case(nx=48, ny=16)


def _boundary_count(tm):
    return int(np.count_nonzero(tm.cut_edge_node | tm.fixed_edge_node))


def _edge_flags(tri, edges):
    want = {(min(a, b), max(a, b)) for a, b in edges}
    fl = np.zeros((len(tri), 3), np.int32)
    for e, t in enumerate(tri):
        for k, (u, v) in enumerate(((t[0], t[1]), (t[1], t[2]), (t[2], t[0]))):
            fl[e, k] = (min(u, v), max(u, v)) in want
    return fl


def _ref_mesh(path, gmsh):
    text = open(path).read()
    return R.read_gmsh(text) if gmsh else R.read_abaqus(text)


@pytest.mark.parametrize("gmsh", [True, False])
def test_round_trip_matches_generated_mesh(tmp_path, case, gmsh):
    tm = case.tool
    path = str(tmp_path / "tool.inp")
    nbd = _boundary_count(tm)
    cases.write_inp(tm, path, gmsh=gmsh, unused_nodes=(nbd, nbd + 7, nbd + 40) if gmsh else ())
    m = S.mesh_read(path, gmsh)
    a = m.arrays()
    tri = tm.elems[:, :3]
    if gmsh:
        assert np.array_equal(a["position"], tm.nodes)
    else:
        # the Abaqus reader divides the written coordinates (metres x 100 of the file) by 100
        exp = np.array([[float(repr(float(x) * 100)) / 100 for x in row] for row in tm.nodes])
        assert np.array_equal(a["position"], exp)
        assert np.max(np.abs(a["position"] - tm.nodes)) < 1e-15
    assert np.array_equal(a["elem"][:, :3], tri)
    assert np.array_equal(a["cut_node"], tm.cut_edge_node)
    assert np.array_equal(a["fixed_node"], tm.fixed_edge_node)
    cut, fixed = cases.tool_mesh_edges(tm)
    assert np.array_equal(a["cut_flags"], _edge_flags(tri, cut))
    assert np.array_equal(a["fixed_flags"], _edge_flags(tri, fixed))
    assert np.array_equal(a["cut_elems"], np.nonzero(_edge_flags(tri, cut).any(1))[0])
    assert np.array_equal(a["fixed_elems"], np.nonzero(_edge_flags(tri, fixed).any(1))[0])
    got = m.tool_mesh(case.const.tool, tm.tl, case.tool_vel)
    # cutting segments in chain order from the top of the rake face, normals into the tool
    assert np.array_equal(got.seg_marks[:, 2:4], tm.seg_marks[:, 2:4])
    assert np.array_equal(got.seg_marks[:, 0], tm.seg_marks[:, 0])
    assert np.array_equal(got.seg_marks[:, 1], a["elem"][got.seg_marks[:, 0], 3])
    if gmsh:
        for f in ("seg_left", "seg_right", "tl", "br"):
            assert np.array_equal(getattr(got, f), getattr(tm, f)), f
        # the generated normals come from the coordinates before the tool was moved into place
        assert np.allclose(got.seg_n, tm.seg_n, rtol=0, atol=1e-12)
        assert got.lowest_point == tm.lowest_point and got.right_point == tm.right_point
    assert got.ref_dist == pytest.approx(tm.ref_dist, rel=1e-13)
    assert np.array_equal(got.seg_left[1:], got.seg_right[:-1])
    d = tm.tl[None, :] - got.seg_left
    assert np.all(np.sum(d * got.seg_n, axis=1) > 0)
    assert np.allclose(np.hypot(got.seg_n[:, 0], got.seg_n[:, 1]), 1.0, rtol=0, atol=1e-15)
    # the same FEM system as the generated mesh, which was assembled before the tool was moved into
    # place and in the order of its own colouring: measured 3e-13 relative on the masses
    row = np.repeat(np.arange(len(tm.M)), np.diff(tm.row_ptr))
    row_scale = np.zeros(len(tm.M))
    np.maximum.at(row_scale, row, np.abs(tm.K))
    assert np.all(np.abs(got.K - tm.K) <= 1e-11 * row_scale[row])
    assert np.allclose(got.M, tm.M, rtol=1e-11, atol=0)
    assert np.all(np.abs(got.f - tm.f) <= 1e-11 * row_scale * case.const.tool["T0"])
    assert np.array_equal(got.T, tm.T)
    m.close()


@pytest.mark.parametrize("gmsh", [True, False])
def test_reader_matches_restatement(tmp_path, case, gmsh):
    tm = case.tool
    path = str(tmp_path / "tool.inp")
    nbd = _boundary_count(tm)
    cases.write_inp(tm, path, gmsh=gmsh, unused_nodes=(nbd, nbd + 2) if gmsh else ())
    m = S.mesh_read(path, gmsh)
    a = m.arrays()
    r = _ref_mesh(path, gmsh)
    _same_mesh(a, r)
    tc = case.const.tool
    got = m.tool_mesh(tc, tm.tl, case.tool_vel)
    col = a["elem"][:, 3].tolist()
    rp, ci = R.csr(r)
    seg = R.segments(r, col, tuple(tm.tl))
    for f, g in (("left", "seg_left"), ("right", "seg_right"), ("n", "seg_n"), ("marks", "seg_marks"),
                 ("tl", "tl"), ("br", "br")):
        assert np.array_equal(getattr(got, g), seg[f]), f
    assert got.ref_dist == seg["ref_dist"] and got.lowest_point == seg["lowest_point"]
    assert got.right_point == seg["right_point"]
    K, M, f, T = R.fem_system(r, col, rp, ci, tc)
    assert np.array_equal(got.K, K) and np.array_equal(got.M, M)
    assert np.array_equal(got.f, f) and np.array_equal(got.T, T)
    m.close()


def _same_mesh(a, r):
    """library arrays == restatement arrays, colours, layers and CSR included"""
    assert np.array_equal(a["position"][:, 0], np.array(r["x"]))
    assert np.array_equal(a["position"][:, 1], np.array(r["y"]))
    assert np.array_equal(a["elem"][:, :3], np.array(r["elem"]))
    for k in ("cut_node", "fixed_node", "cut_flags", "fixed_flags", "cut_elems", "fixed_elems"):
        assert np.array_equal(a[k], np.array(r[k], dtype=np.int32).reshape(a[k].shape)), k
    col, layers = R.colour_and_layers(r)
    assert np.array_equal(a["elem"][:, 3], col)
    assert np.array_equal(a["layers"], np.array(layers))
    rp, ci = R.csr(r)
    assert np.array_equal(a["row_ptr"], rp) and np.array_equal(a["col_ind"], ci)


def test_gmsh_double_renumbering(tmp_path, case):
    """An unused node numbered before boundary nodes: triangles are renumbered once (and stay the
    generated ones), boundary-edge nodes twice, as the reference reader does; both the library and
    the restatement follow the closed form u - 1 - j(u) - j(u - 1 - j(u))."""
    tm = case.tool
    path = str(tmp_path / "tool.inp")
    cases.write_inp(tm, path, gmsh=True, unused_nodes=(3,))
    m = S.mesh_read(path, True)
    a = m.arrays()
    assert np.array_equal(a["elem"][:, :3], tm.elems[:, :3])
    assert np.array_equal(a["position"], tm.nodes)
    exclude = [4]                       # file id of the unused node
    j = lambda v: sum(1 for e in exclude if e < v)
    cut, _ = cases.tool_mesh_edges(tm)
    flagged = set()
    for u, v in cut:
        fu, fv = u + 1 + (u >= 3), v + 1 + (v >= 3)      # file ids
        ru, rv = fu - 1 - j(fu), fv - 1 - j(fv)
        flagged |= {ru - j(ru), rv - j(rv)}
    assert set(np.nonzero(a["cut_node"])[0].tolist()) == flagged
    _same_mesh(a, _ref_mesh(path, True))
    m.close()


def test_colouring_layers_and_csr_invariants(tmp_path, case):
    tm = case.tool
    path = str(tmp_path / "tool.inp")
    cases.write_inp(tm, path, gmsh=True)
    m = S.mesh_read(path, True)
    a = m.arrays()
    elem, col = a["elem"][:, :3], a["elem"][:, 3]
    nn, ne = len(a["position"]), len(elem)
    inc = np.zeros((ne, nn), bool)
    inc[np.repeat(np.arange(ne), 3), elem.ravel()] = True
    share = (inc.astype(np.int32) @ inc.T.astype(np.int32)) > 0
    np.fill_diagonal(share, False)
    same = col[:, None] == col[None, :]
    assert not np.any(share & same), "two elements sharing a node have the same colour"
    assert col.min() >= 1 and col.max() <= 18
    cut = a["cut_elems"]
    assert col[cut].max() <= 6
    assert np.array_equal(np.nonzero(a["layers"][:, 0])[0], cut)
    cut_nodes = set(np.nonzero(a["cut_node"])[0]) | set(elem[cut].ravel())
    l2 = np.array([a["layers"][e, 0] == 0 and bool(cut_nodes & set(elem[e])) for e in range(ne)])
    assert np.array_equal(a["layers"][:, 1].astype(bool), l2)
    adj = (inc.T.astype(np.int32) @ inc.astype(np.int32)) > 0
    for i in range(nn):
        cols = a["col_ind"][a["row_ptr"][i]:a["row_ptr"][i + 1]]
        assert np.array_equal(cols, np.nonzero(adj[i])[0])
    m.close()


def test_fem_system_closed_forms(tmp_path, case):
    tm = case.tool
    tc = case.const.tool
    path = str(tmp_path / "tool.inp")
    cases.write_inp(tm, path, gmsh=True)
    m = S.mesh_read(path, True)
    a = m.arrays()
    got = m.tool_mesh(tc, tm.tl, case.tool_vel)
    nn = len(got.nodes)
    rp, ci, K = got.row_ptr, got.col_ind, got.K
    fixed = a["fixed_node"] == 1
    boundary = fixed | (a["cut_node"] == 1)
    touches_fixed = np.array([fixed[ci[rp[i]:rp[i + 1]]].any() for i in range(nn)])
    x, y = got.nodes[:, 0], got.nodes[:, 1]
    # Dirichlet rows: identity row, zero column, f = T0, M = 1
    for i in np.nonzero(fixed)[0]:
        row = K[rp[i]:rp[i + 1]]
        assert np.array_equal(row, (ci[rp[i]:rp[i + 1]] == i).astype(float))
        assert got.f[i] == tc["T0"] and got.M[i] == 1.0
    for i in range(nn):
        cols, row = ci[rp[i]:rp[i + 1]], K[rp[i]:rp[i + 1]]
        if fixed[i]:
            continue
        assert np.all(row[fixed[cols]] == 0.0)
        scale = np.max(np.abs(row))
        if not touches_fixed[i]:
            assert abs(row.sum()) <= 1e-12 * scale            # constants are in the kernel
            assert got.f[i] == 0.0
        else:
            # f_i = -T0 sum_{j fixed} K_ij (before the rows were cleared) = T0 sum_{j free} K_ij
            assert got.f[i] == pytest.approx(tc["T0"] * row.sum(), rel=1e-10, abs=1e-12 * scale * tc["T0"])
        if not boundary[i] and not touches_fixed[i]:
            for u in (x, y, 3 * x - 2 * y + 1):              # P1 reproduces linear fields
                lin = u - u[i]
                assert abs(row @ lin[cols]) <= 1e-11 * scale * np.max(np.abs(lin[cols]))
    # lumped mass = cp rho / 3 x the area of the triangles around a free node
    tri = a["elem"][:, :3]
    p = got.nodes[tri]
    area = 0.5 * np.abs((p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1])
                        - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1]))
    node_area = np.zeros(nn)
    np.add.at(node_area, tri.ravel(), np.repeat(area, 3))
    free = ~fixed
    assert np.allclose(got.M[free], tc["cp"] * tc["rho"] * node_area[free] / 3, rtol=1e-13, atol=0)
    m.close()


def test_abaqus_all_edges_flagged_correction(tmp_path):
    """Unit square split along 1-3; every node in the CuttingEdge node set and both triangles in
    the element set, so each triangle first gets all three edges flagged.  The correction clears the
    edge opposite a node no other set element shares: node 2 of triangle (1,2,3) -> edge 3-1;
    node 4 of triangle (1,3,4) -> edge 1-3."""
    text = "\n".join([
        "*Node", "1, 0.0, 0.0", "2, 100.0, 0.0", "3, 100.0, 100.0", "4, 0.0, 100.0",
        "*Element, type=CPS3", "1, 1, 2, 3", "2, 1, 3, 4",
        "*Nset, nset=CuttingEdge", "1, 2, 3, 4", "*Elset, elset=CuttingEdge", "1, 2",
        "*Nset, nset=FixedBoundary", "1, 4", "*Elset, elset=FixedBoundary", "2",
        "*End Part", ""])
    path = tmp_path / "square.inp"
    path.write_text(text)
    m = S.mesh_read(str(path), False)
    a = m.arrays()
    assert np.array_equal(a["position"], [[0, 0], [1, 0], [1, 1], [0, 1]])
    assert np.array_equal(a["cut_flags"], [[1, 1, 0], [0, 1, 1]])
    # fixed set: only triangle 2 with nodes 1 and 4 -> edge 4-1 (flag z)
    assert np.array_equal(a["fixed_flags"], [[0, 0, 0], [0, 0, 1]])
    r = R.read_abaqus(text)
    assert r["cut_flags"] == [[1, 1, 0], [0, 1, 1]] and r["fixed_flags"] == [[0, 0, 0], [0, 0, 1]]
    assert np.array_equal(a["elem"][:, 3], [1, 2])
    m.close()


def test_hot_tool_temperature_and_mesh_vtk(tmp_path, case):
    tm = case.tool
    tc = case.const.tool
    path = str(tmp_path / "tool.inp")
    cases.write_inp(tm, path, gmsh=True)
    m = S.mesh_read(path, True)
    vtk = str(tmp_path / "Tool_mesh.vtk")
    m.write_mesh_file(vtk)
    r = _ref_mesh(path, True)
    col, layers = R.colour_and_layers(r)
    assert open(vtk).read() == R.mesh_vtk(r, col, layers)
    T = 300.0 + np.arange(len(tm.nodes)) * 0.25
    m.set_temperature(T)
    got = m.tool_mesh(tc, tm.tl)
    assert np.array_equal(got.T, T)
    # the hot-tool restart: temperatures read back from a tool VTK file
    tool_vtk = tmp_path / "vtk_tool.vtk"
    tool_vtk.write_text("POINT_DATA %d\nSCALARS Temperature float 1\nLOOKUP_TABLE default\n" % len(T)
                        + "".join("%f\n" % v for v in T))
    m.mesh_read_temperature(str(tool_vtk))
    assert np.array_equal(m.tool_mesh(tc, tm.tl).T, T)
    with pytest.raises(_lib.SphfemError):
        m.set_temperature(T[:-1])
    m.set_temperature(None)
    assert np.all(m.tool_mesh(tc, tm.tl).T == tc["T0"])
    m.close()


def test_reader_errors(tmp_path, case):
    with pytest.raises(_lib.SphfemError):
        S.mesh_read(str(tmp_path / "missing.inp"), True)
    path = tmp_path / "tool.inp"
    cases.write_inp(case.tool, str(path), gmsh=True)
    lines = path.read_text().split("\n")
    k = next(i for i, l in enumerate(lines) if "type=CPS3" in l)
    bad = lines[:k + 2] + ["12345, 1"] + lines[k + 2:]
    (tmp_path / "bad.inp").write_text("\n".join(bad))
    with pytest.raises(_lib.SphfemError):
        S.mesh_read(str(tmp_path / "bad.inp"), True)
    bad = lines[:k + 2] + ["99999, 1, 2, 999999"] + lines[k + 2:]
    (tmp_path / "bad2.inp").write_text("\n".join(bad))
    with pytest.raises(_lib.SphfemError):
        S.mesh_read(str(tmp_path / "bad2.inp"), True)
    # a cutting-edge palette of one colour cannot colour neighbouring cutting-edge elements
    with pytest.raises(_lib.SphfemError):
        S.mesh_read(str(path), True, 18, 1)
