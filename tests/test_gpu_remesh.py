"""Tool remeshing on the device (actions_gpu.cu:1022-1038) against the oracle:

* remesh_low_level (mesh_shift, colour-ordered reassembly with the Dirichlet rows, interpolation
  from the mapping) on the same tool state: positions within a few ulp (the profile's pow calls),
  K, M, f and T to 1e-12; the steps after it to 1e-6;
* remesh onto a new mesh read from an .inp file (read_remesh_gmsh): onto the current mesh
  (identity, no wear) the temperatures are kept exactly and the run continues as without the
  remesh, to the rounding of the reassembled system; onto a
  finer mesh the interpolated temperatures equal the oracle's bit for bit and the steps after it
  match the oracle to 1e-6;
* a graph captured before a remesh is refused afterwards; capturing again works;
* the .geo script of remesh_read_segments equals its Python restatement.
"""
import dataclasses

import numpy as np
import pytest

from helpers import max_rel_err
from oracle import mesh_inp as R
from oracle.oracle import Oracle
from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "S", "rho", "T", "eps_pl", "eps_pl_dot", "p")


def _case():
    return cases.small_cutting_case(nx=48, ny=16, wear=True)


def _linear(p):
    """a linear temperature field above 300 K over the tool (x > 0, y > 0.05 cm)"""
    return 300.0 + 1e3 * p[:, 0] + 5e2 * (p[:, 1] - 0.05)


def _row_rel(a, b, tm_rp):
    row = np.repeat(np.arange(len(tm_rp) - 1), np.diff(tm_rp))
    scale = np.zeros(len(tm_rp) - 1)
    np.maximum.at(scale, row, np.abs(b))
    return float(np.max(np.abs(a - b) / np.maximum(scale[row], 1e-300)))


def _sync_tool(sim, o):
    """the oracle's tool state and kinematics onto the device tool"""
    t = o.t
    sim.tool.upload(dict(node_pos=t["node_pos"], T=t["T"], K=t["K"], M=t["M"], k=t["k"], left=t["left"],
                         right=t["right"], n=t["n"], scalars=o.tool_scalars))


def _compare_particles(sim, o, tol, label):
    a, b = sim.state(), o.P
    assert np.array_equal(a["idx"], b["idx"]), label
    live = b["blanked"] != 1.0
    errs = {f: max_rel_err(a[f], b[f], live) for f in STATE}
    assert max(errs.values()) <= tol, (label, errs)


def test_remesh_low_level_matches_oracle():
    case = _case()
    sim, o = S.Simulation(case), Oracle(case)
    sim.step(1)
    o.step(1)
    o.t["T"][:] = _linear(o.t["node_pos"])
    o.Tc.flank_retreat = -1e-6          # 40 x -1e-6 beyond the -6.5e-6 cap
    _sync_tool(sim, o)
    print("remesh_low_level after 1 step, freq 40", flush=True)
    sim.tool.remesh_low_level(40)
    o.remesh_low_level(40)
    sim.ctx.check_error()
    g = sim.tool_state()
    rp = case.tool.row_ptr
    moved = np.abs(o.t["node_pos"] - case.tool.nodes).max()
    assert moved > 1e-6
    assert np.max(np.abs(g["node_pos"] - o.t["node_pos"])) <= 1e-16
    assert _row_rel(g["K"], o.t["K"], rp) <= 1e-12
    assert np.allclose(g["M"], o.t["M"], rtol=1e-12, atol=0)
    assert np.allclose(g["T"], o.t["T"], rtol=1e-12, atol=0)
    lin = _linear(o.t["node_pos"])
    assert np.allclose(g["T"], lin, rtol=0, atol=1e-10 * np.ptp(lin))
    sim.step(5)
    o.step(5)
    sim.ctx.check_error()
    _compare_particles(sim, o, 1e-6, "5 steps after remesh_low_level")
    g = sim.tool_state()
    assert np.allclose(g["T"], o.t["T"], rtol=1e-6, atol=0)
    sim.close()


def _mesh_from(tm, pos, path):
    cases.write_inp(dataclasses.replace(tm, nodes=np.ascontiguousarray(pos)), path, gmsh=True)
    return S.mesh_read(path, True)


def test_remesh_identity_keeps_the_run(tmp_path):
    # without wear the nodes only translate, so the lowest and right points the remesh takes from
    # the new segments are the ones the kinematics carried (with wear they follow the worn edge).
    # The new segments start with zero heat exchange and the new system with a zero temperature
    # rate (tool_FEM.cu:1496, FEM_thermal_gpu.cu:782), which the twin is given at the same step.
    case = cases.small_cutting_case(nx=48, ny=16, wear=False)
    sim, twin = S.Simulation(case), S.Simulation(case)
    sim.step(6)
    twin.step(6)
    g = sim.tool_state()
    m = _mesh_from(case.tool, g["node_pos"], str(tmp_path / "same.inp"))
    n_seg = sim.tool.n_seg
    print("identity remesh after 6 steps", flush=True)
    sim.tool.remesh(m)
    sim.ctx.check_error()
    twin.tool.upload(dict(heat_exchange=np.zeros((n_seg, 2)), fric_heat=np.zeros(n_seg),
                          k=np.zeros(sim.tool.n_nodes)))
    h = sim.tool_state()
    assert sim.tool.n_seg == n_seg and np.array_equal(h["T"], g["T"])
    assert np.array_equal(h["node_pos"], g["node_pos"])
    assert _row_rel(h["K"], g["K"], case.tool.row_ptr) <= 1e-12
    sim.step(10)
    twin.step(10)
    a, b = sim.state(), twin.state()
    assert np.array_equal(a["idx"], b["idx"])
    for f in STATE:
        assert max_rel_err(a[f], b[f], b["blanked"] != 1.0) <= 1e-10, f
    assert np.allclose(sim.tool_state()["T"], twin.tool_state()["T"], rtol=1e-12, atol=0)
    m.close()
    sim.close()
    twin.close()


def test_remesh_onto_finer_mesh_matches_oracle(tmp_path):
    case = _case()
    sim, o = S.Simulation(case), Oracle(case)
    sim.step(3)
    o.step(3)
    o.t["T"][:] = _linear(o.t["node_pos"]) + 2.0 * np.sin(1e3 * o.t["node_pos"][:, 0])
    _sync_tool(sim, o)
    sc = o.tool_scalars
    L = 2.5 * 1.6e-3
    fine = cases.mesh_tool(L, L, 0.0, 7.0, 4e-4, 1200, case.const.tool)
    m = _mesh_from(fine, fine.nodes + (np.array([sc["tl_x"], sc["tl_y"]]) - fine.tl), str(tmp_path / "fine.inp"))
    new = m.tool_mesh(case.const.tool, (sc["tl_x"], sc["tl_y"]), (sc["vel_x"], sc["vel_y"]))
    print(f"remesh onto {len(new.nodes)} nodes (was {len(case.tool.nodes)})", flush=True)
    sim.tool.remesh(m)
    o.remesh(new)
    sim.ctx.check_error()
    g = sim.tool_state()
    assert sim.tool.n_nodes == len(new.nodes) and sim.tool.n_seg == len(new.seg_left)
    assert np.array_equal(g["T"], o.t["T"])
    for f in ("node_pos", "K", "M", "left", "right", "n"):
        assert np.array_equal(g[f], o.t[f]), f
    k = sim.tool.kinematics()
    for f in ("tl_x", "tl_y", "br_x", "br_y", "vel_x", "flank_retreat"):
        assert k[f] == sc[f], f
    assert k["lowest_point"] == new.lowest_point and k["ref_dist"] == new.ref_dist
    sim.step(5)
    o.step(5)
    sim.ctx.check_error()
    _compare_particles(sim, o, 1e-6, "5 steps after the remesh")
    assert np.allclose(sim.tool_state()["T"], o.t["T"], rtol=1e-6, atol=0)
    m.close()
    sim.close()


def test_graph_refused_after_remesh(tmp_path):
    case = _case()
    sim = S.Simulation(case)
    sim.step(2)
    gr = sim.capture(2)
    gr.launch(1)
    m = _mesh_from(case.tool, sim.tool_state()["node_pos"], str(tmp_path / "same.inp"))
    sim.tool.remesh(m)
    with pytest.raises(_lib.SphfemError, match="remeshed"):
        gr.launch(1)
    gr.destroy()
    gr2 = sim.capture(2)
    gr2.launch(2)
    sim.ctx.check_error()
    gr2.destroy()
    # remesh_low_level keeps the arrays: a graph stays valid across it
    gr3 = sim.capture(2)
    sim.tool.remesh_low_level(20)
    gr3.launch(1)
    sim.ctx.check_error()
    gr3.destroy()
    assert np.all(np.isfinite(sim.state(("pos",))["pos"]))
    m.close()
    sim.close()


def test_remesh_geo_file(tmp_path):
    case = _case()
    sim = S.Simulation(case)
    sim.step(4)
    path = str(tmp_path / "tool_remesh.geo")
    sim.tool.write_remesh_geo(path)
    g = sim.tool_state()
    k = sim.tool.kinematics()
    assert open(path).read() == R.remesh_geo(g["left"], g["right"], (k["tl_x"], k["tl_y"]))
    sim.close()


def test_cpp_driver_remesh_low_level(tmp_path):
    """the driver loop of mfree_gpu.cpp:351-357 (remesh_low_level every freq / 20 steps, with wear)
    in the C++ driver and in the Python host, on the same case"""
    import os
    import subprocess
    from sph_fem_2d_cutting_gpu_amd import case_io
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sph_fem_2d_cutting_gpu_amd",
                       "driver", "mfree_driver")
    if not os.path.exists(exe):
        pytest.fail("driver not built (run __graft_entry__.build())")
    case = _case()
    steps, freq = 30, 40
    cdir = case_io.write_case_dir(case, str(tmp_path / "case"), remesh_low_level=True)
    out = tmp_path / "out"
    out.mkdir()
    print(f"driver with remesh_low_level every {freq // 20} steps", flush=True)
    r = subprocess.run([exe, cdir, str(steps), str(freq), "fp64", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = case_io.read_driver_output(str(out), case.N)
    sim = S.Simulation(case)
    for step in range(steps):
        if step % (freq // 20) == 0:
            sim.tool.remesh_low_level(freq)
        sim.step(1)
    sim.ctx.check_error()
    ref = sim.state(("pos", "vel", "S", "T", "idx"))
    assert np.array_equal(got["idx"], ref["idx"])
    for f in ("pos", "vel", "S", "T"):
        scale = max(float(np.max(np.abs(ref[f]))), 1e-300)
        assert float(np.max(np.abs(got[f] - ref[f]))) / scale <= 1e-11, f
    sim.close()


def test_tool_array_views_follow_a_remesh(tmp_path):
    """sphfem_tool_get_arrays (the mesh_GPU_m / FEM_thermal_2D_GPU / segment_FEM_gpu views): the
    pointers read the tool's data, before and after a remesh onto a finer mesh"""
    case = _case()
    sim = S.Simulation(case)
    sim.step(2)
    for stage in ("created", "remeshed"):
        a = sim.tool.arrays()
        g = sim.tool_state()
        nn, ns, nz = a["node_num"], a["seg_num"], a["nz_num"]
        assert (nn, ns, nz) == (sim.tool.n_nodes, sim.tool.n_seg, sim.tool.nnz)
        assert a["fixed_ele_num"] > 0 and a["node_flags"] and a["T_mapping"]
        assert np.array_equal(sim.ctx.dtoh(a["T"], (nn,), np.float64), g["T"])
        assert np.array_equal(sim.ctx.dtoh(a["K"], (nz,), np.float64), g["K"])
        assert np.array_equal(sim.ctx.dtoh(a["position"], (nn, 2), np.float64), g["node_pos"])
        assert np.array_equal(sim.ctx.dtoh(a["left"], (ns, 2), np.float64), g["left"])
        if stage == "created":
            L = 2.5 * 1.6e-3
            k = sim.tool.kinematics()
            fine = cases.mesh_tool(L, L, 0.0, 7.0, 4e-4, 1200, case.const.tool)
            m = _mesh_from(fine, fine.nodes + (np.array([k["tl_x"], k["tl_y"]]) - fine.tl), str(tmp_path / "f.inp"))
            sim.tool.remesh(m)
            m.close()
    sim.close()


def test_remesh_needs_the_remeshing_description(tmp_path):
    """a tool created without node flags and fixed boundary refuses both remesh entry points with
    SPHFEM_ERR_STATE and keeps working"""
    case = _case()
    ctx = S.Context()
    ctx.setup_constants(case.const, case.dt, cases.default_rr2_min(case.dx), case.wear)
    t = S.tool_FEM(ctx, case, remeshable=False)
    assert t.arrays()["fixed_ele_num"] == 0 and not t.arrays()["node_flags"]
    with pytest.raises(_lib.SphfemError, match="remeshing needs"):
        t.remesh_low_level(20)
    m = _mesh_from(case.tool, case.tool.nodes, str(tmp_path / "same.inp"))
    with pytest.raises(_lib.SphfemError, match="remeshing needs"):
        t.remesh(m)
    g = t.download()
    assert np.array_equal(g["node_pos"], case.tool.nodes)
    m.close()
    ctx.call("sphfem_tool_destroy", t.h)
    ctx.close()


def test_long_run_with_remesh_cadence_matches_oracle():
    """300 steps with wear and remesh_low_level every 20 steps (the driver's freq / 20 with
    freq = 400), the tool one lattice spacing into the workpiece, on the device and on the oracle:
    the tool's temperature field and the particles' peak temperature agree within 1e-3 of the
    temperature rise, the tool's nodes within 1e-9 cm (the single-precision libm calls of the
    artificial stress make the two runs drift apart slowly)"""
    case = cases.small_cutting_case(nx=48, ny=16, wear=True, engage=1.0)
    sim, o = S.Simulation(case), Oracle(case)
    freq = 400
    print("300 steps, remesh every 20", flush=True)
    for step in range(300):
        if step % (freq // 20) == 0:
            sim.tool.remesh_low_level(freq)
            o.remesh_low_level(freq)
        sim.step(1)
        o.step(1)
    sim.ctx.check_error()
    g = sim.tool_state()
    T0 = case.const.tool["T0"]
    rise = o.t["T"].max() - T0
    print("tool T rise", g["T"].max() - T0, rise)
    assert rise > 10.0
    assert np.max(np.abs(g["T"] - o.t["T"])) <= 1e-3 * rise
    assert np.max(np.abs(g["node_pos"] - o.t["node_pos"])) <= 1e-9
    a = sim.state(("T", "blanked"))
    assert np.array_equal(a["blanked"] == 1.0, o.P["blanked"] == 1.0)
    assert abs(a["T"].max() - o.P["T"].max()) <= 1e-3 * (o.P["T"].max() - 300.0)
    sim.close()


def test_remesh_geo_with_a_short_segment(tmp_path):
    """the library's remesh geometry on a chain with one segment shortened to 5 % (so the point
    before it is skipped, tool_FEM.cu:1743-1747) equals the restatement"""
    case = _case()
    sim = S.Simulation(case)
    g = sim.tool_state()
    left, right = g["left"].copy(), g["right"].copy()
    k = len(left) // 2
    right[k] = left[k] + 0.05 * (right[k] - left[k])
    left[k + 1] = right[k]
    sim.tool.upload(dict(left=left, right=right))
    path = str(tmp_path / "short.geo")
    sim.tool.write_remesh_geo(path)
    kin = sim.tool.kinematics()
    text = open(path).read()
    assert text == R.remesh_geo(left, right, (kin["tl_x"], kin["tl_y"]))
    assert text.count("Point(") == len(left) + 1     # tl, the chain less one point, the end
    sim.close()


@pytest.mark.parametrize("which", ["low_level", "new_mesh"])
def test_remesh_refuses_a_non_finite_node(which):
    """A tool node at a non-finite position: both remeshes are refused with SPHFEM_ERR_RANGE before the
    element bins are built (binning would cast the coordinate to an integer cell index)."""
    case = _case()
    sim = S.Simulation(case)
    sim.step(1)
    sim.ctx.check_error()
    pos = sim.tool.download()["node_pos"].copy()
    pos[len(pos) // 2] = [np.nan, pos[len(pos) // 2][1]]
    sim.tool.upload(dict(node_pos=pos))
    with pytest.raises(_lib.SphfemError, match=f"status {_lib.ERR_RANGE}.*not finite"):
        if which == "low_level":
            sim.tool.remesh_low_level(20)
        else:
            path = f"/tmp/sphfem_remesh_nonfinite_{id(sim)}.inp"
            cases.write_inp(case.tool, path, gmsh=True)
            sim.tool.remesh(S.mesh_read(path, True))
    sim.close()
