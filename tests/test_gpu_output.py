"""Output formats from device state: the VTK files of particles and of the FEM tool (mesh, shifted
mesh, segments, contact variables) equal, byte for byte, the reference's writers restated in
Python (oracle/vtk_text.py) over the downloaded arrays; the hot-tool temperature reader reads the
FEM tool file back; the Usui wear evaluation (a device reduction) matches the reference's host
loop."""
import math
import time

import numpy as np
import pytest

from oracle import vtk_text
from sph_fem_2d_cutting_gpu_amd import _lib, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu


def _first_diff(got, exp):
    g, e = got.splitlines(), exp.splitlines()
    for i in range(min(len(g), len(e))):
        if g[i] != e[i]:
            return f"line {i}: {g[i]!r} != {e[i]!r}"
    return f"lengths {len(g)} != {len(e)}"


@pytest.fixture(scope="module")
def fem_sim():
    case = cases.small_cutting_case(nx=48, ny=16, wear=True)
    sim = S.Simulation(case)
    sim.step(60)
    sim.ctx.check_error()
    yield sim
    sim.close()


def test_particle_vtk_from_device(tmp_path, fem_sim):
    path = S.vtk_writer_write(fem_sim.p, 60, folder=str(tmp_path))
    assert path.endswith("vtk_out_000060.vtk")
    P = fem_sim.state()
    got = open(path).read()
    exp = vtk_text.particles(P)
    assert got == exp, _first_diff(got, exp)
    assert np.any(P["fc"] != 0) and np.any(P["ft"] != 0) and np.ptp(P["T"]) > 0


def test_fem_tool_vtk_files(tmp_path, fem_sim):
    t = fem_sim.tool
    st = t.download()
    kin = t.kinematics()
    pos, elem, T = st["node_pos"], fem_sim.case.tool.elems, st["T"]
    p1 = S.vtk_writer_write_FEM_tool(t, 60, folder=str(tmp_path))
    assert open(p1).read() == vtk_text.fem_tool(pos, elem, T)
    tl_ref = (kin["tl_x"] - 3e-4, kin["tl_y"] + 1e-4)
    p2 = S.vtk_writer_write_tool_org(t, 60, tl_ref, folder=str(tmp_path))
    shift = (kin["tl_x"] - tl_ref[0], kin["tl_y"] - tl_ref[1])
    assert open(p2).read() == vtk_text.fem_tool(pos, elem, T, shift)
    pp, sl, wn = st["physical_para"], st["sliding_force"], st["wear_nodes"]
    assert np.any(pp[:, 0] != 0)
    p3 = S.vtk_writer_write_FEM_segments(t, 60, folder=str(tmp_path))
    got = open(p3).read()
    exp = vtk_text.fem_segments(st["left"], st["right"], st["n"], pp, sl, wn)
    assert got == exp, _first_diff(got, exp)
    p4 = S.vtk_writer_write_contact_variables(t, 60, folder=str(tmp_path))
    got = open(p4).read()
    exp = vtk_text.contact_variables(st["left"], st["right"], pp, sl, wn)
    assert got == exp, _first_diff(got, exp)
    # the hot-tool restart reader takes the temperatures back (to the six decimals written)
    T_read = S.read_vtk_temperature(p1)
    assert np.array_equal(T_read, np.array([float("%f" % x) for x in T]))


def test_usui_wear_matches_reference_loop():
    case = cases.ref_cut_sph()
    sim = S.Simulation(case)
    sim.step(400)
    K, alpha = 1. / 100 * 7.8, 2500.               # benchmarks.cpp:503-507
    w = S.tool_wear(sim.ctx, K, alpha, case.tool_vel)
    dt_dump = 100 * case.dt
    mn, mx, avg = w.eval_usui(sim.p, dt_dump)
    P = sim.state()
    emn, emx, eavg, cnt = vtk_text.usui(P, K, alpha, case.tool_vel)
    print("usui", mn, mx, avg, "ref", emn, emx, eavg, cnt)
    assert cnt > 3
    assert mn == pytest.approx(emn, rel=1e-14) and mx == pytest.approx(emx, rel=1e-14)
    assert avg == pytest.approx(eavg, rel=1e-13)
    assert w.get_accum_wear() == pytest.approx(dt_dump * eavg, rel=1e-13)
    sim.step(100)
    mn2, mx2, avg2 = w.eval_usui(sim.p, dt_dump)
    assert w.get_accum_wear() == pytest.approx(dt_dump * (avg + avg2), rel=1e-13)
    w.destroy()
    sim.close()


def test_particle_vtk_4m_time(tmp_path):
    """Writing the 4M-particle state: text equal to the reference format on a sample of lines,
    and the time it takes (printed)."""
    case = cases.config3()
    sim = S.Simulation(case, precision=_lib.FP32C)
    sim.step(2)
    sim.ctx.synchronize()
    t0 = time.perf_counter()
    path = S.vtk_writer_write(sim.p, 2, folder=str(tmp_path))
    el = time.perf_counter() - t0
    size = sum(1 for _ in open(path))
    print(f"vtk_out of {case.N} particles: {el:.2f} s, {size} lines")
    assert size == 32 + 11 * case.N
    lines = open(path).read().splitlines()
    P = sim.state(("pos", "T"))
    idx = np.random.default_rng(1).integers(0, case.N, 1000)
    for i in idx:
        assert lines[6 + i] == "%f %f %f" % (P["pos"][i][0], P["pos"][i][1], 0.0)
    sim.close()
