"""A tool read from an .inp file, end to end on the device (benchmarks_NZ.cpp:228-245: mesh_read,
set_up_segments_from_mesh, set_up_FEM_tool_gpu, then the step):

* the step with the tool set up from the mesh file equals the oracle's step with the same tool,
  to 1e-13 after one step and 1e-6 after 25 (the artificial stress's single-precision libm calls),
  for both .inp dialects;
* the C++ driver, reading the file through mesh_read and the tool_FEM(mesh, ...) constructor of
  the C++ host mirror, matches the Python host path on the same file.
"""
import dataclasses
import os
import subprocess

import numpy as np
import pytest

from helpers import assert_close, assert_equal, max_rel_err
from oracle.oracle import Oracle
from sph_fem_2d_cutting_gpu_amd import case_io, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sph_fem_2d_cutting_gpu_amd", "driver", "mfree_driver")

STATE = ("pos", "vel", "S", "rho", "T", "eps_pl", "eps_pl_dot", "fc", "ft", "n", "p", "R")
DERIVS = ("pos_t", "vel_t", "rho_t", "S_t", "T_t")


def _inp_case(tmp_path, gmsh):
    case = cases.small_cutting_case(nx=48, ny=16, wear=True)
    path = str(tmp_path / "tool.inp")
    cases.write_inp(case.tool, path, gmsh=gmsh)
    m = S.mesh_read(path, gmsh)
    tm = m.tool_mesh(case.const.tool, case.tool.tl, case.tool_vel)
    m.close()
    return dataclasses.replace(case, tool=tm), path


@pytest.mark.parametrize("gmsh,nsteps", [(True, 1), (True, 25), (False, 1), (False, 25)])
def test_inp_tool_step_matches_oracle(tmp_path, gmsh, nsteps):
    case, _ = _inp_case(tmp_path, gmsh)
    print(f"inp tool ({'gmsh' if gmsh else 'abaqus'}): {len(case.tool.nodes)} nodes, "
          f"{len(case.tool.seg_left)} segments, {nsteps} steps", flush=True)
    o = Oracle(case)
    sim = S.Simulation(case)
    o.step(nsteps)
    sim.step(nsteps)
    sim.ctx.check_error()
    a, b = sim.state(), o.P
    for f in ("hash", "idx", "blanked"):
        assert_equal(f, a[f], b[f])
    live = b["blanked"] != 1.0
    tol = 1e-13 if nsteps == 1 else 1e-6
    errs = {f: max_rel_err(a[f], b[f], live) for f in STATE + DERIVS}
    assert max(errs.values()) <= tol, errs
    tg = sim.tool_state()
    for f in ("T", "node_pos", "left", "right", "n", "K", "M"):
        assert_close("tool " + f, tg[f], o.t[f], tol)
    if nsteps > 1:
        assert np.ptp(tg["T"]) > 0        # the tool's heat solve has run on the mesh read
    sim.close()


def test_cpp_driver_reads_inp(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("driver not built (run __graft_entry__.build())")
    steps, freq = 30, 10
    base = cases.small_cutting_case(nx=48, ny=16, wear=True)
    cdir = case_io.write_case_dir(base, str(tmp_path / "case"), tool_inp="gmsh")
    case, _ = _inp_case(tmp_path, True)
    out = tmp_path / "out"
    out.mkdir()
    print(f"driver run with tool.inp: {steps} steps, N = {case.N}", flush=True)
    r = subprocess.run([EXE, cdir, str(steps), str(freq), "fp64", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "tool mesh" in r.stdout and "tool.inp" in r.stdout
    # the driver set the tool up as benchmarks_NZ.cpp:228-245 does: tool_FEM(tl, length, height, rake,
    # clearance, r, mu), mesh_CPU / mesh_GPU, set_up_segments_from_mesh(), set_up_FEM_tool_gpu(tc)
    assert "tool_FEM set-up sequence" in r.stdout, r.stdout
    got = case_io.read_driver_output(str(out), case.N)
    sim = S.Simulation(case)
    forces = []
    for step in range(steps):
        if step % freq == 0:
            forces.append([step, *S.report_force(sim.p)])
        sim.step(1)
    forces.append([steps, *S.report_force(sim.p)])
    sim.ctx.check_error()
    ref = sim.state(("pos", "vel", "S", "T", "idx"))
    assert np.array_equal(got["idx"], ref["idx"])
    for f in ("pos", "vel", "S", "T"):
        scale = max(float(np.max(np.abs(ref[f]))), 1e-300)
        err = float(np.max(np.abs(got[f] - ref[f]))) / scale
        assert err <= 1e-11, f"{f}: {err:.3e}"
    F = np.array(forces)
    fscale = max(float(np.max(np.abs(F[:, 1:]))), 1e-300)
    assert float(np.max(np.abs(got["forces"][:, 1:] - F[:, 1:]))) / fscale <= 1e-11
    # the driver read the tool temperatures through the FEM_solver / mesh_GPU / segments_FEM views
    line = next(l for l in r.stdout.splitlines() if l.startswith("tool nodes"))
    _, _, nodes, _, segs, _, _, tmax = line.replace(",", "").split()
    tg = sim.tool_state()
    assert int(nodes) == sim.tool.n_nodes and int(segs) == sim.tool.n_seg
    assert float(tmax) == pytest.approx(float(tg["T"].max()), rel=1e-12)
    sim.close()
