"""The C++ host path: driver/mfree_driver (the reference driver loop over the reference-named
C++ interface in include/sphfem/sphfem.hpp) against the Python host mirror on the same case.

Both drive the same library with the same inputs and the same call sequence, so the sorted
order is identical and the state agrees to the rounding of the order-dependent segment
atomics (friction heat), i.e. far inside 1e-11.
"""
import os
import subprocess

import numpy as np
import pytest

from sph_fem_2d_cutting_gpu_amd import _lib, case_io, cases
from sph_fem_2d_cutting_gpu_amd import solver as S

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sph_fem_2d_cutting_gpu_amd", "driver", "mfree_driver")


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_cpp_driver_matches_python_host(tmp_path, prec):
    if not os.path.exists(EXE):
        pytest.fail("driver not built (run __graft_entry__.build())")
    steps, freq = 30, 10
    case = cases.perturbed(cases.small_cutting_case(nx=48, ny=16, wear=True))
    cdir = case_io.write_case_dir(case, str(tmp_path / "case"))
    out = tmp_path / "out"
    out.mkdir()
    print(f"driver run: {prec}, {steps} steps, N = {case.N}", flush=True)
    r = subprocess.run([EXE, cdir, str(steps), str(freq), prec, str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = case_io.read_driver_output(str(out), case.N)

    sim = S.Simulation(case, precision=_lib.FP32C if prec == "fp32c" else _lib.FP64)
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
    assert np.array_equal(got["forces"][:, 0], F[:, 0])
    fscale = max(float(np.max(np.abs(F[:, 1:]))), 1e-300)
    assert float(np.max(np.abs(got["forces"][:, 1:] - F[:, 1:]))) / fscale <= 1e-11
