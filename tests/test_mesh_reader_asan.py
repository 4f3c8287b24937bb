"""The tool-mesh reader (sph_fem_2d_cutting_gpu_amd/csrc/sphfem_mesh.cpp: .inp parsing, colouring, CSR,
segments and the FEM system of set_up_segments_from_mesh) on damaged files, built for the host with
AddressSanitizer and UndefinedBehaviorSanitizer: seeded mutations of a generated mesh in both dialects
(lines deleted, duplicated or truncated, ids and coordinates replaced by out-of-range, negative,
non-finite or garbage tokens, the file cut at a random byte).  Every file must be either refused with a
status code or read into a mesh whose arrays, tool description and VTK text can be produced -- with no
sanitizer report.  The reader's input is a user file, so it must fail safely on anything."""
import os
import subprocess
import textwrap

import numpy as np
import pytest

from sph_fem_2d_cutting_gpu_amd import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = textwrap.dedent(r'''
    #include <cstdio>
    #include <cstring>
    #include <vector>
    #include "sphfem.h"

    int main(int argc, char **argv) {
        const char *vtk = argv[1];
        int read = 0, refused = 0, described = 0;
        for (int a = 2; a < argc; a++) {
            const int gmsh = std::strstr(argv[a], "gmsh") != nullptr;
            sphfem_mesh *m = nullptr;
            if (sphfem_mesh_read_inp(argv[a], gmsh, 18, 6, &m) != SPHFEM_OK) { refused++; continue; }
            read++;
            sphfem_mesh_info in{};
            if (sphfem_mesh_get_info(m, &in) != SPHFEM_OK) return 3;
            std::vector<sphfem_double2> pos(in.node_num);
            std::vector<int> cut(in.node_num), fix(in.node_num), rp(in.node_num + 1), ci(in.nz_num);
            std::vector<sphfem_int4> el(in.ele_num), cf(in.ele_num), ff(in.ele_num);
            std::vector<int> ce(in.cut_ele_num), fe(in.fixed_ele_num), lay(2 * (size_t)in.ele_num);
            if (sphfem_mesh_get_arrays(m, pos.data(), cut.data(), fix.data(), el.data(), cf.data(), ff.data(),
                                       ce.data(), fe.data(), lay.data(), rp.data(), ci.data()) != SPHFEM_OK)
                return 4;
            sphfem_tool_constants tc{};
            tc.cp = 1.; tc.rho = 1.; tc.k = 1.; tc.T0 = 300.;
            sphfem_tool_desc d{};
            if (sphfem_mesh_tool_desc(m, &tc, sphfem_double2{0., 0.}, sphfem_double2{1., 0.}, &d) == SPHFEM_OK)
                described++;
            sphfem_mesh_write_vtk(m, vtk);
            sphfem_mesh_destroy(m);
        }
        std::printf("read %d refused %d described %d\n", read, refused, described);
        return 0;
    }
''')

GARBAGE = ["", "-1", "0", "2147483647", "-2147483648", "99999999999", "nan", "inf", "-inf", "1e308", "abc", ",",
           "1.5", "*NODE", "*ELEMENT, type=CPS3", "*NSET, NSET=CuttingEdge", "*ELSET, ELSET=FixedBoundary"]


def _mutate(text, rng):
    lines = text.split("\n")
    kind = rng.integers(0, 6)
    i = int(rng.integers(0, len(lines)))
    if kind == 0:                                   # delete a line
        del lines[i]
    elif kind == 1:                                 # duplicate a line
        lines.insert(i, lines[i])
    elif kind == 2:                                 # replace a token
        toks = lines[i].split(",")
        if toks:
            toks[int(rng.integers(0, len(toks)))] = " " + GARBAGE[int(rng.integers(0, len(GARBAGE)))]
        lines[i] = ",".join(toks)
    elif kind == 3:                                 # an id out of range (large or negative)
        toks = lines[i].split(",")
        toks[int(rng.integers(0, len(toks)))] = " " + str(int(rng.choice([-7, 0, 10 ** 6, len(lines) + 3])))
        lines[i] = ",".join(toks)
    elif kind == 4:                                 # a line of garbage tokens
        lines.insert(i, ", ".join(GARBAGE[int(k)] for k in rng.integers(0, len(GARBAGE), 4)))
    else:                                           # cut the file
        return text[: int(rng.integers(0, len(text)))]
    return "\n".join(lines)


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    d = tmp_path_factory.mktemp("mesh_asan")
    src, exe = d / "h.cpp", d / "h"
    src.write_text(HARNESS)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "sph_fem_2d_cutting_gpu_amd", "csrc", "sphfem_mesh.cpp"), str(src),
                        "-o", str(exe)], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + r.stderr[-3000:])
    return str(exe)


@pytest.mark.parametrize("gmsh", [True, False])
def test_damaged_mesh_files_fail_safely(harness, tmp_path, gmsh):
    tm = cases.small_cutting_case(nx=48, ny=16).tool
    base = tmp_path / ("base_gmsh.inp" if gmsh else "base_abaqus.inp")
    cases.write_inp(tm, str(base), gmsh=gmsh)
    text = base.read_text()
    rng = np.random.default_rng(20261203 + int(gmsh))
    files = [str(base)]
    for k in range(160):
        t = text
        for _ in range(int(rng.integers(1, 4))):
            t = _mutate(t, rng)
        f = tmp_path / f"m{k:03d}_{'gmsh' if gmsh else 'abaqus'}.inp"
        f.write_text(t)
        files.append(str(f))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([harness, str(tmp_path / "out.vtk")] + files, capture_output=True, text=True, timeout=600,
                       env=env)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    read, refused, described = (int(v) for v in r.stdout.split()[1::2])
    assert read + refused == len(files) and read >= 1 and described >= 1
