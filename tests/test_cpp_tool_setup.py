"""The reference's tool_FEM set-up sequence through the C++ binding (benchmarks_NZ.cpp:228-245:
tool_FEM(tl, length, height, rake, clearance, r, mu); mesh_CPU = new mesh_read(file); mesh_GPU = new
mesh_GPU_m(mesh_CPU->nodes, mesh_CPU->elements, ...); set_up_segments_from_mesh()), host part only: a
program compiled with g++ against include/sphfem/sphfem.hpp and the library reads a generated tool
mesh and must report the counts and extent (tl_point_ref, br_point, ref_dist, lowest_point,
right_point, tool_FEM.cu:842-873, 1022-1037) of the Python set-up of the same mesh, and the host
copies of mesh_read (nodes.x_init, elements.node_num) must hold the file's mesh.  A mesh_GPU made from
another mesh is refused.  No GPU: the device tool is made by set_up_FEM_tool_gpu, which is not called."""
import os
import subprocess
import textwrap

import numpy as np
import pytest

from sph_fem_2d_cutting_gpu_amd import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sph_fem_2d_cutting_gpu_amd")

PROGRAM = textwrap.dedent(r'''
    #include <cstdio>
    #include <cstdlib>
    #include "sphfem/sphfem.hpp"

    int main(int argc, char **argv) {
        const vec2_t tl(std::atof(argv[2]), std::atof(argv[3]));
        tool_FEM *t = new tool_FEM(tl, std::atof(argv[4]), std::atof(argv[5]), std::atof(argv[6]),
                                   std::atof(argv[7]), std::atof(argv[8]), std::atof(argv[9]));
        t->mesh_CPU = new mesh_read(argv[1], true);
        t->mesh_GPU = new mesh_GPU_m(t->mesh_CPU->nodes, t->mesh_CPU->elements, t->mesh_CPU->color_total,
                                     t->mesh_CPU->color_cutting_edge, 1);
        t->set_up_segments_from_mesh();
        std::printf("counts %d %d %d\n", t->node_number, t->element_number, t->cutting_segment_size);
        std::printf("extent %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", t->tl_point_ref.x, t->tl_point_ref.y,
                    t->br_point.x, t->br_point.y, t->ref_dist, t->lowest_point, t->right_point);
        const nodes_read &nd = t->mesh_CPU->nodes;
        double sx = 0., sy = 0.;
        for (size_t i = 0; i < nd.x_init.size(); i++) { sx += nd.x_init[i]; sy += nd.y_init[i]; }
        long se = 0;
        for (const auto &e : t->mesh_CPU->elements.node_num) se += e.x + e.y + e.z;
        std::printf("host %zu %.17g %.17g %ld\n", nd.x_init.size(), sx, sy, se);
        // a mesh_GPU made from another mesh than mesh_CPU
        mesh_read other(argv[1], true);
        t->mesh_GPU = new mesh_GPU_m(other.nodes, other.elements, other.color_total, other.color_cutting_edge, 1);
        try {
            t->set_up_segments_from_mesh();
            std::printf("mismatch accepted\n");
        } catch (const std::exception &ex) {
            std::printf("mismatch refused: %s\n", ex.what());
        }
        return 0;
    }
''')


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    lib = os.path.join(PKG, "libsphfem_hip.so")
    if not os.path.exists(lib):
        pytest.fail("library not built (run __graft_entry__.build())")
    d = tmp_path_factory.mktemp("tool_setup")
    src, exe = d / "p.cpp", d / "p"
    src.write_text(PROGRAM)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", PKG, "-lsphfem_hip", "-L/opt/rocm/lib", f"-Wl,-rpath,{PKG}",
                        "-Wl,-rpath,/opt/rocm/lib"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def test_set_up_segments_from_mesh_matches_the_python_set_up(program, tmp_path):
    case = cases.small_cutting_case(nx=48, ny=16)
    tm = case.tool
    inp = str(tmp_path / "tool.inp")
    cases.write_inp(tm, inp, gmsh=True)
    length, height, rake, clear, r = tm.outline
    args = [program, inp, repr(float(tm.tl[0])), repr(float(tm.tl[1])), repr(length), repr(height), repr(rake),
            repr(clear), repr(r), "0.35"]
    out = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = {l.split()[0]: l.split()[1:] for l in out.stdout.splitlines() if l and l.split()[0] != "mismatch"}
    assert [int(v) for v in lines["counts"]] == [len(tm.nodes), len(tm.elems), len(tm.seg_left)]
    got = np.array([float(v) for v in lines["extent"]])
    want = np.array([tm.tl[0], tm.tl[1], tm.br[0], tm.br[1], tm.ref_dist, tm.lowest_point, tm.right_point])
    np.testing.assert_allclose(got, want, rtol=1e-15, atol=0)
    n, sx, sy, se = lines["host"]
    assert int(n) == len(tm.nodes)
    np.testing.assert_allclose([float(sx), float(sy)], tm.nodes.sum(axis=0), rtol=1e-14)
    assert int(se) == int(tm.elems[:, :3].sum())
    assert "mismatch refused" in out.stdout
