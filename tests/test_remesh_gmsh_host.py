"""remesh_gmsh of the C++ binding (include/sphfem/sphfem.hpp; actions_gpu.cu:1033-1038): the geometry
is written, the caller's meshing step runs, and only a mesh file newer than that geometry is read; a
missing or stale mesh is refused instead of silently reloading an older mesh.  Host logic only: the
program is compiled with g++ against the header with a stand-in tool type (no GPU, no library)."""
import os
import subprocess
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = textwrap.dedent(r'''
    #include <fstream>
    #include <iostream>
    #include <thread>
    #include <chrono>
    #include "sphfem/sphfem.hpp"

    struct StandIn {                         // records what remesh_gmsh asks of the tool
        std::string log;
        void remesh_read_segments(const std::string &geo) { std::ofstream(geo) << "// geometry\n"; log += "geo;"; }
        void read_remesh_gmsh(const std::string &inp) { log += "read " + inp + ";"; }
    };

    static void touch(const std::string &f) { std::ofstream(f) << "*NODE\n"; }

    int main(int argc, char **argv) {
        const std::string dir = argv[1], geo = dir + "/t.geo", inp = dir + "/t.inp";
        const std::string mode = argv[2];
        StandIn t;
        if (mode == "stale") {                   // an .inp from before: older than the new .geo
            touch(inp);
            std::this_thread::sleep_for(std::chrono::milliseconds(20));
        }
        try {
            remesh_gmsh(&t, geo, inp, [&](const std::string &g, const std::string &i) {
                t.log += "mesh;";
                if (mode == "fresh") {
                    std::this_thread::sleep_for(std::chrono::milliseconds(20));
                    touch(i);
                }
            });
        } catch (const std::exception &e) {
            std::cout << t.log << "refused: " << e.what() << "\n";
            return 3;
        }
        std::cout << t.log << "\n";
        return 0;
    }
''')


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    d = tmp_path_factory.mktemp("remesh_gmsh")
    src = d / "p.cpp"
    src.write_text(PROGRAM)
    exe = d / "p"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def run(program, tmp_path, mode):
    r = subprocess.run([program, str(tmp_path), mode], capture_output=True, text=True, timeout=30)
    return r.returncode, r.stdout.strip()


def test_fresh_mesh_is_read_after_the_meshing_step(program, tmp_path):
    rc, out = run(program, tmp_path, "fresh")
    assert rc == 0 and out == f"geo;mesh;read {tmp_path}/t.inp;", out


def test_stale_mesh_is_refused(program, tmp_path):
    rc, out = run(program, tmp_path, "stale")
    assert rc == 3 and out.startswith("geo;mesh;refused:") and "read" not in out.split("refused")[0], out


def test_missing_mesh_is_refused(program, tmp_path):
    rc, out = run(program, tmp_path, "missing")
    assert rc == 3 and "refused" in out and "read " not in out, out
