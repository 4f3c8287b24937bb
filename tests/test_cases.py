"""Workload generator: determinism, SplitMix64, lattice, tool mesh and FEM set-up invariants."""
import os

import numpy as np
import pytest

from sph_fem_2d_cutting_gpu_amd import cases


def test_splitmix64_known_answers():
    # reference outputs of SplitMix64 (Steele, Lea, Flood) for state 0 and 1234567
    assert int(cases.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF
    got = [int(v) for v in cases.splitmix64(1234567, 5)]
    assert got == [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
                   16408922859458223821]


def test_workload_is_deterministic():
    a, b = cases.config1(), cases.config1()
    assert np.array_equal(a.pos, b.pos)
    assert np.array_equal(a.tool.nodes, b.tool.nodes) and np.array_equal(a.tool.elems, b.tool.elems)
    assert np.array_equal(a.tool.K, b.tool.K)
    c = cases.config1(seed=7)
    assert not np.array_equal(a.pos, c.pos)


def test_lattice_and_constants_of_6400_case():
    c = cases.config1()
    assert c.N == 6400
    dx = 2.5e-4
    lo = np.array([0.005, 0.060 - 39 * dx])                 # benchmarks_NZ.cpp:67-68 placement
    ij = np.round((c.pos - lo) / dx)
    assert ij[:, 0].min() == 0 and ij[:, 0].max() == 159 and ij[:, 1].max() == 39
    assert np.abs(c.pos - (lo + ij * dx)).max() <= 0.0100001 * dx     # +-1 % jitter
    assert c.hh[0] == pytest.approx(1.5 * dx)
    k = c.const
    assert k.phys["G"] == pytest.approx(1.1 / 2.7)
    assert k.joco["A"] == pytest.approx(0.01098) and k.joco["Tmelt"] == 1878.0
    assert k.joco["eps_dot_ref"] == pytest.approx(1e-6)            # 1/s in 1/us
    assert c.tool_vel[0] == pytest.approx(70 / 60 * 1e-4)          # 70 m/min in cm/us
    # CFL step of benchmarks_NZ.cpp:360
    assert c.dt == pytest.approx(0.1 * 1.5 * dx / (np.sqrt(k.phys["K"] / k.phys["rho0"]) + np.sqrt(c.tool_vel[0])))
    assert 1800 <= len(c.tool.elems) <= 2200


def _angles(p):
    a = p[:, 1] - p[:, 0]
    b = p[:, 2] - p[:, 0]
    c = p[:, 2] - p[:, 1]

    def ang(u, v):
        cs = np.sum(u * v, axis=1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1)
        return np.degrees(np.arccos(np.clip(cs, -1, 1)))
    return np.minimum(np.minimum(ang(a, b), ang(-a, c)), ang(-b, -c))


@pytest.mark.parametrize("mk", [cases.config1, cases.config2])
def test_tool_mesh_quality_and_topology(mk):
    tm = mk().tool
    p = tm.nodes[tm.elems[:, :3]]
    det = (p[:, 1, 0] - p[:, 0, 0]) * (p[:, 2, 1] - p[:, 0, 1]) - (p[:, 2, 0] - p[:, 0, 0]) * (p[:, 1, 1] - p[:, 0, 1])
    assert np.all(det > 0)
    assert _angles(p).min() > 10.0
    # chain of cutting segments: consecutive, unit inward normals, marks point at the end nodes
    assert np.array_equal(tm.seg_left[1:], tm.seg_right[:-1])
    assert np.allclose(np.linalg.norm(tm.seg_n, axis=1), 1.0)
    assert np.all(np.sum((tm.tl - tm.seg_left) * tm.seg_n, axis=1) >= -1e-15)
    assert np.array_equal(tm.nodes[tm.seg_marks[:, 2]], tm.seg_left)
    assert np.array_equal(tm.nodes[tm.seg_marks[:, 3]], tm.seg_right)
    # colouring: elements sharing a node never share a colour
    for v in range(len(tm.nodes)):
        es = np.nonzero(np.any(tm.elems[:, :3] == v, axis=1))[0]
        cols = tm.elems[es, 3]
        assert len(set(cols.tolist())) == len(cols)
    assert tm.seg_marks[:, 1].max() <= tm.n_cut_colors
    # CSR: sorted columns with a diagonal in every row
    for i in range(len(tm.nodes)):
        cols = tm.col_ind[tm.row_ptr[i]:tm.row_ptr[i + 1]]
        assert np.all(np.diff(cols) > 0) and i in cols


def test_dirichlet_rows_are_identity():
    tm = cases.config1().tool
    fixed = np.nonzero(tm.fixed_edge_node)[0]
    for i in fixed:
        lo, hi = tm.row_ptr[i], tm.row_ptr[i + 1]
        row = dict(zip(tm.col_ind[lo:hi].tolist(), tm.K[lo:hi].tolist()))
        assert row[i] == 1.0 and all(v == 0.0 for k, v in row.items() if k != i)
        assert tm.M[i] == 1.0 and tm.f[i] == 300.0


def test_grid_geometry_rule():
    c = cases.config1()
    g = cases.grid_geometry(c.grid_bbmin, c.grid_bbmax, c.h)
    assert g["dx"] == c.h
    assert g["nx"] == int(np.ceil((c.grid_bbmax[0] - c.grid_bbmin[0] + 2e-6) / c.h))
    # bbox rule of benchmarks_NZ.cpp:281-288
    hi_x, hi_y = c.pos[:, 0].max(), c.pos[:, 1].max()
    assert c.grid_bbmax[1] >= 0.06 + (0.005 + 160 * 2.5e-4) - 0.005 - 1e-12 or c.grid_bbmax[1] >= hi_y + 500 * 2.5e-4


def test_chip_workload_density_ratio():
    c = cases.chip_case(n_target=200_000)
    assert 0.8e5 < c.N < 4e5
    assert np.allclose(c.S[:, 3], -(c.S[:, 0] + c.S[:, 2]))
    # nearest-neighbour spacing ranges over ~sqrt(10)
    from scipy.spatial import cKDTree
    d, _ = cKDTree(c.pos).query(c.pos, k=2)
    nn = d[:, 1]
    assert nn.max() / nn.min() > 2.5


def test_case_dir_for_the_cpp_driver(tmp_path):
    """case_io writes plain `key float` lines and raw arrays of the right sizes."""
    from sph_fem_2d_cutting_gpu_amd import case_io
    c = cases.small_cutting_case(nx=16, ny=8)
    d = case_io.write_case_dir(c, str(tmp_path))
    sc = {}
    for line in open(tmp_path / "scalars.txt"):
        k, v = line.split()
        sc[k] = float(v)                      # what the driver's `>> double` accepts
    for k in ("N", "dt", "h", "bbmin_x", "bbmax_y", "rr2_min", "phys.mass", "joco.tanh_d", "tool.ref_dist",
              "node_num", "seg_num"):
        assert k in sc, k
    assert sc["N"] == c.N
    assert os.path.getsize(os.path.join(d, "p_pos.bin")) == 16 * c.N
    assert os.path.getsize(os.path.join(d, "t_row_ptr.bin")) == 4 * (int(sc["node_num"]) + 1)
    assert os.path.getsize(os.path.join(d, "t_marks.bin")) == 16 * int(sc["seg_num"])
