"""Case directories for the C++ driver (driver/mfree_driver.cpp).

A case directory holds `scalars.txt` (one `key value` per line: sizes, time step, grid box,
kernel cut-off, wear switch, every constant block field as `block.field`, tool kinematics) and
raw little-endian arrays `<name>.bin`: particle fields (pos, vel, rho, T, h, fixed, S) and the
tool description (node positions, elements, layer flags, CSR pattern, K, M, f, T, segments).
Everything is what `solver.Simulation` uploads for the same case, so a driver run and a
Python run of the same case execute identical device work.
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib, cases

PARTICLE_ARRAYS = (("pos", np.float64), ("vel", np.float64), ("rho", np.float64), ("T", np.float64),
                   ("h", np.float64), ("fixed", np.float64), ("S", np.float64))
TOOL_ARRAYS = (("node_pos", np.float64), ("elem", np.int32), ("layer", np.int32), ("row_ptr", np.int32),
               ("col_ind", np.int32), ("K", np.float64), ("M", np.float64), ("f", np.float64),
               ("T", np.float64), ("left", np.float64), ("right", np.float64), ("n", np.float64),
               ("marks", np.int32))
BLOCKS = (("phys", _lib.PhysConstants), ("corr", _lib.CorrConstants), ("joco", _lib.JocoConstants),
          ("trml_wp", _lib.TrmlConstants), ("trml_tool", _lib.TrmlConstants), ("tool", _lib.ToolConstants),
          ("wear", _lib.WearConstants))


def write_case_dir(case: cases.Case, path: str, rr2_min: float | None = None, wear: bool | None = None,
                   tool_inp: str | None = None, remesh_low_level: bool = False) -> str:
    os.makedirs(path, exist_ok=True)
    init = cases.initial_particles(case)
    sc = {"N": case.N, "dt": case.dt, "h": case.h,
          "bbmin_x": case.grid_bbmin[0], "bbmin_y": case.grid_bbmin[1],
          "bbmax_x": case.grid_bbmax[0], "bbmax_y": case.grid_bbmax[1],
          "rr2_min": cases.default_rr2_min(case.dx) if rr2_min is None else rr2_min,
          "wear": int(case.wear if wear is None else wear),
          "remesh_low_level": int(bool(remesh_low_level))}
    for name, dt in PARTICLE_ARRAYS:
        np.ascontiguousarray(init[name], dtype=dt).tofile(os.path.join(path, f"p_{name}.bin"))
    for block, struct in BLOCKS:
        values = getattr(case.const, block)
        for field, _ in struct._fields_:
            sc[f"{block}.{field}"] = float(values[field])
    if case.tool is not None:
        tm = case.tool
        t = cases.initial_tool(case)
        sc.update(node_num=len(tm.nodes), ele_num=len(tm.elems), nz_num=len(tm.col_ind), seg_num=len(tm.seg_left),
                  color_total=tm.n_colors, color_cutting_edge=tm.n_cut_colors)
        for k, v in t["scalars"].items():
            sc[f"tool.{k}"] = float(v)
        for name, dt in TOOL_ARRAYS:
            np.ascontiguousarray(t[name], dtype=dt).tofile(os.path.join(path, f"t_{name}.bin"))
        # remeshing description: node flags, fixed-boundary elements and their edge flags (int4)
        np.ascontiguousarray(t["node_flags"], dtype=np.int32).tofile(os.path.join(path, "t_node_flags.bin"))
        np.ascontiguousarray(t["fixed_elems"], dtype=np.int32).tofile(os.path.join(path, "t_fixed_elems.bin"))
        ff = np.zeros((len(t["fixed_elems"]), 4), np.int32)
        ff[:, :3] = t["fixed_flags"]
        ff.tofile(os.path.join(path, "t_fixed_flags.bin"))
        sc["fixed_num"] = len(t["fixed_elems"])
        if tool_inp is not None:
            # the driver then reads the mesh from tool.inp and sets the tool up from it
            if tool_inp not in ("gmsh", "abaqus"):
                raise ValueError("tool_inp must be 'gmsh' or 'abaqus'")
            cases.write_inp(tm, os.path.join(path, "tool.inp"), gmsh=tool_inp == "gmsh")
            sc["tool.inp_gmsh"] = int(tool_inp == "gmsh")
            if tm.outline is not None:
                # the tool_FEM(tl, length, height, rake, clearance, r, mu) constructor's arguments
                for k, v in zip(("length", "height", "rake", "clearance", "r"), tm.outline):
                    sc[f"tool_outline.{k}"] = float(v)
    else:
        sc["node_num"] = 0
    with open(os.path.join(path, "scalars.txt"), "w") as f:
        for k, v in sc.items():
            f.write(f"{k} {float(v)!r}\n")
    return path


def read_driver_output(path: str, N: int) -> dict:
    """Final state and force history written by the driver into `path`."""
    out = {"pos": np.fromfile(os.path.join(path, "pos.bin")).reshape(N, 2),
           "vel": np.fromfile(os.path.join(path, "vel.bin")).reshape(N, 2),
           "S": np.fromfile(os.path.join(path, "S.bin")).reshape(N, 4),
           "T": np.fromfile(os.path.join(path, "T.bin")),
           "idx": np.fromfile(os.path.join(path, "idx.bin"), dtype=np.int32)}
    out["forces"] = np.loadtxt(os.path.join(path, "forces.txt"), ndmin=2)
    return out
