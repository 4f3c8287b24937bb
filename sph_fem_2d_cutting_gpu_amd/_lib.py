"""ctypes binding of libsphfem_hip.so (the C ABI declared in include/sphfem.h).

The library is loaded from the package directory only; a missing library raises
immediately -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsphfem_hip.so")

OK = 0
ERR_HIP, ERR_ARG, ERR_STATE, ERR_RANGE = 1, 2, 3, 4      # SPHFEM_ERR_* of include/sphfem.h
FP64 = 0
FP32C = 1

_dbl = C.c_double
_vp = C.c_void_p


class Double2(C.Structure):
    _fields_ = [("x", _dbl), ("y", _dbl)]


def _s(name, fields, t=_dbl):
    return type(name, (C.Structure,), {"_fields_": [(f, t) for f in fields]})


PhysConstants = _s("PhysConstants", ["E", "nu", "rho0", "K", "G", "mass", "mass_tool"])
TrmlConstants = _s("TrmlConstants", ["cp", "tq", "k", "alpha", "T_init", "eta"])
CorrConstants = _s("CorrConstants", ["wdeltap", "stresseps", "xspheps", "alpha", "beta", "eta"])
JocoConstants = _s("JocoConstants", ["A", "B", "C", "n", "m", "Tmelt", "Tref", "eps_dot_ref",
                                     "tanh_a", "tanh_b", "tanh_c", "tanh_d"])
ToolConstants = _s("ToolConstants", ["cp", "tq", "k", "eta", "rho", "T0", "h", "beta", "mu",
                                     "contact_alpha", "slave_mass", "hc"])
WearConstants = _s("WearConstants", ["A", "B", "C", "D"])
ToolKinematics = _s("ToolKinematics", ["vel_x", "vel_y", "tl_x", "tl_y", "br_x", "br_y", "ref_dist",
                                       "lowest_point", "right_point", "flank_retreat", "flank_contact"])


Tuning = _s("Tuning", ["pair_block", "pair_chunk", "update_block", "pair_kernel", "resort_one_block"], C.c_int)
SORT_ONE_BLOCK_MAX = 1 << 19
PAIRS_LIST, PAIRS_STAGED, PAIRS_LIST_FUSED, PAIRS_STAGED_FUSED = 0, 1, 2, 3


class GeomConstants(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("bbmin_x", _dbl), ("bbmin_y", _dbl), ("dx", _dbl)]


# particle_gpu.h:23-83 field order
PARTICLE_FIELDS = ("pos", "vel", "h", "rho", "p", "S", "R", "fc", "ft", "n", "fixed", "blanked", "tool_particle",
                   "eps_pl", "eps_pl_dot", "T", "pos_t", "vel_t", "rho_t", "S_t", "T_t", "S_der", "v_der", "idx",
                   "hash", "on_seg")


class Particles(C.Structure):
    _fields_ = [(f, _vp) for f in PARTICLE_FIELDS] + [("N", C.c_uint)]


class ToolDesc(C.Structure):
    _fields_ = ([(f, C.c_int) for f in ("node_num", "ele_num", "nz_num", "seg_num", "color_total",
                                        "color_cutting_edge")]
                + [(f, _vp) for f in ("position", "node_num_el", "layer", "csr_row_ptr", "csr_col_ind", "K", "M",
                                      "f", "T", "left", "right", "n", "marks")]
                + [("vel", Double2), ("tl_point_ref", Double2), ("br_point", Double2),
                   ("ref_dist", _dbl), ("lowest_point", _dbl), ("right_point", _dbl),
                   ("node_flags", _vp), ("fixed_num", C.c_int), ("fixed_elems", _vp), ("fixed_flags", _vp)])


TOOL_MAX_SEG = 5


class ToolArrays(C.Structure):
    _fields_ = ([(f, C.c_int) for f in ("node_num", "ele_num", "nz_num", "seg_num", "color_total",
                                        "color_cutting_edge", "fixed_ele_num")]
                + [(f, _vp) for f in ("position", "position_mapping", "node_num_el", "csr_row_ptr", "csr_col_ind",
                                      "node_flags", "K", "K_temp", "M", "M_inv", "f", "f_temp", "k", "T", "T_old",
                                      "T_mapping", "left", "right", "n", "heat_exchange", "marks", "fric_heat",
                                      "wear_rate", "wear_nodes", "physical_para", "sliding_force")])

MeshInfo = _s("MeshInfo", ["node_num", "ele_num", "nz_num", "cut_ele_num", "fixed_ele_num", "color_total",
                           "color_cutting_edge"], C.c_int)


class Line(C.Structure):
    _fields_ = [("a", _dbl), ("b", _dbl), ("vertical", C.c_int)]


class Segment(C.Structure):
    _fields_ = [("left", Double2), ("right", Double2), ("l", Line), ("n", Double2)]


class CircleSegment(C.Structure):
    _fields_ = [("r", _dbl), ("t1", _dbl), ("t2", _dbl), ("p", Double2)]


class ToolGeometry(C.Structure):
    """sphfem_tool_geometry: the host state of the reference's class tool (tool.h:109-138)."""
    _fields_ = [("num_seg", C.c_int), ("has_fillet", C.c_int), ("seg", Segment * TOOL_MAX_SEG),
                ("fillet", CircleSegment), ("vel", Double2), ("mu", _dbl), ("front", _dbl)]


BBox = _s("BBox", ["bbmin_x", "bbmax_x", "bbmin_y", "bbmax_y"])

_P = C.POINTER
_ctx = _vp
_SIGS = {
    "sphfem_create": (C.c_int, _P(_vp)),
    "sphfem_destroy": (_ctx,),
    "sphfem_synchronize": (_ctx,),
    "sphfem_check_error": (_ctx,),
    "sphfem_setup_physical_constants": (_ctx, _P(PhysConstants)),
    "sphfem_setup_corrector_constants": (_ctx, _P(CorrConstants)),
    "sphfem_setup_johnson_cook_constants": (_ctx, _P(JocoConstants)),
    "sphfem_setup_thermal_constants_wp": (_ctx, _P(TrmlConstants)),
    "sphfem_setup_thermal_constants_tool": (_ctx, _P(TrmlConstants)),
    "sphfem_setup_tool_constants": (_ctx, _P(ToolConstants)),
    "sphfem_setup_wear_constants": (_ctx, _P(WearConstants)),
    "sphfem_set_dt": (_ctx, _dbl),
    "sphfem_set_wear": (_ctx, C.c_int),
    "sphfem_set_kernel_cutoff": (_ctx, _dbl),
    "sphfem_malloc": (_ctx, _P(_vp), C.c_size_t),
    "sphfem_free": (_ctx, _vp),
    "sphfem_memcpy_htod": (_ctx, _vp, _vp, C.c_size_t),
    "sphfem_memcpy_dtoh": (_ctx, _vp, _vp, C.c_size_t),
    "sphfem_particles_alloc": (_ctx, C.c_uint, _P(Particles)),
    "sphfem_particles_free": (_ctx, _P(Particles)),
    "sphfem_grid_green_create": (_ctx, C.c_int, Double2, Double2, _dbl, _P(_vp)),
    "sphfem_grid_rothlin_create": (_ctx, C.c_int, Double2, Double2, _dbl, _P(_vp)),
    "sphfem_grid_kind": (_vp, _P(C.c_int)),
    "sphfem_grid_green_create_unlocked": (_ctx, C.c_int, C.c_int, _P(_vp)),
    "sphfem_grid_rothlin_create_unlocked": (_ctx, C.c_int, C.c_int, _P(_vp)),
    "sphfem_grid_update_geometry": (_ctx, _vp, _P(Particles), _dbl),
    "sphfem_grid_bounds": (_vp, _P(Double2), _P(C.c_int), _P(C.c_int)),
    "sphfem_grid_destroy": (_ctx, _vp),
    "sphfem_grid_set_resort_limit": (_ctx, _vp, C.c_int),
    "sphfem_leap_frog_graph_create": (_ctx, _vp, _P(Particles), _vp, _vp, C.c_int, _P(_vp)),
    "sphfem_batch_graph_create": (_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, _P(_vp)),
    "sphfem_leap_frog_graph_launch": (_ctx, _vp, C.c_int),
    "sphfem_leap_frog_export_rates": (_ctx, _vp, _P(Particles)),
    "sphfem_leap_frog_import_rates": (_ctx, _vp, _P(Particles)),
    "sphfem_leap_frog_set_tuning": (_ctx, _vp, _P(Tuning)),
    "sphfem_leap_frog_get_tuning": (_ctx, _vp, _P(Tuning)),
    "sphfem_leap_frog_graph_destroy": (_ctx, _vp),
    "sphfem_grid_resort_stats": (_ctx, _P(C.c_int), _P(C.c_int)),
    "sphfem_grid_geometry": (_vp, _P(GeomConstants), _P(C.c_int)),
    "sphfem_grid_assign_hashes": (_ctx, _vp, _P(Particles)),
    "sphfem_grid_sort": (_ctx, _vp, _P(Particles)),
    "sphfem_grid_get_cells": (_ctx, _vp, _P(Particles), _vp, _vp),
    "sphfem_interactions_setup_geometry_constants": (_ctx, _vp),
    "sphfem_interactions_monaghan": (_ctx, _P(Particles), _vp, _vp, C.c_int),
    "sphfem_interactions_heat_pse": (_ctx, _P(Particles), _vp, _vp, C.c_int),
    "sphfem_interactions_monaghan_fp32c": (_ctx, _vp, _P(Particles), _vp, _vp, C.c_int),
    "sphfem_tool_create": (_ctx, _P(ToolDesc), _P(_vp)),
    "sphfem_tool_destroy": (_ctx, _vp),
    "sphfem_tool_set_vel": (_ctx, _vp, Double2),
    "sphfem_tool_get_kinematics": (_ctx, _vp, _P(ToolKinematics)),
    "sphfem_tool_segments_update": (_ctx, _vp, _dbl),
    "sphfem_tool_update_FEM_gpu": (_ctx, _vp, _dbl),
    "sphfem_tool_segments_wear_rate_reset": (_ctx, _vp),
    "sphfem_tool_apply_heat_boundary_condition": (_ctx, _vp),
    "sphfem_FEM_reset_server": (_ctx, _vp),
    "sphfem_FEM_calculate_temp": (_ctx, _vp, _dbl),
    "sphfem_FEM_calculate_temp_rate": (_ctx, _vp),
    "sphfem_tool_download": (_ctx, _vp) + (_vp,) * 13,
    "sphfem_tool_upload": (_ctx, _vp) + (_vp,) * 13,
    "sphfem_tool_set_kinematics": (_ctx, _vp, _P(ToolKinematics)),
    "sphfem_leap_frog_create": (_ctx, C.c_uint, C.c_uint, C.c_int, _P(_vp)),
    "sphfem_leap_frog_destroy": (_ctx, _vp),
    "sphfem_leap_frog_set_precision": (_ctx, _vp, C.c_int),
    "sphfem_leap_frog_init": (_ctx, _vp, _P(Particles)),
    "sphfem_leap_frog_predict": (_ctx, _vp, _P(Particles), _dbl),
    "sphfem_leap_frog_correct": (_ctx, _vp, _P(Particles), _dbl),
    "sphfem_leap_frog_step": (_ctx, _vp, _P(Particles), _vp, _vp),
    "sphfem_leap_frog_step_rigid": (_ctx, _vp, _P(Particles), _vp, _vp),
    "sphfem_leap_frog_graph_create_rigid": (_ctx, _vp, _P(Particles), _vp, _vp, C.c_int, _P(_vp)),
    "sphfem_tool_geometry_from_points": (Double2, Double2, Double2, Double2, _dbl, _dbl, _P(ToolGeometry)),
    "sphfem_tool_geometry_from_angles": (Double2, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _P(ToolGeometry)),
    "sphfem_tool_geometry_move": (_P(ToolGeometry), _dbl),
    "sphfem_tool_geometry_low": (_P(ToolGeometry), _P(_dbl)),
    "sphfem_tool_geometry_safe_bb": (_P(ToolGeometry), _dbl, _P(BBox)),
    "sphfem_tool_geometry_inside": (_P(ToolGeometry), Double2, _P(C.c_int)),
    "sphfem_tool_geometry_sample": (_P(ToolGeometry), _dbl, _vp, C.c_long, _P(C.c_long)),
    "sphfem_tool_gpu_set_up_tool": (_ctx, _P(ToolGeometry), _dbl, _P(PhysConstants), _P(_vp)),
    "sphfem_rigid_tool_destroy": (_ctx, _vp),
    "sphfem_rigid_tool_get_geometry": (_ctx, _vp, _P(ToolGeometry)),
    "sphfem_rigid_tool_set_vel": (_ctx, _vp, Double2),
    "sphfem_tool_gpu_update_tool": (_ctx, _vp, _P(Particles)),
    "sphfem_compute_contact_forces": (_ctx, _P(Particles), _vp),
    "sphfem_material_fric_heat_gen": (_ctx, _P(Particles), Double2),
    "sphfem_interactions_setup_thermal_constants_tool": (_ctx, _P(TrmlConstants)),
    "sphfem_vtk_write_particles": (_ctx, _P(Particles), C.c_char_p, C.c_int),
    "sphfem_vtk_write_particles_host": (_P(Particles), C.c_char_p, C.c_int),
    "sphfem_vtk_write_rigid_tool": (_P(ToolGeometry), C.c_char_p),
    "sphfem_vtk_write_FEM_tool": (_ctx, _vp, C.c_char_p),
    "sphfem_vtk_write_tool_org": (_ctx, _vp, Double2, C.c_char_p),
    "sphfem_vtk_write_FEM_segments": (_ctx, _vp, C.c_char_p),
    "sphfem_vtk_write_contact_variables": (_ctx, _vp, C.c_char_p),
    "sphfem_read_vtk_temperature": (C.c_char_p, _vp, C.c_long, _P(C.c_long)),
    "sphfem_tool_wear_create": (_ctx, _dbl, _dbl, Double2, _P(_vp)),
    "sphfem_tool_wear_destroy": (_ctx, _vp),
    "sphfem_tool_wear_eval_usui": (_ctx, _vp, _P(Particles), _dbl, _P(_dbl), _P(_dbl), _P(_dbl)),
    "sphfem_tool_wear_get_accum_wear": (_vp, _P(_dbl)),
    "sphfem_report_force": (_ctx, _P(Particles), _P(Double2)),
    "sphfem_report_force_noalloc": (_ctx, _P(Particles), _vp),
    "sphfem_mesh_read_inp": (C.c_char_p, C.c_int, C.c_int, C.c_int, _P(_vp)),
    "sphfem_mesh_destroy": (_vp,),
    "sphfem_mesh_get_info": (_vp, _P(MeshInfo)),
    "sphfem_mesh_get_arrays": (_vp,) + (_vp,) * 11,
    "sphfem_mesh_set_temperature": (_vp, _vp, C.c_long),
    "sphfem_mesh_write_vtk": (_vp, C.c_char_p),
    "sphfem_mesh_tool_desc": (_vp, _P(ToolConstants), Double2, Double2, _P(ToolDesc)),
    "sphfem_tool_remesh_low_level": (_ctx, _vp, C.c_int),
    "sphfem_tool_remesh": (_ctx, _vp, _vp),
    "sphfem_tool_write_remesh_geo": (_ctx, _vp, C.c_char_p),
    "sphfem_tool_get_sizes": (_vp, _P(C.c_int), _P(C.c_int), _P(C.c_int), _P(C.c_int)),
    "sphfem_tool_get_arrays": (_vp, _P(ToolArrays)),
    "sphfem_timing_enable": (_ctx, C.c_int),
    "sphfem_timing_reset": (_ctx,),
    "sphfem_timing_select": (_ctx, C.c_char_p),
    "sphfem_timing_get": (_ctx, C.c_char_p, _P(_dbl), _P(C.c_long)),
}
for _name in ("material_eos", "corrector_artificial_stress", "material_stress_rate_jaumann", "contmech_continuity",
              "contmech_momentum", "contmech_advection", "plasticity_johnson_cook", "perform_boundary_conditions",
              "perform_boundary_conditions_thermal", "debug_invalidate"):
    _SIGS["sphfem_" + _name] = (_ctx, _P(Particles))
for _name in ("material_fric_heat_gen_FEM_v2", "compute_contact_forces_FEM_tool_v3"):
    _SIGS["sphfem_" + _name] = (_ctx, _P(Particles), _vp)
_SIGS["sphfem_wear_adjustment"] = (_ctx, _vp)

# symbols that do not return int
_RESTYPES = {"sphfem_last_error": C.c_char_p, "sphfem_timing_names": C.c_char_p}

_lib = None


class SphfemError(RuntimeError):
    pass


def lib():
    """Load libsphfem_hip.so from the package directory (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SphfemError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = list(args)
            f.restype = C.c_int
        for name, rt in _RESTYPES.items():
            f = getattr(L, name)
            f.argtypes = [_ctx]
            f.restype = rt
        _lib = L
    return _lib


def exported_symbols():
    return sorted(set(_SIGS) | set(_RESTYPES))


def check(rc, ctx=None, what=""):
    if rc != OK:
        msg = ""
        if ctx is not None:
            m = lib().sphfem_last_error(ctx)
            msg = m.decode() if m else ""
        raise SphfemError(f"{what} failed with status {rc}: {msg}")
