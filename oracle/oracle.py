"""ctypes binding of the FP64 host oracle -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, bench.py's cpu_baseline leg and __graft_entry__.smoke() as
the checker; the product never imports it.  The state lives in numpy arrays in
the reference layout (cases.PARTICLE_FIELDS); the C structs point into them.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from sph_fem_2d_cutting_gpu_amd import cases

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libsphfem_oracle.so")

_dbl = C.c_double
_P = C.c_void_p


def _struct(name, names, ctype=_dbl):
    return type(name, (C.Structure,), {"_fields_": [(n, ctype) for n in names]})


OPhys = _struct("OPhys", ["E", "nu", "rho0", "K", "G", "mass", "mass_tool"])
OTrml = _struct("OTrml", ["cp", "tq", "k", "alpha", "T_init", "eta"])
OCorr = _struct("OCorr", ["wdeltap", "stresseps", "xspheps", "alpha", "beta", "eta"])
OJoco = _struct("OJoco", ["A", "B", "C", "n", "m", "Tmelt", "Tref", "eps_dot_ref",
                          "tanh_a", "tanh_b", "tanh_c", "tanh_d"])


class OGeom(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("bbmin_x", _dbl), ("bbmin_y", _dbl), ("dx", _dbl)]


OToolc = _struct("OToolc", ["cp", "tq", "k", "eta", "rho", "T0", "h", "beta", "mu", "contact_alpha",
                            "slave_mass", "hc"])
OWearc = _struct("OWearc", ["A", "B", "C", "D"])


class OConsts(C.Structure):
    _fields_ = [("phys", OPhys), ("trml_wp", OTrml), ("trml_tool", OTrml), ("corr", OCorr),
                ("joco", OJoco), ("geom", OGeom), ("tc", OToolc), ("wear", OWearc), ("dt", _dbl),
                ("rr2_min", _dbl), ("wear_pro", C.c_int), ("thermal_bc", C.c_int), ("plastic", C.c_int),
                ("fric_heat", C.c_int), ("trml_itool", OTrml), ("itool_alpha_pse", _dbl)]


class OD2(C.Structure):
    _fields_ = [("x", _dbl), ("y", _dbl)]


class OLine(C.Structure):
    _fields_ = [("a", _dbl), ("b", _dbl), ("vertical", C.c_int)]


class OSeg(C.Structure):
    _fields_ = [("left", OD2), ("right", OD2), ("l", OLine), ("n", OD2)]


class OCirc(C.Structure):
    _fields_ = [("r", _dbl), ("t1", _dbl), ("t2", _dbl), ("p", OD2)]


class ORTool(C.Structure):
    """o_rtool: the rigid tool of tool.h / tool_gpu.cu:16-22."""
    _fields_ = [("num_seg", C.c_int), ("has_fillet", C.c_int), ("seg", OSeg * 5), ("fillet", OCirc), ("vel", OD2),
                ("mu", _dbl), ("front", _dbl), ("contact_alpha", _dbl), ("slave_mass", _dbl)]


_PF = [name for name, _, _ in cases.PARTICLE_FIELDS]


class OParticles(C.Structure):
    _fields_ = [("N", C.c_int)] + [(n, _P) for n in _PF]


class OStepper(C.Structure):
    _fields_ = [(n, _P) for n in ("pos_init", "vel_init", "S_init", "rho_init", "T_init",
                                  "cell_start", "cell_end")] + [("num_cell", C.c_int)]


_TOOL_ARRAYS = ("node_pos", "elem", "layer", "row_ptr", "col_ind", "K", "K_temp", "M", "Minv", "f",
                "f_temp", "T", "T_old", "k", "left", "right", "n", "heat_exchange", "fric_heat",
                "wear_rate", "wear_nodes", "physical_para", "sliding_force", "marks")
_TOOL_SCALARS = ("vel_x", "vel_y", "tl_x", "tl_y", "br_x", "br_y", "ref_dist", "lowest_point",
                 "right_point", "flank_retreat", "flank_contact")


class OTool(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("n_nodes", "n_elem", "nnz", "n_seg", "n_colors", "n_cut_colors")]
                + [(n, _P) for n in _TOOL_ARRAYS] + [(n, _dbl) for n in _TOOL_SCALARS]
                + [("node_flags", _P), ("n_fixed", C.c_int), ("fixed_elems", _P), ("fixed_flags", _P),
                   ("pos_map", _P), ("T_map", _P)])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        pc = C.POINTER(OConsts)
        pp = C.POINTER(OParticles)
        ps = C.POINTER(OStepper)
        pt = C.POINTER(OTool)
        pr = C.POINTER(ORTool)
        sig = {
            "o_compute_hashes": (pc, pp), "o_sort_and_cells": (pc, pp, ps), "o_init": (pp, ps),
            "o_sort_rothlin": (pc, pp, ps), "o_run_rothlin": (pc, pp, ps, pt, C.c_int),
            "o_update_geometry": (pc, pp, _dbl, C.c_int),
            "o_run_unlocked": (pc, pp, ps, pt, C.c_int, C.c_int, C.c_int),
            "o_predict": (pp, ps, _dbl), "o_correct": (pp, ps, _dbl), "o_material_eos": (pc, pp),
            "o_artificial_stress": (pc, pp), "o_interactions_monaghan": (pc, pp, ps),
            "o_interactions_monaghan_range": (pc, pp, ps, C.c_int, C.c_int),
            "o_interactions_heat_pse": (pc, pp, ps),
            "o_stress_rate_jaumann": (pc, pp), "o_continuity": (pp,), "o_momentum": (pc, pp),
            "o_advection": (pp,), "o_fric_heat_v21": (pc, pp, pt), "o_plasticity": (pc, pp),
            "o_contact_s1": (pc, pp, pt), "o_contact_s2": (pc, pp, pt), "o_boundary_conditions": (pp,),
            "o_boundary_conditions_thermal": (pc, pp), "o_invalidate": (pp,),
            "o_fem_reset_server": (pt,), "o_fem_calculate_temp": (pt, _dbl), "o_fem_apply_robin": (pc, pt),
            "o_fem_calculate_temp_rate": (pt,), "o_segments_wear_rate_reset": (pt,),
            "o_segments_update": (pt, _dbl), "o_tool_update": (pt, _dbl), "o_wear_adjustment": (pc, pt),
            "o_remesh_low_level": (pc, pt, C.c_int), "o_interp_new": (pt, C.c_void_p, C.c_int, C.c_void_p),
            "o_step": (pc, pp, ps, pt), "o_run": (pc, pp, ps, pt, C.c_int),
            "o_report_force": (pp, C.POINTER(_dbl), C.POINTER(_dbl)),
            "o_cubic_spline": (_dbl, _dbl, _dbl, _dbl, _dbl, _dbl, C.POINTER(_dbl)),
            "o_sigma_yield": (C.POINTER(OJoco), _dbl, _dbl, _dbl),
            "o_solve_secant": (C.POINTER(OJoco), _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, C.POINTER(C.c_int)),
            "o_set_num_threads": (C.c_int,), "o_num_threads": (),
            "o_rtool_from_points": (pr, OD2, OD2, OD2, OD2, _dbl, _dbl),
            "o_rtool_from_angles": (pr, OD2, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl),
            "o_rtool_move": (pr, _dbl), "o_rtool_low": (pr,), "o_rtool_safe_bb": (pr, _dbl, C.POINTER(_dbl)),
            "o_rtool_inside": (pr, _dbl, _dbl), "o_rtool_sample": (pr, _dbl, C.c_void_p, C.c_long),
            "o_rtool_establish_contact": (pr, _dbl, _dbl, C.POINTER(OD2), C.POINTER(OD2)),
            "o_rtool_contact": (pc, pp, pr), "o_rtool_tool_particles": (pc, pp, pr),
            "o_rtool_fric_heat": (pc, pp, pr), "o_run_rigid": (pc, pp, ps, pr, C.c_int, C.c_int),
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = list(args)
            f.restype = None
        L.o_sigma_yield.restype = _dbl
        L.o_solve_secant.restype = _dbl
        L.o_num_threads.restype = C.c_int
        L.o_update_geometry.restype = C.c_int
        L.o_run_unlocked.restype = C.c_int
        L.o_rtool_low.restype = _dbl
        L.o_rtool_inside.restype = C.c_int
        L.o_rtool_sample.restype = C.c_long
        L.o_rtool_establish_contact.restype = C.c_int
        _lib = L
    return _lib


def _fill(struct, d):
    for k, v in d.items():
        setattr(struct, k, v)
    return struct


def make_consts(case: cases.Case, rr2_min=None, wear_pro=None) -> OConsts:
    k = case.const
    g = cases.grid_geometry(case.grid_bbmin, case.grid_bbmax, case.h)
    c = OConsts()
    _fill(c.phys, k.phys)
    _fill(c.trml_wp, k.trml_wp)
    _fill(c.trml_tool, k.trml_tool)
    _fill(c.corr, k.corr)
    if k.joco is not None:
        _fill(c.joco, k.joco)
    _fill(c.tc, k.tool)
    _fill(c.wear, k.wear)
    c.geom.nx, c.geom.ny = g["nx"], g["ny"]
    c.geom.bbmin_x, c.geom.bbmin_y, c.geom.dx = g["bbmin_x"], g["bbmin_y"], g["dx"]
    c.dt = case.dt
    c.rr2_min = cases.default_rr2_min(case.dx) if rr2_min is None else rr2_min
    c.wear_pro = int(case.wear if wear_pro is None else wear_pro)
    c.thermal_bc = int(k.trml_tool["alpha"] != 0.0)
    c.plastic = int(k.joco is not None)
    c.fric_heat = int(k.trml_wp["eta"] != 0.0)
    if case.rigid is not None:
        _fill(c.trml_itool, case.rigid.trml_itool)
        c.itool_alpha_pse = case.rigid.trml_itool["alpha"]
    return c


def rtool_from_geometry(g, contact_alpha: float, slave_mass: float) -> ORTool:
    """An oracle rigid tool with the fields of a library ToolGeometry (same layout, plus the two
    contact constants of tool_gpu_set_up_tool)."""
    t = ORTool()
    C.memmove(C.addressof(t), C.addressof(g), C.sizeof(g))
    t.contact_alpha = contact_alpha
    t.slave_mass = slave_mass
    return t


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """The reference step on the host, over numpy arrays in the reference layout."""

    def __init__(self, case: cases.Case, rr2_min=None, wear_pro=None, threads=None, grid="green", max_cell=None):
        """grid: "green" or "rothlin"; max_cell: an unlocked grid of that capacity whose geometry
        follows the particles each step (grid_base(max_cell, N)), else the case's locked grid."""
        if grid not in ("green", "rothlin"):
            raise ValueError(grid)
        self.grid = grid
        self.max_cell = case.max_cell if max_cell is None else max_cell
        self.case = case
        self.c = make_consts(case, rr2_min, wear_pro)
        if threads:
            lib().o_set_num_threads(int(threads))
        self.P = cases.initial_particles(case)
        N = case.N
        nc = int(self.max_cell) if self.max_cell else self.c.geom.nx * self.c.geom.ny
        self.st = dict(pos_init=np.zeros((N, 2)), vel_init=np.zeros((N, 2)), S_init=np.zeros((N, 4)),
                       rho_init=np.zeros(N), T_init=np.zeros(N),
                       cell_start=np.full(nc, -1, dtype=np.int32), cell_end=np.zeros(nc, dtype=np.int32))
        self.t = cases.initial_tool(case)
        self.rt = None
        if case.rigid is not None:
            self.rt = rtool_from_geometry(case.rigid.geometry, case.rigid.contact_alpha, case.const.phys["mass"])
        self._bind()

    def _bind(self):
        self.Pc = OParticles()
        self.Pc.N = self.case.N
        for name in _PF:
            setattr(self.Pc, name, _ptr(self.P[name]))
        self.Sc = OStepper()
        for name, a in self.st.items():
            setattr(self.Sc, name, _ptr(a))
        self.Sc.num_cell = len(self.st["cell_start"])
        if self.t is None:
            self.Tc = None
            return
        t = self.t
        self.Tc = OTool()
        self.Tc.n_nodes, self.Tc.n_elem = len(t["node_pos"]), len(t["elem"])
        self.Tc.nnz, self.Tc.n_seg = len(t["col_ind"]), len(t["left"])
        self.Tc.n_colors, self.Tc.n_cut_colors = t["n_colors"], t["n_cut_colors"]
        for name in _TOOL_ARRAYS:
            setattr(self.Tc, name, _ptr(t[name]))
        _fill(self.Tc, t["scalars"])
        if "fixed_flags4" not in t:
            nf = len(t["fixed_elems"])
            t["node_flags"] = np.ascontiguousarray(t["node_flags"], dtype=np.int32)
            t["fixed_elems"] = np.ascontiguousarray(t["fixed_elems"], dtype=np.int32)
            t["fixed_flags4"] = np.zeros((nf, 4), np.int32)
            t["fixed_flags4"][:, :3] = t["fixed_flags"]
            t["pos_map"] = np.zeros_like(t["node_pos"])
            t["T_map"] = np.zeros_like(t["T"])
        self.Tc.node_flags, self.Tc.fixed_elems = _ptr(t["node_flags"]), _ptr(t["fixed_elems"])
        self.Tc.n_fixed = len(t["fixed_elems"])
        self.Tc.fixed_flags = _ptr(t["fixed_flags4"])
        self.Tc.pos_map, self.Tc.T_map = _ptr(t["pos_map"]), _ptr(t["T_map"])

    def clone(self) -> "Oracle":
        """Independent copy of the current state (arrays, tool kinematics and constants)."""
        import copy
        o = Oracle.__new__(Oracle)
        o.grid = self.grid
        o.max_cell = self.max_cell
        o.case = self.case
        o.c = OConsts()
        C.pointer(o.c)[0] = self.c
        o.P = {k: v.copy() for k, v in self.P.items()}
        o.rt = None
        if self.rt is not None:
            o.rt = ORTool()
            C.pointer(o.rt)[0] = self.rt
        o.st = {k: v.copy() for k, v in self.st.items()}
        if self.t is None:
            o.t = None
        else:
            o.t = {k: (v.copy() if isinstance(v, np.ndarray) else copy.deepcopy(v)) for k, v in self.t.items()}
            o.t["scalars"] = self.tool_scalars
        o._bind()
        return o

    @property
    def tool_scalars(self):
        return {k: getattr(self.Tc, k) for k in _TOOL_SCALARS}

    def _args(self, sig):
        m = {"c": C.byref(self.c), "p": C.byref(self.Pc), "s": C.byref(self.Sc),
             "t": C.byref(self.Tc) if self.Tc is not None else None,
             "r": C.byref(self.rt) if self.rt is not None else None}
        return [m[ch] for ch in sig]

    def call(self, fname, sig, *extra):
        getattr(lib(), fname)(*self._args(sig), *extra)

    def step(self, n=1):
        tc = C.byref(self.Tc) if self.Tc is not None else None
        if self.rt is not None or self.case.tool_p is not None:
            # leap_frog::step(particles, grid) (leap_frog.cu:99-160)
            if self.max_cell:
                raise NotImplementedError("rigid-tool steps on an unlocked grid")
            rt = C.byref(self.rt) if self.rt is not None else None
            lib().o_run_rigid(C.byref(self.c), C.byref(self.Pc), C.byref(self.Sc), rt, int(n),
                              int(self.grid == "rothlin"))
            return
        if self.max_cell:
            done = lib().o_run_unlocked(C.byref(self.c), C.byref(self.Pc), C.byref(self.Sc), tc, int(n),
                                        int(self.max_cell), int(self.grid == "rothlin"))
            if done != n:
                raise RuntimeError(f"unlocked grid outgrew max_cell={self.max_cell} at step {done}")
            return
        run = lib().o_run_rothlin if self.grid == "rothlin" else lib().o_run
        run(C.byref(self.c), C.byref(self.Pc), C.byref(self.Sc), tc, int(n))

    def remesh_low_level(self, freq: int):
        """remesh_low_level(tool, freq) (actions_gpu.cu:1022-1031)"""
        lib().o_remesh_low_level(C.byref(self.c), C.byref(self.Tc), int(freq))

    def interp_new(self, new_pos) -> np.ndarray:
        """do_nodal_value_interpolation (tool_FEM.cu:445-517) from the current mesh at new_pos,
        300 where no element holds a point"""
        pos = np.ascontiguousarray(new_pos, dtype=np.float64)
        out = np.full(len(pos), 300.0)
        lib().o_interp_new(C.byref(self.Tc), pos.ctypes.data_as(C.c_void_p), len(pos), out.ctypes.data_as(C.c_void_p))
        return out

    def remesh(self, tm_new):
        """read_remesh_gmsh (tool_FEM.cu:1811-1828) onto the tool mesh tm_new (cases.ToolMesh, built
        from the new mesh file at the current reference point): interpolated temperatures, the new
        mesh's segments and FEM system, kinematics kept except lowest point, right point and
        reference distance"""
        import dataclasses
        T_new = self.interp_new(tm_new.nodes)
        sc = self.tool_scalars
        sc.update(lowest_point=float(tm_new.lowest_point), right_point=float(tm_new.right_point),
                  ref_dist=float(tm_new.ref_dist))
        t = cases.initial_tool(dataclasses.replace(self.case, tool=tm_new))
        t["T"], t["T_old"] = T_new.copy(), T_new.copy()
        t["scalars"] = sc
        self.t = t
        self._bind()

    def report_force(self):
        fx, fy = _dbl(), _dbl()
        lib().o_report_force(C.byref(self.Pc), C.byref(fx), C.byref(fy))
        return fx.value, fy.value


def cubic_spline(pi, pj, h, rr2_min=1e-8):
    out = (_dbl * 3)()
    lib().o_cubic_spline(pi[0], pi[1], pj[0], pj[1], h, rr2_min, out)
    return np.array(out[:])


def joco_struct(joco: dict) -> OJoco:
    return _fill(OJoco(), joco)


def sigma_yield(joco: dict, eps, eps_dot, T):
    return lib().o_sigma_yield(C.byref(joco_struct(joco)), eps, eps_dot, T)


def solve_secant(joco: dict, init, tol, norm_S, eps, T, dt, mu):
    it = C.c_int()
    v = lib().o_solve_secant(C.byref(joco_struct(joco)), init, tol, norm_S, eps, T, dt, mu, C.byref(it))
    return v, it.value
