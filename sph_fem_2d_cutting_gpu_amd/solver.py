"""Host-side mirror of the reference's interface for the per-step path, over the C ABI.

Class and function names follow the reference (particle_gpu, grid_gpu_green, tool_FEM,
leap_frog, material_eos, interactions_monaghan, ..., report_force) so a caller of the
reference finds the same entry points; each call goes straight to libsphfem_hip.
`Simulation` wires a generated workload (cases.Case) the way the reference's setup
(benchmarks_NZ.cpp:51-367) and driver (mfree_gpu.cpp:229-366) do.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, lib
from . import cases

# trailing shapes and dtypes of the particle arrays (particle_gpu.h:23-83)
_FIELD_SPEC = {name: (shp, dt) for name, shp, dt in cases.PARTICLE_FIELDS}
RATE_FIELDS = ("pos_t", "vel_t", "rho_t", "S_t", "T_t")


class Context:
    """A device context: HIP stream plus the constant blocks (replaces the per-file __constant__ state)."""

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        check(lib().sphfem_create(device, C.byref(self.h)), None, "sphfem_create")

    def close(self):
        if self.h:
            lib().sphfem_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        check(getattr(lib(), name)(self.h, *args), self.h, name)

    def synchronize(self):
        self.call("sphfem_synchronize")

    def check_error(self):
        self.call("sphfem_check_error")

    def setup_constants(self, const: cases.Constants, dt: float, rr2_min: float, wear: bool):
        """actions_setup_* / interactions_setup_* of benchmarks_NZ.cpp:341-356."""
        L = _lib
        self.call("sphfem_setup_physical_constants", C.byref(L.PhysConstants(**const.phys)))
        self.call("sphfem_setup_corrector_constants", C.byref(L.CorrConstants(**const.corr)))
        if const.joco is not None:           # setups without plasticity never call it (benchmarks.cpp:148-246)
            self.call("sphfem_setup_johnson_cook_constants", C.byref(L.JocoConstants(**const.joco)))
        self.call("sphfem_setup_thermal_constants_wp", C.byref(L.TrmlConstants(**const.trml_wp)))
        self.call("sphfem_setup_thermal_constants_tool", C.byref(L.TrmlConstants(**const.trml_tool)))
        self.call("sphfem_setup_tool_constants", C.byref(L.ToolConstants(**const.tool)))
        self.call("sphfem_setup_wear_constants", C.byref(L.WearConstants(**const.wear)))
        self.call("sphfem_set_dt", float(dt))
        self.call("sphfem_set_kernel_cutoff", float(rr2_min))
        self.call("sphfem_set_wear", int(bool(wear)))

    def htod(self, dptr, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        self.call("sphfem_memcpy_htod", C.c_void_p(dptr), a.ctypes.data_as(C.c_void_p), a.nbytes)

    def dtoh(self, dptr, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        self.call("sphfem_memcpy_dtoh", out.ctypes.data_as(C.c_void_p), C.c_void_p(dptr), out.nbytes)
        return out

    # measurement
    def timing(self, enable: bool):
        self.call("sphfem_timing_enable", int(enable))

    def timing_select(self, names: str):
        self.call("sphfem_timing_select", names.encode())

    def timing_reset(self):
        self.call("sphfem_timing_reset")

    def timing_get(self, name: str):
        ms, n = C.c_double(), C.c_long()
        self.call("sphfem_timing_get", name.encode(), C.byref(ms), C.byref(n))
        return ms.value, n.value

    def timing_names(self):
        raw = lib().sphfem_timing_names(self.h)
        return [n for n in (raw.decode() if raw else "").split(",") if n]


class particle_gpu:
    """particle_gpu.h:23-83; constructor as particle_gpu(pos, vel, rho, T, h, fixed, N)."""

    def __init__(self, ctx: Context, N: int, **fields):
        self.ctx = ctx
        self.s = _lib.Particles()
        ctx.call("sphfem_particles_alloc", int(N), C.byref(self.s))
        self.N = int(N)
        # a stepper in the FP32C form keeps the rates in FP32 (sphfem_leap_frog_export_rates /
        # import_rates): set by leap_frog.attach, None otherwise
        self.stepper = None
        for k, v in fields.items():
            if v is not None:
                self.upload(k, v)

    def ptr(self, field):
        return getattr(self.s, field)

    def upload(self, field, arr):
        shp, dt = _FIELD_SPEC[field]
        a = np.asarray(arr, dtype=dt).reshape((self.N,) + shp)
        self.ctx.htod(self.ptr(field), a)
        if field in RATE_FIELDS and self.stepper is not None:
            self.ctx.call("sphfem_leap_frog_import_rates", self.stepper.h, C.byref(self.s))

    def download(self, field) -> np.ndarray:
        if field in RATE_FIELDS and self.stepper is not None:
            self.ctx.call("sphfem_leap_frog_export_rates", self.stepper.h, C.byref(self.s))
        shp, dt = _FIELD_SPEC[field]
        return self.ctx.dtoh(self.ptr(field), (self.N,) + shp, dt)

    def state(self, fields=None) -> dict:
        return {f: self.download(f) for f in (fields or _FIELD_SPEC)}

    def free(self):
        if self.s.pos:
            self.ctx.call("sphfem_particles_free", C.byref(self.s))


class grid_base:
    """grid_base (grid.h:23-64).  grid_X(ctx, num_part, bbmin, bbmax, h): locked cell grid of cell
    size h over [bbmin, bbmax] (grid.cu:130-149); grid_X.unlocked(ctx, max_cell, num_part): the
    geometry follows the particles at each update_geometry (grid.cu:88-127), which the step calls."""

    _create = None
    _create_unlocked = None
    kind = None

    def __init__(self, ctx: Context, num_part: int, bbmin, bbmax, h: float):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.call(self._create, int(num_part), _lib.Double2(*map(float, bbmin)),
                 _lib.Double2(*map(float, bbmax)), float(h), C.byref(self.h))

    @classmethod
    def unlocked(cls, ctx: Context, max_cell: int, num_part: int):
        g = cls.__new__(cls)
        g.ctx = ctx
        g.h = C.c_void_p()
        ctx.call(cls._create_unlocked, int(max_cell), int(num_part), C.byref(g.h))
        return g

    def _geometry(self):
        g = _lib.GeomConstants()
        nc = C.c_int()
        check(lib().sphfem_grid_geometry(self.h, C.byref(g), C.byref(nc)), None, "grid_geometry")
        return g, nc.value

    def _bounds(self):
        bb, locked, mc = _lib.Double2(), C.c_int(), C.c_int()
        check(lib().sphfem_grid_bounds(self.h, C.byref(bb), C.byref(locked), C.byref(mc)), None, "grid_bounds")
        return bb, bool(locked.value), mc.value

    @property
    def geom(self):
        return self._geometry()[0]

    @property
    def num_cell(self):
        return self._geometry()[1]

    @property
    def max_cell(self):
        return self._bounds()[2]

    def is_locked(self):
        return self._bounds()[1]

    def nx(self):
        return self.geom.nx

    def ny(self):
        return self.geom.ny

    def bbmin_x(self):
        return self.geom.bbmin_x

    def bbmin_y(self):
        return self.geom.bbmin_y

    def bbmax_x(self):
        return self._bounds()[0].x

    def bbmax_y(self):
        return self._bounds()[0].y

    def dx(self):
        return self.geom.dx

    def update_geometry(self, p: particle_gpu, kernel_width: float = 2.0):
        """grid.cu:88-127 (no-op for a locked grid; synchronises the stream)."""
        self.ctx.call("sphfem_grid_update_geometry", self.h, C.byref(p.s), float(kernel_width))

    def assign_hashes(self, p: particle_gpu):
        self.ctx.call("sphfem_grid_assign_hashes", self.h, C.byref(p.s))

    def sort(self, p: particle_gpu):
        self.ctx.call("sphfem_grid_sort", self.h, C.byref(p.s))

    def resort_stats(self):
        """(mode, changed) of the last step's re-sort: mode 1 = full radix sort, 2 = incremental."""
        m, n = C.c_int(), C.c_int()
        self.ctx.call("sphfem_grid_resort_stats", C.byref(m), C.byref(n))
        return m.value, n.value

    def set_resort_limit(self, max_changed: int):
        """Largest number of cell changes per step re-sorted incrementally (0: always the full sort)."""
        self.ctx.call("sphfem_grid_set_resort_limit", self.h, int(max_changed))

    def get_cells(self, p, cell_start_dptr, cell_end_dptr):
        """grid.h:57 get_cells(particles, cell_start, cell_end); the Green grid does not read p."""
        ps = C.byref(p.s) if p is not None else None
        self.ctx.call("sphfem_grid_get_cells", self.h, ps, C.c_void_p(cell_start_dptr), C.c_void_p(cell_end_dptr))

    def download_cells(self, p=None):
        a, b = C.c_void_p(), C.c_void_p()
        self.ctx.call("sphfem_malloc", C.byref(a), 4 * self.num_cell)
        self.ctx.call("sphfem_malloc", C.byref(b), 4 * self.num_cell)
        try:
            self.get_cells(p, a.value, b.value)
            cs = self.ctx.dtoh(a.value, (self.num_cell,), np.int32)
            ce = self.ctx.dtoh(b.value, (self.num_cell,), np.int32)
        finally:
            self.ctx.call("sphfem_free", a)
            self.ctx.call("sphfem_free", b)
        return cs, ce


class grid_gpu_green(grid_base):
    """grid_gpu_green(num_part, bbmin, bbmax, h) (grid_gpu_green.h:84-85): sort gathers into a double
    buffer and swaps pointers; cell tables mark empty cells with -1."""

    _create = "sphfem_grid_green_create"
    _create_unlocked = "sphfem_grid_green_create_unlocked"
    kind = "green"


class grid_gpu_rothlin(grid_base):
    """grid_gpu_rothlin(num_part, bbmin, bbmax, h) (grid_gpu_rothlin.h:79-80): sort gathers in place
    (idx then carries particle labels; ft and n are not reordered); get_cells gives dense offsets
    computed from the particles' sorted hashes."""

    _create = "sphfem_grid_rothlin_create"
    _create_unlocked = "sphfem_grid_rothlin_create_unlocked"
    kind = "rothlin"


GRIDS = {"green": grid_gpu_green, "rothlin": grid_gpu_rothlin}


def _p(a, dt):
    a = np.ascontiguousarray(a, dtype=dt)
    return a, a.ctypes.data_as(C.c_void_p)


def tool_desc(case: cases.Case, remeshable: bool = True):
    """The sphfem_tool_desc of a case's tool (host arrays of set_up_segments_from_mesh and
    set_up_FEM_tool_gpu) and the arrays it points into, which must outlive its use."""
    tm = case.tool
    init = cases.initial_tool(case)
    keep = []

    def arr(a, dt):
        a2, ptr = _p(a, dt)
        keep.append(a2)
        return ptr

    d = _lib.ToolDesc()
    d.node_num, d.ele_num, d.nz_num = len(tm.nodes), len(tm.elems), len(tm.col_ind)
    d.seg_num, d.color_total, d.color_cutting_edge = len(tm.seg_left), tm.n_colors, tm.n_cut_colors
    d.position = arr(init["node_pos"], np.float64)
    d.node_num_el = arr(init["elem"], np.int32)
    d.layer = arr(init["layer"], np.int32)
    d.csr_row_ptr = arr(init["row_ptr"], np.int32)
    d.csr_col_ind = arr(init["col_ind"], np.int32)
    d.K = arr(init["K"], np.float64)
    d.M = arr(init["M"], np.float64)
    d.f = arr(init["f"], np.float64)
    d.T = arr(init["T"], np.float64)
    d.left = arr(init["left"], np.float64)
    d.right = arr(init["right"], np.float64)
    d.n = arr(init["n"], np.float64)
    d.marks = arr(init["marks"], np.int32)
    sc = init["scalars"]
    d.vel = _lib.Double2(sc["vel_x"], sc["vel_y"])
    d.tl_point_ref = _lib.Double2(sc["tl_x"], sc["tl_y"])
    d.br_point = _lib.Double2(sc["br_x"], sc["br_y"])
    d.ref_dist, d.lowest_point, d.right_point = sc["ref_dist"], sc["lowest_point"], sc["right_point"]
    d.node_flags = arr(init["node_flags"], np.int32) if remeshable else None
    d.fixed_num = len(init["fixed_elems"]) if remeshable else 0
    d.fixed_elems = arr(init["fixed_elems"], np.int32)
    ff = np.zeros((len(init["fixed_elems"]), 4), np.int32)
    ff[:, :3] = init["fixed_flags"]
    d.fixed_flags = arr(ff, np.int32)
    return d, keep


class tool_FEM:
    """Rigid FEM tool with its thermal solver and cutting-edge segments (tool_FEM.h:34-83)."""

    def __init__(self, ctx: Context, case: cases.Case, remeshable: bool = True):
        """remeshable: pass the node flags and fixed boundary the remesh entry points need"""
        self.ctx = ctx
        d, keep = tool_desc(case, remeshable)
        self.h = C.c_void_p()
        ctx.call("sphfem_tool_create", C.byref(d), C.byref(self.h))
        self.n_nodes, self.n_seg, self.nnz = d.node_num, d.seg_num, d.nz_num

    def remesh_low_level(self, freq: int):
        """remesh_low_level(tool, freq) (actions_gpu.cu:1022-1031): interior nodes shifted by
        mesh_shift(freq), the FEM system assembled again, temperatures interpolated"""
        self.ctx.call("sphfem_tool_remesh_low_level", self.h, int(freq))

    def remesh(self, mesh: "mesh_read"):
        """read_remesh_gmsh (tool_FEM.cu:1811-1828) from a mesh already read: temperatures
        interpolated onto the new nodes, segments and FEM system of the new mesh.  Graphs captured
        before must be captured again."""
        self.ctx.call("sphfem_tool_remesh", self.h, mesh.h)
        self.n_nodes, _, self.nnz, self.n_seg = self.sizes()

    def arrays(self) -> dict:
        """device pointers and sizes of the tool's arrays (the reference's mesh_GPU_m,
        FEM_thermal_2D_GPU and segment_FEM_gpu fields); valid until a remesh onto a new mesh"""
        a = _lib.ToolArrays()
        check(lib().sphfem_tool_get_arrays(self.h, C.byref(a)), None, "tool_get_arrays")
        return {f: getattr(a, f) for f, _ in _lib.ToolArrays._fields_}

    def sizes(self):
        """(node_number, element_number, non-zeros, cutting_segment_size) of the current mesh"""
        v = [C.c_int() for _ in range(4)]
        check(lib().sphfem_tool_get_sizes(self.h, *[C.byref(x) for x in v]), None, "tool_get_sizes")
        return tuple(x.value for x in v)

    def write_remesh_geo(self, path: str):
        """remesh_read_segments (tool_FEM.cu:1708-1810) as a Gmsh .geo script"""
        self.ctx.call("sphfem_tool_write_remesh_geo", self.h, str(path).encode())

    def get_vel(self):
        k = self.kinematics()
        return np.array([k["vel_x"], k["vel_y"]])

    def set_vel(self, v):
        self.ctx.call("sphfem_tool_set_vel", self.h, _lib.Double2(float(v[0]), float(v[1])))

    def kinematics(self) -> dict:
        k = _lib.ToolKinematics()
        self.ctx.call("sphfem_tool_get_kinematics", self.h, C.byref(k))
        return {f: getattr(k, f) for f, _ in _lib.ToolKinematics._fields_}

    def segments_update(self, dt):
        self.ctx.call("sphfem_tool_segments_update", self.h, float(dt))

    def tool_update_FEM_gpu(self, dt):
        self.ctx.call("sphfem_tool_update_FEM_gpu", self.h, float(dt))

    def segments_wear_rate_reset(self):
        self.ctx.call("sphfem_tool_segments_wear_rate_reset", self.h)

    def apply_heat_boundary_condition(self):
        self.ctx.call("sphfem_tool_apply_heat_boundary_condition", self.h)

    def reset_FEM_server(self):
        self.ctx.call("sphfem_FEM_reset_server", self.h)

    def calculate_temp(self, dt):
        self.ctx.call("sphfem_FEM_calculate_temp", self.h, float(dt))

    def calculate_temp_rate(self):
        self.ctx.call("sphfem_FEM_calculate_temp_rate", self.h)

    _STATE = ("node_pos", "T", "K", "M", "k", "left", "right", "n", "heat_exchange", "fric_heat",
              "physical_para", "sliding_force", "wear_nodes")

    def upload(self, state: dict):
        """Restore tool arrays (keys as download(); missing keys are left unchanged) and kinematics."""
        keep, ptrs = [], []
        for k in self._STATE:
            if k in state and state[k] is not None:
                a = np.ascontiguousarray(state[k], dtype=np.float64)
                keep.append(a)
                ptrs.append(a.ctypes.data_as(C.c_void_p))
            else:
                ptrs.append(None)
        self.ctx.call("sphfem_tool_upload", self.h, *ptrs)
        if "scalars" in state:
            kin = _lib.ToolKinematics(**{f: float(state["scalars"][f]) for f, _ in _lib.ToolKinematics._fields_})
            self.ctx.call("sphfem_tool_set_kinematics", self.h, C.byref(kin))

    def download(self) -> dict:
        nn, ns, nz = self.n_nodes, self.n_seg, self.nnz
        out = dict(node_pos=np.empty((nn, 2)), T=np.empty(nn), K=np.empty(nz), M=np.empty(nn), k=np.empty(nn),
                   left=np.empty((ns, 2)), right=np.empty((ns, 2)), n=np.empty((ns, 2)),
                   heat_exchange=np.empty((ns, 2)), fric_heat=np.empty(ns), physical_para=np.empty((ns, 4)),
                   sliding_force=np.empty(ns), wear_nodes=np.empty(ns))
        order = ("node_pos", "T", "K", "M", "k", "left", "right", "n", "heat_exchange", "fric_heat",
                 "physical_para", "sliding_force", "wear_nodes")
        self.ctx.call("sphfem_tool_download", self.h, *[out[k].ctypes.data_as(C.c_void_p) for k in order])
        out["scalars"] = self.kinematics()
        return out


def _d2(v):
    return _lib.Double2(float(v[0]), float(v[1]))


class tool:
    """The reference's class tool (tool.h:31-141): a rigid tool on the host, a closed polygon of four
    segments, or five when a fillet of radius r rounds the cutting edge.  The constructors and
    methods run the library's host geometry (no device work)."""

    def __init__(self, geometry: _lib.ToolGeometry | None = None):
        self.g = geometry if geometry is not None else _lib.ToolGeometry()

    @classmethod
    def from_points(cls, tl, tr, br, bl, r: float = 0.0, mu_fric: float = 0.0) -> "tool":
        """tool(tl, tr, br, bl[, r], mu_fric) (tool.cpp:316-331)."""
        t = cls()
        check(lib().sphfem_tool_geometry_from_points(_d2(tl), _d2(tr), _d2(br), _d2(bl), float(r), float(mu_fric),
                                                     C.byref(t.g)), None, "tool(points)")
        return t

    @classmethod
    def from_angles(cls, tl, length: float, height: float, rake_angle: float, clearance_angle: float,
                    r: float = 0.0, mu_fric: float = 0.0) -> "tool":
        """tool(tl, length, height, rake, clearance[, r], mu_fric), angles in degrees (tool.cpp:335-406)."""
        t = cls()
        check(lib().sphfem_tool_geometry_from_angles(_d2(tl), float(length), float(height), float(rake_angle),
                                                     float(clearance_angle), float(r), float(mu_fric),
                                                     C.byref(t.g)), None, "tool(angles)")
        return t

    def copy(self) -> "tool":
        g = _lib.ToolGeometry()
        C.pointer(g)[0] = self.g
        return tool(g)

    def update_tool(self, dt: float):
        check(lib().sphfem_tool_geometry_move(C.byref(self.g), float(dt)), None, "update_tool")

    def set_vel(self, vel):
        self.g.vel = _d2(vel)

    def get_vel(self):
        return np.array([self.g.vel.x, self.g.vel.y])

    def mu(self) -> float:
        return self.g.mu

    def front(self) -> float:
        return self.g.front

    def low(self) -> float:
        v = C.c_double()
        check(lib().sphfem_tool_geometry_low(C.byref(self.g), C.byref(v)), None, "low")
        return v.value

    def safe_bb(self, safety: float = 0.011) -> dict:
        b = _lib.BBox()
        check(lib().sphfem_tool_geometry_safe_bb(C.byref(self.g), float(safety), C.byref(b)), None, "safe_bb")
        return {f: getattr(b, f) for f, _ in _lib.BBox._fields_}

    def inside(self, pos) -> bool:
        v = C.c_int()
        check(lib().sphfem_tool_geometry_inside(C.byref(self.g), _d2(pos), C.byref(v)), None, "inside")
        return bool(v.value)

    def sample_tool(self, dx: float) -> np.ndarray:
        """Lattice points inside the tool (tool.cpp:112-135)."""
        n = C.c_long()
        check(lib().sphfem_tool_geometry_sample(C.byref(self.g), float(dx), None, 0, C.byref(n)), None, "sample")
        out = np.zeros((n.value, 2))
        check(lib().sphfem_tool_geometry_sample(C.byref(self.g), float(dx), out.ctypes.data_as(C.c_void_p),
                                                n.value, C.byref(n)), None, "sample")
        return out

    def get_segments(self) -> list:
        return [dict(left=(s.left.x, s.left.y), right=(s.right.x, s.right.y), a=s.l.a, b=s.l.b,
                     vertical=bool(s.l.vertical), n=(s.n.x, s.n.y)) for s in self.g.seg[: self.g.num_seg]]

    def get_fillet(self) -> dict | None:
        if not self.g.has_fillet:
            return None
        f = self.g.fillet
        return dict(r=f.r, t1=f.t1, t2=f.t2, p=(f.p.x, f.p.y))


class rigid_tool_gpu:
    """The device copy of a rigid tool made by tool_gpu_set_up_tool (tool_gpu.cu:219-266)."""

    def __init__(self, ctx: Context, t: tool, contact_alpha: float, phys: dict):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.call("sphfem_tool_gpu_set_up_tool", C.byref(t.g), float(contact_alpha),
                 C.byref(_lib.PhysConstants(**phys)), C.byref(self.h))

    def geometry(self) -> tool:
        """The device geometry as it stands (moved by every step and tool_gpu_update_tool)."""
        g = _lib.ToolGeometry()
        self.ctx.call("sphfem_rigid_tool_get_geometry", self.h, C.byref(g))
        return tool(g)

    def set_vel(self, vel):
        self.ctx.call("sphfem_rigid_tool_set_vel", self.h, _d2(vel))

    def get_vel(self):
        return self.geometry().get_vel()

    def destroy(self):
        if self.h:
            self.ctx.call("sphfem_rigid_tool_destroy", self.h)
            self.h = C.c_void_p()


def tool_gpu_set_up_tool(ctx: Context, t: tool, alpha: float, phys: dict) -> rigid_tool_gpu:
    """tool_gpu.h:22: the device copy of t with contact alpha and slave mass phys["mass"]."""
    return rigid_tool_gpu(ctx, t, alpha, phys)


def tool_gpu_update_tool(rt: rigid_tool_gpu, p: "particle_gpu | None"):
    """tool_gpu.h:25: move the device tool and the tool particles by global_dt * vel."""
    rt.ctx.call("sphfem_tool_gpu_update_tool", rt.h, C.byref(p.s) if p is not None else None)


def compute_contact_forces(p: "particle_gpu", rt: rigid_tool_gpu):
    """tool_gpu.h:28: penalty contact and friction against the rigid tool."""
    p.ctx.call("sphfem_compute_contact_forces", C.byref(p.s), rt.h)


def material_fric_heat_gen(p: "particle_gpu", vel):
    """actions_gpu.cu:956-965: friction heat of the rigid-tool contact (needs the workpiece's eta)."""
    p.ctx.call("sphfem_material_fric_heat_gen", C.byref(p.s), _d2(vel))


def interactions_setup_thermal_constants_tool(ctx: Context, trml: dict):
    """interactions_gpu.cu:494-502: the thermal block of tool particles in the pair sums."""
    ctx.call("sphfem_interactions_setup_thermal_constants_tool", C.byref(_lib.TrmlConstants(**trml)))


class leap_frog:
    """leap_frog(num_part, num_cell, node_num) and step(particles, grid, tool) (leap_frog.h:46-49)."""

    def __init__(self, ctx: Context, num_part: int, num_cell: int, node_num: int = 0, precision=_lib.FP64):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.call("sphfem_leap_frog_create", int(num_part), int(num_cell), int(node_num), C.byref(self.h))
        ctx.call("sphfem_leap_frog_set_precision", self.h, int(precision))

    def step(self, p: particle_gpu, g: grid_base, tool: tool_FEM | None = None):
        self.ctx.call("sphfem_leap_frog_step", self.h, C.byref(p.s), g.h, tool.h if tool else None)

    def step_rigid(self, p: particle_gpu, g: grid_base, rt: rigid_tool_gpu | None = None):
        """step(particles, grid) with global_tool = rt (leap_frog.cu:99-160)."""
        self.ctx.call("sphfem_leap_frog_step_rigid", self.h, C.byref(p.s), g.h, rt.h if rt else None)

    def set_tuning(self, pair_block: int = 256, pair_chunk: int = 32, update_block: int = 256,
                   pair_kernel: int = _lib.PAIRS_LIST_FUSED, resort_one_block: int = 0):
        t = _lib.Tuning(int(pair_block), int(pair_chunk), int(update_block), int(pair_kernel), int(resort_one_block))
        self.ctx.call("sphfem_leap_frog_set_tuning", self.h, C.byref(t))

    def get_tuning(self) -> dict:
        t = _lib.Tuning()
        self.ctx.call("sphfem_leap_frog_get_tuning", self.h, C.byref(t))
        return {f: getattr(t, f) for f, _ in _lib.Tuning._fields_}

    def capture(self, p: particle_gpu, g: grid_base, tool: tool_FEM | None = None, steps: int = 2):
        """Record `steps` (even) steps into a HIP graph; returns a StepGraph.  Needs one eager step first."""
        gh = C.c_void_p()
        self.ctx.call("sphfem_leap_frog_graph_create", self.h, C.byref(p.s), g.h, tool.h if tool else None,
                      int(steps), C.byref(gh))
        return StepGraph(self.ctx, gh, steps)

    def capture_rigid(self, p: particle_gpu, g: grid_base, rt: rigid_tool_gpu | None = None, steps: int = 2):
        """capture() for step_rigid."""
        gh = C.c_void_p()
        self.ctx.call("sphfem_leap_frog_graph_create_rigid", self.h, C.byref(p.s), g.h, rt.h if rt else None,
                      int(steps), C.byref(gh))
        return StepGraph(self.ctx, gh, steps)

    def init(self, p):
        self.ctx.call("sphfem_leap_frog_init", self.h, C.byref(p.s))

    def predict(self, p, dt):
        self.ctx.call("sphfem_leap_frog_predict", self.h, C.byref(p.s), float(dt))

    def correct(self, p, dt):
        self.ctx.call("sphfem_leap_frog_correct", self.h, C.byref(p.s), float(dt))


class StepGraph:
    """A captured run of an even number of steps; launch(n) advances n * steps steps."""

    def __init__(self, ctx: Context, h, steps: int):
        self.ctx, self.h, self.steps = ctx, h, steps

    def launch(self, repeats: int = 1):
        self.ctx.call("sphfem_leap_frog_graph_launch", self.h, int(repeats))

    def destroy(self):
        if self.h:
            self.ctx.call("sphfem_leap_frog_graph_destroy", self.h)
            self.h = C.c_void_p()


# ---- actions_gpu.h:33-72 / interactions_gpu.h:27-43 / output.h
def _act(name):
    def f(p: particle_gpu):
        p.ctx.call("sphfem_" + name, C.byref(p.s))
    f.__name__ = name
    return f


material_eos = _act("material_eos")
corrector_artificial_stress = _act("corrector_artificial_stress")
material_stress_rate_jaumann = _act("material_stress_rate_jaumann")
contmech_continuity = _act("contmech_continuity")
contmech_momentum = _act("contmech_momentum")
contmech_advection = _act("contmech_advection")
plasticity_johnson_cook = _act("plasticity_johnson_cook")
perform_boundary_conditions = _act("perform_boundary_conditions")
perform_boundary_conditions_thermal = _act("perform_boundary_conditions_thermal")
debug_invalidate = _act("debug_invalidate")


def material_fric_heat_gen_FEM_v2(p: particle_gpu, tool: tool_FEM):
    p.ctx.call("sphfem_material_fric_heat_gen_FEM_v2", C.byref(p.s), tool.h)


def compute_contact_forces_FEM_tool_v3(p: particle_gpu, tool: tool_FEM):
    p.ctx.call("sphfem_compute_contact_forces_FEM_tool_v3", C.byref(p.s), tool.h)


def wear_adjustment(tool: tool_FEM):
    tool.ctx.call("sphfem_wear_adjustment", tool.h)


def interactions_setup_geometry_constants(g: grid_base):
    g.ctx.call("sphfem_interactions_setup_geometry_constants", g.h)


def interactions_monaghan(p: particle_gpu, cell_start_dptr, cell_end_dptr, num_cell):
    p.ctx.call("sphfem_interactions_monaghan", C.byref(p.s), C.c_void_p(cell_start_dptr),
               C.c_void_p(cell_end_dptr), int(num_cell))


def interactions_monaghan_fp32c(lf: "leap_frog", p: particle_gpu, cell_start_dptr, cell_end_dptr, num_cell):
    """interactions_monaghan in the FP32-compensated form of the step's pair sums (lf in FP32C):
    v_der, pos_t and T_t as the reference writes them; stress divergence plus viscosity summed per
    axis in S_der = (ax, 0, ay, 0), vel_t = 0."""
    p.ctx.call("sphfem_interactions_monaghan_fp32c", lf.h, C.byref(p.s), C.c_void_p(cell_start_dptr),
               C.c_void_p(cell_end_dptr), int(num_cell))


def interactions_heat_pse(p: particle_gpu, cell_start_dptr, cell_end_dptr, num_cell):
    """interactions_gpu.cu:434-464: PSE heat conduction into T_t."""
    p.ctx.call("sphfem_interactions_heat_pse", C.byref(p.s), C.c_void_p(cell_start_dptr),
               C.c_void_p(cell_end_dptr), int(num_cell))


def report_force(p: particle_gpu):
    out = _lib.Double2()
    p.ctx.call("sphfem_report_force", C.byref(p.s), C.byref(out))
    return np.array([out.x, out.y])


# ---- vtk_writer.h:26-31, tool_wear.h, mesh_read_temperature: output formats
def _vtk_path(folder, stem, step):
    import os
    return os.path.join(folder, f"{stem}_{int(step):06d}.vtk")


def vtk_writer_write(obj, step: int, folder: str = "results", skip_blanked: bool = False, path: str | None = None):
    """vtk_writer_write(particles, step) -> results/vtk_out_%06d.vtk (vtk_writer.cu:11-219), or, for a
    rigid tool (host `tool` or its device copy), vtk_writer_write(tool, step) -> results/vtk_tool_%06d.vtk
    (vtk_writer.cu:222-302).  Returns the path written."""
    if isinstance(obj, particle_gpu):
        path = path or _vtk_path(folder, "vtk_out", step)
        obj.ctx.call("sphfem_vtk_write_particles", C.byref(obj.s), path.encode(), int(bool(skip_blanked)))
        return path
    t = obj.geometry() if isinstance(obj, rigid_tool_gpu) else obj
    path = path or _vtk_path(folder, "vtk_tool", step)
    check(lib().sphfem_vtk_write_rigid_tool(C.byref(t.g), path.encode()), None, "vtk_writer_write(tool)")
    return path


def vtk_writer_write_FEM_tool(tool: tool_FEM, step: int, folder: str = "results", path: str | None = None):
    """vtk_writer.cu:304-366 -> results/vtk_tool_%06d.vtk"""
    path = path or _vtk_path(folder, "vtk_tool", step)
    tool.ctx.call("sphfem_vtk_write_FEM_tool", tool.h, path.encode())
    return path


def vtk_writer_write_tool_org(tool: tool_FEM, step: int, tl_ref, folder: str = "results", path: str | None = None):
    """vtk_writer.cu:518-587 -> results/vtk_tool_org_%06d.vtk"""
    path = path or _vtk_path(folder, "vtk_tool_org", step)
    tool.ctx.call("sphfem_vtk_write_tool_org", tool.h, _d2(tl_ref), path.encode())
    return path


def vtk_writer_write_FEM_segments(tool: tool_FEM, step: int, folder: str = "results", path: str | None = None):
    """vtk_writer.cu:368-516 -> results/vtk_seg_%06d.vtk"""
    path = path or _vtk_path(folder, "vtk_seg", step)
    tool.ctx.call("sphfem_vtk_write_FEM_segments", tool.h, path.encode())
    return path


def vtk_writer_write_contact_variables(tool: tool_FEM, step: int, folder: str = "results",
                                       path: str | None = None):
    """vtk_writer.cu:589-714 -> results/vtk_phy_wear_%06d.vtk"""
    path = path or _vtk_path(folder, "vtk_phy_wear", step)
    tool.ctx.call("sphfem_vtk_write_contact_variables", tool.h, path.encode())
    return path


def read_vtk_temperature(path: str) -> np.ndarray:
    """mesh_read::mesh_read_temperature (mesh_structure_GPU.cu:1259-1296)."""
    n = C.c_long()
    check(lib().sphfem_read_vtk_temperature(path.encode(), None, 0, C.byref(n)), None, "read_vtk_temperature")
    out = np.zeros(n.value)
    check(lib().sphfem_read_vtk_temperature(path.encode(), out.ctypes.data_as(C.c_void_p), n.value, C.byref(n)),
          None, "read_vtk_temperature")
    return out


class tool_wear:
    """tool_wear(usui_K, usui_alpha, N, phys, tool_vel) (tool_wear.h:17-47): Usui wear of the particles
    in contact with a rigid tool, evaluated on the device."""

    def __init__(self, ctx: Context, usui_K: float, usui_alpha: float, tool_vel):
        self.ctx = ctx
        self.h = C.c_void_p()
        ctx.call("sphfem_tool_wear_create", float(usui_K), float(usui_alpha), _d2(tool_vel), C.byref(self.h))

    def eval_usui(self, p: particle_gpu, dt: float):
        """(min, max, avg) wear rate over the particles in contact; accumulates dt * avg."""
        mn, mx, av = C.c_double(), C.c_double(), C.c_double()
        self.ctx.call("sphfem_tool_wear_eval_usui", self.h, C.byref(p.s), float(dt), C.byref(mn), C.byref(mx),
                      C.byref(av))
        return mn.value, mx.value, av.value

    def get_accum_wear(self) -> float:
        v = C.c_double()
        check(lib().sphfem_tool_wear_get_accum_wear(self.h, C.byref(v)), None, "get_accum_wear")
        return v.value

    def destroy(self):
        if self.h:
            self.ctx.call("sphfem_tool_wear_destroy", self.h)
            self.h = C.c_void_p()


class mesh_read:
    """Tool mesh read from an .inp file (mesh_structure_GPU.h:80-142): mesh_read(file) for the
    Abaqus-style file (coordinates divided by 100, mesh_structure_GPU.cu:784-1257) or
    mesh_read(file, True) for a Gmsh export (:213-783).  Elements are coloured (cutting edge first
    from color_cutting_edge colours, then boundary layer 2 and the rest from color_total colours,
    first free colour counted from 1) and the CSR pattern is built.  Host code only: no device is
    needed until the tool is created."""

    color_total = 18              # mesh_structure_GPU.h:26-27
    color_cutting_edge = 6

    def __init__(self, file_name: str, gmsh_d: bool = False, color_total: int = 18, color_cutting_edge: int = 6):
        self.h = C.c_void_p()
        check(lib().sphfem_mesh_read_inp(str(file_name).encode(), int(bool(gmsh_d)), int(color_total),
                                         int(color_cutting_edge), C.byref(self.h)), None, f"mesh_read({file_name})")
        info = _lib.MeshInfo()
        check(lib().sphfem_mesh_get_info(self.h, C.byref(info)), None, "mesh_get_info")
        self.info = {f: getattr(info, f) for f, _ in _lib.MeshInfo._fields_}
        self.color_total, self.color_cutting_edge = color_total, color_cutting_edge

    def arrays(self) -> dict:
        i = self.info
        nn, ne = i["node_num"], i["ele_num"]
        out = dict(position=np.empty((nn, 2)), cut_node=np.empty(nn, np.int32), fixed_node=np.empty(nn, np.int32),
                   elem=np.empty((ne, 4), np.int32), cut_flags=np.empty((ne, 4), np.int32),
                   fixed_flags=np.empty((ne, 4), np.int32), cut_elems=np.empty(i["cut_ele_num"], np.int32),
                   fixed_elems=np.empty(i["fixed_ele_num"], np.int32), layers=np.empty((ne, 2), np.int32),
                   row_ptr=np.empty(nn + 1, np.int32), col_ind=np.empty(i["nz_num"], np.int32))
        check(lib().sphfem_mesh_get_arrays(self.h, *[a.ctypes.data_as(C.c_void_p) for a in out.values()]), None,
              "mesh_get_arrays")
        out["cut_flags"] = out["cut_flags"][:, :3].copy()
        out["fixed_flags"] = out["fixed_flags"][:, :3].copy()
        return out

    def mesh_read_temperature(self, file_name: str):
        """Initial nodal temperatures of a hot tool from a tool VTK file (mesh_structure_GPU.cu:1259-1296)."""
        T = read_vtk_temperature(file_name)
        self.set_temperature(T)

    def set_temperature(self, T):
        if T is None:
            check(lib().sphfem_mesh_set_temperature(self.h, None, 0), None, "mesh_set_temperature")
            return
        a = np.ascontiguousarray(T, dtype=np.float64)
        check(lib().sphfem_mesh_set_temperature(self.h, a.ctypes.data_as(C.c_void_p), len(a)), None,
              "mesh_set_temperature (the temperature count must equal the node count)")

    def write_mesh_file(self, file_name: str = "Tool_mesh.vtk"):
        """mesh_structure_GPU.cu:1395-1462"""
        check(lib().sphfem_mesh_write_vtk(self.h, str(file_name).encode()), None, "write_mesh_file")

    def tool_mesh(self, tc: dict, tl_point_ref, vel=(0.0, 0.0)) -> cases.ToolMesh:
        """set_up_segments_from_mesh + the FEM system of set_up_FEM_tool_gpu (tool_FEM.cu:771-1039,
        FEM_thermal_gpu.cu:584-659) as the tool description a workload carries."""
        d = _lib.ToolDesc()
        check(lib().sphfem_mesh_tool_desc(self.h, C.byref(_lib.ToolConstants(**tc)), _d2(tl_point_ref), _d2(vel),
                                          C.byref(d)), None, "set_up_segments_from_mesh")
        nn, ne, nz, ns = d.node_num, d.ele_num, d.nz_num, d.seg_num

        def get(ptr, n, dt, shape):
            cnt = n * int(np.prod(shape[1:], dtype=np.int64))
            buf = (C.c_char * (cnt * np.dtype(dt).itemsize)).from_address(ptr)
            return np.frombuffer(buf, dtype=dt, count=cnt).reshape(shape).copy()

        a = self.arrays()
        return cases.ToolMesh(
            nodes=get(d.position, nn, np.float64, (nn, 2)), elems=get(d.node_num_el, ne, np.int32, (ne, 4)),
            layers=a["layers"], cut_edge_node=a["cut_node"], fixed_edge_node=a["fixed_node"],
            row_ptr=get(d.csr_row_ptr, nn + 1, np.int32, (nn + 1,)), col_ind=get(d.csr_col_ind, nz, np.int32, (nz,)),
            K=get(d.K, nz, np.float64, (nz,)), M=get(d.M, nn, np.float64, (nn,)), f=get(d.f, nn, np.float64, (nn,)),
            T=get(d.T, nn, np.float64, (nn,)), seg_left=get(d.left, ns, np.float64, (ns, 2)),
            seg_right=get(d.right, ns, np.float64, (ns, 2)), seg_n=get(d.n, ns, np.float64, (ns, 2)),
            seg_marks=get(d.marks, ns, np.int32, (ns, 4)), n_colors=d.color_total, n_cut_colors=d.color_cutting_edge,
            tl=np.array([d.tl_point_ref.x, d.tl_point_ref.y]), br=np.array([d.br_point.x, d.br_point.y]),
            ref_dist=d.ref_dist, lowest_point=d.lowest_point, right_point=d.right_point)

    def close(self):
        if self.h:
            check(lib().sphfem_mesh_destroy(self.h), None, "mesh_destroy")
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Simulation:
    """A workload on the device: constants, particles, locked grid, FEM tool and stepper."""

    def __init__(self, case: cases.Case, precision=_lib.FP64, device=0, rr2_min=None, wear=None, ctx=None,
                 grid="green", max_cell=None):
        """grid: "green" or "rothlin"; max_cell: an unlocked grid of that capacity (geometry from the
        particles each step) instead of the case's locked grid."""
        self.case = case
        self.ctx = ctx or Context(device)
        rr2 = cases.default_rr2_min(case.dx) if rr2_min is None else rr2_min
        self.ctx.setup_constants(case.const, case.dt, rr2, case.wear if wear is None else wear)
        init = cases.initial_particles(case)
        self.p = particle_gpu(self.ctx, case.N, **{k: init[k] for k in ("pos", "vel", "rho", "T", "h", "fixed", "S",
                                                                         "tool_particle")})
        max_cell = case.max_cell if max_cell is None else max_cell
        if max_cell:
            self.grid = GRIDS[grid].unlocked(self.ctx, max_cell, case.N)
        else:
            self.grid = GRIDS[grid](self.ctx, case.N, case.grid_bbmin, case.grid_bbmax, case.h)
        self.tool = tool_FEM(self.ctx, case) if case.tool is not None else None
        # a rigid tool (benchmarks.cpp:409-415, 518): device copy and the pair sums' tool block
        self.rtool = None
        if case.rigid is not None:
            interactions_setup_thermal_constants_tool(self.ctx, case.rigid.trml_itool)
            self.rtool = tool_gpu_set_up_tool(self.ctx, tool(case.rigid.geometry), case.rigid.contact_alpha,
                                              case.const.phys)
        self.stepper = leap_frog(self.ctx, case.N, self.grid.num_cell,
                                 self.tool.n_nodes if self.tool else 0, precision)
        self.p.stepper = self.stepper      # the rates of an FP32C stepper live in FP32 inside it

    def step(self, n: int = 1):
        for _ in range(n):
            if self.rtool is not None:
                self.stepper.step_rigid(self.p, self.grid, self.rtool)
            else:
                self.stepper.step(self.p, self.grid, self.tool)

    def capture(self, steps: int = 2) -> StepGraph:
        """Graph of `steps` steps (one eager step must have run)."""
        if self.rtool is not None:
            return self.stepper.capture_rigid(self.p, self.grid, self.rtool, steps)
        return self.stepper.capture(self.p, self.grid, self.tool, steps)

    def set_tuning(self, **kw):
        self.stepper.set_tuning(**{**self.stepper.get_tuning(), **kw})

    def get_tuning(self) -> dict:
        return self.stepper.get_tuning()

    def synchronize(self):
        self.ctx.synchronize()

    def state(self, fields=None):
        return self.p.state(fields)

    def tool_state(self):
        return self.tool.download() if self.tool else None

    def close(self):
        """Release the device objects (stepper, tool, grid, particles) and the context."""
        self.ctx.synchronize()
        self.ctx.call("sphfem_leap_frog_destroy", self.stepper.h)
        if self.tool is not None:
            self.ctx.call("sphfem_tool_destroy", self.tool.h)
        if self.rtool is not None:
            self.rtool.destroy()
        self.ctx.call("sphfem_grid_destroy", self.grid.h)
        self.p.free()
        self.ctx.close()


class BatchedSimulation:
    """Independent cases advanced together, each with its own context (HIP stream + constants).
    `capture` records `graph_steps` steps of every case into ONE HIP graph (the cases are parallel
    branches of it, sphfem_batch_graph_create), so one `launch` advances the whole sweep;
    `capture(one_graph=False)` records one graph per case instead (queued case by case).
    `launch_sequential` runs the cases one after another (the baseline a per-case loop would give)."""

    def __init__(self, case_list, precision=_lib.FP64, device=0, graph_steps=2):
        self.cases = list(case_list)
        self.sims = [Simulation(c, precision=precision, device=device) for c in self.cases]
        self.graph_steps = graph_steps
        self.graph = None          # the batch graph
        self.graphs = None         # per-case graphs (one_graph=False, and launch_sequential)
        self.total_particles = sum(c.N for c in self.cases)

    def step(self, n: int = 1):
        """Eager steps, queued case by case (concurrent across the cases' streams)."""
        for _ in range(n):
            for s in self.sims:
                s.step(1)

    def capture(self, one_graph: bool = True):
        """Record the graph(s); every case must have run an eager step."""
        if one_graph:
            n = len(self.sims)
            arr = lambda vals: (C.c_void_p * n)(*vals)
            ctxs = arr([s.ctx.h.value for s in self.sims])
            lfs = arr([s.stepper.h.value for s in self.sims])
            ps = arr([C.addressof(s.p.s) for s in self.sims])
            gs = arr([s.grid.h.value for s in self.sims])
            ts = arr([s.tool.h.value if s.tool else None for s in self.sims])
            gh = C.c_void_p()
            check(lib().sphfem_batch_graph_create(ctxs, lfs, ps, gs, ts, n, int(self.graph_steps), C.byref(gh)),
                  self.sims[0].ctx.h, "sphfem_batch_graph_create")
            self.graph = StepGraph(self.sims[0].ctx, gh, self.graph_steps)
        else:
            self.graphs = [s.capture(self.graph_steps) for s in self.sims]

    def launch(self, repeats: int = 1):
        """Advance every case by repeats * graph_steps steps, all cases in flight together."""
        if self.graph is not None:
            self.graph.launch(repeats)
            return
        for _ in range(repeats):
            for g in self.graphs:
                g.launch(1)

    def launch_sequential(self, repeats: int = 1):
        if self.graphs is None:
            self.graphs = [s.capture(self.graph_steps) for s in self.sims]
        for s, g in zip(self.sims, self.graphs):
            g.launch(repeats)
            s.ctx.synchronize()

    def synchronize(self):
        for s in self.sims:
            s.ctx.synchronize()

    def check_error(self):
        for s in self.sims:
            s.ctx.check_error()

    def release_graphs(self):
        if self.graph is not None:
            self.graph.destroy()
            self.graph = None
        if self.graphs:
            for g in self.graphs:
                g.destroy()
            self.graphs = None

    def close(self):
        self.release_graphs()
        for s in self.sims:
            s.close()
