/*
 * sphfem.hpp -- C++ host mirror of the reference's per-step interface over the C ABI.
 *
 * Names and call shapes follow the reference so that a driver written against
 * particle_gpu.h, grid.h / grid_gpu_green.h, leap_frog.h, tool_FEM.h, actions_gpu.h,
 * interactions_gpu.h and output.h compiles against this header instead:
 *
 *   particle_gpu        particle_gpu.h:23-83 (same field names; the six constructors)
 *   grid_base / grid_gpu_green / grid_gpu_rothlin
 *                       grid.h:23-64, grid_gpu_green.h:84-85, grid_gpu_rothlin.h:79-80
 *   leap_frog           leap_frog.h:31-50 (step(particles, grid[, tool_FEM]))
 *   tool / tool_FEM     tool.h:31-141 (rigid tool: constructors, geometry), tool_gpu.h:19-28,
 *                       tool_FEM.h:34-83 (per-step methods)
 *   free functions      actions_gpu.h:33-72, interactions_gpu.h:27-43, output.h:17,
 *                       vtk_writer.h:26-31, mesh_read_temperature
 *   tool_wear           tool_wear.h:17-47
 *   mesh_read           mesh_structure_GPU.h:80-142 (.inp reader, colouring, CSR, write_mesh_file)
 *   remesh_low_level, remesh_gmsh, tool_FEM::read_remesh_gmsh / remesh_read_segments
 *                       actions_gpu.cu:1022-1038, tool_FEM.cu:1708-1828
 *   global_dt, wear_pro benchmarks.cpp:12, mfree_gpu.cpp:47
 *
 * Differences a caller sees: errors throw std::runtime_error (the reference prints and
 * continues); objects release their device memory in destructors; tool_FEM is built from a
 * mesh_read (an .inp file) or a host mesh description (sphfem_tool_desc); all work is queued on one
 * process-wide context (device 0 unless sphfem_shim::use_device is called first).
 * Header-only, C++17; link with libsphfem_hip.so.
 */
#ifndef SPHFEM_HPP
#define SPHFEM_HPP

#include <sys/stat.h>

#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../sphfem.h"

typedef sphfem_double2 float2_t;
typedef sphfem_double4 float4_t;

struct vec2_t {
    double x = 0., y = 0.;
    vec2_t() = default;
    vec2_t(double a, double b) : x(a), y(b) {}
};

typedef sphfem_phys_constants phys_constants;
typedef sphfem_trml_constants trml_constants;
typedef sphfem_corr_constants corr_constants;
typedef sphfem_joco_constants joco_constants;
typedef sphfem_geom_constants geom_constants;
typedef sphfem_tool_constants tool_constants;
typedef sphfem_wear_constants wear_constants;

namespace sphfem_shim {

inline int &device_slot() {
    static int d = 0;
    return d;
}

inline sphfem_ctx *&ctx_slot() {
    static sphfem_ctx *c = nullptr;
    return c;
}

// the device tool of the last tool_gpu_set_up_tool (the reference's __constant__ tool state,
// tool_gpu.cu:16-22), read by compute_contact_forces
inline sphfem_rigid_tool *&rigid_slot() {
    static sphfem_rigid_tool *t = nullptr;
    return t;
}

inline void check(int rc, const char *what) {
    if (rc != SPHFEM_OK) {
        const char *msg = ctx_slot() ? sphfem_last_error(ctx_slot()) : "no context";
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + msg);
    }
}

// the process-wide context, created on first use
inline sphfem_ctx *ctx() {
    if (!ctx_slot()) check(sphfem_create(device_slot(), &ctx_slot()), "sphfem_create");
    return ctx_slot();
}

inline void use_device(int device) {
    if (ctx_slot()) throw std::runtime_error("sphfem_shim::use_device after the context was created");
    device_slot() = device;
}

template <class T>
inline void upload(T *dst, const T *src, unsigned int n) {
    if (src && n) check(sphfem_memcpy_htod(ctx(), dst, src, sizeof(T) * (size_t)n), "sphfem_memcpy_htod");
}

}  // namespace sphfem_shim

// benchmarks.cpp:12 and mfree_gpu.cpp:47: read by leap_frog::step on every call
inline double global_dt = 0.;
inline double global_t_final = 0.;       // benchmarks.cpp:13
inline int global_step = 0;              // mfree_gpu.cpp:39
inline bool wear_pro = false;

// ---------------------------------------------------------------------------------------------
// particle_gpu.h:23-83 -- the device arrays are the fields of sphfem_particles
// ---------------------------------------------------------------------------------------------
struct particle_gpu : sphfem_particles {
    explicit particle_gpu(unsigned int n) {
        sphfem_shim::check(sphfem_particles_alloc(sphfem_shim::ctx(), n, this), "particle_gpu(N)");
    }
    particle_gpu(const float2_t *pos_h, const float2_t *vel_init, const double *rho_h, const double *h_h,
                 unsigned int n)
        : particle_gpu(n) {
        init(pos_h, vel_init, rho_h, nullptr, h_h, nullptr, nullptr, nullptr);
    }
    particle_gpu(const float2_t *pos_h, const float2_t *vel_init, const double *rho_h, const double *h_h,
                 const double *fixed_h, unsigned int n)
        : particle_gpu(n) {
        init(pos_h, vel_init, rho_h, nullptr, h_h, fixed_h, nullptr, nullptr);
    }
    particle_gpu(const float2_t *pos_h, const float2_t *vel_init, const double *rho_h, const double *T_init,
                 const double *h_h, const double *fixed_h, unsigned int n)
        : particle_gpu(n) {
        init(pos_h, vel_init, rho_h, T_init, h_h, fixed_h, nullptr, nullptr);
    }
    particle_gpu(const float2_t *pos_h, const float2_t *vel_init, const double *rho_h, const double *T_init,
                 const double *h_h, const double *fixed_h, const double *tool_p, unsigned int n)
        : particle_gpu(n) {
        init(pos_h, vel_init, rho_h, T_init, h_h, fixed_h, tool_p, nullptr);
    }
    particle_gpu(const float2_t *pos_h, const float2_t *vel_init, const double *rho_h, const double *h_h,
                 const float4_t *S_init, unsigned int n)
        : particle_gpu(n) {
        init(pos_h, vel_init, rho_h, nullptr, h_h, nullptr, nullptr, S_init);
    }
    ~particle_gpu() {
        if (sphfem_shim::ctx_slot()) (void)sphfem_particles_free(sphfem_shim::ctx_slot(), this);
    }
    particle_gpu(const particle_gpu &) = delete;
    particle_gpu &operator=(const particle_gpu &) = delete;

  private:
    void init(const float2_t *pos_h, const float2_t *vel_h, const double *rho_h, const double *T_h,
              const double *h_h, const double *fixed_h, const double *tool_h, const float4_t *S_h) {
        using sphfem_shim::upload;
        upload(pos, pos_h, N);
        upload(vel, vel_h, N);
        upload(rho, rho_h, N);
        upload(T, T_h, N);
        upload(h, h_h, N);
        upload(fixed, fixed_h, N);
        upload(tool_particle, tool_h, N);
        upload(S, S_h, N);
    }
};

// ---------------------------------------------------------------------------------------------
// tool.h:31-141 (rigid tool: host geometry + the device copy of tool_gpu_set_up_tool) /
// tool_FEM.h:30-83
// ---------------------------------------------------------------------------------------------
typedef sphfem_bbox bbox;

class tool {
  public:
    tool() = default;
    // tool.cpp:316-331
    tool(vec2_t tl, vec2_t tr, vec2_t br, vec2_t bl, double r, double mu_fric) {
        sphfem_shim::check(sphfem_tool_geometry_from_points(d2(tl), d2(tr), d2(br), d2(bl), r, mu_fric, &m_geom),
                           "tool");
    }
    tool(vec2_t tl, vec2_t tr, vec2_t br, vec2_t bl, double mu_fric) : tool(tl, tr, br, bl, 0., mu_fric) {}
    // tool.cpp:335-406 (angles in degrees)
    tool(vec2_t tl, double length, double height, double rake_angle, double clearance_angle, double r,
         double mu_fric) {
        sphfem_shim::check(sphfem_tool_geometry_from_angles(d2(tl), length, height, rake_angle, clearance_angle, r,
                                                            mu_fric, &m_geom),
                           "tool");
    }
    tool(vec2_t tl, double length, double height, double rake_angle, double clearance_angle, double mu_fric)
        : tool(tl, length, height, rake_angle, clearance_angle, 0., mu_fric) {}
    virtual ~tool() {
        if (m_gpu && sphfem_shim::ctx_slot()) {
            if (sphfem_shim::rigid_slot() == m_gpu) sphfem_shim::rigid_slot() = nullptr;
            (void)sphfem_rigid_tool_destroy(sphfem_shim::ctx_slot(), m_gpu);
        }
    }
    tool(const tool &) = delete;
    tool &operator=(const tool &) = delete;

    bbox safe_bb(double safety = 0.011) const {
        bbox b{};
        sphfem_shim::check(sphfem_tool_geometry_safe_bb(&m_geom, safety, &b), "tool::safe_bb");
        return b;
    }
    double low() const {
        double v = 0.;
        sphfem_shim::check(sphfem_tool_geometry_low(&m_geom, &v), "tool::low");
        return v;
    }
    double mu() const { return m_geom.mu; }
    double front() const { return m_geom.front; }
    // tool.cpp:202-212 (the host copy; the device copy moves in the step / tool_gpu_update_tool)
    void update_tool(double dt) { sphfem_shim::check(sphfem_tool_geometry_move(&m_geom, dt), "tool::update_tool"); }
    virtual void set_vel(vec2_t vel) {
        m_velocity = vel;
        m_geom.vel = d2(vel);
        if (m_gpu)
            sphfem_shim::check(sphfem_rigid_tool_set_vel(sphfem_shim::ctx(), m_gpu, m_geom.vel), "tool::set_vel");
    }
    vec2_t get_vel() const { return m_velocity; }
    void set_thermal(bool thermal) { m_thermal = thermal; }
    bool thermal() const { return m_thermal; }
    bool inside(float2_t pos) const {
        int in = 0;
        sphfem_shim::check(sphfem_tool_geometry_inside(&m_geom, pos, &in), "tool::inside");
        return in != 0;
    }
    // tool.cpp:112-135 (T_init, rho and h are not used there either)
    std::vector<float2_t> sample_tool(double dx, double, double, double) const {
        long n = 0;
        sphfem_shim::check(sphfem_tool_geometry_sample(&m_geom, dx, nullptr, 0, &n), "tool::sample_tool");
        std::vector<float2_t> out((size_t)n);
        sphfem_shim::check(sphfem_tool_geometry_sample(&m_geom, dx, out.data(), n, &n), "tool::sample_tool");
        return out;
    }
    const sphfem_tool_geometry &geometry() const { return m_geom; }
    sphfem_rigid_tool *gpu_handle() const { return m_gpu; }

  protected:
    static sphfem_double2 d2(vec2_t v) { return sphfem_double2{v.x, v.y}; }
    vec2_t m_velocity;
    sphfem_tool_geometry m_geom{};
    bool m_thermal = false;
    sphfem_rigid_tool *m_gpu = nullptr;
    friend void tool_gpu_set_up_tool(tool *t, double alpha, phys_constants phys);
};

// mfree_gpu.cpp:40: the rigid tool read by leap_frog::step(particles, grid)
inline tool *global_tool = nullptr;

// tool_gpu.h:19-28
inline void tool_gpu_set_up_tool(tool *t, double alpha, phys_constants phys) {
    sphfem_ctx *c = sphfem_shim::ctx();
    if (t->m_gpu) {
        if (sphfem_shim::rigid_slot() == t->m_gpu) sphfem_shim::rigid_slot() = nullptr;
        sphfem_shim::check(sphfem_rigid_tool_destroy(c, t->m_gpu), "tool_gpu_set_up_tool");
        t->m_gpu = nullptr;
    }
    sphfem_shim::check(sphfem_tool_gpu_set_up_tool(c, &t->m_geom, alpha, &phys, &t->m_gpu), "tool_gpu_set_up_tool");
    sphfem_shim::rigid_slot() = t->m_gpu;
}
inline void tool_gpu_update_tool(tool *t, particle_gpu *particles) {
    if (!t->gpu_handle()) throw std::runtime_error("tool_gpu_update_tool: tool_gpu_set_up_tool was not called");
    sphfem_ctx *c = sphfem_shim::ctx();
    sphfem_shim::check(sphfem_set_dt(c, global_dt), "set_dt");
    t->update_tool(global_dt);
    sphfem_shim::check(sphfem_tool_gpu_update_tool(c, t->gpu_handle(), particles), "tool_gpu_update_tool");
}
inline void compute_contact_forces(particle_gpu *particles) {
    if (!sphfem_shim::rigid_slot()) throw std::runtime_error("compute_contact_forces: no tool set up");
    sphfem_ctx *c = sphfem_shim::ctx();
    sphfem_shim::check(sphfem_set_dt(c, global_dt), "set_dt");
    sphfem_shim::check(sphfem_compute_contact_forces(c, particles, sphfem_shim::rigid_slot()),
                       "compute_contact_forces");
}

// Views of a tool_FEM's device arrays under the reference's names (segment_FEM_gpu.h:31-56,
// FEM_thermal_gpu.h:35-93, mesh_structure_GPU.h:80-142): plain device pointers owned by the tool,
// refreshed when the tool is remeshed.  Kernels and writers that read these fields in the
// reference read the same data here.
struct segment_FEM_gpu {
    int segment_num = 0;
    float2_t *left = nullptr, *right = nullptr, *n = nullptr;
    sphfem_int4 *marks = nullptr;          // element, colour, left node, right node
    float2_t *heat_exchange = nullptr;     // x: q_T * length, y: length
    double *fric_heat = nullptr;
    float2_t *wear_rate = nullptr;
    double *wear_nodes = nullptr;
    float4_t *physical_para = nullptr;     // contact count, force, sliding velocity, temperature
    double *sliding_force = nullptr;
};

struct FEM_thermal_2D_GPU {
    int node_num = 0, ele_num = 0, nz_num = 0, color_total = 0, color_cutting_edge = 0, fixed_ele_num = 0;
    double T_env = 300.;
    int *m_row_ptr = nullptr, *m_col_ind = nullptr;
    double *m_stiffness_matrix = nullptr, *m_stiffness_matrix_temp = nullptr;
    double *m_mass_matrix = nullptr, *m_mass_matrix_inverse = nullptr;
    double *m_f = nullptr, *m_f_temp = nullptr, *m_k = nullptr, *m_T = nullptr, *m_T_old = nullptr;
    double *m_T_mapping = nullptr;
};

// mesh_structure_GPU.h:37-63: host copies of a mesh read from an .inp file (mesh_read fills them;
// int3 flags of the reference are the x, y, z of an int4 here, w unused)
struct elements_read {
    std::vector<sphfem_int4> node_num;                 // 3 node ids + colour
    std::vector<sphfem_int4> cutting_boundary_flags, fixed_boundary_flags;
    std::vector<int> cutting_boundary_elements, fixed_boundary_elements;
    std::vector<int> boundary_layers;                  // 2 per element: layer 1, layer 2
};
struct nodes_read {
    std::vector<double> x_init, y_init;
    std::vector<int> cutting_edge_flag, fixed_edge_flag, row_ptr, col_ind;
    const sphfem_mesh *source = nullptr;               // the mesh object the arrays were read from
};

struct mesh_GPU_m {
    mesh_GPU_m() = default;
    // mesh_structure_GPU.h:135 (benchmarks_NZ.cpp:241): names the mesh whose device copy the tool
    // makes in tool_FEM::set_up_FEM_tool_gpu; the views below are filled then
    mesh_GPU_m(const nodes_read &nodes, const elements_read &elements, int color_number, int cut_edge_col_no,
               bool /*ini*/)
        : no_num((int)nodes.x_init.size()), ele_num_total((int)elements.node_num.size()),
          non_zeros((int)nodes.col_ind.size()), color_total(color_number), cutting_edge_color(cut_edge_col_no),
          ele_num_fixed((int)elements.fixed_boundary_elements.size()), source(nodes.source) {}
    sphfem_int4 *node_num = nullptr;       // 3 node ids + colour per element
    float2_t *position = nullptr;
    float2_t *position_mapping = nullptr;
    int *csr_row_ptr = nullptr, *csr_col_ind = nullptr;
    int *node_flags = nullptr;             // bit 0 cutting edge, bit 1 fixed boundary
    int get_node_number() const { return no_num; }
    int get_element_number() const { return ele_num_total; }
    int get_non_zero_number() const { return non_zeros; }
    int get_color_number() const { return color_total; }
    int get_cutting_edge_color_number() const { return cutting_edge_color; }
    int get_fixed_edge_element_number() const { return ele_num_fixed; }
    int no_num = 0, ele_num_total = 0, non_zeros = 0, color_total = 0, cutting_edge_color = 0, ele_num_fixed = 0;
    const sphfem_mesh *source = nullptr;
};

// mesh_read (mesh_structure_GPU.h:80-142): a tool mesh read from an .inp file on the host --
// mesh_read(file) for the Abaqus-style file, mesh_read(file, true) for a Gmsh export -- with its
// element colours and CSR pattern
class mesh_read {
  public:
    explicit mesh_read(const std::string &file_name, bool gmsh_d = false, int color_total_ = 18,
                       int color_cutting_edge_ = 6)
        : color_total(color_total_), color_cutting_edge(color_cutting_edge_) {
        sphfem_shim::check(sphfem_mesh_read_inp(file_name.c_str(), gmsh_d ? 1 : 0, color_total, color_cutting_edge,
                                                &m_handle),
                           ("mesh_read(" + file_name + ")").c_str());
        sphfem_shim::check(sphfem_mesh_get_info(m_handle, &info), "mesh_read info");
        fill_host_copies();
    }
    ~mesh_read() {
        if (m_handle) (void)sphfem_mesh_destroy(m_handle);
    }
    mesh_read(const mesh_read &) = delete;
    mesh_read &operator=(const mesh_read &) = delete;

    int color_total, color_cutting_edge;
    sphfem_mesh_info info{};
    nodes_read nodes;              // mesh_structure_GPU.h:82-83
    elements_read elements;

    // mesh_structure_GPU.cu:1259-1296: hot-tool nodal temperatures from a tool VTK file
    void mesh_read_temperature(const std::string &file_name) {
        long n = 0;
        sphfem_shim::check(sphfem_read_vtk_temperature(file_name.c_str(), nullptr, 0, &n), "mesh_read_temperature");
        std::vector<double> T((size_t)n);
        sphfem_shim::check(sphfem_read_vtk_temperature(file_name.c_str(), T.data(), n, &n), "mesh_read_temperature");
        sphfem_shim::check(sphfem_mesh_set_temperature(m_handle, T.data(), n),
                           "mesh_read_temperature (node count differs)");
    }
    // mesh_structure_GPU.cu:1395-1462
    void write_mesh_file(const std::string &file_name = "Tool_mesh.vtk") const {
        sphfem_shim::check(sphfem_mesh_write_vtk(m_handle, file_name.c_str()), "write_mesh_file");
    }
    sphfem_mesh *handle() const { return m_handle; }

  private:
    void fill_host_copies() {
        const size_t nn = (size_t)info.node_num, ne = (size_t)info.ele_num;
        std::vector<sphfem_double2> pos(nn);
        nodes.cutting_edge_flag.resize(nn);
        nodes.fixed_edge_flag.resize(nn);
        nodes.row_ptr.resize(nn + 1);
        nodes.col_ind.resize((size_t)info.nz_num);
        elements.node_num.resize(ne);
        elements.cutting_boundary_flags.resize(ne);
        elements.fixed_boundary_flags.resize(ne);
        elements.cutting_boundary_elements.resize((size_t)info.cut_ele_num);
        elements.fixed_boundary_elements.resize((size_t)info.fixed_ele_num);
        elements.boundary_layers.resize(2 * ne);
        sphfem_shim::check(sphfem_mesh_get_arrays(m_handle, pos.data(), nodes.cutting_edge_flag.data(),
                                                  nodes.fixed_edge_flag.data(), elements.node_num.data(),
                                                  elements.cutting_boundary_flags.data(),
                                                  elements.fixed_boundary_flags.data(),
                                                  elements.cutting_boundary_elements.data(),
                                                  elements.fixed_boundary_elements.data(),
                                                  elements.boundary_layers.data(), nodes.row_ptr.data(),
                                                  nodes.col_ind.data()),
                           "mesh_read arrays");
        nodes.x_init.resize(nn);
        nodes.y_init.resize(nn);
        for (size_t i = 0; i < nn; i++) {
            nodes.x_init[i] = pos[i].x;
            nodes.y_init[i] = pos[i].y;
        }
        nodes.source = m_handle;
    }
    sphfem_mesh *m_handle = nullptr;
};

class tool_FEM : public tool {
  public:
    // tool_FEM.cu:623-661, 663-695 (angles in degrees): the tool body's outline and friction
    // coefficient; tl orients the cutting segments.  The mesh is attached afterwards as
    // benchmarks_NZ.cpp:228-245 does: mesh_CPU = new mesh_read(file); mesh_GPU = new
    // mesh_GPU_m(mesh_CPU->nodes, mesh_CPU->elements, ...); set_up_segments_from_mesh();
    // set_up_FEM_tool_gpu(tc)
    tool_FEM(vec2_t tl, double length, double height, double rake_angle, double clearance_angle, double r,
             double mu_fric)
        : tool(tl, length, height, rake_angle, clearance_angle, r, mu_fric), tl_point_ref(tl) {}
    tool_FEM(vec2_t tl, double length, double height, double rake_angle, double clearance_angle, double mu_fric)
        : tool(tl, length, height, rake_angle, clearance_angle, mu_fric), tl_point_ref(tl) {}
    // the same four steps in one call: segments oriented towards tl_point_ref, FEM system with the
    // tool constants tc
    tool_FEM(mesh_read &mesh, const sphfem_tool_constants &tc, vec2_t tl_point_ref_, vec2_t vel)
        : tl_point_ref(tl_point_ref_) {
        m_velocity = vel;
        mesh_CPU = &mesh;
        set_up_FEM_tool_gpu(tc);
    }
    // the device data of set_up_segments_from_mesh + set_up_FEM_tool_gpu, from a host mesh
    explicit tool_FEM(const sphfem_tool_desc &desc) { create(desc); }
    ~tool_FEM() override {
        if (m_handle && sphfem_shim::ctx_slot()) (void)sphfem_tool_destroy(sphfem_shim::ctx_slot(), m_handle);
    }
    tool_FEM(const tool_FEM &) = delete;
    tool_FEM &operator=(const tool_FEM &) = delete;

    int node_number = 0;
    int element_number = 0;
    int cutting_segment_size = 0;
    vec2_t tl_point_ref, br_point;                  // tool_FEM.h:38-48, 54-55: host copies, moved by
    double ref_dist = 0.;                           // tool_update_FEM_gpu like the device's
    double right_point = 0., lowest_point = 0.;
    segment_FEM_gpu *segments_FEM = &m_seg;         // tool_FEM.h:34-47 members
    FEM_thermal_2D_GPU *FEM_solver = &m_fem;
    mesh_GPU_m *mesh_GPU = &m_mesh;
    mesh_read *mesh_CPU = nullptr;                  // set by the caller before set_up_segments_from_mesh

    // tool_FEM.cu:771-1039: cutting segments of mesh_CPU oriented towards tl_point_ref and chained
    // from the highest point; sets the extent (tl_point_ref, br_point, ref_dist, lowest_point,
    // right_point) and the counts.  The device arrays are made by set_up_FEM_tool_gpu.
    void set_up_segments_from_mesh() {
        const sphfem_tool_constants none{};
        m_orient = tl_point_ref;
        m_oriented = true;
        set_up_host(describe(none));
    }
    // tool_FEM.cu:1539-1544: the tool's device data (segments as set up above, FEM system with the
    // tool constants tc, flank retreat 0); a tool set up before is replaced
    void set_up_FEM_tool_gpu(const sphfem_tool_constants &tc) { create(describe(tc)); }

    void set_vel(vec2_t vel) override {
        m_velocity = vel;
        if (m_handle)
            sphfem_shim::check(sphfem_tool_set_vel(sphfem_shim::ctx(), m_handle, sphfem_double2{vel.x, vel.y}),
                               "tool_FEM::set_vel");
    }
    void apply_heat_boundary_condition() {
        sphfem_shim::check(sphfem_tool_apply_heat_boundary_condition(sphfem_shim::ctx(), m_handle),
                           "apply_heat_boundary_condition");
    }
    void tool_update_FEM_gpu(double dt) {
        sphfem_shim::check(sphfem_tool_update_FEM_gpu(sphfem_shim::ctx(), handle_or_throw(), dt),
                           "tool_update_FEM_gpu");
        // tool_FEM.cu:1551-1558
        tl_point_ref = vec2_t(tl_point_ref.x + m_velocity.x * dt, tl_point_ref.y + m_velocity.y * dt);
        br_point = vec2_t(br_point.x + m_velocity.x * dt, br_point.y + m_velocity.y * dt);
        lowest_point += m_velocity.y * dt;
        right_point += m_velocity.x * dt;
    }
    void segments_update(double dt) {
        sphfem_shim::check(sphfem_tool_segments_update(sphfem_shim::ctx(), m_handle, dt), "segments_update");
    }
    void segments_wear_rate_reset() {
        sphfem_shim::check(sphfem_tool_segments_wear_rate_reset(sphfem_shim::ctx(), m_handle),
                           "segments_wear_rate_reset");
    }
    sphfem_tool_kinematics kinematics() const {
        sphfem_tool_kinematics k{};
        sphfem_shim::check(sphfem_tool_get_kinematics(sphfem_shim::ctx(), m_handle, &k), "tool kinematics");
        return k;
    }
    sphfem_tool *handle() const { return m_handle; }

    // read_remesh_gmsh (tool_FEM.cu:1811-1828): the tool takes the mesh of a Gmsh export (by default
    // the file the reference's remesh writes), temperatures interpolated from the current mesh;
    // graphs captured before must be captured again
    void read_remesh_gmsh(const std::string &file_name = "tool_gmsh_remeshed.inp") {
        mesh_read mesh(file_name, true);
        mesh.write_mesh_file();
        remesh(mesh);
    }
    void remesh(mesh_read &mesh) {
        sphfem_shim::check(sphfem_tool_remesh(sphfem_shim::ctx(), m_handle, mesh.handle()), "tool_FEM::remesh");
        sphfem_shim::check(sphfem_tool_get_sizes(m_handle, &node_number, &element_number, nullptr,
                                                 &cutting_segment_size),
                           "tool sizes");
        const sphfem_tool_kinematics k = kinematics();
        right_point = k.right_point;
        lowest_point = k.lowest_point;
        refresh_views();
    }
    // remesh_read_segments (tool_FEM.cu:1708-1810): the Gmsh geometry of the current cutting edge
    void remesh_read_segments(const std::string &geo_file = "tool_remesh.geo") {
        sphfem_shim::check(sphfem_tool_write_remesh_geo(sphfem_shim::ctx(), m_handle, geo_file.c_str()),
                           "remesh_read_segments");
    }

  private:
    sphfem_tool *handle_or_throw() const {
        if (!m_handle) throw std::runtime_error("tool_FEM: set_up_FEM_tool_gpu was not called");
        return m_handle;
    }
    sphfem_tool_desc describe(const sphfem_tool_constants &tc) {
        if (!mesh_CPU) throw std::runtime_error("tool_FEM: mesh_CPU is not set");
        if (mesh_GPU != &m_mesh && mesh_GPU && mesh_GPU->source && mesh_GPU->source != mesh_CPU->handle())
            throw std::runtime_error("tool_FEM: mesh_GPU was made from another mesh than mesh_CPU");
        // the segments are oriented by tl_point_ref as it was when set_up_segments_from_mesh ran
        // (that call replaces tl_point_ref by the mesh's own corner)
        const vec2_t o = m_oriented ? m_orient : tl_point_ref;
        sphfem_tool_desc desc{};
        sphfem_shim::check(sphfem_mesh_tool_desc(mesh_CPU->handle(), &tc, sphfem_double2{o.x, o.y},
                                                 sphfem_double2{m_velocity.x, m_velocity.y}, &desc),
                           "set_up_segments_from_mesh");
        return desc;
    }
    void set_up_host(const sphfem_tool_desc &desc) {
        node_number = desc.node_num;
        element_number = desc.ele_num;
        cutting_segment_size = desc.seg_num;
        m_velocity = vec2_t(desc.vel.x, desc.vel.y);
        tl_point_ref = vec2_t(desc.tl_point_ref.x, desc.tl_point_ref.y);
        br_point = vec2_t(desc.br_point.x, desc.br_point.y);
        ref_dist = desc.ref_dist;
        right_point = desc.right_point;
        lowest_point = desc.lowest_point;
    }
    void create(const sphfem_tool_desc &desc) {
        sphfem_tool *made = nullptr;
        sphfem_shim::check(sphfem_tool_create(sphfem_shim::ctx(), &desc, &made), "tool_FEM");
        if (m_handle) (void)sphfem_tool_destroy(sphfem_shim::ctx(), m_handle);
        m_handle = made;
        set_up_host(desc);
        refresh_views();
    }
    void refresh_views() {
        sphfem_tool_arrays a{};
        sphfem_shim::check(sphfem_tool_get_arrays(m_handle, &a), "tool arrays");
        m_seg.segment_num = a.seg_num;
        m_seg.left = a.left; m_seg.right = a.right; m_seg.n = a.n; m_seg.marks = a.marks;
        m_seg.heat_exchange = a.heat_exchange; m_seg.fric_heat = a.fric_heat; m_seg.wear_rate = a.wear_rate;
        m_seg.wear_nodes = a.wear_nodes; m_seg.physical_para = a.physical_para; m_seg.sliding_force = a.sliding_force;
        m_fem.node_num = a.node_num; m_fem.ele_num = a.ele_num; m_fem.nz_num = a.nz_num;
        m_fem.color_total = a.color_total; m_fem.color_cutting_edge = a.color_cutting_edge;
        m_fem.fixed_ele_num = a.fixed_ele_num;
        m_fem.m_row_ptr = a.csr_row_ptr; m_fem.m_col_ind = a.csr_col_ind;
        m_fem.m_stiffness_matrix = a.K; m_fem.m_stiffness_matrix_temp = a.K_temp;
        m_fem.m_mass_matrix = a.M; m_fem.m_mass_matrix_inverse = a.M_inv;
        m_fem.m_f = a.f; m_fem.m_f_temp = a.f_temp; m_fem.m_k = a.k; m_fem.m_T = a.T; m_fem.m_T_old = a.T_old;
        m_fem.m_T_mapping = a.T_mapping;
        mesh_GPU_m &g = mesh_GPU ? *mesh_GPU : m_mesh;
        g.node_num = a.node_num_el; g.position = a.position; g.position_mapping = a.position_mapping;
        g.csr_row_ptr = a.csr_row_ptr; g.csr_col_ind = a.csr_col_ind; g.node_flags = a.node_flags;
        g.no_num = a.node_num; g.ele_num_total = a.ele_num; g.non_zeros = a.nz_num;
        g.color_total = a.color_total; g.cutting_edge_color = a.color_cutting_edge;
        g.ele_num_fixed = a.fixed_ele_num;
    }
    sphfem_tool *m_handle = nullptr;
    vec2_t m_orient;
    bool m_oriented = false;
    segment_FEM_gpu m_seg;
    FEM_thermal_2D_GPU m_fem;
    mesh_GPU_m m_mesh;
};

// ---------------------------------------------------------------------------------------------
// grid.h:23-64, grid_gpu_green.h:84-85 (locked-geometry constructor)
// ---------------------------------------------------------------------------------------------
class grid_base {
  public:
    virtual ~grid_base() {
        if (m_handle && sphfem_shim::ctx_slot()) (void)sphfem_grid_destroy(sphfem_shim::ctx_slot(), m_handle);
    }
    // the geometry of an unlocked grid changes at every update_geometry: read it from the handle
    int nx() const { return geom().nx; }
    int ny() const { return geom().ny; }
    double bbmin_x() const { return geom().bbmin_x; }
    double bbmin_y() const { return geom().bbmin_y; }
    double bbmax_x() const { return bounds().x; }
    double bbmax_y() const { return bounds().y; }
    double dx() const { return geom().dx; }
    bool is_locked() const {
        int locked = 1;
        sphfem_shim::check(sphfem_grid_bounds(m_handle, nullptr, &locked, nullptr), "grid is_locked");
        return locked != 0;
    }
    int num_cell() const {
        int n = 0;
        sphfem_shim::check(sphfem_grid_geometry(m_handle, nullptr, &n), "grid num_cell");
        return n;
    }

    void assign_hashes(particle_gpu *particles, tool *) const {
        sphfem_shim::check(sphfem_grid_assign_hashes(sphfem_shim::ctx(), m_handle, particles), "assign_hashes");
    }
    // grid.cu:88-127 (no-op for a locked grid)
    void update_geometry(particle_gpu *particles, tool *, double kernel_width = 2.0) {
        sphfem_shim::check(sphfem_grid_update_geometry(sphfem_shim::ctx(), m_handle, particles, kernel_width),
                           "update_geometry");
    }
    virtual void sort(particle_gpu *particles, tool *t) const = 0;
    virtual void get_cells(particle_gpu *particles, int *cell_start, int *cell_end) = 0;
    sphfem_grid *handle() const { return m_handle; }

  protected:
    grid_base() = default;
    void adopt(sphfem_grid *g) { m_handle = g; }
    geom_constants geom() const {
        geom_constants gc{};
        sphfem_shim::check(sphfem_grid_geometry(m_handle, &gc, nullptr), "grid geometry");
        return gc;
    }
    float2_t bounds() const {
        float2_t b{};
        sphfem_shim::check(sphfem_grid_bounds(m_handle, &b, nullptr, nullptr), "grid bounds");
        return b;
    }
    sphfem_grid *m_handle = nullptr;
};

class grid_gpu_green : public grid_base {
  public:
    grid_gpu_green(int num_part, float2_t bbmin, float2_t bbmax, double h) {
        sphfem_grid *g = nullptr;
        sphfem_shim::check(sphfem_grid_green_create(sphfem_shim::ctx(), num_part, bbmin, bbmax, h, &g),
                           "grid_gpu_green");
        adopt(g);
    }
    // grid_gpu_green.cu:258-262: geometry from the particles at each update_geometry
    grid_gpu_green(unsigned int max_cell, unsigned int N) {
        sphfem_grid *g = nullptr;
        sphfem_shim::check(sphfem_grid_green_create_unlocked(sphfem_shim::ctx(), (int)max_cell, (int)N, &g),
                           "grid_gpu_green(max_cell, N)");
        adopt(g);
    }
    void sort(particle_gpu *particles, tool *) const override {
        sphfem_shim::check(sphfem_grid_sort(sphfem_shim::ctx(), m_handle, particles), "grid_gpu_green::sort");
    }
    // incremental re-sort limit of the step (see sphfem_grid_set_resort_limit)
    void set_resort_limit(int max_changed) {
        sphfem_shim::check(sphfem_grid_set_resort_limit(sphfem_shim::ctx(), m_handle, max_changed),
                           "grid_gpu_green::set_resort_limit");
    }
    void get_cells(particle_gpu *particles, int *cell_start, int *cell_end) override {
        sphfem_shim::check(sphfem_grid_get_cells(sphfem_shim::ctx(), m_handle, particles, cell_start, cell_end),
                           "grid_gpu_green::get_cells");
    }
};

// grid_gpu_rothlin.h:79-80: in-place gathers (idx carries labels; ft and n are not reordered) and
// dense cell offsets computed from the particles' sorted hashes
class grid_gpu_rothlin : public grid_base {
  public:
    grid_gpu_rothlin(int num_part, float2_t bbmin, float2_t bbmax, double h) {
        sphfem_grid *g = nullptr;
        sphfem_shim::check(sphfem_grid_rothlin_create(sphfem_shim::ctx(), num_part, bbmin, bbmax, h, &g),
                           "grid_gpu_rothlin");
        adopt(g);
    }
    // grid_gpu_rothlin.cu:277-280
    grid_gpu_rothlin(int max_cell, int N) {
        sphfem_grid *g = nullptr;
        sphfem_shim::check(sphfem_grid_rothlin_create_unlocked(sphfem_shim::ctx(), max_cell, N, &g),
                           "grid_gpu_rothlin(max_cell, N)");
        adopt(g);
    }
    void sort(particle_gpu *particles, tool *) const override {
        sphfem_shim::check(sphfem_grid_sort(sphfem_shim::ctx(), m_handle, particles), "grid_gpu_rothlin::sort");
    }
    void get_cells(particle_gpu *particles, int *cell_start, int *cell_end) override {
        sphfem_shim::check(sphfem_grid_get_cells(sphfem_shim::ctx(), m_handle, particles, cell_start, cell_end),
                           "grid_gpu_rothlin::get_cells");
    }
};

// ---------------------------------------------------------------------------------------------
// leap_frog.h:31-50
// ---------------------------------------------------------------------------------------------
class leap_frog {
  public:
    leap_frog(unsigned int num_part, unsigned int num_cell) : leap_frog(num_part, num_cell, 0) {}
    leap_frog(unsigned int num_part, unsigned int num_cell, int node_num_in_FEM_tool) {
        sphfem_shim::check(sphfem_leap_frog_create(sphfem_shim::ctx(), num_part, num_cell, node_num_in_FEM_tool,
                                                   &m_handle),
                           "leap_frog");
    }
    ~leap_frog() {
        if (m_handle && sphfem_shim::ctx_slot()) (void)sphfem_leap_frog_destroy(sphfem_shim::ctx_slot(), m_handle);
    }
    leap_frog(const leap_frog &) = delete;
    leap_frog &operator=(const leap_frog &) = delete;

    // SPHFEM_FP64 (default) or SPHFEM_FP32C
    // the FP32C form keeps pos_t, vel_t, S_t, rho_t and T_t in FP32 inside the stepper: these write
    // them to / read them from the particle arrays (no-ops in the FP64 form)
    void export_rates(particle_gpu *particles) {
        sphfem_shim::check(sphfem_leap_frog_export_rates(sphfem_shim::ctx(), m_handle, particles),
                           "leap_frog::export_rates");
    }
    void import_rates(particle_gpu *particles) {
        sphfem_shim::check(sphfem_leap_frog_import_rates(sphfem_shim::ctx(), m_handle, particles),
                           "leap_frog::import_rates");
    }
    // launch parameters of the step kernels (no counterpart in the reference): pair block, FP32C list
    // chunk, update block, FP32C pair kernel (list or LDS-staged), one-workgroup sort fallback
    void set_tuning(const sphfem_tuning &tuning) {
        sphfem_shim::check(sphfem_leap_frog_set_tuning(sphfem_shim::ctx(), m_handle, &tuning), "leap_frog::set_tuning");
    }
    sphfem_tuning get_tuning() const {
        sphfem_tuning t{};
        sphfem_shim::check(sphfem_leap_frog_get_tuning(sphfem_shim::ctx(), m_handle, &t), "leap_frog::get_tuning");
        return t;
    }
    void set_precision(int precision) {
        sphfem_shim::check(sphfem_leap_frog_set_precision(sphfem_shim::ctx(), m_handle, precision),
                           "leap_frog::set_precision");
    }
    // leap_frog.cu:162-248, with global_dt and wear_pro read as the reference reads them
    void step(particle_gpu *particles, grid_base *g, tool_FEM *m_tool) {
        sphfem_ctx *c = sphfem_shim::ctx();
        sphfem_shim::check(sphfem_set_dt(c, global_dt), "set_dt");
        sphfem_shim::check(sphfem_set_wear(c, wear_pro ? 1 : 0), "set_wear");
        sphfem_shim::check(sphfem_leap_frog_step(c, m_handle, particles, g->handle(),
                                                 m_tool ? m_tool->handle() : nullptr),
                           "leap_frog::step");
    }
    // leap_frog.cu:99-160 with global_tool (its device copy; a tool never set up on the device, or
    // none, gives the tool-free step).  The host copy of the tool follows, as tool_gpu_update_tool
    // moves it in the reference.
    void step(particle_gpu *particles, grid_base *g) {
        sphfem_ctx *c = sphfem_shim::ctx();
        sphfem_shim::check(sphfem_set_dt(c, global_dt), "set_dt");
        sphfem_rigid_tool *rt = global_tool ? global_tool->gpu_handle() : nullptr;
        sphfem_shim::check(sphfem_leap_frog_step_rigid(c, m_handle, particles, g->handle(), rt), "leap_frog::step");
        if (rt) global_tool->update_tool(global_dt);
    }

  private:
    sphfem_leap_frog *m_handle = nullptr;
};

// ---------------------------------------------------------------------------------------------
// actions_gpu.h:33-72, interactions_gpu.h:27-43: the constant blocks are shared by all kernels
// ---------------------------------------------------------------------------------------------
#define SPHFEM_SHIM_CALL(expr, name) sphfem_shim::check((expr), name)

inline void actions_setup_physical_constants(phys_constants c) {
    SPHFEM_SHIM_CALL(sphfem_setup_physical_constants(sphfem_shim::ctx(), &c), "setup_physical_constants");
}
inline void actions_setup_corrector_constants(corr_constants c) {
    SPHFEM_SHIM_CALL(sphfem_setup_corrector_constants(sphfem_shim::ctx(), &c), "setup_corrector_constants");
}
inline void actions_setup_johnson_cook_constants(joco_constants c) {
    SPHFEM_SHIM_CALL(sphfem_setup_johnson_cook_constants(sphfem_shim::ctx(), &c), "setup_johnson_cook_constants");
}
inline void actions_setup_thermal_constants_wp(trml_constants c) {
    SPHFEM_SHIM_CALL(sphfem_setup_thermal_constants_wp(sphfem_shim::ctx(), &c), "setup_thermal_constants_wp");
}
inline void actions_setup_thermal_constants_tool(trml_constants c) {
    SPHFEM_SHIM_CALL(sphfem_setup_thermal_constants_tool(sphfem_shim::ctx(), &c), "setup_thermal_constants_tool");
}
inline void actions_setup_tool_constants(tool_constants c) {
    SPHFEM_SHIM_CALL(sphfem_setup_tool_constants(sphfem_shim::ctx(), &c), "setup_tool_constants");
}
inline void actions_setup_wear_constants(wear_constants c) {
    SPHFEM_SHIM_CALL(sphfem_setup_wear_constants(sphfem_shim::ctx(), &c), "setup_wear_constants");
}
inline void interactions_setup_physical_constants(phys_constants c) { actions_setup_physical_constants(c); }
inline void interactions_setup_corrector_constants(corr_constants c) { actions_setup_corrector_constants(c); }
inline void interactions_setup_thermal_constants_workpiece(trml_constants c) { actions_setup_thermal_constants_wp(c); }
inline void interactions_setup_tool_constants(tool_constants c) { actions_setup_tool_constants(c); }
// interactions_gpu.cu:494-502
inline void interactions_setup_thermal_constants_tool(trml_constants c, tool *t) {
    SPHFEM_SHIM_CALL(sphfem_interactions_setup_thermal_constants_tool(sphfem_shim::ctx(), &c),
                     "interactions_setup_thermal_constants_tool");
    if (c.alpha != 0. && t) t->set_thermal(true);
}
// actions_gpu.cu:956-965
inline void material_fric_heat_gen(particle_gpu *p, vec2_t vel) {
    sphfem_ctx *c = sphfem_shim::ctx();
    SPHFEM_SHIM_CALL(sphfem_set_dt(c, global_dt), "set_dt");
    SPHFEM_SHIM_CALL(sphfem_material_fric_heat_gen(c, p, sphfem_double2{vel.x, vel.y}), "material_fric_heat_gen");
}
inline void interactions_setup_geometry_constants(grid_base *g) {
    SPHFEM_SHIM_CALL(sphfem_interactions_setup_geometry_constants(sphfem_shim::ctx(), g->handle()),
                     "interactions_setup_geometry_constants");
}
inline void interactions_monaghan(particle_gpu *p, const int *cell_start, const int *cell_end, int num_cell) {
    SPHFEM_SHIM_CALL(sphfem_interactions_monaghan(sphfem_shim::ctx(), p, cell_start, cell_end, num_cell),
                     "interactions_monaghan");
}

inline void interactions_heat_pse(particle_gpu *p, const int *cell_start, const int *cell_end, int num_cell) {
    SPHFEM_SHIM_CALL(sphfem_interactions_heat_pse(sphfem_shim::ctx(), p, cell_start, cell_end, num_cell),
                     "interactions_heat_pse");
}

inline void material_eos(particle_gpu *p) { SPHFEM_SHIM_CALL(sphfem_material_eos(sphfem_shim::ctx(), p), "material_eos"); }
inline void corrector_artificial_stress(particle_gpu *p) {
    SPHFEM_SHIM_CALL(sphfem_corrector_artificial_stress(sphfem_shim::ctx(), p), "corrector_artificial_stress");
}
inline void material_stress_rate_jaumann(particle_gpu *p) {
    SPHFEM_SHIM_CALL(sphfem_material_stress_rate_jaumann(sphfem_shim::ctx(), p), "material_stress_rate_jaumann");
}
inline void contmech_continuity(particle_gpu *p) {
    SPHFEM_SHIM_CALL(sphfem_contmech_continuity(sphfem_shim::ctx(), p), "contmech_continuity");
}
inline void contmech_momentum(particle_gpu *p) {
    SPHFEM_SHIM_CALL(sphfem_contmech_momentum(sphfem_shim::ctx(), p), "contmech_momentum");
}
inline void contmech_advection(particle_gpu *p) {
    SPHFEM_SHIM_CALL(sphfem_contmech_advection(sphfem_shim::ctx(), p), "contmech_advection");
}
inline void plasticity_johnson_cook(particle_gpu *p) {
    SPHFEM_SHIM_CALL(sphfem_plasticity_johnson_cook(sphfem_shim::ctx(), p), "plasticity_johnson_cook");
}
inline void perform_boundary_conditions(particle_gpu *p) {
    SPHFEM_SHIM_CALL(sphfem_perform_boundary_conditions(sphfem_shim::ctx(), p), "perform_boundary_conditions");
}
inline void perform_boundary_conditions_thermal(particle_gpu *p) {
    SPHFEM_SHIM_CALL(sphfem_perform_boundary_conditions_thermal(sphfem_shim::ctx(), p),
                     "perform_boundary_conditions_thermal");
}
inline void debug_invalidate(particle_gpu *p) {
    SPHFEM_SHIM_CALL(sphfem_debug_invalidate(sphfem_shim::ctx(), p), "debug_invalidate");
}
inline void material_fric_heat_gen_FEM_v2(particle_gpu *p, segment_FEM_gpu *, vec2_t, tool_FEM *t) {
    SPHFEM_SHIM_CALL(sphfem_material_fric_heat_gen_FEM_v2(sphfem_shim::ctx(), p, t->handle()),
                     "material_fric_heat_gen_FEM_v2");
}
inline void compute_contact_forces_FEM_tool_v3(particle_gpu *p, segment_FEM_gpu *, tool_FEM *t) {
    SPHFEM_SHIM_CALL(sphfem_compute_contact_forces_FEM_tool_v3(sphfem_shim::ctx(), p, t->handle()),
                     "compute_contact_forces_FEM_tool_v3");
}
inline void wear_adjustment(tool_FEM *t) {
    SPHFEM_SHIM_CALL(sphfem_wear_adjustment(sphfem_shim::ctx(), t->handle()), "wear_adjustment");
}

// actions_gpu.cu:1022-1031
inline void remesh_low_level(tool_FEM *t, int freq) {
    SPHFEM_SHIM_CALL(sphfem_tool_remesh_low_level(sphfem_shim::ctx(), t->handle(), freq), "remesh_low_level");
}

namespace sphfem_shim {
// true if `newer` was modified after `older` (both must exist)
inline bool file_is_newer(const std::string &newer, const std::string &older) {
    struct stat a {}, b {};
    if (stat(newer.c_str(), &a) != 0 || stat(older.c_str(), &b) != 0) return false;
    if (a.st_mtim.tv_sec != b.st_mtim.tv_sec) return a.st_mtim.tv_sec > b.st_mtim.tv_sec;
    return a.st_mtim.tv_nsec > b.st_mtim.tv_nsec;
}
}  // namespace sphfem_shim

// actions_gpu.cu:1033-1038 with the meshing step supplied by the caller (the reference calls the Gmsh
// library between the two halves): the geometry of the current cutting edge goes to geo_file,
// mesh(geo_file, inp_file) must write the new mesh to inp_file, and the tool takes that mesh.  An
// inp_file that is missing or not newer than the geo_file just written is refused (it would be an
// older mesh that undoes the wear).
// (A template over the tool type only so that the host logic can be exercised with a stand-in tool.)
template <class Tool>
inline void remesh_gmsh(Tool *t, const std::string &geo_file, const std::string &inp_file,
                        const std::function<void(const std::string &, const std::string &)> &mesh) {
    t->remesh_read_segments(geo_file);
    if (mesh) mesh(geo_file, inp_file);
    if (!sphfem_shim::file_is_newer(inp_file, geo_file))
        throw std::runtime_error("remesh_gmsh: " + inp_file + " is missing or older than " + geo_file);
    t->read_remesh_gmsh(inp_file);
}

// output.h:17
inline float2_t report_force(particle_gpu *p) {
    float2_t f{0., 0.};
    SPHFEM_SHIM_CALL(sphfem_report_force(sphfem_shim::ctx(), p, &f), "report_force");
    return f;
}

// vtk_writer.h:26-31: files under results/ named as the reference names them
namespace sphfem_shim {
inline std::string vtk_path(const char *stem, int step) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "results/%s_%06d.vtk", stem, step);
    return buf;
}
}  // namespace sphfem_shim
inline void vtk_writer_write(const particle_gpu *particles, int step) {
    SPHFEM_SHIM_CALL(sphfem_vtk_write_particles(sphfem_shim::ctx(), particles,
                                                sphfem_shim::vtk_path("vtk_out", step).c_str(), 0),
                     "vtk_writer_write(particles)");
}
inline void vtk_writer_write(const tool *t, int step) {
    sphfem_tool_geometry g = t->geometry();
    if (t->gpu_handle())   // the device copy is the one the steps move
        SPHFEM_SHIM_CALL(sphfem_rigid_tool_get_geometry(sphfem_shim::ctx(), t->gpu_handle(), &g), "tool geometry");
    SPHFEM_SHIM_CALL(sphfem_vtk_write_rigid_tool(&g, sphfem_shim::vtk_path("vtk_tool", step).c_str()),
                     "vtk_writer_write(tool)");
}
inline void vtk_writer_write_FEM_tool(const tool_FEM *t, int step) {
    SPHFEM_SHIM_CALL(sphfem_vtk_write_FEM_tool(sphfem_shim::ctx(), t->handle(),
                                               sphfem_shim::vtk_path("vtk_tool", step).c_str()),
                     "vtk_writer_write_FEM_tool");
}
inline void vtk_writer_write_FEM_segments(const tool_FEM *t, int step) {
    SPHFEM_SHIM_CALL(sphfem_vtk_write_FEM_segments(sphfem_shim::ctx(), t->handle(),
                                                   sphfem_shim::vtk_path("vtk_seg", step).c_str()),
                     "vtk_writer_write_FEM_segments");
}
inline void vtk_writer_write_tool_org(const tool_FEM *t, int step, float2_t tl_ref) {
    SPHFEM_SHIM_CALL(sphfem_vtk_write_tool_org(sphfem_shim::ctx(), t->handle(), tl_ref,
                                               sphfem_shim::vtk_path("vtk_tool_org", step).c_str()),
                     "vtk_writer_write_tool_org");
}
inline void vtk_writer_write_contact_variables(const tool_FEM *t, int step) {
    SPHFEM_SHIM_CALL(sphfem_vtk_write_contact_variables(sphfem_shim::ctx(), t->handle(),
                                                        sphfem_shim::vtk_path("vtk_phy_wear", step).c_str()),
                     "vtk_writer_write_contact_variables");
}
// mesh_read::mesh_read_temperature (mesh_structure_GPU.cu:1259-1296): the nodal temperatures of a
// hot-tool restart file, for sphfem_tool_desc::T
inline std::vector<double> mesh_read_temperature(const std::string &file_name) {
    long n = 0;
    SPHFEM_SHIM_CALL(sphfem_read_vtk_temperature(file_name.c_str(), nullptr, 0, &n), "mesh_read_temperature");
    std::vector<double> T((size_t)n);
    SPHFEM_SHIM_CALL(sphfem_read_vtk_temperature(file_name.c_str(), T.data(), n, &n), "mesh_read_temperature");
    return T;
}

// tool_wear.h:17-47
class tool_wear {
  public:
    tool_wear(double usui_K, double usui_alpha, unsigned int, phys_constants, vec2_t tool_vel) {
        SPHFEM_SHIM_CALL(sphfem_tool_wear_create(sphfem_shim::ctx(), usui_K, usui_alpha,
                                                 sphfem_double2{tool_vel.x, tool_vel.y}, &m_handle),
                         "tool_wear");
    }
    ~tool_wear() {
        if (m_handle && sphfem_shim::ctx_slot()) (void)sphfem_tool_wear_destroy(sphfem_shim::ctx_slot(), m_handle);
    }
    tool_wear(const tool_wear &) = delete;
    tool_wear &operator=(const tool_wear &) = delete;
    void eval_usui(const particle_gpu *particles, double dt, double &usui_min, double &usui_max, double &usui_avg) {
        SPHFEM_SHIM_CALL(sphfem_tool_wear_eval_usui(sphfem_shim::ctx(), m_handle, particles, dt, &usui_min,
                                                    &usui_max, &usui_avg),
                         "tool_wear::eval_usui");
    }
    double get_accum_wear() const {
        double v = 0.;
        SPHFEM_SHIM_CALL(sphfem_tool_wear_get_accum_wear(m_handle, &v), "tool_wear::get_accum_wear");
        return v;
    }

  private:
    sphfem_tool_wear *m_handle = nullptr;
};

// mfree_gpu.cpp:41-42: the driver's FEM tool and wear objects
inline tool_FEM *global_tool_FEM = nullptr;
inline tool_wear *global_wear = nullptr;

// types.cpp:11-37
inline bool check_cuda_error() { return sphfem_check_error(sphfem_shim::ctx()) == SPHFEM_OK; }

#undef SPHFEM_SHIM_CALL

#endif /* SPHFEM_HPP */
