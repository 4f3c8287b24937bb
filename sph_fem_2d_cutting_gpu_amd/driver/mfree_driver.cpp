// mfree_driver -- the reference driver loop (mfree_gpu.cpp:229-366) written against the C++
// host mirror (include/sphfem/sphfem.hpp): set up the constant blocks, construct
// particle_gpu / grid_gpu_green / tool_FEM / leap_frog, then step and report the cutting force
// every `freq` steps.  Output files replace the VTK dumps: forces.txt (step Fx Fy) and the final
// sorted state (pos, vel, S, T, idx as raw little-endian arrays).
//
// usage: mfree_driver CASE_DIR STEPS FREQ fp64|fp32c OUT_DIR
//   CASE_DIR as written by sph_fem_2d_cutting_gpu_amd.case_io.write_case_dir; with a tool.inp
//   there, the tool is set up from that mesh file
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/sphfem/sphfem.hpp"

namespace {

std::map<std::string, double> read_scalars(const std::string &dir) {
    std::map<std::string, double> m;
    std::ifstream f(dir + "/scalars.txt");
    if (!f) throw std::runtime_error("cannot open " + dir + "/scalars.txt");
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream ss(line);
        std::string k;
        double v;
        if (ss >> k >> v) m[k] = v;
    }
    return m;
}

template <class T>
std::vector<T> read_array(const std::string &path, size_t count) {
    std::vector<T> v(count);
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    size_t got = std::fread(v.data(), sizeof(T), count, f);
    std::fclose(f);
    if (got != count) throw std::runtime_error("short read: " + path);
    return v;
}

template <class T>
void write_array(const std::string &path, const std::vector<T> &v) {
    std::FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

template <class T>
std::vector<T> download(const T *dev, size_t count) {
    std::vector<T> v(count);
    sphfem_shim::check(sphfem_memcpy_dtoh(sphfem_shim::ctx(), v.data(), dev, sizeof(T) * count), "download");
    return v;
}

double get(const std::map<std::string, double> &m, const std::string &k) {
    auto it = m.find(k);
    if (it == m.end()) throw std::runtime_error("missing scalar " + k);
    return it->second;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s CASE_DIR STEPS FREQ fp64|fp32c OUT_DIR\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1], out = argv[5], prec = argv[4];
    const long steps = std::atol(argv[2]);
    const long freq = std::atol(argv[3]);
    if (steps < 0 || freq <= 0 || (prec != "fp64" && prec != "fp32c")) {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    try {
        const auto s = read_scalars(dir);
        const unsigned N = (unsigned)get(s, "N");

        // constant blocks (benchmarks_NZ.cpp:341-356)
        phys_constants phys{get(s, "phys.E"), get(s, "phys.nu"), get(s, "phys.rho0"), get(s, "phys.K"),
                            get(s, "phys.G"), get(s, "phys.mass"), get(s, "phys.mass_tool")};
        corr_constants corr{get(s, "corr.wdeltap"), get(s, "corr.stresseps"), get(s, "corr.xspheps"),
                            get(s, "corr.alpha"), get(s, "corr.beta"), get(s, "corr.eta")};
        joco_constants joco{get(s, "joco.A"), get(s, "joco.B"), get(s, "joco.C"), get(s, "joco.n"),
                            get(s, "joco.m"), get(s, "joco.Tmelt"), get(s, "joco.Tref"), get(s, "joco.eps_dot_ref"),
                            get(s, "joco.tanh_a"), get(s, "joco.tanh_b"), get(s, "joco.tanh_c"),
                            get(s, "joco.tanh_d")};
        auto trml = [&](const char *b) {
            std::string p(b);
            return trml_constants{get(s, p + ".cp"), get(s, p + ".tq"), get(s, p + ".k"), get(s, p + ".alpha"),
                                  get(s, p + ".T_init"), get(s, p + ".eta")};
        };
        tool_constants tc{get(s, "tool.cp"), get(s, "tool.tq"), get(s, "tool.k"), get(s, "tool.eta"),
                          get(s, "tool.rho"), get(s, "tool.T0"), get(s, "tool.h"), get(s, "tool.beta"),
                          get(s, "tool.mu"), get(s, "tool.contact_alpha"), get(s, "tool.slave_mass"),
                          get(s, "tool.hc")};
        wear_constants wc{get(s, "wear.A"), get(s, "wear.B"), get(s, "wear.C"), get(s, "wear.D")};
        actions_setup_physical_constants(phys);
        actions_setup_corrector_constants(corr);
        actions_setup_johnson_cook_constants(joco);
        actions_setup_thermal_constants_wp(trml("trml_wp"));
        actions_setup_thermal_constants_tool(trml("trml_tool"));
        actions_setup_tool_constants(tc);
        actions_setup_wear_constants(wc);
        global_dt = get(s, "dt");
        wear_pro = get(s, "wear") != 0.;
        sphfem_shim::check(sphfem_set_kernel_cutoff(sphfem_shim::ctx(), get(s, "rr2_min")), "set_kernel_cutoff");

        // particles: particle_gpu(pos, vel, rho, T, h, fixed, N) as the cutting setups construct them
        auto pos = read_array<float2_t>(dir + "/p_pos.bin", N);
        auto vel = read_array<float2_t>(dir + "/p_vel.bin", N);
        auto rho = read_array<double>(dir + "/p_rho.bin", N);
        auto T = read_array<double>(dir + "/p_T.bin", N);
        auto h = read_array<double>(dir + "/p_h.bin", N);
        auto fixed = read_array<double>(dir + "/p_fixed.bin", N);
        auto S = read_array<float4_t>(dir + "/p_S.bin", N);
        particle_gpu particles(pos.data(), vel.data(), rho.data(), T.data(), h.data(), fixed.data(), N);
        sphfem_shim::upload(particles.S, S.data(), N);

        grid_gpu_green grid((int)N, float2_t{get(s, "bbmin_x"), get(s, "bbmin_y")},
                            float2_t{get(s, "bbmax_x"), get(s, "bbmax_y")}, get(s, "h"));

        // tool mesh: read from CASE_DIR/tool.inp as benchmarks_NZ.cpp:228-245 does (mesh_read,
        // set_up_segments_from_mesh, set_up_FEM_tool_gpu), else the set-up arrays of the case
        std::unique_ptr<tool_FEM> tool;
        std::unique_ptr<mesh_read> mesh;
        const int nn = (int)get(s, "node_num");
        std::vector<sphfem_double2> t_pos, t_left, t_right, t_n;
        std::vector<sphfem_int4> t_elem, t_marks;
        std::vector<int> t_layer, t_rp, t_ci, t_nflags, t_fixed;
        std::vector<sphfem_int4> t_fflags;
        std::vector<double> t_K, t_M, t_f, t_T;
        const std::string inp = dir + "/tool.inp";
        std::unique_ptr<mesh_GPU_m> mesh_gpu;
        if (nn > 0 && std::ifstream(inp).good() && s.count("tool_outline.length")) {
            // benchmarks_NZ.cpp:228-245 step by step
            tool.reset(new tool_FEM(vec2_t(get(s, "tool.tl_x"), get(s, "tool.tl_y")), get(s, "tool_outline.length"),
                                    get(s, "tool_outline.height"), get(s, "tool_outline.rake"),
                                    get(s, "tool_outline.clearance"), get(s, "tool_outline.r"), tc.mu));
            mesh.reset(new mesh_read(inp, get(s, "tool.inp_gmsh") != 0.));
            tool->mesh_CPU = mesh.get();
            mesh_gpu.reset(new mesh_GPU_m(mesh->nodes, mesh->elements, mesh->color_total, mesh->color_cutting_edge, 1));
            tool->mesh_GPU = mesh_gpu.get();
            tool->set_up_segments_from_mesh();
            tool->set_up_FEM_tool_gpu(tc);
            tool->set_vel(vec2_t(get(s, "tool.vel_x"), get(s, "tool.vel_y")));
            std::printf("tool mesh %s (tool_FEM set-up sequence): %d nodes, %d elements, %d cutting segments\n",
                        inp.c_str(), tool->node_number, tool->element_number, tool->cutting_segment_size);
        } else if (nn > 0 && std::ifstream(inp).good()) {
            mesh.reset(new mesh_read(inp, get(s, "tool.inp_gmsh") != 0.));
            tool.reset(new tool_FEM(*mesh, tc, vec2_t(get(s, "tool.tl_x"), get(s, "tool.tl_y")),
                                    vec2_t(get(s, "tool.vel_x"), get(s, "tool.vel_y"))));
            std::printf("tool mesh %s: %d nodes, %d elements, %d cutting segments\n", inp.c_str(),
                        tool->node_number, tool->element_number, tool->cutting_segment_size);
        } else if (nn > 0) {
            sphfem_tool_desc d{};
            d.node_num = nn;
            d.ele_num = (int)get(s, "ele_num");
            d.nz_num = (int)get(s, "nz_num");
            d.seg_num = (int)get(s, "seg_num");
            d.color_total = (int)get(s, "color_total");
            d.color_cutting_edge = (int)get(s, "color_cutting_edge");
            t_pos = read_array<sphfem_double2>(dir + "/t_node_pos.bin", nn);
            t_elem = read_array<sphfem_int4>(dir + "/t_elem.bin", d.ele_num);
            t_layer = read_array<int>(dir + "/t_layer.bin", d.ele_num);
            t_rp = read_array<int>(dir + "/t_row_ptr.bin", nn + 1);
            t_ci = read_array<int>(dir + "/t_col_ind.bin", d.nz_num);
            t_K = read_array<double>(dir + "/t_K.bin", d.nz_num);
            t_M = read_array<double>(dir + "/t_M.bin", nn);
            t_f = read_array<double>(dir + "/t_f.bin", nn);
            t_T = read_array<double>(dir + "/t_T.bin", nn);
            t_left = read_array<sphfem_double2>(dir + "/t_left.bin", d.seg_num);
            t_right = read_array<sphfem_double2>(dir + "/t_right.bin", d.seg_num);
            t_n = read_array<sphfem_double2>(dir + "/t_n.bin", d.seg_num);
            t_marks = read_array<sphfem_int4>(dir + "/t_marks.bin", d.seg_num);
            d.position = t_pos.data();
            d.node_num_el = t_elem.data();
            d.layer = t_layer.data();
            d.csr_row_ptr = t_rp.data();
            d.csr_col_ind = t_ci.data();
            d.K = t_K.data();
            d.M = t_M.data();
            d.f = t_f.data();
            d.T = t_T.data();
            d.left = t_left.data();
            d.right = t_right.data();
            d.n = t_n.data();
            d.marks = t_marks.data();
            d.vel = sphfem_double2{get(s, "tool.vel_x"), get(s, "tool.vel_y")};
            d.tl_point_ref = sphfem_double2{get(s, "tool.tl_x"), get(s, "tool.tl_y")};
            d.br_point = sphfem_double2{get(s, "tool.br_x"), get(s, "tool.br_y")};
            d.ref_dist = get(s, "tool.ref_dist");
            d.lowest_point = get(s, "tool.lowest_point");
            d.right_point = get(s, "tool.right_point");
            if (s.count("fixed_num")) {      // the remeshing description
                d.fixed_num = (int)get(s, "fixed_num");
                t_nflags = read_array<int>(dir + "/t_node_flags.bin", nn);
                t_fixed = read_array<int>(dir + "/t_fixed_elems.bin", d.fixed_num);
                t_fflags = read_array<sphfem_int4>(dir + "/t_fixed_flags.bin", d.fixed_num);
                d.node_flags = t_nflags.data();
                d.fixed_elems = t_fixed.data();
                d.fixed_flags = t_fflags.data();
            }
            tool.reset(new tool_FEM(d));
        }

        // mfree_gpu.cpp:234
        leap_frog stepper(particles.N, (unsigned)grid.num_cell(), tool ? tool->node_number : 0);
        if (prec == "fp32c") stepper.set_precision(SPHFEM_FP32C);

        // mfree_gpu.cpp:259-260, 351-357: with wear, the tool is remeshed every freq / 20 steps
        // (remesh_low_level; the Gmsh remesh needs the meshing program between its two halves)
        const bool remesh = s.count("remesh_low_level") && get(s, "remesh_low_level") != 0.;
        const long freq_remesh = freq / 20;
        if (remesh && freq_remesh <= 0) throw std::runtime_error("remeshing needs FREQ >= 20 (freq / 20 steps)");

        // mfree_gpu.cpp:289-366 without the dumps
        std::FILE *ff = std::fopen((out + "/forces.txt").c_str(), "w");
        if (!ff) throw std::runtime_error("cannot write " + out + "/forces.txt");
        for (long step = 0; step < steps; step++) {
            if (step % freq == 0) {
                float2_t F = report_force(&particles);
                std::fprintf(ff, "%ld %.17g %.17g\n", step, F.x, F.y);
            }
            if (remesh && tool && wear_pro && step % freq_remesh == 0) remesh_low_level(tool.get(), (int)freq);
            stepper.step(&particles, &grid, tool.get());
        }
        float2_t F = report_force(&particles);
        std::fprintf(ff, "%ld %.17g %.17g\n", steps, F.x, F.y);
        std::fclose(ff);
        if (!check_cuda_error()) throw std::runtime_error(sphfem_last_error(sphfem_shim::ctx()));

        if (tool) {
            // the tool's temperatures through the reference's view (FEM_solver->m_T, FEM_thermal_gpu.h:58)
            const auto T = download(tool->FEM_solver->m_T, (size_t)tool->FEM_solver->node_num);
            double Tmax = T.empty() ? 0. : T[0];
            for (double v : T) Tmax = v > Tmax ? v : Tmax;
            std::printf("tool nodes %d, segments %d, max T %.17g\n", tool->mesh_GPU->get_node_number(),
                        tool->segments_FEM->segment_num, Tmax);
        }
        write_array(out + "/pos.bin", download(particles.pos, N));
        write_array(out + "/vel.bin", download(particles.vel, N));
        write_array(out + "/S.bin", download(particles.S, N));
        write_array(out + "/T.bin", download(particles.T, N));
        write_array(out + "/idx.bin", download(particles.idx, N));
        std::printf("mfree_driver: %ld steps of %u particles (%s), F = (%.9e, %.9e)\n", steps, N, prec.c_str(), F.x,
                    F.y);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "mfree_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
