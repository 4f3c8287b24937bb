// ref_cut_driver -- the rigid-tool cutting setup setup_ref_cut (benchmarks.cpp:248-530) and the
// driver loop (mfree_gpu.cpp:229-366 with leap_frog::step(particles, grid)) written against the
// C++ host mirror (include/sphfem/sphfem.hpp): Ti6Al4V workpiece, tool(tl, length, height, rake,
// clearance, fillet, mu) moved to the feed, sampled with tool particles, tool_gpu_set_up_tool, then
// steps with the cutting force logged every `freq` steps.  Outputs: forces.txt (step Fx Fy), the
// final state (pos, vel, T, fc, idx as raw little-endian arrays) and tool.txt (the device tool's
// segment left ends).
//
// usage: ref_cut_driver NY STEPS FREQ fp64|fp32c OUT_DIR [rake clearance fillet speed_in_ui feed]
// With REF_CUT_RESULTS=1 it also writes the reference's results/ files every FREQ steps.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/sphfem/sphfem.hpp"

namespace {

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

struct Setup {
    particle_gpu *particles = nullptr;
    grid_base *grid = nullptr;
    tool *t = nullptr;
};

// benchmarks.cpp:248-530 (the Usui wear logger of :503-507 is created by main)
Setup setup_ref_cut(int ny, double rake, double clear, double chamfer, double speed, double feed) {
    phys_constants phys{};
    corr_constants corr{};
    corr.wdeltap = 1.;
    joco_constants joco{};
    trml_constants trml_wp{}, trml_tool{};

    const double hi_x = 0.200, hi_y = 0.060;
    const double lo_x = 0.005, lo_y = hi_y - 3 * feed;
    const double dy = (hi_y - lo_y) / (ny - 1);
    const double dx = dy;
    const int nx = (int)((hi_x - lo_x) / dx);
    const double hdx = 1.5;

    phys.E = 1.1;
    phys.nu = 0.35;
    phys.rho0 = 4.429998;
    phys.G = phys.E / (2. * (1. + phys.nu));
    phys.K = 2.0 * phys.G * (1 + phys.nu) / (3 * (1 - 2 * phys.nu));
    phys.mass = dx * dx * phys.rho0;
    joco.A = 0.0085200;
    joco.B = 0.0033890;
    joco.C = 0.02754;
    joco.m = 0.5961;
    joco.n = 0.148;
    joco.Tref = 300.;
    joco.Tmelt = 1836.0000;
    joco.eps_dot_ref = 1e-6;
    joco.tanh_a = 1.1;
    joco.tanh_b = 0.4;
    joco.tanh_c = 12.;
    joco.tanh_d = 1.;
    const double rho_tool = 15.25;
    phys.mass_tool = dx * dx * rho_tool;
    trml_wp.cp = 580 * 1e-8;
    trml_wp.tq = 0.9;
    trml_wp.k = 7.3 * 1e-13;
    trml_wp.alpha = trml_wp.k / (phys.rho0 * trml_wp.cp);
    trml_wp.eta = 0.9;
    trml_wp.T_init = joco.Tref;
    trml_tool.cp = 292 * 1e-08;
    trml_tool.tq = .0;
    trml_tool.k = 88 * 1e-13;
    trml_tool.alpha = trml_tool.k / (rho_tool * trml_tool.cp);
    trml_tool.T_init = joco.Tref;
    corr.alpha = 1.;
    corr.beta = 1.;
    corr.eta = 0.1;
    corr.xspheps = 0.5;
    corr.stresseps = 0.3;
    {
        const double h1 = 1. / (hdx * dx);
        const double q = dx * h1;
        const double fac = 10 * (M_1_PI) / 7.0 * h1 * h1;
        corr.wdeltap = fac * (1 - 1.5 * q * q * (1 - 0.5 * q));
    }

    std::vector<float2_t> pos;
    double min_x = FLT_MAX, max_x = -FLT_MAX, min_y = FLT_MAX;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            pos.push_back(float2_t{i * dx + lo_x, j * dx + lo_y});
            min_x = fmin(min_x, pos.back().x);
            max_x = fmax(max_x, pos.back().x);
            min_y = fmin(min_y, pos.back().y);
        }
    const size_t n = pos.size();
    std::vector<float2_t> vel(n, float2_t{0., 0.});
    std::vector<double> rho(n, phys.rho0), h(n, hdx * dx), T(n, joco.Tref), fixed(n, 0.), tool_p(n, 0.);
    for (size_t i = 0; i < n; i++) {
        bool f = pos[i].y < min_y + 0.5 * dx;
        f = f || (pos[i].x < min_x + dx / 2 && pos[i].y < min_y + 0.5 * (hi_y - lo_y));
        f = f || (pos[i].x > max_x - dx);
        fixed[i] = f ? 1. : 0.;
    }

    const double vel_tool = (speed == 0.) ? 83.333328 * 1e-5 : speed;
    const double mu_fric = 0.35;
    const double nudge = 0.004;
    const vec2_t tl(-0.05 + nudge, 0.0986074);
    const double l_tool = -0.0086824 - -0.05;
    const double h_tool = 0.0986074 - 0.0555074;
    tool *t = new tool(tl, l_tool, h_tool, rake, clear, chamfer, mu_fric);
    {
        const double distance = t->front() - lo_x;
        const double time = distance / vel_tool;
        t->set_vel(vec2_t(vel_tool, 0.));
        t->update_tool(-time);
        const double current_feed = hi_y - t->low();
        const double correction_time = fabs(current_feed - feed) / vel_tool;
        const double sign = (current_feed > feed) ? 1 : -1.;
        t->set_vel(vec2_t(0., vel_tool));
        t->update_tool(correction_time * sign);
    }
    global_tool = t;
    t->set_vel(vec2_t(vel_tool, 0.));
    const double contact_alpha = 0.075;
    tool_gpu_set_up_tool(t, contact_alpha, phys);

    const std::vector<float2_t> samples = t->sample_tool(dx, joco.Tref, phys.rho0, hdx);
    double bbmin_tool_x = FLT_MAX, bbmax_tool_y = -FLT_MAX;
    for (const auto &p : samples) {
        bbmin_tool_x = fmin(bbmin_tool_x, p.x);
        bbmax_tool_y = fmax(bbmax_tool_y, p.y);
    }
    for (const auto &p : samples) {
        pos.push_back(p);
        vel.push_back(float2_t{0., 0.});
        rho.push_back(rho_tool);
        h.push_back(hdx * dx);
        fixed.push_back((p.x < bbmin_tool_x + dx / 2 || p.y > bbmax_tool_y - dx / 2) ? 1. : 0.);
        T.push_back(joco.Tref);
        tool_p.push_back(1.);
    }

    float2_t bbmin{FLT_MAX, FLT_MAX}, bbmax{-FLT_MAX, -FLT_MAX};
    for (const auto &p : pos) {
        bbmin.x = fmin(p.x, bbmin.x);
        bbmin.y = fmin(p.y, bbmin.y);
        bbmax.x = fmax(p.x, bbmax.x);
        bbmax.y = fmax(p.y, bbmax.y);
    }
    bbmin.x -= 1e-8;
    bbmin.y -= 1e-8;
    bbmax.x += 1e-8;
    bbmax.y += 1e-8;
    const bbox toolbb = t->safe_bb(1e-3);
    bbmin.x = fmin(bbmin.x, toolbb.bbmin_x);
    bbmin.y = fmin(bbmin.y, toolbb.bbmin_y);
    bbmax.x = fmax(bbmax.x, toolbb.bbmax_x) + 10 * dx;
    bbmax.y = fmax(bbmax.y, toolbb.bbmax_y);
    bbmax.y = fmax(bbmax.y, hi_y + hi_x - lo_x);

    Setup out;
    out.grid = new grid_gpu_green((int)pos.size(), bbmin, bbmax, hdx * dx);
    actions_setup_corrector_constants(corr);
    actions_setup_physical_constants(phys);
    actions_setup_johnson_cook_constants(joco);
    actions_setup_thermal_constants_wp(trml_wp);
    actions_setup_thermal_constants_tool(trml_tool);
    interactions_setup_geometry_constants(out.grid);
    interactions_setup_corrector_constants(corr);
    interactions_setup_physical_constants(phys);
    interactions_setup_thermal_constants_tool(trml_tool, global_tool);
    interactions_setup_thermal_constants_workpiece(trml_wp);
    global_dt = 0.32 * hdx * dx / (sqrt(phys.K / phys.rho0) + sqrt(vel_tool));
    std::printf("timestep chosen: %e\ncalculating with %zu regular and %zu tool particles\n", global_dt, n,
                samples.size());
    out.particles = new particle_gpu(pos.data(), vel.data(), rho.data(), T.data(), h.data(), fixed.data(),
                                     tool_p.data(), (unsigned)pos.size());
    out.t = t;
    return out;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 6 && argc != 11) {
        std::fprintf(stderr, "usage: %s NY STEPS FREQ fp64|fp32c OUT_DIR [rake clearance fillet speed_in_ui feed]\n",
                     argv[0]);
        return 2;
    }
    const int ny = std::atoi(argv[1]);
    const long steps = std::atol(argv[2]);
    const long freq = std::atol(argv[3]);
    const std::string prec = argv[4], out = argv[5];
    // mfree_gpu.cpp:155-162, 197 defaults
    double rake = 0., clear = 4.5, fillet = 0.00035, speed_in_ui = 300, feed = 0.01;
    if (argc == 11) {
        rake = std::atof(argv[6]);
        clear = std::atof(argv[7]);
        fillet = std::atof(argv[8]);
        speed_in_ui = std::atof(argv[9]);
        feed = std::atof(argv[10]);
    }
    if (ny < 2 || steps < 0 || freq <= 0 || (prec != "fp64" && prec != "fp32c")) {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    const double speed = 83.333328 * 1e-5 / 200 * speed_in_ui;
    try {
        Setup su = setup_ref_cut(ny, rake, clear, fillet, speed, feed);
        particle_gpu &particles = *su.particles;
        const unsigned N = particles.N;
        leap_frog stepper(N, (unsigned)su.grid->num_cell());
        if (prec == "fp32c") stepper.set_precision(SPHFEM_FP32C);
        std::FILE *ff = std::fopen((out + "/forces.txt").c_str(), "w");
        if (!ff) throw std::runtime_error("cannot write " + out + "/forces.txt");
        // mfree_gpu.cpp:269-347: dumps every freq steps into results/ (VTK of particles and tool,
        // force and wear logs) when REF_CUT_RESULTS=1; the Usui logger as benchmarks.cpp:503-507 sets it
        // up (tool velocity (0, v) as there)
        const char *res_env = std::getenv("REF_CUT_RESULTS");
        const bool results = res_env && std::string(res_env) == "1";
        tool_wear wear(1. / 100 * 7.8, 2500., N, phys_constants{}, vec2_t(0., speed));
        std::FILE *fp_force = nullptr, *fp_wear = nullptr;
        if (results) {
            if (std::system("mkdir -p results") != 0) throw std::runtime_error("cannot create results/");
            char buf[256];
            std::snprintf(buf, sizeof(buf), "results/force_%f_%f", fillet, feed);
            fp_force = std::fopen(buf, "w+");
            std::snprintf(buf, sizeof(buf), "results/wear_%f_%f", speed, feed);
            fp_wear = std::fopen(buf, "w+");
            if (!fp_force || !fp_wear) throw std::runtime_error("cannot open the log files in results/");
        }
        int print_iter = 0;
        for (long step = 0; step < steps; step++) {
            if (step % freq == 0) {
                float2_t F = report_force(&particles);
                std::fprintf(ff, "%ld %.17g %.17g\n", step, F.x, F.y);
                if (results) {
                    vtk_writer_write(&particles, print_iter);
                    vtk_writer_write(global_tool, print_iter);
                    std::fprintf(fp_force, "%.9g %.9g\n", F.x, F.y);
                    std::fflush(fp_force);
                    double wmin, wmax, wavg;
                    wear.eval_usui(&particles, freq * global_dt, wmin, wmax, wavg);
                    std::fprintf(fp_wear, "%e %e %e %e\n", wmin, wmax, wavg, wear.get_accum_wear());
                    std::fflush(fp_wear);
                }
                print_iter++;
            }
            stepper.step(&particles, su.grid);
        }
        if (fp_force) std::fclose(fp_force);
        if (fp_wear) std::fclose(fp_wear);
        float2_t F = report_force(&particles);
        std::fprintf(ff, "%ld %.17g %.17g\n", steps, F.x, F.y);
        std::fclose(ff);
        if (!check_cuda_error()) throw std::runtime_error(sphfem_last_error(sphfem_shim::ctx()));
        write_array(out + "/pos.bin", download(particles.pos, N));
        write_array(out + "/vel.bin", download(particles.vel, N));
        write_array(out + "/T.bin", download(particles.T, N));
        write_array(out + "/fc.bin", download(particles.fc, N));
        write_array(out + "/idx.bin", download(particles.idx, N));
        // host copy of the tool (moved by leap_frog::step) against the device copy
        sphfem_tool_geometry dev{};
        sphfem_shim::check(sphfem_rigid_tool_get_geometry(sphfem_shim::ctx(), su.t->gpu_handle(), &dev), "get");
        std::FILE *ft = std::fopen((out + "/tool.txt").c_str(), "w");
        if (!ft) throw std::runtime_error("cannot write " + out + "/tool.txt");
        const sphfem_tool_geometry &host = su.t->geometry();
        for (int i = 0; i < dev.num_seg; i++)
            std::fprintf(ft, "%.17g %.17g %.17g %.17g\n", host.seg[i].left.x, host.seg[i].left.y, dev.seg[i].left.x,
                         dev.seg[i].left.y);
        std::fclose(ft);
        std::printf("ref_cut_driver: %ld steps of %u particles (%s), F = (%.9e, %.9e)\n", steps, N, prec.c_str(), F.x,
                    F.y);
        delete su.particles;
        delete su.grid;
        global_tool = nullptr;
        delete su.t;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "ref_cut_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
