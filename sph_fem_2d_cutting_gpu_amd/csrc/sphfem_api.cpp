// C ABI of libsphfem_hip: context, constants, memory, and the per-step orchestration.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sphfem_internal.h"
#include "sphfem_rigid.h"

using namespace sphfem;

namespace {

enum : unsigned { SET_PHYS = 1, SET_CORR = 2, SET_JOCO = 4, SET_TRML_WP = 8, SET_TOOLC = 16, SET_DT = 32 };

// an input out of the range an entry point can handle (reported as SPHFEM_ERR_RANGE)
struct RangeError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <class F>
int guarded(sphfem_ctx *ctx, F &&f) {
    if (!ctx) return SPHFEM_ERR_ARG;
    try {
        SPHFEM_HIP_CHECK(hipSetDevice(ctx->device));
        int rc = f();
        if (rc == SPHFEM_OK) {
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) {
                ctx->err = std::string("kernel launch: ") + hipGetErrorString(e);
                return SPHFEM_ERR_HIP;
            }
        }
        return rc;
    } catch (const RangeError &e) {
        ctx->err = e.what();
        return SPHFEM_ERR_RANGE;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return SPHFEM_ERR_HIP;
    }
}

int fail(sphfem_ctx *ctx, int code, const std::string &msg) {
    ctx->err = msg;
    return code;
}

template <class T>
T *dalloc(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    SPHFEM_HIP_CHECK(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T *>(p);
}

template <class T>
void dfree(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

template <class T>
T *dupload(hipStream_t s, const T *host, size_t n) {
    T *d = dalloc<T>(n);
    if (host && n) SPHFEM_HIP_CHECK(hipMemcpyAsync(d, host, n * sizeof(T), hipMemcpyHostToDevice, s));
    return d;
}

// allocate the 26 arrays of particle_gpu(N) (particle_gpu.cu:11-104)
void alloc_particles(hipStream_t s, unsigned N, sphfem_particles &p) {
    std::memset(&p, 0, sizeof(p));
    p.N = N;
    p.pos = dalloc<sphfem_double2>(N);
    p.vel = dalloc<sphfem_double2>(N);
    p.h = dalloc<double>(N);
    p.rho = dalloc<double>(N);
    p.p = dalloc<double>(N);
    p.S = dalloc<sphfem_double4>(N);
    p.R = dalloc<sphfem_double4>(N);
    p.fc = dalloc<sphfem_double2>(N);
    p.ft = dalloc<sphfem_double2>(N);
    p.n = dalloc<sphfem_double2>(N);
    p.fixed = dalloc<double>(N);
    p.blanked = dalloc<double>(N);
    p.tool_particle = dalloc<double>(N);
    p.eps_pl = dalloc<double>(N);
    p.eps_pl_dot = dalloc<double>(N);
    p.T = dalloc<double>(N);
    p.pos_t = dalloc<sphfem_double2>(N);
    p.vel_t = dalloc<sphfem_double2>(N);
    p.rho_t = dalloc<double>(N);
    p.S_t = dalloc<sphfem_double4>(N);
    p.T_t = dalloc<double>(N);
    p.S_der = dalloc<sphfem_double4>(N);
    p.v_der = dalloc<sphfem_double4>(N);
    p.idx = dalloc<int>(N);
    p.hash = dalloc<int>(N);
    p.on_seg = dalloc<int>(N);
    // zero everything (doubles and ints alike)
    auto z = [&](void *ptr, size_t bytes) { SPHFEM_HIP_CHECK(hipMemsetAsync(ptr, 0, bytes, s)); };
    z(p.pos, 16ul * N); z(p.vel, 16ul * N); z(p.h, 8ul * N); z(p.rho, 8ul * N); z(p.p, 8ul * N);
    z(p.S, 32ul * N); z(p.R, 32ul * N); z(p.fc, 16ul * N); z(p.ft, 16ul * N); z(p.n, 16ul * N);
    z(p.fixed, 8ul * N); z(p.blanked, 8ul * N); z(p.tool_particle, 8ul * N); z(p.eps_pl, 8ul * N);
    z(p.eps_pl_dot, 8ul * N); z(p.T, 8ul * N); z(p.pos_t, 16ul * N); z(p.vel_t, 16ul * N);
    z(p.rho_t, 8ul * N); z(p.S_t, 32ul * N); z(p.T_t, 8ul * N); z(p.S_der, 32ul * N); z(p.v_der, 32ul * N);
    z(p.hash, 4ul * N); z(p.on_seg, 4ul * N);
    std::vector<int> seq(N);
    for (unsigned i = 0; i < N; i++) seq[i] = (int)i;
    SPHFEM_HIP_CHECK(hipMemcpyAsync(p.idx, seq.data(), 4ul * N, hipMemcpyHostToDevice, s));
    SPHFEM_HIP_CHECK(hipStreamSynchronize(s));
}

void free_particles(sphfem_particles &p) {
    void **f = reinterpret_cast<void **>(&p);
    const int nptr = (int)((offsetof(sphfem_particles, N)) / sizeof(void *));
    for (int k = 0; k < nptr; k++) {
        if (f[k]) (void)hipFree(f[k]);
        f[k] = nullptr;
    }
}

// grid_gpu_green.cu:152-182: swap the 20 gathered arrays with the buffer
void swap_gathered(sphfem_particles &p, sphfem_particles &b) {
    std::swap(p.pos, b.pos); std::swap(p.vel, b.vel); std::swap(p.fc, b.fc); std::swap(p.ft, b.ft);
    std::swap(p.h, b.h); std::swap(p.n, b.n); std::swap(p.rho, b.rho); std::swap(p.S, b.S);
    std::swap(p.eps_pl, b.eps_pl); std::swap(p.eps_pl_dot, b.eps_pl_dot); std::swap(p.T, b.T);
    std::swap(p.fixed, b.fixed); std::swap(p.blanked, b.blanked); std::swap(p.tool_particle, b.tool_particle);
    std::swap(p.pos_t, b.pos_t); std::swap(p.vel_t, b.vel_t); std::swap(p.rho_t, b.rho_t); std::swap(p.S_t, b.S_t);
    std::swap(p.T_t, b.T_t); std::swap(p.on_seg, b.on_seg);
}

// FP32C step: every gathered array but the rates (the FP32 derivative carry holds those)
void swap_gathered_state(sphfem_particles &p, sphfem_particles &b) {
    std::swap(p.pos, b.pos); std::swap(p.vel, b.vel); std::swap(p.fc, b.fc); std::swap(p.ft, b.ft);
    std::swap(p.h, b.h); std::swap(p.n, b.n); std::swap(p.rho, b.rho); std::swap(p.S, b.S);
    std::swap(p.eps_pl, b.eps_pl); std::swap(p.eps_pl_dot, b.eps_pl_dot); std::swap(p.T, b.T);
    std::swap(p.fixed, b.fixed); std::swap(p.blanked, b.blanked); std::swap(p.tool_particle, b.tool_particle);
    std::swap(p.on_seg, b.on_seg);
}

// after the fused step: the arrays written in the sorted buffer become the particle arrays
void swap_fused(sphfem_particles &p, sphfem_particles &b) {
    std::swap(p.pos_t, b.pos_t); std::swap(p.vel_t, b.vel_t); std::swap(p.rho_t, b.rho_t); std::swap(p.S_t, b.S_t);
    std::swap(p.T_t, b.T_t); std::swap(p.fc, b.fc); std::swap(p.ft, b.ft); std::swap(p.n, b.n);
    std::swap(p.on_seg, b.on_seg); std::swap(p.h, b.h); std::swap(p.fixed, b.fixed);
    std::swap(p.tool_particle, b.tool_particle); std::swap(p.p, b.p); std::swap(p.R, b.R);
}

}  // namespace

namespace sphfem {

static bool timed(const sphfem_ctx *ctx, const char *name) {
    if (!ctx->timing.enabled) return false;
    if (ctx->timing.only.empty()) return true;
    return ("," + ctx->timing.only + ",").find("," + std::string(name) + ",") != std::string::npos;
}

void timing_begin(sphfem_ctx *ctx, const char *name, hipEvent_t *ev) {
    *ev = nullptr;
    if (!timed(ctx, name)) return;
    (void)hipEventCreate(ev);
    (void)hipEventRecord(*ev, ctx->stream);
}

void timing_end(sphfem_ctx *ctx, const char *name, hipEvent_t ev) {
    if (!ctx->timing.enabled || !ev) return;
    hipEvent_t b;
    (void)hipEventCreate(&b);
    (void)hipEventRecord(b, ctx->stream);
    ctx->timing.pending[name].push_back({ev, b});
}

}  // namespace sphfem

static void timing_collect(sphfem_ctx *ctx) {
    for (auto &kv : ctx->timing.pending) {
        for (auto &r : kv.second) {
            (void)hipEventSynchronize(r.b);
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, r.a, r.b);
            ctx->timing.total_ms[kv.first] += ms;
            ctx->timing.count[kv.first] += 1;
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        kv.second.clear();
    }
}

// Block partials of sphfem_report_force_noalloc: one (x, y) pair per reduction block, in a
// file-scope device array instead of memory allocated per call.  Each GPU holds its own copy until
// the process ends.  Host code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): it uses the address hipGetSymbolAddress returns for the current
// device, looked up once per device when a context is created there.
static __device__ __align__(256) double scratch[2 * SPHFEM_FORCE_BLOCKS];

static std::mutex g_scratch_mu;
static std::map<int, double *> g_scratch_ptr;   // device -> address of `scratch` on that device

// with ctx->device current
static int lookup_scratch(sphfem_ctx *ctx) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(ctx->device);
    if (it == g_scratch_ptr.end()) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(ctx, SPHFEM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        it = g_scratch_ptr.emplace(ctx->device, static_cast<double *>(ptr)).first;
    }
    ctx->force_scratch = it->second;
    return SPHFEM_OK;
}

extern "C" {

int sphfem_create(int device, sphfem_ctx **out) {
    if (!out) return SPHFEM_ERR_ARG;
    *out = nullptr;
    auto *ctx = new sphfem_ctx();
    ctx->device = device;
    int rc = guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->d_flags = dalloc<int>(SPHFEM_NUM_FLAGS);
        SPHFEM_HIP_CHECK(hipMemsetAsync(ctx->d_flags, 0, SPHFEM_NUM_FLAGS * sizeof(int), ctx->stream));
        ctx->c.rr2_min = 1e-8;
        return lookup_scratch(ctx);
    });
    if (rc != SPHFEM_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return SPHFEM_OK;
}

int sphfem_destroy(sphfem_ctx *ctx) {
    if (!ctx) return SPHFEM_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    timing_collect(ctx);
    dfree(ctx->d_flags);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SPHFEM_OK;
}

const char *sphfem_last_error(sphfem_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int sphfem_synchronize(sphfem_ctx *ctx) {
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return SPHFEM_OK;
    });
}

int sphfem_check_error(sphfem_ctx *ctx) {
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        int f = 0;
        SPHFEM_HIP_CHECK(hipMemcpy(&f, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost));
        if (f & 1) return fail(ctx, SPHFEM_ERR_RANGE, "a particle left the cell grid (hash clamped)");
        if (f & 6) return fail(ctx, SPHFEM_ERR_STATE, "inconsistent sort or cell table detected");
        return SPHFEM_OK;
    });
}

// The bounds of the plasticity's early exit (DevConsts::jc_svm_below): with A > 0 and B, C, n, m >= 0,
// every factor of plasticity.cuh:21-39 is bounded below for eps_pl, eps_pl_dot >= 0 -- A + B eps^n >= A,
// the rate factor >= 1 -- and the thermal factor 1 - theta^m >= 1 - 0.99^m / 2 > 1/2 for theta up to
// 0.99 * 0.5^(1/m), so sigma_yield > A / 2 there; 0.49 A leaves a margin far beyond the rounding of
// the evaluation.  Off for constants outside that range.
static void joco_early_exit_bounds(sphfem::DevConsts &c) {
    const sphfem_joco_constants &j = c.joco;
    c.jc_svm_below = -1.;
    const bool ok = j.A > 0. && j.B >= 0. && j.C >= 0. && j.n > 0. && j.m > 0. && j.eps_dot_ref > 0. &&
                    j.Tmelt > j.Tref && std::isfinite(j.A) && std::isfinite(j.B) && std::isfinite(j.C) &&
                    std::isfinite(j.n) && std::isfinite(j.m) && std::isfinite(j.Tmelt) && std::isfinite(j.Tref);
    if (!ok) return;
    const double theta = 0.99 * std::pow(0.5, 1. / j.m);
    c.jc_T_lo = j.Tref;
    c.jc_T_hi = j.Tref + theta * (j.Tmelt - j.Tref);
    c.jc_svm_below = 0.49 * j.A;
}

#define SETC(fn, field, bit, extra)                                           \
    int fn(sphfem_ctx *ctx, const decltype(sphfem::DevConsts::field) *c) {    \
        if (!ctx || !c) return SPHFEM_ERR_ARG;                                \
        ctx->c.field = *c;                                                    \
        ctx->set_mask |= bit;                                                 \
        extra;                                                                \
        return SPHFEM_OK;                                                     \
    }

SETC(sphfem_setup_physical_constants, phys, SET_PHYS, {
    const double c0 = (double)sqrtf((float)(c->K / c->rho0));
    ctx->c.eos_c02 = c0 * c0;
})
SETC(sphfem_setup_corrector_constants, corr, SET_CORR, )
SETC(sphfem_setup_johnson_cook_constants, joco, SET_JOCO, ctx->c.plastic = 1; joco_early_exit_bounds(ctx->c))
SETC(sphfem_setup_thermal_constants_wp, trml_wp, SET_TRML_WP, if (c->eta != 0.) ctx->c.fric_heat = 1)
SETC(sphfem_setup_thermal_constants_tool, trml_tool, 0, ctx->c.thermal_bc = (c->alpha != 0.) ? 1 : 0)
SETC(sphfem_setup_tool_constants, tc, SET_TOOLC, )
SETC(sphfem_setup_wear_constants, wear, 0, )
#undef SETC

int sphfem_set_dt(sphfem_ctx *ctx, double dt) {
    if (!ctx || !(dt > 0.)) return SPHFEM_ERR_ARG;
    ctx->c.dt = dt;
    ctx->set_mask |= SET_DT;
    return SPHFEM_OK;
}

int sphfem_set_wear(sphfem_ctx *ctx, int wear_pro) {
    if (!ctx) return SPHFEM_ERR_ARG;
    ctx->c.wear_pro = wear_pro ? 1 : 0;
    return SPHFEM_OK;
}

int sphfem_set_kernel_cutoff(sphfem_ctx *ctx, double rr2_min) {
    if (!ctx || !(rr2_min >= 0.)) return SPHFEM_ERR_ARG;
    ctx->c.rr2_min = rr2_min;
    return SPHFEM_OK;
}

int sphfem_malloc(sphfem_ctx *ctx, void **ptr, size_t bytes) {
    if (!ptr) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipMalloc(ptr, bytes ? bytes : 1));
        return SPHFEM_OK;
    });
}

int sphfem_free(sphfem_ctx *ctx, void *ptr) {
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (ptr) SPHFEM_HIP_CHECK(hipFree(ptr));
        return SPHFEM_OK;
    });
}

int sphfem_memcpy_htod(sphfem_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if ((!dst || !src) && bytes) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return SPHFEM_OK;
    });
}

int sphfem_memcpy_dtoh(sphfem_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if ((!dst || !src) && bytes) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return SPHFEM_OK;
    });
}

int sphfem_particles_alloc(sphfem_ctx *ctx, unsigned int N, sphfem_particles *out) {
    if (!out || N == 0) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        alloc_particles(ctx->stream, N, *out);
        return SPHFEM_OK;
    });
}

int sphfem_particles_free(sphfem_ctx *ctx, sphfem_particles *p) {
    if (!p) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        free_particles(*p);
        return SPHFEM_OK;
    });
}

// ---- grid (grid.cu:130-149, grid_gpu_green.cu:220-268, grid_gpu_rothlin.cu:251-285)
// device arrays of a grid with g->max_cell cells at most
static void grid_alloc(sphfem_ctx *ctx, sphfem_grid *g, int kind, int num_part) {
    g->num_part = num_part;
    int bits = 1;
    while (bits < 31 && (1ll << bits) < (long long)g->max_cell) bits++;
    g->key_bits = bits;
    g->cell_start = dalloc<int>(g->max_cell);
    g->cell_end = dalloc<int>(g->max_cell);
    SPHFEM_HIP_CHECK(hipMemsetAsync(g->cell_end, 0, sizeof(int) * (size_t)g->max_cell, ctx->stream));
    launch_fill_int(ctx, g->cell_start, g->max_cell, -1);
    g->keys_alt = dalloc<int>(num_part);
    g->vals_alt = dalloc<int>(num_part);
    g->nblocks = nblk(num_part, 4096);
    g->hist = dalloc<int>(256ul * g->nblocks);
    g->digit_tot = dalloc<int>(256);
    const int nw = nblk(num_part, 64);
    g->cmask = dalloc<unsigned long long>(nw);
    g->wprefix = dalloc<int>(nw + 1);
    g->bprefix = dalloc<int>(g->nblocks + 1);
    g->clist = dalloc<int2>(SPHFEM_RESORT_CAP);
    g->csorted = dalloc<unsigned long long>(SPHFEM_RESORT_CAP);
    g->cdest = dalloc<int4>(SPHFEM_RESORT_CAP);
    alloc_particles(ctx->stream, (unsigned)num_part, g->buf);
    g->first_sort = true;
    g->kind = kind;
    if (kind == SPHFEM_GRID_ROTHLIN) {
        g->perm = dalloc<int>(num_part);
        g->iota = dalloc<int>(num_part);
        launch_iota(ctx, g->iota, num_part);
        g->first_sort = false;   // every Rothlin sort is a full sort
    }
    if (!g->locked) {
        g->bbox_part = dalloc<double>(5 * SPHFEM_BBOX_BLOCKS);
        g->bbox_out = dalloc<double>(5);
    }
}

static int grid_create(sphfem_ctx *ctx, int kind, int num_part, sphfem_double2 bbmin, sphfem_double2 bbmax,
                       double h, sphfem_grid **out) {
    if (!out || num_part <= 0 || !(h > 0.)) return SPHFEM_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        auto *g = new sphfem_grid();
        double bx0 = bbmin.x - 1e-6, by0 = bbmin.y - 1e-6;
        double bx1 = bbmax.x + 1e-6, by1 = bbmax.y + 1e-6;
        double kernel_width = 2.;
        double dx = h * kernel_width / 2.;
        double nxd = std::ceil((bx1 - bx0) / dx), nyd = std::ceil((by1 - by0) / dx);
        if (!(nxd * nyd < 2.0e9)) {
            delete g;
            return fail(ctx, SPHFEM_ERR_ARG, "grid too large");
        }
        g->geom.nx = (int)nxd;
        g->geom.ny = (int)nyd;
        g->geom.bbmin_x = bx0;
        g->geom.bbmin_y = by0;
        g->geom.dx = dx;
        g->bbmax_x = bx1;
        g->bbmax_y = by1;
        g->num_cell = g->geom.nx * g->geom.ny;
        g->max_cell = g->num_cell;   // grid.cu:146
        g->locked = true;
        grid_alloc(ctx, g, kind, num_part);
        ctx->c.geom = g->geom;
        *out = g;
        return SPHFEM_OK;
    });
}

// grid_base(max_cell, num_part) (grid.cu:151-153): geometry set by each update_geometry
static int grid_create_unlocked(sphfem_ctx *ctx, int kind, int max_cell, int num_part, sphfem_grid **out) {
    if (!out || num_part <= 0 || max_cell <= 0) return SPHFEM_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        auto *g = new sphfem_grid();
        g->locked = false;
        g->max_cell = max_cell;
        g->num_cell = max_cell;
        grid_alloc(ctx, g, kind, num_part);
        *out = g;
        return SPHFEM_OK;
    });
}

// grid.cu:88-127: bounding box of all particles less/plus 1e-6, cell size max(h) * kernel_width / 2,
// ceil counts; the reference asserts num_cell < max_cell.  Reads the reduction back to the host
// (as the reference does), so it synchronises the stream.
static int update_geometry(sphfem_ctx *ctx, sphfem_grid *g, const sphfem_particles &p, double kernel_width) {
    if (g->locked) return SPHFEM_OK;
    launch_bbox(ctx, p, g->bbox_part, g->bbox_out);
    double b[5];
    SPHFEM_HIP_CHECK(hipMemcpyAsync(b, g->bbox_out, sizeof(b), hipMemcpyDeviceToHost, ctx->stream));
    SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const double bx0 = b[0] - 1e-6, by0 = b[1] - 1e-6;
    const double bx1 = b[2] + 1e-6, by1 = b[3] + 1e-6;
    const double dx = b[4] * kernel_width / 2;
    const double nxd = std::ceil((bx1 - bx0) / dx), nyd = std::ceil((by1 - by0) / dx);
    if (!(dx > 0.) || !std::isfinite(nxd) || !std::isfinite(nyd))
        return fail(ctx, SPHFEM_ERR_RANGE, "update_geometry: non-finite bounding box or h");
    if (!(nxd * nyd < (double)g->max_cell))
        return fail(ctx, SPHFEM_ERR_RANGE, "update_geometry: the bounding box needs max_cell cells or more");
    sphfem_geom_constants ng{(int)nxd, (int)nyd, bx0, by0, dx};
    if (ng.nx != g->geom.nx || ng.ny != g->geom.ny || ng.bbmin_x != g->geom.bbmin_x ||
        ng.bbmin_y != g->geom.bbmin_y || ng.dx != g->geom.dx) {
        // the previous keys belong to another geometry: the next Green sort is a full one
        if (g->kind == SPHFEM_GRID_GREEN) g->first_sort = true;
    }
    g->geom = ng;
    g->num_cell = ng.nx * ng.ny;
    g->bbmax_x = bx1;
    g->bbmax_y = by1;
    ctx->c.geom = g->geom;
    return SPHFEM_OK;
}

int sphfem_grid_green_create_unlocked(sphfem_ctx *ctx, int max_cell, int num_part, sphfem_grid **out) {
    return grid_create_unlocked(ctx, SPHFEM_GRID_GREEN, max_cell, num_part, out);
}

int sphfem_grid_rothlin_create_unlocked(sphfem_ctx *ctx, int max_cell, int num_part, sphfem_grid **out) {
    return grid_create_unlocked(ctx, SPHFEM_GRID_ROTHLIN, max_cell, num_part, out);
}

int sphfem_grid_update_geometry(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles *p, double kernel_width) {
    if (!g || !p || (int)p->N != g->num_part || !(kernel_width > 0.)) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] { return update_geometry(ctx, g, *p, kernel_width); });
}

int sphfem_grid_bounds(const sphfem_grid *g, sphfem_double2 *bbmax, int *locked, int *max_cell) {
    if (!g) return SPHFEM_ERR_ARG;
    if (bbmax) *bbmax = {g->bbmax_x, g->bbmax_y};
    if (locked) *locked = g->locked ? 1 : 0;
    if (max_cell) *max_cell = g->max_cell;
    return SPHFEM_OK;
}

int sphfem_grid_green_create(sphfem_ctx *ctx, int num_part, sphfem_double2 bbmin, sphfem_double2 bbmax, double h,
                             sphfem_grid **out) {
    return grid_create(ctx, SPHFEM_GRID_GREEN, num_part, bbmin, bbmax, h, out);
}

int sphfem_grid_rothlin_create(sphfem_ctx *ctx, int num_part, sphfem_double2 bbmin, sphfem_double2 bbmax,
                               double h, sphfem_grid **out) {
    return grid_create(ctx, SPHFEM_GRID_ROTHLIN, num_part, bbmin, bbmax, h, out);
}

int sphfem_grid_kind(const sphfem_grid *g, int *kind) {
    if (!g || !kind) return SPHFEM_ERR_ARG;
    *kind = g->kind;
    return SPHFEM_OK;
}

int sphfem_grid_destroy(sphfem_ctx *ctx, sphfem_grid *g) {
    if (!g) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        dfree(g->cell_start); dfree(g->cell_end); dfree(g->keys_alt); dfree(g->vals_alt);
        dfree(g->hist); dfree(g->digit_tot);
        dfree(g->cmask); dfree(g->wprefix); dfree(g->bprefix); dfree(g->clist); dfree(g->csorted); dfree(g->cdest);
        dfree(g->perm); dfree(g->iota); dfree(g->bbox_part); dfree(g->bbox_out);
        free_particles(g->buf);
        delete g;
        return SPHFEM_OK;
    });
}

int sphfem_grid_geometry(const sphfem_grid *g, sphfem_geom_constants *geom, int *num_cell) {
    if (!g) return SPHFEM_ERR_ARG;
    if (geom) *geom = g->geom;
    if (num_cell) *num_cell = g->num_cell;
    return SPHFEM_OK;
}

int sphfem_grid_assign_hashes(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles *p) {
    if (!g || !p || (int)p->N != g->num_part) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (!(g->geom.dx > 0.)) return fail(ctx, SPHFEM_ERR_STATE, "unlocked grid: update_geometry first");
        ctx->c.geom = g->geom;
        launch_hash(ctx, g, *p, false, nullptr);
        // p.hash no longer holds the keys of the (Green) cell table
        if (g->kind == SPHFEM_GRID_GREEN) g->first_sort = true;
        return SPHFEM_OK;
    });
}

int sphfem_grid_sort(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles *p) {
    if (!g || !p || (int)p->N != g->num_part) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (g->kind == SPHFEM_GRID_ROTHLIN) {
            launch_rothlin_sort(ctx, g, *p);
            return SPHFEM_OK;
        }
        launch_radix_sort(ctx, g, *p, false);
        launch_fill_int(ctx, g->cell_start, g->num_cell, -1);
        launch_reorder_full(ctx, g, *p);
        swap_gathered(*p, g->buf);
        g->first_sort = false;
        return SPHFEM_OK;
    });
}

int sphfem_grid_set_resort_limit(sphfem_ctx *ctx, sphfem_grid *g, int max_changed) {
    if (!ctx || !g || max_changed < 0) return SPHFEM_ERR_ARG;
    g->resort_limit = std::min(max_changed, SPHFEM_RESORT_CAP);
    return SPHFEM_OK;
}

int sphfem_grid_resort_stats(sphfem_ctx *ctx, int *mode, int *changed) {
    if (!ctx) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        int h[8];
        SPHFEM_HIP_CHECK(hipMemcpyAsync(h, ctx->d_flags, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (mode) *mode = h[3];
        if (changed) *changed = h[4];
        return SPHFEM_OK;
    });
}

int sphfem_grid_get_cells(sphfem_ctx *ctx, sphfem_grid *g, const sphfem_particles *p, int *cell_start,
                          int *cell_end) {
    if (!g || !cell_start || !cell_end) return SPHFEM_ERR_ARG;
    if (g->kind == SPHFEM_GRID_ROTHLIN && (!p || (int)p->N != g->num_part)) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (g->kind == SPHFEM_GRID_ROTHLIN) {
            launch_cells_dense(ctx, p->hash, (int)p->N, g->num_cell, cell_start, cell_end);
            return SPHFEM_OK;
        }
        SPHFEM_HIP_CHECK(hipMemcpyAsync(cell_start, g->cell_start, sizeof(int) * (size_t)g->num_cell,
                                        hipMemcpyDeviceToDevice, ctx->stream));
        SPHFEM_HIP_CHECK(hipMemcpyAsync(cell_end, g->cell_end, sizeof(int) * (size_t)g->num_cell,
                                        hipMemcpyDeviceToDevice, ctx->stream));
        return SPHFEM_OK;
    });
}

int sphfem_interactions_setup_geometry_constants(sphfem_ctx *ctx, const sphfem_grid *g) {
    if (!ctx || !g) return SPHFEM_ERR_ARG;
    ctx->c.geom = g->geom;
    return SPHFEM_OK;
}

static int need(sphfem_ctx *ctx, unsigned mask) {
    if ((ctx->set_mask & mask) != mask) return fail(ctx, SPHFEM_ERR_STATE, "constants not set up");
    return SPHFEM_OK;
}

int sphfem_interactions_monaghan(sphfem_ctx *ctx, sphfem_particles *p, const int *cell_start, const int *cell_end,
                                 int num_cell) {
    if (!ctx || !p || !cell_start || !cell_end || num_cell != ctx->c.geom.nx * ctx->c.geom.ny) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_PHYS | SET_CORR)) return rc;
        launch_interactions_ref(ctx, *p, cell_start, cell_end);
        return SPHFEM_OK;
    });
}

int sphfem_interactions_monaghan_fp32c(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p,
                                       const int *cell_start, const int *cell_end, int num_cell) {
    if (!ctx || !lf || !p || !cell_start || !cell_end || num_cell != ctx->c.geom.nx * ctx->c.geom.ny)
        return SPHFEM_ERR_ARG;
    if (lf->precision != SPHFEM_FP32C || !lf->rec) return fail(ctx, SPHFEM_ERR_STATE, "leap_frog is not in the FP32C form");
    if (p->N > lf->num_part) return fail(ctx, SPHFEM_ERR_ARG, "more particles than the leap_frog was made for");
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_PHYS | SET_CORR)) return rc;
        launch_interactions_fp32c(ctx, *p, lf, cell_start, cell_end, num_cell);
        return SPHFEM_OK;
    });
}

int sphfem_interactions_heat_pse(sphfem_ctx *ctx, sphfem_particles *p, const int *cell_start, const int *cell_end,
                                 int num_cell) {
    if (!ctx || !p || !cell_start || !cell_end || num_cell != ctx->c.geom.nx * ctx->c.geom.ny) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_PHYS | SET_TRML_WP)) return rc;
        if (ctx->c.trml_wp.alpha == 0.) return SPHFEM_OK;     // the workpiece is not thermal
        launch_interactions_pse(ctx, *p, cell_start, cell_end);
        return SPHFEM_OK;
    });
}

#define ACTION(fn, which, mask)                                   \
    int fn(sphfem_ctx *ctx, sphfem_particles *p) {                \
        if (!p) return SPHFEM_ERR_ARG;                            \
        return guarded(ctx, [&] {                                 \
            if (int rc = need(ctx, mask)) return rc;              \
            launch_actions(ctx, which, *p);                       \
            return SPHFEM_OK;                                     \
        });                                                       \
    }

ACTION(sphfem_material_eos, ACT_EOS, SET_PHYS)
ACTION(sphfem_corrector_artificial_stress, ACT_ARTSTRESS, SET_CORR)
ACTION(sphfem_material_stress_rate_jaumann, ACT_JAUMANN, SET_PHYS)
ACTION(sphfem_contmech_continuity, ACT_CONTINUITY, 0)
ACTION(sphfem_contmech_momentum, ACT_MOMENTUM, SET_PHYS)
ACTION(sphfem_contmech_advection, ACT_ADVECTION, 0)
ACTION(sphfem_perform_boundary_conditions, ACT_BC, 0)
ACTION(sphfem_debug_invalidate, ACT_INVALIDATE, 0)
#undef ACTION

int sphfem_plasticity_johnson_cook(sphfem_ctx *ctx, sphfem_particles *p) {
    if (!p) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (!ctx->c.plastic) return SPHFEM_OK;   // actions_gpu.cu:875
        if (int rc = need(ctx, SET_PHYS | SET_DT)) return rc;
        launch_actions(ctx, ACT_PLASTICITY, *p);
        return SPHFEM_OK;
    });
}

int sphfem_perform_boundary_conditions_thermal(sphfem_ctx *ctx, sphfem_particles *p) {
    if (!p) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (!ctx->c.thermal_bc) return SPHFEM_OK;   // actions_gpu.cu:882
        launch_actions(ctx, ACT_BC_THERMAL, *p);
        return SPHFEM_OK;
    });
}

int sphfem_material_fric_heat_gen_FEM_v2(sphfem_ctx *ctx, sphfem_particles *p, sphfem_tool *t) {
    if (!p || !t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_PHYS | SET_TRML_WP | SET_TOOLC | SET_DT)) return rc;
        launch_fric_heat(ctx, *p, t);
        return SPHFEM_OK;
    });
}

int sphfem_compute_contact_forces_FEM_tool_v3(sphfem_ctx *ctx, sphfem_particles *p, sphfem_tool *t) {
    if (!p || !t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_TRML_WP | SET_TOOLC | SET_DT)) return rc;
        launch_contact(ctx, *p, t);
        return SPHFEM_OK;
    });
}

int sphfem_wear_adjustment(sphfem_ctx *ctx, sphfem_tool *t) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_TOOLC | SET_DT)) return rc;
        launch_wear(ctx, t);
        return SPHFEM_OK;
    });
}

// ---- rigid (SPH) tool: host geometry (tool.cpp), device copy and per-step kernels (tool_gpu.cu)
}  // extern "C"

namespace {

sphfem_double2 d2(double x, double y) { return {x, y}; }

// tool.cpp:11-16
void solve_quad(double a, double b, double c, double &x1, double &x2) {
    x1 = (-b + sqrt(b * b - 4 * a * c)) / (2 * a);
    x2 = (-b - sqrt(b * b - 4 * a * c)) / (2 * a);
}

// tool.cpp:27-53: centre of the fillet of radius r on the bisector lm, at distance r from l1
sphfem_double2 fit_fillet(double r, const sphfem_line &lm, const sphfem_line &l1) {
    if (l1.vertical) {
        const sphfem_line par = {DBL_MAX, l1.b - r, 1};
        return rigid_intersect(par, lm);
    }
    const double A0 = lm.a, B0 = lm.b, a = l1.a, b = l1.b;
    const double A = a - A0, B = b - B0, C = r * sqrt(a * a + 1.);
    double x1, x2;
    solve_quad(A * A, 2 * A * B, B * B - C * C, x1, x2);
    const double xm = fmin(x1, x2);
    return {xm, lm.a * xm + lm.b};
}

// tool.cpp:55-62: closed polygon through the points
void construct_segments(sphfem_tool_geometry &g, const sphfem_double2 *pts, int n) {
    g.num_seg = n;
    for (int i = 0; i < n; i++) g.seg[i] = rigid_segment(pts[i], pts[i + 1 > n - 1 ? 0 : i + 1]);
}

// tool.cpp:64-93: the corner br rounded by a fillet tangent to tr-br and bl-br (5 points)
void construct_points_and_fillet(sphfem_tool_geometry &g, sphfem_double2 tl, sphfem_double2 tr, sphfem_double2 br,
                                 sphfem_double2 bl, double r, sphfem_double2 *pts) {
    const sphfem_double2 pm = br;
    const sphfem_double2 nt = rigid_normalize(d2(tr.x - br.x, tr.y - br.y));
    const sphfem_double2 nl = rigid_normalize(d2(bl.x - br.x, bl.y - br.y));
    const sphfem_double2 nm = rigid_normalize(d2(0.5 * (nt.x + nl.x), 0.5 * (nt.y + nl.y)));
    const sphfem_line lm = rigid_line(pm, d2(pm.x + nm.x, pm.y + nm.y));
    const sphfem_line l1 = rigid_line(tr, br), l2 = rigid_line(bl, br);
    const sphfem_double2 p = fit_fillet(r, lm, l1);
    const sphfem_double2 trc = rigid_closest_point(l1, p), blc = rigid_closest_point(l2, p);
    g.has_fillet = 1;
    g.fillet.r = r;
    g.fillet.t1 = rigid_atan2(p.y - trc.y, p.x - trc.x);
    g.fillet.t2 = rigid_atan2(p.y - blc.y, p.x - blc.x);
    g.fillet.p = p;
    pts[0] = tl; pts[1] = tr; pts[2] = trc; pts[3] = blc; pts[4] = bl;
}

// tool.cpp:316-331 and the tail of :335-369
void build_tool(sphfem_tool_geometry &g, sphfem_double2 tl, sphfem_double2 tr, sphfem_double2 br, sphfem_double2 bl,
                double r, double mu) {
    std::memset(&g, 0, sizeof(g));
    g.mu = mu;
    sphfem_double2 pts[SPHFEM_TOOL_MAX_SEG];
    if (r == 0.) {
        pts[0] = tl; pts[1] = tr; pts[2] = br; pts[3] = bl;
        construct_segments(g, pts, 4);
    } else {
        construct_points_and_fillet(g, tl, tr, br, bl, r, pts);
        construct_segments(g, pts, 5);
    }
    g.front = br.x;
}

bool geometry_ok(const sphfem_tool_geometry *g) {
    return g && g->num_seg >= 0 && g->num_seg <= SPHFEM_TOOL_MAX_SEG;
}

// tool.cpp:178-188
double geometry_low(const sphfem_tool_geometry &g) {
    if (g.has_fillet) return g.fillet.p.y - g.fillet.r;
    double y_min = DBL_MAX;
    for (int i = 0; i < g.num_seg; i++) y_min = fmin(y_min, g.seg[i].left.y);
    return y_min;
}

// tool.cpp:147-176
sphfem_bbox geometry_safe_bb(const sphfem_tool_geometry &g, double safety) {
    double x_min = DBL_MAX, x_max = -DBL_MAX, y_min = DBL_MAX, y_max = -DBL_MAX;
    for (int i = 0; i < g.num_seg; i++) {
        x_min = fmin(g.seg[i].left.x, x_min);
        x_max = fmax(g.seg[i].left.x, x_max);
        y_min = fmin(g.seg[i].left.y, y_min);
        y_max = fmax(g.seg[i].left.y, y_max);
    }
    if (g.has_fillet) {
        y_min = geometry_low(g);
        x_max = g.fillet.p.x + g.fillet.r;
    }
    x_min -= safety;
    y_min -= safety;
    x_max += safety;
    y_max += safety;
    return {x_min, x_max, y_min, y_max};
}

}  // namespace

extern "C" {

int sphfem_tool_geometry_from_points(sphfem_double2 tl, sphfem_double2 tr, sphfem_double2 br, sphfem_double2 bl,
                                     double r, double mu, sphfem_tool_geometry *out) {
    if (!out || !(r >= 0.)) return SPHFEM_ERR_ARG;
    build_tool(*out, tl, tr, br, bl, r, mu);
    return SPHFEM_OK;
}

// tool.cpp:335-406: tr and bl from the length and height, rake and clearance faces through tr and
// bl turned from straight down by the rake angle and by 90 - clearance (glm row-vector products
// down * R = (sin a, -cos a)), br where the two faces meet
int sphfem_tool_geometry_from_angles(sphfem_double2 tl, double length, double height, double rake_angle,
                                     double clearance_angle, double r, double mu, sphfem_tool_geometry *out) {
    if (!out || !(r >= 0.)) return SPHFEM_ERR_ARG;
    const sphfem_double2 tr = d2(tl.x + length, tl.y), bl = d2(tl.x, tl.y - height);
    const double alpha_rake = rake_angle * SPHFEM_RIGID_PI / 180.;
    const double alpha_free = (180 - 90 - clearance_angle) * SPHFEM_RIGID_PI / 180.;
    // (0, -1) * mat2(cos, -sin, sin, cos): x = 0 cos + (-1)(-sin), y = 0 sin + (-1) cos
    const double cr = cos(alpha_rake), sr = sin(alpha_rake), cf = cos(alpha_free), sf = sin(alpha_free);
    const sphfem_double2 dr = d2(0. * cr + (-1.) * (-sr), 0. * sr + (-1.) * cr);
    const sphfem_double2 df = d2(0. * cf + (-1.) * (-sf), 0. * sf + (-1.) * cf);
    const sphfem_double2 trc = d2(tr.x + dr.x, tr.y + dr.y), blc = d2(bl.x + df.x, bl.y + df.y);
    const sphfem_double2 br = rigid_intersect(rigid_line(tr, trc), rigid_line(bl, blc));
    build_tool(*out, tl, tr, br, bl, r, mu);
    return SPHFEM_OK;
}

int sphfem_tool_geometry_move(sphfem_tool_geometry *g, double dt) {
    if (!geometry_ok(g)) return SPHFEM_ERR_ARG;
    rigid_move(*g, dt);
    return SPHFEM_OK;
}

int sphfem_tool_geometry_low(const sphfem_tool_geometry *g, double *low) {
    if (!geometry_ok(g) || !low) return SPHFEM_ERR_ARG;
    *low = geometry_low(*g);
    return SPHFEM_OK;
}

int sphfem_tool_geometry_safe_bb(const sphfem_tool_geometry *g, double safety, sphfem_bbox *out) {
    if (!geometry_ok(g) || !out) return SPHFEM_ERR_ARG;
    *out = geometry_safe_bb(*g, safety);
    return SPHFEM_OK;
}

int sphfem_tool_geometry_inside(const sphfem_tool_geometry *g, sphfem_double2 pos, int *inside) {
    if (!geometry_ok(g) || !inside) return SPHFEM_ERR_ARG;
    *inside = rigid_inside(*g, pos.x, pos.y, g->has_fillet != 0) ? 1 : 0;
    return SPHFEM_OK;
}

// tool.cpp:112-135 (the lattice spans safe_bb(0.1 dx); counts truncate as the int conversion there)
int sphfem_tool_geometry_sample(const sphfem_tool_geometry *g, double dx, sphfem_double2 *out, long cap,
                                long *count) {
    if (!geometry_ok(g) || !count || !(dx > 0.) || cap < 0) return SPHFEM_ERR_ARG;
    const sphfem_bbox box = geometry_safe_bb(*g, 0.1 * dx);
    const double fx = (box.bbmax_x - box.bbmin_x) / dx + 1, fy = (box.bbmax_y - box.bbmin_y) / dx + 1;
    if (!(fx < 2e9 && fy < 2e9)) return SPHFEM_ERR_ARG;
    const int nx = (int)fx, ny = (int)fy;
    long k = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            const sphfem_double2 pos = d2(i * dx + box.bbmin_x, j * dx + box.bbmin_y);
            if (rigid_inside(*g, pos.x, pos.y, g->has_fillet != 0)) {
                if (out && k < cap) out[k] = pos;
                k++;
            }
        }
    *count = k;
    return SPHFEM_OK;
}

int sphfem_tool_gpu_set_up_tool(sphfem_ctx *ctx, const sphfem_tool_geometry *g, double contact_alpha,
                                const sphfem_phys_constants *phys, sphfem_rigid_tool **out) {
    if (!geometry_ok(g) || !phys || !out) return SPHFEM_ERR_ARG;
    if (g->has_fillet && g->num_seg != SPHFEM_TOOL_MAX_SEG)
        return fail(ctx, SPHFEM_ERR_ARG, "a fillet replaces the third of five segments");
    *out = nullptr;
    return guarded(ctx, [&] {
        RigidTool h{};
        h.g = *g;
        h.contact_alpha = contact_alpha;
        h.slave_mass = phys->mass;
        auto *rt = new sphfem_rigid_tool();
        try {
            rt->dev = dupload(ctx->stream, &h, 1);
            SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            dfree(rt->dev);
            delete rt;
            throw;
        }
        *out = rt;
        return SPHFEM_OK;
    });
}

int sphfem_rigid_tool_destroy(sphfem_ctx *ctx, sphfem_rigid_tool *rt) {
    if (!rt) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        dfree(rt->dev);
        delete rt;
        return SPHFEM_OK;
    });
}

int sphfem_rigid_tool_get_geometry(sphfem_ctx *ctx, sphfem_rigid_tool *rt, sphfem_tool_geometry *out) {
    if (!rt || !out) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        RigidTool h;
        SPHFEM_HIP_CHECK(hipMemcpyAsync(&h, rt->dev, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        *out = h.g;
        return SPHFEM_OK;
    });
}

int sphfem_rigid_tool_set_vel(sphfem_ctx *ctx, sphfem_rigid_tool *rt, sphfem_double2 vel) {
    if (!rt) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        const sphfem_double2 v = vel;
        SPHFEM_HIP_CHECK(hipMemcpyAsync(&rt->dev->g.vel, &v, sizeof(v), hipMemcpyHostToDevice, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return SPHFEM_OK;
    });
}

int sphfem_tool_gpu_update_tool(sphfem_ctx *ctx, sphfem_rigid_tool *rt, sphfem_particles *p) {
    if (!rt) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_DT)) return rc;
        launch_rigid_advance(ctx, rt);
        if (p && p->N) launch_rigid_tool_particles(ctx, *p, rt);
        return SPHFEM_OK;
    });
}

int sphfem_compute_contact_forces(sphfem_ctx *ctx, sphfem_particles *p, sphfem_rigid_tool *rt) {
    if (!p || !rt) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_DT)) return rc;
        if (p->N) launch_rigid_contact(ctx, *p, rt);
        return SPHFEM_OK;
    });
}

int sphfem_material_fric_heat_gen(sphfem_ctx *ctx, sphfem_particles *p, sphfem_double2 vel) {
    if (!p) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (!ctx->c.fric_heat) return SPHFEM_OK;   // actions_gpu.cu:957-959
        if (int rc = need(ctx, SET_PHYS | SET_TRML_WP | SET_DT)) return rc;
        if (p->N) launch_rigid_fric_heat(ctx, *p, vel);
        return SPHFEM_OK;
    });
}

int sphfem_interactions_setup_thermal_constants_tool(sphfem_ctx *ctx, const sphfem_trml_constants *c) {
    if (!ctx || !c) return SPHFEM_ERR_ARG;
    ctx->c.itool_alpha_pse = c->alpha;
    if (c->alpha != 0.) ctx->c.trml_itool = *c;
    return SPHFEM_OK;
}

// ---- tool
}  // extern "C"

// every index of a tool description inside its array before anything is copied or indexed: the
// host set-up below and the device kernels index with them unchecked
static bool tool_desc_valid(const sphfem_tool_desc *d, std::string &why) {
    const int nn = d->node_num, ne = d->ele_num, nz = d->nz_num, ns = d->seg_num;
    if (nn <= 0 || ne <= 0 || nz <= 0 || ns <= 0) return why = "empty tool description", false;
    if (!d->position || !d->node_num_el || !d->csr_row_ptr || !d->csr_col_ind || !d->K || !d->M || !d->f ||
        !d->T || !d->left || !d->right || !d->n || !d->marks)
        return why = "tool description: a required array is NULL", false;
    auto node = [&](int v) { return v >= 0 && v < nn; };
    for (int i = 0; i < nn; i++)
        if (!std::isfinite(d->position[i].x) || !std::isfinite(d->position[i].y) || !(d->M[i] > 0.))
            return why = "tool node " + std::to_string(i) + ": position not finite or mass not positive", false;
    for (int e = 0; e < ne; e++) {
        const sphfem_int4 el = d->node_num_el[e];
        if (!node(el.x) || !node(el.y) || !node(el.z) || el.w < 0)
            return why = "tool element " + std::to_string(e) + ": node id or colour out of range", false;
    }
    if (d->csr_row_ptr[0] != 0 || d->csr_row_ptr[nn] != nz) return why = "CSR row pointers do not span nz_num", false;
    for (int i = 0; i < nn; i++)
        if (d->csr_row_ptr[i + 1] < d->csr_row_ptr[i]) return why = "CSR row pointers decrease", false;
    for (int k = 0; k < nz; k++)
        if (!node(d->csr_col_ind[k])) return why = "CSR column index out of range", false;
    for (int q = 0; q < ns; q++) {
        const sphfem_int4 m = d->marks[q];
        if (m.x < 0 || m.x >= ne || !node(m.z) || !node(m.w))
            return why = "cutting segment " + std::to_string(q) + ": element or node out of range", false;
    }
    if (d->fixed_num < 0 || (d->fixed_num > 0 && (!d->fixed_elems || !d->fixed_flags)))
        return why = "fixed-boundary description incomplete", false;
    for (int j = 0; d->fixed_elems && j < d->fixed_num; j++)
        if (d->fixed_elems[j] < 0 || d->fixed_elems[j] >= ne) return why = "fixed-boundary element out of range", false;
    return true;
}

extern "C" {

int sphfem_tool_create(sphfem_ctx *ctx, const sphfem_tool_desc *d, sphfem_tool **out) {
    if (!d || !out) return SPHFEM_ERR_ARG;
    *out = nullptr;
    std::string why;
    if (!tool_desc_valid(d, why)) return ctx ? fail(ctx, SPHFEM_ERR_ARG, why) : SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        auto *t = new sphfem_tool();
        hipStream_t s = ctx->stream;
        const int nn = d->node_num, ne = d->ele_num, nz = d->nz_num, ns = d->seg_num;
        t->n_nodes = nn; t->n_elem = ne; t->nnz = nz; t->n_seg = ns;
        t->n_colors = d->color_total; t->n_cut_colors = d->color_cutting_edge;
        t->pos = dupload(s, d->position, nn);
        t->elem = dupload(s, d->node_num_el, ne);
        t->row_ptr = dupload(s, d->csr_row_ptr, nn + 1);
        t->col_ind = dupload(s, d->csr_col_ind, nz);
        t->K = dupload(s, d->K, nz);
        t->K_temp = dupload(s, d->K, nz);
        t->M = dupload(s, d->M, nn);
        std::vector<double> minv(nn);
        for (int i = 0; i < nn; i++) minv[i] = 1. / d->M[i];
        t->Minv = dupload(s, minv.data(), nn);
        t->f = dupload(s, d->f, nn);
        t->f_temp = dupload(s, d->f, nn);
        t->T = dupload(s, d->T, nn);
        t->T_old = dupload(s, d->T, nn);
        std::vector<double> zeros((size_t)std::max(nn, 4 * ns), 0.);
        t->k = dupload(s, zeros.data(), nn);
        t->left = dupload(s, d->left, ns);
        t->right = dupload(s, d->right, ns);
        t->n = dupload(s, d->n, ns);
        t->marks = dupload(s, d->marks, ns);
        t->heat_exchange = dupload(s, reinterpret_cast<const sphfem_double2 *>(zeros.data()), ns);
        t->fric_heat = dupload(s, zeros.data(), ns);
        t->wear_rate = dupload(s, reinterpret_cast<const sphfem_double2 *>(zeros.data()), ns);
        t->wear_nodes = dupload(s, zeros.data(), ns);
        t->physical_para = dupload(s, reinterpret_cast<const sphfem_double4 *>(zeros.data()), ns);
        t->sliding_force = dupload(s, zeros.data(), ns);
        // diagonal index per row
        std::vector<int> diag(nn, -1);
        for (int i = 0; i < nn; i++)
            for (int k = d->csr_row_ptr[i]; k < d->csr_row_ptr[i + 1]; k++)
                if (d->csr_col_ind[k] == i) diag[i] = k;
        for (int i = 0; i < nn; i++)
            if (diag[i] < 0) throw std::runtime_error("CSR row without diagonal entry");
        t->diag = dupload(s, diag.data(), nn);
        // Robin contributions per node in (colour, segment) order
        std::vector<std::vector<std::pair<long, int>>> rob(nn);
        for (int q = 0; q < ns; q++) {
            const sphfem_int4 m = d->marks[q];
            long key = (long)m.y * (1l << 31) + q;
            rob[m.z].push_back({key, q});
            if (m.w != m.z) rob[m.w].push_back({key, q});
        }
        std::vector<int2> nrob(nn, make_int2(-1, -1));
        for (int i = 0; i < nn; i++) {
            auto &v = rob[i];
            std::sort(v.begin(), v.end());
            if (v.size() > 2) throw std::runtime_error("a tool node belongs to more than two cutting segments");
            if (v.size() > 0) nrob[i].x = v[0].second;
            if (v.size() > 1) nrob[i].y = v[1].second;
        }
        t->node_robin = dupload(s, nrob.data(), nn);
        // nodal-shift contributions per node in segment order
        std::vector<std::vector<int2>> sh(nn);
        for (int q = 0; q < ns; q++) {
            sh[d->marks[q].z].push_back(make_int2(q, 0));
            sh[d->marks[q].w].push_back(make_int2(q, 1));
        }
        std::vector<int> shp(nn + 1, 0);
        std::vector<int2> shl;
        for (int i = 0; i < nn; i++) {
            std::sort(sh[i].begin(), sh[i].end(), [](int2 a, int2 b) { return a.x < b.x || (a.x == b.x && a.y < b.y); });
            for (auto e : sh[i]) shl.push_back(e);
            shp[i + 1] = (int)shl.size();
        }
        t->node_shift_ptr = dupload(s, shp.data(), nn + 1);
        t->node_shift = dupload(s, shl.data(), shl.size());
        // boundary-layer elements in (colour, element) order and their scatter lists
        std::vector<int> lay;
        for (int e = 0; e < ne; e++)
            if (d->layer && d->layer[e]) lay.push_back(e);
        std::stable_sort(lay.begin(), lay.end(),
                         [&](int a, int b) { return d->node_num_el[a].w < d->node_num_el[b].w; });
        t->n_layer = (int)lay.size();
        std::vector<std::vector<int>> nzl(nz), ndl(nn);
        auto find = [&](int row, int col) {
            for (int k = d->csr_row_ptr[row]; k < d->csr_row_ptr[row + 1]; k++)
                if (d->csr_col_ind[k] == col) return k;
            throw std::runtime_error("element entry missing from the CSR pattern");
        };
        for (int j = 0; j < t->n_layer; j++) {
            const sphfem_int4 el = d->node_num_el[lay[j]];
            const int nd[3] = {el.x, el.y, el.z};
            for (int a = 0; a < 3; a++) {
                ndl[nd[a]].push_back((j << 2) | a);
                for (int b = 0; b < 3; b++) nzl[find(nd[a], nd[b])].push_back((j << 4) | (a * 3 + b));
            }
        }
        std::vector<int> nzp(nz + 1, 0), nzv, ndp(nn + 1, 0), ndv;
        for (int k = 0; k < nz; k++) {
            for (int v : nzl[k]) nzv.push_back(v);
            nzp[k + 1] = (int)nzv.size();
        }
        for (int i = 0; i < nn; i++) {
            for (int v : ndl[i]) ndv.push_back(v);
            ndp[i + 1] = (int)ndv.size();
        }
        t->layer_elems = dupload(s, lay.data(), lay.size());
        t->layer_old = dalloc<double>(10ul * std::max(t->n_layer, 1));
        t->layer_new = dalloc<double>(10ul * std::max(t->n_layer, 1));
        t->nnz_layer_ptr = dupload(s, nzp.data(), nz + 1);
        t->nnz_layer = dupload(s, nzv.data(), nzv.size());
        t->node_layer_ptr = dupload(s, ndp.data(), nn + 1);
        t->node_layer = dupload(s, ndv.data(), ndv.size());
        sphfem_tool_kinematics k{};
        k.vel_x = d->vel.x; k.vel_y = d->vel.y;
        k.tl_x = d->tl_point_ref.x; k.tl_y = d->tl_point_ref.y;
        k.br_x = d->br_point.x; k.br_y = d->br_point.y;
        k.ref_dist = d->ref_dist; k.lowest_point = d->lowest_point; k.right_point = d->right_point;
        k.flank_retreat = 0.; k.flank_contact = d->lowest_point;
        t->kin = dupload(s, &k, 1);
        t->host_kin = k;
        if (d->node_flags && d->fixed_num > 0 && d->fixed_elems && d->fixed_flags) {
            // Dirichlet rows in the order do_dirichlet_bdc_p1q3 applies them (colour, fixed element,
            // node); a node met again is a no-op.  Free node x gathers the entries (F, x) of the rows
            // F that reach it, in that order.
            int max_colour = 0;
            for (int j = 0; j < d->fixed_num; j++) {
                const int e = d->fixed_elems[j];
                if (e < 0 || e >= ne) throw std::runtime_error("fixed-boundary element out of range");
                max_colour = std::max(max_colour, d->node_num_el[e].w);
            }
            std::vector<int> order, dir(nn, 0);
            for (int c = 1; c <= max_colour; c++)
                for (int j = 0; j < d->fixed_num; j++) {
                    const sphfem_int4 el = d->node_num_el[d->fixed_elems[j]];
                    if (el.w != c) continue;
                    const sphfem_int4 fl = d->fixed_flags[j];
                    int nf[3] = {0, 0, 0};
                    if (fl.x) nf[0] = nf[1] = 1;
                    if (fl.y) nf[1] = nf[2] = 1;
                    if (fl.z) nf[2] = nf[0] = 1;
                    const int nd[3] = {el.x, el.y, el.z};
                    for (int q = 0; q < 3; q++)
                        if (nf[q] && !dir[nd[q]]) {
                            dir[nd[q]] = 1;
                            order.push_back(nd[q]);
                        }
                }
            // the reassembly writes every element entry: all must be in the pattern
            for (int e = 0; e < ne; e++) {
                const sphfem_int4 el = d->node_num_el[e];
                const int nd[3] = {el.x, el.y, el.z};
                for (int a = 0; a < 3; a++) {
                    if (nd[a] < 0 || nd[a] >= nn) throw std::runtime_error("element node out of range");
                    for (int b = 0; b < 3; b++) (void)find(nd[a], nd[b]);
                }
            }
            std::vector<std::vector<int>> reach(nn);
            for (int F : order)
                for (int q = d->csr_row_ptr[F]; q < d->csr_row_ptr[F + 1]; q++)
                    if (!dir[d->csr_col_ind[q]]) reach[d->csr_col_ind[q]].push_back(q);
            std::vector<int> dp(nn + 1, 0), de;
            for (int i = 0; i < nn; i++) {
                de.insert(de.end(), reach[i].begin(), reach[i].end());
                dp[i + 1] = (int)de.size();
            }
            t->node_flags = dupload(s, d->node_flags, nn);
            t->dir_node = dupload(s, dir.data(), nn);
            t->dir_ptr = dupload(s, dp.data(), nn + 1);
            if (de.empty()) de.push_back(0);
            t->dir_entry = dupload(s, de.data(), de.size());
            t->pos_map = dalloc<sphfem_double2>(nn);
            t->T_map = dalloc<double>(nn);
            t->h_elem = new sphfem_int4[(size_t)ne];
            std::copy(d->node_num_el, d->node_num_el + ne, t->h_elem);
            t->n_fixed = d->fixed_num;
            t->remeshable = true;
        }
        t->sink_keys = dalloc<unsigned long long>(2ul * SPHFEM_SEG_SINK_CAP);
        t->sink_vals = dalloc<double>(8ul * SPHFEM_SEG_SINK_CAP);
        SPHFEM_HIP_CHECK(hipStreamSynchronize(s));
        *out = t;
        return SPHFEM_OK;
    });
}

int sphfem_tool_destroy(sphfem_ctx *ctx, sphfem_tool *t) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        dfree(t->pos); dfree(t->elem); dfree(t->row_ptr); dfree(t->col_ind); dfree(t->diag);
        dfree(t->K); dfree(t->K_temp); dfree(t->M); dfree(t->Minv); dfree(t->f); dfree(t->f_temp);
        dfree(t->T); dfree(t->T_old); dfree(t->k); dfree(t->left); dfree(t->right); dfree(t->n);
        dfree(t->heat_exchange); dfree(t->fric_heat); dfree(t->wear_rate); dfree(t->wear_nodes);
        dfree(t->physical_para); dfree(t->sliding_force); dfree(t->marks); dfree(t->node_robin);
        dfree(t->node_shift_ptr); dfree(t->node_shift); dfree(t->layer_elems); dfree(t->layer_old);
        dfree(t->layer_new); dfree(t->nnz_layer_ptr); dfree(t->nnz_layer); dfree(t->node_layer_ptr);
        dfree(t->node_layer); dfree(t->kin); dfree(t->node_flags); dfree(t->dir_node); dfree(t->dir_ptr);
        dfree(t->dir_entry); dfree(t->pos_map); dfree(t->T_map); dfree(t->bk_ptr); dfree(t->bk_list);
        dfree(t->sink_keys); dfree(t->sink_vals);
        delete[] t->h_elem;
        delete t;
        return SPHFEM_OK;
    });
}

}  // extern "C"

// bins of the tool's elements at its current node positions, in the tool's bin buffers
static ElemBuckets build_elem_buckets(sphfem_ctx *ctx, sphfem_tool *t) {
    const int nn = t->n_nodes, ne = t->n_elem;
    std::vector<sphfem_double2> pos((size_t)nn);
    SPHFEM_HIP_CHECK(hipMemcpyAsync(pos.data(), t->pos, 16ul * nn, hipMemcpyDeviceToHost, ctx->stream));
    SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    for (const auto &p : pos)
        if (!std::isfinite(p.x) || !std::isfinite(p.y)) throw RangeError("a tool node position is not finite");
    ElemBuckets bk;
    if (ne == 0) {                      // no elements: one empty bin
        if (t->bk_ptr_cap < 2) {
            dfree(t->bk_ptr);
            t->bk_ptr = dalloc<int>(2);
            t->bk_ptr_cap = 2;
        }
        if (t->bk_list_cap < 1) {
            dfree(t->bk_list);
            t->bk_list = dalloc<int>(1);
            t->bk_list_cap = 1;
        }
        const int zero[2] = {0, 0};
        SPHFEM_HIP_CHECK(hipMemcpyAsync(t->bk_ptr, zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        bk.nx = bk.ny = 1;
        bk.ptr = t->bk_ptr;
        bk.list = t->bk_list;
        return bk;
    }
    double ext = 0.;
    {
        double x0 = DBL_MAX, y0 = DBL_MAX, x1 = -DBL_MAX, y1 = -DBL_MAX;
        for (const auto &p : pos) {
            x0 = std::min(x0, p.x); y0 = std::min(y0, p.y); x1 = std::max(x1, p.x); y1 = std::max(y1, p.y);
        }
        ext = std::max(x1 - x0, y1 - y0);
    }
    // element boxes padded by 1e-6 of their size and 1e-9 of the tool's: far beyond the 1e-8
    // barycentric tolerance and the 1e-10 coincidence distance of the interpolations
    std::vector<double> box((size_t)4 * ne), diag((size_t)ne);
    double gx0 = DBL_MAX, gy0 = DBL_MAX, gx1 = -DBL_MAX, gy1 = -DBL_MAX;
    for (int e = 0; e < ne; e++) {
        const sphfem_int4 el = t->h_elem[e];
        const sphfem_double2 a = pos[(size_t)el.x], b = pos[(size_t)el.y], c = pos[(size_t)el.z];
        double lx = std::min(a.x, std::min(b.x, c.x)), hx = std::max(a.x, std::max(b.x, c.x));
        double ly = std::min(a.y, std::min(b.y, c.y)), hy = std::max(a.y, std::max(b.y, c.y));
        diag[(size_t)e] = std::hypot(hx - lx, hy - ly);
        const double pad = 1e-6 * diag[(size_t)e] + 1e-9 * ext;
        box[4 * (size_t)e] = lx - pad; box[4 * (size_t)e + 1] = ly - pad;
        box[4 * (size_t)e + 2] = hx + pad; box[4 * (size_t)e + 3] = hy + pad;
        gx0 = std::min(gx0, lx - pad); gy0 = std::min(gy0, ly - pad);
        gx1 = std::max(gx1, hx + pad); gy1 = std::max(gy1, hy + pad);
    }
    std::vector<double> sorted(diag);
    std::nth_element(sorted.begin(), sorted.begin() + ne / 2, sorted.end());
    double h = std::max(2. * sorted[(size_t)ne / 2], 1e-12 * std::max(ext, 1e-300));
    auto cell = [&](double v, double o) { return (int)std::floor((v - o) * bk.inv); };
    long total = 0;
    for (int tries = 0;; tries++) {
        bk.inv = 1. / h;
        bk.x0 = gx0; bk.y0 = gy0;
        const double span = std::max(gx1 - gx0, gy1 - gy0) * bk.inv;
        if (!(span < 1e7)) {            // degenerate extents: widen the bins first
            h = std::max(h * 1.5, std::max(gx1 - gx0, gy1 - gy0) * 1e-6);
            if (!(h > 0.) || tries > 200) {
                bk.inv = 0.;            // one bin holding every element
                bk.nx = bk.ny = 1;
                total = ne;
                break;
            }
            continue;
        }
        bk.nx = cell(gx1, gx0) + 1;
        bk.ny = cell(gy1, gy0) + 1;
        total = 0;
        for (int e = 0; e < ne; e++)
            total += (long)(cell(box[4 * (size_t)e + 2], gx0) - cell(box[4 * (size_t)e], gx0) + 1) *
                     (cell(box[4 * (size_t)e + 3], gy0) - cell(box[4 * (size_t)e + 1], gy0) + 1);
        if ((total <= 16l * ne && (long)bk.nx * bk.ny <= 4l * ne + 16) || tries > 60) break;
        h *= 1.5;
    }
    const size_t ncell = (size_t)bk.nx * bk.ny;
    std::vector<int> ptr(ncell + 1, 0), list((size_t)total), fill;
    for (int pass = 0; pass < 2; pass++) {      // counts, then the lists in element order
        if (pass) fill.assign(ptr.begin(), ptr.end() - 1);
        for (int e = 0; e < ne; e++) {
            const int cx0 = cell(box[4 * (size_t)e], gx0), cx1 = cell(box[4 * (size_t)e + 2], gx0);
            const int cy0 = cell(box[4 * (size_t)e + 1], gy0), cy1 = cell(box[4 * (size_t)e + 3], gy0);
            for (int cy = cy0; cy <= cy1; cy++)
                for (int cx = cx0; cx <= cx1; cx++) {
                    const size_t c = (size_t)cy * bk.nx + cx;
                    if (pass == 0) ptr[c + 1]++;
                    else list[(size_t)fill[c]++] = e;
                }
        }
        if (pass == 0)
            for (size_t c = 0; c < ncell; c++) ptr[c + 1] += ptr[c];
    }
    if (t->bk_ptr_cap < ncell + 1) {
        dfree(t->bk_ptr);
        t->bk_ptr = dalloc<int>(ncell + 1);
        t->bk_ptr_cap = ncell + 1;
    }
    if (t->bk_list_cap < std::max<size_t>(list.size(), 1)) {
        dfree(t->bk_list);
        t->bk_list = dalloc<int>(std::max<size_t>(list.size(), 1));
        t->bk_list_cap = std::max<size_t>(list.size(), 1);
    }
    SPHFEM_HIP_CHECK(hipMemcpyAsync(t->bk_ptr, ptr.data(), 4 * (ncell + 1), hipMemcpyHostToDevice, ctx->stream));
    if (!list.empty())
        SPHFEM_HIP_CHECK(hipMemcpyAsync(t->bk_list, list.data(), 4 * list.size(), hipMemcpyHostToDevice, ctx->stream));
    SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));    // the host vectors go out of scope
    bk.ptr = t->bk_ptr;
    bk.list = t->bk_list;
    return bk;
}

extern "C" {

int sphfem_tool_remesh_low_level(sphfem_ctx *ctx, sphfem_tool *t, int freq) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (!t->remeshable)
            return fail(ctx, SPHFEM_ERR_STATE, "remeshing needs a tool created with node flags and the fixed boundary");
        if (int rc = need(ctx, SET_TOOLC)) return rc;
        const ElemBuckets bk = build_elem_buckets(ctx, t);     // the mesh before the shift
        launch_remesh_low_level(ctx, t, freq, bk);
        return SPHFEM_OK;
    });
}

int sphfem_tool_remesh(sphfem_ctx *ctx, sphfem_tool *t, sphfem_mesh *mesh) {
    if (!t || !mesh) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (!t->remeshable)
            return fail(ctx, SPHFEM_ERR_STATE, "remeshing needs a tool created with node flags and the fixed boundary");
        if (int rc = need(ctx, SET_TOOLC)) return rc;
        sphfem_tool_kinematics k;
        SPHFEM_HIP_CHECK(hipMemcpyAsync(&k, t->kin, sizeof(k), hipMemcpyDeviceToHost, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        sphfem_tool_desc d;
        if (sphfem_mesh_tool_desc(mesh, &ctx->c.tc, sphfem_double2{k.tl_x, k.tl_y}, sphfem_double2{k.vel_x, k.vel_y},
                                  &d) != SPHFEM_OK)
            return fail(ctx, SPHFEM_ERR_ARG, "the new mesh has no single cutting-edge chain");
        sphfem_tool *nt = nullptr;
        if (int rc = sphfem_tool_create(ctx, &d, &nt)) return rc;
        try {
            // mesh_GPU_m(..., false) starts the new nodes at 300 K (mesh_structure_GPU.cu:43-44)
            const std::vector<double> t300((size_t)nt->n_nodes, 300.);
            SPHFEM_HIP_CHECK(hipMemcpyAsync(nt->T, t300.data(), 8ul * nt->n_nodes, hipMemcpyHostToDevice,
                                            ctx->stream));
            const ElemBuckets bk = build_elem_buckets(ctx, t);    // the current mesh
            launch_interp_new(ctx, t, nt->pos, nt->n_nodes, nt->T, bk);
            SPHFEM_HIP_CHECK(hipMemcpyAsync(nt->T_old, nt->T, 8ul * nt->n_nodes, hipMemcpyDeviceToDevice,
                                            ctx->stream));
            sphfem_tool_kinematics nk = k;
            nk.lowest_point = d.lowest_point;
            nk.right_point = d.right_point;
            nk.ref_dist = d.ref_dist;
            SPHFEM_HIP_CHECK(hipMemcpyAsync(nt->kin, &nk, sizeof(nk), hipMemcpyHostToDevice, ctx->stream));
            nt->host_kin = nk;
            SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            (void)sphfem_tool_destroy(ctx, nt);     // the tool keeps its old mesh
            throw;
        }
        const unsigned long gen = t->generation + 1;
        std::swap(*t, *nt);
        t->generation = gen;
        return sphfem_tool_destroy(ctx, nt);      // the old arrays
    });
}

int sphfem_tool_get_arrays(const sphfem_tool *t, sphfem_tool_arrays *o) {
    if (!t || !o) return SPHFEM_ERR_ARG;
    std::memset(o, 0, sizeof(*o));
    o->node_num = t->n_nodes; o->ele_num = t->n_elem; o->nz_num = t->nnz; o->seg_num = t->n_seg;
    o->color_total = t->n_colors; o->color_cutting_edge = t->n_cut_colors; o->fixed_ele_num = t->n_fixed;
    o->position = t->pos; o->position_mapping = t->pos_map; o->node_num_el = t->elem;
    o->csr_row_ptr = t->row_ptr; o->csr_col_ind = t->col_ind; o->node_flags = t->node_flags;
    o->K = t->K; o->K_temp = t->K_temp; o->M = t->M; o->M_inv = t->Minv; o->f = t->f; o->f_temp = t->f_temp;
    o->k = t->k; o->T = t->T; o->T_old = t->T_old; o->T_mapping = t->T_map;
    o->left = t->left; o->right = t->right; o->n = t->n; o->heat_exchange = t->heat_exchange;
    o->marks = t->marks; o->fric_heat = t->fric_heat; o->wear_rate = t->wear_rate; o->wear_nodes = t->wear_nodes;
    o->physical_para = t->physical_para; o->sliding_force = t->sliding_force;
    return SPHFEM_OK;
}

int sphfem_tool_get_sizes(const sphfem_tool *t, int *node_num, int *ele_num, int *nz_num, int *seg_num) {
    if (!t) return SPHFEM_ERR_ARG;
    if (node_num) *node_num = t->n_nodes;
    if (ele_num) *ele_num = t->n_elem;
    if (nz_num) *nz_num = t->nnz;
    if (seg_num) *seg_num = t->n_seg;
    return SPHFEM_OK;
}

int sphfem_tool_get_kinematics(sphfem_ctx *ctx, sphfem_tool *t, sphfem_tool_kinematics *out) {
    if (!t || !out) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipMemcpyAsync(out, t->kin, sizeof(*out), hipMemcpyDeviceToHost, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return SPHFEM_OK;
    });
}

int sphfem_tool_set_vel(sphfem_ctx *ctx, sphfem_tool *t, sphfem_double2 vel) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        sphfem_tool_kinematics k;
        SPHFEM_HIP_CHECK(hipMemcpyAsync(&k, t->kin, sizeof(k), hipMemcpyDeviceToHost, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        k.vel_x = vel.x;
        k.vel_y = vel.y;
        SPHFEM_HIP_CHECK(hipMemcpyAsync(t->kin, &k, sizeof(k), hipMemcpyHostToDevice, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return SPHFEM_OK;
    });
}

int sphfem_tool_segments_update(sphfem_ctx *ctx, sphfem_tool *t, double dt) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        launch_tool_move(ctx, t, dt, true, false);
        return SPHFEM_OK;
    });
}

int sphfem_tool_update_FEM_gpu(sphfem_ctx *ctx, sphfem_tool *t, double dt) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        launch_tool_move(ctx, t, dt, false, true);
        return SPHFEM_OK;
    });
}

int sphfem_tool_segments_wear_rate_reset(sphfem_ctx *ctx, sphfem_tool *t) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        launch_wear_reset(ctx, t);
        return SPHFEM_OK;
    });
}

int sphfem_tool_apply_heat_boundary_condition(sphfem_ctx *ctx, sphfem_tool *t) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_TOOLC)) return rc;
        launch_fem_robin(ctx, t);
        return SPHFEM_OK;
    });
}

int sphfem_FEM_reset_server(sphfem_ctx *ctx, sphfem_tool *t) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        launch_fem_reset(ctx, t);
        return SPHFEM_OK;
    });
}

int sphfem_FEM_calculate_temp(sphfem_ctx *ctx, sphfem_tool *t, double dt) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        launch_fem_temp(ctx, t, dt);
        return SPHFEM_OK;
    });
}

int sphfem_FEM_calculate_temp_rate(sphfem_ctx *ctx, sphfem_tool *t) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        launch_fem_rate(ctx, t);
        return SPHFEM_OK;
    });
}

int sphfem_tool_download(sphfem_ctx *ctx, sphfem_tool *t, sphfem_double2 *position, double *T, double *K, double *M,
                         double *k, sphfem_double2 *left, sphfem_double2 *right, sphfem_double2 *n,
                         sphfem_double2 *heat_exchange, double *fric_heat, sphfem_double4 *physical_para,
                         double *sliding_force, double *wear_nodes) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        auto cp = [&](void *dst, const void *src, size_t bytes) {
            if (dst) SPHFEM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        };
        cp(position, t->pos, 16ul * t->n_nodes);
        cp(T, t->T, 8ul * t->n_nodes);
        cp(K, t->K, 8ul * t->nnz);
        cp(M, t->M, 8ul * t->n_nodes);
        cp(k, t->k, 8ul * t->n_nodes);
        cp(left, t->left, 16ul * t->n_seg);
        cp(right, t->right, 16ul * t->n_seg);
        cp(n, t->n, 16ul * t->n_seg);
        cp(heat_exchange, t->heat_exchange, 16ul * t->n_seg);
        cp(fric_heat, t->fric_heat, 8ul * t->n_seg);
        cp(physical_para, t->physical_para, 32ul * t->n_seg);
        cp(sliding_force, t->sliding_force, 8ul * t->n_seg);
        cp(wear_nodes, t->wear_nodes, 8ul * t->n_seg);
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return SPHFEM_OK;
    });
}

int sphfem_tool_upload(sphfem_ctx *ctx, sphfem_tool *t, const sphfem_double2 *position, const double *T,
                       const double *K, const double *M, const double *k, const sphfem_double2 *left,
                       const sphfem_double2 *right, const sphfem_double2 *n, const sphfem_double2 *heat_exchange,
                       const double *fric_heat, const sphfem_double4 *physical_para, const double *sliding_force,
                       const double *wear_nodes) {
    if (!t) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        auto cp = [&](void *dst, const void *src, size_t bytes) {
            if (src) SPHFEM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        };
        cp(t->pos, position, 16ul * t->n_nodes);
        cp(t->T, T, 8ul * t->n_nodes);
        cp(t->K, K, 8ul * t->nnz);
        cp(t->M, M, 8ul * t->n_nodes);
        cp(t->k, k, 8ul * t->n_nodes);
        cp(t->left, left, 16ul * t->n_seg);
        cp(t->right, right, 16ul * t->n_seg);
        cp(t->n, n, 16ul * t->n_seg);
        cp(t->heat_exchange, heat_exchange, 16ul * t->n_seg);
        cp(t->fric_heat, fric_heat, 8ul * t->n_seg);
        cp(t->physical_para, physical_para, 32ul * t->n_seg);
        cp(t->sliding_force, sliding_force, 8ul * t->n_seg);
        cp(t->wear_nodes, wear_nodes, 8ul * t->n_seg);
        if (M) {
            std::vector<double> minv(t->n_nodes);
            for (int i = 0; i < t->n_nodes; i++) minv[i] = 1. / M[i];
            SPHFEM_HIP_CHECK(hipMemcpyAsync(t->Minv, minv.data(), 8ul * t->n_nodes, hipMemcpyHostToDevice,
                                            ctx->stream));
            SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        }
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return SPHFEM_OK;
    });
}

int sphfem_tool_set_kinematics(sphfem_ctx *ctx, sphfem_tool *t, const sphfem_tool_kinematics *kin) {
    if (!t || !kin) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipMemcpyAsync(t->kin, kin, sizeof(*kin), hipMemcpyHostToDevice, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return SPHFEM_OK;
    });
}

// ---- leap frog
int sphfem_leap_frog_create(sphfem_ctx *ctx, unsigned int num_part, unsigned int num_cell, int node_num,
                            sphfem_leap_frog **out) {
    (void)node_num;
    if (!out || num_part == 0) return SPHFEM_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        auto *lf = new sphfem_leap_frog();
        lf->num_part = num_part;
        lf->num_cell = num_cell;
        lf->pos_init = dalloc<sphfem_double2>(num_part);
        lf->vel_init = dalloc<sphfem_double2>(num_part);
        lf->S_init = dalloc<sphfem_double4>(num_part);
        lf->rho_init = dalloc<double>(num_part);
        lf->T_init = dalloc<double>(num_part);
        lf->pre_vol = dalloc<double>(num_part);
        lf->pre_s = dalloc<sphfem_double4>(num_part);
        lf->psum64 = dalloc<sphfem::PairSums>(num_part);
        lf->cell_start = dalloc<int>(std::max(num_cell, 1u));
        lf->cell_end = dalloc<int>(std::max(num_cell, 1u));
        lf->near_list = dalloc<int>(num_part);
        *out = lf;
        return SPHFEM_OK;
    });
}

int sphfem_leap_frog_destroy(sphfem_ctx *ctx, sphfem_leap_frog *lf) {
    if (!lf) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        dfree(lf->pos_init); dfree(lf->vel_init); dfree(lf->S_init); dfree(lf->rho_init); dfree(lf->T_init);
        dfree(lf->cell_start); dfree(lf->cell_end); dfree(lf->near_list);
        dfree(lf->rec); dfree(lf->psum); dfree(lf->psum_T); dfree(lf->pre_vol); dfree(lf->pre_s); dfree(lf->psum64);
        dfree(lf->rate32[0]); dfree(lf->rate32[1]);
        delete lf;
        return SPHFEM_OK;
    });
}

int sphfem_leap_frog_set_precision(sphfem_ctx *ctx, sphfem_leap_frog *lf, int precision) {
    if (!ctx || !lf) return SPHFEM_ERR_ARG;
    if (precision != SPHFEM_FP64 && precision != SPHFEM_FP32C) return fail(ctx, SPHFEM_ERR_ARG, "unsupported precision");
    return guarded(ctx, [&] {
        if (precision == SPHFEM_FP32C) {
            // neighbour records are addressed by 32-bit byte offsets (16 B each)
            if (lf->num_part >= (1u << 28)) return fail(ctx, SPHFEM_ERR_RANGE, "FP32C supports < 2^28 particles");
            if (!lf->rec) {
                // record blocks r0 [N + 4], r1, r2, r3 [N]; the padding of r0 is zeroed once
                lf->rec = dalloc<float4>(4ul * lf->num_part + 4);
                SPHFEM_HIP_CHECK(hipMemsetAsync(lf->rec + lf->num_part, 0, 4 * sizeof(float4), ctx->stream));
            }
            if (!lf->psum) lf->psum = dalloc<float4>(2ul * lf->num_part);
            if (!lf->psum_T) lf->psum_T = dalloc<float>(lf->num_part);
            for (int k = 0; k < 2; k++)
                if (!lf->rate32[k]) lf->rate32[k] = dalloc<float>(10ul * lf->num_part);
        }
        lf->precision = precision;
        return SPHFEM_OK;
    });
}

int sphfem_leap_frog_set_tuning(sphfem_ctx *ctx, sphfem_leap_frog *lf, const sphfem_tuning *tuning) {
    if (!ctx || !lf || !tuning) return SPHFEM_ERR_ARG;
    const sphfem_tuning &t = *tuning;
    const bool ok = (t.pair_block == 128 || t.pair_block == 256) &&
                    (t.pair_chunk == 16 || t.pair_chunk == 32 || t.pair_chunk == 48) &&
                    (t.update_block == 128 || t.update_block == 256) &&
                    (t.pair_kernel >= SPHFEM_PAIRS_LIST && t.pair_kernel <= SPHFEM_PAIRS_STAGED_FUSED) &&
                    (t.resort_one_block == 0 ||
                     (t.resort_one_block == 1 && lf->num_part <= (unsigned)SPHFEM_SORT_ONE_BLOCK_MAX));
    if (!ok) return fail(ctx, SPHFEM_ERR_ARG, "unsupported tuning");
    lf->tuning = t;
    return SPHFEM_OK;
}

int sphfem_leap_frog_export_rates(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p) {
    if (!lf || !p || p->N != lf->num_part) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (lf->rates_fp32) launch_rates_export(ctx, rate_view(lf->rate32[lf->rate_cur], lf->num_part), *p);
        return SPHFEM_OK;
    });
}

int sphfem_leap_frog_import_rates(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p) {
    if (!lf || !p || p->N != lf->num_part) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (lf->rates_fp32) launch_rates_import(ctx, *p, rate_view(lf->rate32[lf->rate_cur], lf->num_part));
        return SPHFEM_OK;
    });
}

int sphfem_leap_frog_get_tuning(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_tuning *tuning) {
    if (!ctx || !lf || !tuning) return SPHFEM_ERR_ARG;
    *tuning = lf->tuning;
    return SPHFEM_OK;
}

int sphfem_leap_frog_init(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p) {
    if (!lf || !p) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        launch_leapfrog(ctx, LF_INIT, *p, lf, 0.);
        return SPHFEM_OK;
    });
}

int sphfem_leap_frog_predict(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, double dt) {
    if (!lf || !p) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        launch_leapfrog(ctx, LF_PREDICT, *p, lf, dt);
        return SPHFEM_OK;
    });
}

int sphfem_leap_frog_correct(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, double dt) {
    if (!lf || !p) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        launch_leapfrog(ctx, LF_CORRECT, *p, lf, dt);
        return SPHFEM_OK;
    });
}

// leap_frog.cu:162-248, as fused kernels:
//  tool_begin | hash+digit-0 histogram | radix passes | reorder+init+predict+EOS+art. stress |
//  pair sums+Jaumann+continuity+momentum+advection+correct+friction heat+plasticity+far contact+
//  BC+invalidate | FEM rate+temperature | wear | tool move | near contact s1 (+BC) | contact s2
// With a rigid tool (leap_frog.cu:99-160) the tool is moved first, as only the contact of the
// update kernel reads its geometry and that reads the moved one; the update kernel also applies
// the friction heat, moves the tool particles and computes the contact.
// where the rates live follows the stepper's form: the FP32 carry is filled from the particle arrays
// when FP32C steps begin, and written back to them when FP64 steps follow FP32C ones
static void sync_rate_form(sphfem_ctx *ctx, sphfem_leap_frog *lf, const sphfem_particles &p) {
    const bool f32c = lf->precision == SPHFEM_FP32C;
    if (f32c == lf->rates_fp32) return;
    const RateView r = rate_view(lf->rate32[lf->rate_cur], lf->num_part);
    if (f32c) launch_rates_import(ctx, p, r);
    else launch_rates_export(ctx, r, p);
    lf->rates_fp32 = f32c;
}

static int step_body(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, sphfem_grid *g, sphfem_tool *t,
                     sphfem_rigid_tool *rt = nullptr) {
    // leap_frog.cu:175 (a no-op for a locked grid)
    if (int rc = update_geometry(ctx, g, *p, 2.)) return rc;
    ctx->c.geom = g->geom;
    // the FEM tool's step start runs inside the hash pass of the Green re-sort (below), else here
    const bool tool_in_hash = t && g->kind != SPHFEM_GRID_ROTHLIN;
    if (t && !tool_in_hash) launch_tool_begin(ctx, t);
    else if (!t) launch_fill_int(ctx, ctx->d_flags + 1, SPHFEM_FLAG_NONFINITE, 0);   // [1] .. [5]
    if (rt) launch_rigid_advance(ctx, rt);
    // the source of each sorted position: the permutation of the sort (Green), or the identity
    // after the Rothlin sort has gathered in place
    sphfem_particles A = *p;
    const bool f32c = lf->precision == SPHFEM_FP32C;
    const size_t nr = lf->num_part;
    sync_rate_form(ctx, lf, *p);
    if (g->kind == SPHFEM_GRID_ROTHLIN) {
        launch_hash(ctx, g, *p, false, nullptr);
        launch_rothlin_sort(ctx, g, *p);    // sorted keys, in-place gathers, dense cell table
        if (f32c) {                         // the carry follows the in-place gathers
            launch_rates_gather(ctx, rate_view(lf->rate32[lf->rate_cur], nr), rate_view(lf->rate32[lf->rate_cur ^ 1], nr),
                                g->perm, (int)p->N);
            lf->rate_cur ^= 1;
        }
        A.idx = g->iota;
    } else {
        // the cell table is cleared by the hash pass (cells of the previous keys) except when the
        // previous keys are unknown
        if (g->first_sort) launch_fill_int(ctx, g->cell_start, g->num_cell, -1);
        launch_resort(ctx, g, *p, lf->tuning.resort_one_block != 0, tool_in_hash ? t : nullptr);
        g->first_sort = false;
    }
    const RateView r_old = rate_view(lf->rate32[lf->rate_cur], nr);
    const RateView r_new = rate_view(lf->rate32[lf->rate_cur ^ 1], nr);
    launch_reorder_predict(ctx, g, A, lf, f32c ? &r_old : nullptr);
    sphfem_particles B = g->buf;
    B.hash = p->hash;
    B.idx = A.idx;
    B.N = p->N;
    // far from the FEM tool the contact fields are zero and need not be read (not with the Rothlin
    // grid, whose step leaves ft and n ungathered as the reference does, nor with a rigid tool)
    const bool skip_contact = t && !rt && g->kind == SPHFEM_GRID_GREEN;
    launch_fused_update(ctx, B, *p, lf, g->cell_start, g->cell_end, t, rt, f32c ? &r_old : nullptr,
                        f32c ? &r_new : nullptr, skip_contact);
    if (f32c) {
        // the whole new state is in the buffer, the new rates in the other carry buffer; R is not
        // maintained by the FP32C form
        swap_gathered_state(*p, g->buf);
        std::swap(p->p, g->buf.p);
        lf->rate_cur ^= 1;
    } else {
        swap_fused(*p, g->buf);
    }
    if (t) {
        if (ctx->c.wear_pro) {
            launch_tool_fem_end_wear(ctx, t);      // the wear reads the new temperatures, before the motion
            launch_tool_move(ctx, t, ctx->c.dt, true, true);
        } else {
            launch_tool_fem_end_move(ctx, t);
        }
        launch_near_contact(ctx, *p, t, lf);
    }
    return SPHFEM_OK;
}

int sphfem_leap_frog_step(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, sphfem_grid *g,
                          sphfem_tool *t) {
    if (!lf || !p || !g || (int)p->N != g->num_part || p->N != lf->num_part) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_PHYS | SET_CORR | SET_DT | (t ? SET_TOOLC | SET_TRML_WP : 0u))) return rc;
        return step_body(ctx, lf, p, g, t);
    });
}

int sphfem_leap_frog_step_rigid(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, sphfem_grid *g,
                                sphfem_rigid_tool *rt) {
    if (!lf || !p || !g || (int)p->N != g->num_part || p->N != lf->num_part) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_PHYS | SET_CORR | SET_DT)) return rc;
        return step_body(ctx, lf, p, g, nullptr, rt);
    });
}

}  // extern "C"

// what the recorded launches of one case baked in: the constant blocks (passed by value: dt, wear
// switch, cut-off, geometry, ...), the stepper's precision, tuning and intermediates switch, and the
// FEM tool's array generation
struct GraphCase {
    sphfem_ctx *ctx = nullptr;
    sphfem_tool *tool = nullptr;
    unsigned long tool_gen = 0;
    DevConsts consts{};
    sphfem_leap_frog *lf = nullptr;
    int precision = 0;
    sphfem_tuning tuning{};
    bool intermediates = false;
};

struct sphfem_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int steps = 0;
    std::vector<GraphCase> cases;     // one for a step graph, one per case for a batch graph
    // batch graph: a launch on the first context's stream first waits for the work queued on the
    // other cases' streams (join[k]), and those streams then wait for the launch (done)
    std::vector<hipEvent_t> join;
    hipEvent_t done = nullptr;
};

static GraphCase graph_case(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_tool *t) {
    GraphCase g;
    g.ctx = ctx;
    g.tool = t;
    g.tool_gen = t ? t->generation : 0;
    g.consts = ctx->c;
    g.lf = lf;
    g.precision = lf->precision;
    g.tuning = lf->tuning;
    g.intermediates = lf->intermediates;
    return g;
}

// the settings a graph was captured with still hold (same values, whatever was called since)
static bool graph_settings_hold(const GraphCase &g) {
    const sphfem_leap_frog *lf = g.lf;
    return std::memcmp(&g.ctx->c, &g.consts, sizeof(DevConsts)) == 0 && lf->precision == g.precision &&
           lf->tuning.pair_block == g.tuning.pair_block && lf->tuning.pair_chunk == g.tuning.pair_chunk &&
           lf->tuning.update_block == g.tuning.update_block && lf->tuning.pair_kernel == g.tuning.pair_kernel &&
           lf->tuning.resort_one_block == g.tuning.resort_one_block && lf->intermediates == g.intermediates;
}

extern "C" {

static int graph_create(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, sphfem_grid *g, sphfem_tool *t,
                        sphfem_rigid_tool *rt, int steps, sphfem_graph **out) {
    if (!lf || !p || !g || !out || (int)p->N != g->num_part || p->N != lf->num_part) return SPHFEM_ERR_ARG;
    if (steps < 2 || steps % 2) return fail(ctx, SPHFEM_ERR_ARG, "graph steps must be even (pointer swaps)");
    *out = nullptr;
    return guarded(ctx, [&] {
        if (int rc = need(ctx, SET_PHYS | SET_CORR | SET_DT | (t ? SET_TOOLC | SET_TRML_WP : 0u))) return rc;
        if (!g->locked)
            return fail(ctx, SPHFEM_ERR_STATE, "graphs need a locked grid (the geometry update reads back to the host)");
        if (g->first_sort) return fail(ctx, SPHFEM_ERR_STATE, "run one step before capturing (first sort)");
        sync_rate_form(ctx, lf, *p);      // eagerly: the graph records the steady state only
        const bool timing = ctx->timing.enabled;
        ctx->timing.enabled = false;
        auto *gr = new sphfem_graph();
        gr->steps = steps;
        try {
            SPHFEM_HIP_CHECK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
            // recording only: the device state does not advance; after an even number of steps the
            // host-side pointer assignment is the one the graph starts from
            for (int k = 0; k < steps; k++) (void)step_body(ctx, lf, p, g, t, rt);   // locked: cannot fail
            SPHFEM_HIP_CHECK(hipStreamEndCapture(ctx->stream, &gr->graph));
            SPHFEM_HIP_CHECK(hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0));
            gr->cases.push_back(graph_case(ctx, lf, t));
        } catch (...) {
            hipGraph_t dead = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &dead);
            if (dead) (void)hipGraphDestroy(dead);
            if (gr->graph) (void)hipGraphDestroy(gr->graph);
            delete gr;
            ctx->timing.enabled = timing;
            throw;
        }
        ctx->timing.enabled = timing;
        *out = gr;
        return SPHFEM_OK;
    });
}

int sphfem_leap_frog_graph_create(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, sphfem_grid *g,
                                  sphfem_tool *t, int steps, sphfem_graph **out) {
    return graph_create(ctx, lf, p, g, t, nullptr, steps, out);
}

int sphfem_leap_frog_graph_create_rigid(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p,
                                        sphfem_grid *g, sphfem_rigid_tool *rt, int steps, sphfem_graph **out) {
    return graph_create(ctx, lf, p, g, nullptr, rt, steps, out);
}

int sphfem_batch_graph_create(sphfem_ctx *const *ctxs, sphfem_leap_frog *const *lfs, sphfem_particles *const *ps,
                              sphfem_grid *const *gs, sphfem_tool *const *ts, int n, int steps, sphfem_graph **out) {
    if (!ctxs || !lfs || !ps || !gs || !out || n < 1 || !ctxs[0]) return SPHFEM_ERR_ARG;
    sphfem_ctx *ctx0 = ctxs[0];
    if (steps < 2 || steps % 2) return fail(ctx0, SPHFEM_ERR_ARG, "graph steps must be even (pointer swaps)");
    *out = nullptr;
    for (int k = 0; k < n; k++) {
        sphfem_ctx *c = ctxs[k];
        sphfem_leap_frog *lf = lfs[k];
        sphfem_particles *p = ps[k];
        sphfem_grid *g = gs[k];
        if (!c || !lf || !p || !g || (int)p->N != g->num_part || p->N != lf->num_part)
            return fail(ctx0, SPHFEM_ERR_ARG, "batch graph: case " + std::to_string(k) + " is incomplete");
        if (c->device != ctx0->device) return fail(ctx0, SPHFEM_ERR_ARG, "batch graph: cases on different devices");
        for (int j = 0; j < k; j++)
            if (ctxs[j] == c) return fail(ctx0, SPHFEM_ERR_ARG, "batch graph: each case needs its own context");
        sphfem_tool *t = ts ? ts[k] : nullptr;
        if (need(c, SET_PHYS | SET_CORR | SET_DT | (t ? SET_TOOLC | SET_TRML_WP : 0u)) != SPHFEM_OK)
            return fail(ctx0, SPHFEM_ERR_STATE, "batch graph: case " + std::to_string(k) + " lacks constants");
        if (!g->locked || g->first_sort)
            return fail(ctx0, SPHFEM_ERR_STATE,
                        "batch graph: every case needs a locked grid and one eager step before capture");
    }
    return guarded(ctx0, [&] {
        for (int k = 0; k < n; k++) sync_rate_form(ctxs[k], lfs[k], *ps[k]);
        std::vector<bool> timing((size_t)n);
        for (int k = 0; k < n; k++) {
            timing[(size_t)k] = ctxs[k]->timing.enabled;
            ctxs[k]->timing.enabled = false;
        }
        auto *gr = new sphfem_graph();
        gr->steps = steps;
        std::vector<hipEvent_t> ev((size_t)n + 1, nullptr);
        try {
            for (auto &e : ev) SPHFEM_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            // every case's steps as one branch: the other streams join the capture of the first
            // context's stream through an event and it waits for all of them at the end
            SPHFEM_HIP_CHECK(hipStreamBeginCapture(ctx0->stream, hipStreamCaptureModeThreadLocal));
            SPHFEM_HIP_CHECK(hipEventRecord(ev[(size_t)n], ctx0->stream));
            for (int k = 1; k < n; k++) SPHFEM_HIP_CHECK(hipStreamWaitEvent(ctxs[k]->stream, ev[(size_t)n], 0));
            for (int k = 0; k < n; k++)
                for (int st = 0; st < steps; st++)
                    (void)step_body(ctxs[k], lfs[k], ps[k], gs[k], ts ? ts[k] : nullptr);   // locked: cannot fail
            for (int k = 1; k < n; k++) {
                SPHFEM_HIP_CHECK(hipEventRecord(ev[(size_t)k], ctxs[k]->stream));
                SPHFEM_HIP_CHECK(hipStreamWaitEvent(ctx0->stream, ev[(size_t)k], 0));
            }
            SPHFEM_HIP_CHECK(hipStreamEndCapture(ctx0->stream, &gr->graph));
            SPHFEM_HIP_CHECK(hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0));
            for (int k = 0; k < n; k++) gr->cases.push_back(graph_case(ctxs[k], lfs[k], ts ? ts[k] : nullptr));
            gr->join.assign((size_t)n, nullptr);
            for (int k = 1; k < n; k++) SPHFEM_HIP_CHECK(hipEventCreateWithFlags(&gr->join[(size_t)k], hipEventDisableTiming));
            SPHFEM_HIP_CHECK(hipEventCreateWithFlags(&gr->done, hipEventDisableTiming));
        } catch (...) {
            for (auto e : gr->join)
                if (e) (void)hipEventDestroy(e);
            if (gr->done) (void)hipEventDestroy(gr->done);
            hipGraph_t dead = nullptr;
            (void)hipStreamEndCapture(ctx0->stream, &dead);
            if (dead) (void)hipGraphDestroy(dead);
            if (gr->graph) (void)hipGraphDestroy(gr->graph);
            delete gr;
            for (auto e : ev)
                if (e) (void)hipEventDestroy(e);
            for (int k = 0; k < n; k++) ctxs[k]->timing.enabled = timing[(size_t)k];
            throw;
        }
        for (auto e : ev) (void)hipEventDestroy(e);
        for (int k = 0; k < n; k++) ctxs[k]->timing.enabled = timing[(size_t)k];
        *out = gr;
        return SPHFEM_OK;
    });
}

int sphfem_leap_frog_graph_launch(sphfem_ctx *ctx, sphfem_graph *gr, int repeats) {
    if (!gr || repeats < 0) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        if (gr->cases.empty() || gr->cases[0].ctx != ctx)
            return fail(ctx, SPHFEM_ERR_ARG, "launch a graph on the (first) context it was captured with");
        for (const GraphCase &g : gr->cases) {
            if (g.tool && g.tool->generation != g.tool_gen)
                return fail(ctx, SPHFEM_ERR_STATE, "the tool was remeshed after this graph was captured: capture again");
            if (!graph_settings_hold(g))
                return fail(ctx, SPHFEM_ERR_STATE,
                            "constants, dt, wear switch, precision or tuning changed after this graph was captured: "
                            "capture again");
        }
        const size_t nc = gr->cases.size();
        for (size_t k = 1; k < nc; k++) {       // batch: after what the other cases' streams hold
            SPHFEM_HIP_CHECK(hipEventRecord(gr->join[k], gr->cases[k].ctx->stream));
            SPHFEM_HIP_CHECK(hipStreamWaitEvent(ctx->stream, gr->join[k], 0));
        }
        for (int r = 0; r < repeats; r++) SPHFEM_HIP_CHECK(hipGraphLaunch(gr->exec, ctx->stream));
        if (nc > 1) {                            // and before what they are given next
            SPHFEM_HIP_CHECK(hipEventRecord(gr->done, ctx->stream));
            for (size_t k = 1; k < nc; k++) SPHFEM_HIP_CHECK(hipStreamWaitEvent(gr->cases[k].ctx->stream, gr->done, 0));
        }
        return SPHFEM_OK;
    });
}

int sphfem_leap_frog_graph_destroy(sphfem_ctx *ctx, sphfem_graph *gr) {
    if (!gr) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
        if (gr->graph) (void)hipGraphDestroy(gr->graph);
        for (auto e : gr->join)
            if (e) (void)hipEventDestroy(e);
        if (gr->done) (void)hipEventDestroy(gr->done);
        delete gr;
        return SPHFEM_OK;
    });
}

// ---- output formats
}  // extern "C"

struct sphfem_tool_wear {
    double K = 0., alpha = 0.;
    sphfem_double2 vel{0., 0.};
    double accum = 0.;
    double *d_buf = nullptr;    // [4 + 4 * 256]
};

namespace {

template <class T>
std::vector<T> download_vec(hipStream_t s, const T *dev, size_t n) {
    std::vector<T> v(n);
    if (n) SPHFEM_HIP_CHECK(hipMemcpyAsync(v.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost, s));
    return v;
}

FemToolHost download_fem_tool(sphfem_ctx *ctx, sphfem_tool *t, bool mesh, bool segments) {
    FemToolHost h;
    h.n_nodes = t->n_nodes;
    h.n_elem = t->n_elem;
    h.n_seg = t->n_seg;
    hipStream_t s = ctx->stream;
    if (mesh) {
        h.pos = download_vec(s, t->pos, (size_t)t->n_nodes);
        h.elem = download_vec(s, t->elem, (size_t)t->n_elem);
        h.T = download_vec(s, t->T, (size_t)t->n_nodes);
    }
    if (segments) {
        h.left = download_vec(s, t->left, (size_t)t->n_seg);
        h.right = download_vec(s, t->right, (size_t)t->n_seg);
        h.n = download_vec(s, t->n, (size_t)t->n_seg);
        h.physical_para = download_vec(s, t->physical_para, (size_t)t->n_seg);
        h.sliding_force = download_vec(s, t->sliding_force, (size_t)t->n_seg);
        h.wear_nodes = download_vec(s, t->wear_nodes, (size_t)t->n_seg);
    }
    SPHFEM_HIP_CHECK(hipStreamSynchronize(s));
    return h;
}

}  // namespace

extern "C" {

int sphfem_vtk_write_particles_host(const sphfem_particles *host, const char *path, int skip_blanked) {
    if (!host || !path) return SPHFEM_ERR_ARG;
    std::string err;
    try {
        return write_particles_vtk(*host, path, skip_blanked != 0, err);
    } catch (const std::exception &) {
        return SPHFEM_ERR_ARG;
    }
}

int sphfem_vtk_write_particles(sphfem_ctx *ctx, const sphfem_particles *p, const char *path, int skip_blanked) {
    if (!p || !path) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        const size_t N = p->N;
        hipStream_t s = ctx->stream;
        auto pos = download_vec(s, p->pos, N), vel = download_vec(s, p->vel, N);
        auto fc = download_vec(s, p->fc, N), ft = download_vec(s, p->ft, N), nn = download_vec(s, p->n, N);
        auto S = download_vec(s, p->S, N);
        auto pr = download_vec(s, p->p, N), bl = download_vec(s, p->blanked, N), eps = download_vec(s, p->eps_pl, N);
        auto T = download_vec(s, p->T, N);
        auto on_seg = download_vec(s, p->on_seg, N);
        SPHFEM_HIP_CHECK(hipStreamSynchronize(s));
        sphfem_particles h{};
        h.N = p->N;
        h.pos = pos.data(); h.vel = vel.data(); h.fc = fc.data(); h.ft = ft.data(); h.n = nn.data();
        h.S = S.data(); h.p = pr.data(); h.blanked = bl.data(); h.eps_pl = eps.data(); h.T = T.data();
        h.on_seg = on_seg.data();
        return write_particles_vtk(h, path, skip_blanked != 0, ctx->err);
    });
}

int sphfem_vtk_write_rigid_tool(const sphfem_tool_geometry *g, const char *path) {
    if (!g || !path) return SPHFEM_ERR_ARG;
    std::string err;
    return write_rigid_tool_vtk(*g, path, err);
}

// remesh_read_segments (tool_FEM.cu:1708-1810): the gmsh::model::geo calls written as a .geo script
int sphfem_tool_write_remesh_geo(sphfem_ctx *ctx, sphfem_tool *t, const char *path) {
    if (!t || !path) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        FemToolHost h = download_fem_tool(ctx, t, false, true);
        sphfem_tool_kinematics k;
        SPHFEM_HIP_CHECK(hipMemcpyAsync(&k, t->kin, sizeof(k), hipMemcpyDeviceToHost, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        const int ns = h.n_seg;
        // the distance of distance-like helpers there: sqrtf of the double expression
        auto fdist = [](sphfem_double2 a, sphfem_double2 b) {
            return (double)sqrtf((float)((a.x - b.x) * (a.x - b.x) + (a.y - b.y) * (a.y - b.y)));
        };
        struct Pt {
            double x, y, size;
        };
        std::vector<Pt> pts;
        std::vector<int> cut;
        double length = 0.004, dist_prev = 0.;
        const bool down = h.left[0].y > h.right[0].y;    // chain starts at left ends
        for (int i = 0; i < ns; i++) {
            const sphfem_double2 p = down ? h.left[(size_t)i] : h.right[(size_t)i];
            // the reference reads one element past the arrays at the last segment; there the
            // result cannot matter (the loop ends either way), so no distance is taken
            const double judge = i + 1 < ns ? (down ? fdist(h.left[(size_t)i], h.right[(size_t)i + 1])
                                                    : fdist(h.left[(size_t)i + 1], h.right[(size_t)i]))
                                            : HUGE_VAL;
            const double buffer = fdist(h.left[(size_t)i], h.right[(size_t)i]);
            length = 6.0 * buffer;
            pts.push_back({p.x, p.y, length});
            cut.push_back(1 + (int)pts.size());
            if ((judge < 1.25 * buffer && i != ns - 1 && dist_prev > 0.7 * buffer) || judge < 0.00013) i++;
            dist_prev = buffer;
        }
        const sphfem_double2 last = down ? h.right[(size_t)ns - 1] : h.left[(size_t)ns - 1];
        pts.push_back({last.x, last.y, length});
        std::FILE *fp = std::fopen(path, "w");
        if (!fp) return fail(ctx, SPHFEM_ERR_ARG, std::string("cannot write ") + path);
        const int np = (int)pts.size() + 1;     // with the reference point
        std::fprintf(fp, "// tool remesh geometry (remesh_read_segments): mesh with\n"
                         "//   gmsh -2 -format inp -o tool_gmsh_remeshed.inp <this file>\n");
        std::fprintf(fp, "Point(1) = {%.17g, %.17g, 0, %.17g};\n", k.tl_x, k.tl_y, 0.075);
        for (size_t q = 0; q < pts.size(); q++)
            std::fprintf(fp, "Point(%d) = {%.17g, %.17g, 0, %.17g};\n", (int)q + 2, pts[q].x, pts[q].y, pts[q].size);
        for (int q = 1; q < np; q++) std::fprintf(fp, "Line(%d) = {%d, %d};\n", q, q, q + 1);
        std::fprintf(fp, "Line(%d) = {%d, 1};\n", np, np);
        std::fprintf(fp, "Curve Loop(1) = {");
        for (int q = 1; q <= np; q++) std::fprintf(fp, q < np ? "%d, " : "%d};\n", q);
        std::fprintf(fp, "Plane Surface(1) = {1};\n");
        std::fprintf(fp, "Physical Curve(\"CuttingEdge\", 67) = {");
        for (size_t q = 0; q < cut.size(); q++) std::fprintf(fp, q + 1 < cut.size() ? "%d, " : "%d};\n", cut[q]);
        std::fprintf(fp, "Physical Curve(\"FixedBoundary\", 2) = {1, %d};\n", np);
        std::fprintf(fp, "Mesh.MshFileVersion = 4;\nMesh.SaveAll = 1;\n");
        const bool ok = std::ferror(fp) == 0;
        if (std::fclose(fp) != 0 || !ok) return fail(ctx, SPHFEM_ERR_ARG, std::string("cannot write ") + path);
        return SPHFEM_OK;
    });
}

int sphfem_vtk_write_FEM_tool(sphfem_ctx *ctx, sphfem_tool *t, const char *path) {
    if (!t || !path) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        FemToolHost h = download_fem_tool(ctx, t, true, false);
        return write_fem_tool_vtk(h, sphfem_double2{0., 0.}, path, ctx->err);
    });
}

int sphfem_vtk_write_tool_org(sphfem_ctx *ctx, sphfem_tool *t, sphfem_double2 tl_ref, const char *path) {
    if (!t || !path) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        FemToolHost h = download_fem_tool(ctx, t, true, false);
        sphfem_tool_kinematics k;
        SPHFEM_HIP_CHECK(hipMemcpyAsync(&k, t->kin, sizeof(k), hipMemcpyDeviceToHost, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        return write_fem_tool_vtk(h, sphfem_double2{k.tl_x - tl_ref.x, k.tl_y - tl_ref.y}, path, ctx->err);
    });
}

int sphfem_vtk_write_FEM_segments(sphfem_ctx *ctx, sphfem_tool *t, const char *path) {
    if (!t || !path) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        FemToolHost h = download_fem_tool(ctx, t, false, true);
        return write_fem_segments_vtk(h, path, ctx->err);
    });
}

int sphfem_vtk_write_contact_variables(sphfem_ctx *ctx, sphfem_tool *t, const char *path) {
    if (!t || !path) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        FemToolHost h = download_fem_tool(ctx, t, false, true);
        return write_contact_variables_vtk(h, path, ctx->err);
    });
}

int sphfem_read_vtk_temperature(const char *path, double *T, long cap, long *count) {
    if (!path || !count || cap < 0) return SPHFEM_ERR_ARG;
    std::vector<double> v;
    std::string err;
    if (int rc = read_vtk_temperature(path, v, err)) return rc;
    for (long i = 0; T && i < cap && i < (long)v.size(); i++) T[i] = v[(size_t)i];
    *count = (long)v.size();
    return SPHFEM_OK;
}

int sphfem_tool_wear_create(sphfem_ctx *ctx, double usui_K, double usui_alpha, sphfem_double2 tool_vel,
                            sphfem_tool_wear **out) {
    if (!out) return SPHFEM_ERR_ARG;
    *out = nullptr;
    return guarded(ctx, [&] {
        auto *w = new sphfem_tool_wear();
        w->K = usui_K;
        w->alpha = usui_alpha;
        w->vel = tool_vel;
        try {
            w->d_buf = dalloc<double>(4 + 4 * 256);
        } catch (...) {
            delete w;
            throw;
        }
        *out = w;
        return SPHFEM_OK;
    });
}

int sphfem_tool_wear_destroy(sphfem_ctx *ctx, sphfem_tool_wear *w) {
    if (!w) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        dfree(w->d_buf);
        delete w;
        return SPHFEM_OK;
    });
}

// tool_wear.cu:14-57
int sphfem_tool_wear_eval_usui(sphfem_ctx *ctx, sphfem_tool_wear *w, const sphfem_particles *p, double dt,
                               double *usui_min, double *usui_max, double *usui_avg) {
    if (!w || !p || !usui_min || !usui_max || !usui_avg) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        launch_usui(ctx, *p, w->K, w->alpha, w->vel, w->d_buf);
        double r[4];
        SPHFEM_HIP_CHECK(hipMemcpyAsync(r, w->d_buf, sizeof(r), hipMemcpyDeviceToHost, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        const double n = r[3];
        *usui_min = (n == 0.) ? 0. : r[1];
        *usui_max = r[2];
        *usui_avg = (n == 0.) ? 0. : r[0] / n;
        w->accum += dt * *usui_avg;
        return SPHFEM_OK;
    });
}

int sphfem_tool_wear_get_accum_wear(const sphfem_tool_wear *w, double *accum) {
    if (!w || !accum) return SPHFEM_ERR_ARG;
    *accum = w->accum;
    return SPHFEM_OK;
}

int sphfem_report_force(sphfem_ctx *ctx, sphfem_particles *p, sphfem_double2 *out) {
    if (!p || !out) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        double *d = dalloc<double>(2 + 2 * SPHFEM_FORCE_BLOCKS);
        launch_report_force(ctx, *p, d + 2, d);
        SPHFEM_HIP_CHECK(hipMemcpyAsync(out, d, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        dfree(d);
        return SPHFEM_OK;
    });
}

int sphfem_report_force_noalloc(sphfem_ctx *ctx, const sphfem_particles *p, sphfem_double2 *out) {
    if (!ctx || !p || !out) return SPHFEM_ERR_ARG;
    if (p->N > SPHFEM_FORCE_MAX_N)
        return fail(ctx, SPHFEM_ERR_ARG, "report_force_noalloc: more than 2^30 particles (32-bit reduction indices)");
    if (!ctx->force_scratch) return fail(ctx, SPHFEM_ERR_STATE, "report_force_noalloc: no scratch address");
    return guarded(ctx, [&] {
        launch_report_force(ctx, *p, ctx->force_scratch, reinterpret_cast<double *>(out));
        return SPHFEM_OK;
    });
}

int sphfem_timing_enable(sphfem_ctx *ctx, int enable) {
    if (!ctx) return SPHFEM_ERR_ARG;
    ctx->timing.enabled = enable != 0;
    return SPHFEM_OK;
}

int sphfem_timing_select(sphfem_ctx *ctx, const char *names) {
    if (!ctx) return SPHFEM_ERR_ARG;
    ctx->timing.only = names ? names : "";
    return SPHFEM_OK;
}

int sphfem_timing_reset(sphfem_ctx *ctx) {
    return guarded(ctx, [&] {
        SPHFEM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        timing_collect(ctx);
        ctx->timing.total_ms.clear();
        ctx->timing.count.clear();
        return SPHFEM_OK;
    });
}

int sphfem_timing_get(sphfem_ctx *ctx, const char *kernel, double *total_ms, long *launches) {
    if (!kernel) return SPHFEM_ERR_ARG;
    return guarded(ctx, [&] {
        timing_collect(ctx);
        auto it = ctx->timing.total_ms.find(kernel);
        if (total_ms) *total_ms = it == ctx->timing.total_ms.end() ? 0. : it->second;
        if (launches) *launches = ctx->timing.count.count(kernel) ? ctx->timing.count[kernel] : 0;
        return SPHFEM_OK;
    });
}

const char *sphfem_timing_names(sphfem_ctx *ctx) {
    if (!ctx) return "";
    timing_collect(ctx);
    ctx->timing.names.clear();
    for (auto &kv : ctx->timing.total_ms) ctx->timing.names += kv.first + ",";
    return ctx->timing.names.c_str();
}

}  // extern "C"
