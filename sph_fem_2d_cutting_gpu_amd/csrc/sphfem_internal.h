// Internal declarations shared by the HIP translation units of libsphfem_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/sphfem.h"

// largest number of particles that may change cells in one step for the incremental re-sort
#define SPHFEM_RESORT_CAP 4096
// largest N whose full-sort fallback may run in one workgroup (tuning resort_one_block)
#define SPHFEM_SORT_ONE_BLOCK_MAX (1 << 19)
// blocks of the bounding-box reduction of an unlocked grid
#define SPHFEM_BBOX_BLOCKS 1024
// d_flags slot set by the reorder when a live particle carries a non-finite pair input (the pair
// kernels then use the reference's full 5x5-cell window) and cleared by the update kernel
#define SPHFEM_FLAG_NONFINITE 5
// records per step of each ordered segment sum (friction heat, contact heat exchange); the
// reduction sorts them in LDS
#define SPHFEM_SEG_SINK_CAP 4096
#define SPHFEM_FLAG_SINK_FRIC 6
#define SPHFEM_FLAG_SINK_S2 7
#define SPHFEM_NUM_FLAGS 8
// position records a wavefront of the staged FP32C pair kernel copies into LDS (8 KiB)
#define SPHFEM_STAGE_CAP 512
// blocks of the report_force reduction (its partials: 2 doubles per block) and the largest particle
// count of the allocation-free variant
#define SPHFEM_FORCE_BLOCKS 256
#define SPHFEM_FORCE_MAX_N (1u << 30)

namespace sphfem {

// All constant blocks, passed to kernels by value (replaces the __constant__ copies that
// the reference keeps per translation unit: interactions_gpu.cu:15-20, actions_gpu.cu:23-29,
// FEM_thermal_gpu.cu:22, tool_FEM.cu:19-21).
struct DevConsts {
    sphfem_phys_constants phys;
    sphfem_trml_constants trml_wp;
    sphfem_trml_constants trml_tool;
    sphfem_corr_constants corr;
    sphfem_joco_constants joco;
    sphfem_geom_constants geom;
    sphfem_tool_constants tc;
    sphfem_wear_constants wear;
    double dt;
    double rr2_min;
    int wear_pro;
    int thermal_bc;   // actions_gpu.cu:938-943
    int plastic;      // actions_gpu.cu:921-924
    int fric_heat;    // actions_gpu.cu:931-934 (set once the workpiece's eta is non-zero)
    // interactions_gpu.cu:19 / :24: the tool-particle thermal block of the pair sums (device copy,
    // written only for alpha != 0) and the host copy's alpha read by the PSE launch
    sphfem_trml_constants trml_itool;
    double itool_alpha_pse;
    // a von Mises stress below jc_svm_below at a temperature in [jc_T_lo, jc_T_hi] (with eps_pl and
    // its rate >= 0) lies below the Johnson-Cook yield stress whatever the plastic state; -1: off
    double jc_svm_below = -1., jc_T_lo = 0., jc_T_hi = 0.;
    // actions_gpu.cu:222-232's c0 * c0, c0 = sqrtf(K / rho0) (single-precision square root), per
    // physical-constant block
    double eos_c02 = 0.;
};

// Device-resident tool kinematics (the host bookkeeping of tool_FEM.cu:1552-1570 plus the
// __device__ globals flank_retreat / flank_contact of tool_FEM.cu:19-20).
typedef sphfem_tool_kinematics ToolKin;

// Views of the particle arrays passed to kernels.
typedef sphfem_particles Particles;

// FP32 derivative carry of the FP32C form: SoA blocks of float2 pos_t, float2 vel_t, float4 S_t,
// float rho_t, float T_t over n particles from one base pointer (10 n floats)
struct RateView {
    float2 *pos_t, *vel_t;
    float4 *S_t;
    float *rho_t, *T_t;
};
inline RateView rate_view(float *base, size_t n) {
    RateView r;
    r.pos_t = reinterpret_cast<float2 *>(base);
    r.vel_t = reinterpret_cast<float2 *>(base + 2 * n);
    r.S_t = reinterpret_cast<float4 *>(base + 4 * n);
    r.rho_t = base + 8 * n;
    r.T_t = base + 9 * n;
    return r;
}

// FP64 pair sums of one particle (interactions_gpu.cu:143-379 outputs, T as the Laplacian)
struct PairSums {
    sphfem_double4 v_der, S_der;
    sphfem_double2 vel_t, pos_t;
    double T_lapl;
};

struct Timing {
    bool enabled = false;
    std::string only;     // comma-separated kernel names to time (empty: all)
    struct Rec { hipEvent_t a, b; };
    std::map<std::string, std::vector<Rec>> pending;
    std::map<std::string, double> total_ms;
    std::map<std::string, long> count;
    std::string names;
};

}  // namespace sphfem

struct sphfem_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    sphfem::DevConsts c{};
    unsigned set_mask = 0;
    std::string err;
    int *d_flags = nullptr;       // [0]: error bits, [1]: near-list count, [2]: unused,
                                  // [3]: re-sort mode (1 full, 2 incremental), [4]: changed count,
                                  // [5]: non-finite pair inputs present (SPHFEM_FLAG_NONFINITE)
    double *force_scratch = nullptr;  // this device's `scratch` array (sphfem_api.cpp)
    sphfem::Timing timing;
};

struct sphfem_grid {
    sphfem_geom_constants geom{};
    int num_cell = 0;
    int num_part = 0;
    double bbmax_x = 0, bbmax_y = 0;
    int key_bits = 1;
    int *cell_start = nullptr, *cell_end = nullptr;
    // radix sort scratch
    int *keys_alt = nullptr, *vals_alt = nullptr;
    int *hist = nullptr;           // [256 * nblocks]
    int *digit_tot = nullptr;      // [256]
    int nblocks = 0;
    // double buffer of the 20 gathered arrays (grid_gpu_green.h:21-50)
    sphfem_particles buf{};
    bool first_sort = true;
    // incremental re-sort of the step (particles stored in the previous sorted order)
    unsigned long long *cmask = nullptr;   // [nw] bit per position: cell changed this step
    int *wprefix = nullptr;                // [nw] changed positions before each 64-block, within its tile
    int *bprefix = nullptr;                // [nblocks + 1] changed positions before each 4096-tile
    int2 *clist = nullptr;                 // [SPHFEM_RESORT_CAP] changed (key, position), position order
    unsigned long long *csorted = nullptr; // [SPHFEM_RESORT_CAP] changed (key << 32 | position), sorted
    int4 *cdest = nullptr;                 // [SPHFEM_RESORT_CAP] (destination, key, position)
    int resort_limit = SPHFEM_RESORT_CAP;  // max changed particles for the incremental path (0: off)
    // grid_gpu_rothlin: in-place gathers and dense cell offsets
    int kind = SPHFEM_GRID_GREEN;
    int *perm = nullptr;                   // [N] permutation of the last Rothlin sort
    int *iota = nullptr;                   // [N] 0..N-1: the identity gather of the step after an in-place sort
    // geometry from the particles' bounding box each step (grid_base(max_cell, N), grid.cu:88-127)
    bool locked = true;
    int max_cell = 0;
    double *bbox_part = nullptr;           // [5 * SPHFEM_BBOX_BLOCKS] per-block min x, min y, max x, max y, max h
    double *bbox_out = nullptr;            // [5]
};

struct sphfem_tool {
    int n_nodes = 0, n_elem = 0, nnz = 0, n_seg = 0, n_colors = 0, n_cut_colors = 0;
    // mesh and FEM system
    sphfem_double2 *pos = nullptr;
    sphfem_int4 *elem = nullptr;
    int *row_ptr = nullptr, *col_ind = nullptr, *diag = nullptr;
    double *K = nullptr, *K_temp = nullptr, *M = nullptr, *Minv = nullptr, *f = nullptr, *f_temp = nullptr;
    double *T = nullptr, *T_old = nullptr, *k = nullptr;
    // segments (segment_FEM_gpu.h:31-56)
    sphfem_double2 *left = nullptr, *right = nullptr, *n = nullptr, *heat_exchange = nullptr;
    double *fric_heat = nullptr;
    sphfem_double2 *wear_rate = nullptr;
    double *wear_nodes = nullptr;
    sphfem_double4 *physical_para = nullptr;
    double *sliding_force = nullptr;
    sphfem_int4 *marks = nullptr;
    // deterministic gather lists
    int2 *node_robin = nullptr;        // [n_nodes] up to two segments, (colour, segment) order, -1 = none
    int *node_shift_ptr = nullptr;     // [n_nodes + 1] CSR of (segment, side) wear contributions
    int2 *node_shift = nullptr;
    int n_layer = 0;
    int *layer_elems = nullptr;        // [n_layer] element ids in (colour, element) order
    double *layer_old = nullptr, *layer_new = nullptr;   // [n_layer * 10] mass + 3x3 stiffness
    int *nnz_layer_ptr = nullptr;      // [nnz + 1] CSR over layer contributions (layer slot * 16 + a*3+b)
    int *nnz_layer = nullptr;
    int *node_layer_ptr = nullptr;     // [n_nodes + 1]
    int *node_layer = nullptr;         // layer slot * 4 + a
    sphfem::ToolKin *kin = nullptr;    // device
    sphfem_tool_kinematics host_kin{};
    // remeshing (present when the description carried node flags and the fixed boundary)
    bool remeshable = false;
    int n_fixed = 0;                   // fixed-boundary elements
    int *node_flags = nullptr;         // [n_nodes] bit 0 cutting edge, bit 1 fixed boundary
    int *dir_node = nullptr;           // [n_nodes] 1 if a Dirichlet row
    int *dir_ptr = nullptr;            // [n_nodes + 1] CSR: entries (F, x) of the Dirichlet rows F
    int *dir_entry = nullptr;          //   that reach node x, in the order the rows are applied
    sphfem_double2 *pos_map = nullptr; // [n_nodes] positions and temperatures before a remesh
    double *T_map = nullptr;
    unsigned long generation = 0;      // bumped when the device arrays are replaced
    sphfem_int4 *h_elem = nullptr;     // host copy of the elements (new[]), for the element bins
    int *bk_ptr = nullptr, *bk_list = nullptr;   // element bins of the last remesh (device)
    size_t bk_ptr_cap = 0, bk_list_cap = 0;
    // records of the ordered segment sums (SegSink): friction heat (6 values) and the contact heat
    // exchange (2 values), SPHFEM_SEG_SINK_CAP each; their counters are d_flags[6] and d_flags[7]
    unsigned long long *sink_keys = nullptr;     // [2 * cap]
    double *sink_vals = nullptr;                 // [8 * cap]
};

// The tool's elements binned on a uniform grid by their bounding boxes (padded well beyond the
// interpolation tolerances), each bin listing its elements in increasing order: the first
// element of a point's bin that holds the point is the first element of the whole mesh that does.
struct ElemBuckets {
    int nx = 0, ny = 0;
    double x0 = 0., y0 = 0., inv = 0.;
    const int *ptr = nullptr, *list = nullptr;
};

// rigid (SPH) tool of tool_gpu.cu: the geometry and contact constants in device memory
namespace sphfem { struct RigidTool; }
struct sphfem_rigid_tool {
    sphfem::RigidTool *dev = nullptr;
};

struct sphfem_leap_frog {
    unsigned int num_part = 0, num_cell = 0;
    int precision = SPHFEM_FP64;
    sphfem_tuning tuning = {256, 32, 256, SPHFEM_PAIRS_LIST_FUSED, 0};
    sphfem_double2 *pos_init = nullptr, *vel_init = nullptr;
    sphfem_double4 *S_init = nullptr;
    double *rho_init = nullptr, *T_init = nullptr;
    int *cell_start = nullptr, *cell_end = nullptr;   // only used by the unfused path
    // FP64 form: factors of each sorted particle's neighbour terms that depend on that particle
    // alone, computed once by the reorder with the operations the pair loop would repeat per pair:
    // pre_vol = m/rho, pre_s = {(Sxx-p)/rho^2, Sxy/rho^2, (Syy-p)/rho^2, sqrtf(K/rho)}
    double *pre_vol = nullptr;
    sphfem_double4 *pre_s = nullptr;
    sphfem::PairSums *psum64 = nullptr;   // FP64 form: pair sums of the sorted particles
    int *near_list = nullptr;      // [num_part] particles near the tool
    bool intermediates = false;    // also write S_der / v_der in the fused step
    // FP32C neighbour records of the sorted predicted state, SoA blocks r0 [num_part + 4],
    // r1, r2, r3 [num_part]: {x, y relative to the cell corner, cell row, h},
    // {vx, vy, (Sxx-p)/rho^2, Sxy/rho^2}, {(Syy-p)/rho^2, R}, {+-mass/rho, rho, sqrt(K/rho), T - T_init}
    float4 *rec = nullptr;
    // FP32C pair sums of the sorted particles: [0] velocity gradient (xx, xy, yx, yy),
    // [1] stress-divergence + viscous acceleration (x, y), XSPH velocity (x, y); psum_T: Laplacian of T
    float4 *psum = nullptr;
    float *psum_T = nullptr;
    // FP32C form: the derivative carry (pos_t, vel_t, S_t, rho_t, T_t) in FP32, two buffers of 10
    // floats per particle (SoA, see rate_view); rate32[rate_cur] is in the particles' current order.
    // While rates_fp32 is set the carry, not the particle arrays' FP64 rates, holds the rates:
    // sphfem_leap_frog_export_rates writes them out, sphfem_leap_frog_import_rates reads them in.
    float *rate32[2] = {nullptr, nullptr};
    int rate_cur = 0;
    bool rates_fp32 = false;
};

namespace sphfem {

// launch helpers ------------------------------------------------------------
void timing_begin(sphfem_ctx *ctx, const char *name, hipEvent_t *ev);
void timing_end(sphfem_ctx *ctx, const char *name, hipEvent_t ev);

struct KTimer {
    sphfem_ctx *ctx;
    const char *name;
    hipEvent_t ev = nullptr;
    KTimer(sphfem_ctx *c, const char *n) : ctx(c), name(n) { timing_begin(ctx, name, &ev); }
    ~KTimer() { timing_end(ctx, name, ev); }
};

inline int nblk(long n, int bs) { return (int)((n + bs - 1) / bs); }

// Workgroups are dispatched round-robin over the 8 XCDs of the chip, each with its own L2.
// Map the hardware block id to a logical one so that each XCD works on one contiguous run of
// logical blocks (neighbouring particles, which share neighbour records, stay in one L2).
// Bijective for any grid size.
__device__ __forceinline__ int xcd_logical_block() {
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int q = nb >> 3, r = nb & 7;
    const int x = b & 7, local = b >> 3;
    return x < r ? x * (q + 1) + local : r * (q + 1) + (x - r) * q + local;
}

// Streaming (nontemporal) loads for per-particle state that a kernel reads once and that is not
// read again before the step's other streams have passed through the caches: the load does not
// displace cache lines that are reused soon.  Same values as plain loads.
typedef double sphfem_v2d __attribute__((ext_vector_type(2)));
typedef float sphfem_v2f __attribute__((ext_vector_type(2)));
typedef float sphfem_v4f __attribute__((ext_vector_type(4)));
__device__ __forceinline__ double ld_stream(const double *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ld_stream(const float *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ int ld_stream(const int *p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ sphfem_double2 ld_stream(const sphfem_double2 *p) {
    const sphfem_v2d v = __builtin_nontemporal_load(reinterpret_cast<const sphfem_v2d *>(p));
    return {v.x, v.y};
}
__device__ __forceinline__ sphfem_double4 ld_stream(const sphfem_double4 *p) {
    const sphfem_v2d a = __builtin_nontemporal_load(reinterpret_cast<const sphfem_v2d *>(p));
    const sphfem_v2d b = __builtin_nontemporal_load(reinterpret_cast<const sphfem_v2d *>(p) + 1);
    return {a.x, a.y, b.x, b.y};
}
__device__ __forceinline__ float2 ld_stream(const float2 *p) {
    const sphfem_v2f v = __builtin_nontemporal_load(reinterpret_cast<const sphfem_v2f *>(p));
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 ld_stream(const float4 *p) {
    const sphfem_v4f v = __builtin_nontemporal_load(reinterpret_cast<const sphfem_v4f *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
}
// streaming stores, for results not read again before the caches have turned over
__device__ __forceinline__ void st_stream(double *p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_stream(float *p, float v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_stream(int *p, int v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void st_stream(sphfem_double2 *p, sphfem_double2 v) {
    __builtin_nontemporal_store(sphfem_v2d{v.x, v.y}, reinterpret_cast<sphfem_v2d *>(p));
}
__device__ __forceinline__ void st_stream(sphfem_double4 *p, sphfem_double4 v) {
    __builtin_nontemporal_store(sphfem_v2d{v.x, v.y}, reinterpret_cast<sphfem_v2d *>(p));
    __builtin_nontemporal_store(sphfem_v2d{v.z, v.w}, reinterpret_cast<sphfem_v2d *>(p) + 1);
}
__device__ __forceinline__ void st_stream(float2 *p, float2 v) {
    __builtin_nontemporal_store(sphfem_v2f{v.x, v.y}, reinterpret_cast<sphfem_v2f *>(p));
}
__device__ __forceinline__ void st_stream(float4 *p, float4 v) {
    __builtin_nontemporal_store(sphfem_v4f{v.x, v.y, v.z, v.w}, reinterpret_cast<sphfem_v4f *>(p));
}
// a store that streams when NT
template <bool NT, class T> __device__ __forceinline__ void st_maybe_stream(T *p, T v) {
    if constexpr (NT) st_stream(p, v);
    else *p = v;
}

// sort / cells (sphfem_sort.hip)
void launch_hash(sphfem_ctx *ctx, sphfem_grid *g, const sphfem_particles &p, bool clear_old, int *hist0);
void launch_radix_sort(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles &p, bool hash_done_with_hist0,
                       int *vals_out = nullptr);
void launch_rothlin_sort(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles &p);
void launch_cells_dense(sphfem_ctx *ctx, const int *sorted_keys, int N, int num_cell, int *cell_start, int *cell_end);
void launch_iota(sphfem_ctx *ctx, int *ptr, int n);
void launch_bbox(sphfem_ctx *ctx, const sphfem_particles &p, double *part, double *out);
// t: the FEM tool whose step start (tool_begin_item) runs in extra workgroups of the hash pass, or NULL
void launch_resort(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles &p, bool one_block_fallback,
                   sphfem_tool *t = nullptr);
void launch_fill_int(sphfem_ctx *ctx, int *ptr, long n, int value);
void launch_reorder_full(sphfem_ctx *ctx, sphfem_grid *g, const sphfem_particles &p);
// FP32C: the pre-step rates are read from the carry buffer rates (in the order of p)
void launch_reorder_predict(sphfem_ctx *ctx, sphfem_grid *g, const sphfem_particles &p, sphfem_leap_frog *lf,
                            const RateView *rates = nullptr);
void launch_records_state(sphfem_ctx *ctx, const sphfem_particles &p, int num_cell, float4 *rec);

// particles (sphfem_particles.hip)
void launch_actions(sphfem_ctx *ctx, int which, const sphfem_particles &p);
void launch_leapfrog(sphfem_ctx *ctx, int which, const sphfem_particles &p, sphfem_leap_frog *lf, double dt);
void launch_interactions_ref(sphfem_ctx *ctx, const sphfem_particles &p, const int *cs, const int *ce);
void launch_interactions_pse(sphfem_ctx *ctx, const sphfem_particles &p, const int *cs, const int *ce);
void launch_interactions_fp32c(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_leap_frog *lf, const int *cs,
                               const int *ce, int num_cell);
// FP32C: rates_old holds the pre-step rates in the order of A, rates_new receives the new rates in
// sorted order; skip_contact: the contact fields of a particle with on_seg = -1 are zero (FEM tool,
// Green grid) and are not read
void launch_fused_update(sphfem_ctx *ctx, const sphfem_particles &B, const sphfem_particles &A,
                         sphfem_leap_frog *lf, const int *cs, const int *ce, sphfem_tool *t,
                         sphfem_rigid_tool *rt = nullptr, const RateView *rates_old = nullptr,
                         const RateView *rates_new = nullptr, bool skip_contact = false);
// part: 2 * SPHFEM_FORCE_BLOCKS doubles of block partials; out: the 2 sums
void launch_report_force(sphfem_ctx *ctx, const sphfem_particles &p, double *part, double *out);
// FP32C derivative carry: the particle arrays' FP64 rates into / out of a carry buffer; a gather
// dst[i] = src[perm[i]] of a carry buffer
void launch_rates_import(sphfem_ctx *ctx, const sphfem_particles &p, const RateView &dst);
void launch_rates_export(sphfem_ctx *ctx, const RateView &src, const sphfem_particles &p);
void launch_rates_gather(sphfem_ctx *ctx, const RateView &src, const RateView &dst, const int *perm, int n);

// tool (sphfem_tool.hip)
void launch_fric_heat(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_tool *t);
void launch_contact(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_tool *t);
void launch_tool_begin(sphfem_ctx *ctx, sphfem_tool *t);

// the start of a step for the FEM tool, item i of max(nodes, segments): reset_FEM_server's
// T_old = T and calculate_temp(0.5 dt) for the nodes; zero the wear accumulators
// (segments_wear_rate_reset) and, item 0, the near-list and segment-sink counters and the
// window-form switch of the pair kernels
__device__ __forceinline__ void tool_begin_item(const sphfem_tool &t, double half_dt, int *counters, int i) {
    if (i == 0) {
        counters[1] = 0;
        counters[SPHFEM_FLAG_NONFINITE] = 0;
        counters[SPHFEM_FLAG_SINK_FRIC] = 0;
        counters[SPHFEM_FLAG_SINK_S2] = 0;
    }
    if (i < t.n_nodes) {
        double T = t.T[i];
        t.T_old[i] = T;
        t.T[i] = half_dt * t.k[i] + T;
    }
    if (i < t.n_seg) {
        t.wear_rate[i] = {0., 0.};
        t.wear_nodes[i] = 0.;
        t.physical_para[i] = {0., 0., 0., 0.};
        t.sliding_force[i] = 0.;
    }
}
void launch_tool_fem_end(sphfem_ctx *ctx, sphfem_tool *t);
void launch_tool_move(sphfem_ctx *ctx, sphfem_tool *t, double dt, bool segments, bool nodes);
// launch_tool_fem_end then launch_tool_move(dt, segments, nodes), the temperature update and the motion
// in one launch (the step without wear)
void launch_tool_fem_end_move(sphfem_ctx *ctx, sphfem_tool *t);
void launch_tool_fem_end_wear(sphfem_ctx *ctx, sphfem_tool *t);
void launch_wear(sphfem_ctx *ctx, sphfem_tool *t);
void launch_remesh_low_level(sphfem_ctx *ctx, sphfem_tool *t, int freq, const ElemBuckets &b);
void launch_interp_new(sphfem_ctx *ctx, const sphfem_tool *old_t, const sphfem_double2 *new_pos, int n_new,
                       double *T_new, const ElemBuckets &b);
void launch_fem_reset(sphfem_ctx *ctx, sphfem_tool *t);
void launch_fem_temp(sphfem_ctx *ctx, sphfem_tool *t, double dt);
void launch_fem_robin(sphfem_ctx *ctx, sphfem_tool *t);
void launch_fem_rate(sphfem_ctx *ctx, sphfem_tool *t);
void launch_wear_reset(sphfem_ctx *ctx, sphfem_tool *t);
void launch_near_contact(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_tool *t, sphfem_leap_frog *lf);
// the ordered segment sums of the friction-heat records (after the pass that produced them)
void launch_fric_heat_sums(sphfem_ctx *ctx, sphfem_tool *t);

// rigid tool (sphfem_rigid.hip)
void launch_rigid_advance(sphfem_ctx *ctx, sphfem_rigid_tool *rt);
void launch_rigid_tool_particles(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_rigid_tool *rt);
void launch_rigid_contact(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_rigid_tool *rt);
void launch_rigid_fric_heat(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_double2 vel);
// d_out: [4 + 1024] doubles
void launch_usui(sphfem_ctx *ctx, const sphfem_particles &p, double K, double alpha, sphfem_double2 vel,
                 double *d_out);

// output formats (sphfem_output.cpp): host writers; `err` receives the message of a failure
int write_particles_vtk(const sphfem_particles &host, const char *path, bool skip_blanked, std::string &err);
int write_rigid_tool_vtk(const sphfem_tool_geometry &g, const char *path, std::string &err);
struct FemToolHost {
    int n_nodes = 0, n_elem = 0, n_seg = 0;
    std::vector<sphfem_double2> pos, left, right, n;
    std::vector<sphfem_int4> elem;
    std::vector<double> T, sliding_force, wear_nodes;
    std::vector<sphfem_double4> physical_para;
};
int write_fem_tool_vtk(const FemToolHost &t, sphfem_double2 shift, const char *path, std::string &err);
int write_fem_segments_vtk(const FemToolHost &t, const char *path, std::string &err);
int write_contact_variables_vtk(const FemToolHost &t, const char *path, std::string &err);
int read_vtk_temperature(const char *path, std::vector<double> &T, std::string &err);

enum { ACT_EOS, ACT_ARTSTRESS, ACT_JAUMANN, ACT_CONTINUITY, ACT_MOMENTUM, ACT_ADVECTION, ACT_PLASTICITY,
       ACT_BC, ACT_BC_THERMAL, ACT_INVALIDATE };
enum { LF_INIT, LF_PREDICT, LF_CORRECT };

}  // namespace sphfem

#define SPHFEM_HIP_CHECK(expr)                                         \
    do {                                                               \
        hipError_t _e = (expr);                                        \
        if (_e != hipSuccess) {                                        \
            throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
        }                                                              \
    } while (0)
