// Particle kernels: one kernel per reference function (actions_gpu.cu, leap_frog.cu,
// interactions_gpu.cu) for the drop-in entry points, and the fused FP64 step kernel that runs
// pair sums, Jaumann rate, continuity, momentum, advection, corrector, friction heat,
// plasticity, the far-from-tool contact reset, boundary conditions and NaN invalidation in one
// pass over the sorted particles.
#include "sphfem_physics.h"
#include "sphfem_rigid.h"

#pragma clang fp contract(off)

namespace sphfem {

// ---------------------------------------------------------------------------------------------
// pair sums of interactions_gpu.cu:143-379 for particle i, neighbours visited in the reference
// order: column ii, then row jj, then sorted index.  Within a column the occupied cells of
// rows [low_j, high_j) are one contiguous run of sorted indices, so each column is one loop.
// Pairs outside the kernel support contribute exact zeros in the reference and are skipped.
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ void column_range(const int *__restrict__ cs, const int *__restrict__ ce, int col0,
                                             int low_j, int high_j, int &first, int &last) {
    first = -1;
    last = -1;
    for (int jj = low_j; jj < high_j; jj++) {
        int s = cs[col0 + jj];
        if (s != -1) { first = s; break; }
    }
    if (first < 0) return;
    for (int jj = high_j - 1; jj >= low_j; jj--) {
        if (cs[col0 + jj] != -1) { last = ce[col0 + jj]; break; }
    }
}

// PRE: the neighbour factors m/rho_j, (S_j - p_j)/rho_j^2, S_xy,j/rho_j^2 and sqrtf(K/rho_j) come
// from pre_vol/pre_s (computed per particle by the reorder with the same operations, so the sums
// are bit-identical); otherwise they are computed per pair as in the reference.
// LIST_BS > 0: the in-support neighbours go through a per-lane list of CHUNK entries in LDS
// (list = this lane's first entry, stride LIST_BS) before their terms are evaluated.
// WIN: every non-blanked particle of the 5x5 cells is evaluated, as the reference does, instead of
// the in-support ones only.  The kernel's zeros outside the support make the two the same sums for
// finite inputs; a neighbour with a non-finite input turns the sums NaN (0 * NaN) in the reference
// wherever it lies in the window, and only the WIN form reproduces that.
template <bool PRE, int LIST_BS = 0, int CHUNK = 0, bool WIN = false>
__device__ __forceinline__ PairSums pair_sums_fp64(const DevConsts &c, const sphfem_particles &P, int pidx,
                                                   double thermal_alpha, const int *__restrict__ cs,
                                                   const int *__restrict__ ce,
                                                   const double *__restrict__ pre_vol = nullptr,
                                                   const sphfem_double4 *__restrict__ pre_s = nullptr,
                                                   int *__restrict__ list = nullptr) {
    const double mass = c.phys.mass, K = c.phys.K;
    const double wdeltap = c.corr.wdeltap, alpha = c.corr.alpha, beta = c.corr.beta;
    const double eta = c.corr.eta, eps = c.corr.xspheps;
    const int nx = c.geom.nx, ny = c.geom.ny;
    const double rr2_min = c.rr2_min;
    const double is_tool_i = P.tool_particle[pidx];
    const sphfem_double2 pi = P.pos[pidx], vi = P.vel[pidx];
    const sphfem_double4 Si = P.S[pidx], Ri = P.R[pidx];
    const double hi = P.h[pidx], rhoi = P.rho[pidx], prsi = P.p[pidx];
    double Ti = 0.;
    if (thermal_alpha != 0.) Ti = P.T[pidx];
    const double rhoi21 = 1. / (rhoi * rhoi);
    const double fourh2 = 4 * hi * hi;
    const bool unit_wdp = wdeltap == 1.;      // w / 1 == w exactly
    double ci_pre = 0.;
    if constexpr (PRE) ci_pre = pre_s[pidx].w;
    const int hashi = P.hash[pidx];
    const int gi = hashi / ny;
    const int gj = hashi - gi * ny;
    const int low_i = gi - 2 < 0 ? 0 : gi - 2;
    const int low_j = gj - 2 < 0 ? 0 : gj - 2;
    const int high_i = gi + 3 > nx ? nx : gi + 3;
    const int high_j = gj + 3 > ny ? ny : gj + 3;
    PairSums r;
    r.v_der = {0., 0., 0., 0.};
    r.S_der = {0., 0., 0., 0.};
    r.vel_t = {0., 0.};
    r.pos_t = {0., 0.};
    r.T_lapl = 0.;
    // one in-support pair (xij, yij from the sorted positions); the order of the calls is the
    // reference's neighbour order, so the sums are the same whichever loop drives them
    auto pair = [&](int iter, const sphfem_double2 pj, double xij, double yij) {
        double w, w_x, w_y;
        cubic_spline(pi.x, pi.y, pj.x, pj.y, hi, rr2_min, w, w_x, w_y);
        const double rhoj = P.rho[iter];
        const double is_tool_j = P.tool_particle[iter];
        if (!(is_tool_i == 1.) && !(is_tool_j == 1.)) {
            const sphfem_double2 vj = P.vel[iter];
            const sphfem_double4 Rj = P.R[iter];
            const double hj = P.h[iter];
            double volj, sjx, sjy, sjz;
            if constexpr (PRE) {
                volj = pre_vol[iter];
                const sphfem_double4 f = pre_s[iter];
                sjx = f.x;
                sjy = f.y;
                sjz = f.z;
            } else {
                const sphfem_double4 Sj = P.S[iter];
                const double prsj = P.p[iter];
                volj = mass / rhoj;
                const double rhoj21 = 1. / (rhoj * rhoj);
                sjx = (Sj.x - prsj) * rhoj21;
                sjy = Sj.y * rhoj21;
                sjz = (Sj.z - prsj) * rhoj21;
            }
            r.v_der.x += (vj.x - vi.x) * w_x * volj;
            r.v_der.y += (vj.x - vi.x) * w_y * volj;
            r.v_der.z += (vj.y - vi.y) * w_x * volj;
            r.v_der.w += (vj.y - vi.y) * w_y * volj;
            double fab = unit_wdp ? w : w / wdeltap;
            fab *= fab;
            fab *= fab;
            double Rx = fab * (Ri.x + Rj.x);
            double Ry = fab * (Ri.y + Rj.y);
            double Rz = fab * (Ri.z + Rj.z);
            r.S_der.x += mass * ((Si.x - prsi) * rhoi21 + sjx + Rx) * w_x;
            r.S_der.y += mass * (Si.y * rhoi21 + sjy + Ry) * w_y;
            r.S_der.z += mass * (Si.y * rhoi21 + sjy + Ry) * w_x;
            r.S_der.w += mass * ((Si.z - prsi) * rhoi21 + sjz + Rz) * w_y;
            double vijx = vi.x - vj.x;
            double vijy = vi.y - vj.y;
            double vijposij = vijx * xij + vijy * yij;
            double rhoij = 0.5 * (rhoi + rhoj);
            if (vijposij < 0.) {
                double ci, cj;
                if constexpr (PRE) {
                    ci = ci_pre;
                    cj = pre_s[iter].w;
                } else {
                    ci = sqrtf_d(K / rhoi);
                    cj = sqrtf_d(K / rhoj);
                }
                double cij = 0.5 * (ci + cj);
                double hij = 0.5 * (hi + hj);
                double r2ij = xij * xij + yij * yij;
                double muij = (hij * vijposij) / (r2ij + eta * eta * hij * hij);
                double piij = (-alpha * cij * muij + beta * muij * muij) / rhoij;
                r.vel_t.x += -mass * piij * w_x;
                r.vel_t.y += -mass * piij * w_y;
            }
            r.pos_t.x += -eps * w * mass / rhoij * vijx;
            r.pos_t.y += -eps * w * mass / rhoij * vijy;
        }
        if (thermal_alpha != 0.) {
            double rij = sqrt(xij * xij + yij * yij);
            if (rij > 1e-8) {
                const double Tj = P.T[iter];
                double eijx = xij / rij;
                double eijy = yij / rij;
                double rij1 = 1. / rij;
                const double volT = PRE ? pre_vol[iter] : mass / rhoj;
                r.T_lapl += 2.0 * volT * (Ti - Tj) * rij1 * (eijx * w_x + eijy * w_y);
            }
        }
    };
    if constexpr (LIST_BS == 0) {
        for (int ii = low_i; ii < high_i; ii++) {
            int first, last;
            column_range(cs, ce, ii * ny, low_j, high_j, first, last);
            if (first < 0 || last > (int)P.N) continue;   // empty column (or an inconsistent cell table)
            for (int iter = first; iter < last; iter++) {
                if (P.blanked[iter] == 1.) continue;
                const sphfem_double2 pj = P.pos[iter];
                double xij = pi.x - pj.x;
                double yij = pi.y - pj.y;
                double rr2 = xij * xij + yij * yij;
                if (!WIN && (rr2 >= fourh2 || rr2 < rr2_min)) continue;
                pair(iter, pj, xij, yij);
            }
        }
    } else {
        // the in-support neighbours are listed first (in the same order) and the pair terms are
        // evaluated from the list, so the lanes of a wavefront evaluate pairs together instead of
        // idling on candidates outside the support
        int n = 0;
        auto run = [&]() {
            for (int k = 0; k < n; k++) {
                const int iter = list[k * LIST_BS];
                const sphfem_double2 pj = P.pos[iter];
                pair(iter, pj, pi.x - pj.x, pi.y - pj.y);
            }
            n = 0;
        };
        for (int ii = low_i; ii < high_i; ii++) {
            int first, last;
            column_range(cs, ce, ii * ny, low_j, high_j, first, last);
            if (first < 0 || last > (int)P.N) continue;
            for (int iter = first; iter < last; iter++) {
                if (P.blanked[iter] == 1.) continue;
                const sphfem_double2 pj = P.pos[iter];
                double xij = pi.x - pj.x;
                double yij = pi.y - pj.y;
                double rr2 = xij * xij + yij * yij;
                if (!WIN && (rr2 >= fourh2 || rr2 < rr2_min)) continue;
                list[n * LIST_BS] = iter;
                ++n;
                // a full list empties every active lane's list at once: lanes that flushed at
                // their own points would each run their list with the rest of the wavefront idle
                if (__any(n == CHUNK)) run();
            }
        }
        run();
    }
    return r;
}

// ---------------------------------------------------------------------------------------------
// FP32-compensated form
// ---------------------------------------------------------------------------------------------

// FP32C neighbour records (written by the reorder kernel), four SoA blocks of float4:
//   r0[j] = {x - cell corner x, y - cell corner y, cell row, h}   (block padded by 4 records)
//   r1[j] = {vx, vy, (Sxx - p)/rho^2, Sxy/rho^2}     r2[j] = {(Syy - p)/rho^2, Rxx, Rxy, Ryy}
//   r3[j] = {+-mass/rho, rho, sqrt(K/rho), T - T_init}
// The volume is stored negative for tool particles; blanked particles have x = 1e30.  SoA keeps a
// wavefront's gather of one field within few cache lines.
struct RecView {
    const float4 *__restrict__ r0;
    const float4 *__restrict__ r1;
    const float4 *__restrict__ r2;
    const float4 *__restrict__ r3;
};

// record load at a 32-bit byte offset from a uniform base (one VGPR offset, no 64-bit address math)
__device__ __forceinline__ float4 ldrec(const float4 *__restrict__ base, unsigned byte_off) {
    return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(base) + byte_off);
}

// Where the pair loops read the neighbours' position records (r0) from: HBM through L1 (the list
// kernel), or a wavefront's copy of its five columns' windows in LDS (the staged kernel).  Both give
// the same float4 values, so the sums do not depend on the source.
// A neighbour-list entry names the candidate's column d (0..4) and where its records are.
struct GlobalPos {
    using Entry = int;                                      // sorted index | column << 29
    const float4 *__restrict__ r0;
    // candidate j of column d during the scan
    __device__ __forceinline__ float4 scan(int, int j) const { return ldrec(r0, (unsigned)j << 4); }
    __device__ __forceinline__ Entry encode(int d, int j) const { return j | (d << 29); }
    __device__ __forceinline__ int column(Entry e) const { return (int)((unsigned)e >> 29); }
    __device__ __forceinline__ unsigned index(Entry e) const { return (unsigned)(e & 0x0fffffff); }
    __device__ __forceinline__ float4 entry(Entry e) const { return ldrec(r0, index(e) << 4); }
    // the range a lane scans in column d, limited to what the source holds (everything here)
    __device__ __forceinline__ bool clip(int, int &, int &) const { return true; }
};

struct StagedPos {
    using Entry = unsigned short;                           // LDS slot | column << 12
    const float4 *spos;       // LDS: the union windows of columns 0..4, one after the other
    const int *sdelta;        // LDS: per column, sorted index minus LDS slot
    int F[5], L[5], delta[5]; // per column: staged sorted range [F, L) and its delta (uniform)
    __device__ __forceinline__ float4 scan(int d, int j) const { return spos[j - delta[d]]; }
    __device__ __forceinline__ Entry encode(int d, int j) const { return (Entry)((j - delta[d]) | (d << 12)); }
    __device__ __forceinline__ int column(Entry e) const { return (int)(e >> 12); }
    __device__ __forceinline__ unsigned index(Entry e) const { return (unsigned)((e & 0xfff) + sdelta[e >> 12]); }
    __device__ __forceinline__ float4 entry(Entry e) const { return spos[e & 0xfff]; }
    // a lane's window lies inside the wavefront's union by construction when the hashes are sorted;
    // otherwise the range is cut to the staged part (no access outside the LDS copy) and false
    // reports the inconsistent table
    __device__ __forceinline__ bool clip(int d, int &first, int &last) const {
        const bool inside = first >= F[d] && last <= L[d];
        first = first < F[d] ? F[d] : first;
        last = last > L[d] ? L[d] : last;
        if (last < first) last = first;
        return inside;
    }
};

struct PairI {            // particle i and the per-i kernel constants
    float fxi, fyi, rowi, hi, vxi, vyi, sxi, syi, szi, Rxi, Ryi, Rzi, rhoi, ci, Ti;
    float h1, fac, c_out_w, c_out_g, c_in_g;
    float fourh2, rr2min;     // the support test (used by the window form)
    bool thermal;
    bool mech;            // i is not a tool particle
};

struct PairAcc {          // FP32 block sums
    float v0, v1, v2, v3, ax, ay, p0, p1, T;
};

template <bool WIN = false>
__device__ __forceinline__ void pair_term(const PairI &I, float hf, float inv_wdp, float alpha, float beta,
                                          float eta2, const float4 q0, const float4 q1, const float4 q2,
                                          const float4 q3, int d, PairAcc &a) {
#pragma clang fp contract(fast)
    const float dcol = (float)(d - 2);                          // column of j minus column of i
    const float xij = (I.fxi - q0.x) - dcol * hf;
    const float yij = (I.fyi - q0.y) + (I.rowi - q0.z) * hf;
    const float rr2 = xij * xij + yij * yij;
    const float rinv = __builtin_amdgcn_rsqf(rr2);
    const float q = rr2 * rinv * I.h1;
    float w, g;
    if (q > 1.f) {
        const float tq = 2.f - q;
        const float tq2 = tq * tq;
        w = I.c_out_w * (tq2 * tq2 * tq2);
        g = I.c_out_g * tq2 * rinv;
    } else {
        w = I.fac * (1.f - 1.5f * q * q * (1.f - 0.5f * q));
        g = I.c_in_g * (1.f - 0.75f * q);
    }
    if constexpr (WIN) {
        // a window candidate outside the support: the reference's kernel value and gradient are 0
        if (!(rr2 < I.fourh2 && rr2 >= I.rr2min)) {
            w = 0.f;
            g = 0.f;
        }
    }
    const float wx = g * xij, wy = g * yij;
    const float volj = fabsf(q3.x);
    if (I.thermal && rr2 > 1e-16f) a.T += volj * (I.Ti - q3.w) * g;
    if (I.mech && q3.x > 0.f) {                          // neither i nor j is a tool particle
        const float dvx = q1.x - I.vxi, dvy = q1.y - I.vyi;
        const float gx = dvx * volj, gy = dvy * volj;
        a.v0 += gx * wx;
        a.v1 += gx * wy;
        a.v2 += gy * wx;
        a.v3 += gy * wy;
        float f = w * inv_wdp;
        f *= f;
        f *= f;
        float Ax = (I.sxi + q1.z) + f * (I.Rxi + q2.y);
        const float Ay = (I.syi + q1.w) + f * (I.Ryi + q2.z);
        float Az = (I.szi + q2.x) + f * (I.Rzi + q2.w);
        const float inv_rhoij = 2.f * __builtin_amdgcn_rcpf(I.rhoi + q3.y);
        const float vp = -(dvx * xij + dvy * yij);          // (v_i - v_j) . x_ij
        if (vp < 0.f) {
            const float cij = 0.5f * (I.ci + q3.z);
            const float hij = 0.5f * (I.hi + q0.w);
            const float mu = hij * vp * __builtin_amdgcn_rcpf(rr2 + eta2 * hij * hij);
            const float piij = mu * (beta * mu - alpha * cij) * inv_rhoij;
            Ax -= piij;
            Az -= piij;
        }
        a.ax += Ax * wx + Ay * wy;
        a.ay += Ay * wx + Az * wy;
        const float wr = w * inv_rhoij;
        a.p0 -= wr * dvx;                                   // sum w / rho_ij * (v_i - v_j)
        a.p1 -= wr * dvy;
    }
}

// FP32-compensated pair sums (same physics as pair_sums_fp64, different arithmetic), one lane per
// sorted particle:
//  * positions are cell-relative FP32 (x_ij = dx_rel + (cell_i - cell_j) * h): the pair distance
//    carries ~1e-7 of the cell size whatever the absolute coordinates;
//  * one v_rsq per pair (the distance), one v_rcp (2 / (rho_i + rho_j)) and one v_rcp for the
//    viscosity denominator of approaching pairs;
//  * the grad-T term uses e_ij . grad W / r_ij = g (the radial kernel-derivative factor): the
//    reference's algebra without its extra sqrt and divisions;
//  * each lane compacts its in-support candidates (columns in order, sorted index within a
//    column: the reference order) into an LDS list of CHUNK entries, loading candidates four
//    at a time, then evaluates the dense list with the records of the next entry in flight;
//  * summation: blocks of 8 pair terms in FP32, block sums accumulated in FP64, which bounds the
//    rounding of each sum by ~8 units of FP32 roundoff of the sum of |terms|; one final FP32
//    rounding on output.
// Outputs (FP32): psum[i] = velocity gradient (xx, xy, yx, yy); psum[N + i] = (mass * stress
// divergence - mass * viscous term) per axis, XSPH velocity per axis; psum_T[i] = Laplacian of T.
// WIN: every non-blanked record of the 5x5 cells (blanked records sit at x = 1e30) is evaluated with
// the kernel set to 0 outside the support, as the reference does; see pair_sums_fp64.
__device__ __forceinline__ double interaction_alpha(const DevConsts &c, double is_tool) {
    // interactions_gpu.cu:164: the tool block (interactions_gpu.cu:19) is zero unless
    // interactions_setup_thermal_constants_tool set it (the rigid-tool setup does, benchmarks.cpp:518;
    // the FEM setups do not, benchmarks_NZ.cpp:354)
    return (is_tool == 0.) ? c.trml_wp.alpha : c.trml_itool.alpha;
}

// Arguments of the update (everything after the pair sums, leap_frog.cu:71-96 with the actions of
// the step), shared by the update kernel and the FP32C pair kernels that run it after their sums
struct UpdateArgs {
    sphfem_particles B, A;
    const sphfem_double2 *pos_init, *vel_init;
    const sphfem_double4 *S_init;
    const double *rho_init, *T_init;
    ToolView tv;
    int write_intermediates;
    const PairSums *psum64;
    const RigidTool *rt;
    RateView r_old, r_new;
    int skip_contact;
};

// the update of sorted particle i; g_sum, a_sum, T_lapl32 are its FP32C pair sums (the FP64 form
// reads psum64 instead).  NT: streaming stores, for the update run inside a pair kernel, whose
// neighbour records other wavefronts are still reading from L2 (plain stores would displace them)
template <bool F32C, bool RIGID, bool NT = false>
__device__ __forceinline__ void update_item(const DevConsts &c, const UpdateArgs &u, int i, float4 g_sum,
                                            float4 a_sum, float T_lapl32) {
    const sphfem_particles &B = u.B, &A = u.A;
    const sphfem_double2 *__restrict__ pos_init = u.pos_init;
    const sphfem_double2 *__restrict__ vel_init = u.vel_init;
    const sphfem_double4 *__restrict__ S_init = u.S_init;
    const double *__restrict__ rho_init = u.rho_init;
    const double *__restrict__ T_init = u.T_init;
    const ToolView &tv = u.tv;
    const int write_intermediates = u.write_intermediates;
    const PairSums *__restrict__ psum64 = u.psum64;
    const RigidTool *__restrict__ rt = u.rt;
    const RateView &r_old = u.r_old, &r_new = u.r_new;
    const int skip_contact = u.skip_contact;
    // FP32C: the reorder wrote only the neighbour records; every per-particle field is gathered here
    // from the unsorted arrays A through idx, and the sorted step result is written to B.  All loads
    // of the particle are issued before any branch on loaded values, so a wavefront waits for
    // memory once (the kernel holds 3 wavefronts per SIMD); the rates and pair sums of blanked
    // particles are read and not used.
    const int src = F32C ? B.idx[i] : i;
    sphfem_double2 pos, vel;
    sphfem_double4 S;
    double rho, T, eps_pl, eps_pl_dot, pr = 0.;
    sphfem_double2 fc0, ft0, nrm0;
    int on_seg0;
    sphfem_double2 v0 = {0., 0.};
    sphfem_double4 S0 = {0., 0., 0., 0.};
    double rho0 = 0.;
    double blanked0, tool, fixed, h_src = 0.;
    float2 vt_old = make_float2(0.f, 0.f);
    float4 St_old = make_float4(0.f, 0.f, 0.f, 0.f);
    float rt_old = 0.f, Tt_old = 0.f;
    if constexpr (F32C) {
        blanked0 = ld_stream(&A.blanked[src]);
        tool = ld_stream(&A.tool_particle[src]);
        fixed = ld_stream(&A.fixed[src]);
        pos = ld_stream(&A.pos[src]);
        v0 = ld_stream(&A.vel[src]);
        S0 = ld_stream(&A.S[src]);
        rho0 = ld_stream(&A.rho[src]);
        T = ld_stream(&A.T[src]);
        eps_pl = ld_stream(&A.eps_pl[src]);
        eps_pl_dot = ld_stream(&A.eps_pl_dot[src]);
        on_seg0 = ld_stream(&A.on_seg[src]);
        h_src = ld_stream(&A.h[src]);
        vt_old = ld_stream(&r_old.vel_t[src]);          // the FP32 derivative carry
        St_old = ld_stream(&r_old.S_t[src]);
        rt_old = ld_stream(&r_old.rho_t[src]);
        Tt_old = ld_stream(&r_old.T_t[src]);
    } else {
        blanked0 = B.blanked[i];
        tool = B.tool_particle[i];
        fixed = B.fixed[i];
    }
    const bool live = blanked0 != 1.;
    const double thermal_alpha = interaction_alpha(c, tool);
    PairSums r;
    if (live) {
        if constexpr (F32C) {
            // stress divergence and viscous acceleration come summed per axis
            r.v_der = {g_sum.x, g_sum.y, g_sum.z, g_sum.w};
            r.S_der = {a_sum.x, 0., a_sum.y, 0.};
            r.vel_t = {0., 0.};
            r.pos_t = {a_sum.z, a_sum.w};
            r.T_lapl = T_lapl32;
        } else {
            r = psum64[i];
        }
    }
    // FP32C: pos and T are the pre-step values (the corrector starts from them); velocity, stress and
    // density are predicted again from the pre-step state and rates (leap_frog.cu:45-69, the same
    // operations as in the reorder), as is the pressure (actions_gpu.cu:222-232)
    if constexpr (F32C) {
        vel = v0;
        S = S0;
        rho = rho0;
        if (live) {
            const double hdt = 0.5 * c.dt;
            vel.x = vel.x + hdt * vt_old.x;
            vel.y = vel.y + hdt * vt_old.y;
            S.x = S.x + hdt * St_old.x;
            S.y = S.y + hdt * St_old.y;
            S.z = S.z + hdt * St_old.z;
            S.w = S.w + hdt * St_old.w;
            rho = rho + hdt * (double)rt_old;
        }
        if (tool == 0.) pr = eos_pressure(c, rho);
        if (skip_contact && on_seg0 == -1) {
            // every writer of on_seg = -1 (contact s1, the far-from-tool reset) zeroes fc, ft and n
            // with it (actions_gpu.cu:455-491): nothing to read far from the tool
            const sphfem_double2 z = {0., 0.};
            fc0 = z;
            ft0 = z;
            nrm0 = z;
        } else {
            fc0 = ld_stream(&A.fc[src]);
            ft0 = ld_stream(&A.ft[src]);
            nrm0 = ld_stream(&A.n[src]);
        }
    } else {
        pos = B.pos[i];
        vel = B.vel[i];
        S = B.S[i];
        rho = B.rho[i];
        T = B.T[i];
        eps_pl = B.eps_pl[i];
        eps_pl_dot = B.eps_pl_dot[i];
        fc0 = B.fc[i];
        ft0 = B.ft[i];
        nrm0 = B.n[i];
        on_seg0 = B.on_seg[i];
    }
    double blanked = blanked0;
    sphfem_double2 fc = fc0, ft = ft0, nrm = nrm0;
    int on_seg = on_seg0;
    bool contact_written = false;
    bool near = false;
    sphfem_double2 pos_t = {0., 0.}, vel_t = {0., 0.};
    sphfem_double4 S_t = {0., 0., 0., 0.};
    double rho_t = 0., T_t = 0.;
    if (live) {
        pos_t = r.pos_t;
        vel_t = r.vel_t;
        if constexpr (F32C) T_t = thermal_alpha != 0. ? thermal_alpha * r.T_lapl : (double)Tt_old;
        else T_t = thermal_alpha != 0. ? thermal_alpha * r.T_lapl : A.T_t[A.idx[i]];
        if (tool == 0.) {
            S_t = jaumann_rate(c, r.v_der, S);        // actions_gpu.cu:280-314
            rho_t = -rho * (r.v_der.x + r.v_der.w);     // :626-639
            const double mass = c.phys.mass;            // :641-671
            // x / mass is x itself for x = +-0 (mass > 0): far from the tool the contact forces are
            // zero and their divisions are skipped, with the same sums
            sphfem_double2 fcm = fc, ftm = ft;
            if (fc.x != 0. || fc.y != 0. || ft.x != 0. || ft.y != 0.) {
                fcm = {fc.x / mass, fc.y / mass};
                ftm = {ft.x / mass, ft.y / mass};
            }
            vel_t.x += r.S_der.x + r.S_der.y + fcm.x + ftm.x;
            vel_t.y += r.S_der.z + r.S_der.w + fcm.y + ftm.y;
            pos_t.x = pos_t.x + vel.x;                  // :673-687
            pos_t.y = pos_t.y + vel.y;
        }
        if (write_intermediates) {
            A.S_der[i] = r.S_der;
            A.v_der[i] = r.v_der;
        }
        // leap_frog.cu:71-96
        const double dt = c.dt;
        const sphfem_double2 p0 = F32C ? pos : pos_init[i];
        if constexpr (!F32C) {
            v0 = vel_init[i];
            S0 = S_init[i];
            rho0 = rho_init[i];
        }
        const double T0 = F32C ? T : T_init[i];
        pos.x = p0.x + dt * pos_t.x;
        pos.y = p0.y + dt * pos_t.y;
        vel.x = v0.x + dt * vel_t.x;
        vel.y = v0.y + dt * vel_t.y;
        S.x = S0.x + dt * S_t.x;
        S.y = S0.y + dt * S_t.y;
        S.z = S0.z + dt * S_t.z;
        S.w = S0.w + dt * S_t.w;
        rho = rho0 + dt * rho_t;
        T = T0 + dt * T_t;
        if (tv.active) {
            if (!fric_heat_v21(c, tv, *tv.kin, i, eps_pl, ft, on_seg, nrm, vel, fc, rho, T)) blanked = 1.;
        }
        if constexpr (RIGID) {
            if (c.fric_heat && tool == 0.) T = rigid_fric_heat(c, rt->g.vel, vel, ft, nrm, T);
        }
        if (c.plastic && tool == 0. && blanked != 1.) {
            if (plasticity_jc(c, S, F32C ? pr : B.p[i], eps_pl, eps_pl_dot, T, rho)) blanked = 1.;
        }
        // contact s1 for particles outside the 1.05 ref_dist disc of the moved tool
        // (actions_gpu.cu:438-453); the others are finished by the near-tool pass
        if (tv.active && blanked != 1.) {
            ToolKin k = *tv.kin;
            k.tl_x += k.vel_x * dt;
            k.tl_y += k.vel_y * dt;
            if (near_tool(k, pos)) {
                near = true;
                int slot = atomicAdd(tv.near_count, 1);
                tv.near_list[slot] = i;
            } else {
                sphfem_double2 z = {0., 0.};
                fc = z; ft = z; nrm = z;
                on_seg = -1;
                contact_written = true;
            }
        }
    } else if constexpr (RIGID) {
        if (c.fric_heat && tool == 0.) T = rigid_fric_heat(c, rt->g.vel, vel, ft, nrm, T);
    }
    if constexpr (RIGID) {
        if (tool != 0.) {
            pos.x += c.dt * rt->g.vel.x;
            pos.y += c.dt * rt->g.vel.y;
        } else {
            rigid_contact(*rt, c.dt, pos, vel, ft0, fc, ft, nrm);
            contact_written = true;
        }
    }
    if (!near) {
        // actions_gpu.cu:765-781, 756-763, 787-824
        if (tool == 0. && fixed != 0.) {
            sphfem_double2 z = {0., 0.};
            vel = z;
            fc = z;
            pos_t = z;
            vel_t = z;
            contact_written = true;
        }
        if (c.thermal_bc && fixed == 1.) T = c.trml_wp.T_init;
        if (blanked != 1. && live) {
            bool bad = isnaninf(pos_t.x) || isnaninf(pos_t.y) || isnaninf(vel_t.x) || isnaninf(vel_t.y) ||
                       isnaninf(S_t.x) || isnaninf(S_t.y) || isnaninf(S_t.z) || isnaninf(S_t.w) ||
                       isnaninf(rho_t) || isnaninf(T_t);
            if (bad) blanked = 1.;
        }
    }
    if constexpr (F32C) {
        // the new rates in sorted order, FP32 (the pair sums they come from are FP32)
        st_maybe_stream<NT>(&r_new.pos_t[i], make_float2((float)pos_t.x, (float)pos_t.y));
        st_maybe_stream<NT>(&r_new.vel_t[i], make_float2((float)vel_t.x, (float)vel_t.y));
        st_maybe_stream<NT>(&r_new.S_t[i], make_float4((float)S_t.x, (float)S_t.y, (float)S_t.z, (float)S_t.w));
        st_maybe_stream<NT>(&r_new.rho_t[i], (float)rho_t);
        st_maybe_stream<NT>(&r_new.T_t[i], (float)T_t);
    } else {
        B.pos_t[i] = pos_t;
        B.vel_t[i] = vel_t;
        B.S_t[i] = S_t;
        B.rho_t[i] = rho_t;
        B.T_t[i] = T_t;
    }
    const sphfem_particles &O = F32C ? B : A;
    st_maybe_stream<NT>(&O.pos[i], pos);
    st_maybe_stream<NT>(&O.vel[i], vel);
    st_maybe_stream<NT>(&O.S[i], S);
    st_maybe_stream<NT>(&O.rho[i], rho);
    st_maybe_stream<NT>(&O.T[i], T);
    st_maybe_stream<NT>(&O.eps_pl[i], eps_pl);
    st_maybe_stream<NT>(&O.eps_pl_dot[i], eps_pl_dot);
    st_maybe_stream<NT>(&O.blanked[i], blanked);
    if constexpr (F32C) {
        // the sorted copy of every remaining field; p of tool particles stays at its array position,
        // as the reference neither gathers p nor computes it for them
        st_maybe_stream<NT>(&B.fc[i], fc);
        st_maybe_stream<NT>(&B.ft[i], ft);
        st_maybe_stream<NT>(&B.n[i], nrm);
        st_maybe_stream<NT>(&B.on_seg[i], on_seg);
        st_maybe_stream<NT>(&B.h[i], h_src);
        st_maybe_stream<NT>(&B.fixed[i], fixed);
        st_maybe_stream<NT>(&B.tool_particle[i], tool);
        st_maybe_stream<NT>(&B.p[i], tool == 0. ? pr : A.p[i]);
    } else if (contact_written && (fc.x != fc0.x || fc.y != fc0.y || ft.x != ft0.x || ft.y != ft0.y ||
                                   nrm.x != nrm0.x || nrm.y != nrm0.y || on_seg != on_seg0)) {
        // contact fields in place, written only when they change (far from the tool they stay zero)
        B.fc[i] = fc;
        B.ft[i] = ft;
        B.n[i] = nrm;
        B.on_seg[i] = on_seg;
    }
}

struct PairOut {           // one particle's FP32C sums: psum[i], psum[N + i], psum_T[i]
    float4 g, a;
    float T;
};

template <int BS, int CHUNK, bool WIN, class Pos = GlobalPos>
__device__ __forceinline__ PairOut pairs_fp32c_one(const DevConsts &c, const double *__restrict__ blanked,
                                                   const int *__restrict__ hash, int N, const RecView &rv,
                                                   const int *__restrict__ cs, const int *__restrict__ ce, int i,
                                                   typename Pos::Entry *__restrict__ list, const Pos &pos,
                                                   int *__restrict__ flags) {
#pragma clang fp contract(fast)
    double d_v0 = 0., d_v1 = 0., d_v2 = 0., d_v3 = 0.;
    double d_ax = 0., d_ay = 0., d_p0 = 0., d_p1 = 0., d_T = 0.;
    const float4 a3 = rv.r3[i];
    // tool particles (negative volume) sum only the Laplacian of T, and only when the tool's
    // thermal block of the pair sums is set (interactions_gpu.cu:157-164, 260-262); blanked
    // particles are skipped
    const bool tool_i = a3.x < 0.f;
    const double alpha_i = tool_i ? c.trml_itool.alpha : c.trml_wp.alpha;
    if (blanked[i] != 1. && (!tool_i || alpha_i != 0.)) {
        const int nx = c.geom.nx, ny = c.geom.ny;
        const int hashi = hash[i];
        const int gi = hashi / ny;
        const int gj = hashi - gi * ny;
        const int low_j = gj - 2 < 0 ? 0 : gj - 2;
        const int high_j = gj + 3 > ny ? ny : gj + 3;
        // cell-table entries of the five columns, issued together: the lowest and highest row of
        // each column's window decide its sorted range unless those cells are empty
        int cs_lo[5], cs_hi[5], ce_hi[5];
#pragma unroll
        for (int d = 0; d < 5; d++) {
            const int col = gi - 2 + d;
            const bool ok = col >= 0 && col < nx;
            const int c0 = (ok ? col : gi) * ny;
            cs_lo[d] = ok ? cs[c0 + low_j] : -1;
            cs_hi[d] = ok ? cs[c0 + high_j - 1] : -1;
            ce_hi[d] = ok ? ce[c0 + high_j - 1] : -1;
        }
        const float4 a0 = rv.r0[i], a1 = rv.r1[i], a2 = rv.r2[i];
        PairI I;
        I.fxi = a0.x; I.fyi = a0.y; I.rowi = a0.z; I.hi = a0.w;
        I.vxi = a1.x; I.vyi = a1.y; I.sxi = a1.z; I.syi = a1.w;
        I.szi = a2.x; I.Rxi = a2.y; I.Ryi = a2.z; I.Rzi = a2.w;
        I.rhoi = a3.y; I.ci = a3.z; I.Ti = a3.w;
        I.h1 = 1.f / I.hi;
        I.fac = (float)SPHFEM_PI_FAC * I.h1 * I.h1;
        I.c_out_w = 0.25f * I.fac;
        I.c_out_g = -0.75f * I.h1 * I.fac;
        I.c_in_g = -3.f * I.h1 * I.h1 * I.fac;
        I.thermal = alpha_i != 0.;
        I.mech = !tool_i;
        const float fourh2 = 4.f * I.hi * I.hi;
        const float rr2min = (float)c.rr2_min;
        I.fourh2 = fourh2;
        I.rr2min = rr2min;
        const float hf = (float)c.geom.dx;
        const float inv_wdp = (float)(1. / c.corr.wdeltap);
        const float alpha = (float)c.corr.alpha, beta = (float)c.corr.beta;
        const float eta2 = (float)(c.corr.eta * c.corr.eta);

        auto flush = [&](PairAcc &a) {
            d_v0 += a.v0; d_v1 += a.v1; d_v2 += a.v2; d_v3 += a.v3;
            d_ax += a.ax; d_ay += a.ay; d_p0 += a.p0; d_p1 += a.p1; d_T += a.T;
            a = PairAcc{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        };
        // dense list: two entries per iteration, the records of the next entry loaded before the
        // current one is evaluated (ping-pong registers, no copies); each list entry is read from LDS
        // one pair term before its records are loaded, so the LDS latency passes under that term
        auto run_list = [&](int n) {
            if (n <= 0) return;
            PairAcc acc{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            typename Pos::Entry eA = list[0];
            typename Pos::Entry eB = list[(1 < n ? 1 : n - 1) * BS];
            unsigned oA = pos.index(eA) << 4;
            float4 A0 = pos.entry(eA), A1 = ldrec(rv.r1, oA), A2 = ldrec(rv.r2, oA), A3 = ldrec(rv.r3, oA);
            for (int k = 0; k < n; k += 2) {
                const typename Pos::Entry eA_next = list[(k + 2 < n ? k + 2 : n - 1) * BS];
                const unsigned oB = pos.index(eB) << 4;
                const float4 B0 = pos.entry(eB), B1 = ldrec(rv.r1, oB), B2 = ldrec(rv.r2, oB),
                             B3 = ldrec(rv.r3, oB);
                pair_term<WIN>(I, hf, inv_wdp, alpha, beta, eta2, A0, A1, A2, A3, pos.column(eA), acc);
                if (k + 1 < n) {
                    const typename Pos::Entry eB_next = list[(k + 3 < n ? k + 3 : n - 1) * BS];
                    eA = eA_next;
                    oA = pos.index(eA) << 4;
                    A0 = pos.entry(eA); A1 = ldrec(rv.r1, oA); A2 = ldrec(rv.r2, oA); A3 = ldrec(rv.r3, oA);
                    pair_term<WIN>(I, hf, inv_wdp, alpha, beta, eta2, B0, B1, B2, B3, pos.column(eB), acc);
                    eB = eB_next;
                }
                if (((k + 2) & 7) == 0) flush(acc);
            }
            flush(acc);
        };

        int n = 0;
#pragma unroll
        for (int d = 0; d < 5; d++) {
            const int col = gi - 2 + d;
            if (col < 0 || col >= nx) continue;
            const int c0 = col * ny;
            int first = cs_lo[d];
            for (int jr = low_j + 1; first == -1 && jr < high_j; jr++) first = cs[c0 + jr];
            if (first == -1) continue;                                   // empty window
            int last = cs_hi[d] != -1 ? ce_hi[d] : -1;
            for (int jr = high_j - 2; last == -1 && jr >= low_j; jr--)
                if (cs[c0 + jr] != -1) last = ce[c0 + jr];
            if (last > N || first < 0) continue;                        // inconsistent cell table
            if (!pos.clip(d, first, last)) atomicOr(&flags[0], 4);
            const float cofs = (float)(gi - col) * hf;
            for (int j = first; j < last; j += 4) {
                // four scan records; up to three past `last` are read (both sources are padded)
                const float4 b0 = pos.scan(d, j), b1 = pos.scan(d, j + 1), b2 = pos.scan(d, j + 2),
                             b3 = pos.scan(d, j + 3);
                const float4 bb[4] = {b0, b1, b2, b3};
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const float xij = (I.fxi - bb[u].x) + cofs;
                    const float yij = (I.fyi - bb[u].y) + (I.rowi - bb[u].z) * hf;
                    const float rr2 = xij * xij + yij * yij;
                    const bool take = WIN ? rr2 < 1e30f : (rr2 < fourh2 && rr2 >= rr2min);
                    if (j + u < last && take) {
                        list[n * BS] = pos.encode(d, j + u);
                        n++;
                    }
                }
                // dense region: every active lane empties its list at once (a lane emptying its
                // list alone runs it with the rest of the wavefront idle)
                if (__any(n > CHUNK - 4)) {
                    run_list(n);
                    n = 0;
                }
            }
        }
        run_list(n);
    }
    const double mass = c.phys.mass;
    PairOut o;
    o.g = make_float4((float)d_v0, (float)d_v1, (float)d_v2, (float)d_v3);
    o.a = make_float4((float)(mass * d_ax), (float)(mass * d_ay), (float)(-c.corr.xspheps * mass * d_p0),
                      (float)(-c.corr.xspheps * mass * d_p1));
    o.T = (float)(2. * d_T);
    return o;
}

// Where a pair kernel's sums go: psum / psum_T, read by the update kernel (or the stage outputs), or,
// FUSE, straight into the particle's update (the update needs no other particle's sums)
template <bool FUSE, bool RIGID>
__device__ __forceinline__ void emit_pairs(const DevConsts &c, const UpdateArgs &u, float4 *__restrict__ psum,
                                           float *__restrict__ psum_T, int N, int i, const PairOut &o) {
    if constexpr (FUSE) {
        update_item<true, RIGID, true>(c, u, i, o.g, o.a, o.T);
    } else {
        psum[i] = o.g;
        psum[(size_t)N + i] = o.a;
        psum_T[i] = o.T;
    }
}

// in-support pair sums, one lane per sorted particle; when nonfinite[0] is set (a live particle
// with a non-finite input), the window form: every record of the 5x5 cells, the kernel zeroed
// outside the support, so the non-finite values spread as in the reference.  FUSE: each lane then
// runs its particle's update (u) instead of writing psum.
template <int BS, int CHUNK, bool FUSE = false, bool RIGID = false>
__global__ void __launch_bounds__(BS) k_pairs_fp32c(DevConsts c, const double *__restrict__ blanked,
                                                         const int *__restrict__ hash, int N, RecView rv,
                                                         const int *__restrict__ cs, const int *__restrict__ ce,
                                                         float4 *__restrict__ psum, float *__restrict__ psum_T,
                                                         const int *__restrict__ nonfinite, UpdateArgs u) {
    __shared__ int nb_list[CHUNK * BS];
    const int win = *nonfinite;
    const int i = xcd_logical_block() * BS + threadIdx.x;
    if (i >= N) return;
    const PairOut o =
        win != 0 ? pairs_fp32c_one<BS, CHUNK, true>(c, blanked, hash, N, rv, cs, ce, i, nb_list + threadIdx.x,
                                                    GlobalPos{rv.r0}, nullptr)
                 : pairs_fp32c_one<BS, CHUNK, false>(c, blanked, hash, N, rv, cs, ce, i, nb_list + threadIdx.x,
                                                     GlobalPos{rv.r0}, nullptr);
    emit_pairs<FUSE, RIGID>(c, u, psum, psum_T, N, i, o);
}

// LDS-staged form of k_pairs_fp32c, one wavefront per workgroup.  When the wavefront's 64
// particles lie in one cell column and the union of their 5x5-cell windows holds at most CAP
// records, the wavefront first copies the position records of that union -- five sorted ranges,
// one per neighbour column, rows [first row - 2, last row + 2] -- into LDS with coalesced loads;
// the candidate scan and the pair terms then read positions from LDS and only the three other
// records of each pair from L2/HBM.  Otherwise (a wavefront across two columns, a larger union, or
// a long run of empty cells) it runs the list kernel's path.  The per-lane loops, votes and FP32
// summation blocks are those of k_pairs_fp32c and both kernels group particles into wavefronts
// alike, so the two give the same sums bit for bit.  All 64 lanes take part in the staging (lanes
// past N copy records but sum nothing); the union comes from the hashes of the wavefront's first and
// last particles, read by every lane from memory, so no lane depends on another lane's registers.
#define SPHFEM_STAGE_ROW_SCAN 64
template <int CHUNK, int CAP, bool FUSE = false, bool RIGID = false>
__global__ void __launch_bounds__(64) k_pairs_fp32c_staged(DevConsts c, const double *__restrict__ blanked,
                                                           const int *__restrict__ hash, int N, RecView rv,
                                                           const int *__restrict__ cs, const int *__restrict__ ce,
                                                           float4 *__restrict__ psum, float *__restrict__ psum_T,
                                                           const int *__restrict__ nonfinite, int *__restrict__ flags,
                                                           UpdateArgs u) {
    // LDS: staged -- positions [CAP + 4] (+4: the scan reads up to three records past a range) and
    // 16-bit list entries; list path -- 32-bit list entries in the same bytes
    constexpr int POS_BYTES = (CAP + 4) * 16, LIST16 = CHUNK * 64 * 2, LIST32 = CHUNK * 64 * 4;
    constexpr int LDS_BYTES = POS_BYTES + LIST16 > LIST32 ? POS_BYTES + LIST16 : LIST32;
    __shared__ __align__(16) unsigned char lds[LDS_BYTES];
    __shared__ int sdelta[8];
    float4 *spos = reinterpret_cast<float4 *>(lds);
    unsigned short *list16 = reinterpret_cast<unsigned short *>(lds + POS_BYTES);
    int *list32 = reinterpret_cast<int *>(lds);
    const int win = *nonfinite;
    const int base = xcd_logical_block() * 64;
    if (base >= N) return;                              // the whole wavefront
    const int lane = (int)threadIdx.x;
    const int i = base + lane;
    if (win != 0) {                                     // the window form, as k_pairs_fp32c
        if (i < N)
            emit_pairs<FUSE, RIGID>(c, u, psum, psum_T, N, i,
                                    pairs_fp32c_one<64, CHUNK, true>(c, blanked, hash, N, rv, cs, ce, i,
                                                                     list32 + lane, GlobalPos{rv.r0}, nullptr));
        return;
    }
    const int last_i = base + 63 < N ? base + 63 : N - 1;
    const int nx = c.geom.nx, ny = c.geom.ny;
    const int h0 = hash[base], h1 = hash[last_i];
    const int col0 = h0 / ny;
    StagedPos sp;
    sp.spos = spos;
    sp.sdelta = sdelta;
    bool staged = h0 >= 0 && h1 >= h0 && h1 / ny == col0 && col0 < nx;
    int total = 0;
    if (staged) {
        const int r_lo = h0 - col0 * ny - 2 < 0 ? 0 : h0 - col0 * ny - 2;
        const int r_hi = h1 - col0 * ny + 2 > ny - 1 ? ny - 1 : h1 - col0 * ny + 2;
        // the end cells of the five row ranges, loaded together; runs of empty cells are scanned
        // only where an end cell is empty
        int lo_s[5], hi_s[5], hi_e[5];
#pragma unroll
        for (int d = 0; d < 5; d++) {
            const int col = col0 - 2 + d;
            const bool ok = col >= 0 && col < nx;
            const int c0 = (ok ? col : col0) * ny;
            lo_s[d] = ok ? cs[c0 + r_lo] : -1;
            hi_s[d] = ok ? cs[c0 + r_hi] : -1;
            hi_e[d] = ok ? ce[c0 + r_hi] : -1;
        }
#pragma unroll
        for (int d = 0; d < 5; d++) {
            const int col = col0 - 2 + d;
            int F = 0, L = 0;
            if (col >= 0 && col < nx) {
                const int c0 = col * ny;
                // first and last non-empty cell of the rows [r_lo, r_hi] (empty cells hold -1)
                int r = r_lo, f = lo_s[d];
                while (f == -1 && r < r_hi && r - r_lo < SPHFEM_STAGE_ROW_SCAN) f = cs[c0 + ++r];
                if (f != -1) {
                    int l = hi_s[d] != -1 ? hi_e[d] : -1, r2 = r_hi;
                    while (l == -1 && r2 > r && r_hi - r2 < SPHFEM_STAGE_ROW_SCAN)
                        if (cs[c0 + --r2] != -1) l = ce[c0 + r2];
                    if (l == -1) staged = false;                 // the downward scan gave up
                    F = f;
                    L = l;
                } else if (r < r_hi) {
                    staged = false;                              // the upward scan gave up
                }
                if (F < 0 || L > N || L < F) staged = false;     // inconsistent cell table
            }
            sp.F[d] = F;
            sp.L[d] = L;
            sp.delta[d] = F - total;
            total += L - F;
        }
        if (total > CAP) staged = false;
    }
    if (staged) {
#pragma unroll
        for (int d = 0; d < 5; d++) {
            if (lane == d) sdelta[d] = sp.delta[d];
            const int n = sp.L[d] - sp.F[d], off = sp.F[d] - sp.delta[d];
            for (int k = lane; k < n; k += 64) spos[off + k] = rv.r0[sp.F[d] + k];
        }
        __syncthreads();
        if (i < N)
            emit_pairs<FUSE, RIGID>(c, u, psum, psum_T, N, i,
                                    pairs_fp32c_one<64, CHUNK, false, StagedPos>(c, blanked, hash, N, rv, cs, ce, i,
                                                                                 list16 + lane, sp, flags));
    } else if (i < N) {
        emit_pairs<FUSE, RIGID>(c, u, psum, psum_T, N, i,
                                pairs_fp32c_one<64, CHUNK, false>(c, blanked, hash, N, rv, cs, ce, i, list32 + lane,
                                                                  GlobalPos{rv.r0}, nullptr));
    }
}


// interactions_gpu.cu:143-379 as one kernel (drop-in entry point): every particle of the 5x5 cells,
// as the reference visits them
__global__ void __launch_bounds__(256) k_interactions_ref(DevConsts c, sphfem_particles P, const int *__restrict__ cs,
                                                          const int *__restrict__ ce) {
    __shared__ int nb_list[48 * 256];
    int pidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (pidx >= (int)P.N) return;
    if (P.blanked[pidx] == 1.) return;
    double thermal_alpha = interaction_alpha(c, P.tool_particle[pidx]);
    PairSums r = pair_sums_fp64<false, 256, 48, true>(c, P, pidx, thermal_alpha, cs, ce, nullptr, nullptr,
                                                nb_list + threadIdx.x);
    P.S_der[pidx] = r.S_der;
    P.v_der[pidx] = r.v_der;
    P.pos_t[pidx] = r.pos_t;
    P.vel_t[pidx] = r.vel_t;
    if (thermal_alpha != 0.) P.T_t[pidx] = thermal_alpha * r.T_lapl;
}

// interactions_gpu.cu:84-140 with lapl_pse (kernels.cuh:67-84): particle strength exchange heat
// conduction over every particle of the 5x5 cells (no support cut-off; blanked neighbours count),
// the column-run order being the reference's cell order
__global__ void __launch_bounds__(256) k_interactions_pse(DevConsts c, sphfem_particles P, const int *__restrict__ cs,
                                                          const int *__restrict__ ce) {
    const int pidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (pidx >= (int)P.N) return;
    if (P.blanked[pidx] == 1.) return;
    const double mass = c.phys.mass;
    const int nx = c.geom.nx, ny = c.geom.ny;
    const sphfem_double2 pi = P.pos[pidx];
    const double hi = P.h[pidx], Ti = P.T[pidx];
    // the host copies of interactions_gpu.cu:453 (the tool's is zero unless
    // interactions_setup_thermal_constants_tool was called)
    const double alpha = (P.tool_particle[pidx] == 1.) ? c.itool_alpha_pse : c.trml_wp.alpha;
    const int hashi = P.hash[pidx];
    const int gi = hashi / ny;
    const int gj = hashi - gi * ny;
    const int low_i = gi - 2 < 0 ? 0 : gi - 2;
    const int low_j = gj - 2 < 0 ? 0 : gj - 2;
    const int high_i = gi + 3 > nx ? nx : gi + 3;
    const int high_j = gj + 3 > ny ? ny : gj + 3;
    double T_ti = 0.;
    for (int ii = low_i; ii < high_i; ii++) {
        int first, last;
        column_range(cs, ce, ii * ny, low_j, high_j, first, last);
        if (first < 0 || last > (int)P.N) continue;
        for (int iter = first; iter < last; iter++) {
            const sphfem_double2 pj = P.pos[iter];
            const double Tj = P.T[iter], rhoj = P.rho[iter];
            const double xij = pi.x - pj.x, yij = pi.y - pj.y;
            const double xx = sqrt(xij * xij + yij * yij);
            const double h2 = hi * hi;
            const double h4 = h2 * h2;
            const double w2 = +4. / (h4 * 3.14159265358979323846) * exp(-xx * xx / (h2));
            T_ti += (Tj - Ti) * w2 * mass / rhoj;
        }
    }
    P.T_t[pidx] = alpha * T_ti;
}

void launch_interactions_pse(sphfem_ctx *ctx, const sphfem_particles &p, const int *cs, const int *ce) {
    KTimer t(ctx, "interactions_pse");
    k_interactions_pse<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(ctx->c, p, cs, ce);
}

void launch_interactions_ref(sphfem_ctx *ctx, const sphfem_particles &p, const int *cs, const int *ce) {
    KTimer t(ctx, "interactions_ref");
    k_interactions_ref<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(ctx->c, p, cs, ce);
}

// ---------------------------------------------------------------------------------------------
// per-reference-function kernels (actions_gpu.cu)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_action(DevConsts c, sphfem_particles P, int which) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)P.N) return;
    switch (which) {
    case ACT_EOS:   // :222-232
        if (P.tool_particle[i] != 0.) return;
        P.p[i] = eos_pressure(c, P.rho[i]);
        break;
    case ACT_ARTSTRESS:   // :234-278
        if (P.tool_particle[i] != 0.) return;
        P.R[i] = artificial_stress(c, P.rho[i], P.p[i], P.S[i]);
        break;
    case ACT_JAUMANN:   // :280-314
        if (P.tool_particle[i] == 1.) return;
        P.S_t[i] = jaumann_rate(c, P.v_der[i], P.S[i]);
        break;
    case ACT_CONTINUITY:   // :626-639
        if (P.tool_particle[i] != 0.) return;
        P.rho_t[i] = -P.rho[i] * (P.v_der[i].x + P.v_der[i].w);
        break;
    case ACT_MOMENTUM: {   // :641-671
        if (P.tool_particle[i] != 0.) return;
        const double mass = c.phys.mass;
        sphfem_double4 sd = P.S_der[i];
        sphfem_double2 fc = P.fc[i], ft = P.ft[i], vt = P.vel_t[i];
        vt.x += sd.x + sd.y + fc.x / mass + ft.x / mass;
        vt.y += sd.z + sd.w + fc.y / mass + ft.y / mass;
        P.vel_t[i] = vt;
        break;
    }
    case ACT_ADVECTION: {   // :673-687
        if (P.tool_particle[i] != 0.) return;
        sphfem_double2 pt = P.pos_t[i], v = P.vel[i];
        pt.x = pt.x + v.x;
        pt.y = pt.y + v.y;
        P.pos_t[i] = pt;
        break;
    }
    case ACT_PLASTICITY: {   // :689-754
        if (P.tool_particle[i] != 0.) return;
        if (P.blanked[i] == 1.) return;
        sphfem_double4 S = P.S[i];
        double eps = P.eps_pl[i], epsd = P.eps_pl_dot[i], T = P.T[i];
        double S_in[4] = {S.x, S.y, S.z, S.w};
        bool melt = plasticity_jc(c, S, P.p[i], eps, epsd, T, P.rho[i]);
        P.eps_pl_dot[i] = epsd;
        if (S.x != S_in[0] || S.y != S_in[1] || S.z != S_in[2] || S.w != S_in[3] || eps != P.eps_pl[i]) {
            P.eps_pl[i] = eps;
            P.S[i] = S;
        }
        P.T[i] = T;
        if (melt) P.blanked[i] = 1.;
        break;
    }
    case ACT_BC:   // :765-781
        if (P.tool_particle[i] != 0.) return;
        if (P.fixed[i] != 0.) {
            sphfem_double2 z = {0., 0.};
            P.vel[i] = z;
            P.fc[i] = z;
            P.pos_t[i] = z;
            P.vel_t[i] = z;
        }
        break;
    case ACT_BC_THERMAL:   // :756-763
        if (P.fixed[i] == 1.) P.T[i] = c.trml_wp.T_init;
        break;
    case ACT_INVALIDATE: {   // :787-824
        if (P.blanked[i] == 1.) return;
        sphfem_double2 pt = P.pos_t[i], vt = P.vel_t[i];
        sphfem_double4 st = P.S_t[i];
        bool bad = isnaninf(pt.x) || isnaninf(pt.y) || isnaninf(vt.x) || isnaninf(vt.y) || isnaninf(st.x) ||
                   isnaninf(st.y) || isnaninf(st.z) || isnaninf(st.w) || isnaninf(P.rho_t[i]) || isnaninf(P.T_t[i]);
        if (bad) P.blanked[i] = 1.;
        break;
    }
    }
}

void launch_actions(sphfem_ctx *ctx, int which, const sphfem_particles &p) {
    static const char *names[] = {"eos", "artificial_stress", "jaumann", "continuity", "momentum", "advection",
                                  "plasticity", "bc", "bc_thermal", "invalidate"};
    KTimer t(ctx, names[which]);
    k_action<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(ctx->c, p, which);
}

// leap_frog.cu:25-96
__global__ void __launch_bounds__(256) k_leapfrog(sphfem_particles P, int which, double dt,
                                                  sphfem_double2 *pos_init, sphfem_double2 *vel_init,
                                                  sphfem_double4 *S_init, double *rho_init, double *T_init) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)P.N) return;
    if (P.blanked[i] == 1.) return;
    if (which == LF_INIT) {
        pos_init[i] = P.pos[i];
        vel_init[i] = P.vel[i];
        S_init[i] = P.S[i];
        rho_init[i] = P.rho[i];
        T_init[i] = P.T[i];
        return;
    }
    double f = which == LF_PREDICT ? 0.5 * dt : dt;
    sphfem_double2 pt = P.pos_t[i], vt = P.vel_t[i], p0 = pos_init[i], v0 = vel_init[i];
    sphfem_double4 St = P.S_t[i], S0 = S_init[i];
    // predict: x0 + 0.5*dt*x_t, correct: x0 + dt*x_t (same products as the reference)
    P.pos[i] = {p0.x + f * pt.x, p0.y + f * pt.y};
    P.vel[i] = {v0.x + f * vt.x, v0.y + f * vt.y};
    P.S[i] = {S0.x + f * St.x, S0.y + f * St.y, S0.z + f * St.z, S0.w + f * St.w};
    P.rho[i] = rho_init[i] + f * P.rho_t[i];
    P.T[i] = T_init[i] + f * P.T_t[i];
}

void launch_leapfrog(sphfem_ctx *ctx, int which, const sphfem_particles &p, sphfem_leap_frog *lf, double dt) {
    static const char *names[] = {"lf_init", "lf_predict", "lf_correct"};
    KTimer t(ctx, names[which]);
    k_leapfrog<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(p, which, dt, lf->pos_init, lf->vel_init, lf->S_init,
                                                        lf->rho_init, lf->T_init);
}

// ---------------------------------------------------------------------------------------------
// fused step kernel (FP64 form).  B: sorted predicted state (read), A: corrected state (write).
// Fields: A.pos/vel/S/rho/T/eps_pl/eps_pl_dot/blanked are outputs; B.fc/ft/n/on_seg are updated
// in place (own particle only); derivatives go to A.pos_t/vel_t/rho_t/S_t/T_t.
// ---------------------------------------------------------------------------------------------

// B = sorted predicted state (the grid buffer); A = the caller's arrays, whose unsorted content
// is dead after the reorder except the previous derivatives.  Outputs: corrected state into
// A.pos/vel/S/rho/T/eps_pl/eps_pl_dot/blanked, derivatives into B.pos_t/vel_t/rho_t/S_t/T_t,
// contact fields in place in B.fc/ft/n/on_seg; S_der/v_der into A when requested.
// FP64 pair sums of the sorted predicted state (interactions_gpu.cu:143-379) into psum64, for the
// update kernel; a kernel of its own so that the pair loop keeps 3 waves per SIMD instead of the
// update's 2
// nonfinite[0] is set by the reorder when a live particle carries a non-finite pair input: then
// every particle is evaluated with the reference's full window.
template <int BS>
__global__ void __launch_bounds__(BS) k_pairs_fp64(DevConsts c, sphfem_particles B, const int *__restrict__ cs,
                                                   const int *__restrict__ ce, const double *__restrict__ pre_vol,
                                                   const sphfem_double4 *__restrict__ pre_s,
                                                   PairSums *__restrict__ psum64, const int *__restrict__ nonfinite) {
    constexpr int CHUNK = 48;
    __shared__ int nb_list[CHUNK * BS];
    const int N = (int)B.N;
    const int i = xcd_logical_block() * BS + (int)threadIdx.x;
    if (i >= N || B.blanked[i] == 1.) return;
    const double thermal_alpha = interaction_alpha(c, B.tool_particle[i]);
    // the window form when a live particle carries a non-finite input (see pair_sums_fp64)
    if (*nonfinite != 0)
        psum64[i] = pair_sums_fp64<true, BS, CHUNK, true>(c, B, i, thermal_alpha, cs, ce, pre_vol, pre_s,
                                                          nb_list + threadIdx.x);
    else
        psum64[i] = pair_sums_fp64<true, BS, CHUNK, false>(c, B, i, thermal_alpha, cs, ce, pre_vol, pre_s,
                                                           nb_list + threadIdx.x);
}

// RIGID: the rigid-tool step (leap_frog.cu:99-160): friction heat after the corrector
// (actions_gpu.cu:316-343), tool particles moved and contact against the moved tool
// (tool_gpu.cu:146-216) after plasticity, for blanked particles too as in the reference.
// the update as its own kernel, after k_pairs_fp64 or a k_pairs_fp32c that wrote psum
template <bool F32C, int BS, bool RIGID>
__global__ void __launch_bounds__(BS) k_fused_update(DevConsts c, UpdateArgs u, const float4 *__restrict__ psum,
                                                     const float *__restrict__ psum_T, int *__restrict__ nonfinite) {
    const int i = xcd_logical_block() * BS + threadIdx.x;
    if (i == 0) *nonfinite = 0;      // the pair kernels of this step have read it
    if (i >= (int)u.B.N) return;
    float4 g = make_float4(0.f, 0.f, 0.f, 0.f), a = g;
    float tl = 0.f;
    if constexpr (F32C) {
        g = psum[i];                 // sums of k_pairs_fp32c
        a = psum[(size_t)u.B.N + i];
        tl = psum_T[i];
    }
    update_item<F32C, RIGID>(c, u, i, g, a, tl);
}

// FP32C pair sums written into the reference's interaction outputs (interactions_gpu.cu:370-378):
// v_der; the stress-divergence plus viscous acceleration per axis as S_der = {ax, 0, ay, 0} with
// vel_t = 0 (the momentum equation only uses S_der.x + S_der.y + vel_t.x and S_der.z + S_der.w +
// vel_t.y); XSPH in pos_t; alpha times the Laplacian in T_t where the thermal block is set.
// Blanked particles are left untouched, as by the reference.
__global__ void __launch_bounds__(256) k_psum_expand(DevConsts c, sphfem_particles P, const float4 *__restrict__ psum,
                                                     const float *__restrict__ psum_T, int *__restrict__ nonfinite) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = (int)P.N;
    if (i == 0) *nonfinite = 0;
    if (i >= N) return;
    if (P.blanked[i] == 1.) return;
    const double alpha = interaction_alpha(c, P.tool_particle[i]);
    const float4 g = psum[i], a = psum[(size_t)N + i];
    P.v_der[i] = {g.x, g.y, g.z, g.w};
    P.S_der[i] = {a.x, 0., a.y, 0.};
    P.vel_t[i] = {0., 0.};
    P.pos_t[i] = {a.z, a.w};
    if (alpha != 0.) P.T_t[i] = alpha * (double)psum_T[i];
}

template <int BSZ, int CH, bool FUSE, bool RIGID>
static void launch_pairs_fp32c(sphfem_ctx *ctx, const sphfem_particles &B, sphfem_leap_frog *lf, const int *cs,
                               const int *ce, const UpdateArgs &u) {
    const size_t n = B.N;
    RecView rv{lf->rec, lf->rec + n + 4, lf->rec + 2 * n + 4, lf->rec + 3 * n + 4};
    k_pairs_fp32c<BSZ, CH, FUSE, RIGID><<<nblk(B.N, BSZ), BSZ, 0, ctx->stream>>>(
        ctx->c, B.blanked, B.hash, (int)B.N, rv, cs, ce, lf->psum, lf->psum_T, ctx->d_flags + SPHFEM_FLAG_NONFINITE, u);
}

template <int CH, bool FUSE, bool RIGID>
static void launch_pairs_fp32c_staged(sphfem_ctx *ctx, const sphfem_particles &B, sphfem_leap_frog *lf, const int *cs,
                                      const int *ce, const UpdateArgs &u) {
    const size_t n = B.N;
    RecView rv{lf->rec, lf->rec + n + 4, lf->rec + 2 * n + 4, lf->rec + 3 * n + 4};
    k_pairs_fp32c_staged<CH, SPHFEM_STAGE_CAP, FUSE, RIGID><<<nblk(B.N, 64), 64, 0, ctx->stream>>>(
        ctx->c, B.blanked, B.hash, (int)B.N, rv, cs, ce, lf->psum, lf->psum_T, ctx->d_flags + SPHFEM_FLAG_NONFINITE,
        ctx->d_flags, u);
}

template <bool FUSE, bool RIGID>
static void launch_pairs_fp32c_shape(sphfem_ctx *ctx, const sphfem_particles &B, sphfem_leap_frog *lf, const int *cs,
                                     const int *ce, const UpdateArgs &u) {
    const int pb = lf->tuning.pair_block, pc = lf->tuning.pair_chunk;
    const int k = lf->tuning.pair_kernel;
    if (k == SPHFEM_PAIRS_STAGED || k == SPHFEM_PAIRS_STAGED_FUSED) {
        if (pc == 16) launch_pairs_fp32c_staged<16, FUSE, RIGID>(ctx, B, lf, cs, ce, u);
        else if (pc == 48) launch_pairs_fp32c_staged<48, FUSE, RIGID>(ctx, B, lf, cs, ce, u);
        else launch_pairs_fp32c_staged<32, FUSE, RIGID>(ctx, B, lf, cs, ce, u);
    } else if (pb == 128 && pc == 16) launch_pairs_fp32c<128, 16, FUSE, RIGID>(ctx, B, lf, cs, ce, u);
    else if (pb == 128 && pc == 32) launch_pairs_fp32c<128, 32, FUSE, RIGID>(ctx, B, lf, cs, ce, u);
    else if (pb == 128 && pc == 48) launch_pairs_fp32c<128, 48, FUSE, RIGID>(ctx, B, lf, cs, ce, u);
    else if (pb == 256 && pc == 16) launch_pairs_fp32c<256, 16, FUSE, RIGID>(ctx, B, lf, cs, ce, u);
    else if (pb == 256 && pc == 48) launch_pairs_fp32c<256, 48, FUSE, RIGID>(ctx, B, lf, cs, ce, u);
    else launch_pairs_fp32c<256, 32, FUSE, RIGID>(ctx, B, lf, cs, ce, u);
}

// the tuned FP32C pair kernel; fuse: it also runs every particle's update, with these arguments
// (the rigid-tool update when fuse->rt is set)
static void launch_pairs_fp32c_tuned(sphfem_ctx *ctx, const sphfem_particles &B, sphfem_leap_frog *lf, const int *cs,
                                     const int *ce, const UpdateArgs *fuse = nullptr) {
    if (fuse) {
        KTimer tm(ctx, "pairs_update_fp32c");
        if (fuse->rt) launch_pairs_fp32c_shape<true, true>(ctx, B, lf, cs, ce, *fuse);
        else launch_pairs_fp32c_shape<true, false>(ctx, B, lf, cs, ce, *fuse);
    } else {
        KTimer tm(ctx, "pairs_fp32c");
        launch_pairs_fp32c_shape<false, false>(ctx, B, lf, cs, ce, UpdateArgs{});
    }
}

void launch_interactions_fp32c(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_leap_frog *lf, const int *cs,
                               const int *ce, int num_cell) {
    launch_records_state(ctx, p, num_cell, lf->rec);
    launch_pairs_fp32c_tuned(ctx, p, lf, cs, ce);
    KTimer t(ctx, "psum_expand");
    k_psum_expand<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(ctx->c, p, lf->psum, lf->psum_T,
                                                           ctx->d_flags + SPHFEM_FLAG_NONFINITE);
}

void launch_fused_update(sphfem_ctx *ctx, const sphfem_particles &B, const sphfem_particles &A, sphfem_leap_frog *lf,
                         const int *cs, const int *ce, sphfem_tool *t, sphfem_rigid_tool *rt,
                         const RateView *rates_old, const RateView *rates_new, bool skip_contact) {
    const RateView none{};
    const RateView r_old = rates_old ? *rates_old : none, r_new = rates_new ? *rates_new : none;
    const int skip = skip_contact ? 1 : 0;
    ToolView tv{};
    if (t) {
        tv.T_nodes = t->T;
        tv.marks = t->marks;
        tv.fric_heat = t->fric_heat;
        tv.physical_para = t->physical_para;
        tv.sliding_force = t->sliding_force;
        tv.kin = t->kin;
        tv.near_list = lf->near_list;
        tv.near_count = ctx->d_flags + 1;
        tv.active = 1;
        tv.sink = SegSink{ctx->d_flags + SPHFEM_FLAG_SINK_FRIC, t->sink_keys, t->sink_vals, SPHFEM_SEG_SINK_CAP};
    }
    const RigidTool *rtd = rt ? rt->dev : nullptr;
    UpdateArgs u{};
    u.B = B;
    u.A = A;
    u.pos_init = lf->pos_init;
    u.vel_init = lf->vel_init;
    u.S_init = lf->S_init;
    u.rho_init = lf->rho_init;
    u.T_init = lf->T_init;
    u.tv = tv;
    u.write_intermediates = lf->intermediates ? 1 : 0;
    u.rt = rtd;
    u.r_old = r_old;
    u.r_new = r_new;
    int *nf = ctx->d_flags + SPHFEM_FLAG_NONFINITE;
    if (lf->precision == SPHFEM_FP32C) {
        u.skip_contact = skip;
        // the fused pair kernels run the update themselves
        const int pk = lf->tuning.pair_kernel;
        if (pk == SPHFEM_PAIRS_LIST_FUSED || pk == SPHFEM_PAIRS_STAGED_FUSED) {
            launch_pairs_fp32c_tuned(ctx, B, lf, cs, ce, &u);
            if (t) launch_fric_heat_sums(ctx, t);
            return;
        }
        launch_pairs_fp32c_tuned(ctx, B, lf, cs, ce);
        KTimer tm(ctx, "update_fp32c");
#define SPHFEM_UPDATE(BSZ, RG)                                                                              \
    k_fused_update<true, BSZ, RG><<<nblk(B.N, BSZ), BSZ, 0, ctx->stream>>>(ctx->c, u, lf->psum, lf->psum_T, nf)
        if (lf->tuning.update_block == 128) {
            if (rtd) SPHFEM_UPDATE(128, true);
            else SPHFEM_UPDATE(128, false);
        } else {
            if (rtd) SPHFEM_UPDATE(256, true);
            else SPHFEM_UPDATE(256, false);
        }
#undef SPHFEM_UPDATE
    } else {
        u.psum64 = lf->psum64;
        {
            KTimer tp(ctx, "pairs_fp64");
            if (lf->tuning.pair_block == 128)
                k_pairs_fp64<128><<<nblk(B.N, 128), 128, 0, ctx->stream>>>(ctx->c, B, cs, ce, lf->pre_vol, lf->pre_s,
                                                                            lf->psum64, nf);
            else
                k_pairs_fp64<256><<<nblk(B.N, 256), 256, 0, ctx->stream>>>(ctx->c, B, cs, ce, lf->pre_vol, lf->pre_s,
                                                                            lf->psum64, nf);
        }
        KTimer tm(ctx, "fused_update");
#define SPHFEM_UPDATE(BSZ, RG)                                                                              \
    k_fused_update<false, BSZ, RG><<<nblk(B.N, BSZ), BSZ, 0, ctx->stream>>>(ctx->c, u, nullptr, nullptr, nf)
        if (lf->tuning.update_block == 128) {
            if (rtd) SPHFEM_UPDATE(128, true);
            else SPHFEM_UPDATE(128, false);
        } else {
            if (rtd) SPHFEM_UPDATE(256, true);
            else SPHFEM_UPDATE(256, false);
        }
#undef SPHFEM_UPDATE
    }
    if (t) launch_fric_heat_sums(ctx, t);     // the friction-heat segment sums, in particle order
}

// FP32C derivative carry <-> the particle arrays' FP64 rates, and its gather by a permutation
__global__ void __launch_bounds__(256) k_rates_import(sphfem_particles P, RateView d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)P.N) return;
    const sphfem_double2 pt = P.pos_t[i], vt = P.vel_t[i];
    const sphfem_double4 St = P.S_t[i];
    d.pos_t[i] = make_float2((float)pt.x, (float)pt.y);
    d.vel_t[i] = make_float2((float)vt.x, (float)vt.y);
    d.S_t[i] = make_float4((float)St.x, (float)St.y, (float)St.z, (float)St.w);
    d.rho_t[i] = (float)P.rho_t[i];
    d.T_t[i] = (float)P.T_t[i];
}

__global__ void __launch_bounds__(256) k_rates_export(RateView r, sphfem_particles P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)P.N) return;
    const float2 pt = r.pos_t[i], vt = r.vel_t[i];
    const float4 St = r.S_t[i];
    P.pos_t[i] = {pt.x, pt.y};
    P.vel_t[i] = {vt.x, vt.y};
    P.S_t[i] = {St.x, St.y, St.z, St.w};
    P.rho_t[i] = r.rho_t[i];
    P.T_t[i] = r.T_t[i];
}

__global__ void __launch_bounds__(256) k_rates_gather(RateView a, RateView b, const int *__restrict__ perm, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = perm[i];
    if ((unsigned)s >= (unsigned)n) return;
    b.pos_t[i] = a.pos_t[s];
    b.vel_t[i] = a.vel_t[s];
    b.S_t[i] = a.S_t[s];
    b.rho_t[i] = a.rho_t[s];
    b.T_t[i] = a.T_t[s];
}

void launch_rates_import(sphfem_ctx *ctx, const sphfem_particles &p, const RateView &dst) {
    KTimer t(ctx, "rates_import");
    k_rates_import<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(p, dst);
}

void launch_rates_export(sphfem_ctx *ctx, const RateView &src, const sphfem_particles &p) {
    KTimer t(ctx, "rates_export");
    k_rates_export<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(src, p);
}

void launch_rates_gather(sphfem_ctx *ctx, const RateView &src, const RateView &dst, const int *perm, int n) {
    KTimer t(ctx, "rates_gather");
    k_rates_gather<<<nblk(n, 256), 256, 0, ctx->stream>>>(src, dst, perm, n);
}

// ---------------------------------------------------------------------------------------------
// output.cu:22-30: sum of fc and ft over all particles (deterministic two-level reduction)
// ---------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_force_partial(sphfem_particles P, double *part) {
    __shared__ double sx[256], sy[256];
    double x = 0., y = 0.;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < (int)P.N; i += gridDim.x * 256) {
        sphfem_double2 a = P.fc[i], b = P.ft[i];
        x += a.x + b.x;
        y += a.y + b.y;
    }
    sx[threadIdx.x] = x;
    sy[threadIdx.x] = y;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sx[threadIdx.x] += sx[threadIdx.x + s];
            sy[threadIdx.x] += sy[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sx[0];
        part[2 * blockIdx.x + 1] = sy[0];
    }
}

__global__ void k_force_final(const double *part, int n, double *out) {
    if (threadIdx.x != 0) return;
    double x = 0., y = 0.;
    for (int b = 0; b < n; b++) {
        x += part[2 * b];
        y += part[2 * b + 1];
    }
    out[0] = x;
    out[1] = y;
}

void launch_report_force(sphfem_ctx *ctx, const sphfem_particles &p, double *part, double *out) {
    KTimer t(ctx, "report_force");
    k_force_partial<<<SPHFEM_FORCE_BLOCKS, 256, 0, ctx->stream>>>(p, part);
    k_force_final<<<1, 64, 0, ctx->stream>>>(part, SPHFEM_FORCE_BLOCKS, out);
}

}  // namespace sphfem
