// Rigid FEM tool per step: transient heat conduction (FEM_thermal_gpu.cu:342-468, 865-921),
// tool / segment motion (tool_FEM.cu:109-170, 1546-1609), Usui nodal-shift wear
// (FEM_thermal_gpu.cu:476-577, tool_FEM.cu:299-432, 520-587) and the particle-tool contact
// (actions_gpu.cu:345-617).  Where the reference scatters with atomics in colour-ordered
// launches, these kernels gather per destination from lists built once at tool creation, in
// the same (colour, element/segment) order, so results are deterministic.
#include "sphfem_physics.h"

#pragma clang fp contract(off)

namespace sphfem {

// ---------------------------------------------------------------------------------------------
// FEM heat conduction
// ---------------------------------------------------------------------------------------------

// Robin contribution of segment s to the rhs and to the stiffness diagonal (FEM_thermal_gpu.cu:342-384)
__device__ __forceinline__ bool robin_term(const DevConsts &c, const sphfem_tool &t, int s, double &df, double &dk) {
    sphfem_double2 he = t.heat_exchange[s];
    if (he.y == 0) return false;
    double T_cont_particle = he.x / he.y;
    double fric_eng = t.fric_heat[s];
    sphfem_double2 l = t.left[s], r = t.right[s];
    double ddx = l.x - r.x, ddy = l.y - r.y;
    double dist = sqrtf_d(ddx * ddx + ddy * ddy);
    df = 0.5 * c.tc.h * dist * T_cont_particle + 0.5 * fric_eng;
    dk = c.tc.h * (0.5 * dist);
    return true;
}

__global__ void __launch_bounds__(256) k_tool_begin(sphfem_tool t, double half_dt, int *counters) {
    tool_begin_item(t, half_dt, counters, blockIdx.x * blockDim.x + threadIdx.x);
}

void launch_tool_begin(sphfem_ctx *ctx, sphfem_tool *t) {
    int n = std::max(std::max(t->n_nodes, t->n_seg), 1);
    KTimer tm(ctx, "tool_begin");
    k_tool_begin<<<nblk(n, 256), 256, 0, ctx->stream>>>(*t, 0.5 * ctx->c.dt, ctx->d_flags);
}

// k = M^-1 (f + robin - (K + robin_diag) T), one row per thread (FEM_thermal_gpu.cu:443-458 with
// the Robin terms of :342-384 applied to the copies f_temp / K_temp of reset_FEM_server)
__global__ void __launch_bounds__(256) k_fem_rate_fused(DevConsts c, sphfem_tool t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.n_nodes) return;
    double f = t.f[i];
    double kd = t.K[t.diag[i]];
    int2 rs = t.node_robin[i];
    double df, dk;
    if (rs.x >= 0 && robin_term(c, t, rs.x, df, dk)) { f += df; kd += dk; }
    if (rs.y >= 0 && robin_term(c, t, rs.y, df, dk)) { f += df; kd += dk; }
    double buffer = 0.;
    const int dg = t.diag[i];
    for (int k = t.row_ptr[i]; k < t.row_ptr[i + 1]; k++) {
        double kv = (k == dg) ? kd : t.K[k];
        buffer += kv * t.T[t.col_ind[k]];
    }
    t.k[i] = t.Minv[i] * (f - buffer);
}

// calculate_temp (FEM_thermal_gpu.cu:460-468)
__global__ void __launch_bounds__(256) k_fem_temp(sphfem_tool t, double dt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.n_nodes) return;
    t.T[i] = dt * t.k[i] + t.T_old[i];
}

void launch_tool_fem_end(sphfem_ctx *ctx, sphfem_tool *t) {
    {
        KTimer tm(ctx, "fem_rate");
        k_fem_rate_fused<<<nblk(t->n_nodes, 256), 256, 0, ctx->stream>>>(ctx->c, *t);
    }
    KTimer tm(ctx, "fem_temp");
    k_fem_temp<<<nblk(t->n_nodes, 256), 256, 0, ctx->stream>>>(*t, ctx->c.dt);
}

// drop-in pieces: reset_FEM_server, apply_robin_bdc, calculate_temp_rate, calculate_temp
__global__ void k_fem_reset(sphfem_tool t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < t.n_nodes) {
        t.T_old[i] = t.T[i];
        t.f_temp[i] = t.f[i];
    }
    if (i < t.nnz) t.K_temp[i] = t.K[i];
}

__global__ void k_fem_robin(DevConsts c, sphfem_tool t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.n_nodes) return;
    int2 rs = t.node_robin[i];
    double df, dk;
    if (rs.x >= 0 && robin_term(c, t, rs.x, df, dk)) { t.f_temp[i] += df; t.K_temp[t.diag[i]] += dk; }
    if (rs.y >= 0 && robin_term(c, t, rs.y, df, dk)) { t.f_temp[i] += df; t.K_temp[t.diag[i]] += dk; }
}

__global__ void k_fem_rate(sphfem_tool t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.n_nodes) return;
    double buffer = 0.;
    for (int k = t.row_ptr[i]; k < t.row_ptr[i + 1]; k++) buffer += t.K_temp[k] * t.T[t.col_ind[k]];
    t.k[i] = t.Minv[i] * (t.f_temp[i] - buffer);
}

void launch_fem_reset(sphfem_ctx *ctx, sphfem_tool *t) {
    KTimer tm(ctx, "fem_reset");
    k_fem_reset<<<nblk(std::max(t->n_nodes, t->nnz), 256), 256, 0, ctx->stream>>>(*t);
}
void launch_fem_robin(sphfem_ctx *ctx, sphfem_tool *t) {
    KTimer tm(ctx, "fem_robin");
    k_fem_robin<<<nblk(t->n_nodes, 256), 256, 0, ctx->stream>>>(ctx->c, *t);
}
void launch_fem_rate(sphfem_ctx *ctx, sphfem_tool *t) {
    KTimer tm(ctx, "fem_rate_ref");
    k_fem_rate<<<nblk(t->n_nodes, 256), 256, 0, ctx->stream>>>(*t);
}
void launch_fem_temp(sphfem_ctx *ctx, sphfem_tool *t, double dt) {
    KTimer tm(ctx, "fem_temp");
    k_fem_temp<<<nblk(t->n_nodes, 256), 256, 0, ctx->stream>>>(*t, dt);
}

__global__ void k_wear_reset(sphfem_tool t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.n_seg) return;
    t.wear_rate[i] = {0., 0.};
    t.wear_nodes[i] = 0.;
    t.physical_para[i] = {0., 0., 0., 0.};
    t.sliding_force[i] = 0.;
}
void launch_wear_reset(sphfem_ctx *ctx, sphfem_tool *t) {
    KTimer tm(ctx, "wear_reset");
    k_wear_reset<<<nblk(std::max(t->n_seg, 1), 256), 256, 0, ctx->stream>>>(*t);
}

// ---------------------------------------------------------------------------------------------
// tool motion: segments_update (tool_FEM.cu:119-139) and tool_update_FEM_gpu (:109-117, :1552-1570)
// ---------------------------------------------------------------------------------------------
// tool_FEM.cu:1552-1570 for segment and node i with the step's kinematics k (read by every item
// before item 0 rewrites them; the motion uses vel and flank_retreat, which the rewrite keeps)
__device__ __forceinline__ void tool_move_item(const sphfem_tool &t, const ToolKin &k, double dt, int do_segments,
                                               int do_nodes, int i) {
    if (do_segments && i < t.n_seg) {
        sphfem_double2 l = t.left[i], r = t.right[i];
        t.left[i] = {l.x + k.vel_x * dt, l.y + k.vel_y * dt + k.flank_retreat};
        t.right[i] = {r.x + k.vel_x * dt, r.y + k.vel_y * dt + k.flank_retreat};
        t.heat_exchange[i] = {0., 0.};
        t.fric_heat[i] = 0.;
    }
    if (do_nodes && i < t.n_nodes) {
        sphfem_double2 p = t.pos[i];
        p.x += k.vel_x * dt;
        p.y += k.vel_y * dt + k.flank_retreat;
        t.pos[i] = p;
    }
}

__device__ __forceinline__ void tool_move_kinematics(const sphfem_tool &t, const ToolKin &k, double dt) {
    {
        ToolKin n = k;
        n.tl_x += k.vel_x * dt;
        n.tl_y += k.vel_y * dt;
        n.br_x += k.vel_x * dt;
        n.br_y += k.vel_y * dt;
        n.lowest_point += k.vel_y * dt;
        n.right_point += k.vel_x * dt;
        n.flank_contact = n.lowest_point;
        *t.kin = n;
    }
}

__global__ void __launch_bounds__(256) k_tool_move(sphfem_tool t, double dt, int do_segments, int do_nodes) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const ToolKin k = *t.kin;
    tool_move_item(t, k, dt, do_segments, do_nodes, i);
    if (do_nodes && i == 0) tool_move_kinematics(t, k, dt);
}

// k_fem_temp and k_tool_move in one launch (the step without wear: nothing between them reads the
// temperatures or moves the tool otherwise); the same per-item arithmetic
__global__ void __launch_bounds__(256) k_fem_temp_tool_move(sphfem_tool t, double dt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const ToolKin k = *t.kin;
    if (i < t.n_nodes) t.T[i] = dt * t.k[i] + t.T_old[i];
    tool_move_item(t, k, dt, 1, 1, i);
    if (i == 0) tool_move_kinematics(t, k, dt);
}

void launch_tool_fem_end_move(sphfem_ctx *ctx, sphfem_tool *t) {
    {
        KTimer tm(ctx, "fem_rate");
        k_fem_rate_fused<<<nblk(t->n_nodes, 256), 256, 0, ctx->stream>>>(ctx->c, *t);
    }
    KTimer tm(ctx, "fem_temp_tool_move");
    const int n = std::max(std::max(t->n_nodes, t->n_seg), 1);
    k_fem_temp_tool_move<<<nblk(n, 256), 256, 0, ctx->stream>>>(*t, ctx->c.dt);
}

void launch_tool_move(sphfem_ctx *ctx, sphfem_tool *t, double dt, bool segments, bool nodes) {
    int n = std::max(std::max(t->n_nodes, t->n_seg), 1);
    KTimer tm(ctx, "tool_move");
    k_tool_move<<<nblk(n, 256), 256, 0, ctx->stream>>>(*t, dt, segments ? 1 : 0, nodes ? 1 : 0);
}

// ---------------------------------------------------------------------------------------------
// wear (actions_gpu.cu:1009-1020)
// ---------------------------------------------------------------------------------------------

// element matrices of FEM_thermal_gpu.cu:71-109, 126-152 (Ke[a][b] = row a, column b)
__device__ __forceinline__ void element_matrices(const DevConsts &c, const sphfem_tool &t, int e, double *out) {
    sphfem_int4 el = t.elem[e];
    sphfem_double2 p0 = t.pos[el.x], p1 = t.pos[el.y], p2 = t.pos[el.z];
    double j00 = p1.x - p0.x, j01 = p1.y - p0.y;
    double j10 = p2.x - p0.x, j11 = p2.y - p0.y;
    double det = fabs(j00 * j11 - j10 * j01);
    double w = 1. / 3. * 0.5;
    out[0] = w * c.tc.cp * c.tc.rho * det;
    double dd = j00 * j11 - j01 * j10;
    double ood = 1. / dd;
    double i00 = j11 * ood, i01 = -j10 * ood, i10 = -j01 * ood, i11 = j00 * ood;
    const double gx[3] = {-1., 1., 0.}, gy[3] = {-1., 0., 1.};
    double G0[3], G1[3];
    for (int a = 0; a < 3; a++) {
        G0[a] = i00 * gx[a] + i10 * gy[a];
        G1[a] = i01 * gx[a] + i11 * gy[a];
    }
    double s = (w * 3.) * (c.tc.k * det);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) out[1 + 3 * a + b] = (s * G0[a]) * G0[b] + (s * G1[a]) * G1[b];
}

__global__ void k_layer_mats(DevConsts c, sphfem_tool t, double *dst) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= t.n_layer) return;
    double m[10];
    element_matrices(c, t, t.layer_elems[j], m);
    for (int q = 0; q < 10; q++) dst[(long)j * 10 + q] = m[q];
}

// Usui rake/flank wear depth of segment s (tool_FEM.cu:299-432 with USUI_RAKE_FLANK,
// RESTRICT_WEAR_TEMP); false if the segment is skipped
__device__ __forceinline__ bool segment_wear(const DevConsts &c, const sphfem_tool &t, int s, double &wl, double &wr) {
    sphfem_double4 wear = t.physical_para[s];
    if (wear.x == 0.) return false;
    sphfem_int4 mk = t.marks[s];
    double pressure = wear.y / wear.x;
    double vel = wear.z / wear.x;
    double tl = t.T[mk.z], tr = t.T[mk.w];
    if (vel > 0.008) return false;
    tl = fmin(tl, 1350.);
    tr = fmin(tr, 1350.);
    vel = fmin(vel, 0.008);
    pressure = fmin(pressure, 0.04);
    const double dt = c.dt;
    if (fabs(t.n[s].x) > 0.85) {
        wl = dt * c.wear.C * pressure * vel * exp(-c.wear.D / tl);
        wr = dt * c.wear.C * pressure * vel * exp(-c.wear.D / tr);
    } else {
        wl = dt * c.wear.A * pressure * vel * exp(-c.wear.B / tl);
        wr = dt * c.wear.A * pressure * vel * exp(-c.wear.B / tr);
    }
    return true;
}

// nodal shift: each node gathers the shifts of its segments in segment order
__global__ void k_nodal_shift(DevConsts c, sphfem_tool t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < t.n_seg) {
        double wl, wr;
        if (segment_wear(c, t, i, wl, wr)) t.wear_nodes[i] = (wl + wr) / c.dt;
    }
    if (i >= t.n_nodes) return;
    int a = t.node_shift_ptr[i], b = t.node_shift_ptr[i + 1];
    if (a == b) return;
    sphfem_double2 p = t.pos[i];
    for (int q = a; q < b; q++) {
        int2 e = t.node_shift[q];
        double wl, wr;
        if (!segment_wear(c, t, e.x, wl, wr)) continue;
        double w = e.y ? wr : wl;
        sphfem_double2 nn = t.n[e.x];
        p.x += nn.x * w;
        p.y += nn.y * w;
    }
    t.pos[i] = p;
}

// K, M: subtract the old and add the new boundary-layer element matrices in (colour, element)
// order (mesh_retreat_wear / mesh_resume_wear), then M^-1 (do_inverse_mass_matrix_lumped)
__global__ void k_wear_assemble(sphfem_tool t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < t.nnz) {
        int a = t.nnz_layer_ptr[i], b = t.nnz_layer_ptr[i + 1];
        if (a != b) {
            double v = t.K[i];
            for (int q = a; q < b; q++) {
                int code = t.nnz_layer[q];
                v += -1.0 * t.layer_old[(long)(code >> 4) * 10 + 1 + (code & 15)];
            }
            for (int q = a; q < b; q++) {
                int code = t.nnz_layer[q];
                v += 1.0 * t.layer_new[(long)(code >> 4) * 10 + 1 + (code & 15)];
            }
            t.K[i] = v;
        }
    }
    if (i < t.n_nodes) {
        int a = t.node_layer_ptr[i], b = t.node_layer_ptr[i + 1];
        double m = t.M[i];
        if (a != b) {
            for (int q = a; q < b; q++) m += -1.0 * t.layer_old[(long)(t.node_layer[q] >> 2) * 10];
            for (int q = a; q < b; q++) m += 1.0 * t.layer_new[(long)(t.node_layer[q] >> 2) * 10];
            t.M[i] = m;
        }
        t.Minv[i] = 1. / m;
    }
}

// segment_wear_update (tool_FEM.cu:520-548) and correct_flank_contact (:559-587) for any number of
// segments: one block, min over left.y
__global__ void __launch_bounds__(1024) k_segment_wear_update(sphfem_tool t) {
    __shared__ double red[1024];
    double mn = SPHFEM_FLT_MAX;
    for (int s = threadIdx.x; s < t.n_seg; s += blockDim.x) {
        sphfem_int4 mk = t.marks[s];
        sphfem_double2 lp = t.pos[mk.z], rp = t.pos[mk.w];
        t.left[s] = lp;
        t.right[s] = rp;
        double vx = rp.y - lp.y, vy = -(rp.x - lp.x);
        double inv = 1. / sqrt(vx * vx + vy * vy);
        t.n[s] = {vx * inv, vy * inv};
        mn = fmin(mn, lp.y);
    }
    red[threadIdx.x] = mn;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) t.kin->flank_retreat = t.kin->flank_contact - red[0];
}

// the wear step after the old boundary-layer matrices: node shift, new matrices, K/M update,
// segment geometry
static void launch_wear_rest(sphfem_ctx *ctx, sphfem_tool *t) {
    const int nl = std::max(t->n_layer, 1);
    k_nodal_shift<<<nblk(std::max(t->n_nodes, t->n_seg), 256), 256, 0, ctx->stream>>>(ctx->c, *t);
    k_layer_mats<<<nblk(nl, 128), 128, 0, ctx->stream>>>(ctx->c, *t, t->layer_new);
    k_wear_assemble<<<nblk(std::max(t->nnz, t->n_nodes), 256), 256, 0, ctx->stream>>>(*t);
    k_segment_wear_update<<<1, 1024, 0, ctx->stream>>>(*t);
}

void launch_wear(sphfem_ctx *ctx, sphfem_tool *t) {
    KTimer tm(ctx, "wear");
    const int nl = std::max(t->n_layer, 1);
    k_layer_mats<<<nblk(nl, 128), 128, 0, ctx->stream>>>(ctx->c, *t, t->layer_old);
    launch_wear_rest(ctx, t);
}

// k_fem_temp and the wear step's old boundary-layer matrices in one launch: the matrices read the
// node positions and constants only, the temperature update writes the temperatures only (the
// same per-item arithmetic as the two kernels)
__global__ void __launch_bounds__(256) k_fem_temp_layer_mats(DevConsts c, sphfem_tool t, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < t.n_nodes) t.T[i] = dt * t.k[i] + t.T_old[i];
    if (i < t.n_layer) {
        double m[10];
        element_matrices(c, t, t.layer_elems[i], m);
        for (int q = 0; q < 10; q++) t.layer_old[(long)i * 10 + q] = m[q];
    }
}

// the FEM tool's end of a step with wear: temperature rate, then the temperature update with the
// old layer matrices, then the rest of the wear step (launch_tool_fem_end + launch_wear, one launch
// fewer)
void launch_tool_fem_end_wear(sphfem_ctx *ctx, sphfem_tool *t) {
    {
        KTimer tm(ctx, "fem_rate");
        k_fem_rate_fused<<<nblk(t->n_nodes, 256), 256, 0, ctx->stream>>>(ctx->c, *t);
    }
    KTimer tm(ctx, "fem_temp_wear");
    const int n = std::max(std::max(t->n_nodes, t->n_layer), 1);
    k_fem_temp_layer_mats<<<nblk(n, 256), 256, 0, ctx->stream>>>(ctx->c, *t, ctx->c.dt);
    launch_wear_rest(ctx, t);
}

// ---------------------------------------------------------------------------------------------
// remeshing (actions_gpu.cu:1022-1038)
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ double cross2(sphfem_double2 a, sphfem_double2 b) { return a.x * b.y - a.y * b.x; }

// mapping_data_copy (tool_FEM.cu:1611-1622) fused with mesh_shift (:589-616): the interior nodes
// (neither cutting edge nor fixed boundary) move towards the worn flank by a profile of the
// reference points tl / br, scaled by freq times the flank retreat (at least 6.5e-6 inwards)
__global__ void __launch_bounds__(256) k_map_and_shift(sphfem_tool t, int freq) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.n_nodes) return;
    sphfem_double2 p = t.pos[i];
    t.pos_map[i] = p;
    t.T_map[i] = t.T[i];
    if (t.node_flags[i] & 3) return;
    const ToolKin k = *t.kin;
    double shift = k.flank_retreat * 1.0 * freq;
    shift = fmin(shift, -0.0000065);
    const double wy = (p.x - k.tl_x) / (k.br_x - k.tl_x);
    const double wx = fmin(1.0, (p.y - k.tl_y) / (k.br_y - k.tl_y));
    const double c = 0.94;
    if (wx <= c)
        p.x += 1.45 * shift * (pow(c, 0.7) - pow(fabs(wx - c), 0.55)) * wy * wy * wy;
    else if (wx < 1.1 && wx > c)
        p.x += 1.45 * shift * (pow(c, 0.7) - pow(fabs(wx - c), 0.55)) * wy * wy;
    if (wy < 1.01)
        p.y += -0.9 * shift * fmin(pow(wy, 6.0), 1.) * pow(wx, 6.0);
    else if (wy >= 0.99)
        p.y += -0.9 * shift * fmin(pow(wy, 6.0), 1.) * wx * wx * wx * wx;
    t.pos[i] = p;
}

// reset_FEM_matrix (FEM_thermal_gpu.cu:702-720)
__global__ void k_fem_zero(sphfem_tool t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < t.nnz) t.K[i] = 0.;
    if (i < t.n_nodes) {
        t.M[i] = 0.;
        t.f[i] = 0.;
    }
}

__device__ __forceinline__ int csr_entry(const sphfem_tool &t, int row, int col) {
    int k = t.row_ptr[row];
    const int e = t.row_ptr[row + 1];
    while (k < e && t.col_ind[k] != col) k++;
    return k;
}

// do_stiffness_mass_matrix_construction_p1q3 (FEM_thermal_gpu.cu:126-225) for one colour: elements
// of one colour share no node, so every entry takes at most one term per launch
__global__ void __launch_bounds__(256) k_fem_assemble(DevConsts c, sphfem_tool t, int colour) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= t.n_elem) return;
    const sphfem_int4 el = t.elem[e];
    if (el.w != colour) return;
    double m[10];
    element_matrices(c, t, e, m);
    const int nd[3] = {el.x, el.y, el.z};
    for (int a = 0; a < 3; a++) {
        t.M[nd[a]] += m[0];
        for (int b = 0; b < 3; b++) t.K[csr_entry(t, nd[a], nd[b])] += m[1 + 3 * a + b];
    }
}

// do_dirichlet_bdc_p1q3 (FEM_thermal_gpu.cu:234-324) applied row by row in (colour, element, node)
// order.  Two passes so no thread reads an entry another clears: the rhs of a free node x takes
// -K(F, x) T0 of every Dirichlet row F reaching it, in that order (the assembled K(F, x): a free
// node's column is never cleared); then the Dirichlet rows become identity rows and their
// columns zero.
__global__ void __launch_bounds__(256) k_dirichlet_rhs(DevConsts c, sphfem_tool t) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= t.n_nodes) return;
    const double T0 = c.tc.T0;
    if (t.dir_node[x]) {
        t.M[x] = 1.0;
        t.f[x] = T0;
        t.T_old[x] = T0;
        return;
    }
    double f = t.f[x];
    for (int q = t.dir_ptr[x]; q < t.dir_ptr[x + 1]; q++) f += -t.K[t.dir_entry[q]] * T0;
    t.f[x] = f;
}

__global__ void __launch_bounds__(256) k_dirichlet_rows(sphfem_tool t) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= t.n_nodes) return;
    const bool dx = t.dir_node[x] != 0;
    for (int k = t.row_ptr[x]; k < t.row_ptr[x + 1]; k++) {
        const int col = t.col_ind[k];
        if (dx) t.K[k] = col == x ? 1.0 : 0.;
        else if (t.dir_node[col]) t.K[k] = 0.;
    }
}

// inverse_mass_matrix (FEM_thermal_gpu.cu:838-847)
__global__ void k_fem_minv(sphfem_tool t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < t.n_nodes) t.Minv[i] = 1. / t.M[i];
}

__device__ __forceinline__ int bucket_of(const ElemBuckets &b, sphfem_double2 x) {
    int cx = (int)floor((x.x - b.x0) * b.inv), cy = (int)floor((x.y - b.y0) * b.inv);
    cx = cx < 0 ? 0 : (cx >= b.nx ? b.nx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= b.ny ? b.ny - 1 : cy);
    return cy * b.nx + cx;
}

// do_nodal_value_interpolation_old / value_interpolation_in_layer_element (tool_FEM.cu:57-100,
// 434-443): the first element of the mapping (in element order, counter-clockwise, weights
// unnormalised) that holds the node gives its temperature; none keeps the mapped value.  The
// candidates are the elements of the node's bin, in element order.
__global__ void __launch_bounds__(256) k_interp_layer(sphfem_tool t, ElemBuckets b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= t.n_nodes) return;
    const sphfem_double2 x = t.pos[i];
    double T = t.T_map[i];
    const int cell = bucket_of(b, x);
    for (int q = b.ptr[cell]; q < b.ptr[cell + 1]; q++) {
        const int e = b.list[q];
        const sphfem_int4 el = t.elem[e];
        const sphfem_double2 xA = t.pos_map[el.x];
        const sphfem_double2 pb = t.pos_map[el.y], pc = t.pos_map[el.z];
        const sphfem_double2 xB = {pb.x - xA.x, pb.y - xA.y}, xC = {pc.x - xA.x, pc.y - xA.y};
        const sphfem_double2 xP = {x.x - xA.x, x.y - xA.y}, xBC = {xB.x - xC.x, xB.y - xC.y};
        const double d = cross2(xB, xC);
        if (fabs(d) > 1e-13) {
            const double wA = cross2(xB, xC) + cross2(xP, xBC);
            const double wB = cross2(xP, xC);
            const double wC = -cross2(xP, xB);
            if (wA >= 0 && wB >= 0 && wC >= 0 && wA <= 1 && wB <= 1 && wC <= 1) {
                T = (t.T_map[el.x] * wA + t.T_map[el.y] * wB + t.T_map[el.z] * wC) / d;
                break;
            }
        }
    }
    t.T[i] = T;
}

void launch_remesh_low_level(sphfem_ctx *ctx, sphfem_tool *t, int freq, const ElemBuckets &b) {
    KTimer tm(ctx, "remesh");
    hipStream_t s = ctx->stream;
    const int nn = t->n_nodes;
    k_map_and_shift<<<nblk(nn, 256), 256, 0, s>>>(*t, freq);
    k_fem_zero<<<nblk(std::max(t->nnz, nn), 256), 256, 0, s>>>(*t);
    for (int c = 1; c <= t->n_colors; c++) k_fem_assemble<<<nblk(t->n_elem, 256), 256, 0, s>>>(ctx->c, *t, c);
    k_dirichlet_rhs<<<nblk(nn, 256), 256, 0, s>>>(ctx->c, *t);
    k_dirichlet_rows<<<nblk(nn, 256), 256, 0, s>>>(*t);
    k_fem_minv<<<nblk(nn, 256), 256, 0, s>>>(*t);
    k_interp_layer<<<nblk(nn, 256), 256, 0, s>>>(*t, b);
}

// do_nodal_value_interpolation (tool_FEM.cu:445-517): temperature of a new node from the first
// element of the current mesh that holds it (a coincident node within 1e-10 gives its value;
// barycentric weights with a 1e-8 tolerance, the result clamped to the element's range);
// T_new keeps its value where no element holds the node.  Candidates: the node's bin.
__global__ void __launch_bounds__(256) k_interp_new(sphfem_tool t, const sphfem_double2 *new_pos, int n_new,
                                                    double *T_new, ElemBuckets b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_new) return;
    const sphfem_double2 x = new_pos[i];
    auto near = [&](sphfem_double2 a) {
        const double dx = x.x - a.x, dy = x.y - a.y;
        return (double)sqrtf((float)(dx * dx + dy * dy)) < 1e-10;
    };
    const int cell = bucket_of(b, x);
    for (int q = b.ptr[cell]; q < b.ptr[cell + 1]; q++) {
        const int e = b.list[q];
        const sphfem_int4 el = t.elem[e];
        const sphfem_double2 xA = t.pos[el.x], xB = t.pos[el.y], xC = t.pos[el.z];
        if (near(xA)) { T_new[i] = t.T[el.x]; return; }
        if (near(xB)) { T_new[i] = t.T[el.y]; return; }
        if (near(xC)) { T_new[i] = t.T[el.z]; return; }
        const sphfem_double2 xAC = {xA.x - xC.x, xA.y - xC.y}, xBA = {xB.x - xA.x, xB.y - xA.y};
        const sphfem_double2 xCB = {xC.x - xB.x, xC.y - xB.y};
        const sphfem_double2 xPC = {x.x - xC.x, x.y - xC.y}, xPA = {x.x - xA.x, x.y - xA.y};
        const sphfem_double2 xPB = {x.x - xB.x, x.y - xB.y};
        const double d = cross2(xBA, xAC);
        if (fabs(d) > 0) {
            const double wC = cross2(xBA, xPA) / d;
            const double wA = cross2(xCB, xPB) / d;
            const double wB = cross2(xAC, xPC) / d;
            if ((wA >= -1e-8 && wB >= -1e-8 && wC >= -1e-8) || (wA <= 1e-8 && wB <= 1e-8 && wC <= 1e-8)) {
                const double TA = t.T[el.x], TB = t.T[el.y], TC = t.T[el.z];
                double Tn = fabs(TA * wA + TB * wB + TC * wC);
                const double lo = fmin(fmin(TA, TB), TC);
                if (Tn < lo) Tn = lo;
                const double hi = fmax(fmax(TA, TB), TC);
                if (Tn > hi) Tn = hi;
                T_new[i] = Tn;
                return;
            }
        }
    }
}

void launch_interp_new(sphfem_ctx *ctx, const sphfem_tool *old_t, const sphfem_double2 *new_pos, int n_new,
                       double *T_new, const ElemBuckets &b) {
    KTimer tm(ctx, "remesh_interp");
    k_interp_new<<<nblk(std::max(n_new, 1), 256), 256, 0, ctx->stream>>>(*old_t, new_pos, n_new, T_new, b);
}

// ---------------------------------------------------------------------------------------------
// contact
// ---------------------------------------------------------------------------------------------

// material_fric_heat_gen_FEM_v2 (actions_gpu.cu:345-420, 967-981) over all particles
// the two record lists of a tool's ordered segment sums
__device__ __forceinline__ SegSink fric_sink(const sphfem_tool &t, int *counters) {
    return SegSink{counters + SPHFEM_FLAG_SINK_FRIC, t.sink_keys, t.sink_vals, SPHFEM_SEG_SINK_CAP};
}
__device__ __forceinline__ SegSink s2_sink(const sphfem_tool &t, int *counters) {
    return SegSink{counters + SPHFEM_FLAG_SINK_S2, t.sink_keys + SPHFEM_SEG_SINK_CAP,
                   t.sink_vals + 6ul * SPHFEM_SEG_SINK_CAP, SPHFEM_SEG_SINK_CAP};
}

__global__ void __launch_bounds__(256) k_fric_heat(DevConsts c, sphfem_particles P, sphfem_tool t, int *counters) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)P.N) return;
    if (P.blanked[i] == 1.) return;
    ToolView tv{};
    tv.T_nodes = t.T;
    tv.marks = t.marks;
    tv.fric_heat = t.fric_heat;
    tv.physical_para = t.physical_para;
    tv.sliding_force = t.sliding_force;
    tv.sink = fric_sink(t, counters);
    double T = P.T[i];
    if (!fric_heat_v21(c, tv, *t.kin, i, P.eps_pl[i], P.ft[i], P.on_seg[i], P.n[i], P.vel[i], P.fc[i], P.rho[i], T))
        P.blanked[i] = 1.;
    P.T[i] = T;
}

void launch_fric_heat(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_tool *t) {
    launch_fill_int(ctx, ctx->d_flags + SPHFEM_FLAG_SINK_FRIC, 1, 0);
    {
        KTimer tm(ctx, "fric_heat");
        k_fric_heat<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(ctx->c, p, *t, ctx->d_flags);
    }
    launch_fric_heat_sums(ctx, t);
}

// compute_contact_forces_FEM_tool_v3 step 1 (actions_gpu.cu:423-583) over all particles
__global__ void __launch_bounds__(256) k_contact_s1(DevConsts c, sphfem_particles P, sphfem_tool t) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)P.N) return;
    if (P.blanked[i] == 1.) return;
    const ToolKin k = *t.kin;
    sphfem_double2 pos = P.pos[i];
    sphfem_double2 fc, ft = P.ft[i], nrm;
    int seg = -1;
    bool blank = false;
    if (near_tool(k, pos)) {
        seg = contact_s1_near(c, k, pos, P.vel[i], fc, ft, nrm, blank, t.left, t.right, t.n, t.n_seg);
    } else {
        fc = {0., 0.};
        ft = {0., 0.};
        nrm = {0., 0.};
    }
    if (blank) P.blanked[i] = 1.;
    P.fc[i] = fc;
    P.ft[i] = ft;
    P.n[i] = nrm;
    P.on_seg[i] = seg;
}

// step 2 (actions_gpu.cu:585-617) for one particle
__device__ __forceinline__ void contact_s2_one(const DevConsts &c, const sphfem_tool &t, const SegSink &sink, int pidx,
                                               int seg_no, double rho, double &T) {
    sphfem_int4 mk = t.marks[seg_no];
    double T_ = T;
    double T_seg = 0.5 * (t.T[mk.z] + t.T[mk.w]);
    double delta_T = T_seg - T_;
    double length = sqrtf_d(c.tc.slave_mass / rho);
    const double hv[2] = {T_ * length, length};
    double *const hd[2] = {&t.heat_exchange[0].x, &t.heat_exchange[0].y};
    const int hs[2] = {2, 2};
    seg_sink_add<2>(sink, true, seg_no, pidx, hv, hd, hs);
    double dTp = delta_T * c.tc.h * length / c.trml_wp.cp / c.tc.slave_mass * c.dt;
    if (dTp < 0) T += fmax(dTp, delta_T);
    else T += fmin(dTp, delta_T);
}

__global__ void __launch_bounds__(256) k_contact_s2(DevConsts c, sphfem_particles P, sphfem_tool t, int *counters) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)P.N) return;
    if (P.blanked[i] == 1.) return;
    int seg_no = P.on_seg[i];
    if (seg_no == -1) return;
    double T = P.T[i];
    contact_s2_one(c, t, s2_sink(t, counters), i, seg_no, P.rho[i], T);
    P.T[i] = T;
}

// The ordered segment sums: sort the records (segment << 32 | particle index, unique) in LDS and add
// each segment's values in increasing particle index to the segment arrays (dst[q][seg * stride[q]]).
struct SegDst {
    double *p[6];
    int stride[6];
};

constexpr int SEG_REDUCE_BS = 1024;
constexpr int SEG_REDUCE_PER = SPHFEM_SEG_SINK_CAP / SEG_REDUCE_BS;   // records per thread, at most
static_assert(SPHFEM_SEG_SINK_CAP % SEG_REDUCE_BS == 0, "sink capacity: whole records per thread");

template <int NV>
__global__ void __launch_bounds__(SEG_REDUCE_BS) k_segment_reduce(SegSink s, SegDst d) {
    __shared__ unsigned long long sk[SPHFEM_SEG_SINK_CAP];   // keys; after the sort, one value per record
    __shared__ int sl[SPHFEM_SEG_SINK_CAP];                  // record slots
    __shared__ int sseg[SPHFEM_SEG_SINK_CAP];                // segments in sorted order
    const int n = min(*s.count, s.cap);
    if (n <= 0) return;
    int np = 1;
    while (np < n) np <<= 1;
    for (int k = threadIdx.x; k < np; k += SEG_REDUCE_BS) {
        sk[k] = k < n ? s.keys[k] : ~0ull;
        sl[k] = k;
    }
    __syncthreads();
    for (int size = 2; size <= np; size <<= 1) {          // bitonic sort of (key, slot), ascending
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int k = threadIdx.x; k < np; k += SEG_REDUCE_BS) {
                const int j = k ^ stride;
                if (j > k) {
                    const bool up = (k & size) == 0;
                    const unsigned long long a = sk[k], b = sk[j];
                    if ((a > b) == up) {
                        sk[k] = b;
                        sk[j] = a;
                        const int x = sl[k];
                        sl[k] = sl[j];
                        sl[j] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int k = threadIdx.x; k < n; k += SEG_REDUCE_BS) sseg[k] = (int)(unsigned)(sk[k] >> 32);
    __syncthreads();
    // every global load of the additions is issued here, before the first is used: the values of
    // this thread's records (sorted positions threadIdx.x + u * block) and, for a record that
    // starts its segment's run, the segment's current sums (one memory wait instead of one per value)
    double v[NV][SEG_REDUCE_PER], acc0[NV][SEG_REDUCE_PER];
    int seg_of[SEG_REDUCE_PER];
    bool head[SEG_REDUCE_PER];
#pragma unroll
    for (int u = 0; u < SEG_REDUCE_PER; u++) {
        const int k = threadIdx.x + u * SEG_REDUCE_BS;
        const bool in = k < n;
        const int slot = in ? sl[k] : 0;
        const int seg = in ? sseg[k] : 0;
        seg_of[u] = seg;
        head[u] = in && (k == 0 || sseg[k - 1] != seg);
#pragma unroll
        for (int q = 0; q < NV; q++) {
            v[q][u] = in ? s.vals[(size_t)q * s.cap + slot] : 0.;
            acc0[q][u] = head[u] ? d.p[q][(size_t)seg * d.stride[q]] : 0.;
        }
    }
    double *vbuf = reinterpret_cast<double *>(sk);
#pragma unroll
    for (int q = 0; q < NV; q++) {
        // this value of every record in sorted order, then each segment's run added in order by
        // the run's first thread (from LDS: long runs stay cheap)
#pragma unroll
        for (int u = 0; u < SEG_REDUCE_PER; u++) {
            const int k = threadIdx.x + u * SEG_REDUCE_BS;
            if (k < n) vbuf[k] = v[q][u];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < SEG_REDUCE_PER; u++) {
            if (!head[u]) continue;
            const int k = threadIdx.x + u * SEG_REDUCE_BS;
            double acc = acc0[q][u];
            for (int r = k; r < n && sseg[r] == seg_of[u]; r++) acc += vbuf[r];
            d.p[q][(size_t)seg_of[u] * d.stride[q]] = acc;
        }
        __syncthreads();
    }
}

static SegSink host_sink(sphfem_ctx *ctx, sphfem_tool *t, bool s2) {
    if (s2)
        return SegSink{ctx->d_flags + SPHFEM_FLAG_SINK_S2, t->sink_keys + SPHFEM_SEG_SINK_CAP,
                       t->sink_vals + 6ul * SPHFEM_SEG_SINK_CAP, SPHFEM_SEG_SINK_CAP};
    return SegSink{ctx->d_flags + SPHFEM_FLAG_SINK_FRIC, t->sink_keys, t->sink_vals, SPHFEM_SEG_SINK_CAP};
}

void launch_fric_heat_sums(sphfem_ctx *ctx, sphfem_tool *t) {
    SegDst d{};
    d.p[0] = t->fric_heat;
    d.p[1] = &t->physical_para[0].x;
    d.p[2] = &t->physical_para[0].y;
    d.p[3] = &t->physical_para[0].z;
    d.p[4] = &t->physical_para[0].w;
    d.p[5] = t->sliding_force;
    const int st[6] = {1, 4, 4, 4, 4, 1};
    for (int q = 0; q < 6; q++) d.stride[q] = st[q];
    KTimer tm(ctx, "fric_heat_sums");
    k_segment_reduce<6><<<1, SEG_REDUCE_BS, 0, ctx->stream>>>(host_sink(ctx, t, false), d);
}

static void launch_s2_sums(sphfem_ctx *ctx, sphfem_tool *t) {
    SegDst d{};
    d.p[0] = &t->heat_exchange[0].x;
    d.p[1] = &t->heat_exchange[0].y;
    d.stride[0] = d.stride[1] = 2;
    KTimer tm(ctx, "contact_s2_sums");
    k_segment_reduce<2><<<1, SEG_REDUCE_BS, 0, ctx->stream>>>(host_sink(ctx, t, true), d);
}

void launch_contact(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_tool *t) {
    {
        KTimer tm(ctx, "contact_s1");
        k_contact_s1<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(ctx->c, p, *t);
    }
    launch_fill_int(ctx, ctx->d_flags + SPHFEM_FLAG_SINK_S2, 1, 0);
    {
        KTimer tm(ctx, "contact_s2");
        k_contact_s2<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(ctx->c, p, *t, ctx->d_flags);
    }
    launch_s2_sums(ctx, t);
}

// Near-tool particles of the fused step: contact s1, mechanical boundary condition, NaN check, then
// for particles left in contact contact s2 (its heat exchange reads only the particle itself and
// the tool's temperatures, so it follows s1 in the same lane), and the thermal boundary condition
// and the invalidation.
// use32: the rates are the FP32 derivative carry r32 of the FP32C form (the particle arrays' FP64
// rates are not maintained by its steps)
__global__ void __launch_bounds__(256) k_near_contact(DevConsts c, sphfem_particles P, sphfem_tool t,
                                                      const int *__restrict__ near_list, int *counters,
                                                      RateView r32, int use32) {
    // one particle of the near list per wavefront (wave-uniform loop); the segment loops of
    // contact s1 run over the lanes
    const int count = counters[1];
    const ToolKin k = *t.kin;
    const int lane = threadIdx.x & 63;
    const int nwaves = (gridDim.x * blockDim.x) >> 6;
    for (int q = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; q < count; q += nwaves) {
        const int i = near_list[q];
        sphfem_double2 fc, ft = P.ft[i], nrm;
        bool blank = false;
        const int seg = contact_s1_near_wave(c, k, P.pos[i], P.vel[i], fc, ft, nrm, blank, t.left, t.right, t.n,
                                             t.n_seg, lane);
        if (lane != 0) continue;
        double blanked = blank ? 1. : P.blanked[i];
        sphfem_double2 pos_t, vel_t;
        if (use32) {
            const float2 a = r32.pos_t[i], b = r32.vel_t[i];
            pos_t = {a.x, a.y};
            vel_t = {b.x, b.y};
        } else {
            pos_t = P.pos_t[i];
            vel_t = P.vel_t[i];
        }
        if (P.tool_particle[i] == 0. && P.fixed[i] != 0.) {
            sphfem_double2 z = {0., 0.};
            P.vel[i] = z;
            fc = z;
            pos_t = z;
            vel_t = z;
            if (use32) {
                r32.pos_t[i] = make_float2(0.f, 0.f);
                r32.vel_t[i] = make_float2(0.f, 0.f);
            } else {
                P.pos_t[i] = z;
                P.vel_t[i] = z;
            }
        }
        P.fc[i] = fc;
        P.ft[i] = ft;
        P.n[i] = nrm;
        P.on_seg[i] = seg;
        bool bad = false;
        if (blanked != 1.) {
            sphfem_double4 st;
            double rt, Tt;
            if (use32) {
                const float4 a = r32.S_t[i];
                st = {a.x, a.y, a.z, a.w};
                rt = r32.rho_t[i];
                Tt = r32.T_t[i];
            } else {
                st = P.S_t[i];
                rt = P.rho_t[i];
                Tt = P.T_t[i];
            }
            bad = isnaninf(pos_t.x) || isnaninf(pos_t.y) || isnaninf(vel_t.x) || isnaninf(vel_t.y) ||
                  isnaninf(st.x) || isnaninf(st.y) || isnaninf(st.z) || isnaninf(st.w) || isnaninf(rt) ||
                  isnaninf(Tt);
        }
        if (seg != -1) {
            // contact s2 (actions_gpu.cu:585-617; one lane, its sink record slot from a one-lane ballot)
            double T = P.T[i];
            contact_s2_one(c, t, s2_sink(t, counters), i, seg, P.rho[i], T);
            if (c.thermal_bc && P.fixed[i] == 1.) T = c.trml_wp.T_init;
            P.T[i] = T;
        } else if (c.thermal_bc && P.fixed[i] == 1.) {
            P.T[i] = c.trml_wp.T_init;
        }
        if (bad) blanked = 1.;
        if (blanked == 1.) P.blanked[i] = 1.;
    }
}

void launch_near_contact(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_tool *t, sphfem_leap_frog *lf) {
    int blocks = std::max(1, std::min(nblk(p.N, 256), 1024));
    {
        KTimer tm(ctx, "near_contact");
        const bool use32 = lf->rates_fp32;
        const RateView r32 = use32 ? rate_view(lf->rate32[lf->rate_cur], lf->num_part) : RateView{};
        k_near_contact<<<blocks, 256, 0, ctx->stream>>>(ctx->c, p, *t, lf->near_list, ctx->d_flags, r32,
                                                        use32 ? 1 : 0);
    }
    launch_s2_sums(ctx, t);
}

}  // namespace sphfem
