// Rigid (SPH) tool kernels: the tool move of tool_gpu_update_tool (tool_gpu.cu:205-216, 270-311),
// the contact of compute_contact_forces (tool_gpu.cu:146-202, 314-319) and the friction heat of
// material_fric_heat_gen (actions_gpu.cu:316-343, 956-965).  The tool geometry stays in device
// memory and is moved there, so that a step needs no host round trip and can be recorded in a graph.
#include "sphfem_rigid.h"

#pragma clang fp contract(off)

namespace sphfem {

// tool.cpp:202-212 on the device copy (one lane)
__global__ void k_rigid_advance(RigidTool *t, double dt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) rigid_move(t->g, dt);
}

// tool_gpu.cu:205-216: tool particles move with the tool (blanked ones too, as there)
__global__ void __launch_bounds__(256) k_rigid_tool_particles(sphfem_particles P, const RigidTool *__restrict__ t,
                                                              double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)P.N) return;
    if (P.tool_particle[i] == 0.) return;
    sphfem_double2 pos = P.pos[i];
    pos.x += dt * t->g.vel.x;
    pos.y += dt * t->g.vel.y;
    P.pos[i] = pos;
}

// tool_gpu.cu:146-202 (every non-tool particle, blanked ones included, as there)
__global__ void __launch_bounds__(256) k_rigid_contact(sphfem_particles P, const RigidTool *__restrict__ t, double dt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)P.N) return;
    if (P.tool_particle[i] != 0.) return;
    sphfem_double2 fc, ft, n;
    rigid_contact(*t, dt, P.pos[i], P.vel[i], P.ft[i], fc, ft, n);
    P.fc[i] = fc;
    P.ft[i] = ft;
    P.n[i] = n;
}

// actions_gpu.cu:316-343 (every non-tool particle, as there)
__global__ void __launch_bounds__(256) k_rigid_fric_heat(DevConsts c, sphfem_particles P, sphfem_double2 vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int)P.N) return;
    if (P.tool_particle[i] != 0.) return;
    P.T[i] = rigid_fric_heat(c, vel, P.vel[i], P.ft[i], P.n[i], P.T[i]);
}

// tool_wear.cu:14-72: Usui wear rate of the particles in contact, reduced per block in a fixed
// order (sum, min, max, count) and then over the blocks by one lane
constexpr int USUI_BLOCKS = 256;

__global__ void __launch_bounds__(256) k_usui_partial(sphfem_particles P, double K, double alpha,
                                                      sphfem_double2 vel_tool, double *__restrict__ part) {
    __shared__ double s_sum[256], s_min[256], s_max[256], s_cnt[256];
    double sum = 0., mn = (double)FLT_MAX, mx = 0., cnt = 0.;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < (int)P.N; i += gridDim.x * 256) {
        const sphfem_double2 fc = P.fc[i], ft = P.ft[i];
        // the fourth test repeats fc.y where ft.y is meant (tool_wear.cu:31), kept as written
        const bool in_contact = fabs(fc.x) > 0. || fabs(fc.y) > 0. || fabs(ft.x) > 0. || fabs(fc.y) > 0.;
        if (!in_contact) continue;
        const sphfem_double2 v = P.vel[i], n = P.n[i];
        const double vx = v.x - vel_tool.x, vy = v.y - vel_tool.y;
        const double vrx = vx - vx * n.x, vry = vy - vy * n.y;
        const double vmag = sqrt(vrx * vrx + vry * vry);
        const double fmag = sqrt(fc.x * fc.x + fc.y * fc.y);
        const double A = 2 * P.h[i] * SPHFEM_RIGID_PI;
        const double p = fmag / A;
        const double w = K * exp(alpha / P.T[i]) * p * vmag;
        sum += w;
        mn = fmin(mn, w);
        mx = fmax(mx, w);
        cnt += 1.;
    }
    s_sum[threadIdx.x] = sum;
    s_min[threadIdx.x] = mn;
    s_max[threadIdx.x] = mx;
    s_cnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_sum[threadIdx.x] += s_sum[threadIdx.x + s];
            s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + s]);
            s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + s]);
            s_cnt[threadIdx.x] += s_cnt[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x + 0] = s_sum[0];
        part[4 * blockIdx.x + 1] = s_min[0];
        part[4 * blockIdx.x + 2] = s_max[0];
        part[4 * blockIdx.x + 3] = s_cnt[0];
    }
}

__global__ void k_usui_final(const double *__restrict__ part, int nb, double *__restrict__ out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double sum = 0., mn = (double)FLT_MAX, mx = 0., cnt = 0.;
    for (int b = 0; b < nb; b++) {
        sum += part[4 * b];
        mn = fmin(mn, part[4 * b + 1]);
        mx = fmax(mx, part[4 * b + 2]);
        cnt += part[4 * b + 3];
    }
    out[0] = sum;
    out[1] = mn;
    out[2] = mx;
    out[3] = cnt;
}

// d_out: [4 + 4 * USUI_BLOCKS] doubles; out[0..3] = sum, min, max, count
void launch_usui(sphfem_ctx *ctx, const sphfem_particles &p, double K, double alpha, sphfem_double2 vel,
                 double *d_out) {
    KTimer t(ctx, "usui");
    k_usui_partial<<<USUI_BLOCKS, 256, 0, ctx->stream>>>(p, K, alpha, vel, d_out + 4);
    k_usui_final<<<1, 64, 0, ctx->stream>>>(d_out + 4, USUI_BLOCKS, d_out);
}

void launch_rigid_advance(sphfem_ctx *ctx, sphfem_rigid_tool *rt) {
    KTimer t(ctx, "rigid_advance");
    k_rigid_advance<<<1, 64, 0, ctx->stream>>>(rt->dev, ctx->c.dt);
}

void launch_rigid_tool_particles(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_rigid_tool *rt) {
    KTimer t(ctx, "rigid_tool_particles");
    k_rigid_tool_particles<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(p, rt->dev, ctx->c.dt);
}

void launch_rigid_contact(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_rigid_tool *rt) {
    KTimer t(ctx, "rigid_contact");
    k_rigid_contact<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(p, rt->dev, ctx->c.dt);
}

void launch_rigid_fric_heat(sphfem_ctx *ctx, const sphfem_particles &p, sphfem_double2 vel) {
    KTimer t(ctx, "rigid_fric_heat");
    k_rigid_fric_heat<<<nblk(p.N, 256), 256, 0, ctx->stream>>>(ctx->c, p, vel);
}

}  // namespace sphfem
