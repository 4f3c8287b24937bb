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
