/*
 * sphfem.h -- C ABI of the MI355X (gfx950) SPH-FEM cutting hot path.
 *
 * Drop-in boundary for the per-timestep path of nyzhang1227/SPH_FEM_2D_cutting_gpu:
 * every entry point replaces one reference function (cited per declaration) and
 * operates on device arrays laid out as the reference's structs.  All calls are
 * asynchronous on the context's HIP stream; only sphfem_synchronize and the
 * download helpers block.  Errors are returned as status codes (0 = OK) and
 * described by sphfem_last_error(); nothing throws across this boundary.
 *
 * Units are the reference's (cm, g, us, K; stress in units of 100 GPa).
 */
#ifndef SPHFEM_H
#define SPHFEM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPHFEM_OK 0
#define SPHFEM_ERR_HIP 1      /* a HIP runtime call failed */
#define SPHFEM_ERR_ARG 2      /* invalid argument */
#define SPHFEM_ERR_STATE 3    /* call out of order (e.g. constants missing) */
#define SPHFEM_ERR_RANGE 4    /* a particle left the cell grid */

#define SPHFEM_FP64 0         /* plain FP64 form (reference arithmetic) */
#define SPHFEM_FP32C 1        /* FP32 pair sums with compensated accumulation */

#define SPHFEM_GRID_GREEN 0    /* grid_gpu_green: gather into a double buffer, pointer swap */
#define SPHFEM_GRID_ROTHLIN 1  /* grid_gpu_rothlin: in-place gathers, dense cell offsets */

typedef struct { double x, y; } sphfem_double2;
typedef struct { double x, y, z, w; } sphfem_double4;
typedef struct { int x, y, z, w; } sphfem_int4;

/* constants_structs.h:18-123 (the friction-model fields under SHEAR_LIMIT_FRI etc. are off) */
typedef struct { double E, nu, rho0, K, G, mass, mass_tool; } sphfem_phys_constants;
typedef struct { double cp, tq, k, alpha, T_init, eta; } sphfem_trml_constants;
typedef struct { double wdeltap, stresseps, xspheps, alpha, beta, eta; } sphfem_corr_constants;
typedef struct { double A, B, C, n, m, Tmelt, Tref, eps_dot_ref, tanh_a, tanh_b, tanh_c, tanh_d; } sphfem_joco_constants;
typedef struct { int nx, ny; double bbmin_x, bbmin_y, dx; } sphfem_geom_constants;
typedef struct { double cp, tq, k, eta, rho, T0, h, beta, mu, contact_alpha, slave_mass, hc; } sphfem_tool_constants;
typedef struct { double A, B, C, D; } sphfem_wear_constants;

/* particle_gpu.h:23-83: device pointers in the reference's field order.  Like the
 * reference, sorting and stepping swap these pointers with internal buffers, so
 * callers re-read them after every call that takes the struct by pointer. */
typedef struct sphfem_particles {
    sphfem_double2 *pos, *vel;
    double *h, *rho, *p;
    sphfem_double4 *S, *R;
    sphfem_double2 *fc, *ft, *n;
    double *fixed, *blanked, *tool_particle;
    double *eps_pl, *eps_pl_dot;
    double *T;
    sphfem_double2 *pos_t, *vel_t;
    double *rho_t;
    sphfem_double4 *S_t;
    double *T_t;
    sphfem_double4 *S_der, *v_der;
    int *idx, *hash, *on_seg;
    unsigned int N;
} sphfem_particles;

/* Host-side description of a rigid FEM tool, the data set_up_segments_from_mesh and
 * set_up_FEM_tool_gpu leave on the device (tool_FEM.cu:771-1039, 1539-1544;
 * FEM_thermal_gpu.cu:584-659; segment_FEM_gpu.h:31-56; mesh_structure_GPU.h:80-142).
 * All pointers are host pointers; sphfem_tool_create copies them. */
typedef struct sphfem_tool_desc {
    int node_num, ele_num, nz_num, seg_num, color_total, color_cutting_edge;
    const sphfem_double2 *position;      /* [node_num] */
    const sphfem_int4 *node_num_el;      /* [ele_num] 3 node ids + colour */
    const int *layer;                    /* [ele_num] 1 if in cutting boundary layer 1 or 2 */
    const int *csr_row_ptr;              /* [node_num + 1] */
    const int *csr_col_ind;              /* [nz_num] */
    const double *K, *M, *f, *T;         /* stiffness [nz_num], lumped mass, rhs, temperature [node_num] */
    const sphfem_double2 *left, *right, *n;   /* [seg_num] */
    const sphfem_int4 *marks;            /* [seg_num] element, colour, left node, right node */
    sphfem_double2 vel;                  /* tool velocity (tool.h:48-58) */
    sphfem_double2 tl_point_ref, br_point;
    double ref_dist, lowest_point, right_point;
    /* optional, needed by the remesh entry points (mesh_structure_GPU.h:88-115): per node bit 0 on
     * the cutting edge, bit 1 on the fixed boundary [node_num]; the fixed-boundary elements and
     * their edge flags (x: nodes 0-1, y: 1-2, z: 2-0) [fixed_num].  NULL / 0 when not given. */
    const int *node_flags;
    int fixed_num;
    const int *fixed_elems;
    const sphfem_int4 *fixed_flags;
} sphfem_tool_desc;

/* tool kinematic state kept on the device (tool_FEM.h:34-60 host fields + tool_FEM.cu:19-20) */
typedef struct {
    double vel_x, vel_y, tl_x, tl_y, br_x, br_y, ref_dist, lowest_point, right_point;
    double flank_retreat, flank_contact;
} sphfem_tool_kinematics;

typedef struct sphfem_ctx sphfem_ctx;
typedef struct sphfem_grid sphfem_grid;
typedef struct sphfem_tool sphfem_tool;
typedef struct sphfem_leap_frog sphfem_leap_frog;

/* ---- context (replaces the per-translation-unit __constant__ state and the legacy stream) */
int sphfem_create(int device, sphfem_ctx **out);
int sphfem_destroy(sphfem_ctx *ctx);
const char *sphfem_last_error(sphfem_ctx *ctx);
int sphfem_synchronize(sphfem_ctx *ctx);
/* check_cuda_error (types.cpp:11-37): synchronise and report device-side errors */
int sphfem_check_error(sphfem_ctx *ctx);

/* ---- constants (actions_gpu.cu:913-954, interactions_gpu.cu:476-507, FEM_thermal_gpu.cu:597) */
int sphfem_setup_physical_constants(sphfem_ctx *ctx, const sphfem_phys_constants *c);
int sphfem_setup_corrector_constants(sphfem_ctx *ctx, const sphfem_corr_constants *c);
int sphfem_setup_johnson_cook_constants(sphfem_ctx *ctx, const sphfem_joco_constants *c);
int sphfem_setup_thermal_constants_wp(sphfem_ctx *ctx, const sphfem_trml_constants *c);
int sphfem_setup_thermal_constants_tool(sphfem_ctx *ctx, const sphfem_trml_constants *c);
int sphfem_setup_tool_constants(sphfem_ctx *ctx, const sphfem_tool_constants *c);
int sphfem_setup_wear_constants(sphfem_ctx *ctx, const sphfem_wear_constants *c);
/* global_dt (benchmarks.cpp:12) and wear_pro (mfree_gpu.cpp:47) */
int sphfem_set_dt(sphfem_ctx *ctx, double dt);
int sphfem_set_wear(sphfem_ctx *ctx, int wear_pro);
/* lower pair cut-off of cubic_spline (kernels.cuh:38, 1e-8 cm^2 in the reference) */
int sphfem_set_kernel_cutoff(sphfem_ctx *ctx, double rr2_min);

/* ---- device memory helpers */
int sphfem_malloc(sphfem_ctx *ctx, void **ptr, size_t bytes);
int sphfem_free(sphfem_ctx *ctx, void *ptr);
int sphfem_memcpy_htod(sphfem_ctx *ctx, void *dst, const void *src, size_t bytes);
int sphfem_memcpy_dtoh(sphfem_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- particles (particle_gpu.cu:11-104: zeroed arrays, idx = 0..N-1, on_seg = 0) */
int sphfem_particles_alloc(sphfem_ctx *ctx, unsigned int N, sphfem_particles *out);
int sphfem_particles_free(sphfem_ctx *ctx, sphfem_particles *p);

/* ---- cell list, grid_gpu_green (grid.h:23-64, grid_gpu_green.h:84-85) */
int sphfem_grid_green_create(sphfem_ctx *ctx, int num_part, sphfem_double2 bbmin, sphfem_double2 bbmax,
                             double h, sphfem_grid **out);
/* grid_gpu_rothlin.h:79-80 (locked-geometry constructor, grid.cu:130-149 geometry).  Same cell
 * geometry and hashes as the Green grid; sort and get_cells follow grid_gpu_rothlin.cu (below). */
int sphfem_grid_rothlin_create(sphfem_ctx *ctx, int num_part, sphfem_double2 bbmin, sphfem_double2 bbmax,
                               double h, sphfem_grid **out);
/* grid_gpu_green(max_cell, N) (grid_gpu_green.cu:258-262) and grid_gpu_rothlin(max_cell, N)
 * (grid_gpu_rothlin.cu:277-280): geometry not locked; sphfem_grid_update_geometry (called by the
 * step) sets it from the particles, for up to max_cell - 1 cells. */
int sphfem_grid_green_create_unlocked(sphfem_ctx *ctx, int max_cell, int num_part, sphfem_grid **out);
int sphfem_grid_rothlin_create_unlocked(sphfem_ctx *ctx, int max_cell, int num_part, sphfem_grid **out);
/* grid.cu:88-127: no-op for a locked grid; otherwise bounding box of all particle positions
 * less/plus 1e-6, cell size max(h) * kernel_width / 2, ceil counts.  SPHFEM_ERR_RANGE when that
 * needs max_cell cells or more (the reference asserts).  Synchronises the stream (the box is read
 * back to the host, as in the reference). */
int sphfem_grid_update_geometry(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles *p, double kernel_width);
int sphfem_grid_destroy(sphfem_ctx *ctx, sphfem_grid *g);
int sphfem_grid_geometry(const sphfem_grid *g, sphfem_geom_constants *geom, int *num_cell);
/* grid.h:45-50: bbmax_x/bbmax_y, is_locked, and the cell capacity */
int sphfem_grid_bounds(const sphfem_grid *g, sphfem_double2 *bbmax, int *locked, int *max_cell);
/* SPHFEM_GRID_GREEN or SPHFEM_GRID_ROTHLIN */
int sphfem_grid_kind(const sphfem_grid *g, int *kind);
/* grid.cu:80-86 */
int sphfem_grid_assign_hashes(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles *p);
/* Green (grid_gpu_green.cu:184-213): stable radix sort, cell boundaries, gather of 20 arrays,
 * pointer swap (p's array pointers change).
 * Rothlin (grid_gpu_rothlin.cu:129-249): stable radix sort, in-place gathers (p's pointers are
 * kept) of idx (which then carries particle labels), pos, vel, fc, h, rho, S, eps_pl, eps_pl_dot,
 * fixed, blanked, tool_particle, T, pos_t, vel_t, rho_t, S_t, T_t and on_seg; ft and n stay in
 * place as in the reference.  on_seg is gathered, where the reference copies the gathered idx
 * into it (:244-245). */
int sphfem_grid_sort(sphfem_ctx *ctx, sphfem_grid *g, sphfem_particles *p);
/* The step re-sorts incrementally when at most max_changed particles changed cells since the
 * previous step (the result equals the full stable sort); otherwise, and on the first step, it
 * runs the full radix sort.  0 always runs the full sort.  Values above 4096 are clamped.
 * No counterpart in the reference (a tuning knob of this implementation). */
int sphfem_grid_set_resort_limit(sphfem_ctx *ctx, sphfem_grid *g, int max_changed);
/* the last step's re-sort: mode 1 = full radix sort, 2 = incremental; changed = particles that
 * changed cells (blocks until the stream is idle) */
int sphfem_grid_resort_stats(sphfem_ctx *ctx, int *mode, int *changed);
/* grid.h:57 get_cells(particles, cell_start, cell_end) into caller arrays [num_cell].
 * Green (grid_gpu_green.cu:215-218): copy of the tables of the last sort (-1 = empty cell; p is
 * not read and may be NULL).  Rothlin (grid_gpu_rothlin.cu:19-126): dense offsets from p's sorted
 * hashes, cell_start[c] = #{hash < c}, cell_end[c] = #{hash < c + 1} (p required). */
int sphfem_grid_get_cells(sphfem_ctx *ctx, sphfem_grid *g, const sphfem_particles *p, int *cell_start,
                          int *cell_end);

/* ---- pair interactions (interactions_gpu.cu:381-432, 466-474) */
int sphfem_interactions_setup_geometry_constants(sphfem_ctx *ctx, const sphfem_grid *g);
int sphfem_interactions_monaghan(sphfem_ctx *ctx, sphfem_particles *p, const int *cell_start,
                                 const int *cell_end, int num_cell);
/* interactions_gpu.cu:466-474 in the FP32-compensated form (the pair sums of
 * sphfem_leap_frog_step with SPHFEM_FP32C, lf set to that precision): FP32 terms on cell-relative
 * positions, FP32 blocks of 8 terms accumulated in FP64, from the sorted state p (p and R as left
 * by material_eos and corrector_artificial_stress).  Writes v_der, pos_t and T_t as the reference
 * does; the stress-divergence and viscous accelerations come summed per axis as S_der = {ax, 0,
 * ay, 0} with vel_t = 0, the two sums the momentum equation (actions_gpu.cu:641-671) uses. */
int sphfem_interactions_monaghan_fp32c(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p,
                                       const int *cell_start, const int *cell_end, int num_cell);
/* interactions_gpu.cu:84-140, 434-464 (PSE heat conduction, compiled out of the reference's
 * flag set): T_t = alpha sum (T_j - T_i) 4/(pi h^4) exp(-r^2/h^2) m/rho_j over the 5x5 cells */
int sphfem_interactions_heat_pse(sphfem_ctx *ctx, sphfem_particles *p, const int *cell_start,
                                 const int *cell_end, int num_cell);

/* ---- per-particle actions (actions_gpu.h:33-72) */
int sphfem_material_eos(sphfem_ctx *ctx, sphfem_particles *p);
int sphfem_corrector_artificial_stress(sphfem_ctx *ctx, sphfem_particles *p);
int sphfem_material_stress_rate_jaumann(sphfem_ctx *ctx, sphfem_particles *p);
int sphfem_contmech_continuity(sphfem_ctx *ctx, sphfem_particles *p);
int sphfem_contmech_momentum(sphfem_ctx *ctx, sphfem_particles *p);
int sphfem_contmech_advection(sphfem_ctx *ctx, sphfem_particles *p);
int sphfem_plasticity_johnson_cook(sphfem_ctx *ctx, sphfem_particles *p);
int sphfem_perform_boundary_conditions(sphfem_ctx *ctx, sphfem_particles *p);
int sphfem_perform_boundary_conditions_thermal(sphfem_ctx *ctx, sphfem_particles *p);
int sphfem_debug_invalidate(sphfem_ctx *ctx, sphfem_particles *p);
int sphfem_material_fric_heat_gen_FEM_v2(sphfem_ctx *ctx, sphfem_particles *p, sphfem_tool *t);
int sphfem_compute_contact_forces_FEM_tool_v3(sphfem_ctx *ctx, sphfem_particles *p, sphfem_tool *t);
int sphfem_wear_adjustment(sphfem_ctx *ctx, sphfem_tool *t);

/* ---- FEM tool (tool_FEM.h:57-83, FEM_thermal_gpu.h:67-91) */
int sphfem_tool_create(sphfem_ctx *ctx, const sphfem_tool_desc *desc, sphfem_tool **out);
int sphfem_tool_destroy(sphfem_ctx *ctx, sphfem_tool *t);
int sphfem_tool_set_vel(sphfem_ctx *ctx, sphfem_tool *t, sphfem_double2 vel);
int sphfem_tool_get_kinematics(sphfem_ctx *ctx, sphfem_tool *t, sphfem_tool_kinematics *out);
int sphfem_tool_segments_update(sphfem_ctx *ctx, sphfem_tool *t, double dt);
int sphfem_tool_update_FEM_gpu(sphfem_ctx *ctx, sphfem_tool *t, double dt);
int sphfem_tool_segments_wear_rate_reset(sphfem_ctx *ctx, sphfem_tool *t);
int sphfem_tool_apply_heat_boundary_condition(sphfem_ctx *ctx, sphfem_tool *t);
int sphfem_FEM_reset_server(sphfem_ctx *ctx, sphfem_tool *t);
int sphfem_FEM_calculate_temp(sphfem_ctx *ctx, sphfem_tool *t, double dt);
int sphfem_FEM_calculate_temp_rate(sphfem_ctx *ctx, sphfem_tool *t);
/* copy the per-step tool state to host arrays (NULL pointers are skipped) */
int sphfem_tool_download(sphfem_ctx *ctx, sphfem_tool *t, sphfem_double2 *position, double *T, double *K,
                         double *M, double *k, sphfem_double2 *left, sphfem_double2 *right, sphfem_double2 *n,
                         sphfem_double2 *heat_exchange, double *fric_heat, sphfem_double4 *physical_para,
                         double *sliding_force, double *wear_nodes);

/* restore a per-step tool state from host arrays (same layout as sphfem_tool_download; NULL skips) */
int sphfem_tool_upload(sphfem_ctx *ctx, sphfem_tool *t, const sphfem_double2 *position, const double *T,
                       const double *K, const double *M, const double *k, const sphfem_double2 *left,
                       const sphfem_double2 *right, const sphfem_double2 *n, const sphfem_double2 *heat_exchange,
                       const double *fric_heat, const sphfem_double4 *physical_para, const double *sliding_force,
                       const double *wear_nodes);
int sphfem_tool_set_kinematics(sphfem_ctx *ctx, sphfem_tool *t, const sphfem_tool_kinematics *kin);

/* ---- rigid (SPH) tool: tool.h:31-141, tool_gpu.h:19-28, tool_gpu.cuh:23-45 */
#define SPHFEM_TOOL_MAX_SEG 5
/* geometric_primitives.h:15-57: y = a x + b, or x = b when vertical (a = DBL_MAX) */
typedef struct { double a, b; int vertical; } sphfem_line;
/* left/right ends, line, normal normalize(dy, -dx) of right - left (pointing into the tool) */
typedef struct { sphfem_double2 left, right; sphfem_line l; sphfem_double2 n; } sphfem_segment;
/* fillet of radius r about p between the angles t1 and t2 (radians in (0, 2 pi]) */
typedef struct { double r, t1, t2; sphfem_double2 p; } sphfem_circle_segment;
/* the host state of class tool (tool.h:109-138): a closed polygon of 4 segments, or 5 when a
 * fillet replaces the third one (has_fillet), velocity, friction coefficient and front */
typedef struct {
    int num_seg, has_fillet;
    sphfem_segment seg[SPHFEM_TOOL_MAX_SEG];
    sphfem_circle_segment fillet;
    sphfem_double2 vel;
    double mu, front;
} sphfem_tool_geometry;
/* bbox (geometric_primitives.h:59-72) */
typedef struct { double bbmin_x, bbmax_x, bbmin_y, bbmax_y; } sphfem_bbox;

/* Host-only geometry (no device work, no context): the constructors of tool.cpp:316-406,
 * tool(tl, tr, br, bl[, r], mu) and tool(tl, length, height, rake, clearance[, r], mu) with
 * angles in degrees; r = 0 builds the sharp tool.  Velocity starts at zero. */
int sphfem_tool_geometry_from_points(sphfem_double2 tl, sphfem_double2 tr, sphfem_double2 br, sphfem_double2 bl,
                                     double r, double mu, sphfem_tool_geometry *out);
int sphfem_tool_geometry_from_angles(sphfem_double2 tl, double length, double height, double rake_angle,
                                     double clearance_angle, double r, double mu, sphfem_tool_geometry *out);
/* tool::update_tool (tool.cpp:202-212): ends and fillet centre move by dt * vel, lines rebuilt */
int sphfem_tool_geometry_move(sphfem_tool_geometry *g, double dt);
/* tool::low (tool.cpp:178-188), tool::safe_bb (:147-176), tool::inside (:95-110) */
int sphfem_tool_geometry_low(const sphfem_tool_geometry *g, double *low);
int sphfem_tool_geometry_safe_bb(const sphfem_tool_geometry *g, double safety, sphfem_bbox *out);
int sphfem_tool_geometry_inside(const sphfem_tool_geometry *g, sphfem_double2 pos, int *inside);
/* tool::sample_tool (tool.cpp:112-135): lattice points of spacing dx inside the tool, column by
 * column.  Writes at most cap points to out (may be NULL) and the full count to *count. */
int sphfem_tool_geometry_sample(const sphfem_tool_geometry *g, double dx, sphfem_double2 *out, long cap,
                                long *count);

typedef struct sphfem_rigid_tool sphfem_rigid_tool;
/* tool_gpu_set_up_tool (tool_gpu.cu:219-266): the device copy of the geometry, velocity and mu,
 * with contact_alpha and slave_mass = phys->mass */
int sphfem_tool_gpu_set_up_tool(sphfem_ctx *ctx, const sphfem_tool_geometry *g, double contact_alpha,
                                const sphfem_phys_constants *phys, sphfem_rigid_tool **out);
int sphfem_rigid_tool_destroy(sphfem_ctx *ctx, sphfem_rigid_tool *rt);
/* the device geometry as it stands (blocks until the stream is idle) */
int sphfem_rigid_tool_get_geometry(sphfem_ctx *ctx, sphfem_rigid_tool *rt, sphfem_tool_geometry *out);
/* tool::set_vel (tool.cpp:215-217) on the device copy */
int sphfem_rigid_tool_set_vel(sphfem_ctx *ctx, sphfem_rigid_tool *rt, sphfem_double2 vel);
/* tool_gpu_update_tool (tool_gpu.cu:270-311): the segments and fillet move by global_dt * vel on
 * the device (the reference moves its host copy and uploads it) and the tool particles of p by
 * global_dt * vel (update_tool_particles, :205-216).  p may be NULL (geometry only). */
int sphfem_tool_gpu_update_tool(sphfem_ctx *ctx, sphfem_rigid_tool *rt, sphfem_particles *p);
/* compute_contact_forces (tool_gpu.cu:82-202, 314-319): penalty contact and LS-DYNA-style
 * Coulomb friction against the polygon and fillet, for every non-tool particle */
int sphfem_compute_contact_forces(sphfem_ctx *ctx, sphfem_particles *p, sphfem_rigid_tool *rt);
/* material_fric_heat_gen (actions_gpu.cu:316-343, 956-965): T += eta dt |ft| |v_rel| / (cp m)
 * for non-tool particles; a no-op unless the workpiece's eta is non-zero */
int sphfem_material_fric_heat_gen(sphfem_ctx *ctx, sphfem_particles *p, sphfem_double2 vel);
/* interactions_setup_thermal_constants_tool (interactions_gpu.cu:494-502): the thermal block of
 * tool particles in the pair sums.  Brookshaw conduction reads the device block, which is
 * written only when alpha != 0; the PSE variant reads the host copy, always written.  Both start
 * at zero (tool particles then conduct no heat). */
int sphfem_interactions_setup_thermal_constants_tool(sphfem_ctx *ctx, const sphfem_trml_constants *c);

/* ---- time integration (leap_frog.h:46-49, leap_frog.cu:25-96, 162-248) */
int sphfem_leap_frog_create(sphfem_ctx *ctx, unsigned int num_part, unsigned int num_cell, int node_num,
                            sphfem_leap_frog **out);
int sphfem_leap_frog_destroy(sphfem_ctx *ctx, sphfem_leap_frog *lf);
int sphfem_leap_frog_set_precision(sphfem_ctx *ctx, sphfem_leap_frog *lf, int precision);
/* The FP32C form keeps the derivative carry (pos_t, vel_t, S_t, rho_t, T_t) inside the stepper in
 * FP32 (the pair sums it comes from are FP32; positions, velocities, stresses, density and
 * temperature are integrated in FP64 from it): its steps neither read nor write those five arrays
 * of sphfem_particles.  export_rates writes the current carry into them (sorted order, as after
 * the reference's step); import_rates reads them in after a caller changed them between steps.
 * The first FP32C step imports them, and the first FP64 step after FP32C steps exports them, by
 * itself.  Both calls do nothing in the FP64 form.  No counterpart in the reference. */
int sphfem_leap_frog_export_rates(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p);
int sphfem_leap_frog_import_rates(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p);
int sphfem_leap_frog_init(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p);
int sphfem_leap_frog_predict(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, double dt);
int sphfem_leap_frog_correct(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, double dt);
/* Launch parameters of the step kernels, chosen by the autotuner (sph_fem_2d_cutting_gpu_amd/
 * autotune.py): pair kernel block size (128 or 256, both forms), FP32C neighbour-list chunk (16,
 * 32 or 48 entries per lane), update kernel block size (128 or 256).  Other values are rejected.
 * Results do not depend on them beyond the FP32 summation blocks of the FP32C form.  No
 * counterpart in the reference (fixed launch shapes there). */
/* pair_kernel (FP32C): SPHFEM_PAIRS_LIST gathers every neighbour record through L1;
 * SPHFEM_PAIRS_STAGED first copies each wavefront's window positions into LDS (pair_block unused,
 * one wavefront per workgroup).  At the same pair_chunk both give the same sums bit for bit.
 * The _FUSED forms are the same kernels running each particle's update right after its sums (no
 * pair-sum round trip through memory, one launch fewer; update_block unused); the same step
 * result bit for bit, with the FEM tool and with a rigid tool.  The FP64 form always runs its
 * pair and update kernels separately (pair_kernel does not apply to it). */
enum { SPHFEM_PAIRS_LIST = 0, SPHFEM_PAIRS_STAGED = 1, SPHFEM_PAIRS_LIST_FUSED = 2, SPHFEM_PAIRS_STAGED_FUSED = 3 };
/* resort_one_block: the step's full-sort fallback (first step, or more particles changed cells than
 * the incremental re-sort takes) by one workgroup in one launch instead of the multi-block radix
 * sort's per-pass launches, which are queued every step and return at once on the incremental
 * path; the same permutation.  Only for N <= 2^19 (rejected above). */
typedef struct { int pair_block, pair_chunk, update_block, pair_kernel, resort_one_block; } sphfem_tuning;
int sphfem_leap_frog_set_tuning(sphfem_ctx *ctx, sphfem_leap_frog *lf, const sphfem_tuning *tuning);
int sphfem_leap_frog_get_tuning(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_tuning *tuning);
/* leap_frog.cu:162-248: one full step (fused kernels); tool may be NULL for a tool-free workload.
 * With a Rothlin grid the step sorts as grid_gpu_rothlin::sort does (ft and n not reordered, idx
 * carrying labels) and then runs the same fused kernels over the sorted arrays. */
int sphfem_leap_frog_step(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, sphfem_grid *g,
                          sphfem_tool *t);
/* leap_frog::step(particles, grid) (leap_frog.cu:99-160) with global_tool = rt: the same fused
 * kernels, with friction heat after the corrector, then plasticity, the tool moved by global_dt *
 * vel (segments, fillet, tool particles) and contact against the moved tool.  rt may be NULL (no
 * tool: the step of leap_frog.cu:99-160 with global_tool = 0). */
int sphfem_leap_frog_step_rigid(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, sphfem_grid *g,
                                sphfem_rigid_tool *rt);

/* Step graphs: `steps` (even) steps of sphfem_leap_frog_step recorded into one HIP graph and
 * replayed with sphfem_leap_frog_graph_launch (each launch advances `steps` steps).  Capture
 * records only (the state does not advance) and needs a locked grid and, for a Green grid, one
 * eager step first; the graph bakes in
 * the constants, dt, wear switch, precision and tuning at capture, and a launch after any of them
 * changed (or after the tool was remeshed onto a new mesh) is refused with SPHFEM_ERR_STATE.  No counterpart in the reference
 * (its step is a sequence of launches with host synchronisations). */
typedef struct sphfem_graph sphfem_graph;
int sphfem_leap_frog_graph_create(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p, sphfem_grid *g,
                                  sphfem_tool *t, int steps, sphfem_graph **out);
/* the same for sphfem_leap_frog_step_rigid (rt may be NULL) */
int sphfem_leap_frog_graph_create_rigid(sphfem_ctx *ctx, sphfem_leap_frog *lf, sphfem_particles *p,
                                        sphfem_grid *g, sphfem_rigid_tool *rt, int steps, sphfem_graph **out);
/* Independent cases (a parameter sweep) in one HIP graph: `steps` (even) steps of every case k --
 * its own context ctxs[k] (stream and constants), stepper, particles, locked grid and FEM tool (ts
 * may be NULL, or hold NULL entries, for tool-free cases) -- recorded as parallel branches of a
 * single graph, which sphfem_leap_frog_graph_launch(ctxs[0], ...) replays: one launch advances all
 * cases.  Every case needs its own context and one eager step first; the same settings checks as
 * for a step graph apply to every case.  No counterpart in the reference. */
int sphfem_batch_graph_create(sphfem_ctx *const *ctxs, sphfem_leap_frog *const *lfs, sphfem_particles *const *ps,
                              sphfem_grid *const *gs, sphfem_tool *const *ts, int n, int steps, sphfem_graph **out);
int sphfem_leap_frog_graph_launch(sphfem_ctx *ctx, sphfem_graph *graph, int repeats);
int sphfem_leap_frog_graph_destroy(sphfem_ctx *ctx, sphfem_graph *graph);

/* ---- observables (output.cu:22-30) */
int sphfem_report_force(sphfem_ctx *ctx, sphfem_particles *p, sphfem_double2 *out);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops): the sums of sphfem_report_force, bit for bit, written
 * asynchronously on the context's stream to `out`, a device pointer to 2 doubles, without
 * allocating anything.  The block partials live in a fixed device array of 256 x 2 doubles per GPU
 * (4 KiB, held until the process ends), shared by every call on that GPU: calls that overlap on
 * one GPU (two contexts on different streams) race.  N above 2^30 is rejected (SPHFEM_ERR_ARG). */
int sphfem_report_force_noalloc(sphfem_ctx *ctx, const sphfem_particles *p, sphfem_double2 *out);

/* ---- output formats (vtk_writer.cu, tool_wear.cu, mesh_structure_GPU.cu:1259-1296).
 * Legacy ASCII VTK with the reference's headers, field names, order, scaling and printf
 * formats (%f, %d), text identical to what vtk_writer.cu writes for the same values. */
/* vtk_writer_write(particles, step) (vtk_writer.cu:11-219) to `path`: points, vertex cells,
 * Temperature, on_seg, EquivAccumPlasticStrain, Velocity, fc and ft (x 1e5), normal, SvM.  The
 * reference never downloads `blanked` and so writes every particle; skip_blanked = 1 leaves
 * blanked particles out (cells then index the written points).  Downloads and synchronises. */
int sphfem_vtk_write_particles(sphfem_ctx *ctx, const sphfem_particles *p, const char *path, int skip_blanked);
/* the same writer over host arrays laid out as the device ones (no device access; the fields
 * read are pos, vel, p, S, fc, ft, n, blanked, eps_pl, T, on_seg) */
int sphfem_vtk_write_particles_host(const sphfem_particles *host, const char *path, int skip_blanked);
/* vtk_writer_write(tool, step) (vtk_writer.cu:222-302): the rigid tool as triangles (two or
 * three for the body, a 19-triangle fan for the fillet); nothing is written for a tool without
 * segments.  Host only. */
int sphfem_vtk_write_rigid_tool(const sphfem_tool_geometry *g, const char *path);
/* vtk_writer_write_FEM_tool (vtk_writer.cu:304-366): mesh and nodal temperature */
int sphfem_vtk_write_FEM_tool(sphfem_ctx *ctx, sphfem_tool *t, const char *path);
/* vtk_writer_write_tool_org (vtk_writer.cu:518-587): the mesh shifted back by tl - tl_ref */
int sphfem_vtk_write_tool_org(sphfem_ctx *ctx, sphfem_tool *t, sphfem_double2 tl_ref, const char *path);
/* vtk_writer_write_FEM_segments (vtk_writer.cu:368-516): one triangle per cutting segment with
 * contact pressure (GPa), sliding velocity (m/s), temperature, frictional stress and wear rate */
int sphfem_vtk_write_FEM_segments(sphfem_ctx *ctx, sphfem_tool *t, const char *path);
/* vtk_writer_write_contact_variables (vtk_writer.cu:589-714): segment midpoints with the same
 * variables; as in the reference, Wear_rate carries three values per point */
int sphfem_vtk_write_contact_variables(sphfem_ctx *ctx, sphfem_tool *t, const char *path);
/* mesh_read::mesh_read_temperature (mesh_structure_GPU.cu:1259-1296): the numbers on every
 * non-empty line after the first "LOOKUP_TABLE default" line of a VTK file (a file written by
 * sphfem_vtk_write_FEM_tool).  Writes at most cap values to T (may be NULL) and the count to
 * *count; SPHFEM_ERR_ARG if a line is not a number (the reference throws). */
int sphfem_read_vtk_temperature(const char *path, double *T, long cap, long *count);

/* tool_wear (tool_wear.h:17-47, tool_wear.cu:14-72): Usui wear rate K exp(alpha / T) p |v_r| of
 * the particles in contact (p = |fc| / (2 pi h)), min / max / mean over them, and the wear
 * accumulated as dt * mean per call.  Evaluated by a deterministic device reduction. */
typedef struct sphfem_tool_wear sphfem_tool_wear;
int sphfem_tool_wear_create(sphfem_ctx *ctx, double usui_K, double usui_alpha, sphfem_double2 tool_vel,
                            sphfem_tool_wear **out);
int sphfem_tool_wear_destroy(sphfem_ctx *ctx, sphfem_tool_wear *w);
/* tool_wear::eval_usui; blocks until the result is on the host */
int sphfem_tool_wear_eval_usui(sphfem_ctx *ctx, sphfem_tool_wear *w, const sphfem_particles *p, double dt,
                               double *usui_min, double *usui_max, double *usui_avg);
int sphfem_tool_wear_get_accum_wear(const sphfem_tool_wear *w, double *accum);

/* ---- tool mesh (mesh_structure_GPU.cu, tool_FEM.cu:771-1039, FEM_thermal_gpu.cu:126-330, 584-659).
 * Host-side: the .inp reader, element colouring, CSR pattern, mesh VTK file, cutting segments and
 * the FEM system, ending in a sphfem_tool_desc for sphfem_tool_create. */
typedef struct sphfem_mesh sphfem_mesh;
typedef struct {
    int node_num, ele_num, nz_num, cut_ele_num, fixed_ele_num, color_total, color_cutting_edge;
} sphfem_mesh_info;
/* mesh_read(file) (mesh_structure_GPU.cu:784-1257, gmsh_export 0: Abaqus-style *Node / *Element /
 * CuttingEdge and FixedBoundary node and element sets, coordinates divided by 100) or
 * mesh_read(file, true) (:213-783, gmsh_export 1: Gmsh export with T3D2 boundary edges and CPS3
 * triangles, nodes outside the triangles dropped and renumbered).  Colours the cutting-edge
 * elements from a palette of color_cutting_edge colours, then boundary layer 2, then the rest from
 * color_total colours (mesh_structure_GPU.h:26-27 uses 18 and 6), each with the first free colour
 * counted from 1; builds the CSR pattern.  SPHFEM_ERR_ARG on an unreadable file, a malformed line
 * or a palette too small for the mesh (the reference asserts). */
int sphfem_mesh_read_inp(const char *path, int gmsh_export, int color_total, int color_cutting_edge,
                         sphfem_mesh **out);
int sphfem_mesh_destroy(sphfem_mesh *m);
int sphfem_mesh_get_info(const sphfem_mesh *m, sphfem_mesh_info *info);
/* copies of the mesh arrays (any pointer may be NULL): position [node_num], cut/fixed node flags
 * [node_num], elements (3 nodes + colour) and cut/fixed edge flags (x: edge 0-1, y: 1-2, z: 2-0)
 * [ele_num], cut/fixed element lists, boundary layers 1 and 2 [2 ele_num], CSR row_ptr
 * [node_num + 1] and col_ind [nz_num] */
int sphfem_mesh_get_arrays(const sphfem_mesh *m, sphfem_double2 *position, int *cut_node, int *fixed_node,
                           sphfem_int4 *elem, sphfem_int4 *cut_flags, sphfem_int4 *fixed_flags, int *cut_elems,
                           int *fixed_elems, int *layers, int *row_ptr, int *col_ind);
/* mesh_read_temperature (mesh_structure_GPU.cu:1259-1296): initial nodal temperatures of a hot
 * tool, n == node_num; T NULL goes back to the uniform T0 */
int sphfem_mesh_set_temperature(sphfem_mesh *m, const double *T, long n);
/* write_mesh_file (mesh_structure_GPU.cu:1395-1462): mesh with colour and layer flags */
int sphfem_mesh_write_vtk(const sphfem_mesh *m, const char *path);
/* set_up_segments_from_mesh (tool_FEM.cu:771-1039) and the FEM system of FEM_thermal_2D_GPU
 * (FEM_thermal_gpu.cu:584-659): cutting segments oriented towards tl_point_ref and chained from
 * the highest point, P1 stiffness and lumped mass assembled colour by colour, Dirichlet rows at
 * tc->T0 on the fixed boundary.  The arrays `out` points to belong to the mesh object and stay
 * valid until the next call or sphfem_mesh_destroy. */
int sphfem_mesh_tool_desc(sphfem_mesh *m, const sphfem_tool_constants *tc, sphfem_double2 tl_point_ref,
                          sphfem_double2 vel, sphfem_tool_desc *out);

/* ---- tool remeshing (actions_gpu.cu:1022-1038).  Both need a tool created with node_flags and
 * fixed_elems / fixed_flags (SPHFEM_ERR_STATE otherwise). */
/* remesh_low_level(tool, freq): the nodal temperatures and positions are kept as the mapping
 * (tool_FEM.cu:1611-1622), the interior nodes are moved by mesh_shift(freq) (tool_FEM.cu:589-616),
 * K, M and f are assembled again colour by colour with the Dirichlet rows and the inverse mass
 * (FEM_thermal_gpu.cu:702-720, 797-847), and the temperatures are interpolated from the mapping
 * (tool_FEM.cu:57-100, 434-443).  Arrays keep their addresses: captured graphs stay valid. */
int sphfem_tool_remesh_low_level(sphfem_ctx *ctx, sphfem_tool *t, int freq);
/* read_remesh_gmsh (tool_FEM.cu:1811-1828) from a mesh already read (the .inp Gmsh writes from
 * the geometry of sphfem_tool_write_remesh_geo): nodal temperatures interpolated from the current
 * mesh at the new node positions (tool_FEM.cu:445-517; 300 where no element holds the node),
 * cutting segments rebuilt from the new mesh and oriented towards the current reference point
 * (tool_FEM.cu:1245-1539), the FEM system of the new mesh (FEM_thermal_gpu.cu:722-795).  The
 * tool's velocity, reference points and flank retreat are kept; lowest point, right point and
 * reference distance come from the new segments; the segment accumulators start at zero.  The
 * tool's device arrays are replaced: graphs captured before are refused afterwards. */
int sphfem_tool_remesh(sphfem_ctx *ctx, sphfem_tool *t, sphfem_mesh *mesh);
/* remesh_read_segments (tool_FEM.cu:1708-1810): the Gmsh geometry of the current cutting edge --
 * the reference point, the chain's points with mesh sizes of six segment lengths (a point is
 * skipped after a short edge), the closing lines, physical groups CuttingEdge (67) and
 * FixedBoundary (2) -- written as a .geo script for `gmsh -2 -format inp` */
int sphfem_tool_write_remesh_geo(sphfem_ctx *ctx, sphfem_tool *t, const char *path);
/* the tool's device arrays under the reference's struct views (mesh_structure_GPU.h:80-142 mesh_GPU_m,
 * FEM_thermal_gpu.h:35-93 FEM_thermal_2D_GPU, segment_FEM_gpu.h:31-56): device pointers owned by
 * the tool, valid until it is remeshed onto a new mesh or destroyed.  The mapping arrays and the
 * node flags are NULL for a tool created without the remeshing description; fixed_ele_num is 0
 * then. */
typedef struct {
    int node_num, ele_num, nz_num, seg_num, color_total, color_cutting_edge, fixed_ele_num;
    sphfem_double2 *position, *position_mapping;
    sphfem_int4 *node_num_el;
    int *csr_row_ptr, *csr_col_ind, *node_flags;
    double *K, *K_temp, *M, *M_inv, *f, *f_temp, *k, *T, *T_old, *T_mapping;
    sphfem_double2 *left, *right, *n, *heat_exchange;
    sphfem_int4 *marks;
    double *fric_heat;
    sphfem_double2 *wear_rate;
    double *wear_nodes;
    sphfem_double4 *physical_para;
    double *sliding_force;
} sphfem_tool_arrays;
int sphfem_tool_get_arrays(const sphfem_tool *t, sphfem_tool_arrays *out);
/* node, element, non-zero and segment counts of the tool's current mesh (tool_FEM.h:39-41 node_number,
 * element_number, cutting_segment_size); any pointer may be NULL */
int sphfem_tool_get_sizes(const sphfem_tool *t, int *node_num, int *ele_num, int *nz_num, int *seg_num);

/* ---- measurement: per-kernel HIP-event timing on the context stream */
int sphfem_timing_enable(sphfem_ctx *ctx, int enable);
int sphfem_timing_reset(sphfem_ctx *ctx);
/* restrict timing to a comma-separated list of kernel names (NULL or "" = all kernels) */
int sphfem_timing_select(sphfem_ctx *ctx, const char *names);
/* accumulated milliseconds and launch count of one kernel (names listed by sphfem_timing_names) */
int sphfem_timing_get(sphfem_ctx *ctx, const char *kernel, double *total_ms, long *launches);
const char *sphfem_timing_names(sphfem_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
