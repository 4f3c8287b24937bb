/*
 * FP64 host oracle of the SPH-FEM cutting step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the per-timestep particle hot path of
 * nyzhang1227/SPH_FEM_2D_cutting_gpu (FP64 build, USE_FEM_TOOL, WEAR_NODE_SHIFT,
 * USE_USUI_RAKE_FLANK, RESTRICT_WEAR_TEMP, SMOOTH_TEMP_WEAR,
 * Thermal_Conduction_Brookshaw; types.h:22-122).  Every function cites the
 * reference lines it follows.  It is the checker of the HIP library and the
 * CPU baseline of the benchmark; nothing in the product links or calls it.
 *
 * Arrays use the reference layout: double2 = {x, y}, double4 = {x, y, z, w}.
 */
#ifndef SPHFEM_ORACLE_H
#define SPHFEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double x, y; } o_d2;
typedef struct { double x, y, z, w; } o_d4;
typedef struct { int x, y, z, w; } o_i4;

/* constants_structs.h:18-123 (SHEAR_LIMIT/VELOCITY/TEMP/PRES friction models off) */
typedef struct { double E, nu, rho0, K, G, mass, mass_tool; } o_phys;
typedef struct { double cp, tq, k, alpha, T_init, eta; } o_trml;
typedef struct { double wdeltap, stresseps, xspheps, alpha, beta, eta; } o_corr;
typedef struct { double A, B, C, n, m, Tmelt, Tref, eps_dot_ref, tanh_a, tanh_b, tanh_c, tanh_d; } o_joco;
typedef struct { int nx, ny; double bbmin_x, bbmin_y, dx; } o_geom;
typedef struct { double cp, tq, k, eta, rho, T0, h, beta, mu, contact_alpha, slave_mass, hc; } o_toolc;
typedef struct { double A, B, C, D; } o_wearc;

typedef struct {
    o_phys phys; o_trml trml_wp; o_trml trml_tool; o_corr corr; o_joco joco;
    o_geom geom; o_toolc tc; o_wearc wear;
    double dt;
    double rr2_min;      /* kernels.cuh:38 lower cut-off (1e-8 in the reference) */
    int wear_pro;        /* mfree_gpu.cpp:300 */
    int thermal_bc;      /* actions_gpu.cu:938-943 (tool alpha != 0) */
    int plastic;         /* actions_gpu.cu:921-924 */
    int fric_heat;       /* actions_gpu.cu:931-934 (workpiece eta != 0) */
    o_trml trml_itool;   /* interactions_gpu.cu:19 (device block, written for alpha != 0) */
    double itool_alpha_pse;   /* interactions_gpu.cu:24 host copy's alpha (PSE launch, :453) */
} o_consts;

/* particle_gpu.h:23-83 */
typedef struct {
    int N;
    o_d2 *pos, *vel; double *h, *rho, *p; o_d4 *S, *R; o_d2 *fc, *ft, *n;
    double *fixed, *blanked, *tool_particle, *eps_pl, *eps_pl_dot, *T;
    o_d2 *pos_t, *vel_t; double *rho_t; o_d4 *S_t; double *T_t; o_d4 *S_der, *v_der;
    int *idx, *hash, *on_seg;
} o_particles;

/* leap_frog.h:28-38 */
typedef struct {
    o_d2 *pos_init, *vel_init; o_d4 *S_init; double *rho_init, *T_init;
    int *cell_start, *cell_end; int num_cell;
} o_stepper;

/* tool_FEM + FEM_thermal_2D_GPU + segment_FEM_gpu + mesh_GPU_m (per-step state) */
typedef struct {
    int n_nodes, n_elem, nnz, n_seg, n_colors, n_cut_colors;
    o_d2 *node_pos; o_i4 *elem; int *layer;          /* layer: element is in boundary layer 1 or 2 */
    int *row_ptr, *col_ind;
    double *K, *K_temp, *M, *Minv, *f, *f_temp, *T, *T_old, *k;
    o_d2 *left, *right, *n, *heat_exchange; double *fric_heat;
    o_d2 *wear_rate; double *wear_nodes; o_d4 *physical_para; double *sliding_force; o_i4 *marks;
    /* host-side kinematics of tool_FEM (tool_FEM.cu:1552-1570) */
    double vel_x, vel_y, tl_x, tl_y, br_x, br_y, ref_dist, lowest_point, right_point;
    double flank_retreat, flank_contact;   /* __device__ globals tool_FEM.cu:19-20 */
    /* remeshing: node flags (bit 0 cutting edge, bit 1 fixed), fixed-boundary elements and their
     * edge flags (mesh_structure_GPU.h:88-115), mapping buffers (tool_FEM.cu:1611-1622) */
    int *node_flags; int n_fixed; int *fixed_elems; o_i4 *fixed_flags;
    o_d2 *pos_map; double *T_map;
} o_tool;

/* rigid (SPH) tool: geometric_primitives.h:15-57, tool.h:109-138, tool_gpu.cu:16-22 */
typedef struct { double a, b; int vertical; } o_line;
typedef struct { o_d2 left, right; o_line l; o_d2 n; } o_seg;
typedef struct { double r, t1, t2; o_d2 p; } o_circ;
typedef struct {
    int num_seg, has_fillet;
    o_seg seg[5];
    o_circ fillet;
    o_d2 vel;
    double mu, front;
    double contact_alpha, slave_mass;    /* tool_gpu_set_up_tool (tool_gpu.cu:219-266) */
} o_rtool;

/* kernels.cuh:25-65; out[0..2] = w, w_x, w_y */
void o_cubic_spline(double pix, double piy, double pjx, double pjy, double hi, double rr2_min, double *out);
/* plasticity.cuh:21-39, :117-162 */
double o_sigma_yield(const o_joco *jc, double eps_pl, double eps_pl_dot, double t);
double o_solve_secant(const o_joco *jc, double init, double tol, double norm_Strial, double eps_pl_init,
                      double t, double dt, double mu, int *iters);

/* grid.cu:25-42 / grid_gpu_green.cu:13-115, 184-209 */
void o_compute_hashes(const o_consts *c, o_particles *P);
void o_sort_and_cells(const o_consts *c, o_particles *P, o_stepper *st);
/* grid_gpu_rothlin.cu:19-126, 129-249 */
void o_sort_rothlin(const o_consts *c, o_particles *P, o_stepper *st);

/* leap_frog.cu:25-96 */
void o_init(o_particles *P, o_stepper *st);
void o_predict(o_particles *P, o_stepper *st, double dt);
void o_correct(o_particles *P, o_stepper *st, double dt);

/* actions_gpu.cu */
void o_material_eos(const o_consts *c, o_particles *P);
void o_artificial_stress(const o_consts *c, o_particles *P);
void o_interactions_monaghan(const o_consts *c, o_particles *P, const o_stepper *st);
void o_interactions_monaghan_range(const o_consts *c, o_particles *P, const o_stepper *st, int i0, int i1);
void o_interactions_heat_pse(const o_consts *c, o_particles *P, const o_stepper *st);
void o_stress_rate_jaumann(const o_consts *c, o_particles *P);
void o_continuity(o_particles *P);
void o_momentum(const o_consts *c, o_particles *P);
void o_advection(o_particles *P);
void o_fric_heat_v21(const o_consts *c, o_particles *P, o_tool *t);
void o_plasticity(const o_consts *c, o_particles *P);
void o_contact_s1(const o_consts *c, o_particles *P, o_tool *t);
void o_contact_s2(const o_consts *c, o_particles *P, o_tool *t);
void o_boundary_conditions(o_particles *P);
void o_boundary_conditions_thermal(const o_consts *c, o_particles *P);
void o_invalidate(o_particles *P);

/* FEM_thermal_gpu.cu / tool_FEM.cu per-step methods */
void o_fem_reset_server(o_tool *t);
void o_fem_calculate_temp(o_tool *t, double dt);
void o_fem_apply_robin(const o_consts *c, o_tool *t);
void o_fem_calculate_temp_rate(o_tool *t);
void o_segments_wear_rate_reset(o_tool *t);
void o_segments_update(o_tool *t, double dt);
void o_tool_update(o_tool *t, double dt);
void o_wear_adjustment(const o_consts *c, o_tool *t);
/* remesh_low_level (actions_gpu.cu:1022-1031) */
void o_remesh_low_level(const o_consts *c, o_tool *t, int freq);
/* do_nodal_value_interpolation (tool_FEM.cu:445-517) from the tool's current mesh onto new_pos */
void o_interp_new(const o_tool *t, const o_d2 *new_pos, int n_new, double *T_new);

/* tool.cpp:316-406 constructors (r = 0: sharp), :202-212 update_tool, :178-188 low,
 * :147-176 safe_bb (out = bbmin_x, bbmax_x, bbmin_y, bbmax_y), :95-110 inside,
 * :112-135 sample_tool (writes up to cap points, returns the count) */
void o_rtool_from_points(o_rtool *t, o_d2 tl, o_d2 tr, o_d2 br, o_d2 bl, double r, double mu);
void o_rtool_from_angles(o_rtool *t, o_d2 tl, double length, double height, double rake, double clear, double r,
                         double mu);
void o_rtool_move(o_rtool *t, double dt);
double o_rtool_low(const o_rtool *t);
void o_rtool_safe_bb(const o_rtool *t, double safety, double *out);
int o_rtool_inside(const o_rtool *t, double x, double y);
long o_rtool_sample(const o_rtool *t, double dx, o_d2 *out, long cap);
/* tool_gpu.cu:82-134 for one point: 1 if inside, contact point and normal */
int o_rtool_establish_contact(const o_rtool *t, double qx, double qy, o_d2 *cp, o_d2 *n);
/* tool_gpu.cu:146-202, 205-216 and actions_gpu.cu:316-343 over all particles */
void o_rtool_contact(const o_consts *c, o_particles *P, const o_rtool *t);
void o_rtool_tool_particles(const o_consts *c, o_particles *P, const o_rtool *t);
void o_rtool_fric_heat(const o_consts *c, o_particles *P, const o_rtool *t);

/* output.cu:22-30 */
void o_report_force(const o_particles *P, double *fx, double *fy);

/* leap_frog.cu:162-248: one full step in the reference order */
void o_step(const o_consts *c, o_particles *P, o_stepper *st, o_tool *t);
void o_run(const o_consts *c, o_particles *P, o_stepper *st, o_tool *t, int nsteps);
void o_run_rothlin(const o_consts *c, o_particles *P, o_stepper *st, o_tool *t, int nsteps);
/* grid.cu:88-127 and a step on an unlocked grid */
int o_update_geometry(o_consts *c, const o_particles *P, double kernel_width, int max_cell);
int o_run_unlocked(o_consts *c, o_particles *P, o_stepper *st, o_tool *t, int nsteps, int max_cell, int rothlin);

/* leap_frog.cu:99-160: the step with a rigid tool (t may be NULL: no tool); rothlin selects the
 * Rothlin sort */
void o_run_rigid(const o_consts *c, o_particles *P, o_stepper *st, o_rtool *t, int nsteps, int rothlin);

int o_num_threads(void);
void o_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
