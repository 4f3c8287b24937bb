/*
 * FP64 host oracle of the SPH-FEM cutting step -- TEST INFRASTRUCTURE ONLY.
 * Checker of the HIP library and CPU baseline of the benchmark; never linked
 * into the product.  Each function restates the reference lines it cites, in
 * the same evaluation order, including the reference's FP32 rounding sites
 * (sqrtf / fmaxf / fminf / atan2f / cosf / sinf applied to doubles).
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp (no FMA contraction, so the
 * evaluation order written here is the one executed).
 *
 * Loops whose iterations write only their own particle run under OpenMP;
 * loops that accumulate into shared tool/segment arrays run sequentially in
 * index order (the reference uses atomicAdd there, whose order is unspecified).
 */
#include "sphfem_oracle.h"

#include <float.h>
#include <math.h>

#define O_PI 3.14159265358979323846   /* M_PI */
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_FAC 0.454728408833987            /* kernels.cuh:19 */
#define ORACLE_FLT_MAX ((double)FLT_MAX)

int o_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void o_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* float rounding of the reference's single-precision calls on doubles */
static inline double sqrtf_d(double x) { return (double)sqrtf((float)x); }
static inline double fmaxf_d(double a, double b) { return (double)fmaxf((float)a, (float)b); }
static inline double fminf_d(double a, double b) { return (double)fminf((float)a, (float)b); }

/* ------------------------------------------------------------------------- */
/* kernels.cuh:25-65 (cubic spline; the q>1 value is 0.25*(2-q)^6 as written) */
void o_cubic_spline(double pix, double piy, double pjx, double pjy, double hi, double rr2_min, double *w) {
    w[0] = 0.; w[1] = 0.; w[2] = 0.;
    double xij = pix - pjx;
    double yij = piy - pjy;
    double rr2 = xij * xij + yij * yij;
    double h1 = 1 / hi;
    double fourh2 = 4 * hi * hi;
    if (rr2 >= fourh2 || rr2 < rr2_min) return;
    double der, val;
    double rad = sqrtf_d(rr2);
    double q = rad * h1;
    double fac = PI_FAC * h1 * h1;
    if (q > 1) {
        double _2mq = 2 - q;
        double _2mq2 = _2mq * _2mq;
        val = 0.25 * _2mq2 * _2mq2 * _2mq2;
        der = -0.75 * _2mq2 * h1 / rad;
    } else {
        val = 1 - 1.5 * q * q * (1 - 0.5 * q);
        der = -3.0 * q * (1 - 0.75 * q) * h1 / rad;
    }
    w[0] = val * fac;
    w[1] = der * xij * fac;
    w[2] = der * yij * fac;
}

/* plasticity.cuh:21-39 */
double o_sigma_yield(const o_joco *jc, double eps_pl, double eps_pl_dot, double t) {
    double theta = (t - jc->Tref) / (jc->Tmelt - jc->Tref);
    if (theta > 1.) theta = 1.;
    double Term_A = jc->A + jc->B * pow(eps_pl, jc->n);
    double Term_B = 1.0;
    double eps_dot = eps_pl_dot / jc->eps_dot_ref;
    if (eps_dot > 1.0) Term_B = 1.0 + jc->C * log(eps_dot);
    else Term_B = pow((1.0 + eps_dot), jc->C);
    double Term_C = 1.0 - pow(theta, jc->m);
    return Term_A * Term_B * Term_C;
}

/* plasticity.cuh:117-126 */
static double opt_fun(const o_joco *jc, double delta_lambda, double norm_Strial, double eps_pl_init,
                      double t, double delta_t, double mu) {
    double eps_pl_equiv = eps_pl_init + sqrt(2.0 / 3.0) * fmax(delta_lambda, 0.);
    double eps_pl_equiv_dot = sqrt(2.0 / 3.0) * fmax(delta_lambda, 0.) / delta_t;
    double sigmaY = o_sigma_yield(jc, eps_pl_equiv, eps_pl_equiv_dot, t);
    return norm_Strial - 2 * mu * delta_lambda - sqrt(2.0 / 3.0) * sigmaY;
}

/* plasticity.cuh:128-162 */
double o_solve_secant(const o_joco *jc, double init, double tol, double norm_Strial, double eps_pl_init,
                      double t, double delta_t, double mu, int *iters) {
    double delta_lambda = init;
    double delta_lambda_old = delta_lambda;
    int iter = 0;
    int max_iter = 100;
    for (;;) {
        double g = opt_fun(jc, delta_lambda, norm_Strial, eps_pl_init, t, delta_t, mu);
        double delta_lambda_1 = delta_lambda * 1.01;
        double g_1 = opt_fun(jc, delta_lambda_1, norm_Strial, eps_pl_init, t, delta_t, mu);
        double slope = (g - g_1) / (delta_lambda - delta_lambda_1);
        delta_lambda_old = delta_lambda;
        delta_lambda = delta_lambda - g / slope;
        if (delta_lambda < 0.) delta_lambda = -0.1 * delta_lambda;
        if (fabs(delta_lambda - delta_lambda_old) < tol) { if (iters) *iters = iter; return delta_lambda; }
        if (iter > max_iter) { if (iters) *iters = iter; return delta_lambda; }
        iter++;
    }
}

/* ------------------------------------------------------------------------- */
/* grid.cu:25-42: hash of non-blanked particles (blanked keep their old hash) */
void o_compute_hashes(const o_consts *c, o_particles *P) {
    const double dx = c->geom.dx, bx = c->geom.bbmin_x, by = c->geom.bbmin_y;
    const int ny = c->geom.ny;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->blanked[i] == 1.) continue;
        int ix = (int)((P->pos[i].x - bx) / dx);
        int iy = (int)((P->pos[i].y - by) / dx);
        P->hash[i] = ix * ny + iy;
    }
}

#define GATHER(field, type)                                                   \
    do {                                                                      \
        type *tmp = (type *)malloc(sizeof(type) * (size_t)N);                 \
        for (int i = 0; i < N; i++) tmp[i] = P->field[P->idx[i]];             \
        memcpy(P->field, tmp, sizeof(type) * (size_t)N);                      \
        free(tmp);                                                            \
    } while (0)

/* grid_gpu_green.cu:184-218: stable sort of (hash, idx), cell_start = -1,
 * boundaries from neighbouring hashes, gather of the 20 state arrays
 * (:53-113), copy of the cell arrays into the stepper (get_cells). */
void o_sort_and_cells(const o_consts *c, o_particles *P, o_stepper *st) {
    const int N = P->N;
    const int num_cell = c->geom.nx * c->geom.ny;
    /* stable counting sort: order of equal keys is the current order */
    int *count = (int *)calloc((size_t)num_cell + 1, sizeof(int));
    for (int i = 0; i < N; i++) {
        int h = P->hash[i];
        if (h < 0 || h >= num_cell) abort();
        count[h + 1]++;
    }
    for (int k = 0; k < num_cell; k++) count[k + 1] += count[k];
    int *keys = (int *)malloc(sizeof(int) * (size_t)N);
    for (int i = 0; i < N; i++) {
        int h = P->hash[i];
        int dst = count[h]++;
        keys[dst] = h;
        P->idx[dst] = i;
    }
    memcpy(P->hash, keys, sizeof(int) * (size_t)N);
    free(keys);
    free(count);

    for (int k = 0; k < num_cell; k++) st->cell_start[k] = -1;
    for (int i = 0; i < N; i++) {
        int h = P->hash[i];
        if (i == 0 || h != P->hash[i - 1]) {
            st->cell_start[h] = i;
            if (i > 0) st->cell_end[P->hash[i - 1]] = i;
        }
        if (i == N - 1) st->cell_end[h] = i + 1;
    }
    GATHER(pos, o_d2); GATHER(vel, o_d2); GATHER(fc, o_d2); GATHER(ft, o_d2);
    GATHER(h, double); GATHER(n, o_d2); GATHER(rho, double); GATHER(S, o_d4);
    GATHER(eps_pl, double); GATHER(eps_pl_dot, double); GATHER(T, double);
    GATHER(fixed, double); GATHER(blanked, double); GATHER(tool_particle, double);
    GATHER(pos_t, o_d2); GATHER(vel_t, o_d2); GATHER(rho_t, double); GATHER(S_t, o_d4);
    GATHER(T_t, double); GATHER(on_seg, int);
}

/* grid_gpu_rothlin.cu:129-249 and :19-126.  The sort: stable sort of the keys with the
 * sequence as values, then in-place gathers of idx (so idx keeps the particle label rather than
 * the permutation), pos, vel, fc, h, rho, S, eps_pl, eps_pl_dot, fixed, blanked, tool_particle, T,
 * pos_t, vel_t, rho_t, S_t, T_t and on_seg; ft and n are not gathered (no gather for them in the
 * reference).  The reference's on_seg line copies the gathered idx buffer instead of the gathered
 * on_seg (:244-245); on_seg is gathered here.  The cells: unique_by_key_copy + differences +
 * scatter + inclusive scan with the last cell fixed up (:19-78) give cell offsets equal to the
 * number of keys below each cell, so cell_start[c] = #{keys < c}, cell_end[c] = #{keys < c+1}:
 * dense, with start == end for empty cells. */
void o_sort_rothlin(const o_consts *c, o_particles *P, o_stepper *st) {
    const int N = P->N;
    const int num_cell = c->geom.nx * c->geom.ny;
    int *count = (int *)calloc((size_t)num_cell + 1, sizeof(int));
    for (int i = 0; i < N; i++) {
        int h = P->hash[i];
        if (h < 0 || h >= num_cell) abort();
        count[h + 1]++;
    }
    for (int k = 0; k < num_cell; k++) count[k + 1] += count[k];
    /* count[k] = #{keys < k}: the dense offsets */
    for (int k = 0; k < num_cell; k++) {
        st->cell_start[k] = count[k];
        st->cell_end[k] = count[k + 1];
    }
    int *keys = (int *)malloc(sizeof(int) * (size_t)N);
    int *perm = (int *)malloc(sizeof(int) * (size_t)N);
    int *label = (int *)malloc(sizeof(int) * (size_t)N);
    for (int i = 0; i < N; i++) {
        int h = P->hash[i];
        int dst = count[h]++;
        keys[dst] = h;
        perm[dst] = i;
    }
    memcpy(P->hash, keys, sizeof(int) * (size_t)N);
    for (int i = 0; i < N; i++) label[i] = P->idx[perm[i]];
    /* GATHER reads the permutation from P->idx */
    memcpy(P->idx, perm, sizeof(int) * (size_t)N);
    GATHER(pos, o_d2); GATHER(vel, o_d2); GATHER(fc, o_d2); GATHER(h, double);
    GATHER(rho, double); GATHER(S, o_d4); GATHER(eps_pl, double); GATHER(eps_pl_dot, double);
    GATHER(fixed, double); GATHER(blanked, double); GATHER(tool_particle, double); GATHER(T, double);
    GATHER(pos_t, o_d2); GATHER(vel_t, o_d2); GATHER(rho_t, double); GATHER(S_t, o_d4);
    GATHER(T_t, double); GATHER(on_seg, int);
    memcpy(P->idx, label, sizeof(int) * (size_t)N);
    free(label);
    free(perm);
    free(keys);
    free(count);
}

/* leap_frog.cu:25-37 */
void o_init(o_particles *P, o_stepper *st) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->blanked[i] == 1.) continue;
        st->pos_init[i] = P->pos[i];
        st->vel_init[i] = P->vel[i];
        st->S_init[i] = P->S[i];
        st->rho_init[i] = P->rho[i];
        st->T_init[i] = P->T[i];
    }
}

/* leap_frog.cu:45-69 */
void o_predict(o_particles *P, o_stepper *st, double dt) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->blanked[i] == 1.) continue;
        P->pos[i].x = st->pos_init[i].x + 0.5 * dt * P->pos_t[i].x;
        P->pos[i].y = st->pos_init[i].y + 0.5 * dt * P->pos_t[i].y;
        P->vel[i].x = st->vel_init[i].x + 0.5 * dt * P->vel_t[i].x;
        P->vel[i].y = st->vel_init[i].y + 0.5 * dt * P->vel_t[i].y;
        P->S[i].x = st->S_init[i].x + 0.5 * dt * P->S_t[i].x;
        P->S[i].y = st->S_init[i].y + 0.5 * dt * P->S_t[i].y;
        P->S[i].z = st->S_init[i].z + 0.5 * dt * P->S_t[i].z;
        P->S[i].w = st->S_init[i].w + 0.5 * dt * P->S_t[i].w;
        P->rho[i] = st->rho_init[i] + 0.5 * dt * P->rho_t[i];
        P->T[i] = st->T_init[i] + 0.5 * dt * P->T_t[i];
    }
}

/* leap_frog.cu:71-96 */
void o_correct(o_particles *P, o_stepper *st, double dt) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->blanked[i] == 1.) continue;
        P->pos[i].x = st->pos_init[i].x + dt * P->pos_t[i].x;
        P->pos[i].y = st->pos_init[i].y + dt * P->pos_t[i].y;
        P->vel[i].x = st->vel_init[i].x + dt * P->vel_t[i].x;
        P->vel[i].y = st->vel_init[i].y + dt * P->vel_t[i].y;
        P->S[i].x = st->S_init[i].x + dt * P->S_t[i].x;
        P->S[i].y = st->S_init[i].y + dt * P->S_t[i].y;
        P->S[i].z = st->S_init[i].z + dt * P->S_t[i].z;
        P->S[i].w = st->S_init[i].w + dt * P->S_t[i].w;
        P->rho[i] = st->rho_init[i] + dt * P->rho_t[i];
        P->T[i] = st->T_init[i] + dt * P->T_t[i];
    }
}

/* actions_gpu.cu:222-232 */
void o_material_eos(const o_consts *c, o_particles *P) {
    const double rho0 = c->phys.rho0;
    const double c0 = sqrtf_d(c->phys.K / rho0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->tool_particle[i] != 0.) continue;
        P->p[i] = c0 * c0 * (P->rho[i] - rho0);
    }
}

/* actions_gpu.cu:45-49, 234-278 */
void o_artificial_stress(const o_consts *c, o_particles *P) {
    const double eps = c->corr.stresseps;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->tool_particle[i] != 0.) continue;
        double rhoi = P->rho[i];
        double pi = P->p[i];
        o_d4 Si = P->S[i];
        double sxx = Si.x, sxy = Si.y, syy = Si.z;
        sxx -= pi;
        syy -= pi;
        double rhoi21 = 1. / (rhoi * rhoi);
        double numer = 2. * sxy;
        double denom = sxx - syy + 0.;
        double theta = 0.5 * (double)atan2f((float)numer, (float)denom);
        double cos_theta = (double)cosf((float)theta);
        double sin_theta = (double)sinf((float)theta);
        double cos_theta2 = cos_theta * cos_theta;
        double sin_theta2 = sin_theta * sin_theta;
        double rot_sxx = cos_theta2 * sxx + 2.0 * cos_theta * sin_theta * sxy + sin_theta2 * syy;
        double rot_syy = sin_theta2 * sxx - 2.0 * cos_theta * sin_theta * sxy + cos_theta2 * syy;
        double rot_rxx = 0., rot_ryy = 0.;
        if (rot_sxx > 0) rot_rxx = -eps * rot_sxx * rhoi21;
        if (rot_syy > 0) rot_ryy = -eps * rot_syy * rhoi21;
        o_d4 Ri;
        Ri.x = cos_theta2 * rot_rxx + sin_theta2 * rot_ryy;
        Ri.y = cos_theta * sin_theta * (rot_rxx - rot_ryy);
        Ri.z = sin_theta2 * rot_rxx + cos_theta2 * rot_ryy;
        Ri.w = 0.;
        P->R[i] = Ri;
    }
}

/* interactions_gpu.cu:84-140, 434-464 with lapl_pse of kernels.cuh:67-84: particle strength
 * exchange heat conduction (compiled out of the reference's flag set, kept as an entry point).
 * Every particle of the 5x5 cells (no support cut-off, blanked neighbours included),
 * T_t = alpha * sum (T_j - T_i) 4/(pi h^4) exp(-|x_ij|^2/h^2) m/rho_j, alpha of the tool or the
 * workpiece block.  Off when the workpiece is not thermal. */
void o_interactions_heat_pse(const o_consts *c, o_particles *P, const o_stepper *st) {
    if (c->trml_wp.alpha == 0.) return;
    const double mass = c->phys.mass;
    const int nx = c->geom.nx, ny = c->geom.ny;
#pragma omp parallel for schedule(dynamic, 128)
    for (int pidx = 0; pidx < P->N; pidx++) {
        if (P->blanked[pidx] == 1.) continue;
        o_d2 pi = P->pos[pidx];
        double hi = P->h[pidx], Ti = P->T[pidx];
        double alpha = (P->tool_particle[pidx] == 1.) ? c->itool_alpha_pse : c->trml_wp.alpha;
        int hashi = P->hash[pidx];
        int gi = hashi / ny;
        int gj = hashi - gi * ny;
        int low_i = gi - 2 < 0 ? 0 : gi - 2;
        int low_j = gj - 2 < 0 ? 0 : gj - 2;
        int high_i = gi + 3 > nx ? nx : gi + 3;
        int high_j = gj + 3 > ny ? ny : gj + 3;
        double T_ti = 0.;
        for (int ii = low_i; ii < high_i; ii++) {
            for (int jj = low_j; jj < high_j; jj++) {
                int idx = ii * ny + jj;
                int c_start = st->cell_start[idx];
                if (c_start == -1) continue;
                int c_end = st->cell_end[idx];
                for (int iter = c_start; iter < c_end; iter++) {
                    o_d2 pj = P->pos[iter];
                    double Tj = P->T[iter], rhoj = P->rho[iter];
                    double xij = pi.x - pj.x, yij = pi.y - pj.y;
                    double xx = sqrt(xij * xij + yij * yij);
                    double h2 = hi * hi;
                    double h4 = h2 * h2;
                    double w2 = +4. / (h4 * O_PI) * exp(-xx * xx / (h2));
                    T_ti += (Tj - Ti) * w2 * mass / rhoj;
                }
            }
        }
        P->T_t[pidx] = alpha * T_ti;
    }
}

/* interactions_gpu.cu:143-379 (Brookshaw on, SURFACE off); neighbours in
 * the reference's cell order ii (column), jj (row), then sorted index. */
void o_interactions_monaghan(const o_consts *c, o_particles *P, const o_stepper *st) {
    o_interactions_monaghan_range(c, P, st, 0, P->N);
}

/* the same sums for the sorted particles i0 <= pidx < i1 only (sub-windows of large inputs) */
void o_interactions_monaghan_range(const o_consts *c, o_particles *P, const o_stepper *st, int i0, int i1) {
    const double mass = c->phys.mass, K = c->phys.K;
    const double wdeltap = c->corr.wdeltap, alpha = c->corr.alpha, beta = c->corr.beta;
    const double eta = c->corr.eta, eps = c->corr.xspheps;
    const int nx = c->geom.nx, ny = c->geom.ny;
    const double rr2_min = c->rr2_min;
#pragma omp parallel for schedule(dynamic, 128)
    for (int pidx = i0; pidx < i1; pidx++) {
        if (P->blanked[pidx] == 1.) continue;
        double is_tool_i = P->tool_particle[pidx];
        /* interactions_gpu.cu:164: the tool block (:19) is zero unless interactions_setup_thermal_constants_tool
         * set it (benchmarks.cpp:518; the FEM setups do not, benchmarks_NZ.cpp:354) */
        double thermal_alpha = (is_tool_i == 0.) ? c->trml_wp.alpha : c->trml_itool.alpha;
        o_d2 pi = P->pos[pidx], vi = P->vel[pidx];
        o_d4 Si = P->S[pidx], Ri = P->R[pidx];
        double hi = P->h[pidx], rhoi = P->rho[pidx], prsi = P->p[pidx];
        double Ti = 0.;
        if (thermal_alpha != 0.) Ti = P->T[pidx];
        double rhoi21 = 1. / (rhoi * rhoi);
        int hashi = P->hash[pidx];
        int gi = hashi / ny;
        int gj = hashi - gi * ny;
        int low_i = gi - 2 < 0 ? 0 : gi - 2;
        int low_j = gj - 2 < 0 ? 0 : gj - 2;
        int high_i = gi + 3 > nx ? nx : gi + 3;
        int high_j = gj + 3 > ny ? ny : gj + 3;
        o_d4 vi_der = {0., 0., 0., 0.}, Si_der = {0., 0., 0., 0.};
        o_d2 vi_t = {0., 0.}, xi_t = {0., 0.};
        double T_lapl = 0.;
        for (int ii = low_i; ii < high_i; ii++) {
            for (int jj = low_j; jj < high_j; jj++) {
                int idx = ii * ny + jj;
                int c_start = st->cell_start[idx];
                int c_end = st->cell_end[idx];
                if (c_start == -1) continue;
                for (int iter = c_start; iter < c_end; iter++) {
                    if (P->blanked[iter] == 1.) continue;
                    o_d2 pj = P->pos[iter], vj = P->vel[iter];
                    o_d4 Sj = P->S[iter], Rj = P->R[iter];
                    double hj = P->h[iter], rhoj = P->rho[iter], prsj = P->p[iter];
                    double Tj = 0.;
                    if (thermal_alpha != 0.) Tj = P->T[iter];
                    double is_tool_j = P->tool_particle[iter];
                    double volj = mass / rhoj;
                    double rhoj21 = 1. / (rhoj * rhoj);
                    double ww[3];
                    o_cubic_spline(pi.x, pi.y, pj.x, pj.y, hi, rr2_min, ww);
                    double w = ww[0], w_x = ww[1], w_y = ww[2];
                    if (!(is_tool_i == 1.) && !(is_tool_j == 1.)) {
                        vi_der.x += (vj.x - vi.x) * w_x * volj;
                        vi_der.y += (vj.x - vi.x) * w_y * volj;
                        vi_der.z += (vj.y - vi.y) * w_x * volj;
                        vi_der.w += (vj.y - vi.y) * w_y * volj;
                        double fab = w / wdeltap;
                        fab *= fab;
                        fab *= fab;
                        double Rx = fab * (Ri.x + Rj.x);
                        double Ry = fab * (Ri.y + Rj.y);
                        double Rz = fab * (Ri.z + Rj.z);
                        Si_der.x += mass * ((Si.x - prsi) * rhoi21 + (Sj.x - prsj) * rhoj21 + Rx) * w_x;
                        Si_der.y += mass * (Si.y * rhoi21 + Sj.y * rhoj21 + Ry) * w_y;
                        Si_der.z += mass * (Si.y * rhoi21 + Sj.y * rhoj21 + Ry) * w_x;
                        Si_der.w += mass * ((Si.z - prsi) * rhoi21 + (Sj.z - prsj) * rhoj21 + Rz) * w_y;
                        double xij = pi.x - pj.x;
                        double yij = pi.y - pj.y;
                        double vijx = vi.x - vj.x;
                        double vijy = vi.y - vj.y;
                        double vijposij = vijx * xij + vijy * yij;
                        double rhoij = 0.5 * (rhoi + rhoj);
                        if (vijposij < 0.) {
                            double ci = sqrtf_d(K / rhoi);
                            double cj = sqrtf_d(K / rhoj);
                            double cij = 0.5 * (ci + cj);
                            double hij = 0.5 * (hi + hj);
                            double r2ij = xij * xij + yij * yij;
                            double muij = (hij * vijposij) / (r2ij + eta * eta * hij * hij);
                            double piij = (-alpha * cij * muij + beta * muij * muij) / rhoij;
                            vi_t.x += -mass * piij * w_x;
                            vi_t.y += -mass * piij * w_y;
                        }
                        xi_t.x += -eps * w * mass / rhoij * vijx;
                        xi_t.y += -eps * w * mass / rhoij * vijy;
                    }
                    if (thermal_alpha != 0.) {
                        double xij = pi.x - pj.x;
                        double yij = pi.y - pj.y;
                        double rij = sqrt(xij * xij + yij * yij);
                        if (rij > 1e-8) {
                            double eijx = xij / rij;
                            double eijy = yij / rij;
                            double rij1 = 1. / rij;
                            T_lapl += 2.0 * (mass / rhoj) * (Ti - Tj) * rij1 * (eijx * w_x + eijy * w_y);
                        }
                    }
                }
            }
        }
        P->S_der[pidx] = Si_der;
        P->v_der[pidx] = vi_der;
        P->pos_t[pidx] = xi_t;
        P->vel_t[pidx] = vi_t;
        if (thermal_alpha != 0.) P->T_t[pidx] = thermal_alpha * T_lapl;
    }
}

/* actions_gpu.cu:280-314 (GLM 3x3 products reduced to their single non-zero terms) */
void o_stress_rate_jaumann(const o_consts *c, o_particles *P) {
    const double G = c->phys.G;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->tool_particle[i] == 1.) continue;
        o_d4 d = P->v_der[i], S = P->S[i];
        double vx_x = d.x, vx_y = d.y, vy_x = d.z, vy_y = d.w;
        double Sxx = S.x, Sxy = S.y, Syy = S.z;
        double exy = 0.5 * (vx_y + vy_x);
        double w01 = 0.5 * (vx_y - vy_x);   /* omega row 0, column 1 */
        double w10 = 0.5 * (vy_x - vx_y);   /* omega row 1, column 0 */
        double tr = vx_x + vy_y + 0.;
        double t3 = (1. / 3.) * tr;
        double g2 = 2. * G;
        P->S_t[i].x = (g2 * (vx_x - t3) + w01 * Sxy) + Sxy * w01;
        P->S_t[i].y = (g2 * (exy - 0.) + w10 * Sxx) + Syy * w01;
        P->S_t[i].z = (g2 * (vy_y - t3) + w10 * Sxy) + Sxy * w10;
        P->S_t[i].w = g2 * (0. - t3);
    }
}

/* actions_gpu.cu:626-639 */
void o_continuity(o_particles *P) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->tool_particle[i] != 0.) continue;
        P->rho_t[i] = -P->rho[i] * (P->v_der[i].x + P->v_der[i].w);
    }
}

/* actions_gpu.cu:641-671 */
void o_momentum(const o_consts *c, o_particles *P) {
    const double mass = c->phys.mass;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->tool_particle[i] != 0.) continue;
        o_d4 sd = P->S_der[i];
        P->vel_t[i].x += sd.x + sd.y + P->fc[i].x / mass + P->ft[i].x / mass;
        P->vel_t[i].y += sd.z + sd.w + P->fc[i].y / mass + P->ft[i].y / mass;
    }
}

/* actions_gpu.cu:673-687 */
void o_advection(o_particles *P) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->tool_particle[i] != 0.) continue;
        P->pos_t[i].x = P->pos_t[i].x + P->vel[i].x;
        P->pos_t[i].y = P->pos_t[i].y + P->vel[i].y;
    }
}

/* actions_gpu.cu:345-420 (WEAR_NODE_SHIFT on); segment sums in particle order */
void o_fric_heat_v21(const o_consts *c, o_particles *P, o_tool *t) {
    const double dt = c->dt;
    const double vtx = t->vel_x, vty = t->vel_y;
    for (int i = 0; i < P->N; i++) {
        if (P->blanked[i] == 1.) continue;
        if (P->eps_pl[i] > 20.) continue;
        o_d2 f_T = P->ft[i];
        double f_fric_mag = sqrtf_d(f_T.x * f_T.x + f_T.y * f_T.y);
        int seg = P->on_seg[i];
        if (seg == -1) continue;
        const double partition_wp = c->tc.beta;
        o_d2 normal = P->n[i];
        o_d2 vp = P->vel[i];
        double vdx = vp.x - vtx, vdy = vp.y - vty;
        double vdn = vdx * normal.x + vdy * normal.y;
        double vrx = vdx - vdn * normal.x, vry = vdy - vdn * normal.y;
        double v_rel_mag = sqrtf_d(vrx * vrx + vry * vry);
        double value = c->tc.eta * f_fric_mag * v_rel_mag;
        if (isnan(v_rel_mag)) { P->blanked[i] = 1.; continue; }
        o_d2 f_C = P->fc[i];
        double f_normal = sqrtf_d(f_C.x * f_C.x + f_C.y * f_C.y);
        double rho_ = P->rho[i];
        double length = sqrtf_d(c->tc.slave_mass / rho_);
        double pressure_trial = f_normal / length;
        if (v_rel_mag > 0.03) continue;
        double T_buffer = P->T[i] + value * (1 - partition_wp) * dt / c->trml_wp.cp / c->phys.mass;
        if (T_buffer > 1550.) { P->T[i] = 1550.; continue; }
        P->T[i] = T_buffer;
        t->fric_heat[seg] += value * partition_wp;
        int left_no = t->marks[seg].z, right_no = t->marks[seg].w;
        double T_seg = 0.5 * (t->T[left_no] + t->T[right_no]);
        double vel_tool_norm = sqrt(vtx * vtx + vty * vty);
        if (v_rel_mag > 0.008) continue;
        else if (v_rel_mag > vel_tool_norm * 8.0 && fabs(normal.y) > 0.5) continue;
        t->physical_para[seg].x += 1.;
        t->physical_para[seg].y += pressure_trial;
        t->physical_para[seg].z += v_rel_mag;
        t->physical_para[seg].w += T_seg;
        t->sliding_force[seg] += f_fric_mag / length;
    }
}

/* actions_gpu.cu:689-754 */
void o_plasticity(const o_consts *c, o_particles *P) {
    const double mu = c->phys.G, dt = c->dt;
    const o_joco *jc = &c->joco;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < P->N; i++) {
        if (P->tool_particle[i] != 0.) continue;
        if (P->blanked[i] == 1.) continue;
        o_d4 S = P->S[i];
        double Sxx = S.x, Syy = S.z, Szz = S.w, Sxy = S.y;
        double norm_Strial = sqrtf_d(Sxx * Sxx + Syy * Syy + Szz * Szz + 2. * Sxy * Sxy);
        double p = P->p[i];
        double cxx = Sxx - p, cyy = Syy - p, czz = Szz - p, cxy = Sxy;
        double eps_pl = P->eps_pl[i], eps_pl_dot = P->eps_pl_dot[i], T = P->T[i];
        double svm = sqrtf_d((cxx * cxx + cyy * cyy + czz * czz) - cxx * cyy - cxx * czz - cyy * czz + 3.0 * cxy * cxy);
        double sigma_Y = o_sigma_yield(jc, eps_pl, eps_pl_dot, T);
        if (svm < sigma_Y) { P->eps_pl_dot[i] = 0.; continue; }
        double delta_lambda = o_solve_secant(jc, fmax(eps_pl_dot * dt * sqrt(2. / 3.), 1e-8), 1e-6,
                                             norm_Strial, eps_pl, T, dt, c->phys.G, NULL);
        double s23 = (double)sqrtf(2.0f / 3.0f);
        double dl = (double)fmaxf((float)delta_lambda, 0.f);
        double eps_pl_new = eps_pl + s23 * dl;
        double eps_pl_dot_new = s23 * dl / dt;
        P->eps_pl[i] = eps_pl_new;
        P->eps_pl_dot[i] = eps_pl_dot_new;
        o_d4 Sn;
        Sn.x = Sxx - Sxx / norm_Strial * delta_lambda * 2. * mu;
        Sn.z = Syy - Syy / norm_Strial * delta_lambda * 2. * mu;
        Sn.w = Szz - Szz / norm_Strial * delta_lambda * 2. * mu;
        Sn.y = Sxy - Sxy / norm_Strial * delta_lambda * 2. * mu;
        P->S[i] = Sn;
        if (c->trml_wp.tq != 0.) {
            double delta_eps_pl = eps_pl_new - eps_pl;
            double sY = o_sigma_yield(jc, eps_pl_new, eps_pl_dot_new, T);
            double rho = P->rho[i];
            P->T[i] += c->trml_wp.tq / (c->trml_wp.cp * rho) * delta_eps_pl * sY;
        }
        if (P->T[i] > jc->Tmelt) { P->T[i] = jc->Tmelt; P->blanked[i] = 1.; }
    }
}

/* actions_gpu.cu:33-43 */
static inline double find_max(double a, double b) { return a >= b ? a : b; }
static inline double find_min(double a, double b) { return a >= b ? b : a; }

/* actions_gpu.cu:51-63 ("abs" on a double read as fabs) */
static int on_segment_mesh(o_d2 left, o_d2 right, o_d2 point) {
    if (point.y <= fmaxf_d(left.y, right.y) && point.y >= fminf_d(left.y, right.y)) {
        if (left.x == right.x && point.x < left.x) return 1;
        if (fabs(left.y - right.y) != 0) {
            double x_on_seg = right.x + (left.x - right.x) / (left.y - right.y) * (point.y - right.y);
            if (x_on_seg > point.x && x_on_seg != left.x) return 1;
        } else {
            if (point.x <= fmax(left.x, right.x)) return 1;
        }
    }
    return 0;
}

/* actions_gpu.cu:83-93 */
static int inside_judgement(o_d2 qp, const o_d2 *left, const o_d2 *right, int num_seg) {
    int intersections = 0;
    for (int i = 0; i < num_seg; i++) intersections += on_segment_mesh(left[i], right[i], qp);
    return (intersections % 2 == 0) ? 0 : 1;
}

/* actions_gpu.cu:139-220 */
static int find_shortest_distance_v1(o_d2 qp, double *dN, double *nre_x, double *nre_y, const o_d2 *left,
                                     const o_d2 *right, const o_d2 *n, int num_seg) {
    int seg = -1;
    double depth = ORACLE_FLT_MAX;
    const double limit_value = 2.5e-4;
    for (int i = 0; i < num_seg; i++) {
        double ymax = find_max(left[i].y, right[i].y);
        double ymin = find_min(left[i].y, right[i].y);
        double xmax = find_max(left[i].x, right[i].x);
        double xmin = find_min(left[i].x, right[i].x);
        if (qp.x > xmin - 0.0008 && qp.x < xmax + 0.0008 && qp.y > ymin - 0.0008 && qp.y < ymax + 0.0008) {
            double d = (qp.x - left[i].x) * n[i].x + (qp.y - left[i].y) * n[i].y;
            if (d >= 0. && d < depth) {
                double ix = qp.x - d * n[i].x;
                double iy = qp.y - d * n[i].y;
                if (iy <= ymax && iy >= ymin && ix <= xmax && ix >= xmin) { depth = d; seg = i; }
            }
        }
    }
    if (seg != -1 && depth <= limit_value) {
        *nre_x = -n[seg].x;
        *nre_y = -n[seg].y;
        *dN = -depth;
    } else {
        depth = ORACLE_FLT_MAX;
        for (int i = 0; i < num_seg; i++) {
            double d_trial = sqrtf_d((qp.x - right[i].x) * (qp.x - right[i].x) + (qp.y - right[i].y) * (qp.y - right[i].y));
            if (d_trial < depth) { depth = d_trial; seg = i; }
        }
        if (seg == -1) return seg;
        double nsx = n[seg].x, nsy = n[seg].y;
        for (int j = 0; j < num_seg; j++) {
            if (left[j].y == right[seg].y && left[j].x == right[seg].x) {
                nsx = (n[j].x + n[seg].x) / 2.;
                nsy = (n[j].y + n[seg].y) / 2.;
                break;
            }
        }
        double n_length = sqrtf_d(nsx * nsx + nsy * nsy);
        *nre_x = -nsx / n_length;
        *nre_y = -nsy / n_length;
        *dN = -1.0 * depth;
    }
    return seg;
}

/* actions_gpu.cu:423-583 (friction models off); all particles are covered
 * (the reference grid of :988 skips a tail when N mod 256 is in [1, 128]). */
void o_contact_s1(const o_consts *c, o_particles *P, o_tool *t) {
    const double dt = c->dt;
    const double m = c->tc.slave_mass, ca = c->tc.contact_alpha, cof = c->tc.mu;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < P->N; i++) {
        if (P->blanked[i] == 1.) continue;
        o_d2 pos = P->pos[i], vel = P->vel[i];
        double ddx = pos.x - t->tl_x, ddy = pos.y - t->tl_y;
        int inside = 0;
        if (sqrt(ddx * ddx + ddy * ddy) < 1.05 * t->ref_dist)
            inside = inside_judgement(pos, t->left, t->right, t->n_seg);
        if (inside == 0) {
            P->fc[i].x = 0.; P->fc[i].y = 0.; P->ft[i].x = 0.; P->ft[i].y = 0.;
            P->n[i].x = 0.; P->n[i].y = 0.; P->on_seg[i] = -1;
            continue;
        }
        double gN = 0., nx_ = 0., ny_ = 0.;
        int seg_no = find_shortest_distance_v1(pos, &gN, &nx_, &ny_, t->left, t->right, t->n, t->n_seg);
        P->on_seg[i] = seg_no;
        if (seg_no == -1 || gN < -0.00025) {
            if (seg_no != -1) P->blanked[i] = 1.;
            P->fc[i].x = 0.; P->fc[i].y = 0.; P->ft[i].x = 0.; P->ft[i].y = 0.;
            P->n[i].x = 0.; P->n[i].y = 0.; P->on_seg[i] = -1;
            continue;
        }
        double dt2 = dt * dt;
        double fNx = -m * gN * nx_ / dt2 * ca;
        double fNy = -m * gN * ny_ / dt2 * ca;
        double vx = vel.x - t->vel_x, vy = vel.y - t->vel_y;
        double vdn = vx * nx_ + vy * ny_;
        double vrx = vx - nx_ * vdn, vry = vy - ny_ * vdn;
        double fox = P->ft[i].x, foy = P->ft[i].y;
        double kx = ca * m * vrx / dt, ky = ca * m * vry / dt;
        double fy = cof * sqrt(fNx * fNx + fNy * fNy);
        double fsx = fox - kx, fsy = foy - ky;
        double lfs = sqrt(fsx * fsx + fsy * fsy);
        double fTx, fTy;
        if (lfs > fy) { fTx = fy * fsx / lfs; fTy = fy * fsy / lfs; }
        else { fTx = fsx; fTy = fsy; }
        P->fc[i].x = fNx; P->fc[i].y = fNy;
        P->ft[i].x = fTx; P->ft[i].y = fTy;
        P->n[i].x = nx_; P->n[i].y = ny_;
    }
}

/* actions_gpu.cu:585-617; heat_exchange sums in particle order */
void o_contact_s2(const o_consts *c, o_particles *P, o_tool *t) {
    const double dt = c->dt;
    for (int i = 0; i < P->N; i++) {
        if (P->blanked[i] == 1.) continue;
        int seg_no = P->on_seg[i];
        if (seg_no == -1) continue;
        int left_no = t->marks[seg_no].z, right_no = t->marks[seg_no].w;
        double T_ = P->T[i];
        double rho_ = P->rho[i];
        double T_seg = 0.5 * (t->T[left_no] + t->T[right_no]);
        double delta_T = T_seg - T_;
        double length = sqrtf_d(c->tc.slave_mass / rho_);
        t->heat_exchange[seg_no].x += T_ * length;
        t->heat_exchange[seg_no].y += length;
        double dTp = delta_T * c->tc.h * length / c->trml_wp.cp / c->tc.slave_mass * dt;
        if (dTp < 0) P->T[i] += fmax(dTp, delta_T);
        else P->T[i] += fmin(dTp, delta_T);
    }
}

/* actions_gpu.cu:765-781 */
void o_boundary_conditions(o_particles *P) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->tool_particle[i] != 0.) continue;
        if (P->fixed[i] != 0.) {
            P->vel[i].x = 0.; P->vel[i].y = 0.;
            P->fc[i].x = 0.; P->fc[i].y = 0.;
            P->pos_t[i].x = 0.; P->pos_t[i].y = 0.;
            P->vel_t[i].x = 0.; P->vel_t[i].y = 0.;
        }
    }
}

/* actions_gpu.cu:756-763 */
void o_boundary_conditions_thermal(const o_consts *c, o_particles *P) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++)
        if (P->fixed[i] == 1.) P->T[i] = c->trml_wp.T_init;
}

static inline int isnaninf(double v) { return isnan(v) || isinf(v); }

/* actions_gpu.cu:787-824 */
void o_invalidate(o_particles *P) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->blanked[i] == 1.) continue;
        int bad = isnaninf(P->pos_t[i].x) || isnaninf(P->pos_t[i].y) || isnaninf(P->vel_t[i].x) ||
                  isnaninf(P->vel_t[i].y) || isnaninf(P->S_t[i].x) || isnaninf(P->S_t[i].y) ||
                  isnaninf(P->S_t[i].z) || isnaninf(P->S_t[i].w) || isnaninf(P->rho_t[i]) ||
                  isnaninf(P->T_t[i]);
        if (bad) P->blanked[i] = 1.;
    }
}

/* ------------------------------------------------------------------------- */
/* FEM_thermal_gpu.cu:902-921 */
void o_fem_reset_server(o_tool *t) {
    memcpy(t->T_old, t->T, sizeof(double) * (size_t)t->n_nodes);
    memcpy(t->f_temp, t->f, sizeof(double) * (size_t)t->n_nodes);
    memcpy(t->K_temp, t->K, sizeof(double) * (size_t)t->nnz);
}

/* FEM_thermal_gpu.cu:460-468, 891-900 */
void o_fem_calculate_temp(o_tool *t, double dt) {
    for (int i = 0; i < t->n_nodes; i++) t->T[i] = dt * t->k[i] + t->T_old[i];
}

static int diag_index(const o_tool *t, int row) {
    for (int k = t->row_ptr[row]; k < t->row_ptr[row + 1]; k++)
        if (t->col_ind[k] == row) return k;
    return -1;
}

/* FEM_thermal_gpu.cu:342-384, 865-877: colour by colour, segment order */
void o_fem_apply_robin(const o_consts *c, o_tool *t) {
    for (int colour = 1; colour < t->n_cut_colors + 1; colour++) {
        for (int s = 0; s < t->n_seg; s++) {
            if (t->marks[s].y != colour) continue;
            if (t->heat_exchange[s].y == 0) continue;
            double T_cont_particle = t->heat_exchange[s].x / t->heat_exchange[s].y;
            double fric_eng = t->fric_heat[s];
            int l = t->marks[s].z, r = t->marks[s].w;
            double ddx = t->left[s].x - t->right[s].x, ddy = t->left[s].y - t->right[s].y;
            double dist = sqrtf_d(ddx * ddx + ddy * ddy);
            t->f_temp[l] += 0.5 * c->tc.h * dist * T_cont_particle + 0.5 * fric_eng;
            t->f_temp[r] += 0.5 * c->tc.h * dist * T_cont_particle + 0.5 * fric_eng;
            double em = 0.5 * dist;
            int kl = diag_index(t, l), kr = diag_index(t, r);
            if (kl >= 0) t->K_temp[kl] += c->tc.h * em;
            if (kr >= 0) t->K_temp[kr] += c->tc.h * em;
        }
    }
}

/* FEM_thermal_gpu.cu:443-458 */
void o_fem_calculate_temp_rate(o_tool *t) {
    for (int i = 0; i < t->n_nodes; i++) {
        double buffer = 0.;
        for (int k = t->row_ptr[i]; k < t->row_ptr[i + 1]; k++) buffer += t->K_temp[k] * t->T[t->col_ind[k]];
        t->k[i] = t->Minv[i] * (t->f_temp[i] - buffer);
    }
}

/* tool_FEM.cu:1595-1609 */
void o_segments_wear_rate_reset(o_tool *t) {
    for (int s = 0; s < t->n_seg; s++) {
        t->wear_rate[s].x = 0.; t->wear_rate[s].y = 0.;
        t->wear_nodes[s] = 0.;
        t->physical_para[s].x = 0.; t->physical_para[s].y = 0.; t->physical_para[s].z = 0.; t->physical_para[s].w = 0.;
        t->sliding_force[s] = 0.;
    }
}

/* tool_FEM.cu:119-139, 1572-1582 */
void o_segments_update(o_tool *t, double dt) {
    for (int s = 0; s < t->n_seg; s++) {
        double lx = t->left[s].x, ly = t->left[s].y, rx = t->right[s].x, ry = t->right[s].y;
        t->left[s].x = lx + t->vel_x * dt;
        t->left[s].y = ly + t->vel_y * dt + t->flank_retreat;
        t->right[s].x = rx + t->vel_x * dt;
        t->right[s].y = ry + t->vel_y * dt + t->flank_retreat;
        t->heat_exchange[s].x = 0; t->heat_exchange[s].y = 0;
        t->fric_heat[s] = 0.;
    }
}

/* tool_FEM.cu:109-117, 1552-1570 */
void o_tool_update(o_tool *t, double dt) {
    t->tl_x += t->vel_x * dt; t->tl_y += t->vel_y * dt;
    t->br_x += t->vel_x * dt; t->br_y += t->vel_y * dt;
    t->lowest_point += t->vel_y * dt;
    t->right_point += t->vel_x * dt;
    for (int i = 0; i < t->n_nodes; i++) {
        t->node_pos[i].x += t->vel_x * dt;
        t->node_pos[i].y += t->vel_y * dt + t->flank_retreat;
    }
    t->flank_contact = t->lowest_point;
}

/* FEM_thermal_gpu.cu:33-109: lumped mass and P1 stiffness (K[a][b] = row a, column b) */
static void element_matrices(const o_tool *t, int e, const o_toolc *tc, double *mass, double Ke[3][3]) {
    o_d2 p0 = t->node_pos[t->elem[e].x], p1 = t->node_pos[t->elem[e].y], p2 = t->node_pos[t->elem[e].z];
    double j00 = p1.x - p0.x, j01 = p1.y - p0.y;
    double j10 = p2.x - p0.x, j11 = p2.y - p0.y;
    double det = fabs(j00 * j11 - j10 * j01);
    double w = 1. / 3. * 0.5;
    *mass = w * tc->cp * tc->rho * det;
    double dd = j00 * j11 - j01 * j10;
    double ood = 1. / dd;
    double i00 = j11 * ood, i01 = -j10 * ood, i10 = -j01 * ood, i11 = j00 * ood;
    static const double gref[3][2] = {{-1., -1.}, {1., 0.}, {0., 1.}};
    double G[3][2];
    for (int a = 0; a < 3; a++) {
        G[a][0] = i00 * gref[a][0] + i10 * gref[a][1];
        G[a][1] = i01 * gref[a][0] + i11 * gref[a][1];
    }
    double s = (w * 3.) * (tc->k * det);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) Ke[a][b] = (s * G[a][0]) * G[b][0] + (s * G[a][1]) * G[b][1];
}

/* FEM_thermal_gpu.cu:476-577, 923-951: +/- boundary-layer element matrices, colour by colour */
static void mesh_adjust_wear(const o_consts *c, o_tool *t, double sign) {
    for (int colour = 1; colour < t->n_colors + 1; colour++) {
        for (int e = 0; e < t->n_elem; e++) {
            if (!t->layer[e]) continue;
            if (t->elem[e].w != colour) continue;
            double mass, Ke[3][3];
            element_matrices(t, e, &c->tc, &mass, Ke);
            int nodes[3] = {t->elem[e].x, t->elem[e].y, t->elem[e].z};
            for (int a = 0; a < 3; a++) t->M[nodes[a]] += sign * mass;
            for (int a = 0; a < 3; a++) {
                int row = nodes[a];
                int cnt = 3;
                for (int k = t->row_ptr[row]; k < t->row_ptr[row + 1]; k++) {
                    int col = t->col_ind[k];
                    if (col == nodes[0]) { t->K[k] += sign * Ke[a][0]; cnt--; continue; }
                    if (col == nodes[1]) { t->K[k] += sign * Ke[a][1]; cnt--; continue; }
                    if (col == nodes[2]) { t->K[k] += sign * Ke[a][2]; cnt--; continue; }
                    if (cnt == 0) break;
                }
            }
        }
    }
}

static double o_cross(o_d2 a, o_d2 b) { return a.x * b.y - a.y * b.x; }

static int o_find(const o_tool *t, int row, int col) {
    int k = t->row_ptr[row];
    while (k < t->row_ptr[row + 1] && t->col_ind[k] != col) k++;
    return k;
}

/* mesh_shift (tool_FEM.cu:589-616) of one interior node */
static void o_mesh_shift_node(const o_tool *t, int freq, o_d2 *pp) {
    o_d2 p = *pp;
    double shift = t->flank_retreat * 1.0 * freq;
    shift = fmin(shift, -0.0000065);
    double wy = (p.x - t->tl_x) / (t->br_x - t->tl_x);
    double wx = fmin(1.0, (p.y - t->tl_y) / (t->br_y - t->tl_y));
    double c = 0.94;
    if (wx <= c)
        p.x += 1.45 * shift * (pow(c, 0.7) - pow(fabs(wx - c), 0.55)) * wy * wy * wy;
    else if (wx < 1.1 && wx > c)
        p.x += 1.45 * shift * (pow(c, 0.7) - pow(fabs(wx - c), 0.55)) * wy * wy;
    if (wy < 1.01)
        p.y += -0.9 * shift * fmin(pow(wy, 6), 1.) * pow(wx, 6);
    else if (wy >= 0.99)
        p.y += -0.9 * shift * fmin(pow(wy, 6), 1.) * wx * wx * wx * wx;
    *pp = p;
}

/* actions_gpu.cu:1022-1031: mapping_data_copy, mesh_shift, reset_FEM_matrix,
 * stiffness_mass_matrix_construction (colour by colour), dirichlet_bdc (colour by colour, fixed
 * element by fixed element: FEM_thermal_gpu.cu:234-324), inverse_mass_matrix,
 * nodal_value_interpolation (the layer variant, tool_FEM.cu:57-100, 434-443) */
void o_remesh_low_level(const o_consts *c, o_tool *t, int freq) {
    const int nn = t->n_nodes;
    for (int i = 0; i < nn; i++) {
        t->pos_map[i] = t->node_pos[i];
        t->T_map[i] = t->T[i];
    }
    for (int i = 0; i < nn; i++)
        if (!(t->node_flags[i] & 3)) o_mesh_shift_node(t, freq, &t->node_pos[i]);
    for (int k = 0; k < t->nnz; k++) t->K[k] = 0.;
    for (int i = 0; i < nn; i++) t->M[i] = t->f[i] = 0.;
    int ncol = 0;
    for (int e = 0; e < t->n_elem; e++) ncol = t->elem[e].w > ncol ? t->elem[e].w : ncol;
    for (int colour = 1; colour <= ncol; colour++)
        for (int e = 0; e < t->n_elem; e++) {
            if (t->elem[e].w != colour) continue;
            double mass, Ke[3][3];
            element_matrices(t, e, &c->tc, &mass, Ke);
            int nd[3] = {t->elem[e].x, t->elem[e].y, t->elem[e].z};
            for (int a = 0; a < 3; a++) {
                t->M[nd[a]] += mass;
                for (int b = 0; b < 3; b++) t->K[o_find(t, nd[a], nd[b])] += Ke[a][b];
            }
        }
    const double T0 = c->tc.T0;
    for (int colour = 1; colour <= ncol; colour++)
        for (int j = 0; j < t->n_fixed; j++) {
            o_i4 el = t->elem[t->fixed_elems[j]];
            if (el.w != colour) continue;
            o_i4 fl = t->fixed_flags[j];
            int nf[3] = {0, 0, 0};
            if (fl.x) nf[0] = nf[1] = 1;
            if (fl.y) nf[1] = nf[2] = 1;
            if (fl.z) nf[2] = nf[0] = 1;
            int nd[3] = {el.x, el.y, el.z};
            for (int q = 0; q < 3; q++) {
                if (!nf[q]) continue;
                int node = nd[q];
                for (int k = t->row_ptr[node]; k < t->row_ptr[node + 1]; k++) {
                    int loc = t->col_ind[k];
                    t->f[loc] += -t->K[k] * T0;
                    t->K[k] = 0.;
                    t->K[o_find(t, loc, node)] = 0.;
                    if (loc == node) t->K[k] = 1.0;
                }
                t->M[node] = 1.0;
                t->f[node] = T0;
                t->T_old[node] = T0;
            }
        }
    for (int i = 0; i < nn; i++) t->Minv[i] = 1. / t->M[i];
    for (int i = 0; i < nn; i++) {
        o_d2 x = t->node_pos[i];
        double T = t->T_map[i];
        for (int e = 0; e < t->n_elem; e++) {
            o_i4 el = t->elem[e];
            o_d2 xA = t->pos_map[el.x];
            o_d2 xB = {t->pos_map[el.y].x - xA.x, t->pos_map[el.y].y - xA.y};
            o_d2 xC = {t->pos_map[el.z].x - xA.x, t->pos_map[el.z].y - xA.y};
            o_d2 xP = {x.x - xA.x, x.y - xA.y};
            o_d2 xBC = {xB.x - xC.x, xB.y - xC.y};
            double d = o_cross(xB, xC);
            if (fabs(d) > 1e-13) {
                double wA = o_cross(xB, xC) + o_cross(xP, xBC);
                double wB = o_cross(xP, xC);
                double wC = -o_cross(xP, xB);
                if (wA >= 0 && wB >= 0 && wC >= 0 && wA <= 1 && wB <= 1 && wC <= 1) {
                    T = (t->T_map[el.x] * wA + t->T_map[el.y] * wB + t->T_map[el.z] * wC) / d;
                    break;
                }
            }
        }
        t->T[i] = T;
    }
}

void o_interp_new(const o_tool *t, const o_d2 *new_pos, int n_new, double *T_new) {
    for (int i = 0; i < n_new; i++) {
        o_d2 x = new_pos[i];
        for (int e = 0; e < t->n_elem; e++) {
            o_i4 el = t->elem[e];
            o_d2 xA = t->node_pos[el.x], xB = t->node_pos[el.y], xC = t->node_pos[el.z];
            o_d2 pts[3] = {xA, xB, xC};
            int ids[3] = {el.x, el.y, el.z}, hit = -1;
            for (int q = 0; q < 3 && hit < 0; q++) {
                double dx = x.x - pts[q].x, dy = x.y - pts[q].y;
                if ((double)sqrtf((float)(dx * dx + dy * dy)) < 1e-10) hit = ids[q];
            }
            if (hit >= 0) {
                T_new[i] = t->T[hit];
                break;
            }
            o_d2 xAC = {xA.x - xC.x, xA.y - xC.y}, xBA = {xB.x - xA.x, xB.y - xA.y};
            o_d2 xCB = {xC.x - xB.x, xC.y - xB.y};
            o_d2 xPC = {x.x - xC.x, x.y - xC.y}, xPA = {x.x - xA.x, x.y - xA.y}, xPB = {x.x - xB.x, x.y - xB.y};
            double d = o_cross(xBA, xAC);
            if (fabs(d) > 0) {
                double wC = o_cross(xBA, xPA) / d;
                double wA = o_cross(xCB, xPB) / d;
                double wB = o_cross(xAC, xPC) / d;
                if ((wA >= -1e-8 && wB >= -1e-8 && wC >= -1e-8) || (wA <= 1e-8 && wB <= 1e-8 && wC <= 1e-8)) {
                    double TA = t->T[el.x], TB = t->T[el.y], TC = t->T[el.z];
                    double Tn = fabs(TA * wA + TB * wB + TC * wC);
                    double lo = fmin(fmin(TA, TB), TC);
                    if (Tn < lo) Tn = lo;
                    double hi = fmax(fmax(TA, TB), TC);
                    if (Tn > hi) Tn = hi;
                    T_new[i] = Tn;
                    break;
                }
            }
        }
    }
}

/* actions_gpu.cu:1009-1020 -> FEM_thermal_gpu.cu:923-951, tool_FEM.cu:299-432 (USUI_RAKE_FLANK,
 * RESTRICT_WEAR_TEMP), :520-548, :559-587, FEM_thermal_gpu.cu:326-332 */
void o_wear_adjustment(const o_consts *c, o_tool *t) {
    const double dt = c->dt;
    mesh_adjust_wear(c, t, -1.0);
    for (int s = 0; s < t->n_seg; s++) {
        o_d4 wear = t->physical_para[s];
        if (wear.x == 0.) continue;
        int left = t->marks[s].z, right = t->marks[s].w;
        double pressure = wear.y / wear.x;
        double vel = wear.z / wear.x;
        double tl = t->T[left], tr = t->T[right];
        if (vel > 0.008) continue;
        tl = fmin(tl, 1350.);
        tr = fmin(tr, 1350.);
        vel = fmin(vel, 0.008);
        pressure = fmin(pressure, 0.04);
        double wl, wr;
        if (fabs(t->n[s].x) > 0.85) {
            wl = dt * c->wear.C * pressure * vel * exp(-c->wear.D / tl);
            wr = dt * c->wear.C * pressure * vel * exp(-c->wear.D / tr);
        } else {
            wl = dt * c->wear.A * pressure * vel * exp(-c->wear.B / tl);
            wr = dt * c->wear.A * pressure * vel * exp(-c->wear.B / tr);
        }
        t->wear_nodes[s] = (wl + wr) / dt;
        t->node_pos[left].x += t->n[s].x * wl;
        t->node_pos[left].y += t->n[s].y * wl;
        t->node_pos[right].x += t->n[s].x * wr;
        t->node_pos[right].y += t->n[s].y * wr;
    }
    mesh_adjust_wear(c, t, 1.0);
    for (int s = 0; s < t->n_seg; s++) {
        o_d2 lp = t->node_pos[t->marks[s].z], rp = t->node_pos[t->marks[s].w];
        t->left[s] = lp;
        t->right[s] = rp;
        double vx = rp.y - lp.y, vy = -(rp.x - lp.x);
        double inv = 1. / sqrt(vx * vx + vy * vy);
        t->n[s].x = vx * inv;
        t->n[s].y = vy * inv;
    }
    double mn = ORACLE_FLT_MAX;
    for (int s = 0; s < t->n_seg; s++) mn = fmin(mn, t->left[s].y);
    t->flank_retreat = t->flank_contact - mn;
    for (int i = 0; i < t->n_nodes; i++) t->Minv[i] = 1. / t->M[i];
}

/* output.cu:22-30 */
void o_report_force(const o_particles *P, double *fx, double *fy) {
    double x = 0., y = 0.;
    for (int i = 0; i < P->N; i++) { x += P->fc[i].x; y += P->fc[i].y; }
    for (int i = 0; i < P->N; i++) { x += P->ft[i].x; y += P->ft[i].y; }
    *fx = x;
    *fy = y;
}

/* leap_frog.cu:162-248 */
/* ---------------------------------------------------------------------------------------------
 * rigid (SPH) tool: tool.cpp, geometric_primitives.cpp, tool_gpu.cu
 * ------------------------------------------------------------------------------------------- */

/* geometric_primitives.cpp:61-76 */
static o_line r_line(o_d2 p1, o_d2 p2) {
    o_line l;
    double det = p1.x * 1. - 1. * p2.x;
    if (fabs(det) < 1e-12) {
        l.vertical = 1; l.a = DBL_MAX; l.b = p1.x;
        return l;
    }
    l.vertical = 0;
    l.a = (p1.y * 1. - 1. * p2.y) / det;
    l.b = (p2.y * p1.x - p2.x * p1.y) / det;
    return l;
}

/* geometric_primitives.cpp:40-55 */
static o_d2 r_closest(o_line l, o_d2 q) {
    o_d2 r;
    if (l.vertical) { r.x = l.b; r.y = q.y; return r; }
    double bb = -1., den = l.a * l.a + bb * bb;
    r.x = (bb * (bb * q.x - l.a * q.y) - l.a * l.b) / den;
    r.y = (l.a * (-bb * q.x + l.a * q.y) - bb * l.b) / den;
    return r;
}

/* geometric_primitives.cpp:11-38 */
static o_d2 r_intersect(o_line m, o_line l) {
    o_d2 r = {DBL_MAX, DBL_MAX};
    if (m.vertical && l.vertical) return r;
    if (fabs(m.a - l.a) < 1e-12) return r;
    if (m.vertical) { r.x = m.b; r.y = l.a * m.b + l.b; return r; }
    if (l.vertical) { r.x = l.b; r.y = m.a * l.b + m.b; return r; }
    r.x = (l.b - m.b) / (m.a - l.a);
    r.y = m.a * r.x + m.b;
    return r;
}

static o_d2 r_norm(double x, double y) {        /* glm::normalize */
    double s = 1. / sqrt(x * x + y * y);
    o_d2 r = {x * s, y * s};
    return r;
}

static double r_atan2(double y, double x) {     /* tool.cpp:18-25 */
    double t = atan2(y, x);
    return t > 0. ? t : t + 2 * O_PI;
}

static void r_segments(o_rtool *t, const o_d2 *pts, int n) {   /* tool.cpp:55-62 + geometric_primitives.cpp:103-113 */
    t->num_seg = n;
    for (int i = 0; i < n; i++) {
        o_d2 a = pts[i], b = pts[(i + 1) % n];
        t->seg[i].left = a;
        t->seg[i].right = b;
        t->seg[i].n = r_norm(b.y - a.y, -(b.x - a.x));
        t->seg[i].l = r_line(a, b);
    }
}

static void r_build(o_rtool *t, o_d2 tl, o_d2 tr, o_d2 br, o_d2 bl, double r, double mu) {
    memset(t, 0, sizeof(*t));
    t->mu = mu;
    t->front = br.x;
    if (r == 0.) {
        o_d2 pts[4] = {tl, tr, br, bl};
        r_segments(t, pts, 4);
        return;
    }
    /* tool.cpp:64-93 with fit_fillet :27-53 and solve_quad :11-16 */
    o_d2 nt = r_norm(tr.x - br.x, tr.y - br.y), nl = r_norm(bl.x - br.x, bl.y - br.y);
    o_d2 nm = r_norm(0.5 * (nt.x + nl.x), 0.5 * (nt.y + nl.y));
    o_d2 q = {br.x + nm.x, br.y + nm.y};
    o_line lm = r_line(br, q), l1 = r_line(tr, br), l2 = r_line(bl, br);
    o_d2 p;
    if (l1.vertical) {
        o_line par = {DBL_MAX, l1.b - r, 1};
        p = r_intersect(par, lm);
    } else {
        double A = l1.a - lm.a, B = l1.b - lm.b, C = r * sqrt(l1.a * l1.a + 1.);
        double qa = A * A, qb = 2 * A * B, qc = B * B - C * C;
        double x1 = (-qb + sqrt(qb * qb - 4 * qa * qc)) / (2 * qa);
        double x2 = (-qb - sqrt(qb * qb - 4 * qa * qc)) / (2 * qa);
        p.x = fmin(x1, x2);
        p.y = lm.a * p.x + lm.b;
    }
    o_d2 trc = r_closest(l1, p), blc = r_closest(l2, p);
    t->has_fillet = 1;
    t->fillet.r = r;
    t->fillet.t1 = r_atan2(p.y - trc.y, p.x - trc.x);
    t->fillet.t2 = r_atan2(p.y - blc.y, p.x - blc.x);
    t->fillet.p = p;
    o_d2 pts[5] = {tl, tr, trc, blc, bl};
    r_segments(t, pts, 5);
}

void o_rtool_from_points(o_rtool *t, o_d2 tl, o_d2 tr, o_d2 br, o_d2 bl, double r, double mu) {
    r_build(t, tl, tr, br, bl, r, mu);
}

/* tool.cpp:335-406: faces through tr and bl along (0, -1) * R(a) = (sin a, -cos a) */
void o_rtool_from_angles(o_rtool *t, o_d2 tl, double length, double height, double rake, double clear, double r,
                         double mu) {
    o_d2 tr = {tl.x + length, tl.y}, bl = {tl.x, tl.y - height};
    double ar = rake * O_PI / 180., af = (90. - clear) * O_PI / 180.;
    o_d2 trc = {tr.x + (0. * cos(ar) + -1. * -sin(ar)), tr.y + (0. * sin(ar) + -1. * cos(ar))};
    o_d2 blc = {bl.x + (0. * cos(af) + -1. * -sin(af)), bl.y + (0. * sin(af) + -1. * cos(af))};
    o_d2 br = r_intersect(r_line(tr, trc), r_line(bl, blc));
    r_build(t, tl, tr, br, bl, r, mu);
}

void o_rtool_move(o_rtool *t, double dt) {     /* tool.cpp:202-212 */
    for (int i = 0; i < t->num_seg; i++) {
        o_seg *g = &t->seg[i];
        g->left.x += dt * t->vel.x; g->left.y += dt * t->vel.y;
        g->right.x += dt * t->vel.x; g->right.y += dt * t->vel.y;
        g->l = r_line(g->left, g->right);
    }
    if (t->has_fillet) { t->fillet.p.x += dt * t->vel.x; t->fillet.p.y += dt * t->vel.y; }
}

double o_rtool_low(const o_rtool *t) {
    if (t->has_fillet) return t->fillet.p.y - t->fillet.r;
    double y = DBL_MAX;
    for (int i = 0; i < t->num_seg; i++) y = fmin(y, t->seg[i].left.y);
    return y;
}

void o_rtool_safe_bb(const o_rtool *t, double safety, double *out) {
    double x0 = DBL_MAX, x1 = -DBL_MAX, y0 = DBL_MAX, y1 = -DBL_MAX;
    for (int i = 0; i < t->num_seg; i++) {
        x0 = fmin(t->seg[i].left.x, x0); x1 = fmax(t->seg[i].left.x, x1);
        y0 = fmin(t->seg[i].left.y, y0); y1 = fmax(t->seg[i].left.y, y1);
    }
    if (t->has_fillet) { y0 = o_rtool_low(t); x1 = t->fillet.p.x + t->fillet.r; }
    out[0] = x0 - safety; out[1] = x1 + safety; out[2] = y0 - safety; out[3] = y1 + safety;
}

static int r_inside(const o_rtool *t, double x, double y, int fillet) {
    int in = 1;
    for (int i = 0; i < t->num_seg; i++)
        in = in && ((t->seg[i].left.x - x) * t->seg[i].n.x + (t->seg[i].left.y - y) * t->seg[i].n.y) < 0.;
    if (fillet) {
        double ex = t->fillet.p.x - x, ey = t->fillet.p.y - y;
        in = in || sqrt(ex * ex + ey * ey) < t->fillet.r;
    }
    return in;
}

int o_rtool_inside(const o_rtool *t, double x, double y) { return r_inside(t, x, y, t->has_fillet); }

long o_rtool_sample(const o_rtool *t, double dx, o_d2 *out, long cap) {
    double bb[4];
    o_rtool_safe_bb(t, 0.1 * dx, bb);
    int nx = (int)((bb[1] - bb[0]) / dx + 1), ny = (int)((bb[3] - bb[2]) / dx + 1);
    long k = 0;
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            double x = i * dx + bb[0], y = j * dx + bb[2];
            if (!o_rtool_inside(t, x, y)) continue;
            if (out && k < cap) { out[k].x = x; out[k].y = y; }
            k++;
        }
    return k;
}

/* tool_gpu.cu:82-134 (the fillet counts when there are five segments) */
int o_rtool_establish_contact(const o_rtool *t, double qx, double qy, o_d2 *cp, o_d2 *n) {
    int fil = t->num_seg == 5;
    cp->x = cp->y = n->x = n->y = 0.;
    if (!r_inside(t, qx, qy, fil)) return 0;
    double depth = ORACLE_FLT_MAX;
    for (int i = 0; i < t->num_seg; i++) {
        const o_seg *g = &t->seg[i];
        double d;
        if (g->l.vertical) {
            d = fabs(g->left.x - qx);
            if (d < depth) {
                depth = d;
                *n = r_norm(g->left.x - qx, 0.);
                cp->x = g->left.x; cp->y = qy;
            }
        } else if (fil && i == 2) {
            /* circle_segment_distance, tool_gpu.cu:33-58 */
            const o_circ *f = &t->fillet;
            double ex = qx - f->p.x, ey = qy - f->p.y;
            double Cx = f->p.x + f->r * ex / sqrt(ex * ex + ey * ey);
            double Cy = f->p.y + f->r * ey / sqrt(ex * ex + ey * ey);
            double a = r_atan2(f->p.y - Cy, f->p.x - Cx);
            d = (a > fmin(f->t1, f->t2) && a < fmax(f->t1, f->t2))
                    ? sqrt((Cx - qx) * (Cx - qx) + (Cy - qy) * (Cy - qy)) : ORACLE_FLT_MAX;
            if (d < depth) {
                depth = d;
                *n = r_norm(qx - f->p.x, qy - f->p.y);
                cp->x = f->p.x + n->x * f->r; cp->y = f->p.y + n->y * f->r;
            }
        } else {
            o_d2 q = {qx, qy}, p = r_closest(g->l, q);
            d = sqrt((p.x - qx) * (p.x - qx) + (p.y - qy) * (p.y - qy));
            if (d < depth) {
                depth = d;
                n->x = -g->n.x; n->y = -g->n.y;
                *cp = p;
            }
        }
    }
    return 1;
}

/* tool_gpu.cu:146-202 */
void o_rtool_contact(const o_consts *c, o_particles *P, const o_rtool *t) {
    const double dt = c->dt, ms = t->slave_mass, ca = t->contact_alpha;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P->N; i++) {
        if (P->tool_particle[i] != 0.) continue;
        o_d2 xs = P->pos[i], vs = P->vel[i], xm, n;
        if (!o_rtool_establish_contact(t, xs.x, xs.y, &xm, &n)) {
            P->fc[i].x = P->fc[i].y = P->ft[i].x = P->ft[i].y = P->n[i].x = P->n[i].y = 0.;
            continue;
        }
        double dt2 = dt * dt;
        double gN = (xs.x - xm.x) * n.x + (xs.y - xm.y) * n.y;
        o_d2 fN = {-ms * gN * n.x / dt2 * ca, -ms * gN * n.y / dt2 * ca}, fT = {0., 0.};
        if (t->mu != 0.) {
            double vx = vs.x - t->vel.x, vy = vs.y - t->vel.y;
            double vrx = vx - vx * n.x, vry = vy - vy * n.y;         /* glm v - v*n, per component */
            double kx = ca * ms * vrx / dt, ky = ca * ms * vry / dt;
            double fy = t->mu * sqrt(fN.x * fN.x + fN.y * fN.y);
            double sx = P->ft[i].x - kx, sy = P->ft[i].y - ky;
            double ls = sqrt(sx * sx + sy * sy);
            if (ls > fy) { fT.x = fy * sx / ls; fT.y = fy * sy / ls; }
            else { fT.x = sx; fT.y = sy; }
        }
        P->fc[i] = fN;
        P->ft[i] = fT;
        P->n[i] = n;
    }
}

/* tool_gpu.cu:205-216 */
void o_rtool_tool_particles(const o_consts *c, o_particles *P, const o_rtool *t) {
    for (int i = 0; i < P->N; i++) {
        if (P->tool_particle[i] == 0.) continue;
        P->pos[i].x += c->dt * t->vel.x;
        P->pos[i].y += c->dt * t->vel.y;
    }
}

/* actions_gpu.cu:316-343 (sqrtf of a double: single-precision magnitudes) */
void o_rtool_fric_heat(const o_consts *c, o_particles *P, const o_rtool *t) {
    for (int i = 0; i < P->N; i++) {
        if (P->tool_particle[i] != 0.) continue;
        o_d2 f = P->ft[i];
        double fm = (double)sqrtf((float)(f.x * f.x + f.y * f.y));
        if (fm == 0.) continue;
        o_d2 nn = P->n[i];
        double vdx = P->vel[i].x - t->vel.x, vdy = P->vel[i].y - t->vel.y;
        double vdn = vdx * nn.x + vdy * nn.y;
        double rx = vdx - vdn * nn.x, ry = vdy - vdn * nn.y;
        double vm = (double)sqrtf((float)(rx * rx + ry * ry));
        P->T[i] += c->trml_wp.eta * c->dt * fm * vm / (c->trml_wp.cp * c->phys.mass);
    }
}

/* leap_frog.cu:99-160 */
static void step_rigid(const o_consts *c, o_particles *P, o_stepper *st, o_rtool *t, int rothlin) {
    o_compute_hashes(c, P);
    if (rothlin) o_sort_rothlin(c, P, st);
    else o_sort_and_cells(c, P, st);
    o_init(P, st);
    o_predict(P, st, c->dt);
    o_material_eos(c, P);
    o_artificial_stress(c, P);
    o_interactions_monaghan(c, P, st);
    o_stress_rate_jaumann(c, P);
    o_continuity(P);
    o_momentum(c, P);
    o_advection(P);
    o_correct(P, st, c->dt);
    if (t && c->fric_heat) o_rtool_fric_heat(c, P, t);
    if (c->plastic) o_plasticity(c, P);
    if (t) {
        o_rtool_move(t, c->dt);
        o_rtool_tool_particles(c, P, t);
        o_rtool_contact(c, P, t);
    }
    o_boundary_conditions(P);
    if (c->thermal_bc) o_boundary_conditions_thermal(c, P);
    o_invalidate(P);
}

void o_run_rigid(const o_consts *c, o_particles *P, o_stepper *st, o_rtool *t, int nsteps, int rothlin) {
    for (int s = 0; s < nsteps; s++) step_rigid(c, P, st, t, rothlin);
}

static void step_impl(const o_consts *c, o_particles *P, o_stepper *st, o_tool *t, int rothlin) {
    o_compute_hashes(c, P);
    if (rothlin) o_sort_rothlin(c, P, st);
    else o_sort_and_cells(c, P, st);
    o_init(P, st);
    if (t) { o_fem_reset_server(t); o_segments_wear_rate_reset(t); }
    o_predict(P, st, c->dt);
    if (t) o_fem_calculate_temp(t, 0.5 * c->dt);
    o_material_eos(c, P);
    o_artificial_stress(c, P);
    o_interactions_monaghan(c, P, st);
    o_stress_rate_jaumann(c, P);
    o_continuity(P);
    o_momentum(c, P);
    o_advection(P);
    o_correct(P, st, c->dt);
    if (t) {
        o_fric_heat_v21(c, P, t);
        o_fem_apply_robin(c, t);
        o_fem_calculate_temp_rate(t);
        o_fem_calculate_temp(t, c->dt);
    }
    if (c->plastic) o_plasticity(c, P);
    if (t) {
        if (c->wear_pro) o_wear_adjustment(c, t);
        o_segments_update(t, c->dt);
        o_tool_update(t, c->dt);
        o_contact_s1(c, P, t);
        o_contact_s2(c, P, t);
    }
    o_boundary_conditions(P);
    if (c->thermal_bc) o_boundary_conditions_thermal(c, P);
    o_invalidate(P);
}

void o_step(const o_consts *c, o_particles *P, o_stepper *st, o_tool *t) { step_impl(c, P, st, t, 0); }

void o_run(const o_consts *c, o_particles *P, o_stepper *st, o_tool *t, int nsteps) {
    for (int s = 0; s < nsteps; s++) step_impl(c, P, st, t, 0);
}

/* grid.cu:88-127 for an unlocked grid: bounding box of all positions (min_element /
 * max_element by x and by y), margins of 1e-6, cell size max(h) * kernel_width / 2, ceil counts.
 * Returns 0, or -1 when num_cell >= max_cell (the reference's assert). */
int o_update_geometry(o_consts *c, const o_particles *P, double kernel_width, int max_cell) {
    double mnx = P->pos[0].x, mny = P->pos[0].y, mxx = P->pos[0].x, mxy = P->pos[0].y, mh = P->h[0];
    for (int i = 1; i < P->N; i++) {
        if (P->pos[i].x < mnx) mnx = P->pos[i].x;
        if (P->pos[i].y < mny) mny = P->pos[i].y;
        if (P->pos[i].x > mxx) mxx = P->pos[i].x;
        if (P->pos[i].y > mxy) mxy = P->pos[i].y;
        if (P->h[i] > mh) mh = P->h[i];
    }
    const double bx0 = mnx - 1e-6, by0 = mny - 1e-6, bx1 = mxx + 1e-6, by1 = mxy + 1e-6;
    const double dx = mh * kernel_width / 2;
    const double nx = ceil((bx1 - bx0) / dx), ny = ceil((by1 - by0) / dx);
    if (!(nx * ny < (double)max_cell)) return -1;
    c->geom.nx = (int)nx;
    c->geom.ny = (int)ny;
    c->geom.bbmin_x = bx0;
    c->geom.bbmin_y = by0;
    c->geom.dx = dx;
    return 0;
}

/* leap_frog::step with an unlocked grid (leap_frog.cu:175): geometry from the particles first.
 * Returns the number of steps done (fewer than nsteps if the grid outgrew max_cell). */
int o_run_unlocked(o_consts *c, o_particles *P, o_stepper *st, o_tool *t, int nsteps, int max_cell, int rothlin) {
    for (int s = 0; s < nsteps; s++) {
        if (o_update_geometry(c, P, 2., max_cell)) return s;
        step_impl(c, P, st, t, rothlin);
    }
    return nsteps;
}

/* leap_frog::step with a grid_gpu_rothlin (the same step, the Rothlin sort and cells) */
void o_run_rothlin(const o_consts *c, o_particles *P, o_stepper *st, o_tool *t, int nsteps) {
    for (int s = 0; s < nsteps; s++) step_impl(c, P, st, t, 1);
}
