// Per-particle physics of the cutting step as device functions, shared by the
// one-kernel-per-reference-function entry points and the fused step kernels.
// Each function keeps the reference's evaluation order and its single-precision
// rounding sites (sqrtf / fmaxf / fminf / atan2f / cosf / sinf applied to doubles),
// and FP contraction is off so the FP64 form executes the written operations.
#pragma once

#include <hip/hip_runtime.h>

#include "sphfem_internal.h"

#pragma clang fp contract(off)

namespace sphfem {

#define SPHFEM_PI_FAC 0.454728408833987   // kernels.cuh:19
#define SPHFEM_FLT_MAX 3.40282346638528859811704183484516925e+38

// Correctly rounded single-precision sqrt of (float)x: the device's sqrtf/__fsqrt_rn are not
// correctly rounded on gfx950, while the double sqrt is; rounding the double root of a float
// to float gives the correctly rounded float root (53 >= 2*24 + 2 bits).
__device__ __forceinline__ double sqrtf_d(double x) { return (double)(float)sqrt((double)(float)x); }
__device__ __forceinline__ double fmaxf_d(double a, double b) { return (double)fmaxf((float)a, (float)b); }
__device__ __forceinline__ double fminf_d(double a, double b) { return (double)fminf((float)a, (float)b); }

// kernels.cuh:25-65 (the q > 1 value is 0.25*(2-q)^6 as written there).  Returns false when the
// pair is outside the support or closer than the cut-off, where the reference returns zeros.
__device__ __forceinline__ bool cubic_spline(double pix, double piy, double pjx, double pjy, double hi,
                                             double rr2_min, double &w, double &w_x, double &w_y) {
    double xij = pix - pjx;
    double yij = piy - pjy;
    double rr2 = xij * xij + yij * yij;
    double h1 = 1 / hi;
    double fourh2 = 4 * hi * hi;
    if (rr2 >= fourh2 || rr2 < rr2_min) {
        w = 0.; w_x = 0.; w_y = 0.;
        return false;
    }
    double der, val;
    double rad = sqrtf_d(rr2);
    double q = rad * h1;
    double fac = SPHFEM_PI_FAC * h1 * h1;
    if (q > 1) {
        double _2mq = 2 - q;
        double _2mq2 = _2mq * _2mq;
        val = 0.25 * _2mq2 * _2mq2 * _2mq2;
        der = -0.75 * _2mq2 * h1 / rad;
    } else {
        val = 1 - 1.5 * q * q * (1 - 0.5 * q);
        der = -3.0 * q * (1 - 0.75 * q) * h1 / rad;
    }
    w = val * fac;
    w_x = der * xij * fac;
    w_y = der * yij * fac;
    return true;
}

// actions_gpu.cu:222-232
__device__ __forceinline__ double eos_pressure(const DevConsts &c, double rho) {
    // c0 * c0 with c0 = sqrtf(K / rho0), computed once when the constants are set (DevConsts::eos_c02)
    return c.eos_c02 * (rho - c.phys.rho0);
}

// actions_gpu.cu:45-49, 234-278
__device__ __forceinline__ sphfem_double4 artificial_stress(const DevConsts &c, double rhoi, double pi,
                                                             sphfem_double4 Si) {
    double eps = c.corr.stresseps;
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
    sphfem_double4 R;
    R.x = cos_theta2 * rot_rxx + sin_theta2 * rot_ryy;
    R.y = cos_theta * sin_theta * (rot_rxx - rot_ryy);
    R.z = sin_theta2 * rot_rxx + cos_theta2 * rot_ryy;
    R.w = 0.;
    return R;
}

// actions_gpu.cu:280-314: 2G(eps_dot - tr/3 I) + Omega S + S Omega^T with the GLM 3x3 products
// reduced to their single non-zero term per entry.
__device__ __forceinline__ sphfem_double4 jaumann_rate(const DevConsts &c, sphfem_double4 d, sphfem_double4 S) {
    double vx_x = d.x, vx_y = d.y, vy_x = d.z, vy_y = d.w;
    double Sxx = S.x, Sxy = S.y, Syy = S.z;
    double exy = 0.5 * (vx_y + vy_x);
    double w01 = 0.5 * (vx_y - vy_x);
    double w10 = 0.5 * (vy_x - vx_y);
    double tr = vx_x + vy_y + 0.;
    double t3 = (1. / 3.) * tr;
    double g2 = 2. * c.phys.G;
    sphfem_double4 St;
    St.x = (g2 * (vx_x - t3) + w01 * Sxy) + Sxy * w01;
    St.y = (g2 * (exy - 0.) + w10 * Sxx) + Syy * w01;
    St.z = (g2 * (vy_y - t3) + w10 * Sxy) + Sxy * w10;
    St.w = g2 * (0. - t3);
    return St;
}

// plasticity.cuh:21-39
__device__ __forceinline__ double sigma_yield(const sphfem_joco_constants &jc, double eps_pl, double eps_pl_dot,
                                              double t) {
    double theta = (t - jc.Tref) / (jc.Tmelt - jc.Tref);
    if (theta > 1.) theta = 1.;
    double Term_A = jc.A + jc.B * pow(eps_pl, jc.n);
    double Term_B = 1.0;
    double eps_dot = eps_pl_dot / jc.eps_dot_ref;
    if (eps_dot > 1.0) Term_B = 1.0 + jc.C * log(eps_dot);
    else Term_B = pow((1.0 + eps_dot), jc.C);
    double Term_C = 1.0 - pow(theta, jc.m);
    return Term_A * Term_B * Term_C;
}

// plasticity.cuh:117-126
__device__ __forceinline__ double opt_fun(const sphfem_joco_constants &jc, double delta_lambda, double norm_Strial,
                                          double eps_pl_init, double t, double delta_t, double mu) {
    double eps_pl_equiv = eps_pl_init + sqrt(2.0 / 3.0) * fmax(delta_lambda, 0.);
    double eps_pl_equiv_dot = sqrt(2.0 / 3.0) * fmax(delta_lambda, 0.) / delta_t;
    double sigmaY = sigma_yield(jc, eps_pl_equiv, eps_pl_equiv_dot, t);
    return norm_Strial - 2 * mu * delta_lambda - sqrt(2.0 / 3.0) * sigmaY;
}

// plasticity.cuh:128-162
__device__ __forceinline__ double solve_secant(const sphfem_joco_constants &jc, double init, double tol,
                                               double norm_Strial, double eps_pl_init, double t, double delta_t,
                                               double mu) {
    double delta_lambda = init;
    double delta_lambda_old = delta_lambda;
    int iter = 0;
    const int max_iter = 100;
    for (;;) {
        double g = opt_fun(jc, delta_lambda, norm_Strial, eps_pl_init, t, delta_t, mu);
        double delta_lambda_1 = delta_lambda * 1.01;
        double g_1 = opt_fun(jc, delta_lambda_1, norm_Strial, eps_pl_init, t, delta_t, mu);
        double slope = (g - g_1) / (delta_lambda - delta_lambda_1);
        delta_lambda_old = delta_lambda;
        delta_lambda = delta_lambda - g / slope;
        if (delta_lambda < 0.) delta_lambda = -0.1 * delta_lambda;
        if (fabs(delta_lambda - delta_lambda_old) < tol) return delta_lambda;
        if (iter > max_iter) return delta_lambda;
        iter++;
    }
}

// actions_gpu.cu:689-754 on one particle's state; returns true if the particle melts (is blanked).
__device__ __forceinline__ bool plasticity_jc(const DevConsts &c, sphfem_double4 &S, double p, double &eps_pl,
                                              double &eps_pl_dot, double &T, double rho) {
    const double mu = c.phys.G, dt = c.dt;
    double Sxx = S.x, Syy = S.z, Szz = S.w, Sxy = S.y;
    double norm_Strial = sqrtf_d(Sxx * Sxx + Syy * Syy + Szz * Szz + 2. * Sxy * Sxy);
    double cxx = Sxx - p, cyy = Syy - p, czz = Szz - p, cxy = Sxy;
    double svm = sqrtf_d((cxx * cxx + cyy * cyy + czz * czz) - cxx * cyy - cxx * czz - cyy * czz + 3.0 * cxy * cxy);
    // well below yield (a lower bound of the yield stress, DevConsts::jc_svm_below): the same outcome
    // as the comparison below, without the three pow/log evaluations of the yield stress
    if (svm < c.jc_svm_below && eps_pl >= 0. && eps_pl_dot >= 0. && T >= c.jc_T_lo && T <= c.jc_T_hi) {
        eps_pl_dot = 0.;
        return false;
    }
    double sigma_Y = sigma_yield(c.joco, eps_pl, eps_pl_dot, T);
    if (svm < sigma_Y) {
        eps_pl_dot = 0.;
        return false;
    }
    double delta_lambda = solve_secant(c.joco, fmax(eps_pl_dot * dt * sqrt(2. / 3.), 1e-8), 1e-6, norm_Strial,
                                       eps_pl, T, dt, c.phys.G);
    double s23 = sqrtf_d((double)(2.0f / 3.0f));
    double dl = (double)fmaxf((float)delta_lambda, 0.f);
    double eps_pl_new = eps_pl + s23 * dl;
    double eps_pl_dot_new = s23 * dl / dt;
    double eps_old = eps_pl;
    eps_pl = eps_pl_new;
    eps_pl_dot = eps_pl_dot_new;
    sphfem_double4 Sn;
    Sn.x = Sxx - Sxx / norm_Strial * delta_lambda * 2. * mu;
    Sn.z = Syy - Syy / norm_Strial * delta_lambda * 2. * mu;
    Sn.w = Szz - Szz / norm_Strial * delta_lambda * 2. * mu;
    Sn.y = Sxy - Sxy / norm_Strial * delta_lambda * 2. * mu;
    S = Sn;
    if (c.trml_wp.tq != 0.) {
        double delta_eps_pl = eps_pl_new - eps_old;
        double sY = sigma_yield(c.joco, eps_pl_new, eps_pl_dot_new, T);
        T += c.trml_wp.tq / (c.trml_wp.cp * rho) * delta_eps_pl * sY;
    }
    if (T > c.joco.Tmelt) {
        T = c.joco.Tmelt;
        return true;
    }
    return false;
}

// actions_gpu.cu:33-43
__device__ __forceinline__ double find_max(double a, double b) { return a >= b ? a : b; }
__device__ __forceinline__ double find_min(double a, double b) { return a >= b ? b : a; }

// actions_gpu.cu:51-63 ("abs" on a double read as fabs)
__device__ __forceinline__ int on_segment_mesh(sphfem_double2 left, sphfem_double2 right, sphfem_double2 point) {
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

// actions_gpu.cu:83-93
__device__ __forceinline__ int inside_judgement(sphfem_double2 qp, const sphfem_double2 *left,
                                                const sphfem_double2 *right, int num_seg) {
    int intersections = 0;
    for (int i = 0; i < num_seg; i++) intersections += on_segment_mesh(left[i], right[i], qp);
    return (intersections % 2 == 0) ? 0 : 1;
}

// actions_gpu.cu:139-220; succ[s] is the first j with left[j] == right[s] (precomputed, -1 if none)
__device__ __forceinline__ int find_shortest_distance_v1(sphfem_double2 qp, double &dN, double &nre_x,
                                                         double &nre_y, const sphfem_double2 *left,
                                                         const sphfem_double2 *right, const sphfem_double2 *n,
                                                         int num_seg) {
    int seg = -1;
    double depth = SPHFEM_FLT_MAX;
    const double limit_value = 2.5e-4;
    for (int i = 0; i < num_seg; i++) {
        sphfem_double2 l = left[i], r = right[i];
        double ymax = find_max(l.y, r.y);
        double ymin = find_min(l.y, r.y);
        double xmax = find_max(l.x, r.x);
        double xmin = find_min(l.x, r.x);
        if (qp.x > xmin - 0.0008 && qp.x < xmax + 0.0008 && qp.y > ymin - 0.0008 && qp.y < ymax + 0.0008) {
            sphfem_double2 ni = n[i];
            double d = (qp.x - l.x) * ni.x + (qp.y - l.y) * ni.y;
            if (d >= 0. && d < depth) {
                double ix = qp.x - d * ni.x;
                double iy = qp.y - d * ni.y;
                if (iy <= ymax && iy >= ymin && ix <= xmax && ix >= xmin) {
                    depth = d;
                    seg = i;
                }
            }
        }
    }
    if (seg != -1 && depth <= limit_value) {
        nre_x = -n[seg].x;
        nre_y = -n[seg].y;
        dN = -depth;
    } else {
        depth = SPHFEM_FLT_MAX;
        for (int i = 0; i < num_seg; i++) {
            sphfem_double2 r = right[i];
            double d_trial = sqrtf_d((qp.x - r.x) * (qp.x - r.x) + (qp.y - r.y) * (qp.y - r.y));
            if (d_trial < depth) {
                depth = d_trial;
                seg = i;
            }
        }
        if (seg == -1) return seg;
        double nsx = n[seg].x, nsy = n[seg].y;
        sphfem_double2 rs = right[seg];
        for (int j = 0; j < num_seg; j++) {
            if (left[j].y == rs.y && left[j].x == rs.x) {
                nsx = (n[j].x + n[seg].x) / 2.;
                nsy = (n[j].y + n[seg].y) / 2.;
                break;
            }
        }
        double n_length = sqrtf_d(nsx * nsx + nsy * nsy);
        nre_x = -nsx / n_length;
        nre_y = -nsy / n_length;
        dN = -1.0 * depth;
    }
    return seg;
}

// actions_gpu.cu:423-583 for one particle that passed the blanked check and whose distance to the
// tool reference point is below 1.05 ref_dist.  Returns the segment (or -1) and writes fc, ft, n;
// sets `blank` for penetrations deeper than 2.5e-4.
__device__ __forceinline__ int contact_s1_near(const DevConsts &c, const ToolKin &kin, sphfem_double2 pos,
                                               sphfem_double2 vel, sphfem_double2 &fc, sphfem_double2 &ft,
                                               sphfem_double2 &nrm, bool &blank, const sphfem_double2 *left,
                                               const sphfem_double2 *right, const sphfem_double2 *n, int n_seg) {
    blank = false;
    int inside = inside_judgement(pos, left, right, n_seg);
    sphfem_double2 z = {0., 0.};
    if (inside == 0) {
        fc = z; ft = z; nrm = z;
        return -1;
    }
    double gN = 0., nx_ = 0., ny_ = 0.;
    int seg_no = find_shortest_distance_v1(pos, gN, nx_, ny_, left, right, n, n_seg);
    if (seg_no == -1 || gN < -0.00025) {
        if (seg_no != -1) blank = true;
        fc = z; ft = z; nrm = z;
        return -1;
    }
    const double dt = c.dt, m = c.tc.slave_mass, ca = c.tc.contact_alpha, cof = c.tc.mu;
    double dt2 = dt * dt;
    double fNx = -m * gN * nx_ / dt2 * ca;
    double fNy = -m * gN * ny_ / dt2 * ca;
    double vx = vel.x - kin.vel_x, vy = vel.y - kin.vel_y;
    double vdn = vx * nx_ + vy * ny_;
    double vrx = vx - nx_ * vdn, vry = vy - ny_ * vdn;
    double kx = ca * m * vrx / dt, ky = ca * m * vry / dt;
    double fy = cof * sqrt(fNx * fNx + fNy * fNy);
    double fsx = ft.x - kx, fsy = ft.y - ky;
    double lfs = sqrt(fsx * fsx + fsy * fsy);
    double fTx, fTy;
    if (lfs > fy) {
        fTx = fy * fsx / lfs;
        fTy = fy * fsy / lfs;
    } else {
        fTx = fsx;
        fTy = fsy;
    }
    fc.x = fNx; fc.y = fNy;
    ft.x = fTx; ft.y = fTy;
    nrm.x = nx_; nrm.y = ny_;
    return seg_no;
}

// ---- wavefront-cooperative form of contact s1: one particle per wavefront, segments over lanes.
// Every lane must execute these (uniform control flow over the wavefront).
__device__ __forceinline__ int wave_sum_int(int v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// lexicographic minimum of (d, i) over the wavefront; i = -1 sorts last among equal d
__device__ __forceinline__ void wave_argmin(double &d, int &i) {
    for (int off = 32; off > 0; off >>= 1) {
        const double od = __shfl_xor(d, off);
        const int oi = __shfl_xor(i, off);
        if (od < d || (od == d && (unsigned)oi < (unsigned)i)) {
            d = od;
            i = oi;
        }
    }
}

__device__ __forceinline__ int wave_min_int(int v) {
    for (int off = 32; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off);
        v = o < v ? o : v;
    }
    return v;
}

// find_shortest_distance_v1 (actions_gpu.cu:139-220) over the lanes.  Each loop of the reference
// keeps the first segment that attains the minimum (strict <), which is the lexicographic
// (distance, index) minimum computed here; the normal-averaging loop takes the lowest index j with
// left[j] == right[seg].  Same arithmetic per segment, so the result is identical.
__device__ __forceinline__ int find_shortest_distance_wave(sphfem_double2 qp, double &dN, double &nre_x,
                                                           double &nre_y, const sphfem_double2 *left,
                                                           const sphfem_double2 *right, const sphfem_double2 *n,
                                                           int num_seg, int lane) {
    const double limit_value = 2.5e-4;
    double depth = SPHFEM_FLT_MAX;
    int seg = -1;
    for (int i = lane; i < num_seg; i += 64) {
        sphfem_double2 l = left[i], r = right[i];
        double ymax = find_max(l.y, r.y);
        double ymin = find_min(l.y, r.y);
        double xmax = find_max(l.x, r.x);
        double xmin = find_min(l.x, r.x);
        if (qp.x > xmin - 0.0008 && qp.x < xmax + 0.0008 && qp.y > ymin - 0.0008 && qp.y < ymax + 0.0008) {
            sphfem_double2 ni = n[i];
            double d = (qp.x - l.x) * ni.x + (qp.y - l.y) * ni.y;
            if (d >= 0. && d < depth) {
                double ix = qp.x - d * ni.x;
                double iy = qp.y - d * ni.y;
                if (iy <= ymax && iy >= ymin && ix <= xmax && ix >= xmin) {
                    depth = d;
                    seg = i;
                }
            }
        }
    }
    wave_argmin(depth, seg);
    if (seg != -1 && depth <= limit_value) {
        nre_x = -n[seg].x;
        nre_y = -n[seg].y;
        dN = -depth;
        return seg;
    }
    depth = SPHFEM_FLT_MAX;
    seg = -1;
    for (int i = lane; i < num_seg; i += 64) {
        sphfem_double2 r = right[i];
        double d_trial = sqrtf_d((qp.x - r.x) * (qp.x - r.x) + (qp.y - r.y) * (qp.y - r.y));
        if (d_trial < depth) {
            depth = d_trial;
            seg = i;
        }
    }
    wave_argmin(depth, seg);
    if (seg == -1) return seg;
    const sphfem_double2 rs = right[seg];
    int jmatch = 0x7fffffff;
    for (int j = lane; j < num_seg; j += 64) {
        if (left[j].y == rs.y && left[j].x == rs.x) {
            jmatch = j;
            break;
        }
    }
    jmatch = wave_min_int(jmatch);
    double nsx = n[seg].x, nsy = n[seg].y;
    if (jmatch != 0x7fffffff) {
        nsx = (n[jmatch].x + n[seg].x) / 2.;
        nsy = (n[jmatch].y + n[seg].y) / 2.;
    }
    double n_length = sqrtf_d(nsx * nsx + nsy * nsy);
    nre_x = -nsx / n_length;
    nre_y = -nsy / n_length;
    dN = -1.0 * depth;
    return seg;
}

// contact_s1_near with the segment loops spread over the wavefront's lanes; every lane returns the
// same result
__device__ __forceinline__ int contact_s1_near_wave(const DevConsts &c, const ToolKin &kin, sphfem_double2 pos,
                                                    sphfem_double2 vel, sphfem_double2 &fc, sphfem_double2 &ft,
                                                    sphfem_double2 &nrm, bool &blank, const sphfem_double2 *left,
                                                    const sphfem_double2 *right, const sphfem_double2 *n, int n_seg,
                                                    int lane) {
    blank = false;
    int cnt = 0;
    for (int i = lane; i < n_seg; i += 64) cnt += on_segment_mesh(left[i], right[i], pos);
    cnt = wave_sum_int(cnt);
    sphfem_double2 z = {0., 0.};
    if (cnt % 2 == 0) {
        fc = z; ft = z; nrm = z;
        return -1;
    }
    double gN = 0., nx_ = 0., ny_ = 0.;
    int seg_no = find_shortest_distance_wave(pos, gN, nx_, ny_, left, right, n, n_seg, lane);
    if (seg_no == -1 || gN < -0.00025) {
        if (seg_no != -1) blank = true;
        fc = z; ft = z; nrm = z;
        return -1;
    }
    const double dt = c.dt, m = c.tc.slave_mass, ca = c.tc.contact_alpha, cof = c.tc.mu;
    double dt2 = dt * dt;
    double fNx = -m * gN * nx_ / dt2 * ca;
    double fNy = -m * gN * ny_ / dt2 * ca;
    double vx = vel.x - kin.vel_x, vy = vel.y - kin.vel_y;
    double vdn = vx * nx_ + vy * ny_;
    double vrx = vx - nx_ * vdn, vry = vy - ny_ * vdn;
    double kx = ca * m * vrx / dt, ky = ca * m * vry / dt;
    double fy = cof * sqrt(fNx * fNx + fNy * fNy);
    double fsx = ft.x - kx, fsy = ft.y - ky;
    double lfs = sqrt(fsx * fsx + fsy * fsy);
    double fTx, fTy;
    if (lfs > fy) {
        fTx = fy * fsx / lfs;
        fTy = fy * fsy / lfs;
    } else {
        fTx = fsx;
        fTy = fsy;
    }
    fc.x = fNx; fc.y = fNy;
    ft.x = fTx; ft.y = fTy;
    nrm.x = nx_; nrm.y = ny_;
    return seg_no;
}

__device__ __forceinline__ bool near_tool(const ToolKin &kin, sphfem_double2 pos) {
    double ddx = pos.x - kin.tl_x, ddy = pos.y - kin.tl_y;
    const double d2 = ddx * ddx + ddy * ddy, r = 1.05 * kin.ref_dist;
    // beyond the radius by a relative 1e-10 (far above the rounding of the square root and of r * r)
    // the comparison below is false: skip the square root there
    if (d2 > r * r * (1. + 1e-10)) return false;
    return sqrt(d2) < r;
}

__device__ __forceinline__ bool isnaninf(double v) { return isnan(v) || isinf(v); }

// Atomic add of v[q] into dst[q][key * stride] for the executing lanes with `active`, aggregated
// per distinct key within the wavefront: the values of lanes with equal keys are summed in lane
// order (reading only those lanes) and the lowest of them issues one atomic per value.  Safe in
// divergent code: only executing lanes take part.  Keeps the first step cheap, when every
// particle still carries segment 0 (particle_gpu.cu:98) and would hit one address.
template <int NV>
__device__ __forceinline__ void wave_atomic_add(bool active, int key, const double (&v)[NV], double *const (&dst)[NV],
                                                const int (&stride)[NV]) {
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int lkey = __shfl(key, leader);
        const bool mine = active && key == lkey && ((pending >> __lane_id()) & 1ull);
        const uint64_t group = __ballot(mine);
        double s[NV];
        for (int q = 0; q < NV; q++) s[q] = 0.0;
        uint64_t g = group;
        while (g) {
            const int l = __ffsll((unsigned long long)g) - 1;
            for (int q = 0; q < NV; q++) s[q] += __shfl(v[q], l);
            g &= g - 1;
        }
        if (__lane_id() == leader)
            for (int q = 0; q < NV; q++) unsafeAtomicAdd(dst[q] + (long)lkey * stride[q], s[q]);
        pending &= ~group;
    }
}

// Segment sums in a fixed order.  Each contributing particle appends one record (segment, sorted
// particle index, values); k_segment_reduce then adds the values to the segment arrays in
// increasing particle index per segment, the order of the reference's particle loop
// (actions_gpu.cu:398-418, 607-608), which its atomics leave open.  Contributions beyond `cap` in
// one step (the first step of a large case, where every particle still carries segment 0) are
// added directly with wave-aggregated atomics: still correct, no longer reproducible bit for bit.
struct SegSink {
    int *count;
    unsigned long long *keys;   // [cap] segment << 32 | particle index
    double *vals;               // [NV][cap]
    int cap;
};

template <int NV>
__device__ __forceinline__ void seg_sink_add(const SegSink &s, bool active, int seg, int i, const double (&v)[NV],
                                             double *const (&dst)[NV], const int (&stride)[NV]) {
    const uint64_t m = __ballot(active);
    if (!m) return;
    const int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (__lane_id() == leader) base = atomicAdd(s.count, __popcll(m));
    base = __shfl(base, leader);
    if (!active) return;
    const int slot = base + __popcll(m & ((1ull << __lane_id()) - 1ull));
    const bool fits = slot < s.cap;
    if (fits) {
        s.keys[slot] = ((unsigned long long)(unsigned)seg << 32) | (unsigned)i;
        for (int q = 0; q < NV; q++) s.vals[(size_t)q * s.cap + slot] = v[q];
    }
    wave_atomic_add<NV>(!fits, seg, v, dst, stride);
}

// tool-side arrays the particle kernels accumulate into
struct ToolView {
    const double *T_nodes;
    const sphfem_int4 *marks;
    double *fric_heat;
    sphfem_double4 *physical_para;
    double *sliding_force;
    const ToolKin *kin;
    int *near_list;
    int *near_count;
    int active;
    SegSink sink;          // the friction-heat segment sums (fric_heat, physical_para, sliding_force)
};

// actions_gpu.cu:345-420 for one particle; returns false if the particle was blanked
__device__ __forceinline__ bool fric_heat_v21(const DevConsts &c, const ToolView &tv, const ToolKin &kin, int pidx,
                                              double eps_pl, sphfem_double2 ft, int seg, sphfem_double2 normal,
                                              sphfem_double2 vp, sphfem_double2 fcv, double rho, double &T) {
    bool keep = true, add_heat = false, add_wear = false;
    double heat = 0., pp[4] = {0., 0., 0., 0.}, slide = 0.;
    do {
        if (eps_pl > 20.) break;
        double f_fric_mag = sqrtf_d(ft.x * ft.x + ft.y * ft.y);
        if (seg == -1) break;
        const double partition_wp = c.tc.beta;
        const double dt = c.dt;
        double vdx = vp.x - kin.vel_x, vdy = vp.y - kin.vel_y;
        double vdn = vdx * normal.x + vdy * normal.y;
        double vrx = vdx - vdn * normal.x, vry = vdy - vdn * normal.y;
        double v_rel_mag = sqrtf_d(vrx * vrx + vry * vry);
        double value = c.tc.eta * f_fric_mag * v_rel_mag;
        if (isnan(v_rel_mag)) { keep = false; break; }
        double f_normal = sqrtf_d(fcv.x * fcv.x + fcv.y * fcv.y);
        double length = sqrtf_d(c.tc.slave_mass / rho);
        double pressure_trial = f_normal / length;
        if (v_rel_mag > 0.03) break;
        double T_buffer = T + value * (1 - partition_wp) * dt / c.trml_wp.cp / c.phys.mass;
        if (T_buffer > 1550.) {
            T = 1550.;
            break;
        }
        T = T_buffer;
        add_heat = true;
        heat = value * partition_wp;
        sphfem_int4 mk = tv.marks[seg];
        double T_seg = 0.5 * (tv.T_nodes[mk.z] + tv.T_nodes[mk.w]);
        double vel_tool_norm = sqrt(kin.vel_x * kin.vel_x + kin.vel_y * kin.vel_y);
        if (v_rel_mag > 0.008) break;
        else if (v_rel_mag > vel_tool_norm * 8.0 && fabs(normal.y) > 0.5) break;
        add_wear = true;
        pp[0] = 1.;
        pp[1] = pressure_trial;
        pp[2] = v_rel_mag;
        pp[3] = T_seg;
        slide = f_fric_mag / length;
    } while (false);
    // one record per heated particle; the Usui inputs and the sliding force are zero unless add_wear
    // (adding +0 to the segment sums leaves them unchanged)
    const double v[6] = {heat, pp[0], pp[1], pp[2], pp[3], slide};
    double *const dst[6] = {tv.fric_heat, &tv.physical_para[0].x, &tv.physical_para[0].y, &tv.physical_para[0].z,
                            &tv.physical_para[0].w, tv.sliding_force};
    const int stride[6] = {1, 4, 4, 4, 4, 1};
    (void)add_wear;
    seg_sink_add<6>(tv.sink, add_heat, seg, pidx, v, dst, stride);
    return keep;
}

}  // namespace sphfem
