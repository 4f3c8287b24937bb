// Rigid (SPH) tool: polygon-plus-fillet geometry shared by the host constructors and the device
// contact, the contact/friction law of tool_gpu.cu and the friction heat of actions_gpu.cu:316-343.
//
// Every routine performs the reference's operations in the reference's order (FP contraction
// off), so that host (glibc) and device results agree except where atan2 differs in the last
// place; the only consumer of atan2 on the device is the validity test of the fillet arc.
#pragma once

#include <float.h>
#include <math.h>

#include "sphfem_internal.h"

#pragma clang fp contract(off)

namespace sphfem {

// the device copy of a rigid tool (tool_gpu.cu:16-22: segments, num_seg, fillet, friction_mu,
// contact_alpha, slave_mass, tool_vel)
struct RigidTool {
    sphfem_tool_geometry g;
    double contact_alpha;
    double slave_mass;
};

#define SPHFEM_RIGID_PI 3.14159265358979323846

// geometric_primitives.cpp:61-76: y = a x + b through p1, p2 by Cramer's rule on
// [[p1.x, 1], [p2.x, 1]]; vertical (x = p1.x) when the determinant is below 1e-12
__host__ __device__ inline sphfem_line rigid_line(sphfem_double2 p1, sphfem_double2 p2) {
    sphfem_line l;
    const double Mxx = p1.x, Mxy = 1., Myx = p2.x, Myy = 1.;
    const double detM = Mxx * Myy - Mxy * Myx;
    if (fabs(detM) < 1e-12) {
        l.vertical = 1;
        l.a = DBL_MAX;
        l.b = p1.x;
        return l;
    }
    l.vertical = 0;
    l.a = (p1.y * Myy - Mxy * p2.y) / detM;
    l.b = (p2.y * Mxx - Myx * p1.y) / detM;
    return l;
}

// geometric_primitives.cpp:40-55 (and tool_gpu.cu:60-76): foot of the perpendicular from xq on
// the line a x - y + b = 0
__host__ __device__ inline sphfem_double2 rigid_closest_point(const sphfem_line &l, sphfem_double2 xq) {
    if (l.vertical) return {l.b, xq.y};
    const double bb = -1., cc = l.b, aa = l.a;
    const double px = (bb * (bb * xq.x - aa * xq.y) - aa * cc) / (aa * aa + bb * bb);
    const double py = (aa * (-bb * xq.x + aa * xq.y) - bb * cc) / (aa * aa + bb * bb);
    return {px, py};
}

// geometric_primitives.cpp:11-38 (DBL_MAX for parallel lines)
__host__ __device__ inline sphfem_double2 rigid_intersect(const sphfem_line &m, const sphfem_line &l) {
    if (m.vertical && l.vertical) return {DBL_MAX, DBL_MAX};
    if (fabs(m.a - l.a) < 1e-12) return {DBL_MAX, DBL_MAX};
    if (m.vertical) return {m.b, l.a * m.b + l.b};
    if (l.vertical) return {l.b, m.a * l.b + m.b};
    const double x = (l.b - m.b) / (m.a - l.a);
    return {x, m.a * x + m.b};
}

// glm::normalize (v * inversesqrt(dot(v, v)), inversesqrt = 1 / sqrt)
__host__ __device__ inline sphfem_double2 rigid_normalize(sphfem_double2 v) {
    const double s = 1. / sqrt(v.x * v.x + v.y * v.y);
    return {v.x * s, v.y * s};
}

// glm::length
__host__ __device__ inline double rigid_length(double x, double y) { return sqrt(x * x + y * y); }

// geometric_primitives.cpp:103-113
__host__ __device__ inline sphfem_segment rigid_segment(sphfem_double2 left, sphfem_double2 right) {
    sphfem_segment s;
    s.left = left;
    s.right = right;
    const sphfem_double2 d = {right.x - left.x, right.y - left.y};
    s.n = rigid_normalize({d.y, -d.x});
    s.l = rigid_line(left, right);
    return s;
}

// tool.cpp:18-25 / tool_gpu.cu:24-31: atan2 moved to (0, 2 pi]
__host__ __device__ inline double rigid_atan2(double y, double x) {
    const double t = atan2(y, x);
    return t > 0. ? t : t + 2 * SPHFEM_RIGID_PI;
}

// tool.cpp:202-212: both ends and the fillet centre advance by dt * vel, lines rebuilt from the
// moved ends (normals unchanged)
__host__ __device__ inline void rigid_move(sphfem_tool_geometry &g, double dt) {
    for (int i = 0; i < g.num_seg; i++) {
        sphfem_segment &s = g.seg[i];
        s.left.x += dt * g.vel.x;
        s.left.y += dt * g.vel.y;
        s.right.x += dt * g.vel.x;
        s.right.y += dt * g.vel.y;
        s.l = rigid_line(s.left, s.right);
    }
    if (g.has_fillet) {
        g.fillet.p.x += dt * g.vel.x;
        g.fillet.p.y += dt * g.vel.y;
    }
}

// tool.cpp:95-110 and tool_gpu.cu:85-97: strictly inside every half-plane dot(left - q, n) < 0,
// or within the fillet circle (the host tests the fillet when one is set, the device when there
// are five segments)
__host__ __device__ inline bool rigid_inside(const sphfem_tool_geometry &g, double qx, double qy, bool fillet) {
    bool in = true;
    for (int i = 0; i < g.num_seg; i++) {
        const sphfem_segment &s = g.seg[i];
        in = in && ((s.left.x - qx) * s.n.x + (s.left.y - qy) * s.n.y) < 0.;
    }
    if (fillet) in = in || rigid_length(g.fillet.p.x - qx, g.fillet.p.y - qy) < g.fillet.r;
    return in;
}

// tool_gpu.cu:33-58: distance from qp to the fillet arc when the radial projection of qp falls
// between t1 and t2 (measured as the angle of centre - projection), FLT_MAX otherwise
__host__ __device__ inline double rigid_circle_distance(const sphfem_circle_segment &f, double qx, double qy) {
    const double Ax = f.p.x, Ay = f.p.y, Bx = qx, By = qy, r = f.r;
    const double Cx = Ax + r * (Bx - Ax) / sqrt((Bx - Ax) * (Bx - Ax) + (By - Ay) * (By - Ay));
    const double Cy = Ay + r * (By - Ay) / sqrt((Bx - Ax) * (Bx - Ax) + (By - Ay) * (By - Ay));
    const double t = rigid_atan2(f.p.y - Cy, f.p.x - Cx);
    const bool valid = t > fmin(f.t1, f.t2) && t < fmax(f.t1, f.t2);
    return valid ? rigid_length(Cx - qx, Cy - qy) : (double)FLT_MAX;
}

// tool_gpu.cu:82-134: inside test, then the closest primitive (vertical segment, fillet arc in
// place of the third segment, or the line of the segment) gives the contact point cp and the
// outward normal n.  The reference leaves cp and n unset when no primitive is closer than FLT_MAX
// (a non-finite position); they are zero here.
__host__ __device__ inline bool rigid_establish_contact(const sphfem_tool_geometry &g, double qx, double qy,
                                                        sphfem_double2 &cp, sphfem_double2 &n) {
    cp = {0., 0.};
    n = {0., 0.};
    const bool has_fillet = g.num_seg == SPHFEM_TOOL_MAX_SEG;
    if (!rigid_inside(g, qx, qy, has_fillet)) return false;
    double depth = (double)FLT_MAX;
    for (int i = 0; i < g.num_seg; i++) {
        const sphfem_segment &s = g.seg[i];
        if (s.l.vertical) {
            const double d = fabs(s.left.x - qx);
            if (d < depth) {
                depth = d;
                n = rigid_normalize({s.left.x - qx, 0.});
                cp = {s.left.x, qy};
            }
        } else if (has_fillet && i == 2) {
            const double d = rigid_circle_distance(g.fillet, qx, qy);
            if (d < depth) {
                depth = d;
                n = rigid_normalize({qx - g.fillet.p.x, qy - g.fillet.p.y});
                cp = {g.fillet.p.x + n.x * g.fillet.r, g.fillet.p.y + n.y * g.fillet.r};
            }
        } else {
            const sphfem_double2 p = rigid_closest_point(s.l, {qx, qy});
            const double d = rigid_length(p.x - qx, p.y - qy);
            if (d < depth) {
                depth = d;
                n = {-s.n.x, -s.n.y};
                cp = p;
            }
        }
    }
    return true;
}

// tool_gpu.cu:146-202 for one non-tool particle: penalty normal force (slave mass, gap, dt^2,
// contact alpha) and the LS-DYNA return-mapped friction from the previous friction force.  The
// tangential velocity is v - v*n taken component by component, as written in the reference.
__host__ __device__ inline void rigid_contact(const RigidTool &t, double dt, sphfem_double2 xs, sphfem_double2 vs,
                                              sphfem_double2 ft_old, sphfem_double2 &fc, sphfem_double2 &ft,
                                              sphfem_double2 &nrm) {
    sphfem_double2 xm, n;
    if (!rigid_establish_contact(t.g, xs.x, xs.y, xm, n)) {
        fc = {0., 0.};
        ft = {0., 0.};
        nrm = {0., 0.};
        return;
    }
    const double dt2 = dt * dt;
    const double gN = (xs.x - xm.x) * n.x + (xs.y - xm.y) * n.y;
    const double s = -t.slave_mass * gN;
    const sphfem_double2 fN = {s * n.x / dt2 * t.contact_alpha, s * n.y / dt2 * t.contact_alpha};
    sphfem_double2 fT = {0., 0.};
    if (t.g.mu != 0.) {
        const double vx = vs.x - t.g.vel.x, vy = vs.y - t.g.vel.y;
        const double vrx = vx - vx * n.x, vry = vy - vy * n.y;
        const double k = t.contact_alpha * t.slave_mass;
        const double kdx = k * vrx / dt, kdy = k * vry / dt;
        const double fy = t.g.mu * rigid_length(fN.x, fN.y);
        const sphfem_double2 fstar = {ft_old.x - kdx, ft_old.y - kdy};
        const double ls = rigid_length(fstar.x, fstar.y);
        if (ls > fy) {
            fT = {fy * fstar.x / ls, fy * fstar.y / ls};
        } else {
            fT = fstar;
        }
    }
    fc = fN;
    ft = fT;
    nrm = n;
}

// actions_gpu.cu:316-343 for one non-tool particle: friction power eta |ft| |v_rel| into T
// (|ft| and |v_rel| rounded to single precision as sqrtf does there); unchanged when ft = 0
__device__ inline double rigid_fric_heat(const DevConsts &c, sphfem_double2 vel_tool, sphfem_double2 v,
                                         sphfem_double2 ft, sphfem_double2 normal, double T) {
    const double f_mag = (double)(float)sqrt((double)(float)(ft.x * ft.x + ft.y * ft.y));
    if (f_mag == 0.) return T;
    const double vdx = v.x - vel_tool.x, vdy = v.y - vel_tool.y;
    const double vdn = vdx * normal.x + vdy * normal.y;
    const double vrx = vdx - vdn * normal.x, vry = vdy - vdn * normal.y;
    const double vr_mag = (double)(float)sqrt((double)(float)(vrx * vrx + vry * vry));
    return T + c.trml_wp.eta * c.dt * f_mag * vr_mag / (c.trml_wp.cp * c.phys.mass);
}

}  // namespace sphfem
