// Output formats of the reference (vtk_writer.cu, mesh_structure_GPU.cu:1259-1296): legacy ASCII
// VTK files whose text is what the reference's fprintf calls produce for the same values.
//
// "%f" is rendered by an exact integer conversion (below) instead of printf: the binary value
// x = m 2^e is scaled by 10^6 in 128-bit integer arithmetic and rounded to nearest, ties to even,
// which is what glibc's printf does; non-finite and very large values fall back to snprintf.
// Large particle sets are formatted in parallel chunks (std::thread) and written in order.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "sphfem_internal.h"

#pragma clang fp contract(off)

namespace sphfem {
namespace {

// decimal digits of v (v < 10^20) into out, returns the count
int put_u64(char *out, uint64_t v) {
    char tmp[24];
    int n = 0;
    do {
        tmp[n++] = (char)('0' + v % 10);
        v /= 10;
    } while (v);
    for (int i = 0; i < n; i++) out[i] = tmp[n - 1 - i];
    return n;
}

// printf("%f", x): six decimals, correctly rounded (ties to even), "-" for any negative sign bit
int fmt_f(char *out, double x) {
    uint64_t bits;
    std::memcpy(&bits, &x, sizeof(bits));
    const bool neg = (bits >> 63) != 0;
    const int be = (int)((bits >> 52) & 0x7ff);
    const uint64_t frac = bits & ((1ull << 52) - 1);
    if (be == 0x7ff) return std::snprintf(out, 48, "%f", x);
    uint64_t m;
    int e;
    if (be == 0) {
        m = frac;
        e = -1074;
    } else {
        m = frac | (1ull << 52);
        e = be - 1075;
    }
    unsigned __int128 q;
    if (e >= 0) {
        if (e > 10) return std::snprintf(out, 400, "%f", x);   // |x| >= 2^63: not in these files
        q = ((unsigned __int128)m << e) * 1000000u;
    } else {
        const int k = -e;
        if (k >= 100) {
            q = 0;                                              // m 10^6 < 2^73 < 2^(k-1)
        } else {
            const unsigned __int128 N = (unsigned __int128)m * 1000000u;
            q = N >> k;
            const unsigned __int128 r = N - (q << k);
            const unsigned __int128 half = (unsigned __int128)1 << (k - 1);
            if (r > half || (r == half && (q & 1))) q++;
        }
    }
    int n = 0;
    if (neg) out[n++] = '-';
    const uint64_t ip = (uint64_t)(q / 1000000u);
    uint32_t fp = (uint32_t)(q % 1000000u);
    n += put_u64(out + n, ip);
    out[n++] = '.';
    for (int d = 5; d >= 0; d--) {
        out[n + d] = (char)('0' + fp % 10);
        fp /= 10;
    }
    return n + 6;
}

int fmt_d(char *out, long v) {
    int n = 0;
    uint64_t u;
    if (v < 0) {
        out[n++] = '-';
        u = (uint64_t)(-(v + 1)) + 1;
    } else {
        u = (uint64_t)v;
    }
    return n + put_u64(out + n, u);
}

// a growing text buffer
struct Text {
    std::string s;
    char tmp[512];
    void f(double x) { s.append(tmp, (size_t)fmt_f(tmp, x)); }
    void d(long v) { s.append(tmp, (size_t)fmt_d(tmp, v)); }
    void c(char ch) { s.push_back(ch); }
    void lit(const char *t) { s.append(t); }
    void f3(double a, double b, double cc) {      // "%f %f %f\n"
        f(a); c(' '); f(b); c(' '); f(cc); c('\n');
    }
};

// rows [0, n) formatted by `row(Text&, i)` in parallel chunks, concatenated in order
template <class Row>
std::string rows(long n, Row row) {
    unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)(hw == 0 ? 4 : (hw > 16 ? 16 : hw));
    if (n < 20000) nt = 1;
    std::vector<Text> parts((size_t)nt);
    auto work = [&](int t) {
        const long a = n * t / nt, b = n * (t + 1) / nt;
        parts[(size_t)t].s.reserve((size_t)(b - a) * 40);
        for (long i = a; i < b; i++) row(parts[(size_t)t], i);
    };
    if (nt == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nt; t++) th.emplace_back(work, t);
        for (auto &x : th) x.join();
    }
    size_t total = 0;
    for (auto &p : parts) total += p.s.size();
    std::string out;
    out.reserve(total);
    for (auto &p : parts) out += p.s;
    return out;
}

struct File {
    std::FILE *fp = nullptr;
    explicit File(const char *path) : fp(std::fopen(path, "w+")) {}
    ~File() {
        if (fp) std::fclose(fp);
    }
    void put(const std::string &s) { std::fwrite(s.data(), 1, s.size(), fp); }
    void put(const char *s) { std::fputs(s, fp); }
    bool close() {
        bool ok = std::ferror(fp) == 0;
        ok = (std::fclose(fp) == 0) && ok;
        fp = nullptr;
        return ok;
    }
};

const char *HEADER = "# vtk DataFile Version 2.0\nmfree iwf\nASCII\n\nDATASET UNSTRUCTURED_GRID\n";

int open_fail(const char *path, std::string &err) {
    err = std::string("cannot open ") + path + " for writing";
    return SPHFEM_ERR_ARG;
}

int write_fail(const char *path, std::string &err) {
    err = std::string("write error on ") + path;
    return SPHFEM_ERR_ARG;
}

std::string header_line(const char *fmt, long a, long b = -1) {
    char buf[128];
    if (b < 0) std::snprintf(buf, sizeof(buf), fmt, a);
    else std::snprintf(buf, sizeof(buf), fmt, a, b);
    return buf;
}

}  // namespace

// vtk_writer.cu:11-219
int write_particles_vtk(const sphfem_particles &P, const char *path, bool skip_blanked, std::string &err) {
    const long N = (long)P.N;
    std::vector<long> sel;
    sel.reserve((size_t)N);
    for (long i = 0; i < N; i++)
        if (!(skip_blanked && P.blanked[i] == 1.)) sel.push_back(i);
    const long n = (long)sel.size();
    File f(path);
    if (!f.fp) return open_fail(path, err);
    f.put(HEADER);
    f.put(header_line("POINTS %ld float\n", n));
    f.put(rows(n, [&](Text &t, long k) { const long i = sel[(size_t)k]; t.f3(P.pos[i].x, P.pos[i].y, 0.); }));
    f.put("\n");
    f.put(header_line("CELLS %ld %ld\n", n, 2 * n));
    f.put(rows(n, [&](Text &t, long k) { t.lit("1 "); t.d(k); t.c('\n'); }));
    f.put("\n");
    f.put(header_line("CELL_TYPES %ld\n", n));
    f.put(rows(n, [&](Text &t, long) { t.lit("1\n"); }));
    f.put("\n");
    f.put(header_line("POINT_DATA %ld\n", n));
    f.put("SCALARS Temperature float 1\nLOOKUP_TABLE default\n");
    f.put(rows(n, [&](Text &t, long k) { t.f(P.T[sel[(size_t)k]]); t.c('\n'); }));
    f.put("\n");
    f.put("SCALARS on_seg int 1\nLOOKUP_TABLE default\n");
    f.put(rows(n, [&](Text &t, long k) { t.d(P.on_seg[sel[(size_t)k]]); t.c('\n'); }));
    f.put("\n");
    f.put("SCALARS EquivAccumPlasticStrain float 1\nLOOKUP_TABLE default\n");
    f.put(rows(n, [&](Text &t, long k) { t.f(P.eps_pl[sel[(size_t)k]]); t.c('\n'); }));
    f.put("\n");
    f.put("VECTORS Velocity float\n");
    f.put(rows(n, [&](Text &t, long k) { const long i = sel[(size_t)k]; t.f3(P.vel[i].x, P.vel[i].y, 0.); }));
    f.put("\n");
    // "%2f": a minimum width of 2 never pads a six-decimal number
    f.put("VECTORS fc float\n");
    f.put(rows(n, [&](Text &t, long k) {
        const long i = sel[(size_t)k];
        t.f3(100000 * P.fc[i].x, 100000 * P.fc[i].y, 0.);
    }));
    f.put("\n");
    f.put("VECTORS ft float\n");
    f.put(rows(n, [&](Text &t, long k) {
        const long i = sel[(size_t)k];
        t.f3(100000 * P.ft[i].x, 100000 * P.ft[i].y, 0.);
    }));
    f.put("\n");
    f.put("VECTORS normal float\n");
    f.put(rows(n, [&](Text &t, long k) { const long i = sel[(size_t)k]; t.f3(P.n[i].x, P.n[i].y, 0.); }));
    f.put("\n");
    f.put("SCALARS SvM float 1\nLOOKUP_TABLE default\n");
    f.put(rows(n, [&](Text &t, long k) {
        const long i = sel[(size_t)k];
        const double p = P.p[i];
        const sphfem_double4 S = P.S[i];
        const double cxx = S.x - p, cyy = S.z - p, czz = S.w - p, cxy = S.y;
        const double svm2 = (cxx * cxx + cyy * cyy + czz * czz) - cxx * cyy - cxx * czz - cyy * czz + 3.0 * cxy * cxy;
        t.f(svm2 > 0 ? std::sqrt(svm2) : 0.);
        t.c('\n');
    }));
    f.put("\n");
    if (!f.close()) return write_fail(path, err);
    return SPHFEM_OK;
}

// vtk_writer.cu:222-302
int write_rigid_tool_vtk(const sphfem_tool_geometry &g, const char *path, std::string &err) {
    if (g.num_seg == 0) return SPHFEM_OK;
    if (g.num_seg != 4 && g.num_seg != 5) {
        err = "a rigid tool has 4 or 5 segments";
        return SPHFEM_ERR_ARG;
    }
    struct Tri { sphfem_double2 a, b, c; };
    std::vector<Tri> tris;
    const sphfem_segment *s = g.seg;
    if (g.num_seg == 4) {
        tris.push_back({s[0].left, s[0].right, s[1].right});
        tris.push_back({s[2].left, s[2].right, s[3].right});
    } else {
        tris.push_back({s[0].left, s[0].right, s[2].right});
        tris.push_back({s[1].left, s[1].right, s[2].right});
        tris.push_back({s[3].left, s[3].right, s[4].right});
    }
    if (g.has_fillet) {
        // fan of 19 triangles from the centre; the start angle lo = t1 - 0.1 t1 and the (sin, cos)
        // order of the offsets are the reference's
        const int num_discr = 20;
        const double t1 = std::fmin(g.fillet.t1, g.fillet.t2), t2 = std::fmax(g.fillet.t1, g.fillet.t2);
        const double lo = t1 - 0.1 * t1;
        const double d_angle = (t2 - t1) / (num_discr - 1);
        const double r = g.fillet.r;
        for (int i = 0; i < num_discr - 1; i++) {
            const double a1 = lo + (i + 0) * d_angle, a2 = lo + (i + 1) * d_angle;
            const sphfem_double2 p1 = g.fillet.p;
            tris.push_back({p1, {p1.x + r * std::sin(a1), p1.y + r * std::cos(a1)},
                            {p1.x + r * std::sin(a2), p1.y + r * std::cos(a2)}});
        }
    }
    const long nt = (long)tris.size();
    File f(path);
    if (!f.fp) return open_fail(path, err);
    Text t;
    t.lit(HEADER);
    t.lit(header_line("POINTS %ld float\n", 3 * nt).c_str());
    for (const Tri &x : tris) {
        t.f3(x.a.x, x.a.y, 0.);
        t.f3(x.b.x, x.b.y, 0.);
        t.f3(x.c.x, x.c.y, 0.);
    }
    t.c('\n');
    t.lit(header_line("CELLS %ld %ld\n", nt, 3 * nt + nt).c_str());
    for (long i = 0; i < nt; i++) {
        t.lit("3 "); t.d(3 * i + 0); t.c(' '); t.d(3 * i + 1); t.c(' '); t.d(3 * i + 2); t.c('\n');
    }
    t.c('\n');
    t.lit(header_line("CELL_TYPES %ld\n", nt).c_str());
    for (long i = 0; i < nt; i++) t.lit("5\n");
    f.put(t.s);
    if (!f.close()) return write_fail(path, err);
    return SPHFEM_OK;
}

// vtk_writer.cu:304-366 (shift = 0) and :518-587 (shift = tl - tl_ref subtracted from the nodes)
int write_fem_tool_vtk(const FemToolHost &T, sphfem_double2 shift, const char *path, std::string &err) {
    const long nn = T.n_nodes, ne = T.n_elem;
    File f(path);
    if (!f.fp) return open_fail(path, err);
    f.put(HEADER);
    f.put(header_line("POINTS %ld float\n", nn));
    f.put(rows(nn, [&](Text &t, long i) { t.f3(T.pos[(size_t)i].x - shift.x, T.pos[(size_t)i].y - shift.y, 0.); }));
    f.put("\n");
    f.put(header_line("CELLS %ld %ld\n", ne, 3 * ne + ne));
    f.put(rows(ne, [&](Text &t, long i) {
        const sphfem_int4 e = T.elem[(size_t)i];
        t.lit("3 "); t.d(e.x); t.c(' '); t.d(e.y); t.c(' '); t.d(e.z); t.c('\n');
    }));
    f.put("\n");
    f.put(header_line("CELL_TYPES %ld\n", ne));
    f.put(rows(ne, [&](Text &t, long) { t.lit("5\n"); }));
    f.put(header_line("POINT_DATA %ld\n", nn));
    f.put("SCALARS temperature float 1\nLOOKUP_TABLE default\n");
    f.put(rows(nn, [&](Text &t, long i) { t.f(T.T[(size_t)i]); t.c('\n'); }));
    f.put("\n");
    if (!f.close()) return write_fail(path, err);
    return SPHFEM_OK;
}

namespace {
// one variable of the segment files: per-segment value when the contact count pp.x is non-zero
template <class V>
void seg_scalar(Text &t, const FemToolHost &T, const char *name, int zeros, V value) {
    t.lit("SCALARS ");
    t.lit(name);
    t.lit(" float 1\nLOOKUP_TABLE default\n");
    for (int i = 0; i < T.n_seg; i++) {
        for (int z = 0; z < zeros; z++) t.lit("0.000000\n");
        t.f(T.physical_para[(size_t)i].x != 0. ? value(i) : 0.);
        t.c('\n');
    }
    t.c('\n');
}
}  // namespace

// vtk_writer.cu:368-516
int write_fem_segments_vtk(const FemToolHost &T, const char *path, std::string &err) {
    const long ns = T.n_seg;
    Text t;
    t.lit(HEADER);
    t.lit(header_line("POINTS %ld float\n", 3 * ns).c_str());
    for (long i = 0; i < ns; i++) {
        const sphfem_double2 l = T.left[(size_t)i], r = T.right[(size_t)i], n = T.n[(size_t)i];
        t.f3(l.x, l.y, 0.);
        t.f3(r.x, r.y, 0.);
        const double dis = std::sqrt((l.x - r.x) * (l.x - r.x) + (l.y - r.y) * (l.y - r.y));
        t.f3(r.x + dis * n.x, r.y + dis * n.y, 0.);
    }
    t.c('\n');
    t.lit(header_line("CELLS %ld %ld\n", ns, 3 * ns + ns).c_str());
    for (long i = 0; i < ns; i++) {
        t.lit("3 "); t.d(3 * i); t.c(' '); t.d(3 * i + 1); t.c(' '); t.d(3 * i + 2); t.lit(" \n");
    }
    t.c('\n');
    t.lit(header_line("CELL_TYPES %ld\n", ns).c_str());
    for (long i = 0; i < ns; i++) t.lit("5\n");
    t.lit(header_line("POINT_DATA %ld\n", 3 * ns).c_str());
    const auto &pp = T.physical_para;
    seg_scalar(t, T, "PressureGPa", 2, [&](int i) { return pp[(size_t)i].y / pp[(size_t)i].x * 100.; });
    seg_scalar(t, T, "Velocity", 2, [&](int i) { return pp[(size_t)i].z / pp[(size_t)i].x * 10000.; });
    seg_scalar(t, T, "Temperature", 2, [&](int i) { return pp[(size_t)i].w / pp[(size_t)i].x; });
    seg_scalar(t, T, "Frictional_stress", 2,
               [&](int i) { return T.sliding_force[(size_t)i] / pp[(size_t)i].x * 100.; });
    seg_scalar(t, T, "Wear_rate", 2, [&](int i) { return T.wear_nodes[(size_t)i] * 1e10; });
    File f(path);
    if (!f.fp) return open_fail(path, err);
    f.put(t.s);
    if (!f.close()) return write_fail(path, err);
    return SPHFEM_OK;
}

// vtk_writer.cu:589-714
int write_contact_variables_vtk(const FemToolHost &T, const char *path, std::string &err) {
    const long ns = T.n_seg;
    Text t;
    t.lit(HEADER);
    t.lit(header_line("POINTS %ld float\n", ns).c_str());
    for (long i = 0; i < ns; i++) {
        const sphfem_double2 l = T.left[(size_t)i], r = T.right[(size_t)i];
        t.f3(0.5 * (l.x + r.x), 0.5 * (l.y + r.y), 0.);
    }
    t.c('\n');
    t.lit(header_line("CELLS %ld %ld\n", ns, 2 * ns).c_str());
    for (long i = 0; i < ns; i++) {
        t.lit("1 "); t.d(i); t.c('\n');
    }
    t.c('\n');
    t.lit(header_line("CELL_TYPES %ld\n", ns).c_str());
    for (long i = 0; i < ns; i++) t.lit("1\n");
    t.c('\n');
    t.lit(header_line("POINT_DATA %ld\n", ns).c_str());
    const auto &pp = T.physical_para;
    seg_scalar(t, T, "PressureGPa", 0, [&](int i) { return 100. * pp[(size_t)i].y / pp[(size_t)i].x; });
    seg_scalar(t, T, "Velocity", 0, [&](int i) { return pp[(size_t)i].z / pp[(size_t)i].x * 10000.; });
    seg_scalar(t, T, "Temperature", 0, [&](int i) { return pp[(size_t)i].w / pp[(size_t)i].x; });
    seg_scalar(t, T, "Frictional_stress", 0,
               [&](int i) { return T.sliding_force[(size_t)i] / pp[(size_t)i].x * 100.; });
    seg_scalar(t, T, "Wear_rate", 2, [&](int i) { return T.wear_nodes[(size_t)i] * 1e10; });
    File f(path);
    if (!f.fp) return open_fail(path, err);
    f.put(t.s);
    if (!f.close()) return write_fail(path, err);
    return SPHFEM_OK;
}

// mesh_structure_GPU.cu:1259-1296
int read_vtk_temperature(const char *path, std::vector<double> &T, std::string &err) {
    std::ifstream in(path);
    if (!in) {
        err = std::string("cannot open ") + path;
        return SPHFEM_ERR_ARG;
    }
    std::string line;
    bool started = false;
    T.clear();
    while (std::getline(in, line)) {
        if (line.find("LOOKUP_TABLE default") != std::string::npos) {
            started = true;
            continue;
        }
        if (!started || line.empty()) continue;
        const char *b = line.c_str();
        char *e = nullptr;
        const double v = std::strtod(b, &e);
        if (e == b) {
            err = std::string("not a number in ") + path + ": " + line;
            return SPHFEM_ERR_ARG;
        }
        T.push_back(v);
    }
    return SPHFEM_OK;
}

}  // namespace sphfem
