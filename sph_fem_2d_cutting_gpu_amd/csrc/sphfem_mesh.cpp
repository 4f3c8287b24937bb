// Tool mesh pipeline on the host: the .inp readers of mesh_structure_GPU.cu (Abaqus-style
// mesh_read(file), :784-1257, and the Gmsh export mesh_read(file, true), :213-783), element
// colouring and the CSR pattern, the mesh VTK file (:1395-1462), the cutting segments of
// tool_FEM::set_up_segments_from_mesh (tool_FEM.cu:771-1039) and the FEM system that
// FEM_thermal_2D_GPU sets up (FEM_thermal_gpu.cu:126-330, 584-659): P1 stiffness and lumped mass
// assembled colour by colour, Dirichlet rows on the fixed boundary.
//
// Colouring follows the reference's first-free-colour search, cycling through the palette, but
// starts every search at colour 1 instead of rand() % palette + 1, so the result is reproducible.
// Neighbour lookups go through a node -> element table (the reference scans every element).
#include <algorithm>
#include <array>
#include <cerrno>
#include <cfloat>
#include <climits>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/sphfem.h"

#pragma clang fp contract(off)

struct sphfem_mesh {
    std::vector<double> x, y;
    std::vector<int> cut_node, fixed_node;
    std::vector<sphfem_int4> elem;                 // 3 node ids + colour
    std::vector<std::array<int, 3>> cut_flags, fixed_flags;
    std::vector<int> cut_elems, fixed_elems;
    std::vector<std::array<int, 2>> layers;        // boundary layer 1 / 2
    std::vector<int> row_ptr, col_ind;
    int color_total = 18, color_cutting_edge = 6;
    std::vector<double> T_init;                    // hot tool (mesh_read_temperature), empty if none
    // arrays of the last sphfem_mesh_tool_desc
    std::vector<sphfem_double2> d_pos, d_left, d_right, d_n;
    std::vector<sphfem_int4> d_marks, d_fixed_flags;
    std::vector<int> d_layer, d_node_flags;
    std::vector<double> d_K, d_M, d_f, d_T;
};

namespace {

// the tokens std::getline(ss, tok, sep) yields: no empty token after a final separator
std::vector<std::string> split(const std::string &line, char sep) {
    std::vector<std::string> out;
    size_t a = 0;
    while (a < line.size()) {
        const size_t b = line.find(sep, a);
        if (b == std::string::npos) {
            out.push_back(line.substr(a));
            break;
        }
        out.push_back(line.substr(a, b - a));
        a = b + 1;
    }
    return out;
}

// the integer at the start of a token (as `stringstream >> int`: leading blanks, a sign, digits);
// false if there is none or it does not fit an int
// ids and counts of a mesh file stay within +-2^30, so that id arithmetic (id - 1 - jumps) cannot
// overflow; larger values are parse errors
bool to_int(const std::string &tok, int &v) {
    const char *b = tok.c_str();
    char *e = nullptr;
    errno = 0;
    const long x = std::strtol(b, &e, 10);
    if (e == b || errno == ERANGE || x < -(1L << 30) || x > (1L << 30)) return false;
    v = (int)x;
    return true;
}

bool is_blank(const std::string &tok) {
    return tok.find_first_not_of(" \t\r") == std::string::npos;
}

// the double at the start of a string (as std::stod)
bool to_double(const std::string &tok, double &v) {
    const char *b = tok.c_str();
    char *e = nullptr;
    v = std::strtod(b, &e);
    return e != b;
}

struct ParseError {
    std::string msg;
};

// mesh_structure_GPU.cu:857-876 / 922-941: an element whose three edges are all flagged keeps only
// the edges next to nodes it shares with another flagged element
void correct_flags(sphfem_mesh &m, std::vector<std::array<int, 3>> &flags, const std::vector<int> &corr,
                   const std::vector<int> &list) {
    for (int a : corr) {
        const int na[3] = {m.elem[a].x, m.elem[a].y, m.elem[a].z};
        const int clear_edge[3] = {1, 2, 0};      // node x -> flag y, node y -> flag z, node z -> flag x
        for (int k = 0; k < 3; k++) {
            bool shared = false;
            for (int b : list) {
                if (b == a) continue;
                const sphfem_int4 e = m.elem[b];
                if (na[k] == e.x || na[k] == e.y || na[k] == e.z) {
                    shared = true;
                    break;
                }
            }
            if (!shared) flags[a][clear_edge[k]] = 0;
        }
    }
}

// mesh_structure_GPU.cu:784-1050
void read_abaqus(sphfem_mesh &m, std::istream &in) {
    bool node_start = false, elem_start = false, cut_nset = false, fixed_nset = false, cut_elset = false,
         fixed_elset = false;
    std::vector<double> x_, y_;
    std::vector<int> cut_corr, fixed_corr;
    std::string line;
    while (std::getline(in, line)) {
        if (line.find('#') != std::string::npos) continue;
        if (line.find("*Node") != std::string::npos) {
            node_start = true;
            continue;
        }
        if (line.find("Element") != std::string::npos) {
            elem_start = true;
            node_start = false;
            const size_t n = x_.size();
            m.x.resize(n);
            m.y.resize(n);
            m.cut_node.assign(n, 0);
            m.fixed_node.assign(n, 0);
            for (size_t i = 0; i < n; i++) {
                m.x[i] = x_[i] / 100;
                m.y[i] = y_[i] / 100;
            }
            continue;
        }
        const bool has_cut = line.find("CuttingEdge") != std::string::npos;
        const bool has_fixed = line.find("FixedBoundary") != std::string::npos;
        if (has_cut && line.find("Nset") != std::string::npos) {
            cut_nset = true;
            elem_start = false;
            continue;
        }
        if (has_cut && line.find("Elset") != std::string::npos) {
            cut_elset = true;
            cut_nset = false;
            continue;
        }
        if (has_fixed && line.find("Nset") != std::string::npos) {
            fixed_nset = true;
            cut_elset = false;
            correct_flags(m, m.cut_flags, cut_corr, m.cut_elems);
            continue;
        }
        if (has_fixed && line.find("Elset") != std::string::npos) {
            fixed_elset = true;
            fixed_nset = false;
            continue;
        }
        if (line.find("End") != std::string::npos) {
            correct_flags(m, m.fixed_flags, fixed_corr, m.fixed_elems);
            break;
        }
        if (node_start) {
            const auto tok = split(line, ',');
            double x, y;
            int id;
            if (tok.size() < 3 || !to_int(tok[0], id) || !to_double(tok[1], x)) throw ParseError{"node line: " + line};
            std::string rest = line.substr(tok[0].size() + 1 + tok[1].size() + 1);
            if (!to_double(rest, y)) throw ParseError{"node line: " + line};
            x_.push_back(x);
            y_.push_back(y);
        }
        if (elem_start) {
            const auto tok = split(line, ',');
            int a[3];
            if (tok.size() < 4 || !to_int(tok[1], a[0]) || !to_int(tok[2], a[1]) || !to_int(tok[3], a[2]))
                throw ParseError{"element line: " + line};
            for (int k = 0; k < 3; k++)
                if (a[k] < 1 || a[k] > (int)m.x.size()) throw ParseError{"element node out of range: " + line};
            m.elem.push_back({a[0] - 1, a[1] - 1, a[2] - 1, 0});
            m.cut_flags.push_back({0, 0, 0});
            m.fixed_flags.push_back({0, 0, 0});
        }
        // node and element sets: comma-separated ids (empty tokens, e.g. after a trailing comma, skipped)
        auto ids = [&](const std::string &l) {
            std::vector<int> v;
            for (const auto &t : split(l, ',')) {
                if (is_blank(t)) continue;
                int a;
                if (!to_int(t, a)) throw ParseError{"set id: " + l};
                v.push_back(a);
            }
            return v;
        };
        if (cut_nset || fixed_nset) {
            for (int a : ids(line)) {
                if (a < 1 || a > (int)m.x.size()) throw ParseError{"node set id out of range: " + line};
                (cut_nset ? m.cut_node : m.fixed_node)[a - 1] = 1;
            }
        }
        if (cut_elset || fixed_elset) {
            auto &flags = cut_elset ? m.cut_flags : m.fixed_flags;
            auto &list = cut_elset ? m.cut_elems : m.fixed_elems;
            auto &corr = cut_elset ? cut_corr : fixed_corr;
            const auto &on = cut_elset ? m.cut_node : m.fixed_node;
            for (int a : ids(line)) {
                if (a < 1 || a > (int)m.elem.size()) throw ParseError{"element set id out of range: " + line};
                const int e = a - 1;
                list.push_back(e);
                const sphfem_int4 el = m.elem[e];
                const bool i1 = on[el.x] != 0, i2 = on[el.y] != 0, i3 = on[el.z] != 0;
                if (i1 && i2) flags[e][0] = 1;
                if (i2 && i3) flags[e][1] = 1;
                if (i3 && i1) flags[e][2] = 1;
                if (flags[e][0] == 1 && flags[e][1] == 1 && flags[e][2] == 1) corr.push_back(e);
            }
        }
    }
}

// mesh_structure_GPU.cu:12-19
int number_to_jump(int a, const std::vector<int> &exclude) {
    int j = 0;
    for (int e : exclude)
        if (e < a) j++;
    return j;
}

// mesh_structure_GPU.cu:213-594 (three passes over the file)
void read_gmsh(sphfem_mesh &m, const std::string &text) {
    std::set<int> used;     // node ids (0-based) referenced by the triangles
    {
        std::istringstream in(text);
        std::string line;
        bool elem_start = false;
        while (std::getline(in, line)) {
            if (line.find("type=CPS3") != std::string::npos) {
                elem_start = true;
                continue;
            }
            if (line.find("FixedBoundary") != std::string::npos) break;
            if (elem_start) {
                const auto tok = split(line, ',');
                for (int k = 1; k < 4; k++) {
                    int a;
                    if ((int)tok.size() <= k || !to_int(tok[k], a)) throw ParseError{"element line: " + line};
                    used.insert(a - 1);
                }
            }
        }
    }
    std::vector<int> exclude, seq;
    std::vector<double> x_, y_;
    std::vector<std::array<int, 2>> boundary_edge, cut_edges, fixed_edges;
    {
        std::istringstream in(text);
        std::string line;
        bool node_start = false, edge_start = false, cut_start = false, fixed_start = false;
        auto all_off = [&] { node_start = edge_start = cut_start = fixed_start = false; };
        while (std::getline(in, line)) {
            if (line.find('#') != std::string::npos) continue;
            if (line.find("*NODE") != std::string::npos) {
                node_start = true;
                continue;
            }
            if (line.find("T3D2") != std::string::npos) {
                all_off();
                edge_start = true;
                continue;
            }
            if (line.find("type=CPS3") != std::string::npos) {
                all_off();
                continue;
            }
            if (line.find("*******") != std::string::npos) {
                all_off();
                const size_t n = x_.size();
                m.x.assign(n, 0.);
                m.y.assign(n, 0.);
                m.cut_node.assign(n, 0);
                m.fixed_node.assign(n, 0);
                for (size_t i = 0; i < n; i++) {
                    if (seq[i] < 0 || seq[i] >= (int)n) throw ParseError{"node numbering is not contiguous"};
                    m.x[(size_t)seq[i]] = x_[i];
                    m.y[(size_t)seq[i]] = y_[i];
                }
                continue;
            }
            if (line.find("CuttingEdge") != std::string::npos) {
                all_off();
                cut_start = true;
                continue;
            }
            if (line.find("FixedBoundary") != std::string::npos) {
                all_off();
                fixed_start = true;
                continue;
            }
            if (node_start) {
                const auto tok = split(line, ',');
                int a;
                double x, y;
                if (tok.size() < 3 || !to_int(tok[0], a)) throw ParseError{"node line: " + line};
                if (!used.count(a - 1)) {
                    exclude.push_back(a);
                    continue;
                }
                if (!to_double(tok[1], x) || !to_double(tok[2], y)) throw ParseError{"node line: " + line};
                seq.push_back(a - 1 - number_to_jump(a, exclude));
                x_.push_back(x);
                y_.push_back(y);
            }
            if (edge_start) {
                const auto tok = split(line, ',');
                int a, b;
                if (tok.size() < 3 || !to_int(tok[1], a) || !to_int(tok[2], b)) throw ParseError{"edge line: " + line};
                boundary_edge.push_back({a - 1 - number_to_jump(a, exclude), b - 1 - number_to_jump(b, exclude)});
            }
            if (cut_start || fixed_start) {
                for (const auto &t : split(line, ',')) {
                    if (is_blank(t)) continue;
                    int a;
                    if (!to_int(t, a)) throw ParseError{"edge set id: " + line};
                    if (a < 1 || a > (int)boundary_edge.size()) throw ParseError{"edge set id out of range: " + line};
                    const auto be = boundary_edge[(size_t)a - 1];
                    // the ids are renumbered a second time here, as in the reference (:424-427); a no-op
                    // when no node was left out
                    const std::array<int, 2> e = {be[0] - number_to_jump(be[0], exclude),
                                                  be[1] - number_to_jump(be[1], exclude)};
                    if (e[0] < 0 || e[1] < 0 || e[0] >= (int)m.x.size() || e[1] >= (int)m.x.size())
                        throw ParseError{"edge node out of range: " + line};
                    (cut_start ? m.cut_node : m.fixed_node)[(size_t)e[0]] = 1;
                    (cut_start ? m.cut_node : m.fixed_node)[(size_t)e[1]] = 1;
                    (cut_start ? cut_edges : fixed_edges).push_back(e);
                }
            }
        }
    }
    {
        std::istringstream in(text);
        std::string line;
        bool elem_start = false;
        // the boundary edges as unordered node pairs
        auto key = [](int u, int v) { return ((long long)std::min(u, v) << 32) | (unsigned)std::max(u, v); };
        std::set<long long> cut_set, fixed_set;
        for (const auto &e : cut_edges) cut_set.insert(key(e[0], e[1]));
        for (const auto &e : fixed_edges) fixed_set.insert(key(e[0], e[1]));
        auto matches = [&](const std::vector<std::array<int, 2>> &edges, int u, int v) {
            return (&edges == &cut_edges ? cut_set : fixed_set).count(key(u, v)) > 0;
        };
        while (std::getline(in, line)) {
            if (line.find("type=CPS3") != std::string::npos) {
                elem_start = true;
                continue;
            }
            if (line.find("FixedBoundary") != std::string::npos) break;
            if (!elem_start) continue;
            const auto tok = split(line, ',');
            if (tok.size() < 4) throw ParseError{"element line: " + line};
            int a[3];
            for (int k = 0; k < 3; k++) {
                if (!to_int(tok[(size_t)k + 1], a[k])) throw ParseError{"element line: " + line};
                a[k] = a[k] - 1 - number_to_jump(a[k], exclude);
                if (a[k] < 0 || a[k] >= (int)m.x.size()) throw ParseError{"element node out of range: " + line};
            }
            const int e = (int)m.elem.size();
            m.elem.push_back({a[0], a[1], a[2], 0});
            std::array<int, 3> fc = {matches(cut_edges, a[0], a[1]) ? 1 : 0, matches(cut_edges, a[1], a[2]) ? 1 : 0,
                                     matches(cut_edges, a[2], a[0]) ? 1 : 0};
            std::array<int, 3> ff = {matches(fixed_edges, a[0], a[1]) ? 1 : 0, matches(fixed_edges, a[1], a[2]) ? 1 : 0,
                                     matches(fixed_edges, a[2], a[0]) ? 1 : 0};
            m.cut_flags.push_back(fc);
            m.fixed_flags.push_back(ff);
            if (fc[0] || fc[1] || fc[2]) m.cut_elems.push_back(e);
            if (ff[0] || ff[1] || ff[2]) m.fixed_elems.push_back(e);
        }
    }
}

// mesh_structure_GPU.cu:1059-1225: boundary layers, colours, CSR
void finish_mesh(sphfem_mesh &m) {
    const int N = (int)m.elem.size(), nn = (int)m.x.size();
    m.layers.assign((size_t)N, {0, 0});
    std::vector<std::vector<int>> node_elems((size_t)nn);
    for (int e = 0; e < N; e++) {
        const sphfem_int4 t = m.elem[(size_t)e];
        for (int v : {t.x, t.y, t.z}) node_elems[(size_t)v].push_back(e);
    }
    auto shares_colour = [&](int e, int c, const std::vector<char> *only) {
        const sphfem_int4 t = m.elem[(size_t)e];
        for (int v : {t.x, t.y, t.z})
            for (int o : node_elems[(size_t)v])
                if (o != e && (!only || (*only)[(size_t)o]) && m.elem[(size_t)o].w == c) return true;
        return false;
    };
    auto colour = [&](int e, int palette, const std::vector<char> *only) {
        int c = 1, loops = 0;
        while (shares_colour(e, c, only)) {
            c = c + 1 > palette ? 1 : c + 1;
            if (++loops >= palette) throw ParseError{"the colour palette is too small for this mesh"};
        }
        m.elem[(size_t)e].w = c;
    };
    std::vector<char> is_cut((size_t)N, 0);
    for (int e : m.cut_elems) is_cut[(size_t)e] = 1;
    for (int e : m.cut_elems) {
        m.layers[(size_t)e][0] = 1;
        colour(e, m.color_cutting_edge, &is_cut);
    }
    // layer 2: elements touching a cutting-edge node or a node of a layer-1 element
    std::vector<char> layer1_node((size_t)nn, 0);
    for (int e : m.cut_elems) {
        const sphfem_int4 t = m.elem[(size_t)e];
        layer1_node[(size_t)t.x] = layer1_node[(size_t)t.y] = layer1_node[(size_t)t.z] = 1;
    }
    for (int e = 0; e < N; e++) {
        if (m.layers[(size_t)e][0] == 1) continue;
        const sphfem_int4 t = m.elem[(size_t)e];
        bool touch = false;
        for (int v : {t.x, t.y, t.z}) touch = touch || m.cut_node[(size_t)v] || layer1_node[(size_t)v];
        if (!touch) continue;
        m.layers[(size_t)e][1] = 1;
        colour(e, m.color_total, nullptr);
    }
    for (int e = 0; e < N; e++)
        if (m.elem[(size_t)e].w == 0) colour(e, m.color_total, nullptr);
    // CSR: the sorted set of nodes sharing an element with each node
    m.row_ptr.assign((size_t)nn + 1, 0);
    m.col_ind.clear();
    for (int k = 0; k < nn; k++) {
        std::vector<int> cols;
        for (int e : node_elems[(size_t)k]) {
            const sphfem_int4 t = m.elem[(size_t)e];
            cols.push_back(t.x);
            cols.push_back(t.y);
            cols.push_back(t.z);
        }
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
        m.col_ind.insert(m.col_ind.end(), cols.begin(), cols.end());
        m.row_ptr[(size_t)k + 1] = (int)m.col_ind.size();
    }
}

// FEM_thermal_gpu.cu:71-152 for one triangle (the same operations as the device assembly)
void element_matrices(const sphfem_mesh &m, const sphfem_int4 &t, double cp, double rho, double k, double mass[3],
                      double Ke[3][3]) {
    const double x0 = m.x[(size_t)t.x], y0 = m.y[(size_t)t.x];
    const double j00 = m.x[(size_t)t.y] - x0, j01 = m.y[(size_t)t.y] - y0;
    const double j10 = m.x[(size_t)t.z] - x0, j11 = m.y[(size_t)t.z] - y0;
    const double det = std::fabs(j00 * j11 - j10 * j01);
    const double w = 1.0 / 3.0 * 0.5;
    const double ms = w * cp * rho * det;
    mass[0] = mass[1] = mass[2] = ms;
    const double dd = j00 * j11 - j01 * j10;
    const double ood = 1.0 / dd;
    const double i00 = j11 * ood, i01 = -j10 * ood, i10 = -j01 * ood, i11 = j00 * ood;
    const double gref[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
    double G[3][2];
    for (int a = 0; a < 3; a++) {
        G[a][0] = i00 * gref[a][0] + i10 * gref[a][1];
        G[a][1] = i01 * gref[a][0] + i11 * gref[a][1];
    }
    const double s = (w * 3.0) * (k * det);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) Ke[a][b] = (s * G[a][0]) * G[b][0] + (s * G[a][1]) * G[b][1];
}

int csr_find(const sphfem_mesh &m, int row, int col) {
    for (int i = m.row_ptr[(size_t)row]; i < m.row_ptr[(size_t)row + 1]; i++)
        if (m.col_ind[(size_t)i] == col) return i;
    return m.row_ptr[(size_t)row + 1];
}

// segment(left, right, tl) (geometric_primitives.cpp:115-137): normal (dy, -dx) normalised, both
// flipped (with the ends) when it points away from tl
struct Seg {
    sphfem_double2 left, right, n;
};

Seg make_seg(sphfem_double2 l, sphfem_double2 r, sphfem_double2 tl) {
    const double dx = r.x - l.x, dy = r.y - l.y;
    const double s = 1. / std::sqrt(dy * dy + (-dx) * (-dx));
    const sphfem_double2 n = {dy * s, -dx * s};
    const double refx = tl.x - l.x, refy = tl.y - l.y;
    if ((refx * n.x + refy * n.y) < 0) return {r, l, {-n.x, -n.y}};
    return {l, r, n};
}

double dist(sphfem_double2 a, sphfem_double2 b) {
    const double dx = a.x - b.x, dy = a.y - b.y;
    return std::sqrt(dx * dx + dy * dy);
}

bool eq(sphfem_double2 a, sphfem_double2 b) { return a.x == b.x && a.y == b.y; }

}  // namespace

extern "C" {

int sphfem_mesh_read_inp(const char *path, int gmsh_export, int color_total, int color_cutting_edge,
                         sphfem_mesh **out) {
    if (!path || !out || color_total < 1 || color_cutting_edge < 1) return SPHFEM_ERR_ARG;
    *out = nullptr;
    std::ifstream f(path);
    if (!f) return SPHFEM_ERR_ARG;
    std::stringstream buf;
    buf << f.rdbuf();
    auto *m = new sphfem_mesh();
    m->color_total = color_total;
    m->color_cutting_edge = color_cutting_edge;
    try {
        if (gmsh_export) {
            read_gmsh(*m, buf.str());
        } else {
            std::istringstream in(buf.str());
            read_abaqus(*m, in);
        }
        if (m->x.empty() || m->elem.empty()) throw ParseError{"no nodes or elements"};
        finish_mesh(*m);
    } catch (const ParseError &e) {
        std::fprintf(stderr, "sphfem_mesh_read_inp: %s: %s\n", path, e.msg.c_str());
        delete m;
        return SPHFEM_ERR_ARG;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "sphfem_mesh_read_inp: %s: %s\n", path, e.what());
        delete m;
        return SPHFEM_ERR_ARG;
    }
    *out = m;
    return SPHFEM_OK;
}

int sphfem_mesh_destroy(sphfem_mesh *m) {
    if (!m) return SPHFEM_ERR_ARG;
    delete m;
    return SPHFEM_OK;
}

int sphfem_mesh_get_info(const sphfem_mesh *m, sphfem_mesh_info *info) {
    if (!m || !info) return SPHFEM_ERR_ARG;
    info->node_num = (int)m->x.size();
    info->ele_num = (int)m->elem.size();
    info->nz_num = (int)m->col_ind.size();
    info->cut_ele_num = (int)m->cut_elems.size();
    info->fixed_ele_num = (int)m->fixed_elems.size();
    info->color_total = m->color_total;
    info->color_cutting_edge = m->color_cutting_edge;
    return SPHFEM_OK;
}

int sphfem_mesh_get_arrays(const sphfem_mesh *m, sphfem_double2 *position, int *cut_node, int *fixed_node,
                           sphfem_int4 *elem, sphfem_int4 *cut_flags, sphfem_int4 *fixed_flags, int *cut_elems,
                           int *fixed_elems, int *layers, int *row_ptr, int *col_ind) {
    if (!m) return SPHFEM_ERR_ARG;
    const size_t nn = m->x.size(), ne = m->elem.size();
    for (size_t i = 0; i < nn; i++) {
        if (position) position[i] = {m->x[i], m->y[i]};
        if (cut_node) cut_node[i] = m->cut_node[i];
        if (fixed_node) fixed_node[i] = m->fixed_node[i];
    }
    for (size_t e = 0; e < ne; e++) {
        if (elem) elem[e] = m->elem[e];
        if (cut_flags) cut_flags[e] = {m->cut_flags[e][0], m->cut_flags[e][1], m->cut_flags[e][2], 0};
        if (fixed_flags) fixed_flags[e] = {m->fixed_flags[e][0], m->fixed_flags[e][1], m->fixed_flags[e][2], 0};
        if (layers) {
            layers[2 * e] = m->layers[e][0];
            layers[2 * e + 1] = m->layers[e][1];
        }
    }
    if (cut_elems) std::copy(m->cut_elems.begin(), m->cut_elems.end(), cut_elems);
    if (fixed_elems) std::copy(m->fixed_elems.begin(), m->fixed_elems.end(), fixed_elems);
    if (row_ptr) std::copy(m->row_ptr.begin(), m->row_ptr.end(), row_ptr);
    if (col_ind) std::copy(m->col_ind.begin(), m->col_ind.end(), col_ind);
    return SPHFEM_OK;
}

int sphfem_mesh_set_temperature(sphfem_mesh *m, const double *T, long n) {
    if (!m) return SPHFEM_ERR_ARG;
    if (!T) {
        m->T_init.clear();
        return SPHFEM_OK;
    }
    if (n != (long)m->x.size()) return SPHFEM_ERR_ARG;    // the reference asserts n == node count
    m->T_init.assign(T, T + n);
    return SPHFEM_OK;
}

// mesh_structure_GPU.cu:1395-1462
int sphfem_mesh_write_vtk(const sphfem_mesh *m, const char *path) {
    if (!m || !path) return SPHFEM_ERR_ARG;
    std::FILE *fp = std::fopen(path, "w+");
    if (!fp) return SPHFEM_ERR_ARG;
    const int nn = (int)m->x.size(), nt = (int)m->elem.size();
    std::fprintf(fp, "# vtk DataFile Version 2.0\nmfree iwf\nASCII\n\nDATASET UNSTRUCTURED_GRID\n");
    std::fprintf(fp, "POINTS %d float\n", nn);
    for (int i = 0; i < nn; i++) std::fprintf(fp, "%f %f %f\n", m->x[(size_t)i], m->y[(size_t)i], 0.);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "CELLS %d %d\n", nt, 3 * nt + nt);
    for (const auto &e : m->elem) std::fprintf(fp, "3 %d %d %d\n", e.x, e.y, e.z);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "CELL_TYPES %d\n", nt);
    for (int i = 0; i < nt; i++) std::fprintf(fp, "5\n");
    std::fprintf(fp, "CELL_DATA %d\n", nt);
    std::fprintf(fp, "SCALARS color int 1\nLOOKUP_TABLE default\n");
    for (const auto &e : m->elem) std::fprintf(fp, "%u\n", (unsigned)e.w);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "SCALARS layerf int 1\nLOOKUP_TABLE default\n");
    for (const auto &l : m->layers) std::fprintf(fp, "%u\n", (unsigned)l[0]);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "SCALARS layers int 1\nLOOKUP_TABLE default\n");
    for (const auto &l : m->layers) std::fprintf(fp, "%u\n", (unsigned)l[1]);
    std::fprintf(fp, "\n");
    const bool ok = std::ferror(fp) == 0;
    return (std::fclose(fp) == 0 && ok) ? SPHFEM_OK : SPHFEM_ERR_ARG;
}

// tool_FEM::set_up_segments_from_mesh (tool_FEM.cu:771-1039) and the FEM system of
// FEM_thermal_2D_GPU's constructor (FEM_thermal_gpu.cu:584-659): fills `out` with arrays owned by
// the mesh object
int sphfem_mesh_tool_desc(sphfem_mesh *mp, const sphfem_tool_constants *tc, sphfem_double2 tl_point_ref,
                          sphfem_double2 vel, sphfem_tool_desc *out) {
    if (!mp || !tc || !out) return SPHFEM_ERR_ARG;
    sphfem_mesh &m = *mp;
    const int nn = (int)m.x.size(), ne = (int)m.elem.size();
    // cutting segments, oriented by the reference point (tool_FEM.cu:781-839)
    std::vector<Seg> d;
    std::vector<sphfem_int4> mark;
    for (int e : m.cut_elems) {
        const auto fl = m.cut_flags[(size_t)e];
        const sphfem_int4 t = m.elem[(size_t)e];
        const int pairs[3][2] = {{t.x, t.y}, {t.y, t.z}, {t.z, t.x}};
        for (int k = 0; k < 3; k++) {
            if (fl[(size_t)k] == 0) continue;
            const sphfem_double2 l = {m.x[(size_t)pairs[k][0]], m.y[(size_t)pairs[k][0]]};
            const sphfem_double2 r = {m.x[(size_t)pairs[k][1]], m.y[(size_t)pairs[k][1]]};
            const Seg s = make_seg(l, r, tl_point_ref);
            d.push_back(s);
            sphfem_int4 mk = {e, t.w, 0, 0};
            if (eq(s.left, l)) {
                mk.z = pairs[k][0];
                mk.w = pairs[k][1];
            } else if (eq(s.left, r)) {
                mk.z = pairs[k][1];
                mk.w = pairs[k][0];
            }
            mark.push_back(mk);
        }
    }
    const int ns = (int)d.size();
    if (ns == 0) return SPHFEM_ERR_ARG;
    // extent and reference distance (:842-873)
    double y_min = DBL_MAX, x_max = -DBL_MAX, y_max = -DBL_MAX, x_min = DBL_MAX, dist_max = -DBL_MAX;
    sphfem_double2 tl = tl_point_ref;
    for (const Seg &s : d) {
        y_min = std::fmin(y_min, s.left.y);
        y_min = std::fmin(y_min, s.right.y);
        x_max = std::fmax(x_max, s.left.x);
        x_max = std::fmax(x_max, s.right.x);
        x_min = std::fmin(x_min, s.left.x);
        x_min = std::fmin(x_min, s.right.x);
        y_max = std::fmax(y_max, s.left.y);
        y_max = std::fmax(y_max, s.right.y);
        dist_max = std::fmax(dist(tl, s.left), dist_max);
        dist_max = std::fmax(dist(tl, s.right), dist_max);
    }
    tl = {x_min, y_max};
    for (const Seg &s : d) dist_max = std::fmax(dist(tl, s.left), dist_max);
    // chain order from the top end (:875-948)
    bool is_left = false;
    int num_buffer = 0;
    for (int i = 0; i < ns; i++) {
        if (d[(size_t)i].left.y == y_max) {
            is_left = true;
            num_buffer = i;
            break;
        }
        if (d[(size_t)i].right.y == y_max) {
            is_left = false;
            num_buffer = i;
            break;
        }
    }
    std::vector<Seg> chain;
    std::vector<sphfem_int4> cmark;
    if (is_left) {
        sphfem_double2 b = {d[(size_t)num_buffer].left.x, y_max};
        int rep = 0;
        for (int i = 0; i < ns; i++) {
            bool done = false;
            for (int j = 0; j < ns && !done; j++) {
                if (d[(size_t)j].left.y == b.y && d[(size_t)j].left.x == b.x) {
                    b = d[(size_t)j].right;
                    rep = j;
                    chain.push_back(d[(size_t)j]);
                    cmark.push_back(mark[(size_t)j]);
                    done = true;
                }
            }
            for (int j = 0; j < ns && !done; j++) {
                if (d[(size_t)j].right.y == b.y && d[(size_t)j].right.x == b.x && j != rep) {
                    Seg &s = d[(size_t)j];
                    s.right = s.left;
                    s.left = b;
                    s.n = {-s.n.x, -s.n.y};
                    b = s.right;
                    std::swap(mark[(size_t)j].z, mark[(size_t)j].w);
                    rep = j;
                    chain.push_back(s);
                    cmark.push_back(mark[(size_t)j]);
                    done = true;
                }
            }
        }
    } else {
        sphfem_double2 b = {d[(size_t)num_buffer].right.x, y_max};
        int rep = 0;
        for (int i = 0; i < ns; i++) {
            bool done = false;
            for (int j = 0; j < ns && !done; j++) {
                if (d[(size_t)j].right.y == b.y && d[(size_t)j].right.x == b.x) {
                    b = d[(size_t)j].left;
                    chain.push_back(d[(size_t)j]);
                    cmark.push_back(mark[(size_t)j]);
                    done = true;
                }
            }
            for (int j = 0; j < ns && !done; j++) {
                if (d[(size_t)j].left.y == b.y && d[(size_t)j].left.x == b.x && j != rep) {
                    Seg &s = d[(size_t)j];
                    s.left = s.right;
                    s.right = b;
                    s.n = {-s.n.x, -s.n.y};
                    b = s.left;
                    std::swap(mark[(size_t)j].z, mark[(size_t)j].w);
                    rep = j;
                    chain.push_back(s);
                    cmark.push_back(mark[(size_t)j]);
                    done = true;
                }
            }
        }
    }
    if ((int)chain.size() != ns) {
        std::fprintf(stderr, "sphfem_mesh_tool_desc: the cutting edge is not one chain (%zu of %d segments)\n",
                     chain.size(), ns);
        return SPHFEM_ERR_ARG;
    }
    // tl and br nodes (:1021-1037)
    double tl_x = FLT_MAX, tl_y = -FLT_MAX, br_x = -FLT_MAX, br_y = FLT_MAX;
    for (int i = 0; i < nn; i++) {
        if (m.x[(size_t)i] <= tl_x && m.y[(size_t)i] >= tl_y) {
            tl_x = m.x[(size_t)i];
            tl_y = m.y[(size_t)i];
        }
        if (m.x[(size_t)i] >= br_x && m.y[(size_t)i] <= br_y) {
            br_x = m.x[(size_t)i];
            br_y = m.y[(size_t)i];
        }
    }
    // FEM system: assembly colour by colour (element order within a colour), Dirichlet rows of the
    // fixed elements colour by colour, nodal temperatures T0 or the hot-tool values
    m.d_K.assign(m.col_ind.size(), 0.);
    m.d_M.assign((size_t)nn, 0.);
    m.d_f.assign((size_t)nn, 0.);
    if (m.T_init.empty()) m.d_T.assign((size_t)nn, tc->T0);
    else m.d_T = m.T_init;
    int max_colour = 0;
    for (const auto &e : m.elem) max_colour = std::max(max_colour, e.w);
    for (int c = 1; c <= max_colour; c++) {
        for (int e = 0; e < ne; e++) {
            const sphfem_int4 t = m.elem[(size_t)e];
            if (t.w != c) continue;
            double mass[3], Ke[3][3];
            element_matrices(m, t, tc->cp, tc->rho, tc->k, mass, Ke);
            const int nd[3] = {t.x, t.y, t.z};
            for (int a = 0; a < 3; a++) {
                m.d_M[(size_t)nd[a]] += mass[a];
                for (int b = 0; b < 3; b++) m.d_K[(size_t)csr_find(m, nd[a], nd[b])] += Ke[a][b];
            }
        }
    }
    for (int c = 1; c <= max_colour; c++) {
        for (int e : m.fixed_elems) {
            const sphfem_int4 t = m.elem[(size_t)e];
            if (t.w != c) continue;
            const auto fl = m.fixed_flags[(size_t)e];
            int nf[3] = {0, 0, 0};
            if (fl[0]) nf[0] = nf[1] = 1;
            if (fl[1]) nf[1] = nf[2] = 1;
            if (fl[2]) nf[2] = nf[0] = 1;
            const int nd[3] = {t.x, t.y, t.z};
            for (int k = 0; k < 3; k++) {
                if (!nf[k]) continue;
                const int node = nd[k];
                for (int i = m.row_ptr[(size_t)node]; i < m.row_ptr[(size_t)node + 1]; i++) {
                    const int loc = m.col_ind[(size_t)i];
                    m.d_f[(size_t)loc] += -m.d_K[(size_t)i] * tc->T0;
                    m.d_K[(size_t)i] = 0.;
                    m.d_K[(size_t)csr_find(m, loc, node)] = 0.;
                    if (loc == node) m.d_K[(size_t)i] = 1.0;
                }
                m.d_M[(size_t)node] = 1.0;
                m.d_f[(size_t)node] = tc->T0;
            }
        }
    }
    m.d_pos.resize((size_t)nn);
    for (int i = 0; i < nn; i++) m.d_pos[(size_t)i] = {m.x[(size_t)i], m.y[(size_t)i]};
    m.d_layer.resize((size_t)ne);
    for (int e = 0; e < ne; e++) m.d_layer[(size_t)e] = (m.layers[(size_t)e][0] | m.layers[(size_t)e][1]);
    m.d_left.resize((size_t)ns);
    m.d_right.resize((size_t)ns);
    m.d_n.resize((size_t)ns);
    m.d_marks = cmark;
    for (int i = 0; i < ns; i++) {
        m.d_left[(size_t)i] = chain[(size_t)i].left;
        m.d_right[(size_t)i] = chain[(size_t)i].right;
        m.d_n[(size_t)i] = chain[(size_t)i].n;
    }
    std::memset(out, 0, sizeof(*out));
    out->node_num = nn;
    out->ele_num = ne;
    out->nz_num = (int)m.col_ind.size();
    out->seg_num = ns;
    out->color_total = max_colour;
    int max_cut = 0;
    for (int e : m.cut_elems) max_cut = std::max(max_cut, m.elem[(size_t)e].w);
    out->color_cutting_edge = max_cut;
    out->position = m.d_pos.data();
    out->node_num_el = m.elem.data();
    out->layer = m.d_layer.data();
    out->csr_row_ptr = m.row_ptr.data();
    out->csr_col_ind = m.col_ind.data();
    out->K = m.d_K.data();
    out->M = m.d_M.data();
    out->f = m.d_f.data();
    out->T = m.d_T.data();
    out->left = m.d_left.data();
    out->right = m.d_right.data();
    out->n = m.d_n.data();
    out->marks = m.d_marks.data();
    out->vel = vel;
    out->tl_point_ref = {tl_x, tl_y};
    out->br_point = {br_x, br_y};
    out->ref_dist = dist_max;
    out->lowest_point = y_min;
    out->right_point = x_max;
    m.d_node_flags.resize((size_t)nn);
    for (int i = 0; i < nn; i++)
        m.d_node_flags[(size_t)i] = (m.cut_node[(size_t)i] ? 1 : 0) | (m.fixed_node[(size_t)i] ? 2 : 0);
    m.d_fixed_flags.resize(m.fixed_elems.size());
    for (size_t j = 0; j < m.fixed_elems.size(); j++) {
        const auto &fl = m.fixed_flags[(size_t)m.fixed_elems[j]];
        m.d_fixed_flags[j] = {fl[0], fl[1], fl[2], 0};
    }
    out->node_flags = m.d_node_flags.data();
    out->fixed_num = (int)m.fixed_elems.size();
    out->fixed_elems = m.fixed_elems.data();
    out->fixed_flags = m.d_fixed_flags.data();
    return SPHFEM_OK;
}

}  // extern "C"
