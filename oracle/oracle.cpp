// oracle.cpp -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).
//
// A from-scratch restatement of bio-phys/trimem's hot-path arithmetic, used as
// the checker for the HIP library and as the CPU baseline of bench.py.  Every
// function names the reference file:line it follows.  The OpenMesh behaviour
// trimem depends on (half-edge construction, circulators, calc_* helpers,
// is_flip_ok, flip) is restated from OpenMesh's documented semantics because
// the OpenMesh submodule is absent from the reference checkout.
//
// Structure mirrors the reference's vertex-centric OpenMP design
// (src/kernel.h:140-262): one parallel_for over vertices for the properties,
// one declared reduction, one parallel_for for the property gradients and one
// for the combination.  Build flags: -O3 -fopenmp -ffp-contract=off.

#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- vectors --
struct V3 {
    double x, y, z;
};
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline V3 operator*(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 operator/(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
static inline V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
static inline V3& operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// OpenMesh VectorT::norm() = sqrt(sqrnorm()), sqrnorm = x*x + y*y + z*z
static inline double norm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }

// ------------------------------------------------------------------ mesh ---
// Half-edge store with OpenMesh ArrayKernel numbering: edge e owns halfedges
// 2e and 2e+1, opposite(h) = h^1, from(h) = to(opposite(h)).
struct Mesh {
    std::vector<V3> p;
    std::vector<int> v_he;    // outgoing halfedge per vertex (-1: isolated)
    std::vector<int> he_to, he_next, he_prev, he_face;
    std::vector<int> f_he;

    int nv() const { return (int)p.size(); }
    int nh() const { return (int)he_to.size(); }
    int ne() const { return nh() / 2; }
    int nf() const { return (int)f_he.size(); }
    int to(int h) const { return he_to[h]; }
    int from(int h) const { return he_to[h ^ 1]; }
    int next(int h) const { return he_next[h]; }
    int prev(int h) const { return he_prev[h]; }
    int opp(int h) const { return h ^ 1; }
    bool he_boundary(int h) const { return he_face[h] < 0; }
    bool edge_boundary(int e) const { return he_face[2 * e] < 0 || he_face[2 * e + 1] < 0; }
    // OpenMesh: a vertex is boundary if it is isolated or its halfedge has
    // no face (OpenMesh keeps boundary vertices' halfedge on the boundary).
    bool v_boundary(int v) const { int h = v_he[v]; return h < 0 || he_face[h] < 0; }
    // clockwise rotation of an outgoing halfedge: next(opposite(h))
    int cw(int h) const { return he_next[h ^ 1]; }
};

// Outgoing halfedges of v in OpenMesh voh_range order (clockwise from
// halfedge_handle(v)).  A bound guards against corrupt topology.
template <class F>
static inline void for_voh(const Mesh& m, int v, F&& f)
{
    int h0 = m.v_he[v];
    if (h0 < 0) return;
    int h = h0;
    int guard = 0;
    do {
        f(h);
        h = m.cw(h);
        if (++guard > 1 << 20) throw std::runtime_error("voh circulation does not close");
    } while (h != h0);
}

template <class F>
static inline void for_vv(const Mesh& m, int v, F&& f)
{
    for_voh(m, v, [&](int h) { f(m.to(h)); });
}

// Mesh construction following trimem's from_points_cells (mesh_py.cpp:12-69)
// and OpenMesh PolyConnectivity::add_face: faces with out-of-range ids are
// dropped first, new edges are numbered in creation order with halfedge 2e
// pointing v_i -> v_{i+1}, the face halfedge is v_{n-1} -> v_0, and faces
// that would make a complex vertex or complex edge are rejected.
static std::unique_ptr<Mesh> build_mesh(const double* pts, int n, const int* cells, int m)
{
    auto mesh = std::make_unique<Mesh>();
    Mesh& M = *mesh;
    M.p.resize(n);
    for (int i = 0; i < n; i++) M.p[i] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
    M.v_he.assign(n, -1);
    if (n < 3 || m == 0) return mesh;

    std::unordered_map<uint64_t, int> he_of;    // (from,to) -> halfedge
    auto key = [](int a, int b) { return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; };
    auto find_he = [&](int a, int b) {
        auto it = he_of.find(key(a, b));
        return it == he_of.end() ? -1 : it->second;
    };
    // per-vertex count of boundary outgoing halfedges and of outgoing halfedges;
    // a vertex with outgoing halfedges and no boundary outgoing halfedge is a
    // closed (non-boundary) vertex, to which OpenMesh refuses to add faces.
    std::vector<int> n_out(n, 0), n_bout(n, 0);

    for (int f = 0; f < m; f++) {
        int vh[3];
        int nvh = 0;
        for (int j = 0; j < 3; j++) {
            int c = cells[3 * f + j];
            if (c >= 0 && c < n) vh[nvh++] = c;
        }
        if (nvh < 3) continue;
        // add_face checks: complex vertex / complex edge
        bool ok = true;
        int hh[3];
        for (int i = 0; i < 3; i++) {
            int a = vh[i], b = vh[(i + 1) % 3];
            if (n_out[a] > 0 && n_bout[a] == 0) { ok = false; break; }
            hh[i] = find_he(a, b);
            if (hh[i] >= 0 && M.he_face[hh[i]] >= 0) { ok = false; break; }
        }
        if (!ok) continue;
        int fid = (int)M.f_he.size();
        for (int i = 0; i < 3; i++) {
            int a = vh[i], b = vh[(i + 1) % 3];
            if (hh[i] < 0) {
                int e = M.ne();
                M.he_to.push_back(b); M.he_to.push_back(a);
                M.he_next.push_back(-1); M.he_next.push_back(-1);
                M.he_prev.push_back(-1); M.he_prev.push_back(-1);
                M.he_face.push_back(-1); M.he_face.push_back(-1);
                he_of[key(a, b)] = 2 * e;
                he_of[key(b, a)] = 2 * e + 1;
                hh[i] = 2 * e;
                n_out[a]++; n_bout[a]++;
                n_out[b]++; n_bout[b]++;
                if (M.v_he[a] < 0) M.v_he[a] = 2 * e;
                if (M.v_he[b] < 0) M.v_he[b] = 2 * e + 1;
            }
        }
        for (int i = 0; i < 3; i++) {
            int h = hh[i];
            M.he_face[h] = fid;
            M.he_next[h] = hh[(i + 1) % 3];
            M.he_prev[hh[(i + 1) % 3]] = h;
            n_bout[vh[i]]--;    // h = v_i -> v_{i+1} is no longer boundary
        }
        M.f_he.push_back(hh[2]);
    }
    // Link the remaining boundary halfedges into loops: next(h) is the
    // boundary halfedge leaving to(h).
    std::vector<int> bout(n, -1);
    for (int h = 0; h < M.nh(); h++)
        if (M.he_face[h] < 0) bout[M.from(h)] = h;
    for (int h = 0; h < M.nh(); h++) {
        if (M.he_face[h] < 0) {
            int nx = bout[M.to(h)];
            M.he_next[h] = nx;
            if (nx >= 0) M.he_prev[nx] = h;
        }
    }
    // boundary vertices keep a boundary outgoing halfedge (adjust_outgoing_halfedge)
    for (int v = 0; v < n; v++)
        if (bout[v] >= 0) M.v_he[v] = bout[v];
    return mesh;
}

// OpenMesh TriConnectivity::is_flip_ok
static bool is_flip_ok(const Mesh& m, int e)
{
    if (m.edge_boundary(e)) return false;
    int hh = 2 * e, oh = 2 * e + 1;
    int ah = m.to(m.next(hh));
    int bh = m.to(m.next(oh));
    if (ah == bh) return false;
    bool adj = false;
    for_vv(m, ah, [&](int w) { if (w == bh) adj = true; });
    return !adj;
}

// OpenMesh TriConnectivity::flip
static void flip(Mesh& m, int e)
{
    int a0 = 2 * e, b0 = 2 * e + 1;
    int a1 = m.next(a0), a2 = m.next(a1);
    int b1 = m.next(b0), b2 = m.next(b1);
    int va0 = m.to(a0), va1 = m.to(a1);
    int vb0 = m.to(b0), vb1 = m.to(b1);
    int fa = m.he_face[a0], fb = m.he_face[b0];
    m.he_to[a0] = va1;
    m.he_to[b0] = vb1;
    auto link = [&](int h, int n) { m.he_next[h] = n; m.he_prev[n] = h; };
    link(a0, a2); link(a2, b1); link(b1, a0);
    link(b0, b2); link(b2, a1); link(a1, b0);
    m.he_face[a1] = fb;
    m.he_face[b1] = fa;
    m.f_he[fa] = a0;
    m.f_he[fb] = b0;
    if (m.v_he[va0] == b0) m.v_he[va0] = a1;
    if (m.v_he[vb0] == a0) m.v_he[vb0] = b1;
}

// ---------------------------------------------- OpenMesh calc_* helpers ----
static inline V3 calc_edge_vector(const Mesh& m, int h) { return m.p[m.to(h)] - m.p[m.from(h)]; }

static inline V3 calc_sector_normal(const Mesh& m, int h)
{
    V3 v0 = m.p[m.to(m.next(h))] - m.p[m.to(h)];
    V3 v1 = m.p[m.from(h)] - m.p[m.to(h)];
    return cross(v0, v1);
}

static inline double sane_aarg(double a) { return a < -1 ? -1 : (a > 1 ? 1 : a); }
static inline double om_angle(double c, double s)
{
    c = std::min(1.0, std::max(-1.0, c));
    return s >= 0 ? std::acos(c) : -std::acos(c);
}
// default epsilon of OpenMesh's is_zero for doubles (parity unpinned: only
// degenerate sectors reach it)
static inline bool is_zero(double a) { return std::fabs(a) < 100 * std::numeric_limits<double>::epsilon(); }

static V3 calc_face_normal(const Mesh& m, int f)
{
    int h = m.f_he[f];
    V3 p0 = m.p[m.to(h)], p1 = m.p[m.to(m.next(h))], p2 = m.p[m.to(m.next(m.next(h)))];
    V3 n = cross(p1 - p0, p2 - p0);
    double l = norm(n);
    return l != 0 ? n * (1.0 / l) : V3{0, 0, 0};
}

static double calc_sector_angle(const Mesh& m, int h)
{
    V3 v0 = calc_edge_vector(m, m.next(h));
    V3 v1 = calc_edge_vector(m, m.opp(h));
    double denom = norm(v0) * norm(v1);
    if (is_zero(denom)) return 0;
    double c = dot(v0, v1) / denom;
    if (m.he_boundary(h)) {
        int f = m.he_face[m.opp(h)];
        V3 fn = f >= 0 ? calc_face_normal(m, f) : V3{0, 0, 0};
        double s = dot(cross(v0, v1), fn);
        return om_angle(c, s);
    }
    return std::acos(sane_aarg(c));
}

static double calc_dihedral_angle(const Mesh& m, int h)
{
    if (m.edge_boundary(h >> 1)) return 0;
    V3 n0 = calc_sector_normal(m, h);
    V3 n1 = calc_sector_normal(m, m.opp(h));
    V3 he = calc_edge_vector(m, h);
    double denom = norm(n0) * norm(n1);
    if (denom == 0) return 0;
    double c = dot(n0, n1) / denom;
    double s = dot(cross(n0, n1), he);
    return om_angle(c, s);
}

// ------------------------------------------------- mesh_util.h:15-185 ------
static inline V3 face_normal(const Mesh& m, int h)    // mesh_util.h:23-29
{
    V3 n = calc_sector_normal(m, h);
    return n / norm(n);
}
static inline double edge_length(const Mesh& m, int h) { return norm(calc_edge_vector(m, h)); }

// edge_length_grad<N> (mesh_util.h:36-56): bit0 -> from, bit1 -> to
static inline void edge_length_grad(const Mesh& m, int h, V3* g_from, V3* g_to)
{
    V3 e = calc_edge_vector(m, h);
    e = e / norm(e);
    if (g_from) *g_from = -e;
    if (g_to) *g_to = e;
}

static inline double face_area(const Mesh& m, int h) { return norm(calc_sector_normal(m, h)) / 2; }

// face_area_grad<N> (mesh_util.h:65-92): [from, to, to(next)]
static inline void face_area_grad(const Mesh& m, int h, V3* g0, V3* g1, V3* g2)
{
    V3 n = face_normal(m, h);
    n = n / norm(n);
    if (g0) *g0 = 0.5 * cross(n, calc_edge_vector(m, m.next(h)));
    if (g1) *g1 = 0.5 * cross(n, calc_edge_vector(m, m.prev(h)));
    if (g2) *g2 = 0.5 * cross(n, calc_edge_vector(m, h));
}

static inline double face_volume(const Mesh& m, int h)    // mesh_util.h:94-101
{
    V3 p0 = m.p[m.from(h)], p1 = m.p[m.to(h)], p2 = m.p[m.to(m.next(h))];
    return dot(cross(p1, p2), p0) / 6;
}

static inline void face_volume_grad(const Mesh& m, int h, V3* g0, V3* g1, V3* g2)
{
    V3 p0 = m.p[m.from(h)], p1 = m.p[m.to(h)], p2 = m.p[m.to(m.next(h))];
    if (g0) *g0 = cross(p1, p2) / 6;
    if (g1) *g1 = cross(p2, p0) / 6;
    if (g2) *g2 = cross(p0, p1) / 6;
}

static inline double dihedral_angle(const Mesh& m, int h) { return calc_dihedral_angle(m, h); }

// dihedral_angle_grad<N> (mesh_util.h:135-185), bit order
// [from, to, opposite-in-face, opposite-in-twin-face]
static void dihedral_angle_grad(const Mesh& m, int h, int bits, V3* out)
{
    V3 n0 = face_normal(m, h);
    V3 n1 = face_normal(m, m.opp(h));
    double l = edge_length(m, h);
    double a01 = 0, a02 = 0, a03 = 0, a04 = 0;
    bool b0 = bits & 1, b1 = bits & 2, b2 = bits & 4, b3 = bits & 8;
    if (b1 || b2) a01 = calc_sector_angle(m, m.prev(h));
    if (b0 || b2) a03 = calc_sector_angle(m, h);
    if (b1 || b3) a02 = calc_sector_angle(m, m.opp(h));
    if (b0 || b3) a04 = calc_sector_angle(m, m.prev(m.opp(h)));
    int k = 0;
    if (b0) out[k++] = (1. / std::tan(a03) * n0 + 1. / std::tan(a04) * n1) / l;
    if (b1) out[k++] = (1. / std::tan(a01) * n0 + 1. / std::tan(a02) * n1) / l;
    if (b2) out[k++] = -(1. / std::tan(a01) + 1. / std::tan(a03)) * n0 / l;
    if (b3) out[k++] = -(1. / std::tan(a02) + 1. / std::tan(a04)) * n1 / l;
}

// ------------------------------------------------------------- params -----
struct Params {
    orc_params p;
};

// ---------------------------------------------------- tether potentials ----
// mesh_tether.h:29-121
struct Bonds {
    int type;    // 0 edge, 1 area
    double lc0, lc1, a0;
    int r;
    int valence() const { return type == 0 ? 2 : 3; }

    double vertex_property(const Mesh& m, int h) const
    {
        if (type == 0) {    // mesh_tether.h:40-55
            double l = edge_length(m, h);
            if (l > lc0) return std::pow(r, r + 1) * std::pow(l - lc0, r);
            if (l < lc1) return std::exp(l / (l - lc1)) / std::pow(l, r);
            return 0.0;
        }
        double a = face_area(m, h);    // mesh_tether.h:94-102
        double d = a / a0 - 1.0;
        return d * d;
    }
    // gradient w.r.t. from(h) only (mesh_tether.h:58-82, 104-117)
    V3 vertex_property_grad(const Mesh& m, int h) const
    {
        if (type == 0) {
            double l = edge_length(m, h);
            V3 lg;
            edge_length_grad(m, h, &lg, nullptr);
            double fac = 0.0;
            if (l > lc0) fac = std::pow(r, r + 2) * std::pow(l - lc0, r - 1);
            if (l < lc1) {
                double lr = std::pow(l, r);
                double lmc = l - lc1;
                fac -= lc1 / (lmc * lmc) + r / l;
                fac /= lr;
                fac *= std::exp(l / lmc);
            }
            return lg * fac;
        }
        double a = face_area(m, h);
        V3 ag;
        face_area_grad(m, h, &ag, nullptr, nullptr);
        double d = 2.0 * (a / a0 - 1.0) / a0;
        return ag * d;
    }
};

// ------------------------------------------------------ neighbour lists ----
static inline double squ_distance(const V3& a, const V3& b)    // cell_list.h:11-20
{
    double dist = 0.0;
    double d0 = a.x - b.x, d1 = a.y - b.y, d2 = a.z - b.z;
    dist += d0 * d0;
    dist += d1 * d1;
    dist += d2 * d2;
    return dist;
}
static inline double distance(const V3& a, const V3& b) { return std::sqrt(squ_distance(a, b)); }

// exclusion test of cell_list.h:170-202 (vv_range of i, and of its neighbours)
static bool excluded(const Mesh& m, int i, int j, int excl)
{
    if (excl >= 0 && i == j) return true;
    if (excl > 0) {
        bool in_ring = false;
        int h0 = m.v_he[i];
        if (h0 < 0) return false;
        int h = h0;
        do {
            int vi = m.to(h);
            if (vi == j) { in_ring = true; break; }
            if (excl > 1) {
                for_vv(m, vi, [&](int w) { if (w == j) in_ring = true; });
                if (in_ring) break;
            }
            h = m.cw(h);
        } while (h != h0);
        return in_ring;
    }
    return false;
}

struct NList {
    virtual ~NList() = default;
    virtual void distance_matrix(const Mesh& m, double dmax, std::vector<double>& d,
                                 std::vector<int>& i, std::vector<int>& j) const = 0;
    virtual void point_distances(const Mesh& m, int pid, double dmax, std::vector<V3>& dist,
                                 std::vector<int>& jdx) const = 0;
};

// CellList<exclusion> (cell_list.h:28-313)
struct CellList : NList {
    int excl;
    std::map<int, std::vector<int>> cells;
    std::array<int, 3> shape, strides;
    std::array<double, 3> r_list;
    std::array<double, 3> box_min, box_max;
    std::vector<std::pair<int, int>> cell_pairs;

    CellList(const Mesh& m, double rlist, int excl_, double box_eps = 1.0e-6) : excl(excl_)
    {
        // box init reproduces cell_list.h:46-47: only component 0 of each
        // bound is initialised to (max, -1); components 1 and 2 start at 0.
        box_min = {std::numeric_limits<double>::max(), 0.0, 0.0};
        box_max = {-1.0, 0.0, 0.0};
        for (int i = 0; i < m.nv(); i++) {
            const double* p = &m.p[i].x;
            for (int k = 0; k < 3; k++) {
                if (p[k] < box_min[k]) box_min[k] = p[k];
                if (p[k] > box_max[k]) box_max[k] = p[k];
            }
        }
        for (int k = 0; k < 3; k++) {    // cell_list.h:59-66
            double dim = box_max[k] - box_min[k] + box_eps;
            int nk = int(dim / rlist);
            r_list[k] = (nk > 0) ? dim / nk : dim;
            shape[k] = (nk > 0) ? nk : 1;
        }
        strides = {1, shape[0], shape[0] * shape[1]};
        for (int i = 0; i < m.nv(); i++) {    // cell_list.h:70-79
            const double* p = &m.p[i].x;
            int id = 0;
            for (int k = 0; k < 3; k++) id += int((p[k] - box_min[k]) / r_list[k]) * strides[k];
            cells[id].push_back(i);
        }
        compute_cell_pairs();
    }

    void compute_cell_pairs()    // cell_list.h:89-129
    {
        for (auto& it : cells) {
            int icell = it.first;
            int kicell = icell / strides[2];
            int rem = icell % strides[2];
            int jicell = rem / strides[1];
            int iicell = rem % strides[1];
            for (int i = -1; i < 2; i++) {
                int ii = iicell + i;
                if (ii < 0 || ii >= shape[0]) continue;
                for (int j = -1; j < 2; j++) {
                    int jj = jicell + j;
                    if (jj < 0 || jj >= shape[1]) continue;
                    for (int k = -1; k < 2; k++) {
                        int kk = kicell + k;
                        if (kk < 0 || kk >= shape[2]) continue;
                        int ocell = ii * strides[0] + jj * strides[1] + kk * strides[2];
                        if (cells.find(ocell) == cells.end()) continue;
                        if (icell <= ocell) cell_pairs.push_back({icell, ocell});
                    }
                }
            }
        }
    }

    void point_distances(const Mesh& m, int pid, double dmax, std::vector<V3>& out,
                         std::vector<int>& nb) const override    // cell_list.h:133-219
    {
        V3 point = m.p[pid];
        const double* pp = &point.x;
        int coords[3];
        for (int k = 0; k < 3; k++) coords[k] = int((pp[k] - box_min[k]) / r_list[k]);
        for (int i = -1; i < 2; i++) {
            int ii = coords[0] + i;
            if (ii < 0 || ii >= shape[0]) continue;
            for (int j = -1; j < 2; j++) {
                int jj = coords[1] + j;
                if (jj < 0 || jj >= shape[1]) continue;
                for (int k = -1; k < 2; k++) {
                    int kk = coords[2] + k;
                    if (kk < 0 || kk >= shape[2]) continue;
                    int ocell = ii * strides[0] + jj * strides[1] + kk * strides[2];
                    auto oit = cells.find(ocell);
                    if (oit == cells.end()) continue;
                    for (int jid : oit->second) {
                        if (excluded(m, pid, jid, excl)) continue;
                        V3 dist = point - m.p[jid];
                        if (norm(dist) < dmax) {
                            out.push_back(dist);
                            nb.push_back(jid);
                        }
                    }
                }
            }
        }
    }

    void distance_matrix(const Mesh& m, double rlist, std::vector<double>& dists,
                         std::vector<int>& idx, std::vector<int>& jdx) const override    // cell_list.h:221-312
    {
        for (auto& cp : cell_pairs) {
            const std::vector<int>& ic = cells.at(cp.first);
            const std::vector<int>& oc = cells.at(cp.second);
            bool same = cp.first == cp.second;
            for (int a : ic) {
                for (int b : oc) {
                    if (excluded(m, a, b, excl)) continue;
                    if (same && a > b) continue;
                    double dist = distance(m.p[a], m.p[b]);
                    if (dist <= rlist) {
                        dists.push_back(dist);
                        idx.push_back(std::min(a, b));
                        jdx.push_back(std::max(a, b));
                    }
                }
            }
        }
    }
};

// VerletList<exclusion> (verlet_list.h:14-103)
struct VerletList : NList {
    std::vector<std::vector<int>> neighbours;

    VerletList(const Mesh& m, double rlist, int excl, double box_eps = 1.e-6)
    {
        CellList cl(m, rlist, excl, box_eps);
        std::vector<double> d;
        std::vector<int> i, j;
        cl.distance_matrix(m, rlist, d, i, j);
        neighbours.resize(m.nv());
        for (size_t k = 0; k < i.size(); k++) {
            neighbours[i[k]].push_back(j[k]);
            neighbours[j[k]].push_back(i[k]);
        }
    }

    void distance_matrix(const Mesh& m, double dmax, std::vector<double>& dists,
                         std::vector<int>& idx, std::vector<int>& jdx) const override
    {
        for (int i = 0; i < m.nv(); i++) {
            for (int j : neighbours.at(i)) {
                double dist = distance(m.p[i], m.p[j]);
                if (dist < dmax) {
                    dists.push_back(dist);
                    idx.push_back(i);
                    jdx.push_back(j);
                }
            }
        }
    }

    void point_distances(const Mesh& m, int pid, double dmax, std::vector<V3>& dist,
                         std::vector<int>& jdx) const override
    {
        const auto& nb = neighbours.at(pid);
        V3 point = m.p[pid];
        for (int jid : nb) {
            V3 di = point - m.p[jid];
            if (norm(di) < dmax) {
                dist.push_back(di);
                jdx.push_back(jid);
            }
        }
    }
};

static std::unique_ptr<NList> make_nlist(const Mesh& m, const orc_params& p)    // nlist.cpp:10-53
{
    if (p.rep_exclusion < 0 || p.rep_exclusion > 2) throw std::runtime_error("Unsupported exclusion level");
    if (p.rep_nlist == 0) return std::make_unique<CellList>(m, p.rep_rlist, p.rep_exclusion);
    if (p.rep_nlist == 1) return std::make_unique<VerletList>(m, p.rep_rlist, p.rep_exclusion);
    throw std::runtime_error("Unknown neighbour search algorithm.");
}

// ------------------------------------------------- surface repulsion -------
// SurfaceRepulsionFlatBottom (mesh_repulsion.h:28-79)
struct Repulsion {
    int r;
    double lc1;
    const NList* nl;

    double vertex_property(const Mesh& m, int pid) const
    {
        std::vector<V3> dij;
        std::vector<int> jdx;
        nl->point_distances(m, pid, lc1, dij, jdx);
        double pot = 0;
        for (auto& d : dij) {
            double l = norm(d);
            pot += std::exp(l / (l - lc1)) / std::pow(l, r);
        }
        return pot;
    }

    void vertex_property_grad(const Mesh& m, int pid, std::vector<V3>& dij) const
    {
        std::vector<int> jdx;
        nl->point_distances(m, pid, lc1, dij, jdx);
        for (auto& d : dij) {
            double fac = 0.0;
            double l = norm(d);
            double lr = std::pow(l, r + 1);
            double lmc = l - lc1;
            fac -= lc1 / (lmc * lmc) + r / l;
            fac /= lr;
            fac *= std::exp(l / lmc);
            d = d * fac;
        }
    }
};

// -------------------------------------------------- vertex properties ------
struct VP {
    double area = 0, volume = 0, curvature = 0, bending = 0, tethering = 0, repulsion = 0;
    VP& operator+=(const VP& o)
    {
        area += o.area; volume += o.volume; curvature += o.curvature;
        bending += o.bending; tethering += o.tethering; repulsion += o.repulsion;
        return *this;
    }
    VP& operator-=(const VP& o)
    {
        area -= o.area; volume -= o.volume; curvature -= o.curvature;
        bending -= o.bending; tethering -= o.tethering; repulsion -= o.repulsion;
        return *this;
    }
};
struct VPG {
    V3 area{0, 0, 0}, volume{0, 0, 0}, curvature{0, 0, 0}, bending{0, 0, 0}, tethering{0, 0, 0},
        repulsion{0, 0, 0};
};

// vertex_properties (mesh_properties.cpp:12-49)
static VP vertex_properties(const Mesh& m, const Bonds* bonds, const Repulsion& rep, int ve)
{
    VP p;
    for_voh(m, ve, [&](int he) {
        if (!m.he_boundary(he)) {
            double el = edge_length(m, he);
            double da = dihedral_angle(m, he);
            p.curvature += 0.5 * el * da;
            p.area += face_area(m, he);
            p.volume += face_volume(m, he);
            if (bonds) p.tethering += bonds->vertex_property(m, he);
        }
    });
    p.area /= 3;
    p.volume /= 3;
    p.curvature /= 2;
    p.tethering /= bonds ? bonds->valence() : 1;
    p.bending = 2 * p.curvature * p.curvature / p.area;
    p.repulsion = rep.vertex_property(m, ve);
    return p;
}

// vertex_properties_grad (mesh_properties.cpp:51-148)
static void vertex_properties_grad(const Mesh& m, const Bonds* bonds, const Repulsion& rep, int ve,
                                   const std::vector<VP>& props, VPG& d)
{
    double c_to_a[3];
    int idx = ve;
    c_to_a[0] = props[idx].curvature / props[idx].area;
    for_voh(m, ve, [&](int he) {
        if (m.he_boundary(he)) return;
        int n_he = m.next(he);
        int jdx = m.to(he);
        int kdx = m.to(n_he);
        c_to_a[1] = props[jdx].curvature / props[jdx].area;
        c_to_a[2] = props[kdx].curvature / props[kdx].area;

        double el = edge_length(m, he);
        double ea = dihedral_angle(m, he);
        V3 d_length;
        edge_length_grad(m, he, &d_length, nullptr);
        V3 d_angle;
        dihedral_angle_grad(m, he, 1, &d_angle);
        for (int i = 0; i < 2; i++) {
            V3 val = 0.25 * (ea * d_length + el * d_angle);
            d.curvature += val;
            d.bending += 4.0 * c_to_a[i] * val;
        }
        V3 d_face_area;
        face_area_grad(m, he, &d_face_area, nullptr, nullptr);
        for (int i = 0; i < 3; i++) {
            V3 val = d_face_area / 3;
            d.area += val;
            d.bending -= 2.0 * c_to_a[i] * c_to_a[i] * val;
        }
        V3 d_face_volume;
        face_volume_grad(m, he, &d_face_volume, nullptr, nullptr);
        for (int i = 0; i < 3; i++) {
            V3 val = d_face_volume / 3;
            d.volume += val;
        }
        el = edge_length(m, n_he);
        ea = dihedral_angle(m, n_he);
        (void)ea;
        dihedral_angle_grad(m, n_he, 4, &d_angle);
        for (int i = 1; i < 3; i++) {
            V3 val = 0.25 * el * d_angle;
            d.curvature += val;
            d.bending += 4.0 * c_to_a[i] * val;
        }
        if (bonds) d.tethering += bonds->vertex_property_grad(m, he);
    });
    std::vector<V3> d_repulse;
    rep.vertex_property_grad(m, idx, d_repulse);
    for (auto& v : d_repulse) d.repulsion += 2 * v;
}

// ------------------------------------------------------ energy manager -----
struct Energy {
    orc_params params;
    VP initial_props;
    std::unique_ptr<Bonds> bonds;
    std::unique_ptr<NList> nlist;
    Repulsion repulse;

    Energy(const Mesh& m, const orc_params& p) : params(p)    // energy.cpp:13-28
    {
        if (p.bond_type == 0 || p.bond_type == 1) {
            bonds = std::make_unique<Bonds>();
            bonds->type = p.bond_type;
            bonds->lc0 = p.bond_lc0;
            bonds->lc1 = p.bond_lc1;
            bonds->a0 = p.bond_a0;
            bonds->r = p.bond_r;
        } else {
            throw std::runtime_error("Unknown bond potential");
        }
        nlist = make_nlist(m, params);
        repulse = Repulsion{params.rep_r, params.rep_lc1, nlist.get()};
        initial_props = properties(m);
    }

    VP ref_props() const    // energy.cpp:30-45
    {
        double i_af = 1.0 - params.area_frac;
        double i_vf = 1.0 - params.volume_frac;
        double i_cf = 1.0 - params.curvature_frac;
        double lam = params.cont_lambda;
        VP r;
        r.area = (1.0 - lam * i_af) * initial_props.area;
        r.volume = (1.0 - lam * i_vf) * initial_props.volume;
        r.curvature = (1.0 - lam * i_cf) * initial_props.curvature;
        return r;
    }

    void update_reference_properties()    // energy.cpp:47-55
    {
        if (params.cont_lambda < 1.0) params.cont_lambda += params.cont_delta;
    }

    void update_repulsion(const Mesh& m)    // energy.cpp:57-61
    {
        nlist = make_nlist(m, params);
        repulse = Repulsion{params.rep_r, params.rep_lc1, nlist.get()};
    }

    void vertex_props(const Mesh& m, std::vector<VP>& vp) const
    {
        const int n = m.nv();
        vp.assign(n, VP());
#pragma omp parallel for
        for (int i = 0; i < n; i++) vp[i] = vertex_properties(m, bonds.get(), repulse, i);
    }

    static VP reduce(const std::vector<VP>& vp)    // kernel.h:244-254
    {
        VP s;
        const int n = (int)vp.size();
#pragma omp declare reduction(tred:VP : omp_out += omp_in) initializer(omp_priv = VP())
#pragma omp parallel for reduction(tred : s)
        for (int i = 0; i < n; i++) s += vp[i];
        return s;
    }

    VP properties(const Mesh& m) const    // energy.cpp:63-77
    {
        std::vector<VP> vp;
        vertex_props(m, vp);
        return reduce(vp);
    }

    double energy(const VP& props) const    // kernel.h:110-122
    {
        VP ref = ref_props();
        const orc_params& P = params;
        double e = 0.0;
        double d;
        d = props.area / ref.area - 1.0;
        e += P.kappa_a * d * d;
        d = props.volume / ref.volume - 1.0;
        e += P.kappa_v * d * d;
        d = props.curvature / ref.curvature - 1.0;
        e += P.kappa_c * d * d;
        e += P.kappa_t * props.tethering;
        e += P.kappa_r * props.repulsion;
        e += P.kappa_b * props.bending;
        return e;
    }

    void gradient(const Mesh& m, std::vector<V3>& grad) const    // energy.cpp:93-124
    {
        const int n = m.nv();
        std::vector<VP> vp;
        vertex_props(m, vp);
        VP props = reduce(vp);
        VP ref = ref_props();
        std::vector<VPG> gp(n);
#pragma omp parallel for
        for (int i = 0; i < n; i++) vertex_properties_grad(m, bonds.get(), repulse, i, vp, gp[i]);
        grad.assign(n, V3{0, 0, 0});
        const orc_params& P = params;
#pragma omp parallel for
        for (int i = 0; i < n; i++) {    // kernel.h:124-137
            V3 g{0, 0, 0};
            double d, fac;
            d = props.area / ref.area - 1.0;
            fac = 2.0 * P.kappa_a / ref.area * d;
            g += fac * gp[i].area;
            d = props.volume / ref.volume - 1.0;
            fac = 2.0 * P.kappa_v / ref.volume * d;
            g += fac * gp[i].volume;
            d = props.curvature / ref.curvature - 1.0;
            fac = 2.0 * P.kappa_c / ref.curvature * d;
            g += fac * gp[i].curvature;
            g += P.kappa_t * gp[i].tethering;
            g += P.kappa_r * gp[i].repulsion;
            g += P.kappa_b * gp[i].bending;
            grad[i] += g;
        }
    }
};

// edge_vertex_properties (flip_utils.cpp:19-41)
static VP edge_vertex_properties(const Mesh& m, int e, const Energy& em)
{
    VP props;
    for (int i = 0; i < 2; i++) {
        int heh = 2 * e + i;
        int ve = m.to(heh);
        props += vertex_properties(m, em.bonds.get(), em.repulse, ve);
        int nh = m.next(heh);
        ve = m.to(nh);
        props += vertex_properties(m, em.bonds.get(), em.repulse, ve);
    }
    return props;
}

// flip_serial (flips.cpp:22-80) with the proposal and acceptance streams
// injected.  A uniform is consumed only for proposals that pass
// is_flip_ok and the boundary test (flips.cpp:45, 64).
static int flip_serial_injected(Mesh& m, Energy& em, int nflips, const int* proposals,
                                const double* uniforms, int n_uniforms, int* n_used)
{
    VP props = em.properties(m);
    double e0 = em.energy(props);
    int acc = 0;
    int k = 0;
    for (int i = 0; i < nflips; i++) {
        int idx = proposals[i];
        if (idx < 0 || idx >= m.ne()) throw std::runtime_error("proposal out of range");
        if (is_flip_ok(m, idx) && !m.edge_boundary(idx)) {
            VP oprops = edge_vertex_properties(m, idx, em);
            props -= oprops;
            flip(m, idx);
            VP nprops = edge_vertex_properties(m, idx, em);
            props += nprops;
            double en = em.energy(props);
            double de = en - e0;
            double alpha = de < 0.0 ? 1.0 : std::exp(-de);
            if (k >= n_uniforms) throw std::runtime_error("uniform stream exhausted");
            double u = uniforms[k++];
            if (u <= alpha) {
                e0 = en;
                acc += 1;
            } else {
                flip(m, idx);
                props -= nprops;
                props += oprops;
            }
        }
    }
    if (n_used) *n_used = k;
    return acc;
}

// util.cpp:9-41
static double mesh_area(const Mesh& m)
{
    double a = 0.0;
    const int nf = m.nf();
#pragma omp parallel for reduction(+ : a)
    for (int i = 0; i < nf; i++) a += face_area(m, m.f_he[i]);
    return a;
}
static double mesh_edges_length(const Mesh& m)
{
    double l = 0.0;
    const int ne = m.ne();
#pragma omp parallel for reduction(+ : l)
    for (int i = 0; i < ne; i++) l += edge_length(m, 2 * i);
    return l;
}

}    // namespace orc

// ================================================================ C API =====
using namespace orc;

static thread_local std::string g_err;
static void set_err(const char* s) { g_err = s; }

#define GUARD_BEGIN try {
#define GUARD_END(ret)                 \
    }                                  \
    catch (const std::exception& ex) { \
        set_err(ex.what());            \
        return ret;                    \
    }

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }
int orc_num_threads(void) { return omp_get_max_threads(); }

void* orc_mesh_new(const double* points, int n_points, const int* cells, int n_cells)
{
    GUARD_BEGIN
    return build_mesh(points, n_points, cells, n_cells).release();
    GUARD_END(nullptr)
}
void orc_mesh_free(void* m) { delete (Mesh*)m; }
int orc_mesh_nv(void* m) { return ((Mesh*)m)->nv(); }
int orc_mesh_ne(void* m) { return ((Mesh*)m)->ne(); }
int orc_mesh_nf(void* m) { return ((Mesh*)m)->nf(); }
void orc_mesh_get_points(void* mm, double* out)
{
    Mesh& m = *(Mesh*)mm;
    std::memcpy(out, m.p.data(), sizeof(V3) * m.p.size());
}
void orc_mesh_set_points(void* mm, const double* in)
{
    Mesh& m = *(Mesh*)mm;
    std::memcpy(m.p.data(), in, sizeof(V3) * m.p.size());
}
// fv_indices (mesh_py.cpp:73-93): face vertices starting at to(halfedge(f))
void orc_mesh_fv_indices(void* mm, int* out)
{
    Mesh& m = *(Mesh*)mm;
    for (int f = 0; f < m.nf(); f++) {
        int h = m.f_he[f];
        out[3 * f + 0] = m.to(h);
        h = m.next(h);
        out[3 * f + 1] = m.to(h);
        h = m.next(h);
        out[3 * f + 2] = m.to(h);
    }
}
void orc_mesh_topology(void* mm, int* he_to, int* he_next, int* he_face, int* v_he, int* f_he)
{
    Mesh& m = *(Mesh*)mm;
    if (he_to) std::memcpy(he_to, m.he_to.data(), sizeof(int) * m.nh());
    if (he_next) std::memcpy(he_next, m.he_next.data(), sizeof(int) * m.nh());
    if (he_face) std::memcpy(he_face, m.he_face.data(), sizeof(int) * m.nh());
    if (v_he) std::memcpy(v_he, m.v_he.data(), sizeof(int) * m.nv());
    if (f_he) std::memcpy(f_he, m.f_he.data(), sizeof(int) * m.nf());
}
int orc_mesh_is_flip_ok(void* m, int e) { return is_flip_ok(*(Mesh*)m, e) ? 1 : 0; }
void orc_mesh_flip(void* m, int e) { flip(*(Mesh*)m, e); }
int orc_mesh_next_halfedge(void* m, int he) { return ((Mesh*)m)->next(he); }

static void put(double* out, int k, V3 v)
{
    out[3 * k] = v.x;
    out[3 * k + 1] = v.y;
    out[3 * k + 2] = v.z;
}
double orc_edge_length(void* m, int he) { return edge_length(*(Mesh*)m, he); }
void orc_edge_length_grad(void* m, int he, double* out)
{
    V3 a, b;
    edge_length_grad(*(Mesh*)m, he, &a, &b);
    put(out, 0, a);
    put(out, 1, b);
}
double orc_face_area(void* m, int he) { return face_area(*(Mesh*)m, he); }
void orc_face_area_grad(void* m, int he, double* out)
{
    V3 a, b, c;
    face_area_grad(*(Mesh*)m, he, &a, &b, &c);
    put(out, 0, a);
    put(out, 1, b);
    put(out, 2, c);
}
double orc_face_volume(void* m, int he) { return face_volume(*(Mesh*)m, he); }
void orc_face_volume_grad(void* m, int he, double* out)
{
    V3 a, b, c;
    face_volume_grad(*(Mesh*)m, he, &a, &b, &c);
    put(out, 0, a);
    put(out, 1, b);
    put(out, 2, c);
}
double orc_dihedral_angle(void* m, int he) { return dihedral_angle(*(Mesh*)m, he); }
void orc_dihedral_angle_grad(void* m, int he, double* out)
{
    V3 g[4];
    dihedral_angle_grad(*(Mesh*)m, he, 15, g);
    for (int k = 0; k < 4; k++) put(out, k, g[k]);
}

void* orc_energy_new(void* mesh, const orc_params* p)
{
    GUARD_BEGIN
    return new Energy(*(Mesh*)mesh, *p);
    GUARD_END(nullptr)
}
void orc_energy_free(void* em) { delete (Energy*)em; }

static void put_vp(double* out, const VP& v)
{
    out[0] = v.area; out[1] = v.volume; out[2] = v.curvature;
    out[3] = v.bending; out[4] = v.tethering; out[5] = v.repulsion;
}

int orc_properties(void* em, void* mesh, double* out6)
{
    GUARD_BEGIN
    put_vp(out6, ((Energy*)em)->properties(*(Mesh*)mesh));
    return 0;
    GUARD_END(-1)
}
int orc_vertex_properties(void* em, void* mesh, double* out)
{
    GUARD_BEGIN
    std::vector<VP> vp;
    ((Energy*)em)->vertex_props(*(Mesh*)mesh, vp);
    for (size_t i = 0; i < vp.size(); i++) put_vp(out + 6 * i, vp[i]);
    return 0;
    GUARD_END(-1)
}
double orc_energy(void* em, void* mesh)
{
    GUARD_BEGIN
    Energy& E = *(Energy*)em;
    return E.energy(E.properties(*(Mesh*)mesh));
    GUARD_END(std::nan(""))
}
double orc_energy_props(void* em, const double* p)
{
    VP v;
    v.area = p[0]; v.volume = p[1]; v.curvature = p[2];
    v.bending = p[3]; v.tethering = p[4]; v.repulsion = p[5];
    return ((Energy*)em)->energy(v);
}
int orc_gradient(void* em, void* mesh, double* out)
{
    GUARD_BEGIN
    std::vector<V3> g;
    ((Energy*)em)->gradient(*(Mesh*)mesh, g);
    std::memcpy(out, g.data(), sizeof(V3) * g.size());
    return 0;
    GUARD_END(-1)
}
void orc_update_reference_properties(void* em) { ((Energy*)em)->update_reference_properties(); }
int orc_update_repulsion(void* em, void* mesh)
{
    GUARD_BEGIN
    ((Energy*)em)->update_repulsion(*(Mesh*)mesh);
    return 0;
    GUARD_END(-1)
}
void orc_initial_props(void* em, double* out6) { put_vp(out6, ((Energy*)em)->initial_props); }
void orc_get_params(void* em, orc_params* out) { *out = ((Energy*)em)->params; }

// finite-difference gradient (bindings.cpp:25-50)
int orc_fd_gradient(void* emp, void* mesh, double* out, double eps)
{
    GUARD_BEGIN
    Energy& E = *(Energy*)emp;
    Mesh& m = *(Mesh*)mesh;
    double e0 = E.energy(E.properties(m));
    for (int i = 0; i < m.nv(); i++) {
        double* pt = &m.p[i].x;
        for (int j = 0; j < 3; j++) {
            pt[j] += eps;
            out[3 * i + j] = (E.energy(E.properties(m)) - e0) / eps;
            pt[j] -= eps;
        }
    }
    return 0;
    GUARD_END(-1)
}

void* orc_nlist_new(void* mesh, const orc_params* p)
{
    GUARD_BEGIN
    return make_nlist(*(Mesh*)mesh, *p).release();
    GUARD_END(nullptr)
}
void orc_nlist_free(void* nl) { delete (NList*)nl; }
void* orc_energy_nlist(void* em) { return ((Energy*)em)->nlist.get(); }

long orc_nlist_distance_matrix(void* nl, void* mesh, double dmax, double* d, int* i, int* j, long cap)
{
    GUARD_BEGIN
    std::vector<double> dd;
    std::vector<int> ii, jj;
    ((NList*)nl)->distance_matrix(*(Mesh*)mesh, dmax, dd, ii, jj);
    long n = (long)dd.size();
    for (long k = 0; k < n && k < cap; k++) {
        d[k] = dd[k];
        i[k] = ii[k];
        j[k] = jj[k];
    }
    return n;
    GUARD_END(-1)
}

long orc_nlist_point_distances(void* nl, void* mesh, int pid, double dmax, double* dist3, int* jdx, long cap)
{
    GUARD_BEGIN
    std::vector<V3> dd;
    std::vector<int> jj;
    ((NList*)nl)->point_distances(*(Mesh*)mesh, pid, dmax, dd, jj);
    long n = (long)dd.size();
    for (long k = 0; k < n && k < cap; k++) {
        put(dist3, (int)k, dd[k]);
        jdx[k] = jj[k];
    }
    return n;
    GUARD_END(-1)
}

int orc_flip_serial_injected(void* em, void* mesh, int nflips, const int* proposals,
                             const double* uniforms, int n_uniforms, int* n_used)
{
    GUARD_BEGIN
    return flip_serial_injected(*(Mesh*)mesh, *(Energy*)em, nflips, proposals, uniforms, n_uniforms,
                                n_used);
    GUARD_END(-1)
}

int orc_edge_vertex_properties(void* em, void* mesh, int edge, double* out6)
{
    GUARD_BEGIN
    put_vp(out6, edge_vertex_properties(*(Mesh*)mesh, edge, *(Energy*)em));
    return 0;
    GUARD_END(-1)
}

double orc_area(void* m) { return mesh_area(*(Mesh*)m); }
double orc_edges_length(void* m) { return mesh_edges_length(*(Mesh*)m); }

}    // extern "C"
