// tm_om.h -- half-edge (OpenMesh-semantics) device functions.
//
// These walk the halfedge arrays exactly as trimem's C++ walks OpenMesh, and
// are used where the fast ring-table passes do not apply: boundary vertices,
// the serial and coloured flip kernels (which rewrite the topology), and the
// single-halfedge geometry entry point.  FP64 throughout.
//
// Reference semantics restated here:
//   OpenMesh calc_edge_vector / calc_sector_normal / calc_sector_angle /
//   calc_dihedral_angle / calc_face_normal, voh circulation (clockwise,
//   cw(h) = next(opposite(h)));
//   trimem mesh_util.h:15-185 (face_normal, edge_length(_grad),
//   face_area(_grad), face_volume(_grad), dihedral_angle(_grad));
//   trimem mesh_properties.cpp:12-148 (vertex_properties, vertex_properties_grad);
//   trimem mesh_tether.h:29-121 (bond potentials);
//   trimem mesh_repulsion.h:28-79 (flat-bottom repulsion).
#pragma once

#include "tm_common.h"

namespace tmh {

// Mutable or immutable topology view used by the walkers.
struct Topo {
    const double* x;
    const int* he_to;
    const int* he_next;
    const int* he_prev;
    const int* he_face;
    const int* v_he;
    const int* f_he;

    __device__ __forceinline__ d3 P(int v) const { return d3{x[3 * v], x[3 * v + 1], x[3 * v + 2]}; }
    __device__ __forceinline__ int to(int h) const { return he_to[h]; }
    __device__ __forceinline__ int from(int h) const { return he_to[h ^ 1]; }
    __device__ __forceinline__ int next(int h) const { return he_next[h]; }
    __device__ __forceinline__ int prev(int h) const { return he_prev[h]; }
    __device__ __forceinline__ bool he_boundary(int h) const { return he_face[h] < 0; }
    __device__ __forceinline__ bool edge_boundary(int e) const
    {
        return he_face[2 * e] < 0 || he_face[2 * e + 1] < 0;
    }
    __device__ __forceinline__ int cw(int h) const { return he_next[h ^ 1]; }
    __device__ __forceinline__ int face(int h) const { return he_face[h]; }
    __device__ __forceinline__ int vhe(int v) const { return v_he[v]; }
    __device__ __forceinline__ int fhe(int f) const { return f_he[f]; }
};

__device__ __forceinline__ Topo topo_of(const MeshView& m)
{
    return Topo{m.x, m.he_to, m.he_next, m.he_prev, m.he_face, m.v_he, m.f_he};
}

// ------------------------------------------------ OpenMesh calc_* helpers --
template <class TP>
__device__ __forceinline__ d3 om_edge_vector(const TP& t, int h) { return t.P(t.to(h)) - t.P(t.from(h)); }

template <class TP>
__device__ __forceinline__ d3 om_sector_normal(const TP& t, int h)
{
    d3 pt = t.P(t.to(h));
    return cross(t.P(t.to(t.next(h))) - pt, t.P(t.from(h)) - pt);
}

__device__ __forceinline__ double om_angle(double c, double s)
{
    c = fmin(1.0, fmax(-1.0, c));
    return s >= 0 ? acos(c) : -acos(c);
}

template <class TP>
__device__ __forceinline__ d3 om_face_normal(const TP& t, int f)
{
    int h = t.fhe(f);
    d3 p0 = t.P(t.to(h));
    d3 p1 = t.P(t.to(t.next(h)));
    d3 p2 = t.P(t.to(t.next(t.next(h))));
    d3 n = cross(p1 - p0, p2 - p0);
    double l = norm(n);
    return l != 0 ? n * (1.0 / l) : d3{0, 0, 0};
}

// OpenMesh is_zero default tolerance for doubles
__device__ __forceinline__ bool om_is_zero(double a) { return fabs(a) < 100 * 2.220446049250313e-16; }

template <class TP>
__device__ __forceinline__ double om_sector_angle(const TP& t, int h)
{
    d3 v0 = om_edge_vector(t, t.next(h));
    d3 v1 = om_edge_vector(t, h ^ 1);
    double denom = norm(v0) * norm(v1);
    if (om_is_zero(denom)) return 0;
    double c = dot(v0, v1) / denom;
    if (t.he_boundary(h)) {
        int f = t.face(h ^ 1);
        d3 fn = f >= 0 ? om_face_normal(t, f) : d3{0, 0, 0};
        return om_angle(c, dot(cross(v0, v1), fn));
    }
    return acos(fmin(1.0, fmax(-1.0, c)));
}

template <class TP>
__device__ __forceinline__ double om_dihedral_angle(const TP& t, int h)
{
    if (t.edge_boundary(h >> 1)) return 0;
    d3 n0 = om_sector_normal(t, h);
    d3 n1 = om_sector_normal(t, h ^ 1);
    d3 he = om_edge_vector(t, h);
    double denom = norm(n0) * norm(n1);
    if (denom == 0) return 0;
    return om_angle(dot(n0, n1) / denom, dot(cross(n0, n1), he));
}

// --------------------------------------------------- mesh_util.h:15-185 ----
template <class TP>
__device__ __forceinline__ d3 mu_face_normal(const TP& t, int h)
{
    d3 n = om_sector_normal(t, h);
    return n / norm(n);
}
template <class TP>
__device__ __forceinline__ double mu_edge_length(const TP& t, int h) { return norm(om_edge_vector(t, h)); }
template <class TP>
__device__ __forceinline__ d3 mu_edge_length_grad_from(const TP& t, int h)
{
    d3 e = om_edge_vector(t, h);
    e = e / norm(e);
    return -e;
}
template <class TP>
__device__ __forceinline__ double mu_face_area(const TP& t, int h) { return norm(om_sector_normal(t, h)) / 2; }
// face_area_grad<N>: [from, to, to(next)] (mesh_util.h:65-92)
template <class TP>
__device__ __forceinline__ d3 mu_face_area_grad(const TP& t, int h, int which)
{
    d3 n = mu_face_normal(t, h);
    n = n / norm(n);
    int g = which == 0 ? t.next(h) : (which == 1 ? t.prev(h) : h);
    return 0.5 * cross(n, om_edge_vector(t, g));
}
template <class TP>
__device__ __forceinline__ double mu_face_volume(const TP& t, int h)
{
    d3 p0 = t.P(t.from(h)), p1 = t.P(t.to(h)), p2 = t.P(t.to(t.next(h)));
    return dot(cross(p1, p2), p0) / 6;
}
template <class TP>
__device__ __forceinline__ d3 mu_face_volume_grad(const TP& t, int h, int which)
{
    d3 p0 = t.P(t.from(h)), p1 = t.P(t.to(h)), p2 = t.P(t.to(t.next(h)));
    if (which == 0) return cross(p1, p2) / 6.0;
    if (which == 1) return cross(p2, p0) / 6.0;
    return cross(p0, p1) / 6.0;
}
// dihedral_angle_grad<N> for a single bit b in [from, to, opp-in-face,
// opp-in-twin-face] (mesh_util.h:135-185)
template <class TP>
__device__ __forceinline__ d3 mu_dihedral_angle_grad(const TP& t, int h, int b)
{
    d3 n0 = mu_face_normal(t, h);
    d3 n1 = mu_face_normal(t, h ^ 1);
    double l = mu_edge_length(t, h);
    if (b == 0) {
        double a03 = om_sector_angle(t, h);
        double a04 = om_sector_angle(t, t.prev(h ^ 1));
        return (1. / tan(a03) * n0 + 1. / tan(a04) * n1) / l;
    }
    if (b == 1) {
        double a01 = om_sector_angle(t, t.prev(h));
        double a02 = om_sector_angle(t, h ^ 1);
        return (1. / tan(a01) * n0 + 1. / tan(a02) * n1) / l;
    }
    if (b == 2) {
        double a01 = om_sector_angle(t, t.prev(h));
        double a03 = om_sector_angle(t, h);
        return -(1. / tan(a01) + 1. / tan(a03)) * n0 / l;
    }
    double a02 = om_sector_angle(t, h ^ 1);
    double a04 = om_sector_angle(t, t.prev(h ^ 1));
    return -(1. / tan(a02) + 1. / tan(a04)) * n1 / l;
}

// x^r for the integer exponents of the potentials (std::pow with an int
// exponent in the reference); exact for the small powers used in practice.
__device__ __forceinline__ double ipow(double x, int r)
{
    double y = 1.0;
    for (int i = 0; i < r; i++) y *= x;
    return y;
}

// ------------------------------------------------- bond potentials --------
// mesh_tether.h:40-55 (edge) and 94-102 (area), per halfedge.  Below lc1 the
// reference divides by l - lc1, by l^r and (in the derivative) by (l - lc1)^2
// and l; here one FP64 division gives w = 1 / (l (l - lc1)), from which
// 1 / l = (l - lc1) w and 1 / (l - lc1) = l w, and the rest are products
// (each a few ulp from the divided form).
__device__ __forceinline__ double bond_edge_value(double l, const EParams& p)
{
    int r = p.bond_r;
    if (l > p.bond_lc0) return ipow((double)r, r + 1) * ipow(l - p.bond_lc0, r);
    if (l < p.bond_lc1) {
        const double lmc = l - p.bond_lc1, w = 1.0 / (l * lmc);
        return exp(l * (l * w)) * ipow(lmc * w, r);
    }
    return 0.0;
}
// derivative factor dV/dl (mesh_tether.h:58-82; the two branches accumulate)
__device__ __forceinline__ double bond_edge_dfac(double l, const EParams& p)
{
    int r = p.bond_r;
    double fac = 0.0;
    if (l > p.bond_lc0) fac = ipow((double)r, r + 2) * ipow(l - p.bond_lc0, r - 1);
    if (l < p.bond_lc1) {
        const double lmc = l - p.bond_lc1, w = 1.0 / (l * lmc);
        const double il = lmc * w, iq = l * w;    // 1 / l, 1 / (l - lc1)
        fac -= p.bond_lc1 * (iq * iq) + r * il;
        fac *= ipow(il, r);
        fac *= exp(l * iq);
    }
    return fac;
}
__device__ __forceinline__ double bond_area_value(double a, const EParams& p)
{
    double d = a / p.bond_a0 - 1.0;
    return d * d;
}
__device__ __forceinline__ double bond_area_dfac(double a, const EParams& p)
{
    return 2.0 * (a / p.bond_a0 - 1.0) / p.bond_a0;
}

// ---------------------------------------------- surface repulsion ---------
// mesh_repulsion.h:38-54 potential and 56-79 gradient factor (per pair)
__device__ __forceinline__ double rep_value(double l, double lc1, int r)
{
    return exp(l / (l - lc1)) / ipow(l, r);
}
__device__ __forceinline__ double rep_dfac(double l, double lc1, int r)
{
    double fac = 0.0;
    double lr = ipow(l, r + 1);
    double lmc = l - lc1;
    fac -= lc1 / (lmc * lmc) + r / l;
    fac /= lr;
    fac *= exp(l / lmc);
    return fac;
}

// Is j inside the exclusion neighbourhood of i (cell_list.h:170-202)?
// Uses the halfedge circulators, so it sees the current topology.
template <class TP>
__device__ __forceinline__ bool om_excluded(const TP& t, int i, int j, int excl)
{
    if (i == j) return true;
    if (excl <= 0) return false;
    int h0 = t.vhe(i);
    if (h0 < 0) return false;
    int h = h0;
    int guard = 0;
    do {
        int vi = t.to(h);
        if (vi == j) return true;
        if (excl > 1) {
            int g0 = t.vhe(vi);
            int g = g0;
            int guard2 = 0;
            do {
                if (t.to(g) == j) return true;
                g = t.cw(g);
            } while (g != g0 && ++guard2 < 4 * kMaxValence);
        }
        h = t.cw(h);
    } while (h != h0 && ++guard < 4 * kMaxValence);
    return false;
}

// Cell coordinates of a position (cell_list.h:76, 141), relative to the
// occupied sub-grid.  The quotient is clamped to two cells beyond the
// sub-grid before the conversion: every stencil from there on is empty, as
// it is for the reference's unclamped coordinate, and the conversion stays
// defined for a vertex that has moved arbitrarily far (or is not finite).
__device__ __forceinline__ int cell_coord(const NListView& nl, double p, int k)
{
    const double q = (p - nl.box_min[k]) / nl.r_list[k];
    return (int)fmin(fmax(q, (double)(nl.lo[k] - 2)), (double)(nl.lo[k] + nl.shape[k] + 1)) - nl.lo[k];
}

__device__ __forceinline__ void cell_coords(const NListView& nl, d3 p, int c[3])
{
    c[0] = cell_coord(nl, p.x, 0);
    c[1] = cell_coord(nl, p.y, 1);
    c[2] = cell_coord(nl, p.z, 2);
}

// Visit every neighbour j of vertex i within dmax (point_distances of
// cell_list.h:133-219 / verlet_list.h:76-101), as a set.  The exclusion
// test is applied only to candidates inside dmax, which leaves the set
// unchanged and skips most ring walks.  The distance test is the
// reference's norm(x_i - x_j) < dmax, evaluated exactly in FP64.  f(j, d)
// gets d = x_i - x_j.
template <class Excl, class F>
__device__ __forceinline__ void nlist_visit(const NListView& nl, const double* __restrict__ x, int i, d3 pi,
                                            double dmax, Excl&& excluded, F&& f)
{
    // squared-distance prefilter; every pair with sqrt(s) < dmax (correctly
    // rounded) has s below dmax^2 (1 + 1e-15), so only the exact test decides
    const double dmax2 = dmax * dmax * (1.0 + 1e-15);
    if (nl.kind == 0) {
        int c[3];
        cell_coords(nl, pi, c);
        for (int dk = -1; dk < 2; dk++) {
            int kk = c[2] + dk;
            if (kk < 0 || kk >= nl.shape[2]) continue;
            for (int dj = -1; dj < 2; dj++) {
                int jj = c[1] + dj;
                if (jj < 0 || jj >= nl.shape[1]) continue;
                // cells (ii-1 .. ii+1, jj, kk) are consecutive flat ids
                int i0 = max(c[0] - 1, 0);
                int i1 = min(c[0] + 1, nl.shape[0] - 1);
                if (i0 > i1) continue;
                int base = jj * nl.strides[1] + kk * nl.strides[2];
                int s = nl.cell_start[base + i0];
                int e = nl.cell_start[base + i1 + 1];
                for (int q = s; q < e; q++) {
                    int j = nl.cell_verts[q];
                    d3 d = pi - d3{x[3 * j], x[3 * j + 1], x[3 * j + 2]};
                    double sq = sqnorm(d);
                    if (sq <= dmax2 && sqrt(sq) < dmax && j != i && !excluded(i, j)) f(j, d);
                }
            }
        }
    } else if (nl.kind == 1) {
        int s = nl.adj_start[i], e = nl.adj_start[i + 1];
        for (int q = s; q < e; q++) {
            int j = nl.adj[q];
            d3 d = pi - d3{x[3 * j], x[3 * j + 1], x[3 * j + 2]};
            double sq = sqnorm(d);
            if (sq <= dmax2 && sqrt(sq) < dmax) f(j, d);
        }
    }
}

// Per-vertex properties via halfedge circulation (mesh_properties.cpp:12-49).
// Returns area, volume, curvature, bending, tethering, repulsion.
struct VP {
    double v[6];
};

template <class TP>
__device__ inline VP om_vertex_properties(const TP& t, const NListView& nl, const EParams& p, int ve)
{
    VP r;
    double area = 0, vol = 0, curv = 0, teth = 0;
    int h0 = t.vhe(ve);
    if (h0 >= 0) {
        int h = h0;
        int guard = 0;
        do {
            if (!t.he_boundary(h)) {
                double el = mu_edge_length(t, h);
                double da = om_dihedral_angle(t, h);
                curv += 0.5 * el * da;
                area += mu_face_area(t, h);
                vol += mu_face_volume(t, h);
                if (p.bond_type == 0) teth += bond_edge_value(el, p);
                else teth += bond_area_value(mu_face_area(t, h), p);
            }
            h = t.cw(h);
        } while (h != h0 && ++guard < 4 * kMaxValence);
    }
    area /= 3;
    vol /= 3;
    curv /= 2;
    teth /= (p.bond_type == 0 ? 2 : 3);
    r.v[0] = area;
    r.v[1] = vol;
    r.v[2] = curv;
    r.v[3] = 2 * curv * curv / area;
    r.v[4] = teth;
    double rep = 0;
    if (p.rep_lc1 > 0 && nl.kind >= 0) {
        d3 pi = t.P(ve);
        auto ex = [&](int a, int b) { return nl.kind == 0 && om_excluded(t, a, b, nl.exclusion); };
        nlist_visit(nl, t.x, ve, pi, p.rep_lc1, ex,
                    [&](int, d3 d) { rep += rep_value(norm(d), p.rep_lc1, p.rep_r); });
    }
    r.v[5] = rep;
    return r;
}

// Gradient of trimem's energy at vertex ve gathered as in
// mesh_properties.cpp:51-148 and combined as in kernel.h:124-137.
// q[v] = curvature/area of vertex v.
template <class TP>
__device__ inline d3 om_vertex_gradient(const TP& t, const NListView& nl, const EParams& p, const Globals& g,
                                        const double* __restrict__ q, int ve)
{
    d3 dA{0, 0, 0}, dV{0, 0, 0}, dC{0, 0, 0}, dB{0, 0, 0}, dT{0, 0, 0}, dR{0, 0, 0};
    double c0 = q[ve];
    int h0 = t.vhe(ve);
    if (h0 >= 0) {
        int he = h0;
        int guard = 0;
        do {
            if (!t.he_boundary(he)) {
                int n_he = t.next(he);
                double c1 = q[t.to(he)];
                double c2 = q[t.to(n_he)];
                double c_to_a[3] = {c0, c1, c2};
                double el = mu_edge_length(t, he);
                double ea = om_dihedral_angle(t, he);
                d3 dl = mu_edge_length_grad_from(t, he);
                d3 dang = mu_dihedral_angle_grad(t, he, 0);
                for (int i = 0; i < 2; i++) {
                    d3 val = 0.25 * (ea * dl + el * dang);
                    dC += val;
                    dB += 4.0 * c_to_a[i] * val;
                }
                d3 dfa = mu_face_area_grad(t, he, 0);
                for (int i = 0; i < 3; i++) {
                    d3 val = dfa / 3.0;
                    dA += val;
                    dB -= 2.0 * c_to_a[i] * c_to_a[i] * val;
                }
                d3 dfv = mu_face_volume_grad(t, he, 0);
                for (int i = 0; i < 3; i++) dV += dfv / 3.0;
                el = mu_edge_length(t, n_he);
                dang = mu_dihedral_angle_grad(t, n_he, 2);
                for (int i = 1; i < 3; i++) {
                    d3 val = 0.25 * el * dang;
                    dC += val;
                    dB += 4.0 * c_to_a[i] * val;
                }
                if (p.bond_type == 0) {
                    double l = mu_edge_length(t, he);
                    dT += bond_edge_dfac(l, p) * mu_edge_length_grad_from(t, he);
                } else {
                    dT += bond_area_dfac(mu_face_area(t, he), p) * mu_face_area_grad(t, he, 0);
                }
            }
            he = t.cw(he);
        } while (he != h0 && ++guard < 4 * kMaxValence);
    }
    if (p.rep_lc1 > 0 && nl.kind >= 0 && p.kappa_r != 0) {
        d3 pi = t.P(ve);
        auto ex = [&](int a, int b) { return nl.kind == 0 && om_excluded(t, a, b, nl.exclusion); };
        nlist_visit(nl, t.x, ve, pi, p.rep_lc1, ex,
                    [&](int, d3 d) { dR += 2.0 * (rep_dfac(norm(d), p.rep_lc1, p.rep_r) * d); });
    }
    d3 out{0, 0, 0};
    out += g.fac_a * dA;
    out += g.fac_v * dV;
    out += g.fac_c * dC;
    out += p.kappa_t * dT;
    out += p.kappa_r * dR;
    out += p.kappa_b * dB;
    return out;
}

}    // namespace tmh
