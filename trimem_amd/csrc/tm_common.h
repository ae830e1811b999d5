// tm_common.h -- types shared by the host code and the gfx950 kernels.
//
// Device layout of one mesh (all arrays in HBM, structure of arrays):
//   x        double[3N]  positions, row-major (N,3) like the numpy points view
//   he_to    int[H]      target vertex of halfedge h   (edge e owns 2e, 2e+1)
//   he_next  int[H]      next halfedge in its face (or boundary loop)
//   he_prev  int[H]      previous halfedge
//   he_face  int[H]      face of h, -1 on the boundary
//   v_he     int[N]      an outgoing halfedge (the boundary one on boundaries)
//   f_he     int[F]      a halfedge of face f
//   ring     int[W*N]    one-ring in clockwise order, column-major:
//                        ring[k*N + v] = k-th neighbour of v
//   meta     int[N]      valence | (boundary slot + 1) << 16
// The ring table is what the two energy passes read; the halfedge arrays are
// what flips rewrite and what the OpenMesh-semantics fallback walks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define TM_HD __host__ __device__ __forceinline__

namespace tmh {

constexpr int kMaxValence = 32;    // ring table width W
constexpr int kRingPad = 8;        // ring columns [valence, kRingPad) hold -1

template <class T>
struct V3 {
    T x, y, z;
};
template <class T> TM_HD V3<T> mk(T a, T b, T c) { return V3<T>{a, b, c}; }
template <class T> TM_HD V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> TM_HD V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> TM_HD V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> TM_HD V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> TM_HD V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> TM_HD V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> TM_HD V3<T>& operator+=(V3<T>& a, V3<T> b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
template <class T> TM_HD V3<T>& operator-=(V3<T>& a, V3<T> b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
template <class T> TM_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> TM_HD V3<T> cross(V3<T> a, V3<T> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> TM_HD T sqnorm(V3<T> a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
template <class T> TM_HD T norm(V3<T> a) { return sqrt(sqnorm(a)); }
template <class D, class S> TM_HD V3<D> cvt(V3<S> a) { return {(D)a.x, (D)a.y, (D)a.z}; }

using d3 = V3<double>;
using f3 = V3<float>;

// Flattened trimem EnergyParams (src/params.h:19-71).
struct EParams {
    double kappa_b, kappa_a, kappa_v, kappa_c, kappa_t, kappa_r;
    double bond_lc0, bond_lc1, bond_a0;
    int bond_r, bond_type;    // bond_type 0 edge, 1 area
    double rep_lc1;
    int rep_r, rep_exclusion;
    // reference properties (energy.cpp:30-45), set per call by the host
    double ref_area, ref_volume, ref_curvature;
};

// Global properties and derived gradient factors, written by the reduction.
struct Globals {
    double props[6];    // area, volume, curvature, bending, tethering, repulsion
    double energy;
    double fac_a, fac_v, fac_c;    // dE/dA, dE/dV, dE/dC (kernel.h:27-71)
    double fac_v6;                 // fac_v / 6 (the volume gradient's factor per sector normal)
};

// Read-only device view of a mesh.
struct MeshView {
    const double* __restrict__ x;
    const int* __restrict__ he_to;
    const int* __restrict__ he_next;
    const int* __restrict__ he_prev;
    const int* __restrict__ he_face;
    const int* __restrict__ v_he;
    const int* __restrict__ f_he;
    const int* __restrict__ ring;
    const int* __restrict__ meta;
    int nv, nh, nf;
};

// Neighbour-list view.  kind 0: cell list (cells frozen at build time, queried
// with current positions); kind 1: Verlet list (CSR adjacency); kind -1: none.
struct NListView {
    int kind;
    int exclusion;
    // cell list.  Cell coordinates are the reference's, int((p - box_min) /
    // r_list) per dimension; the table covers only the occupied sub-grid,
    // global coordinates lo .. lo + shape - 1, and stores them relative to lo
    double box_min[3];
    double r_list[3];
    int lo[3];
    int shape[3];
    int strides[3];
    const int* __restrict__ cell_start;    // [ncells + 1]
    const int* __restrict__ cell_verts;    // vertices sorted by (cell, id)
    const int* __restrict__ vert_cell;     // build-time cell of each vertex
    const int* __restrict__ perm;          // internal -> user vertex id
    // per occupied cell, the vertices of its 27-cell stencil, stored once
    // per cell; vcand[v] is the list of the build-time cell of v
    const int2* __restrict__ vcand;        // [N] range in cand of the build-time cell of each vertex
    const int* __restrict__ cand;
    const int* __restrict__ cand_start;    // [N + 1] per slot of cell_verts; a cell's list starts at its first slot
    // per vertex, bit i set if the i-th candidate of its list is excluded
    // (topology only; 2 x 64 bits, the first 128 candidates); null if not
    // built. Read only by the vertex-group query of the energy passes, which
    // run on a fixed topology; the flip kernels' walkers ignore it
    const unsigned long long* __restrict__ excl_mask;
    // verlet list
    const int* __restrict__ adj_start;     // [N + 1]
    const int* __restrict__ adj;           // neighbour ids, ascending per vertex
};

TM_HD int meta_valence(int m) { return m & 0xffff; }
TM_HD int meta_bslot(int m) { return (m >> 16) - 1; }

}    // namespace tmh
