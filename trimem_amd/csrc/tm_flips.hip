// tm_flips.hip -- ring-table rebuild and the edge-flip sweeps: the serial
// sweep (one workgroup, k_flip_serial, or windows of proposals evaluated at
// once and committed in order, k_sf_*) and the parallel sweep (k_pflip_*).
//
// flip_serial (src/flips.cpp:22-80) is a Metropolis loop in which every
// proposal depends on the topology left by the previous one.  The
// one-workgroup kernel walks the proposal stream in order.  For each proposal:
//   * lane 0 tests OpenMesh is_flip_ok and the boundary condition;
//   * the workgroup evaluates vertex_properties of the four patch vertices
//     (flip_utils.cpp:19-41), 16 lanes per vertex (coop_patch), including
//     their repulsion queries;
//   * lane 0 flips (OpenMesh TriConnectivity::flip rewiring), the lanes
//     re-evaluate the patch, and lane 0 applies the FP64 Metropolis test
//     u <= (dE < 0 ? 1 : exp(-dE)) and flips back on rejection.
// A uniform is consumed only for proposals that pass the validity test
// (flips.cpp:45, 64).  Proposals and uniforms are either injected by the
// caller or drawn from a counter-based generator keyed by a seed.
#include "tm_internal.h"
#include "tm_om.h"
#include "tm_rng.h"

#include <algorithm>

namespace tmh {

// ------------------------------------------------------- ring rebuild -----
__global__ void k_build_ring(MeshView m, int* __restrict__ ring, int* __restrict__ meta, int* __restrict__ overflow)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.nv) return;
    const int N = m.nv;
    int h0 = m.v_he[v];
    int n = 0, b = -1;
    if (h0 >= 0) {
        int h = h0;
        do {
            if (n < kMaxValence) ring[n * N + v] = m.he_to[h];
            if (m.he_face[h] < 0 && b < 0) b = n;
            n++;
            h = m.he_next[h ^ 1];
        } while (h != h0 && n <= kMaxValence);
    }
    if (n > kMaxValence) {
        atomicOr(overflow, 1);
        n = kMaxValence;
        b = 0;    // route to the halfedge walker
    }
    for (int k = n; k < kRingPad; k++) ring[k * N + v] = -1;
    meta[v] = n | ((b + 1) << 16);
}

void launch_build_ring(const MeshView& m, int* ring, int* meta, int* overflow, hipStream_t s)
{
    if (m.nv == 0) return;
    hipLaunchKernelGGL(k_build_ring, dim3((m.nv + 255) / 256), dim3(256), 0, s, m, ring, meta, overflow);
}

// ------------------------------------------------------- permutations -----
__global__ void k_gather_rows(const double* __restrict__ src, const int* __restrict__ idx, double* __restrict__ dst,
                              int n, int w)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * w) return;
    int i = t / w, c = t - i * w;
    dst[t] = src[(size_t)idx[i] * w + c];
}

void launch_gather_rows(const double* src, const int* idx, double* dst, int n, int w, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather_rows, dim3((n * w + 255) / 256), dim3(256), 0, s, src, idx, dst, n, w);
}

// ---------------------------------------------------------------- RNG -----
__device__ __forceinline__ double rng_uniform(uint64_t seed, uint64_t i) { return rng_unit(rng_draw(seed, 1, i)); }
__device__ __forceinline__ int rng_edge(uint64_t seed, uint64_t i, int ne)
{
    return (int)(((rng_draw(seed, 0, i) >> 32) * (uint64_t)ne) >> 32);
}

// ------------------------------------------------------------ helpers -----
// OpenMesh TriConnectivity::is_flip_ok (not a boundary edge, the two
// opposite vertices distinct and not yet neighbours), plus one condition the
// reference does not have: a flip adds a neighbour to each opposite vertex,
// and it is refused when either would exceed the ring-table width
// kMaxValence.  Meshes are created with valences of at most kMaxValence, so
// every topology a sweep can reach stays within the table.  The two one-rings
// are walked in one loop so that their dependent loads overlap.
__device__ __forceinline__ bool flip_ok(const Topo& t, int e)
{
    if (t.edge_boundary(e)) return false;
    int ah = t.to(t.next(2 * e));
    int bh = t.to(t.next(2 * e + 1));
    if (ah == bh) return false;
    const int ha0 = t.vhe(ah), hb0 = t.vhe(bh);
    int ha = ha0, hb = hb0, na = 0, nb = 0;
    bool da = false, db = false, adjacent = false;
    do {
        if (!da) {
            adjacent = adjacent || t.to(ha) == bh;
            ha = t.cw(ha);
            da = ha == ha0 || ++na >= 4 * kMaxValence;
        }
        if (!db) {
            hb = t.cw(hb);
            db = hb == hb0 || ++nb >= 4 * kMaxValence;
        }
    } while (!(da && db));
    // na, nb count the steps before the walk closed: valence - 1
    return !adjacent && na + 1 < kMaxValence && nb + 1 < kMaxValence;
}

__device__ __forceinline__ void do_flip(const FlipMeshView& fm, int e)
{
    // OpenMesh TriConnectivity::flip
    int a0 = 2 * e, b0 = 2 * e + 1;
    int a1 = fm.he_next[a0], a2 = fm.he_next[a1];
    int b1 = fm.he_next[b0], b2 = fm.he_next[b1];
    int va0 = fm.he_to[a0], va1 = fm.he_to[a1];
    int vb0 = fm.he_to[b0], vb1 = fm.he_to[b1];
    int fa = fm.he_face[a0], fb = fm.he_face[b0];
    fm.he_to[a0] = va1;
    fm.he_to[b0] = vb1;
    fm.he_next[a0] = a2; fm.he_prev[a2] = a0;
    fm.he_next[a2] = b1; fm.he_prev[b1] = a2;
    fm.he_next[b1] = a0; fm.he_prev[a0] = b1;
    fm.he_next[b0] = b2; fm.he_prev[b2] = b0;
    fm.he_next[b2] = a1; fm.he_prev[a1] = b2;
    fm.he_next[a1] = b0; fm.he_prev[b0] = a1;
    fm.he_face[a1] = fb;
    fm.he_face[b1] = fa;
    fm.f_he[fa] = a0;
    fm.f_he[fb] = b0;
    if (fm.v_he[va0] == b0) fm.v_he[va0] = a1;
    if (fm.v_he[vb0] == a0) fm.v_he[vb0] = b1;
}

// The topology after flipping interior edge e, without writing it: every
// accessor returns what do_flip stores, so a workgroup can evaluate the
// flipped patch while others read the unflipped mesh.
struct FlippedTopo {
    Topo t;
    const double* x;
    int a0, a1, a2, b0, b1, b2;
    int va1, vb1, va0, vb0, fa, fb;
    int vhe_a, vhe_b;    // v_he of va0 and vb0 after the flip

    __device__ FlippedTopo(const Topo& base, int e) : t(base), x(base.x)
    {
        a0 = 2 * e;
        b0 = 2 * e + 1;
        a1 = t.next(a0);
        a2 = t.next(a1);
        b1 = t.next(b0);
        b2 = t.next(b1);
        va0 = t.to(a0);
        va1 = t.to(a1);
        vb0 = t.to(b0);
        vb1 = t.to(b1);
        fa = t.face(a0);
        fb = t.face(b0);
        const int ha = t.vhe(va0), hb = t.vhe(vb0);
        vhe_a = ha == b0 ? a1 : ha;
        vhe_b = hb == a0 ? b1 : hb;
    }
    __device__ __forceinline__ d3 P(int v) const { return t.P(v); }
    __device__ __forceinline__ int to(int h) const { return h == a0 ? va1 : (h == b0 ? vb1 : t.to(h)); }
    __device__ __forceinline__ int from(int h) const { return to(h ^ 1); }
    __device__ __forceinline__ int next(int h) const
    {
        if (h == a0) return a2;
        if (h == a2) return b1;
        if (h == b1) return a0;
        if (h == b0) return b2;
        if (h == b2) return a1;
        if (h == a1) return b0;
        return t.next(h);
    }
    __device__ __forceinline__ int prev(int h) const
    {
        if (h == a2) return a0;
        if (h == b1) return a2;
        if (h == a0) return b1;
        if (h == b2) return b0;
        if (h == a1) return b2;
        if (h == b0) return a1;
        return t.prev(h);
    }
    __device__ __forceinline__ int face(int h) const { return h == a1 ? fb : (h == b1 ? fa : t.face(h)); }
    __device__ __forceinline__ bool he_boundary(int h) const { return face(h) < 0; }
    __device__ __forceinline__ bool edge_boundary(int e) const { return face(2 * e) < 0 || face(2 * e + 1) < 0; }
    __device__ __forceinline__ int cw(int h) const { return next(h ^ 1); }
    __device__ __forceinline__ int vhe(int v) const { return v == va0 ? vhe_a : (v == vb0 ? vhe_b : t.vhe(v)); }
    __device__ __forceinline__ int fhe(int f) const { return f == fa ? a0 : (f == fb ? b0 : t.fhe(f)); }
};

// do_flip(fm, e) twice (a rejected proposal), stored in one pass.  The
// second flip undoes the connectivity but not the ids: the edge's two
// halfedges swap direction, the two faces swap triangles, and the v_he of
// the edge's endpoints move off the swapped halfedges exactly as the first
// flip moved them (the second flip's v_he tests cannot hold, since its
// endpoints' outgoing halfedges are not the edge's).
__device__ __forceinline__ void do_flip_twice(const FlipMeshView& fm, int e)
{
    int a0 = 2 * e, b0 = 2 * e + 1;
    int a1 = fm.he_next[a0], a2 = fm.he_next[a1];
    int b1 = fm.he_next[b0], b2 = fm.he_next[b1];
    int va0 = fm.he_to[a0], vb0 = fm.he_to[b0];
    int ta2 = fm.he_to[a2], tb2 = fm.he_to[b2];
    int fa = fm.he_face[a0], fb = fm.he_face[b0];
    int ha = fm.v_he[va0], hb = fm.v_he[vb0];
    fm.he_to[a0] = ta2;
    fm.he_to[b0] = tb2;
    fm.he_next[a0] = b1; fm.he_prev[b1] = a0;
    fm.he_next[b1] = b2; fm.he_prev[b2] = b1;
    fm.he_next[b2] = a0; fm.he_prev[a0] = b2;
    fm.he_next[b0] = a1; fm.he_prev[a1] = b0;
    fm.he_next[a1] = a2; fm.he_prev[a2] = a1;
    fm.he_next[a2] = b0; fm.he_prev[b0] = a2;
    fm.he_face[a1] = fb;
    fm.he_face[a2] = fb;
    fm.he_face[b1] = fa;
    fm.he_face[b2] = fa;
    fm.f_he[fa] = a0;
    fm.f_he[fb] = b0;
    if (ha == b0) fm.v_he[va0] = a1;
    if (hb == a0) fm.v_he[vb0] = b1;
}

__device__ __forceinline__ double energy_of(const double pr[6], const EParams& p)
{
    // trimem_energy (kernel.h:110-122)
    double e = 0.0, d;
    d = pr[0] / p.ref_area - 1.0;
    e += p.kappa_a * d * d;
    d = pr[1] / p.ref_volume - 1.0;
    e += p.kappa_v * d * d;
    d = pr[2] / p.ref_curvature - 1.0;
    e += p.kappa_c * d * d;
    e += p.kappa_t * pr[4];
    e += p.kappa_r * pr[5];
    e += p.kappa_b * pr[3];
    return e;
}

// ------------------------------------------ cooperative patch evaluation --
// vertex_properties (mesh_properties.cpp:12-49) of the four vertices of an
// edge's flip patch (flip_utils.cpp:19-41), to(2e), to(next(2e)), to(2e+1),
// to(next(2e+1)), by one 64-lane workgroup: lanes 16g .. 16g+15 evaluate
// patch vertex g.
//   phase 1  lane 0 of the group walks the one-ring (clockwise from v_he,
//            as the circulator does) and locates the candidate runs of the
//            repulsion query;
//   phase 2  the lanes compute the per-halfedge terms;
//   phase 3  the lanes test the candidates in chunks of 64, the exclusion
//            test reading the one-ring from LDS (coop_excluded);
//   phase 4  lane 0 adds the terms of each accumulator in the order the
//            single-thread walker (om_vertex_properties) adds them, so every
//            value is bit-identical to it; thread 0 adds the four vertices.
// A vertex whose valence exceeds kMaxValence is evaluated by its lane 0
// with om_vertex_properties.
constexpr int kCoopLanes = 16;
// The parallel sweep's patch kernels keep to 128 registers (four waves per
// SIMD, 16 workgroups per CU): a sweep over a replica group has far more
// patches than the GPU holds at once, and occupancy, not one patch's
// latency, then bounds it (at 153 registers, 12 workgroups per CU, a round
// over 32 meshes of 491520 edges took 21% longer).
#define TM_COOP_OCCUPANCY __attribute__((amdgpu_waves_per_eu(4)))
constexpr int kCoopChunk = 4 * kCoopLanes;

struct CoopShared {
    int h[4][kMaxValence];              // outgoing halfedges, circulator order
    int nb[4][kMaxValence];             // their target vertices (the one-ring)
    int bnd[4][kMaxValence];            // halfedge on the boundary
    int n[4];                           // valence; -1 above kMaxValence
    int total[4];                       // repulsion candidates
    double slot[4][kMaxValence][4];     // area, volume, edge curvature, tether
    int run_s[4][9], run_len[4][9];     // candidate runs (CandRuns) of the repulsion query
    double pos[4][3];                   // the patch vertex's position
    int ve[4];                          // the patch vertices
    int flip_ok;                        // flip_ok(e), when coop_patch tests it
    int lost[4];                        // a patch edge of the vertex is owned by another proposal
    double rep[4][kCoopChunk];
    double vp[4][6];
    double sum[6];
};

template <class TP>
__device__ __forceinline__ int patch_vertex(const TP& t, int e, int g)
{
    int h = 2 * e + (g >> 1);
    return (g & 1) ? t.to(t.next(h)) : t.to(h);
}

// The candidate runs of a vertex's repulsion query, in the order
// nlist_visit walks them: the nine (dj, dk) rows of three consecutive cells
// of cell_verts, or the vertex's adj range of the Verlet list.  Rows that
// fall outside the grid are empty runs, so every index is a constant after
// unrolling and the runs stay in registers.
struct CandRuns {
    int s[9], len[9];
    int total;
};

__device__ __forceinline__ CandRuns cand_runs(const NListView& nl, int i, d3 pi)
{
    CandRuns r;
#pragma unroll
    for (int k = 0; k < 9; k++) r.s[k] = r.len[k] = 0;
    if (nl.kind == 0) {
        int c[3];
        cell_coords(nl, pi, c);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            int kk = c[2] + k / 3 - 1, jj = c[1] + k % 3 - 1;
            int i0 = max(c[0] - 1, 0), i1 = min(c[0] + 1, nl.shape[0] - 1);
            if (kk >= 0 && kk < nl.shape[2] && jj >= 0 && jj < nl.shape[1] && i0 <= i1) {
                int base = jj * nl.strides[1] + kk * nl.strides[2];
                r.s[k] = nl.cell_start[base + i0];
                r.len[k] = nl.cell_start[base + i1 + 1] - r.s[k];
            }
        }
    } else if (nl.kind == 1) {
        r.s[0] = nl.adj_start[i];
        r.len[0] = nl.adj_start[i + 1] - r.s[0];
    }
    r.total = 0;
#pragma unroll
    for (int k = 0; k < 9; k++) r.total += r.len[k];
    return r;
}

// Slot of the c-th candidate (0 <= c < total) in cell_verts / adj.
__device__ __forceinline__ int cand_slot(const CandRuns& r, int c)
{
    int q = -1;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        if (q < 0 && c < r.len[k]) q = r.s[k] + c;
        if (q < 0) c -= r.len[k];
    }
    return q;
}

__device__ __forceinline__ bool coop_in_ring(const CoopShared& S, int g, int n, int j)
{
    for (int k = 0; k < n; k++)
        if (S.nb[g][k] == j) return true;
    return false;
}

// om_excluded(ve, j) for patch vertex g (valence n <= kMaxValence, one-ring
// in S.nb).  Adjacency is symmetric, so j lies in the one-ring of a ring
// vertex of ve exactly when j's own one-ring meets ve's: the walk runs
// around j, and only for candidates within the cutoff that are not in ve's
// one-ring, instead of around every ring vertex of every patch vertex.  A
// candidate of valence above kMaxValence takes om_excluded.
template <class TP>
__device__ __forceinline__ bool coop_excluded(const TP& t, const CoopShared& S, int g, int n, int ve, int excl,
                                              int j)
{
    if (excl <= 0) return false;
    if (coop_in_ring(S, g, n, j)) return true;
    if (excl == 1) return false;
    const int h0 = t.vhe(j);
    if (h0 < 0) return false;
    int h = h0, m = 0;
    do {
        if (m == kMaxValence) return om_excluded(t, ve, j, excl);
        if (coop_in_ring(S, g, n, t.to(h))) return true;
        m++;
        h = t.cw(h);
    } while (h != h0);
    return false;
}

// Called by all 64 threads of the workgroup; S.vp and S.sum hold the result
// on return.  test_flip (on the topology before the flip, e interior): S.flip_ok
// = flip_ok(t, e) from the one-rings of phase 1, which include those of the
// two vertices opposite the edge (patch vertices 1 and 3), and nothing more
// is evaluated when it is false.  owner (same topology): S.lost = some edge
// of the parallel sweep's patch, the spokes h >> 1 and rims next(h) >> 1 of
// every outgoing halfedge h of the four vertices (the flip patch), is
// owned by another key, tested on phase 2's halfedges; nothing more is
// evaluated then.
template <class TP>
__device__ void coop_patch(const TP& t, const NListView& nl, const EParams& p, int e, CoopShared& S,
                           bool test_flip = false, const unsigned long long* __restrict__ owner = nullptr,
                           unsigned long long key = 0)
{
    const int g = threadIdx.x / kCoopLanes, l = threadIdx.x % kCoopLanes;
    const int ve = patch_vertex(t, e, g);
    const bool rep_on = p.rep_lc1 > 0 && nl.kind >= 0;
    const int excl = nl.kind == 0 ? nl.exclusion : 0;

    // phase 1 (the candidate runs wait in LDS while phase 2 needs the registers)
    if (l == 0) {
        const d3 pv = t.P(ve);
        S.pos[g][0] = pv.x;
        S.pos[g][1] = pv.y;
        S.pos[g][2] = pv.z;
        const CandRuns runs = cand_runs(nl, ve, pv);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            S.run_s[g][k] = runs.s[k];
            S.run_len[g][k] = runs.len[k];
        }
        S.total[g] = rep_on ? runs.total : 0;
        int n = 0;
        int h0 = t.vhe(ve);
        if (h0 >= 0) {
            int h = h0;
            do {
                if (n == kMaxValence) {
                    n = -1;
                    break;
                }
                S.h[g][n] = h;
                S.nb[g][n] = t.to(h);
                n++;
                h = t.cw(h);
            } while (h != h0);
        }
        S.n[g] = n;
        S.ve[g] = ve;
        if (n < 0) S.total[g] = 0;
        // a ring too long for LDS: its patch edges by a walk of their own
        bool lost = false;
        if (owner && n < 0 && h0 >= 0) {
            int h = h0, guard = 0;
            do {
                lost = lost || owner[h >> 1] != key || owner[t.next(h) >> 1] != key;
                h = t.cw(h);
            } while (h != h0 && ++guard < 4 * kMaxValence);
        }
        S.lost[g] = lost;
    }
    __syncthreads();
    if (test_flip) {
        // flip_ok's walks: the opposite vertices differ, are not adjacent
        // and have valence below kMaxValence (n is -1 above it)
        if (threadIdx.x == 0) {
            const int a = S.ve[1], b = S.ve[3], na = S.n[1], nb = S.n[3];
            bool ok = a != b && na >= 0 && na < kMaxValence && nb >= 0 && nb < kMaxValence;
            for (int k = 0; ok && k < na; k++) ok = S.nb[1][k] != b;
            S.flip_ok = ok;
        }
        __syncthreads();
        if (!S.flip_ok) return;
    }
    const int n = S.n[g];

    // phase 2
    for (int s = l; s < n; s += kCoopLanes) {
        const int h = S.h[g][s];
        const int b = t.he_boundary(h);
        S.bnd[g][s] = b;
        if (!b) {
            double el = mu_edge_length(t, h);
            double da = om_dihedral_angle(t, h);
            S.slot[g][s][0] = mu_face_area(t, h);
            S.slot[g][s][1] = mu_face_volume(t, h);
            S.slot[g][s][2] = 0.5 * el * da;
            S.slot[g][s][3] = p.bond_type == 0 ? bond_edge_value(el, p) : bond_area_value(mu_face_area(t, h), p);
        }
        if (owner && (owner[h >> 1] != key || owner[t.next(h) >> 1] != key)) S.lost[g] = 1;
    }
    __syncthreads();
    if (owner && (S.lost[0] | S.lost[1] | S.lost[2] | S.lost[3])) return;

    // phase 3
    const d3 pi{S.pos[g][0], S.pos[g][1], S.pos[g][2]};
    CandRuns runs;
#pragma unroll
    for (int k = 0; k < 9; k++) {
        runs.s[k] = S.run_s[g][k];
        runs.len[k] = S.run_len[g][k];
    }
    const int total = S.total[g];
    const int tmax = max(max(S.total[0], S.total[1]), max(S.total[2], S.total[3]));
    const double dmax = p.rep_lc1;
    const double dmax2 = dmax * dmax * (1.0 + 1e-15);    // prefilter, as in nlist_visit
    double rep = 0.0;
    for (int base = 0; base < tmax; base += kCoopChunk) {
        double val[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int c = base + k * kCoopLanes + l;
            val[k] = 0.0;
            if (c < total) {
                const int q = cand_slot(runs, c);
                const int j = nl.kind == 0 ? nl.cell_verts[q] : nl.adj[q];
                const d3 d = pi - t.P(j);
                const double sq = sqnorm(d);
                bool in = sq <= dmax2 && sqrt(sq) < dmax;
                if (in && nl.kind == 0)
                    in = j != ve && !coop_excluded(t, S, g, n, ve, excl, j);
                if (in) val[k] = rep_value(norm(d), p.rep_lc1, p.rep_r);
            }
            S.rep[g][k * kCoopLanes + l] = val[k];
        }
        // a zero term leaves the (non-negative) sum unchanged: lane 0 adds
        // only the lanes whose term is non-zero, in candidate order
        unsigned long long mask[4];
#pragma unroll
        for (int k = 0; k < 4; k++) mask[k] = __ballot(val[k] != 0.0);
        __syncthreads();
        if (l == 0) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                unsigned bits = (unsigned)(mask[k] >> (kCoopLanes * g)) & 0xffffu;
                while (bits) {
                    const int b = __ffs(bits) - 1;
                    bits &= bits - 1;
                    rep += S.rep[g][k * kCoopLanes + b];
                }
            }
        }
        __syncthreads();
    }

    // phase 4
    if (l == 0) {
        VP r;
        if (n < 0) {
            r = om_vertex_properties(t, nl, p, ve);
        } else {
            double area = 0, vol = 0, curv = 0, teth = 0;
            for (int s = 0; s < n; s++) {
                if (S.bnd[g][s]) continue;
                area += S.slot[g][s][0];
                vol += S.slot[g][s][1];
                curv += S.slot[g][s][2];
                teth += S.slot[g][s][3];
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
            r.v[5] = rep;
        }
        for (int k = 0; k < 6; k++) S.vp[g][k] = r.v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 6; k++) S.sum[k] = 0.0;
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 6; k++) S.sum[k] += S.vp[i][k];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(64) k_flip_serial(FlipMeshView fm, NListView nl, EParams p, int nflips,
                                                    const int* __restrict__ proposals,
                                                    const double* __restrict__ uniforms, int n_uniforms,
                                                    uint64_t seed, int injected, const Globals* __restrict__ g0,
                                                    int* __restrict__ result)
{
    __shared__ CoopShared S;
    __shared__ int s_go;
    __shared__ int s_stop;
    const int ne = fm.nh / 2;
    const Topo t{fm.x, fm.he_to, fm.he_next, fm.he_prev, fm.he_face, fm.v_he, fm.f_he};
    double props[6];
    for (int k = 0; k < 6; k++) props[k] = g0->props[k];
    double e0 = energy_of(props, p);
    int acc = 0, used = 0, err = 0;
    if (threadIdx.x == 0) s_stop = 0;
    __syncthreads();
    for (int i = 0; i < nflips; i++) {
        __syncthreads();
        if (threadIdx.x == 0) {
            int idx = injected ? proposals[i] : rng_edge(seed, (uint64_t)i, ne);
            if (idx < 0 || idx >= ne) {
                err = 1;
                s_stop = 1;
                s_go = -1;
            } else {
                s_go = (flip_ok(t, idx) && !t.edge_boundary(idx)) ? idx : -1;
            }
        }
        __syncthreads();
        if (s_stop) break;
        const int e = s_go;
        if (e < 0) continue;
        coop_patch(t, nl, p, e, S);
        double old_sum[6];
        if (threadIdx.x == 0) {
            for (int k = 0; k < 6; k++) {
                old_sum[k] = S.sum[k];
                props[k] -= old_sum[k];
            }
            do_flip(fm, e);
        }
        __threadfence_block();
        __syncthreads();
        coop_patch(t, nl, p, e, S);
        if (threadIdx.x == 0) {
            for (int k = 0; k < 6; k++) props[k] += S.sum[k];
            double en = energy_of(props, p);
            double de = en - e0;
            double alpha = de < 0.0 ? 1.0 : exp(-de);
            double u;
            if (injected) {
                if (used >= n_uniforms) {
                    err = 2;
                    s_stop = 1;
                    u = 2.0;
                } else {
                    u = uniforms[used];
                }
            } else {
                u = rng_uniform(seed, (uint64_t)used);
            }
            used++;
            if (u <= alpha) {
                e0 = en;
                acc += 1;
            } else {
                do_flip(fm, e);
                for (int k = 0; k < 6; k++) {
                    props[k] -= S.sum[k];
                    props[k] += old_sum[k];
                }
            }
        }
        __threadfence_block();
        __syncthreads();
        if (s_stop) break;
    }
    if (threadIdx.x == 0) {
        result[0] = acc;
        result[1] = used;
        result[2] = err;
    }
}

void launch_flip_serial(const FlipMeshView& fm, const NListView& nl, const EParams& p, int nflips,
                        const int* proposals, const double* uniforms, int n_uniforms, uint64_t seed, int injected,
                        const Globals* g0, int* result, hipStream_t s)
{
    hipLaunchKernelGGL(k_flip_serial, dim3(1), dim3(64), 0, s, fm, nl, p, nflips, proposals, uniforms, n_uniforms,
                       seed, injected, g0, result);
}

// ================================================ windowed serial sweep ==
// The serial sweep's decisions and topology, bit for bit, from windows of W
// proposals evaluated at once.  Per window:
//   k_sf_eval    one workgroup per proposal and side: is_flip_ok and the
//                patch properties before the flip (side 0) and after it on
//                FlippedTopo (side 1), on the topology at the window start.
//   k_sf_commit  one workgroup commits the window in order up to the first
//                proposal whose evaluation an earlier one may have changed.
// What an earlier proposal j of the window changes for a later one k:
//   * accepted: the one-rings of j's four patch vertices V(j).  k read the
//     one-rings of its patch vertices P_k (in circulator order) and, for
//     exclusion level 2, the neighbour sets of their one-ring vertices
//     N(P_k); so V(j) meeting P_k (or N(P_k)) makes k stale.
//   * rejected (do_flip twice): every vertex keeps its neighbours in the
//     same cyclic order, and each halfedge a walk reaches leads to the same
//     vertex; only ids move.  The edge's two halfedges swap direction (a
//     later proposal of the same edge sees its patch in another order), and
//     an endpoint whose v_he was one of them starts its circulator
//     elsewhere (a later proposal with that vertex in P_k sums its ring in
//     another order).  The faces' f_he move too; only boundary sector
//     angles read them, and no flip evaluation does.
// Valid proposals stamp their edge and those endpoints with their window
// position (atomicMax of a window-keyed 64-bit key keeps the smallest);
// accepted ones are checked against the later proposals' P_k and N(P_k) as
// they are found.  A rejected proposal can still share a triangle with a
// later committed one, so flips whose faces an earlier valid proposal of the
// window touched (stamped like the edges) are applied after the others, in
// order, by one lane.  Up to the first stale proposal, lanes 0-5 run the serial
// sweep's rounding path (props -= old, += new; on rejection -= new, += old)
// through the rejected branch, one component each; the lanes then take the
// Metropolis decisions at once, and the first acceptance ends the segment:
// its state is the serial one, and the walk continues after it.
// A window commits at least its first proposal, which cannot be stale.
__device__ __forceinline__ unsigned long long sf_key(int window, int W, int k)
{
    return ((unsigned long long)(unsigned)(window + 1) << 32) | (unsigned)(W - k);
}

// bitmap position (12 bits) of a vertex in the commit's staleness filter
__device__ __forceinline__ unsigned sf_hash(int u) { return ((unsigned)u * 2654435761u) >> 20; }

// window position recorded in a stamp, or W if the stamp is older
__device__ __forceinline__ int sf_pos(unsigned long long key, unsigned wid, int W)
{
    return (unsigned)(key >> 32) == wid ? W - (int)(unsigned)(key & 0xffffffffu) : W;
}

__global__ void __launch_bounds__(64) k_sf_init(const Globals* __restrict__ g0, EParams p, int nflips,
                                                SerialState* __restrict__ st)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double props[6];
    for (int k = 0; k < 6; k++) props[k] = st->props[k] = g0->props[k];
    st->e0 = energy_of(props, p);
    st->pos = 0;
    st->done = nflips <= 0;
    st->used = st->acc = st->err = st->window = 0;
}

__global__ void __launch_bounds__(64) k_sf_eval(FlipMeshView fm, NListView nl, EParams p, int nflips,
                                                const int* __restrict__ proposals, uint64_t seed, int injected,
                                                int W, int need_nb, const SerialState* __restrict__ st,
                                                SerialWindow w, unsigned long long* __restrict__ vstamp,
                                                unsigned long long* __restrict__ estamp,
                                                unsigned long long* __restrict__ fstamp)
{
    __shared__ CoopShared S;
    __shared__ int s_ok;
    if (st->done) return;
    const int k = blockIdx.x, side = blockIdx.y, tid = threadIdx.x;
    const int i = st->pos + k;
    if (i >= nflips) return;
    const Topo t{fm.x, fm.he_to, fm.he_next, fm.he_prev, fm.he_face, fm.v_he, fm.f_he};
    const int ne = fm.nh / 2;
    const int e = injected ? proposals[i] : rng_edge(seed, (uint64_t)i, ne);
    if (e < 0 || e >= ne) {
        if (side == 0 && tid == 0) w.valid[k] = -1;
        return;
    }
    // the vertices whose one-rings decide validity and the patch terms (for
    // a boundary edge, is_flip_ok reads only the edge's own halfedges)
    if (side == 0 && tid < 4)
        w.pv[4 * k + tid] = t.edge_boundary(e) ? t.to(2 * e + (tid & 1)) : patch_vertex(t, e, tid);
    if (side == 0 && tid == 0) w.e[k] = e;
    if (tid == 0) s_ok = !t.edge_boundary(e);
    __syncthreads();
    if (!s_ok) {
        if (side == 0 && tid == 0) w.valid[k] = 0;
        return;
    }
    // side 0 decides flip_ok from its one-rings; side 1 evaluates the
    // flipped patch without waiting for it (FlippedTopo is a valid view of
    // any interior edge; its values are read only for valid proposals)
    if (side == 0) {
        coop_patch(t, nl, p, e, S, true);
        if (!S.flip_ok) {
            if (tid == 0) w.valid[k] = 0;
            return;
        }
        const unsigned long long key = sf_key(st->window, W, k);
        if (tid == 0) {
            for (int c = 0; c < 6; c++) w.dO[6 * k + c] = S.sum[c];
            w.valid[k] = 1;
            atomicMax(&estamp[e], key);
        } else if (tid >= 3 && tid < 5) {
            const int f = t.face(2 * e + tid - 3);
            w.fc[2 * k + tid - 3] = f;
            atomicMax(&fstamp[f], key);
        } else if (tid < 3) {
            // an endpoint whose v_he is one of the edge's halfedges
            const int h = 2 * e + tid - 1;
            const int v = t.to(h);
            if (t.vhe(v) == (h ^ 1)) atomicMax(&vstamp[v], key);
        }
        if (need_nb) {
            const int g = tid / kCoopLanes, l = tid % kCoopLanes;
            const int n = S.n[g];
            if (l == 0) w.nbn[4 * k + g] = n;
            for (int s = l; s < n; s += kCoopLanes) w.nb[(size_t)(4 * k + g) * kMaxValence + s] = S.nb[g][s];
        }
    } else {
        const FlippedTopo ft(t, e);
        coop_patch(ft, nl, p, e, S);
        if (tid == 0)
            for (int c = 0; c < 6; c++) w.dN[6 * k + c] = S.sum[c];
    }
}

// The reads of do_flip for up to kPer edges, then the writes: the edges'
// halfedges are disjoint, so the loads of all of them overlap.
template <int N>
__device__ __forceinline__ void flip_batch(const FlipMeshView& fm, const int (&e)[N], const int (&mode)[N])
{
    // mode: 0 none, 1 do_flip, 2 do_flip twice
    int a1[N], a2[N], b1[N], b2[N], va0[N], vb0[N], va1[N], vb1[N], ta2[N], tb2[N], fa[N], fb[N], ha[N], hb[N];
#pragma unroll
    for (int r = 0; r < N; r++) {
        if (!mode[r]) continue;
        const int a0 = 2 * e[r], b0 = a0 + 1;
        a1[r] = fm.he_next[a0];
        b1[r] = fm.he_next[b0];
        va0[r] = fm.he_to[a0];
        vb0[r] = fm.he_to[b0];
        fa[r] = fm.he_face[a0];
        fb[r] = fm.he_face[b0];
    }
#pragma unroll
    for (int r = 0; r < N; r++) {
        if (!mode[r]) continue;
        a2[r] = fm.he_next[a1[r]];
        b2[r] = fm.he_next[b1[r]];
        va1[r] = fm.he_to[a1[r]];
        vb1[r] = fm.he_to[b1[r]];
        ha[r] = fm.v_he[va0[r]];
        hb[r] = fm.v_he[vb0[r]];
    }
#pragma unroll
    for (int r = 0; r < N; r++) {
        if (!mode[r]) continue;
        ta2[r] = fm.he_to[a2[r]];
        tb2[r] = fm.he_to[b2[r]];
    }
#pragma unroll
    for (int r = 0; r < N; r++) {
        if (!mode[r]) continue;
        const int a0 = 2 * e[r], b0 = a0 + 1;
        if (mode[r] == 1) {
            // do_flip
            fm.he_to[a0] = va1[r];
            fm.he_to[b0] = vb1[r];
            fm.he_next[a0] = a2[r]; fm.he_prev[a2[r]] = a0;
            fm.he_next[a2[r]] = b1[r]; fm.he_prev[b1[r]] = a2[r];
            fm.he_next[b1[r]] = a0; fm.he_prev[a0] = b1[r];
            fm.he_next[b0] = b2[r]; fm.he_prev[b2[r]] = b0;
            fm.he_next[b2[r]] = a1[r]; fm.he_prev[a1[r]] = b2[r];
            fm.he_next[a1[r]] = b0; fm.he_prev[b0] = a1[r];
            fm.he_face[a1[r]] = fb[r];
            fm.he_face[b1[r]] = fa[r];
        } else {
            // do_flip_twice
            fm.he_to[a0] = ta2[r];
            fm.he_to[b0] = tb2[r];
            fm.he_next[a0] = b1[r]; fm.he_prev[b1[r]] = a0;
            fm.he_next[b1[r]] = b2[r]; fm.he_prev[b2[r]] = b1[r];
            fm.he_next[b2[r]] = a0; fm.he_prev[a0] = b2[r];
            fm.he_next[b0] = a1[r]; fm.he_prev[a1[r]] = b0;
            fm.he_next[a1[r]] = a2[r]; fm.he_prev[a2[r]] = a1[r];
            fm.he_next[a2[r]] = b0; fm.he_prev[b0] = a2[r];
            fm.he_face[a1[r]] = fb[r];
            fm.he_face[a2[r]] = fb[r];
            fm.he_face[b1[r]] = fa[r];
            fm.he_face[b2[r]] = fa[r];
        }
        fm.f_he[fa[r]] = a0;
        fm.f_he[fb[r]] = b0;
        if (ha[r] == b0) fm.v_he[va0[r]] = a1[r];
        if (hb[r] == a0) fm.v_he[vb0[r]] = b1[r];
    }
}

__global__ void __launch_bounds__(64) k_sf_commit(FlipMeshView fm, EParams p, int nflips,
                                                  const double* __restrict__ uniforms, int n_uniforms,
                                                  uint64_t seed, int injected, int W, int need_nb,
                                                  SerialState* __restrict__ st, SerialWindow w,
                                                  const unsigned long long* __restrict__ vstamp,
                                                  const unsigned long long* __restrict__ estamp,
                                                  const unsigned long long* __restrict__ fstamp)
{
    constexpr int kPer = kSerialWindowMax / 64;
    constexpr int kChunk = 64;                  // valid proposals per decision round
    __shared__ int s_valid[kSerialWindowMax];
    __shared__ int s_stale[kSerialWindowMax];
    __shared__ int s_defer[kSerialWindowMax];   // an earlier valid proposal has one of its faces
    __shared__ int s_accf[kSerialWindowMax];    // accepted
    __shared__ int s_vk[kSerialWindowMax];      // window position of the m-th valid proposal
    __shared__ int s_rank[kSerialWindowMax + 1];    // valid proposals before position k
    __shared__ int s_pv[4 * kSerialWindowMax];
    __shared__ int s_nbn[4 * kSerialWindowMax];
    __shared__ int s_ring[4 * kMaxValence];     // one-rings of an accepted proposal's patch
    __shared__ int s_dlist[kSerialWindowMax];   // positions of the deferred flips, ascending
    __shared__ unsigned s_bits[128];            // hash bitmap of s_ring and the accepted patch
    // per proposal: the bitmap positions of its patch vertices (12 bits each)
    // and, in bit 63, whether one of its one-rings was too long to record
    __shared__ unsigned long long s_ph[kSerialWindowMax];
    // by rank among the valid proposals: uniforms, the rounding path's
    // states and the patch properties before and after the flip
    __shared__ double s_u[kSerialWindowMax];
    __shared__ double s_q[6 * kSerialWindowMax];
    __shared__ double s_o[6 * kSerialWindowMax];
    __shared__ double s_n[6 * kSerialWindowMax];
    __shared__ double s_P[6], s_Prej[6];
    __shared__ double s_e0;
    __shared__ int s_acc, s_err, s_end, s_first, s_ovf, s_ndef;
    const int tid = threadIdx.x;
    if (st->done) return;
    const int pos0 = st->pos;
    const int cnt = min(W, nflips - pos0);
    const unsigned wid = (unsigned)st->window + 1u;
    const int used0 = st->used;

    // ---- stage; position k = 64 r + tid
    int v[kPer], e[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int k = 64 * r + tid;
        v[r] = k < cnt ? w.valid[k] : -2;
        e[r] = v[r] >= 0 ? w.e[k] : 0;
    }
    int pv[kPer][4], fc[kPer][2];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int k = 64 * r + tid;
        for (int g = 0; g < 4; g++) pv[r][g] = v[r] >= 0 ? w.pv[4 * k + g] : 0;
        for (int g = 0; g < 2; g++) fc[r][g] = v[r] == 1 ? w.fc[2 * k + g] : 0;
    }
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int k = 64 * r + tid;
        int stale = 0, defer = 0;
        if (v[r] >= 0) {
            for (int g = 0; g < 4; g++) stale |= sf_pos(vstamp[pv[r][g]], wid, W) < k;
            stale |= sf_pos(estamp[e[r]], wid, W) < k;
            for (int g = 0; g < 4; g++) s_pv[4 * k + g] = pv[r][g];
        }
        unsigned long long ph = 0ull;
        if (v[r] >= 0)
            for (int g = 0; g < 4; g++) ph |= (unsigned long long)sf_hash(pv[r][g]) << (12 * g);
        if (v[r] == 1) {
            defer = sf_pos(fstamp[fc[r][0]], wid, W) < k || sf_pos(fstamp[fc[r][1]], wid, W) < k;
            if (need_nb)
                for (int g = 0; g < 4; g++) {
                    const int nn = w.nbn[4 * k + g];
                    s_nbn[4 * k + g] = nn;
                    if (nn < 0) ph |= 1ull << 63;
                }
        }
        s_ph[k] = ph;
        s_valid[k] = v[r];
        s_stale[k] = stale;
        s_defer[k] = defer;
        s_accf[k] = 0;
    }
    // the patch properties of the valid proposals, all rows in flight at once
    // (one workgroup: registers are not the limit)
    double po[kPer][6], pn[kPer][6];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int k = 64 * r + tid;
#pragma unroll
        for (int c = 0; c < 6; c++) {
            po[r][c] = v[r] == 1 ? w.dO[6 * k + c] : 0.0;
            pn[r][c] = v[r] == 1 ? w.dN[6 * k + c] : 0.0;
        }
    }
    // ranks among the valid proposals (their uniform index is used0 + rank)
    // and the list of deferred flips
    int base = 0, dbase = 0;
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const int k = 64 * r + tid;
        const unsigned long long below = (1ull << tid) - 1ull;
        const unsigned long long bd = __ballot(s_defer[k] != 0);
        if (s_defer[k]) s_dlist[dbase + __popcll(bd & below)] = k;
        dbase += __popcll(bd);
        const unsigned long long b = __ballot(v[r] == 1);
        const int m = base + __popcll(b & ((1ull << tid) - 1ull));
        s_rank[k] = m;
        if (v[r] == 1) {
            const int u = used0 + m;
            s_vk[m] = k;
            s_u[m] = injected ? (u < n_uniforms ? uniforms[u] : 2.0) : rng_uniform(seed, (uint64_t)u);
            for (int c = 0; c < 6; c++) {
                s_o[6 * m + c] = po[r][c];
                s_n[6 * m + c] = pn[r][c];
            }
        }
        base += __popcll(b);
    }
    if (tid == 0) {
        s_rank[kSerialWindowMax] = base;
        s_ndef = dbase;
        for (int c = 0; c < 6; c++) s_P[c] = st->props[c];
        s_e0 = st->e0;
        s_acc = st->acc;
        s_err = 0;
        // an injected stream too short: the proposal that finds it empty is
        // rejected (u = 2) and the sweep stops after it
        s_ovf = (injected && base > n_uniforms - used0) ? max(n_uniforms - used0, 0) : kSerialWindowMax;
    }
    __syncthreads();
    const int movf = s_ovf;    // rank of that proposal
    const int ovf = movf < base ? s_vk[movf] : cnt;

    // ---- decide, chunk by chunk, up to the first stale proposal
    int pos = 0;
    while (pos < cnt) {
        if (tid == 0) s_end = cnt;
        __syncthreads();
        for (int k = pos + tid; k < cnt; k += 64)
            if (s_stale[k] || s_valid[k] < 0) atomicMin(&s_end, k);
        __syncthreads();
        int end = s_end;
        if (end == pos) break;
        if (ovf < end) end = ovf + 1;
        const int m0 = s_rank[pos], mend = s_rank[end];
        const int m1 = min(mend, m0 + kChunk);
        const int cend = m1 < mend ? s_vk[m1] : end;    // positions [pos, cend) hold valid m0 .. m1-1
        if (tid < 6) {
            // the serial rounding path through the rejected branch, one lane
            // per component
            const int c = tid;
            double P = s_P[c];
#pragma unroll 4
            for (int m = m0; m < m1; m++) {
                const double o = s_o[6 * m + c], nw = s_n[6 * m + c];
                double q = P - o;
                q += nw;
                s_q[6 * m + c] = q;
                P = q - nw;
                P += o;
            }
            s_Prej[c] = P;
        }
        if (tid == 0) s_first = cend;
        __syncthreads();
        const double e0 = s_e0;
        for (int m = m0 + tid; m < m1; m += 64) {
            const int k = s_vk[m];
            if (k == ovf) continue;
            const double en = energy_of(s_q + 6 * m, p);
            const double de = en - e0;
            const double alpha = de < 0.0 ? 1.0 : exp(-de);
            if (s_u[m] <= alpha) atomicMin(&s_first, k);
        }
        __syncthreads();
        const int kf = s_first;
        if (kf < cend) {
            if (tid == 0) {
                const int mf = s_rank[kf];
                for (int c = 0; c < 6; c++) s_P[c] = s_q[6 * mf + c];
                s_e0 = energy_of(s_q + 6 * mf, p);
                s_acc += 1;
                s_accf[kf] = 1;
            }
            // later proposals whose patch meets the accepted patch are stale;
            // for exclusion level 2 also those with a patch vertex adjacent
            // to it (adjacency is symmetric: the accepted patch meets their
            // patch's one-rings), found in the accepted patch's one-rings
            const int f0 = s_pv[4 * kf], f1 = s_pv[4 * kf + 1], f2 = s_pv[4 * kf + 2], f3 = s_pv[4 * kf + 3];
            auto in_v = [&](int u) { return u == f0 || u == f1 || u == f2 || u == f3; };
            int rn[4] = {0, 0, 0, 0};
            bool all = false;
            for (int t = tid; t < 128; t += 64) s_bits[t] = 0u;
            if (need_nb) {
                for (int g = 0; g < 4; g++) {
                    rn[g] = s_nbn[4 * kf + g];
                    all = all || rn[g] < 0;
                }
                for (int t = tid; t < 4 * kMaxValence; t += 64) {
                    const int g = t / kMaxValence, s = t % kMaxValence;
                    if (s < rn[g]) s_ring[t] = w.nb[(size_t)(4 * kf + g) * kMaxValence + s];
                }
            }
            __syncthreads();
            if (tid < 4) {
                const unsigned h = sf_hash(s_pv[4 * kf + tid]);
                atomicOr(&s_bits[h >> 5], 1u << (h & 31));
            }
            for (int t = tid; t < 4 * kMaxValence; t += 64) {
                const int g = t / kMaxValence, s = t % kMaxValence;
                if (s < rn[g]) {
                    const unsigned h = sf_hash(s_ring[t]);
                    atomicOr(&s_bits[h >> 5], 1u << (h & 31));
                }
            }
            __syncthreads();
            auto maybe = [&](unsigned long long ph, int g) {
                const unsigned h = (unsigned)(ph >> (12 * g)) & 0xfffu;
                return (s_bits[h >> 5] >> (h & 31)) & 1u;
            };
            for (int k = kf + 1 + tid; k < cnt; k += 64) {
                if (s_valid[k] < 0) continue;
                const unsigned long long ph = s_ph[k];
                const bool own = need_nb && s_valid[k] == 1 && (all || (ph >> 63));
                if (!own && !(maybe(ph, 0) | maybe(ph, 1) | maybe(ph, 2) | maybe(ph, 3))) continue;
                const int q0 = s_pv[4 * k], q1 = s_pv[4 * k + 1], q2 = s_pv[4 * k + 2], q3 = s_pv[4 * k + 3];
                bool hit = own || in_v(q0) || in_v(q1) || in_v(q2) || in_v(q3);
                if (!hit && need_nb && s_valid[k] == 1) {
                    for (int g = 0; g < 4 && !hit; g++)
                        for (int s = 0; s < rn[g] && !hit; s++) {
                            const int u = s_ring[g * kMaxValence + s];
                            hit = u == q0 || u == q1 || u == q2 || u == q3;
                        }
                }
                if (hit) s_stale[k] = 1;
            }
            pos = kf + 1;
        } else {
            if (tid == 0) {
                for (int c = 0; c < 6; c++) s_P[c] = s_Prej[c];
                if (ovf < cend) s_err = 2;
            }
            pos = cend;
        }
        __syncthreads();
        if (s_err) break;
    }

    // ---- the committed flips: proposals whose faces no earlier one touched
    // are disjoint from every earlier flip; the others follow, in order
    {
        int fe[kPer], mode[kPer];
#pragma unroll
        for (int r = 0; r < kPer; r++) {
            const int k = 64 * r + tid;
            fe[r] = e[r];
            mode[r] = (k < pos && s_valid[k] == 1 && !s_defer[k]) ? (s_accf[k] ? 1 : 2) : 0;
        }
        flip_batch(fm, fe, mode);
    }
    __syncthreads();
    if (tid == 0)
        for (int d = 0; d < s_ndef; d++) {
            const int k = s_dlist[d];
            if (k >= pos) break;
            if (s_accf[k]) do_flip(fm, w.e[k]);
            else do_flip_twice(fm, w.e[k]);
        }
    if (tid == 0) {
        int err = s_err;
        if (!err && pos < cnt && s_valid[pos] == -1) err = 1;
        for (int c = 0; c < 6; c++) st->props[c] = s_P[c];
        st->e0 = s_e0;
        st->used = used0 + s_rank[pos];
        st->acc = s_acc;
        st->err = err;
        st->pos = pos0 + pos;
        st->window += 1;
        if (err || pos0 + pos >= nflips) st->done = 1;
    }
}

void launch_serial_init(const Globals* g0, const EParams& p, int nflips, SerialState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_sf_init, dim3(1), dim3(64), 0, s, g0, p, nflips, st);
}

void launch_serial_windows(const FlipMeshView& fm, const NListView& nl, const EParams& p, int nflips,
                           const int* proposals, const double* uniforms, int n_uniforms, uint64_t seed, int injected,
                           int W, SerialState* st, const SerialWindow& w, unsigned long long* vstamp,
                           unsigned long long* estamp, unsigned long long* fstamp, int n_windows, hipStream_t s)
{
    const int need_nb = nl.kind == 0 && nl.exclusion > 1 && p.rep_lc1 > 0;
    for (int r = 0; r < n_windows; r++) {
        hipLaunchKernelGGL(k_sf_eval, dim3(W, 2), dim3(64), 0, s, fm, nl, p, nflips, proposals, seed, injected, W,
                           need_nb, st, w, vstamp, estamp, fstamp);
        hipLaunchKernelGGL(k_sf_commit, dim3(1), dim3(64), 0, s, fm, p, nflips, uniforms, n_uniforms, seed, injected,
                           W, need_nb, st, w, vstamp, estamp, fstamp);
    }
}

// ============================================ conflict-free parallel sweep ==
// pflip (flip_parallel_batches, flips.cpp:82-166) lets OpenMP threads lock
// the edge patch of their proposal (flip_patch, flip_utils.cpp:43-68), skip
// proposals whose patch is taken, flip and evaluate in parallel and apply the
// Metropolis test one flip at a time.  Here a round of B proposals is one
// batch:
//   claim   every proposal writes its priority key into all edges of its
//           patch with a 64-bit atomicMax (later rounds have larger keys, so
//           the owner array never needs clearing);
//   eval    a proposal whose key survived on its whole patch is a winner;
//           winners have pairwise disjoint patches.  Winners that pass
//           is_flip_ok evaluate the patch properties before the flip and,
//           on FlippedTopo, after it: delta = new - old.  The evaluation
//           after the flip sees the proposal's own flip only (the
//           reference's threads see whichever other flips have happened);
//   scan    the FP64 Metropolis test in proposal order on the running
//           global properties;
//   apply   accepted winners flip.
// Proposals that lose the claim are skipped, like the reference's failed
// locks.  Every phase is its own kernel, so no phase reads topology that
// another thread is rewriting; the result depends only on the seed.

// The flip patch of edge e (flip_utils.cpp:43-68) is the set of edges of h
// and of next(h) over the outgoing halfedges h of its four vertices (the two
// endpoints and the two opposite vertices, patch_vertex(t, e, 0..3)); the
// claim phase and coop_patch's ownership test visit exactly these.

__device__ __forceinline__ int pflip_edge(uint64_t seed, int round, int B, int i, int ne)
{
    return rng_edge(seed ^ 0x5bd1e9955bd1e995ull, (uint64_t)round * (uint64_t)B + (uint64_t)i, ne);
}

__device__ __forceinline__ unsigned long long pflip_key(int round, int B, int i)
{
    return ((unsigned long long)(round + 1) << 32) | (unsigned)(B - i);
}

// Eight lanes per proposal, two per patch vertex, walking the vertex's
// outgoing halfedges: lane 0 of the pair clockwise from v_he, lane 1
// counter-clockwise (the inverse of cw, prev(h) ^ 1), each claiming its
// halfedges' edges and next edges until they meet, so the chain of
// dependent loads is half as long.  The patch's edges get the same
// atomicMax as in a one-lane walk, and the maximum does not depend on the
// order of the atomics.
__device__ __forceinline__ void pflip_claim_body(const FlipMeshView& fm, uint64_t seed, int round, int B, int nprop,
                                                 unsigned long long* __restrict__ owner, int lb)
{
    const int i = lb * kClaimPerBlock + threadIdx.x / 8;
    if (i >= nprop) return;
    const Topo t{fm.x, fm.he_to, fm.he_next, fm.he_prev, fm.he_face, fm.v_he, fm.f_he};
    const int e = pflip_edge(seed, round, B, i, fm.nh / 2);
    if (t.edge_boundary(e)) return;    // both lanes of every pair leave together
    const unsigned long long key = pflip_key(round, B, i);
    const int w = (threadIdx.x / 2) & 3, back = threadIdx.x & 1;
    const int h0 = t.vhe(patch_vertex(t, e, w));
    if (h0 < 0) return;
    // a: the clockwise walker's next halfedge, b: the counter-clockwise
    // walker's; both lanes hold both, so they leave the loop together
    int a = h0, b = t.prev(h0) ^ 1, pb = -1;
#pragma unroll 1
    for (int it = 0; it < 2 * kMaxValence && a != pb; it++) {
        const int h = back ? b : a;
        if (!back || a != b) {
            atomicMax(&owner[h >> 1], key);
            atomicMax(&owner[t.next(h) >> 1], key);
        }
        if (a == b) break;
        const int nxt = back ? (t.prev(b) ^ 1) : t.cw(a);
        const int oth = __shfl_xor(nxt, 1, 64);
        pb = b;
        a = back ? oth : nxt;
        b = back ? nxt : oth;
    }
}

__global__ void __launch_bounds__(64) k_pflip_claim(FlipMeshView fm, uint64_t seed, int round, int B, int nprop,
                                                    unsigned long long* __restrict__ owner)
{
    pflip_claim_body(fm, seed, round, B, nprop, owner, blockIdx.x);
}

// status: 0 skipped, 1 winner pending, 2 accepted, 3 rejected.
// One workgroup per proposal and side.  Side 0 evaluates the patch before
// the flip (dold = -old) and decides: the proposal's key survived on every
// patch edge and is_flip_ok holds, both read off the one-rings coop_patch
// walks anyway.  Side 1 evaluates the patch after the flip on FlippedTopo
// (dnew = new) without waiting for the decision: FlippedTopo is a valid view
// of any interior edge, every walk on it is bounded, and dnew is read only
// for the proposals side 0 marks pending (a loser's evaluation is wasted,
// a few percent of them at E / 1024 proposals per round).
__device__ __forceinline__ void pflip_eval_body(const FlipMeshView& fm, const NListView& nl, const EParams& p,
                                                uint64_t seed, int round, int B, int nprop,
                                                const unsigned long long* __restrict__ owner,
                                                int* __restrict__ status, double* __restrict__ dold,
                                                double* __restrict__ dnew, int i)
{
    __shared__ CoopShared S;
    const int side = blockIdx.y;
    const Topo t{fm.x, fm.he_to, fm.he_next, fm.he_prev, fm.he_face, fm.v_he, fm.f_he};
    const int e = pflip_edge(seed, round, B, i, fm.nh / 2);
    if (threadIdx.x == 0 && side == 0) status[i] = 0;
    if (t.edge_boundary(e)) return;
    if (side == 0) {
        coop_patch(t, nl, p, e, S, true, owner, pflip_key(round, B, i));
        if (!S.flip_ok || (S.lost[0] | S.lost[1] | S.lost[2] | S.lost[3])) return;
        if (threadIdx.x == 0) {
            for (int k = 0; k < 6; k++) dold[6 * i + k] = -S.sum[k];
            status[i] = 1;
        }
    } else {
        const FlippedTopo ft(t, e);
        coop_patch(ft, nl, p, e, S);
        if (threadIdx.x == 0)
            for (int k = 0; k < 6; k++) dnew[6 * i + k] = S.sum[k];
    }
}

__global__ void __launch_bounds__(64) TM_COOP_OCCUPANCY
k_pflip_eval(FlipMeshView fm, NListView nl, EParams p, uint64_t seed, int round, int B, int nprop,
             const unsigned long long* __restrict__ owner, int* __restrict__ status, double* __restrict__ dold,
             double* __restrict__ dnew)
{
    pflip_eval_body(fm, nl, p, seed, round, B, nprop, owner, status, dold, dnew, blockIdx.x);
}

__global__ void k_pflip_apply(FlipMeshView fm, uint64_t seed, int round, int B, int nprop,
                              const int* __restrict__ status, int want)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nprop || status[i] != want) return;
    do_flip(fm, pflip_edge(seed, round, B, i, fm.nh / 2));
}

// state: props[6], e0, accepted count (as double).  The Metropolis test
// walks the winners in proposal order on the running properties
// (flips.cpp:146-163), with a rejected flip restoring the running
// properties exactly.  Most winners are rejected and a rejection leaves the
// state unchanged, so the workgroup tests the next `window` proposals
// against the current state at once, takes the first acceptance in
// proposal order (an LDS atomicMin), applies it and continues after it:
// every decision is the one the sequential walk makes, and the walk takes
// one parallel step per acceptance instead of one dependent FP64 chain per
// winner.  Statuses, deltas and uniforms are staged in LDS per chunk.
constexpr int kScanThreads = 256;
constexpr int kScanChunk = 1024;

__device__ __forceinline__ void pflip_scan_body(const EParams& p, uint64_t seed, int round, int B, int nprop,
                                                int window, int* __restrict__ status, const double* __restrict__ dold,
                                                const double* __restrict__ dnew, double* __restrict__ state)
{
    __shared__ int s_st[kScanChunk];
    __shared__ double s_dp[6 * kScanChunk];
    __shared__ double s_u[kScanChunk];
    __shared__ double s_state[8];
    __shared__ int s_first[2];
    const int tid = threadIdx.x;
    if (tid < 8) s_state[tid] = state[tid];
    for (int base = 0; base < nprop; base += kScanChunk) {
        const int cnt = min(kScanChunk, nprop - base);
        for (int k = tid; k < cnt; k += kScanThreads) {
            const int st = status[base + k];
            s_st[k] = st;
            if (st == 1)
                s_u[k] = rng_uniform(seed ^ 0x5bd1e9955bd1e995ull,
                                     (uint64_t)round * (uint64_t)B + (uint64_t)(base + k));
        }
        for (int k = tid; k < 6 * cnt; k += kScanThreads) {
            double d = dold[6 * (size_t)base + k];
            d += dnew[6 * (size_t)base + k];
            s_dp[k] = d;
        }
        if (tid == 0) s_first[0] = s_first[1] = cnt;
        __syncthreads();
        // s_first[it & 1] collects iteration it's first acceptance; thread 0
        // resets the other slot, whose readers finished before the last
        // barrier and whose writers start after the next one
        int it = 0;
        for (int pos = 0; pos < cnt; it++) {
            const int k = pos + tid;
            double q[6], en = 0.0;
            if (tid < window && k < cnt && s_st[k] == 1) {
                for (int c = 0; c < 6; c++) q[c] = s_state[c] + s_dp[6 * k + c];
                en = energy_of(q, p);
                const double de = en - s_state[6];
                const double alpha = de < 0.0 ? 1.0 : exp(-de);
                if (s_u[k] <= alpha) atomicMin(&s_first[it & 1], k);
            }
            if (tid == 0) s_first[(it + 1) & 1] = cnt;
            __syncthreads();
            const int kf = s_first[it & 1];
            if (kf < cnt && k == kf) {
                for (int c = 0; c < 6; c++) s_state[c] = q[c];
                s_state[6] = en;
                s_state[7] += 1.0;
                s_st[k] = 2;
            }
            pos = kf < cnt ? kf + 1 : pos + window;
            __syncthreads();
        }
        for (int k = tid; k < cnt; k += kScanThreads)
            if (s_st[k] >= 1) status[base + k] = s_st[k] == 2 ? 2 : 3;
        __syncthreads();
    }
    if (tid < 8) state[tid] = s_state[tid];
}

__global__ void __launch_bounds__(kScanThreads) k_pflip_scan(EParams p, uint64_t seed, int round, int B, int nprop,
                                                             int window, int* __restrict__ status,
                                                             const double* __restrict__ dold,
                                                             const double* __restrict__ dnew,
                                                             double* __restrict__ state)
{
    pflip_scan_body(p, seed, round, B, nprop, window, status, dold, dnew, state);
}

__device__ __forceinline__ void pflip_init_body(const Globals* __restrict__ g0, const EParams& p,
                                                double* __restrict__ state)
{
    double props[6];
    for (int k = 0; k < 6; k++) props[k] = state[k] = g0->props[k];
    state[6] = energy_of(props, p);
    state[7] = 0.0;
}

__global__ void k_pflip_init(const Globals* __restrict__ g0, EParams p, double* __restrict__ state)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    pflip_init_body(g0, p, state);
}

// ---- many meshes: round r of every replica per launch ----
// proposals of replica a in round r (<= 0 once its sweep is done)
__device__ __forceinline__ int pflip_nprop(const PflipArgs& a, int round)
{
    return min(a.batch, a.nflips - round * a.batch);
}

__global__ void k_pflip_init_grouped(const PflipArgs* __restrict__ A)
{
    if (threadIdx.x == 0) pflip_init_body(A[blockIdx.x].g0, A[blockIdx.x].p, A[blockIdx.x].state);
}

__global__ void __launch_bounds__(64) k_pflip_claim_grouped(const PflipArgs* __restrict__ A,
                                                            const int* __restrict__ block0, int R, int round)
{
    const int r = find_replica(block0, R, blockIdx.x);
    const PflipArgs& a = A[r];
    const int nprop = pflip_nprop(a, round);
    if (nprop <= 0) return;
    pflip_claim_body(a.fm, a.seed, round, a.batch, nprop, a.owner, blockIdx.x - block0[r]);
}

__global__ void __launch_bounds__(64) TM_COOP_OCCUPANCY
k_pflip_eval_grouped(const PflipArgs* __restrict__ A, const int* __restrict__ block0, int R, int round)
{
    const int r = find_replica(block0, R, blockIdx.x);
    const PflipArgs& a = A[r];
    const int nprop = pflip_nprop(a, round), i = blockIdx.x - block0[r];
    if (i >= nprop) return;
    pflip_eval_body(a.fm, a.nl, a.p, a.seed, round, a.batch, nprop, a.owner, a.status, a.dold, a.dnew, i);
}

__global__ void __launch_bounds__(kScanThreads) k_pflip_scan_grouped(const PflipArgs* __restrict__ A, int round,
                                                                     int window)
{
    const PflipArgs& a = A[blockIdx.x];
    const int nprop = pflip_nprop(a, round);
    if (nprop <= 0) return;
    pflip_scan_body(a.p, a.seed, round, a.batch, nprop, window, a.status, a.dold, a.dnew, a.state);
}

__global__ void k_pflip_apply_grouped(const PflipArgs* __restrict__ A, const int* __restrict__ block0, int R,
                                      int round)
{
    const int r = find_replica(block0, R, blockIdx.x);
    const PflipArgs& a = A[r];
    const int i = (blockIdx.x - block0[r]) * blockDim.x + threadIdx.x;
    if (i >= pflip_nprop(a, round) || a.status[i] != 2) return;
    do_flip(a.fm, pflip_edge(a.seed, round, a.batch, i, a.fm.nh / 2));
}

void launch_flip_parallel(const FlipMeshView& fm, const NListView& nl, const EParams& p, int nflips, int batch,
                          int window, uint64_t seed, const Globals* g0, unsigned long long* owner, int* status,
                          double* dprops, double* state, hipStream_t s)
{
    hipLaunchKernelGGL(k_pflip_init, dim3(1), dim3(64), 0, s, g0, p, state);
    const int rounds = (nflips + batch - 1) / batch;
    for (int r = 0; r < rounds; r++) {
        int nprop = std::min(batch, nflips - r * batch);
        dim3 grid((nprop + 127) / 128), block(128);
        const dim3 claim_grid((nprop + kClaimPerBlock - 1) / kClaimPerBlock);
        hipLaunchKernelGGL(k_pflip_claim, claim_grid, dim3(64), 0, s, fm, seed, r, batch, nprop, owner);
        hipLaunchKernelGGL(k_pflip_eval, dim3(nprop, 2), dim3(64), 0, s, fm, nl, p, seed, r, batch, nprop, owner,
                           status, dprops, dprops + 6 * (size_t)batch);
        hipLaunchKernelGGL(k_pflip_scan, dim3(1), dim3(kScanThreads), 0, s, p, seed, r, batch, nprop,
                           std::min(std::max(window, 1), kScanThreads), status, dprops, dprops + 6 * (size_t)batch,
                           state);
        hipLaunchKernelGGL(k_pflip_apply, grid, block, 0, s, fm, seed, r, batch, nprop, status, 2);
    }
}

void launch_flip_parallel_grouped(const PflipGroup& g, hipStream_t s)
{
    if (g.R <= 0) return;
    hipLaunchKernelGGL(k_pflip_init_grouped, dim3(g.R), dim3(64), 0, s, g.d_args);
    const int window = std::min(std::max(g.window, 1), kScanThreads);
    for (int r = 0; r < g.rounds; r++) {
        hipLaunchKernelGGL(k_pflip_claim_grouped, dim3(g.total_claim), dim3(64), 0, s, g.d_args, g.d_claim0, g.R, r);
        hipLaunchKernelGGL(k_pflip_eval_grouped, dim3(g.total_eval, 2), dim3(64), 0, s, g.d_args, g.d_eval0, g.R, r);
        hipLaunchKernelGGL(k_pflip_scan_grouped, dim3(g.R), dim3(kScanThreads), 0, s, g.d_args, r, window);
        hipLaunchKernelGGL(k_pflip_apply_grouped, dim3(g.total_apply), dim3(128), 0, s, g.d_args, g.d_apply0, g.R, r);
    }
}

}    // namespace tmh
