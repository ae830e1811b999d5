// tm_flips.hip -- ring-table rebuild and the serial edge-flip sweep.
//
// flip_serial (src/flips.cpp:22-80) is a Metropolis loop in which every
// proposal depends on the topology left by the previous one, so it runs as
// one workgroup walking the proposal stream in order.  For each proposal:
//   * lane 0 tests OpenMesh is_flip_ok and the boundary condition;
//   * lanes 0-3 evaluate vertex_properties of the four patch vertices
//     (flip_utils.cpp:19-41) in parallel with the OpenMesh-semantics walker,
//     including their repulsion queries;
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
__device__ __forceinline__ bool flip_ok(const Topo& t, int e)
{
    // OpenMesh TriConnectivity::is_flip_ok
    if (t.edge_boundary(e)) return false;
    int ah = t.to(t.next(2 * e));
    int bh = t.to(t.next(2 * e + 1));
    if (ah == bh) return false;
    int h0 = t.v_he[ah];
    int h = h0;
    int guard = 0;
    do {
        if (t.to(h) == bh) return false;
        h = t.cw(h);
    } while (h != h0 && ++guard < 4 * kMaxValence);
    return true;
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

// Patch properties: lanes 0-3 each evaluate one of the four vertices
// to(2e), to(next(2e)), to(2e+1), to(next(2e+1)); the sum is accumulated in
// that order in s_sum by thread 0.
__device__ __forceinline__ void patch_properties(const FlipMeshView& fm, const NListView& nl, const EParams& p,
                                                 int e, double (*s_vp)[6], double* s_sum)
{
    const Topo t{fm.x, fm.he_to, fm.he_next, fm.he_prev, fm.he_face, fm.v_he, fm.f_he};
    int lane = threadIdx.x;
    if (lane < 4) {
        int h = 2 * e + (lane >> 1);
        int ve = (lane & 1) ? t.to(t.next(h)) : t.to(h);
        VP r = om_vertex_properties(t, nl, p, ve);
        for (int k = 0; k < 6; k++) s_vp[lane][k] = r.v[k];
    }
    __syncthreads();
    if (lane == 0) {
        for (int k = 0; k < 6; k++) s_sum[k] = 0.0;
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 6; k++) s_sum[k] += s_vp[i][k];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(64) k_flip_serial(FlipMeshView fm, NListView nl, EParams p, int nflips,
                                                    const int* __restrict__ proposals,
                                                    const double* __restrict__ uniforms, int n_uniforms,
                                                    uint64_t seed, int injected, const Globals* __restrict__ g0,
                                                    int* __restrict__ result)
{
    __shared__ double s_vp[4][6];
    __shared__ double s_old[6];
    __shared__ double s_new[6];
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
        patch_properties(fm, nl, p, e, s_vp, s_old);
        if (threadIdx.x == 0) {
            for (int k = 0; k < 6; k++) props[k] -= s_old[k];
            do_flip(fm, e);
        }
        __threadfence_block();
        __syncthreads();
        patch_properties(fm, nl, p, e, s_vp, s_new);
        if (threadIdx.x == 0) {
            for (int k = 0; k < 6; k++) props[k] += s_new[k];
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
                    props[k] -= s_new[k];
                    props[k] += s_old[k];
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
//           is_flip_ok evaluate the patch properties before the flip;
//   apply   winners flip;
//   eval2   winners evaluate the patch after the flip: delta = new - old;
//   scan    one lane walks the winners in proposal order and applies the
//           FP64 Metropolis test to the running global properties;
//   revert  rejected winners flip back.
// Proposals that lose the claim are skipped, like the reference's failed
// locks.  Every phase is its own kernel, so no phase reads topology that
// another thread is rewriting; the result depends only on the seed.

// Edges of the flip patch of edge e (flip_utils.cpp:43-68): for the two
// endpoints' and the two opposite vertices' outgoing halfedges h, the edges
// of h and of next(h).
template <class F>
__device__ __forceinline__ void for_patch_edges(const Topo& t, int e, F&& f)
{
    for (int i = 0; i < 2; i++) {
        int heh = 2 * e + i;
        for (int w = 0; w < 2; w++) {
            int vh = t.to(heh);
            int h0 = t.v_he[vh];
            int h = h0;
            int guard = 0;
            do {
                f(h >> 1);
                f(t.next(h) >> 1);
                h = t.cw(h);
            } while (h != h0 && ++guard < 4 * kMaxValence);
            heh = t.next(heh);
        }
    }
}

__device__ __forceinline__ int pflip_edge(uint64_t seed, int round, int B, int i, int ne)
{
    return rng_edge(seed ^ 0x5bd1e9955bd1e995ull, (uint64_t)round * (uint64_t)B + (uint64_t)i, ne);
}

__global__ void k_pflip_claim(FlipMeshView fm, uint64_t seed, int round, int B, int nprop,
                              unsigned long long* __restrict__ owner)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nprop) return;
    const Topo t{fm.x, fm.he_to, fm.he_next, fm.he_prev, fm.he_face, fm.v_he, fm.f_he};
    int e = pflip_edge(seed, round, B, i, fm.nh / 2);
    if (t.edge_boundary(e)) return;
    unsigned long long key = ((unsigned long long)(round + 1) << 32) | (unsigned)(B - i);
    for_patch_edges(t, e, [&](int pe) { atomicMax(&owner[pe], key); });
}

// status: 0 skipped, 1 winner pending, 2 accepted, 3 rejected
__global__ void k_pflip_eval(FlipMeshView fm, NListView nl, EParams p, uint64_t seed, int round, int B, int nprop,
                             const unsigned long long* __restrict__ owner, int* __restrict__ status,
                             double* __restrict__ dprops)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nprop) return;
    const Topo t{fm.x, fm.he_to, fm.he_next, fm.he_prev, fm.he_face, fm.v_he, fm.f_he};
    int e = pflip_edge(seed, round, B, i, fm.nh / 2);
    status[i] = 0;
    if (t.edge_boundary(e)) return;
    unsigned long long key = ((unsigned long long)(round + 1) << 32) | (unsigned)(B - i);
    bool win = true;
    for_patch_edges(t, e, [&](int pe) { win = win && owner[pe] == key; });
    if (!win || !flip_ok(t, e)) return;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int s = 0; s < 4; s++) {
        int h = 2 * e + (s >> 1);
        int ve = (s & 1) ? t.to(t.next(h)) : t.to(h);
        VP r = om_vertex_properties(t, nl, p, ve);
        for (int k = 0; k < 6; k++) acc[k] += r.v[k];
    }
    for (int k = 0; k < 6; k++) dprops[6 * i + k] = -acc[k];
    status[i] = 1;
}

__global__ void k_pflip_apply(FlipMeshView fm, uint64_t seed, int round, int B, int nprop,
                              const int* __restrict__ status, int want)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nprop || status[i] != want) return;
    do_flip(fm, pflip_edge(seed, round, B, i, fm.nh / 2));
}

__global__ void k_pflip_eval2(FlipMeshView fm, NListView nl, EParams p, uint64_t seed, int round, int B, int nprop,
                              const int* __restrict__ status, double* __restrict__ dprops)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nprop || status[i] != 1) return;
    const Topo t{fm.x, fm.he_to, fm.he_next, fm.he_prev, fm.he_face, fm.v_he, fm.f_he};
    int e = pflip_edge(seed, round, B, i, fm.nh / 2);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int s = 0; s < 4; s++) {
        int h = 2 * e + (s >> 1);
        int ve = (s & 1) ? t.to(t.next(h)) : t.to(h);
        VP r = om_vertex_properties(t, nl, p, ve);
        for (int k = 0; k < 6; k++) acc[k] += r.v[k];
    }
    for (int k = 0; k < 6; k++) dprops[6 * i + k] += acc[k];
}

// state: props[6], e0, accepted count (as double)
__global__ void k_pflip_scan(EParams p, uint64_t seed, int round, int B, int nprop, int* __restrict__ status,
                             const double* __restrict__ dprops, double* __restrict__ state)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double props[6];
    for (int k = 0; k < 6; k++) props[k] = state[k];
    double e0 = state[6];
    double acc = state[7];
    for (int i = 0; i < nprop; i++) {
        if (status[i] != 1) continue;
        for (int k = 0; k < 6; k++) props[k] += dprops[6 * i + k];
        double en = energy_of(props, p);
        double de = en - e0;
        double alpha = de < 0.0 ? 1.0 : exp(-de);
        double u = rng_uniform(seed ^ 0x5bd1e9955bd1e995ull, (uint64_t)round * (uint64_t)B + (uint64_t)i);
        if (u <= alpha) {
            e0 = en;
            acc += 1.0;
            status[i] = 2;
        } else {
            for (int k = 0; k < 6; k++) props[k] -= dprops[6 * i + k];
            status[i] = 3;
        }
    }
    for (int k = 0; k < 6; k++) state[k] = props[k];
    state[6] = e0;
    state[7] = acc;
}

__global__ void k_pflip_init(const Globals* __restrict__ g0, EParams p, double* __restrict__ state)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double props[6];
    for (int k = 0; k < 6; k++) props[k] = state[k] = g0->props[k];
    state[6] = energy_of(props, p);
    state[7] = 0.0;
}

void launch_flip_parallel(const FlipMeshView& fm, const NListView& nl, const EParams& p, int nflips, int batch,
                          uint64_t seed, const Globals* g0, unsigned long long* owner, int* status, double* dprops,
                          double* state, hipStream_t s)
{
    hipLaunchKernelGGL(k_pflip_init, dim3(1), dim3(64), 0, s, g0, p, state);
    const int rounds = (nflips + batch - 1) / batch;
    for (int r = 0; r < rounds; r++) {
        int nprop = std::min(batch, nflips - r * batch);
        dim3 grid((nprop + 127) / 128), block(128);
        hipLaunchKernelGGL(k_pflip_claim, grid, block, 0, s, fm, seed, r, batch, nprop, owner);
        hipLaunchKernelGGL(k_pflip_eval, grid, block, 0, s, fm, nl, p, seed, r, batch, nprop, owner, status, dprops);
        hipLaunchKernelGGL(k_pflip_apply, grid, block, 0, s, fm, seed, r, batch, nprop, status, 1);
        hipLaunchKernelGGL(k_pflip_eval2, grid, block, 0, s, fm, nl, p, seed, r, batch, nprop, status, dprops);
        hipLaunchKernelGGL(k_pflip_scan, dim3(1), dim3(64), 0, s, p, seed, r, batch, nprop, status, dprops, state);
        hipLaunchKernelGGL(k_pflip_apply, grid, block, 0, s, fm, seed, r, batch, nprop, status, 3);
    }
}

}    // namespace tmh
