// tm_energy.hip -- the two energy passes and the global reduction.
//
// trimem evaluates the functional vertex by vertex (src/kernel.h:140-262,
// src/energy.cpp:63-124): a properties pass, a reduction of six globals, a
// gradient-gather pass and a combination.  The same two-pass structure is
// forced here (the gradient at v needs curvature/area of v's neighbours and
// the global A, V, C), but each pass is one kernel over vertices that reads
// the clockwise one-ring table:
//
//   pass 1  per vertex v: area, volume, curvature, tether, bending and
//           repulsion from the faces (v, j_m, j_{m-1}) and edges (v, j_m) of
//           its ring (mesh_properties.cpp:12-49); writes q_v = c_v / a_v and
//           block partial sums of the six globals in FP64; the last block to
//           finish sums the partials in a fixed order and writes the energy
//           and the penalty factors (kernel.h:19-122), so the result is
//           deterministic and needs no float atomics.
//   pass 2  per vertex v: the gather of mesh_properties.cpp:51-148 (spoke
//           edges, faces and rim edges of the ring, tether, repulsion) fused
//           with the combination of kernel.h:124-137; writes grad_v.
//
// Precision: T = double gives the FP64 path.  T = float gives the mixed path:
// edge vectors, face normals and the arguments of every angle are formed in
// FP64 (they cancel), the square roots, reciprocals, cotangents and dihedral
// angles (atan2) built from them run in FP32, and accumulation, the global
// reduction and the penalty factors stay FP64.
//
// Boundary vertices (ring with a faceless slot) take the OpenMesh-semantics
// walker in FP64 in both paths, which reproduces the reference's boundary
// branches of calc_sector_angle / calc_dihedral_angle.
#include "tm_internal.h"
#include "tm_om.h"

namespace tmh {

// Scalar kernels of the two precision modes.  Geometry vectors (edge vectors,
// face normals and the inputs of every angle) are always FP64; T is the type
// of the square roots, reciprocals, cotangents and angles built from them.
template <class T>
struct Prec;
template <>
struct Prec<double> {
    __device__ static __forceinline__ double sqrt_(double a) { return sqrt(a); }
    __device__ static __forceinline__ double dihedral(const d3& n0, const d3& n1, const d3& e, double)
    {
        // OpenMesh calc_dihedral_angle: acos of the normalised dot, signed
        double denom = sqrt(sqnorm(n0)) * sqrt(sqnorm(n1));
        if (denom == 0) return 0;
        return om_angle(dot(n0, n1) / denom, dot(cross(n0, n1), e));
    }
};
template <>
struct Prec<float> {
    __device__ static __forceinline__ float sqrt_(double a) { return sqrtf((float)a); }
    __device__ static __forceinline__ float dihedral(const d3& n0, const d3& n1, const d3& e, float l)
    {
        // the same angle as atan2(|n0 x n1|, n0.n1), signed by (n0 x n1).e;
        // n0 x n1 is parallel to e, so (n0 x n1).e / |e| is the signed sine
        // term.  Both arguments are formed in FP64 (they cancel for small
        // angles) and rounded once.
        double s = dot(cross(n0, n1), e);
        double c = dot(n0, n1);
        return atan2f((float)s / l, (float)c);
    }
};

__device__ __forceinline__ d3 load_p(const double* __restrict__ x, int v)
{
    return d3{x[3 * v], x[3 * v + 1], x[3 * v + 2]};
}

// ------------------------------------------------------------ reduction ---
// Sum six per-thread doubles over the block; result valid in s_out[0..5].
__device__ __forceinline__ void block_sum6(double v[6], double* s_red /* [nwaves*6] */, double* s_out)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double a = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        v[k] = a;
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 6; k++) s_red[wave * 6 + k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double a = 0;
        for (int w = 0; w < nw; w++) a += s_red[w * 6 + threadIdx.x];
        s_out[threadIdx.x] = a;
    }
    __syncthreads();
}

// Last block: fixed-order sum of all partials, energy and factors.
__device__ void finalize_globals(const double* __restrict__ partials, int nblocks, const EParams& p,
                                 Globals* g, double* s_red, double* s_out)
{
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < 6; k++) v[k] += partials[b * 6 + k];
    }
    block_sum6(v, s_red, s_out);
    if (threadIdx.x == 0) {
        double A = s_out[0], V = s_out[1], C = s_out[2], B = s_out[3], T = s_out[4], R = s_out[5];
        for (int k = 0; k < 6; k++) g->props[k] = s_out[k];
        double e = 0.0, d;
        d = A / p.ref_area - 1.0;
        e += p.kappa_a * d * d;
        g->fac_a = 2.0 * p.kappa_a / p.ref_area * d;
        d = V / p.ref_volume - 1.0;
        e += p.kappa_v * d * d;
        g->fac_v = 2.0 * p.kappa_v / p.ref_volume * d;
        d = C / p.ref_curvature - 1.0;
        e += p.kappa_c * d * d;
        g->fac_c = 2.0 * p.kappa_c / p.ref_curvature * d;
        e += p.kappa_t * T;
        e += p.kappa_r * R;
        e += p.kappa_b * B;
        g->energy = e;
    }
}

// --------------------------------------------------------------- pass 1 ---
// Ring slot k of an interior vertex v: halfedge v -> j_k, its face
// F_k = (v, j_k, j_{k-1}) and the twin face (j_k, v, j_{k+1}) = F_{k+1}.
// With e_k = x_{j_k} - x_v, the OpenMesh sector normal of v -> j_k is
// n_k = e_k x e_{k-1} (|n_k| = 2 area(F_k)) and that of j_k -> v is n_{k+1}.
template <class T>
__device__ __forceinline__ void ring_properties(const MeshView& m, const EParams& p, int v, int n,
                                                double out[6])
{
    const int N = m.nv;
    const d3 xv = load_p(m.x, v);
    d3 e_prev = load_p(m.x, m.ring[(n - 1) * N + v]) - xv;
    d3 e_cur = load_p(m.x, m.ring[v]) - xv;
    d3 n_cur = cross(e_cur, e_prev);
    double area = 0, vol = 0, curv = 0, teth = 0;
    for (int k = 0; k < n; k++) {
        int kn = (k + 1 == n) ? 0 : k + 1;
        d3 e_next = load_p(m.x, m.ring[kn * N + v]) - xv;
        d3 n_next = cross(e_next, e_cur);
        T af = Prec<T>::sqrt_(sqnorm(n_cur)) * (T)0.5;
        T l = Prec<T>::sqrt_(sqnorm(e_cur));
        T th = Prec<T>::dihedral(n_cur, n_next, e_cur, l);
        curv += (double)((T)0.5 * l * th);
        area += (double)af;
        vol += dot(xv, n_cur) / 6;    // = x_v . (x_j x x_k) / 6 (mesh_util.h:94-101)
        if (p.bond_type == 0) teth += bond_edge_value((double)l, p);
        else teth += bond_area_value((double)af, p);
        e_prev = e_cur;
        e_cur = e_next;
        n_cur = n_next;
    }
    double a = area / 3, c = curv / 2;
    out[0] = a;
    out[1] = vol / 3;
    out[2] = c;
    out[3] = 2 * c * c / a;
    out[4] = teth / (p.bond_type == 0 ? 2 : 3);
}

__device__ __forceinline__ double vertex_repulsion(const MeshView& m, const NListView& nl, const EParams& p, int v)
{
    if (!(p.rep_lc1 > 0) || nl.kind < 0) return 0.0;
    const Topo t = topo_of(m);
    const d3 pi = load_p(m.x, v);
    double rep = 0;
    auto ex = [&](int a, int b) { return nl.kind == 0 && om_excluded(t, a, b, nl.exclusion); };
    nlist_visit(nl, m.x, v, pi, p.rep_lc1, ex, [&](int, d3 d) { rep += rep_value(norm(d), p.rep_lc1, p.rep_r); });
    return rep;
}

template <class T>
__global__ void __launch_bounds__(256) k_pass1(MeshView m, NListView nl, EParams p, double* __restrict__ q,
                                               double* __restrict__ vprops, double* __restrict__ partials,
                                               unsigned* __restrict__ counter, Globals* __restrict__ g)
{
    __shared__ double s_red[4 * 6];
    __shared__ double s_out[6];
    __shared__ int s_last;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    double pr[6] = {0, 0, 0, 0, 0, 0};
    if (v < m.nv) {
        const int meta = m.meta[v];
        const int n = meta_valence(meta);
        if (meta_bslot(meta) < 0 && n >= 3) {
            ring_properties<T>(m, p, v, n, pr);
            pr[5] = vertex_repulsion(m, nl, p, v);
        } else {
            VP r = om_vertex_properties(topo_of(m), nl, p, v);
            for (int k = 0; k < 6; k++) pr[k] = r.v[k];
        }
        q[v] = pr[2] / pr[0];
        if (vprops) {
            for (int k = 0; k < 6; k++) vprops[6 * v + k] = pr[k];
        }
    }
    block_sum6(pr, s_red, s_out);
    if (threadIdx.x < 6) partials[blockIdx.x * 6 + threadIdx.x] = s_out[threadIdx.x];
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = (atomicAdd(counter, 1u) == gridDim.x - 1);
    __syncthreads();
    if (s_last) {
        __threadfence();
        finalize_globals(partials, gridDim.x, p, g, s_red, s_out);
        if (threadIdx.x == 0) *counter = 0u;
    }
}

// --------------------------------------------------------------- pass 2 ---
// Gradient at an interior vertex v.  Mathematically this is the gather of
// mesh_properties.cpp:51-148 combined as in kernel.h:124-137, regrouped so
// that no term relies on cancellation of large contributions:
//
//  * The l * dtheta parts of the spoke edges (v, j_k) and rim edges
//    (j_k, j_{k-1}) meet face by face.  For F_k with interior angles
//    alpha_j at j_k and alpha_i at j_{k-1} and unit normal u_k, the spoke
//    terms give w_{v j_k} cot(alpha_j) u_k + w_{v j_{k-1}} cot(alpha_i) u_k and
//    the rim term -w_{j_k j_{k-1}} (cot(alpha_j) + cot(alpha_i)) u_k, so with
//    edge weights w_{ab} = q_a + q_b (bending, mesh_properties.cpp:88-98,
//    121-129) the face contributes
//        [cot(alpha_j) (q_v - q_{k-1}) + cot(alpha_i) (q_v - q_k)] u_k
//    and with the constant weight of the curvature penalty it contributes 0
//    (Schlaefli's identity).  Only curvature/area differences remain.
//  * The theta * dl parts of the spokes stay as they are: -(q_v + q_k) theta e_hat.
//  * The face-volume gradients cross(x_j, x_k)/6 sum to sum_k n_k / 6 around a
//    closed ring (the x_v x rim terms telescope to zero).
template <class T>
__device__ __forceinline__ d3 ring_gradient(const MeshView& m, const NListView& nl, const EParams& p,
                                            const Globals& g, const double* __restrict__ q, int v, int n)
{
    const int N = m.nv;
    const d3 xv = load_p(m.x, v);
    const double qv = q[v];
    int j = m.ring[(n - 1) * N + v];
    d3 e_prev = load_p(m.x, j) - xv;
    double q_prev = q[j];
    j = m.ring[v];
    d3 e_cur = load_p(m.x, j) - xv;
    double q_cur = q[j];
    d3 n_cur = cross(e_cur, e_prev);
    d3 gA{0, 0, 0}, gN{0, 0, 0}, gB{0, 0, 0}, gC{0, 0, 0}, gT{0, 0, 0};
    const double kb = p.kappa_b;
    for (int k = 0; k < n; k++) {
        int kn = (k + 1 == n) ? 0 : k + 1;
        j = m.ring[kn * N + v];
        d3 e_next = load_p(m.x, j) - xv;
        double q_next = q[j];
        d3 n_next = cross(e_next, e_cur);
        // face F_k = (v, j_k, j_{k-1})
        T n0l = Prec<T>::sqrt_(sqnorm(n_cur));
        T inv0 = (T)1 / n0l;
        d3 rim = e_prev - e_cur;                               // x_{k-1} - x_k
        T cot_j = (T)(-dot(rim, e_cur)) * inv0;                // angle at j_k
        T cot_i = (T)dot(e_prev, rim) * inv0;                  // angle at j_{k-1}
        d3 u = (double)inv0 * n_cur;
        d3 dA = 0.5 * cross(u, rim);                           // d area(F_k) / d x_v
        gA += dA;
        gN += n_cur;
        double wq = (double)cot_j * (qv - q_prev) + (double)cot_i * (qv - q_cur);
        double q2 = qv * qv + q_cur * q_cur + q_prev * q_prev;
        gB += wq * u;
        gB -= (2.0 / 3.0 * q2) * dA;
        // spoke (v, j_k) with faces F_k and F_{k+1}
        T l = Prec<T>::sqrt_(sqnorm(e_cur));
        T th = Prec<T>::dihedral(n_cur, n_next, e_cur, l);
        d3 eh = (double)((T)1 / l) * e_cur;
        gB -= ((qv + q_cur) * (double)th) * eh;
        gC -= (0.5 * (double)th) * eh;
        if (p.bond_type == 0) gT -= bond_edge_dfac((double)l, p) * eh;
        else gT += bond_area_dfac((double)(n0l * (T)0.5), p) * dA;
        e_prev = e_cur;
        q_prev = q_cur;
        e_cur = e_next;
        q_cur = q_next;
        n_cur = n_next;
    }
    d3 gR{0, 0, 0};
    if (p.rep_lc1 > 0 && nl.kind >= 0 && p.kappa_r != 0) {
        const Topo t = topo_of(m);
        auto ex = [&](int a, int b) { return nl.kind == 0 && om_excluded(t, a, b, nl.exclusion); };
        nlist_visit(nl, m.x, v, xv, p.rep_lc1, ex,
                    [&](int, d3 d) { gR += (2.0 * rep_dfac(norm(d), p.rep_lc1, p.rep_r)) * d; });
    }
    d3 out{0, 0, 0};
    out += g.fac_a * gA;
    out += (g.fac_v / 6.0) * gN;
    out += g.fac_c * gC;
    out += p.kappa_t * gT;
    out += p.kappa_r * gR;
    out += kb * gB;
    return out;
}

template <class T>
__global__ void __launch_bounds__(256) k_pass2(MeshView m, NListView nl, EParams p, const double* __restrict__ q,
                                               const Globals* __restrict__ gp, double* __restrict__ grad)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.nv) return;
    const Globals g = *gp;
    const int meta = m.meta[v];
    const int n = meta_valence(meta);
    d3 out;
    if (meta_bslot(meta) < 0 && n >= 3) out = ring_gradient<T>(m, nl, p, g, q, v, n);
    else out = om_vertex_gradient(topo_of(m), nl, p, g, q, v);
    grad[3 * v] = out.x;
    grad[3 * v + 1] = out.y;
    grad[3 * v + 2] = out.z;
}

// ------------------------------------------------------------ launchers ---
static inline int nblocks(int n, int b) { return (n + b - 1) / b; }

int pass1_blocks(int nv) { return nblocks(nv, 256); }

void launch_pass1(const MeshView& m, const NListView& nl, const EParams& p, int mixed, double* q, double* vprops,
                  double* partials, unsigned* counter, Globals* g, hipStream_t s)
{
    dim3 grid(nblocks(m.nv, 256)), block(256);
    if (mixed) hipLaunchKernelGGL(k_pass1<float>, grid, block, 0, s, m, nl, p, q, vprops, partials, counter, g);
    else hipLaunchKernelGGL(k_pass1<double>, grid, block, 0, s, m, nl, p, q, vprops, partials, counter, g);
}

void launch_pass2(const MeshView& m, const NListView& nl, const EParams& p, int mixed, const double* q,
                  const Globals* g, double* grad, hipStream_t s)
{
    dim3 grid(nblocks(m.nv, 256)), block(256);
    if (mixed) hipLaunchKernelGGL(k_pass2<float>, grid, block, 0, s, m, nl, p, q, g, grad);
    else hipLaunchKernelGGL(k_pass2<double>, grid, block, 0, s, m, nl, p, q, g, grad);
}

// --------------------------------------------- single-halfedge geometry ---
// out: [l, dl(2x3), A, dA(3x3), V, dV(3x3), theta, dtheta(4x3)]  (47 doubles)
__global__ void k_halfedge_geometry(MeshView m, int h, double* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const Topo t = topo_of(m);
    int k = 0;
    out[k++] = mu_edge_length(t, h);
    d3 g0 = mu_edge_length_grad_from(t, h);
    d3 gs[4];
    gs[0] = g0;
    gs[1] = -g0;
    for (int i = 0; i < 2; i++) { out[k++] = gs[i].x; out[k++] = gs[i].y; out[k++] = gs[i].z; }
    out[k++] = mu_face_area(t, h);
    for (int i = 0; i < 3; i++) { d3 a = mu_face_area_grad(t, h, i); out[k++] = a.x; out[k++] = a.y; out[k++] = a.z; }
    out[k++] = mu_face_volume(t, h);
    for (int i = 0; i < 3; i++) { d3 a = mu_face_volume_grad(t, h, i); out[k++] = a.x; out[k++] = a.y; out[k++] = a.z; }
    out[k++] = om_dihedral_angle(t, h);
    for (int i = 0; i < 4; i++) { d3 a = mu_dihedral_angle_grad(t, h, i); out[k++] = a.x; out[k++] = a.y; out[k++] = a.z; }
}

void launch_halfedge_geometry(const MeshView& m, int h, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_halfedge_geometry, dim3(1), dim3(64), 0, s, m, h, out);
}

// -------------------------------------------------- mesh-wide utilities ---
// sum of face areas (util.cpp:9-21) or edge lengths (util.cpp:23-35)
__global__ void k_sum_geometry(MeshView m, int what, double* partials, unsigned* counter, double* result)
{
    __shared__ double s_red[4 * 6];
    __shared__ double s_out[6];
    __shared__ int s_last;
    const Topo t = topo_of(m);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (what == 0 && i < m.nf) v[0] = mu_face_area(t, m.f_he[i]);
    if (what == 1 && i < m.nh / 2) v[0] = mu_edge_length(t, 2 * i);
    block_sum6(v, s_red, s_out);
    if (threadIdx.x == 0) partials[blockIdx.x] = s_out[0];
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = (atomicAdd(counter, 1u) == gridDim.x - 1);
    __syncthreads();
    if (s_last) {
        __threadfence();
        double a[6] = {0, 0, 0, 0, 0, 0};
        for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) a[0] += partials[b];
        block_sum6(a, s_red, s_out);
        if (threadIdx.x == 0) {
            *result = s_out[0];
            *counter = 0u;
        }
    }
}

void launch_sum_geometry(const MeshView& m, int what, double* partials, unsigned* counter, double* result,
                         hipStream_t s)
{
    int n = what == 0 ? m.nf : m.nh / 2;
    hipLaunchKernelGGL(k_sum_geometry, dim3(nblocks(n > 0 ? n : 1, 256)), dim3(256), 0, s, m, what, partials,
                       counter, result);
}

}    // namespace tmh
