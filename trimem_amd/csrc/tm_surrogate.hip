// tm_surrogate.hip -- a one-ring message-passing network that predicts the
// membrane forces, trained on the GPU from exact (positions, gradient) pairs.
//
// No counterpart in the reference.  The network drives HMC trajectories in
// place of the exact gradient; every trajectory is still accepted or rejected
// with the exact energy, so the chain samples the exact distribution (a
// position-dependent force field keeps the leapfrog map volume preserving and
// reversible, whether or not it is a gradient).
//
// Network (width W = 64, closed meshes, valence <= kSlots):
//   inputs   f_v   (8, normalised): vertex area, mean curvature c_v/a_v,
//                  valence, mean spoke length, the relative deviations of
//                  the global area, volume and curvature from their reference
//                  values, angle deficit; eps_vk (2): spoke length
//                  (l - lbar)/lbar and dihedral angle of edge (v, j_k)
//   layer 0  Z0 = f W_in + b_in,  H0 = silu(Z0)
//   layer l  P = H U_l,  M_v = (1/6) sum_k silu(P_{j_k} + eps_vk E_l),
//            Z = [H | M] [W_l; V_l] + b_l,  H <- H + silu(Z)       (l = 1..3)
//   output   s_vk = w_s . (H_v * H_{j_k}) + b_s,  t_v = w_t . H_v + b_t,
//            F_v = sigma_F (sum_k s_vk e_vk / |e_vk| + t_v nu_v)
// F is rotation equivariant (edge directions and the vertex normal nu_v carry
// all directions; every learned quantity is invariant) and s is symmetric in
// (v, j).  Dense products run on MFMA (bf16 operands, FP32 accumulation,
// v_mfma_f32_16x16x16_bf16); activations and all element-wise work are FP32.
// The training loss is mean |F/sigma_F - F*/sigma_F|^2 over vertices and
// components, F* = -dE/dx from the exact passes.
#include "tm_internal.h"
#include "tm_om.h"
#include "tm_surrogate.h"

namespace tmh {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// FP32 -> bf16, round to nearest even (finite inputs)
__device__ __forceinline__ short to_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float silu(float x) { return x / (1.0f + __expf(-x)); }
__device__ __forceinline__ float silu_grad(float x)
{
    const float s = 1.0f / (1.0f + __expf(-x));
    return s * (1.0f + x * (1.0f - s));
}

// ------------------------------------------------------------ packing -----
// B operands of the GEMMs in bf16, as Bt[col][k] (k contiguous): forward
// products take W^T of the stored [k_in][c_out] weights, backward data
// products take the weights as stored.
__global__ void k_sg_pack(const float* __restrict__ theta, SgPacked pk)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    // input layer forward: in_t[c][k] = W_in[k][c]  (64 x 16)
    if (t < kSgW * kSgFeatPad) {
        const int c = t / kSgFeatPad, k = t % kSgFeatPad;
        pk.in_t[t] = to_bf16(theta[kSgOffWin + k * kSgW + c]);
    }
    for (int l = 0; l < kSgLayers; l++) {
        const float* U = theta + sg_off_U(l);
        const float* Wm = theta + sg_off_W(l);
        const float* V = theta + sg_off_V(l);
        if (t < kSgW * kSgW) {
            const int c = t / kSgW, k = t % kSgW;
            pk.u_t[l][t] = to_bf16(U[k * kSgW + c]);
            pk.u_n[l][t] = to_bf16(U[t]);
            pk.w_n[l][t] = to_bf16(Wm[t]);
            pk.v_n[l][t] = to_bf16(V[t]);
        }
        if (t < kSgW * 2 * kSgW) {
            const int c = t / (2 * kSgW), k = t % (2 * kSgW);
            pk.wv_t[l][t] = to_bf16(k < kSgW ? Wm[k * kSgW + c] : V[(k - kSgW) * kSgW + c]);
        }
    }
}

// --------------------------------------------------------------- GEMMs -----
// Y[n][c] = sum_k A[n][k] Bt[c][k] over K = KA + KB (A = [A1 | A2]), c < 64.
// Block: 4 waves, 64 rows, staged in LDS with the weights; wave w owns rows
// 16w..16w+15 and all 64 columns (four 16x16 tiles).  MFMA 16x16x16 operand lanes: A lane i holds row i % 16,
// k = 4 (i / 16) + 0..3; B lane i holds column i % 16, the same k; the
// accumulator lane i holds rows 4 (i / 16) + 0..3 of column i % 16.
// Epilogues: kSgStore Y = acc; kSgAccum Y += acc; kSgAct Z = acc + bias and
// H = (res ? res : 0) + silu(Z).
enum { kSgStore = 0, kSgAccum = 1, kSgAct = 2 };
template <int KA, int KB, int MODE>
__global__ void __launch_bounds__(256) k_sg_gemm(int N, const float* __restrict__ A1, const float* __restrict__ A2,
                                                 const short* __restrict__ Bt, const float* __restrict__ bias,
                                                 const float* __restrict__ res, float* __restrict__ Y,
                                                 float* __restrict__ H)
{
    constexpr int K = KA + KB;
    constexpr int KP = K + 8;    // padded LDS row (bf16): spreads the fragment reads over banks
    // the block's 64 rows of A and all of B, staged as bf16 with coalesced
    // 16-byte global loads
    __shared__ short s_a[64 * KP];
    __shared__ short s_b[64 * KP];
    const int row_base = blockIdx.x * 64;
    for (int q = threadIdx.x; q < 64 * (K / 4); q += 256) {
        const int r = q / (K / 4), k = 4 * (q % (K / 4));
        const int row = row_base + r;
        float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
        if (row < N)
            x = k < KA ? *reinterpret_cast<const float4*>(A1 + (size_t)row * KA + k)
                       : *reinterpret_cast<const float4*>(A2 + (size_t)row * KB + (k - KA));
        short* d = s_a + r * KP + k;
        d[0] = to_bf16(x.x);
        d[1] = to_bf16(x.y);
        d[2] = to_bf16(x.z);
        d[3] = to_bf16(x.w);
    }
    for (int q = threadIdx.x; q < 64 * (K / 4); q += 256) {
        const int c = q / (K / 4), k = 4 * (q % (K / 4));
        const short4 x = *reinterpret_cast<const short4*>(Bt + c * K + k);
        short* d = s_b + c * KP + k;
        d[0] = x.x;
        d[1] = x.y;
        d[2] = x.z;
        d[3] = x.w;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int kq = 4 * (lane >> 4);
    const short* sa = s_a + (w * 16 + (lane & 15)) * KP + kq;
    floatx4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = floatx4{0, 0, 0, 0};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a;
#pragma unroll
        for (int i = 0; i < 4; i++) a[i] = sa[k0 + i];
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const short* sb = s_b + (16 * t + (lane & 15)) * KP + kq + k0;
            bf16x4 b;
#pragma unroll
            for (int i = 0; i < 4; i++) b[i] = sb[i];
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
    const int row0 = row_base + w * 16;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int col = 16 * t + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int row = row0 + kq + r;
            if (row >= N) continue;
            const size_t o = (size_t)row * kSgW + col;
            float y = acc[t][r];
            if (MODE == kSgStore) Y[o] = y;
            else if (MODE == kSgAccum) Y[o] += y;
            else {
                y += bias[col];
                Y[o] = y;
                H[o] = (res ? res[o] : 0.0f) + silu(y);
            }
        }
    }
}

// Weight gradient partials: G[k][c] = sum_n A[n][k] D[n][c] over the block's
// rows, K = KA + KB, c < 64.  Here the MFMA "k" dimension runs over rows n:
// A operand lane i holds feature row kt*16 + i % 16 at n = n0 + 4 (i/16) + 0..3,
// B operand lane i holds column i % 16 (+16 t) at the same n.  Wave w owns
// the feature tiles w, w + 4, ... of the K/16 tiles.  Partials
// [block][K][64] are summed in block order by k_sg_reduce.
template <int KA, int KB>
__global__ void __launch_bounds__(256) k_sg_wgrad(int N, int rows_per_block, const float* __restrict__ A1,
                                                  const float* __restrict__ A2, const float* __restrict__ D,
                                                  float* __restrict__ partial)
{
    constexpr int K = KA + KB;
    constexpr int KT = K / 16;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n_begin = blockIdx.x * rows_per_block;
    const int n_end = min(N, n_begin + rows_per_block);
    const int kq = 4 * (lane >> 4);
    for (int kt = w; kt < KT; kt += 4) {
        floatx4 acc[4];
#pragma unroll
        for (int t = 0; t < 4; t++) acc[t] = floatx4{0, 0, 0, 0};
        const int fk = kt * 16 + (lane & 15);
        for (int n0 = n_begin; n0 < n_end; n0 += 16) {
            bf16x4 a;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int n = n0 + kq + i;
                float x = 0.0f;
                if (n < n_end) x = fk < KA ? A1[(size_t)n * KA + fk] : A2[(size_t)n * KB + (fk - KA)];
                a[i] = to_bf16(x);
            }
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int col = 16 * t + (lane & 15);
                bf16x4 b;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int n = n0 + kq + i;
                    b[i] = to_bf16(n < n_end ? D[(size_t)n * kSgW + col] : 0.0f);
                }
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
            }
        }
        // acc[t] lane i: rows (feature) kt*16 + 4 (i/16) + r, column 16 t + i % 16
#pragma unroll
        for (int t = 0; t < 4; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int k = kt * 16 + kq + r;
                const int col = 16 * t + (lane & 15);
                partial[((size_t)blockIdx.x * K + k) * kSgW + col] = acc[t][r];
            }
    }
}

// column sums of D[N][64] per block (bias gradients)
__global__ void __launch_bounds__(256) k_sg_colsum(int N, int rows_per_block, const float* __restrict__ D,
                                                   float* __restrict__ partial)
{
    __shared__ float s[4][kSgW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n_begin = blockIdx.x * rows_per_block;
    const int n_end = min(N, n_begin + rows_per_block);
    float a = 0.0f;
    for (int n = n_begin + w; n < n_end; n += 4) a += D[(size_t)n * kSgW + lane];
    s[w][lane] = a;
    __syncthreads();
    if (w == 0) partial[(size_t)blockIdx.x * kSgW + lane] = s[0][lane] + s[1][lane] + s[2][lane] + s[3][lane];
}

// out[i] (+)= scale * sum_b partial[b][i]: one block per output element,
// each thread a fixed stride of the partials, then a fixed tree
__global__ void __launch_bounds__(256) k_sg_reduce(int nblocks, int count, const float* __restrict__ partial,
                                                   float scale, int accumulate, float* __restrict__ out)
{
    __shared__ float s[256];
    const int i = blockIdx.x;
    float a = 0.0f;
    for (int b = threadIdx.x; b < nblocks; b += 256) a += partial[(size_t)b * count + i];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = accumulate ? out[i] + scale * s[0] : scale * s[0];
}

// ------------------------------------------------------------ features -----
// One thread per vertex: geometry of the ring (FP64 differences, FP32
// results), the normalised input features and the per-spoke edge data.
__global__ void k_sg_features(MeshView m, const double* __restrict__ vprops, const Globals* __restrict__ g,
                              SgNorm nrm, SgGeom geo, float* __restrict__ f0, int* __restrict__ err)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.nv) return;
    const int N = m.nv;
    const int meta = m.meta[v];
    const int n = meta_valence(meta);
    if (meta_bslot(meta) >= 0 || n < 3 || n > kSgSlots) {
        atomicOr(err, 1);
        return;
    }
    const d3 xv = d3{m.x[3 * v], m.x[3 * v + 1], m.x[3 * v + 2]};
    d3 e[kSgSlots];
    for (int k = 0; k < n; k++) {
        const int j = m.ring[k * N + v];
        e[k] = d3{m.x[3 * j], m.x[3 * j + 1], m.x[3 * j + 2]} - xv;
    }
    d3 nsum{0, 0, 0};
    double lsum = 0.0, angles = 0.0;
    for (int k = 0; k < n; k++) {
        const d3 nk = cross(e[k], e[(k + n - 1) % n]);           // face (v, j_k, j_{k-1})
        const d3 nk1 = cross(e[(k + 1) % n], e[k]);              // face (v, j_{k+1}, j_k)
        nsum += nk;
        angles += atan2(sqrt(sqnorm(nk)), dot(e[k], e[(k + n - 1) % n]));    // corner angle at v
        const double l = sqrt(sqnorm(e[k]));
        lsum += l;
        // dihedral of edge (v, j_k) between the two faces, signed by
        // (n_k x n_{k+1}) . e_k like the properties pass
        const float s = (float)(dot(cross(nk, nk1), e[k]) / l);
        const float c = (float)dot(nk, nk1);
        const int slot = k * N + v;
        geo.eps[2 * slot] = (float)((l - nrm.lbar) / nrm.lbar);
        geo.eps[2 * slot + 1] = atan2f(s, c);
        geo.ehat[3 * slot] = (float)(e[k].x / l);
        geo.ehat[3 * slot + 1] = (float)(e[k].y / l);
        geo.ehat[3 * slot + 2] = (float)(e[k].z / l);
    }
    const double nn = sqrt(sqnorm(nsum));
    geo.nu[3 * v] = (float)(nsum.x / nn);
    geo.nu[3 * v + 1] = (float)(nsum.y / nn);
    geo.nu[3 * v + 2] = (float)(nsum.z / nn);
    const double a = vprops[6 * v], c = vprops[6 * v + 2];
    // relative deviations of the global area, volume and curvature from
    // their reference values (the penalty factors without their kappas), and
    // the angle deficit (the discrete Gaussian curvature of v)
    const double raw[kSgFeat] = {a, c / a, (double)n, lsum / n, g->props[0] / nrm.ref[0] - 1.0,
                                 g->props[1] / nrm.ref[1] - 1.0, g->props[2] / nrm.ref[2] - 1.0,
                                 6.283185307179586 - angles};
    for (int i = 0; i < kSgFeat; i++) f0[v * kSgFeatPad + i] = (float)((raw[i] - nrm.mu[i]) / nrm.sig[i]);
    for (int i = kSgFeat; i < kSgFeatPad; i++) f0[v * kSgFeatPad + i] = 0.0f;
}

// slot of v in the ring of its k-th neighbour (closed meshes)
__global__ void k_sg_twin(MeshView m, int* __restrict__ twin)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= m.nv) return;
    const int N = m.nv;
    const int n = meta_valence(m.meta[v]);
    for (int k = 0; k < n && k < kSgSlots; k++) {
        const int j = m.ring[k * N + v];
        const int nj = meta_valence(m.meta[j]);
        int t = -1;
        for (int a = 0; a < nj && a < kSgSlots; a++)
            if (m.ring[a * N + j] == v) t = a;
        twin[k * N + v] = t;
    }
}

// ------------------------------------------------------------- messages -----
// One wave per vertex (grid-stride), lane = channel.
__global__ void __launch_bounds__(256) k_sg_message(MeshView m, const float* __restrict__ P,
                                                    const float* __restrict__ eps, const float* __restrict__ E,
                                                    float* __restrict__ M)
{
    const int c = threadIdx.x & 63;
    const int N = m.nv;
    const float e0 = E[c], e1 = E[kSgW + c];
    for (int v = blockIdx.x * 4 + (threadIdx.x >> 6); v < N; v += gridDim.x * 4) {
        const int n = meta_valence(m.meta[v]);
        float acc = 0.0f;
        for (int k = 0; k < n; k++) {
            const int j = m.ring[k * N + v];
            const int s = k * N + v;
            acc += silu(P[(size_t)j * kSgW + c] + eps[2 * s] * e0 + eps[2 * s + 1] * e1);
        }
        M[(size_t)v * kSgW + c] = acc * (1.0f / 6.0f);
    }
}

// Pull form of the message backward: the spoke (u -> j) feeds M_u through
// P_j, so dP_j = sum over the neighbours u of j of (1/6) dM_u silu'(A_uj),
// A_uj = P_j + eps_uj E.  Each directed spoke is visited once, from j, so
// dE needs no atomics: per-block partials of sum dA eps.
__global__ void __launch_bounds__(256) k_sg_message_bwd(MeshView m, const int* __restrict__ twin,
                                                        const float* __restrict__ P, const float* __restrict__ eps,
                                                        const float* __restrict__ E, const float* __restrict__ dM,
                                                        float* __restrict__ dP, float* __restrict__ dE_partial)
{
    __shared__ float s_e[4][2 * kSgW];
    const int w = threadIdx.x >> 6, c = threadIdx.x & 63;
    float de0 = 0.0f, de1 = 0.0f;
    for (int j = blockIdx.x * 4 + w; j < m.nv; j += gridDim.x * 4) {
        const int N = m.nv;
        const int n = meta_valence(m.meta[j]);
        const float e0 = E[c], e1 = E[kSgW + c];
        const float pj = P[(size_t)j * kSgW + c];
        float acc = 0.0f;
        for (int k = 0; k < n; k++) {
            const int u = m.ring[k * N + j];
            const int s = twin[k * N + j] * N + u;    // the spoke (u -> j) in u's ring
            const float x0 = eps[2 * s], x1 = eps[2 * s + 1];
            const float da = (1.0f / 6.0f) * dM[(size_t)u * kSgW + c] * silu_grad(pj + x0 * e0 + x1 * e1);
            acc += da;
            de0 += da * x0;
            de1 += da * x1;
        }
        dP[(size_t)j * kSgW + c] = acc;
    }
    s_e[w][c] = de0;
    s_e[w][kSgW + c] = de1;
    __syncthreads();
    if (w == 0) {
        float a0 = 0.0f, a1 = 0.0f;
        for (int q = 0; q < 4; q++) {
            a0 += s_e[q][c];
            a1 += s_e[q][kSgW + c];
        }
        dE_partial[(size_t)blockIdx.x * 2 * kSgW + c] = a0;
        dE_partial[(size_t)blockIdx.x * 2 * kSgW + kSgW + c] = a1;
    }
}

// dZ = dH * silu'(Z)
__global__ void k_sg_dz(int count, const float* __restrict__ dH, const float* __restrict__ Z, float* __restrict__ dZ)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dZ[i] = dH[i] * silu_grad(Z[i]);
}

__device__ __forceinline__ float wave_sum(float x)
{
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

// --------------------------------------------------------------- output -----
// One wave per vertex: s_vk, t_v, the force; with a target also the loss
// gradient g_v = d loss / d F_hat_v (F_hat = F / sigma_F) and per-block loss
// partials.  The target is the exact gradient dE/dx (so F* = -target); the
// force is written as its negative (the gradient the HMC step expects).
__global__ void __launch_bounds__(256) k_sg_output(MeshView m, const float* __restrict__ H, SgGeom geo,
                                                   const float* __restrict__ theta, float sigma_F,
                                                   const double* __restrict__ target, double* __restrict__ neg_force,
                                                   float* __restrict__ ghat, float* __restrict__ loss_partial,
                                                   float loss_scale)
{
    __shared__ float s_loss[4];
    const int w = threadIdx.x >> 6, c = threadIdx.x & 63;
    float lw = 0.0f;
    for (int v = blockIdx.x * 4 + w; v < m.nv; v += gridDim.x * 4) {
        const int N = m.nv;
        const int n = meta_valence(m.meta[v]);
        const float* ws = theta + kSgOffWs;
        const float hv = H[(size_t)v * kSgW + c];
        float F[3] = {0, 0, 0};
        for (int k = 0; k < n; k++) {
            const int j = m.ring[k * N + v];
            const float s = wave_sum(ws[c] * hv * H[(size_t)j * kSgW + c]) + theta[kSgOffBs];
            const int sl = k * N + v;
            F[0] += s * geo.ehat[3 * sl];
            F[1] += s * geo.ehat[3 * sl + 1];
            F[2] += s * geo.ehat[3 * sl + 2];
        }
        const float t = wave_sum(theta[kSgOffWt + c] * hv) + theta[kSgOffBt];
        for (int d = 0; d < 3; d++) F[d] += t * geo.nu[3 * v + d];
        if (c < 3) {
            if (neg_force) neg_force[3 * v + c] = -(double)sigma_F * F[c];
            if (target) {
                const float r = F[c] + (float)(target[3 * v + c] / sigma_F);
                ghat[3 * v + c] = 2.0f * loss_scale * r;
                lw += loss_scale * r * r;
            }
        }
    }
    lw = wave_sum(lw);
    if (c == 0) s_loss[w] = lw;
    __syncthreads();
    if (threadIdx.x == 0 && loss_partial) loss_partial[blockIdx.x] = s_loss[0] + s_loss[1] + s_loss[2] + s_loss[3];
}

// Backward of the output layer, one wave per vertex:
//   dH_v = w_s * sum_k ((g_v - g_{j_k}) . e_vk) H_{j_k} + (g_v . nu_v) w_t
// (the spoke (j -> v) contributes through H_v with e_jv = -e_vj), and block
// partials of d w_s (64), d b_s, d w_t (64), d b_t.
__global__ void __launch_bounds__(256) k_sg_output_bwd(MeshView m, const float* __restrict__ H, SgGeom geo,
                                                       const float* __restrict__ theta,
                                                       const float* __restrict__ ghat, float* __restrict__ dH,
                                                       float* __restrict__ partial)
{
    __shared__ float s_p[4][kSgOutGrad];
    const int w = threadIdx.x >> 6, c = threadIdx.x & 63;
    float dws = 0.0f, dbs = 0.0f, dwt = 0.0f, dbt = 0.0f;
    for (int v = blockIdx.x * 4 + w; v < m.nv; v += gridDim.x * 4) {
        const int N = m.nv;
        const int n = meta_valence(m.meta[v]);
        const float hv = H[(size_t)v * kSgW + c];
        const float gv0 = ghat[3 * v], gv1 = ghat[3 * v + 1], gv2 = ghat[3 * v + 2];
        float acc = 0.0f;
        for (int k = 0; k < n; k++) {
            const int j = m.ring[k * N + v];
            const int sl = k * N + v;
            const float ex = geo.ehat[3 * sl], ey = geo.ehat[3 * sl + 1], ez = geo.ehat[3 * sl + 2];
            const float ds_v = gv0 * ex + gv1 * ey + gv2 * ez;
            const float ds_j = ghat[3 * j] * ex + ghat[3 * j + 1] * ey + ghat[3 * j + 2] * ez;
            const float hj = H[(size_t)j * kSgW + c];
            acc += (ds_v - ds_j) * hj;
            dws += ds_v * hv * hj;
            dbs += ds_v;
        }
        const float dt = gv0 * geo.nu[3 * v] + gv1 * geo.nu[3 * v + 1] + gv2 * geo.nu[3 * v + 2];
        dH[(size_t)v * kSgW + c] = theta[kSgOffWs + c] * acc + dt * theta[kSgOffWt + c];
        dwt += dt * hv;
        dbt += dt;
    }
    // the parameter layout: w_s (64), b_s, w_t (64), b_t
    s_p[w][c] = dws;
    s_p[w][kSgW + 1 + c] = dwt;
    if (c == 0) {
        s_p[w][kSgW] = dbs;
        s_p[w][2 * kSgW + 1] = dbt;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kSgOutGrad; i += blockDim.x)
        partial[(size_t)blockIdx.x * kSgOutGrad + i] = s_p[0][i] + s_p[1][i] + s_p[2][i] + s_p[3][i];
}

// ------------------------------------------------------------------ Adam -----
__global__ void k_sg_adam(int n, float* __restrict__ theta, const float* __restrict__ grad, float* __restrict__ m1,
                          float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.0f - b1) * g;
    const float b = b2 * m2[i] + (1.0f - b2) * g * g;
    m1[i] = a;
    m2[i] = b;
    theta[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// ------------------------------------------------------------ launchers -----
static inline int cdiv(int a, int b) { return (a + b - 1) / b; }
// blocks of the one-wave-per-vertex kernels (4 vertices per block and
// step, grid-stride beyond): a capped grid keeps their partials few
int sg_vertex_blocks(int nv) { return std::min(cdiv(nv, 4), 4096); }

// the largest reduction: K = 128 weight-gradient partials of every block
size_t sg_partial_floats(int nv) { return (size_t)cdiv(nv, kSgRowsPerBlock) * 2 * kSgW * kSgW + (size_t)cdiv(nv, 4) * kSgOutGrad + 1024; }

void sg_pack(const float* theta, const SgPacked& pk, hipStream_t s)
{
    hipLaunchKernelGGL(k_sg_pack, dim3(cdiv(kSgW * 2 * kSgW, 256)), dim3(256), 0, s, theta, pk);
}

void sg_features(const MeshView& m, const double* vprops, const Globals* g, const SgNorm& nrm, const SgGeom& geo,
                 float* f0, int* err, hipStream_t s)
{
    hipLaunchKernelGGL(k_sg_features, dim3(cdiv(m.nv, 128)), dim3(128), 0, s, m, vprops, g, nrm, geo, f0, err);
}

void sg_twin(const MeshView& m, int* twin, hipStream_t s)
{
    hipLaunchKernelGGL(k_sg_twin, dim3(cdiv(m.nv, 128)), dim3(128), 0, s, m, twin);
}

template <int KA, int KB, int MODE>
static void gemm(int N, const float* A1, const float* A2, const short* Bt, const float* bias, const float* res,
                 float* Y, float* H, hipStream_t s)
{
    hipLaunchKernelGGL((k_sg_gemm<KA, KB, MODE>), dim3(cdiv(N, 64)), dim3(256), 0, s, N, A1, A2, Bt, bias, res, Y,
                       H);
}

template <int KA, int KB>
static void wgrad(int N, const float* A1, const float* A2, const float* D, float* partial, float* out, float scale,
                  hipStream_t s)
{
    const int rows = kSgRowsPerBlock;
    const int nb = cdiv(N, rows);
    hipLaunchKernelGGL((k_sg_wgrad<KA, KB>), dim3(nb), dim3(256), 0, s, N, rows, A1, A2, D, partial);
    const int count = (KA + KB) * kSgW;
    hipLaunchKernelGGL(k_sg_reduce, dim3(count), dim3(256), 0, s, nb, count, partial, scale, 0, out);
}

static void colsum(int N, const float* D, float* partial, float* out, hipStream_t s)
{
    const int nb = cdiv(N, kSgRowsPerBlock);
    hipLaunchKernelGGL(k_sg_colsum, dim3(nb), dim3(256), 0, s, N, kSgRowsPerBlock, D, partial);
    hipLaunchKernelGGL(k_sg_reduce, dim3(kSgW), dim3(256), 0, s, nb, kSgW, partial, 1.0f, 0, out);
}

// Forward: features already in b.f0; with a target also the loss gradient.
void sg_forward(const MeshView& m, const SgBuffers& b, const SgGeom& geo, const float* theta, const SgPacked& pk,
                float sigma_F, const double* target, double* neg_force, hipStream_t s)
{
    const int N = m.nv;
    gemm<kSgFeatPad, 0, kSgAct>(N, b.f0, nullptr, pk.in_t, theta + kSgOffBin, nullptr, b.Z[0], b.H[0], s);
    for (int l = 0; l < kSgLayers; l++) {
        gemm<kSgW, 0, kSgStore>(N, b.H[l], nullptr, pk.u_t[l], nullptr, nullptr, b.P[l], nullptr, s);
        hipLaunchKernelGGL(k_sg_message, dim3(cdiv(N, 4)), dim3(256), 0, s, m, b.P[l], geo.eps,
                           theta + sg_off_E(l), b.M[l]);
        gemm<kSgW, kSgW, kSgAct>(N, b.H[l], b.M[l], pk.wv_t[l], theta + sg_off_b(l), b.H[l], b.Z[l + 1],
                                 b.H[l + 1], s);
    }
    const float loss_scale = 1.0f / (3.0f * (float)N);
    hipLaunchKernelGGL(k_sg_output, dim3(sg_vertex_blocks(N)), dim3(256), 0, s, m, b.H[kSgLayers], geo, theta, sigma_F,
                       target, neg_force, b.ghat, b.loss_partial, loss_scale);
    if (target)
        hipLaunchKernelGGL(k_sg_reduce, dim3(1), dim3(256), 0, s, sg_vertex_blocks(N), 1, b.loss_partial, 1.0f, 0,
                           b.loss);
}

// Backward after a forward with a target: parameter gradient into grad.
void sg_backward(const MeshView& m, const SgBuffers& b, const SgGeom& geo, const float* theta, const SgPacked& pk,
                 const int* twin, float* grad, hipStream_t s)
{
    const int N = m.nv;
    const int nb4 = sg_vertex_blocks(N);
    const int cnt = N * kSgW;
    // output layer
    hipLaunchKernelGGL(k_sg_output_bwd, dim3(nb4), dim3(256), 0, s, m, b.H[kSgLayers], geo, theta, b.ghat, b.dH,
                       b.partial);
    hipLaunchKernelGGL(k_sg_reduce, dim3(kSgOutGrad), dim3(256), 0, s, nb4, kSgOutGrad, b.partial, 1.0f, 0,
                       grad + kSgOffWs);
    // message-passing layers, last to first; b.dH holds dL/dH_l on entry
    for (int l = kSgLayers - 1; l >= 0; l--) {
        hipLaunchKernelGGL(k_sg_dz, dim3(cdiv(cnt, 256)), dim3(256), 0, s, cnt, b.dH, b.Z[l + 1], b.dZ);
        wgrad<kSgW, kSgW>(N, b.H[l], b.M[l], b.dZ, b.partial, grad + sg_off_W(l), 1.0f, s);    // [W; V]
        colsum(N, b.dZ, b.partial, grad + sg_off_b(l), s);
        // residual: dH_l = dH_{l+1} + dZ W^T + dP U^T
        gemm<kSgW, 0, kSgAccum>(N, b.dZ, nullptr, pk.w_n[l], nullptr, nullptr, b.dH, nullptr, s);
        gemm<kSgW, 0, kSgStore>(N, b.dZ, nullptr, pk.v_n[l], nullptr, nullptr, b.dM, nullptr, s);
        hipLaunchKernelGGL(k_sg_message_bwd, dim3(nb4), dim3(256), 0, s, m, twin, b.P[l], geo.eps,
                           theta + sg_off_E(l), b.dM, b.dP, b.partial);
        hipLaunchKernelGGL(k_sg_reduce, dim3(2 * kSgW), dim3(256), 0, s, nb4, 2 * kSgW, b.partial, 1.0f, 0,
                           grad + sg_off_E(l));
        wgrad<kSgW, 0>(N, b.H[l], nullptr, b.dP, b.partial, grad + sg_off_U(l), 1.0f, s);
        gemm<kSgW, 0, kSgAccum>(N, b.dP, nullptr, pk.u_n[l], nullptr, nullptr, b.dH, nullptr, s);
    }
    // input layer: H0 = silu(Z0)
    hipLaunchKernelGGL(k_sg_dz, dim3(cdiv(cnt, 256)), dim3(256), 0, s, cnt, b.dH, b.Z[0], b.dZ);
    wgrad<kSgFeatPad, 0>(N, b.f0, nullptr, b.dZ, b.partial, grad + kSgOffWin, 1.0f, s);
    colsum(N, b.dZ, b.partial, grad + kSgOffBin, s);
}

void sg_adam(int n, float* theta, const float* grad, float* m1, float* m2, float lr, int step, hipStream_t s)
{
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_sg_adam, dim3(cdiv(n, 256)), dim3(256), 0, s, n, theta, grad, m1, m2, lr, b1, b2, 1e-8f,
                       c1, c2);
}

}    // namespace tmh
