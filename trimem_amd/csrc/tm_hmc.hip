// tm_hmc.hip -- one Hamiltonian Monte Carlo step on the device.
//
// The step of trimem's HMC sampler (src/trimem/mc/hmc.py:133-158): momenta
// p ~ N(0, m T), a velocity-Verlet trajectory of L steps
// (hmc.py:14-26), then a Metropolis test on the FP64 Hamiltonian
// H = E(x) + 0.5 p.p / m.  Positions, momenta and forces stay in HBM; the
// energy passes of tm_energy.hip run between the element-wise kernels
// below, and the decision is taken on the device, so a step needs no host
// synchronisation.
//
// Element-wise arithmetic is written in the operation order of hmc.py:21,23
// (x + (p dt + 0.5 a dt^2) / m and p + 0.5 (a + an) dt) and the library is
// built without FP contraction, so a trajectory equals, bit for bit, the
// same trajectory computed on the host from the same gradients.
#include "tm_internal.h"
#include "tm_rng.h"

namespace tmh {

constexpr int kHmcThreads = 256;

int hmc_blocks(int n3) { return n3 > 0 ? (n3 + kHmcThreads - 1) / kHmcThreads : 1; }

// deterministic block sum (fixed butterfly, then waves in order); the result
// is valid in thread 0
__device__ __forceinline__ double hmc_block_sum(double v, double* s_w)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += s_w[w];
    __syncthreads();
    return t;
}

__device__ __forceinline__ uint64_t hmc_key(uint64_t seed, uint64_t step)
{
    return splitmix64(seed ^ splitmix64(step + 0x632be59bd9b4e019ull));
}

// standard normal for element k (user vertex id * 3 + component): the two
// Box-Muller outputs of pair k/2 (Box & Muller, Ann. Math. Stat. 29, 1958)
__device__ __forceinline__ double hmc_normal(uint64_t key, uint64_t k)
{
    const uint64_t pair = k >> 1;
    const double u1 = (double)((rng_draw(key, 2, pair) >> 11) + 1) * (1.0 / 9007199254740992.0);    // (0, 1]
    const double u2 = rng_unit(rng_draw(key, 3, pair));
    const double r = sqrt(-2.0 * log(u1));
    const double t = 6.283185307179586 * u2;
    return (k & 1) ? r * sin(t) : r * cos(t);
}

__device__ __forceinline__ void begin_body(int n3, const double* x, double* x0, double* p, const int* perm, int draw,
                                           uint64_t key, double sigma, double* kpart, int lb)
{
    __shared__ double s_w[kHmcThreads / 64];
    const int e = lb * blockDim.x + threadIdx.x;
    double pp = 0.0;
    if (e < n3) {
        x0[e] = x[e];
        double pe;
        if (draw) {
            const int i = e / 3, c = e - 3 * i;
            pe = hmc_normal(key, (uint64_t)perm[i] * 3 + c) * sigma;
            p[e] = pe;
        } else {
            pe = p[e];
        }
        pp = pe * pe;
    }
    const double t = hmc_block_sum(pp, s_w);
    if (threadIdx.x == 0) kpart[lb] = t;
}

// After a force evaluation g = dE/dx at the current x (an = -g):
//   kKick:  p = p + 0.5 (a + an) dt      (hmc.py:23)
//   always: a = an                        (hmc.py:24)
//   kDrift: x = x + (p dt + 0.5 an dt^2) / m   (hmc.py:21, next step)
//   kLast:  partial sums of p.p
//   kFirst: record E(x0) and the global properties at x0
// g may alias x (harmonic test model): each thread reads its g before it
// writes its x.
__device__ __forceinline__ void leap_body(int n3, double* x, double* p, double* a, const double* g, double dt,
                                          double dt2, double mass, int mode, const Globals* globals, HmcState* st,
                                          double* kpart, int lb)
{
    __shared__ double s_w[kHmcThreads / 64];
    const int e = lb * blockDim.x + threadIdx.x;
    double pp = 0.0;
    if (e < n3) {
        const double an = -g[e];
        double pe = p[e];
        if (mode & kHmcKick) {
            pe = pe + 0.5 * (a[e] + an) * dt;
            p[e] = pe;
        }
        a[e] = an;
        if (mode & kHmcDrift) x[e] = x[e] + (pe * dt + 0.5 * an * dt2) / mass;
        pp = pe * pe;
    }
    if (mode & kHmcLast) {
        const double t = hmc_block_sum(pp, s_w);
        if (threadIdx.x == 0) kpart[lb] = t;
    }
    if ((mode & kHmcFirst) && globals && lb == 0 && threadIdx.x == 0) {
        st->energy0 = globals->energy;
        for (int k = 0; k < 6; k++) st->props0[k] = globals->props[k];
    }
}

// Metropolis decision (hmc.py:146-155) in FP64, one block.  Harmonic model
// (globals == null): E = 0.5 x.x of the start and end points (the sampler
// test's energy, tests/test_sampling.py:40-44); else E(x0) as recorded and
// E(xn) from the last evaluation.
__device__ __forceinline__ void accept_body(int nblk, const double* kpart0, const double* kpart1, int n3,
                                            const double* x0, const double* x, const Globals* globals, double mass,
                                            double temperature, double uniform, uint64_t key, HmcState* st)
{
    __shared__ double s_w[kHmcThreads / 64];
    const bool harmonic = globals == nullptr;
    double k0 = 0.0, k1 = 0.0, e0 = 0.0, e1 = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        k0 += kpart0[i];
        k1 += kpart1[i];
    }
    if (harmonic)
        for (int i = threadIdx.x; i < n3; i += blockDim.x) {
            e0 += x0[i] * x0[i];
            e1 += x[i] * x[i];
        }
    k0 = hmc_block_sum(k0, s_w);
    k1 = hmc_block_sum(k1, s_w);
    e0 = hmc_block_sum(e0, s_w);
    e1 = hmc_block_sum(e1, s_w);
    if (threadIdx.x != 0) return;
    const double E0 = harmonic ? 0.5 * e0 : st->energy0;
    const double E1 = harmonic ? 0.5 * e1 : globals->energy;
    const double K0 = 0.5 * k0 / mass;
    const double K1 = 0.5 * k1 / mass;
    const double dh = ((E1 + K1) - (E0 + K0)) / temperature;
    const double alpha = dh <= 0.0 ? 1.0 : exp(-dh);
    const double u = uniform >= 0.0 ? uniform : rng_unit(rng_draw(key, 4, 0));
    const int acc = u <= alpha ? 1 : 0;
    st->energy0 = E0;
    st->energy1 = E1;
    st->kinetic0 = K0;
    st->kinetic1 = K1;
    st->dh = dh;
    st->uniform = u;
    st->accepted = acc;
    if (!harmonic)
        for (int k = 0; k < 6; k++) st->props[k] = acc ? globals->props[k] : st->props0[k];
    st->n_accepted += (unsigned long long)acc;
    st->n_steps += 1ull;
}

__device__ __forceinline__ void restore_body(int n3, double* x, const double* x0, const HmcState* st, int lb)
{
    const int e = lb * blockDim.x + threadIdx.x;
    if (e < n3 && !st->accepted) x[e] = x0[e];
}

// ------------------------------------------------------ one mesh -----
__global__ void __launch_bounds__(kHmcThreads) k_hmc_begin(int n3, const double* x, double* x0, double* p,
                                                           const int* perm, int draw, uint64_t seed, uint64_t step,
                                                           double sigma, double* kpart)
{
    begin_body(n3, x, x0, p, perm, draw, hmc_key(seed, step), sigma, kpart, blockIdx.x);
}

__global__ void __launch_bounds__(kHmcThreads) k_hmc_leap(int n3, double* x, double* p, double* a, const double* g,
                                                          double dt, double dt2, double mass, int mode,
                                                          const Globals* globals, HmcState* st, double* kpart)
{
    leap_body(n3, x, p, a, g, dt, dt2, mass, mode, globals, st, kpart, blockIdx.x);
}

__global__ void __launch_bounds__(kHmcThreads) k_hmc_accept(int nblk, const double* kpart0, const double* kpart1,
                                                            int n3, const double* x0, const double* x,
                                                            const Globals* globals, double mass, double temperature,
                                                            double uniform, uint64_t seed, uint64_t step,
                                                            HmcState* st)
{
    accept_body(nblk, kpart0, kpart1, n3, x0, x, globals, mass, temperature, uniform, hmc_key(seed, step), st);
}

__global__ void __launch_bounds__(kHmcThreads) k_hmc_restore(int n3, double* x, const double* x0, const HmcState* st)
{
    restore_body(n3, x, x0, st, blockIdx.x);
}

// ---------------------------------------------------- replica groups -----
// block b of a grouped launch is local block b - block0[r] of replica r
__device__ __forceinline__ int hmc_replica(const int* __restrict__ block0, int R, int b)
{
    int lo = 0, hi = R - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (block0[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(kHmcThreads) k_hmc_begin_grouped(const HmcArgs* __restrict__ H,
                                                                   const int* __restrict__ block0, int R)
{
    const int r = hmc_replica(block0, R, blockIdx.x);
    const HmcArgs& h = H[r];
    begin_body(h.n3, h.x, h.x0, h.p, h.perm, 1, hmc_key(h.seed, h.st->n_steps), sqrt(h.mass * h.temperature),
               h.kpart0, blockIdx.x - block0[r]);
}

__global__ void __launch_bounds__(kHmcThreads) k_hmc_leap_grouped(const HmcArgs* __restrict__ H,
                                                                  const int* __restrict__ block0, int R, int mode)
{
    const int r = hmc_replica(block0, R, blockIdx.x);
    const HmcArgs& h = H[r];
    leap_body(h.n3, h.x, h.p, h.a, h.grad, h.dt, h.dt2, h.mass, mode, h.globals, h.st, h.kpart1,
              blockIdx.x - block0[r]);
}

__global__ void __launch_bounds__(kHmcThreads) k_hmc_accept_grouped(const HmcArgs* __restrict__ H)
{
    const HmcArgs& h = H[blockIdx.x];
    accept_body(h.nblocks, h.kpart0, h.kpart1, h.n3, h.x0, h.x, h.globals, h.mass, h.temperature, -1.0,
                hmc_key(h.seed, h.st->n_steps), h.st);
}

__global__ void __launch_bounds__(kHmcThreads) k_hmc_restore_grouped(const HmcArgs* __restrict__ H,
                                                                     const int* __restrict__ block0, int R)
{
    const int r = hmc_replica(block0, R, blockIdx.x);
    const HmcArgs& h = H[r];
    restore_body(h.n3, h.x, h.x0, h.st, blockIdx.x - block0[r]);
}

// Replica-exchange step over a kappa_b ladder (one block): even rounds pair
// slots (0,1), (2,3), ..., odd rounds (1,2), (3,4), ...  The energy is
// kappa_b B + (terms that stay with the replica), so exchanging the ladder
// values of the replicas a (slot s) and b (slot s+1) changes the total by
// (k[s+1] - k[s]) (B_a - B_b).
__global__ void __launch_bounds__(1024) k_pt_swap(PassArgs* A, const HmcArgs* __restrict__ H, int* owner,
                                                  const double* __restrict__ ladder, int R, uint64_t seed,
                                                  unsigned long long* round, unsigned long long* swaps_accepted,
                                                  unsigned long long* swaps_tried)
{
    const unsigned long long rd = *round;
    const int parity = (int)(rd & 1ull);
    for (int i = threadIdx.x; 2 * i + parity + 1 < R; i += blockDim.x) {
        const int s = 2 * i + parity;
        const int ra = owner[s], rb = owner[s + 1];
        const double dE = (ladder[s + 1] - ladder[s]) * (H[ra].st->props[3] - H[rb].st->props[3]);
        const double alpha = dE <= 0.0 ? 1.0 : exp(-dE / H[ra].temperature);
        const double u = rng_unit(rng_draw(seed, 5, rd * (unsigned long long)R + s));
        swaps_tried[s] += 1ull;
        if (u <= alpha) {
            owner[s] = rb;
            owner[s + 1] = ra;
            A[ra].p.kappa_b = ladder[s + 1];
            A[rb].p.kappa_b = ladder[s];
            swaps_accepted[s] += 1ull;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) *round = rd + 1ull;
}

void launch_hmc_begin(int n3, const double* x, double* x0, double* p, const int* perm, int draw, uint64_t seed,
                      uint64_t step, double sigma, double* kpart, hipStream_t s)
{
    hipLaunchKernelGGL(k_hmc_begin, dim3(hmc_blocks(n3)), dim3(kHmcThreads), 0, s, n3, x, x0, p, perm, draw, seed,
                       step, sigma, kpart);
}

void launch_hmc_leap(int n3, double* x, double* p, double* a, const double* g, double dt, double dt2, double mass,
                     int mode, const Globals* globals, HmcState* st, double* kpart, hipStream_t s)
{
    hipLaunchKernelGGL(k_hmc_leap, dim3(hmc_blocks(n3)), dim3(kHmcThreads), 0, s, n3, x, p, a, g, dt, dt2, mass,
                       mode, globals, st, kpart);
}

void launch_hmc_accept(int n3, const double* kpart0, const double* kpart1, const double* x0, const double* x,
                       int harmonic, const Globals* globals, double mass, double temperature, double uniform,
                       uint64_t seed, uint64_t step, HmcState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_hmc_accept, dim3(1), dim3(kHmcThreads), 0, s, hmc_blocks(n3), kpart0, kpart1, n3, x0, x,
                       harmonic ? nullptr : globals, mass, temperature, uniform, seed, step, st);
}

void launch_hmc_restore(int n3, double* x, const double* x0, const HmcState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_hmc_restore, dim3(hmc_blocks(n3)), dim3(kHmcThreads), 0, s, n3, x, x0, st);
}

void launch_hmc_begin_grouped(const HmcArgs* d, const int* d_block0, int R, int total_blocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_hmc_begin_grouped, dim3(total_blocks), dim3(kHmcThreads), 0, s, d, d_block0, R);
}

void launch_hmc_leap_grouped(const HmcArgs* d, const int* d_block0, int R, int total_blocks, int mode,
                             hipStream_t s)
{
    hipLaunchKernelGGL(k_hmc_leap_grouped, dim3(total_blocks), dim3(kHmcThreads), 0, s, d, d_block0, R, mode);
}

void launch_hmc_accept_grouped(const HmcArgs* d, int R, hipStream_t s)
{
    hipLaunchKernelGGL(k_hmc_accept_grouped, dim3(R), dim3(kHmcThreads), 0, s, d);
}

void launch_hmc_restore_grouped(const HmcArgs* d, const int* d_block0, int R, int total_blocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_hmc_restore_grouped, dim3(total_blocks), dim3(kHmcThreads), 0, s, d, d_block0, R);
}

void launch_pt_swap(PassArgs* d_args, const HmcArgs* d_hmc, int* owner, const double* ladder, int R, uint64_t seed,
                    unsigned long long* round, unsigned long long* swaps_accepted, unsigned long long* swaps_tried,
                    hipStream_t s)
{
    hipLaunchKernelGGL(k_pt_swap, dim3(1), dim3(1024), 0, s, d_args, d_hmc, owner, ladder, R, seed, round,
                       swaps_accepted, swaps_tried);
}

}    // namespace tmh
