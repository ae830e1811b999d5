// tm_surrogate.h -- layout and launchers of the surrogate force network
// (tm_surrogate.hip).
#pragma once

#include "tm_common.h"

namespace tmh {

constexpr int kSgW = 64;          // hidden width
constexpr int kSgLayers = 3;      // message-passing layers
constexpr int kSgFeat = 8;        // vertex input features
constexpr int kSgFeatPad = 16;    // padded to one MFMA k-step
constexpr int kSgSlots = 10;      // largest valence handled
constexpr int kSgRowsPerBlock = 256;

// Parameter vector (FP32): W_in [16][64], b_in [64], per layer U [64][64],
// W [64][64], V [64][64] (W and V adjacent: [W; V] is one [128][64] block),
// E [2][64], b [64]; then w_s [64], b_s, w_t [64], b_t.
constexpr int kSgOffWin = 0;
constexpr int kSgOffBin = kSgFeatPad * kSgW;
constexpr int kSgOffLayer0 = kSgOffBin + kSgW;
constexpr int kSgLayerSize = 3 * kSgW * kSgW + 2 * kSgW + kSgW;
__host__ __device__ constexpr int sg_off_U(int l) { return kSgOffLayer0 + l * kSgLayerSize; }
__host__ __device__ constexpr int sg_off_W(int l) { return sg_off_U(l) + kSgW * kSgW; }
__host__ __device__ constexpr int sg_off_V(int l) { return sg_off_W(l) + kSgW * kSgW; }
__host__ __device__ constexpr int sg_off_E(int l) { return sg_off_V(l) + kSgW * kSgW; }
__host__ __device__ constexpr int sg_off_b(int l) { return sg_off_E(l) + 2 * kSgW; }
constexpr int kSgOffWs = kSgOffLayer0 + kSgLayers * kSgLayerSize;
constexpr int kSgOffBs = kSgOffWs + kSgW;
constexpr int kSgOffWt = kSgOffBs + 1;
constexpr int kSgOffBt = kSgOffWt + kSgW;
constexpr int kSgParams = kSgOffBt + 1;
constexpr int kSgOutGrad = 2 * kSgW + 2;    // w_s, b_s, w_t, b_t

// input normalisation (fixed at creation) and the mean edge length
struct SgNorm {
    double mu[kSgFeat];
    double sig[kSgFeat];
    double lbar;
    double ref[3];    // reference area, volume, curvature of the energy
};

// per-spoke and per-vertex geometry written by the feature kernel:
// eps [2 * kSgSlots * N], ehat [3 * kSgSlots * N] (ring layout k * N + v),
// nu [3 N]
struct SgGeom {
    float* eps;
    float* ehat;
    float* nu;
};

// bf16 weights as GEMM B operands
struct SgPacked {
    short* in_t;                  // [64][16]
    short* u_t[kSgLayers];        // [64][64]  U^T
    short* wv_t[kSgLayers];       // [64][128] [W; V]^T
    short* u_n[kSgLayers];        // [64][64]  U
    short* w_n[kSgLayers];        // W
    short* v_n[kSgLayers];        // V
};

// activations and work arrays (FP32, N x 64 unless noted)
struct SgBuffers {
    float* f0;                    // [N][16]
    float* Z[kSgLayers + 1];
    float* H[kSgLayers + 1];
    float* P[kSgLayers];
    float* M[kSgLayers];
    float* dH;
    float* dZ;
    float* dP;
    float* dM;
    float* ghat;                  // [N][3]
    float* loss_partial;          // [ceil(N/4)]
    float* loss;                  // [1]
    float* partial;               // reduction scratch
};
size_t sg_partial_floats(int nv);

void sg_pack(const float* theta, const SgPacked& pk, hipStream_t s);
void sg_features(const MeshView& m, const double* vprops, const Globals* g, const SgNorm& nrm, const SgGeom& geo,
                 float* f0, int* err, hipStream_t s);
void sg_twin(const MeshView& m, int* twin, hipStream_t s);
void sg_forward(const MeshView& m, const SgBuffers& b, const SgGeom& geo, const float* theta, const SgPacked& pk,
                float sigma_F, const double* target, double* neg_force, hipStream_t s);
void sg_backward(const MeshView& m, const SgBuffers& b, const SgGeom& geo, const float* theta, const SgPacked& pk,
                 const int* twin, float* grad, hipStream_t s);
void sg_adam(int n, float* theta, const float* grad, float* m1, float* m2, float lr, int step, hipStream_t s);

}    // namespace tmh
