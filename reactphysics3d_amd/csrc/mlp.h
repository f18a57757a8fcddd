// Learned warm start: a 24 -> 128 -> 128 -> 3 SiLU network evaluated and trained in BF16 on the
// matrix cores (v_mfma_f32_16x16x16_bf16, FP32 accumulation).  It has no counterpart in the
// reference, whose solver starts new contacts from zero impulse (ContactSolverSystem.cpp:344-376).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

struct rp3dg_mlp;

namespace rp3dg {

constexpr int kMlpIn = 24;
constexpr int kMlpInPad = 32;
constexpr int kMlpHid = 128;
constexpr int kMlpOut = 3;
constexpr int kMlpOutPad = 16;

// fp32 master parameters, flat: W1[in][hid], b1, W2[hid][hid], b2, W3[hid][out], b3
constexpr int kP_W1 = 0;
constexpr int kP_B1 = kP_W1 + kMlpIn * kMlpHid;
constexpr int kP_W2 = kP_B1 + kMlpHid;
constexpr int kP_B2 = kP_W2 + kMlpHid * kMlpHid;
constexpr int kP_W3 = kP_B2 + kMlpHid;
constexpr int kP_B3 = kP_W3 + kMlpHid * kMlpOut;
constexpr int kMlpParams = kP_B3 + kMlpOut;   // 20,099

// Device view of a network: bf16 weight copies in the layouts the MFMA operands load from, fp32
// biases, input normalisation (x - mu) * inv_sigma and output scale.
struct MlpDev {
    const uint16_t* w1t;   // [hid][in_pad]   W1^T (zero padded)
    const uint16_t* w2t;   // [hid][hid]      W2^T
    const uint16_t* w2n;   // [hid][hid]      W2
    const uint16_t* w3t;   // [out_pad][hid]  W3^T (zero padded)
    const uint16_t* w3n;   // [hid][out_pad]  W3 (zero padded)
    const float* params;   // fp32 master (biases are read from here)
    const float* mu;       // [in]
    const float* inv_sigma;// [in]
    const float* out_scale;// [out]
};

// rows = feature vectors [n][kMlpIn]; out = impulses [n][kMlpOut] (already scaled), or, with row_slot,
// written as float4 (x, y, z, 0) to pred[row_slot[r]].
__global__ void mlp_infer(MlpDev M, const float* feats, const int* n_ptr, int n_cap, float* out, const int* row_slot,
                          float4* pred);
// Forward + backward over samples [0, n) (sample order idx[s] when idx is given); writes per-block
// partial gradients [gridDim.x][kMlpParams] and per-block loss sums.  Loss = 1/(2n) sum |y - t/s|^2.
__global__ void mlp_grad(MlpDev M, const float* feats, const float* targets, const int* idx, int n, float* partial,
                         float* loss_partial);
// Sums the partial gradients in block order; Adam step on the fp32 master (lr > 0) or plain gradient
// output (lr == 0, into grad); refreshes the bf16 copies.
__global__ void mlp_reduce_adam(float* params, const float* partial, int nparts, float* grad, float* m, float* v,
                                float lr, float beta1, float beta2, float eps, float bc1, float bc2);
MlpDev mlp_view(const rp3dg_mlp* m);

__global__ void mlp_refresh(const float* params, uint16_t* w1t, uint16_t* w2t, uint16_t* w2n, uint16_t* w3t,
                            uint16_t* w3n);

}  // namespace rp3dg
