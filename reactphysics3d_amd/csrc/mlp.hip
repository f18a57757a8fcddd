// Learned warm start network on the matrix cores, and its C ABI.
//
// Everything is computed transposed (features along rows, the 16 samples of a wavefront along
// columns) so that the FP32 result tile of one v_mfma_f32_16x16x16_bf16 is, lane for lane, the B
// operand of the next layer: lane l holds rows 4*(l/16)+i, column l%16 of D, and the B operand wants
// rows (k) 4*(l/16)+j, column l%16.  Only the weight gradients need the sample dimension as the
// contraction index; those go through LDS.
#include <hip/hip_runtime.h>

#include <mutex>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rp3d_mi355x.h"
#include "kcommon.h"
#include "mlp.h"
#include "world.h"

namespace rp3dg {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// fp32 -> bf16, round to nearest even (NaN kept quiet)
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ v4s pack4(float a, float b, float c, float d) {
    v4s r;
    r.x = (short)f2bf(a); r.y = (short)f2bf(b); r.z = (short)f2bf(c); r.w = (short)f2bf(d);
    return r;
}
__device__ __forceinline__ v4s ld4(const uint16_t* p) {   // 4 consecutive bf16, 8-byte aligned
    const uint2 u = *reinterpret_cast<const uint2*>(p);
    v4s r;
    r.x = (short)(u.x & 0xffffu); r.y = (short)(u.x >> 16); r.z = (short)(u.y & 0xffffu); r.w = (short)(u.y >> 16);
    return r;
}
__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

struct Fwd {
    v4f z1[8], z2[8];
    v4s h1[8], h2[8];
    v4f y;
};

// Forward pass of the 16 samples of one wavefront.  xb[kt]: lane l holds the normalised features
// 16*kt + 4*(l/16) + j (j = 0..3) of sample l%16.
__device__ __forceinline__ void mlp_fwd16(const MlpDev& M, const v4s xb[2], Fwd& F) {
    const int l = lane_id(), r = l & 15, g = l >> 4;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 2; kt++) acc = mfma(ld4(M.w1t + (16 * t + r) * kMlpInPad + 16 * kt + 4 * g), xb[kt], acc);
        float h[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float z = acc[i] + M.params[kP_B1 + 16 * t + 4 * g + i];
            F.z1[t][i] = z;
            h[i] = z * sigm(z);
        }
        F.h1[t] = pack4(h[0], h[1], h[2], h[3]);
    }
#pragma unroll
    for (int t = 0; t < 8; t++) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 8; kt++) acc = mfma(ld4(M.w2t + (16 * t + r) * kMlpHid + 16 * kt + 4 * g), F.h1[kt], acc);
        float h[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const float z = acc[i] + M.params[kP_B2 + 16 * t + 4 * g + i];
            F.z2[t][i] = z;
            h[i] = z * sigm(z);
        }
        F.h2[t] = pack4(h[0], h[1], h[2], h[3]);
    }
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < 8; kt++) acc = mfma(ld4(M.w3t + r * kMlpHid + 16 * kt + 4 * g), F.h2[kt], acc);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int o = 4 * g + i;
        F.y[i] = acc[i] + (o < kMlpOut ? M.params[kP_B3 + o] : 0.0f);
    }
}

// normalised, bf16-packed features of sample s (zero when s is out of range)
__device__ __forceinline__ void load_x(const MlpDev& M, const float* feats, int s, bool valid, v4s xb[2], float xn[8]) {
    const int g = lane_id() >> 4;
#pragma unroll
    for (int kt = 0; kt < 2; kt++) {
        float v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int f = 16 * kt + 4 * g + j;
            v[j] = (valid && f < kMlpIn) ? (feats[(size_t)s * kMlpIn + f] - M.mu[f]) * M.inv_sigma[f] : 0.0f;
            xn[4 * kt + j] = v[j];
        }
        xb[kt] = pack4(v[0], v[1], v[2], v[3]);
    }
}

__global__ __launch_bounds__(256) void mlp_infer(MlpDev M, const float* feats, const int* n_ptr, int n_cap, float* out,
                                                 const int* row_slot, float4* pred) {
    const int n = n_ptr ? min(*n_ptr, n_cap) : n_cap;
    const int l = lane_id(), r = l & 15, g = l >> 4;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const int nwaves = gridDim.x * blockDim.x / kWave;
    for (int base = wave * 16; base < n; base += nwaves * 16) {
        const int s = base + r;
        const bool valid = s < n;
        v4s xb[2];
        float xn[8];
        load_x(M, feats, s, valid, xb, xn);
        Fwd F;
        mlp_fwd16(M, xb, F);
        if (g == 0 && valid) {
            const float a = F.y[0] * M.out_scale[0], b = F.y[1] * M.out_scale[1], c = F.y[2] * M.out_scale[2];
            if (pred) pred[row_slot[s]] = make_float4(a, b, c, 0.0f);
            if (out) { out[3 * s] = a; out[3 * s + 1] = b; out[3 * s + 2] = c; }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Training: forward + backward, one chunk of 64 samples (4 wavefronts x 16) per block; the weight
// gradients are MFMA GEMMs over the chunk's samples staged in LDS.  Each block writes one partial
// gradient; the reduction sums the partials in block order (deterministic).
// ---------------------------------------------------------------------------------------------
constexpr int kChunk = 64;
constexpr int kGradTiles = 16 + 64 + 8;   // dW1 (2 x 8), dW2 (8 x 8), dW3 (8 x 1) tiles of 16 x 16
constexpr int kTilesPerWave = kGradTiles / 4;

struct GradLds {
    uint16_t x[kChunk][kMlpInPad];
    uint16_t h1[kChunk][kMlpHid];
    uint16_t h2[kChunk][kMlpHid];
    uint16_t dz1[kChunk][kMlpHid];
    uint16_t dz2[kChunk][kMlpHid];
    uint16_t dy[kChunk][kMlpOutPad];
    float red[4];
};

__device__ __forceinline__ void tile_of(int tidx, int& layer, int& ti, int& to) {
    if (tidx < 16) { layer = 0; ti = tidx >> 3; to = tidx & 7; }
    else if (tidx < 80) { layer = 1; ti = (tidx - 16) >> 3; to = (tidx - 16) & 7; }
    else { layer = 2; ti = tidx - 80; to = 0; }
}

__global__ __launch_bounds__(256) void mlp_grad(MlpDev M, const float* feats, const float* targets, const int* idx,
                                                int n, float* partial, float* loss_partial) {
    __shared__ GradLds L;
    const int tid = threadIdx.x, w = tid / kWave, l = lane_id(), r = l & 15, g = l >> 4;
    const float inv_n = 1.0f / (float)n;
    const int srow = 16 * w + r;
    const int s = blockIdx.x * kChunk + srow;
    const bool valid = s < n;
    const int si = valid ? (idx ? idx[s] : s) : 0;
    float loss = 0.0f;
    {
        v4s xb[2];
        float xn[8];
        load_x(M, feats, si, valid, xb, xn);
        Fwd F;
        mlp_fwd16(M, xb, F);
#pragma unroll
        for (int kt = 0; kt < 2; kt++)
#pragma unroll
            for (int j = 0; j < 4; j++) L.x[srow][16 * kt + 4 * g + j] = (uint16_t)xb[kt][j];
#pragma unroll
        for (int t = 0; t < 8; t++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                L.h1[srow][16 * t + 4 * g + i] = (uint16_t)F.h1[t][i];
                L.h2[srow][16 * t + 4 * g + i] = (uint16_t)F.h2[t][i];
            }
        // output error (lanes with g == 0 hold outputs 0..3 of sample r)
        float dyv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int o = 4 * g + i;
            float d = 0.0f;
            if (valid && o < kMlpOut) {
                const float t = targets[(size_t)si * kMlpOut + o] / M.out_scale[o];
                const float e = F.y[i] - t;
                loss += 0.5f * e * e;
                d = e * inv_n;
            }
            dyv[i] = d;
        }
        const v4s dyb = pack4(dyv[0], dyv[1], dyv[2], dyv[3]);
#pragma unroll
        for (int i = 0; i < 4; i++) L.dy[srow][4 * g + i] = (uint16_t)dyb[i];
        // dz2 = (W3 dy) * silu'(z2)
        v4s dz2b[8];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            v4f a = {0.f, 0.f, 0.f, 0.f};
            a = mfma(ld4(M.w3n + (16 * t + r) * kMlpOutPad + 4 * g), dyb, a);
            float d[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float z = F.z2[t][i], sg = sigm(z);
                d[i] = a[i] * (sg * (1.0f + z * (1.0f - sg)));
            }
            dz2b[t] = pack4(d[0], d[1], d[2], d[3]);
#pragma unroll
            for (int i = 0; i < 4; i++) L.dz2[srow][16 * t + 4 * g + i] = (uint16_t)dz2b[t][i];
        }
        // dz1 = (W2 dz2) * silu'(z1)
#pragma unroll
        for (int t = 0; t < 8; t++) {
            v4f a = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 8; kt++) a = mfma(ld4(M.w2n + (16 * t + r) * kMlpHid + 16 * kt + 4 * g), dz2b[kt], a);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float z = F.z1[t][i], sg = sigm(z);
                L.dz1[srow][16 * t + 4 * g + i] = f2bf(a[i] * (sg * (1.0f + z * (1.0f - sg))));
            }
        }
    }
    __syncthreads();
    // weight gradients: dW[in][out] = sum_s IN[s][in] * DOUT[s][out]
    float* P = partial + (size_t)blockIdx.x * kMlpParams;
    for (int k = 0; k < kTilesPerWave; k++) {
        int layer, ti, to;
        tile_of(w + 4 * k, layer, ti, to);
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < kChunk / 16; kc++) {
            v4s a, b;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int ss = 16 * kc + 4 * g + j;
                uint16_t av, bv;
                if (layer == 0) { av = L.x[ss][16 * ti + r]; bv = L.dz1[ss][16 * to + r]; }
                else if (layer == 1) { av = L.h1[ss][16 * ti + r]; bv = L.dz2[ss][16 * to + r]; }
                else { av = L.h2[ss][16 * ti + r]; bv = L.dy[ss][r]; }
                a[j] = (short)av;
                b[j] = (short)bv;
            }
            acc = mfma(a, b, acc);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int in = 16 * ti + 4 * g + i, out = 16 * to + r;
            if (layer == 0) { if (in < kMlpIn) P[kP_W1 + in * kMlpHid + out] = acc[i]; }
            else if (layer == 1) P[kP_W2 + in * kMlpHid + out] = acc[i];
            else if (out < kMlpOut) P[kP_W3 + in * kMlpOut + out] = acc[i];
        }
    }
    // bias gradients from the same bf16 errors, fixed sample order
    if (tid < kMlpHid) {
        float db = 0.0f;
        for (int ss = 0; ss < kChunk; ss++) db += bf2f(L.dz1[ss][tid]);
        P[kP_B1 + tid] = db;
        if (tid < kMlpOut) {
            float db3 = 0.0f;
            for (int ss = 0; ss < kChunk; ss++) db3 += bf2f(L.dy[ss][tid]);
            P[kP_B3 + tid] = db3;
        }
    } else {
        float db = 0.0f;
        for (int ss = 0; ss < kChunk; ss++) db += bf2f(L.dz2[ss][tid - kMlpHid]);
        P[kP_B2 + tid - kMlpHid] = db;
    }
    // loss: block sum in a fixed order
    loss = wave_sumf(loss);
    if (l == 0) L.red[w] = loss;
    __syncthreads();
    if (tid == 0) loss_partial[blockIdx.x] = (L.red[0] + L.red[1]) + (L.red[2] + L.red[3]);
}

__global__ void mlp_reduce_adam(float* params, const float* partial, int nparts, float* grad, float* m, float* v,
                                float lr, float beta1, float beta2, float eps, float bc1, float bc2) {
    RP3DG_GRID_STRIDE(p, kMlpParams) {
        float gsum = 0.0f;
        for (int b = 0; b < nparts; b++) gsum += partial[(size_t)b * kMlpParams + p];
        if (grad) grad[p] = gsum;
        if (lr > 0.0f) {
            const float mm = beta1 * m[p] + (1.0f - beta1) * gsum;
            const float vv = beta2 * v[p] + (1.0f - beta2) * gsum * gsum;
            m[p] = mm;
            v[p] = vv;
            params[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
        }
    }
}

__global__ void mlp_refresh(const float* P, uint16_t* w1t, uint16_t* w2t, uint16_t* w2n, uint16_t* w3t, uint16_t* w3n) {
    RP3DG_GRID_STRIDE(e, kMlpHid * kMlpHid) {
        const int a = e / kMlpHid, b = e % kMlpHid;   // a = row, b = column of a [128][128] view
        w2t[e] = f2bf(P[kP_W2 + b * kMlpHid + a]);
        w2n[e] = f2bf(P[kP_W2 + e]);
        if (e < kMlpHid * kMlpInPad) {
            const int h = e / kMlpInPad, i = e % kMlpInPad;
            w1t[e] = i < kMlpIn ? f2bf(P[kP_W1 + i * kMlpHid + h]) : 0;
        }
        if (e < kMlpOutPad * kMlpHid) {
            const int o = e / kMlpHid, i = e % kMlpHid;
            w3t[e] = o < kMlpOut ? f2bf(P[kP_W3 + i * kMlpOut + o]) : 0;
            const int i2 = e / kMlpOutPad, o2 = e % kMlpOutPad;
            w3n[e] = o2 < kMlpOut ? f2bf(P[kP_W3 + i2 * kMlpOut + o2]) : 0;
        }
    }
}

}  // namespace rp3dg

using namespace rp3dg;

// ---------------------------------------------------------------------------------------------
// host object
// ---------------------------------------------------------------------------------------------
struct rp3dg_mlp {
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *norm = nullptr;
    uint16_t *w1t = nullptr, *w2t = nullptr, *w2n = nullptr, *w3t = nullptr, *w3n = nullptr;
    float *partial = nullptr, *loss_partial = nullptr;
    int nparts_cap = 0;
    float* noalloc_scratch = nullptr;  // device address of the fixed scratch array (device of creation)
    int64_t adam_t = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;

    MlpDev view() const {
        MlpDev d;
        d.w1t = w1t; d.w2t = w2t; d.w2n = w2n; d.w3t = w3t; d.w3n = w3n;
        d.params = params;
        d.mu = norm;
        d.inv_sigma = norm + kMlpIn;
        d.out_scale = norm + 2 * kMlpIn;
        return d;
    }
    ~rp3dg_mlp() {
        for (void* p : allocs) hipFree(p);
        if (stream) hipStreamDestroy(stream);
    }
};

namespace rp3dg {
MlpDev mlp_view(const rp3dg_mlp* m) { return m->view(); }
}  // namespace rp3dg

#define MLPCHECK(x)                                                                         \
    do {                                                                                    \
        hipError_t e__ = (x);                                                               \
        if (e__ != hipSuccess) {                                                            \
            set_error(std::string("HIP error ") + hipGetErrorString(e__) + " at " #x);      \
            return RP3DG_EHIP;                                                              \
        }                                                                                   \
    } while (0)

template <typename T>
static T* mlp_alloc(rp3dg_mlp* m, size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(1, n) * sizeof(T)) != hipSuccess) return nullptr;
    m->allocs.push_back(p);
    return (T*)p;
}

static int mlp_refresh_copies(rp3dg_mlp* m) {
    mlp_refresh<<<64, 256, 0, m->stream>>>(m->params, m->w1t, m->w2t, m->w2n, m->w3t, m->w3n);
    MLPCHECK(hipGetLastError());
    return RP3DG_OK;
}

constexpr int kMaxParts = 1024;   // samples per gradient pass <= 65,536

// Allocation-free gradient (rp3dg_mlp_gradient_noalloc): the partial gradients of its blocks and their
// losses live in this fixed array instead of memory the library allocates.  Each GPU holds its own copy
// until the process ends; host code uses the address hipGetSymbolAddress returns for the current device
// (the array's name is not a device address on the host).  Every element a call reads is written by the
// same call first, so nothing needs clearing; calls that overlap on one GPU share it and race.
constexpr int kNoAllocParts = RP3DG_MLP_NOALLOC_MAX_SAMPLES / kChunk;
static_assert(kNoAllocParts * kChunk == RP3DG_MLP_NOALLOC_MAX_SAMPLES, "sample limit must be whole chunks");
static __device__ __align__(256) float scratch[(size_t)kNoAllocParts * kMlpParams + kNoAllocParts];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static float* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device.
static float* noalloc_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        set_error("no current HIP device for the allocation-free gradient");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
            return nullptr;
        }
        g_scratch_addr[dev] = (float*)ptr;
    }
    return g_scratch_addr[dev];
}

// Mean loss from the per-block losses, summed in block order in double precision (as the host does for
// rp3dg_mlp_gradient).
__global__ void mlp_loss_mean(const float* loss_partial, int nparts, int n, float* loss) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        for (int b = 0; b < nparts; b++) s += loss_partial[b];
        loss[0] = (float)(s / n);
    }
}

static int mlp_ensure_parts(rp3dg_mlp* m, int nparts) {
    if (nparts <= m->nparts_cap) return RP3DG_OK;
    m->partial = mlp_alloc<float>(m, (size_t)nparts * kMlpParams);
    m->loss_partial = mlp_alloc<float>(m, nparts);
    if (!m->partial || !m->loss_partial) { set_error("device allocation failed (mlp)"); return RP3DG_EHIP; }
    m->nparts_cap = nparts;
    return RP3DG_OK;
}

// Gradient (lr == 0) or one Adam step (lr > 0) over device samples.
static int mlp_grad_step(rp3dg_mlp* m, const float* dfeats, const float* dtargets, const int* didx, int n, float lr,
                         float* host_loss) {
    const int nparts = (n + kChunk - 1) / kChunk;   // one block per chunk of samples
    if (nparts > kMaxParts) { set_error("too many samples for one gradient pass"); return RP3DG_EINVAL; }
    int r = mlp_ensure_parts(m, nparts);
    if (r) return r;
    mlp_grad<<<nparts, 256, 0, m->stream>>>(m->view(), dfeats, dtargets, didx, n, m->partial, m->loss_partial);
    float bc1 = 1.0f, bc2 = 1.0f;
    const float b1 = 0.9f, b2 = 0.999f;
    if (lr > 0.0f) {
        m->adam_t++;
        bc1 = 1.0f - std::pow(b1, (float)m->adam_t);
        bc2 = 1.0f - std::pow(b2, (float)m->adam_t);
    }
    mlp_reduce_adam<<<(kMlpParams + 255) / 256, 256, 0, m->stream>>>(m->params, m->partial, nparts, m->grad, m->m, m->v,
                                                                      lr, b1, b2, 1e-8f, bc1, bc2);
    if (lr > 0.0f && (r = mlp_refresh_copies(m))) return r;
    MLPCHECK(hipGetLastError());
    if (host_loss) {
        std::vector<float> lp(nparts);
        MLPCHECK(hipMemcpyAsync(lp.data(), m->loss_partial, sizeof(float) * nparts, hipMemcpyDeviceToHost, m->stream));
        MLPCHECK(hipStreamSynchronize(m->stream));
        double s = 0.0;
        for (float x : lp) s += x;
        *host_loss = (float)(s / n);
    }
    return RP3DG_OK;
}

extern "C" {

rp3dg_mlp* rp3dg_mlp_create(uint64_t seed) {
    rp3dg_mlp* m = new rp3dg_mlp();
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) { set_error("stream"); delete m; return nullptr; }
    m->params = mlp_alloc<float>(m, kMlpParams);
    m->m = mlp_alloc<float>(m, kMlpParams);
    m->v = mlp_alloc<float>(m, kMlpParams);
    m->grad = mlp_alloc<float>(m, kMlpParams);
    m->norm = mlp_alloc<float>(m, 2 * kMlpIn + kMlpOut);
    m->w1t = mlp_alloc<uint16_t>(m, kMlpHid * kMlpInPad);
    m->w2t = mlp_alloc<uint16_t>(m, kMlpHid * kMlpHid);
    m->w2n = mlp_alloc<uint16_t>(m, kMlpHid * kMlpHid);
    m->w3t = mlp_alloc<uint16_t>(m, kMlpOutPad * kMlpHid);
    m->w3n = mlp_alloc<uint16_t>(m, kMlpHid * kMlpOutPad);
    if (!m->w3n || mlp_ensure_parts(m, 16)) { set_error("device allocation failed (mlp)"); delete m; return nullptr; }
    m->noalloc_scratch = noalloc_scratch();
    if (!m->noalloc_scratch) { delete m; return nullptr; }
    // Glorot-uniform weights from a splitmix64 stream, zero biases, identity normalisation
    std::vector<float> P(kMlpParams, 0.0f);
    uint64_t st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (float)(z >> 40) / 16777216.0f;
    };
    auto fill = [&](int off, int fin, int fout) {
        const float a = std::sqrt(6.0f / (float)(fin + fout));
        for (int k = 0; k < fin * fout; k++) P[off + k] = (2.0f * next() - 1.0f) * a;
    };
    fill(kP_W1, kMlpIn, kMlpHid);
    fill(kP_W2, kMlpHid, kMlpHid);
    fill(kP_W3, kMlpHid, kMlpOut);
    std::vector<float> norm(2 * kMlpIn + kMlpOut, 1.0f);
    for (int k = 0; k < kMlpIn; k++) norm[k] = 0.0f;
    std::vector<float> zero(kMlpParams, 0.0f);
    if (hipMemcpy(m->params, P.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->m, zero.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->v, zero.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->norm, norm.data(), sizeof(float) * norm.size(), hipMemcpyHostToDevice) != hipSuccess ||
        mlp_refresh_copies(m) != RP3DG_OK || hipStreamSynchronize(m->stream) != hipSuccess) {
        set_error("mlp initialisation failed");
        delete m;
        return nullptr;
    }
    return m;
}

void rp3dg_mlp_destroy(rp3dg_mlp* m) { delete m; }

int rp3dg_mlp_set_params(rp3dg_mlp* m, const float* p, int64_t n) {
    if (!m || !p || n != kMlpParams) { set_error("mlp params: expected 20099 values"); return RP3DG_EINVAL; }
    MLPCHECK(hipMemcpyAsync(m->params, p, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, m->stream));
    MLPCHECK(hipMemsetAsync(m->m, 0, sizeof(float) * kMlpParams, m->stream));
    MLPCHECK(hipMemsetAsync(m->v, 0, sizeof(float) * kMlpParams, m->stream));
    m->adam_t = 0;
    const int r = mlp_refresh_copies(m);
    if (r) return r;
    MLPCHECK(hipStreamSynchronize(m->stream));
    return RP3DG_OK;
}

int rp3dg_mlp_get_params(rp3dg_mlp* m, float* p, int64_t n) {
    if (!m || !p || n < kMlpParams) { set_error("mlp params: buffer too small"); return RP3DG_EINVAL; }
    MLPCHECK(hipMemcpyAsync(p, m->params, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->stream));
    MLPCHECK(hipStreamSynchronize(m->stream));
    return RP3DG_OK;
}

int rp3dg_mlp_set_normalization(rp3dg_mlp* m, const float* mu, const float* inv_sigma, const float* out_scale) {
    if (!m || !mu || !inv_sigma || !out_scale) return RP3DG_EINVAL;
    std::vector<float> norm(2 * kMlpIn + kMlpOut);
    for (int k = 0; k < kMlpIn; k++) { norm[k] = mu[k]; norm[kMlpIn + k] = inv_sigma[k]; }
    for (int k = 0; k < kMlpOut; k++) {
        if (!(out_scale[k] > 0.0f)) { set_error("output scales must be > 0"); return RP3DG_EINVAL; }
        norm[2 * kMlpIn + k] = out_scale[k];
    }
    MLPCHECK(hipMemcpyAsync(m->norm, norm.data(), sizeof(float) * norm.size(), hipMemcpyHostToDevice, m->stream));
    MLPCHECK(hipStreamSynchronize(m->stream));
    return RP3DG_OK;
}

static int upload_samples(rp3dg_mlp* m, const float* feats, const float* targets, int64_t n, float** df, float** dt) {
    *df = nullptr;
    *dt = nullptr;
    MLPCHECK(hipMalloc(df, sizeof(float) * kMlpIn * std::max<int64_t>(1, n)));
    if (targets) MLPCHECK(hipMalloc(dt, sizeof(float) * kMlpOut * std::max<int64_t>(1, n)));
    MLPCHECK(hipMemcpyAsync(*df, feats, sizeof(float) * kMlpIn * n, hipMemcpyHostToDevice, m->stream));
    if (targets) MLPCHECK(hipMemcpyAsync(*dt, targets, sizeof(float) * kMlpOut * n, hipMemcpyHostToDevice, m->stream));
    return RP3DG_OK;
}

int rp3dg_mlp_forward(rp3dg_mlp* m, const float* feats, int64_t n, float* out) {
    if (!m || n < 0 || n > (1 << 28) || (n && (!feats || !out))) return RP3DG_EINVAL;
    if (n == 0) return RP3DG_OK;
    float *df, *dt, *dout = nullptr;
    int r = upload_samples(m, feats, nullptr, n, &df, &dt);
    if (!r && hipMalloc(&dout, sizeof(float) * kMlpOut * n) != hipSuccess) { set_error("hipMalloc"); r = RP3DG_EHIP; }
    if (!r) {
        const int grid = (int)std::min<int64_t>(4096, (n + 63) / 64);
        mlp_infer<<<grid, 256, 0, m->stream>>>(m->view(), df, nullptr, (int)n, dout, nullptr, nullptr);
        if (hipMemcpyAsync(out, dout, sizeof(float) * kMlpOut * n, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess) { set_error("mlp forward failed"); r = RP3DG_EHIP; }
    }
    hipFree(df);
    hipFree(dout);
    return r;
}

int rp3dg_mlp_gradient(rp3dg_mlp* m, const float* feats, const float* targets, int64_t n, float* grad, float* loss) {
    if (!m || n <= 0 || n > kMaxParts * kChunk || !feats || !targets || !grad) return RP3DG_EINVAL;
    float *df, *dt;
    int r = upload_samples(m, feats, targets, n, &df, &dt);
    if (!r) r = mlp_grad_step(m, df, dt, nullptr, (int)n, 0.0f, loss);
    if (!r && (hipMemcpyAsync(grad, m->grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, m->stream) != hipSuccess ||
               hipStreamSynchronize(m->stream) != hipSuccess)) { set_error("mlp gradient readback failed"); r = RP3DG_EHIP; }
    hipFree(df);
    hipFree(dt);
    return r;
}

int rp3dg_mlp_gradient_noalloc(rp3dg_mlp* m, const float* feats, const float* targets, int64_t n, float* grad,
                               float* loss) {
    if (!m || !feats || !targets || !grad || !loss) { set_error("mlp gradient (no alloc): null argument"); return RP3DG_EINVAL; }
    if (n <= 0 || n > RP3DG_MLP_NOALLOC_MAX_SAMPLES) {
        set_error("mlp gradient (no alloc): sample count must be in [1, " +
                  std::to_string(RP3DG_MLP_NOALLOC_MAX_SAMPLES) + "], got " + std::to_string((long long)n));
        return RP3DG_EINVAL;
    }
    float* partial = m->noalloc_scratch;
    float* loss_partial = partial + (size_t)kNoAllocParts * kMlpParams;
    const int nparts = (int)((n + kChunk - 1) / kChunk);
    mlp_grad<<<nparts, 256, 0, m->stream>>>(m->view(), feats, targets, nullptr, (int)n, partial, loss_partial);
    mlp_reduce_adam<<<(kMlpParams + 255) / 256, 256, 0, m->stream>>>(m->params, partial, nparts, grad, nullptr, nullptr,
                                                                      0.0f, 0.9f, 0.999f, 1e-8f, 1.0f, 1.0f);
    mlp_loss_mean<<<1, 64, 0, m->stream>>>(loss_partial, nparts, (int)n, loss);
    MLPCHECK(hipGetLastError());
    MLPCHECK(hipStreamSynchronize(m->stream));
    return RP3DG_OK;
}

int rp3dg_mlp_train(rp3dg_mlp* m, const float* feats, const float* targets, int64_t n, int32_t epochs,
                    int32_t minibatch, float lr, uint64_t seed, float* last_loss) {
    if (!m || n <= 0 || n > (1 << 28) || !feats || !targets || epochs < 0 || minibatch < 16 ||
        minibatch > kMaxParts * kChunk || !(lr > 0.0f))
        return RP3DG_EINVAL;
    float *df, *dt;
    int* didx = nullptr;
    int r = upload_samples(m, feats, targets, n, &df, &dt);
    if (!r && hipMalloc(&didx, sizeof(int) * n) != hipSuccess) { set_error("hipMalloc"); r = RP3DG_EHIP; }
    std::vector<int> perm(n);
    uint64_t st = seed;
    auto next = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    float loss = 0.0f;
    for (int e = 0; e < epochs && !r; e++) {
        for (int64_t i = 0; i < n; i++) perm[i] = (int)i;
        for (int64_t i = n - 1; i > 0; i--) std::swap(perm[i], perm[next() % (uint64_t)(i + 1)]);
        if (hipMemcpyAsync(didx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, m->stream) != hipSuccess) {
            set_error("upload");
            r = RP3DG_EHIP;
            break;
        }
        for (int64_t b = 0; b < n && !r; b += minibatch) {
            const int nb = (int)std::min<int64_t>(minibatch, n - b);
            const bool last = b + minibatch >= n;
            r = mlp_grad_step(m, df, dt, didx + b, nb, lr, last ? &loss : nullptr);
        }
    }
    if (!r && hipStreamSynchronize(m->stream) != hipSuccess) { set_error("mlp train failed"); r = RP3DG_EHIP; }
    if (last_loss) *last_loss = loss;
    hipFree(df);
    hipFree(dt);
    hipFree(didx);
    return r;
}

}  // extern "C"
