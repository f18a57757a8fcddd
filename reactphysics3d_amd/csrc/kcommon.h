// Small device utilities shared by the step kernels (wave-64 scans, float4 <-> V3).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rp3d_internal.h"

namespace rp3dg {

constexpr int kWave = 64;

__device__ __forceinline__ V3 ld3(const float4* a, int i) { float4 t = a[i]; return v3(t.x, t.y, t.z); }
__device__ __forceinline__ Q4 ldq(const float4* a, int i) { float4 t = a[i]; return q4(t.x, t.y, t.z, t.w); }
__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
// stores x, y, z of a float4 slot and leaves its w lane as it is
__device__ __forceinline__ void st3(float4* a, int i, V3 v) {
    float4& r = a[i];
    r.x = v.x;
    r.y = v.y;
    r.z = v.z;
}
__device__ __forceinline__ float4 f4q(Q4 q) { return make_float4(q.x, q.y, q.z, q.w); }
__device__ __forceinline__ V3 xyz(float4 t) { return v3(t.x, t.y, t.z); }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Inclusive prefix sum across the 64 lanes of a wavefront.
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        int t = __shfl_up(v, off, kWave);
        if (lane >= off) v += t;
    }
    return v;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}
__device__ __forceinline__ float wave_sumf(float v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const float o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// Block-wide exclusive scan (blockDim.x multiple of 64, <= 1024); returns the exclusive
// prefix of v and writes the block total to *total.  smem must hold 17 ints.
__device__ __forceinline__ int block_excl_scan(int v, int* smem, int* total) {
    const int lane = lane_id();
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    int inc = wave_incl_scan(v);
    if (lane == kWave - 1) smem[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        int s = lane < nw ? smem[lane] : 0;
        int si = wave_incl_scan(s);
        if (lane < nw) smem[lane] = si - s;
        if (lane == nw - 1) smem[16] = si;
    }
    __syncthreads();
    const int res = inc - v + smem[wid];
    *total = smem[16];
    __syncthreads();
    return res;
}

#define RP3DG_GRID_STRIDE(i, n) for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)

}  // namespace rp3dg
