// Broadphase of the MI355X step path.
//
// The reference keeps an incremental DynamicAABBTree of *fat* AABBs and only
// re-inserts a collider when its tight AABB leaves its fat AABB
// (DynamicAABBTree.cpp:186-214).  Because a fat AABB changes only on such a
// move, the reference pair set after computeBroadPhase() is exactly
//   { collider pairs whose current fat AABBs overlap (closed test, AABB.h:179) }
// filtered by different-body / category masks / no-collision body pairs
// (CollisionDetectionSystem.cpp:270-343).  We therefore keep the persistent
// fat AABB per collider (same FP32 refit as BroadPhaseSystem.cpp:155-185) and
// rebuild a linear BVH over them every step:
//   bp_bounds -> bp_morton -> radix sort (LDS, stable, 8-bit digits)
//   -> bp_karras (Karras 2012 radix tree) -> node boxes as unions over leaf ranges
//      (bp_leaves, bp_blocks, bp_internal; no inter-thread synchronisation)
//   -> bp_traverse (count, first 8 partners sorted into a slab) -> scan
//   -> bp_copy_slabs (slab -> per-collider sorted segment) + bp_traverse again
//      for the few colliders with more partners
//   -> bp_merge (carry orientation, temporal-coherence record and contact cache
//      of surviving pairs; orient new pairs like the reference).
#include <hip/hip_runtime.h>

#include "kcommon.h"
#include "rp3d_internal.h"
#include "refit_dev.h"

namespace rp3dg {

__global__ void bp_refit(DevBodies B, DevColliders C, const float4* hull_bounds, int init) {
    RP3DG_GRID_STRIDE(i, C.n) refit_one(B, C, hull_bounds, init, i);
}

// ---------------------------------------------------------------------------------------------
// Morton keys over fat-AABB centres; the world id sits above the 30 Morton bits so that every
// world occupies a contiguous key range and owns one subtree of the radix tree.
// ---------------------------------------------------------------------------------------------

__global__ void bp_bounds(DevColliders C, float4* partial) {
    __shared__ float smn[3][256], smx[3][256];
    V3 mn = v3(kDecimalLargest, kDecimalLargest, kDecimalLargest);
    V3 mx = v3(kDecimalSmallest, kDecimalSmallest, kDecimalSmallest);
    RP3DG_GRID_STRIDE(i, C.n) {
        const V3 a = ld3(C.amin, i), b = ld3(C.amax, i);
        const V3 c = v3((a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f);
        mn = vmin(mn, c);
        mx = vmax(mx, c);
    }
    const int t = threadIdx.x;
    smn[0][t] = mn.x; smn[1][t] = mn.y; smn[2][t] = mn.z;
    smx[0][t] = mx.x; smx[1][t] = mx.y; smx[2][t] = mx.z;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            for (int k = 0; k < 3; k++) {
                smn[k][t] = fminf(smn[k][t], smn[k][t + s]);
                smx[k][t] = fmaxf(smx[k][t], smx[k][t + s]);
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        partial[2 * blockIdx.x] = make_float4(smn[0][0], smn[1][0], smn[2][0], 0.f);
        partial[2 * blockIdx.x + 1] = make_float4(smx[0][0], smx[1][0], smx[2][0], 0.f);
    }
}

__device__ inline uint32_t expand10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void bp_morton(DevColliders C, const float4* partial, int nparts, uint64_t* keys, int* vals) {
    __shared__ float4 bmn, bmx;
    if (threadIdx.x == 0) {
        float4 mn = partial[0], mx = partial[1];
        for (int p = 1; p < nparts; p++) {
            const float4 a = partial[2 * p], b = partial[2 * p + 1];
            mn.x = fminf(mn.x, a.x); mn.y = fminf(mn.y, a.y); mn.z = fminf(mn.z, a.z);
            mx.x = fmaxf(mx.x, b.x); mx.y = fmaxf(mx.y, b.y); mx.z = fmaxf(mx.z, b.z);
        }
        bmn = mn;
        bmx = mx;
    }
    __syncthreads();
    const float sx = bmx.x > bmn.x ? 1023.0f / (bmx.x - bmn.x) : 0.f;
    const float sy = bmx.y > bmn.y ? 1023.0f / (bmx.y - bmn.y) : 0.f;
    const float sz = bmx.z > bmn.z ? 1023.0f / (bmx.z - bmn.z) : 0.f;
    RP3DG_GRID_STRIDE(i, C.n) {
        const V3 a = ld3(C.amin, i), b = ld3(C.amax, i);
        const float cx = (a.x + b.x) * 0.5f, cy = (a.y + b.y) * 0.5f, cz = (a.z + b.z) * 0.5f;
        const uint32_t qx = (uint32_t)fminf(fmaxf((cx - bmn.x) * sx, 0.f), 1023.f);
        const uint32_t qy = (uint32_t)fminf(fmaxf((cy - bmn.y) * sy, 0.f), 1023.f);
        const uint32_t qz = (uint32_t)fminf(fmaxf((cz - bmn.z) * sz, 0.f), 1023.f);
        const uint32_t m = (expand10(qx) << 2) | (expand10(qy) << 1) | expand10(qz);
        keys[i] = ((uint64_t)(uint32_t)C.world[i] << 30) | m;
        vals[i] = i;
    }
}

// ---------------------------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits.  Tiles of kSortTile keys per block; inside a tile keys are
// ranked by wave-level multisplit (8 ballots give the lanes sharing a digit) and cross-wave
// prefix counts in LDS, so equal digits keep their input order.
// ---------------------------------------------------------------------------------------------

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kSortThreads * kSortItems;

// n: key count, or its upper bound when nptr (the count in device memory) is given.
__global__ __launch_bounds__(kSortThreads) void rs_hist(const uint64_t* keys, int n, const int* nptr, int shift, int* hist,
                                                        int nblocks) {
    if (nptr) n = min(n, *nptr);
    __shared__ int h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
    for (int k = 0; k < kSortItems; k++) {
        const int i = base + k * kSortThreads + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xff], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__global__ __launch_bounds__(kSortThreads) void rs_scatter(const uint64_t* kin, const int* vin, uint64_t* kout,
                                                           int* vout, int n, const int* nptr, int shift, const int* offs,
                                                           int nblocks) {
    if (nptr) n = min(n, *nptr);
    __shared__ int base[256];
    __shared__ int wcnt[kSortThreads / kWave][256];
    const int t = threadIdx.x;
    base[t] = offs[t * nblocks + blockIdx.x];
    const int wid = t / kWave;
    const int blk = blockIdx.x * kSortTile;
    for (int k = 0; k < kSortItems; k++) {
        for (int w = 0; w < kSortThreads / kWave; w++) wcnt[w][t] = 0;
        __syncthreads();
        const int i = blk + k * kSortThreads + t;
        const bool valid = i < n;
        uint64_t key = valid ? kin[i] : 0;
        const int d = valid ? (int)((key >> shift) & 0xff) : 0;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1;
            const uint64_t bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        const int rank = __popcll(peers & lanemask_lt());
        if (valid && rank == 0) wcnt[wid][d] = __popcll(peers);
        __syncthreads();
        {  // exclusive prefix over waves for digit t, advance base
            int run = 0;
            for (int w = 0; w < kSortThreads / kWave; w++) {
                const int c = wcnt[w][t];
                wcnt[w][t] = run;
                run += c;
            }
            wcnt[0][t] += 0;
            __syncthreads();
            if (valid) {
                const int pos = base[d] + wcnt[wid][d] + rank;
                kout[pos] = key;
                vout[pos] = vin[i];
            }
            __syncthreads();
            base[t] += run;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------------------------
// Karras (2012) radix tree: internal nodes [0, n-1), leaf of sorted position s is node n-1+s.
// ---------------------------------------------------------------------------------------------

__device__ __forceinline__ int delta(const uint64_t* k, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

__global__ void bp_karras(const uint64_t* k, int n, int2* child, int* parent, int2* range) {
    RP3DG_GRID_STRIDE(i, n - 1) {
        const int d = (delta(k, n, i, i + 1) - delta(k, n, i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = delta(k, n, i, i - d);
        int lmax = 2;
        while (delta(k, n, i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (delta(k, n, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = delta(k, n, i, j);
        int s = 0;
        int t = l;
        do {
            t = (t + 1) >> 1;
            if (delta(k, n, i, i + (s + t) * d) > dnode) s += t;
        } while (t > 1);
        const int g = i + s * d + (d < 0 ? d : 0);
        const int lo = min(i, j), hi = max(i, j);
        const int left = (lo == g) ? (n - 1 + g) : g;
        const int right = (hi == g + 1) ? (n - 1 + g + 1) : (g + 1);
        child[i] = make_int2(left, right);
        parent[left] = i;
        parent[right] = i;
        range[i] = make_int2(lo, hi);
    }
}

// Synchronisation-free fit.  An internal node's box is the union of the fat AABBs of its leaf range
// [lo, hi] (range[i] from bp_karras).  Min / max are exact, so any union order gives the same bits
// as a bottom-up fit.  Unions over aligned blocks of 64, 4096 and 262144 sorted leaves are
// precomputed, so a range costs at most ~64 reads per level.
constexpr int kFitBlock = 64;

__device__ __forceinline__ void box_union(float4& mn, float4& mx, float4 bmn, float4 bmx) {
    mn = make_float4(fminf(mn.x, bmn.x), fminf(mn.y, bmn.y), fminf(mn.z, bmn.z), 0.f);
    mx = make_float4(fmaxf(mx.x, bmx.x), fmaxf(mx.y, bmx.y), fmaxf(mx.z, bmx.z), 0.f);
}

__device__ __forceinline__ void wave_box_union(float4& mn, float4& mx) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) {
        float4 omn, omx;
        omn.x = __shfl_xor(mn.x, off, kWave); omn.y = __shfl_xor(mn.y, off, kWave); omn.z = __shfl_xor(mn.z, off, kWave);
        omx.x = __shfl_xor(mx.x, off, kWave); omx.y = __shfl_xor(mx.y, off, kWave); omx.z = __shfl_xor(mx.z, off, kWave);
        box_union(mn, mx, omn, omx);
    }
}

// Leaves (node n-1+s = sorted collider s) and the unions of aligned blocks of 64 leaves.
__global__ void bp_leaves(DevColliders C, const int* sorted, int n, float4* nmin, float4* nmax, float4* bmin,
                          float4* bmax) {
    const int nb = (n + kWave - 1) / kWave;
    for (int base = (blockIdx.x * blockDim.x + threadIdx.x) & ~(kWave - 1); base < nb * kWave;
         base += gridDim.x * blockDim.x) {
        const int s = base + lane_id();
        float4 mn = make_float4(INFINITY, INFINITY, INFINITY, 0.f), mx = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
        if (s < n) {
            const int c = sorted[s];
            mn = C.amin[c];
            mx = C.amax[c];
            nmin[n - 1 + s] = mn;
            nmax[n - 1 + s] = mx;
        }
        wave_box_union(mn, mx);
        if (lane_id() == 0) {
            bmin[base / kWave] = mn;
            bmax[base / kWave] = mx;
        }
    }
}

// Unions of aligned groups of 64 entries of the level below.
__global__ void bp_blocks(const float4* imin, const float4* imax, int nin, float4* omin, float4* omax) {
    const int nout = (nin + kWave - 1) / kWave;
    for (int base = (blockIdx.x * blockDim.x + threadIdx.x) & ~(kWave - 1); base < nout * kWave;
         base += gridDim.x * blockDim.x) {
        const int s = base + lane_id();
        float4 mn = make_float4(INFINITY, INFINITY, INFINITY, 0.f), mx = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
        if (s < nin) { mn = imin[s]; mx = imax[s]; }
        wave_box_union(mn, mx);
        if (lane_id() == 0) {
            omin[base / kWave] = mn;
            omax[base / kWave] = mx;
        }
    }
}

// Internal nodes: union over the leaf range with the largest aligned blocks that fit.
__global__ void bp_internal(int n, const int2* range, float4* nmin, float4* nmax, const float4* b1min,
                            const float4* b1max, const float4* b2min, const float4* b2max, const float4* b3min,
                            const float4* b3max) {
    RP3DG_GRID_STRIDE(i, n - 1) {
        const int2 r = range[i];
        int s = r.x;
        const int e = r.y + 1;   // exclusive
        float4 mn = make_float4(INFINITY, INFINITY, INFINITY, 0.f), mx = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
        while (s < e) {
            if ((s & (262144 - 1)) == 0 && s + 262144 <= e) {
                box_union(mn, mx, b3min[s >> 18], b3max[s >> 18]);
                s += 262144;
            } else if ((s & (4096 - 1)) == 0 && s + 4096 <= e) {
                box_union(mn, mx, b2min[s >> 12], b2max[s >> 12]);
                s += 4096;
            } else if ((s & (kFitBlock - 1)) == 0 && s + kFitBlock <= e) {
                box_union(mn, mx, b1min[s >> 6], b1max[s >> 6]);
                s += kFitBlock;
            } else {
                box_union(mn, mx, nmin[n - 1 + s], nmax[n - 1 + s]);
                s += 1;
            }
        }
        nmin[i] = mn;
        nmax[i] = mx;
    }
}

// Root node of every world: the node whose leaf range is exactly the world's key range.
__global__ void bp_world_roots(const uint64_t* k, int n, const int2* range, int* root) {
    RP3DG_GRID_STRIDE(i, 2 * n - 1) {
        int lo, hi;
        if (i < n - 1) { const int2 r = range[i]; lo = r.x; hi = r.y; }
        else { lo = hi = i - (n - 1); }
        const uint32_t w = (uint32_t)(k[lo] >> 30);
        if ((uint32_t)(k[hi] >> 30) != w) continue;
        if (lo > 0 && (uint32_t)(k[lo - 1] >> 30) == w) continue;
        if (hi < n - 1 && (uint32_t)(k[hi + 1] >> 30) == w) continue;
        root[w] = i;
    }
}

__device__ __forceinline__ bool aabb_overlap(float4 amn, float4 amx, float4 bmn, float4 bmx) {
    // AABB::testCollision (AABB.h:179-187), closed intervals
    if (amx.x < bmn.x || bmx.x < amn.x) return false;
    if (amx.y < bmn.y || bmx.y < amn.y) return false;
    if (amx.z < bmn.z || bmx.z < amn.z) return false;
    return true;
}

// Body pairs that must not collide (joints with collideConnected = false): per-body partner lists
__device__ __forceinline__ bool no_collision(const NoColLists& L, int ba, int bb) {
    const int e = L.offs[ba + 1];
    for (int k = L.offs[ba]; k < e; k++)
        if (L.list[k] == bb) return true;
    return false;
}

__device__ __forceinline__ bool filters_pass(uint32_t fa, uint32_t fb) {
    const uint32_t cata = fa & 0xffffu, maska = fa >> 16, catb = fb & 0xffffu, maskb = fb >> 16;
    return (maska & catb) != 0 && (cata & maskb) != 0;
}

// Pair filter of CollisionDetectionSystem::updateOverlappingPairs (CD.cpp:284-331): different bodies, no
// joint excluding the pair, collision category / mask, at least one convex shape (:324-331)
__device__ __forceinline__ bool pair_allowed(const DevColliders& C, int a, int b, const NoColLists& L) {
    const int ba = C.body[a], bb = C.body[b];
    if (ba == bb) return false;
    if (no_collision(L, ba, bb)) return false;
    if (shape_is_concave(C.type[a]) && shape_is_concave(C.type[b])) return false;
    return filters_pass(C.filter[a], C.filter[b]);
}

constexpr int kTravStack = 64;   // traversal stack depth (per thread)

// Each overlapping pair is emitted by exactly one of its colliders (the pair key's high word):
//   moving vs moving -> the lower collider index; moving vs static -> the moving one;
//   static vs static -> the lower index, from the constant list built at creation (statics never
//   traverse: a floor's AABB overlaps most of the tree and would serialise one thread over it).
// Pass 0 counts, pass 1 writes each emitter's partner list sorted ascending into its segment, so the
// pair table is sorted by key.
// Pass 0 traverses once (colliders in Morton order, so neighbouring lanes walk similar paths),
// counts every collider's partners and keeps the first kTravSlab of them, sorted, in a
// per-collider slab.  Pass 1 (after the scan of the counts) runs over colliders in index order,
// copies the slabs into the emitters' segments of the pair-key table (coalesced), and traverses
// again only for colliders with more partners than the slab holds.
__global__ __launch_bounds__(kTravBlock) void bp_traverse(DevColliders C, const int* sorted, int n, const int2* child,
                                                          const float4* nmin, const float4* nmax, const int* root,
                                                          NoColLists nc, int* count, const int* offset,
                                                          uint64_t* out_keys, uint64_t* slab, int capacity, int pass,
                                                          int* walk_list, int* walk_count, int* status) {
    const int nw = pass ? *walk_count : n;
    RP3DG_GRID_STRIDE(s, nw) {
        const int c = pass ? walk_list[s] : sorted[s];
        const int base = pass ? offset[c] : 0;
        if (!pass && C.is_static[c]) {
            count[c] = C.ss_offs[c + 1] - C.ss_offs[c];
            continue;
        }
        uint64_t* sl = slab + (size_t)c * kTravSlab;
        const float4 amn = C.amin[c], amx = C.amax[c];
        int stack[kTravStack];
        int sp = 0;
        stack[sp++] = root[C.world[c]];
        int cnt = 0;
        while (sp > 0) {
            const int node = stack[--sp];
            if (!aabb_overlap(amn, amx, nmin[node], nmax[node])) continue;
            if (node >= n - 1) {
                const int o = sorted[node - (n - 1)];
                if (o == c) continue;
                if (!C.is_static[o] && o < c) continue;
                if (!pair_allowed(C, c, o, nc)) continue;
                const uint64_t key = ((uint64_t)(uint32_t)c << 32) | (uint32_t)o;
                if (pass) {
                    int pos = base + cnt;
                    if (pos < capacity) {
                        // insertion into the sorted segment
                        while (pos > base && out_keys[pos - 1] > key) {
                            out_keys[pos] = out_keys[pos - 1];
                            pos--;
                        }
                        out_keys[pos] = key;
                    } else {
                        atomicOr(status, STATUS_PAIR_OVERFLOW);
                    }
                } else if (cnt < kTravSlab) {
                    int pos = cnt;
                    while (pos > 0 && sl[pos - 1] > key) {
                        sl[pos] = sl[pos - 1];
                        pos--;
                    }
                    sl[pos] = key;
                }
                cnt++;
            } else {
                if (sp + 2 > kTravStack) { atomicOr(status, STATUS_STACK_OVERFLOW); break; }
                const int2 ch = child[node];
                stack[sp++] = ch.x;
                stack[sp++] = ch.y;
            }
        }
        if (!pass) {
            count[c] = cnt;
            if (cnt > kTravSlab) walk_list[atomicAdd(walk_count, 1)] = c;   // walked again after the scan
        }
    }
}

// Second traversal pass, one wavefront per collider whose partners overflowed its slab (a ground or
// wall overlapping hundreds of bodies).  The overlapping top of the tree is expanded breadth-first
// into at most 64 subtrees, each lane walks one with the filters of bp_traverse, the keys go to the
// collider's segment, and the segment is sorted: in LDS by a bitonic network when it fits, else by
// lane 0.  Keys are unique (each leaf is reached once), so the segment equals the one pass 1 of
// bp_traverse produces.
constexpr int kWalkSort = 4096;   // keys sorted in LDS (32 KB)
__global__ __launch_bounds__(kWave) void bp_walk_wave(DevColliders C, const int* sorted, int n, const int2* child,
                                                      const float4* nmin, const float4* nmax, const int* root,
                                                      NoColLists ncl, const int* offset, uint64_t* out_keys,
                                                      int capacity, const int* walk_list, const int* walk_count,
                                                      int* status) {
    __shared__ uint64_t s_keys[kWalkSort];
    __shared__ int s_front[2 * kWave];
    __shared__ int s_nf, s_cnt;
    const int lane = threadIdx.x;
    const int nw = *walk_count;
    for (int s = blockIdx.x; s < nw; s += gridDim.x) {
        const int c = walk_list[s];
        const int base = offset[c];
        const float4 amn = C.amin[c], amx = C.amax[c];
        if (lane == 0) { s_front[0] = root[C.world[c]]; s_nf = 1; s_cnt = 0; }
        __syncthreads();
        // breadth-first expansion while the frontier holds at most kWave / 2 nodes and some are internal
        for (int level = 0; level < 64; level++) {
            const int nf = s_nf;
            const int node = lane < nf ? s_front[lane] : -1;
            const bool internal = node >= 0 && node < n - 1;
            if (nf > kWave / 2 || !__ballot(internal)) break;
            const bool ov = node >= 0 && aabb_overlap(amn, amx, nmin[node], nmax[node]);
            const int emit = !ov ? 0 : (internal ? 2 : 1);
            const int incl = wave_incl_scan(emit);
            const int off = incl - emit;
            const int tot = __shfl(incl, kWave - 1, kWave);
            __syncthreads();   // every lane has read the frontier
            if (emit == 1) s_front[off] = node;
            if (emit == 2) {
                const int2 ch = child[node];
                s_front[off] = ch.x;
                s_front[off + 1] = ch.y;
            }
            if (lane == 0) s_nf = tot;
            __syncthreads();
        }
        const int nf = s_nf;
        for (int f = lane; f < nf; f += kWave) {
            int stack[kTravStack];
            int sp = 0;
            stack[sp++] = s_front[f];
            while (sp > 0) {
                const int node = stack[--sp];
                if (!aabb_overlap(amn, amx, nmin[node], nmax[node])) continue;
                if (node >= n - 1) {
                    const int o = sorted[node - (n - 1)];
                    if (o == c) continue;
                    if (!C.is_static[o] && o < c) continue;
                    if (!pair_allowed(C, c, o, ncl)) continue;
                    const uint64_t key = ((uint64_t)(uint32_t)c << 32) | (uint32_t)o;
                    const int pos = atomicAdd(&s_cnt, 1);
                    if (pos < kWalkSort) s_keys[pos] = key;
                    if (base + pos < capacity) out_keys[base + pos] = key;
                    else atomicOr(status, STATUS_PAIR_OVERFLOW);
                } else {
                    if (sp + 2 > kTravStack) { atomicOr(status, STATUS_STACK_OVERFLOW); break; }
                    const int2 ch = child[node];
                    stack[sp++] = ch.x;
                    stack[sp++] = ch.y;
                }
            }
        }
        __syncthreads();
        const int nk = s_cnt;
        const int nsort = min(nk, capacity - base);   // keys present in the segment
        if (nk <= kWalkSort) {
            int P = 1;
            while (P < nk) P <<= 1;
            for (int i = nk + lane; i < P; i += kWave) s_keys[i] = ~0ull;
            __syncthreads();
            for (int k = 2; k <= P; k <<= 1)
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int i = lane; i < P; i += kWave) {
                        const int ixj = i ^ j;
                        if (ixj > i) {
                            const uint64_t a = s_keys[i], b = s_keys[ixj];
                            if ((a > b) == ((i & k) == 0)) { s_keys[i] = b; s_keys[ixj] = a; }
                        }
                    }
                    __syncthreads();
                }
            for (int i = lane; i < nsort; i += kWave) out_keys[base + i] = s_keys[i];
        } else if (lane == 0) {
            for (int i = 1; i < nsort; i++) {
                const uint64_t key = out_keys[base + i];
                int pos = base + i;
                while (pos > base && out_keys[pos - 1] > key) {
                    out_keys[pos] = out_keys[pos - 1];
                    pos--;
                }
                out_keys[pos] = key;
            }
        }
        __syncthreads();   // the frontier and keys of this collider are done before the next one
    }
}

// Second traversal pass for the colliders with at most kTravSlab partners: their sorted partners are
// in the slab; eight lanes per collider copy them to the collider's segment of the key array.  Static
// colliders copy their static-static partner list.
__global__ void bp_copy_slabs(DevColliders C, int n, const int* count, const int* offset, const uint64_t* slab,
                              uint64_t* out_keys, int capacity, int* status) {
    RP3DG_GRID_STRIDE(t, n * kTravSlab) {
        const int c = t / kTravSlab, k = t % kTravSlab;
        const int base = offset[c];
        if (C.is_static[c]) {
            if (k != 0) continue;
            const int b = C.ss_offs[c], e = C.ss_offs[c + 1];
            for (int q = b; q < e; q++)
                if (base + q - b < capacity) out_keys[base + q - b] = ((uint64_t)(uint32_t)c << 32) | (uint32_t)C.ss_list[q];
            if (base + e - b > capacity) atomicOr(status, STATUS_PAIR_OVERFLOW);
            continue;
        }
        const int total = count[c];
        if (total > kTravSlab || k >= total) continue;
        // past the capacity the table is truncated (and the step reports the overflow), but every entry
        // below it is written, since the merge reads all of them
        if (base + k >= capacity) {
            atomicOr(status, STATUS_PAIR_OVERFLOW);
            continue;
        }
        out_keys[base + k] = slab[(size_t)c * kTravSlab + k];
    }
}

// ---------------------------------------------------------------------------------------------
// Batches of small worlds: no tree.  One workgroup per world group stages the fat AABBs of its
// colliders in LDS and every dynamic collider tests the boxes of all colliders of its own world in
// index order, so its partners come out sorted; the scan runs twice (count, then write after the
// prefix sum of the counts).  The pair set is the tree query's (the same
// AABB::testCollision on the same boxes, the same emitter rule and filters); the cost is
// sum over worlds of (colliders)^2 box tests, cheaper than building and walking a tree per world up to
// about a thousand colliders.  LDS per collider: min (w = body | static << 31), max (w = filter bits).
// ---------------------------------------------------------------------------------------------
// pass 0 writes the partner count of every collider, pass 1 (after the scan) its sorted keys.
__global__ __launch_bounds__(256) void bp_world_scan(DevColliders C, const int* gworld, const int* wcoll,
                                                     const int* block_group, NoColLists L, int* count,
                                                     const int* offset, uint64_t* out_keys, int capacity, int pass,
                                                     int* status) {
    extern __shared__ float4 s_box[];   // [group colliders] min boxes, then [group colliders] max boxes
    const int g = block_group[blockIdx.x];
    const int gc0 = wcoll[gworld[g]], gc1 = wcoll[gworld[g + 1]];
    const int ncg = gc1 - gc0;
    float4* smn = s_box;
    float4* smx = s_box + ncg;
    for (int i = threadIdx.x; i < ncg; i += blockDim.x) {
        const int c = gc0 + i;
        float4 mn = C.amin[c], mx = C.amax[c];
        mn.w = __int_as_float(C.body[c] | (C.is_static[c] ? (int)0x80000000u : 0));
        mx.w = __uint_as_float(C.filter[c]);
        smn[i] = mn;
        smx[i] = mx;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < ncg; i += blockDim.x) {
        const int c = gc0 + i;
        const float4 amn = smn[i], amx = smx[i];
        const int bc = __float_as_int(amn.w);
        const int base = pass ? offset[c] : 0;
        if (bc < 0) {   // static colliders do not query (their static-static pairs are listed at upload)
            const int b = C.ss_offs[c], e = C.ss_offs[c + 1];
            if (!pass) {
                count[c] = e - b;
            } else {
                for (int q = b; q < e; q++)
                    if (base + q - b < capacity) out_keys[base + q - b] = ((uint64_t)(uint32_t)c << 32) | (uint32_t)C.ss_list[q];
                if (base + e - b > capacity) atomicOr(status, STATUS_PAIR_OVERFLOW);
            }
            continue;
        }
        const int w = C.world[c];
        const int j0 = wcoll[w] - gc0, j1 = wcoll[w + 1] - gc0;
        const uint32_t fc = __float_as_uint(amx.w);
        const bool concave = shape_is_concave(C.type[c]);   // concave vs concave pairs are not created
        const int n0 = L.offs[bc], n1 = L.offs[bc + 1];   // bodies this collider's body must not collide with
        int cnt = 0;
        for (int j = j0; j < j1; j++) {
            const float4 bmn = smn[j], bmx = smx[j];
            const int bo = __float_as_int(bmn.w);
            if (j == i || (bo >= 0 && j < i)) continue;
            if (!aabb_overlap(amn, amx, bmn, bmx)) continue;
            const int bb = bo & 0x7fffffff;
            if (bb == bc) continue;
            bool excluded = false;
            for (int k = n0; k < n1; k++) excluded |= L.list[k] == bb;
            if (excluded || !filters_pass(fc, __float_as_uint(bmx.w))) continue;
            if (concave && shape_is_concave(C.type[gc0 + j])) continue;
            if (pass) {
                if (base + cnt < capacity) out_keys[base + cnt] = ((uint64_t)(uint32_t)c << 32) | (uint32_t)(gc0 + j);
                else atomicOr(status, STATUS_PAIR_OVERFLOW);
            }
            cnt++;
        }
        if (!pass) count[c] = cnt;
    }
}

// ---------------------------------------------------------------------------------------------
// Pair table update.  Surviving pairs keep collider order, LastFrameCollisionInfo and the contact
// cache.  A new pair is oriented as OverlappingPairs::addPair receives it: collider1 is the moved
// shape whose broadphase query reported the pair first, i.e. the moved one, or the one that comes
// first in the moved-shape set iteration order (rank) when both moved this step.
// ---------------------------------------------------------------------------------------------

// CollisionDispatch::selectAlgorithm (CollisionDispatch.cpp:80-115) on CollisionShapeType
__device__ __forceinline__ int shape_kind(int t) { return t == SHAPE_SPHERE ? 0 : (t == SHAPE_CAPSULE ? 1 : 2); }
__device__ __forceinline__ int select_algo(int t1, int t2) {
    if (shape_is_concave(t1) || shape_is_concave(t2)) return ALGO_CONCAVE;   // OverlappingPairs.cpp:427-440
    int a = shape_kind(t1), b = shape_kind(t2);
    if (a > b) { const int tmp = a; a = b; b = tmp; }
    if (a == 0 && b == 0) return ALGO_SPHERE_SPHERE;
    if (a == 0 && b == 1) return ALGO_SPHERE_CAPSULE;
    if (a == 1 && b == 1) return ALGO_CAPSULE_CAPSULE;
    if (a == 0 && b == 2) return ALGO_SPHERE_POLY;
    if (a == 1 && b == 2) return ALGO_CAPSULE_POLY;
    return ALGO_POLY_POLY;
}

// nkeys_ptr: overlapping pairs found (clamped to the capacity; an overflow is flagged by the traversal).
// nold_ptr: size of the previous table O.  nnew_ptr / nb_pairs_ptr receive the size of the new table N.
// old_offs: first entry of every emitter's segment in the previous table (its keys are sorted and
// grouped by emitter, key >> 32), so a key is searched only within its emitter's few old entries.
__global__ void bp_merge(DevColliders C, const uint64_t* keys, const int* nkeys_ptr, int capacity, DevPairs O,
                         const int* nold_ptr, const int* old_offs, DevPairs N, int* nnew_ptr, int* nb_pairs_ptr) {
    const int nkeys = min(*nkeys_ptr, capacity);
    const int nold = *nold_ptr;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *nnew_ptr = nkeys;
        *nb_pairs_ptr = nkeys;
    }
    RP3DG_GRID_STRIDE(k, nkeys) {
        const uint64_t key = keys[k];
        N.key[k] = key;
        const int e = (int)(key >> 32);
        int lo = min(old_offs[e], nold), hi = min(old_offs[e + 1], nold) - 1, found = -1;
        while (lo <= hi) {
            const int mid = (lo + hi) >> 1;
            const uint64_t v = O.key[mid];
            if (v == key) { found = mid; break; }
            if (v < key) lo = mid + 1;
            else hi = mid - 1;
        }
        N.prev[k] = found;
        if (found >= 0) {
            // the contact cache stays in O; the narrowphase reads it there through N.prev
            N.col[k] = O.col[found];
            N.algo[k] = O.algo[found];
            N.lf[k] = O.lf[found];
            N.lfaxis[k] = O.lfaxis[found];
            N.evf[k] = O.evf[found];
            O.kept[found] = 1;
        } else {
            N.evf[k] = 0;
            const int a = (int)(key >> 32), b = (int)(key & 0xffffffffu);
            const int ma = C.moved[a], mb = C.moved[b];
            bool a_first;
            if (ma && !mb) a_first = true;
            else if (mb && !ma) a_first = false;
            else a_first = C.rank[a] < C.rank[b];
            const int c1 = a_first ? a : b, c2 = a_first ? b : a;
            N.col[k] = make_int2(c1, c2);
            N.algo[k] = select_algo(C.type[c1], C.type[c2]);
            N.lf[k] = 0u;
            N.lfaxis[k] = make_float4(0.f, 1.f, 0.f, 0.f);   // LastFrameCollisionInfo() default axis
            N.ncp[k] = 0;
            N.fimp[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            N.fv1[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            N.fv2[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

}  // namespace rp3dg
