// Manifold list, graph colouring and the per-colour solver kernels.
//
// Colouring is Jones-Plassmann with hashed priorities: a manifold takes the
// smallest colour unused by its higher-priority neighbours once they are all
// coloured, which equals sequential greedy colouring in priority order and is
// therefore deterministic.  A fixed number of parallel rounds is launched (each
// returns at once when nothing is left) and a single-workgroup finisher
// completes whatever remains, so the step needs no host round trip.
#include <hip/hip_runtime.h>

#include "kcommon.h"
#include "rp3d_internal.h"
#include "solver.h"
#include "solver_dev.h"

namespace rp3dg {

// ------------------------------------------------------------------------------------------------
// Integration and per-colour rows (bodies in solver_dev.h)
// ------------------------------------------------------------------------------------------------

__global__ void dyn_integrate_velocities(DevBodies B, Settings S, float dt) {
    RP3DG_GRID_STRIDE(i, B.n) dyn_integrate_velocities_one(B, S, dt, i);
}

__global__ void dyn_integrate_positions(DevBodies B, float dt) {
    RP3DG_GRID_STRIDE(i, B.n) dyn_integrate_positions_one(B, dt, i);
}

__global__ void dyn_update_state(DevBodies B) {
    RP3DG_GRID_STRIDE(i, B.n) dyn_update_state_one(B, i);
}

__global__ void sv_prepare(SvArgs A) {
    const int nm = *A.nm;
    RP3DG_GRID_STRIDE(i, nm) sv_prepare_one(A, i);
}

__global__ void sv_warmstart(SvArgs A, int color) {
    if (color >= *A.ncolors) return;
    const int lo = A.cstart[color], hi = A.cstart[color + 1];
    RP3DG_GRID_STRIDE(t, hi - lo) sv_warmstart_one(A, lo + t);
}

__global__ void sv_solve_color(SvArgs A, int color, float beta_dt) {
    if (color >= *A.ncolors) return;
    const int lo = A.cstart[color], hi = A.cstart[color + 1];
    RP3DG_GRID_STRIDE(t, hi - lo) solve_manifold(A, lo + t, beta_dt);
}

__global__ void sv_store(SvArgs A) {
    const int nm = *A.nm;
    RP3DG_GRID_STRIDE(i, nm) sv_store_one(A, i);
}

// ------------------------------------------------------------------------------------------------
// Manifold list: pairs with at least one contact point, in pair order
// ------------------------------------------------------------------------------------------------

__global__ void sv_flag_manifolds(const int* ncp, const int* nb_pairs, int* flag) {
    const int np = *nb_pairs;
    RP3DG_GRID_STRIDE(k, np) flag[k] = ncp[k] > 0 ? 1 : 0;
}

__global__ void sv_compact_manifolds(const int* flag, const int* offs, const int* nb_pairs, const int* ncp, int* list,
                                     int* counters_manifolds, int* counters_points,
                                     unsigned long long* total_points) {
    const int np = *nb_pairs;
    if (blockIdx.x == 0 && threadIdx.x == 0) *counters_manifolds = offs[np];
    int pts = 0;
    RP3DG_GRID_STRIDE(k, np) {
        if (flag[k]) {
            list[offs[k]] = k;
            pts += ncp[k];
        }
    }
    pts = wave_sum(pts);
    if (lane_id() == 0 && pts) {
        atomicAdd(counters_points, pts);
        atomicAdd(total_points, (unsigned long long)pts);
    }
}

// ------------------------------------------------------------------------------------------------
// body -> manifold adjacency (counts, then fill after a scan)
// ------------------------------------------------------------------------------------------------

__global__ void sv_adj_count(DevBodies B, DevColliders C, DevPairs P, const int* list, const int* nm_ptr, int* cnt) {
    const int nm = *nm_ptr;
    RP3DG_GRID_STRIDE(m, nm) {
        const int2 c = P.col[list[m]];
        const int b1 = C.body[c.x], b2 = C.body[c.y];
        if (dyn_body(B, b1)) atomicAdd(&cnt[b1], 1);
        if (dyn_body(B, b2)) atomicAdd(&cnt[b2], 1);
    }
}

__global__ void sv_adj_fill(DevBodies B, DevColliders C, DevPairs P, const int* list, const int* nm_ptr, const int* offs,
                            int* cursor, int* adj, int2* mb) {
    const int nm = *nm_ptr;
    RP3DG_GRID_STRIDE(m, nm) {
        const int2 c = P.col[list[m]];
        const int b1 = C.body[c.x], b2 = C.body[c.y];
        const int d1 = dyn_body(B, b1) ? b1 : -1;
        const int d2 = dyn_body(B, b2) ? b2 : -1;
        mb[m] = make_int2(d1, d2);
        if (d1 >= 0) adj[offs[d1] + atomicAdd(&cursor[d1], 1)] = m;
        if (d2 >= 0) adj[offs[d2] + atomicAdd(&cursor[d2], 1)] = m;
    }
}

// ------------------------------------------------------------------------------------------------
// Jones-Plassmann colouring
// ------------------------------------------------------------------------------------------------

__device__ __forceinline__ uint32_t prio(uint32_t m) {
    uint32_t x = m * 0x9E3779B1u;
    x ^= x >> 16;
    x *= 0x85EBCA6Bu;
    x ^= x >> 13;
    return x;
}
__device__ __forceinline__ bool higher(uint32_t a, uint32_t b) {  // priority of a > priority of b
    const uint32_t pa = prio(a), pb = prio(b);
    return pa > pb || (pa == pb && a > b);
}

// Tries to colour manifold m (color[m] == -1 while uncoloured).  Returns false when a
// higher-priority neighbour is still uncoloured.  A neighbour coloured concurrently is read either
// as -1 or as its final colour, so the result does not depend on timing.
__device__ __forceinline__ bool jp_try(int m, const int2* mb, const int* offs, const int* adj, int* color, int* status) {
    const int2 bb = mb[m];
    uint64_t used = 0;
    for (int s = 0; s < 2; s++) {
        const int b = s == 0 ? bb.x : bb.y;
        if (b < 0) continue;
        const int e = offs[b + 1];
        for (int k = offs[b]; k < e; k++) {
            const int o = adj[k];
            if (o == m) continue;
            if (higher((uint32_t)o, (uint32_t)m)) {
                const int co = __atomic_load_n(&color[o], __ATOMIC_RELAXED);
                if (co < 0) return false;
                if (co < 64) used |= 1ull << co;
            }
        }
    }
    const uint64_t freem = ~used;
    int c = freem ? __ffsll((unsigned long long)freem) - 1 : 64;
    if (c >= kMaxColors) { atomicOr(status, STATUS_COLOR_OVERFLOW); c = kMaxColors - 1; }
    __atomic_store_n(&color[m], c, __ATOMIC_RELAXED);
    return true;
}

// One parallel round; left_prev = manifolds left after the previous round (nullptr for the first).
__global__ void sv_color_round(const int2* mb, const int* nm_ptr, const int* offs, const int* adj, int* color,
                               const int* left_prev, int* left, int* status) {
    if (left_prev && *left_prev == 0) return;
    const int nm = *nm_ptr;
    int n_left = 0;
    RP3DG_GRID_STRIDE(m, nm) {
        if (color[m] >= 0) continue;
        if (!jp_try(m, mb, offs, adj, color, status)) n_left++;
    }
    n_left = wave_sum(n_left);
    if (lane_id() == 0 && n_left) atomicAdd(left, n_left);
}

// Single workgroup: rounds until every manifold is coloured.  Each round colours at least the
// highest-priority uncoloured manifold, so nm + 1 rounds always suffice.
__global__ __launch_bounds__(1024) void sv_color_finish(const int2* mb, const int* nm_ptr, const int* offs,
                                                        const int* adj, int* color, const int* left_prev, int* rounds,
                                                        int* status) {
    if (left_prev && *left_prev == 0) return;
    __shared__ int left;
    const int nm = *nm_ptr;
    for (int r = 0; r <= nm; r++) {
        if (threadIdx.x == 0) left = 0;
        __syncthreads();
        int mine = 0;
        for (int m = threadIdx.x; m < nm; m += blockDim.x) {
            if (color[m] >= 0) continue;
            if (!jp_try(m, mb, offs, adj, color, status)) mine++;
        }
        if (mine) atomicAdd(&left, mine);
        __syncthreads();
        const int l = left;
        __syncthreads();
        if (l == 0) {
            if (threadIdx.x == 0) *rounds = r + 1;
            return;
        }
    }
    if (threadIdx.x == 0) atomicOr(status, STATUS_COLOR_STALL);
}

// Segment key of every manifold: its colour, or (world, colour) when the solver runs one workgroup per
// world.  Counts per segment and the number of colours are accumulated for the segment table.
__global__ void sv_color_keys(const int* color, const int* nm_ptr, DevColliders C, DevPairs P, const int* mlist,
                              int world_major, uint64_t* keys, int* vals, int* seg_cnt, int* ncolors) {
    __shared__ int hist[kMaxColors];
    for (int i = threadIdx.x; i < kMaxColors; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int nm = *nm_ptr;
    int cmax = -1;
    RP3DG_GRID_STRIDE(m, nm) {
        const int c = color[m];
        uint64_t seg = (uint64_t)c;
        if (world_major) {
            const int w = C.world[P.col[mlist[m]].x];
            seg = (uint64_t)w * kMaxColors + c;
            atomicAdd(&seg_cnt[seg], 1);
        } else {
            atomicAdd(&hist[c], 1);
        }
        keys[m] = seg;
        vals[m] = m;
        cmax = c > cmax ? c : cmax;
    }
    cmax = wave_max(cmax);
    if (lane_id() == 0 && cmax >= 0) atomicMax(ncolors, cmax + 1);
    __syncthreads();
    if (!world_major)
        for (int i = threadIdx.x; i < kMaxColors; i += blockDim.x)
            if (hist[i]) atomicAdd(&seg_cnt[i], hist[i]);
}

}  // namespace rp3dg
