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

__global__ void dyn_integrate_positions(DevBodies B, float dt, float split_factor) {
    RP3DG_GRID_STRIDE(i, B.n) dyn_integrate_positions_one(B, dt, i, split_factor);
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
    if (A.split) {
        RP3DG_GRID_STRIDE(t, hi - lo) solve_manifold<true>(A, lo + t, beta_dt);
    } else {
        RP3DG_GRID_STRIDE(t, hi - lo) solve_manifold<false>(A, lo + t, beta_dt);
    }
}

__global__ void sv_store(SvArgs A) {
    const int nm = *A.nm;
    RP3DG_GRID_STRIDE(i, nm) sv_store_one(A, i);
}

// ------------------------------------------------------------------------------------------------
// Manifold list: pairs with at least one contact point, in pair order
// ------------------------------------------------------------------------------------------------

// a pair's manifolds: its entry, then (concave pairs, xm non-null) its xm tail entries from xs
__global__ void sv_flag_manifolds(const int* ncp, const int* xm, const int* nb_pairs, int* flag) {
    const int np = *nb_pairs;
    RP3DG_GRID_STRIDE(k, np) flag[k] = ncp[k] > 0 ? 1 + (xm ? xm[k] : 0) : 0;
}

__global__ void sv_compact_manifolds(const int* flag, const int* offs, const int* nb_pairs, const int* ncp, const int* xm,
                                     const int* xs, int* list, int* counters_manifolds, int* counters_points,
                                     unsigned long long* total_points) {
    const int np = *nb_pairs;
    if (blockIdx.x == 0 && threadIdx.x == 0) *counters_manifolds = offs[np];
    int pts = 0;
    RP3DG_GRID_STRIDE(k, np) {
        if (flag[k]) {
            list[offs[k]] = k;
            pts += ncp[k];
            for (int j = 1; j < flag[k]; j++) {
                const int slot = xs[k] + j - 1;
                list[offs[k] + j] = slot;
                pts += ncp[slot];
            }
        }
    }
    // one atomic per workgroup: same-address atomics from every wave would serialise in the L2
    __shared__ int s_pts;
    if (threadIdx.x == 0) s_pts = 0;
    __syncthreads();
    pts = wave_sum(pts);
    if (lane_id() == 0 && pts) atomicAdd(&s_pts, pts);
    __syncthreads();
    if (threadIdx.x == 0 && s_pts) {
        atomicAdd(counters_points, s_pts);
        atomicAdd(total_points, (unsigned long long)s_pts);
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

// ------------------------------------------------------------------------------------------------
// Colouring per world group (fused layout)
// ------------------------------------------------------------------------------------------------

// The manifold list is in pair-table order (sorted by the lower collider index) and colliders are
// numbered world by world, so the manifolds of a world form one contiguous range.  One pass writes the
// dynamic bodies of every manifold (-1 for static or kinematic) and the first manifold of every world
// (wmstart[nw] = number of manifolds).
__global__ void sv_world_ranges(DevBodies B, DevColliders C, DevPairs P, const int* list, const int* nm_ptr, int nw,
                                int2* mb, int* wmstart) {
    const int nm = *nm_ptr;
    if (nm == 0) {
        RP3DG_GRID_STRIDE(w, nw + 1) wmstart[w] = 0;
        return;
    }
    RP3DG_GRID_STRIDE(m, nm) {
        const int2 c = P.col[list[m]];
        const int b1 = C.body[c.x], b2 = C.body[c.y];
        mb[m] = make_int2(dyn_body(B, b1) ? b1 : -1, dyn_body(B, b2) ? b2 : -1);
        const int w = C.world[c.x];
        const int wp = m ? C.world[P.col[list[m - 1]].x] : -1;
        for (int k = wp + 1; k <= w; k++) wmstart[k] = m;
        if (m == nm - 1)
            for (int k = w + 1; k <= nw; k++) wmstart[k] = nm;
    }
}

// One 64-lane workgroup per world group.  Lane k colours the manifolds of world k of the group
// greedily in manifold order (smallest colour unused at either dynamic body; worlds share no body, so
// this equals greedy colouring of the whole group in manifold order), then the group's manifolds are
// placed colour by colour, in manifold order within a colour: order[row] = manifold, keys[row] =
// group * kMaxColors + colour, seg_start[group * kMaxColors + c] = first row of colour c.
// Dynamic LDS: one 64-bit colour mask per body of the largest group.
__global__ __launch_bounds__(kWave) void sv_group_color(const int2* mb, const int* wmstart, const int* gworld,
                                                        const int* wbody, int ngroups, int* color, int* order,
                                                        uint64_t* keys, int* seg_start, int* ncolors, int* status) {
    extern __shared__ unsigned long long s_mask[];
    __shared__ int s_hist[kMaxColors];
    const int g = blockIdx.x;
    const int lane = threadIdx.x;
    const int w0 = gworld[g], w1 = gworld[g + 1];
    const int b0 = wbody[w0], nbg = wbody[w1] - b0;
    const int ma = wmstart[w0], mz = wmstart[w1];
    for (int i = lane; i < nbg; i += kWave) s_mask[i] = 0ull;
    s_hist[lane] = 0;
    __syncthreads();
    // greedy colouring, one world per lane.  The chain of colour decisions is serial; the body pairs of
    // the next kAhead manifolds are loaded while the current kAhead are coloured, so a global load's
    // latency is paid once per world rather than once per batch.
    constexpr int kAhead = 16;
    int cmax = -1;
    if (w1 - w0 == 1 && mz - ma > 2 * kWave) {
        // A group of one large world: the wave colours 64 manifolds at a time.  A manifold's colour is
        // the smallest one used neither at its bodies by earlier chunks (the masks) nor by the earlier
        // manifolds of its chunk that share a body with it.  Those dependencies point backwards, so
        // iterating every lane's choice to a fixed point (at most one round per manifold of the chunk,
        // plus one that sees no change) gives the colours of the sequential greedy pass.
        for (int i0 = ma; i0 < mz; i0 += kWave) {
            const int i = i0 + lane;
            const int2 bb = i < mz ? mb[i] : make_int2(-1, -1);
            unsigned long long pre = 0ull;
            if (bb.x >= 0) pre |= s_mask[bb.x - b0];
            if (bb.y >= 0) pre |= s_mask[bb.y - b0];
            unsigned long long dep = 0ull;   // earlier lanes of the chunk sharing a dynamic body
            for (int k = 0; k < kWave; k++) {
                const int kx = __shfl(bb.x, k, kWave), ky = __shfl(bb.y, k, kWave);
                const bool share = (kx >= 0 && (kx == bb.x || kx == bb.y)) || (ky >= 0 && (ky == bb.x || ky == bb.y));
                if (k < lane && share) dep |= 1ull << k;
            }
            const int nd = wave_max(__popcll(dep));
            unsigned long long used = pre;
            int c = ~used ? __ffsll(~used) - 1 : kMaxColors - 1;
            for (int round = 0; round <= kWave; round++) {
                used = pre;
                unsigned long long d = dep;
                for (int t = 0; t < nd; t++) {   // every lane takes part in each shuffle
                    const int k = d ? __ffsll(d) - 1 : lane;
                    const int ck = __shfl(c, k, kWave);
                    if (d) { used |= 1ull << ck; d &= d - 1; }
                }
                const int cn = ~used ? __ffsll(~used) - 1 : kMaxColors - 1;
                const bool changed = cn != c;
                c = cn;
                if (!__ballot(changed)) break;
            }
            if (i < mz) {
                if (!~used) atomicOr(status, STATUS_COLOR_OVERFLOW);
                const unsigned long long bit = 1ull << c;
                if (bb.x >= 0) atomicOr(&s_mask[bb.x - b0], bit);
                if (bb.y >= 0) atomicOr(&s_mask[bb.y - b0], bit);
                color[i] = c;
                cmax = c > cmax ? c : cmax;
            }
            __syncthreads();
        }
    } else if (lane < w1 - w0) {
        const int s = wmstart[w0 + lane], e = wmstart[w0 + lane + 1];
        int2 cur[kAhead], nxt[kAhead];
#pragma unroll
        for (int u = 0; u < kAhead; u++) cur[u] = s + u < e ? mb[s + u] : make_int2(-1, -1);
        for (int i = s; i < e; i += kAhead) {
#pragma unroll
            for (int u = 0; u < kAhead; u++) nxt[u] = i + kAhead + u < e ? mb[i + kAhead + u] : make_int2(-1, -1);
#pragma unroll
            for (int u = 0; u < kAhead; u++) {
                if (i + u >= e) break;
                const int2 bb = cur[u];
                unsigned long long used = 0ull;
                if (bb.x >= 0) used |= s_mask[bb.x - b0];
                if (bb.y >= 0) used |= s_mask[bb.y - b0];
                int c = ~used ? __ffsll(~used) - 1 : kMaxColors;
                if (c >= kMaxColors) {
                    atomicOr(status, STATUS_COLOR_OVERFLOW);
                    c = kMaxColors - 1;
                }
                const unsigned long long bit = 1ull << c;
                if (bb.x >= 0) s_mask[bb.x - b0] |= bit;
                if (bb.y >= 0) s_mask[bb.y - b0] |= bit;
                color[i + u] = c;
                cmax = c > cmax ? c : cmax;
            }
#pragma unroll
            for (int u = 0; u < kAhead; u++) cur[u] = nxt[u];
        }
    }
    cmax = wave_max(cmax);
    if (lane == 0 && cmax >= 0) atomicMax(ncolors, cmax + 1);
    __syncthreads();   // the colours written above are read back by other lanes of this workgroup
    for (int i = ma + lane; i < mz; i += kWave) atomicAdd(&s_hist[color[i]], 1);
    __syncthreads();
    // segment starts: exclusive prefix of the colour counts
    const int cnt = s_hist[lane];
    const int base = ma + wave_incl_scan(cnt) - cnt;
    seg_start[(size_t)g * kMaxColors + lane] = base;
    if (g == ngroups - 1 && lane == 0) seg_start[(size_t)ngroups * kMaxColors] = wmstart[gworld[ngroups]];
    __syncthreads();
    s_hist[lane] = base;   // running cursor of each colour
    __syncthreads();
    // stable placement, 64 manifolds at a time: lanes of equal colour take consecutive rows
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int i0 = ma; i0 < mz; i0 += kWave) {
        const int i = i0 + lane;
        const int c = i < mz ? color[i] : -1;
        unsigned long long pending = __ballot(c >= 0);
        while (pending) {
            const int leader = __ffsll(pending) - 1;
            const int cl = __shfl(c, leader, kWave);
            const unsigned long long same = __ballot(c == cl);
            if (c == cl) {
                const int row = s_hist[cl] + __popcll(same & below);
                order[row] = i;
                keys[row] = (uint64_t)g * kMaxColors + cl;
            }
            pending &= ~same;
            __syncthreads();
            if (lane == leader) s_hist[cl] += __popcll(same);
            __syncthreads();
        }
    }
}

// Segment key of every manifold: its colour, or (world group, colour) when the solver runs one workgroup
// per group of consecutive worlds (world_group != null).  Counts per segment and the number of colours
// are accumulated for the segment table.
__global__ void sv_color_keys(const int* color, const int* nm_ptr, DevColliders C, DevPairs P, const int* mlist,
                              const int* world_group, uint64_t* keys, int* vals, int* seg_cnt, int* ncolors) {
    __shared__ int hist[kMaxColors];
    for (int i = threadIdx.x; i < kMaxColors; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const int nm = *nm_ptr;
    int cmax = -1;
    RP3DG_GRID_STRIDE(m, nm) {
        const int c = color[m];
        uint64_t seg = (uint64_t)c;
        if (world_group) {
            const int g = world_group[C.world[P.col[mlist[m]].x]];
            seg = (uint64_t)g * kMaxColors + c;
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
    if (!world_group)
        for (int i = threadIdx.x; i < kMaxColors; i += blockDim.x)
            if (hist[i]) atomicAdd(&seg_cnt[i], hist[i]);
}

}  // namespace rp3dg
