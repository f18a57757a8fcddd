// Kernel argument block and launch declarations of the solver / integration kernels.
#pragma once

#include "rp3d_internal.h"

namespace rp3dg {

constexpr int kMaxColors = 64;

struct SvArgs {
    DevBodies B;
    DevColliders C;
    DevPairs P;
    DevSolver S;
    const int* nm;        // number of manifolds
    const int* mlist;     // manifold -> pair slot (pair order)
    const int* order;     // colour-sorted position -> manifold
    const int* cstart;    // colour ranges [cstart[c], cstart[c+1]) (colour-major layout)
    const int* ncolors;
    float restitution_threshold;
    int learned;          // learned warm start for points without a cached impulse
    const int* adj_offs;  // body -> manifold adjacency offsets (manifold count of a dynamic body)
    int split;            // split impulses (1) or Baumgarte contacts (0), PhysicsWorld.h:599-607
};

__global__ void dyn_integrate_velocities(DevBodies B, Settings S, float dt);
__global__ void dyn_integrate_positions(DevBodies B, float dt, float split_factor);
__global__ void dyn_update_state(DevBodies B);
__global__ void sv_flag_manifolds(const int* ncp, const int* xm, const int* nb_pairs, int* flag);
__global__ void sv_compact_manifolds(const int* flag, const int* offs, const int* nb_pairs, const int* ncp, const int* xm,
                                     const int* xs, int* list, int* counters_manifolds, int* counters_points,
                                     unsigned long long* total_points);
__global__ void sv_adj_count(DevBodies B, DevColliders C, DevPairs P, const int* list, const int* nm_ptr, int* cnt);
__global__ void sv_adj_fill(DevBodies B, DevColliders C, DevPairs P, const int* list, const int* nm_ptr, const int* offs,
                            int* cursor, int* adj, int2* mb);
__global__ void sv_color_round(const int2* mb, const int* nm_ptr, const int* offs, const int* adj, int* color,
                               const int* left_prev, int* left, int* status);
__global__ void sv_color_finish(const int2* mb, const int* nm_ptr, const int* offs, const int* adj, int* color,
                                const int* left_prev, int* rounds, int* status);
__global__ void sv_world_ranges(DevBodies B, DevColliders C, DevPairs P, const int* list, const int* nm_ptr, int nw,
                                int2* mb, int* wmstart);
__global__ void sv_group_color(const int2* mb, const int* wmstart, const int* gworld, const int* wbody, int ngroups,
                               int* color, int* order, uint64_t* keys, int* seg_start, int* ncolors, int* status);
__global__ void sv_color_keys(const int* color, const int* nm_ptr, DevColliders C, DevPairs P, const int* mlist,
                              const int* world_group, uint64_t* keys, int* vals, int* seg_cnt, int* ncolors);
__global__ void sv_prepare(SvArgs A);
__global__ void sv_warmstart(SvArgs A, int color);
__global__ void sv_solve_color(SvArgs A, int color, float beta_dt);
__global__ void sv_store(SvArgs A);

}  // namespace rp3dg
