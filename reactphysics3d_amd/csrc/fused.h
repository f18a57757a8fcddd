// Argument block of the fused per-world kernels (fused.hip).
#pragma once

#include "joints.h"
#include "rp3d_internal.h"
#include "solver.h"

namespace rp3dg {

struct FusedArgs {
    SvArgs A;                  // rows in (world group, colour)-major order; A.cstart unused
    DevJoints J;               // rows sorted by (world group, kind, colour, creation index)
    Settings settings;
    const int* seg_start;      // [g * kMaxColors + c]: first row of colour c of group g; [ng * kMaxColors] = total
    const int2* jt_wr;         // [g * jt_slots + s]: joint rows of slot s (kind, colour) of group g
    int jt_slots;
    const int* wbody;          // [nw + 1] body ranges of the worlds
    const int* wcoll;          // [nw + 1] collider ranges of the worlds
    const int* gworld;         // [ng + 1] world ranges of the groups (consecutive worlds)
    const int* block_world;    // workgroup -> group (largest groups first)
    const float4* hull_bounds;
    float dt, beta_dt;
    int max_iters;             // velocity iterations (fixed), or the cap in residual mode
    float res_thr;             // > 0: iterate to this velocity residual
    float4 *pv, *pw;           // velocities after the previous sweep (residual mode)
    int* world_iters;          // iterations each world ran in the last step
    float* world_res;          // its last residual (residual mode)
    // learned warm start / sample logging
    int phase;                 // 0: integration + rows (+ features), 1: warm start .. store, 2: both
    float* feats;              // [feat_cap][24] features of the points without a cached impulse
    int* row_slot;             // solver point slot (4 * row + point) of each feature row
    int* nrows;                // feature rows written this step
    int feat_cap;
    int lds_bodies;            // worlds with at most this many bodies keep cv, cw, sv, sw, the world inverse
                               // inertia rows and the flags in LDS (100 * lds_bodies bytes); 0 = HBM
    int lds_residual;          // residual mode: the previous sweep's velocities in LDS too (32 B per body
                               // more) instead of pv / pw
};

// LDS bytes of a fused solver workgroup staging `bodies` bodies (with the residual's previous velocities)
__host__ __device__ constexpr size_t fused_lds_bytes(int bodies, bool residual) {
    return (size_t)bodies * (6 * sizeof(float4) + sizeof(int)) + (residual ? (size_t)bodies * 2 * sizeof(float4) : 0);
}

constexpr int kFeatures = 24;

constexpr int kFusedMaxBlock = 256;   // threads per group at most (register budget of the fused kernels)
constexpr int kFusedMaxLds = 512;     // bodies per group staged in LDS at most (the classes use 64 .. 512)
constexpr int kGroupMaxWorlds = 64;   // worlds per group at most (one bit each in the convergence mask)
constexpr int kGroupShift = 16;       // staged body flags carry the body's world index in its group here

// one instance per contact position-correction technique: SPLIT = split impulses, else Baumgarte contacts
template <bool SPLIT>
__global__ void sv_world_solve(FusedArgs F);
__global__ void world_position(FusedArgs F);
// Appends (features, converged impulses) of this step's feature rows to the sample log at *logged.
__global__ void sv_log_samples(FusedArgs F, float* log_feats, float* log_targets, const unsigned long long* logged,
                               long long log_cap);
__global__ void sv_log_advance(const int* nrows, int feat_cap, unsigned long long* logged, long long log_cap);

}  // namespace rp3dg
