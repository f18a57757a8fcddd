// Kernel argument block of the narrowphase kernels.
#pragma once

#include "rp3d_internal.h"

namespace rp3dg {

struct NpArgs {
    DevBodies B;
    DevColliders C;
    DevPairs P;                // pair table of this step
    DevPairs O;                // previous table: contact caches of surviving pairs (P.prev)
    const int* nb_pairs;
    const HullHeader* hulls;
    const float4* hull_verts;
    const float4* hull_normals;
    const int4* hull_edges;
    const int* hull_face_edge;
    const int* hull_face_verts;
    const int* hull_face_vidx;
    float persistent_threshold;
    float cos_similar;
    int* status;               // device status word (STATUS_*)
    // world queries (PhysicsWorld::testOverlap / testCollision, CD.cpp:1775-1896) run the same kernels:
    int query;                 // NP_STEP, NP_QUERY_ALL (both colliders world-query colliders) or NP_QUERY_BODIES
    int qbody1, qbody2;        // NP_QUERY_BODIES: pairs with body qbody1 (and qbody2 when >= 0)
    int report;                // NP_REPORT_NO_TRIGGER, NP_REPORT_ALL or NP_REPORT_NONE (reportContacts)
    int clip_prev;             // SATAlgorithm's clipWithPreviousAxisIfStillColliding (1 in steps, 0 in queries)
    int lf_update;             // processPotentialContacts updates wasColliding / isValid (0: overlap queries)
};
constexpr int NP_STEP = 0, NP_QUERY_ALL = 1, NP_QUERY_BODIES = 2;
constexpr int NP_REPORT_NO_TRIGGER = 0, NP_REPORT_ALL = 1, NP_REPORT_NONE = 2;

// Narrowphase test lists: bins 1..5 closed forms, ALGO_POLY_POLY pairs with a convex mesh,
// ALGO_COUNT box-box pairs colliding last step, ALGO_COUNT + 1 box-box pairs separated last step.
// List slot of each bin (kNpListSlots lists of the pair capacity each are allocated).
constexpr int kNpBins = ALGO_COUNT + 2;
constexpr int kNpListSlots = ALGO_COUNT + 1;
__host__ __device__ constexpr int np_list_slot(int bin) {
    return bin == ALGO_COUNT ? 0 : (bin == ALGO_COUNT + 1 ? ALGO_COUNT : bin);
}
static_assert(np_list_slot(ALGO_COUNT + 1) < kNpListSlots && np_list_slot(ALGO_COUNT) < kNpListSlots &&
                  np_list_slot(ALGO_POLY_POLY) < kNpListSlots,
              "every narrowphase bin needs an allocated list slot");

__global__ void np_classify(NpArgs A, int* lists, int list_stride, int* counts, int* box_count, int* box_sep_count,
                            int* hull_count);
__global__ void np_sphere_sphere(NpArgs A, const int* list, const int* count);
__global__ void np_sphere_capsule(NpArgs A, const int* list, const int* count);
__global__ void np_capsule_capsule(NpArgs A, const int* list, const int* count);
__global__ void np_sphere_poly(NpArgs A, const int* list, const int* count);
__global__ void np_capsule_poly(NpArgs A, const int* list, const int* count);
__global__ void np_poly_poly(NpArgs A, const int* list, const int* count);
// np_box_box runs in workgroups of exactly this many threads (its clip buffers are sliced from LDS by
// threadIdx.x)
constexpr int kNpBoxThreads = 128;
__global__ void np_box_box(NpArgs A, const int* list, const int* count);
void upload_box_topology(const int* edges4, const int* face_edge, const int* face_verts);

}  // namespace rp3dg
