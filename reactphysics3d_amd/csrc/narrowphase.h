// Kernel argument block of the narrowphase kernels.
#pragma once

#include "rp3d_internal.h"

namespace rp3dg {

// The triangle tests of one step's convex vs concave pairs, pair by pair in middle-phase order.
struct ConcaveTests {
    int* pair;                 // pair slot
    int* tri;                  // triangle id (mesh: triangle index; height field: HeightField::computeTriangleShapeId)
    uint32_t* lf;              // the triangle's LastFrameCollisionInfo after the test
    float4* lfaxis;
    int* hit;                  // NarrowPhaseInfo::isColliding
    int2* pts;                 // its contact points in the point pool (first, count)
    int cap;
    int* count;
};

// The triangles tested this step as hulls: test t's TriangleShape has vertices [3t, 3t+3) and faces
// [2t, 2t+2) (normal, opposite normal), over the six half-edges and face vertex lists that every triangle
// shares (PhysicsCommon.cpp initTriangleShapeHalfEdgeStructure; face_edge0 = each face's first half-edge).
struct TriArena {
    float4* verts;       // 3 per test slot
    float4* normals;     // 2 per test slot
    int* face_edge;      // 2 per test slot
    int* face_verts;     // 2 x 2 per test slot: (offset into vidx, 3)
    const int4* edges;   // 6 shared half-edges
    const int* vidx;     // 6 shared face vertex indices
    int face_edge0[2];
};

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
    // convex vs concave pairs (concave.cpp arenas; null pointers when the batch has no concave collider)
    const ConcaveInst* cinst;
    const float4* cverts;      // mesh vertices (unscaled)
    const int4* ctris;         // mesh triangles: vertex indices
    const float4* cnorm;       // scaled vertex normals of each concave collider
    const float4* cnodes;      // mesh AABB tree: (min, child0) (max, child1); a leaf has child0 = -1, child1 = triangle
    const float* cheights;     // height field samples
    const float4* hull_bounds; // local bounds of convex meshes (the convex shape's AABB in the concave space)
    ConcaveTests CT, CTo;      // this step's triangle tests, the previous step's (the pairs' last-frame infos)
    float4 *cp_nd, *cp_l1, *cp_l2;   // contact points of the tests: normal + depth, local points
    int cp_cap;
    int* cp_count;
    TriArena tri;
    int tail_base, tail_cap;   // tail entries of P: [tail_base, tail_base + tail_cap)
    int* tail_count;
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
// convex vs concave pairs: middle phase (lists the triangle tests), the tests, then per pair the potential
// manifolds, their reduction and the warm-start transfer
__global__ void np_concave_mid(NpArgs A, const int* list, const int* count);
__global__ void np_concave_tri(NpArgs A);
__global__ void np_concave_reduce(NpArgs A, const int* list, const int* count);

}  // namespace rp3dg
