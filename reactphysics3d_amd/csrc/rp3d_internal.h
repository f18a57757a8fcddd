// Internal host/device layout of the MI355X step path.
//
// HBM layout (all SoA, one entry per body / collider / pair / manifold of the
// whole batch; worlds are concatenated and every record carries its world id):
//   bodies     : float4 streams (x = centre of mass, q, v, w, ...) so each
//                per-body kernel reads/writes 16-byte aligned vectors.
//   colliders  : shape, local/world transform, fat AABB, material, filter.
//   pairs      : persistent overlapping-pair table sorted by (lo, hi) collider
//                index, double buffered across steps.  Each pair carries the
//                SAT/GJK temporal-coherence record and its contact cache (one
//                manifold of <= 4 points, the only case for convex pairs,
//                CollisionDetectionSystem.cpp:1247-1290).
//   manifolds  : per-step compacted, colour-sorted list of colliding pairs
//                with the solver rows (ContactSolverSystem.h:133-260).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "rp3d_math.h"

namespace rp3dg {

enum ShapeType : int { SHAPE_SPHERE = 0, SHAPE_CAPSULE = 1, SHAPE_BOX = 2, SHAPE_HULL = 3, SHAPE_CONCAVE_MESH = 4, SHAPE_HEIGHTFIELD = 5 };
__host__ __device__ constexpr bool shape_is_concave(int t) { return t >= SHAPE_CONCAVE_MESH; }

// NarrowPhaseAlgorithmType (include/reactphysics3d/collision/narrowphase/CollisionDispatch.h)
enum Algo : int {
    ALGO_NONE = 0,
    ALGO_SPHERE_SPHERE = 1,
    ALGO_SPHERE_CAPSULE = 2,
    ALGO_CAPSULE_CAPSULE = 3,
    ALGO_SPHERE_POLY = 4,
    ALGO_CAPSULE_POLY = 5,
    ALGO_POLY_POLY = 6,
    ALGO_CONCAVE = 7,      // convex vs concave: middle phase over the triangles (CD.cpp:522-615)
    ALGO_COUNT = 8
};

// Body flags
constexpr int BF_DYNAMIC = 1;   // BodyType::DYNAMIC
constexpr int BF_ENABLED = 2;   // in the enabled partition (not static, not sleeping)
constexpr int BF_GRAVITY = 4;   // gravity enabled
constexpr int BF_KINEMATIC = 8;
constexpr int BF_UNIT_LOCKS = 16;   // linear and angular lock factors are (1, 1, 1): the solver skips them
constexpr int BF_ALLOW_SLEEP = 32;  // RigidBody::isAllowedToSleep
constexpr int BF_WOKEN = 64;        // woken up this step: its colliders enter the moved set (RigidBody.cpp:1045-1051)

// LastFrameCollisionInfo (include/reactphysics3d/engine/OverlappingPairs.h:55-93) packed in a word:
constexpr uint32_t LF_VALID = 1u << 0;
constexpr uint32_t LF_WAS_COLLIDING = 1u << 1;
constexpr uint32_t LF_WAS_GJK = 1u << 2;
constexpr uint32_t LF_WAS_SAT = 1u << 3;
constexpr uint32_t LF_SAT_FACE1 = 1u << 4;
constexpr uint32_t LF_SAT_FACE2 = 1u << 5;
// bits 8..15 satMinAxisFaceIndex, 16..23 satMinEdge1Index, 24..31 satMinEdge2Index
__host__ __device__ inline uint32_t lf_face(uint32_t w) { return (w >> 8) & 0xffu; }
__host__ __device__ inline uint32_t lf_edge1(uint32_t w) { return (w >> 16) & 0xffu; }
__host__ __device__ inline uint32_t lf_edge2(uint32_t w) { return (w >> 24) & 0xffu; }

constexpr int kTravBlock = 128;       // threads per block of the broadphase traversal
constexpr int kTravSlab = 8;          // partners kept per collider by the first traversal pass
constexpr int kMaxRawContacts = 16;   // NB_MAX_CONTACT_POINTS_IN_NARROWPHASE_INFO (configuration.h:134)
constexpr int kMaxManifoldPoints = 4; // MAX_CONTACT_POINTS_IN_MANIFOLD
// Hull faces have at most this many vertices: clipping an incident face by the side planes of a
// reference face adds at most one vertex per plane, so the clip buffers hold 2 * 24 = 48 vertices.
constexpr int kMaxHullFaceVertices = 24;

// Convex hull arena (ConvexMeshShape + HalfEdgeStructure), one record per hull.
struct HullHeader {
    int32_t vbase, nv;     // vertices (float4 xyz) [vbase, vbase+nv)
    int32_t fbase, nf;     // faces: normal (float4) + first half-edge + vertex list
    int32_t ebase, ne;     // half-edges (int4: vertex, twin, face, next)
    int32_t fvbase, nfv;   // concatenated face vertex indices
    float cx, cy, cz, pad; // centroid
};

struct DevBodies {
    float4 *x, *q, *p;          // centre of mass (world), orientation, body origin
    float4 *v, *w;              // velocities (RigidBodyComponents::mLinear/AngularVelocities)
    float4 *cv, *cw;            // constrained velocities, indexed [body - vbase]
    float4 *sv, *sw;            // split velocities, indexed [body - vbase]
    float4 *cx, *cq;            // constrained position / orientation
    float4 *force, *torque;     // external force / torque
    float4 *ii;                 // inverse local inertia xyz, w = inverse mass
    float4 *iw0, *iw1, *iw2;    // world inverse inertia rows
    float4 *com;                // local centre of mass xyz, w = mass
    float4 *llock, *alock;      // lock factors xyz; llock.w = linear damping, alock.w = angular damping
    int *flags;                 // indexed [body - vbase] like the velocities
    int *world;
    int n;
    int vbase;                  // 0, or the first body of a world whose velocities are staged in LDS
};

// Body pairs excluded from collision (joints with collideConnected = false), per body: partners of body
// b are list[offs[b] .. offs[b + 1])
struct NoColLists {
    const int* offs;
    const int* list;
};

// Collider kinds (Collider::getIsTrigger / getIsSimulationCollider / getIsWorldQueryCollider, Collider.cpp:240-300);
// the values are the RP3DG_COLLIDER_* bits of the C ABI.
constexpr int CK_TRIGGER = 1;
constexpr int CK_SIMULATION = 2;
constexpr int CK_WORLD_QUERY = 4;

struct DevColliders {
    int *body;
    int *type;
    float4 *shape;             // sphere r | capsule r, halfHeight | box hx,hy,hz | hull: x = hull index
    float4 *lpos, *lq;         // local-to-body transform
    float4 *wpos, *wq;         // local-to-world transform
    float4 *amin, *amax;       // fat AABB (broadphase tree node AABB)
    float2 *mat;               // sqrt(friction), bounciness
    uint32_t *filter;          // category | mask << 16
    int *kind;                 // CK_* bits: trigger, simulation collider, world-query collider
    int *moved;                // fat AABB recomputed by the last refit (BroadPhaseSystem moved set)
    int *rank;                 // position of the collider in the moved-set iteration order
    int *world;
    int *is_static;            // collider of a static body (never moves, never traverses the BVH)
    int *ss_offs;              // static-static overlapping partners (CSR, constant after creation)
    int *ss_list;
    int n;
};

struct DevPairs {
    uint64_t *key;             // (lo << 32) | hi, collider indices
    int2 *col;                 // collider1, collider2 as the reference orients them
    int *algo;
    uint32_t *lf;              // LastFrameCollisionInfo flags + SAT indices
    float4 *lfaxis;            // gjkSeparatingAxis
    int *ncp;                  // contact points in the cache (0 = no contact pair)
    int *prev;                 // entry of the same pair in the previous table (-1: new pair); its contact
                               // cache (ncp, points, friction) is read there, not copied
    float4 *fv1, *fv2;         // manifold friction vectors
    float4 *fimp;              // friction1, friction2, twist impulses
    float4 *pn;                // [4*slot+k] normal xyz, depth
    float4 *pl1;               // [4*slot+k] local point on shape 1, w = accumulated normal impulse
    float4 *pl2;               // [4*slot+k] local point on shape 2, w = resting flag (0/1)
    int *evf;                  // EV_CUR / EV_PREV: OverlappingPair::collidingInCurrentFrame / InPreviousFrame
    int *kept;                 // set by the merge on every entry of the previous table that survives
    // convex vs concave pairs (null when the batch has no concave collider): the step's triangle tests
    // [ct0, ct0 + ctn) of the test list, and the contact manifolds beyond the first one (a concave pair
    // has up to NB_MAX_CONTACT_MANIFOLDS = 4): xm of them in the table's tail entries [xs, xs + xm),
    // which carry the pair's col and their own contact cache
    int *ct0, *ctn, *xm, *xs;
};

// A concave collider's shape (ConcaveMeshShape / HeightFieldShape) with its scale; the collider's
// shape.x is its local-bounds slot (after the hulls' slots), shape.y the index of this record.
struct ConcaveInst {
    int kind;                  // SHAPE_CONCAVE_MESH / SHAPE_HEIGHTFIELD
    int vbase, tbase, nbase;   // mesh: first vertex, triangle and scaled vertex normal in the arenas
    int node_base, root;       // mesh: its AABB tree nodes (two float4 each) and root
    int columns, rows, hbase;  // height field: grid and first height
    int pad;
    float sx, sy, sz;          // scale
    float isx, isy, isz;       // decimal(1.0) / scale
    float width, length, origin;              // height field (HeightField.cpp:56-99)
    float bminx, bminy, bminz, bmaxx, bmaxy, bmaxz;   // height field bounds, unscaled
};

constexpr int EV_CUR = 1;
constexpr int EV_PREV = 2;
// contact events (CollisionCallback::ContactPair::EventType, CollisionCallback.h)
constexpr int EV_START = 0;
constexpr int EV_STAY = 1;
constexpr int EV_EXIT = 2;
// trigger events (OverlapCallback::OverlapPair::EventType + 3)
constexpr int EV_TRIGGER_START = 3;
constexpr int EV_TRIGGER_STAY = 4;
constexpr int EV_TRIGGER_EXIT = 5;

// Solver rows of one contact manifold (ContactManifoldSolver + ContactPointSolver), colour order.
struct DevSolver {
    int *pair;                 // pair slot of each manifold in solve order
    int2 *bodies;              // rigid body indices
    int *npts;
    float4 *n;                 // manifold normal xyz, w = friction coefficient
    float4 *r1f, *r2f;         // r1Friction, r2Friction; w: inverse masses body1 / body2
    float4 *t1, *t2;           // friction vectors; w: inverse friction masses 1 / 2
    float4 *imp;               // friction1, friction2, twist impulses; w: inverse twist mass
    float4 *oldt1, *oldt2;     // previous friction vectors
    // per point [4*m + k]
    float4 *pn;                // normal xyz, penetration depth
    float4 *r1, *r2;           // w: inverse penetration mass | restitution bias
    float4 *lam;               // penetration impulse, split impulse, resting flag (warm-start transfer)
    float4 *pred;              // learned warm start: predicted normal impulse, friction 1, friction 2
};

struct DevJoints;  // joint solver rows (joints.hip)

struct Settings {
    V3 gravity;
    float persistent_threshold;
    float restitution_threshold;
    float cos_similar;
    int vel_iters, pos_iters;
    int gravity_enabled;
    int sleeping_enabled;
    float sleep_lin_sq, sleep_ang_sq;   // squared sleep velocities (PhysicsWorld.cpp:875-876)
    float time_before_sleep;
    int split_impulse;                  // contacts: SPLIT_IMPULSES (1, the default) or BAUMGARTE_CONTACTS (0)
};

// Counters written by kernels (one device block, read back after each step).
constexpr int kColorRoundsMax = 48;   // parallel colouring rounds launched at most per step

// Per-step counters, cleared by a fill kernel at the start of every step and read back after it.
struct DevCounters {
    int nb_pairs;              // pairs in the current table
    int nb_overlaps;           // overlapping pairs found by the traversal (may exceed the capacity)
    int nb_tests[ALGO_COUNT];
    int nb_manifolds;
    int nb_points;
    int nb_colors;
    int status;
    int finish_rounds;         // rounds taken by the single-workgroup colouring finisher
    int nb_new_points;         // contact points without a cached impulse
    int color_left[kColorRoundsMax];   // manifolds still uncoloured after each parallel round
    int nb_box_tests;          // box-box subset of nb_tests[ALGO_POLY_POLY] that collided last step
    int nb_box_sep_tests;      // box-box pairs that did not (cheap temporal-coherence rejects, listed apart)
    int nb_hull_tests;         // the rest of nb_tests[ALGO_POLY_POLY] (a convex mesh involved)
    int nb_walk_again;         // colliders with more partners than the traversal slab holds
    int nb_events;             // contact events of the step (may exceed the event buffer)
    int nb_concave_tests;      // triangle tests of the convex vs concave pairs
    int nb_concave_points;     // contact points of those tests
    int nb_tails;              // contact manifolds of concave pairs beyond their first
};

constexpr int STATUS_PAIR_OVERFLOW = 1;
constexpr int STATUS_COLOR_OVERFLOW = 2;
constexpr int STATUS_STACK_OVERFLOW = 4;
constexpr int STATUS_COLOR_STALL = 8;
constexpr int STATUS_CLIP_OVERFLOW = 16;   // a face clip produced more vertices than its buffer holds
constexpr int STATUS_CONCAVE_OVERFLOW = 32;   // concave pairs: triangle tests, their points or manifolds over capacity

}  // namespace rp3dg
