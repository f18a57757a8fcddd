// Collider world transform and fat-AABB refit, one collider per call.  Used by the refit kernel
// (broadphase.hip) and by the fused per-world position kernel (fused.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kcommon.h"
#include "rp3d_internal.h"

namespace rp3dg {

// ---------------------------------------------------------------------------------------------
// Collider world transforms and fat AABB refit (DynamicsSystem.cpp:104-109, BroadPhaseSystem.cpp
// :155-185, CollisionShape.cpp:54-100, SphereShape.cpp:49-58, DynamicAABBTree.cpp:146-214).
// ---------------------------------------------------------------------------------------------

__device__ inline void local_bounds(int type, float4 s, const float4* hull_bounds, V3& lo, V3& hi) {
    if (type == SHAPE_CAPSULE) {  // CapsuleShape.cpp:90-94
        lo = v3(-s.x, -s.y - s.x, -s.x);
        hi = v3(s.x, s.y + s.x, s.x);
    } else if (type == SHAPE_BOX) {  // BoxShape.cpp:149-151
        lo = v3(-s.x, -s.y, -s.z);
        hi = v3(s.x, s.y, s.z);
    } else {  // ConvexMeshShape.cpp:208-212 (scale 1)
        const int h = (int)s.x;
        lo = xyz(hull_bounds[2 * h]);
        hi = xyz(hull_bounds[2 * h + 1]);
    }
}

__device__ inline void transformed_aabb(int type, float4 s, const float4* hull_bounds, const Tf& t, V3& mn, V3& mx) {
    if (type == SHAPE_SPHERE) {  // SphereShape::computeTransformedAABB
        const V3 e = v3(s.x, s.x, s.x);
        mn = t.p - e;
        mx = t.p + e;
        return;
    }
    V3 lo, hi;
    local_bounds(type, s, hull_bounds, lo, hi);
    const M3 m = qmat(t.q);
    for (int i = 0; i < 3; i++) {
        float rmin = comp(t.p, i), rmax = comp(t.p, i);
        for (int j = 0; j < 3; j++) {
            const float e = mel(m, i, j) * comp(lo, j);
            const float f = mel(m, i, j) * comp(hi, j);
            if (e < f) { rmin += e; rmax += f; }
            else { rmin += f; rmax += e; }
        }
        setcomp(mn, i, rmin);
        setcomp(mx, i, rmax);
    }
}

__device__ inline void fatten(V3 mn, V3 mx, V3& fmn, V3& fmx) {
    // DynamicAABBTree.cpp:153 / :203: gap = extent * 0.08 * 0.5
    const V3 ext = mx - mn;
    const V3 gap = (ext * 0.08f) * 0.5f;
    fmn = mn - gap;
    fmx = mx + gap;
}

// init = 1: collider creation (RigidBody::addCollider -> BroadPhaseSystem::addCollider), every collider
// gets a fat AABB and enters the moved set.  init = 0: end-of-step refit of enabled colliders.
// init = 2: a body moved or woken between two steps (Body::updateBroadPhaseState, Body.cpp:226-241):
// the colliders of enabled bodies are refit and join the moved set, which keeps what the last step's
// refit put there.
__device__ __forceinline__ void refit_one(const DevBodies& B, const DevColliders& C, const float4* hull_bounds, int init,
                                          int i) {
    {
        const int b = C.body[i];
        const int bf = B.flags[b];
        int moved = 0;
        if (init == 1 || (bf & BF_ENABLED)) {
            const Tf bt = tf(ld3(B.p, b), ldq(B.q, b));
            const Tf lt = tf(ld3(C.lpos, i), ldq(C.lq, i));
            const Tf wt = tmul(bt, lt);
            C.wpos[i] = f4(wt.p, 0.f);
            C.wq[i] = f4q(wt.q);
            V3 mn, mx;
            transformed_aabb(C.type[i], C.shape[i], hull_bounds, wt, mn, mx);
            bool inside = false;
            if (init != 1) {
                const V3 fmn = ld3(C.amin, i), fmx = ld3(C.amax, i);
                // AABB::contains(const AABB&) (AABB.h:246-256)
                inside = fmn.x <= mn.x && fmn.y <= mn.y && fmn.z <= mn.z && fmx.x >= mx.x && fmx.y >= mx.y &&
                         fmx.z >= mx.z;
            }
            if (!inside) {
                V3 fmn, fmx;
                fatten(mn, mx, fmn, fmx);
                C.amin[i] = f4(fmn, 0.f);
                C.amax[i] = f4(fmx, 0.f);
                moved = 1;
            }
            // a body woken this step put its colliders into the moved set without a new fat AABB
            // (RigidBody::askForBroadPhaseCollisionCheck)
            if (bf & BF_WOKEN) moved = 1;
        }
        if (init == 2) {
            if (moved) C.moved[i] = 1;
        } else {
            C.moved[i] = moved;
        }
    }
}

}  // namespace rp3dg
