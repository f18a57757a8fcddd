// Islands and sleeping (PhysicsWorld::createIslands / updateSleepingBodies, PhysicsWorld.cpp:695-926).
//
// The reference grows islands by a depth-first search from every awake non-static body over its contact
// pairs of this step and all its joints, never crossing a static body, and wakes every sleeping body
// the search reaches.  Which bodies end up in one island does not depend on the search order: an
// island is a connected component of the graph whose nodes are the non-static bodies and whose edges
// are the colliding pairs and the joints, and it is searched (and woken) iff it holds a body that was
// awake.  The device builds the components with a lock-free union-find (hooking the larger root under
// the smaller with a compare-and-swap), marks the components that hold an awake body and wakes their
// sleeping bodies.  The sleep test at the end of the step takes each island's minimum sleep time with
// an atomic minimum on the float bits (sleep times are non-negative, so their bit patterns order like
// the floats).
#include <hip/hip_runtime.h>

#include "islands.h"
#include "kcommon.h"

namespace rp3dg {

__device__ __forceinline__ bool non_static(int f) { return (f & (BF_DYNAMIC | BF_KINEMATIC)) != 0; }

__device__ __forceinline__ int uf_load(const int* p, int i) {
    return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// root of x, halving the path on the way (a node that is not a root never becomes one again, so
// pointing it at its grandparent keeps it in the same tree whatever other threads do)
__device__ int uf_find(int* parent, int x) {
    while (true) {
        const int p = uf_load(parent, x);
        if (p == x) return x;
        const int gp = uf_load(parent, p);
        if (gp != p) __hip_atomic_store(parent + x, gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = gp;
    }
}

// every failed compare-and-swap means another thread hooked the root first, and every hook removes one
// root, so the loop ends
__device__ void uf_union(int* parent, int a, int b) {
    while (true) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(parent + a, a, b) == a) return;
    }
}

__global__ void isl_init(DevBodies B, DevIslands I) {
    RP3DG_GRID_STRIDE(b, B.n) {
        I.parent[b] = b;
        I.minst[b] = 0x7f7fffffu;   // DECIMAL_LARGEST (PhysicsWorld.cpp:883)
        I.active[b] = 0;
        B.flags[b] &= ~BF_WOKEN;
    }
}

// contact pairs of this step (manifold list) between two non-static bodies
__global__ void isl_link_manifolds(DevBodies B, DevColliders C, DevPairs P, const int* mlist, const int* nm, DevIslands I) {
    const int n = *nm;
    RP3DG_GRID_STRIDE(i, n) {
        const int2 c = P.col[mlist[i]];
        const int b1 = C.body[c.x], b2 = C.body[c.y];
        if (non_static(B.flags[b1]) && non_static(B.flags[b2])) uf_union(I.parent, b1, b2);
    }
}

__global__ void isl_link_joints(DevBodies B, DevJoints J, DevIslands I) {
    RP3DG_GRID_STRIDE(j, J.n) {
        const int2 bb = J.bodies[j];
        if (non_static(B.flags[bb.x]) && non_static(B.flags[bb.y])) uf_union(I.parent, bb.x, bb.y);
    }
}

// after all unions: point every body at its root and flag the roots of islands with an awake body
__global__ void isl_mark(DevBodies B, DevIslands I) {
    RP3DG_GRID_STRIDE(b, B.n) {
        const int f = B.flags[b];
        if (!non_static(f)) continue;
        const int r = uf_find(I.parent, b);
        if (f & BF_ENABLED) I.active[r] = 1;
    }
}

// RigidBody::setIsSleeping(false) for the sleeping bodies of searched islands (RigidBody.cpp:1006-1057):
// sleep time 0, enabled again, colliders into the moved set
__global__ void isl_wake(DevBodies B, DevIslands I) {
    RP3DG_GRID_STRIDE(b, B.n) {
        const int f = B.flags[b];
        if (!non_static(f)) continue;
        const int r = uf_find(I.parent, b);
        __hip_atomic_store(I.parent + b, r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!(f & BF_ENABLED) && I.active[r]) {
            I.sleep_time[b] = 0.0f;
            B.flags[b] = f | BF_ENABLED | BF_WOKEN;
        }
    }
}

// PhysicsWorld::updateSleepingBodies (PhysicsWorld.cpp:871-910): the bodies of this step's islands are
// the enabled ones; static bodies are skipped
__global__ void isl_sleep_time(DevBodies B, DevIslands I, Settings S, float dt) {
    RP3DG_GRID_STRIDE(b, B.n) {
        const int f = B.flags[b];
        if (!(f & BF_ENABLED)) continue;
        const V3 v = ld3(B.v, b), w = ld3(B.w, b);
        float t;
        if (len2(v) > S.sleep_lin_sq || len2(w) > S.sleep_ang_sq || !(f & BF_ALLOW_SLEEP)) {
            t = 0.0f;
        } else {
            t = I.sleep_time[b] + dt;
        }
        I.sleep_time[b] = t;
        atomicMin(I.minst + I.parent[b], __float_as_uint(t));
    }
}

// islands whose minimum sleep time reached the time before sleep go to sleep (PhysicsWorld.cpp:912-925,
// RigidBody.cpp:1021-1044): sleep time 0, disabled, velocities zero (forces are cleared with everyone's)
__global__ void isl_sleep(DevBodies B, DevIslands I, Settings S) {
    RP3DG_GRID_STRIDE(b, B.n) {
        const int f = B.flags[b];
        if (!(f & BF_ENABLED)) continue;
        if (__uint_as_float(I.minst[I.parent[b]]) >= S.time_before_sleep) {
            I.sleep_time[b] = 0.0f;
            B.flags[b] = f & ~BF_ENABLED;
            B.v[b] = make_float4(0.f, 0.f, 0.f, 0.f);
            B.w[b] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

// A pair whose two bodies are now disabled (static or asleep) is moved into the disabled pair array,
// and moved back when one of them wakes; both moves rebuild the pair, so its LastFrameCollisionInfo
// restarts from the default (OverlappingPairs.cpp:301-346).  The narrowphase does not touch a pair
// while both bodies are disabled, so clearing it once at the end of every step covers both moves.
__global__ void isl_reset_disabled_pairs(DevBodies B, DevColliders C, DevPairs P, const int* nb_pairs) {
    const int n = *nb_pairs;
    RP3DG_GRID_STRIDE(k, n) {
        const int2 c = P.col[k];
        if (!((B.flags[C.body[c.x]] | B.flags[C.body[c.y]]) & BF_ENABLED)) {
            P.lf[k] = 0u;
            P.lfaxis[k] = make_float4(0.f, 1.f, 0.f, 0.f);   // LastFrameCollisionInfo() default axis
        }
    }
}

}  // namespace rp3dg
