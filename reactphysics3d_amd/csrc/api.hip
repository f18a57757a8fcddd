// RigidBody state changes between steps, on the device copy of a batch (RigidBody.cpp:151-239, 298-311,
// 794-866, 938-987, 1006-1057).  The host queues the calls in the order the application makes them and
// one thread applies them in that order before the next step or read-back, so a sequence of calls on one
// body composes exactly as the reference's calls do (each reads the state the previous one left).
#include <hip/hip_runtime.h>

#include "api.h"
#include "kcommon.h"
#include "refit_dev.h"

namespace rp3dg {

// RigidBody::setIsSleeping (RigidBody.cpp:1006-1057) for a non-static, active body.
__device__ void api_set_sleeping(const ApiArgs& A, int b, bool sleep) {
    const DevBodies& B = A.B;
    const int f = B.flags[b];
    const bool asleep = !(f & BF_ENABLED);
    if (asleep == sleep) return;
    if (!(f & (BF_DYNAMIC | BF_KINEMATIC))) return;   // a static body is always awake
    if (A.sleep_time) A.sleep_time[b] = 0.0f;
    if (sleep) {
        B.flags[b] = f & ~BF_ENABLED;
        B.v[b] = make_float4(0.f, 0.f, 0.f, 0.f);
        B.w[b] = make_float4(0.f, 0.f, 0.f, 0.f);
        B.force[b] = make_float4(0.f, 0.f, 0.f, 0.f);
        B.torque[b] = make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        B.flags[b] = f | BF_ENABLED;
        // askForBroadPhaseCollisionCheck: the body's colliders enter the moved set
        for (int k = A.coll_offs[b]; k < A.coll_offs[b + 1]; k++) A.C.moved[A.coll_list[k]] = 1;
    }
}

__global__ void api_apply_ops(ApiArgs A, const ApiOp* ops, int n) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const DevBodies& B = A.B;
    for (int k = 0; k < n; k++) {
        const ApiOp o = ops[k];
        const int b = o.body;
        const int f = B.flags[b];
        const bool dynamic = (f & BF_DYNAMIC) != 0;
        const bool is_static = !(f & (BF_DYNAMIC | BF_KINEMATIC));
        const Q4 q = ldq(B.q, b);
        V3 a = xyz(o.a);
        switch (o.op) {
            case API_SET_TRANSFORM: {   // RigidBody::setTransform (RigidBody.cpp:837-866)
                const Tf t = tf(a, q4(o.b.x, o.b.y, o.b.z, o.b.w));
                const V3 old = ld3(B.x, b);
                const V3 com = tapply(t, ld3(B.com, b));
                B.x[b] = f4(com, 0.f);
                V3 v = ld3(B.v, b);
                v += cross(ld3(B.w, b), com - old);
                st3(B.v, b, v);
                api_set_sleeping(A, b, false);
                B.p[b] = f4(t.p, 0.f);
                B.q[b] = f4q(t.q);
                // Body::setTransform -> updateBroadPhaseState (Body.cpp:226-241, 397-407): new collider
                // transforms, a new fat AABB and a place in the moved set for a shape that left its fat AABB
                for (int j = A.coll_offs[b]; j < A.coll_offs[b + 1]; j++) refit_one(B, A.C, A.hull_bounds, 2, A.coll_list[j]);
                break;
            }
            case API_SET_LINEAR_VELOCITY:   // RigidBody::setLinearVelocity (RigidBody.cpp:794-808)
            case API_SET_ANGULAR_VELOCITY:  // RigidBody::setAngularVelocity (RigidBody.cpp:814-828)
                if (is_static) break;
                if (o.op == API_SET_LINEAR_VELOCITY) st3(B.v, b, a);
                else st3(B.w, b, a);
                if (len2(a) > 0.0f) api_set_sleeping(A, b, false);
                break;
            case API_APPLY_FORCE: {   // RigidBody::apply{Local,World}ForceAt{CenterOfMass,LocalPosition,WorldPosition}
                if (o.flags & API_FORCE_LOCAL) a = qrot(q, a);   // the body orientation times the local force
                if (!dynamic) break;
                api_set_sleeping(A, b, false);
                st3(B.force, b, ld3(B.force, b) + a);
                if (o.flags & API_AT_POINT) {
                    V3 p = xyz(o.b);
                    if (o.flags & API_POINT_LOCAL) p = tapply(tf(ld3(B.p, b), q), p);
                    st3(B.torque, b, ld3(B.torque, b) + cross(p - ld3(B.x, b), a));
                }
                break;
            }
            case API_APPLY_TORQUE:   // RigidBody::applyWorldTorque / applyLocalTorque (RigidBody.cpp:938-967)
                if (o.flags & API_FORCE_LOCAL) a = qrot(q, a);
                if (!dynamic) break;
                api_set_sleeping(A, b, false);
                st3(B.torque, b, ld3(B.torque, b) + a);
                break;
            case API_RESET_FORCE:    // RigidBody::resetForce / resetTorque (RigidBody.cpp:970-987)
                if (dynamic) B.force[b] = make_float4(0.f, 0.f, 0.f, 0.f);
                break;
            case API_RESET_TORQUE:
                if (dynamic) B.torque[b] = make_float4(0.f, 0.f, 0.f, 0.f);
                break;
            case API_SET_SLEEPING:
                api_set_sleeping(A, b, o.flags != 0);
                break;
            default:
                break;
        }
    }
}

}  // namespace rp3dg
