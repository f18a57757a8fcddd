// RigidBody state changes between steps (api.hip): the queued operation and the state it touches.
#pragma once

#include "rp3d_internal.h"

namespace rp3dg {

enum ApiOpCode : int {
    API_SET_TRANSFORM = 1,         // a = position, b = orientation
    API_SET_LINEAR_VELOCITY = 2,   // a
    API_SET_ANGULAR_VELOCITY = 3,  // a
    API_APPLY_FORCE = 4,           // a = force, b = point (flags API_AT_POINT)
    API_APPLY_TORQUE = 5,          // a = torque
    API_RESET_FORCE = 6,
    API_RESET_TORQUE = 7,
    API_SET_SLEEPING = 8,          // flags = 1: sleep, 0: wake
};
constexpr int API_FORCE_LOCAL = 1;   // force / torque given in the body's local frame
constexpr int API_AT_POINT = 2;      // the force acts at point b (else at the centre of mass)
constexpr int API_POINT_LOCAL = 4;   // b is in the body's local frame

struct ApiOp {
    int op, body, flags, pad;
    float4 a, b;
};

struct ApiArgs {
    DevBodies B;
    DevColliders C;
    const float4* hull_bounds;
    const int* coll_offs;   // colliders of body b: coll_list[coll_offs[b] .. coll_offs[b + 1])
    const int* coll_list;
    float* sleep_time;      // RigidBodyComponents::mSleepTimes, or nullptr when sleeping is disabled
};

__global__ void api_apply_ops(ApiArgs A, const ApiOp* ops, int n);

}  // namespace rp3dg
