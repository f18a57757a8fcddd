// Islands and sleeping (PhysicsWorld::createIslands / updateSleepingBodies, PhysicsWorld.cpp:695-926).
#pragma once

#include "joints.h"
#include "rp3d_internal.h"

namespace rp3dg {

// Per-body island state, allocated when the worlds have sleeping enabled.
struct DevIslands {
    int* parent;          // union-find forest over the non-static bodies; a root is its island's id
    unsigned int* minst;  // per root: minimum sleep time of the island (float bits, non-negative)
    int* active;          // per root: the island holds a body that was awake when the step began
    float* sleep_time;    // RigidBodyComponents::mSleepTimes
};

__global__ void isl_init(DevBodies B, DevIslands I);
__global__ void isl_link_manifolds(DevBodies B, DevColliders C, DevPairs P, const int* mlist, const int* nm, DevIslands I);
__global__ void isl_link_joints(DevBodies B, DevJoints J, DevIslands I);
__global__ void isl_mark(DevBodies B, DevIslands I);
__global__ void isl_wake(DevBodies B, DevIslands I);
__global__ void isl_sleep_time(DevBodies B, DevIslands I, Settings S, float dt);
__global__ void isl_sleep(DevBodies B, DevIslands I, Settings S);
__global__ void isl_reset_disabled_pairs(DevBodies B, DevColliders C, DevPairs P, const int* nb_pairs);

}  // namespace rp3dg
