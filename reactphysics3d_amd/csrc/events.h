// Contact events: the CollisionCallback::CallbackData an EventListener::onContact receives each step
// (CollisionDetectionSystem::reportContacts, CD.cpp:1708-1741; CollisionCallback.cpp:69-76).
#pragma once

#include "rp3d_internal.h"

namespace rp3dg {

// One event: collider1, collider2 (as the pair orients them), EV_START / EV_STAY / EV_EXIT (contact pairs) or
// EV_TRIGGER_START / _STAY / _EXIT (pairs with a trigger collider), contact points.
struct DevEvents {
    int4* ev;
    int cap;
    int* count;
    int record;   // 0: only the per-pair colliding flags are kept up to date
};

__global__ void ev_removed(DevColliders C, DevPairs O, const int* nold, DevEvents E);
__global__ void ev_update(DevBodies B, DevColliders C, DevPairs N, const int* nb_pairs, DevEvents E, int shift);
__global__ void ev_shift(DevBodies B, DevColliders C, DevPairs N, const int* nb_pairs);
// the exits recorded by world queries since the last step open this step's events
__global__ void ev_take_pending(const int4* pend, int* pend_count, int4* ev, int cap, int* count);

}  // namespace rp3dg
