// Contact events (CollisionDetectionSystem::reportContacts and the colliding flags of
// OverlappingPairs, CD.cpp:174-234, :380-441, :975-1011, :1224-1303, :1708-1741;
// OverlappingPairs.cpp:465-482).
//
// Pairs with a trigger collider report the same Start / Stay / Exit sequence to onTrigger
// (CollisionDetectionSystem::reportTriggers, CD.cpp:1745-1757) as EV_TRIGGER_* events, without points.
// Every pair keeps the reference's two flags: collidingInCurrentFrame (EV_CUR, set by this step's
// narrowphase on the pairs it tests) and collidingInPreviousFrame (EV_PREV, copied from EV_CUR by
// updateCollidingInPreviousFrame on the enabled pairs after the islands are built).  Events of a
// step: a colliding pair reports ContactStart or ContactStay by EV_PREV; a "lost" pair reports
// ContactExit, either because the broadphase removed it while EV_PREV was set or because the
// narrowphase tested it and found no contact while EV_PREV was set.  Events are appended in any
// order (the reference lists them in island order, then the lost pairs).
#include <hip/hip_runtime.h>

#include "events.h"
#include "kcommon.h"

namespace rp3dg {

// wave-aggregated slot reservation: one atomic per wave
__device__ __forceinline__ void ev_push(const DevEvents& E, bool want, int4 e) {
    const unsigned long long m = __ballot(want);
    if (m == 0ull) return;
    const int lane = lane_id();
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(E.count, __popcll(m));
    base = __shfl(base, leader);
    if (want) {
        const int slot = base + __popcll(m & ((1ull << lane) - 1ull));
        if (slot < E.cap) E.ev[slot] = e;
    }
}

__device__ __forceinline__ bool is_trigger_pair(const DevColliders& C, int2 c) {
    return ((C.kind[c.x] | C.kind[c.y]) & CK_TRIGGER) != 0;
}

// pairs of the previous table that the broadphase removed this step (CD.cpp:189-202); a trigger pair
// reports OverlapExit to onTrigger instead (OverlapCallback.cpp:79-95)
__global__ void ev_removed(DevColliders C, DevPairs O, const int* nold, DevEvents E) {
    const int n = *nold;
    RP3DG_GRID_STRIDE(k, n) {
        const bool lost = !O.kept[k] && (O.evf[k] & EV_PREV);
        const int2 c = O.col[k];
        const int type = lost && is_trigger_pair(C, c) ? EV_TRIGGER_EXIT : EV_EXIT;
        ev_push(E, lost, make_int4(c.x, c.y, type, 0));
    }
}

// A world query's broadphase removes pairs like a step's does; the reference keeps them in its lost contact
// pairs until the next step reports them (CD.cpp:158-170, 1723-1725).  Run right after the step clears
// its counters, in one workgroup.
__global__ void ev_take_pending(const int4* pend, int* pend_count, int4* ev, int cap, int* count) {
    const int n = min(*pend_count, cap);
    for (int i = threadIdx.x; i < n; i += blockDim.x) ev[i] = pend[i];
    __syncthreads();
    if (threadIdx.x == 0) {
        *count = *pend_count;
        *pend_count = 0;
    }
}

// after the narrowphase, before the islands: the tested pairs (a body enabled) get EV_CUR; colliding
// pairs report Start / Stay, tested pairs that stopped colliding report Exit (CD.cpp:975-996)
// shift = 1 (no sleeping: the enabled pairs cannot change before the islands) also does ev_shift's work
__global__ void ev_update(DevBodies B, DevColliders C, DevPairs N, const int* nb_pairs, DevEvents E, int shift) {
    const int n = *nb_pairs;
    RP3DG_GRID_STRIDE(k, n) {
        const int2 c = N.col[k];
        // the middle phase clears collidingInCurrentFrame of every pair with an enabled body (CD.cpp:389) and
        // tests those whose colliders are both simulation colliders or triggers
        const bool tested = ((B.flags[C.body[c.x]] | B.flags[C.body[c.y]]) & BF_ENABLED) != 0;
        const int k1 = C.kind[c.x], k2 = C.kind[c.y];
        const bool active = (k1 & (CK_SIMULATION | CK_TRIGGER)) && (k2 & (CK_SIMULATION | CK_TRIGGER));
        const bool trig = ((k1 | k2) & CK_TRIGGER) != 0;
        int np = N.ncp[k];
        if (N.xm && np > 0)   // a concave pair's points: its first manifold's and its tail manifolds'
            for (int j = 0; j < N.xm[k]; j++) np += N.ncp[N.xs[k] + j];
        const int f = N.evf[k];
        // a trigger pair collides without contact points: its test result is the wasColliding flag
        const bool hit = trig ? (N.lf[k] & LF_WAS_COLLIDING) != 0 : np > 0;
        const bool cur = tested ? (active && hit) : (f & EV_CUR) != 0;
        const bool prev = (f & EV_PREV) != 0;
        if (tested) N.evf[k] = shift ? (cur ? EV_CUR | EV_PREV : 0) : ((prev ? EV_PREV : 0) | (cur ? EV_CUR : 0));
        if (E.record) {
            const bool emit = tested && (cur || prev);
            int type = cur ? (prev ? EV_STAY : EV_START) : EV_EXIT;
            if (trig) type += EV_TRIGGER_START;   // onTrigger: OverlapStart / OverlapStay / OverlapExit
            ev_push(E, emit, make_int4(c.x, c.y, type, cur && !trig ? np : 0));
        }
    }
}

// updateCollidingInPreviousFrame on the pairs enabled after the islands woke their bodies
__global__ void ev_shift(DevBodies B, DevColliders C, DevPairs N, const int* nb_pairs) {
    const int n = *nb_pairs;
    RP3DG_GRID_STRIDE(k, n) {
        const int2 c = N.col[k];
        if ((B.flags[C.body[c.x]] | B.flags[C.body[c.y]]) & BF_ENABLED)
            N.evf[k] = (N.evf[k] & EV_CUR) ? (EV_CUR | EV_PREV) : 0;
    }
}

}  // namespace rp3dg
