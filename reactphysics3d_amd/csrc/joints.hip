// Per-colour joint kernels (one launch per colour of one joint type).  The rows
// themselves are in joints_dev.h; the fused per-world solver calls the same rows.
#include <hip/hip_runtime.h>

#include "joints.h"
#include "joints_dev.h"
#include "kcommon.h"

namespace rp3dg {

__global__ void jt_init(DevBodies B, DevJoints J, float dt) {
    RP3DG_GRID_STRIDE(j, J.n) jt_init_one(B, J, j, dt);
}

__global__ void jt_warmstart(DevBodies B, DevJoints J, const int* range, int slot) {
    const int lo = range[2 * slot], hi = range[2 * slot + 1];
    RP3DG_GRID_STRIDE(t, hi - lo) jt_warmstart_one(B, J, lo + t);
}

__global__ void jt_solve_velocity(DevBodies B, DevJoints J, const int* range, int slot, float dt) {
    const int lo = range[2 * slot], hi = range[2 * slot + 1];
    RP3DG_GRID_STRIDE(t, hi - lo) jt_solve_velocity_one(B, J, lo + t, dt);
}

__global__ void jt_solve_position(DevBodies B, DevJoints J, const int* range, int slot) {
    const int lo = range[2 * slot], hi = range[2 * slot + 1];
    RP3DG_GRID_STRIDE(t, hi - lo) jt_solve_position_one(B, J, lo + t);
}

}  // namespace rp3dg
