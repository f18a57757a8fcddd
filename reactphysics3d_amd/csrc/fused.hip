// Fused per-world solver for batches of many small worlds.
//
// One workgroup owns one group of consecutive worlds for the whole solver phase
// of a step: velocity integration of its bodies, contact rows, joint rows, warm
// start, every velocity iteration (colours and joint slots separated by
// workgroup barriers), impulse store; a second kernel does position
// integration, the joint position iterations, the state update and the
// fat-AABB refit.  Groups share nothing, so no grid-wide synchronisation and no
// per-colour launches are needed.  Colour c of a group is colour c of each of
// its worlds side by side (worlds share no body), so every world runs its
// sweeps in the same order as the per-colour kernels (joint slots, then contact
// colours) and results are identical; packing worlds fills the wavefronts of
// the colour phases, which a single small world leaves mostly idle.
//
// In residual mode a world iterates until the velocity residual
//   r_k = max over its dynamic bodies of max(|dv|_inf, |dw|_inf) between sweeps
// drops below the threshold (or the iteration cap is reached); a converged
// world of a group skips the remaining sweeps of the others.  This mode has no
// counterpart in the reference, which always runs a fixed iteration count
// (PhysicsWorld.cpp:334).
#include <hip/hip_runtime.h>

#include "fused.h"
#include "joints_dev.h"
#include "kcommon.h"
#include "refit_dev.h"
#include "solver_dev.h"

namespace rp3dg {

__device__ __forceinline__ float block_maxf(float v, float* red) {
    v = wave_maxf(v);
    const int wid = threadIdx.x / kWave;
    const int nwv = blockDim.x / kWave;
    if (lane_id() == 0) red[wid] = v;
    __syncthreads();
    float r = red[0];
    for (int k = 1; k < nwv; k++) r = red[k] > r ? red[k] : r;
    __syncthreads();
    return r;
}

// Warm start, velocity iterations and impulse store of one group.  Instantiated twice: with the body
// state in LDS (A.B's velocity, inertia and flag arrays point into the workgroup's LDS, so every access
// compiles to an LDS instruction) and with it in HBM (groups larger than the LDS budget).
// PV / PW: the velocities after the previous sweep (residual mode), indexed by body - b0, in LDS or HBM.
template <bool SPLIT>
__device__ __forceinline__ void group_sweeps(const FusedArgs& F, const SvArgs& A, const DevBodies& B, bool staged,
                                             int w0, int ngw, int b0, int b1, int m1, const int* seg, const int2* jr,
                                             float* red, unsigned int* s_wres, unsigned long long* s_done,
                                             float4* PV, float4* PW) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int nbw = b1 - b0;
    const DevBodies& BS = A.B;
    // contact warm start, colour by colour (CS.cpp:326-476)
    for (int c = 0; c < kMaxColors; c++) {
        const int lo = seg[c], hi = seg[c + 1];
        if (lo == m1) break;
        for (int i = lo + tid; i < hi; i += nt) sv_warmstart_one(A, i);
        __syncthreads();
    }
    // joint warm start, slot by slot
    for (int s = 0; s < F.jt_slots; s++) {
        const int2 r = jr[s];
        if (r.x == r.y) continue;
        for (int j = r.x + tid; j < r.y; j += nt) jt_warmstart_one(BS, F.J, j);
        __syncthreads();
    }
    const bool residual = F.res_thr > 0.0f;
    // residual mode with several worlds: each world stops on its own residual (mask of converged worlds)
    const bool masked = residual && ngw > 1;
    const unsigned long long all = ngw >= 64 ? ~0ull : (1ull << ngw) - 1ull;
    unsigned long long done = 0;
    // world of body b within the group
    auto wk = [&](int b) { return staged ? (BS.flags[b - BS.vbase] >> kGroupShift) & 63 : B.world[b] - w0; };
    if (residual) {
        for (int i = b0 + tid; i < b1; i += nt) {
            PV[i - b0] = BS.cv[i - BS.vbase];
            PW[i - b0] = BS.cw[i - BS.vbase];
        }
        if (masked)
            for (int k = tid; k < ngw; k += nt) {
                s_wres[k] = 0u;
                F.world_iters[w0 + k] = 0;
                F.world_res[w0 + k] = 0.0f;
            }
        __syncthreads();
    }
    int it = 0;
    float res = 0.0f;
    while (it < F.max_iters) {
        for (int s = 0; s < F.jt_slots; s++) {
            const int2 r = jr[s];
            if (r.x == r.y) continue;
            for (int j = r.x + tid; j < r.y; j += nt) {
                if (masked && ((done >> wk(F.J.bodies[j].x)) & 1ull)) continue;
                jt_solve_velocity_one(BS, F.J, j, F.dt);
            }
            __syncthreads();
        }
        for (int c = 0; c < kMaxColors; c++) {
            const int lo = seg[c], hi = seg[c + 1];
            if (lo == m1) break;
            for (int i = lo + tid; i < hi; i += nt) {
                if (masked && ((done >> wk(A.S.bodies[i].x)) & 1ull)) continue;
                solve_manifold<SPLIT>(A, i, F.beta_dt);
            }
            __syncthreads();
        }
        it++;
        if (residual) {
            float r = 0.0f;
            for (int i = b0 + tid; i < b1; i += nt) {
                if (!dyn_body(B, i)) continue;
                int k = 0;
                if (masked) {
                    k = wk(i);
                    if ((done >> k) & 1ull) continue;
                }
                const float4 v = BS.cv[i - BS.vbase], wv = BS.cw[i - BS.vbase], pv = PV[i - b0], pw = PW[i - b0];
                float d = fabsf(v.x - pv.x);
                d = fmaxf(d, fabsf(v.y - pv.y));
                d = fmaxf(d, fabsf(v.z - pv.z));
                d = fmaxf(d, fabsf(wv.x - pw.x));
                d = fmaxf(d, fabsf(wv.y - pw.y));
                d = fmaxf(d, fabsf(wv.z - pw.z));
                if (masked) atomicMax(&s_wres[k], __float_as_uint(d));   // d >= 0: bit order = value order
                else r = fmaxf(r, d);
                PV[i - b0] = v;
                PW[i - b0] = wv;
            }
            if (!masked) {
                res = block_maxf(r, red);
                if (res < F.res_thr) break;
            } else {
                __syncthreads();
                if (tid < kWave) {
                    bool fin = false;
                    if (tid < ngw && !((done >> tid) & 1ull)) {
                        const float rk = __uint_as_float(s_wres[tid]);
                        s_wres[tid] = 0u;
                        fin = rk < F.res_thr;
                        F.world_iters[w0 + tid] = it;
                        F.world_res[w0 + tid] = rk;
                    }
                    const unsigned long long bal = __ballot(fin);
                    if (tid == 0) *s_done = done | bal;
                }
                __syncthreads();
                done = *s_done;
                if (done == all) break;
            }
        }
    }
    // storeImpulses (CS.cpp:763-785)
    for (int i = seg[0] + tid; i < m1; i += nt) sv_store_one(A, i);
    if (staged) {
        __syncthreads();
        for (int i = tid; i < nbw; i += nt) {
            B.cv[b0 + i] = BS.cv[i];
            B.cw[b0 + i] = BS.cw[i];
            B.sv[b0 + i] = BS.sv[i];
            B.sw[b0 + i] = BS.sw[i];
        }
    }
    if (!masked)
        for (int k = tid; k < ngw; k += nt) {
            F.world_iters[w0 + k] = it;
            F.world_res[w0 + k] = res;
        }
}

template <bool SPLIT>
__global__ __launch_bounds__(kFusedMaxBlock) void sv_world_solve(FusedArgs F) {
    __shared__ float red[16];
    __shared__ unsigned int s_wres[kGroupMaxWorlds];   // residual of each world of the group (float bits)
    __shared__ unsigned long long s_done;               // converged worlds of the group
    extern __shared__ float4 s_vel[];   // [6][lds_bodies] cv, cw, sv, sw (w: third inertia row), iw0, iw1;
                                        // [lds_bodies] flags; with lds_residual [2][lds_bodies] previous cv, cw
    const SvArgs& AG = F.A;
    const DevBodies& B = AG.B;
    const int g = F.block_world[blockIdx.x];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int w0 = F.gworld[g], ngw = F.gworld[g + 1] - w0;
    const int b0 = F.wbody[w0], b1 = F.wbody[w0 + ngw];
    const int* seg = F.seg_start + (size_t)g * kMaxColors;
    const int m0 = seg[0], m1 = seg[kMaxColors];
    const int2* jr = F.jt_wr + (size_t)g * F.jt_slots;
    const int j0 = F.jt_slots ? jr[0].x : 0, j1 = F.jt_slots ? jr[F.jt_slots - 1].y : 0;

    if (F.phase != 1) {
        // PhysicsWorld.cpp:266-272: inverse world inertia tensors and velocity integration
        for (int i = b0 + tid; i < b1; i += nt) dyn_integrate_velocities_one(B, F.settings, F.dt, i);
        if (F.J.bfirst)
            for (int w = tid; w < ngw; w += nt) F.J.bfirst[w0 + w] = INT_MAX;
        __syncthreads();
        // ContactSolverSystem::init (rows); ConstraintSolverSystem::initialize (initBeforeSolve)
        const SvArgs& A = AG;
        for (int i = m0 + tid; i < m1; i += nt) sv_prepare_one(A, i);
        for (int j = j0 + tid; j < j1; j += nt) jt_init_one(B, F.J, j, F.dt);
        if (F.feats) {
            // feature rows of the points without a cached impulse; same manifold -> thread mapping as
            // above, so every thread reads only rows it wrote itself
            const int lane = lane_id();
            for (int base = m0 + (tid & ~(kWave - 1)); base < m1; base += nt) {
                const int i = base + lane;
                int k = 0;
                if (i < m1) {
                    const int np = A.S.npts[i];
                    for (int p = 0; p < np; p++) k += A.S.lam[4 * i + p].z == 0.0f ? 1 : 0;
                }
                const int incl = wave_incl_scan(k);
                int start = 0;
                if (lane == kWave - 1 && incl > 0) start = atomicAdd(F.nrows, incl);
                start = __shfl(start, kWave - 1, kWave) + incl - k;
                if (i < m1) {
                    const int np = A.S.npts[i];
                    for (int p = 0; p < np; p++) {
                        const int q = 4 * i + p;
                        if (A.S.lam[q].z != 0.0f) continue;
                        A.S.pred[q] = make_float4(0.f, 0.f, 0.f, 0.f);
                        if (start < F.feat_cap) {
                            float f[kFeatures];
                            sv_point_features(A, i, p, f);
                            for (int e = 0; e < kFeatures; e++) F.feats[(size_t)start * kFeatures + e] = f[e];
                            F.row_slot[start] = q;
                        }
                        start++;
                    }
                }
            }
        }
        __syncthreads();
        if (F.phase == 0) return;
    }
    // stage the group's solver velocities, world inverse inertia rows and flags in LDS for the sweeps
    // (the rows and the rest of the body data stay in HBM); the third inertia row rides in the w lanes of
    // cw, sv and sw: 100 bytes a body, so six 256-body groups fit the 160 KB of a CU (five at 116 bytes)
    const int nbw = b1 - b0;
    if (nbw <= F.lds_bodies) {
        const int L = F.lds_bodies;
        float4 *lcv = s_vel, *lcw = s_vel + L, *lsv = s_vel + 2 * L, *lsw = s_vel + 3 * L;
        float4 *li0 = s_vel + 4 * L, *li1 = s_vel + 5 * L;
        int* lfl = reinterpret_cast<int*>(s_vel + 6 * L);
        for (int i = tid; i < nbw; i += nt) {
            lfl[i] = B.flags[b0 + i] | ((B.world[b0 + i] - w0) << kGroupShift);
            const float4 r2 = B.iw2[b0 + i];
            float4 t = B.cw[b0 + i];
            t.w = r2.x;
            lcw[i] = t;
            t = B.sv[b0 + i];
            t.w = r2.y;
            lsv[i] = t;
            t = B.sw[b0 + i];
            t.w = r2.z;
            lsw[i] = t;
            lcv[i] = B.cv[b0 + i];
            li0[i] = B.iw0[b0 + i];
            li1[i] = B.iw1[b0 + i];
        }
        SvArgs A = AG;
        A.B.cv = lcv; A.B.cw = lcw; A.B.sv = lsv; A.B.sw = lsw;
        A.B.iw0 = li0; A.B.iw1 = li1; A.B.iw2 = nullptr;
        A.B.flags = lfl;
        A.B.vbase = b0;
        float4* lpv = F.lds_residual ? reinterpret_cast<float4*>(lfl + L) : F.pv + b0;
        float4* lpw = F.lds_residual ? lpv + L : F.pw + b0;
        __syncthreads();
        group_sweeps<SPLIT>(F, A, B, true, w0, ngw, b0, b1, m1, seg, jr, red, s_wres, &s_done, lpv, lpw);
    } else {
        group_sweeps<SPLIT>(F, AG, B, false, w0, ngw, b0, b1, m1, seg, jr, red, s_wres, &s_done, F.pv + b0, F.pw + b0);
    }
}
template __global__ void sv_world_solve<true>(FusedArgs F);
template __global__ void sv_world_solve<false>(FusedArgs F);

__global__ __launch_bounds__(kFusedMaxBlock) void world_position(FusedArgs F) {
    const DevBodies& B = F.A.B;
    const int g = F.block_world[blockIdx.x];
    const int tid = threadIdx.x, nt = blockDim.x;
    const int w0 = F.gworld[g], w1 = F.gworld[g + 1];
    const int b0 = F.wbody[w0], b1 = F.wbody[w1];
    const int c0 = F.wcoll[w0], c1 = F.wcoll[w1];
    const int2* jr = F.jt_wr + (size_t)g * F.jt_slots;
    // PhysicsWorld.cpp:275-288: integrate positions, joint position correction, update body state
    const float split_factor = F.settings.split_impulse ? 1.0f : 0.0f;
    for (int i = b0 + tid; i < b1; i += nt) dyn_integrate_positions_one(B, F.dt, i, split_factor);
    __syncthreads();
    for (int it = 0; it < F.settings.pos_iters; it++) {
        for (int s = 0; s < F.jt_slots; s++) {
            const int2 r = jr[s];
            if (r.x == r.y) continue;
            for (int j = r.x + tid; j < r.y; j += nt) jt_solve_position_one(B, F.J, j);
            __syncthreads();
        }
    }
    for (int i = b0 + tid; i < b1; i += nt) dyn_update_state_one(B, i);
    __syncthreads();
    // collider transforms and fat AABBs for the next broadphase
    for (int i = c0 + tid; i < c1; i += nt) refit_one(B, F.A.C, F.hull_bounds, 0, i);
}

__global__ void sv_log_samples(FusedArgs F, float* log_feats, float* log_targets, const unsigned long long* logged,
                               long long log_cap) {
    const int n = min(*F.nrows, F.feat_cap);
    const long long base = (long long)*logged;
    RP3DG_GRID_STRIDE(r, n) {
        const long long dst = base + r;
        if (dst >= log_cap) continue;
        const int q = F.row_slot[r], i = q >> 2;
        const float4 imp = F.A.S.imp[i];
        const float inv_np = 1.0f / (float)F.A.S.npts[i];
        for (int e = 0; e < kFeatures; e++) log_feats[dst * kFeatures + e] = F.feats[(size_t)r * kFeatures + e];
        log_targets[dst * 3 + 0] = F.A.S.lam[q].x;
        log_targets[dst * 3 + 1] = imp.x * inv_np;
        log_targets[dst * 3 + 2] = imp.y * inv_np;
    }
}

__global__ void sv_log_advance(const int* nrows, int feat_cap, unsigned long long* logged, long long log_cap) {
    const long long n = min(*nrows, feat_cap);
    const long long v = (long long)*logged + n;
    *logged = (unsigned long long)(v < log_cap ? v : log_cap);
}

}  // namespace rp3dg
