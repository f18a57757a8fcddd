// Contact solver rows and body integration of the MI355X step path, one item per call.
//
// The reference solves contact manifolds sequentially in island order
// (ContactSolverSystem.cpp:479-759).  Here manifolds are graph-coloured so
// that no two manifolds of a colour share a dynamic body; all manifolds of a
// colour are then solved in parallel, which is exactly the sequential
// Gauss-Seidel sweep in colour order (static and kinematic bodies have zero
// inverse mass and are never written, so they do not constrain the colouring).
// Used by the per-colour kernels (solver.hip) and the fused per-world solver
// (fused.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "kcommon.h"
#include "rp3d_internal.h"
#include "solver.h"

namespace rp3dg {

__device__ __forceinline__ int body_flags(const DevBodies& B, int b) { return B.flags[b - B.vbase]; }
__device__ __forceinline__ bool dyn_body(const DevBodies& B, int b) { return (body_flags(B, b) & BF_DYNAMIC) != 0; }
// lock factors of body b (absolute index; multiplying by 1 is exact, so unit factors are not loaded)
__device__ __forceinline__ void body_locks(const DevBodies& B, int b, V3& ll, V3& al) {
    if (body_flags(B, b) & BF_UNIT_LOCKS) {
        ll = v3(1.0f, 1.0f, 1.0f);
        al = v3(1.0f, 1.0f, 1.0f);
    } else {
        ll = ld3(B.llock, b);
        al = ld3(B.alock, b);
    }
}

// x * f, or x when the lock factor is known to be 1
template <bool LK>
__device__ __forceinline__ float lk(float x, float f) {
    if constexpr (LK) return x * f;
    else return x;
}
template <bool LK>
__device__ __forceinline__ V3 lkv(V3 f, V3 x) {
    if constexpr (LK) return cmul(f, x);
    else return x;
}

// ------------------------------------------------------------------------------------------------
// Integration (DynamicsSystem.cpp, PhysicsWorld.cpp:310-318)
// ------------------------------------------------------------------------------------------------

// updateBodiesInverseWorldInertiaTensors + integrateRigidBodiesVelocities (DynamicsSystem.cpp:117-183)
__device__ __forceinline__ void dyn_integrate_velocities_one(const DevBodies& B, const Settings& S, float dt, int i) {
    {
        const int f = B.flags[i];
        if (!(f & BF_ENABLED)) return;
        const Q4 q = ldq(B.q, i);
        const float4 ii = B.ii[i];
        const M3 Iw = world_inv_inertia(qmat(q), v3(ii.x, ii.y, ii.z));
        B.iw0[i] = f4(Iw.r[0], 0.f);
        B.iw1[i] = f4(Iw.r[1], 0.f);
        B.iw2[i] = f4(Iw.r[2], 0.f);
        B.sv[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        B.sw[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        const float4 ll4 = B.llock[i], al4 = B.alock[i];
        const V3 ll = xyz(ll4), al = xyz(al4);
        const float im = ii.w;
        const V3 v = ld3(B.v, i), w = ld3(B.w, i);
        V3 cv = v + cmul((dt * im) * ll, ld3(B.force, i));
        V3 cw = w + cmul(dt * al, mvmul(Iw, ld3(B.torque, i)));
        if (S.gravity_enabled && (f & BF_GRAVITY)) {
            const float mass = B.com[i].w;
            cv = cv + cmul(((dt * im) * ll) * mass, S.gravity);
        }
        const float ld = 1.0f / (1.0f + ll4.w * dt);
        const float ad = 1.0f / (1.0f + al4.w * dt);
        cv = cv * ld;
        cw = cw * ad;
        B.cv[i] = f4(cv, 0.f);
        B.cw[i] = f4(cw, 0.f);
    }
}

// integrateRigidBodiesPositions (DynamicsSystem.cpp:44-71): the split velocities enter with the factor
// 1 (split impulses) or 0 (Baumgarte contacts), multiplied as the reference does
__device__ __forceinline__ void dyn_integrate_positions_one(const DevBodies& B, float dt, int i, float split_factor) {
    {
        if (!(B.flags[i] & BF_ENABLED)) return;
        const V3 nv = ld3(B.cv, i) + split_factor * ld3(B.sv, i);
        const V3 nw = ld3(B.cw, i) + split_factor * ld3(B.sw, i);
        const V3 x = ld3(B.x, i);
        const Q4 q = ldq(B.q, i);
        const V3 cx = x + nv * dt;
        const Q4 dq = qscale(qscale(qmul(q4(nw.x, nw.y, nw.z, 0.0f), q), 0.5f), dt);
        B.cx[i] = f4(cx, 0.f);
        B.cq[i] = f4q(qadd(q, dq));
    }
}

// updateBodiesState (DynamicsSystem.cpp:74-101) and resetBodiesForceAndTorque (:186-195)
__device__ __forceinline__ void dyn_update_state_one(const DevBodies& B, int i) {
    {
        if (B.flags[i] & BF_ENABLED) {
            B.v[i] = B.cv[i];
            B.w[i] = B.cw[i];
            const V3 x = ld3(B.cx, i);
            const Q4 q = qunit(ldq(B.cq, i));
            B.x[i] = f4(x, 0.f);
            B.q[i] = f4q(q);
            B.p[i] = f4(x - qrot(q, ld3(B.com, i)), 0.f);
        }
        B.force[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        B.torque[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}


// ------------------------------------------------------------------------------------------------
// Solver rows (ContactSolverSystem::initializeForIsland, CS.cpp:115-320) in colour order
// ------------------------------------------------------------------------------------------------

// world inverse inertia rows, indexed [body - vbase] like the velocities (staged in LDS per world)
// World inverse inertia of body b.  Staged in LDS (iw2 == nullptr), its third row lives in the w lanes of
// cw, sv and sw, which the solver rows therefore store with st3 (x, y, z only).
__device__ __forceinline__ M3 body_iw(const DevBodies& B, int b) {
    const int u = b - B.vbase;
    M3 m;
    m.r[0] = ld3(B.iw0, u); m.r[1] = ld3(B.iw1, u);
    m.r[2] = B.iw2 ? ld3(B.iw2, u) : v3(B.cw[u].w, B.sv[u].w, B.sw[u].w);
    return m;
}

// computeFrictionVectors (CS.cpp:789-818)
__device__ inline void friction_vectors(V3 dv, V3 n, V3& t1, V3& t2) {
    const float dvn = dot(dv, n);
    const V3 nv = dvn * n;
    const V3 tv = v3(dv.x - nv.x, dv.y - nv.y, dv.z - nv.z);
    const float lt = len(tv);
    if (lt > kMachineEps) t1 = tv / lt;
    else t1 = one_unit_orthogonal(n);
    t2 = cross(n, t1);
}

__device__ __forceinline__ V3 rel_vel(V3 v1, V3 w1, V3 v2, V3 w2, V3 r1, V3 r2) {
    return v3(v2.x + w2.y * r2.z - w2.z * r2.y - v1.x - w1.y * r1.z + w1.z * r1.y,
              v2.y + w2.z * r2.x - w2.x * r2.z - v1.y - w1.z * r1.x + w1.x * r1.z,
              v2.z + w2.x * r2.y - w2.y * r2.x - v1.z - w1.x * r1.y + w1.y * r1.x);
}

__device__ __forceinline__ void sv_prepare_one(const SvArgs& A, int i) {
    {
        const int m = A.order[i];
        const int k = A.mlist[m];
        const DevBodies& B = A.B;
        const DevColliders& C = A.C;
        const DevPairs& P = A.P;
        const DevSolver& S = A.S;
        const int2 c = P.col[k];
        const int b1 = C.body[c.x], b2 = C.body[c.y];
        const V3 x1 = ld3(B.x, b1), x2 = ld3(B.x, b2);
        const M3 I1 = body_iw(B, b1), I2 = body_iw(B, b2);
        const float im1 = B.ii[b1].w, im2 = B.ii[b2].w;
        const float2 m1 = C.mat[c.x], m2 = C.mat[c.y];
        const float mu = m1.x * m2.x;
        const float rest = (m1.y > m2.y) ? m1.y : m2.y;
        const V3 v1 = ld3(B.v, b1), w1 = ld3(B.w, b1), v2 = ld3(B.v, b2), w2 = ld3(B.w, b2);
        const Tf T1 = tf(ld3(C.wpos, c.x), ldq(C.wq, c.x));
        const Tf T2 = tf(ld3(C.wpos, c.y), ldq(C.wq, c.y));
        const int np = P.ncp[k];
        V3 nsum = v3zero(), fp1 = v3zero(), fp2 = v3zero();
        for (int p = 0; p < np; p++) {
            const float4 pn = P.pn[4 * k + p];
            const float4 l1 = P.pl1[4 * k + p];
            const float4 l2 = P.pl2[4 * k + p];
            const V3 n = xyz(pn);
            const V3 p1 = tapply(T1, xyz(l1));
            const V3 p2 = tapply(T2, xyz(l2));
            const V3 r1 = v3(p1.x - x1.x, p1.y - x1.y, p1.z - x1.z);
            const V3 r2 = v3(p2.x - x2.x, p2.y - x2.y, p2.z - x2.z);
            fp1.x += p1.x; fp1.y += p1.y; fp1.z += p1.z;
            fp2.x += p2.x; fp2.y += p2.y; fp2.z += p2.z;
            const V3 dv = rel_vel(v1, w1, v2, w2, r1, r2);
            const V3 i1n = mvmul(I1, cross(r1, n));
            const V3 i2n = mvmul(I2, cross(r2, n));
            const float mp = im1 + im2 + dot(cross(i1n, r1), n) + dot(cross(i2n, r2), n);
            const float imp = mp > 0.0f ? 1.0f / mp : 0.0f;
            float rb = 0.0f;
            const float dvn = dv.x * n.x + dv.y * n.y + dv.z * n.z;
            if (dvn < -A.restitution_threshold) rb = rest * dvn;
            nsum.x += n.x; nsum.y += n.y; nsum.z += n.z;
            S.pn[4 * i + p] = pn;
            S.r1[4 * i + p] = f4(r1, imp);
            S.r2[4 * i + p] = f4(r2, rb);
            // lam: cached normal impulse, split impulse, resting flag from the warm-start transfer
            S.lam[4 * i + p] = make_float4(l1.w, 0.0f, l2.w, 0.f);
        }
        const float fn = (float)np;
        fp1 = v3(fp1.x / fn, fp1.y / fn, fp1.z / fn);
        fp2 = v3(fp2.x / fn, fp2.y / fn, fp2.z / fn);
        const V3 r1f = v3(fp1.x - x1.x, fp1.y - x1.y, fp1.z - x1.z);
        const V3 r2f = v3(fp2.x - x2.x, fp2.y - x2.y, fp2.z - x2.z);
        const V3 n = normalized(nsum);
        const V3 dvf = rel_vel(v1, w1, v2, w2, r1f, r2f);
        V3 t1, t2;
        friction_vectors(dvf, n, t1, t2);
        const V3 r1ct1 = cross(r1f, t1), r1ct2 = cross(r1f, t2);
        const V3 r2ct1 = cross(r2f, t1), r2ct2 = cross(r2f, t2);
        const float f1m = im1 + im2 + dot(cross(mvmul(I1, r1ct1), r1f), t1) + dot(cross(mvmul(I2, r2ct1), r2f), t1);
        const float f2m = im1 + im2 + dot(cross(mvmul(I1, r1ct2), r1f), t2) + dot(cross(mvmul(I2, r2ct2), r2f), t2);
        const float twm = dot(n, mvmul(I1, n)) + dot(n, mvmul(I2, n));
        S.pair[i] = k;
        S.bodies[i] = make_int2(b1, b2);
        S.npts[i] = np;
        S.n[i] = f4(n, mu);
        S.r1f[i] = f4(r1f, im1);
        S.r2f[i] = f4(r2f, im2);
        S.t1[i] = f4(t1, f1m > 0.0f ? 1.0f / f1m : 0.0f);
        S.t2[i] = f4(t2, f2m > 0.0f ? 1.0f / f2m : 0.0f);
        // r x t and I (r x n) are recomputed by the sweeps (same operations, same bits) instead of
        // being stored: every iteration re-reads the rows, so smaller rows mean less traffic
        const float4 fi = P.fimp[k];
        S.imp[i] = make_float4(fi.x, fi.y, fi.z, twm > 0.0f ? 1.0f / twm : 0.0f);   // w: inverse twist mass
        S.oldt1[i] = P.fv1[k];
        S.oldt2[i] = P.fv2[k];
    }
}

// ------------------------------------------------------------------------------------------------
// Per-colour sweeps
// ------------------------------------------------------------------------------------------------

struct BodyVel {
    V3 v, w;
};

__device__ __forceinline__ void load_body(const DevBodies& B, int b, const float4* lin, const float4* ang, V3& v, V3& w) {
    v = ld3(lin, b);
    w = ld3(ang, b);
}

// ContactSolverSystem::warmStart (CS.cpp:326-476) for the manifolds [lo, hi) of one colour
__device__ __forceinline__ void sv_warmstart_one(const SvArgs& A, int i) {
    {
        const DevBodies& B = A.B;
        const DevSolver& S = A.S;
        const int2 bb = S.bodies[i];
        const int b1 = bb.x, b2 = bb.y;
        const bool d1 = dyn_body(B, b1), d2 = dyn_body(B, b2);
        const int u1 = b1 - B.vbase, u2 = b2 - B.vbase;
        V3 v1 = ld3(B.cv, u1), w1 = ld3(B.cw, u1), v2 = ld3(B.cv, u2), w2 = ld3(B.cw, u2);
        const float im1 = S.r1f[i].w, im2 = S.r2f[i].w;
        V3 ll1, al1, ll2, al2;
        body_locks(B, b1, ll1, al1);
        body_locks(B, b2, ll2, al2);
        const int np = S.npts[i];
        const M3 I1 = body_iw(B, b1), I2 = body_iw(B, b2);
        bool any = false;
        float plx = 0.0f, ply = 0.0f;   // learned friction impulses summed over the new points
        for (int p = 0; p < np; p++) {
            const int q = 4 * i + p;
            const float4 lam4 = S.lam[q];
            if (lam4.z != 0.0f || A.learned) {
                float lam;
                if (lam4.z != 0.0f) {
                    any = true;
                    lam = lam4.x;
                } else {   // learned warm start of a point with no cached impulse
                    const float4 pr = S.pred[q];
                    lam = fmax_s(pr.x, 0.0f);
                    S.lam[q].x = lam;
                    plx += pr.y;
                    ply += pr.z;
                }
                const V3 n = xyz(S.pn[q]);
                const V3 i1n = mvmul(I1, cross(ld3(S.r1, q), n)), i2n = mvmul(I2, cross(ld3(S.r2, q), n));
                const V3 ip = v3(n.x * lam, n.y * lam, n.z * lam);
                v1.x -= im1 * ip.x * ll1.x; v1.y -= im1 * ip.y * ll1.y; v1.z -= im1 * ip.z * ll1.z;
                w1.x -= i1n.x * al1.x * lam; w1.y -= i1n.y * al1.y * lam; w1.z -= i1n.z * al1.z * lam;
                v2.x += im2 * ip.x * ll2.x; v2.y += im2 * ip.y * ll2.y; v2.z += im2 * ip.z * ll2.z;
                w2.x += i2n.x * al2.x * lam; w2.y += i2n.y * al2.y * lam; w2.z += i2n.z * al2.z * lam;
            } else {
                S.lam[q].x = 0.0f;
            }
        }
        float4 imp = S.imp[i];
        // learned friction only for a manifold with no resting point (the reference then starts from zero)
        const bool learned_friction = !any && A.learned && np > 0;
        if (any || learned_friction) {
            const V3 t1 = ld3(S.t1, i), t2 = ld3(S.t2, i);
            if (any) {
                const V3 ot1 = ld3(S.oldt1, i), ot2 = ld3(S.oldt2, i);
                const V3 oi = v3(imp.x * ot1.x + imp.y * ot2.x, imp.x * ot1.y + imp.y * ot2.y, imp.x * ot1.z + imp.y * ot2.z);
                imp.x = dot(oi, t1);
                imp.y = dot(oi, t2);
            } else {
                imp.x = plx;
                imp.y = ply;
                imp.z = 0.0f;
            }
            const V3 r1f = ld3(S.r1f, i), r2f = ld3(S.r2f, i);
            const V3 r1ct1 = cross(r1f, t1), r2ct1 = cross(r2f, t1), r1ct2 = cross(r1f, t2), r2ct2 = cross(r2f, t2);
            V3 a1 = v3(-r1ct1.x * imp.x, -r1ct1.y * imp.x, -r1ct1.z * imp.x);
            V3 l2 = v3(t1.x * imp.x, t1.y * imp.x, t1.z * imp.x);
            V3 a2 = v3(r2ct1.x * imp.x, r2ct1.y * imp.x, r2ct1.z * imp.x);
            v1 -= cmul(im1 * l2, ll1);
            w1 += cmul(al1, mvmul(I1, a1));
            v2 += cmul(im2 * l2, ll2);
            w2 += cmul(al2, mvmul(I2, a2));
            a1 = v3(-r1ct2.x * imp.y, -r1ct2.y * imp.y, -r1ct2.z * imp.y);
            l2 = v3(t2.x * imp.y, t2.y * imp.y, t2.z * imp.y);
            a2 = v3(r2ct2.x * imp.y, r2ct2.y * imp.y, r2ct2.z * imp.y);
            v1.x -= im1 * l2.x * ll1.x; v1.y -= im1 * l2.y * ll1.y; v1.z -= im1 * l2.z * ll1.z;
            w1 += cmul(al1, mvmul(I1, a1));
            v2.x += im2 * l2.x * ll2.x; v2.y += im2 * l2.y * ll2.y; v2.z += im2 * l2.z * ll2.z;
            w2 += cmul(al2, mvmul(I2, a2));
            const V3 n = ld3(S.n, i);
            a1 = v3(-n.x * imp.z, -n.y * imp.z, -n.z * imp.z);
            a2 = v3(n.x * imp.z, n.y * imp.z, n.z * imp.z);
            w1 += cmul(al1, mvmul(I1, a1));
            w2 += cmul(al2, mvmul(I2, a2));
            // the reference applies the twist impulse a second time (CS.cpp:457-466)
            w1 -= cmul(al1, mvmul(I1, a2));
            w2 += cmul(al2, mvmul(I2, a2));
        } else {
            imp.x = imp.y = imp.z = 0.0f;
        }
        S.imp[i] = imp;
        if (d1) { B.cv[u1] = f4(v1, 0.f); st3(B.cw, u1, w1); }
        if (d2) { B.cv[u2] = f4(v2, 0.f); st3(B.cw, u2, w2); }
    }
}

// Learned warm start: the 24 features of contact point p of solver row i (after sv_prepare_one and
// the velocity integration): normal, r1, r2, relative velocity in the (t1, t2, n) frame, inverse
// masses, world inverse inertia (xx of each body), the manifold's cached normal impulses (sum over
// its resting points, their count), the manifold counts of both bodies, depth, friction
// coefficient, inverse normal mass, 1 / points in the manifold.
__device__ __forceinline__ void sv_point_features(const SvArgs& A, int i, int p, float* f) {
    const DevBodies& B = A.B;
    const DevSolver& S = A.S;
    const int q = 4 * i + p;
    const int2 bb = S.bodies[i];
    const float4 pn = S.pn[q], r14 = S.r1[q], r24 = S.r2[q];
    const V3 n = xyz(pn), r1 = xyz(r14), r2 = xyz(r24);
    const int u1 = bb.x - B.vbase, u2 = bb.y - B.vbase;
    const V3 dv = rel_vel(ld3(B.cv, u1), ld3(B.cw, u1), ld3(B.cv, u2), ld3(B.cw, u2), r1, r2);
    const V3 t1 = ld3(S.t1, i), t2 = ld3(S.t2, i);
    f[0] = n.x; f[1] = n.y; f[2] = n.z;
    f[3] = r1.x; f[4] = r1.y; f[5] = r1.z;
    f[6] = r2.x; f[7] = r2.y; f[8] = r2.z;
    f[9] = dot(dv, t1); f[10] = dot(dv, t2); f[11] = dot(dv, n);
    f[12] = S.r1f[i].w; f[13] = S.r2f[i].w;
    float cached = 0.0f, resting = 0.0f;
    for (int k = 0; k < S.npts[i]; k++) {
        const float4 l = S.lam[4 * i + k];
        if (l.z != 0.0f) {
            cached += l.x;
            resting += 1.0f;
        }
    }
    const int* ao = A.adj_offs;
    f[14] = B.iw0[bb.x].x; f[15] = cached; f[16] = resting;
    f[17] = B.iw0[bb.y].x;
    f[18] = dyn_body(B, bb.x) ? (float)(ao[bb.x + 1] - ao[bb.x]) : 0.0f;
    f[19] = dyn_body(B, bb.y) ? (float)(ao[bb.y + 1] - ao[bb.y]) : 0.0f;
    f[20] = pn.w;
    f[21] = S.n[i].w;
    f[22] = r14.w;
    f[23] = 1.0f / (float)S.npts[i];
}

// ContactSolverSystem::solve (CS.cpp:479-759), for one colour: with split impulses the penetration bias
// drives the split row; with Baumgarte contacts it joins the restitution bias in the main row
// LK = false: both bodies have unit lock factors, and every multiplication by a factor (x * 1 = x
// exactly) is left out; LK = true multiplies as the reference does.  SPLIT: split impulses (the
// default technique) or Baumgarte contacts, a separate copy each so that the default's sweep keeps
// its own registers and schedule.
template <bool LK, bool SPLIT>
__device__ __forceinline__ void solve_manifold_t(const SvArgs& A, int i, float beta_dt) {
    const DevBodies& B = A.B;
    const DevSolver& S = A.S;
    const int2 bb = S.bodies[i];
    const int b1 = bb.x, b2 = bb.y;
    const bool d1 = dyn_body(B, b1), d2 = dyn_body(B, b2);
    const int u1 = b1 - B.vbase, u2 = b2 - B.vbase;
    V3 v1 = ld3(B.cv, u1), w1 = ld3(B.cw, u1), v2 = ld3(B.cv, u2), w2 = ld3(B.cw, u2);
    V3 sv1 = ld3(B.sv, u1), sw1 = ld3(B.sw, u1), sv2 = ld3(B.sv, u2), sw2 = ld3(B.sw, u2);
    const float im1 = S.r1f[i].w, im2 = S.r2f[i].w;
    V3 ll1 = v3(1.0f, 1.0f, 1.0f), al1 = ll1, ll2 = ll1, al2 = ll1;
    if constexpr (LK) {
        body_locks(B, b1, ll1, al1);
        body_locks(B, b2, ll2, al2);
    }
    const int np = S.npts[i];
    const M3 I1 = body_iw(B, b1), I2 = body_iw(B, b2);
    float sumPen = 0.0f;
    for (int p = 0; p < np; p++) {
        const int q = 4 * i + p;
        const float4 pn = S.pn[q];
        const V3 n = xyz(pn);
        const float4 r14 = S.r1[q], r24 = S.r2[q];
        const V3 r1 = xyz(r14), r2 = xyz(r24);
        const V3 i1n = mvmul(I1, cross(r1, n)), i2n = mvmul(I2, cross(r2, n));
        const float ipm = r14.w, rb = r24.w;
        float4 lam = S.lam[q];
        const V3 dv = rel_vel(v1, w1, v2, w2, r1, r2);
        const float Jv = dv.x * n.x + dv.y * n.y + dv.z * n.z;
        float biasPen = 0.0f;
        if (pn.w > 0.01f) biasPen = -(beta_dt) * fmax_s(0.0f, pn.w - 0.01f);
        float dl = SPLIT ? -(Jv + rb) * ipm : -(Jv + (biasPen + rb)) * ipm;
        float lt = lam.x;
        lam.x = fmax_s(lam.x + dl, 0.0f);
        dl = lam.x - lt;
        const V3 li = v3(n.x * dl, n.y * dl, n.z * dl);
        v1.x -= lk<LK>(im1 * li.x, ll1.x); v1.y -= lk<LK>(im1 * li.y, ll1.y); v1.z -= lk<LK>(im1 * li.z, ll1.z);
        w1.x -= lk<LK>(i1n.x, al1.x) * dl; w1.y -= lk<LK>(i1n.y, al1.y) * dl; w1.z -= lk<LK>(i1n.z, al1.z) * dl;
        v2.x += lk<LK>(im2 * li.x, ll2.x); v2.y += lk<LK>(im2 * li.y, ll2.y); v2.z += lk<LK>(im2 * li.z, ll2.z);
        w2.x += lk<LK>(i2n.x, al2.x) * dl; w2.y += lk<LK>(i2n.y, al2.y) * dl; w2.z += lk<LK>(i2n.z, al2.z) * dl;
        sumPen += lam.x;
        if constexpr (!SPLIT) {
            S.lam[q] = lam;
            continue;
        }
        // split impulse row
        const V3 dvs = rel_vel(sv1, sw1, sv2, sw2, r1, r2);
        const float Jvs = dvs.x * n.x + dvs.y * n.y + dvs.z * n.z;
        float dls = -(Jvs + biasPen) * ipm;
        const float lts = lam.y;
        lam.y = fmax_s(lam.y + dls, 0.0f);
        dls = lam.y - lts;
        const V3 ls = v3(n.x * dls, n.y * dls, n.z * dls);
        sv1.x -= lk<LK>(im1 * ls.x, ll1.x); sv1.y -= lk<LK>(im1 * ls.y, ll1.y); sv1.z -= lk<LK>(im1 * ls.z, ll1.z);
        sw1.x -= lk<LK>(i1n.x, al1.x) * dls; sw1.y -= lk<LK>(i1n.y, al1.y) * dls; sw1.z -= lk<LK>(i1n.z, al1.z) * dls;
        sv2.x += lk<LK>(im2 * ls.x, ll2.x); sv2.y += lk<LK>(im2 * ls.y, ll2.y); sv2.z += lk<LK>(im2 * ls.z, ll2.z);
        sw2.x += lk<LK>(i2n.x, al2.x) * dls; sw2.y += lk<LK>(i2n.y, al2.y) * dls; sw2.z += lk<LK>(i2n.z, al2.z) * dls;
        S.lam[q] = lam;
    }
    const V3 r1f = ld3(S.r1f, i), r2f = ld3(S.r2f, i);
    const float4 t14 = S.t1[i], t24 = S.t2[i], n4 = S.n[i];
    const V3 t1 = xyz(t14), t2 = xyz(t24), n = xyz(n4);
    const float mu = n4.w;
    float4 imp = S.imp[i];
    // friction 1
    {
        const V3 dv = rel_vel(v1, w1, v2, w2, r1f, r2f);
        const float Jv = dv.x * t1.x + dv.y * t1.y + dv.z * t1.z;
        float dl = -Jv * t14.w;
        const float lim = mu * sumPen;
        const float lt = imp.x;
        imp.x = fmax_s(-lim, fmin_s(imp.x + dl, lim));
        dl = imp.x - lt;
        const V3 r1ct1 = cross(r1f, t1), r2ct1 = cross(r2f, t1);
        const V3 a1 = v3(-r1ct1.x * dl, -r1ct1.y * dl, -r1ct1.z * dl);
        const V3 l2 = v3(t1.x * dl, t1.y * dl, t1.z * dl);
        const V3 a2 = v3(r2ct1.x * dl, r2ct1.y * dl, r2ct1.z * dl);
        v1.x -= lk<LK>(im1 * l2.x, ll1.x); v1.y -= lk<LK>(im1 * l2.y, ll1.y); v1.z -= lk<LK>(im1 * l2.z, ll1.z);
        const V3 aw1 = lkv<LK>(al1, mvmul(I1, a1));
        w1.x += aw1.x; w1.y += aw1.y; w1.z += aw1.z;
        v2.x += lk<LK>(im2 * l2.x, ll2.x); v2.y += lk<LK>(im2 * l2.y, ll2.y); v2.z += lk<LK>(im2 * l2.z, ll2.z);
        const V3 aw2 = lkv<LK>(al2, mvmul(I2, a2));
        w2.x += aw2.x; w2.y += aw2.y; w2.z += aw2.z;
    }
    // friction 2
    {
        const V3 dv = rel_vel(v1, w1, v2, w2, r1f, r2f);
        const float Jv = dv.x * t2.x + dv.y * t2.y + dv.z * t2.z;
        float dl = -Jv * t24.w;
        const float lim = mu * sumPen;
        const float lt = imp.y;
        imp.y = fmax_s(-lim, fmin_s(imp.y + dl, lim));
        dl = imp.y - lt;
        const V3 r1ct2 = cross(r1f, t2), r2ct2 = cross(r2f, t2);
        const V3 a1 = v3(-r1ct2.x * dl, -r1ct2.y * dl, -r1ct2.z * dl);
        const V3 l2 = v3(t2.x * dl, t2.y * dl, t2.z * dl);
        const V3 a2 = v3(r2ct2.x * dl, r2ct2.y * dl, r2ct2.z * dl);
        v1.x -= lk<LK>(im1 * l2.x, ll1.x); v1.y -= lk<LK>(im1 * l2.y, ll1.y); v1.z -= lk<LK>(im1 * l2.z, ll1.z);
        const V3 aw1 = lkv<LK>(al1, mvmul(I1, a1));
        w1.x += aw1.x; w1.y += aw1.y; w1.z += aw1.z;
        v2.x += lk<LK>(im2 * l2.x, ll2.x); v2.y += lk<LK>(im2 * l2.y, ll2.y); v2.z += lk<LK>(im2 * l2.z, ll2.z);
        const V3 aw2 = lkv<LK>(al2, mvmul(I2, a2));
        w2.x += aw2.x; w2.y += aw2.y; w2.z += aw2.z;
    }
    // twist
    {
        const V3 dv = w2 - w1;
        const float Jv = dv.x * n.x + dv.y * n.y + dv.z * n.z;
        float dl = -Jv * imp.w;   // inverse twist mass
        const float lim = mu * sumPen;
        const float lt = imp.z;
        imp.z = fmax_s(-lim, fmin_s(imp.z + dl, lim));
        dl = imp.z - lt;
        const V3 a2 = v3(n.x * dl, n.y * dl, n.z * dl);
        const V3 aw1 = lkv<LK>(al1, mvmul(I1, a2));
        w1.x -= aw1.x; w1.y -= aw1.y; w1.z -= aw1.z;
        const V3 aw2 = lkv<LK>(al2, mvmul(I2, a2));
        w2.x += aw2.x; w2.y += aw2.y; w2.z += aw2.z;
    }
    S.imp[i] = imp;
    if (d1) {
        B.cv[u1] = f4(v1, 0.f); st3(B.cw, u1, w1);
        st3(B.sv, u1, sv1); st3(B.sw, u1, sw1);
    }
    if (d2) {
        B.cv[u2] = f4(v2, 0.f); st3(B.cw, u2, w2);
        st3(B.sv, u2, sv2); st3(B.sw, u2, sw2);
    }
}

// SPLIT = A.split, fixed by the caller's kernel instance (the fused solver kernel has one per technique)
template <bool SPLIT>
__device__ __forceinline__ void solve_manifold(const SvArgs& A, int i, float beta_dt) {
    const int2 bb = A.S.bodies[i];
    if (body_flags(A.B, bb.x) & body_flags(A.B, bb.y) & BF_UNIT_LOCKS) solve_manifold_t<false, SPLIT>(A, i, beta_dt);
    else solve_manifold_t<true, SPLIT>(A, i, beta_dt);
}

// ContactSolverSystem::storeImpulses (CS.cpp:763-785); initializeForIsland marks every point resting (:161)
__device__ __forceinline__ void sv_store_one(const SvArgs& A, int i) {
    {
        const int k = A.S.pair[i];
        const int np = A.S.npts[i];
        for (int p = 0; p < np; p++) {
            float4 l1 = A.P.pl1[4 * k + p];
            l1.w = A.S.lam[4 * i + p].x;
            A.P.pl1[4 * k + p] = l1;
            float4 l2 = A.P.pl2[4 * k + p];
            l2.w = 1.0f;
            A.P.pl2[4 * k + p] = l2;
        }
        const float4 imp = A.S.imp[i];
        A.P.fimp[k] = make_float4(imp.x, imp.y, imp.z, 0.f);
        A.P.fv1[k] = f4(ld3(A.S.t1, i), 0.f);
        A.P.fv2[k] = f4(ld3(A.S.t2, i), 0.f);
    }
}

}  // namespace rp3dg
