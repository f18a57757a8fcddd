// Joint rows of the MI355X step path, one joint per call: ball-and-socket (with its cone limit), fixed,
// hinge and slider joints (src/systems/Solve{BallAndSocket,Fixed,Hinge,Slider}JointSystem.cpp) with the
// default non-linear Gauss-Seidel position correction.
//
// Joints never change during a run, so the host colours every joint type once
// (no two joints of a colour share a body, static bodies included: the position
// solver rewrites and renormalises the constrained orientation of both bodies,
// static ones too).  Joints of one colour of one type are solved in parallel;
// types follow the reference's order BallAndSocket -> Fixed -> Hinge -> Slider
// (ConstraintSolverSystem.cpp:94-113).  Used by the per-colour kernels (joints.hip)
// and by the fused per-world solver (fused.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "joints.h"
#include "kcommon.h"
#include "solver_dev.h"

namespace rp3dg {

// ---- small matrix helpers (Matrix3x3.h / Matrix2x2.h) ----
__device__ inline M3 m3(float a, float b, float c, float d, float e, float f, float g, float h, float i) {
    M3 m;
    m.r[0] = v3(a, b, c); m.r[1] = v3(d, e, f); m.r[2] = v3(g, h, i);
    return m;
}
__device__ inline M3 madd(const M3& a, const M3& b) {
    M3 r;
    for (int i = 0; i < 3; i++) r.r[i] = a.r[i] + b.r[i];
    return r;
}
__device__ inline M3 skew(V3 v) { return m3(0, -v.z, v.y, v.z, 0, -v.x, -v.y, v.x, 0); }   // Matrix3x3.h:251-253
__device__ inline float mdet(const M3& m) {   // Matrix3x3.h:211-217
    return (m.r[0].x * (m.r[1].y * m.r[2].z - m.r[2].y * m.r[1].z) -
            m.r[0].y * (m.r[1].x * m.r[2].z - m.r[2].x * m.r[1].z) +
            m.r[0].z * (m.r[1].x * m.r[2].y - m.r[2].x * m.r[1].y));
}
__device__ inline M3 minv(const M3& m, float det) {   // Matrix3x3.cpp:33-52
    const float id = 1.0f / det;
    M3 t = m3((m.r[1].y * m.r[2].z - m.r[2].y * m.r[1].z), -(m.r[0].y * m.r[2].z - m.r[2].y * m.r[0].z),
              (m.r[0].y * m.r[1].z - m.r[0].z * m.r[1].y), -(m.r[1].x * m.r[2].z - m.r[2].x * m.r[1].z),
              (m.r[0].x * m.r[2].z - m.r[2].x * m.r[0].z), -(m.r[0].x * m.r[1].z - m.r[1].x * m.r[0].z),
              (m.r[1].x * m.r[2].y - m.r[2].x * m.r[1].y), -(m.r[0].x * m.r[2].y - m.r[2].x * m.r[0].y),
              (m.r[0].x * m.r[1].y - m.r[0].y * m.r[1].x));
    for (int i = 0; i < 3; i++) t.r[i] = v3(t.r[i].x * id, t.r[i].y * id, t.r[i].z * id);
    return t;
}
__device__ inline M3 mzero() { return m3(0, 0, 0, 0, 0, 0, 0, 0, 0); }

__device__ inline M3 ldm(const float4* r0, const float4* r1, const float4* r2, int i) {
    M3 m;
    m.r[0] = ld3(r0, i); m.r[1] = ld3(r1, i); m.r[2] = ld3(r2, i);
    return m;
}
__device__ inline void stm(float4* r0, float4* r1, float4* r2, int i, const M3& m) {
    r0[i] = f4(m.r[0], 0.f); r1[i] = f4(m.r[1], 0.f); r2[i] = f4(m.r[2], 0.f);
}

// K = (m1+m2) I + U1 I1 U1^T + U2 I2 U2^T, inverted when invertible and a body is dynamic
__device__ inline M3 translation_inverse_mass(float imsum, V3 r1, V3 r2, const M3& I1, const M3& I2, bool anyDyn, bool& ok) {
    const M3 U1 = skew(r1), U2 = skew(r2);
    const M3 K = madd(madd(m3(imsum, 0, 0, 0, imsum, 0, 0, 0, imsum), mmul(mmul(U1, I1), mtranspose(U1))),
                      mmul(mmul(U2, I2), mtranspose(U2)));
    const float det = mdet(K);
    ok = fabsf(det) > kMachineEps;
    if (ok && anyDyn) return minv(K, det);
    return mzero();
}

__device__ inline V3 lock_mul(V3 lock, V3 v) { return cmul(lock, v); }

// q += (0, w) q * 0.5 ; q.normalize()
__device__ inline Q4 nudge(Q4 q, V3 w) {
    const Q4 d = qscale(qmul(q4(w.x, w.y, w.z, 0.0f), q), 0.5f);
    Q4 r = qadd(q, d);
    const float l = qlen(r);
    return q4(r.x / l, r.y / l, r.z / l, r.w / l);
}

// ---- hinge angle (SolveHingeJointSystem.cpp:733-799) ----
__device__ inline float normalized_angle(float a) {
    const float twopi = 2.0f * kPi;
    a = fmodf(a, twopi);
    if (a < -kPi) return a + twopi;
    if (a > kPi) return a - twopi;
    return a;
}
__device__ inline float angle_near_limits(float a, float lo, float hi) {
    const float twopi = 2.0f * kPi;
    if (hi <= lo) return a;
    if (a > hi) {
        const float du = fabsf(normalized_angle(a - hi));
        const float dl = fabsf(normalized_angle(a - lo));
        return (du > dl) ? (a - twopi) : a;
    }
    if (a < lo) {
        const float du = fabsf(normalized_angle(hi - a));
        const float dl = fabsf(normalized_angle(lo - a));
        return (du > dl) ? a : (a + twopi);
    }
    return a;
}
__device__ inline float hinge_angle(Q4 q1, Q4 q2, Q4 qinit, V3 a1, float lo, float hi) {
    Q4 diff = qmul(q2, qconj(q1));
    float l = qlen(diff);
    diff = q4(diff.x / l, diff.y / l, diff.z / l, diff.w / l);
    Q4 rel = qmul(diff, qinit);
    l = qlen(rel);
    rel = q4(rel.x / l, rel.y / l, rel.z / l, rel.w / l);
    const float c = rel.w;
    const V3 vv = v3(rel.x, rel.y, rel.z);
    const float s = len(vv);
    const float d = dot(vv, a1);
    float ang = d >= 0.0f ? 2.0f * ref_atan2f(s, c) : 2.0f * ref_atan2f(s, -c);
    ang = normalized_angle(ang);
    return angle_near_limits(ang, lo, hi);
}

// PhysicsWorld::enableDisableJoints (PhysicsWorld.cpp:363-393): a joint whose bodies are both disabled
// (static or sleeping) is not solved and keeps its impulses
__device__ __forceinline__ bool joint_enabled(const DevBodies& B, int b1, int b2) {
    return ((body_flags(B, b1) | body_flags(B, b2)) & BF_ENABLED) != 0;
}

// ------------------------------------------------------------------------------------------------
// initBeforeSolve (all joints, one launch; no body state is written)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void jt_init_one(const DevBodies& B, const DevJoints& J, int j, float dt) {
    {
        const int2 bb = J.bodies[j];
        const int b1 = bb.x, b2 = bb.y;
        if (!joint_enabled(B, b1, b2)) return;
        const int t = J.type[j];
        const M3 I1 = ldm(B.iw0, B.iw1, B.iw2, b1), I2 = ldm(B.iw0, B.iw1, B.iw2, b2);
        stm(J.i1r0, J.i1r1, J.i1r2, j, I1);
        stm(J.i2r0, J.i2r1, J.i2r2, j, I2);
        const Q4 q1 = ldq(B.q, b1), q2 = ldq(B.q, b2);
        const V3 r1 = qrot(q1, ld3(J.la1, j) - ld3(B.com, b1));
        const V3 r2 = qrot(q2, ld3(J.la2, j) - ld3(B.com, b2));
        J.r1[j] = f4(r1, 0.f);
        J.r2[j] = f4(r2, 0.f);
        const float im1 = B.ii[b1].w, im2 = B.ii[b2].w;
        const bool anyDyn = ((body_flags(B, b1) | body_flags(B, b2)) & BF_DYNAMIC) != 0;
        bool ok;
        const M3 Kt = translation_inverse_mass(im1 + im2, r1, r2, I1, I2, anyDyn, ok);
        stm(J.kt0, J.kt1, J.kt2, j, Kt);
        // Baumgarte biases (BETA = 0.2, biasFactor = BETA / mTimeStep in every Solve*JointSystem::initBeforeSolve)
        const bool bg = (J.flags[j] & JF_BAUMGARTE) != 0;
        const float bf = 0.2f / dt;
        if (t == JT_BALLSOCKET) {
            // cone-limit axis a x b with a = unit(r1), b = -unit(r2) (SolveBallAndSocketJointSystem.cpp:113-115)
            const V3 r1u = unit(r1), r2u = -unit(r2);
            const V3 acb = cross(r1u, r2u);
            J.a1[j] = f4(acb, 0.f);
            int fl = J.flags[j];
            float cone_bias = 0.0f;
            if (fl & JF_LIMIT) {
                // SolveBallAndSocketJointSystem.cpp:117-146.  A disabled limit keeps the half angle pi
                // (PhysicsWorld.cpp:500), which acos never exceeds: never violated, impulse stays 0.
                const float err = J.lim[j].x - ref_acosf(dot(r1u, r2u));
                const bool old = (fl & JF_LOWER_VIOLATED) != 0;
                const bool vio = err < 0.0f;
                fl = vio ? (fl | JF_LOWER_VIOLATED) : (fl & ~JF_LOWER_VIOLATED);
                if (!vio || vio != old) {
                    float4 il = J.impL[j];
                    il.x = 0.0f;
                    J.impL[j] = il;
                }
                float m = dot(acb, mvmul(I1, acb)) + dot(acb, mvmul(I2, acb));
                J.imlm[j] = m > 0.0f ? 1.0f / m : 0.0f;
                J.flags[j] = fl;
                cone_bias = bf * err;   // SolveBallAndSocketJointSystem.cpp:146-149
            }
            // SolveBallAndSocketJointSystem.cpp:113-117
            if (bg) J.bias0[j] = f4(bf * (ld3(B.x, b2) + r2 - ld3(B.x, b1) - r1), cone_bias);
        } else if (t == JT_FIXED) {
            // SolveFixedJointSystem.cpp:120-129: keeps I1 + I2 when it cannot be inverted
            M3 Kr = madd(I1, I2);
            const float det = mdet(Kr);
            if (fabsf(det) > kMachineEps && anyDyn) Kr = minv(Kr, det);
            stm(J.kr0, J.kr1, J.kr2, j, Kr);
            if (bg) {   // SolveFixedJointSystem.cpp:110-132
                J.bias0[j] = f4(bf * (ld3(B.x, b2) + r2 - ld3(B.x, b1) - r1), 0.f);
                const Q4 qe = qmul(qmul(q2, ldq(J.qinit, j)), qconj(q1));
                J.bias1[j] = f4((bf * 2.0f) * v3(qe.x, qe.y, qe.z), 0.f);
            }
        } else if (t == JT_HINGE) {
            const V3 a1 = normalized(qrot(q1, ld3(J.ax1, j)));
            const V3 a2 = normalized(qrot(q2, ld3(J.ax2, j)));
            const V3 b2v = one_unit_orthogonal(a2);
            const V3 c2v = cross(a2, b2v);
            const V3 b2xa1 = cross(b2v, a1), c2xa1 = cross(c2v, a1);
            J.a1[j] = f4(a1, 0.f);
            J.b2xa1[j] = f4(b2xa1, 0.f);
            J.c2xa1[j] = f4(c2xa1, 0.f);
            const V3 i1b = mvmul(I1, b2xa1), i1c = mvmul(I1, c2xa1), i2b = mvmul(I2, b2xa1), i2c = mvmul(I2, c2xa1);
            const float e11 = dot(b2xa1, i1b) + dot(b2xa1, i2b);
            const float e12 = dot(b2xa1, i1c) + dot(b2xa1, i2c);
            const float e21 = dot(c2xa1, i1b) + dot(c2xa1, i2b);
            const float e22 = dot(c2xa1, i1c) + dot(c2xa1, i2c);
            float4 kr = make_float4(0.f, 0.f, 0.f, 0.f);
            const float det = e11 * e22 - e21 * e12;
            if (fabsf(det) > kMachineEps && anyDyn) {
                const float id = 1.0f / det;
                kr = make_float4(e22 * id, -e12 * id, -e21 * id, e11 * id);
            }
            J.kr0[j] = kr;
            // limits / motor (SolveHingeJointSystem.cpp:159-212)
            int fl = J.flags[j];
            const float4 lim = J.lim[j];
            float4 il = J.impL[j];
            // SolveHingeJointSystem.cpp:91-95, 124-128, 198-209 (the limit biases are read only while a
            // limit is violated, a step in which they are computed)
            float4 brot = make_float4(bf * dot(a1, b2v), bf * dot(a1, c2v), 0.f, 0.f);
            if ((fl & JF_LIMIT) || (fl & JF_MOTOR)) {
                const float ang = hinge_angle(q1, q2, ldq(J.qinit, j), a1, lim.x, lim.y);
                const float le = ang - lim.x, ue = lim.y - ang;
                brot.z = bf * le;
                brot.w = bf * ue;
                const bool oldL = (fl & JF_LOWER_VIOLATED) != 0, oldU = (fl & JF_UPPER_VIOLATED) != 0;
                const bool vl = le <= 0, vu = ue <= 0;
                fl = vl ? (fl | JF_LOWER_VIOLATED) : (fl & ~JF_LOWER_VIOLATED);
                fl = vu ? (fl | JF_UPPER_VIOLATED) : (fl & ~JF_UPPER_VIOLATED);
                if (!vl || vl != oldL) il.x = 0.0f;
                if (!vu || vu != oldU) il.y = 0.0f;
                if ((fl & JF_MOTOR) || ((fl & JF_LIMIT) && (vl || vu))) {
                    float m = dot(a1, mvmul(I1, a1)) + dot(a1, mvmul(I2, a1));
                    m = m > 0.0f ? 1.0f / m : 0.0f;
                    J.imlm[j] = m;
                }
                J.flags[j] = fl;
                J.impL[j] = il;
            }
            if (bg) {
                J.bias0[j] = f4(bf * (ld3(B.x, b2) + r2 - ld3(B.x, b1) - r1), 0.f);
                J.bias1[j] = brot;
            }
        } else {   // slider, SolveSliderJointSystem.cpp:52-237
            const V3 ax = normalized(qrot(q1, ld3(J.ax1, j)));
            const V3 n1 = one_unit_orthogonal(ax);
            const V3 n2 = cross(ax, n1);
            const V3 u = ld3(B.x, b2) + r2 - ld3(B.x, b1) - r1;
            const V3 rpu = r1 + u;
            const V3 un1 = cross(rpu, n1), un2 = cross(rpu, n2), ua = cross(rpu, ax);
            J.a1[j] = f4(ax, 0.f);
            J.b2xa1[j] = f4(n1, 0.f);
            J.c2xa1[j] = f4(n2, 0.f);
            J.un1[j] = f4(un1, 0.f);
            J.un2[j] = f4(un2, 0.f);
            J.ua[j] = f4(ua, 0.f);
            // limit state is tracked whether or not the limit is enabled
            const float ud = dot(u, ax);
            const float4 lim = J.lim[j];
            const float le = ud - lim.x, ue = lim.y - ud;
            int fl = J.flags[j];
            float4 il = J.impL[j];
            const bool oldL = (fl & JF_LOWER_VIOLATED) != 0, oldU = (fl & JF_UPPER_VIOLATED) != 0;
            const bool vl = le <= 0, vu = ue <= 0;
            fl = vl ? (fl | JF_LOWER_VIOLATED) : (fl & ~JF_LOWER_VIOLATED);
            fl = vu ? (fl | JF_UPPER_VIOLATED) : (fl & ~JF_UPPER_VIOLATED);
            if (!vl || vl != oldL) il.x = 0.0f;
            if (!vu || vu != oldU) il.y = 0.0f;
            J.flags[j] = fl;
            J.impL[j] = il;
            if (bg) {   // SolveSliderJointSystem.cpp:116-122, 142-152, 208-213
                J.bias0[j] = make_float4(dot(u, n1) * bf, dot(u, n2) * bf, bf * le, bf * ue);
                const Q4 qe = qmul(qmul(q2, ldq(J.qinit, j)), qconj(q1));
                J.bias1[j] = f4((bf * 2.0f) * v3(qe.x, qe.y, qe.z), 0.f);
                if (J.bfirst) atomicMin(J.bfirst + B.world[b1], J.srank[j]);
            }
            const float ims = im1 + im2;
            if ((fl & JF_LIMIT) && (vl || vu)) {
                // reads r2 x axis of the previous step (recomputed only below)
                const V3 ra_old = ld3(J.ra, j);
                const float m = ims + dot(ua, mvmul(I1, ua)) + dot(ra_old, mvmul(I2, ra_old));
                J.imlm[j] = m > 0.0f ? 1.0f / m : 0.0f;
            }
            const V3 rn1 = cross(r2, n1), rn2 = cross(r2, n2);
            J.rn1[j] = f4(rn1, 0.f);
            J.rn2[j] = f4(rn2, 0.f);
            J.ra[j] = f4(cross(r2, ax), 0.f);
            const V3 i1n1 = mvmul(I1, un1), i1n2 = mvmul(I1, un2), i2n1 = mvmul(I2, rn1), i2n2 = mvmul(I2, rn2);
            const float e11 = ims + dot(un1, i1n1) + dot(rn1, i2n1);
            const float e12 = dot(un1, i1n2) + dot(rn1, i2n2);
            const float e21 = dot(un2, i1n1) + dot(rn2, i2n1);
            const float e22 = ims + dot(un2, i1n2) + dot(rn2, i2n2);
            float4 kt = make_float4(0.f, 0.f, 0.f, 0.f);
            const float det = e11 * e22 - e21 * e12;
            if (fabsf(det) > kMachineEps && anyDyn) {
                const float id = 1.0f / det;
                kt = make_float4(e22 * id, -e12 * id, -e21 * id, e11 * id);
            }
            J.kt0[j] = kt;
            M3 Kr = madd(I1, I2);
            const float detr = mdet(Kr);
            if (fabsf(detr) > kMachineEps && anyDyn) Kr = minv(Kr, detr);
            stm(J.kr0, J.kr1, J.kr2, j, Kr);
            if (fl & JF_MOTOR) J.imm[j] = ims > 0.0f ? 1.0f / ims : 0.0f;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// warm start (per colour)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void jt_warmstart_one(const DevBodies& B, const DevJoints& J, int j) {
    {
        const int2 bb = J.bodies[j];
        const int b1 = bb.x, b2 = bb.y;
        if (!joint_enabled(B, b1, b2)) return;
        const int type = J.type[j];
        const int u1 = b1 - B.vbase, u2 = b2 - B.vbase;
        V3 v1 = ld3(B.cv, u1), w1 = ld3(B.cw, u1), v2 = ld3(B.cv, u2), w2 = ld3(B.cw, u2);
        const float im1 = B.ii[b1].w, im2 = B.ii[b2].w;
        V3 ll1, al1, ll2, al2;
        body_locks(B, b1, ll1, al1);
        body_locks(B, b2, ll2, al2);
        const M3 I1 = ldm(J.i1r0, J.i1r1, J.i1r2, j), I2 = ldm(J.i2r0, J.i2r1, J.i2r2, j);
        const V3 r1 = ld3(J.r1, j), r2 = ld3(J.r2, j);
        const V3 it = ld3(J.impT, j);
        if (type == JT_BALLSOCKET) {   // SolveBallAndSocketJointSystem.cpp:196-241
            // the cone impulse stays 0 while the limit is disabled; 0 * (a x b) is still added (keeps the
            // sign of zero components identical to the reference)
            const V3 cone = J.impL[j].x * ld3(J.a1, j);
            const V3 li1 = -it;
            V3 ai1 = cross(it, r1);
            ai1 += cone;
            v1 += lock_mul((im1 * ll1), li1);
            w1 += cmul(al1, mvmul(I1, ai1));
            V3 ai2 = -cross(it, r2);
            ai2 += -cone;
            v2 += lock_mul((im2 * ll2), it);
            w2 += cmul(al2, mvmul(I2, ai2));
        } else if (type == JT_FIXED) {   // SolveFixedJointSystem.cpp:142-186
            const V3 ir = ld3(J.impR, j);
            const V3 li1 = -it;
            V3 ai1 = cross(it, r1);
            ai1 += -ir;
            v1 += lock_mul((im1 * ll1), li1);
            w1 += cmul(al1, mvmul(I1, ai1));
            V3 ai2 = -cross(it, r2);
            ai2 += ir;
            v2 += lock_mul((im2 * ll2), it);
            w2 += cmul(al2, mvmul(I2, ai2));
        } else if (type == JT_HINGE) {   // SolveHingeJointSystem.cpp:216-270
            const float4 ir = J.impR[j];
            const float4 il = J.impL[j];
            const V3 a1 = ld3(J.a1, j), bxa = ld3(J.b2xa1, j), cxa = ld3(J.c2xa1, j);
            const V3 rot = (-bxa) * ir.x - cxa * ir.y;
            const V3 limI = (il.y - il.x) * a1;
            const V3 motI = (-il.z) * a1;
            const V3 li1 = -it;
            V3 ai1 = cross(it, r1);
            ai1 += rot;
            ai1 += limI;
            ai1 += motI;
            v1 += lock_mul((im1 * ll1), li1);
            w1 += cmul(al1, mvmul(I1, ai1));
            V3 ai2 = -cross(it, r2);
            ai2 += -rot;
            ai2 += -limI;
            ai2 += -motI;
            v2 += lock_mul((im2 * ll2), it);
            w2 += cmul(al2, mvmul(I2, ai2));
        } else {   // slider, SolveSliderJointSystem.cpp:240-314
            const float4 ir = J.impR[j];
            const float4 il = J.impL[j];
            const V3 ax = ld3(J.a1, j), n1 = ld3(J.b2xa1, j), n2 = ld3(J.c2xa1, j);
            const V3 un1 = ld3(J.un1, j), un2 = ld3(J.un2, j);
            const float lims = il.y - il.x;
            const V3 linLim = lims * ax;
            const V3 motI = il.z * ax;
            V3 li1 = (-n1) * it.x - n2 * it.y;
            V3 ai1 = (-un1) * it.x - un2 * it.y;
            ai1 += -v3(ir.x, ir.y, ir.z);
            li1 += linLim;
            ai1 += lims * ld3(J.ua, j);
            li1 += motI;
            v1 += lock_mul((im1 * ll1), li1);
            w1 += cmul(al1, mvmul(I1, ai1));
            V3 li2 = n1 * it.x + n2 * it.y;
            V3 ai2 = ld3(J.rn1, j) * it.x + ld3(J.rn2, j) * it.y;
            ai2 += v3(ir.x, ir.y, ir.z);
            li2 += -linLim;
            ai2 += (-lims) * ld3(J.ra, j);
            li2 += -motI;
            v2 += lock_mul((im2 * ll2), li2);
            w2 += cmul(al2, mvmul(I2, ai2));
        }
        B.cv[u1] = f4(v1, 0.f); st3(B.cw, u1, w1);
        B.cv[u2] = f4(v2, 0.f); st3(B.cw, u2, w2);
    }
}

// ------------------------------------------------------------------------------------------------
// velocity constraints (per colour)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void jt_solve_velocity_one(const DevBodies& B, const DevJoints& J, int j, float dt) {
    {
        const int2 bb = J.bodies[j];
        const int b1 = bb.x, b2 = bb.y;
        if (!joint_enabled(B, b1, b2)) return;
        const int type = J.type[j];
        const int u1 = b1 - B.vbase, u2 = b2 - B.vbase;
        V3 v1 = ld3(B.cv, u1), w1 = ld3(B.cw, u1), v2 = ld3(B.cv, u2), w2 = ld3(B.cw, u2);
        const float im1 = B.ii[b1].w, im2 = B.ii[b2].w;
        V3 ll1, al1, ll2, al2;
        body_locks(B, b1, ll1, al1);
        body_locks(B, b2, ll2, al2);
        const M3 I1 = ldm(J.i1r0, J.i1r1, J.i1r2, j), I2 = ldm(J.i2r0, J.i2r1, J.i2r2, j);
        const V3 r1 = ld3(J.r1, j), r2 = ld3(J.r2, j);
        const M3 Kt = ldm(J.kt0, J.kt1, J.kt2, j);
        // Baumgarte biases; zero under non-linear Gauss-Seidel position correction
        float4 bias0 = make_float4(0.f, 0.f, 0.f, 0.f), bias1 = bias0;
        if (J.flags[j] & JF_BAUMGARTE) {
            bias0 = J.bias0[j];
            bias1 = J.bias1[j];
        }
        const V3 bt = v3(bias0.x, bias0.y, bias0.z);
        if (type == JT_BALLSOCKET) {   // SolveBallAndSocketJointSystem.cpp:245-308
            if ((J.flags[j] & (JF_LIMIT | JF_LOWER_VIOLATED)) == (JF_LIMIT | JF_LOWER_VIOLATED)) {
                const V3 acb = ld3(J.a1, j);
                const float JvC = dot(acb, w1 - w2);
                float d = J.imlm[j] * (-JvC - bias0.w);
                float4 il = J.impL[j];
                const float lt = il.x;
                il.x = fmax_s(il.x + d, 0.0f);
                d = il.x - lt;
                J.impL[j] = il;
                w1 += cmul(al1, mvmul(I1, d * acb));
                w2 += cmul(al2, mvmul(I2, (-d) * acb));
            }
            const V3 Jv = v2 + cross(w2, r2) - v1 - cross(w1, r1);
            const V3 dl = mvmul(Kt, -Jv - bt);
            J.impT[j] = f4(ld3(J.impT, j) + dl, 0.f);
            const V3 li1 = -dl;
            const V3 ai1 = cross(dl, r1);
            v1 += lock_mul((im1 * ll1), li1);
            w1 += cmul(al1, mvmul(I1, ai1));
            const V3 ai2 = -cross(dl, r2);
            v2 += lock_mul((im2 * ll2), dl);
            w2 += cmul(al2, mvmul(I2, ai2));
        } else if (type == JT_FIXED) {   // SolveFixedJointSystem.cpp:189-262
            const V3 Jvt = v2 + cross(w2, r2) - v1 - cross(w1, r1);
            const V3 dl = mvmul(Kt, -Jvt - bt);
            J.impT[j] = f4(ld3(J.impT, j) + dl, 0.f);
            const V3 li1 = -dl;
            V3 ai1 = cross(dl, r1);
            v1 += lock_mul((im1 * ll1), li1);
            w1 += cmul(al1, mvmul(I1, ai1));
            const V3 ai2 = -cross(dl, r2);
            v2 += lock_mul((im2 * ll2), dl);
            w2 += cmul(al2, mvmul(I2, ai2));
            const V3 Jvr = w2 - w1;
            const M3 Kr = ldm(J.kr0, J.kr1, J.kr2, j);
            const V3 dl2 = mvmul(Kr, -Jvr - v3(bias1.x, bias1.y, bias1.z));
            J.impR[j] = f4(ld3(J.impR, j) + dl2, 0.f);
            ai1 = -dl2;
            w1 += cmul(al1, mvmul(I1, ai1));
            w2 += cmul(al2, mvmul(I2, dl2));
        } else if (type == JT_HINGE) {   // SolveHingeJointSystem.cpp:273-440
            const V3 a1 = ld3(J.a1, j);
            const int fl = J.flags[j];
            float4 il = J.impL[j];
            const float imlm = J.imlm[j];
            if (fl & JF_LIMIT) {
                if (fl & JF_LOWER_VIOLATED) {
                    const float JvL = dot(w2 - w1, a1);
                    float d = imlm * (-JvL - bias1.z);
                    const float lt = il.x;
                    il.x = fmax_s(il.x + d, 0.0f);
                    d = il.x - lt;
                    const V3 ai1 = (-d) * a1;
                    w1 += cmul(al1, mvmul(I1, ai1));
                    const V3 ai2 = d * a1;
                    w2 += cmul(al2, mvmul(I2, ai2));
                }
                if (fl & JF_UPPER_VIOLATED) {
                    const float JvU = -dot(w2 - w1, a1);
                    float d = imlm * (-JvU - bias1.w);
                    const float lt = il.y;
                    il.y = fmax_s(il.y + d, 0.0f);
                    d = il.y - lt;
                    const V3 ai1 = d * a1;
                    w1 += cmul(al1, mvmul(I1, ai1));
                    const V3 ai2 = (-d) * a1;
                    w2 += cmul(al2, mvmul(I2, ai2));
                }
            }
            if (fl & JF_MOTOR) {
                const float4 lim = J.lim[j];
                const float JvM = dot(a1, w1 - w2);
                const float maxI = lim.w * dt;
                float d = imlm * (-JvM - lim.z);
                const float lt = il.z;
                il.z = clampf(il.z + d, -maxI, maxI);
                d = il.z - lt;
                const V3 ai1 = (-d) * a1;
                w1 += cmul(al1, mvmul(I1, ai1));
                const V3 ai2 = d * a1;
                w2 += cmul(al2, mvmul(I2, ai2));
            }
            J.impL[j] = il;
            const V3 bxa = ld3(J.b2xa1, j), cxa = ld3(J.c2xa1, j);
            const float jr0 = -dot(bxa, w1) + dot(bxa, w2);
            const float jr1 = -dot(cxa, w1) + dot(cxa, w2);
            const float4 kr = J.kr0[j];
            const float x0 = -jr0 - bias1.x, x1 = -jr1 - bias1.y;
            const float d0 = kr.x * x0 + kr.y * x1;
            const float d1 = kr.z * x0 + kr.w * x1;
            float4 ir = J.impR[j];
            ir.x += d0;
            ir.y += d1;
            J.impR[j] = ir;
            V3 ai1 = (-bxa) * d0 - cxa * d1;
            w1 += cmul(al1, mvmul(I1, ai1));
            V3 ai2 = bxa * d0 + cxa * d1;
            w2 += cmul(al2, mvmul(I2, ai2));
            const V3 Jvt = v2 + cross(w2, r2) - v1 - cross(w1, r1);
            const V3 dl = mvmul(Kt, -Jvt - bt);
            J.impT[j] = f4(ld3(J.impT, j) + dl, 0.f);
            const V3 li1 = -dl;
            ai1 = cross(dl, r1);
            v1 += lock_mul((im1 * ll1), li1);
            w1 += cmul(al1, mvmul(I1, ai1));
            ai2 = -cross(dl, r2);
            v2 += lock_mul((im2 * ll2), dl);
            w2 += cmul(al2, mvmul(I2, ai2));
        } else {   // slider, SolveSliderJointSystem.cpp:317-510
            const V3 ax = ld3(J.a1, j), ua = ld3(J.ua, j), ra = ld3(J.ra, j);
            const int fl = J.flags[j];
            float4 il = J.impL[j];
            if (fl & JF_LIMIT) {
                const float imlm = J.imlm[j];
                if (fl & JF_LOWER_VIOLATED) {
                    const float JvL = dot(ax, v2) + dot(ra, w2) - dot(ax, v1) - dot(ua, w1);
                    float d = imlm * (-JvL - bias0.z);
                    const float lt = il.x;
                    il.x = fmax_s(il.x + d, 0.0f);
                    d = il.x - lt;
                    v1 += lock_mul((im1 * ll1), (-d) * ax);
                    w1 += cmul(al1, mvmul(I1, (-d) * ua));
                    v2 += lock_mul((im2 * ll2), d * ax);
                    w2 += cmul(al2, mvmul(I2, d * ra));
                }
                if (fl & JF_UPPER_VIOLATED) {
                    const float JvU = dot(ax, v1) + dot(ua, w1) - dot(ax, v2) - dot(ra, w2);
                    float d = imlm * (-JvU - bias0.w);
                    const float lt = il.y;
                    il.y = fmax_s(il.y + d, 0.0f);
                    d = il.y - lt;
                    v1 += lock_mul((im1 * ll1), d * ax);
                    w1 += cmul(al1, mvmul(I1, d * ua));
                    v2 += lock_mul((im2 * ll2), (-d) * ax);
                    w2 += cmul(al2, mvmul(I2, (-d) * ra));
                }
            }
            if (fl & JF_MOTOR) {
                const float4 lim = J.lim[j];
                const float JvM = dot(ax, v1) - dot(ax, v2);
                const float maxI = lim.w * dt;
                float d = J.imm[j] * (-JvM - lim.z);
                const float lt = il.z;
                il.z = clampf(il.z + d, -maxI, maxI);
                d = il.z - lt;
                v1 += lock_mul((im1 * ll1), d * ax);
                v2 += lock_mul((im2 * ll2), (-d) * ax);
            }
            J.impL[j] = il;
            // 3 rotation rows
            const V3 Jvr = w2 - w1;
            const M3 Kr = ldm(J.kr0, J.kr1, J.kr2, j);
            const V3 dr = mvmul(Kr, -Jvr - v3(bias1.x, bias1.y, bias1.z));
            J.impR[j] = f4(ld3(J.impR, j) + dr, 0.f);
            w1 += cmul(al1, mvmul(I1, -dr));
            w2 += cmul(al2, mvmul(I2, dr));
            // 2 translation rows
            const V3 n1 = ld3(J.b2xa1, j), n2 = ld3(J.c2xa1, j);
            const V3 un1 = ld3(J.un1, j), un2 = ld3(J.un2, j), rn1 = ld3(J.rn1, j), rn2 = ld3(J.rn2, j);
            const float el1 = -dot(n1, v1) - dot(w1, un1) + dot(n1, v2) + dot(w2, rn1);
            const float el2 = -dot(n2, v1) - dot(w1, un2) + dot(n2, v2) + dot(w2, rn2);
            const float x0 = -el1 - bias0.x, x1 = -el2 - bias0.y;
            const float4 kt = J.kt0[j];
            const float d0 = kt.x * x0 + kt.y * x1;
            const float d1 = kt.z * x0 + kt.w * x1;
            float4 it = J.impT[j];
            it.x += d0;
            it.y += d1;
            J.impT[j] = it;
            const V3 li1 = (-n1) * d0 - n2 * d1;
            const V3 ai1 = (-un1) * d0 - un2 * d1;
            v1 += lock_mul((im1 * ll1), li1);
            w1 += cmul(al1, mvmul(I1, ai1));
            const V3 ai2 = rn1 * d0 + rn2 * d1;
            v2 += lock_mul((im2 * ll2), -li1);
            w2 += cmul(al2, mvmul(I2, ai2));
        }
        B.cv[u1] = f4(v1, 0.f); st3(B.cw, u1, w1);
        B.cv[u2] = f4(v2, 0.f); st3(B.cw, u2, w2);
    }
}

// ------------------------------------------------------------------------------------------------
// non-linear Gauss-Seidel position correction (per colour)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void jt_solve_position_one(const DevBodies& B, const DevJoints& J, int j) {
    {
        const int2 bb = J.bodies[j];
        const int b1 = bb.x, b2 = bb.y;
        if (!joint_enabled(B, b1, b2)) return;
        const int type = J.type[j];
        // only non-linear Gauss-Seidel joints (SolveBallAndSocketJointSystem.cpp:309,
        // SolveFixedJointSystem.cpp:299, SolveHingeJointSystem.cpp:479); the slider system stops at its
        // first Baumgarte joint (SolveSliderJointSystem.cpp:527)
        if (J.flags[j] & JF_BAUMGARTE) return;
        if (type == JT_SLIDER && J.bfirst && J.srank[j] >= J.bfirst[B.world[b1]]) return;
        Q4 q1 = ldq(B.cq, b1), q2 = ldq(B.cq, b2);
        V3 x1 = ld3(B.cx, b1), x2 = ld3(B.cx, b2);
        const float4 ii1 = B.ii[b1], ii2 = B.ii[b2];
        const float im1 = ii1.w, im2 = ii2.w;
        V3 ll1, al1, ll2, al2;
        body_locks(B, b1, ll1, al1);
        body_locks(B, b2, ll2, al2);
        const bool anyDyn = ((body_flags(B, b1) | body_flags(B, b2)) & BF_DYNAMIC) != 0;
        M3 I1 = world_inv_inertia(qmat(q1), v3(ii1.x, ii1.y, ii1.z));
        M3 I2 = world_inv_inertia(qmat(q2), v3(ii2.x, ii2.y, ii2.z));
        const V3 r1 = qrot(q1, ld3(J.la1, j) - ld3(B.com, b1));
        const V3 r2 = qrot(q2, ld3(J.la2, j) - ld3(B.com, b2));
        if (type == JT_BALLSOCKET) {   // SolveBallAndSocketJointSystem.cpp:312-445
            int fl = J.flags[j];
            if (fl & JF_LIMIT) {
                const V3 r1u = unit(r1), r2u = -unit(r2);
                const V3 acb = cross(r1u, r2u);
                const float err = J.lim[j].x - ref_acosf(dot(r1u, r2u));
                fl = err < 0.0f ? (fl | JF_LOWER_VIOLATED) : (fl & ~JF_LOWER_VIOLATED);
                J.flags[j] = fl;
                if (fl & JF_LOWER_VIOLATED) {
                    const float m = dot(acb, mvmul(I1, acb)) + dot(acb, mvmul(I2, acb));
                    const float l = (m > 0.0f ? 1.0f / m : 0.0f) * (-err);
                    q1 = nudge(q1, cmul(al1, mvmul(I1, l * acb)));
                    q2 = nudge(q2, cmul(al2, mvmul(I2, (-l) * acb)));
                }
            }
            bool ok;
            const M3 Kt = translation_inverse_mass(im1 + im2, r1, r2, I1, I2, anyDyn, ok);
            if (ok) {
                const V3 err = x2 + r2 - x1 - r1;
                const V3 lam = mvmul(Kt, -err);
                const V3 li1 = -lam;
                const V3 ai1 = cross(lam, r1);
                const V3 dv1 = lock_mul((im1 * ll1), li1);
                const V3 dw1 = cmul(al1, mvmul(I1, ai1));
                x1 += dv1;
                q1 = nudge(q1, dw1);
                const V3 ai2 = -cross(lam, r2);
                const V3 dv2 = lock_mul((im2 * ll2), lam);
                const V3 dw2 = cmul(al2, mvmul(I2, ai2));
                x2 += dv2;
                q2 = nudge(q2, dw2);
            }
        } else if (type == JT_FIXED) {   // SolveFixedJointSystem.cpp:266-448
            bool ok;
            const M3 Kt = translation_inverse_mass(im1 + im2, r1, r2, I1, I2, anyDyn, ok);
            if (ok) {
                const V3 err = x2 + r2 - x1 - r1;
                const V3 lam = mvmul(Kt, -err);
                const V3 li1 = -lam;
                const V3 ai1 = cross(lam, r1);
                const V3 dv1 = lock_mul((im1 * ll1), li1);
                const V3 dw1 = cmul(al1, mvmul(I1, ai1));
                x1 += dv1;
                q1 = nudge(q1, dw1);
                const V3 ai2 = -cross(lam, r2);
                const V3 dv2 = lock_mul((im2 * ll2), lam);
                const V3 dw2 = cmul(al2, mvmul(I2, ai2));
                x2 += dv2;
                q2 = nudge(q2, dw2);
            }
            M3 Kr = madd(I1, I2);
            const float det = mdet(Kr);
            if (fabsf(det) > kMachineEps) {
                if (anyDyn) Kr = minv(Kr, det);
                const Q4 qe = qmul(qmul(q2, ldq(J.qinit, j)), qconj(q1));
                const V3 er = 2.0f * v3(qe.x, qe.y, qe.z);
                const V3 lr = mvmul(Kr, -er);
                const V3 ai1 = -lr;
                const V3 dw1 = cmul(al1, mvmul(I1, ai1));
                q1 = nudge(q1, dw1);
                const V3 dw2 = cmul(al2, mvmul(I2, lr));
                q2 = nudge(q2, dw2);
            }
        } else if (type == JT_HINGE) {   // SolveHingeJointSystem.cpp:444-730
            V3 a1 = normalized(qrot(q1, ld3(J.ax1, j)));
            const V3 a2 = normalized(qrot(q2, ld3(J.ax2, j)));
            const V3 b2v = one_unit_orthogonal(a2);
            const V3 c2v = cross(a2, b2v);
            const V3 bxa = cross(b2v, a1), cxa = cross(c2v, a1);
            J.a1[j] = f4(a1, 0.f);
            J.b2xa1[j] = f4(bxa, 0.f);
            J.c2xa1[j] = f4(cxa, 0.f);
            int fl = J.flags[j];
            const float4 lim = J.lim[j];
            if (fl & (JF_LIMIT | JF_MOTOR)) {
                const float ang = hinge_angle(q1, q2, ldq(J.qinit, j), a1, lim.x, lim.y);
                const float le = ang - lim.x, ue = lim.y - ang;
                fl = le <= 0 ? (fl | JF_LOWER_VIOLATED) : (fl & ~JF_LOWER_VIOLATED);
                fl = ue <= 0 ? (fl | JF_UPPER_VIOLATED) : (fl & ~JF_UPPER_VIOLATED);
                J.flags[j] = fl;
                if (fl & JF_LIMIT) {
                    const float imlm_old = J.imlm[j];
                    if (fl & (JF_LOWER_VIOLATED | JF_UPPER_VIOLATED)) {
                        const float m = dot(a1, mvmul(I1, a1)) + dot(a1, mvmul(I2, a1));
                        J.imlm[j] = imlm_old > 0.0f ? 1.0f / m : 0.0f;   // reference tests the old value
                    }
                    if (fl & JF_LOWER_VIOLATED) {
                        const float l = imlm_old * (-le);
                        q1 = nudge(q1, cmul(al1, mvmul(I1, (-l) * a1)));
                        q2 = nudge(q2, cmul(al2, mvmul(I2, l * a1)));
                    }
                    if (fl & JF_UPPER_VIOLATED) {
                        const float l = imlm_old * (-ue);
                        q1 = nudge(q1, cmul(al1, mvmul(I1, l * a1)));
                        q2 = nudge(q2, cmul(al2, mvmul(I2, (-l) * a1)));
                    }
                }
            }
            const V3 i1b = mvmul(I1, bxa), i1c = mvmul(I1, cxa), i2b = mvmul(I2, bxa), i2c = mvmul(I2, cxa);
            const float e11 = dot(bxa, i1b) + dot(bxa, i2b);
            const float e12 = dot(bxa, i1c) + dot(bxa, i2c);
            const float e21 = dot(cxa, i1b) + dot(cxa, i2b);
            const float e22 = dot(cxa, i1c) + dot(cxa, i2c);
            const float det = e11 * e22 - e21 * e12;
            if (fabsf(det) > kMachineEps) {
                float k00 = 0, k01 = 0, k10 = 0, k11 = 0;
                if (anyDyn) {
                    const float id = 1.0f / det;
                    k00 = e22 * id; k01 = -e12 * id; k10 = -e21 * id; k11 = e11 * id;
                }
                const float er0 = dot(a1, b2v), er1 = dot(a1, c2v);
                const float x0 = -er0, x1v = -er1;
                const float l0 = k00 * x0 + k01 * x1v;
                const float l1 = k10 * x0 + k11 * x1v;
                const V3 ai1 = (-bxa) * l0 - cxa * l1;
                q1 = nudge(q1, cmul(al1, mvmul(I1, ai1)));
                const V3 ai2 = bxa * l0 + cxa * l1;
                q2 = nudge(q2, cmul(al2, mvmul(I2, ai2)));
            }
            bool ok;
            const M3 Kt = translation_inverse_mass(im1 + im2, r1, r2, I1, I2, anyDyn, ok);
            if (ok) {
                const V3 err = x2 + r2 - x1 - r1;
                const V3 lam = mvmul(Kt, -err);
                const V3 li1 = -lam;
                const V3 ai1 = cross(lam, r1);
                const V3 dv1 = lock_mul((im1 * ll1), li1);
                const V3 dw1 = cmul(al1, mvmul(I1, ai1));
                x1 += dv1;
                q1 = nudge(q1, dw1);
                const V3 ai2 = -cross(lam, r2);
                const V3 dv2 = lock_mul((im2 * ll2), lam);
                const V3 dw2 = cmul(al2, mvmul(I2, ai2));
                x2 += dv2;
                q2 = nudge(q2, dw2);
            }
        } else {   // slider, SolveSliderJointSystem.cpp:514-826
            const V3 u = x2 + r2 - x1 - r1;
            const V3 ax = normalized(qrot(q1, ld3(J.ax1, j)));
            const V3 n1 = one_unit_orthogonal(ax);
            const V3 n2 = cross(ax, n1);
            const float4 lim = J.lim[j];
            const float ud = dot(u, ax);
            int fl = J.flags[j];
            fl = (ud - lim.x) <= 0 ? (fl | JF_LOWER_VIOLATED) : (fl & ~JF_LOWER_VIOLATED);
            fl = (lim.y - ud) <= 0 ? (fl | JF_UPPER_VIOLATED) : (fl & ~JF_UPPER_VIOLATED);
            J.flags[j] = fl;
            const V3 rn1 = cross(r2, n1), rn2 = cross(r2, n2), ra = cross(r2, ax);
            J.ra[j] = f4(ra, 0.f);   // read by the next step's limit inverse mass
            const V3 rpu = r1 + u;
            const V3 un1 = cross(rpu, n1), un2 = cross(rpu, n2), ua = cross(rpu, ax);
            const float ims = im1 + im2;
            if (fl & JF_LIMIT) {
                float imlm = J.imlm[j];
                if (fl & (JF_LOWER_VIOLATED | JF_UPPER_VIOLATED)) {
                    const float m = ims + dot(ua, mvmul(I1, ua)) + dot(ra, mvmul(I2, ra));
                    imlm = m > 0.0f ? 1.0f / m : 0.0f;
                    J.imlm[j] = imlm;
                }
                if (fl & JF_LOWER_VIOLATED) {
                    const float err = dot(x2 + r2 - x1 - r1, ax) - lim.x;
                    const float l = imlm * (-err);
                    x1 += lock_mul((im1 * ll1), (-l) * ax);
                    q1 = nudge(q1, cmul(al1, mvmul(I1, (-l) * ua)));
                    x2 += lock_mul((im2 * ll2), l * ax);
                    q2 = nudge(q2, cmul(al2, mvmul(I2, l * ra)));
                }
                if (fl & JF_UPPER_VIOLATED) {
                    const float err = lim.y - dot(x2 + r2 - x1 - r1, ax);
                    const float l = imlm * (-err);
                    x1 += lock_mul((im1 * ll1), l * ax);
                    q1 = nudge(q1, cmul(al1, mvmul(I1, l * ua)));
                    x2 += lock_mul((im2 * ll2), (-l) * ax);
                    q2 = nudge(q2, cmul(al2, mvmul(I2, (-l) * ra)));
                }
            }
            // 3 rotation rows
            M3 Kr = madd(I1, I2);
            const float detr = mdet(Kr);
            if (fabsf(detr) > kMachineEps) {
                if (anyDyn) Kr = minv(Kr, detr);
                const Q4 qe = qmul(qmul(q2, ldq(J.qinit, j)), qconj(q1));
                const V3 er = 2.0f * v3(qe.x, qe.y, qe.z);
                const V3 lr = mvmul(Kr, -er);
                q1 = nudge(q1, cmul(al1, mvmul(I1, -lr)));
                q2 = nudge(q2, cmul(al2, mvmul(I2, lr)));
            }
            // 2 translation rows (with u and the cross products from before the corrections above)
            const V3 i1n1 = mvmul(I1, un1), i1n2 = mvmul(I1, un2), i2n1 = mvmul(I2, rn1), i2n2 = mvmul(I2, rn2);
            const float e11 = ims + dot(un1, i1n1) + dot(rn1, i2n1);
            const float e12 = dot(un1, i1n2) + dot(rn1, i2n2);
            const float e21 = dot(un2, i1n1) + dot(rn2, i2n1);
            const float e22 = ims + dot(un2, i1n2) + dot(rn2, i2n2);
            const float det = e11 * e22 - e21 * e12;
            if (fabsf(det) > kMachineEps) {
                float k00 = 0, k01 = 0, k10 = 0, k11 = 0;
                if (anyDyn) {
                    const float id = 1.0f / det;
                    k00 = e22 * id; k01 = -e12 * id; k10 = -e21 * id; k11 = e11 * id;
                }
                const float t0 = -dot(u, n1), t1 = -dot(u, n2);
                const float l0 = k00 * t0 + k01 * t1;
                const float l1 = k10 * t0 + k11 * t1;
                const V3 li1 = (-n1) * l0 - n2 * l1;
                const V3 ai1 = (-un1) * l0 - un2 * l1;
                x1 += lock_mul((im1 * ll1), li1);
                q1 = nudge(q1, cmul(al1, mvmul(I1, ai1)));
                const V3 li2 = n1 * l0 + n2 * l1;
                const V3 ai2 = rn1 * l0 + rn2 * l1;
                x2 += lock_mul((im2 * ll2), li2);
                q2 = nudge(q2, cmul(al2, mvmul(I2, ai2)));
            }
        }
        B.cx[b1] = f4(x1, 0.f); B.cq[b1] = f4q(q1);
        B.cx[b2] = f4(x2, 0.f); B.cq[b2] = f4q(q2);
    }
}

}  // namespace rp3dg
