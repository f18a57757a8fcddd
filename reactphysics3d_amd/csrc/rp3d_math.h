// FP32 vector / quaternion / transform arithmetic for the step path.
//
// Every expression keeps the operand order of the ReactPhysics3D 0.10.2
// mathematics headers so that, compiled with -ffp-contract=off and IEEE
// division/sqrt, the device produces the same bits as the x86-64 reference
// build.  Citations point at the reference function each helper matches.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define RD __host__ __device__ __forceinline__

namespace rp3dg {

constexpr float kMachineEps = 1.1920928955078125e-07f;        // configuration.h:120
constexpr float kDecimalLargest = 3.40282346638528859812e+38f;  // configuration.h:117
constexpr float kDecimalSmallest = -3.40282346638528859812e+38f;  // configuration.h:114
constexpr float kPi = 3.141592653589f;                          // configuration.h:123

// ---- acos / atan2 with the bits of the reference's C library ----
// The reference calls std::acos / std::atan2 on floats (SolveBallAndSocketJointSystem.h:153,
// SolveHingeJointSystem.cpp:786-789), i.e. acosf / atan2f of the x86-64 glibc it links, whose float
// versions follow fdlibm's e_acosf.c, s_atanf.c and e_atan2f.c (Sun Microsystems, 1993).  The GPU's own
// acosf / atan2f differ from those by 1-3 ulp on about a third of the arguments, which the hinge and
// cone limits turn into different limit decisions, so the step uses this restatement of the fdlibm
// method instead: same constants (as the float values the decimal literals round to), same operation
// order.  Compiled with -ffp-contract=off and IEEE division / sqrt it returns glibc's bits (checked on
// every float in [-1, 1] for acos and on 6e8 arguments for atan / atan2).
//
// ref_acosf, ref_atanf and ref_atan2f restate fdlibm's float routines, which carry this notice:
//
// ====================================================
// Copyright (C) 1993 by Sun Microsystems, Inc. All rights reserved.
//
// Developed at SunPro, a Sun Microsystems, Inc. business.
// Permission to use, copy, modify, and distribute this
// software is freely granted, provided that this notice
// is preserved.
// ====================================================
//
// (float conversion of e_acosf.c, s_atanf.c and e_atan2f.c by Ian Lance Taylor, Cygnus Support.)
RD float fbits_as(int32_t i) { return __builtin_bit_cast(float, i); }
RD int32_t fbits_of(float f) { return __builtin_bit_cast(int32_t, f); }

RD float ref_acosf(float x) {
    const float one = 1.0f, pi = fbits_as(0x40490fda), pio2_hi = fbits_as(0x3fc90fda), pio2_lo = fbits_as(0x33a22168);
    const float pS0 = fbits_as(0x3e2aaaab), pS1 = fbits_as((int32_t)0xbea6b090), pS2 = fbits_as(0x3e4e0aa8),
                pS3 = fbits_as((int32_t)0xbd241146), pS4 = fbits_as(0x3a4f7f04), pS5 = fbits_as(0x3811ef08);
    const float qS1 = fbits_as((int32_t)0xc019d139), qS2 = fbits_as(0x4001572d), qS3 = fbits_as((int32_t)0xbf303361),
                qS4 = fbits_as(0x3d9dc62e);
    const int32_t hx = fbits_of(x), ix = hx & 0x7fffffff;
    if (ix == 0x3f800000) return hx > 0 ? 0.0f : pi + 2.0f * pio2_lo;   // |x| = 1
    if (ix > 0x3f800000) return (x - x) / (x - x);                        // |x| > 1: NaN
    if (ix < 0x3f000000) {                                                 // |x| < 0.5
        if (ix <= 0x23000000) return pio2_hi + pio2_lo;
        const float z = x * x;
        const float p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
        const float q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
        const float r = p / q;
        return pio2_hi - (x - (pio2_lo - x * r));
    }
    if (hx < 0) {                                                          // x <= -0.5
        const float z = (one + x) * 0.5f;
        const float p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
        const float q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
        const float s = sqrtf(z);
        const float r = p / q;
        const float w = r * s - pio2_lo;
        return pi - 2.0f * (s + w);
    }
    const float z = (one - x) * 0.5f;                                      // x >= 0.5
    const float s = sqrtf(z);
    const float df = fbits_as(fbits_of(s) & (int32_t)0xfffff000);
    const float c = (z - df * df) / (s + df);
    const float p = z * (pS0 + z * (pS1 + z * (pS2 + z * (pS3 + z * (pS4 + z * pS5)))));
    const float q = one + z * (qS1 + z * (qS2 + z * (qS3 + z * qS4)));
    const float r = p / q;
    const float w = r * s + c;
    return 2.0f * (df + w);
}

RD float ref_atanf(float x) {
    const float one = 1.0f;
    const int32_t hx = fbits_of(x), ix = hx & 0x7fffffff;
    int id;
    if (ix >= 0x4c000000) {                                                // |x| >= 2^25
        if (ix > 0x7f800000) return x + x;
        const float hi = fbits_as(0x3fc90fda), lo = fbits_as(0x33a22168);
        return hx > 0 ? hi + lo : -hi - lo;
    }
    if (ix < 0x3ee00000) {                                                 // |x| < 7/16
        if (ix < 0x31000000) return x;
        id = -1;
    } else {
        x = fabsf(x);
        if (ix < 0x3f980000) {
            if (ix < 0x3f300000) { id = 0; x = (2.0f * x - one) / (2.0f + x); }
            else { id = 1; x = (x - one) / (x + one); }
        } else if (ix < 0x401c0000) {
            id = 2; x = (x - 1.5f) / (one + 1.5f * x);
        } else {
            id = 3; x = -1.0f / x;
        }
    }
    // atan(x) = x - x^3 a0 + x^5 a1 - ...: odd and even halves of the series in w = x^4
    const float a0 = fbits_as(0x3eaaaaab), a1 = fbits_as((int32_t)0xbe4ccccd), a2 = fbits_as(0x3e124925),
                a3 = fbits_as((int32_t)0xbde38e38), a4 = fbits_as(0x3dba2e6e), a5 = fbits_as((int32_t)0xbd9d8795),
                a6 = fbits_as(0x3d886b35), a7 = fbits_as((int32_t)0xbd6ef16b), a8 = fbits_as(0x3d4bda59),
                a9 = fbits_as((int32_t)0xbd15a221), a10 = fbits_as(0x3c8569d7);
    float z = x * x;
    const float w = z * z;
    const float s1 = z * (a0 + w * (a2 + w * (a4 + w * (a6 + w * (a8 + w * a10)))));
    const float s2 = w * (a1 + w * (a3 + w * (a5 + w * (a7 + w * a9))));
    if (id < 0) return x - x * (s1 + s2);
    // atan(0.5), atan(1), atan(1.5), atan(inf) split into high and low parts
    const float hi = id == 0 ? fbits_as(0x3eed6338) : id == 1 ? fbits_as(0x3f490fda)
                   : id == 2 ? fbits_as(0x3f7b985e) : fbits_as(0x3fc90fda);
    const float lo = id == 0 ? fbits_as(0x31ac3769) : id == 1 ? fbits_as(0x33222168)
                   : id == 2 ? fbits_as(0x33140fb4) : fbits_as(0x33a22168);
    z = hi - ((x * (s1 + s2) - lo) - x);
    return hx < 0 ? -z : z;
}

RD float ref_atan2f(float y, float x) {
    const float tiny = 1.0e-30f, pi_o_4 = fbits_as(0x3f490fdb), pi_o_2 = fbits_as(0x3fc90fdb),
                pi = fbits_as(0x40490fdb), pi_lo = fbits_as((int32_t)0xb3bbbd2e);
    const int32_t hx = fbits_of(x), ix = hx & 0x7fffffff, hy = fbits_of(y), iy = hy & 0x7fffffff;
    if (ix > 0x7f800000 || iy > 0x7f800000) return x + y;                // NaN
    if (hx == 0x3f800000) return ref_atanf(y);                            // x = 1
    const int m = ((hy >> 31) & 1) | ((hx >> 30) & 2);                    // 2 sign(x) + sign(y)
    if (iy == 0) return m <= 1 ? y : (m == 2 ? pi + tiny : -pi - tiny);
    if (ix == 0) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
    if (ix == 0x7f800000) {
        if (iy == 0x7f800000)
            return m == 0 ? pi_o_4 + tiny : m == 1 ? -pi_o_4 - tiny : m == 2 ? 3.0f * pi_o_4 + tiny : -3.0f * pi_o_4 - tiny;
        return m == 0 ? 0.0f : m == 1 ? -0.0f : m == 2 ? pi + tiny : -pi - tiny;
    }
    if (iy == 0x7f800000) return hy < 0 ? -pi_o_2 - tiny : pi_o_2 + tiny;
    const int k = (iy - ix) >> 23;
    float z;
    if (k > 60) z = pi_o_2 + 0.5f * pi_lo;                                // |y / x| > 2^60
    else if (hx < 0 && k < -60) z = 0.0f;
    else z = ref_atanf(fabsf(y / x));
    switch (m) {
        case 0: return z;
        case 1: return -z;
        case 2: return pi - (z - pi_lo);
        default: return (z - pi_lo) - pi;
    }
}

struct V3 { float x, y, z; };
struct Q4 { float x, y, z, w; };
struct Tf { V3 p; Q4 q; };
struct M3 { V3 r[3]; };

RD V3 v3(float x, float y, float z) { V3 r; r.x = x; r.y = y; r.z = z; return r; }
RD V3 v3zero() { return v3(0.f, 0.f, 0.f); }
RD Q4 q4(float x, float y, float z, float w) { Q4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
RD V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
RD V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
RD V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
// Vector3.h: operator*(Vector3, decimal) multiplies number * component.
RD V3 operator*(V3 a, float s) { return v3(s * a.x, s * a.y, s * a.z); }
RD V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
RD V3 operator/(V3 a, float s) { return v3(a.x / s, a.y / s, a.z / s); }
RD V3 cmul(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
RD V3& operator+=(V3& a, V3 b) { a.x += b.x; a.y += b.y; a.z += b.z; return a; }
RD V3& operator-=(V3& a, V3 b) { a.x -= b.x; a.y -= b.y; a.z -= b.z; return a; }
RD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RD V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
RD float len2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
RD float len(V3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
// std::min / std::max semantics: min(a,b) = (b < a) ? b : a.
RD float fmin_s(float a, float b) { return (b < a) ? b : a; }
RD float fmax_s(float a, float b) { return (a < b) ? b : a; }
RD float clampf(float v, float lo, float hi) { return fmin_s(fmax_s(v, lo), hi); }  // mathematics_functions.h:52
RD V3 vmin(V3 a, V3 b) { return v3(fmin_s(a.x, b.x), fmin_s(a.y, b.y), fmin_s(a.z, b.z)); }
RD V3 vmax(V3 a, V3 b) { return v3(fmax_s(a.x, b.x), fmax_s(a.y, b.y), fmax_s(a.z, b.z)); }
RD float comp(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
RD void setcomp(V3& a, int i, float v) { if (i == 0) a.x = v; else if (i == 1) a.y = v; else a.z = v; }

// Vector3.cpp:33-43 (getUnit multiplies by the reciprocal of the length).
RD V3 unit(V3 a) {
    float l = len(a);
    if (l < kMachineEps) return a;
    float inv = 1.0f / l;
    return v3(a.x * inv, a.y * inv, a.z * inv);
}
// Vector3.h normalize(): divides each component by the length.
RD V3 normalized(V3 a) {
    float l = len(a);
    if (l < kMachineEps) return a;
    return v3(a.x / l, a.y / l, a.z / l);
}
RD int min_axis(V3 a) { return (a.x < a.y ? (a.x < a.z ? 0 : 2) : (a.y < a.z ? 1 : 2)); }
// Vector3.cpp:46-64
RD V3 one_unit_orthogonal(V3 a) {
    V3 ab = v3(fabsf(a.x), fabsf(a.y), fabsf(a.z));
    int m = min_axis(ab);
    if (m == 0) return v3(0.0f, -a.z, a.y) / sqrtf(a.y * a.y + a.z * a.z);
    if (m == 1) return v3(-a.z, 0.0f, a.x) / sqrtf(a.x * a.x + a.z * a.z);
    return v3(-a.y, a.x, 0.0f) / sqrtf(a.x * a.x + a.y * a.y);
}

// ---- quaternions (Quaternion.h) ----
RD Q4 qconj(Q4 q) { return q4(-q.x, -q.y, -q.z, q.w); }
RD float qlen(Q4 q) { return sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w); }
RD Q4 qunit(Q4 q) { float l = qlen(q); return q4(q.x / l, q.y / l, q.z / l, q.w / l); }
RD Q4 qadd(Q4 a, Q4 b) { return q4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
RD Q4 qscale(Q4 a, float s) { return q4(s * a.x, s * a.y, s * a.z, s * a.w); }
// Quaternion.h:349-361
RD Q4 qmul(Q4 a, Q4 b) {
    return q4(a.w * b.x + b.w * a.x + a.y * b.z - a.z * b.y,
              a.w * b.y + b.w * a.y + a.z * b.x - a.x * b.z,
              a.w * b.z + b.w * a.z + a.x * b.y - a.y * b.x,
              a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z);
}
// Quaternion.h:365-379 (rotate a vector)
RD V3 qrot(Q4 q, V3 p) {
    const float px = q.w * p.x + q.y * p.z - q.z * p.y;
    const float py = q.w * p.y + q.z * p.x - q.x * p.z;
    const float pz = q.w * p.z + q.x * p.y - q.y * p.x;
    const float pw = -q.x * p.x - q.y * p.y - q.z * p.z;
    return v3(q.w * px - py * q.z + pz * q.y - pw * q.x,
              q.w * py - pz * q.x + px * q.z - pw * q.y,
              q.w * pz - px * q.y + py * q.x - pw * q.z);
}
// Quaternion.cpp:138-165
RD M3 qmat(Q4 q) {
    float nQ = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    float s = 0.0f;
    if (nQ > 0.0f) s = 2.0f / nQ;
    float xs = q.x * s, ys = q.y * s, zs = q.z * s;
    float wxs = q.w * xs, wys = q.w * ys, wzs = q.w * zs;
    float xxs = q.x * xs, xys = q.x * ys, xzs = q.x * zs;
    float yys = q.y * ys, yzs = q.y * zs, zzs = q.z * zs;
    M3 m;
    m.r[0] = v3(1.0f - yys - zzs, xys - wzs, xzs + wys);
    m.r[1] = v3(xys + wzs, 1.0f - xxs - zzs, yzs - wxs);
    m.r[2] = v3(xzs - wys, yzs + wxs, 1.0f - xxs - yys);
    return m;
}

// ---- 3x3 matrices (Matrix3x3.h) ----
RD float mel(const M3& m, int i, int j) { return comp(m.r[i], j); }
RD V3 mvmul(const M3& m, V3 v) {  // Matrix3x3.h:324-331
    return v3(m.r[0].x * v.x + m.r[0].y * v.y + m.r[0].z * v.z,
              m.r[1].x * v.x + m.r[1].y * v.y + m.r[1].z * v.z,
              m.r[2].x * v.x + m.r[2].y * v.y + m.r[2].z * v.z);
}
RD M3 mmul(const M3& a, const M3& b) {  // Matrix3x3.h:302-320
    M3 r;
    r.r[0] = v3(a.r[0].x * b.r[0].x + a.r[0].y * b.r[1].x + a.r[0].z * b.r[2].x,
                a.r[0].x * b.r[0].y + a.r[0].y * b.r[1].y + a.r[0].z * b.r[2].y,
                a.r[0].x * b.r[0].z + a.r[0].y * b.r[1].z + a.r[0].z * b.r[2].z);
    r.r[1] = v3(a.r[1].x * b.r[0].x + a.r[1].y * b.r[1].x + a.r[1].z * b.r[2].x,
                a.r[1].x * b.r[0].y + a.r[1].y * b.r[1].y + a.r[1].z * b.r[2].y,
                a.r[1].x * b.r[0].z + a.r[1].y * b.r[1].z + a.r[1].z * b.r[2].z);
    r.r[2] = v3(a.r[2].x * b.r[0].x + a.r[2].y * b.r[1].x + a.r[2].z * b.r[2].x,
                a.r[2].x * b.r[0].y + a.r[2].y * b.r[1].y + a.r[2].z * b.r[2].y,
                a.r[2].x * b.r[0].z + a.r[2].y * b.r[1].z + a.r[2].z * b.r[2].z);
    return r;
}
RD M3 mtranspose(const M3& m) {
    M3 r;
    r.r[0] = v3(m.r[0].x, m.r[1].x, m.r[2].x);
    r.r[1] = v3(m.r[0].y, m.r[1].y, m.r[2].y);
    r.r[2] = v3(m.r[0].z, m.r[1].z, m.r[2].z);
    return r;
}
// RigidBody.h:246-262: I_w^-1 = R * (diag(I_l^-1) * R^T)
RD M3 world_inv_inertia(const M3& R, V3 il) {
    M3 t;
    t.r[0] = v3(R.r[0].x * il.x, R.r[1].x * il.x, R.r[2].x * il.x);
    t.r[1] = v3(R.r[0].y * il.y, R.r[1].y * il.y, R.r[2].y * il.y);
    t.r[2] = v3(R.r[0].z * il.z, R.r[1].z * il.z, R.r[2].z * il.z);
    return mmul(R, t);
}

// ---- transforms (Transform.h) ----
RD Tf tf(V3 p, Q4 q) { Tf t; t.p = p; t.q = q; return t; }
RD V3 tapply(const Tf& t, V3 v) { return qrot(t.q, v) + t.p; }  // Transform.h operator*(Vector3)
RD Tf tinv(const Tf& t) {  // Transform.h:162-165
    Q4 iq = qconj(t.q);
    return tf(qrot(iq, -t.p), iq);
}
// Transform.h operator*(Transform)
RD Tf tmul(const Tf& a, const Tf& b) {
    const Q4& q = a.q;
    const V3& p2 = b.p;
    const float px = q.w * p2.x + q.y * p2.z - q.z * p2.y;
    const float py = q.w * p2.y + q.z * p2.x - q.x * p2.z;
    const float pz = q.w * p2.z + q.x * p2.y - q.y * p2.x;
    const float pw = -q.x * p2.x - q.y * p2.y - q.z * p2.z;
    V3 p = v3(a.p.x + q.w * px - py * q.z + pz * q.y - pw * q.x,
              a.p.y + q.w * py - pz * q.x + px * q.z - pw * q.y,
              a.p.z + q.w * pz - px * q.y + py * q.x - pw * q.z);
    const Q4& r = b.q;
    Q4 o = q4(q.w * r.x + r.w * q.x + q.y * r.z - q.z * r.y,
              q.w * r.y + r.w * q.y + q.z * r.x - q.x * r.z,
              q.w * r.z + r.w * q.z + q.x * r.y - q.y * r.x,
              q.w * r.w - q.x * r.x - q.y * r.y - q.z * r.z);
    return tf(p, o);
}

// ---- geometric helpers (mathematics_functions.h) ----
RD bool parallel_vectors(V3 a, V3 b) { return len2(cross(a, b)) < 0.00001f; }  // :60
RD V3 closest_point_on_segment(V3 a, V3 b, V3 c) {  // :76-90
    const V3 ab = b - a;
    float l2 = len2(ab);
    if (l2 < kMachineEps) return a;
    float t = dot(c - a, ab) / l2;
    if (t < 0.0f) t = 0.0f;
    if (t > 1.0f) t = 1.0f;
    return a + t * ab;
}
// :94-160
RD void closest_points_between_segments(V3 s1a, V3 s1b, V3 s2a, V3 s2b, V3& c1, V3& c2) {
    const V3 d1 = s1b - s1a;
    const V3 d2 = s2b - s2a;
    const V3 r = s1a - s2a;
    float a = len2(d1);
    float e = len2(d2);
    float f = dot(d2, r);
    float s, t;
    if (a <= kMachineEps && e <= kMachineEps) { c1 = s1a; c2 = s2a; return; }
    if (a <= kMachineEps) {
        s = 0.0f;
        t = clampf(f / e, 0.0f, 1.0f);
    } else {
        float c = dot(d1, r);
        if (e <= kMachineEps) {
            t = 0.0f;
            s = clampf(-c / a, 0.0f, 1.0f);
        } else {
            float b = dot(d1, d2);
            float denom = a * e - b * b;
            if (denom != 0.0f) s = clampf((b * f - c * e) / denom, 0.0f, 1.0f);
            else s = 0.0f;
            t = (b * s + f) / e;
            if (t < 0.0f) { t = 0.0f; s = clampf(-c / a, 0.0f, 1.0f); }
            else if (t > 1.0f) { t = 1.0f; s = clampf((b - c) / a, 0.0f, 1.0f); }
        }
    }
    c1 = s1a + d1 * s;
    c2 = s2a + d2 * t;
}
// :200-212
RD float plane_segment_intersection(V3 a, V3 b, float planeD, V3 n) {
    const float eps = 0.0001f;
    float t = -1.0f;
    const float nab = dot(n, b - a);
    if (fabsf(nab) > eps) t = (planeD - dot(n, a)) / nab;
    return t;
}
RD V3 project_point_onto_plane(V3 p, V3 n, V3 pp) { return p - dot(n, p - pp) * n; }  // :394

RD uint64_t pair_numbers(uint32_t a, uint32_t b) { uint64_t x = a, y = b; return x * x + x + y; }  // :428

}  // namespace rp3dg
