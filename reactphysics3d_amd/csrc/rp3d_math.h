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
