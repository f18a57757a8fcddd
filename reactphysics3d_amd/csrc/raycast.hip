// World raycasts (PhysicsWorld::raycast -> CollisionDetectionSystem::raycast -> BroadPhaseSystem::raycast
// -> DynamicAABBTree::raycast -> Collider::raycast -> <Shape>::raycast), batched: one thread per ray,
// with the callback that keeps the closest hit (RaycastCallback::notifyRaycastHit returning the hit
// fraction, which clips the ray), or the one that keeps every hit (notifyRaycastHit returning 1: the
// ray is never clipped, so the hits are the colliders whose shape the ray crosses within its own max
// fraction, a set that does not depend on the visiting order).  Only world-query colliders whose
// category passes the mask are tested (BroadPhaseRaycastCallback, BroadPhaseSystem.cpp:224-241).
//
// A thread walks the colliders of its ray's world in index order instead of the tree: every collider
// whose fat AABB the (clipped) ray crosses (AABB::testRayIntersect, AABB.h:285-313) and whose category
// passes the mask is tested against its shape in the collider's space, and each reported hit clips the
// ray.  The closest hit is the same in any visiting order; only hits at exactly equal fractions (the
// box and convex-mesh tests accept a hit at the current fraction, the sphere and capsule tests do not)
// may be reported in a different order than the tree visits them.  A hit at fraction 0 ends the ray,
// as in DynamicAABBTree.cpp:711-714.
#include <hip/hip_runtime.h>

#include "kcommon.h"
#include "raycast.h"

namespace rp3dg {

namespace {

struct Hit {
    float t;
    V3 p, n;
};

// SphereShape::raycast (SphereShape.cpp:61-104); also the capsule end caps
// (CapsuleShape::raycastWithSphereEndCap, CapsuleShape.cpp:243-292)
__device__ bool ray_sphere(V3 p1, V3 p2, V3 centre, float r, float maxf, Hit& h) {
    const V3 m = p1 - centre;
    const float c = dot(m, m) - r * r;
    if (c < 0.0f) return false;
    const V3 dir = p2 - p1;
    const float b = dot(m, dir);
    if (b > 0.0f) return false;
    const float rsl = len2(dir);
    const float disc = b * b - rsl * c;
    if (disc < 0.0f || rsl < kMachineEps) return false;
    float t = -b - sqrtf(disc);
    if (t < maxf * rsl) {
        t /= rsl;
        h.t = t;
        h.p = p1 + t * dir;
        h.n = h.p - centre;
        return true;
    }
    return false;
}

__device__ __forceinline__ float vcomp(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }

// BoxShape::raycast (BoxShape.cpp:62-121)
__device__ bool ray_box(V3 p1, V3 p2, V3 he, float maxf, Hit& h) {
    const V3 dir = p2 - p1;
    float tmin = kDecimalSmallest, tmax = kDecimalLargest;
    V3 nd = v3(0.f, 0.f, 0.f);
    for (int i = 0; i < 3; i++) {
        const float di = vcomp(dir, i), pi = vcomp(p1, i), hi = vcomp(he, i);
        if (fabsf(di) < kMachineEps) {
            if (pi > hi || pi < -hi) return false;
        } else {
            const float inv = 1.0f / di;
            float t1 = (-hi - pi) * inv;
            float t2 = (hi - pi) * inv;
            V3 cn = v3(i == 0 ? -hi : 0.0f, i == 1 ? -hi : 0.0f, i == 2 ? -hi : 0.0f);
            if (t1 > t2) {
                const float s = t1;
                t1 = t2;
                t2 = s;
                cn = -cn;
            }
            if (t1 > tmin) {
                tmin = t1;
                nd = cn;
            }
            tmax = fmin_s(tmax, t2);
            if (tmin > maxf) return false;
            if (tmin > tmax) return false;
        }
    }
    if (tmin < 0.0f || tmin > maxf) return false;
    h.t = tmin;
    h.p = p1 + tmin * dir;
    h.n = nd;
    return true;
}

// CapsuleShape::raycast (CapsuleShape.cpp:97-240); capsule axis along local y
__device__ bool ray_capsule(V3 p1, V3 p2, float r, float hh, float maxf, Hit& h) {
    const V3 n = p2 - p1;
    const float epsilon = 0.01f;
    const V3 p = v3(0.0f, -hh, 0.0f), q = v3(0.0f, hh, 0.0f);
    const V3 d = q - p;
    const V3 m = p1 - p;
    const float mDotD = dot(m, d), nDotD = dot(n, d), dDotD = dot(d, d);
    const float vec1DotD = dot(p1 - v3(0.0f, -hh - r, 0.0f), d);
    if (vec1DotD < 0.0f && vec1DotD + nDotD < 0.0f) return false;
    const float extra = 2.0f * r * d.y;
    if (vec1DotD > dDotD + extra && vec1DotD + nDotD > dDotD + extra) return false;
    const float nDotN = dot(n, n), mDotN = dot(m, n);
    const float a = dDotD * nDotN - nDotD * nDotD;
    const float k = dot(m, m) - r * r;
    const float c = dDotD * k - mDotD * mDotD;
    if (fabsf(a) < epsilon) {
        if (c > 0.0f) return false;
        if (mDotD < 0.0f) return ray_sphere(p1, p2, p, r, maxf, h);
        if (mDotD > dDotD) return ray_sphere(p1, p2, q, r, maxf, h);
        return false;
    }
    const float b = dDotD * mDotN - nDotD * mDotD;
    const float disc = b * b - a * c;
    if (disc < 0.0f) return false;
    const float t = (-b - sqrtf(disc)) / a;
    const float value = mDotD + t * nDotD;
    if (value < 0.0f) return ray_sphere(p1, p2, p, r, maxf, h);
    if (value > dDotD) return ray_sphere(p1, p2, q, r, maxf, h);
    if (t < 0.0f || t > maxf) return false;
    const V3 hp = p1 + t * n;
    const V3 v = hp - p;
    const V3 w = (dot(v, d) / len2(d)) * d;
    h.t = t;
    h.p = hp;
    h.n = unit(hp - (p + w));
    return true;
}

// ConvexMeshShape::raycast (ConvexMeshShape.cpp:107-170): clip the ray against every face plane
__device__ bool ray_hull(const RayArgs& A, int hull, V3 p1, V3 p2, float maxf, Hit& h) {
    const HullHeader H = A.hulls[hull];
    const V3 dir = p2 - p1;
    float tmin = 0.0f, tmax = maxf;
    V3 fn = v3(0.f, 0.f, 0.f);
    bool found = false;
    for (int f = 0; f < H.nf; f++) {
        const V3 nrm = ld3(A.hnormals, H.fbase + f);
        const V3 fp = ld3(A.hverts, H.vbase + A.hfvidx[A.hfverts[2 * (H.fbase + f)]]);
        const float denom = dot(nrm, dir);
        const float planeD = dot(nrm, fp);
        const float dist = planeD - dot(nrm, p1);
        if (denom == 0.0f) {
            if (dist < 0.0f) return false;
        } else {
            const float t = dist / denom;
            if (denom < 0.0f) {
                if (t > tmin) {
                    tmin = t;
                    fn = nrm;
                    found = true;
                }
            } else if (t < tmax) {
                tmax = t;
            }
            if (tmin > tmax) return false;
        }
    }
    if (!found) return false;
    h.t = tmin;
    h.p = p1 + tmin * dir;
    h.n = fn;
    return true;
}

// AABB::testRayIntersect (AABB.h:285-313)
__device__ __forceinline__ bool ray_aabb(V3 mn, V3 mx, V3 o, V3 inv, float maxf) {
    float t1 = (mn.x - o.x) * inv.x, t2 = (mx.x - o.x) * inv.x;
    float tmin = fmin_s(t1, t2);
    float tmax = fmax_s(t1, t2);
    tmax = fmin_s(tmax, maxf);
    t1 = (mn.y - o.y) * inv.y;
    t2 = (mx.y - o.y) * inv.y;
    tmin = fmax_s(tmin, fmin_s(t1, t2));
    tmax = fmin_s(tmax, fmax_s(t1, t2));
    t1 = (mn.z - o.z) * inv.z;
    t2 = (mx.z - o.z) * inv.z;
    tmin = fmax_s(tmin, fmin_s(t1, t2));
    tmax = fmin_s(tmax, fmax_s(t1, t2));
    return tmax >= fmax_s(tmin, 0.0f);
}

// One collider against a ray (BroadPhaseRaycastCallback::raycastBroadPhaseShape + Collider::raycast,
// Collider.cpp:173-193): the ray in the collider's space, the hit back in world space.
__device__ bool ray_collider(const RayArgs& A, int c, uint32_t mask, V3 p1, V3 p2, V3 inv, float maxf, Hit& hw) {
    if ((A.C.filter[c] & mask & 0xffffu) == 0u || !(A.C.kind[c] & CK_WORLD_QUERY)) return false;
    if (!ray_aabb(ld3(A.C.amin, c), ld3(A.C.amax, c), p1, inv, maxf)) return false;
    const Tf l2w = tf(ld3(A.C.wpos, c), ldq(A.C.wq, c));
    const Tf w2l = tinv(l2w);
    const V3 lp1 = tapply(w2l, p1), lp2 = tapply(w2l, p2);
    const float4 s = A.C.shape[c];
    Hit h;
    bool hit;
    switch (A.C.type[c]) {
        case SHAPE_SPHERE: hit = ray_sphere(lp1, lp2, v3(0.f, 0.f, 0.f), s.x, maxf, h); break;
        case SHAPE_BOX: hit = ray_box(lp1, lp2, v3(s.x, s.y, s.z), maxf, h); break;
        case SHAPE_CAPSULE: hit = ray_capsule(lp1, lp2, s.x, s.y, maxf, h); break;
        default: hit = ray_hull(A, (int)s.x, lp1, lp2, maxf, h); break;
    }
    if (!hit) return false;
    hw.t = h.t;
    hw.p = tapply(l2w, h.p);
    hw.n = normalized(qrot(l2w.q, h.n));
    return true;
}

}  // namespace

// Every hit of every ray: hits are appended at an atomically reserved slot (those past cap are counted,
// not written); *count ends as the number of hits.
__global__ void rc_all(RayArgs A, const int* ray_world, const float4* p1s, const float4* p2s, int n, uint32_t mask,
                       int2* hit_ray_collider, float4* hit_point, float4* hit_normal, int cap, int* count) {
    RP3DG_GRID_STRIDE(i, n) {
        const float4 a = p1s[i], b = p2s[i];
        const V3 p1 = v3(a.x, a.y, a.z), p2 = v3(b.x, b.y, b.z);
        const float maxf = b.w;
        const V3 dir = p2 - p1;
        const V3 inv = v3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
        const int w = ray_world[i];
        for (int c = A.wcoll[w]; c < A.wcoll[w + 1]; c++) {
            Hit h;
            if (!ray_collider(A, c, mask, p1, p2, inv, maxf, h)) continue;
            const int slot = atomicAdd(count, 1);
            if (slot < cap) {
                hit_ray_collider[slot] = make_int2(i, c);
                hit_point[slot] = make_float4(h.p.x, h.p.y, h.p.z, h.t);
                hit_normal[slot] = make_float4(h.n.x, h.n.y, h.n.z, 0.f);
            }
        }
    }
}

__global__ void rc_closest(RayArgs A, const int* ray_world, const float4* p1s, const float4* p2s, int n, uint32_t mask,
                           int* hit_collider, float4* hit_point, float4* hit_normal) {
    RP3DG_GRID_STRIDE(i, n) {
        const float4 a = p1s[i], b = p2s[i];
        const V3 p1 = v3(a.x, a.y, a.z), p2 = v3(b.x, b.y, b.z);
        float maxf = b.w;
        const V3 dir = p2 - p1;
        const V3 inv = v3(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
        const int w = ray_world[i];
        const int c0 = A.wcoll[w], c1 = A.wcoll[w + 1];
        int best = -1;
        Hit hb{0.f, v3(0.f, 0.f, 0.f), v3(0.f, 0.f, 0.f)};
        for (int c = c0; c < c1; c++) {
            Hit h;
            if (!ray_collider(A, c, mask, p1, p2, inv, maxf, h)) continue;
            hb = h;
            best = c;
            // DynamicAABBTree::raycast (DynamicAABBTree.cpp:705-721) with the closest-hit callback
            if (h.t == 0.0f) break;
            if (h.t > 0.0f && h.t < maxf) maxf = h.t;
        }
        hit_collider[i] = best;
        hit_point[i] = make_float4(hb.p.x, hb.p.y, hb.p.z, hb.t);
        hit_normal[i] = make_float4(hb.n.x, hb.n.y, hb.n.z, 0.f);
    }
}

}  // namespace rp3dg
