// Batched world raycasts (raycast.hip).
#pragma once

#include "rp3d_internal.h"

namespace rp3dg {

struct RayArgs {
    DevColliders C;
    const HullHeader* hulls;
    const float4* hverts;
    const float4* hnormals;
    const int* hfverts;   // [2*f] = start, [2*f+1] = count into hfvidx
    const int* hfvidx;
    const int* wcoll;     // colliders of world w: [wcoll[w], wcoll[w+1])
};

// p2s[i].w = the ray's max fraction; hit_point.w = hit fraction; hit_collider = -1 when nothing is hit
__global__ void rc_closest(RayArgs A, const int* ray_world, const float4* p1s, const float4* p2s, int n, uint32_t mask,
                           int* hit_collider, float4* hit_point, float4* hit_normal);

// every hit: (ray, collider), point (w = fraction) and normal per hit, at most cap written, *count = all
__global__ void rc_all(RayArgs A, const int* ray_world, const float4* p1s, const float4* p2s, int n, uint32_t mask,
                       int2* hit_ray_collider, float4* hit_point, float4* hit_normal, int cap, int* count);

}  // namespace rp3dg
