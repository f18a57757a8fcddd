// Convex hull polygons of a point cloud (the reference's QuickHull, quickhull.cpp).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "rp3d_math.h"

namespace rp3dg {

// Hull of `points` as polygons: vertices xyz (hull vertices numbered in order of first use), vertex count
// of each face, concatenated face vertex indices (CCW seen from outside), in the order
// QuickHull::computeFinalPolygonVertexArray emits them.  False with the reference's message in *err for
// fewer than 4 distinct points or a degenerate (flat or collinear) set.
bool quickhull_polygons(const std::vector<V3>& points, std::vector<float>& verts, std::vector<int32_t>& counts,
                        std::vector<int32_t>& idx, std::string* err);

}  // namespace rp3dg
