// Concave collision shapes on the host: TriangleMesh (vertex clean-up, vertex normals, the mesh's AABB
// tree), HeightField, and the C ABI that creates them.  The device reads the arenas uploaded from here
// (batch.hip) and walks the same tree in the same order in its middle phase (concave.hip).
#include <algorithm>
#include <cmath>
#include <limits>
#include <string>

#include "world.h"

namespace rp3dg {

namespace {

// AABB helpers of the tree (AABB.h:179-243)
inline float volume(V3 mn, V3 mx) {
    const V3 d = mx - mn;
    return d.x * d.y * d.z;
}
inline void merge(V3 amn, V3 amx, V3 bmn, V3 bmx, V3& mn, V3& mx) {
    mn = v3(std::min(amn.x, bmn.x), std::min(amn.y, bmn.y), std::min(amn.z, bmn.z));
    mx = v3(std::max(amx.x, bmx.x), std::max(amx.y, bmx.y), std::max(amx.z, bmx.z));
}
inline bool overlap(V3 amn, V3 amx, V3 bmn, V3 bmx) {
    if (amx.x < bmn.x || bmx.x < amn.x) return false;
    if (amx.y < bmn.y || bmx.y < amn.y) return false;
    if (amx.z < bmn.z || bmx.z < amn.z) return false;
    return true;
}

// DynamicAABBTree::balanceSubTreeAtNode (DynamicAABBTree.cpp:420-587): one AVL-style rotation
int balance(HostTriMesh& m, int a) {
    auto& N = m.nodes;
    if (N[a].child0 < 0 || N[a].height < 2) return a;
    const int b = N[a].child0, c = N[a].child1;
    const int bal = N[c].height - N[b].height;
    auto relink = [&](int up, int old_child) {
        N[up].parent = N[old_child].parent;
        N[old_child].parent = up;
        const int p = N[up].parent;
        if (p >= 0) {
            if (N[p].child0 == old_child) N[p].child0 = up;
            else N[p].child1 = up;
        } else {
            m.root = up;
        }
    };
    if (bal > 1) {   // C is higher: C goes up
        const int f = N[c].child0, g = N[c].child1;
        N[c].child0 = a;
        relink(c, a);
        if (N[f].height > N[g].height) {
            N[c].child1 = f;
            N[a].child1 = g;
            N[g].parent = a;
            merge(N[b].mn, N[b].mx, N[g].mn, N[g].mx, N[a].mn, N[a].mx);
            merge(N[a].mn, N[a].mx, N[f].mn, N[f].mx, N[c].mn, N[c].mx);
            N[a].height = std::max(N[b].height, N[g].height) + 1;
            N[c].height = std::max(N[a].height, N[f].height) + 1;
        } else {
            N[c].child1 = g;
            N[a].child1 = f;
            N[f].parent = a;
            merge(N[b].mn, N[b].mx, N[f].mn, N[f].mx, N[a].mn, N[a].mx);
            merge(N[a].mn, N[a].mx, N[g].mn, N[g].mx, N[c].mn, N[c].mx);
            N[a].height = std::max(N[b].height, N[f].height) + 1;
            N[c].height = std::max(N[a].height, N[g].height) + 1;
        }
        return c;
    }
    if (bal < -1) {   // B is higher: B goes up
        const int f = N[b].child0, g = N[b].child1;
        N[b].child0 = a;
        relink(b, a);
        if (N[f].height > N[g].height) {
            N[b].child1 = f;
            N[a].child0 = g;
            N[g].parent = a;
            merge(N[c].mn, N[c].mx, N[g].mn, N[g].mx, N[a].mn, N[a].mx);
            merge(N[a].mn, N[a].mx, N[f].mn, N[f].mx, N[b].mn, N[b].mx);
            N[a].height = std::max(N[c].height, N[g].height) + 1;
            N[b].height = std::max(N[a].height, N[f].height) + 1;
        } else {
            N[b].child1 = g;
            N[a].child0 = f;
            N[f].parent = a;
            merge(N[c].mn, N[c].mx, N[f].mn, N[f].mx, N[a].mn, N[a].mx);
            merge(N[a].mn, N[a].mx, N[g].mn, N[g].mx, N[b].mn, N[b].mx);
            N[a].height = std::max(N[c].height, N[f].height) + 1;
            N[b].height = std::max(N[a].height, N[g].height) + 1;
        }
        return b;
    }
    return a;
}

// DynamicAABBTree::addObject + insertLeafNode (DynamicAABBTree.cpp:146-166, 219-346) with a zero fat
// margin; nodes are numbered in allocation order, as the reference's free list hands them out when
// nothing is ever released
void tree_insert(HostTriMesh& m, V3 mn, V3 mx, int data) {
    auto& N = m.nodes;
    const int leaf = (int)N.size();
    N.push_back({mn, mx, -1, -1, -1, 0, data});
    if (m.root < 0) {
        m.root = leaf;
        return;
    }
    int cur = m.root;
    while (N[cur].child0 >= 0) {
        const int l = N[cur].child0, r = N[cur].child1;
        const float vol = volume(N[cur].mn, N[cur].mx);
        V3 a, b;
        merge(N[cur].mn, N[cur].mx, mn, mx, a, b);
        const float merged = volume(a, b);
        const float costS = 2.0f * merged;
        const float costI = 2.0f * (merged - vol);
        merge(mn, mx, N[l].mn, N[l].mx, a, b);
        const float costL = N[l].child0 < 0 ? volume(a, b) + costI : costI + volume(a, b) - volume(N[l].mn, N[l].mx);
        merge(mn, mx, N[r].mn, N[r].mx, a, b);
        const float costR = N[r].child0 < 0 ? volume(a, b) + costI : costI + volume(a, b) - volume(N[r].mn, N[r].mx);
        if (costS < costL && costS < costR) break;
        cur = costL < costR ? l : r;
    }
    const int sib = cur, oldp = N[sib].parent;
    const int np = (int)N.size();
    N.push_back({v3zero(), v3zero(), oldp, sib, leaf, N[sib].height + 1, -1});
    merge(N[sib].mn, N[sib].mx, mn, mx, N[np].mn, N[np].mx);
    if (oldp >= 0) {
        if (N[oldp].child0 == sib) N[oldp].child0 = np;
        else N[oldp].child1 = np;
    } else {
        m.root = np;
    }
    N[sib].parent = np;
    N[leaf].parent = np;
    cur = N[leaf].parent;
    while (cur >= 0) {
        cur = balance(m, cur);
        const int l = N[cur].child0, r = N[cur].child1;
        N[cur].height = std::max(N[l].height, N[r].height) + 1;
        merge(N[l].mn, N[l].mx, N[r].mn, N[r].mx, N[cur].mn, N[cur].mx);
        cur = N[cur].parent;
    }
}

// TriangleMesh::computeVerticesNormals (TriangleMesh.cpp:268-318): angle-weighted face normals
void compute_vertex_normals(HostTriMesh& m) {
    const int nt = (int)m.tris.size() / 3;
    for (int f = 0; f < nt; f++) {
        V3 t[3];
        for (int k = 0; k < 3; k++) t[k] = m.vertices[m.tris[3 * f + k]];
        float el[3];
        el[0] = len(t[1] - t[0]);
        el[1] = len(t[2] - t[1]);
        el[2] = len(t[0] - t[2]);
        for (int v = 0; v < 3; v++) {
            const int prev = v == 0 ? 2 : v - 1, next = v == 2 ? 0 : v + 1;
            const V3 a = t[next] - t[v];
            const V3 b = t[prev] - t[v];
            const float dp = dot(a, b);
            const float lab = el[prev] * el[v];
            float cosA = dp / lab;
            cosA = std::min(std::max(cosA, 0.0f), 1.0f);
            const float angle = std::acos(cosA);
            const V3 n = normalized(cross(a, b));
            m.normals[m.tris[3 * f + v]] += angle * n;
        }
    }
    for (V3& n : m.normals) n = normalized(n);
}

}  // namespace

std::vector<V3> concave_scaled_normals(const HostTriMesh& m, V3 scale) {
    std::vector<V3> out;
    out.reserve(m.normals.size());
    // Vector3(1.0 / mScale.x, ...): a double quotient rounded to float
    const V3 inv = v3((float)(1.0 / (double)scale.x), (float)(1.0 / (double)scale.y), (float)(1.0 / (double)scale.z));
    for (V3 n : m.normals) {
        n = cmul(inv, n);
        const float l = len(n);
        out.push_back(v3(n.x / l, n.y / l, n.z / l));
    }
    return out;
}

void concave_local_bounds(const HostWorld& w, const HostCollider& c, V3& lo, V3& hi) {
    const V3 s = v3(c.params[1], c.params[2], c.params[3]);
    if (c.type == SHAPE_CONCAVE_MESH) {
        const HostTriMesh& m = w.meshes[(int)c.params[0]];
        lo = cmul(m.nodes[m.root].mn, s);
        hi = cmul(m.nodes[m.root].mx, s);
    } else {
        const HostHeightField& h = w.heightfields[(int)c.params[0]];
        lo = cmul(h.bmin, s);
        hi = cmul(h.bmax, s);
    }
}

// HeightField::getVertexAt (HeightField.cpp:399-409)
static V3 hf_vertex(const HostHeightField& h, int x, int y) {
    const float height = h.heights[(size_t)y * h.columns + x];
    return v3(-h.width * 0.5f + (float)(uint32_t)x, h.origin + height, -h.length * 0.5f + (float)(uint32_t)y);
}

void concave_overlapping_triangles(const HostWorld& w, const HostCollider& c, V3 amin, V3 amax, std::vector<int>& ids,
                                   std::vector<V3>& vertices, std::vector<V3>& normals) {
    const V3 s = v3(c.params[1], c.params[2], c.params[3]);
    const V3 inv = v3(1.0f / s.x, 1.0f / s.y, 1.0f / s.z);
    if (c.type == SHAPE_CONCAVE_MESH) {
        const HostTriMesh& m = w.meshes[(int)c.params[0]];
        const std::vector<V3> sn = concave_scaled_normals(m, s);
        const V3 mn = cmul(amin, inv), mx = cmul(amax, inv);   // AABB::applyScale
        std::vector<int> stack{m.root};
        while (!stack.empty()) {
            const int id = stack.back();
            stack.pop_back();
            if (id < 0) continue;
            const HostTriMesh::Node& n = m.nodes[id];
            if (!overlap(mn, mx, n.mn, n.mx)) continue;
            if (n.child0 < 0) {
                const int t = n.data;
                ids.push_back(t);
                for (int k = 0; k < 3; k++) {
                    vertices.push_back(cmul(m.vertices[m.tris[3 * t + k]], s));
                    normals.push_back(sn[m.tris[3 * t + k]]);
                }
            } else {
                stack.push_back(n.child0);
                stack.push_back(n.child1);
            }
        }
        return;
    }
    const HostHeightField& h = w.heightfields[(int)c.params[0]];
    const V3 mn = cmul(amin, inv), mx = cmul(amax, inv);
    // HeightField::computeMinMaxGridCoordinates (HeightField.cpp:183-216)
    V3 lo = vmin(vmax(mn, h.bmin), h.bmax);
    V3 hi = vmax(vmin(mx, h.bmax), h.bmin);
    const V3 tr = (h.bmax - h.bmin) * 0.5f;
    lo += tr;
    hi += tr;
    const int minc[3] = {(int)(lo.x + 0.5) - 1, (int)(lo.y + 0.5) - 1, (int)(lo.z + 0.5) - 1};
    const int maxc[3] = {(int)(hi.x + 0.5) + 1, (int)(hi.y + 0.5) + 1, (int)(hi.z + 0.5) + 1};
    // clamp(int value, int lower, int upper) (mathematics_functions.h:48-51)
    auto clampi = [](int v, int hi_) { return (uint32_t)std::min(std::max(v, 0), hi_); };
    const uint32_t iMin = clampi(minc[0], h.columns - 1), iMax = clampi(maxc[0], h.columns - 1);
    const uint32_t jMin = clampi(minc[2], h.rows - 1), jMax = clampi(maxc[2], h.rows - 1);
    for (uint32_t i = iMin; i < iMax; i++) {
        for (uint32_t j = jMin; j < jMax; j++) {
            const V3 p1 = cmul(hf_vertex(h, i, j), s), p2 = cmul(hf_vertex(h, i, j + 1), s);
            const V3 p3 = cmul(hf_vertex(h, i + 1, j), s), p4 = cmul(hf_vertex(h, i + 1, j + 1), s);
            const V3 n1 = unit(cross(p2 - p1, p3 - p1));
            ids.push_back((int)((j * (h.columns - 1) + i) * 2));
            vertices.push_back(p1); vertices.push_back(p2); vertices.push_back(p3);
            normals.push_back(n1); normals.push_back(n1); normals.push_back(n1);
            const V3 n2 = unit(cross(p2 - p3, p4 - p3));
            ids.push_back((int)((j * (h.columns - 1) + i) * 2 + 1));
            vertices.push_back(p3); vertices.push_back(p2); vertices.push_back(p4);
            normals.push_back(n2); normals.push_back(n2); normals.push_back(n2);
        }
    }
}

}  // namespace rp3dg

using namespace rp3dg;

extern "C" {

// PhysicsCommon::createTriangleMesh(TriangleVertexArray, messages) (TriangleMesh.cpp:39-200)
int32_t rp3dg_triangle_mesh_create(rp3dg_world* w, const float* vertices, int32_t nv, const int32_t* indices,
                                   int32_t nt, const float* normals, int32_t* nb_skipped) {
    if (!w || !vertices || !indices || nv < 3 || nt < 1) { set_error("invalid triangle mesh arguments"); return RP3DG_EINVAL; }
    if (w->uploaded) { set_error("world already uploaded to the device; edit before the first update()"); return RP3DG_ESTATE; }
    HostTriMesh m;
    auto vtx = [&](int i) { return v3(vertices[3 * i], vertices[3 * i + 1], vertices[3 * i + 2]); };
    // TriangleMesh::computeEpsilon (:72-97)
    V3 mxabs = v3zero();
    for (int i = 0; i < nv; i++) {
        const V3 p = vtx(i);
        if (std::fabs(p.x) > mxabs.x) mxabs.x = std::fabs(p.x);
        if (std::fabs(p.y) > mxabs.y) mxabs.y = std::fabs(p.y);
        if (std::fabs(p.z) > mxabs.z) mxabs.z = std::fabs(p.z);
    }
    const float eps = 3 * (mxabs.x + mxabs.y + mxabs.z) * std::numeric_limits<float>::epsilon();
    const float eps2 = eps * eps;
    // TriangleMesh::copyData (:100-200)
    std::vector<bool> used(nv, false);
    m.vertices.assign(nv, v3zero());
    m.normals.assign(nv, v3zero());
    for (int t = 0; t < nt; t++) {
        const int i0 = indices[3 * t], i1 = indices[3 * t + 1], i2 = indices[3 * t + 2];
        if (i0 < 0 || i1 < 0 || i2 < 0 || i0 >= nv || i1 >= nv || i2 >= nv) {
            set_error("the face with index " + std::to_string(t) + " has a vertex index out of range");
            return RP3DG_EINVAL;
        }
        const V3 v1 = vtx(i0), v2 = vtx(i1), v3_ = vtx(i2);
        const bool zero_area = len2(cross(v3_ - v1, v2 - v1)) < eps2;
        const bool zero_edge = len2(v2 - v1) < eps2 || len2(v3_ - v2) < eps2 || len2(v1 - v3_) < eps2;
        if (zero_area || zero_edge) {
            m.nb_skipped++;
            continue;
        }
        V3 vn[3] = {v3zero(), v3zero(), v3zero()};
        if (normals) {
            const int ix[3] = {i0, i1, i2};
            for (int k = 0; k < 3; k++) {
                vn[k] = v3(normals[3 * ix[k]], normals[3 * ix[k] + 1], normals[3 * ix[k] + 2]);
                if (len2(vn[k]) < eps2) {
                    set_error("the length of the provided normal for vertex " + std::to_string(ix[k]) + " is too small");
                    return RP3DG_EINVAL;
                }
            }
        }
        m.normals[i0] = vn[0]; m.vertices[i0] = v1;
        m.normals[i1] = vn[1]; m.vertices[i1] = v2;
        m.normals[i2] = vn[2]; m.vertices[i2] = v3_;
        m.tris.push_back(i0); m.tris.push_back(i1); m.tris.push_back(i2);
        used[i0] = used[i1] = used[i2] = true;
    }
    // TriangleMesh::removeUnusedVertices (:203-228): vertex 0 always stays
    for (int i = nv - 1; i > 0; i--) {
        if (used[i]) continue;
        m.vertices.erase(m.vertices.begin() + i);
        m.normals.erase(m.normals.begin() + i);
        for (int& x : m.tris)
            if (x > i) x--;
    }
    if (m.tris.empty()) { set_error("the mesh does not have any valid triangle faces"); return RP3DG_EINVAL; }
    if (!normals) compute_vertex_normals(m);
    // TriangleMesh::initBVHTree (:231-249)
    const int ntv = (int)m.tris.size() / 3;
    for (int f = 0; f < ntv; f++) {
        const V3 p0 = m.vertices[m.tris[3 * f]], p1 = m.vertices[m.tris[3 * f + 1]], p2 = m.vertices[m.tris[3 * f + 2]];
        // AABB::createAABBForTriangle (AABB.h:260-283)
        V3 mn = p0, mx = p0;
        if (p1.x < mn.x) mn.x = p1.x;
        if (p1.y < mn.y) mn.y = p1.y;
        if (p1.z < mn.z) mn.z = p1.z;
        if (p2.x < mn.x) mn.x = p2.x;
        if (p2.y < mn.y) mn.y = p2.y;
        if (p2.z < mn.z) mn.z = p2.z;
        if (p1.x > mx.x) mx.x = p1.x;
        if (p1.y > mx.y) mx.y = p1.y;
        if (p1.z > mx.z) mx.z = p1.z;
        if (p2.x > mx.x) mx.x = p2.x;
        if (p2.y > mx.y) mx.y = p2.y;
        if (p2.z > mx.z) mx.z = p2.z;
        tree_insert(m, mn, mx, f);
    }
    if (nb_skipped) *nb_skipped = m.nb_skipped;
    w->meshes.push_back(std::move(m));
    return (int32_t)w->meshes.size() - 1;
}

// PhysicsCommon::createHeightField with HeightDataType::HEIGHT_FLOAT_TYPE (HeightField.cpp:43-130)
int32_t rp3dg_heightfield_create(rp3dg_world* w, int32_t columns, int32_t rows, const float* heights) {
    if (!w || !heights) { set_error("invalid height field arguments"); return RP3DG_EINVAL; }
    if (w->uploaded) { set_error("world already uploaded to the device; edit before the first update()"); return RP3DG_ESTATE; }
    if (columns < 2 || rows < 2) { set_error("the number of grid columns and grid rows must be at least two"); return RP3DG_EINVAL; }
    HostHeightField h;
    h.columns = columns;
    h.rows = rows;
    h.width = (float)(columns - 1);
    h.length = (float)(rows - 1);
    h.heights.assign(heights, heights + (size_t)columns * rows);
    for (int x = 0; x < columns; x++) {
        for (int y = 0; y < rows; y++) {
            const float v = h.heights[(size_t)y * columns + x];
            if (!std::isfinite(v)) { set_error("height field heights must be finite"); return RP3DG_EINVAL; }
            if (x == 0 && y == 0) h.min_h = h.max_h = v;
            if (v < h.min_h) h.min_h = v;
            if (v > h.max_h) h.max_h = v;
        }
    }
    h.origin = -(h.max_h - h.min_h) * 0.5f - h.min_h;
    const float half = (h.max_h - h.min_h) * 0.5f;
    h.bmin = v3(-h.width * 0.5f, -half, -h.length * 0.5f);
    h.bmax = v3(h.width * 0.5f, half, h.length * 0.5f);
    w->heightfields.push_back(std::move(h));
    return (int32_t)w->heightfields.size() - 1;
}

int32_t rp3dg_concave_overlapping_triangles(rp3dg_world* w, int32_t collider, const float aabb_min[3],
                                            const float aabb_max[3], int32_t* ids, float* vertices, float* normals,
                                            int32_t capacity) {
    if (!w || collider < 0 || collider >= (int32_t)w->colliders.size() || !aabb_min || !aabb_max) {
        set_error("invalid collider");
        return RP3DG_EINVAL;
    }
    const HostCollider& c = w->colliders[collider];
    if (c.type != SHAPE_CONCAVE_MESH && c.type != SHAPE_HEIGHTFIELD) { set_error("not a concave collider"); return RP3DG_EINVAL; }
    std::vector<int> id;
    std::vector<V3> vs, ns;
    concave_overlapping_triangles(*w, c, v3(aabb_min[0], aabb_min[1], aabb_min[2]), v3(aabb_max[0], aabb_max[1], aabb_max[2]),
                                  id, vs, ns);
    for (int i = 0; i < (int)id.size() && i < capacity; i++) {
        if (ids) ids[i] = id[i];
        for (int k = 0; k < 3; k++) {
            const V3 p = vs[3 * i + k], n = ns[3 * i + k];
            if (vertices) { vertices[9 * i + 3 * k] = p.x; vertices[9 * i + 3 * k + 1] = p.y; vertices[9 * i + 3 * k + 2] = p.z; }
            if (normals) { normals[9 * i + 3 * k] = n.x; normals[9 * i + 3 * k + 1] = n.y; normals[9 * i + 3 * k + 2] = n.z; }
        }
    }
    return (int32_t)id.size();
}

}  // extern "C"
