// Host-side world model behind the C ABI (the part of PhysicsWorld / RigidBody /
// Collider that lives on the CPU: object creation, materials, mass properties,
// hull topology).  Nothing here steps the simulation; rp3dg_world_update()
// uploads the world to the device and the HIP kernels do the step.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/rp3d_mi355x.h"
#include "rp3d_internal.h"

namespace rp3dg {

struct HalfEdgeMesh {
    // HalfEdgeStructure (include/reactphysics3d/collision/HalfEdgeStructure.h:46-72)
    struct Edge { int vertex, twin, face, next; };
    std::vector<std::vector<int>> faces;   // vertex indices per face
    std::vector<int> face_edge;            // Face::edgeIndex
    std::vector<int> vertex_edge;          // Vertex::edgeIndex
    std::vector<Edge> edges;
    bool build(int nb_vertices, std::string* err);
};

struct HostHull {
    std::vector<V3> vertices;
    std::vector<V3> face_normals;
    HalfEdgeMesh mesh;
    V3 centroid;
    V3 local_min, local_max;
    float volume;
};

// TriangleMesh (src/collision/TriangleMesh.cpp): vertices and triangles after the reference's clean-up
// (degenerate triangles dropped, unused vertices removed), vertex normals (given or angle-weighted), and
// the mesh's DynamicAABBTree built by inserting the triangles' AABBs in order (DynamicAABBTree.cpp:146-
// 348, 420-587), node for node.  Leaves: child0 = -1, data = triangle index.
struct HostTriMesh {
    struct Node { V3 mn, mx; int parent, child0, child1, height, data; };
    std::vector<V3> vertices, normals;
    std::vector<int> tris;        // 3 vertex indices per triangle
    std::vector<Node> nodes;
    int root = -1;
    int nb_skipped = 0;           // triangles dropped as degenerate (Message::Type::Warning)
};

// HeightField (src/collision/HeightField.cpp:43-130) with float heights: grid of nb_columns x nb_rows
// samples, height of (x, y) at heights[y * nb_columns + x], centred on the origin with y up.
struct HostHeightField {
    int columns = 0, rows = 0;
    std::vector<float> heights;
    float width = 0.f, length = 0.f, min_h = 0.f, max_h = 0.f, origin = 0.f;
    V3 bmin, bmax;
};

struct HostBody {
    Tf transform;               // body origin + orientation (TransformComponents)
    V3 com_world, com_local;
    V3 v, w;
    V3 force, torque;
    V3 llock{1, 1, 1}, alock{1, 1, 1};
    float ldamp = 0.f, adamp = 0.f;
    float mass = 1.f, inv_mass = 1.f;         // RigidBodyComponents.cpp:225-226
    V3 inertia{1, 1, 1}, inv_inertia{1, 1, 1};
    int type = 2;                              // DYNAMIC
    bool gravity = true;
    bool allowed_to_sleep = true;
    std::vector<int> colliders;
};

struct HostCollider {
    int body;
    int type;
    float params[4];
    Tf local;
    float friction_sqrt, bounciness, density;
    uint16_t category = 0x0001, mask = 0xFFFF;
    int kind = CK_SIMULATION | CK_WORLD_QUERY;   // ColliderComponents defaults (ColliderComponents.h)
};

struct HostJoint {
    int type;
    int body1, body2;
    V3 anchor_world, axis_world;
    bool collide_connected;
    int flags;
    float limits[4];
};

struct HostWorld {
    rp3dg_world_settings cfg;
    std::vector<HostBody> bodies;
    std::vector<HostCollider> colliders;
    std::vector<HostHull> hulls;
    std::vector<HostTriMesh> meshes;
    std::vector<HostHeightField> heightfields;
    std::vector<HostJoint> joints;
    rp3dg_batch* own = nullptr;   // device batch of one, created on first update()
    bool events = false;          // record contact events in the own batch
    bool split_impulse = true;    // ContactsPositionCorrectionTechnique::SPLIT_IMPULSES (the default)
    bool uploaded = false;
    // host mirror of the device body state, refreshed lazily after update()
    std::vector<float> mirror_pos, mirror_q, mirror_v, mirror_w;
    std::vector<int32_t> mirror_sleeping;
    bool mirror_dirty = false;
};

// Shape formulas (SphereShape.h, CapsuleShape.{h,cpp}, BoxShape.{h,cpp}, ConvexMeshShape.cpp)
float shape_volume(const HostWorld& w, const HostCollider& c);
V3 shape_local_inertia(const HostWorld& w, const HostCollider& c, float mass);

void set_error(const std::string& msg);

// Concave shapes (concave.cpp): triangles overlapping an AABB given in the shape's local space, in the
// order of ConcaveShape::computeOverlappingTriangles (ConcaveMeshShape.cpp:148-176, HeightField.cpp:
// 133-180): triangle ids (mesh: triangle index; height field: HeightField::computeTriangleShapeId),
// vertices and vertex normals in the shape's scaled local space.
void concave_overlapping_triangles(const HostWorld& w, const HostCollider& c, V3 aabb_min, V3 aabb_max,
                                   std::vector<int>& ids, std::vector<V3>& vertices, std::vector<V3>& normals);
// ConcaveShape local bounds (ConcaveMeshShape.cpp:246-250, HeightFieldShape.cpp:40-44), scale applied
void concave_local_bounds(const HostWorld& w, const HostCollider& c, V3& lo, V3& hi);
// ConcaveMeshShape::computeScaledVerticesNormals (ConcaveMeshShape.cpp:53-70)
std::vector<V3> concave_scaled_normals(const HostTriMesh& m, V3 scale);

}  // namespace rp3dg

struct rp3dg_world : rp3dg::HostWorld {};
