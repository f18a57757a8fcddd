// C++ mirror of the ReactPhysics3D 0.10.2 public API for the step path, over the rp3d_mi355x C-ABI.
//
// An application written against the reference (include/reactphysics3d/engine/PhysicsCommon.h,
// PhysicsWorld.h, body/RigidBody.h, collision/Collider.h, collision/shapes/*.h,
// constraint/*Joint.h) includes this header instead and keeps its code: the same class names, the
// same factory calls and the same PhysicsWorld::update(dt).  Objects are thin handles (world +
// index) into the host world behind the C-ABI; getters after update() read a host mirror that the
// library refreshes from the device once per step.  Only what the step path needs is provided:
// sphere / box / capsule / convex-mesh colliders, ball-and-socket / hinge / slider / fixed joints with
// their limits and motors, materials, collision filters, triggers, velocities, damping, gravity,
// sleeping, and the RigidBody calls between steps (setTransform, velocities, forces and torques at the
// centre of mass or at a point, resets, setIsSleeping).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <utility>
#include <stdexcept>
#include <string>
#include <vector>

#include "rp3d_mi355x.h"

namespace reactphysics3d {

using decimal = float;
using uint32 = uint32_t;
using uint64 = uint64_t;
using uint16 = uint16_t;

inline void check(int status, const char* what) {
    if (status < 0) throw std::runtime_error(std::string(what) + ": " + rp3dg_last_error());
}

struct Vector3 {
    decimal x = 0, y = 0, z = 0;
    Vector3() = default;
    Vector3(decimal a, decimal b, decimal c) : x(a), y(b), z(c) {}
    static Vector3 zero() { return Vector3(0, 0, 0); }
    Vector3 operator+(const Vector3& o) const { return Vector3(x + o.x, y + o.y, z + o.z); }
    Vector3 operator-(const Vector3& o) const { return Vector3(x - o.x, y - o.y, z - o.z); }
    Vector3 operator*(decimal s) const { return Vector3(x * s, y * s, z * s); }
    decimal length() const { return std::sqrt(x * x + y * y + z * z); }
};

struct Quaternion {
    decimal x = 0, y = 0, z = 0, w = 1;
    Quaternion() = default;
    Quaternion(decimal a, decimal b, decimal c, decimal d) : x(a), y(b), z(c), w(d) {}
    static Quaternion identity() { return Quaternion(0, 0, 0, 1); }
};

class Transform {
   public:
    Transform() = default;
    Transform(const Vector3& p, const Quaternion& q) : mPosition(p), mOrientation(q) {}
    static Transform identity() { return Transform(); }
    const Vector3& getPosition() const { return mPosition; }
    const Quaternion& getOrientation() const { return mOrientation; }
    void setPosition(const Vector3& p) { mPosition = p; }
    void setOrientation(const Quaternion& q) { mOrientation = q; }

   private:
    Vector3 mPosition;
    Quaternion mOrientation;
};

enum class BodyType { STATIC, KINEMATIC, DYNAMIC };
// configuration.h:95-109
enum class JointsPositionCorrectionTechnique { BAUMGARTE_JOINTS, NON_LINEAR_GAUSS_SEIDEL };
enum class ContactsPositionCorrectionTechnique { BAUMGARTE_CONTACTS, SPLIT_IMPULSES };

// ---- shapes (collision/shapes/*.h): parameter holders, one collider each ----
class CollisionShape {
   public:
    virtual ~CollisionShape() = default;
    int32_t type() const { return mType; }
    const float* params() const { return mParams; }

   protected:
    int32_t mType = RP3DG_SHAPE_SPHERE;
    float mParams[4] = {0, 0, 0, 0};
};
class SphereShape : public CollisionShape {
   public:
    explicit SphereShape(decimal radius) { mType = RP3DG_SHAPE_SPHERE; mParams[0] = radius; }
    decimal getRadius() const { return mParams[0]; }
};
class CapsuleShape : public CollisionShape {
   public:
    CapsuleShape(decimal radius, decimal height) { mType = RP3DG_SHAPE_CAPSULE; mParams[0] = radius; mParams[1] = height; }
};
class BoxShape : public CollisionShape {
   public:
    explicit BoxShape(const Vector3& halfExtents) {
        mType = RP3DG_SHAPE_BOX;
        mParams[0] = halfExtents.x; mParams[1] = halfExtents.y; mParams[2] = halfExtents.z;
    }
};

// Message (utils/Message.h): what createConvexMesh reports when it refuses a mesh.
struct Message {
    enum class Type { Error = 1, Warning = 2, Information = 4 };
    std::string text;
    Type type;
    Message(std::string t, Type ty = Type::Error) : text(std::move(t)), type(ty) {}
};

// PolygonVertexArray (collision/PolygonVertexArray.h:45-140): strided float or double vertices,
// strided int or short indices, faces as (nbVertices, indexBase).  The data is copied at construction.
class PolygonVertexArray {
   public:
    enum class VertexDataType { VERTEX_FLOAT_TYPE, VERTEX_DOUBLE_TYPE };
    enum class IndexDataType { INDEX_INTEGER_TYPE, INDEX_SHORT_TYPE };
    struct PolygonFace { uint32 nbVertices; uint32 indexBase; };
    PolygonVertexArray() = default;
    PolygonVertexArray(uint32 nbVertices, const void* verticesStart, uint32 verticesStride, const void* indexesStart,
                       uint32 indexesStride, uint32 nbFaces, PolygonFace* facesStart, VertexDataType vertexDataType,
                       IndexDataType indexDataType) {
        init(nbVertices, verticesStart, verticesStride, indexesStart, indexesStride, nbFaces, facesStart,
             vertexDataType, indexDataType);
    }
    void init(uint32 nbVertices, const void* verticesStart, uint32 verticesStride, const void* indexesStart,
              uint32 indexesStride, uint32 nbFaces, PolygonFace* facesStart, VertexDataType vertexDataType,
              IndexDataType indexDataType) {
        mVertexType = vertexDataType;
        mIndexType = indexDataType;
        mVertices.clear();
        mCounts.clear();
        mIndices.clear();
        mFaces.assign(facesStart, facesStart + nbFaces);
        const unsigned char* v = static_cast<const unsigned char*>(verticesStart);
        for (uint32 i = 0; i < nbVertices; i++)
            for (int k = 0; k < 3; k++)
                mVertices.push_back(vertexDataType == VertexDataType::VERTEX_FLOAT_TYPE
                                        ? reinterpret_cast<const float*>(v + i * verticesStride)[k]
                                        : (float)reinterpret_cast<const double*>(v + i * verticesStride)[k]);
        const unsigned char* ix = static_cast<const unsigned char*>(indexesStart);
        for (uint32 f = 0; f < nbFaces; f++) {
            mCounts.push_back((int32_t)facesStart[f].nbVertices);
            for (uint32 k = 0; k < facesStart[f].nbVertices; k++) {
                const unsigned char* at = ix + (facesStart[f].indexBase + k) * indexesStride;
                mIndices.push_back(indexDataType == IndexDataType::INDEX_INTEGER_TYPE
                                       ? (int32_t) * reinterpret_cast<const uint32*>(at)
                                       : (int32_t) * reinterpret_cast<const uint16*>(at));
            }
        }
    }
    VertexDataType getVertexDataType() const { return mVertexType; }
    IndexDataType getIndexDataType() const { return mIndexType; }
    uint32 getNbVertices() const { return (uint32)(mVertices.size() / 3); }
    uint32 getNbFaces() const { return (uint32)mCounts.size(); }
    Vector3 getVertex(uint32 i) const { return Vector3(mVertices[3 * i], mVertices[3 * i + 1], mVertices[3 * i + 2]); }
    uint32 getVertexIndexInFace(uint32 face, uint32 k) const { return (uint32)mIndices[mFaces[face].indexBase + k]; }
    const PolygonFace* getPolygonFace(uint32 face) const { return &mFaces[face]; }

    std::vector<float> mVertices;            // xyz per vertex
    std::vector<int32_t> mCounts, mIndices;  // per face: vertex count; concatenated face vertex indices
    std::vector<PolygonFace> mFaces;

   private:
    VertexDataType mVertexType = VertexDataType::VERTEX_FLOAT_TYPE;
    IndexDataType mIndexType = IndexDataType::INDEX_INTEGER_TYPE;
};

// VertexArray (collision/VertexArray.h:40-90): a point cloud for createConvexMesh's QuickHull.
class VertexArray {
   public:
    enum class DataType { VERTEX_FLOAT_TYPE, VERTEX_DOUBLE_TYPE };
    VertexArray(const void* start, uint32 stride, uint32 nbVertices, DataType dataType)
        : mStart(start), mStride(stride), mNbVertices(nbVertices), mDataType(dataType) {}
    DataType getDataType() const { return mDataType; }
    uint32 getNbVertices() const { return mNbVertices; }
    uint32 getStride() const { return mStride; }
    const void* getStart() const { return mStart; }

   private:
    const void* mStart;
    uint32 mStride, mNbVertices;
    DataType mDataType;
};

// ConvexMesh (collision/ConvexMesh.h): the polygons of a hull; each collider that uses it uploads it
// into its world.
class ConvexMesh {
   public:
    explicit ConvexMesh(PolygonVertexArray a) : mArray(std::move(a)) {}
    uint32 getNbVertices() const { return mArray.getNbVertices(); }
    uint32 getNbFaces() const { return mArray.getNbFaces(); }
    const Vector3 getVertex(uint32 i) const { return mArray.getVertex(i); }
    PolygonVertexArray mArray;
};
class ConvexMeshShape : public CollisionShape {
   public:
    explicit ConvexMeshShape(ConvexMesh* mesh) : mMesh(mesh) { mType = RP3DG_SHAPE_CONVEX_MESH; }
    ConvexMesh* mesh() const { return mMesh; }

   private:
    ConvexMesh* mMesh;
};

// TriangleVertexArray (collision/TriangleVertexArray.h:45-200): strided float or double vertices (and
// optional vertex normals) and int or short triangle indices, copied as floats and int32.
class TriangleVertexArray {
   public:
    enum class VertexDataType { VERTEX_FLOAT_TYPE, VERTEX_DOUBLE_TYPE };
    enum class NormalDataType { NORMAL_FLOAT_TYPE, NORMAL_DOUBLE_TYPE };
    enum class IndexDataType { INDEX_INTEGER_TYPE, INDEX_SHORT_TYPE };
    TriangleVertexArray(uint32 nbVertices, const void* verticesStart, uint32 verticesStride, uint32 nbTriangles,
                        const void* indexesStart, uint32 indexesStride, VertexDataType vertexDataType,
                        IndexDataType indexDataType) {
        copy_vertices(mVertices, nbVertices, verticesStart, verticesStride, vertexDataType == VertexDataType::VERTEX_DOUBLE_TYPE);
        copy_indices(nbTriangles, indexesStart, indexesStride, indexDataType == IndexDataType::INDEX_SHORT_TYPE);
    }
    TriangleVertexArray(uint32 nbVertices, const void* verticesStart, uint32 verticesStride, const void* normalsStart,
                        uint32 normalsStride, uint32 nbTriangles, const void* indexesStart, uint32 indexesStride,
                        VertexDataType vertexDataType, NormalDataType normalDataType, IndexDataType indexDataType) {
        copy_vertices(mVertices, nbVertices, verticesStart, verticesStride, vertexDataType == VertexDataType::VERTEX_DOUBLE_TYPE);
        copy_vertices(mNormals, nbVertices, normalsStart, normalsStride, normalDataType == NormalDataType::NORMAL_DOUBLE_TYPE);
        copy_indices(nbTriangles, indexesStart, indexesStride, indexDataType == IndexDataType::INDEX_SHORT_TYPE);
    }
    uint32 getNbVertices() const { return (uint32)(mVertices.size() / 3); }
    uint32 getNbTriangles() const { return (uint32)(mIndices.size() / 3); }
    bool getHasNormals() const { return !mNormals.empty(); }
    std::vector<float> mVertices, mNormals;
    std::vector<int32_t> mIndices;

   private:
    static void copy_vertices(std::vector<float>& out, uint32 n, const void* start, uint32 stride, bool dbl) {
        out.resize(3 * (size_t)n);
        for (uint32 i = 0; i < n; i++) {
            const unsigned char* p = static_cast<const unsigned char*>(start) + (size_t)i * stride;
            for (int k = 0; k < 3; k++)
                out[3 * i + k] = dbl ? (float)reinterpret_cast<const double*>(p)[k] : reinterpret_cast<const float*>(p)[k];
        }
    }
    void copy_indices(uint32 n, const void* start, uint32 stride, bool shrt) {
        mIndices.resize(3 * (size_t)n);
        for (uint32 i = 0; i < n; i++) {
            const unsigned char* p = static_cast<const unsigned char*>(start) + (size_t)i * stride;
            for (int k = 0; k < 3; k++)
                mIndices[3 * i + k] = shrt ? (int32_t)reinterpret_cast<const unsigned short*>(p)[k]
                                           : (int32_t)reinterpret_cast<const uint32*>(p)[k];
        }
    }
};
// TriangleMesh (collision/TriangleMesh.h); each collider that uses it builds it in its world.
class TriangleMesh {
   public:
    explicit TriangleMesh(TriangleVertexArray a) : mArray(std::move(a)) {}
    uint32 getNbVertices() const { return mArray.getNbVertices(); }
    uint32 getNbTriangles() const { return mArray.getNbTriangles(); }
    TriangleVertexArray mArray;
};
class ConcaveMeshShape : public CollisionShape {
   public:
    ConcaveMeshShape(TriangleMesh* mesh, const Vector3& scaling) : mMesh(mesh), mScale(scaling) {
        mType = RP3DG_SHAPE_CONCAVE_MESH;
        mParams[1] = scaling.x; mParams[2] = scaling.y; mParams[3] = scaling.z;
    }
    TriangleMesh* mesh() const { return mMesh; }
    const Vector3& getScale() const { return mScale; }

   private:
    TriangleMesh* mMesh;
    Vector3 mScale;
};
// HeightField (collision/HeightField.h): float, double or integer samples (HeightField.cpp:72-102), kept as
// floats; heights[y * columns + x].
class HeightField {
   public:
    enum class HeightDataType { HEIGHT_FLOAT_TYPE, HEIGHT_DOUBLE_TYPE, HEIGHT_INT_TYPE };
    int getNbColumns() const { return mColumns; }
    int getNbRows() const { return mRows; }
    int mColumns = 0, mRows = 0;
    std::vector<float> mHeights;
};
class HeightFieldShape : public CollisionShape {
   public:
    HeightFieldShape(HeightField* field, const Vector3& scaling) : mField(field), mScale(scaling) {
        mType = RP3DG_SHAPE_HEIGHTFIELD;
        mParams[1] = scaling.x; mParams[2] = scaling.y; mParams[3] = scaling.z;
    }
    HeightField* field() const { return mField; }
    const Vector3& getScale() const { return mScale; }

   private:
    HeightField* mField;
    Vector3 mScale;
};

class PhysicsWorld;

class Material {
   public:
    void setFrictionCoefficient(decimal f) { mFriction = f; push(); }
    void setBounciness(decimal b) { mBounciness = b; push(); }
    void setMassDensity(decimal d) { mDensity = d; push(); }
    decimal getFrictionCoefficient() const { return mFriction; }
    decimal getBounciness() const { return mBounciness; }
    decimal getMassDensity() const { return mDensity; }

   private:
    friend class RigidBody;
    void push() { check(rp3dg_collider_set_material(mWorld, mCollider, mFriction, mBounciness, mDensity), "material"); }
    rp3dg_world* mWorld = nullptr;
    int32_t mCollider = -1;
    decimal mFriction = 0.3f, mBounciness = 0.5f, mDensity = 1.0f;
};

class Collider {
   public:
    Material& getMaterial() { return mMaterial; }
    void setCollisionCategoryBits(uint16 bits) { mCategory = bits; check(rp3dg_collider_set_filter(mWorld, mIndex, mCategory, mMask), "filter"); }
    void setCollideWithMaskBits(uint16 bits) { mMask = bits; check(rp3dg_collider_set_filter(mWorld, mIndex, mCategory, mMask), "filter"); }
    // Collider.cpp:240-300
    bool getIsTrigger() const { return (rp3dg_collider_get_kind(mWorld, mIndex) & RP3DG_COLLIDER_TRIGGER) != 0; }
    void setIsTrigger(bool on) const { check(rp3dg_collider_set_is_trigger(mWorld, mIndex, on ? 1 : 0), "setIsTrigger"); }
    bool getIsSimulationCollider() const { return (rp3dg_collider_get_kind(mWorld, mIndex) & RP3DG_COLLIDER_SIMULATION) != 0; }
    void setIsSimulationCollider(bool on) const {
        check(rp3dg_collider_set_is_simulation_collider(mWorld, mIndex, on ? 1 : 0), "setIsSimulationCollider");
    }
    bool getIsWorldQueryCollider() const { return (rp3dg_collider_get_kind(mWorld, mIndex) & RP3DG_COLLIDER_WORLD_QUERY) != 0; }
    void setIsWorldQueryCollider(bool on) const {
        check(rp3dg_collider_set_is_world_query_collider(mWorld, mIndex, on ? 1 : 0), "setIsWorldQueryCollider");
    }
    int32_t index() const { return mIndex; }

   private:
    friend class RigidBody;
    rp3dg_world* mWorld = nullptr;
    int32_t mIndex = -1;
    uint16 mCategory = 0x0001, mMask = 0xFFFF;
    Material mMaterial;
};

class RigidBody {
   public:
    void setType(BodyType t) {
        check(rp3dg_body_set_type(mWorld, mIndex, t == BodyType::STATIC ? RP3DG_STATIC
                                                  : t == BodyType::KINEMATIC ? RP3DG_KINEMATIC : RP3DG_DYNAMIC), "setType");
    }
    Collider* addCollider(CollisionShape* shape, const Transform& t) {
        const float p[3] = {t.getPosition().x, t.getPosition().y, t.getPosition().z};
        const Quaternion& q = t.getOrientation();
        const float o[4] = {q.x, q.y, q.z, q.w};
        float params[4] = {shape->params()[0], shape->params()[1], shape->params()[2], shape->params()[3]};
        if (shape->type() == RP3DG_SHAPE_CONVEX_MESH) {
            const PolygonVertexArray& a = static_cast<ConvexMeshShape*>(shape)->mesh()->mArray;
            const int32_t h = rp3dg_hull_create(mWorld, a.mVertices.data(), (int32_t)(a.mVertices.size() / 3),
                                                a.mCounts.data(), a.mIndices.data(), (int32_t)a.getNbFaces());
            check(h, "createConvexMesh");
            params[0] = (float)h;
        } else if (shape->type() == RP3DG_SHAPE_CONCAVE_MESH) {
            const TriangleVertexArray& a = static_cast<ConcaveMeshShape*>(shape)->mesh()->mArray;
            const int32_t m = rp3dg_triangle_mesh_create(mWorld, a.mVertices.data(), (int32_t)a.getNbVertices(),
                                                         a.mIndices.data(), (int32_t)a.getNbTriangles(),
                                                         a.getHasNormals() ? a.mNormals.data() : nullptr, nullptr);
            check(m, "createTriangleMesh");
            params[0] = (float)m;
        } else if (shape->type() == RP3DG_SHAPE_HEIGHTFIELD) {
            const HeightField& f = *static_cast<HeightFieldShape*>(shape)->field();
            const int32_t h = rp3dg_heightfield_create(mWorld, f.mColumns, f.mRows, f.mHeights.data());
            check(h, "createHeightField");
            params[0] = (float)h;
        }
        const int32_t c = rp3dg_collider_add(mWorld, mIndex, shape->type(), params, p, o);
        check(c, "addCollider");
        auto col = std::make_unique<Collider>();
        col->mWorld = mWorld;
        col->mIndex = c;
        col->mMaterial.mWorld = mWorld;
        col->mMaterial.mCollider = c;
        col->mMaterial.mFriction = mDefaultFriction;     // the world's default material
        col->mMaterial.mBounciness = mDefaultBounciness;
        mColliders.push_back(std::move(col));
        return mColliders.back().get();
    }
    void updateMassPropertiesFromColliders() { check(rp3dg_body_update_mass_properties(mWorld, mIndex), "updateMassProperties"); }
    // RigidBody::setLinearVelocity / setAngularVelocity (RigidBody.cpp:794-828)
    void setLinearVelocity(const Vector3& v) {
        const float a[3] = {v.x, v.y, v.z};
        check(rp3dg_body_set_velocity(mWorld, mIndex, a, nullptr), "setLinearVelocity");
    }
    // RigidBody::setTransform (RigidBody.cpp:837-866)
    void setTransform(const Transform& t) {
        const float p[3] = {t.getPosition().x, t.getPosition().y, t.getPosition().z};
        const Quaternion& q = t.getOrientation();
        const float o[4] = {q.x, q.y, q.z, q.w};
        check(rp3dg_body_set_transform(mWorld, mIndex, p, o), "setTransform");
    }
    // RigidBody::apply{World,Local}ForceAt{CenterOfMass,WorldPosition,LocalPosition} (RigidBody.cpp:151-311)
    void applyWorldForceAtCenterOfMass(const Vector3& f) { force(f, nullptr, 0); }
    void applyLocalForceAtCenterOfMass(const Vector3& f) { force(f, nullptr, RP3DG_FORCE_LOCAL); }
    void applyWorldForceAtWorldPosition(const Vector3& f, const Vector3& p) { force(f, &p, RP3DG_FORCE_AT_POINT); }
    void applyLocalForceAtWorldPosition(const Vector3& f, const Vector3& p) {
        force(f, &p, RP3DG_FORCE_AT_POINT | RP3DG_FORCE_LOCAL);
    }
    void applyWorldForceAtLocalPosition(const Vector3& f, const Vector3& p) {
        force(f, &p, RP3DG_FORCE_AT_POINT | RP3DG_FORCE_POINT_LOCAL);
    }
    void applyLocalForceAtLocalPosition(const Vector3& f, const Vector3& p) {
        force(f, &p, RP3DG_FORCE_AT_POINT | RP3DG_FORCE_POINT_LOCAL | RP3DG_FORCE_LOCAL);
    }
    // RigidBody::applyWorldTorque / applyLocalTorque / resetForce / resetTorque (RigidBody.cpp:938-987)
    void applyWorldTorque(const Vector3& t) { torque(t, 0); }
    void applyLocalTorque(const Vector3& t) { torque(t, RP3DG_FORCE_LOCAL); }
    void resetForce() { check(rp3dg_body_reset_force_torque(mWorld, mIndex, 1, 0), "resetForce"); }
    void resetTorque() { check(rp3dg_body_reset_force_torque(mWorld, mIndex, 0, 1), "resetTorque"); }
    // RigidBody::getForce / getTorque (RigidBody.cpp:993-1003)
    Vector3 getForce() const {
        float f[3];
        check(rp3dg_body_get_force_torque(mWorld, mIndex, f, nullptr), "getForce");
        return Vector3(f[0], f[1], f[2]);
    }
    Vector3 getTorque() const {
        float t[3];
        check(rp3dg_body_get_force_torque(mWorld, mIndex, nullptr, t), "getTorque");
        return Vector3(t[0], t[1], t[2]);
    }
    // RigidBody::setIsSleeping (RigidBody.cpp:1006-1057)
    void setIsSleeping(bool sleeping) { check(rp3dg_body_set_sleeping(mWorld, mIndex, sleeping ? 1 : 0), "setIsSleeping"); }
    // RigidBody::setIsAllowedToSleep / isSleeping (RigidBody.cpp:1006-1170)
    void setIsAllowedToSleep(bool allowed) {
        check(rp3dg_body_set_allowed_to_sleep(mWorld, mIndex, allowed ? 1 : 0), "setIsAllowedToSleep");
    }
    bool isSleeping() const {
        const int r = rp3dg_world_get_body_sleeping(mWorld, mIndex);
        check(r, "isSleeping");
        return r == 1;
    }
    void setAngularVelocity(const Vector3& w) {
        const float a[3] = {w.x, w.y, w.z};
        check(rp3dg_body_set_velocity(mWorld, mIndex, nullptr, a), "setAngularVelocity");
    }
    void enableGravity(bool on) { check(rp3dg_body_enable_gravity(mWorld, mIndex, on ? 1 : 0), "enableGravity"); }
    void setLinearDamping(decimal d) { mLinDamp = d; check(rp3dg_body_set_damping(mWorld, mIndex, mLinDamp, mAngDamp), "damping"); }
    void setAngularDamping(decimal d) { mAngDamp = d; check(rp3dg_body_set_damping(mWorld, mIndex, mLinDamp, mAngDamp), "damping"); }
    Transform getTransform() const {
        float p[3], q[4];
        check(rp3dg_world_get_body_state(mWorld, mIndex, p, q, nullptr, nullptr), "getTransform");
        return Transform(Vector3(p[0], p[1], p[2]), Quaternion(q[0], q[1], q[2], q[3]));
    }
    Vector3 getLinearVelocity() const {
        float v[3];
        check(rp3dg_world_get_body_state(mWorld, mIndex, nullptr, nullptr, v, nullptr), "getLinearVelocity");
        return Vector3(v[0], v[1], v[2]);
    }
    Vector3 getAngularVelocity() const {
        float w[3];
        check(rp3dg_world_get_body_state(mWorld, mIndex, nullptr, nullptr, nullptr, w), "getAngularVelocity");
        return Vector3(w[0], w[1], w[2]);
    }
    int32_t index() const { return mIndex; }

   private:
    friend class PhysicsWorld;
    void force(const Vector3& f, const Vector3* p, int32_t flags) {
        const float a[3] = {f.x, f.y, f.z};
        const float c[3] = {p ? p->x : 0.f, p ? p->y : 0.f, p ? p->z : 0.f};
        check(rp3dg_body_apply_force(mWorld, mIndex, a, p ? c : nullptr, flags), "applyForce");
    }
    void torque(const Vector3& t, int32_t flags) {
        const float a[3] = {t.x, t.y, t.z};
        check(rp3dg_body_apply_torque(mWorld, mIndex, a, flags), "applyTorque");
    }
    rp3dg_world* mWorld = nullptr;
    int32_t mIndex = -1;
    decimal mDefaultFriction = decimal(0.3), mDefaultBounciness = decimal(0.5);
    decimal mLinDamp = 0, mAngDamp = 0;
    std::vector<std::unique_ptr<Collider>> mColliders;
};

// ---- joints (constraint/*Joint.h), world-space anchor constructors ----
struct JointInfo {
    RigidBody* body1;
    RigidBody* body2;
    Vector3 anchorPointWorldSpace;
    bool isCollisionEnabled = true;
    JointsPositionCorrectionTechnique positionCorrectionTechnique = JointsPositionCorrectionTechnique::NON_LINEAR_GAUSS_SEIDEL;
};
struct BallAndSocketJointInfo : JointInfo {
    BallAndSocketJointInfo(RigidBody* b1, RigidBody* b2, const Vector3& anchor) : JointInfo{b1, b2, anchor} {}
};
struct FixedJointInfo : JointInfo {
    FixedJointInfo(RigidBody* b1, RigidBody* b2, const Vector3& anchor) : JointInfo{b1, b2, anchor} {}
};
struct HingeJointInfo : JointInfo {
    HingeJointInfo(RigidBody* b1, RigidBody* b2, const Vector3& anchor, const Vector3& axis)
        : JointInfo{b1, b2, anchor}, rotationAxisWorld(axis) {}
    Vector3 rotationAxisWorld;
    bool isLimitEnabled = false, isMotorEnabled = false;
    decimal minAngleLimit = -1, maxAngleLimit = 1, motorSpeed = 0, maxMotorTorque = 0;
};
struct SliderJointInfo : JointInfo {   // SliderJoint.h:93-101
    SliderJointInfo(RigidBody* b1, RigidBody* b2, const Vector3& anchor, const Vector3& axis)
        : JointInfo{b1, b2, anchor}, sliderAxisWorldSpace(axis) {}
    Vector3 sliderAxisWorldSpace;
    bool isLimitEnabled = false, isMotorEnabled = false;
    decimal minTranslationLimit = -1, maxTranslationLimit = 1, motorSpeed = 0, maxMotorForce = 0;
};

// Joint handles: the limit / motor setters of the reference's joint classes, applied before the first
// update() (the device uploads joints once).  createJoint returns the derived class as Joint*, which the
// caller casts as with the reference (static_cast<HingeJoint*>(world->createJoint(info))).
class Joint {
   public:
    virtual ~Joint() = default;
    int32_t index() const { return mIndex; }

   protected:
    friend class PhysicsWorld;
    void apply() {
        const float lim[4] = {mLim[0], mLim[1], mLim[2], mLim[3]};
        check(rp3dg_joint_set_limits(mWorld, mIndex, (mLimit ? 1 : 0) | (mMotor ? 2 : 0), lim), "joint limits");
    }
    rp3dg_world* mWorld = nullptr;
    int32_t mIndex = -1;
    bool mLimit = false, mMotor = false;
    decimal mLim[4] = {0, 0, 0, 0};
};
class BallAndSocketJoint : public Joint {   // BallAndSocketJoint.cpp:71-105
   public:
    BallAndSocketJoint() { mLim[0] = decimal(3.141592653589); }
    void enableConeLimit(bool on) { mLimit = on; apply(); }
    bool isConeLimitEnabled() const { return mLimit; }
    void setConeLimitHalfAngle(decimal a) { mLim[0] = a; apply(); }
    decimal getConeLimitHalfAngle() const { return mLim[0]; }
};
class FixedJoint : public Joint {};
class HingeJoint : public Joint {   // HingeJoint.cpp:90-195
   public:
    void enableLimit(bool on) { mLimit = on; apply(); }
    void enableMotor(bool on) { mMotor = on; apply(); }
    void setMinAngleLimit(decimal v) { mLim[0] = v; apply(); }
    void setMaxAngleLimit(decimal v) { mLim[1] = v; apply(); }
    void setMotorSpeed(decimal v) { mLim[2] = v; apply(); }
    void setMaxMotorTorque(decimal v) { mLim[3] = v; apply(); }
    bool isLimitEnabled() const { return mLimit; }
    bool isMotorEnabled() const { return mMotor; }
};
class SliderJoint : public Joint {   // SliderJoint.cpp:98-235
   public:
    void enableLimit(bool on) { mLimit = on; apply(); }
    void enableMotor(bool on) { mMotor = on; apply(); }
    void setMinTranslationLimit(decimal v) { mLim[0] = v; apply(); }
    void setMaxTranslationLimit(decimal v) { mLim[1] = v; apply(); }
    void setMotorSpeed(decimal v) { mLim[2] = v; apply(); }
    void setMaxMotorForce(decimal v) { mLim[3] = v; apply(); }
    bool isLimitEnabled() const { return mLimit; }
    bool isMotorEnabled() const { return mMotor; }
};

// ---- contact events (collision/CollisionCallback.h, engine/EventListener.h) ----
// The callback data is built after update() from the device's event list and contact manifolds; it
// is valid during onContact only, as with the reference.
class CollisionCallback {
   public:
    class ContactPoint {
       public:
        const Vector3& getWorldNormal() const { return mNormal; }
        decimal getPenetrationDepth() const { return mDepth; }
        const Vector3& getLocalPointOnCollider1() const { return mLocal1; }
        const Vector3& getLocalPointOnCollider2() const { return mLocal2; }
        Vector3 mNormal, mLocal1, mLocal2;
        decimal mDepth = 0;
    };
    class ContactPair {
       public:
        enum class EventType { ContactStart, ContactStay, ContactExit };
        Collider* getCollider1() const { return mCollider1; }
        Collider* getCollider2() const { return mCollider2; }
        RigidBody* getBody1() const { return mBody1; }
        RigidBody* getBody2() const { return mBody2; }
        uint32 getNbContactPoints() const { return (uint32)mPoints.size(); }
        ContactPoint getContactPoint(uint32 i) const { return mPoints[i]; }
        EventType getEventType() const { return mType; }
        Collider *mCollider1 = nullptr, *mCollider2 = nullptr;
        RigidBody *mBody1 = nullptr, *mBody2 = nullptr;
        EventType mType = EventType::ContactStart;
        std::vector<ContactPoint> mPoints;
    };
    class CallbackData {
       public:
        uint32 getNbContactPairs() const { return (uint32)mPairs.size(); }
        ContactPair getContactPair(uint64 i) const { return mPairs[i]; }
        std::vector<ContactPair> mPairs;
    };
    virtual ~CollisionCallback() = default;
    virtual void onContact(const CallbackData& data) = 0;
};
// Overlaps of trigger colliders (collision/OverlapCallback.h), delivered to EventListener::onTrigger.
class OverlapCallback {
   public:
    class OverlapPair {
       public:
        enum class EventType { OverlapStart, OverlapStay, OverlapExit };
        Collider* getCollider1() const { return mCollider1; }
        Collider* getCollider2() const { return mCollider2; }
        RigidBody* getBody1() const { return mBody1; }
        RigidBody* getBody2() const { return mBody2; }
        EventType getEventType() const { return mType; }
        Collider *mCollider1 = nullptr, *mCollider2 = nullptr;
        RigidBody *mBody1 = nullptr, *mBody2 = nullptr;
        EventType mType = EventType::OverlapStart;
    };
    class CallbackData {
       public:
        uint32 getNbOverlappingPairs() const { return (uint32)mPairs.size(); }
        OverlapPair getOverlappingPair(uint32 i) const { return mPairs[i]; }
        std::vector<OverlapPair> mPairs;
    };
    virtual ~OverlapCallback() = default;
    virtual void onOverlap(CallbackData& data) = 0;
};
// engine/EventListener.h:45-72: contacts of simulation colliders to onContact, trigger overlaps to onTrigger.
class EventListener : public CollisionCallback {
   public:
    void onContact(const CollisionCallback::CallbackData&) override {}
    virtual void onTrigger(const OverlapCallback::CallbackData&) {}
};

// ---- raycasts (mathematics/Ray.h, collision/RaycastInfo.h) ----
struct Ray {
    Ray(const Vector3& p1, const Vector3& p2, decimal maxFrac = decimal(1.0)) : point1(p1), point2(p2), maxFraction(maxFrac) {}
    Vector3 point1, point2;
    decimal maxFraction;
};
struct RaycastInfo {
    Vector3 worldPoint, worldNormal;
    decimal hitFraction = 0;
    RigidBody* body = nullptr;
    Collider* collider = nullptr;
};
class RaycastCallback {
   public:
    virtual ~RaycastCallback() = default;
    virtual decimal notifyRaycastHit(const RaycastInfo& info) = 0;
};

class PhysicsWorld {
   public:
    struct WorldSettings {
        Vector3 gravity{0, decimal(-9.81), 0};
        decimal persistentContactDistanceThreshold = decimal(0.03);
        decimal defaultFrictionCoefficient = decimal(0.3);
        decimal defaultBounciness = decimal(0.5);
        decimal restitutionVelocityThreshold = decimal(0.5);
        decimal cosAngleSimilarContactManifold = decimal(0.95);
        uint16 defaultVelocitySolverNbIterations = 6;
        uint16 defaultPositionSolverNbIterations = 3;
        bool isSleepingEnabled = true;
        bool isGravityEnabled = true;
        float defaultTimeBeforeSleep = 1.0f;
        decimal defaultSleepLinearVelocity = decimal(0.02);
        decimal defaultSleepAngularVelocity = decimal(3.0) * (decimal(3.141592653589) / decimal(180.0));
    };

    explicit PhysicsWorld(const WorldSettings& ws) : mSettings(ws) {
        rp3dg_world_settings s;
        rp3dg_world_settings_default(&s);
        s.gravity[0] = ws.gravity.x; s.gravity[1] = ws.gravity.y; s.gravity[2] = ws.gravity.z;
        s.persistent_contact_distance_threshold = ws.persistentContactDistanceThreshold;
        s.default_friction_coefficient = ws.defaultFrictionCoefficient;
        s.default_bounciness = ws.defaultBounciness;
        s.restitution_velocity_threshold = ws.restitutionVelocityThreshold;
        s.cos_angle_similar_contact_manifold = ws.cosAngleSimilarContactManifold;
        s.default_velocity_solver_iterations = ws.defaultVelocitySolverNbIterations;
        s.default_position_solver_iterations = ws.defaultPositionSolverNbIterations;
        s.is_sleeping_enabled = ws.isSleepingEnabled ? 1 : 0;
        s.is_gravity_enabled = ws.isGravityEnabled ? 1 : 0;
        s.time_before_sleep = ws.defaultTimeBeforeSleep;
        s.sleep_linear_velocity = ws.defaultSleepLinearVelocity;
        s.sleep_angular_velocity = ws.defaultSleepAngularVelocity;
        mWorld = rp3dg_world_create(&s);
        if (!mWorld) throw std::runtime_error(std::string("createPhysicsWorld: ") + rp3dg_last_error());
    }
    ~PhysicsWorld() { rp3dg_world_destroy(mWorld); }
    PhysicsWorld(const PhysicsWorld&) = delete;
    PhysicsWorld& operator=(const PhysicsWorld&) = delete;

    RigidBody* createRigidBody(const Transform& t) {
        const float p[3] = {t.getPosition().x, t.getPosition().y, t.getPosition().z};
        const Quaternion& q = t.getOrientation();
        const float o[4] = {q.x, q.y, q.z, q.w};
        const int32_t b = rp3dg_body_create(mWorld, p, o);
        check(b, "createRigidBody");
        auto body = std::make_unique<RigidBody>();
        body->mWorld = mWorld;
        body->mIndex = b;
        body->mDefaultFriction = mSettings.defaultFrictionCoefficient;
        body->mDefaultBounciness = mSettings.defaultBounciness;
        mBodies.push_back(std::move(body));
        return mBodies.back().get();
    }
    Joint* createJoint(const BallAndSocketJointInfo& i) {
        return addJoint(std::make_unique<BallAndSocketJoint>(), RP3DG_JOINT_BALLSOCKET, i, Vector3(), false, false, nullptr);
    }
    Joint* createJoint(const FixedJointInfo& i) {
        return addJoint(std::make_unique<FixedJoint>(), RP3DG_JOINT_FIXED, i, Vector3(), false, false, nullptr);
    }
    Joint* createJoint(const HingeJointInfo& i) {
        const decimal lim[4] = {i.minAngleLimit, i.maxAngleLimit, i.motorSpeed, i.maxMotorTorque};
        return addJoint(std::make_unique<HingeJoint>(), RP3DG_JOINT_HINGE, i, i.rotationAxisWorld, i.isLimitEnabled,
                        i.isMotorEnabled, lim);
    }
    Joint* createJoint(const SliderJointInfo& i) {
        const decimal lim[4] = {i.minTranslationLimit, i.maxTranslationLimit, i.motorSpeed, i.maxMotorForce};
        return addJoint(std::make_unique<SliderJoint>(), RP3DG_JOINT_SLIDER, i, i.sliderAxisWorldSpace, i.isLimitEnabled,
                        i.isMotorEnabled, lim);
    }
    void setNbIterationsVelocitySolver(uint16 n) { mVel = n; check(rp3dg_world_set_iterations(mWorld, mVel, mPos), "iterations"); }
    void setNbIterationsPositionSolver(uint16 n) { mPos = n; check(rp3dg_world_set_iterations(mWorld, mVel, mPos), "iterations"); }
    // PhysicsWorld::setContactsPositionCorrectionTechnique (PhysicsWorld.h:599-607)
    void setContactsPositionCorrectionTechnique(ContactsPositionCorrectionTechnique t) {
        check(rp3dg_world_set_contacts_position_correction(mWorld, t == ContactsPositionCorrectionTechnique::BAUMGARTE_CONTACTS
                                                                       ? RP3DG_CONTACTS_BAUMGARTE
                                                                       : RP3DG_CONTACTS_SPLIT_IMPULSES),
              "setContactsPositionCorrectionTechnique");
    }
    void setGravity(const Vector3& g) {
        const float v[3] = {g.x, g.y, g.z};
        check(rp3dg_world_set_gravity(mWorld, v), "setGravity");
    }
    // PhysicsWorld::update (src/engine/PhysicsWorld.cpp:239-307), on the GPU; the contact events of the
    // step go to the event listener (CollisionDetectionSystem::reportContactsAndTriggers)
    void update(decimal timeStep) {
        check(rp3dg_world_update(mWorld, timeStep), "update");
        if (mListener) report();
    }
    // PhysicsWorld::raycast (PhysicsWorld.h:486-490): the device finds every hit of the ray (world-query
    // colliders whose category meets the mask), and the callback sees them from the closest on, with
    // DynamicAABBTree::raycast's use of its return value (DynamicAABBTree.cpp:705-721): 0 stops, a
    // fraction in (0, max) clips the ray (farther hits are not reported), 1 or -1 go on.  A callback that
    // keeps the closest hit, or every hit, sees what the reference reports; one that stops at its first
    // hit gets the closest (the reference gives the first its tree visits).
    void raycast(const Ray& ray, RaycastCallback* callback, unsigned short mask = 0xFFFF) const {
        rp3dg_batch* b = device();
        const float p1[3] = {ray.point1.x, ray.point1.y, ray.point1.z}, p2[3] = {ray.point2.x, ray.point2.y, ray.point2.z};
        const float mf = ray.maxFraction;
        const int64_t n = rp3dg_batch_raycast_all(b, 1, nullptr, p1, p2, &mf, mask, nullptr, nullptr, nullptr, nullptr,
                                                  nullptr, 0);
        check((int)std::min<int64_t>(n, 0), "raycast");
        if (n == 0) return;
        std::vector<int32_t> ray_index(n), col(n);
        std::vector<float> pt(3 * n), nr(3 * n), fr(n);
        check((int)std::min<int64_t>(rp3dg_batch_raycast_all(b, 1, nullptr, p1, p2, &mf, mask, ray_index.data(), col.data(),
                                                              pt.data(), nr.data(), fr.data(), n), 0), "raycast");
        std::vector<int64_t> order(n);
        for (int64_t i = 0; i < n; i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t c) { return fr[a] < fr[c]; });
        const auto owners = owner_map();
        decimal maxf = ray.maxFraction;
        for (int64_t i : order) {
            if (fr[i] > maxf) break;
            RaycastInfo info;
            info.worldPoint = Vector3(pt[3 * i], pt[3 * i + 1], pt[3 * i + 2]);
            info.worldNormal = Vector3(nr[3 * i], nr[3 * i + 1], nr[3 * i + 2]);
            info.hitFraction = fr[i];
            auto it = owners.find(col[i]);
            if (it != owners.end()) { info.collider = it->second.first; info.body = it->second.second; }
            const decimal r = callback->notifyRaycastHit(info);
            if (r == decimal(0)) break;
            if (r > decimal(0) && r < maxf) maxf = r;
        }
    }
    // PhysicsWorld::testOverlap / testCollision (PhysicsWorld.h:319-334, CollisionDetectionSystem.cpp:
    // 1775-1896), on the device; like the reference's, each runs the broadphase first.
    bool testOverlap(RigidBody* body1, RigidBody* body2) {
        const int r = rp3dg_batch_test_overlap_bodies(device(), body1->index(), body2->index());
        check(r, "testOverlap");
        return r == 1;
    }
    void testOverlap(RigidBody* body, OverlapCallback& callback) { overlap(body ? body->index() : -1, callback); }
    void testOverlap(OverlapCallback& callback) { overlap(-1, callback); }
    void testCollision(RigidBody* body1, RigidBody* body2, CollisionCallback& callback) {
        collision(body1->index(), body2->index(), callback);
    }
    void testCollision(RigidBody* body, CollisionCallback& callback) { collision(body->index(), -1, callback); }
    void testCollision(CollisionCallback& callback) { collision(-1, -1, callback); }
    // PhysicsWorld::setEventListener (src/engine/PhysicsWorld.h)
    void setEventListener(EventListener* listener) {
        mListener = listener;
        check(rp3dg_world_set_contact_events(mWorld, listener ? 1 : 0), "setEventListener");
    }
    rp3dg_world* handle() const { return mWorld; }

   private:
    rp3dg_batch* device() const {
        rp3dg_batch* b = rp3dg_world_upload(mWorld);
        if (!b) throw std::runtime_error(std::string("world query: ") + rp3dg_last_error());
        return b;
    }
    std::map<int32_t, std::pair<Collider*, RigidBody*>> owner_map() const {
        std::map<int32_t, std::pair<Collider*, RigidBody*>> owner;
        for (auto& body : mBodies)
            for (auto& col : body->mColliders) owner[col->index()] = {col.get(), body.get()};
        return owner;
    }
    struct QueryRecords {
        std::vector<int32_t> cols, npts, ev;
        std::vector<float> nd, l1, l2;
    };
    QueryRecords query_records(rp3dg_batch* b, int64_t n) const {
        QueryRecords q;
        q.cols.resize(2 * n + 2); q.npts.resize(n + 1); q.ev.resize(n + 1);
        q.nd.resize(16 * n + 16); q.l1.resize(12 * n + 12); q.l2.resize(12 * n + 12);
        std::vector<float> imp(4 * n + 4), fr(3 * n + 3);
        rp3dg_contact_view v{q.cols.data(), q.npts.data(), q.ev.data(), q.nd.data(), q.l1.data(), q.l2.data(), imp.data(),
                             fr.data()};
        check((int)std::min<int64_t>(rp3dg_batch_get_query_results(b, &v, n), 0), "query results");
        return q;
    }
    void overlap(int32_t body, OverlapCallback& callback) {
        rp3dg_batch* b = device();
        const int64_t n = rp3dg_batch_test_overlap(b, body);
        check((int)std::min<int64_t>(n, 0), "testOverlap");
        if (n == 0) return;
        const QueryRecords q = query_records(b, n);
        const auto owners = owner_map();
        OverlapCallback::CallbackData data;
        for (int64_t m = 0; m < n; m++) {
            OverlapCallback::OverlapPair o;
            o.mCollider1 = owners.at(q.cols[2 * m]).first;
            o.mBody1 = owners.at(q.cols[2 * m]).second;
            o.mCollider2 = owners.at(q.cols[2 * m + 1]).first;
            o.mBody2 = owners.at(q.cols[2 * m + 1]).second;
            o.mType = q.ev[m] ? OverlapCallback::OverlapPair::EventType::OverlapStay
                              : OverlapCallback::OverlapPair::EventType::OverlapStart;
            data.mPairs.push_back(o);
        }
        callback.onOverlap(data);
    }
    void collision(int32_t body1, int32_t body2, CollisionCallback& callback) {
        rp3dg_batch* b = device();
        const int64_t n = rp3dg_batch_test_collision(b, body1, body2);
        check((int)std::min<int64_t>(n, 0), "testCollision");
        if (n == 0) return;
        const QueryRecords q = query_records(b, n);
        const auto owners = owner_map();
        CollisionCallback::CallbackData data;
        for (int64_t m = 0; m < n; m++) {
            CollisionCallback::ContactPair p;
            p.mCollider1 = owners.at(q.cols[2 * m]).first;
            p.mBody1 = owners.at(q.cols[2 * m]).second;
            p.mCollider2 = owners.at(q.cols[2 * m + 1]).first;
            p.mBody2 = owners.at(q.cols[2 * m + 1]).second;
            p.mType = q.ev[m] ? CollisionCallback::ContactPair::EventType::ContactStay
                              : CollisionCallback::ContactPair::EventType::ContactStart;
            for (int k = 0; k < q.npts[m]; k++) {
                CollisionCallback::ContactPoint cp;
                const float* d = &q.nd[16 * m + 4 * k];
                cp.mNormal = Vector3(d[0], d[1], d[2]);
                cp.mDepth = d[3];
                cp.mLocal1 = Vector3(q.l1[12 * m + 3 * k], q.l1[12 * m + 3 * k + 1], q.l1[12 * m + 3 * k + 2]);
                cp.mLocal2 = Vector3(q.l2[12 * m + 3 * k], q.l2[12 * m + 3 * k + 1], q.l2[12 * m + 3 * k + 2]);
                p.mPoints.push_back(cp);
            }
            data.mPairs.push_back(p);
        }
        callback.onContact(data);
    }
    // events of the last step with their contact points (one manifold per convex pair), delivered
    // when there is at least one (CollisionDetectionSystem.cpp:1733-1741)
    void report() {
        rp3dg_batch* b = rp3dg_world_get_batch(mWorld);
        const int64_t n = rp3dg_batch_get_contact_events(b, nullptr, nullptr, nullptr, nullptr, 0);
        check((int)n, "contact events");
        if (n == 0) return;
        std::vector<int32_t> c1(n), c2(n), type(n), np(n);
        check((int)rp3dg_batch_get_contact_events(b, c1.data(), c2.data(), type.data(), np.data(), n), "contact events");
        rp3dg_counts counts;
        check(rp3dg_batch_get_counts(b, &counts), "counts");
        const int64_t nm = counts.nb_manifolds;
        std::vector<int32_t> mc(2 * nm + 2), mn(nm + 1), color(nm + 1);
        std::vector<float> nd(16 * nm + 16), l1(12 * nm + 12), l2(12 * nm + 12), imp(4 * nm + 4), fr(3 * nm + 3);
        rp3dg_contact_view v{mc.data(), mn.data(), color.data(), nd.data(), l1.data(), l2.data(), imp.data(), fr.data()};
        check(rp3dg_batch_get_contacts(b, &v, nm), "contacts");
        std::map<std::pair<int32_t, int32_t>, int64_t> manifold;
        for (int64_t m = 0; m < nm; m++) manifold[{mc[2 * m], mc[2 * m + 1]}] = m;
        std::map<int32_t, std::pair<Collider*, RigidBody*>> owner;
        for (auto& body : mBodies)
            for (auto& col : body->mColliders) owner[col->index()] = {col.get(), body.get()};
        CollisionCallback::CallbackData data;
        OverlapCallback::CallbackData overlaps;
        for (int64_t e = 0; e < n; e++) {
            if (type[e] >= RP3DG_TRIGGER_START) {   // onTrigger (CollisionDetectionSystem.cpp:1745-1757)
                OverlapCallback::OverlapPair o;
                o.mCollider1 = owner[c1[e]].first;
                o.mBody1 = owner[c1[e]].second;
                o.mCollider2 = owner[c2[e]].first;
                o.mBody2 = owner[c2[e]].second;
                o.mType = type[e] == RP3DG_TRIGGER_START ? OverlapCallback::OverlapPair::EventType::OverlapStart
                          : type[e] == RP3DG_TRIGGER_STAY ? OverlapCallback::OverlapPair::EventType::OverlapStay
                                                          : OverlapCallback::OverlapPair::EventType::OverlapExit;
                overlaps.mPairs.push_back(o);
                continue;
            }
            CollisionCallback::ContactPair p;
            p.mCollider1 = owner[c1[e]].first;
            p.mBody1 = owner[c1[e]].second;
            p.mCollider2 = owner[c2[e]].first;
            p.mBody2 = owner[c2[e]].second;
            p.mType = type[e] == RP3DG_CONTACT_START ? CollisionCallback::ContactPair::EventType::ContactStart
                      : type[e] == RP3DG_CONTACT_STAY ? CollisionCallback::ContactPair::EventType::ContactStay
                                                      : CollisionCallback::ContactPair::EventType::ContactExit;
            auto it = manifold.find({c1[e], c2[e]});
            if (type[e] != RP3DG_CONTACT_EXIT && it != manifold.end()) {
                const int64_t m = it->second;
                for (int k = 0; k < mn[m]; k++) {
                    CollisionCallback::ContactPoint cp;
                    const float* d = &nd[16 * m + 4 * k];
                    cp.mNormal = Vector3(d[0], d[1], d[2]);
                    cp.mDepth = d[3];
                    cp.mLocal1 = Vector3(l1[12 * m + 3 * k], l1[12 * m + 3 * k + 1], l1[12 * m + 3 * k + 2]);
                    cp.mLocal2 = Vector3(l2[12 * m + 3 * k], l2[12 * m + 3 * k + 1], l2[12 * m + 3 * k + 2]);
                    p.mPoints.push_back(cp);
                }
            }
            data.mPairs.push_back(std::move(p));
        }
        // both callbacks run whenever the step has a colliding or a lost pair (CollisionDetectionSystem.cpp:1708-1757)
        mListener->onContact(data);
        mListener->onTrigger(overlaps);
    }
    Joint* addJoint(std::unique_ptr<Joint> joint, int32_t type, const JointInfo& i, const Vector3& axis, bool limit,
                    bool motor, const decimal* lim) {
        const float a[3] = {i.anchorPointWorldSpace.x, i.anchorPointWorldSpace.y, i.anchorPointWorldSpace.z};
        const float ax[3] = {axis.x, axis.y, axis.z};
        if (lim)
            for (int k = 0; k < 4; k++) joint->mLim[k] = lim[k];
        joint->mLimit = limit;
        joint->mMotor = motor;
        const float l[4] = {joint->mLim[0], joint->mLim[1], joint->mLim[2], joint->mLim[3]};
        const int32_t j = rp3dg_joint_create(mWorld, type, i.body1->index(), i.body2->index(), a, ax,
                                             i.isCollisionEnabled ? 1 : 0,
                                             (limit ? 1 : 0) | (motor ? 2 : 0) |
                                                 (i.positionCorrectionTechnique == JointsPositionCorrectionTechnique::BAUMGARTE_JOINTS
                                                      ? RP3DG_JOINT_BAUMGARTE : 0),
                                             l);
        check(j, "createJoint");
        joint->mWorld = mWorld;
        joint->mIndex = j;
        mJoints.push_back(std::move(joint));
        return mJoints.back().get();
    }
    rp3dg_world* mWorld = nullptr;
    WorldSettings mSettings;
    uint16 mVel = 6, mPos = 3;
    std::vector<std::unique_ptr<RigidBody>> mBodies;
    std::vector<std::unique_ptr<Joint>> mJoints;
    EventListener* mListener = nullptr;
};

// PhysicsCommon (engine/PhysicsCommon.h:176-257): owns worlds and shapes.
class PhysicsCommon {
   public:
    PhysicsWorld* createPhysicsWorld(const PhysicsWorld::WorldSettings& s = PhysicsWorld::WorldSettings()) {
        mWorlds.push_back(std::make_unique<PhysicsWorld>(s));
        return mWorlds.back().get();
    }
    void destroyPhysicsWorld(PhysicsWorld* w) {
        for (auto it = mWorlds.begin(); it != mWorlds.end(); ++it)
            if (it->get() == w) { mWorlds.erase(it); return; }
    }
    SphereShape* createSphereShape(decimal r) { return keep(std::make_unique<SphereShape>(r)); }
    BoxShape* createBoxShape(const Vector3& he) { return keep(std::make_unique<BoxShape>(he)); }
    CapsuleShape* createCapsuleShape(decimal r, decimal h) { return keep(std::make_unique<CapsuleShape>(r, h)); }
    // PhysicsCommon::createConvexMesh(PolygonVertexArray, messages) (PhysicsCommon.cpp:600-630): the mesh
    // is checked as the library builds it (closed, Euler's formula, face size); nullptr and an error
    // message when it is refused.
    ConvexMesh* createConvexMesh(const PolygonVertexArray& a, std::vector<Message>& messages) {
        rp3dg_world* probe = rp3dg_world_create(nullptr);
        const int32_t h = rp3dg_hull_create(probe, a.mVertices.data(), (int32_t)a.getNbVertices(), a.mCounts.data(),
                                            a.mIndices.data(), (int32_t)a.getNbFaces());
        if (h < 0) messages.emplace_back(std::string("Error when creating a ConvexMesh: ") + rp3dg_last_error());
        rp3dg_world_destroy(probe);
        if (h < 0) return nullptr;
        mMeshes.push_back(std::make_unique<ConvexMesh>(a));
        return mMeshes.back().get();
    }
    // PhysicsCommon::createConvexMesh(VertexArray, messages) (PhysicsCommon.cpp:638-660): the convex hull
    // of the points, polygon for polygon the reference's QuickHull result.
    ConvexMesh* createConvexMesh(const VertexArray& va, std::vector<Message>& messages) {
        const int is_double = va.getDataType() == VertexArray::DataType::VERTEX_DOUBLE_TYPE ? 1 : 0;
        int32_t nv = 0, nf = 0;
        rp3dg_quickhull(va.getStart(), (int32_t)va.getNbVertices(), (int32_t)va.getStride(), is_double, nullptr, 0,
                        nullptr, 0, nullptr, 0, &nv, &nf);
        if (nv <= 0 || nf <= 0) {
            messages.emplace_back(std::string("Error during the convex hull construction: ") + rp3dg_last_error());
            return nullptr;
        }
        std::vector<float> v(3 * (size_t)nv);
        std::vector<int32_t> counts((size_t)nf), idx(3 * (size_t)nv + 6 * (size_t)nf);
        int32_t nv2 = 0, nf2 = 0;
        if (rp3dg_quickhull(va.getStart(), (int32_t)va.getNbVertices(), (int32_t)va.getStride(), is_double, v.data(),
                            nv, counts.data(), nf, idx.data(), (int32_t)idx.size(), &nv2, &nf2) < 0) {
            messages.emplace_back(std::string("Error during the convex hull construction: ") + rp3dg_last_error());
            return nullptr;
        }
        std::vector<PolygonVertexArray::PolygonFace> faces((size_t)nf);
        uint32 base = 0;
        for (int32_t f = 0; f < nf; f++) { faces[f] = {(uint32)counts[f], base}; base += (uint32)counts[f]; }
        return createConvexMesh(PolygonVertexArray((uint32)nv, v.data(), 3 * sizeof(float), idx.data(), sizeof(int32_t),
                                                   (uint32)nf, faces.data(),
                                                   PolygonVertexArray::VertexDataType::VERTEX_FLOAT_TYPE,
                                                   PolygonVertexArray::IndexDataType::INDEX_INTEGER_TYPE),
                                messages);
    }
    // createConvexMeshShape (PhysicsCommon.h:206): a scaled mesh is not supported (the scaling must be 1).
    ConvexMeshShape* createConvexMeshShape(ConvexMesh* m, const Vector3& scaling = Vector3(1, 1, 1)) {
        if (scaling.x != 1 || scaling.y != 1 || scaling.z != 1)
            throw std::runtime_error("createConvexMeshShape: scaled convex meshes are not supported");
        return keep(std::make_unique<ConvexMeshShape>(m));
    }
    // PhysicsCommon::createTriangleMesh(TriangleVertexArray, messages) (TriangleMesh.cpp:39-200): nullptr
    // and an error message when the reference refuses the array; a warning per dropped degenerate triangle.
    TriangleMesh* createTriangleMesh(const TriangleVertexArray& a, std::vector<Message>& messages) {
        rp3dg_world* probe = rp3dg_world_create(nullptr);
        int32_t skipped = 0;
        const int32_t m = rp3dg_triangle_mesh_create(probe, a.mVertices.data(), (int32_t)a.getNbVertices(),
                                                     a.mIndices.data(), (int32_t)a.getNbTriangles(),
                                                     a.getHasNormals() ? a.mNormals.data() : nullptr, &skipped);
        if (m < 0) messages.emplace_back(std::string("Error when creating a TriangleMesh: ") + rp3dg_last_error());
        rp3dg_world_destroy(probe);
        if (m < 0) return nullptr;
        for (int32_t i = 0; i < skipped; i++)
            messages.emplace_back("A face with an almost zero area or edge is not part of the final collision shape",
                                  Message::Type::Warning);
        mTriMeshes.push_back(std::make_unique<TriangleMesh>(a));
        return mTriMeshes.back().get();
    }
    ConcaveMeshShape* createConcaveMeshShape(TriangleMesh* m, const Vector3& scaling = Vector3(1, 1, 1)) {
        return keep(std::make_unique<ConcaveMeshShape>(m, scaling));
    }
    // PhysicsCommon::createHeightField (HeightField.cpp:43-130): integer samples times integerHeightScale
    HeightField* createHeightField(int columns, int rows, const void* data, HeightField::HeightDataType type,
                                   std::vector<Message>& messages, decimal integerHeightScale = 1.0f) {
        if (columns < 2 || rows < 2) {
            messages.emplace_back("The number of grid columns and grid rows must be at least two", Message::Type::Error);
            return nullptr;
        }
        auto f = std::make_unique<HeightField>();
        f->mColumns = columns;
        f->mRows = rows;
        f->mHeights.resize((size_t)columns * rows);
        for (size_t i = 0; i < f->mHeights.size(); i++) {
            if (type == HeightField::HeightDataType::HEIGHT_FLOAT_TYPE) f->mHeights[i] = static_cast<const float*>(data)[i];
            else if (type == HeightField::HeightDataType::HEIGHT_DOUBLE_TYPE) f->mHeights[i] = (float)static_cast<const double*>(data)[i];
            else f->mHeights[i] = (float)(static_cast<const int*>(data)[i] * integerHeightScale);
        }
        mHeightFields.push_back(std::move(f));
        return mHeightFields.back().get();
    }
    HeightFieldShape* createHeightFieldShape(HeightField* f, const Vector3& scaling = Vector3(1, 1, 1)) {
        return keep(std::make_unique<HeightFieldShape>(f, scaling));
    }

   private:
    template <typename T>
    T* keep(std::unique_ptr<T> p) {
        T* r = p.get();
        mShapes.push_back(std::move(p));
        return r;
    }
    std::vector<std::unique_ptr<PhysicsWorld>> mWorlds;
    std::vector<std::unique_ptr<CollisionShape>> mShapes;
    std::vector<std::unique_ptr<ConvexMesh>> mMeshes;
    std::vector<std::unique_ptr<TriangleMesh>> mTriMeshes;
    std::vector<std::unique_ptr<HeightField>> mHeightFields;
};

}  // namespace reactphysics3d
