// ORACLE / TEST INFRASTRUCTURE — never linked into the product library.
//
// C-ABI harness over the *unmodified* ReactPhysics3D 0.10.2 library compiled
// from /root/reference (see oracle/Makefile).  It builds a reference
// PhysicsWorld from the same flat scene description the tests give the MI355X
// library and exposes:
//   * rp3dref_step           : PhysicsWorld::update(dt) (src/engine/PhysicsWorld.cpp:239-307)
//   * rp3dref_step_ordered   : the same phases, but the contact solver rows are
//                              permuted into a caller-given manifold order before
//                              warm starting (the "coloured-order oracle"): the
//                              solver then runs exactly the sequence the GPU's
//                              colour sweeps run.
//   * dumps of bodies, overlapping pairs and contacts.
// Compiled with -fno-access-control so that private members can be read without
// editing any reference file.
#include <reactphysics3d/reactphysics3d.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <tuple>
#include <map>
#include <unordered_map>
#include <vector>

using namespace reactphysics3d;

extern "C" {

struct RefSettings {
    float gravity[3];
    int32_t vel_iters, pos_iters;
    float persistent, restitution_threshold, cos_similar;
    float default_friction, default_bounciness;
    int32_t sleeping;
};

struct RefBody {
    int32_t type;
    float pos[3];
    float quat[4];
    float lin[3];
    float ang[3];
    int32_t gravity;
    float ldamp, adamp;
    float llock[3], alock[3];
};

struct RefCollider {
    int32_t body;
    int32_t shape;    // 0 sphere, 1 capsule, 2 box, 3 convex mesh
    float params[4];
    float lpos[3];
    float lquat[4];
    float friction, bounciness, density;
    int32_t category, mask;
    int32_t kind;     // bit0 trigger, bit1 simulation collider, bit2 world-query collider
};

struct RefJoint {
    int32_t type;     // 0 ball-socket, 1 slider, 2 hinge, 3 fixed
    int32_t body1, body2;
    float anchor[3];
    float axis[3];
    int32_t collide;
    int32_t flags;    // bit0 limit, bit1 motor, bit2 Baumgarte position correction
    float limits[4];  // min, max, motor speed, max motor force/torque
};

}  // extern "C"

struct RefWorld;

// EventListener that records every contact event of a step as (collider1, collider2, event type,
// contact points).
struct EventRecorder : public EventListener {
    RefWorld* R = nullptr;
    std::vector<int32_t> ev;
    void onContact(const CollisionCallback::CallbackData& data) override;
    void onTrigger(const OverlapCallback::CallbackData& data) override;   // recorded as type 3 + OverlapPair::EventType
};

struct RefWorld {
    PhysicsCommon common;
    PhysicsWorld* world = nullptr;
    std::vector<RigidBody*> bodies;
    std::vector<Collider*> colliders;
    std::vector<Joint*> joints;
    std::unordered_map<uint32_t, int> collider_of_entity;
    std::unordered_map<uint32_t, int> body_of_entity;
    std::vector<int32_t> joint_order;   // solve order requested by rp3dref_set_joint_order (joint indices)
    std::vector<TriangleMesh*> tri_meshes;   // concave shapes' data (PhysicsCommon::createTriangleMesh)
    std::vector<HeightField*> heightfields;  // (PhysicsCommon::createHeightField)
    EventRecorder* events = nullptr;    // set by rp3dref_enable_events
    float last_dt = 0.f;
};

static Transform make_tf(const float* p, const float* q) {
    return Transform(Vector3(p[0], p[1], p[2]), Quaternion(q[0], q[1], q[2], q[3]));
}

extern "C" {

int rp3dref_precision_bytes() { return (int)sizeof(decimal); }

void* rp3dref_create(const RefSettings* s, int nb, const RefBody* bodies, int nc, const RefCollider* cols, int nj,
                     const RefJoint* joints, const float* hull_vertices, const int32_t* hull_offsets) {
    RefWorld* R = new RefWorld();
    PhysicsWorld::WorldSettings ws;
    ws.gravity = Vector3(s->gravity[0], s->gravity[1], s->gravity[2]);
    ws.persistentContactDistanceThreshold = s->persistent;
    ws.restitutionVelocityThreshold = s->restitution_threshold;
    ws.cosAngleSimilarContactManifold = s->cos_similar;
    ws.defaultFrictionCoefficient = s->default_friction;
    ws.defaultBounciness = s->default_bounciness;
    ws.isSleepingEnabled = s->sleeping != 0;
    ws.defaultVelocitySolverNbIterations = (uint16)s->vel_iters;
    ws.defaultPositionSolverNbIterations = (uint16)s->pos_iters;
    R->world = R->common.createPhysicsWorld(ws);
    // convex meshes from explicit polygons (PhysicsCommon::createConvexMesh(PolygonVertexArray)):
    // hull_offsets = [nb hulls, then per hull: nv, nf, vertex count of each face, face vertex indices];
    // nf = -1: the hull's nv vertices are a point cloud for createConvexMesh(VertexArray)
    std::vector<ConvexMesh*> meshes;
    if (hull_offsets && hull_vertices) {
        const int nh = hull_offsets[0];
        int k = 1;
        size_t voff = 0;
        for (int h = 0; h < nh; h++) {
            const int nv = hull_offsets[k++];
            const int nf = hull_offsets[k++];
            if (nf < 0) {   // a point cloud: PhysicsCommon::createConvexMesh(VertexArray) (QuickHull)
                VertexArray va(hull_vertices + 3 * voff, 3 * sizeof(float), (uint32)nv, VertexArray::DataType::VERTEX_FLOAT_TYPE);
                std::vector<Message> msgs;
                ConvexMesh* mesh = R->common.createConvexMesh(va, msgs);
                if (!mesh) { delete R; return nullptr; }
                meshes.push_back(mesh);
                voff += (size_t)nv;
                continue;
            }
            std::vector<PolygonVertexArray::PolygonFace> faces(nf);
            uint32 base = 0;
            for (int f = 0; f < nf; f++) {
                faces[f].nbVertices = (uint32)hull_offsets[k + f];
                faces[f].indexBase = base;
                base += faces[f].nbVertices;
            }
            k += nf;
            const int32_t* idx = hull_offsets + k;
            k += (int)base;
            PolygonVertexArray pva((uint32)nv, hull_vertices + 3 * voff, 3 * sizeof(float), idx, sizeof(int32_t),
                                   (uint32)nf, faces.data(), PolygonVertexArray::VertexDataType::VERTEX_FLOAT_TYPE,
                                   PolygonVertexArray::IndexDataType::INDEX_INTEGER_TYPE);
            std::vector<Message> msgs;
            ConvexMesh* mesh = R->common.createConvexMesh(pva, msgs);
            if (!mesh) { delete R; return nullptr; }
            meshes.push_back(mesh);
            voff += (size_t)nv;
        }
        // concave shapes follow the hulls: [nb triangle meshes, then per mesh nv, nt, has normals, 3 nt
        // indices] with nv vertices (and nv normals) in the float array, then [nb height fields, then per
        // field columns, rows] with columns * rows heights in the float array
        const int nm = hull_offsets[k++];
        for (int m = 0; m < nm; m++) {
            const int nv = hull_offsets[k++], nt = hull_offsets[k++], hasn = hull_offsets[k++];
            const int32_t* idx = hull_offsets + k;
            k += 3 * nt;
            const float* vs = hull_vertices + 3 * voff;
            voff += (size_t)nv;
            std::vector<Message> msgs;
            TriangleMesh* tm = nullptr;
            if (hasn) {
                const float* ns = hull_vertices + 3 * voff;
                voff += (size_t)nv;
                TriangleVertexArray tva((uint32)nv, vs, 3 * sizeof(float), ns, 3 * sizeof(float), (uint32)nt, idx,
                                        3 * sizeof(int32_t), TriangleVertexArray::VertexDataType::VERTEX_FLOAT_TYPE,
                                        TriangleVertexArray::NormalDataType::NORMAL_FLOAT_TYPE,
                                        TriangleVertexArray::IndexDataType::INDEX_INTEGER_TYPE);
                tm = R->common.createTriangleMesh(tva, msgs);
            } else {
                TriangleVertexArray tva((uint32)nv, vs, 3 * sizeof(float), (uint32)nt, idx, 3 * sizeof(int32_t),
                                        TriangleVertexArray::VertexDataType::VERTEX_FLOAT_TYPE,
                                        TriangleVertexArray::IndexDataType::INDEX_INTEGER_TYPE);
                tm = R->common.createTriangleMesh(tva, msgs);
            }
            if (!tm) { delete R; return nullptr; }
            R->tri_meshes.push_back(tm);
        }
        const int nhf = hull_offsets[k++];
        for (int f = 0; f < nhf; f++) {
            const int cols = hull_offsets[k++], rows = hull_offsets[k++];
            std::vector<Message> msgs;
            HeightField* hf = R->common.createHeightField(cols, rows, hull_vertices + 3 * voff,
                                                          HeightField::HeightDataType::HEIGHT_FLOAT_TYPE, msgs);
            // heights take cols * rows floats, padded to whole xyz triplets
            voff += ((size_t)cols * rows + 2) / 3;
            if (!hf) { delete R; return nullptr; }
            R->heightfields.push_back(hf);
        }
    }
    // canonical creation order (identical to the device-side scene loader):
    // body, setType, colliders (+material, filter), mass properties, velocities
    int ci = 0;
    for (int b = 0; b < nb; b++) {
        const RefBody& B = bodies[b];
        RigidBody* body = R->world->createRigidBody(make_tf(B.pos, B.quat));
        body->setType(B.type == 0 ? BodyType::STATIC : (B.type == 1 ? BodyType::KINEMATIC : BodyType::DYNAMIC));
        R->bodies.push_back(body);
        R->body_of_entity[body->getEntity().id] = b;
        while (ci < nc && cols[ci].body == b) {
            const RefCollider& C = cols[ci];
            CollisionShape* shape = nullptr;
            if (C.shape == 0) shape = R->common.createSphereShape(C.params[0]);
            else if (C.shape == 1) shape = R->common.createCapsuleShape(C.params[0], C.params[1]);
            else if (C.shape == 2) shape = R->common.createBoxShape(Vector3(C.params[0], C.params[1], C.params[2]));
            else if (C.shape == 3 && (size_t)C.params[0] < meshes.size())
                shape = R->common.createConvexMeshShape(meshes[(size_t)C.params[0]]);
            else if (C.shape == 4 && (size_t)C.params[0] < R->tri_meshes.size())
                shape = R->common.createConcaveMeshShape(R->tri_meshes[(size_t)C.params[0]],
                                                         Vector3(C.params[1], C.params[2], C.params[3]));
            else if (C.shape == 5 && (size_t)C.params[0] < R->heightfields.size())
                shape = R->common.createHeightFieldShape(R->heightfields[(size_t)C.params[0]],
                                                         Vector3(C.params[1], C.params[2], C.params[3]));
            else { delete R; return nullptr; }
            Collider* col = body->addCollider(shape, make_tf(C.lpos, C.lquat));
            col->getMaterial().setFrictionCoefficient(C.friction);
            col->getMaterial().setBounciness(C.bounciness);
            col->getMaterial().setMassDensity(C.density);
            if (C.category != 0x0001) col->setCollisionCategoryBits((unsigned short)C.category);
            if (C.mask != 0xFFFF) col->setCollideWithMaskBits((unsigned short)C.mask);
            if (C.kind & 1) col->setIsTrigger(true);
            else if (!(C.kind & 2)) col->setIsSimulationCollider(false);
            if (!(C.kind & 4)) col->setIsWorldQueryCollider(false);
            R->colliders.push_back(col);
            R->collider_of_entity[col->getEntity().id] = ci;
            ci++;
        }
        if (B.type == 2) body->updateMassPropertiesFromColliders();
        if (B.type != 0) {
            body->setLinearVelocity(Vector3(B.lin[0], B.lin[1], B.lin[2]));
            body->setAngularVelocity(Vector3(B.ang[0], B.ang[1], B.ang[2]));
        }
        if (!B.gravity) body->enableGravity(false);
        if (B.ldamp != 0.f) body->setLinearDamping(B.ldamp);
        if (B.adamp != 0.f) body->setAngularDamping(B.adamp);
        if (B.llock[0] != 1.f || B.llock[1] != 1.f || B.llock[2] != 1.f)
            body->setLinearLockAxisFactor(Vector3(B.llock[0], B.llock[1], B.llock[2]));
        if (B.alock[0] != 1.f || B.alock[1] != 1.f || B.alock[2] != 1.f)
            body->setAngularLockAxisFactor(Vector3(B.alock[0], B.alock[1], B.alock[2]));
    }
    for (int j = 0; j < nj; j++) {
        const RefJoint& J = joints[j];
        RigidBody* b1 = R->bodies[J.body1];
        RigidBody* b2 = R->bodies[J.body2];
        const Vector3 anchor(J.anchor[0], J.anchor[1], J.anchor[2]);
        const Vector3 axis(J.axis[0], J.axis[1], J.axis[2]);
        Joint* joint = nullptr;
        if (J.type == 0) {
            BallAndSocketJointInfo info(b1, b2, anchor);
            info.isCollisionEnabled = J.collide != 0;
            if (J.flags & 4) info.positionCorrectionTechnique = JointsPositionCorrectionTechnique::BAUMGARTE_JOINTS;
            joint = R->world->createJoint(info);
            if (J.flags & 1) {   // cone limit with half angle limits[0]
                BallAndSocketJoint* bs = static_cast<BallAndSocketJoint*>(joint);
                bs->enableConeLimit(true);
                bs->setConeLimitHalfAngle(J.limits[0]);
            }
        } else if (J.type == 1) {
            SliderJointInfo info(b1, b2, anchor, axis);
            info.isCollisionEnabled = J.collide != 0;
            if (J.flags & 4) info.positionCorrectionTechnique = JointsPositionCorrectionTechnique::BAUMGARTE_JOINTS;
            if (J.flags & 1) { info.isLimitEnabled = true; info.minTranslationLimit = J.limits[0]; info.maxTranslationLimit = J.limits[1]; }
            if (J.flags & 2) { info.isMotorEnabled = true; info.motorSpeed = J.limits[2]; info.maxMotorForce = J.limits[3]; }
            joint = R->world->createJoint(info);
        } else if (J.type == 2) {
            HingeJointInfo info(b1, b2, anchor, axis);
            info.isCollisionEnabled = J.collide != 0;
            if (J.flags & 4) info.positionCorrectionTechnique = JointsPositionCorrectionTechnique::BAUMGARTE_JOINTS;
            if (J.flags & 1) { info.isLimitEnabled = true; info.minAngleLimit = J.limits[0]; info.maxAngleLimit = J.limits[1]; }
            if (J.flags & 2) { info.isMotorEnabled = true; info.motorSpeed = J.limits[2]; info.maxMotorTorque = J.limits[3]; }
            joint = R->world->createJoint(info);
        } else {
            FixedJointInfo info(b1, b2, anchor);
            info.isCollisionEnabled = J.collide != 0;
            if (J.flags & 4) info.positionCorrectionTechnique = JointsPositionCorrectionTechnique::BAUMGARTE_JOINTS;
            joint = R->world->createJoint(info);
        }
        R->joints.push_back(joint);
    }
    return R;
}

void rp3dref_destroy(void* h) {
    RefWorld* R = (RefWorld*)h;
    if (!R) return;
    R->common.destroyPhysicsWorld(R->world);
    delete R->events;
    delete R;
}

// PhysicsWorld::setNbIterationsVelocitySolver / setNbIterationsPositionSolver (PhysicsWorld.cpp:663,957).
int rp3dref_set_iterations(void* h, int vel, int pos) {
    RefWorld* R = (RefWorld*)h;
    R->world->setNbIterationsVelocitySolver((uint16)vel);
    R->world->setNbIterationsPositionSolver((uint16)pos);
    return 0;
}

// PhysicsWorld::setEventListener with a recorder; rp3dref_get_events returns the last step's events.
int rp3dref_enable_events(void* h) {
    RefWorld* R = (RefWorld*)h;
    if (!R->events) {
        R->events = new EventRecorder();
        R->events->R = R;
        R->world->setEventListener(R->events);
    }
    return 0;
}

int rp3dref_get_events(void* h, int32_t* out, int cap) {
    RefWorld* R = (RefWorld*)h;
    if (!R->events) return -1;
    const int n = (int)R->events->ev.size() / 4;
    if (n > cap) return -n;
    std::memcpy(out, R->events->ev.data(), sizeof(int32_t) * 4 * n);
    return n;
}

int rp3dref_step(void* h, float dt) {
    RefWorld* R = (RefWorld*)h;
    if (R->events) R->events->ev.clear();
    R->world->update(dt);
    R->last_dt = dt;
    return 0;
}

static int collider_index(RefWorld* R, Entity e) {
    auto it = R->collider_of_entity.find(e.id);
    return it == R->collider_of_entity.end() ? -1 : it->second;
}

void EventRecorder::onContact(const CollisionCallback::CallbackData& data) {
    for (uint32 i = 0; i < data.getNbContactPairs(); i++) {
        const CollisionCallback::ContactPair p = data.getContactPair(i);
        ev.push_back(collider_index(R, p.getCollider1()->getEntity()));
        ev.push_back(collider_index(R, p.getCollider2()->getEntity()));
        ev.push_back((int32_t)p.getEventType());
        ev.push_back((int32_t)p.getNbContactPoints());
    }
}

void EventRecorder::onTrigger(const OverlapCallback::CallbackData& data) {
    for (uint32 i = 0; i < data.getNbOverlappingPairs(); i++) {
        const OverlapCallback::OverlapPair p = data.getOverlappingPair(i);
        ev.push_back(collider_index(R, p.getCollider1()->getEntity()));
        ev.push_back(collider_index(R, p.getCollider2()->getEntity()));
        ev.push_back(3 + (int32_t)p.getEventType());
        ev.push_back(0);
    }
}

static int apply_joint_order(RefWorld* R);

// ContactSolverSystem::init (src/systems/ContactSolverSystem.cpp:66-105) with the rows of each
// manifold permuted into the order of (order_c1[i], order_c2[i]) before warmStart().
static int solver_init_ordered(RefWorld* R, float dt, const int32_t* oc1, const int32_t* oc2, int n, bool warm_start) {
    PhysicsWorld& W = *R->world;
    ContactSolverSystem& CS = W.mContactSolverSystem;
    CollisionDetectionSystem& CD = W.mCollisionDetection;
    Array<ContactManifold>* manifolds = CD.mCurrentContactManifolds;
    Array<ContactPoint>* points = CD.mCurrentContactPoints;
    CS.mAllContactManifolds = manifolds;
    CS.mAllContactPoints = points;
    CS.mTimeStep = dt;
    const uint32 nbM = (uint32)manifolds->size();
    const uint32 nbP = (uint32)points->size();
    CS.mNbContactManifolds = 0;
    CS.mNbContactPoints = 0;
    CS.mContactConstraints = nullptr;
    CS.mContactPoints = nullptr;
    if (nbM == 0 || nbP == 0) return 0;
    CS.mContactPoints = static_cast<ContactSolverSystem::ContactPointSolver*>(
        W.mMemoryManager.allocate(MemoryManager::AllocationType::Frame, sizeof(ContactSolverSystem::ContactPointSolver) * nbP));
    CS.mContactConstraints = static_cast<ContactSolverSystem::ContactManifoldSolver*>(
        W.mMemoryManager.allocate(MemoryManager::AllocationType::Frame, sizeof(ContactSolverSystem::ContactManifoldSolver) * nbM));
    const uint32 nbIslands = W.mIslands.getNbIslands();
    for (uint32 i = 0; i < nbIslands; i++)
        if (W.mIslands.nbContactManifolds[i] > 0) CS.initializeForIsland(i);
    // permutation; a pair of a convex and a concave shape has up to four manifolds, created one after the
    // other: oc2's bits 28-31 carry the manifold's index among its pair's (its sub-index)
    std::map<std::tuple<int, int, int>, int> pos;
    for (int i = 0; i < n; i++) pos[std::make_tuple(oc1[i], oc2[i] & 0x0fffffff, (int)((uint32_t)oc2[i] >> 28))] = i;
    std::vector<int> msub(nbM, 0);
    {
        std::map<std::pair<int, int>, int> seen;
        for (uint32 c = 0; c < nbM; c++) {
            const ContactManifold& cm = (*manifolds)[c];
            msub[c] = seen[{collider_index(R, cm.colliderEntity1), collider_index(R, cm.colliderEntity2)}]++;
        }
    }
    const uint32 m = CS.mNbContactManifolds;
    std::vector<uint32> first(m);
    uint32 acc = 0;
    for (uint32 c = 0; c < m; c++) { first[c] = acc; acc += CS.mContactConstraints[c].nbContacts; }
    std::vector<std::pair<int, uint32>> keyed(m);
    bool missing = false;
    for (uint32 c = 0; c < m; c++) {
        const ContactManifold* em = CS.mContactConstraints[c].externalContactManifold;
        const int a = collider_index(R, em->colliderEntity1), b = collider_index(R, em->colliderEntity2);
        auto it = pos.find(std::make_tuple(a, b, msub[(size_t)(em - &(*manifolds)[0])]));
        if (it == pos.end()) missing = true;
        keyed[c] = {it == pos.end() ? 0 : it->second, c};
    }
    // a manifold missing from the given order: keep the reference's own order so that the step still
    // completes (the world stays consistent and can be destroyed) and report the mismatch
    if (missing)
        for (uint32 c = 0; c < m; c++) keyed[c] = {(int)c, c};
    std::sort(keyed.begin(), keyed.end());
    std::vector<ContactSolverSystem::ContactManifoldSolver> ms(CS.mContactConstraints, CS.mContactConstraints + m);
    std::vector<ContactSolverSystem::ContactPointSolver> ps(CS.mContactPoints, CS.mContactPoints + CS.mNbContactPoints);
    uint32 wp = 0;
    for (uint32 c = 0; c < m; c++) {
        const uint32 src = keyed[c].second;
        CS.mContactConstraints[c] = ms[src];
        for (int k = 0; k < ms[src].nbContacts; k++) CS.mContactPoints[wp++] = ps[first[src] + k];
    }
    if (warm_start) CS.warmStart();
    return missing ? 1 : 0;
}

// Learned warm start, the project's extension of ContactSolverSystem::warmStart
// (src/systems/ContactSolverSystem.cpp:326-476): a point without a resting (cached) impulse starts from
// max(pred_n, 0) instead of 0 and is applied like a resting point; a manifold with no resting point
// starts its two friction impulses from the sums of its points' predicted friction impulses (twist 0)
// and applies them like a projected cached impulse.  Everything else, and every operation order, is the
// reference's.  pred: 12 floats per solver manifold (4 point slots x {normal, friction 1, friction 2}).
static void warm_start_learned(ContactSolverSystem& CS, const float* pred) {
    RigidBodyComponents& RB = CS.mRigidBodyComponents;
    uint32 cpi = 0;
    for (uint32 c = 0; c < CS.mNbContactManifolds; c++) {
        ContactSolverSystem::ContactManifoldSolver& M = CS.mContactConstraints[c];
        const uint32 b1 = M.rigidBodyComponentIndexBody1, b2 = M.rigidBodyComponentIndexBody2;
        bool resting = false;
        decimal plx = 0, ply = 0;
        for (short int i = 0; i < M.nbContacts; i++) {
            ContactSolverSystem::ContactPointSolver& P = CS.mContactPoints[cpi];
            if (!P.isRestingContact) {
                const float* pr = pred + 12 * c + 3 * i;
                P.penetrationImpulse = pr[0] > 0.0f ? pr[0] : 0.0f;
                plx += pr[1];
                ply += pr[2];
            } else {
                resting = true;
            }
            const decimal lam = P.penetrationImpulse;
            const Vector3 ip(P.normal.x * lam, P.normal.y * lam, P.normal.z * lam);
            RB.mConstrainedLinearVelocities[b1].x -= M.massInverseBody1 * ip.x * M.linearLockAxisFactorBody1.x;
            RB.mConstrainedLinearVelocities[b1].y -= M.massInverseBody1 * ip.y * M.linearLockAxisFactorBody1.y;
            RB.mConstrainedLinearVelocities[b1].z -= M.massInverseBody1 * ip.z * M.linearLockAxisFactorBody1.z;
            RB.mConstrainedAngularVelocities[b1].x -= P.i1TimesR1CrossN.x * M.angularLockAxisFactorBody1.x * lam;
            RB.mConstrainedAngularVelocities[b1].y -= P.i1TimesR1CrossN.y * M.angularLockAxisFactorBody1.y * lam;
            RB.mConstrainedAngularVelocities[b1].z -= P.i1TimesR1CrossN.z * M.angularLockAxisFactorBody1.z * lam;
            RB.mConstrainedLinearVelocities[b2].x += M.massInverseBody2 * ip.x * M.linearLockAxisFactorBody2.x;
            RB.mConstrainedLinearVelocities[b2].y += M.massInverseBody2 * ip.y * M.linearLockAxisFactorBody2.y;
            RB.mConstrainedLinearVelocities[b2].z += M.massInverseBody2 * ip.z * M.linearLockAxisFactorBody2.z;
            RB.mConstrainedAngularVelocities[b2].x += P.i2TimesR2CrossN.x * M.angularLockAxisFactorBody2.x * lam;
            RB.mConstrainedAngularVelocities[b2].y += P.i2TimesR2CrossN.y * M.angularLockAxisFactorBody2.y * lam;
            RB.mConstrainedAngularVelocities[b2].z += P.i2TimesR2CrossN.z * M.angularLockAxisFactorBody2.z * lam;
            cpi++;
        }
        if (resting) {
            const Vector3 old(M.friction1Impulse * M.oldFrictionVector1.x + M.friction2Impulse * M.oldFrictionVector2.x,
                              M.friction1Impulse * M.oldFrictionVector1.y + M.friction2Impulse * M.oldFrictionVector2.y,
                              M.friction1Impulse * M.oldFrictionVector1.z + M.friction2Impulse * M.oldFrictionVector2.z);
            M.friction1Impulse = old.dot(M.frictionVector1);
            M.friction2Impulse = old.dot(M.frictionVector2);
        } else if (M.nbContacts > 0) {
            M.friction1Impulse = plx;
            M.friction2Impulse = ply;
            M.frictionTwistImpulse = 0;
        } else {
            M.friction1Impulse = M.friction2Impulse = M.frictionTwistImpulse = 0;
            continue;
        }
        const decimal f1 = M.friction1Impulse, f2 = M.friction2Impulse, tw = M.frictionTwistImpulse;
        Vector3 a1(-M.r1CrossT1.x * f1, -M.r1CrossT1.y * f1, -M.r1CrossT1.z * f1);
        Vector3 l2(M.frictionVector1.x * f1, M.frictionVector1.y * f1, M.frictionVector1.z * f1);
        Vector3 a2(M.r2CrossT1.x * f1, M.r2CrossT1.y * f1, M.r2CrossT1.z * f1);
        RB.mConstrainedLinearVelocities[b1] -= M.massInverseBody1 * l2 * M.linearLockAxisFactorBody1;
        RB.mConstrainedAngularVelocities[b1] += M.angularLockAxisFactorBody1 * (M.inverseInertiaTensorBody1 * a1);
        RB.mConstrainedLinearVelocities[b2] += M.massInverseBody2 * l2 * M.linearLockAxisFactorBody2;
        RB.mConstrainedAngularVelocities[b2] += M.angularLockAxisFactorBody2 * (M.inverseInertiaTensorBody2 * a2);
        a1 = Vector3(-M.r1CrossT2.x * f2, -M.r1CrossT2.y * f2, -M.r1CrossT2.z * f2);
        l2 = Vector3(M.frictionVector2.x * f2, M.frictionVector2.y * f2, M.frictionVector2.z * f2);
        a2 = Vector3(M.r2CrossT2.x * f2, M.r2CrossT2.y * f2, M.r2CrossT2.z * f2);
        RB.mConstrainedLinearVelocities[b1].x -= M.massInverseBody1 * l2.x * M.linearLockAxisFactorBody1.x;
        RB.mConstrainedLinearVelocities[b1].y -= M.massInverseBody1 * l2.y * M.linearLockAxisFactorBody1.y;
        RB.mConstrainedLinearVelocities[b1].z -= M.massInverseBody1 * l2.z * M.linearLockAxisFactorBody1.z;
        RB.mConstrainedAngularVelocities[b1] += M.angularLockAxisFactorBody1 * (M.inverseInertiaTensorBody1 * a1);
        RB.mConstrainedLinearVelocities[b2].x += M.massInverseBody2 * l2.x * M.linearLockAxisFactorBody2.x;
        RB.mConstrainedLinearVelocities[b2].y += M.massInverseBody2 * l2.y * M.linearLockAxisFactorBody2.y;
        RB.mConstrainedLinearVelocities[b2].z += M.massInverseBody2 * l2.z * M.linearLockAxisFactorBody2.z;
        RB.mConstrainedAngularVelocities[b2] += M.angularLockAxisFactorBody2 * (M.inverseInertiaTensorBody2 * a2);
        a1 = Vector3(-M.normal.x * tw, -M.normal.y * tw, -M.normal.z * tw);
        a2 = Vector3(M.normal.x * tw, M.normal.y * tw, M.normal.z * tw);
        RB.mConstrainedAngularVelocities[b1] += M.angularLockAxisFactorBody1 * (M.inverseInertiaTensorBody1 * a1);
        RB.mConstrainedAngularVelocities[b2] += M.angularLockAxisFactorBody2 * (M.inverseInertiaTensorBody2 * a2);
        // the twist impulse is applied a second time, as in the reference (ContactSolverSystem.cpp:457-466)
        RB.mConstrainedAngularVelocities[b1] -= M.angularLockAxisFactorBody1 * (M.inverseInertiaTensorBody1 * a2);
        RB.mConstrainedAngularVelocities[b2] += M.angularLockAxisFactorBody2 * (M.inverseInertiaTensorBody2 * a2);
    }
}

// How the velocity solver of one ordered step runs.
struct SolveMode {
    float res_thr = 0.0f;        // > 0: iterate until the velocity residual is below it (else fixed count)
    int max_iters = 0;           // iteration cap in residual mode
    const float* pred = nullptr; // learned warm start predictions (12 per manifold, in the given order)
    int iters = 0;               // out: velocity iterations run
    float res = 0.0f;            // out: last residual (residual mode)
};

// Velocity residual between two sweeps: max over the enabled dynamic bodies of
// max(|v_k - v_{k-1}|_inf, |w_k - w_{k-1}|_inf).  The reference has no such rule (it runs
// mNbVelocitySolverIterations, PhysicsWorld.cpp:334); this is the stopping rule of the residual mode.
static float sweep_residual(RigidBodyComponents& RB, std::vector<Vector3>& pv, std::vector<Vector3>& pw) {
    const uint32 n = RB.getNbEnabledComponents();
    float r = 0.0f;
    for (uint32 i = 0; i < n; i++) {
        const Vector3 v = RB.mConstrainedLinearVelocities[i], w = RB.mConstrainedAngularVelocities[i];
        if (RB.mBodyTypes[i] == BodyType::DYNAMIC) {
            float d = std::fabs((float)(v.x - pv[i].x));
            d = std::fmax(d, std::fabs((float)(v.y - pv[i].y)));
            d = std::fmax(d, std::fabs((float)(v.z - pv[i].z)));
            d = std::fmax(d, std::fabs((float)(w.x - pw[i].x)));
            d = std::fmax(d, std::fabs((float)(w.y - pw[i].y)));
            d = std::fmax(d, std::fabs((float)(w.z - pw[i].z)));
            r = std::fmax(r, d);
        }
        pv[i] = v;
        pw[i] = w;
    }
    return r;
}

// PhysicsWorld::update (src/engine/PhysicsWorld.cpp:239-307), phase by phase, with the ordered solver init.
static int step_ordered_mode(RefWorld* R, float dt, const int32_t* oc1, const int32_t* oc2, int n, SolveMode* mode) {
    PhysicsWorld& W = *R->world;
    if (R->events) R->events->ev.clear();
    W.mCollisionDetection.computeCollisionDetection();
    W.createIslands();
    W.mCollisionDetection.createContacts();
    W.mCollisionDetection.reportContactsAndTriggers();
    W.updateBodiesInverseWorldInertiaTensors();
    W.enableDisableJoints();
    if (!R->joint_order.empty() && apply_joint_order(R) != 0) return -1;
    W.mDynamicsSystem.integrateRigidBodiesVelocities(dt);
    // solveContactsAndConstraints (PhysicsWorld.cpp:321-345)
    const float* pred = mode ? mode->pred : nullptr;
    const int order_rc = solver_init_ordered(R, dt, oc1, oc2, n, pred == nullptr);
    if (pred) warm_start_learned(W.mContactSolverSystem, pred);
    W.mConstraintSolverSystem.initialize(dt);
    if (mode && mode->res_thr > 0.0f) {
        RigidBodyComponents& RB = W.mRigidBodyComponents;
        const uint32 nb = RB.getNbEnabledComponents();
        std::vector<Vector3> pv(nb), pw(nb);
        for (uint32 i = 0; i < nb; i++) { pv[i] = RB.mConstrainedLinearVelocities[i]; pw[i] = RB.mConstrainedAngularVelocities[i]; }
        int it = 0;
        float res = 0.0f;
        while (it < mode->max_iters) {
            W.mConstraintSolverSystem.solveVelocityConstraints();
            W.mContactSolverSystem.solve();
            it++;
            res = sweep_residual(RB, pv, pw);
            if (res < mode->res_thr) break;
        }
        mode->iters = it;
        mode->res = res;
    } else {
        for (uint32 i = 0; i < W.mNbVelocitySolverIterations; i++) {
            W.mConstraintSolverSystem.solveVelocityConstraints();
            W.mContactSolverSystem.solve();
        }
        if (mode) { mode->iters = (int)W.mNbVelocitySolverIterations; mode->res = 0.0f; }
    }
    W.mContactSolverSystem.storeImpulses();
    W.mContactSolverSystem.reset();
    W.mDynamicsSystem.integrateRigidBodiesPositions(dt, W.mContactSolverSystem.isSplitImpulseActive());
    W.solvePositionCorrection();
    W.mDynamicsSystem.updateBodiesState();
    W.mCollisionDetection.updateColliders();
    if (W.mIsSleepingEnabled) W.updateSleepingBodies(dt);
    W.mDynamicsSystem.resetBodiesForceAndTorque();
    W.mIslands.clear();
    W.mProcessContactPairsOrderIslands.clear(true);
    W.mMemoryManager.resetFrameAllocator();
    R->last_dt = dt;
    return order_rc != 0 ? -1 : 0;
}

int rp3dref_step_ordered(void* h, float dt, const int32_t* oc1, const int32_t* oc2, int n) {
    return step_ordered_mode((RefWorld*)h, dt, oc1, oc2, n, nullptr);
}

// rp3dref_step_ordered with the residual stopping rule (res_thr > 0, at most max_iters sweeps) and / or
// the learned warm start (pred != NULL); iters / res (may be NULL) receive the sweeps run and the last
// residual.
int rp3dref_step_ordered_mode(void* h, float dt, const int32_t* oc1, const int32_t* oc2, int n, float res_thr,
                              int max_iters, const float* pred, int32_t* iters, float* res) {
    SolveMode m;
    m.res_thr = res_thr;
    m.max_iters = max_iters;
    m.pred = pred;
    const int rc = step_ordered_mode((RefWorld*)h, dt, oc1, oc2, n, &m);
    if (iters) *iters = m.iters;
    if (res) *res = m.res;
    return rc;
}

}  // extern "C"

// Runs fn(w) for w in [0, n) on `threads` host threads (worlds are independent).
template <class F>
static void parallel_worlds(int n, int threads, F fn) {
    threads = std::max(1, std::min(threads, n));
    std::atomic<int> next(0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&]() {
            for (int w = next.fetch_add(1); w < n; w = next.fetch_add(1)) fn(w);
        });
    for (auto& th : pool) th.join();
}

extern "C" {

// One ordered step of nw reference worlds on `threads` threads.  The manifold order of world w is
// (oc1, oc2)[moff[w] .. moff[w + 1]) in world-local collider indices; pred (NULL: plain warm start)
// holds 12 floats per manifold at the same offsets.  iters / res receive each world's sweeps and last
// residual.  Returns 0, or -(1 + w) for the first world whose manifolds are missing from its order.
int rp3dref_batch_step_ordered(void** hs, int nw, float dt, const int32_t* oc1, const int32_t* oc2,
                               const int64_t* moff, float res_thr, int max_iters, const float* pred, int32_t* iters,
                               float* res, int threads) {
    std::vector<int> rc(nw, 0);
    parallel_worlds(nw, threads, [&](int w) {
        SolveMode m;
        m.res_thr = res_thr;
        m.max_iters = max_iters;
        m.pred = pred ? pred + 12 * moff[w] : nullptr;
        rc[w] = step_ordered_mode((RefWorld*)hs[w], dt, oc1 + moff[w], oc2 + moff[w], (int)(moff[w + 1] - moff[w]), &m);
        if (iters) iters[w] = m.iters;
        if (res) res[w] = m.res;
    });
    for (int w = 0; w < nw; w++)
        if (rc[w] != 0) return -(1 + w);
    return 0;
}

// Reorders the joint components of every joint kind so that the reference solves them in the given
// order (global joint creation indices, as the device solves them).  Components::swapComponents keeps
// each component's data (and warm-start impulses) with its entity.
// Only the enabled components are reordered (the solvers iterate those; a joint whose two bodies are
// static or sleeping sits in the disabled part, PhysicsWorld.cpp:363-393).
static void reorder(Components& comps, const std::vector<Entity>& want) {
    const uint32 nen = comps.getNbEnabledComponents();
    uint32 i = 0;
    for (const Entity& e : want) {
        const uint32 cur = comps.getEntityIndex(e);
        if (cur >= nen) continue;
        if (cur != i) comps.swapComponents(i, cur);
        i++;
    }
}

static int apply_joint_order(RefWorld* R) {
    PhysicsWorld& W = *R->world;
    std::vector<Entity> bs, fx, hg, sl;
    for (int32_t o : R->joint_order) {
        Joint* j = R->joints[o];
        switch (j->getType()) {
            case JointType::BALLSOCKETJOINT: bs.push_back(j->getEntity()); break;
            case JointType::FIXEDJOINT: fx.push_back(j->getEntity()); break;
            case JointType::HINGEJOINT: hg.push_back(j->getEntity()); break;
            case JointType::SLIDERJOINT: sl.push_back(j->getEntity()); break;
            default: return -1;
        }
    }
    reorder(W.mBallAndSocketJointsComponents, bs);
    reorder(W.mFixedJointsComponents, fx);
    reorder(W.mHingeJointsComponents, hg);
    reorder(W.mSliderJointsComponents, sl);
    return 0;
}

// ConcaveShape::computeOverlappingTriangles of a concave collider's shape for an AABB in its local space:
// shape ids, vertices and vertex normals of the overlapping triangles in the middle phase's order.
int rp3dref_concave_overlapping_triangles(void* h, int collider, const float* aabb_min, const float* aabb_max,
                                          int32_t* ids, float* vertices, float* normals, int capacity) {
    RefWorld* R = (RefWorld*)h;
    if (collider < 0 || collider >= (int)R->colliders.size()) return -1;
    CollisionShape* cs = R->colliders[collider]->getCollisionShape();
    if (cs->isConvex()) return -1;
    ConcaveShape* shape = static_cast<ConcaveShape*>(cs);
    MemoryAllocator& alloc = R->common.mMemoryManager.getHeapAllocator();
    Array<Vector3> tv(alloc), tn(alloc);
    Array<uint32> sid(alloc);
    const AABB aabb(Vector3(aabb_min[0], aabb_min[1], aabb_min[2]), Vector3(aabb_max[0], aabb_max[1], aabb_max[2]));
    shape->computeOverlappingTriangles(aabb, tv, tn, sid, alloc);
    const int n = (int)sid.size();
    for (int i = 0; i < n && i < capacity; i++) {
        ids[i] = (int32_t)sid[i];
        for (int k = 0; k < 3; k++) {
            const Vector3& v = tv[3 * i + k];
            const Vector3& m = tn[3 * i + k];
            vertices[9 * i + 3 * k] = v.x; vertices[9 * i + 3 * k + 1] = v.y; vertices[9 * i + 3 * k + 2] = v.z;
            normals[9 * i + 3 * k] = m.x; normals[9 * i + 3 * k + 1] = m.y; normals[9 * i + 3 * k + 2] = m.z;
        }
    }
    return n;
}

// Sets the joint solve order; with sleeping enabled it is re-applied to the enabled joints after every
// enableDisableJoints (disabling and enabling a component moves it, Components.cpp:159-194).
int rp3dref_set_joint_order(void* h, const int32_t* order, int n) {
    RefWorld* R = (RefWorld*)h;
    for (int i = 0; i < n; i++)
        if (order[i] < 0 || order[i] >= (int)R->joints.size()) return -1;
    R->joint_order.assign(order, order + n);
    return apply_joint_order(R);
}

// PhysicsWorld::raycast of n rays with a callback that keeps the closest hit (returns the hit fraction,
// which clips the ray; the last reported hit is the closest).  collider[i] = -1 when nothing is hit.
struct ClosestHit : public RaycastCallback {
    RefWorld* R = nullptr;
    int collider = -1;
    RaycastInfo info;
    decimal notifyRaycastHit(const RaycastInfo& i) override {
        info.worldPoint = i.worldPoint;
        info.worldNormal = i.worldNormal;
        info.hitFraction = i.hitFraction;
        collider = collider_index(R, i.collider->getEntity());
        return i.hitFraction;
    }
};

int rp3dref_raycast(void* h, int n, const float* p1, const float* p2, const float* maxf, int mask, int32_t* collider,
                    float* point, float* normal, float* fraction) {
    RefWorld* R = (RefWorld*)h;
    for (int i = 0; i < n; i++) {
        const Ray ray(Vector3(p1[3 * i], p1[3 * i + 1], p1[3 * i + 2]), Vector3(p2[3 * i], p2[3 * i + 1], p2[3 * i + 2]),
                      maxf ? maxf[i] : decimal(1.0));
        ClosestHit cb;
        cb.R = R;
        R->world->raycast(ray, &cb, (unsigned short)mask);
        collider[i] = cb.collider;
        for (int k = 0; k < 3; k++) {
            point[3 * i + k] = cb.collider >= 0 ? (float)cb.info.worldPoint[k] : 0.f;
            normal[3 * i + k] = cb.collider >= 0 ? (float)cb.info.worldNormal[k] : 0.f;
        }
        fraction[i] = cb.collider >= 0 ? (float)cb.info.hitFraction : 0.f;
    }
    return 0;
}

// PhysicsWorld::raycast of n rays with a callback that keeps every hit (notifyRaycastHit returns 1):
// the hits in the order the reference reports them, (ray, collider, point, normal, fraction) each.
// Returns the number of hits (only counted past cap).
struct AllHits : public RaycastCallback {
    RefWorld* R = nullptr;
    int ray = 0;
    std::vector<int32_t> rc;
    std::vector<float> pnf;
    decimal notifyRaycastHit(const RaycastInfo& i) override {
        rc.push_back(ray);
        rc.push_back(collider_index(R, i.collider->getEntity()));
        for (int k = 0; k < 3; k++) pnf.push_back((float)i.worldPoint[k]);
        for (int k = 0; k < 3; k++) pnf.push_back((float)i.worldNormal[k]);
        pnf.push_back((float)i.hitFraction);
        return decimal(1.0);
    }
};

int rp3dref_raycast_all(void* h, int n, const float* p1, const float* p2, const float* maxf, int mask,
                        int32_t* ray_collider, float* point_normal_fraction, int cap) {
    RefWorld* R = (RefWorld*)h;
    AllHits cb;
    cb.R = R;
    for (int i = 0; i < n; i++) {
        const Ray ray(Vector3(p1[3 * i], p1[3 * i + 1], p1[3 * i + 2]), Vector3(p2[3 * i], p2[3 * i + 1], p2[3 * i + 2]),
                      maxf ? maxf[i] : decimal(1.0));
        cb.ray = i;
        R->world->raycast(ray, &cb, (unsigned short)mask);
    }
    const int total = (int)(cb.rc.size() / 2);
    for (int k = 0; k < total && k < cap; k++) {
        ray_collider[2 * k] = cb.rc[2 * k];
        ray_collider[2 * k + 1] = cb.rc[2 * k + 1];
        for (int j = 0; j < 7; j++) point_normal_fraction[7 * k + j] = cb.pnf[7 * k + j];
    }
    return total;
}

// PhysicsWorld::setContactsPositionCorrectionTechnique: 1 Baumgarte contacts, 0 split impulses.
int rp3dref_set_contacts_baumgarte(void* h, int baumgarte) {
    RefWorld* R = (RefWorld*)h;
    R->world->setContactsPositionCorrectionTechnique(baumgarte ? ContactsPositionCorrectionTechnique::BAUMGARTE_CONTACTS
                                                               : ContactsPositionCorrectionTechnique::SPLIT_IMPULSES);
    return 0;
}

// World queries.  rp3dref_test_overlap_bodies: PhysicsWorld::testOverlap(body1, body2).
int rp3dref_test_overlap_bodies(void* h, int b1, int b2) {
    RefWorld* R = (RefWorld*)h;
    return R->world->testOverlap(R->bodies[b1], R->bodies[b2]) ? 1 : 0;
}

struct OverlapRecorder : public OverlapCallback {
    RefWorld* R = nullptr;
    std::vector<int32_t> cols, evs;
    void onOverlap(CallbackData& data) override {
        for (uint32 i = 0; i < data.getNbOverlappingPairs(); i++) {
            const OverlapPair p = data.getOverlappingPair(i);
            cols.push_back(collider_index(R, p.getCollider1()->getEntity()));
            cols.push_back(collider_index(R, p.getCollider2()->getEntity()));
            evs.push_back((int32_t)p.getEventType());
        }
    }
};

// testOverlap(callback) (body < 0) or testOverlap(body, callback): the reported collider pairs and their
// OverlapPair::EventType in report order; returns their count (pairs past cap are counted only).
int rp3dref_test_overlap(void* h, int body, int32_t* c1, int32_t* c2, int32_t* ev, int cap) {
    RefWorld* R = (RefWorld*)h;
    OverlapRecorder cb;
    cb.R = R;
    if (body < 0) R->world->testOverlap(cb);
    else R->world->testOverlap(R->bodies[body], cb);
    const int n = (int)cb.cols.size() / 2;
    for (int i = 0; i < n && i < cap; i++) { c1[i] = cb.cols[2 * i]; c2[i] = cb.cols[2 * i + 1]; ev[i] = cb.evs[i]; }
    return n;
}

struct CollisionRecorder : public CollisionCallback {
    RefWorld* R = nullptr;
    std::vector<int32_t> cols, npts, evs;
    std::vector<float> nd, l1, l2;
    void onContact(const CallbackData& data) override {
        for (uint32 i = 0; i < data.getNbContactPairs(); i++) {
            const ContactPair p = data.getContactPair(i);
            cols.push_back(collider_index(R, p.getCollider1()->getEntity()));
            cols.push_back(collider_index(R, p.getCollider2()->getEntity()));
            evs.push_back((int32_t)p.getEventType());
            const uint32 n = p.getNbContactPoints();
            npts.push_back((int32_t)n);
            for (uint32 k = 0; k < 4; k++) {
                if (k < n) {
                    const ContactPoint c = p.getContactPoint(k);
                    for (int j = 0; j < 3; j++) nd.push_back((float)c.getWorldNormal()[j]);
                    nd.push_back((float)c.getPenetrationDepth());
                    for (int j = 0; j < 3; j++) l1.push_back((float)c.getLocalPointOnCollider1()[j]);
                    for (int j = 0; j < 3; j++) l2.push_back((float)c.getLocalPointOnCollider2()[j]);
                } else {
                    for (int j = 0; j < 4; j++) nd.push_back(0.f);
                    for (int j = 0; j < 3; j++) { l1.push_back(0.f); l2.push_back(0.f); }
                }
            }
        }
    }
};

// testCollision(callback) (b1 < 0), testCollision(b1, callback) (b2 < 0) or testCollision(b1, b2,
// callback): one record per reported contact pair (colliders, point count, ContactPair::EventType, 4 x
// normal + depth, 4 x local point on collider 1 and 2), in report order; returns the count (records past cap are counted only).
int rp3dref_test_collision(void* h, int b1, int b2, int32_t* cols, int32_t* npts, int32_t* ev, float* nd, float* l1,
                           float* l2, int cap) {
    RefWorld* R = (RefWorld*)h;
    CollisionRecorder cb;
    cb.R = R;
    if (b1 < 0) R->world->testCollision(cb);
    else if (b2 < 0) R->world->testCollision(R->bodies[b1], cb);
    else R->world->testCollision(R->bodies[b1], R->bodies[b2], cb);
    const int n = (int)cb.npts.size();
    for (int i = 0; i < n && i < cap; i++) {
        cols[2 * i] = cb.cols[2 * i];
        cols[2 * i + 1] = cb.cols[2 * i + 1];
        npts[i] = cb.npts[i];
        ev[i] = cb.evs[i];
        for (int j = 0; j < 16; j++) nd[16 * i + j] = cb.nd[16 * i + j];
        for (int j = 0; j < 12; j++) { l1[12 * i + j] = cb.l1[12 * i + j]; l2[12 * i + j] = cb.l2[12 * i + j]; }
    }
    return n;
}

// CollisionDetectionSystem::computeBroadPhase alone: what a query about bodies of another world leaves in
// this one (a query runs the broadphase of its world before filtering pairs by body).
int rp3dref_broadphase_only(void* h) {
    RefWorld* R = (RefWorld*)h;
    R->world->mCollisionDetection.computeBroadPhase();
    return 0;
}

// RigidBody::applyWorldForceAtCenterOfMass / applyWorldTorque on every body whose force / torque (xyz
// each, NULL = none) is not zero.
int rp3dref_apply_forces(void* h, const float* f, const float* t) {
    RefWorld* R = (RefWorld*)h;
    auto nonzero = [](const float* a, size_t i) { return a && (a[3 * i] != 0.f || a[3 * i + 1] != 0.f || a[3 * i + 2] != 0.f); };
    for (size_t i = 0; i < R->bodies.size(); i++) {
        if (nonzero(f, i)) R->bodies[i]->applyWorldForceAtCenterOfMass(Vector3(f[3 * i], f[3 * i + 1], f[3 * i + 2]));
        if (nonzero(t, i)) R->bodies[i]->applyWorldTorque(Vector3(t[3 * i], t[3 * i + 1], t[3 * i + 2]));
    }
    return 0;
}

// One RigidBody call on body b, by op code: 1 setTransform(a = position, c = orientation xyzw),
// 2 setLinearVelocity(a), 3 setAngularVelocity(a), 4 apply*Force* (a = force, c = point; flags 1 = local
// force, 2 = at a point, 4 = point in the body frame), 5 applyWorldTorque / applyLocalTorque (flags 1),
// 6 resetForce, 7 resetTorque, 8 setIsSleeping(flags != 0).
int rp3dref_body_op(void* h, int b, int op, int flags, const float* a, const float* c) {
    RefWorld* R = (RefWorld*)h;
    if (b < 0 || b >= (int)R->bodies.size()) return -1;
    RigidBody* B = R->bodies[b];
    const Vector3 va = a ? Vector3(a[0], a[1], a[2]) : Vector3::zero();
    const Vector3 vc = c ? Vector3(c[0], c[1], c[2]) : Vector3::zero();
    switch (op) {
        case 1: B->setTransform(make_tf(a, c)); break;
        case 2: B->setLinearVelocity(va); break;
        case 3: B->setAngularVelocity(va); break;
        case 4:
            if (!(flags & 2)) {
                if (flags & 1) B->applyLocalForceAtCenterOfMass(va);
                else B->applyWorldForceAtCenterOfMass(va);
            } else if (flags & 4) {
                if (flags & 1) B->applyLocalForceAtLocalPosition(va, vc);
                else B->applyWorldForceAtLocalPosition(va, vc);
            } else {
                if (flags & 1) B->applyLocalForceAtWorldPosition(va, vc);
                else B->applyWorldForceAtWorldPosition(va, vc);
            }
            break;
        case 5:
            if (flags & 1) B->applyLocalTorque(va);
            else B->applyWorldTorque(va);
            break;
        case 6: B->resetForce(); break;
        case 7: B->resetTorque(); break;
        case 8: B->setIsSleeping(flags != 0); break;
        default: return -1;
    }
    return 0;
}

// RigidBody::getForce / getTorque of every body (xyz each).
int rp3dref_get_force_torque(void* h, float* f, float* t) {
    RefWorld* R = (RefWorld*)h;
    for (size_t i = 0; i < R->bodies.size(); i++) {
        const Vector3 fi = R->bodies[i]->getForce(), ti = R->bodies[i]->getTorque();
        for (int k = 0; k < 3; k++) {
            if (f) f[3 * i + k] = (float)fi[k];
            if (t) t[3 * i + k] = (float)ti[k];
        }
    }
    return (int)R->bodies.size();
}

// RigidBody::isSleeping of every body.
int rp3dref_get_sleeping(void* h, int32_t* out) {
    RefWorld* R = (RefWorld*)h;
    for (size_t i = 0; i < R->bodies.size(); i++) out[i] = R->bodies[i]->isSleeping() ? 1 : 0;
    return (int)R->bodies.size();
}

// Bodies: origin position, orientation (x,y,z,w), centre-of-mass linear velocity, angular velocity.
int rp3dref_get_bodies(void* h, float* pos, float* q, float* v, float* w) {
    RefWorld* R = (RefWorld*)h;
    for (size_t i = 0; i < R->bodies.size(); i++) {
        const Transform& t = R->bodies[i]->getTransform();
        const Vector3 lv = R->bodies[i]->getLinearVelocity(), av = R->bodies[i]->getAngularVelocity();
        if (pos) { pos[3 * i] = (float)t.getPosition().x; pos[3 * i + 1] = (float)t.getPosition().y; pos[3 * i + 2] = (float)t.getPosition().z; }
        if (q) {
            q[4 * i] = (float)t.getOrientation().x; q[4 * i + 1] = (float)t.getOrientation().y;
            q[4 * i + 2] = (float)t.getOrientation().z; q[4 * i + 3] = (float)t.getOrientation().w;
        }
        if (v) { v[3 * i] = (float)lv.x; v[3 * i + 1] = (float)lv.y; v[3 * i + 2] = (float)lv.z; }
        if (w) { w[3 * i] = (float)av.x; w[3 * i + 1] = (float)av.y; w[3 * i + 2] = (float)av.z; }
    }
    return (int)R->bodies.size();
}

// Mass properties as computed by RigidBody::updateMassPropertiesFromColliders.
int rp3dref_get_mass_properties(void* h, float* mass, float* inertia, float* com_local) {
    RefWorld* R = (RefWorld*)h;
    for (size_t i = 0; i < R->bodies.size(); i++) {
        RigidBody* b = R->bodies[i];
        if (mass) mass[i] = (float)b->getMass();
        const Vector3 I = b->getLocalInertiaTensor(), c = b->getLocalCenterOfMass();
        if (inertia) { inertia[3 * i] = (float)I.x; inertia[3 * i + 1] = (float)I.y; inertia[3 * i + 2] = (float)I.z; }
        if (com_local) { com_local[3 * i] = (float)c.x; com_local[3 * i + 1] = (float)c.y; com_local[3 * i + 2] = (float)c.z; }
    }
    return (int)R->bodies.size();
}

// Fat AABBs of the broadphase tree (DynamicAABBTree node of every collider).
int rp3dref_get_fat_aabbs(void* h, float* mn, float* mx) {
    RefWorld* R = (RefWorld*)h;
    for (size_t i = 0; i < R->colliders.size(); i++) {
        const int id = R->colliders[i]->getBroadPhaseId();
        const AABB& a = R->world->mCollisionDetection.mBroadPhaseSystem.mDynamicAABBTree.getFatAABB(id);
        mn[3 * i] = (float)a.getMin().x; mn[3 * i + 1] = (float)a.getMin().y; mn[3 * i + 2] = (float)a.getMin().z;
        mx[3 * i] = (float)a.getMax().x; mx[3 * i + 1] = (float)a.getMax().y; mx[3 * i + 2] = (float)a.getMax().z;
    }
    return (int)R->colliders.size();
}

int rp3dref_get_broadphase_ids(void* h, int32_t* ids) {
    RefWorld* R = (RefWorld*)h;
    for (size_t i = 0; i < R->colliders.size(); i++) ids[i] = R->colliders[i]->getBroadPhaseId();
    return (int)R->colliders.size();
}

// Convex overlapping pairs: collider1 / collider2 as the reference oriented them.
// Overlapping convex pairs, the enabled ones first, then the disabled ones (both bodies static or
// sleeping, kept aside untested: OverlappingPairs::mDisabledConvexPairs).
int rp3dref_get_pairs(void* h, int32_t* c1, int32_t* c2, int cap) {
    RefWorld* R = (RefWorld*)h;
    auto& OP = R->world->mCollisionDetection.mOverlappingPairs;
    const int ne = (int)OP.mConvexPairs.size(), nd = (int)OP.mDisabledConvexPairs.size();
    const int nce = (int)OP.mConcavePairs.size(), ncd = (int)OP.mDisabledConcavePairs.size();
    const int n = ne + nd + nce + ncd;
    if (n > cap) return -n;
    int o = 0;
    auto put = [&](Entity a, Entity b) { c1[o] = collider_index(R, a); c2[o] = collider_index(R, b); o++; };
    for (int i = 0; i < ne; i++) put(OP.mConvexPairs[i].collider1, OP.mConvexPairs[i].collider2);
    for (int i = 0; i < nd; i++) put(OP.mDisabledConvexPairs[i].collider1, OP.mDisabledConvexPairs[i].collider2);
    for (int i = 0; i < nce; i++) put(OP.mConcavePairs[i].collider1, OP.mConcavePairs[i].collider2);
    for (int i = 0; i < ncd; i++) put(OP.mDisabledConcavePairs[i].collider1, OP.mDisabledConcavePairs[i].collider2);
    return n;
}

// Contacts of the last step: per manifold (island order) collider1/2, nb points, friction impulses
// (friction1, friction2, twist); per point (4 slots) normal+depth, local points, normal impulse.
int rp3dref_get_contacts(void* h, int32_t* cols, int32_t* npts, float* nd, float* l1, float* l2, float* imp,
                         float* fric, int cap) {
    RefWorld* R = (RefWorld*)h;
    Array<ContactManifold>& M = *R->world->mCollisionDetection.mCurrentContactManifolds;
    Array<ContactPoint>& P = *R->world->mCollisionDetection.mCurrentContactPoints;
    const int n = (int)M.size();
    if (n > cap) return -n;
    for (int m = 0; m < n; m++) {
        const ContactManifold& cm = M[m];
        cols[2 * m] = collider_index(R, cm.colliderEntity1);
        cols[2 * m + 1] = collider_index(R, cm.colliderEntity2);
        npts[m] = cm.nbContactPoints;
        fric[3 * m] = (float)cm.frictionImpulse1;
        fric[3 * m + 1] = (float)cm.frictionImpulse2;
        fric[3 * m + 2] = (float)cm.frictionTwistImpulse;
        for (int k = 0; k < 4; k++) {
            float* a = nd + 16 * m + 4 * k;
            float* b = l1 + 12 * m + 3 * k;
            float* c = l2 + 12 * m + 3 * k;
            if (k < cm.nbContactPoints) {
                const ContactPoint& p = P[cm.contactPointsIndex + k];
                a[0] = (float)p.getNormal().x; a[1] = (float)p.getNormal().y; a[2] = (float)p.getNormal().z;
                a[3] = (float)p.getPenetrationDepth();
                b[0] = (float)p.getLocalPointOnShape1().x; b[1] = (float)p.getLocalPointOnShape1().y; b[2] = (float)p.getLocalPointOnShape1().z;
                c[0] = (float)p.getLocalPointOnShape2().x; c[1] = (float)p.getLocalPointOnShape2().y; c[2] = (float)p.getLocalPointOnShape2().z;
                imp[4 * m + k] = (float)p.getPenetrationImpulse();
            } else {
                a[0] = a[1] = a[2] = a[3] = 0.f;
                b[0] = b[1] = b[2] = 0.f;
                c[0] = c[1] = c[2] = 0.f;
                imp[4 * m + k] = 0.f;
            }
        }
    }
    return n;
}

// Batch readback over nw worlds, concatenated in world order; collider indices are offset by
// collider_base[w] (the batch's global numbering).  Each returns the total count, or -(needed) when cap
// is too small.
int rp3dref_batch_get_bodies(void** hs, int nw, float* pos, float* q, float* v, float* w) {
    int64_t off = 0;
    for (int k = 0; k < nw; k++) {
        RefWorld* R = (RefWorld*)hs[k];
        rp3dref_get_bodies(R, pos ? pos + 3 * off : nullptr, q ? q + 4 * off : nullptr, v ? v + 3 * off : nullptr,
                           w ? w + 3 * off : nullptr);
        off += (int64_t)R->bodies.size();
    }
    return (int)off;
}

int64_t rp3dref_batch_get_pairs(void** hs, int nw, const int32_t* collider_base, int32_t* c1, int32_t* c2, int64_t cap) {
    int64_t tot = 0;
    for (int k = 0; k < nw; k++) {
        auto& OP = ((RefWorld*)hs[k])->world->mCollisionDetection.mOverlappingPairs;
        tot += (int64_t)OP.mConvexPairs.size() + (int64_t)OP.mDisabledConvexPairs.size();
    }
    if (tot > cap) return -tot;
    int64_t o = 0;
    for (int k = 0; k < nw; k++) {
        RefWorld* R = (RefWorld*)hs[k];
        const int n = rp3dref_get_pairs(R, c1 + o, c2 + o, (int)(cap - o));
        for (int i = 0; i < n; i++) { c1[o + i] += collider_base[k]; c2[o + i] += collider_base[k]; }
        o += n;
    }
    return o;
}

int64_t rp3dref_batch_get_contacts(void** hs, int nw, const int32_t* collider_base, int32_t* cols, int32_t* npts,
                                   float* nd, float* l1, float* l2, float* imp, float* fric, int64_t cap) {
    int64_t tot = 0;
    for (int k = 0; k < nw; k++) tot += (int64_t)((RefWorld*)hs[k])->world->mCollisionDetection.mCurrentContactManifolds->size();
    if (tot > cap) return -tot;
    int64_t o = 0;
    for (int k = 0; k < nw; k++) {
        RefWorld* R = (RefWorld*)hs[k];
        const int n = rp3dref_get_contacts(R, cols + 2 * o, npts + o, nd + 16 * o, l1 + 12 * o, l2 + 12 * o, imp + 4 * o,
                                           fric + 3 * o, (int)(cap - o));
        for (int i = 0; i < 2 * n; i++) cols[2 * o + i] += collider_base[k];
        o += n;
    }
    return o;
}

// PhysicsCommon::createConvexMesh(const VertexArray&, messages) (src/engine/PhysicsCommon.cpp:638-660, QuickHull
// then ConvexMesh::init): the mesh's vertices and its faces' vertex lists.  points: n x 3 floats.  Returns 0
// and the counts, -1 when the reference refuses the points, -2 when a capacity is short.
int rp3dref_quickhull(const float* points, int n, float* verts, int cap_v, int32_t* counts, int cap_f, int32_t* idx,
                      int cap_i, int32_t* nv, int32_t* nf) {
    PhysicsCommon common;
    VertexArray va(points, 3 * sizeof(float), (uint32)n, VertexArray::DataType::VERTEX_FLOAT_TYPE);
    std::vector<Message> msgs;
    ConvexMesh* mesh = common.createConvexMesh(va, msgs);
    if (!mesh) return -1;
    const HalfEdgeStructure& H = mesh->getHalfEdgeStructure();
    *nv = (int32_t)mesh->getNbVertices();
    *nf = (int32_t)mesh->getNbFaces();
    int ni = 0;
    for (uint32 f = 0; f < mesh->getNbFaces(); f++) ni += (int)H.getFace(f).faceVertices.size();
    if (*nv > cap_v || *nf > cap_f || ni > cap_i) return -2;
    for (uint32 i = 0; i < mesh->getNbVertices(); i++) {
        const Vector3& p = mesh->getVertex(i);
        verts[3 * i] = (float)p.x; verts[3 * i + 1] = (float)p.y; verts[3 * i + 2] = (float)p.z;
    }
    int o = 0;
    for (uint32 f = 0; f < mesh->getNbFaces(); f++) {
        const auto& fv = H.getFace(f).faceVertices;
        counts[f] = (int32_t)fv.size();
        for (uint32 k = 0; k < fv.size(); k++) idx[o++] = (int32_t)fv[k];
    }
    return 0;
}

// CPU baseline: `threads` identical reference worlds, one per thread, each stepped `warmup` untimed
// then `steps` timed steps of PhysicsWorld::update(dt).  Worlds are created serially (the static
// PhysicsWorld::mNbWorlds counter, src/engine/PhysicsWorld.cpp:46,108, is not atomic).  Wall time is
// measured from a start barrier to the last join.  out[0] = seconds, out[1] = contact points summed
// over the timed steps of all worlds.
int rp3dref_bench(const RefSettings* s, int nb, const RefBody* bodies, int nc, const RefCollider* cols, int nj,
                  const RefJoint* joints, const float* hull_vertices, const int32_t* hull_offsets, int threads,
                  int warmup, int steps, float dt, double* out) {
    std::vector<RefWorld*> ws(threads, nullptr);
    for (int t = 0; t < threads; t++) {
        ws[t] = (RefWorld*)rp3dref_create(s, nb, bodies, nc, cols, nj, joints, hull_vertices, hull_offsets);
        if (!ws[t]) {
            for (RefWorld* w : ws) rp3dref_destroy(w);
            return -1;
        }
        for (int k = 0; k < warmup; k++) ws[t]->world->update(dt);
    }
    std::atomic<int> ready(0);
    std::atomic<bool> go(false);
    std::vector<double> contacts(threads, 0.0);
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        pool.emplace_back([&, t]() {
            ready.fetch_add(1);
            while (!go.load()) std::this_thread::yield();
            for (int k = 0; k < steps; k++) {
                ws[t]->world->update(dt);
                contacts[t] += (double)ws[t]->world->mCollisionDetection.mCurrentContactPoints->size();
            }
        });
    }
    while (ready.load() < threads) std::this_thread::yield();
    const auto t0 = std::chrono::steady_clock::now();
    go.store(true);
    for (auto& th : pool) th.join();
    const auto t1 = std::chrono::steady_clock::now();
    out[0] = std::chrono::duration<double>(t1 - t0).count();
    out[1] = 0.0;
    for (double c : contacts) out[1] += c;
    for (RefWorld* w : ws) rp3dref_destroy(w);
    return 0;
}

}  // extern "C"
