// An application that drives its bodies between PhysicsWorld::update calls through the C++ mirror of
// the reference's RigidBody API: a box pushed at a point of its own frame, spun by a torque and put to
// sleep, a ball kicked, teleported and given a spin, every call between two steps.  Prints the state of
// every body after the last step, and the force and torque sums of the box after the calls of the
// frame given as second argument (before that frame's step).
// Usage: driven [steps] [probe_frame]
#include <cstdio>
#include <cstdlib>

#include "reactphysics3d_mi355x.hpp"

using namespace reactphysics3d;

int main(int argc, char** argv) {
    const int steps = argc > 1 ? std::atoi(argv[1]) : 150;
    const int probe = argc > 2 ? std::atoi(argv[2]) : 40;
    try {
        PhysicsCommon common;
        PhysicsWorld::WorldSettings settings;
        settings.gravity = Vector3(0, 0, decimal(-9.81));
        settings.defaultVelocitySolverNbIterations = 10;
        settings.defaultPositionSolverNbIterations = 5;
        PhysicsWorld* world = common.createPhysicsWorld(settings);

        RigidBody* ground = world->createRigidBody(Transform(Vector3(0, 0, decimal(-0.5)), Quaternion::identity()));
        ground->setType(BodyType::STATIC);
        Collider* g = ground->addCollider(common.createBoxShape(Vector3(20, 20, decimal(0.5))), Transform::identity());
        g->getMaterial().setFrictionCoefficient(decimal(0.5));
        g->getMaterial().setBounciness(0);

        RigidBody* box = world->createRigidBody(Transform(Vector3(0, 0, decimal(0.6)), Quaternion::identity()));
        Collider* b = box->addCollider(common.createBoxShape(Vector3(decimal(0.5), decimal(0.5), decimal(0.5))),
                                       Transform::identity());
        b->getMaterial().setFrictionCoefficient(decimal(0.5));
        b->getMaterial().setBounciness(0);
        box->updateMassPropertiesFromColliders();

        RigidBody* ball = world->createRigidBody(Transform(Vector3(6, 0, decimal(0.5)), Quaternion::identity()));
        Collider* s = ball->addCollider(common.createSphereShape(decimal(0.4)), Transform::identity());
        s->getMaterial().setFrictionCoefficient(decimal(0.5));
        s->getMaterial().setBounciness(0);
        ball->updateMassPropertiesFromColliders();

        const decimal dt = decimal(1.0) / decimal(60.0);
        float probe_ft[6] = {0, 0, 0, 0, 0, 0};
        for (int i = 0; i < steps; i++) {
            if (i >= 20 && i < 40) box->applyLocalForceAtLocalPosition(Vector3(12, 0, 0), Vector3(0, decimal(0.3), decimal(0.4)));
            if (i == 30) box->applyWorldTorque(Vector3(0, 0, 3));
            if (i == 35) box->applyWorldForceAtWorldPosition(Vector3(0, 4, 0), Vector3(0, 0, 1));
            if (i == 38) box->resetTorque();
            if (i == 25) ball->setLinearVelocity(Vector3(-1, 2, 0));
            if (i == 50) ball->setTransform(Transform(Vector3(4, 4, 2), Quaternion(0, decimal(0.6), 0, decimal(0.8))));
            if (i == 51) ball->setAngularVelocity(Vector3(0, 5, 0));
            if (i == 52) ball->applyLocalForceAtCenterOfMass(Vector3(0, 0, 30));
            if (i == 53) ball->applyLocalTorque(Vector3(1, 0, 0));
            if (i == 60) box->setIsSleeping(true);
            if (i == 90) box->applyWorldForceAtCenterOfMass(Vector3(0, 0, 20));   // wakes it
            if (i == 91) box->applyLocalForceAtWorldPosition(Vector3(5, 0, 0), Vector3(0, 0, 0));
            if (i == 92) box->resetForce();
            if (i == probe) {
                const Vector3 f = box->getForce(), t = box->getTorque();
                probe_ft[0] = f.x; probe_ft[1] = f.y; probe_ft[2] = f.z;
                probe_ft[3] = t.x; probe_ft[4] = t.y; probe_ft[5] = t.z;
            }
            world->update(dt);
        }

        RigidBody* bodies[3] = {ground, box, ball};
        for (int i = 0; i < 3; i++) {
            const Transform t = bodies[i]->getTransform();
            const Vector3 v = bodies[i]->getLinearVelocity(), w = bodies[i]->getAngularVelocity();
            std::printf("body %d %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", i,
                        t.getPosition().x, t.getPosition().y, t.getPosition().z, t.getOrientation().x,
                        t.getOrientation().y, t.getOrientation().z, t.getOrientation().w, v.x, v.y, v.z, w.x, w.y,
                        w.z);
        }
        std::printf("probe %d %.9g %.9g %.9g %.9g %.9g %.9g\n", probe, probe_ft[0], probe_ft[1], probe_ft[2],
                    probe_ft[3], probe_ft[4], probe_ft[5]);
        std::printf("sleeping %d %d\n", box->isSleeping() ? 1 : 0, ball->isSleeping() ? 1 : 0);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
