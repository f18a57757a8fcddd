// A helloworld-style program written against the C++ mirror of the reference's API (the reference's
// own helloworld/Main.cpp steps one body with no collider; this one gives its bodies colliders so that
// the step has contacts): a box and a sphere dropped on a static ground, stepped with
// PhysicsWorld::update on the GPU.
// Usage: hello_world [steps]   (steps = 0 builds the world and prints the creation state only)
#include <cstdio>
#include <cstdlib>

#include "reactphysics3d_mi355x.hpp"

using namespace reactphysics3d;

int main(int argc, char** argv) {
    const int steps = argc > 1 ? std::atoi(argv[1]) : 120;
    try {
        PhysicsCommon common;
        PhysicsWorld::WorldSettings settings;
        settings.gravity = Vector3(0, 0, decimal(-9.81));
        settings.defaultVelocitySolverNbIterations = 10;
        settings.defaultPositionSolverNbIterations = 5;
        PhysicsWorld* world = common.createPhysicsWorld(settings);

        RigidBody* ground = world->createRigidBody(Transform(Vector3(0, 0, decimal(-0.5)), Quaternion::identity()));
        ground->setType(BodyType::STATIC);
        Collider* g = ground->addCollider(common.createBoxShape(Vector3(10, 10, decimal(0.5))), Transform::identity());
        g->getMaterial().setFrictionCoefficient(decimal(0.5));
        g->getMaterial().setBounciness(0);

        RigidBody* box = world->createRigidBody(
            Transform(Vector3(0, 0, 2), Quaternion(decimal(0.0998334), 0, 0, decimal(0.9950042))));
        Collider* b = box->addCollider(common.createBoxShape(Vector3(decimal(0.5), decimal(0.5), decimal(0.5))),
                                       Transform::identity());
        b->getMaterial().setFrictionCoefficient(decimal(0.5));
        b->getMaterial().setBounciness(0);
        b->getMaterial().setMassDensity(1);
        box->updateMassPropertiesFromColliders();

        RigidBody* ball = world->createRigidBody(Transform(Vector3(decimal(3.5), 0, 3), Quaternion::identity()));
        Collider* s = ball->addCollider(common.createSphereShape(decimal(0.4)), Transform::identity());
        s->getMaterial().setFrictionCoefficient(decimal(0.5));
        s->getMaterial().setBounciness(0);
        s->getMaterial().setMassDensity(1);
        ball->updateMassPropertiesFromColliders();
        ball->setLinearVelocity(Vector3(-1, 0, 0));

        const decimal dt = decimal(1.0) / decimal(60.0);
        for (int i = 0; i < steps; i++) world->update(dt);

        RigidBody* bodies[3] = {ground, box, ball};
        for (int i = 0; i < 3; i++) {
            const Transform t = bodies[i]->getTransform();
            const Vector3 v = bodies[i]->getLinearVelocity(), w = bodies[i]->getAngularVelocity();
            std::printf("body %d %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", i,
                        t.getPosition().x, t.getPosition().y, t.getPosition().z, t.getOrientation().x,
                        t.getOrientation().y, t.getOrientation().z, t.getOrientation().w, v.x, v.y, v.z, w.x, w.y,
                        w.z);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
