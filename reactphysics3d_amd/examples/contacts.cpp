// Contact events through the C++ mirror, the reference's way: an EventListener subclass handed to
// PhysicsWorld::setEventListener receives onContact after every update().  Three boxes dropped apart
// on a static ground bounce (default bounciness 0.5) and settle; every step prints the number of
// ContactStart / ContactStay / ContactExit pairs and of contact points.
// Usage: contacts [steps]
#include <cstdio>
#include <cstdlib>

#include "reactphysics3d_mi355x.hpp"

using namespace reactphysics3d;

class Listener : public EventListener {
   public:
    int start = 0, stay = 0, exit_ = 0, points = 0;
    void onContact(const CollisionCallback::CallbackData& data) override {
        for (uint32 i = 0; i < data.getNbContactPairs(); i++) {
            const CollisionCallback::ContactPair pair = data.getContactPair(i);
            switch (pair.getEventType()) {
                case CollisionCallback::ContactPair::EventType::ContactStart: start++; break;
                case CollisionCallback::ContactPair::EventType::ContactStay: stay++; break;
                case CollisionCallback::ContactPair::EventType::ContactExit: exit_++; break;
            }
            points += (int)pair.getNbContactPoints();
        }
    }
};

int main(int argc, char** argv) {
    const int steps = argc > 1 ? std::atoi(argv[1]) : 120;
    try {
        PhysicsCommon common;
        PhysicsWorld::WorldSettings settings;
        settings.gravity = Vector3(0, 0, decimal(-9.81));
        PhysicsWorld* world = common.createPhysicsWorld(settings);
        RigidBody* ground = world->createRigidBody(Transform(Vector3(0, 0, decimal(-0.5)), Quaternion::identity()));
        ground->setType(BodyType::STATIC);
        ground->addCollider(common.createBoxShape(Vector3(10, 10, decimal(0.5))), Transform::identity());
        const Vector3 at[3] = {Vector3(0, 0, 1), Vector3(3, 0, decimal(2.5)), Vector3(-3, decimal(0.5), 4)};
        for (const Vector3& p : at) {
            RigidBody* b = world->createRigidBody(Transform(p, Quaternion::identity()));
            b->addCollider(common.createBoxShape(Vector3(decimal(0.5), decimal(0.5), decimal(0.5))), Transform::identity());
            b->updateMassPropertiesFromColliders();
        }
        // PhysicsWorld::raycast with a closest-hit callback, before and after the boxes fell
        struct Closest : public RaycastCallback {
            decimal fraction = -1;
            decimal notifyRaycastHit(const RaycastInfo& info) override {
                fraction = info.hitFraction;
                return info.hitFraction;
            }
        };
        Listener listener;
        world->setEventListener(&listener);
        const decimal dt = decimal(1.0) / decimal(60.0);
        for (int i = 0; i < steps; i++) {
            listener = Listener();
            world->update(dt);
            std::printf("step %d %d %d %d %d\n", i, listener.start, listener.stay, listener.exit_, listener.points);
        }
        Closest hit;
        world->raycast(Ray(Vector3(0, 0, 10), Vector3(0, 0, -10)), &hit);
        std::printf("ray %.9g\n", hit.fraction);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
