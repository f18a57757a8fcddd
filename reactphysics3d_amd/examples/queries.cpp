// World queries through the C++ mirror of the reference's API: three boxes dropped on a ground, stepped
// with PhysicsWorld::update, asked testOverlap / testCollision and raycasts (closest hit and every hit)
// between steps, then stepped on.  Prints each answer and the final body states.
// Usage: queries [steps_before] [steps_after]
#include <cstdio>
#include <cstdlib>

#include "reactphysics3d_mi355x.hpp"

using namespace reactphysics3d;

struct Overlaps : public OverlapCallback {
    void onOverlap(CallbackData& data) override {
        for (uint32 i = 0; i < data.getNbOverlappingPairs(); i++) {
            const OverlapPair p = data.getOverlappingPair(i);
            std::printf("overlap %d %d %d\n", p.getCollider1()->index(), p.getCollider2()->index(), (int)p.getEventType());
        }
    }
};

struct Contacts : public CollisionCallback {
    void onContact(const CallbackData& data) override {
        for (uint32 i = 0; i < data.getNbContactPairs(); i++) {
            const ContactPair p = data.getContactPair(i);
            std::printf("contact %d %d %d %u", p.getCollider1()->index(), p.getCollider2()->index(), (int)p.getEventType(),
                        p.getNbContactPoints());
            for (uint32 k = 0; k < p.getNbContactPoints(); k++) {
                const ContactPoint c = p.getContactPoint(k);
                std::printf(" %.9g %.9g %.9g %.9g", c.getWorldNormal().x, c.getWorldNormal().y, c.getWorldNormal().z,
                            c.getPenetrationDepth());
            }
            std::printf("\n");
        }
    }
};

struct Closest : public RaycastCallback {
    decimal notifyRaycastHit(const RaycastInfo& info) override {
        std::printf("closest %d %.9g\n", info.collider->index(), info.hitFraction);
        return info.hitFraction;
    }
};

struct Every : public RaycastCallback {
    decimal notifyRaycastHit(const RaycastInfo& info) override {
        std::printf("hit %d %.9g\n", info.collider->index(), info.hitFraction);
        return decimal(1.0);
    }
};

int main(int argc, char** argv) {
    const int before = argc > 1 ? std::atoi(argv[1]) : 60;
    const int after = argc > 2 ? std::atoi(argv[2]) : 30;
    try {
        PhysicsCommon common;
        PhysicsWorld::WorldSettings settings;
        settings.gravity = Vector3(0, 0, decimal(-9.81));
        PhysicsWorld* world = common.createPhysicsWorld(settings);
        RigidBody* ground = world->createRigidBody(Transform(Vector3(0, 0, decimal(-0.5)), Quaternion::identity()));
        ground->setType(BodyType::STATIC);
        ground->addCollider(common.createBoxShape(Vector3(10, 10, decimal(0.5))), Transform::identity());
        RigidBody* boxes[3];
        for (int i = 0; i < 3; i++) {
            boxes[i] = world->createRigidBody(Transform(Vector3(decimal(3 * i - 3), 0, decimal(1 + i)), Quaternion::identity()));
            boxes[i]->addCollider(common.createBoxShape(Vector3(decimal(0.5), decimal(0.5), decimal(0.5))), Transform::identity());
            boxes[i]->updateMassPropertiesFromColliders();
        }
        const decimal dt = decimal(1.0) / decimal(60.0);
        for (int i = 0; i < before; i++) world->update(dt);

        std::printf("touching %d %d\n", world->testOverlap(boxes[0], ground) ? 1 : 0,
                    world->testOverlap(boxes[0], boxes[1]) ? 1 : 0);
        Overlaps overlaps;
        world->testOverlap(boxes[1], overlaps);
        world->testOverlap(overlaps);
        Contacts contacts;
        world->testCollision(contacts);
        world->testCollision(boxes[2], ground, contacts);
        Closest closest;
        Every every;
        world->raycast(Ray(Vector3(-8, 0, decimal(0.3)), Vector3(8, 0, decimal(0.3))), &closest);
        world->raycast(Ray(Vector3(-8, 0, decimal(0.3)), Vector3(8, 0, decimal(0.3))), &every);

        for (int i = 0; i < after; i++) world->update(dt);
        RigidBody* bodies[4] = {ground, boxes[0], boxes[1], boxes[2]};
        for (int i = 0; i < 4; i++) {
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
