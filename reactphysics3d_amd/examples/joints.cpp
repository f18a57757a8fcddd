// Joints through the C++ mirror of the reference API, set up the reference's way: a slider with a limit
// and a motor from SliderJointInfo fields, a hinge whose limits are switched on through the joint after
// createJoint (HingeJoint::enableLimit / setMinAngleLimit / setMaxAngleLimit), and a ball-and-socket
// pendulum with a cone limit (BallAndSocketJoint::enableConeLimit / setConeLimitHalfAngle).
// Usage: joints [steps] [baumgarte]   (steps = 0 builds the world and prints the creation state only;
// baumgarte = 1 gives every joint JointsPositionCorrectionTechnique::BAUMGARTE_JOINTS)
#include <cstdio>
#include <cstdlib>

#include "reactphysics3d_mi355x.hpp"

using namespace reactphysics3d;

static RigidBody* box_body(PhysicsCommon& common, PhysicsWorld* world, const Vector3& p, const Vector3& half) {
    RigidBody* b = world->createRigidBody(Transform(p, Quaternion::identity()));
    b->addCollider(common.createBoxShape(half), Transform::identity());
    b->updateMassPropertiesFromColliders();
    return b;
}

int main(int argc, char** argv) {
    const int steps = argc > 1 ? std::atoi(argv[1]) : 120;
    const JointsPositionCorrectionTechnique technique = argc > 2 && std::atoi(argv[2]) != 0
                                                            ? JointsPositionCorrectionTechnique::BAUMGARTE_JOINTS
                                                            : JointsPositionCorrectionTechnique::NON_LINEAR_GAUSS_SEIDEL;
    try {
        PhysicsCommon common;
        PhysicsWorld::WorldSettings settings;
        settings.gravity = Vector3(0, 0, decimal(-9.81));
        settings.defaultVelocitySolverNbIterations = 10;
        settings.defaultPositionSolverNbIterations = 5;
        PhysicsWorld* world = common.createPhysicsWorld(settings);

        // slider: a carriage below a static post, sliding on a tilted axis, motor against gravity
        RigidBody* post = world->createRigidBody(Transform(Vector3(0, 0, 5), Quaternion::identity()));
        post->setType(BodyType::STATIC);
        RigidBody* car = box_body(common, world, Vector3(0, 0, decimal(4.2)), Vector3(decimal(0.3), decimal(0.3), decimal(0.3)));
        SliderJointInfo si(post, car, Vector3(0, 0, decimal(4.2)), Vector3(decimal(0.6), 0, decimal(0.8)));
        si.isCollisionEnabled = false;
        si.positionCorrectionTechnique = technique;
        si.isLimitEnabled = true;
        si.minTranslationLimit = decimal(-0.4);
        si.maxTranslationLimit = decimal(0.6);
        si.isMotorEnabled = true;
        si.motorSpeed = decimal(0.5);
        si.maxMotorForce = decimal(1.0);
        world->createJoint(si);

        // hinge: a paddle on a horizontal axis, limits set on the joint after creation
        RigidBody* pivot = world->createRigidBody(Transform(Vector3(4, 0, 5), Quaternion::identity()));
        pivot->setType(BodyType::STATIC);
        RigidBody* paddle = box_body(common, world, Vector3(decimal(4.6), 0, 5), Vector3(decimal(0.5), decimal(0.1), decimal(0.3)));
        HingeJointInfo hi(pivot, paddle, Vector3(decimal(4.05), 0, 5), Vector3(0, 1, 0));
        hi.isCollisionEnabled = false;
        hi.positionCorrectionTechnique = technique;
        HingeJoint* hinge = static_cast<HingeJoint*>(world->createJoint(hi));
        hinge->enableLimit(true);
        hinge->setMinAngleLimit(decimal(-0.6));
        hinge->setMaxAngleLimit(decimal(0.4));

        // ball-and-socket: a sphere hanging sideways from a static body, held by a 0.3 rad cone
        RigidBody* hook = world->createRigidBody(Transform(Vector3(8, 0, 5), Quaternion::identity()));
        hook->setType(BodyType::STATIC);
        RigidBody* ball = world->createRigidBody(Transform(Vector3(9, 0, 5), Quaternion::identity()));
        ball->addCollider(common.createSphereShape(decimal(0.2)), Transform::identity());
        ball->updateMassPropertiesFromColliders();
        BallAndSocketJointInfo bi(hook, ball, Vector3(decimal(8.5), 0, 5));
        bi.isCollisionEnabled = false;
        bi.positionCorrectionTechnique = technique;
        BallAndSocketJoint* cone = static_cast<BallAndSocketJoint*>(world->createJoint(bi));
        cone->enableConeLimit(true);
        cone->setConeLimitHalfAngle(decimal(0.3));

        const decimal dt = decimal(1.0) / decimal(60.0);
        for (int i = 0; i < steps; i++) world->update(dt);

        RigidBody* bodies[6] = {post, car, pivot, paddle, hook, ball};
        for (int i = 0; i < 6; i++) {
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
