// Concave shapes through the C++ mirror of the reference API: a triangle-mesh terrain
// (PhysicsCommon::createTriangleMesh from a TriangleVertexArray, createConcaveMeshShape) and a height field
// (createHeightField, createHeightFieldShape, turned so that its y axis points up the world's z) on static
// bodies, with spheres, boxes and a capsule dropped on them.  Prints the final body states.
// Usage: concave [steps]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "reactphysics3d_mi355x.hpp"

using namespace reactphysics3d;

int main(int argc, char** argv) {
    const int steps = argc > 1 ? std::atoi(argv[1]) : 150;
    try {
        PhysicsCommon common;
        PhysicsWorld::WorldSettings settings;
        settings.gravity = Vector3(0, 0, decimal(-9.81));
        PhysicsWorld* world = common.createPhysicsWorld(settings);
        std::vector<Message> messages;

        // 8 x 8 cells, heights 0.1 * ((7i + 3j) mod 5) - 0.2, two triangles per cell
        const int n = 8;
        std::vector<float> v;
        for (int j = 0; j <= n; j++)
            for (int i = 0; i <= n; i++) {
                v.push_back((float)(i - n / 2));
                v.push_back((float)(j - n / 2));
                v.push_back(0.1f * (float)((7 * i + 3 * j) % 5) - 0.2f);
            }
        std::vector<int> t;
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                const int a = j * (n + 1) + i, b = a + 1, c = a + n + 2, d = a + n + 1;
                if ((i + j) % 2 == 0) { t.insert(t.end(), {a, b, c, a, c, d}); }
                else { t.insert(t.end(), {a, b, d, b, c, d}); }
            }
        TriangleVertexArray tva((uint32)(v.size() / 3), v.data(), 3 * sizeof(float), (uint32)(t.size() / 3), t.data(),
                                3 * sizeof(int), TriangleVertexArray::VertexDataType::VERTEX_FLOAT_TYPE,
                                TriangleVertexArray::IndexDataType::INDEX_INTEGER_TYPE);
        TriangleMesh* mesh = common.createTriangleMesh(tva, messages);
        RigidBody* terrain = world->createRigidBody(Transform::identity());
        terrain->setType(BodyType::STATIC);
        terrain->addCollider(common.createConcaveMeshShape(mesh), Transform::identity());

        // 9 x 9 height field, 0.05 * ((x + 2y) mod 4)
        std::vector<float> h;
        for (int y = 0; y < 9; y++)
            for (int x = 0; x < 9; x++) h.push_back(0.05f * (float)((x + 2 * y) % 4));
        HeightField* field = common.createHeightField(9, 9, h.data(), HeightField::HeightDataType::HEIGHT_FLOAT_TYPE, messages);
        RigidBody* hill = world->createRigidBody(Transform(Vector3(12, 0, 0), Quaternion::identity()));
        hill->setType(BodyType::STATIC);
        hill->addCollider(common.createHeightFieldShape(field, Vector3(1, decimal(1.5), 1)),
                          Transform(Vector3(0, 0, 0), Quaternion(decimal(0.70710677), 0, 0, decimal(0.70710677))));

        std::vector<RigidBody*> bodies = {terrain, hill};
        const float drops[6][3] = {{-2, -1, 2}, {1, 2, 3}, {2, -2, 2.5f}, {-1, 1.5f, 4}, {11, 1, 2}, {13, -1, 3}};
        for (int k = 0; k < 6; k++) {
            RigidBody* b = world->createRigidBody(Transform(Vector3(drops[k][0], drops[k][1], drops[k][2]), Quaternion::identity()));
            if (k < 2) b->addCollider(common.createSphereShape(decimal(0.4)), Transform::identity());
            else if (k == 3) b->addCollider(common.createCapsuleShape(decimal(0.25), decimal(0.8)), Transform::identity());
            else b->addCollider(common.createBoxShape(Vector3(decimal(0.4), decimal(0.3), decimal(0.35))), Transform::identity());
            b->updateMassPropertiesFromColliders();
            bodies.push_back(b);
        }
        const decimal dt = decimal(1.0) / decimal(60.0);
        for (int i = 0; i < steps; i++) world->update(dt);
        for (size_t i = 0; i < bodies.size(); i++) {
            const Transform tr = bodies[i]->getTransform();
            const Vector3 lv = bodies[i]->getLinearVelocity(), av = bodies[i]->getAngularVelocity();
            std::printf("body %zu %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", i,
                        tr.getPosition().x, tr.getPosition().y, tr.getPosition().z, tr.getOrientation().x,
                        tr.getOrientation().y, tr.getOrientation().z, tr.getOrientation().w, lv.x, lv.y, lv.z, av.x,
                        av.y, av.z);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
