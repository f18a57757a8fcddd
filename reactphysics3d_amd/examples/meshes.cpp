// Convex meshes through the C++ mirror of the reference's API, host side only (no step):
// PolygonVertexArray with strided double vertices and short indices, VertexArray + QuickHull, and a
// mesh that createConvexMesh refuses with a message.  Prints what it built.
// Usage: meshes
#include <cstdio>

#include "reactphysics3d_mi355x.hpp"

using namespace reactphysics3d;

int main() {
    try {
        PhysicsCommon common;
        std::vector<Message> messages;

        // a 2 x 1 x 1 box: vertices as doubles in a 4-double stride, indices as shorts in a 2-short stride
        const double v[8][4] = {{-1, -0.5, -0.5, 9}, {1, -0.5, -0.5, 9}, {1, 0.5, -0.5, 9}, {-1, 0.5, -0.5, 9},
                                {-1, -0.5, 0.5, 9},  {1, -0.5, 0.5, 9},  {1, 0.5, 0.5, 9},  {-1, 0.5, 0.5, 9}};
        const short quads[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6}, {0, 4, 7, 3}};
        short idx[48];
        for (int f = 0; f < 6; f++)
            for (int k = 0; k < 4; k++) { idx[2 * (4 * f + k)] = quads[f][k]; idx[2 * (4 * f + k) + 1] = -1; }
        PolygonVertexArray::PolygonFace faces[6];
        for (int f = 0; f < 6; f++) faces[f] = {4, (uint32)(4 * f)};
        PolygonVertexArray box(8, v, 4 * sizeof(double), idx, 2 * sizeof(short), 6, faces,
                               PolygonVertexArray::VertexDataType::VERTEX_DOUBLE_TYPE,
                               PolygonVertexArray::IndexDataType::INDEX_SHORT_TYPE);
        ConvexMesh* boxMesh = common.createConvexMesh(box, messages);
        std::printf("box %d %u %u\n", boxMesh ? 1 : 0, boxMesh ? boxMesh->getNbVertices() : 0,
                    boxMesh ? boxMesh->getNbFaces() : 0);

        // QuickHull of a point cloud: the corners of a unit cube, the centre and points on the faces
        float cloud[14][4];
        int n = 0;
        for (int i = 0; i < 8; i++) {
            cloud[n][0] = (i & 1) ? 1.f : -1.f; cloud[n][1] = (i & 2) ? 1.f : -1.f; cloud[n][2] = (i & 4) ? 1.f : -1.f;
            cloud[n][3] = 0.f;
            n++;
        }
        const float extra[6][3] = {{0, 0, 0}, {0.5f, 0.25f, 1}, {-0.25f, 1, 0.5f}, {1, 0, 0}, {0.3f, -1, 0.1f}, {0, 0, -1}};
        for (int i = 0; i < 6; i++, n++)
            for (int k = 0; k < 4; k++) cloud[n][k] = k < 3 ? extra[i][k] : 0.f;
        VertexArray points(cloud, 4 * sizeof(float), (uint32)n, VertexArray::DataType::VERTEX_FLOAT_TYPE);
        ConvexMesh* hull = common.createConvexMesh(points, messages);
        std::printf("hull %d %u %u\n", hull ? 1 : 0, hull ? hull->getNbVertices() : 0, hull ? hull->getNbFaces() : 0);
        if (hull) {
            for (uint32 f = 0; f < hull->getNbFaces(); f++) {
                const PolygonVertexArray::PolygonFace* pf = hull->mArray.getPolygonFace(f);
                std::printf("face %u", f);
                for (uint32 k = 0; k < pf->nbVertices; k++) {
                    const Vector3 p = hull->getVertex(hull->mArray.getVertexIndexInFace(f, k));
                    std::printf(" %.9g %.9g %.9g", p.x, p.y, p.z);
                }
                std::printf("\n");
            }
        }

        // an open mesh (a tetrahedron without its base) is refused with a message
        const float tv[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        const int ti[9] = {0, 2, 1, 0, 1, 3, 0, 3, 2};
        PolygonVertexArray::PolygonFace tf[3] = {{3, 0}, {3, 3}, {3, 6}};
        PolygonVertexArray open(4, tv, 3 * sizeof(float), ti, sizeof(int), 3, tf,
                                PolygonVertexArray::VertexDataType::VERTEX_FLOAT_TYPE,
                                PolygonVertexArray::IndexDataType::INDEX_INTEGER_TYPE);
        ConvexMesh* bad = common.createConvexMesh(open, messages);
        std::printf("open %d %zu\n", bad ? 1 : 0, messages.size());
        for (const Message& m : messages) std::printf("message %d %s\n", (int)m.type, m.text.c_str());

        // the meshes go into a world as colliders
        PhysicsWorld* world = common.createPhysicsWorld();
        RigidBody* body = world->createRigidBody(Transform(Vector3(0, 0, 3), Quaternion::identity()));
        Collider* c1 = body->addCollider(common.createConvexMeshShape(boxMesh), Transform::identity());
        Collider* c2 = body->addCollider(common.createConvexMeshShape(hull), Transform(Vector3(0, 0, 2), Quaternion::identity()));
        std::printf("colliders %d %d\n", c1->index(), c2->index());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
