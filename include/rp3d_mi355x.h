/*
 * rp3d_mi355x — C ABI of the MI355X-native PhysicsWorld::update() step path.
 *
 * The ABI is the drop-in boundary: the C++ facade in
 * include/reactphysics3d_mi355x.hpp (namespace reactphysics3d,
 * PhysicsCommon / PhysicsWorld / RigidBody / Collider / shapes / joints) and
 * the Python tests and bench call only these functions.  Plain pointers,
 * sizes and POD structs cross it; no C++ or torch types.
 *
 * Each entry point names the reference interface it replaces
 * (reference = ReactPhysics3D 0.10.2, paths relative to its repository).
 *
 * Status convention: functions returning int return RP3DG_OK (0) on success
 * and a negative RP3DG_E* code otherwise; rp3dg_last_error() returns the
 * message of the last failure on the calling thread.  Object-creating
 * functions return the new index (>= 0) or a negative code.
 * Threading: one world / batch handle per thread; handles are not
 * internally synchronised (the reference's PhysicsWorld is not either).
 */
#ifndef RP3D_MI355X_H
#define RP3D_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RP3DG_OK 0
#define RP3DG_EINVAL (-1)      /* invalid argument / index */
#define RP3DG_EHIP (-2)        /* HIP runtime error */
#define RP3DG_EOVERFLOW (-3)   /* device arena overflow that could not be grown */
#define RP3DG_ESTATE (-4)      /* call not allowed in the current state */
#define RP3DG_EUNSUPPORTED (-5)

/* Body types: same order as reactphysics3d::BodyType (include/reactphysics3d/body/RigidBody.h). */
#define RP3DG_STATIC 0
#define RP3DG_KINEMATIC 1
#define RP3DG_DYNAMIC 2

/* Collision shape kinds (reactphysics3d::CollisionShapeName, CollisionShape.h). */
#define RP3DG_SHAPE_SPHERE 0
#define RP3DG_SHAPE_CAPSULE 1
#define RP3DG_SHAPE_BOX 2
#define RP3DG_SHAPE_CONVEX_MESH 3
#define RP3DG_SHAPE_CONCAVE_MESH 4   /* params = {mesh index, scale x, y, z} (ConcaveMeshShape) */
#define RP3DG_SHAPE_HEIGHTFIELD 5    /* params = {height field index, scale x, y, z} (HeightFieldShape) */

/* Joint kinds (reactphysics3d::JointType, constraint/Joint.h). */
#define RP3DG_JOINT_BALLSOCKET 0
#define RP3DG_JOINT_SLIDER 1
#define RP3DG_JOINT_HINGE 2
#define RP3DG_JOINT_FIXED 3
/* rp3dg_joint_create flags bit2: JointInfo::positionCorrectionTechnique = BAUMGARTE_JOINTS
 * (include/reactphysics3d/constraint/Joint.h:66, configuration.h:101); clear: NON_LINEAR_GAUSS_SEIDEL */
#define RP3DG_JOINT_BAUMGARTE 4

typedef struct rp3dg_world rp3dg_world;
typedef struct rp3dg_batch rp3dg_batch;

/* Mirrors PhysicsWorld::WorldSettings (include/reactphysics3d/engine/PhysicsWorld.h:77-137). */
typedef struct rp3dg_world_settings {
    float gravity[3];                            /* default (0, -9.81, 0) */
    float persistent_contact_distance_threshold; /* 0.03 */
    float default_friction_coefficient;          /* 0.3 */
    float default_bounciness;                    /* 0.5 */
    float restitution_velocity_threshold;        /* 0.5 */
    float cos_angle_similar_contact_manifold;    /* 0.95 */
    uint32_t default_velocity_solver_iterations; /* 6 */
    uint32_t default_position_solver_iterations; /* 3 */
    int32_t is_sleeping_enabled;                 /* 1: islands of still bodies fall asleep */
    int32_t is_gravity_enabled;                  /* 1 */
    float time_before_sleep;                     /* 1.0 s */
    float sleep_linear_velocity;                 /* 0.02 m/s */
    float sleep_angular_velocity;                /* 3 * (pi / 180) rad/s */
} rp3dg_world_settings;

/* ---- world construction (host side, replaces PhysicsCommon / PhysicsWorld factories) ---- */

/* WorldSettings() defaults (PhysicsWorld.h:122-137). */
void rp3dg_world_settings_default(rp3dg_world_settings* out);
/* PhysicsCommon::createPhysicsWorld (src/engine/PhysicsCommon.cpp:221-240). */
rp3dg_world* rp3dg_world_create(const rp3dg_world_settings* settings);
/* PhysicsCommon::destroyPhysicsWorld (src/engine/PhysicsCommon.cpp:243). */
void rp3dg_world_destroy(rp3dg_world* world);
/* PhysicsWorld::setNbIterationsVelocitySolver / PositionSolver (src/engine/PhysicsWorld.cpp:663,957). */
int rp3dg_world_set_iterations(rp3dg_world* world, uint32_t velocity_iterations, uint32_t position_iterations);
/* PhysicsWorld::setGravity (src/engine/PhysicsWorld.cpp:969). */
int rp3dg_world_set_gravity(rp3dg_world* world, const float gravity[3]);
/* PhysicsWorld::setContactsPositionCorrectionTechnique (PhysicsWorld.h:599-607,
 * ContactSolverSystem.cpp:486-590, DynamicsSystem.cpp:44-71): RP3DG_CONTACTS_SPLIT_IMPULSES (the default)
 * or RP3DG_CONTACTS_BAUMGARTE (the penetration bias joins the normal impulse; no split velocities).  Any
 * time; worlds of one batch share it (rp3dg_batch_set_contacts_position_correction). */
#define RP3DG_CONTACTS_BAUMGARTE 0
#define RP3DG_CONTACTS_SPLIT_IMPULSES 1
int rp3dg_world_set_contacts_position_correction(rp3dg_world* world, int32_t technique);
int rp3dg_batch_set_contacts_position_correction(rp3dg_batch* batch, int32_t technique);

/* PhysicsWorld::createRigidBody (src/engine/PhysicsWorld.cpp:401-442); returns the body index. */
int32_t rp3dg_body_create(rp3dg_world* world, const float position[3], const float orientation[4]);
/* RigidBody::setType (src/body/RigidBody.cpp:66-130). */
int rp3dg_body_set_type(rp3dg_world* world, int32_t body, int32_t type);
/* RigidBody::setLinearVelocity / setAngularVelocity (src/body/RigidBody.cpp:794-828); NULL leaves one as
 * is.  This and the other RigidBody calls below (velocity, transform, forces, torques, sleeping) may be
 * made between update() calls of a world stepped with rp3dg_world_update: they are queued for its
 * device batch (see rp3dg_batch_body_set_transform).  For a world in a shared batch they return
 * RP3DG_ESTATE after the batch exists; use the rp3dg_batch_body_* calls instead. */
int rp3dg_body_set_velocity(rp3dg_world* world, int32_t body, const float linear[3], const float angular[3]);
/* RigidBody::setTransform (src/body/RigidBody.cpp:837-866). */
int rp3dg_body_set_transform(rp3dg_world* world, int32_t body, const float position[3], const float orientation[4]);
/* RigidBody::apply{World,Local}ForceAt{CenterOfMass,WorldPosition,LocalPosition}
 * (src/body/RigidBody.cpp:151-239, 298-311), flags RP3DG_FORCE_* as for rp3dg_batch_body_apply_force. */
int rp3dg_body_apply_force(rp3dg_world* world, int32_t body, const float force[3], const float point[3], int32_t flags);
/* RigidBody::applyWorldTorque / applyLocalTorque (src/body/RigidBody.cpp:938-967). */
int rp3dg_body_apply_torque(rp3dg_world* world, int32_t body, const float torque[3], int32_t flags);
/* RigidBody::resetForce / resetTorque (src/body/RigidBody.cpp:970-987). */
int rp3dg_body_reset_force_torque(rp3dg_world* world, int32_t body, int32_t force, int32_t torque);
/* RigidBody::getForce / getTorque (src/body/RigidBody.cpp:993-1003). */
int rp3dg_body_get_force_torque(rp3dg_world* world, int32_t body, float force[3], float torque[3]);
/* RigidBody::setIsSleeping (src/body/RigidBody.cpp:1006-1057), after the first update (bodies are
 * created awake; RP3DG_EUNSUPPORTED to put one to sleep before). */
int rp3dg_body_set_sleeping(rp3dg_world* world, int32_t body, int32_t sleeping);
/* RigidBody::setLinearDamping / setAngularDamping. */
int rp3dg_body_set_damping(rp3dg_world* world, int32_t body, float linear, float angular);
/* RigidBody::enableGravity. */
int rp3dg_body_enable_gravity(rp3dg_world* world, int32_t body, int32_t enabled);
/* RigidBody::applyWorldForceAtCenterOfMass + applyWorldTorque: accumulate until the next update(). */
int rp3dg_body_apply_force_torque(rp3dg_world* world, int32_t body, const float force[3], const float torque[3]);
/* RigidBody::setLinearLockAxisFactor / setAngularLockAxisFactor. */
int rp3dg_body_set_lock_factors(rp3dg_world* world, int32_t body, const float linear[3], const float angular[3]);
/* RigidBody::updateMassPropertiesFromColliders (src/body/RigidBody.cpp:548-612). */
int rp3dg_body_update_mass_properties(rp3dg_world* world, int32_t body);
/* RigidBody::setMass / setLocalInertiaTensor / setLocalCenterOfMass (explicit values). */
int rp3dg_body_set_mass_properties(rp3dg_world* world, int32_t body, float mass, const float inertia_local[3],
                                   const float com_local[3]);

/* PhysicsCommon::create{Sphere,Capsule,Box}Shape + RigidBody::addCollider (src/body/RigidBody.cpp:661-721).
 * params: sphere {radius}; capsule {radius, height}; box {hx, hy, hz}; convex mesh {hull index}.
 * Returns the collider index. */
int32_t rp3dg_collider_add(rp3dg_world* world, int32_t body, int32_t shape_type, const float params[4],
                           const float local_position[3], const float local_orientation[4]);
/* Collider::getMaterial().set{FrictionCoefficient,Bounciness,MassDensity} (engine/Material.h). */
int rp3dg_collider_set_material(rp3dg_world* world, int32_t collider, float friction, float bounciness,
                                float mass_density);
/* RigidBody::setIsAllowedToSleep (src/body/RigidBody.cpp:1154-1163); bodies are allowed by default. */
int rp3dg_body_set_allowed_to_sleep(rp3dg_world* world, int32_t body, int32_t allowed);
/* Collider::setCollisionCategoryBits / setCollideWithMaskBits (src/collision/Collider.cpp). */
int rp3dg_collider_set_filter(rp3dg_world* world, int32_t collider, uint16_t category, uint16_t mask);

/* Collider kinds (Collider::getIsTrigger / getIsSimulationCollider / getIsWorldQueryCollider,
 * src/collision/Collider.cpp:240-300).  A new collider is a simulation and world-query collider. */
#define RP3DG_COLLIDER_TRIGGER 1
#define RP3DG_COLLIDER_SIMULATION 2
#define RP3DG_COLLIDER_WORLD_QUERY 4
/* Collider::setIsTrigger (src/collision/Collider.cpp:250-257): a trigger reports overlaps (onTrigger) and
 * takes no part in the simulation; setting it clears the simulation flag. */
int rp3dg_collider_set_is_trigger(rp3dg_world* world, int32_t collider, int32_t is_trigger);
/* Collider::setIsSimulationCollider (src/collision/Collider.cpp:272-287); setting it clears the trigger flag.
 * A collider that is neither is never tested in the step. */
int rp3dg_collider_set_is_simulation_collider(rp3dg_world* world, int32_t collider, int32_t is_simulation);
/* Collider::setIsWorldQueryCollider (src/collision/Collider.cpp:298-300): whether testOverlap /
 * testCollision queries see the collider. */
int rp3dg_collider_set_is_world_query_collider(rp3dg_world* world, int32_t collider, int32_t is_world_query);
/* The RP3DG_COLLIDER_* bits of a collider (or a negative code). */
int rp3dg_collider_get_kind(rp3dg_world* world, int32_t collider);

/* PhysicsCommon::createConvexMesh(PolygonVertexArray) (src/engine/PhysicsCommon.cpp:600-630):
 * vertices xyz[nv], faces given as counts[nf] + concatenated vertex indices (CCW seen from outside).
 * Returns the hull index used as params[0] of RP3DG_SHAPE_CONVEX_MESH.  Limits: at most 255 half-edges
 * and 24 vertices per face (RP3DG_EUNSUPPORTED beyond). */
int32_t rp3dg_hull_create(rp3dg_world* world, const float* vertices, int32_t nb_vertices,
                          const int32_t* face_counts, const int32_t* face_indices, int32_t nb_faces);

/* PhysicsCommon::createConvexMesh(const VertexArray&, std::vector<Message>&) (src/engine/PhysicsCommon.cpp:638-660):
 * the convex hull of nb_vertices points, computed as the reference's QuickHull computes it
 * (src/utils/quickhull/QuickHull.cpp:44-131: same faces, face order and vertex numbering), then built as
 * rp3dg_hull_create builds a polygon mesh.  Vertex i is 3 floats (is_double = 0, VERTEX_FLOAT_TYPE) or
 * 3 doubles (VERTEX_DOUBLE_TYPE) at vertices + i * stride bytes (VertexArray.h).  Returns the hull
 * index, or RP3DG_EINVAL with the reference's message in rp3dg_last_error() (fewer than 4 distinct
 * points, points too close to each other). */
int32_t rp3dg_hull_create_from_points(rp3dg_world* world, const void* vertices, int32_t nb_vertices, int32_t stride,
                                      int32_t is_double);
/* QuickHull::computeConvexHull alone (src/utils/quickhull/QuickHull.cpp:44-131): the polygons that
 * createConvexMesh(VertexArray) builds its mesh from.  Writes up to cap_vertices vertices (xyz),
 * cap_faces face vertex counts and cap_indices face vertex indices; *nb_out_vertices / *nb_out_faces
 * receive the counts even when a capacity is short (then RP3DG_EINVAL). */
int rp3dg_quickhull(const void* vertices, int32_t nb_vertices, int32_t stride, int32_t is_double, float* out_vertices,
                    int32_t cap_vertices, int32_t* face_counts, int32_t cap_faces, int32_t* face_indices,
                    int32_t cap_indices, int32_t* nb_out_vertices, int32_t* nb_out_faces);

/* PhysicsCommon::createTriangleMesh(TriangleVertexArray, messages) (src/collision/TriangleMesh.cpp:39-249):
 * nb_vertices xyz vertices, nb_triangles triplets of vertex indices (counter-clockwise seen from the front),
 * optional xyz vertex normals (NULL: TriangleMesh::computeVerticesNormals, angle-weighted).  Triangles
 * with an almost zero area or edge are dropped (the reference's warning messages; their number goes to
 * *nb_skipped when not NULL) and unused vertices removed; the mesh's AABB tree is built as the reference
 * builds it.  Returns the mesh index used as params[0] of RP3DG_SHAPE_CONCAVE_MESH, or RP3DG_EINVAL (an
 * index out of range, a too short given normal, no valid triangle). */
int32_t rp3dg_triangle_mesh_create(rp3dg_world* world, const float* vertices, int32_t nb_vertices,
                                   const int32_t* indices, int32_t nb_triangles, const float* normals,
                                   int32_t* nb_skipped);
/* PhysicsCommon::createHeightField(columns, rows, heights, HEIGHT_FLOAT_TYPE, messages)
 * (src/collision/HeightField.cpp:43-130): columns x rows samples, height of grid point (x, y) at
 * heights[y * columns + x]; the field is centred on the origin with y up, one unit between grid
 * points before the shape's scale.  Returns the index used as params[0] of RP3DG_SHAPE_HEIGHTFIELD. */
int32_t rp3dg_heightfield_create(rp3dg_world* world, int32_t columns, int32_t rows, const float* heights);
/* ConcaveShape::computeOverlappingTriangles (ConcaveMeshShape.cpp:148-176, HeightField.cpp:133-180) of a
 * concave collider for an AABB in the shape's local space: the triangles in the order of the middle
 * phase, as ids (mesh: triangle index; height field: (j * (columns - 1) + i) * 2 + k), 9 vertex and 9
 * vertex-normal floats each (scaled local space).  Returns the number of triangles (outputs hold at
 * most capacity of them; NULL outputs are skipped). */
int32_t rp3dg_concave_overlapping_triangles(rp3dg_world* world, int32_t collider, const float aabb_min[3],
                                            const float aabb_max[3], int32_t* ids, float* vertices,
                                            float* normals, int32_t capacity);

/* PhysicsWorld::createJoint (src/engine/PhysicsWorld.cpp:483-600) with anchor / axis given in world space
 * (the JointInfo "world space" constructors, e.g. BallAndSocketJointInfo(body1, body2, anchor)).
 * axis is used by hinge and slider; limits[4] = {min, max, motor speed, max motor force/torque}
 * with flags bit0 limit enabled, bit1 motor enabled (HingeJointInfo / SliderJointInfo fields) and
 * RP3DG_JOINT_BAUMGARTE (bit2) for JointInfo::positionCorrectionTechnique = BAUMGARTE_JOINTS: the position
 * error enters the velocity rows as a bias of 0.2 / dt times the error and the joint gets no non-linear
 * Gauss-Seidel position correction (Solve*JointSystem::initBeforeSolve / solvePositionConstraint).  As in
 * the reference, the sliders of a world solved after a Baumgarte slider get no position correction either
 * (SolveSliderJointSystem.cpp:527 returns).
 * Ball-and-socket: flags bit0 enables the cone limit with half angle limits[0] in [0, pi]
 * (BallAndSocketJoint::enableConeLimit / setConeLimitHalfAngle, BallAndSocketJoint.cpp:71-96).
 * Slider limits must satisfy min <= 0 <= max and the max motor force must be >= 0
 * (SliderJoint.cpp:41-43); RP3DG_EINVAL otherwise.  Returns the joint index. */
int32_t rp3dg_joint_create(rp3dg_world* world, int32_t type, int32_t body1, int32_t body2,
                           const float anchor_world[3], const float axis_world[3], int32_t collide_connected,
                           int32_t flags, const float limits[4]);
/* Replaces the limit / motor state of a joint before the world's first update (flags bits 0-1 and limits
 * as in rp3dg_joint_create, the position correction technique stays; RP3DG_ESTATE afterwards): BallAndSocketJoint::enableConeLimit /
 * setConeLimitHalfAngle (BallAndSocketJoint.cpp:71-96), HingeJoint::enableLimit / enableMotor /
 * setMinAngleLimit / setMaxAngleLimit / setMotorSpeed / setMaxMotorTorque (HingeJoint.cpp:90-195),
 * SliderJoint::enableLimit / enableMotor / setMinTranslationLimit / setMaxTranslationLimit /
 * setMotorSpeed / setMaxMotorForce (SliderJoint.cpp:98-235). */
int rp3dg_joint_set_limits(rp3dg_world* world, int32_t joint, int32_t flags, const float limits[4]);

/* ---- stepping (replaces PhysicsWorld::update, src/engine/PhysicsWorld.cpp:239-307) ---- */

/* Steps one world on the GPU (uploads it on first use, as a batch of one). */
int rp3dg_world_update(rp3dg_world* world, float time_step);
/* RigidBody::getTransform / getLinearVelocity / getAngularVelocity (src/body/RigidBody.cpp) of a world
 * stepped with rp3dg_world_update: a host mirror refreshed from the device once after each update.
 * Before the first update it returns the creation state.  Any output pointer may be NULL. */
int rp3dg_world_get_body_state(rp3dg_world* world, int32_t body, float position[3], float orientation[4],
                               float linear_velocity[3], float angular_velocity[3]);
/* RigidBody::isSleeping (src/body/RigidBody.cpp:1006-1057) after the last update: 1 asleep, 0 awake. */
int rp3dg_world_get_body_sleeping(rp3dg_world* world, int32_t body);

/* Many independent worlds stepped together: one launch sequence per step for all of them.  The
 * reference has one PhysicsWorld per update() call (src/engine/PhysicsWorld.cpp:239); a batch holds
 * many, which must share one WorldSettings. */
rp3dg_batch* rp3dg_batch_create(rp3dg_world* const* worlds, int32_t nb_worlds);

/* Solver layout of a batch. */
#define RP3DG_SOLVER_AUTO 0   /* FUSED when every world has at most 4096 bodies, else PER_COLOR */
#define RP3DG_SOLVER_PER_COLOR 1  /* one launch per colour over all worlds (large worlds) */
#define RP3DG_BROADPHASE_AUTO 0        /* world scan when it applies and pays (>= 64 worlds, or worlds of <= 256 colliders), else the tree */
#define RP3DG_BROADPHASE_TREE 1        /* linear BVH over all colliders, rebuilt every step */
#define RP3DG_BROADPHASE_WORLD_SCAN 2  /* fused layout, worlds of <= 1024 colliders: every collider tests its
                                          world's boxes in LDS (no tree) */
#define RP3DG_SOLVER_FUSED 2  /* one workgroup per group of small worlds runs the whole solver phase (worlds up to 4096 bodies) */

typedef struct rp3dg_batch_options {
    int32_t solver;           /* RP3DG_SOLVER_* */
    int32_t fused_block;      /* threads per world in the fused solver (64, 128, 192 or 256; 0 = per world size
                                 class: bodies bound / 2, clamped to [64, 256], autotunable) */
    int32_t pairs_per_collider; /* pair-table capacity per collider (0 = 16) */
    int32_t group_bodies;     /* fused solver: consecutive worlds are packed into one workgroup up to this many
                                 bodies together (at most 64 worlds, 512 bodies; 0 = 256, 1 = one world per
                                 workgroup); results do not depend on it */
    int32_t broadphase;       /* RP3DG_BROADPHASE_*; the pair set does not depend on it */
    int32_t reserved[3];
} rp3dg_batch_options;
/* As rp3dg_batch_create, with options (NULL = defaults). */
rp3dg_batch* rp3dg_batch_create_ex(rp3dg_world* const* worlds, int32_t nb_worlds, const rp3dg_batch_options* options);
void rp3dg_batch_destroy(rp3dg_batch* batch);
int rp3dg_batch_step(rp3dg_batch* batch, float time_step, int32_t nb_steps);
/* Captures one step into a hipGraph (once per time step value) and replays it nb_steps times: no
 * allocation and no host synchronisation inside the step.  Requires the fused solver layout. */
int rp3dg_batch_step_graph(rp3dg_batch* batch, float time_step, int32_t nb_steps);
int rp3dg_batch_synchronize(rp3dg_batch* batch);

/* ---- readback (host mirror of RigidBody getters / contact data) ---- */

typedef struct rp3dg_counts {
    int64_t nb_bodies, nb_colliders, nb_joints, nb_worlds;
    int64_t nb_pairs;          /* overlapping pairs after the broadphase of the last step */
    int64_t nb_manifolds;      /* contact manifolds of the last step */
    int64_t nb_contact_points; /* contact points of the last step */
    int64_t nb_colors;         /* solver colours used by the last step */
    int64_t status;            /* device status word (0 = OK) */
    int64_t nb_tests[7];       /* narrowphase tests of the last step per algorithm (index 1..6:
                                  sphere-sphere, sphere-capsule, capsule-capsule, sphere-poly,
                                  capsule-poly, poly-poly) */
    int64_t total_contact_points; /* contact points summed over every step since creation */
    int64_t steps;             /* steps run since creation */
    int64_t nb_new_points;     /* contact points of the last step without a cached impulse (counted when the
                                  learned warm start or sample logging is on, else 0) */
    int64_t nb_box_tests;      /* box-box subset of nb_tests[6] */
} rp3dg_counts;
int rp3dg_batch_get_counts(rp3dg_batch* batch, rp3dg_counts* out);

/* Body state, indices are global over the batch (world order, then body creation order).
 * position = body origin (Transform position), orientation xyzw, velocities of the centre of mass. */
int rp3dg_batch_get_bodies(rp3dg_batch* batch, float* position3, float* orientation4, float* linear3,
                           float* angular3);
int rp3dg_batch_set_bodies(rp3dg_batch* batch, const float* position3, const float* orientation4,
                           const float* linear3, const float* angular3);
/* Overlapping pairs (global collider indices, collider1 = reference's pair collider1), sorted by (min,max). */
int rp3dg_batch_get_pairs(rp3dg_batch* batch, int32_t* collider1, int32_t* collider2, int64_t capacity);

/* Contacts of the last step, one record per manifold in solver order.  Per manifold m:
 * colliders[2m..2m+1], nb_points[m], color[m]; per point p (4 slots per manifold):
 * normal[3], depth, local1[3], local2[3], impulse (accumulated normal impulse after the step). */
typedef struct rp3dg_contact_view {
    int32_t* colliders;     /* 2 per manifold */
    int32_t* nb_points;     /* 1 per manifold */
    int32_t* color;         /* 1 per manifold */
    float* normal_depth;    /* 16 per manifold (4 points x {nx,ny,nz,depth}) */
    float* local1;          /* 12 per manifold */
    float* local2;          /* 12 per manifold */
    float* impulse;         /* 4 per manifold */
    float* friction;        /* 3 per manifold: friction1, friction2, twist impulses */
    int32_t* manifold;      /* 1 per manifold: its index among its pair's manifolds (0 but for a pair of a
                             * convex and a concave shape, which has up to 4, ContactPair.h) */
} rp3dg_contact_view;
int rp3dg_batch_get_contacts(rp3dg_batch* batch, rp3dg_contact_view* out, int64_t capacity);
/* Contact events, what PhysicsWorld::setEventListener + EventListener::onContact deliver
 * (CollisionDetectionSystem::reportContacts, CollisionDetectionSystem.cpp:1708-1741): once enabled,
 * every step records one event per colliding pair, RP3DG_CONTACT_START or RP3DG_CONTACT_STAY (by
 * whether the pair collided in the previous step), and RP3DG_CONTACT_EXIT for every pair that stopped
 * colliding or that the broadphase removed while colliding (CollisionCallback.cpp:69-76).  Enabling or
 * disabling drops captured step graphs. */
#define RP3DG_CONTACT_START 0
#define RP3DG_CONTACT_STAY 1
#define RP3DG_CONTACT_EXIT 2
/* Pairs with a trigger collider report to EventListener::onTrigger instead
 * (CollisionDetectionSystem::reportTriggers, CollisionDetectionSystem.cpp:1745-1757): OverlapStart /
 * OverlapStay / OverlapExit (OverlapCallback::OverlapPair::EventType), without contact points. */
#define RP3DG_TRIGGER_START 3
#define RP3DG_TRIGGER_STAY 4
#define RP3DG_TRIGGER_EXIT 5
int rp3dg_batch_set_contact_events(rp3dg_batch* batch, int32_t enable);
/* The events of the last step (of the last step of a graph replay), in no particular order: colliders
 * (batch indices, oriented as the pair), event type and contact points (0 for an exit).  Returns the
 * count (only the count when every output pointer is NULL), or a negative status (RP3DG_ESTATE when
 * events are off, RP3DG_EINVAL if capacity is short). */
int64_t rp3dg_batch_get_contact_events(rp3dg_batch* batch, int32_t* collider1, int32_t* collider2, int32_t* type,
                                       int32_t* nb_points, int64_t capacity);
/* PhysicsWorld::setEventListener for a world stepped with rp3dg_world_update: events are recorded from
 * the next update on; read them (and the contact points) through rp3dg_world_get_batch. */
int rp3dg_world_set_contact_events(rp3dg_world* world, int32_t enable);
/* The batch of one that rp3dg_world_update steps (NULL before the first update). */
rp3dg_batch* rp3dg_world_get_batch(rp3dg_world* world);
/* Creates that batch now, as the first update() would, without stepping (for raycasts and world queries
 * before the first step); returns it (NULL with rp3dg_last_error() on failure). */
rp3dg_batch* rp3dg_world_upload(rp3dg_world* world);
/* World queries (PhysicsWorld::testOverlap / testCollision, PhysicsWorld.h:319-334,
 * CollisionDetectionSystem.cpp:1775-1896) on the current state.  As in the reference, a query first runs
 * the broadphase (computeBroadPhase): pairs are created and removed, the moved set empties, and pairs it
 * removes while colliding report ContactExit at the next step; the pairs it tests update their
 * LastFrameCollisionInfo (SAT without clipWithPreviousAxisIfStillColliding), the contacts it finds are
 * reported here and do not replace the last step's contacts.  A query applies to every world of the
 * batch, as if each world made it (bodies are batch indices and belong to one world); pairs are listed in
 * no particular order.
 * testOverlap(body1, body2): 1 when a pair of the two bodies' colliders collides, else 0. */
int rp3dg_batch_test_overlap_bodies(rp3dg_batch* batch, int32_t body1, int32_t body2);
/* testOverlap(OverlapCallback&) (body < 0: pairs of world-query colliders) or testOverlap(body,
 * OverlapCallback&) (pairs of the body): returns the number of overlapping collider pairs; read them with
 * rp3dg_batch_get_query_results (OverlapPair::getCollider1/2, no points). */
int64_t rp3dg_batch_test_overlap(rp3dg_batch* batch, int32_t body);
/* testCollision(CollisionCallback&) (body1 < 0: pairs of world-query colliders), testCollision(body1,
 * CollisionCallback&) (body2 < 0) or testCollision(body1, body2, CollisionCallback&): returns the number
 * of colliding pairs; read them with rp3dg_batch_get_query_results.  Pairs with a trigger collider are not
 * listed (CollisionCallback::CallbackData keeps contact pairs only, CollisionCallback.cpp:60-70). */
int64_t rp3dg_batch_test_collision(rp3dg_batch* batch, int32_t body1, int32_t body2);
/* What the CallbackData of the last rp3dg_batch_test_overlap / test_collision holds, one record per pair
 * in the contact view layout: colliders, nb_points, normal_depth, local1, local2
 * (ContactPoint::getWorldNormal / getPenetrationDepth / getLocalPointOnCollider1/2), and in color the
 * event type: testCollision gives 0 ContactStart or 1 ContactStay by whether the pair collided in the
 * previous step, testOverlap always 0 OverlapStart (CollisionDetectionSystem.cpp:831); impulse and
 * friction are written 0.  Reading does not run the query again.  Returns
 * the count. */
int64_t rp3dg_batch_get_query_results(rp3dg_batch* batch, rp3dg_contact_view* out, int64_t capacity);
/* RigidBody::applyWorldForceAtCenterOfMass / applyWorldTorque (src/body/RigidBody.cpp:298-311, 938-951)
 * on every dynamic body of the batch whose force / torque (xyz per body, NULL = none) is not zero,
 * before the next step (the step clears the accumulated forces); a sleeping body is woken. */
int rp3dg_batch_apply_forces(rp3dg_batch* batch, const float* force, const float* torque);
/* RigidBody calls on one body of a stepped batch (global body index), for applications that drive
 * bodies between update() calls.  Calls are queued in the order they are made and applied on the
 * device, in that order, before the next step, graph replay, raycast or read-back, so a sequence of
 * calls on one body composes as the reference's calls do.  rp3dg_world_* setters of a world stepped
 * with rp3dg_world_update go through the same queue.
 * RigidBody::setTransform (src/body/RigidBody.cpp:837-866, Body.cpp:397-407): the centre of mass
 * follows, the linear velocity gains angular x (new - old centre of mass), the body wakes and its
 * colliders are refit (a shape that left its fat AABB gets a new one and joins the moved set).  A static
 * body's transform is fixed once the batch exists (RP3DG_EUNSUPPORTED): its static-static pairs are
 * listed at creation. */
int rp3dg_batch_body_set_transform(rp3dg_batch* batch, int32_t body, const float position[3],
                                   const float orientation[4]);
/* RigidBody::setLinearVelocity / setAngularVelocity (src/body/RigidBody.cpp:794-828): ignored for a
 * static body, a non-zero velocity wakes the body; NULL leaves that velocity as is. */
int rp3dg_batch_body_set_velocity(rp3dg_batch* batch, int32_t body, const float linear[3], const float angular[3]);
/* RigidBody::apply{World,Local}ForceAt{CenterOfMass,WorldPosition,LocalPosition}
 * (src/body/RigidBody.cpp:151-239, 298-311): the force (world frame, or body frame with
 * RP3DG_FORCE_LOCAL) adds to the body's external force; with RP3DG_FORCE_AT_POINT the point (world
 * frame, or body frame with RP3DG_FORCE_POINT_LOCAL) adds (point - centre of mass) x force to the
 * torque.  Dynamic bodies only; a sleeping body wakes.  The step clears the sums. */
#define RP3DG_FORCE_LOCAL 1
#define RP3DG_FORCE_AT_POINT 2
#define RP3DG_FORCE_POINT_LOCAL 4
int rp3dg_batch_body_apply_force(rp3dg_batch* batch, int32_t body, const float force[3], const float point[3],
                                 int32_t flags);
/* RigidBody::applyWorldTorque / applyLocalTorque (src/body/RigidBody.cpp:938-967); flags 0 or
 * RP3DG_FORCE_LOCAL. */
int rp3dg_batch_body_apply_torque(rp3dg_batch* batch, int32_t body, const float torque[3], int32_t flags);
/* RigidBody::resetForce / resetTorque (src/body/RigidBody.cpp:970-987), each when its flag is set. */
int rp3dg_batch_body_reset_force_torque(rp3dg_batch* batch, int32_t body, int32_t force, int32_t torque);
/* RigidBody::setIsSleeping (src/body/RigidBody.cpp:1006-1057): putting a body to sleep zeroes its
 * velocities and external force/torque, waking it puts its colliders into the moved set; a static body
 * stays awake.  Pairs whose two bodies are disabled (asleep or static) once the queued calls have run
 * restart their LastFrameCollisionInfo, as the reference's disabled-pair move does
 * (OverlappingPairs.cpp:301-346).  RP3DG_ESTATE to put a body to sleep when sleeping is disabled. */
int rp3dg_batch_body_set_sleeping(rp3dg_batch* batch, int32_t body, int32_t sleeping);
/* RigidBody::getForce / getTorque (src/body/RigidBody.cpp:993-1003): the external force and torque
 * accumulated since the last step, xyz per body (NULL = skip). */
int rp3dg_batch_get_force_torque(rp3dg_batch* batch, float* force3, float* torque3);
/* PhysicsWorld::raycast (PhysicsWorld.h:486-490, CollisionDetectionSystem.cpp:1186-1195)
 * with a RaycastCallback that keeps the closest hit (notifyRaycastHit returns the hit fraction), for
 * n rays at once: ray i runs from point1 to point2 (xyz each) in world world[i] (NULL: world 0) with
 * max fraction max_fraction[i] (NULL: 1) against the world-query colliders whose category bits meet
 * category_mask, in the state after the last step.  Out: collider[i] (batch index, -1 for no hit),
 * hit point and unit normal in world space (xyz each) and hit fraction; NULL outputs are skipped. */
int rp3dg_batch_raycast(rp3dg_batch* batch, int64_t n, const int32_t* world, const float* point1, const float* point2,
                        const float* max_fraction, uint16_t category_mask, int32_t* collider, float* point,
                        float* normal, float* fraction);
/* PhysicsWorld::raycast (PhysicsWorld.h:486-490) with a RaycastCallback whose notifyRaycastHit returns 1
 * (DynamicAABBTree.cpp:684-747: the ray is never clipped): every world-query collider whose category
 * passes the mask and whose shape the ray crosses within its max fraction, for n rays at once (arguments
 * as rp3dg_batch_raycast).  Hits are written sorted by (ray, collider): ray index, collider, world point,
 * world normal and fraction (NULL = skip).  Returns the hit count, or RP3DG_EINVAL when it exceeds
 * capacity (capacity 0 with NULL ray and collider returns the count). */
int64_t rp3dg_batch_raycast_all(rp3dg_batch* batch, int64_t n, const int32_t* world, const float* point1,
                                const float* point2, const float* max_fraction, uint16_t category_mask, int32_t* ray,
                                int32_t* collider, float* point, float* normal, float* fraction, int64_t capacity);
/* RigidBody::isSleeping of every body of the batch (batch order), 0 / 1; capacity >= body count. */
int rp3dg_batch_get_sleeping(rp3dg_batch* batch, int32_t* sleeping, int64_t capacity);
/* Joint solve order (joint indices in the order the device solved them this step, per type block). */
int rp3dg_batch_get_joint_order(rp3dg_batch* batch, int32_t* order, int64_t capacity);

/* Per-kernel timing with HIP events on the batch's stream (for the roofline figures), eager steps only
 * (rp3dg_batch_step_graph refuses to run while timing is on).  rp3dg_batch_profile(batch, 1, NULL)
 * times every kernel, (batch, 1, "name") only that kernel (two events per step), (batch, 0, NULL)
 * disables and clears.
 * get returns the mean launch time in ms of each named kernel. */
int rp3dg_batch_profile(rp3dg_batch* batch, int32_t enable, const char* only_kernel);
int rp3dg_batch_get_kernel_times(rp3dg_batch* batch, const char* kernel, double* mean_ms, int64_t* launches);

/* Autotuner: times each candidate launch shape of the solver over steps_per_trial real steps of the
 * batch (the steps advance the simulation) and keeps the fastest; returns the choice and its time.
 * Fused layout: worlds run in size classes (bodies in (L/2, L], L = 64 .. 4096), one concurrent launch
 * each; the candidates are the divisor d of the class workgroup clamp(L / d, 64, 256), d = 1, 2, 4, 8, 16.
 * Per-colour layout: threads per block of the colour kernels, 64..512. */
int rp3dg_batch_autotune(rp3dg_batch* batch, float time_step, int32_t steps_per_trial, int32_t* chosen,
                         double* best_ms_per_step);
/* Sets the solver launch shape to an autotuner choice (fused layout: the class workgroup divisor 1, 2, 4, 8
 * or 16; per-colour layout: threads per block, a multiple of 64 in 64..512), e.g. to time two batches
 * with the same shape.  Drops captured step graphs. */
int rp3dg_batch_set_solver_shape(rp3dg_batch* batch, int32_t choice);
/* Fused layout, residual mode: each step dispatches the workgroups of a size class longest expected
 * first (the largest sweep count of a group's worlds in the previous step), so groups that stop early
 * fill the tail (by_sweeps = 1, the default), or in the static order of group sizes (0).  Results do not
 * depend on it. */
int rp3dg_batch_set_group_order(rp3dg_batch* batch, int32_t by_sweeps);
/* The trials of the last rp3dg_batch_autotune call: up to capacity (choice, ms per step) pairs;
 * returns the number of trials. */
int32_t rp3dg_batch_get_tune_table(rp3dg_batch* batch, int32_t* choices, double* ms_per_step, int32_t capacity);

/* Solver convergence mode (no counterpart in the reference, which runs a fixed iteration count,
 * PhysicsWorld.cpp:334): every world iterates until the velocity residual
 *   max over its dynamic bodies of max(|v_k - v_{k-1}|_inf, |w_k - w_{k-1}|_inf)
 * between successive sweeps is below residual_threshold, at most max_iterations sweeps.
 * residual_threshold <= 0 restores the fixed WorldSettings iteration count.  Fused layout only. */
int rp3dg_batch_set_residual_mode(rp3dg_batch* batch, float residual_threshold, int32_t max_iterations);
/* Mean and max over worlds of the velocity iterations run by the last step, and the largest final
 * residual (0 in fixed-iteration mode). */
int rp3dg_batch_get_iteration_stats(rp3dg_batch* batch, double* mean_iterations, double* max_iterations,
                                    double* max_residual);

/* Velocity iterations each world ran in the last step and its last residual (0 in fixed-iteration
 * mode), world order; capacity >= number of worlds.  Returns the number of worlds. */
int rp3dg_batch_get_world_iterations(rp3dg_batch* batch, int32_t* iterations, float* residuals, int64_t capacity);

/* ---- learned warm start (no counterpart in the reference) ----
 * A 24 -> 128 -> 128 -> 3 SiLU network, BF16 on the matrix cores with FP32 accumulation, predicts the
 * normal impulse of every contact point that has no cached impulse (the reference starts those from 0,
 * ContactSolverSystem.cpp:344-376) and the two friction impulses of a manifold with no resting point
 * (the reference zeroes them, :468-474).  Features per point: normal (3), r1 (3), r2 (3), relative
 * velocity along t1, t2, n (3), inverse masses (2), world inverse inertia xx of body 1, the
 * manifold's cached normal impulse sum and resting point count, world inverse inertia xx of body 2,
 * manifold counts of both bodies, depth, friction coefficient, inverse normal mass, 1 / points in
 * the manifold.  Targets:
 * converged normal impulse, manifold friction impulses / points.  Parameters (20,099 floats, fp32
 * master): W1[24][128], b1[128], W2[128][128], b2[128], W3[128][3], b3[3]. */
typedef struct rp3dg_mlp rp3dg_mlp;
#define RP3DG_WARM_START_PLAIN 0    /* cached impulses, zero for new points (the reference) */
#define RP3DG_WARM_START_LEARNED 1  /* cached impulses, network prediction for new points */

/* Glorot-uniform weights from a splitmix64 stream seeded with seed, zero biases, identity input
 * normalisation and unit output scale. */
rp3dg_mlp* rp3dg_mlp_create(uint64_t seed);
void rp3dg_mlp_destroy(rp3dg_mlp* mlp);
int rp3dg_mlp_set_params(rp3dg_mlp* mlp, const float* params, int64_t count);
int rp3dg_mlp_get_params(rp3dg_mlp* mlp, float* params, int64_t capacity);
/* x_normalised = (x - mu) * inv_sigma; impulse = network output * out_scale. */
int rp3dg_mlp_set_normalization(rp3dg_mlp* mlp, const float mu[24], const float inv_sigma[24],
                                const float out_scale[3]);
/* Host arrays: features [n][24] -> impulses [n][3]. */
int rp3dg_mlp_forward(rp3dg_mlp* mlp, const float* features, int64_t n, float* impulses);
/* Gradient of loss = 1/(2n) sum_s |y_s - target_s / out_scale|^2 over n <= 65536 samples. */
int rp3dg_mlp_gradient(rp3dg_mlp* mlp, const float* features, const float* targets, int64_t n, float* gradient,
                       float* loss);
/* No counterpart in the reference.  rp3dg_mlp_gradient for callers that cannot allocate device memory
 * (code under HIP graph capture, tight training or serving loops, harnesses that forbid hipMalloc):
 * features [n][24], targets [n][3], gradient [20099] and loss [1] are DEVICE pointers, and the
 * per-block partial gradients go to a fixed device array of RP3DG_MLP_NOALLOC_MAX_SAMPLES / 64 partials
 * (5.1 MB on each GPU the library runs on, held until the process ends) instead of library allocations.
 * n must be in [1, RP3DG_MLP_NOALLOC_MAX_SAMPLES]; RP3DG_EINVAL with a message otherwise.  Same
 * kernels and summation order as rp3dg_mlp_gradient, so the gradient and loss are identical.  Runs on
 * the network's stream and returns when the outputs are written; the inputs must be ready on the device.
 * Calls that overlap on one GPU (any networks) share the array and race. */
#define RP3DG_MLP_NOALLOC_MAX_SAMPLES 4096
int rp3dg_mlp_gradient_noalloc(rp3dg_mlp* mlp, const float* features, const float* targets, int64_t n,
                               float* gradient, float* loss);
/* Adam (beta 0.9 / 0.999) over shuffled minibatches, on the device. */
int rp3dg_mlp_train(rp3dg_mlp* mlp, const float* features, const float* targets, int64_t n, int32_t epochs,
                    int32_t minibatch, float learning_rate, uint64_t seed, float* last_loss);

/* Warm start of a batch (fused layout); the network must outlive its use by the batch. */
int rp3dg_batch_set_warm_start(rp3dg_batch* batch, int32_t mode, rp3dg_mlp* mlp);
/* Logs (features, converged impulses) of every point without a cached impulse, up to capacity samples
 * (0 disables and frees the log). */
int rp3dg_batch_set_logging(rp3dg_batch* batch, int64_t capacity);
/* Copies up to capacity logged samples; returns how many were copied. */
int64_t rp3dg_batch_get_log(rp3dg_batch* batch, float* features, float* targets, int64_t capacity);

/* The learned warm start predictions of the last step, one record of 12 floats per manifold in solver
 * order (the order of rp3dg_batch_get_contacts): 4 point slots x {normal impulse, friction 1, friction 2}
 * as the network predicted them for the points that had no cached impulse (other slots hold stale
 * values of earlier steps; all are 0 when the warm start is plain).  Returns the manifold count (only
 * the count when pred is NULL). */
int64_t rp3dg_batch_get_warm_start_predictions(rp3dg_batch* batch, float* pred, int64_t capacity);

/* ---- Snapshot and restore of a batch.  No counterpart in the reference, which keeps no checkpoint; a
 * snapshot covers the state that PhysicsWorld::update evolves: the bodies (RigidBodyComponents.h:70-160),
 * the broadphase's fat AABBs, the overlapping pairs with their LastFrameCollisionInfo
 * (OverlappingPairs.h:55-93), the previous step's contact cache (CollisionDetectionSystem.cpp:881-903),
 * joints' accumulated impulses and limit states, sleep timers and islands. ---- */
typedef struct rp3dg_snapshot rp3dg_snapshot;
/* Copies the batch's state into a new snapshot held in device memory (RigidBody calls still queued are
 * applied first): every device array of the batch and the host-side state of the step sequence (pair
 * table parity, counters of the last step, step count).  RP3DG_EHIP when device memory runs out,
 * RP3DG_ESTATE during a graph capture. */
int rp3dg_batch_snapshot(rp3dg_batch* batch, rp3dg_snapshot** snapshot);
/* Puts the batch back into a snapshot taken from it: the steps that follow equal, bit for bit, those
 * that followed the snapshot, and captured step graphs stay valid.  RigidBody calls queued since are
 * dropped; body and joint properties changed since return to the snapshot's.  Batch-level settings
 * (iterations, position correction techniques, residual mode, warm start, contact events on or off)
 * and the sample log stay as they are now.  RP3DG_EINVAL for a snapshot of another batch.  After a step
 * that failed with RP3DG_EOVERFLOW, a restore gives the batch back in the state of the snapshot. */
int rp3dg_batch_restore(rp3dg_batch* batch, const rp3dg_snapshot* snapshot);
/* Device bytes held by a snapshot. */
int64_t rp3dg_snapshot_bytes(const rp3dg_snapshot* snapshot);
void rp3dg_snapshot_destroy(rp3dg_snapshot* snapshot);

const char* rp3dg_last_error(void);
const char* rp3dg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RP3D_MI355X_H */
