// Host-side world construction: the C ABI functions that mirror the
// reference's object factories and RigidBody mass-property code.
#include "world.h"

#include <cstring>
#include <map>
#include <utility>

#include "../../include/rp3d_mi355x.h"
#include "quickhull.h"

namespace rp3dg {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

// HalfEdgeStructure::computeHalfEdges (src/collision/HalfEdgeStructure.cpp:35-121):
// half-edges are appended in twin pairs (earlier half first) the moment the
// second half of an edge is met, which fixes the edge numbering SAT iterates in.
bool HalfEdgeMesh::build(int nb_vertices, std::string* err) {
    typedef std::pair<int, int> Key;
    std::map<Key, Edge> pending;
    std::map<Key, Key> next_key;
    std::map<Key, int> index_of;
    std::map<int, Key> key_of;
    std::vector<Key> face_last_key(faces.size());
    edges.clear();
    vertex_edge.assign(nb_vertices, 0);
    face_edge.assign(faces.size(), 0);
    for (size_t f = 0; f < faces.size(); f++) {
        const std::vector<int>& fv = faces[f];
        const size_t n = fv.size();
        Key first(0, 0);
        Key prev(0, 0);
        for (size_t v = 0; v < n; v++) {
            const int a = fv[v];
            const int b = fv[v == n - 1 ? 0 : v + 1];
            if (a < 0 || a >= nb_vertices || b < 0 || b >= nb_vertices) {
                if (err) *err = "face vertex index out of range";
                return false;
            }
            const Key k(a, b);
            Edge e{a, 0, (int)f, 0};
            if (v == 0) first = k;
            else next_key[prev] = k;
            if (v == n - 1) next_key[k] = first;
            pending[k] = e;
            face_last_key[f] = k;
            const Key tk(b, a);
            auto it = pending.find(tk);
            if (it != pending.end()) {
                const int idx = (int)edges.size();
                Edge twin = it->second;
                twin.twin = idx + 1;
                e.twin = idx;
                key_of[idx] = tk;
                key_of[idx + 1] = k;
                vertex_edge[a] = idx + 1;
                vertex_edge[b] = idx;
                index_of[k] = idx + 1;
                index_of[tk] = idx;
                edges.push_back(twin);
                edges.push_back(e);
            }
            prev = k;
        }
    }
    for (size_t i = 0; i < edges.size(); i++) {
        auto kit = key_of.find((int)i);
        auto nit = next_key.find(kit->second);
        auto iit = (nit == next_key.end()) ? index_of.end() : index_of.find(nit->second);
        if (nit == next_key.end() || iit == index_of.end()) {
            if (err) *err = "mesh is not closed (half-edge without successor)";
            return false;
        }
        edges[i].next = iit->second;
    }
    for (size_t f = 0; f < faces.size(); f++) {
        auto it = index_of.find(face_last_key[f]);
        if (it == index_of.end()) {
            if (err) *err = "mesh is not closed (face edge without twin)";
            return false;
        }
        face_edge[f] = it->second;
    }
    return true;
}

float shape_volume(const HostWorld& w, const HostCollider& c) {
    switch (c.type) {
        case SHAPE_SPHERE: {  // SphereShape.h:176-178
            const float r = c.params[0];
            return 4.0f / 3.0f * kPi * r * r * r;
        }
        case SHAPE_CAPSULE: {  // CapsuleShape.h:175-177
            const float r = c.params[0];
            const float hh = c.params[1] * 0.5f;
            return kPi * r * r * (4.0f * r / 3.0f + 2.0f * hh);
        }
        case SHAPE_BOX:  // BoxShape.h:233-235
            return 8 * c.params[0] * c.params[1] * c.params[2];
        case SHAPE_HULL:
            return w.hulls[(int)c.params[0]].volume;
        case SHAPE_CONCAVE_MESH:
        case SHAPE_HEIGHTFIELD: {  // ConcaveShape::getVolume (ConcaveShape.cpp:41-51): volume of the local AABB
            V3 lo, hi;
            concave_local_bounds(w, c, lo, hi);
            const float lx = hi.x - lo.x, ly = hi.y - lo.y, lz = hi.z - lo.z;
            return lx * ly * lz;
        }
    }
    return 0.f;
}

V3 shape_local_inertia(const HostWorld& w, const HostCollider& c, float mass) {
    switch (c.type) {
        case SHAPE_SPHERE: {  // SphereShape.h:170-173
            const float r = c.params[0];
            const float d = 0.4f * mass * r * r;
            return v3(d, d, d);
        }
        case SHAPE_CAPSULE: {  // CapsuleShape.cpp:51-67
            const float r = c.params[0];
            const float hh = c.params[1] * 0.5f;
            const float height = hh + hh;
            const float r2 = r * r;
            const float h2 = height * height;
            const float r2d = r2 + r2;
            const float f1 = 2.0f * r / (4.0f * r + 3.0f * height);
            const float f2 = 3.0f * height / (4.0f * r + 3.0f * height);
            const float s1 = 0.4f * r2d;
            const float s2 = 0.75f * height * r + 0.5f * h2;
            const float s3 = 0.25f * r2 + (float)(1.0 / 12.0) * h2;
            const float ixz = f1 * mass * (s1 + s2) + f2 * mass * s3;
            const float iyy = f1 * mass * s1 + f2 * mass * 0.25f * r2d;
            return v3(ixz, iyy, ixz);
        }
        case SHAPE_BOX: {  // BoxShape.cpp:53-59
            const float factor = (1.0f / 3.0f) * mass;
            const float x2 = c.params[0] * c.params[0];
            const float y2 = c.params[1] * c.params[1];
            const float z2 = c.params[2] * c.params[2];
            return v3(factor * (y2 + z2), factor * (x2 + z2), factor * (x2 + y2));
        }
        case SHAPE_HULL: {  // ConvexMeshShape::getLocalInertiaTensor: inertia of the local bounds box
            const HostHull& h = w.hulls[(int)c.params[0]];
            const float factor = (1.0f / 3.0f) * mass;
            const V3 realExtent = 0.5f * (h.local_max - h.local_min);
            const float x2 = realExtent.x * realExtent.x;
            const float y2 = realExtent.y * realExtent.y;
            const float z2 = realExtent.z * realExtent.z;
            return v3(factor * (y2 + z2), factor * (x2 + z2), factor * (x2 + y2));
        }
        case SHAPE_CONCAVE_MESH:
        case SHAPE_HEIGHTFIELD:   // ConcaveShape::getLocalInertiaTensor (ConcaveShape.h:168-175)
            return v3(mass, mass, mass);
    }
    return v3zero();
}

// RigidBody::updateMassPropertiesFromColliders (src/body/RigidBody.cpp:548-612) with
// computeCenterOfMass (:390-413) and computeMassAndInertiaTensorLocal (:418-469).
static void update_mass_properties(HostWorld& w, HostBody& b) {
    const V3 old_com_world = b.com_world;
    float total = 0.0f;
    V3 com = v3zero();
    for (int ci : b.colliders) {
        const HostCollider& c = w.colliders[ci];
        const float m = shape_volume(w, c) * c.density;
        total += m;
        com += m * c.local.p;
    }
    if (total > 0.0f) com = v3(com.x / total, com.y / total, com.z / total);
    const V3 com_world = tapply(b.transform, com);
    b.com_local = com;
    b.com_world = com_world;
    if (b.type == RP3DG_DYNAMIC) b.v += cross(b.w, com_world - old_com_world);

    M3 acc;
    acc.r[0] = acc.r[1] = acc.r[2] = v3zero();
    float total_mass = 0.0f;
    for (int ci : b.colliders) {
        const HostCollider& c = w.colliders[ci];
        const float m = shape_volume(w, c) * c.density;
        total_mass += m;
        const V3 sl = shape_local_inertia(w, c, m);
        const M3 R = qmat(c.local.q);
        M3 Rt = mtranspose(R);
        Rt.r[0] = Rt.r[0] * sl.x;
        Rt.r[1] = Rt.r[1] * sl.y;
        Rt.r[2] = Rt.r[2] * sl.z;
        const M3 I = mmul(R, Rt);
        const V3 off = c.local.p - com;
        const float off2 = len2(off);
        M3 O;
        O.r[0] = v3(off2, 0.0f, 0.0f);
        O.r[1] = v3(0.0f, off2, 0.0f);
        O.r[2] = v3(0.0f, 0.0f, off2);
        O.r[0] += off * (-off.x);
        O.r[1] += off * (-off.y);
        O.r[2] += off * (-off.z);
        for (int i = 0; i < 3; i++) O.r[i] = O.r[i] * m;
        for (int i = 0; i < 3; i++) acc.r[i] += (I.r[i] + O.r[i]);
    }
    b.inertia = v3(acc.r[0].x, acc.r[1].y, acc.r[2].z);
    if (b.type == RP3DG_DYNAMIC) {
        b.inv_inertia = v3(b.inertia.x != 0.0f ? 1.0f / b.inertia.x : 0.0f,
                           b.inertia.y != 0.0f ? 1.0f / b.inertia.y : 0.0f,
                           b.inertia.z != 0.0f ? 1.0f / b.inertia.z : 0.0f);
    }
    b.mass = total_mass;
    if (b.type == RP3DG_DYNAMIC) b.inv_mass = total_mass > 0.0f ? 1.0f / total_mass : 0.0f;
}

static bool valid_body(HostWorld* w, int32_t b) {
    if (!w || b < 0 || b >= (int32_t)w->bodies.size()) { set_error("invalid body index"); return false; }
    return true;
}
// SliderJoint.cpp:41-43: lower limit <= 0 <= upper limit, max motor force >= 0
static bool valid_limits(int32_t type, int32_t flags, const float* limits) {
    if (type != RP3DG_JOINT_SLIDER || !limits) return true;
    if ((flags & 1) && !(limits[0] <= 0.0f && limits[1] >= 0.0f)) {
        set_error("slider limits must satisfy min <= 0 <= max");
        return false;
    }
    if ((flags & 2) && !(limits[3] >= 0.0f)) { set_error("slider max motor force must be >= 0"); return false; }
    return true;
}
static bool mutable_world(HostWorld* w) {
    if (w->uploaded) { set_error("world already uploaded to the device; edit before the first update()"); return false; }
    return true;
}
// RigidBody calls after the first update go to the device batch that rp3dg_world_update steps; a world
// that shares a batch with others takes them through rp3dg_batch_body_* (global body indices).
static rp3dg_batch* stepping_batch(HostWorld* w) {
    if (!w->own) set_error("world belongs to a shared batch; use rp3dg_batch_body_* with the batch's body index");
    else w->mirror_dirty = true;
    return w->own;
}

}  // namespace rp3dg

using namespace rp3dg;

extern "C" {

const char* rp3dg_last_error(void) { return g_last_error.c_str(); }
const char* rp3dg_version(void) { return "rp3d_mi355x 0.1 (ReactPhysics3D 0.10.2 step path)"; }

void rp3dg_world_settings_default(rp3dg_world_settings* s) {
    if (!s) return;
    s->gravity[0] = 0.0f; s->gravity[1] = -9.81f; s->gravity[2] = 0.0f;
    s->persistent_contact_distance_threshold = 0.03f;
    s->default_friction_coefficient = 0.3f;
    s->default_bounciness = 0.5f;
    s->restitution_velocity_threshold = 0.5f;
    s->cos_angle_similar_contact_manifold = 0.95f;
    s->default_velocity_solver_iterations = 6;
    s->default_position_solver_iterations = 3;
    s->is_sleeping_enabled = 1;
    s->is_gravity_enabled = 1;
    s->time_before_sleep = 1.0f;
    s->sleep_linear_velocity = 0.02f;
    s->sleep_angular_velocity = 3.0f * (3.141592653589f / 180.0f);
}

rp3dg_world* rp3dg_world_create(const rp3dg_world_settings* s) {
    rp3dg_world* w = new rp3dg_world();
    if (s) w->cfg = *s;
    else rp3dg_world_settings_default(&w->cfg);
    return w;
}

int rp3dg_world_set_iterations(rp3dg_world* w, uint32_t vi, uint32_t pi) {
    if (!w) return RP3DG_EINVAL;
    w->cfg.default_velocity_solver_iterations = vi;
    w->cfg.default_position_solver_iterations = pi;
    return RP3DG_OK;
}

// PhysicsWorld::setContactsPositionCorrectionTechnique (PhysicsWorld.h:599-607)
int rp3dg_world_set_contacts_position_correction(rp3dg_world* w, int32_t technique) {
    if (!w) return RP3DG_EINVAL;
    if (technique != RP3DG_CONTACTS_BAUMGARTE && technique != RP3DG_CONTACTS_SPLIT_IMPULSES) {
        set_error("unknown contacts position correction technique");
        return RP3DG_EINVAL;
    }
    if (w->uploaded && !w->own) {
        set_error("world belongs to a shared batch; use rp3dg_batch_set_contacts_position_correction");
        return RP3DG_ESTATE;
    }
    w->split_impulse = technique == RP3DG_CONTACTS_SPLIT_IMPULSES;
    return w->own ? rp3dg_batch_set_contacts_position_correction(w->own, technique) : RP3DG_OK;
}

int rp3dg_world_set_gravity(rp3dg_world* w, const float g[3]) {
    if (!w || !g) return RP3DG_EINVAL;
    w->cfg.gravity[0] = g[0]; w->cfg.gravity[1] = g[1]; w->cfg.gravity[2] = g[2];
    return RP3DG_OK;
}

int32_t rp3dg_body_create(rp3dg_world* w, const float pos[3], const float q[4]) {
    if (!w || !pos || !q) { set_error("null argument"); return RP3DG_EINVAL; }
    if (!mutable_world(w)) return RP3DG_ESTATE;
    HostBody b;
    b.transform = tf(v3(pos[0], pos[1], pos[2]), q4(q[0], q[1], q[2], q[3]));
    b.com_world = b.transform.p;
    b.com_local = v3zero();
    b.v = b.w = b.force = b.torque = v3zero();
    // PhysicsWorld.cpp:431 sets the inverse mass from the default mass 1
    b.inv_mass = 1.0f / b.mass;
    w->bodies.push_back(b);
    return (int32_t)w->bodies.size() - 1;
}

int rp3dg_body_set_type(rp3dg_world* w, int32_t bi, int32_t type) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (!mutable_world(w)) return RP3DG_ESTATE;
    if (type < RP3DG_STATIC || type > RP3DG_DYNAMIC) { set_error("invalid body type"); return RP3DG_EINVAL; }
    HostBody& b = w->bodies[bi];
    if (b.type == type) return RP3DG_OK;
    b.type = type;
    // RigidBody::setType (src/body/RigidBody.cpp:66-130)
    if (type == RP3DG_STATIC) b.v = b.w = v3zero();
    if (type == RP3DG_STATIC || type == RP3DG_KINEMATIC) {
        b.inv_mass = 0.0f;
        b.inv_inertia = v3zero();
    } else {
        b.inv_mass = b.mass > 0.0f ? 1.0f / b.mass : 0.0f;
        b.inv_inertia = v3(b.inertia.x != 0.0f ? 1.0f / b.inertia.x : 0.0f,
                           b.inertia.y != 0.0f ? 1.0f / b.inertia.y : 0.0f,
                           b.inertia.z != 0.0f ? 1.0f / b.inertia.z : 0.0f);
    }
    b.force = b.torque = v3zero();
    return RP3DG_OK;
}

int rp3dg_body_set_velocity(rp3dg_world* w, int32_t bi, const float lin[3], const float ang[3]) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (w->uploaded) {
        rp3dg_batch* b = stepping_batch(w);
        return b ? rp3dg_batch_body_set_velocity(b, bi, lin, ang) : RP3DG_ESTATE;
    }
    HostBody& b = w->bodies[bi];
    if (b.type == RP3DG_STATIC) return RP3DG_OK;  // RigidBody::setLinearVelocity ignores static bodies
    if (lin) b.v = v3(lin[0], lin[1], lin[2]);
    if (ang) b.w = v3(ang[0], ang[1], ang[2]);
    return RP3DG_OK;
}

int rp3dg_body_set_damping(rp3dg_world* w, int32_t bi, float l, float a) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (!mutable_world(w)) return RP3DG_ESTATE;
    if (l < 0.f || a < 0.f) { set_error("damping must be >= 0"); return RP3DG_EINVAL; }
    w->bodies[bi].ldamp = l;
    w->bodies[bi].adamp = a;
    return RP3DG_OK;
}

int rp3dg_body_set_allowed_to_sleep(rp3dg_world* w, int32_t bi, int32_t allowed) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (!mutable_world(w)) return RP3DG_ESTATE;
    w->bodies[bi].allowed_to_sleep = allowed != 0;
    return RP3DG_OK;
}

int rp3dg_body_enable_gravity(rp3dg_world* w, int32_t bi, int32_t e) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (!mutable_world(w)) return RP3DG_ESTATE;
    w->bodies[bi].gravity = e != 0;
    return RP3DG_OK;
}

int rp3dg_body_apply_force_torque(rp3dg_world* w, int32_t bi, const float f[3], const float t[3]) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (w->uploaded) {
        rp3dg_batch* b = stepping_batch(w);
        if (!b) return RP3DG_ESTATE;
        int r = f ? rp3dg_batch_body_apply_force(b, bi, f, nullptr, 0) : RP3DG_OK;
        if (!r && t) r = rp3dg_batch_body_apply_torque(b, bi, t, 0);
        return r;
    }
    HostBody& b = w->bodies[bi];
    if (b.type != RP3DG_DYNAMIC) return RP3DG_OK;
    if (f) b.force += v3(f[0], f[1], f[2]);
    if (t) b.torque += v3(t[0], t[1], t[2]);
    return RP3DG_OK;
}

// RigidBody::apply{World,Local}ForceAt{CenterOfMass,WorldPosition,LocalPosition} (RigidBody.cpp:151-239,
// 298-311); before the first update the body is awake and only its sums change.
int rp3dg_body_apply_force(rp3dg_world* w, int32_t bi, const float f[3], const float point[3], int32_t flags) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (w->uploaded) {
        rp3dg_batch* b = stepping_batch(w);
        return b ? rp3dg_batch_body_apply_force(b, bi, f, point, flags) : RP3DG_ESTATE;
    }
    if (!f) { set_error("null force"); return RP3DG_EINVAL; }
    if (flags & ~(RP3DG_FORCE_LOCAL | RP3DG_FORCE_AT_POINT | RP3DG_FORCE_POINT_LOCAL)) { set_error("unknown force flags"); return RP3DG_EINVAL; }
    if ((flags & RP3DG_FORCE_AT_POINT) && !point) { set_error("null point"); return RP3DG_EINVAL; }
    HostBody& b = w->bodies[bi];
    V3 force = v3(f[0], f[1], f[2]);
    if (flags & RP3DG_FORCE_LOCAL) force = qrot(b.transform.q, force);
    if (b.type != RP3DG_DYNAMIC) return RP3DG_OK;
    b.force += force;
    if (flags & RP3DG_FORCE_AT_POINT) {
        V3 p = v3(point[0], point[1], point[2]);
        if (flags & RP3DG_FORCE_POINT_LOCAL) p = tapply(b.transform, p);
        b.torque += cross(p - b.com_world, force);
    }
    return RP3DG_OK;
}

// RigidBody::applyWorldTorque / applyLocalTorque (RigidBody.cpp:938-967).
int rp3dg_body_apply_torque(rp3dg_world* w, int32_t bi, const float t[3], int32_t flags) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (w->uploaded) {
        rp3dg_batch* b = stepping_batch(w);
        return b ? rp3dg_batch_body_apply_torque(b, bi, t, flags) : RP3DG_ESTATE;
    }
    if (!t) { set_error("null torque"); return RP3DG_EINVAL; }
    if (flags & ~RP3DG_FORCE_LOCAL) { set_error("unknown torque flags"); return RP3DG_EINVAL; }
    HostBody& b = w->bodies[bi];
    V3 torque = v3(t[0], t[1], t[2]);
    if (flags & RP3DG_FORCE_LOCAL) torque = qrot(b.transform.q, torque);
    if (b.type == RP3DG_DYNAMIC) b.torque += torque;
    return RP3DG_OK;
}

// RigidBody::resetForce / resetTorque (RigidBody.cpp:970-987).
int rp3dg_body_reset_force_torque(rp3dg_world* w, int32_t bi, int32_t force, int32_t torque) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (w->uploaded) {
        rp3dg_batch* b = stepping_batch(w);
        return b ? rp3dg_batch_body_reset_force_torque(b, bi, force, torque) : RP3DG_ESTATE;
    }
    HostBody& b = w->bodies[bi];
    if (b.type != RP3DG_DYNAMIC) return RP3DG_OK;
    if (force) b.force = v3zero();
    if (torque) b.torque = v3zero();
    return RP3DG_OK;
}

// RigidBody::setTransform (RigidBody.cpp:837-866): before the first update the centre of mass follows
// and the linear velocity gains angular x (new - old centre of mass); colliders are placed at upload.
int rp3dg_body_set_transform(rp3dg_world* w, int32_t bi, const float pos[3], const float q[4]) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (!pos || !q) { set_error("null argument"); return RP3DG_EINVAL; }
    if (w->uploaded) {
        rp3dg_batch* b = stepping_batch(w);
        return b ? rp3dg_batch_body_set_transform(b, bi, pos, q) : RP3DG_ESTATE;
    }
    HostBody& b = w->bodies[bi];
    const V3 old = b.com_world;
    b.transform = tf(v3(pos[0], pos[1], pos[2]), q4(q[0], q[1], q[2], q[3]));
    b.com_world = tapply(b.transform, b.com_local);
    b.v += cross(b.w, b.com_world - old);
    return RP3DG_OK;
}

// RigidBody::setIsSleeping (RigidBody.cpp:1006-1057) after the first update (bodies are created awake).
int rp3dg_body_set_sleeping(rp3dg_world* w, int32_t bi, int32_t sleeping) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (!w->uploaded) {
        if (!sleeping) return RP3DG_OK;
        set_error("bodies are awake until the first update(); put a body to sleep after it");
        return RP3DG_EUNSUPPORTED;
    }
    rp3dg_batch* b = stepping_batch(w);
    return b ? rp3dg_batch_body_set_sleeping(b, bi, sleeping) : RP3DG_ESTATE;
}

int rp3dg_body_get_force_torque(rp3dg_world* w, int32_t bi, float f[3], float t[3]) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (!w->uploaded) {
        const HostBody& b = w->bodies[bi];
        if (f) { f[0] = b.force.x; f[1] = b.force.y; f[2] = b.force.z; }
        if (t) { t[0] = b.torque.x; t[1] = b.torque.y; t[2] = b.torque.z; }
        return RP3DG_OK;
    }
    if (!w->own) { set_error("world belongs to a shared batch; use rp3dg_batch_get_force_torque"); return RP3DG_ESTATE; }
    const size_t n = w->bodies.size();
    std::vector<float> F(3 * n), T(3 * n);
    const int r = rp3dg_batch_get_force_torque(w->own, F.data(), T.data());
    if (r) return r;
    for (int k = 0; k < 3; k++) {
        if (f) f[k] = F[3 * bi + k];
        if (t) t[k] = T[3 * bi + k];
    }
    return RP3DG_OK;
}

int rp3dg_body_set_lock_factors(rp3dg_world* w, int32_t bi, const float l[3], const float a[3]) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (!mutable_world(w)) return RP3DG_ESTATE;
    if (l) w->bodies[bi].llock = v3(l[0], l[1], l[2]);
    if (a) w->bodies[bi].alock = v3(a[0], a[1], a[2]);
    return RP3DG_OK;
}

int rp3dg_body_update_mass_properties(rp3dg_world* w, int32_t bi) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (!mutable_world(w)) return RP3DG_ESTATE;
    update_mass_properties(*w, w->bodies[bi]);
    return RP3DG_OK;
}

int rp3dg_body_set_mass_properties(rp3dg_world* w, int32_t bi, float mass, const float il[3], const float com[3]) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (!mutable_world(w)) return RP3DG_ESTATE;
    HostBody& b = w->bodies[bi];
    if (mass <= 0.f) { set_error("mass must be > 0"); return RP3DG_EINVAL; }
    b.mass = mass;
    if (b.type == RP3DG_DYNAMIC) b.inv_mass = 1.0f / mass;
    if (il) {
        b.inertia = v3(il[0], il[1], il[2]);
        if (b.type == RP3DG_DYNAMIC)
            b.inv_inertia = v3(il[0] != 0.f ? 1.0f / il[0] : 0.f, il[1] != 0.f ? 1.0f / il[1] : 0.f,
                               il[2] != 0.f ? 1.0f / il[2] : 0.f);
    }
    if (com) {
        const V3 old = b.com_world;
        b.com_local = v3(com[0], com[1], com[2]);
        b.com_world = tapply(b.transform, b.com_local);
        if (b.type == RP3DG_DYNAMIC) b.v += cross(b.w, b.com_world - old);
    }
    return RP3DG_OK;
}

int32_t rp3dg_collider_add(rp3dg_world* w, int32_t bi, int32_t shape, const float params[4], const float lp[3],
                           const float lq[4]) {
    if (!valid_body(w, bi)) return RP3DG_EINVAL;
    if (!mutable_world(w)) return RP3DG_ESTATE;
    if (!params) { set_error("null shape parameters"); return RP3DG_EINVAL; }
    HostCollider c;
    c.body = bi;
    c.type = shape;
    for (int i = 0; i < 4; i++) c.params[i] = params[i];
    switch (shape) {
        case SHAPE_SPHERE:
            if (!(params[0] > 0.f)) { set_error("sphere radius must be > 0"); return RP3DG_EINVAL; }
            break;
        case SHAPE_CAPSULE:
            if (!(params[0] > 0.f) || !(params[1] > 0.f)) { set_error("capsule radius/height must be > 0"); return RP3DG_EINVAL; }
            break;
        case SHAPE_BOX:
            if (!(params[0] > 0.f) || !(params[1] > 0.f) || !(params[2] > 0.f)) {
                set_error("box half-extents must be > 0");
                return RP3DG_EINVAL;
            }
            break;
        case SHAPE_HULL:
            if ((int)params[0] < 0 || (int)params[0] >= (int)w->hulls.size()) { set_error("invalid hull index"); return RP3DG_EINVAL; }
            break;
        case SHAPE_CONCAVE_MESH:
        case SHAPE_HEIGHTFIELD: {
            const int n = shape == SHAPE_CONCAVE_MESH ? (int)w->meshes.size() : (int)w->heightfields.size();
            if ((int)params[0] < 0 || (int)params[0] >= n || params[0] != (float)(int)params[0]) {
                set_error(shape == SHAPE_CONCAVE_MESH ? "invalid triangle mesh index" : "invalid height field index");
                return RP3DG_EINVAL;
            }
            if (!(params[1] > 0.f) || !(params[2] > 0.f) || !(params[3] > 0.f)) {
                set_error("concave shape scale must be > 0");
                return RP3DG_EINVAL;
            }
            // a concave shape collides only with convex ones and is meant for bodies that do not respond to
            // contacts (the reference's user manual: static bodies)
            if (w->bodies[bi].type == 2) {
                set_error("concave shapes (triangle mesh, height field) go on static or kinematic bodies only");
                return RP3DG_EUNSUPPORTED;
            }
            break;
        }
        default:
            set_error("unsupported collision shape");
            return RP3DG_EUNSUPPORTED;
    }
    c.local = tf(lp ? v3(lp[0], lp[1], lp[2]) : v3zero(), lq ? q4(lq[0], lq[1], lq[2], lq[3]) : q4(0, 0, 0, 1));
    // RigidBody::addCollider: Material(defaultFriction, defaultBounciness), density 1 (Material.cpp:32)
    c.friction_sqrt = sqrtf(w->cfg.default_friction_coefficient);
    c.bounciness = w->cfg.default_bounciness;
    c.density = 1.0f;
    w->colliders.push_back(c);
    const int32_t ci = (int32_t)w->colliders.size() - 1;
    w->bodies[bi].colliders.push_back(ci);
    return ci;
}

int rp3dg_collider_set_material(rp3dg_world* w, int32_t ci, float f, float b, float d) {
    if (!w || ci < 0 || ci >= (int32_t)w->colliders.size()) { set_error("invalid collider"); return RP3DG_EINVAL; }
    if (!mutable_world(w)) return RP3DG_ESTATE;
    if (f < 0.f || b < 0.f || b > 1.f || d <= 0.f) { set_error("invalid material"); return RP3DG_EINVAL; }
    w->colliders[ci].friction_sqrt = sqrtf(f);  // Material::setFrictionCoefficient (Material.h:130-133)
    w->colliders[ci].bounciness = b;
    w->colliders[ci].density = d;
    return RP3DG_OK;
}

int rp3dg_collider_set_filter(rp3dg_world* w, int32_t ci, uint16_t cat, uint16_t mask) {
    if (!w || ci < 0 || ci >= (int32_t)w->colliders.size()) { set_error("invalid collider"); return RP3DG_EINVAL; }
    if (!mutable_world(w)) return RP3DG_ESTATE;
    w->colliders[ci].category = cat;
    w->colliders[ci].mask = mask;
    return RP3DG_OK;
}

// Collider::setIsTrigger (src/collision/Collider.cpp:250-257): a trigger stops being a simulation collider.
int rp3dg_collider_set_is_trigger(rp3dg_world* w, int32_t ci, int32_t is_trigger) {
    if (!w || ci < 0 || ci >= (int32_t)w->colliders.size()) { set_error("invalid collider"); return RP3DG_EINVAL; }
    if (!mutable_world(w)) return RP3DG_ESTATE;
    int& k = w->colliders[ci].kind;
    k = is_trigger ? (k | CK_TRIGGER) : (k & ~CK_TRIGGER);
    if (is_trigger) k &= ~CK_SIMULATION;
    return RP3DG_OK;
}

// Collider::setIsSimulationCollider (src/collision/Collider.cpp:272-287): a simulation collider stops being a
// trigger.
int rp3dg_collider_set_is_simulation_collider(rp3dg_world* w, int32_t ci, int32_t is_simulation) {
    if (!w || ci < 0 || ci >= (int32_t)w->colliders.size()) { set_error("invalid collider"); return RP3DG_EINVAL; }
    if (!mutable_world(w)) return RP3DG_ESTATE;
    int& k = w->colliders[ci].kind;
    k = is_simulation ? (k | CK_SIMULATION) : (k & ~CK_SIMULATION);
    if (is_simulation) k &= ~CK_TRIGGER;
    return RP3DG_OK;
}

// Collider::setIsWorldQueryCollider (src/collision/Collider.cpp:298-300).
int rp3dg_collider_set_is_world_query_collider(rp3dg_world* w, int32_t ci, int32_t is_query) {
    if (!w || ci < 0 || ci >= (int32_t)w->colliders.size()) { set_error("invalid collider"); return RP3DG_EINVAL; }
    if (!mutable_world(w)) return RP3DG_ESTATE;
    int& k = w->colliders[ci].kind;
    k = is_query ? (k | CK_WORLD_QUERY) : (k & ~CK_WORLD_QUERY);
    return RP3DG_OK;
}

// Collider::getIsTrigger / getIsSimulationCollider / getIsWorldQueryCollider as RP3DG_COLLIDER_* bits.
int rp3dg_collider_get_kind(rp3dg_world* w, int32_t ci) {
    if (!w || ci < 0 || ci >= (int32_t)w->colliders.size()) { set_error("invalid collider"); return RP3DG_EINVAL; }
    return w->colliders[ci].kind;
}

// VertexArray (src/collision/VertexArray.cpp) read as QuickHull::extractPoints does: 3 floats or 3 doubles
// per vertex at `stride` bytes, doubles converted to float.
static bool read_points(const void* data, int32_t n, int32_t stride, int32_t is_double, std::vector<V3>& out) {
    const int32_t need = is_double ? 3 * (int32_t)sizeof(double) : 3 * (int32_t)sizeof(float);
    if (!data || n < 0 || stride < need) { set_error("invalid vertex array (null data, negative count or stride too small)"); return false; }
    const unsigned char* base = static_cast<const unsigned char*>(data);
    out.resize(n);
    for (int32_t i = 0; i < n; i++) {
        const unsigned char* p = base + (size_t)i * (size_t)stride;
        if (is_double) {
            double d[3];
            std::memcpy(d, p, sizeof d);
            out[i] = v3((float)d[0], (float)d[1], (float)d[2]);
        } else {
            float f[3];
            std::memcpy(f, p, sizeof f);
            out[i] = v3(f[0], f[1], f[2]);
        }
    }
    return true;
}

int rp3dg_quickhull(const void* vertices, int32_t nb_vertices, int32_t stride, int32_t is_double, float* out_vertices,
                    int32_t cap_vertices, int32_t* face_counts, int32_t cap_faces, int32_t* face_indices,
                    int32_t cap_indices, int32_t* nb_out_vertices, int32_t* nb_out_faces) {
    std::vector<V3> pts;
    if (!read_points(vertices, nb_vertices, stride, is_double, pts)) return RP3DG_EINVAL;
    std::vector<float> v;
    std::vector<int32_t> c, idx;
    std::string err;
    if (!quickhull_polygons(pts, v, c, idx, &err)) { set_error(err); return RP3DG_EINVAL; }
    const int32_t nv = (int32_t)(v.size() / 3), nf = (int32_t)c.size(), ni = (int32_t)idx.size();
    if (nb_out_vertices) *nb_out_vertices = nv;
    if (nb_out_faces) *nb_out_faces = nf;
    if (nv > cap_vertices || nf > cap_faces || ni > cap_indices || !out_vertices || !face_counts || !face_indices) {
        set_error("quickhull: output capacity too small (" + std::to_string(nv) + " vertices, " + std::to_string(nf) +
                  " faces, " + std::to_string(ni) + " indices)");
        return RP3DG_EINVAL;
    }
    std::memcpy(out_vertices, v.data(), sizeof(float) * v.size());
    std::memcpy(face_counts, c.data(), sizeof(int32_t) * c.size());
    std::memcpy(face_indices, idx.data(), sizeof(int32_t) * idx.size());
    return RP3DG_OK;
}

int32_t rp3dg_hull_create_from_points(rp3dg_world* w, const void* vertices, int32_t nb_vertices, int32_t stride,
                                      int32_t is_double) {
    if (!w) { set_error("null world"); return RP3DG_EINVAL; }
    std::vector<V3> pts;
    if (!read_points(vertices, nb_vertices, stride, is_double, pts)) return RP3DG_EINVAL;
    std::vector<float> v;
    std::vector<int32_t> c, idx;
    std::string err;
    if (!quickhull_polygons(pts, v, c, idx, &err)) { set_error(err); return RP3DG_EINVAL; }
    return rp3dg_hull_create(w, v.data(), (int32_t)(v.size() / 3), c.data(), idx.data(), (int32_t)c.size());
}

// ConvexMesh::init from a PolygonVertexArray (src/collision/ConvexMesh.cpp:55-264): centroid,
// bounds, half-edge structure, Newell face normals, volume.
int32_t rp3dg_hull_create(rp3dg_world* w, const float* verts, int32_t nv, const int32_t* counts,
                          const int32_t* idx, int32_t nf) {
    if (!w || !verts || !counts || !idx || nv < 4 || nf < 4) { set_error("invalid hull arguments"); return RP3DG_EINVAL; }
    if (!mutable_world(w)) return RP3DG_ESTATE;
    HostHull h;
    h.centroid = v3zero();
    h.local_min = h.local_max = v3(verts[0], verts[1], verts[2]);
    for (int i = 0; i < nv; i++) {
        const V3 p = v3(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
        h.vertices.push_back(p);
        h.centroid += p;
        if (p.x > h.local_max.x) h.local_max.x = p.x;
        if (p.x < h.local_min.x) h.local_min.x = p.x;
        if (p.y > h.local_max.y) h.local_max.y = p.y;
        if (p.y < h.local_min.y) h.local_min.y = p.y;
        if (p.z > h.local_max.z) h.local_max.z = p.z;
        if (p.z < h.local_min.z) h.local_min.z = p.z;
    }
    h.centroid = v3(h.centroid.x / (float)nv, h.centroid.y / (float)nv, h.centroid.z / (float)nv);
    int off = 0;
    int nedges = 0;
    for (int f = 0; f < nf; f++) {
        if (counts[f] < 3) { set_error("hull face with less than three vertices"); return RP3DG_EINVAL; }
        if (counts[f] > kMaxHullFaceVertices) {
            // the face-contact clipping keeps at most 2 * kMaxHullFaceVertices vertices per polygon
            set_error("hull face with more than " + std::to_string(kMaxHullFaceVertices) + " vertices");
            return RP3DG_EUNSUPPORTED;
        }
        std::vector<int> fv(idx + off, idx + off + counts[f]);
        off += counts[f];
        nedges += counts[f];
        h.mesh.faces.push_back(fv);
    }
    if (2 + nedges / 2 - nf != nv) { set_error("hull violates Euler's formula (duplicated vertices?)"); return RP3DG_EINVAL; }
    std::string err;
    if (!h.mesh.build(nv, &err)) { set_error("hull topology: " + err); return RP3DG_EINVAL; }
    if ((int)h.mesh.edges.size() > 255) { set_error("hull has more than 255 half-edges"); return RP3DG_EUNSUPPORTED; }
    auto newell = [&](int f) {
        const std::vector<int>& fv = h.mesh.faces[f];
        V3 n = v3zero();
        for (size_t i = fv.size() - 1, j = 0; j < fv.size(); i = j, j++) {
            const V3& a = h.vertices[fv[i]];
            const V3& b = h.vertices[fv[j]];
            n += v3((a.y - b.y) * (a.z + b.z), (a.z - b.z) * (a.x + b.x), (a.x - b.x) * (a.y + b.y));
        }
        return n;
    };
    for (int f = 0; f < nf; f++) {
        V3 n = newell(f);
        const float l = len(n);
        if (!(l > kMachineEps)) { set_error("hull face with zero area"); return RP3DG_EINVAL; }
        n = v3(n.x / l, n.y / l, n.z / l);
        // ConvexMeshShape::computeScaledFacesNormals with scale 1: n * (1/1), renormalised
        const V3 s = v3((float)(1.0 / 1.0) * n.x, (float)(1.0 / 1.0) * n.y, (float)(1.0 / 1.0) * n.z);
        const float sl = len(s);
        h.face_normals.push_back(v3(s.x / sl, s.y / sl, s.z / sl));
    }
    float sum = 0.0f;
    for (int f = 0; f < nf; f++) {
        const float area = len(newell(f)) * 0.5f;
        V3 n = newell(f);
        const float l = len(n);
        n = v3(n.x / l, n.y / l, n.z / l);
        sum += dot(h.vertices[h.mesh.faces[f][0]], n) * area;
    }
    h.volume = fabsf(sum) / 3.0f;
    w->hulls.push_back(h);
    return (int32_t)w->hulls.size() - 1;
}

int32_t rp3dg_joint_create(rp3dg_world* w, int32_t type, int32_t b1, int32_t b2, const float anchor[3],
                           const float axis[3], int32_t collide, int32_t flags, const float limits[4]) {
    if (!valid_body(w, b1) || !valid_body(w, b2)) return RP3DG_EINVAL;
    if (!mutable_world(w)) return RP3DG_ESTATE;
    if (b1 == b2) { set_error("a joint needs two different bodies"); return RP3DG_EINVAL; }
    if (type < RP3DG_JOINT_BALLSOCKET || type > RP3DG_JOINT_FIXED) { set_error("invalid joint type"); return RP3DG_EINVAL; }
    if (!valid_limits(type, flags, limits)) return RP3DG_EINVAL;
    HostJoint j;
    j.type = type;
    j.body1 = b1;
    j.body2 = b2;
    j.anchor_world = anchor ? v3(anchor[0], anchor[1], anchor[2]) : v3zero();
    j.axis_world = axis ? v3(axis[0], axis[1], axis[2]) : v3(0, 0, 1);
    j.collide_connected = collide != 0;
    j.flags = flags;
    for (int i = 0; i < 4; i++) j.limits[i] = limits ? limits[i] : 0.0f;
    w->joints.push_back(j);
    return (int32_t)w->joints.size() - 1;
}

int rp3dg_joint_set_limits(rp3dg_world* w, int32_t joint, int32_t flags, const float limits[4]) {
    if (!w || joint < 0 || joint >= (int32_t)w->joints.size()) { set_error("invalid joint index"); return RP3DG_EINVAL; }
    if (!mutable_world(w)) return RP3DG_ESTATE;
    HostJoint& j = w->joints[joint];
    if (!valid_limits(j.type, flags, limits)) return RP3DG_EINVAL;
    j.flags = (flags & 3) | (j.flags & RP3DG_JOINT_BAUMGARTE);
    for (int i = 0; i < 4; i++) j.limits[i] = limits ? limits[i] : 0.0f;
    return RP3DG_OK;
}

}  // extern "C"
