// world.cpp -- the drop-in b3World over the C ABI of libbxstep (include/bxstep.h).
//
// The world keeps a host mirror of its bodies, fixtures and joints in creation order. When the set
// of bodies, fixtures or joints changes, Step uploads the mirror as one bx_world_desc (the current
// poses and velocities as the creation state); when only velocities were set, it writes them to
// the resident world; then it steps on the GPU and reads back the state the getters return.
// Uploading again after the world has stepped starts its contacts afresh (warm starting restarts),
// which the reference does not do; worlds built before their first step match it bit for bit.
#include <cstring>
#include <string>
#include <unordered_map>
#include "bounce/bounce.h"
#include "bxstep.h"

// ------------------------------------------------------------------ b3BoxHull
// The box topology of the reference's b3BoxHull::SetIdentity (box_hull.h:41-121): 8 vertices,
// 24 half-edges {origin, twin, face, prev, next}, 6 faces. Feature ids are half-edge indices, so the
// table is part of the output and must match.
namespace {
const float kBoxV[8][3] = {{-1, -1, 1}, {1, -1, -1}, {1, 1, 1}, {-1, 1, -1}, {1, 1, -1}, {-1, 1, 1}, {1, -1, 1}, {-1, -1, -1}};
const uint32 kBoxE[24][5] = {
    {0, 1, 0, 6, 2},   {5, 0, 1, 12, 8},  {5, 3, 0, 0, 4},   {3, 2, 5, 19, 13}, {3, 5, 0, 2, 6},   {7, 4, 3, 15, 18},
    {7, 7, 0, 4, 0},   {0, 6, 2, 9, 14},  {0, 9, 1, 1, 10},  {6, 8, 2, 16, 7},  {6, 11, 1, 8, 12}, {2, 10, 4, 22, 17},
    {2, 13, 1, 10, 1}, {5, 12, 5, 3, 23}, {7, 15, 2, 7, 16}, {1, 14, 3, 20, 5}, {1, 17, 2, 14, 9}, {6, 16, 4, 11, 21},
    {3, 19, 3, 5, 20}, {4, 18, 5, 23, 3}, {4, 21, 3, 18, 15}, {1, 20, 4, 17, 22}, {4, 23, 4, 21, 11}, {2, 22, 5, 13, 19}};
const uint32 kBoxF[6] = {0, 8, 7, 15, 21, 23};
const float kBoxP[6][4] = {{-1, 0, 0, 1}, {0, 0, 1, 1}, {0, -1, 0, 1}, {0, 0, -1, 1}, {1, 0, 0, 1}, {0, 1, 0, 1}};
}  // namespace

void b3BoxHull::SetIdentity() { SetExtents(scalar(1), scalar(1), scalar(1)); }

void b3BoxHull::SetExtents(scalar ex, scalar ey, scalar ez) {
    // the identity box scaled by (ex, ey, ez) (b3Hull::Scale, hull.cpp:120-150): vertices scale,
    // plane normals scale by the inverse and are renormalized, offsets divide by the length.
    const scalar s[3] = {ex, ey, ez};
    const scalar inv[3] = {scalar(1) / ex, scalar(1) / ey, scalar(1) / ez};
    for (int i = 0; i < 8; ++i) boxVertices[i].Set(kBoxV[i][0] * s[0], kBoxV[i][1] * s[1], kBoxV[i][2] * s[2]);
    for (int i = 0; i < 24; ++i) {
        boxEdges[i].origin = kBoxE[i][0];
        boxEdges[i].twin = kBoxE[i][1];
        boxEdges[i].face = kBoxE[i][2];
        boxEdges[i].prev = kBoxE[i][3];
        boxEdges[i].next = kBoxE[i][4];
    }
    for (int i = 0; i < 6; ++i) {
        boxFaces[i].edge = kBoxF[i];
        b3Vec3 n(inv[0] * kBoxP[i][0], inv[1] * kBoxP[i][1], inv[2] * kBoxP[i][2]);
        scalar len = n.Normalize();
        boxPlanes[i].normal = n;
        boxPlanes[i].offset = kBoxP[i][3] / len;
    }
    centroid.SetZero();
    vertexCount = 8;
    vertices = boxVertices;
    edgeCount = 24;
    edges = boxEdges;
    faceCount = 6;
    faces = boxFaces;
    planes = boxPlanes;
}

// ------------------------------------------------------------------ device state of a world
struct b3WorldDevice {
    bx_context* ctx = nullptr;
    bx_batch* batch = nullptr;
    std::string error;
};

namespace {
void Copy3(float* out, const b3Vec3& v) { out[0] = v.x; out[1] = v.y; out[2] = v.z; }
void Copy4(float* out, const b3Quat& q) { out[0] = q.v.x; out[1] = q.v.y; out[2] = q.v.z; out[3] = q.s; }
}  // namespace

// ------------------------------------------------------------------ b3Fixture / b3Body
b3Fixture::~b3Fixture() { delete m_shape; }

b3Body::b3Body(const b3BodyDef& def, b3World* world, uint32 index)
    : m_def(def), m_type(def.type), m_world(world), m_index(index), m_xf(def.position, def.orientation),
      m_worldCenter(def.position), m_linearVelocity(def.linearVelocity), m_angularVelocity(def.angularVelocity),
      m_awake(def.awake), m_sleepTime(0), m_fixtureList(nullptr), m_next(nullptr), m_userData(def.userData) {
    if (m_type == e_staticBody) {
        m_linearVelocity.SetZero();
        m_angularVelocity.SetZero();
    }
}

b3Body::~b3Body() {
    b3Fixture* f = m_fixtureList;
    while (f) {
        b3Fixture* n = f->m_next;
        delete f;
        f = n;
    }
}

b3Fixture* b3Body::CreateFixture(const b3FixtureDef& def) {
    if (!def.shape) return nullptr;
    b3Fixture* f = new b3Fixture();
    switch (def.shape->GetType()) {
    case b3Shape::e_sphere: f->m_shape = new b3SphereShape(*static_cast<const b3SphereShape*>(def.shape)); break;
    case b3Shape::e_capsule: f->m_shape = new b3CapsuleShape(*static_cast<const b3CapsuleShape*>(def.shape)); break;
    case b3Shape::e_hull: f->m_shape = new b3HullShape(*static_cast<const b3HullShape*>(def.shape)); break;
    case b3Shape::e_triangle:
        f->m_shape = new b3TriangleShape(*static_cast<const b3TriangleShape*>(def.shape));
        break;
    case b3Shape::e_mesh: f->m_shape = new b3MeshShape(*static_cast<const b3MeshShape*>(def.shape)); break;
    default: delete f; return nullptr;
    }
    f->m_type = def.shape->GetType();
    f->m_body = this;
    f->m_isSensor = def.isSensor;
    f->m_density = def.density;
    f->m_friction = def.friction;
    f->m_restitution = def.restitution;
    f->m_userData = def.userData;
    f->m_index = uint32(m_world->m_fixtures.size());
    f->m_next = m_fixtureList;  // the reference prepends (body.cpp:120-122)
    m_fixtureList = f;
    m_world->m_fixtures.push_back(f);
    m_world->MarkTopologyChanged();
    return f;
}

void b3Body::SetLinearVelocity(const b3Vec3& v) {
    if (m_type == e_staticBody) return;
    m_linearVelocity = v;
    if (b3Dot(v, v) > scalar(0)) m_awake = true;
    m_world->MarkStateChanged();
}

void b3Body::SetAngularVelocity(const b3Vec3& w) {
    if (m_type == e_staticBody) return;
    m_angularVelocity = w;
    if (b3Dot(w, w) > scalar(0)) m_awake = true;
    m_world->MarkStateChanged();
}

// ------------------------------------------------------------------ b3World
b3World::b3World()
    : m_gravity(scalar(0), scalar(-9.8), scalar(0)), m_sleeping(false), m_warmStarting(true), m_autoClearForces(true),
      m_bodyList(nullptr), m_jointList(nullptr), m_contactCount(0), m_topologyChanged(true), m_stateChanged(false),
      m_device(new b3WorldDevice()) {}

b3World::~b3World() {
    if (m_device->batch) bx_batch_destroy(m_device->batch);
    if (m_device->ctx) bx_context_destroy(m_device->ctx);
    delete m_device;
    for (b3Joint* j : m_joints) delete j;
    for (b3Body* b : m_bodies) delete b;
}

void b3World::SetSleeping(bool flag) { m_sleeping = flag; MarkTopologyChanged(); }
void b3World::SetWarmStarting(bool flag) { m_warmStarting = flag; MarkTopologyChanged(); }
void b3World::SetGravity(const b3Vec3& g) { m_gravity = g; MarkTopologyChanged(); }
void b3World::SetAutoClearForces(bool flag) { m_autoClearForces = flag; MarkTopologyChanged(); }
const char* b3World::GetLastError() const { return m_device->error.c_str(); }

b3Body* b3World::CreateBody(const b3BodyDef& def) {
    b3Body* b = new b3Body(def, this, uint32(m_bodies.size()));
    b->m_next = m_bodyList;  // world.cpp:94-110 prepends
    m_bodyList = b;
    m_bodies.push_back(b);
    MarkTopologyChanged();
    return b;
}

void b3World::DestroyBody(b3Body* body) {
    if (!body) return;
    // joints attached to the body go with it (world.cpp DestroyBody)
    for (size_t i = 0; i < m_joints.size();) {
        if (m_joints[i]->m_def.bodyA == body || m_joints[i]->m_def.bodyB == body) DestroyJoint(m_joints[i]);
        else ++i;
    }
    for (size_t i = 0; i < m_fixtures.size();) {
        if (m_fixtures[i]->m_body == body) m_fixtures.erase(m_fixtures.begin() + i);
        else ++i;
    }
    b3Body** link = &m_bodyList;
    while (*link && *link != body) link = &(*link)->m_next;
    if (*link) *link = body->m_next;
    for (size_t i = 0; i < m_bodies.size(); ++i)
        if (m_bodies[i] == body) { m_bodies.erase(m_bodies.begin() + i); break; }
    delete body;
    for (uint32 i = 0; i < m_bodies.size(); ++i) m_bodies[i]->m_index = i;
    for (uint32 i = 0; i < m_fixtures.size(); ++i) m_fixtures[i]->m_index = i;
    MarkTopologyChanged();
}

b3Joint* b3World::CreateJoint(const b3JointDef& def) {
    if (def.type != e_revoluteJoint || !def.bodyA || !def.bodyB) return nullptr;
    b3Joint* j = new b3Joint();
    j->m_def = static_cast<const b3RevoluteJointDef&>(def);
    j->m_next = m_jointList;
    m_jointList = j;
    m_joints.push_back(j);
    MarkTopologyChanged();
    return j;
}

void b3World::DestroyJoint(b3Joint* joint) {
    b3Joint** link = &m_jointList;
    while (*link && *link != joint) link = &(*link)->m_next;
    if (*link) *link = joint->m_next;
    for (size_t i = 0; i < m_joints.size(); ++i)
        if (m_joints[i] == joint) { m_joints.erase(m_joints.begin() + i); break; }
    delete joint;
    MarkTopologyChanged();
}

int b3World::Upload() {
    // shape library: every distinct b3Hull once
    std::vector<const b3Hull*> hulls;
    std::unordered_map<const b3Hull*, uint32> hullIndex;
    for (b3Fixture* f : m_fixtures) {
        if (f->m_type != b3Shape::e_hull) continue;
        const b3Hull* h = static_cast<const b3HullShape*>(f->m_shape)->m_hull;
        if (hullIndex.emplace(h, uint32(hulls.size())).second) hulls.push_back(h);
    }
    std::vector<bx_hull_desc> hd(hulls.size());
    for (size_t i = 0; i < hulls.size(); ++i) {
        const b3Hull* h = hulls[i];
        hd[i].vertex_count = h->vertexCount;
        hd[i].edge_count = h->edgeCount;
        hd[i].face_count = h->faceCount;
        hd[i].vertices = &h->vertices[0].x;
        hd[i].edges = &h->edges[0].origin;
        hd[i].faces = &h->faces[0].edge;
        hd[i].planes = &h->planes[0].normal.x;
        Copy3(hd[i].centroid, h->centroid);
    }
    // every distinct b3Mesh once; the descriptor takes the triangles' vertex indices only
    std::vector<const b3Mesh*> meshes;
    std::unordered_map<const b3Mesh*, uint32> meshIndex;
    for (b3Fixture* f : m_fixtures) {
        if (f->m_type != b3Shape::e_mesh) continue;
        const b3Mesh* m = static_cast<const b3MeshShape*>(f->m_shape)->m_mesh;
        if (meshIndex.emplace(m, uint32(meshes.size())).second) meshes.push_back(m);
    }
    std::vector<std::vector<uint32>> meshTris(meshes.size());
    std::vector<bx_mesh_desc> md(meshes.size());
    for (size_t i = 0; i < meshes.size(); ++i) {
        const b3Mesh* m = meshes[i];
        std::vector<uint32>& t = meshTris[i];
        t.resize(size_t(3) * m->triangleCount);
        for (uint32 k = 0; k < m->triangleCount; ++k) {
            t[3 * k] = m->triangles[k].v1;
            t[3 * k + 1] = m->triangles[k].v2;
            t[3 * k + 2] = m->triangles[k].v3;
        }
        md[i].vertex_count = m->vertexCount;
        md[i].triangle_count = m->triangleCount;
        md[i].vertices = m->vertexCount ? &m->vertices[0].x : nullptr;
        md[i].triangles = t.data();
    }
    bx_shape_library lib;
    std::memset(&lib, 0, sizeof(lib));
    lib.hull_count = uint32(hd.size());
    lib.hulls = hd.data();
    lib.mesh_count = uint32(md.size());
    lib.meshes = md.data();
    std::vector<bx_body_desc> bd(m_bodies.size());
    for (size_t i = 0; i < m_bodies.size(); ++i) {
        const b3Body* b = m_bodies[i];
        bx_body_desc& d = bd[i];
        std::memset(&d, 0, sizeof(d));
        d.type = uint32(b->m_type);
        d.flags = (b->m_awake ? BX_BODY_AWAKE : 0u) | (b->m_def.allowSleep ? BX_BODY_AUTO_SLEEP : 0u) |
                  (b->m_def.fixedRotationX ? BX_BODY_FIXED_ROT_X : 0u) |
                  (b->m_def.fixedRotationY ? BX_BODY_FIXED_ROT_Y : 0u) |
                  (b->m_def.fixedRotationZ ? BX_BODY_FIXED_ROT_Z : 0u);
        Copy3(d.position, b->m_xf.translation);
        Copy4(d.orientation, b->m_xf.rotation);
        Copy3(d.linear_velocity, b->m_linearVelocity);
        Copy3(d.angular_velocity, b->m_angularVelocity);
        Copy3(d.linear_damping, b->m_def.linearDamping);
        Copy3(d.angular_damping, b->m_def.angularDamping);
        Copy3(d.gravity_scale, b->m_def.gravityScale);
        d.linear_sleep_tolerance = b->m_def.linearSleepTolerance;
        d.angular_sleep_tolerance = b->m_def.angularSleepTolerance;
    }
    std::vector<bx_fixture_desc> fd(m_fixtures.size());
    for (size_t i = 0; i < m_fixtures.size(); ++i) {
        const b3Fixture* f = m_fixtures[i];
        bx_fixture_desc& d = fd[i];
        std::memset(&d, 0, sizeof(d));
        d.body = f->m_body->m_index;
        d.shape_type = uint32(f->m_type);
        d.is_sensor = f->m_isSensor ? 1u : 0u;
        d.radius = f->m_shape->m_radius;
        d.density = f->m_density;
        d.friction = f->m_friction;
        d.restitution = f->m_restitution;
        d.mesh_scale[0] = d.mesh_scale[1] = d.mesh_scale[2] = 1.0f;
        if (f->m_type == b3Shape::e_sphere) {
            Copy3(d.vertex1, static_cast<const b3SphereShape*>(f->m_shape)->m_center);
        } else if (f->m_type == b3Shape::e_capsule) {
            Copy3(d.vertex1, static_cast<const b3CapsuleShape*>(f->m_shape)->m_vertex1);
            Copy3(d.vertex2, static_cast<const b3CapsuleShape*>(f->m_shape)->m_vertex2);
        } else if (f->m_type == b3Shape::e_triangle) {
            const b3TriangleShape* t = static_cast<const b3TriangleShape*>(f->m_shape);
            Copy3(d.vertex1, t->m_vertex1);
            Copy3(d.vertex2, t->m_vertex2);
            Copy3(d.vertex3, t->m_vertex3);
        } else if (f->m_type == b3Shape::e_mesh) {
            const b3MeshShape* m = static_cast<const b3MeshShape*>(f->m_shape);
            d.shape_index = meshIndex[m->m_mesh];
            Copy3(d.mesh_scale, m->m_scale);
        } else {
            d.shape_index = hullIndex[static_cast<const b3HullShape*>(f->m_shape)->m_hull];
        }
    }
    std::vector<bx_joint_desc> jd(m_joints.size());
    for (size_t i = 0; i < m_joints.size(); ++i) {
        const b3RevoluteJointDef& j = m_joints[i]->m_def;
        bx_joint_desc& d = jd[i];
        std::memset(&d, 0, sizeof(d));
        d.type = BX_JOINT_REVOLUTE;
        d.body_a = j.bodyA->m_index;
        d.body_b = j.bodyB->m_index;
        d.collide_connected = j.collideConnected ? 1u : 0u;
        d.use_initialize = j.m_initialize ? 1u : 0u;
        Copy3(d.axis, j.m_axis);
        Copy3(d.anchor, j.m_anchor);
        Copy3(d.local_anchor_a, j.localAnchorA);
        Copy4(d.local_rotation_a, j.localRotationA);
        Copy3(d.local_anchor_b, j.localAnchorB);
        Copy4(d.local_rotation_b, j.localRotationB);
        Copy4(d.reference_rotation, j.referenceRotation);
        d.enable_limit = j.enableLimit ? 1u : 0u;
        d.lower_angle = j.lowerAngle;
        d.upper_angle = j.upperAngle;
        d.enable_motor = j.enableMotor ? 1u : 0u;
        d.motor_speed = j.motorSpeed;
        d.max_motor_torque = j.maxMotorTorque;
    }
    bx_world_desc wd;
    std::memset(&wd, 0, sizeof(wd));
    Copy3(wd.gravity, m_gravity);
    wd.flags = (m_sleeping ? BX_WORLD_SLEEPING : 0u) | (m_warmStarting ? 0u : BX_WORLD_NO_WARM_START) |
               (m_autoClearForces ? 0u : BX_WORLD_NO_CLEAR_FORCES);
    wd.body_count = uint32(bd.size());
    wd.bodies = bd.data();
    wd.fixture_count = uint32(fd.size());
    wd.fixtures = fd.data();
    wd.joint_count = uint32(jd.size());
    wd.joints = jd.data();
    if (m_device->batch) {
        bx_batch_destroy(m_device->batch);
        m_device->batch = nullptr;
    }
    return bx_batch_create(m_device->ctx, &lib, &wd, 1, nullptr, &m_device->batch);
}

int b3World::PushState() {
    std::vector<bx_body_state> s(m_bodies.size());
    for (size_t i = 0; i < m_bodies.size(); ++i) {
        std::memset(&s[i], 0, sizeof(s[i]));
        Copy3(s[i].linear_velocity, m_bodies[i]->m_linearVelocity);
        Copy3(s[i].angular_velocity, m_bodies[i]->m_angularVelocity);
    }
    return bx_batch_set_bodies(m_device->batch, 0, s.data(), uint32(s.size()));
}

int b3World::Pull() {
    std::vector<bx_body_state> s(m_bodies.size());
    int rc = bx_batch_get_bodies(m_device->batch, 0, s.data(), uint32(s.size()));
    if (rc < 0) return rc;
    for (size_t i = 0; i < m_bodies.size(); ++i) {
        b3Body* b = m_bodies[i];
        const bx_body_state& x = s[i];
        b->m_xf.translation.Set(x.position[0], x.position[1], x.position[2]);
        b->m_xf.rotation = b3Quat(b3Vec3(x.orientation[0], x.orientation[1], x.orientation[2]), x.orientation[3]);
        b->m_worldCenter.Set(x.world_center[0], x.world_center[1], x.world_center[2]);
        b->m_linearVelocity.Set(x.linear_velocity[0], x.linear_velocity[1], x.linear_velocity[2]);
        b->m_angularVelocity.Set(x.angular_velocity[0], x.angular_velocity[1], x.angular_velocity[2]);
        b->m_awake = (x.flags & BX_BODY_AWAKE) != 0;
        b->m_sleepTime = x.sleep_time;
    }
    bx_world_counts c;
    rc = bx_batch_get_counts(m_device->batch, 0, &c);
    if (rc < 0) return rc;
    m_contactCount = c.contact_count;
    return BX_OK;
}

void b3World::Step(scalar dt, uint32 velocityIterations, uint32 positionIterations) {
    m_device->error.clear();
    int rc = BX_OK;
    if (!m_device->ctx) rc = bx_context_create(0, &m_device->ctx);
    if (rc == BX_OK && (m_topologyChanged || !m_device->batch)) {
        rc = Upload();
        if (rc == BX_OK) { m_topologyChanged = false; m_stateChanged = false; }
    }
    if (rc == BX_OK && m_stateChanged) {
        rc = PushState();
        if (rc >= 0) { rc = BX_OK; m_stateChanged = false; }
    }
    bx_step_cfg cfg;
    cfg.dt = dt;
    cfg.velocity_iterations = velocityIterations;
    cfg.position_iterations = positionIterations;
    if (rc == BX_OK) rc = bx_batch_step(m_device->batch, &cfg, 1);
    if (rc == BX_OK) rc = bx_batch_sync(m_device->batch);
    if (rc == BX_OK) rc = Pull();
    if (rc < 0) m_device->error = bx_last_error();
}
