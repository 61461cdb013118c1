// ref_harness.cpp -- test infrastructure, NOT product code.
//
// Drives the reference build of Bounce (compiled from the sources under the reference tree
// into oracle/_ref by oracle/Makefile) through its public C++ API, from the same
// bx_world_desc descriptors the device library takes, and dumps the state the parity tests
// compare: bodies, contacts with manifolds and feature ids, the broadphase pair buffer and the
// island solve order. It is compiled twice: FP32 (scalar = float) and FP64 (B3_USE_DOUBLE).
//
// Only tests/, the smoke check and bench.py's CPU baseline load the resulting library.
// Internal fields (m_flags, m_sleepTime, m_islandIndex, the broadphase pair buffer) are read by
// compiling this file with the class access keywords opened; the reference objects are built
// without that, and access specifiers do not change the layout of these classes.

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <new>
#include <vector>
#include <algorithm>
#include <unordered_map>
#include <thread>
#include <atomic>
#include <chrono>
#include <stdarg.h>
#include <float.h>
#include <assert.h>
#include <stddef.h>

#include "../include/bxstep.h"

#define private public
#define protected public
#include <bounce/dynamics/contacts/contact.h>
#include <bounce/dynamics/joints/joint.h>
#include <bounce/bounce.h>
#include <bounce/dynamics/contacts/hull_contact.h>
#include <bounce/collision/time_of_impact.h>
#undef private
#undef protected

#ifdef B3_USE_DOUBLE
#define BXREF(name) bxref64_##name
#else
#define BXREF(name) bxref32_##name
#endif

namespace {

struct RefHull {
    std::vector<b3Vec3> vertices;
    std::vector<b3HalfEdge> edges;
    std::vector<b3Face> faces;
    std::vector<b3Plane> planes;
    b3Hull hull;
};

struct RefMesh {
    std::vector<b3Vec3> vertices;
    std::vector<b3MeshTriangle> triangles;
    b3Mesh mesh;
};

// Records the island order (PostSolve) and every listener call of a step. EndContact is called from
// b3ContactManager::Destroy while the contact is still flagged touching, and from b3Contact::Update after
// the flag was cleared, which tells the two apart.
struct IslandRecorder : public b3ContactListener {
    struct Event { uint32_t type; const b3Fixture* a; const b3Fixture* b; b3Contact* contact; };
    std::vector<b3Contact*> order;
    std::vector<Event> events;
    void BeginContact(b3Contact* c) override { events.push_back(Event{1u, c->GetFixtureA(), c->GetFixtureB(), c}); }
    void EndContact(b3Contact* c) override {
        events.push_back(Event{2u, c->GetFixtureA(), c->GetFixtureB(), c->IsTouching() ? nullptr : c});
    }
    void PreSolve(b3Contact* c) override { events.push_back(Event{3u, c->GetFixtureA(), c->GetFixtureB(), c}); }
    void PostSolve(b3Contact* c) override {
        order.push_back(c);
        events.push_back(Event{4u, c->GetFixtureA(), c->GetFixtureB(), c});
    }
};

// A b3ContactFilter answering from two bit tables over the fixtures in creation order.
struct TableFilter : public b3ContactFilter {
    std::unordered_map<const b3Fixture*, uint32_t> index;
    std::vector<uint8_t> collide, respond;
    size_t n = 0;
    bool Bit(const std::vector<uint8_t>& t, const b3Fixture* a, const b3Fixture* b) const {
        const size_t k = size_t(index.at(a)) * n + index.at(b);
        return (t[k >> 3] >> (k & 7)) & 1u;
    }
    bool ShouldCollide(b3Fixture* a, b3Fixture* b) override { return Bit(collide, a, b); }
    bool ShouldRespond(b3Fixture* a, b3Fixture* b) override { return Bit(respond, a, b); }
};

struct RefWorld {
    b3World world;
    std::vector<RefHull*> hulls;
    std::vector<RefMesh*> meshes;
    std::vector<b3Body*> bodies;
    std::vector<b3Fixture*> fixtures;
    std::vector<b3Joint*> joints;
    IslandRecorder recorder;
    std::vector<b3Contact*> lastIslandOrder;
    TableFilter filter;
};

b3Vec3 V3(const float* p) { return b3Vec3(scalar(p[0]), scalar(p[1]), scalar(p[2])); }
b3Quat Q4(const float* p) { return b3Quat(b3Vec3(scalar(p[0]), scalar(p[1]), scalar(p[2])), scalar(p[3])); }

void BuildHull(RefHull* h, const bx_hull_desc& d) {
    h->vertices.resize(d.vertex_count);
    for (uint32 i = 0; i < d.vertex_count; ++i) h->vertices[i] = V3(d.vertices + 3 * i);
    h->edges.resize(d.edge_count);
    for (uint32 i = 0; i < d.edge_count; ++i) {
        const uint32_t* e = d.edges + 5 * i;
        h->edges[i] = b3MakeEdge(e[0], e[1], e[2], e[3], e[4]);
    }
    h->faces.resize(d.face_count);
    h->planes.resize(d.face_count);
    for (uint32 i = 0; i < d.face_count; ++i) {
        h->faces[i].edge = d.faces[i];
        h->planes[i].normal = V3(d.planes + 4 * i);
        h->planes[i].offset = scalar(d.planes[4 * i + 3]);
    }
    h->hull.centroid = V3(d.centroid);
    h->hull.vertexCount = d.vertex_count;
    h->hull.vertices = h->vertices.data();
    h->hull.edgeCount = d.edge_count;
    h->hull.edges = h->edges.data();
    h->hull.faceCount = d.face_count;
    h->hull.faces = h->faces.data();
    h->hull.planes = h->planes.data();
}

void BuildMesh(RefMesh* m, const bx_mesh_desc& d) {
    m->vertices.resize(d.vertex_count);
    for (uint32 i = 0; i < d.vertex_count; ++i) m->vertices[i] = V3(d.vertices + 3 * i);
    m->triangles.resize(d.triangle_count);
    for (uint32 i = 0; i < d.triangle_count; ++i) {
        m->triangles[i].v1 = d.triangles[3 * i + 0];
        m->triangles[i].v2 = d.triangles[3 * i + 1];
        m->triangles[i].v3 = d.triangles[3 * i + 2];
    }
    m->mesh.vertexCount = d.vertex_count;
    m->mesh.vertices = m->vertices.data();
    m->mesh.triangleCount = d.triangle_count;
    m->mesh.triangles = m->triangles.data();
    m->mesh.BuildTree();
    m->mesh.BuildAdjacency();
}

void InitRegistersOnce() {
    static std::atomic<int> done{0};
    if (done.load()) return;
    static std::atomic_flag lock = ATOMIC_FLAG_INIT;
    while (lock.test_and_set()) {}
    if (!b3Contact::s_initialized) {
        b3Contact::InitializeRegisters();
        b3Contact::s_initialized = true;
    }
    done.store(1);
    lock.clear();
}

void AddBody(RefWorld* w, const bx_body_desc& bd) {
    b3BodyDef def;
    def.type = b3BodyType(bd.type);
    def.awake = (bd.flags & BX_BODY_AWAKE) != 0;
    def.allowSleep = (bd.flags & BX_BODY_AUTO_SLEEP) != 0;
    def.fixedRotationX = (bd.flags & BX_BODY_FIXED_ROT_X) != 0;
    def.fixedRotationY = (bd.flags & BX_BODY_FIXED_ROT_Y) != 0;
    def.fixedRotationZ = (bd.flags & BX_BODY_FIXED_ROT_Z) != 0;
    def.userData = (void*)(uintptr_t)w->bodies.size();
    def.position = V3(bd.position);
    def.orientation = Q4(bd.orientation);
    def.linearVelocity = V3(bd.linear_velocity);
    def.angularVelocity = V3(bd.angular_velocity);
    def.linearDamping = V3(bd.linear_damping);
    def.angularDamping = V3(bd.angular_damping);
    def.gravityScale = V3(bd.gravity_scale);
    def.linearSleepTolerance = scalar(bd.linear_sleep_tolerance);
    def.angularSleepTolerance = scalar(bd.angular_sleep_tolerance);
    w->bodies.push_back(w->world.CreateBody(def));
}

bool AddFixture(RefWorld* w, const bx_fixture_desc& fd) {
    b3FixtureDef def;
    def.userData = (void*)(uintptr_t)w->fixtures.size();
    def.isSensor = fd.is_sensor != 0;
    def.density = scalar(fd.density);
    def.friction = scalar(fd.friction);
    def.restitution = scalar(fd.restitution);
    b3SphereShape sphere;
    b3CapsuleShape capsule;
    b3TriangleShape triangle;
    b3HullShape hull;
    b3MeshShape mesh;
    switch (fd.shape_type) {
    case BX_SHAPE_SPHERE:
        sphere.m_center = V3(fd.vertex1);
        sphere.m_radius = scalar(fd.radius);
        def.shape = &sphere;
        break;
    case BX_SHAPE_CAPSULE:
        capsule.m_vertex1 = V3(fd.vertex1);
        capsule.m_vertex2 = V3(fd.vertex2);
        capsule.m_radius = scalar(fd.radius);
        def.shape = &capsule;
        break;
    case BX_SHAPE_TRIANGLE:
        triangle.Set(V3(fd.vertex1), V3(fd.vertex2), V3(fd.vertex3));
        def.shape = &triangle;
        break;
    case BX_SHAPE_HULL:
        hull.m_hull = &w->hulls[fd.shape_index]->hull;
        def.shape = &hull;
        break;
    case BX_SHAPE_MESH:
        mesh.m_mesh = &w->meshes[fd.shape_index]->mesh;
        mesh.m_scale = V3(fd.mesh_scale);
        def.shape = &mesh;
        break;
    default:
        return false;
    }
    w->fixtures.push_back(w->bodies[fd.body]->CreateFixture(def));
    return true;
}

// The other joint types through their own defs and Initialize (bx_joint_desc, bxstep.h).
static bool AddOtherJoint(RefWorld* w, const bx_joint_desc& jd) {
    b3Body* a = w->bodies[jd.body_a];
    b3Body* b = w->bodies[jd.body_b];
    b3Joint* j = nullptr;
    const b3Vec3 anchor = V3(jd.anchor);
    auto common = [&](b3JointDef& def) {
        def.bodyA = a;
        def.bodyB = b;
        def.collideConnected = jd.collide_connected != 0;
    };
    switch (jd.type) {
    case BX_JOINT_SPHERE: {
        b3SphereJointDef def;
        if (jd.use_initialize) def.Initialize(a, b, anchor);
        else { def.localAnchorA = V3(jd.local_anchor_a); def.localAnchorB = V3(jd.local_anchor_b); }
        common(def);
        j = w->world.CreateJoint(def);
        break;
    }
    case BX_JOINT_WELD: {
        b3WeldJointDef def;
        if (jd.use_initialize) def.Initialize(a, b, anchor);
        else {
            def.localAnchorA = V3(jd.local_anchor_a);
            def.localAnchorB = V3(jd.local_anchor_b);
            def.referenceRotation = Q4(jd.reference_rotation);
        }
        def.frequencyHz = scalar(jd.frequency_hz);
        def.dampingRatio = scalar(jd.damping_ratio);
        def.correctionFactor = scalar(jd.correction_factor);
        common(def);
        j = w->world.CreateJoint(def);
        break;
    }
    case BX_JOINT_SPRING: {
        b3SpringJointDef def;
        if (jd.use_initialize) def.Initialize(a, b, anchor, V3(jd.anchor_b));
        else {
            def.localAnchorA = V3(jd.local_anchor_a);
            def.localAnchorB = V3(jd.local_anchor_b);
            def.length = scalar(jd.length);
        }
        def.frequencyHz = scalar(jd.frequency_hz);
        def.dampingRatio = scalar(jd.damping_ratio);
        common(def);
        j = w->world.CreateJoint(def);
        break;
    }
    case BX_JOINT_MOUSE: {
        b3MouseJointDef def;
        def.target = V3(jd.target);
        def.maxForce = scalar(jd.max_force);
        def.frequencyHz = scalar(jd.frequency_hz);
        def.dampingRatio = scalar(jd.damping_ratio);
        common(def);
        j = w->world.CreateJoint(def);
        break;
    }
    case BX_JOINT_FRICTION: {
        b3FrictionJointDef def;
        if (jd.use_initialize) def.Initialize(a, b, anchor);
        else { def.localAnchorA = V3(jd.local_anchor_a); def.localAnchorB = V3(jd.local_anchor_b); }
        def.maxForce = scalar(jd.max_force);
        def.maxTorque = scalar(jd.max_torque);
        common(def);
        j = w->world.CreateJoint(def);
        break;
    }
    case BX_JOINT_MOTOR: {
        b3MotorJointDef def;
        if (jd.use_initialize) def.Initialize(a, b);
        else { def.linearOffset = V3(jd.linear_offset); def.angularOffset = Q4(jd.angular_offset); }
        def.maxForce = scalar(jd.max_force);
        def.maxTorque = scalar(jd.max_torque);
        def.correctionFactor = scalar(jd.correction_factor);
        common(def);
        j = w->world.CreateJoint(def);
        break;
    }
    case BX_JOINT_CONE: {
        b3ConeJointDef def;
        if (jd.use_initialize) def.Initialize(a, b, V3(jd.axis), anchor, scalar(jd.cone_angle));
        else {
            def.localAnchorA = V3(jd.local_anchor_a);
            def.localRotationA = Q4(jd.local_rotation_a);
            def.localAnchorB = V3(jd.local_anchor_b);
            def.localRotationB = Q4(jd.local_rotation_b);
            def.referenceRotation = Q4(jd.reference_rotation);
            def.coneAngle = scalar(jd.cone_angle);
        }
        def.enableConeLimit = jd.enable_limit != 0;
        common(def);
        j = w->world.CreateJoint(def);
        break;
    }
    case BX_JOINT_PRISMATIC: {
        b3PrismaticJointDef def;
        if (jd.use_initialize) def.Initialize(a, b, anchor, V3(jd.axis));
        else {
            def.localAnchorA = V3(jd.local_anchor_a);
            def.localAnchorB = V3(jd.local_anchor_b);
            def.localAxisA = V3(jd.local_axis_a);
            def.referenceRotation = Q4(jd.reference_rotation);
        }
        def.enableLimit = jd.enable_limit != 0;
        def.lowerTranslation = scalar(jd.lower_translation);
        def.upperTranslation = scalar(jd.upper_translation);
        def.enableMotor = jd.enable_motor != 0;
        def.maxMotorForce = scalar(jd.max_force);
        def.motorSpeed = scalar(jd.motor_speed);
        def.correctionFactor = scalar(jd.correction_factor);
        common(def);
        j = w->world.CreateJoint(def);
        break;
    }
    case BX_JOINT_WHEEL: {
        b3WheelJointDef def;
        if (jd.use_initialize) def.Initialize(a, b, anchor, V3(jd.axis), V3(jd.axis_b));
        else {
            def.localAnchorA = V3(jd.local_anchor_a);
            def.localAnchorB = V3(jd.local_anchor_b);
            def.localAxisA = V3(jd.local_axis_a);
            def.localAxisB = V3(jd.local_axis_b);
        }
        def.enableMotor = jd.enable_motor != 0;
        def.maxMotorTorque = scalar(jd.max_motor_torque);
        def.motorSpeed = scalar(jd.motor_speed);
        def.frequencyHz = scalar(jd.frequency_hz);
        def.dampingRatio = scalar(jd.damping_ratio);
        common(def);
        j = w->world.CreateJoint(def);
        break;
    }
    default: return false;
    }
    w->joints.push_back(j);
    return true;
}

bool AddJoint(RefWorld* w, const bx_joint_desc& jd) {
    if (jd.type != BX_JOINT_REVOLUTE) return AddOtherJoint(w, jd);
    b3RevoluteJointDef def;
    b3Body* a = w->bodies[jd.body_a];
    b3Body* b = w->bodies[jd.body_b];
    if (jd.use_initialize) {
        def.Initialize(a, b, V3(jd.axis), V3(jd.anchor), scalar(jd.lower_angle), scalar(jd.upper_angle));
    } else {
        def.bodyA = a;
        def.bodyB = b;
        def.localAnchorA = V3(jd.local_anchor_a);
        def.localRotationA = Q4(jd.local_rotation_a);
        def.localAnchorB = V3(jd.local_anchor_b);
        def.localRotationB = Q4(jd.local_rotation_b);
        def.referenceRotation = Q4(jd.reference_rotation);
        def.lowerAngle = scalar(jd.lower_angle);
        def.upperAngle = scalar(jd.upper_angle);
    }
    def.collideConnected = jd.collide_connected != 0;
    def.enableLimit = jd.enable_limit != 0;
    def.enableMotor = jd.enable_motor != 0;
    def.motorSpeed = scalar(jd.motor_speed);
    def.maxMotorTorque = scalar(jd.max_motor_torque);
    w->joints.push_back(w->world.CreateJoint(def));
    return true;
}

void AddLibrary(RefWorld* w, const bx_shape_library* lib) {
    for (uint32 i = uint32(w->hulls.size()); lib && i < lib->hull_count; ++i) {
        RefHull* h = new RefHull();
        BuildHull(h, lib->hulls[i]);
        w->hulls.push_back(h);
    }
    for (uint32 i = uint32(w->meshes.size()); lib && i < lib->mesh_count; ++i) {
        RefMesh* m = new RefMesh();
        BuildMesh(m, lib->meshes[i]);
        w->meshes.push_back(m);
    }
}

RefWorld* CreateWorld(const bx_shape_library* lib, const bx_world_desc* wd) {
    InitRegistersOnce();
    RefWorld* w = new RefWorld();
    // b3World's constructor writes the process-global b3_convexCache = true (world.cpp:68).
    w->world.SetGravity(V3(wd->gravity));
    w->world.SetSleeping((wd->flags & BX_WORLD_SLEEPING) != 0);
    w->world.SetWarmStarting((wd->flags & BX_WORLD_NO_WARM_START) == 0);
    w->world.SetAutoClearForces((wd->flags & BX_WORLD_NO_CLEAR_FORCES) == 0);
    w->world.SetContactListener(&w->recorder);
    AddLibrary(w, lib);
    for (uint32 i = 0; i < wd->body_count; ++i) AddBody(w, wd->bodies[i]);
    for (uint32 i = 0; i < wd->fixture_count; ++i)
        if (!AddFixture(w, wd->fixtures[i])) { delete w; return nullptr; }
    for (uint32 i = 0; i < wd->joint_count; ++i)
        if (!AddJoint(w, wd->joints[i])) { delete w; return nullptr; }
    return w;
}

// Creation indices in the user data of every body and fixture, after a destroy moved them.
void Renumber(RefWorld* w) {
    for (size_t i = 0; i < w->bodies.size(); ++i) w->bodies[i]->m_userData = (void*)(uintptr_t)i;
    for (size_t i = 0; i < w->fixtures.size(); ++i) w->fixtures[i]->m_userData = (void*)(uintptr_t)i;
}

void Step(RefWorld* w, scalar dt, uint32 vit, uint32 pit) {
    w->recorder.order.clear();
    w->recorder.events.clear();
    w->world.Step(dt, vit, pit);
    w->lastIslandOrder = w->recorder.order;
}

uint32 FixtureIndex(const b3Fixture* f) { return uint32(uintptr_t(f->m_userData)); }

void FillContact(const b3Contact* c, bx_contact_state* out) {
    memset(out, 0, sizeof(*out));
    out->fixture_a = FixtureIndex(c->m_fixtureA);
    out->fixture_b = FixtureIndex(c->m_fixtureB);
    out->touching = c->IsTouching() ? 1u : 0u;
    out->manifold_count = c->m_manifoldCount;
    if (c->m_fixtureA->GetType() == b3Shape::e_hull && c->m_fixtureB->GetType() == b3Shape::e_hull) {
        const b3HullContact* hc = static_cast<const b3HullContact*>(c);
        out->cache[0] = uint32(hc->m_cache.cacheType);
        out->cache[1] = uint32(hc->m_cache.featureType);
        out->cache[2] = hc->m_cache.index1;
        out->cache[3] = hc->m_cache.index2;
    }
    for (uint32 i = 0; i < c->m_manifoldCount && i < BX_MAX_MANIFOLDS; ++i) {
        const b3Manifold& m = c->m_manifolds[i];
        bx_manifold_state& ms = out->manifolds[i];
        ms.point_count = m.pointCount;
        ms.tangent_impulse[0] = float(m.tangentImpulse.x);
        ms.tangent_impulse[1] = float(m.tangentImpulse.y);
        ms.motor_impulse = float(m.motorImpulse);
        for (uint32 k = 0; k < m.pointCount; ++k) {
            const b3ManifoldPoint& p = m.points[k];
            bx_manifold_point_state& ps = ms.points[k];
            for (int a = 0; a < 3; ++a) {
                ps.local_normal1[a] = float(p.localNormal1[a]);
                ps.local_point1[a] = float(p.localPoint1[a]);
                ps.local_point2[a] = float(p.localPoint2[a]);
            }
            ps.normal_impulse = float(p.normalImpulse);
            ps.triangle_key = p.id.triangleKey;
            ps.key[0] = p.id.key.pair.inEdge1;
            ps.key[1] = p.id.key.pair.outEdge1;
            ps.key[2] = p.id.key.pair.inEdge2;
            ps.key[3] = p.id.key.pair.outEdge2;
            ps.persisting = p.persisting ? 1u : 0u;
            ps.edge_contact = p.edgeContact ? 1u : 0u;
        }
    }
}

}  // namespace

extern "C" {

void* BXREF(create)(const bx_shape_library* lib, const bx_world_desc* wd) {
    return CreateWorld(lib, wd);
}

void BXREF(destroy)(void* h) {
    RefWorld* w = (RefWorld*)h;
    if (!w) return;
    std::vector<RefHull*> hulls = w->hulls;
    std::vector<RefMesh*> meshes = w->meshes;
    delete w;
    for (RefHull* p : hulls) delete p;
    for (RefMesh* p : meshes) delete p;
}

int BXREF(set_contact_filter)(void* h, const uint8_t* collide, const uint8_t* respond) {
    RefWorld* w = (RefWorld*)h;
    if (!collide) {
        w->world.SetContactFilter(nullptr);
        return 0;
    }
    TableFilter& f = w->filter;
    f.n = w->fixtures.size();
    f.index.clear();
    for (size_t i = 0; i < w->fixtures.size(); ++i) f.index[w->fixtures[i]] = uint32_t(i);
    const size_t bytes = (f.n * f.n + 7) / 8;
    f.collide.assign(collide, collide + bytes);
    if (respond) f.respond.assign(respond, respond + bytes);
    else f.respond.assign(bytes, 0xFF);
    w->world.SetContactFilter(&f);
    return 0;
}

int BXREF(get_events)(void* h, bx_contact_event* out, uint32_t cap, uint32_t* count) {
    RefWorld* w = (RefWorld*)h;
    std::unordered_map<const b3Fixture*, uint32_t> fi;
    for (size_t i = 0; i < w->fixtures.size(); ++i) fi[w->fixtures[i]] = uint32_t(i);
    std::unordered_map<const b3Contact*, uint32_t> ci;
    uint32_t k = 0;
    for (b3Contact* c = w->world.m_contactManager.m_contactList; c; c = c->m_next) ci[c] = k++;
    uint32_t n = 0;
    for (const IslandRecorder::Event& e : w->recorder.events) {
        uint32_t idx = BX_NULL_INDEX;
        if (e.contact && ci.count(e.contact)) idx = ci[e.contact];
        if (n < cap) out[n] = bx_contact_event{e.type, fi[e.a], fi[e.b], idx};
        ++n;
    }
    if (count) *count = n;
    return 0;
}

void BXREF(step)(void* h, float dt, uint32_t vit, uint32_t pit, uint32_t n) {
    RefWorld* w = (RefWorld*)h;
    for (uint32_t i = 0; i < n; ++i) Step(w, scalar(dt), vit, pit);
}

// Body edits between steps through the reference's own b3Body methods (bx_body_edit, bxstep.h).
int BXREF(edit_bodies)(void* h, const bx_body_edit* edits, uint32_t count) {
    RefWorld* w = (RefWorld*)h;
    for (uint32_t k = 0; k < count; ++k) {
        const bx_body_edit& e = edits[k];
        const b3Vec3 v(scalar(e.v[0]), scalar(e.v[1]), scalar(e.v[2]));
        const b3Vec3 p(scalar(e.p[0]), scalar(e.p[1]), scalar(e.p[2]));
        switch (e.op) {
        case BX_EDIT_WORLD_GRAVITY: w->world.SetGravity(v); continue;
        case BX_EDIT_WORLD_SLEEPING: w->world.SetSleeping(e.wake != 0); continue;
        case BX_EDIT_WORLD_WARM_STARTING: w->world.SetWarmStarting(e.wake != 0); continue;
        case BX_EDIT_WORLD_AUTO_CLEAR_FORCES: w->world.SetAutoClearForces(e.wake != 0); continue;
        case BX_EDIT_WORLD_CLEAR_FORCES: w->world.ClearForces(); continue;
        default: break;
        }
        if (e.body >= w->bodies.size()) return -1;
        b3Body* b = w->bodies[e.body];
        switch (e.op) {
        case BX_EDIT_LINEAR_VELOCITY: b->SetLinearVelocity(v); break;
        case BX_EDIT_ANGULAR_VELOCITY: b->SetAngularVelocity(v); break;
        case BX_EDIT_FORCE: b->ApplyForce(v, p, e.wake != 0); break;
        case BX_EDIT_FORCE_TO_CENTER: b->ApplyForceToCenter(v, e.wake != 0); break;
        case BX_EDIT_TORQUE: b->ApplyTorque(v, e.wake != 0); break;
        case BX_EDIT_LINEAR_IMPULSE: b->ApplyLinearImpulse(v, p, e.wake != 0); break;
        case BX_EDIT_ANGULAR_IMPULSE: b->ApplyAngularImpulse(v, e.wake != 0); break;
        case BX_EDIT_TRANSFORM: {
            b3Transform xf;
            xf.translation = v;
            xf.rotation = b3Quat(p, scalar(e.p[3]));
            b->SetTransform(xf);
            break;
        }
        case BX_EDIT_AWAKE: b->SetAwake(e.wake != 0); break;
        case BX_EDIT_SLEEPING_ALLOWED: b->SetSleepingAllowed(e.wake != 0); break;
        case BX_EDIT_GRAVITY_SCALE: b->SetGravityScale(v); break;
        case BX_EDIT_LINEAR_DAMPING: b->SetLinearDamping(v); break;
        case BX_EDIT_ANGULAR_DAMPING: b->SetAngularDamping(v); break;
        case BX_EDIT_LINEAR_SLEEP_TOLERANCE: b->SetLinearSleepTolerance(scalar(e.v[0])); break;
        case BX_EDIT_ANGULAR_SLEEP_TOLERANCE: b->SetAngularSleepTolerance(scalar(e.v[0])); break;
        case BX_EDIT_FIXED_ROTATION: b->SetFixedRotation(e.v[0] != 0.0f, e.v[1] != 0.0f, e.v[2] != 0.0f); break;
        default: return -1;
        }
    }
    return 0;
}

int BXREF(get_counts)(void* h, bx_world_counts* out) {
    RefWorld* w = (RefWorld*)h;
    memset(out, 0, sizeof(*out));
    out->body_count = uint32_t(w->bodies.size());
    out->fixture_count = uint32_t(w->fixtures.size());
    out->joint_count = uint32_t(w->joints.size());
    out->contact_count = w->world.m_contactManager.m_contactCount;
    for (b3Contact* c = w->world.m_contactManager.m_contactList; c; c = c->m_next)
        out->touching_count += c->IsTouching() ? 1u : 0u;
    const b3BroadPhase& bp = w->world.m_contactManager.m_broadPhase;
    uint32_t unique = 0;
    for (uint32_t i = 0; i < bp.m_pairCount; ++i) {
        if (i == 0 || bp.m_pairs[i].proxy1 != bp.m_pairs[i - 1].proxy1 ||
            bp.m_pairs[i].proxy2 != bp.m_pairs[i - 1].proxy2)
            ++unique;
    }
    out->pair_count = unique;
    for (b3Body* b : w->bodies) out->awake_count += b->IsAwake() ? 1u : 0u;
    return 0;
}

int BXREF(get_bodies)(void* h, bx_body_state* out, uint32_t cap) {
    RefWorld* w = (RefWorld*)h;
    uint32_t n = std::min<uint32_t>(cap, uint32_t(w->bodies.size()));
    for (uint32_t i = 0; i < n; ++i) {
        const b3Body* b = w->bodies[i];
        bx_body_state& s = out[i];
        memset(&s, 0, sizeof(s));
        for (int a = 0; a < 3; ++a) {
            s.position[a] = float(b->m_xf.translation[a]);
            s.world_center[a] = float(b->m_sweep.worldCenter[a]);
            s.linear_velocity[a] = float(b->m_linearVelocity[a]);
            s.angular_velocity[a] = float(b->m_angularVelocity[a]);
        }
        s.orientation[0] = float(b->m_xf.rotation.v.x);
        s.orientation[1] = float(b->m_xf.rotation.v.y);
        s.orientation[2] = float(b->m_xf.rotation.v.z);
        s.orientation[3] = float(b->m_xf.rotation.s);
        s.sleep_time = float(b->m_sleepTime);
        s.flags = b->m_flags;
        s.island_index = b->m_islandIndex;
    }
    return int(n);
}

// 16 doubles per body: position 3, orientation 4, world centre 3, linear velocity 3, angular 3.
int BXREF(get_bodies_f64)(void* h, double* out, uint32_t cap) {
    RefWorld* w = (RefWorld*)h;
    uint32_t n = std::min<uint32_t>(cap, uint32_t(w->bodies.size()));
    for (uint32_t i = 0; i < n; ++i) {
        const b3Body* b = w->bodies[i];
        double* o = out + 16 * i;
        for (int a = 0; a < 3; ++a) o[a] = double(b->m_xf.translation[a]);
        o[3] = double(b->m_xf.rotation.v.x);
        o[4] = double(b->m_xf.rotation.v.y);
        o[5] = double(b->m_xf.rotation.v.z);
        o[6] = double(b->m_xf.rotation.s);
        for (int a = 0; a < 3; ++a) o[7 + a] = double(b->m_sweep.worldCenter[a]);
        for (int a = 0; a < 3; ++a) o[10 + a] = double(b->m_linearVelocity[a]);
        for (int a = 0; a < 3; ++a) o[13 + a] = double(b->m_angularVelocity[a]);
    }
    return int(n);
}

int BXREF(get_contacts)(void* h, bx_contact_state* out, uint32_t cap) {
    RefWorld* w = (RefWorld*)h;
    uint32_t n = 0;
    for (b3Contact* c = w->world.m_contactManager.m_contactList; c && n < cap; c = c->m_next)
        FillContact(c, out + n++);
    return int(n);
}

int BXREF(get_pairs)(void* h, bx_pair* out, uint32_t cap) {
    RefWorld* w = (RefWorld*)h;
    const b3BroadPhase& bp = w->world.m_contactManager.m_broadPhase;
    uint32_t n = 0;
    for (uint32_t i = 0; i < bp.m_pairCount; ++i) {
        if (i > 0 && bp.m_pairs[i].proxy1 == bp.m_pairs[i - 1].proxy1 &&
            bp.m_pairs[i].proxy2 == bp.m_pairs[i - 1].proxy2)
            continue;
        if (n < cap) {
            out[n].proxy1 = bp.m_pairs[i].proxy1;
            out[n].proxy2 = bp.m_pairs[i].proxy2;
        }
        ++n;
    }
    return int(n);
}

// Proxy id of each fixture (creation order).
int BXREF(get_proxy_ids)(void* h, uint32_t* out, uint32_t cap) {
    RefWorld* w = (RefWorld*)h;
    uint32_t n = std::min<uint32_t>(cap, uint32_t(w->fixtures.size()));
    for (uint32_t i = 0; i < n; ++i) out[i] = w->fixtures[i]->m_proxyId;
    return int(n);
}

// Fat AABBs of each fixture's proxy (creation order), 6 floats each.
int BXREF(get_fat_aabbs)(void* h, float* out, uint32_t cap) {
    RefWorld* w = (RefWorld*)h;
    uint32_t n = std::min<uint32_t>(cap, uint32_t(w->fixtures.size()));
    const b3BroadPhase& bp = w->world.m_contactManager.m_broadPhase;
    for (uint32_t i = 0; i < n; ++i) {
        const b3AABB& a = bp.GetFatAABB(w->fixtures[i]->m_proxyId);
        for (int k = 0; k < 3; ++k) {
            out[6 * i + k] = float(a.lowerBound[k]);
            out[6 * i + 3 + k] = float(a.upperBound[k]);
        }
    }
    return int(n);
}

// Island solve order of the last step: indices into the contact list of get_contacts.
int BXREF(get_island_contacts)(void* h, uint32_t* out, uint32_t cap, uint32_t* count) {
    RefWorld* w = (RefWorld*)h;
    std::unordered_map<const b3Contact*, uint32_t> index;
    uint32_t k = 0;
    for (b3Contact* c = w->world.m_contactManager.m_contactList; c; c = c->m_next) index[c] = k++;
    uint32_t n = 0;
    for (const b3Contact* c : w->lastIslandOrder) {
        auto it = index.find(c);
        if (n < cap) out[n] = it == index.end() ? BX_NULL_INDEX : it->second;
        ++n;
    }
    *count = n;
    return 0;
}

// Mass properties after creation (creation order): mass, invMass, localCenter 3, I 9, invI 9.
int BXREF(get_mass)(void* h, float* out, uint32_t cap) {
    RefWorld* w = (RefWorld*)h;
    uint32_t n = std::min<uint32_t>(cap, uint32_t(w->bodies.size()));
    for (uint32_t i = 0; i < n; ++i) {
        const b3Body* b = w->bodies[i];
        float* o = out + 23 * i;
        o[0] = float(b->m_mass);
        o[1] = float(b->m_invMass);
        for (int a = 0; a < 3; ++a) o[2 + a] = float(b->m_sweep.localCenter[a]);
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) {
                o[5 + 3 * c + r] = float(b->m_I[c][r]);
                o[14 + 3 * c + r] = float(b->m_invI[c][r]);
            }
    }
    return int(n);
}

// Topology changes through the reference's own b3World / b3Body calls, in order (the operations of
// bx_batch_edit_topology).
int BXREF(edit_topology)(void* h, const bx_shape_library* lib, const bx_world_desc* adds, const bx_topology_op* ops,
                         uint32_t count) {
    RefWorld* w = static_cast<RefWorld*>(h);
    AddLibrary(w, lib);
    for (uint32_t k = 0; k < count; ++k) {
        const bx_topology_op& op = ops[k];
        switch (op.op) {
        case BX_TOPO_CREATE_BODY: AddBody(w, adds->bodies[op.index]); break;
        case BX_TOPO_CREATE_FIXTURE:
            if (!AddFixture(w, adds->fixtures[op.index])) return -1;
            break;
        case BX_TOPO_CREATE_JOINT:
            if (!AddJoint(w, adds->joints[op.index])) return -1;
            break;
        case BX_TOPO_DESTROY_JOINT:
            w->world.DestroyJoint(w->joints[op.index]);
            w->joints.erase(w->joints.begin() + op.index);
            break;
        case BX_TOPO_DESTROY_FIXTURE: {
            b3Fixture* f = w->fixtures[op.index];
            f->GetBody()->DestroyFixture(f);
            w->fixtures.erase(w->fixtures.begin() + op.index);
            break;
        }
        case BX_TOPO_DESTROY_BODY: {
            b3Body* b = w->bodies[op.index];
            for (size_t j = w->joints.size(); j-- > 0;)
                if (w->joints[j]->GetBodyA() == b || w->joints[j]->GetBodyB() == b) w->joints.erase(w->joints.begin() + j);
            for (size_t f = w->fixtures.size(); f-- > 0;)
                if (w->fixtures[f]->GetBody() == b) w->fixtures.erase(w->fixtures.begin() + f);
            w->world.DestroyBody(b);
            w->bodies.erase(w->bodies.begin() + op.index);
            break;
        }
        default: return -1;
        }
        Renumber(w);
    }
    w->lastIslandOrder.clear();
    return 0;
}

// World queries through the reference's own b3World::RayCast / RayCastSingle / QueryAABB. The ray
// listener never clips (returns 1), so every hit is reported; hits and query results are sorted
// here (fraction, then fixture creation index) since the dynamic tree's visit order is not part of
// what is compared. Filters answer from an optional byte per fixture.
struct MaskRayFilter : public b3RayCastFilter {
    const uint8_t* mask = nullptr;
    bool ShouldRayCast(b3Fixture* f) override { return !mask || mask[FixtureIndex(f)] != 0; }
};
struct CollectRays : public b3RayCastListener {
    std::vector<bx_ray_hit> hits;
    scalar ReportFixture(b3Fixture* f, const b3Vec3& point, const b3Vec3& normal, scalar fraction) override {
        bx_ray_hit h;
        h.fixture = FixtureIndex(f);
        h.fraction = float(fraction);
        for (int a = 0; a < 3; ++a) {
            h.point[a] = float(point[a]);
            h.normal[a] = float(normal[a]);
        }
        hits.push_back(h);
        return scalar(1);
    }
};
struct MaskQueryFilter : public b3QueryFilter {
    const uint8_t* mask = nullptr;
    bool ShouldReport(b3Fixture* f) override { return !mask || mask[FixtureIndex(f)] != 0; }
};
struct CollectFixtures : public b3QueryListener {
    std::vector<uint32_t> fixtures;
    bool ReportFixture(b3Fixture* f) override {
        fixtures.push_back(FixtureIndex(f));
        return true;
    }
};

int BXREF(raycast)(void* h, const float* p1, const float* p2, const uint8_t* mask, bx_ray_hit* out, uint32_t cap,
                   uint32_t* count) {
    RefWorld* w = static_cast<RefWorld*>(h);
    MaskRayFilter filter;
    filter.mask = mask;
    CollectRays rays;
    w->world.RayCast(&rays, &filter, V3(p1), V3(p2));
    std::stable_sort(rays.hits.begin(), rays.hits.end(), [](const bx_ray_hit& a, const bx_ray_hit& b) {
        return a.fraction < b.fraction || (a.fraction == b.fraction && a.fixture < b.fixture);
    });
    *count = uint32_t(rays.hits.size());
    for (uint32_t i = 0; i < cap && i < rays.hits.size(); ++i) out[i] = rays.hits[i];
    return 0;
}

int BXREF(raycast_single)(void* h, const float* p1, const float* p2, const uint8_t* mask, bx_ray_hit* out) {
    RefWorld* w = static_cast<RefWorld*>(h);
    MaskRayFilter filter;
    filter.mask = mask;
    b3RayCastSingleOutput o;
    std::memset(out, 0, sizeof(*out));
    out->fixture = BX_NULL_INDEX;
    if (!w->world.RayCastSingle(&o, &filter, V3(p1), V3(p2))) return 0;
    out->fixture = FixtureIndex(o.fixture);
    out->fraction = float(o.fraction);
    for (int a = 0; a < 3; ++a) {
        out->point[a] = float(o.point[a]);
        out->normal[a] = float(o.normal[a]);
    }
    return 1;
}

// Shape casts through the reference's own b3World::ShapeCast / ShapeCastSingle, and time-of-impact
// problems through its b3TimeOfImpact. A cast shape is built from a bx_cast_shape; a hull comes from the
// world's library (shape casts) or from the given library (time of impact).
struct CastShapeHolder {
    b3SphereShape sphere;
    b3CapsuleShape capsule;
    b3TriangleShape triangle;
    b3HullShape hull;
    RefHull ownHull;
    const b3Shape* shape = nullptr;
};

static bool MakeCastShape(const bx_cast_shape& c, RefWorld* w, const bx_shape_library* lib, CastShapeHolder& out) {
    switch (c.shape_type) {
    case BX_SHAPE_SPHERE:
        out.sphere.m_center = V3(c.vertex1);
        out.sphere.m_radius = scalar(c.radius);
        out.shape = &out.sphere;
        return true;
    case BX_SHAPE_CAPSULE:
        out.capsule.m_vertex1 = V3(c.vertex1);
        out.capsule.m_vertex2 = V3(c.vertex2);
        out.capsule.m_radius = scalar(c.radius);
        out.shape = &out.capsule;
        return true;
    case BX_SHAPE_TRIANGLE:
        out.triangle.Set(V3(c.vertex1), V3(c.vertex2), V3(c.vertex3));
        out.triangle.m_radius = scalar(c.radius);
        out.shape = &out.triangle;
        return true;
    case BX_SHAPE_HULL:
        if (w) {
            if (c.shape_index >= w->hulls.size()) return false;
            out.hull.m_hull = &w->hulls[c.shape_index]->hull;
        } else {
            if (!lib || c.shape_index >= lib->hull_count) return false;
            BuildHull(&out.ownHull, lib->hulls[c.shape_index]);
            out.hull.m_hull = &out.ownHull.hull;
        }
        out.hull.m_radius = scalar(c.radius);
        out.shape = &out.hull;
        return true;
    }
    return false;
}

static b3Transform XF(const float* p, const float* q) {
    b3Transform xf;
    xf.translation = V3(p);
    xf.rotation = b3Quat(b3Vec3(scalar(q[0]), scalar(q[1]), scalar(q[2])), scalar(q[3]));
    return xf;
}

static b3Sweep SW(const bx_sweep& s) {
    b3Sweep w;
    w.localCenter = V3(s.local_center);
    w.worldCenter0 = V3(s.center0);
    w.worldCenter = V3(s.center);
    w.orientation0 = b3Quat(b3Vec3(scalar(s.orientation0[0]), scalar(s.orientation0[1]), scalar(s.orientation0[2])),
                            scalar(s.orientation0[3]));
    w.orientation = b3Quat(b3Vec3(scalar(s.orientation[0]), scalar(s.orientation[1]), scalar(s.orientation[2])),
                           scalar(s.orientation[3]));
    w.t0 = scalar(s.t0);
    return w;
}

int BXREF(time_of_impact)(const bx_shape_library* lib, const bx_toi_query* q, uint32_t count, bx_toi_result* out) {
    for (uint32_t i = 0; i < count; ++i) {
        const bx_toi_query& x = q[i];
        b3TOIOutput o;
        if (x.kind == BX_TOI_AABB) {
            b3AABB a, b;
            a.lowerBound = V3(x.lower1);
            a.upperBound = V3(x.upper1);
            b.lowerBound = V3(x.lower2);
            b.upperBound = V3(x.upper2);
            o = b3TimeOfImpact(a, V3(x.displacement1), b, V3(x.displacement2));
        } else {
            CastShapeHolder s1, s2;
            if (!MakeCastShape(x.shape1, nullptr, lib, s1) || !MakeCastShape(x.shape2, nullptr, lib, s2)) return -1;
            b3GJKProxy p1(s1.shape, 0), p2(s2.shape, 0);
            if (x.kind == BX_TOI_SWEEP) {
                b3TOIInput in;
                in.proxy1 = p1;
                in.proxy2 = p2;
                in.sweep1 = SW(x.sweep1);
                in.sweep2 = SW(x.sweep2);
                in.tMax = scalar(x.t_max);
                o = b3TimeOfImpact(in);
            } else if (x.kind == BX_TOI_TRANSLATION) {
                o = b3TimeOfImpact(XF(x.position1, x.orientation1), p1, V3(x.displacement1),
                                   XF(x.position2, x.orientation2), p2, V3(x.displacement2), x.max_iterations);
            } else {
                return -1;
            }
        }
        out[i].state = uint32_t(o.state);
        out[i].t = float(o.t);
        out[i].iterations = o.iterations;
    }
    return 0;
}

struct MaskShapeFilter : public b3ShapeCastFilter {
    const uint8_t* mask = nullptr;
    bool ShouldShapeCast(b3Fixture* f) override { return !mask || mask[FixtureIndex(f)] != 0; }
};
struct CollectShapeHits : public b3ShapeCastListener {
    std::vector<bx_shape_hit> hits;
    scalar ReportFixture(b3Fixture* f, const b3Vec3& point, const b3Vec3& normal, scalar fraction) override {
        bx_shape_hit h;
        h.fixture = FixtureIndex(f);
        h.child = 0;  // the listener is not told the mesh triangle
        h.fraction = float(fraction);
        for (int a = 0; a < 3; ++a) {
            h.point[a] = float(point[a]);
            h.normal[a] = float(normal[a]);
        }
        hits.push_back(h);
        return scalar(1);
    }
};

// Hits sorted by (fraction, fixture, point, normal), as the tree's visit order is not compared.
int BXREF(shape_cast)(void* h, const bx_shape_cast_query* q, const uint8_t* mask, bx_shape_hit* out, uint32_t cap,
                      uint32_t* count) {
    RefWorld* w = static_cast<RefWorld*>(h);
    CastShapeHolder s;
    if (!MakeCastShape(q->shape, w, nullptr, s)) return -1;
    MaskShapeFilter filter;
    filter.mask = mask;
    CollectShapeHits hits;
    w->world.ShapeCast(&hits, &filter, s.shape, XF(q->position, q->orientation), V3(q->displacement));
    std::stable_sort(hits.hits.begin(), hits.hits.end(), [](const bx_shape_hit& a, const bx_shape_hit& b) {
        if (a.fraction != b.fraction) return a.fraction < b.fraction;
        if (a.fixture != b.fixture) return a.fixture < b.fixture;
        return std::memcmp(a.point, b.point, sizeof(float) * 6) < 0;
    });
    *count = uint32_t(hits.hits.size());
    for (uint32_t i = 0; i < cap && i < hits.hits.size(); ++i) out[i] = hits.hits[i];
    return 0;
}

int BXREF(shape_cast_single)(void* h, const bx_shape_cast_query* q, const uint8_t* mask, bx_shape_hit* out) {
    RefWorld* w = static_cast<RefWorld*>(h);
    CastShapeHolder s;
    std::memset(out, 0, sizeof(*out));
    out->fixture = BX_NULL_INDEX;
    if (!MakeCastShape(q->shape, w, nullptr, s)) return -1;
    MaskShapeFilter filter;
    filter.mask = mask;
    b3ShapeCastSingleOutput o;
    if (!w->world.ShapeCastSingle(&o, &filter, s.shape, XF(q->position, q->orientation), V3(q->displacement)))
        return 0;
    out->fixture = FixtureIndex(o.fixture);
    out->fraction = float(o.fraction);
    for (int a = 0; a < 3; ++a) {
        out->point[a] = float(o.point[a]);
        out->normal[a] = float(o.normal[a]);
    }
    return 0;
}

int BXREF(query_aabb)(void* h, const float* lower, const float* upper, const uint8_t* mask, uint32_t* out,
                      uint32_t cap, uint32_t* count) {
    RefWorld* w = static_cast<RefWorld*>(h);
    MaskQueryFilter filter;
    filter.mask = mask;
    CollectFixtures q;
    b3AABB box;
    box.lowerBound = V3(lower);
    box.upperBound = V3(upper);
    w->world.QueryAABB(&q, &filter, box);
    std::sort(q.fixtures.begin(), q.fixtures.end());
    *count = uint32_t(q.fixtures.size());
    for (uint32_t i = 0; i < cap && i < q.fixtures.size(); ++i) out[i] = q.fixtures[i];
    return 0;
}

// b3BoxHull(ex, ey, ez) data in bx_hull_desc layout: 8 vertices, 24 edges, 6 faces.
void BXREF(box_hull)(float ex, float ey, float ez, float* vertices, uint32_t* edges,
                     uint32_t* faces, float* planes, float* centroid) {
    b3BoxHull box{scalar(ex), scalar(ey), scalar(ez)};
    for (uint32 i = 0; i < 8; ++i)
        for (int a = 0; a < 3; ++a) vertices[3 * i + a] = float(box.vertices[i][a]);
    for (uint32 i = 0; i < 24; ++i) {
        const b3HalfEdge& e = box.edges[i];
        edges[5 * i + 0] = e.origin;
        edges[5 * i + 1] = e.twin;
        edges[5 * i + 2] = e.face;
        edges[5 * i + 3] = e.prev;
        edges[5 * i + 4] = e.next;
    }
    for (uint32 i = 0; i < 6; ++i) {
        faces[i] = box.faces[i].edge;
        for (int a = 0; a < 3; ++a) planes[4 * i + a] = float(box.planes[i].normal[a]);
        planes[4 * i + 3] = float(box.planes[i].offset);
    }
    for (int a = 0; a < 3; ++a) centroid[a] = float(box.centroid[a]);
}

// CPU baseline: step `world_count` worlds, one world at a time per thread, `threads` threads.
// Returns wall seconds of the timed steps (after `warmup` untimed steps per world).
double BXREF(bench)(const bx_shape_library* lib, const bx_world_desc* worlds, uint32_t world_count,
                    float dt, uint32_t vit, uint32_t pit, uint32_t warmup, uint32_t steps,
                    uint32_t threads) {
    InitRegistersOnce();
    if (threads == 0) threads = 1;
    std::atomic<uint32_t> next{0};
    std::vector<double> busy(threads, 0.0);
    auto worker = [&](uint32_t t) {
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= world_count) break;
            RefWorld* w = CreateWorld(lib, worlds + i);
            w->world.SetContactListener(nullptr);
            for (uint32_t s = 0; s < warmup; ++s) w->world.Step(scalar(dt), vit, pit);
            auto t0 = std::chrono::steady_clock::now();
            for (uint32_t s = 0; s < steps; ++s) w->world.Step(scalar(dt), vit, pit);
            auto t1 = std::chrono::steady_clock::now();
            busy[t] += std::chrono::duration<double>(t1 - t0).count();
            BXREF(destroy)(w);
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t) pool.emplace_back(worker, t);
    for (auto& th : pool) th.join();
    // Aggregate throughput over threads: timed work runs concurrently, so the wall time of
    // the timed regions is the busy time of the most loaded thread.
    double wall = 0.0;
    for (double b : busy) wall = std::max(wall, b);
    return wall;
}

}  // extern "C"
