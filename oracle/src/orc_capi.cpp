// orc_capi.cpp -- oracle (test infrastructure): C ABI of the CPU restatement, with the same
// signatures as the reference harness (oracle/ref_harness.cpp), so tests drive both alike.
// World construction restates b3World::CreateBody (world.cpp:94-110, body.cpp:25-103),
// b3Body::CreateFixture (body.cpp:105-136, fixture.cpp:43-56) and b3World::CreateJoint.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include "orc_world.h"
#include "../../include/bxstep.h"

namespace orc {
void CreateJoint(World& w, const bx_joint_desc& jd);
bool BuildMesh(Mesh& m, const bx_mesh_desc& d);
void InitTriangleShape(Shape& s, const Vec3& a, const Vec3& b, const Vec3& c);
struct RayHit { u32 fixture; scalar fraction; Vec3 point, normal; };
void RayCastAll(const World& w, const Vec3& p1, const Vec3& p2, const uint8_t* mask, std::vector<RayHit>& hits);
void QueryAABB(const World& w, const AABB& aabb, const uint8_t* mask, std::vector<u32>& out);
struct ShapeHit { u32 fixture, child; scalar fraction; Vec3 point, normal; };
void ShapeCastAll(const World& w, const Shape& shape, const Transform& xf, const Vec3& d, const uint8_t* mask,
                  std::vector<ShapeHit>& hits);
}

using namespace orc;

static Vec3 V3(const float* p) { return Vec3(p[0], p[1], p[2]); }
static Quat Q4(const float* p) { return Quat(Vec3(p[0], p[1], p[2]), p[3]); }

static bool BuildHull(Hull& h, const bx_hull_desc& d) {
    if (d.vertex_count < 4 || d.edge_count < 6 || d.face_count < 4) return false;
    h.vertices.resize(d.vertex_count);
    for (u32 i = 0; i < d.vertex_count; ++i) h.vertices[i] = V3(d.vertices + 3 * i);
    h.edges.resize(d.edge_count);
    for (u32 i = 0; i < d.edge_count; ++i) {
        const uint32_t* e = d.edges + 5 * i;
        h.edges[i] = HalfEdge{e[0], e[1], e[2], e[3], e[4]};
    }
    h.faces.assign(d.faces, d.faces + d.face_count);
    h.planes.resize(d.face_count);
    for (u32 i = 0; i < d.face_count; ++i) h.planes[i] = Plane(V3(d.planes + 4 * i), d.planes[4 * i + 3]);
    h.centroid = V3(d.centroid);
    return true;
}

static void AddLibrary(World* w, const bx_shape_library* lib, bool* ok) {
    // hulls and meshes are held by value and fixtures point at them: grow only before any fixture
    // refers to the vectors' storage again (reserve keeps the pointers of existing entries)
    if (!lib) return;
    if (lib->hull_count > w->hulls.size()) {
        std::vector<Hull> grown(lib->hull_count);
        for (u32 i = 0; i < w->hulls.size(); ++i) grown[i] = w->hulls[i];
        for (u32 i = u32(w->hulls.size()); i < lib->hull_count; ++i)
            if (!BuildHull(grown[i], lib->hulls[i])) *ok = false;
        w->hulls.swap(grown);
        for (Fixture& f : w->fixtures)
            if (f.shape.type == e_hull) f.shape.hull = &w->hulls[f.shape.hullIndex];
    }
    if (lib->mesh_count > w->meshes.size()) {
        std::vector<Mesh> grown(lib->mesh_count);
        for (u32 i = 0; i < w->meshes.size(); ++i) grown[i] = w->meshes[i];
        for (u32 i = u32(w->meshes.size()); i < lib->mesh_count; ++i)
            if (!BuildMesh(grown[i], lib->meshes[i])) *ok = false;
        w->meshes.swap(grown);
        for (Fixture& f : w->fixtures)
            if (f.shape.type == e_mesh) f.shape.mesh = &w->meshes[f.shape.meshIndex];
    }
}

static void AddBody(World* w, const bx_body_desc& d) {  // b3Body::b3Body (body.cpp:25-103)
    w->bodies.emplace_back();
    Body& b = w->bodies.back();
    b.type = d.type;
    b.flags = 0;
    if (d.flags & BX_BODY_AWAKE) b.flags |= e_awakeFlag;
    if (d.flags & BX_BODY_AUTO_SLEEP) b.flags |= e_autoSleepFlag;
    if (b.type == e_dynamicBody) {
        b.mass = 1; b.invMass = 1;
        b.flags |= d.flags & (BX_BODY_FIXED_ROT_X | BX_BODY_FIXED_ROT_Y | BX_BODY_FIXED_ROT_Z);
    } else {
        b.mass = 0; b.invMass = 0;
    }
    b.I.SetZero(); b.invI.SetZero(); b.worldInvI.SetZero();
    b.force.SetZero(); b.torque.SetZero();
    b.linearVelocity = V3(d.linear_velocity);
    b.angularVelocity = V3(d.angular_velocity);
    b.sweep.localCenter.SetZero();
    b.sweep.worldCenter = V3(d.position);
    b.sweep.orientation = Q4(d.orientation);
    b.sweep.worldCenter0 = V3(d.position);
    b.sweep.orientation0 = Q4(d.orientation);
    b.sweep.t0 = 0;
    b.xf.translation = b.sweep.worldCenter;
    b.xf.rotation = b.sweep.orientation;
    b.linearDamping = V3(d.linear_damping);
    b.angularDamping = V3(d.angular_damping);
    b.gravityScale = V3(d.gravity_scale);
    b.linearSleepTolerance = d.linear_sleep_tolerance;
    b.angularSleepTolerance = d.angular_sleep_tolerance;
    b.sleepTime = 0;
}

// b3Body::CreateFixture (body.cpp:105-136, fixture.cpp:43-56).
static bool AddFixture(World* w, const bx_fixture_desc& d) {
    if (d.body >= w->bodies.size()) return false;
    const u32 i = u32(w->fixtures.size());
    w->fixtures.emplace_back();
    Fixture& f = w->fixtures.back();
    f.body = d.body;
    f.friction = d.friction;
    f.restitution = d.restitution;
    f.isSensor = d.is_sensor != 0;
    f.density = d.density;
    Shape& s = f.shape;
    s.type = d.shape_type;
    switch (d.shape_type) {
    case BX_SHAPE_SPHERE: s.v1 = V3(d.vertex1); s.radius = d.radius; break;
    case BX_SHAPE_CAPSULE: s.v1 = V3(d.vertex1); s.v2 = V3(d.vertex2); s.radius = d.radius; break;
    case BX_SHAPE_TRIANGLE: InitTriangleShape(s, V3(d.vertex1), V3(d.vertex2), V3(d.vertex3)); break;
    case BX_SHAPE_HULL:
        if (d.shape_index >= w->hulls.size()) return false;
        s.hull = &w->hulls[d.shape_index];
        s.hullIndex = d.shape_index;
        s.radius = kHullRadius;
        break;
    case BX_SHAPE_MESH:
        if (d.shape_index >= w->meshes.size()) return false;
        s.mesh = &w->meshes[d.shape_index];
        s.meshIndex = d.shape_index;
        s.scale = V3(d.mesh_scale);
        s.radius = kHullRadius;
        break;
    default: return false;
    }
    Body& b = w->bodies[d.body];
    AABB aabb;
    ComputeAABB(s, b.xf, &aabb);
    f.proxyId = w->CreateProxy(aabb, i);
    if (f.proxyId == kNullU32) return false;  // a node whose id depends on the tree's shape
    b.fixtures.insert(b.fixtures.begin(), i);
    if (f.density > scalar(0)) w->ResetMass(d.body);
    w->newContacts = true;
    return true;
}

static World* Create(const bx_shape_library* lib, const bx_world_desc* wd) {
    World* w = new World();
    w->gravity = V3(wd->gravity);
    w->sleeping = (wd->flags & BX_WORLD_SLEEPING) != 0;
    w->warmStarting = (wd->flags & BX_WORLD_NO_WARM_START) == 0;
    w->clearForces = (wd->flags & BX_WORLD_NO_CLEAR_FORCES) == 0;
    w->recordEvents = (wd->flags & BX_WORLD_EVENTS) != 0;
    bool ok = true;
    AddLibrary(w, lib, &ok);
    if (!ok) { delete w; return nullptr; }
    for (u32 i = 0; i < wd->body_count; ++i) AddBody(w, wd->bodies[i]);
    for (u32 i = 0; i < wd->fixture_count; ++i)
        if (!AddFixture(w, wd->fixtures[i])) { delete w; return nullptr; }
    for (u32 i = 0; i < wd->joint_count; ++i) CreateJoint(*w, wd->joints[i]);
    return w;
}

// ------------------------------------------------------------------ topology changes
// b3JointManager::Destroy (joint_manager.cpp:75-141): both bodies wake, the edges go.
static void DestroyJoint(World* w, u32 j) {
    const Joint J = w->joints[j];
    w->bodies[J.bodyA].SetAwake(true);
    w->bodies[J.bodyB].SetAwake(true);
    for (Body& b : w->bodies) {
        b.joints.erase(std::remove_if(b.joints.begin(), b.joints.end(), [&](const JointEdge& e) { return e.joint == j; }),
                       b.joints.end());
        for (JointEdge& e : b.joints)
            if (e.joint > j) --e.joint;
    }
    w->joints.erase(w->joints.begin() + j);
}

// The fixture's contacts (b3ContactManager::Destroy) and proxy (b3BroadPhase::DestroyProxy) go;
// fixture indices above it move down.
static void RemoveFixture(World* w, u32 f) {
    while (!w->fixtures[f].contacts.empty()) w->DestroyContact(w->fixtures[f].contacts.front());
    w->DestroyProxy(w->fixtures[f].proxyId);
    Body& b = w->bodies[w->fixtures[f].body];
    b.fixtures.erase(std::find(b.fixtures.begin(), b.fixtures.end(), f));
    w->fixtures.erase(w->fixtures.begin() + f);
    for (Body& x : w->bodies)
        for (u32& k : x.fixtures)
            if (k > f) --k;
    for (Contact* c : w->contactList) {
        if (c->fixtureA > f) --c->fixtureA;
        if (c->fixtureB > f) --c->fixtureB;
    }
    for (u32& k : w->proxyFixture)
        if (k != kNullU32 && k > f) --k;
}

static int EditTopology(World* w, const bx_world_desc* adds, const bx_topology_op& op) {
    switch (op.op) {
    case BX_TOPO_CREATE_BODY: AddBody(w, adds->bodies[op.index]); return 0;
    case BX_TOPO_CREATE_FIXTURE: return AddFixture(w, adds->fixtures[op.index]) ? 0 : -1;
    case BX_TOPO_CREATE_JOINT: CreateJoint(*w, adds->joints[op.index]); return 0;
    case BX_TOPO_DESTROY_JOINT: DestroyJoint(w, op.index); return 0;
    case BX_TOPO_DESTROY_FIXTURE: {  // b3Body::DestroyFixture (body.cpp:138-191)
        const u32 body = w->fixtures[op.index].body;
        RemoveFixture(w, op.index);
        w->ResetMass(body);
        return 0;
    }
    case BX_TOPO_DESTROY_BODY: {  // b3World::DestroyBody (world.cpp:112-175)
        const u32 b = op.index;
        while (!w->bodies[b].joints.empty()) DestroyJoint(w, w->bodies[b].joints.front().joint);
        while (!w->bodies[b].fixtures.empty()) RemoveFixture(w, w->bodies[b].fixtures.front());
        w->bodies.erase(w->bodies.begin() + b);
        for (Fixture& f : w->fixtures)
            if (f.body > b) --f.body;
        for (Joint& j : w->joints) {
            if (j.bodyA > b) --j.bodyA;
            if (j.bodyB > b) --j.bodyB;
        }
        for (Body& x : w->bodies)
            for (JointEdge& e : x.joints)
                if (e.other > b) --e.other;
        return 0;
    }
    }
    return -1;
}

static void FillContact(const World& w, const Contact* c, bx_contact_state* out) {
    std::memset(out, 0, sizeof(*out));
    out->fixture_a = c->fixtureA;
    out->fixture_b = c->fixtureB;
    out->touching = (c->flags & 2u) ? 1u : 0u;
    out->manifold_count = c->manifoldCount;
    if (w.fixtures[c->fixtureA].shape.type == e_hull && w.fixtures[c->fixtureB].shape.type == e_hull) {
        out->cache[0] = c->cache.cacheType;
        out->cache[1] = c->cache.featureType;
        out->cache[2] = c->cache.index1;
        out->cache[3] = c->cache.index2;
    }
    for (u32 i = 0; i < c->manifoldCount && i < kMaxManifolds; ++i) {
        const Manifold& m = c->manifolds[i];
        bx_manifold_state& ms = out->manifolds[i];
        ms.point_count = m.pointCount;
        ms.tangent_impulse[0] = m.tangentImpulse.x;
        ms.tangent_impulse[1] = m.tangentImpulse.y;
        ms.motor_impulse = m.motorImpulse;
        for (u32 k = 0; k < m.pointCount; ++k) {
            const ManifoldPoint& p = m.points[k];
            bx_manifold_point_state& ps = ms.points[k];
            for (int a = 0; a < 3; ++a) {
                ps.local_normal1[a] = p.localNormal1[a];
                ps.local_point1[a] = p.localPoint1[a];
                ps.local_point2[a] = p.localPoint2[a];
            }
            ps.normal_impulse = p.normalImpulse;
            ps.triangle_key = p.id.triangleKey;
            ps.key[0] = p.id.pair.inEdge1;
            ps.key[1] = p.id.pair.outEdge1;
            ps.key[2] = p.id.pair.inEdge2;
            ps.key[3] = p.id.pair.outEdge2;
            ps.persisting = p.persisting ? 1u : 0u;
            ps.edge_contact = p.edgeContact ? 1u : 0u;
        }
    }
}

extern "C" {

void* orc_create(const bx_shape_library* lib, const bx_world_desc* wd) { return Create(lib, wd); }
void orc_destroy(void* h) { delete (World*)h; }

void orc_step(void* h, float dt, uint32_t vit, uint32_t pit, uint32_t n) {
    World* w = (World*)h;
    for (uint32_t i = 0; i < n; ++i) w->Step(dt, vit, pit);
}

// Body edits between steps (bx_body_edit, bxstep.h), restating b3Body's setters and force / impulse
// methods (body.h:588-840) and b3Body::SetTransform (body.cpp:194-214).
int orc_edit_bodies(void* h, const bx_body_edit* edits, uint32_t count) {
    World* w = (World*)h;
    for (uint32_t k = 0; k < count; ++k) {
        const bx_body_edit& e = edits[k];
        const Vec3 v(e.v[0], e.v[1], e.v[2]), p(e.p[0], e.p[1], e.p[2]);
        switch (e.op) {  // world edits (world.h:95-170, world.cpp:82-92, 244-251)
        case BX_EDIT_WORLD_GRAVITY: w->gravity = v; continue;
        case BX_EDIT_WORLD_SLEEPING:
            w->sleeping = e.wake != 0;
            if (!w->sleeping)
                for (Body& b : w->bodies) b.SetAwake(true);
            continue;
        case BX_EDIT_WORLD_WARM_STARTING: w->warmStarting = e.wake != 0; continue;
        case BX_EDIT_WORLD_AUTO_CLEAR_FORCES: w->clearForces = e.wake != 0; continue;
        case BX_EDIT_WORLD_CLEAR_FORCES:
            for (Body& b : w->bodies) {
                b.force.SetZero();
                b.torque.SetZero();
            }
            continue;
        default: break;
        }
        if (e.body >= w->bodies.size()) return -1;
        Body& b = w->bodies[e.body];
        const bool dynamic = b.type == e_dynamicBody;
        auto wakeFirst = [&]() {  // the Apply* methods: wake first if asked, act only on an awake body
            if (e.wake && !b.IsAwake()) b.SetAwake(true);
            return b.IsAwake();
        };
        auto setVelocity = [&](Vec3& dst) {  // body.h:652-685
            if (b.type == e_staticBody) return;
            if (Dot(v, v) > scalar(0)) b.SetAwake(true);
            dst = v;
        };
        switch (e.op) {
        case BX_EDIT_LINEAR_VELOCITY: setVelocity(b.linearVelocity); break;
        case BX_EDIT_ANGULAR_VELOCITY: setVelocity(b.angularVelocity); break;
        case BX_EDIT_FORCE:
            if (dynamic && wakeFirst()) {
                b.force += v;
                b.torque += Cross(p - b.sweep.worldCenter, v);
            }
            break;
        case BX_EDIT_FORCE_TO_CENTER:
            if (dynamic && wakeFirst()) b.force += v;
            break;
        case BX_EDIT_TORQUE:
            if (dynamic && wakeFirst()) b.torque += v;
            break;
        case BX_EDIT_LINEAR_IMPULSE:
            if (dynamic && wakeFirst()) {
                b.linearVelocity += b.invMass * v;
                b.angularVelocity += b.worldInvI * Cross(p - b.sweep.worldCenter, v);
            }
            break;
        case BX_EDIT_ANGULAR_IMPULSE:
            if (dynamic && wakeFirst()) b.angularVelocity += b.worldInvI * v;
            break;
        case BX_EDIT_TRANSFORM: {
            Transform xf;
            xf.translation = v;
            xf.rotation = Quat(p, e.p[3]);
            b.xf = xf;
            b.sweep.worldCenter = Mul(xf, b.sweep.localCenter);
            b.sweep.orientation = xf.rotation;
            b.sweep.worldCenter0 = b.sweep.worldCenter;
            b.sweep.orientation0 = b.sweep.orientation;
            b.worldInvI = RotateToFrame(b.invI, xf.rotation);
            for (u32 f : b.fixtures) w->SynchronizeFixture(f, xf, xf);
            w->newContacts = true;
            break;
        }
        case BX_EDIT_AWAKE: b.SetAwake(e.wake != 0); break;
        case BX_EDIT_SLEEPING_ALLOWED:
            if (e.wake) {
                b.flags |= e_autoSleepFlag;
            } else {
                b.flags &= ~u32(e_autoSleepFlag);
                b.SetAwake(true);
            }
            break;
        case BX_EDIT_GRAVITY_SCALE:
            if (b.type != e_staticBody) b.gravityScale = v;
            break;
        case BX_EDIT_LINEAR_DAMPING: b.linearDamping = v; break;
        case BX_EDIT_ANGULAR_DAMPING: b.angularDamping = v; break;
        case BX_EDIT_LINEAR_SLEEP_TOLERANCE: b.linearSleepTolerance = e.v[0]; break;
        case BX_EDIT_ANGULAR_SLEEP_TOLERANCE: b.angularSleepTolerance = e.v[0]; break;
        case BX_EDIT_FIXED_ROTATION: {  // body.cpp:507-548
            const u32 mask = e_fixedRotationX | e_fixedRotationY | e_fixedRotationZ;
            const u32 want = (e.v[0] != 0.0f ? u32(e_fixedRotationX) : 0u) |
                             (e.v[1] != 0.0f ? u32(e_fixedRotationY) : 0u) |
                             (e.v[2] != 0.0f ? u32(e_fixedRotationZ) : 0u);
            if ((b.flags & mask) == want) break;
            b.flags = (b.flags & ~mask) | want;
            b.angularVelocity.SetZero();
            w->ResetMass(e.body);
            break;
        }
        default: return -1;
        }
    }
    return 0;
}

// The filter as two bit tables over the fixtures (bx_batch_set_contact_filter's layout); null removes it.
int orc_set_contact_filter(void* h, const uint8_t* collide, const uint8_t* respond) {
    World* w = (World*)h;
    const size_t bytes = (w->fixtures.size() * w->fixtures.size() + 7) / 8;
    w->filterCollide.assign(collide, collide ? collide + bytes : collide);
    w->filterRespond.assign(respond, respond ? respond + bytes : respond);
    if (collide && !respond) w->filterRespond.assign(bytes, 0xFF);
    return 0;
}

// Events of the last step with contacts as positions in the world contact list after the step.
int orc_get_events(void* h, bx_contact_event* out, uint32_t cap, uint32_t* count) {
    World* w = (World*)h;
    uint32_t n = 0;
    for (const World::Event& e : w->events) {
        uint32_t idx = BX_NULL_INDEX;
        if (e.contact) {
            auto it = std::find(w->contactList.begin(), w->contactList.end(), e.contact);
            if (it != w->contactList.end()) idx = uint32_t(it - w->contactList.begin());
        }
        if (n < cap) out[n] = bx_contact_event{e.type, e.fixtureA, e.fixtureB, idx};
        ++n;
    }
    if (count) *count = n;
    return 0;
}

int orc_get_counts(void* h, bx_world_counts* out) {
    World* w = (World*)h;
    std::memset(out, 0, sizeof(*out));
    out->body_count = u32(w->bodies.size());
    out->fixture_count = u32(w->fixtures.size());
    out->joint_count = u32(w->joints.size());
    out->contact_count = u32(w->contactList.size());
    for (Contact* c : w->contactList) out->touching_count += (c->flags & 2u) ? 1u : 0u;
    out->pair_count = u32(w->lastPairs.size());
    out->island_count = w->lastIslandCount;
    for (const Body& b : w->bodies) out->awake_count += b.IsAwake() ? 1u : 0u;
    return 0;
}

int orc_get_bodies(void* h, bx_body_state* out, uint32_t cap) {
    World* w = (World*)h;
    u32 n = std::min<u32>(cap, u32(w->bodies.size()));
    for (u32 i = 0; i < n; ++i) {
        const Body& b = w->bodies[i];
        bx_body_state& s = out[i];
        std::memset(&s, 0, sizeof(s));
        for (int a = 0; a < 3; ++a) {
            s.position[a] = b.xf.translation[a];
            s.world_center[a] = b.sweep.worldCenter[a];
            s.linear_velocity[a] = b.linearVelocity[a];
            s.angular_velocity[a] = b.angularVelocity[a];
        }
        s.orientation[0] = b.xf.rotation.v.x;
        s.orientation[1] = b.xf.rotation.v.y;
        s.orientation[2] = b.xf.rotation.v.z;
        s.orientation[3] = b.xf.rotation.s;
        s.sleep_time = b.sleepTime;
        s.flags = b.flags;
        s.island_index = b.islandIndex;
    }
    return int(n);
}

int orc_get_contacts(void* h, bx_contact_state* out, uint32_t cap) {
    World* w = (World*)h;
    u32 n = 0;
    for (Contact* c : w->contactList) {
        if (n >= cap) break;
        FillContact(*w, c, out + n++);
    }
    return int(n);
}

int orc_get_pairs(void* h, bx_pair* out, uint32_t cap) {
    World* w = (World*)h;
    u32 n = 0;
    for (auto& p : w->lastPairs) {
        if (n < cap) { out[n].proxy1 = p.first; out[n].proxy2 = p.second; }
        ++n;
    }
    return int(n);
}

int orc_get_island_contacts(void* h, uint32_t* out, uint32_t cap, uint32_t* count) {
    World* w = (World*)h;
    u32 n = 0;
    for (Contact* c : w->lastIslandOrder) {
        auto it = std::find(w->contactList.begin(), w->contactList.end(), c);
        if (n < cap) out[n] = it == w->contactList.end() ? BX_NULL_INDEX : u32(it - w->contactList.begin());
        ++n;
    }
    *count = n;
    return 0;
}

int orc_get_proxy_ids(void* h, uint32_t* out, uint32_t cap) {
    World* w = (World*)h;
    u32 n = std::min<u32>(cap, u32(w->fixtures.size()));
    for (u32 i = 0; i < n; ++i) out[i] = w->fixtures[i].proxyId;
    return int(n);
}

int orc_get_fat_aabbs(void* h, float* out, uint32_t cap) {
    World* w = (World*)h;
    u32 n = std::min<u32>(cap, u32(w->fixtures.size()));
    for (u32 i = 0; i < n; ++i) {
        const AABB& a = w->fatAABB[w->fixtures[i].proxyId];
        for (int k = 0; k < 3; ++k) { out[6 * i + k] = a.lowerBound[k]; out[6 * i + 3 + k] = a.upperBound[k]; }
    }
    return int(n);
}

int orc_get_mass(void* h, float* out, uint32_t cap) {
    World* w = (World*)h;
    u32 n = std::min<u32>(cap, u32(w->bodies.size()));
    for (u32 i = 0; i < n; ++i) {
        const Body& b = w->bodies[i];
        float* o = out + 23 * i;
        o[0] = b.mass; o[1] = b.invMass;
        for (int a = 0; a < 3; ++a) o[2 + a] = b.sweep.localCenter[a];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) { o[5 + 3 * c + r] = b.I[c][r]; o[14 + 3 * c + r] = b.invI[c][r]; }
    }
    return int(n);
}

int orc_edit_topology(void* h, const bx_shape_library* lib, const bx_world_desc* adds, const bx_topology_op* ops,
                      uint32_t count) {
    World* w = (World*)h;
    bool ok = true;
    AddLibrary(w, lib, &ok);
    if (!ok) return -1;
    for (uint32_t k = 0; k < count; ++k)
        if (EditTopology(w, adds, ops[k])) return -1;
    w->lastIslandOrder.clear();
    return 0;
}

// World queries (orc_query.cpp), same contract as the reference harness's entries.
int orc_raycast(void* h, const float* p1, const float* p2, const uint8_t* mask, bx_ray_hit* out, uint32_t cap,
                uint32_t* count) {
    World* w = (World*)h;
    std::vector<RayHit> hits;
    RayCastAll(*w, V3(p1), V3(p2), mask, hits);
    *count = u32(hits.size());
    for (u32 i = 0; i < cap && i < hits.size(); ++i) {
        out[i].fixture = hits[i].fixture;
        out[i].fraction = hits[i].fraction;
        for (int a = 0; a < 3; ++a) {
            out[i].point[a] = hits[i].point[a];
            out[i].normal[a] = hits[i].normal[a];
        }
    }
    return 0;
}

int orc_raycast_single(void* h, const float* p1, const float* p2, const uint8_t* mask, bx_ray_hit* out) {
    uint32_t n = 0;
    std::memset(out, 0, sizeof(*out));
    out->fixture = BX_NULL_INDEX;
    orc_raycast(h, p1, p2, mask, out, 1, &n);
    if (n == 0) {
        out->fixture = BX_NULL_INDEX;
        return 0;
    }
    return 1;
}

int orc_query_aabb(void* h, const float* lower, const float* upper, const uint8_t* mask, uint32_t* out, uint32_t cap,
                   uint32_t* count) {
    World* w = (World*)h;
    AABB box;
    box.lowerBound = V3(lower);
    box.upperBound = V3(upper);
    std::vector<u32> found;
    QueryAABB(*w, box, mask, found);
    *count = u32(found.size());
    for (u32 i = 0; i < cap && i < found.size(); ++i) out[i] = found[i];
    return 0;
}

// ------------------------------------------------------------------ time of impact and shape casts
// A cast shape (bx_cast_shape) as a Shape; a hull names one of `hulls`. False when the type or the hull
// index is out of range.
static bool CastShape(const bx_cast_shape& c, const std::vector<Hull>* hulls, const bx_shape_library* lib,
                      Shape& s, GJKProxy& proxy) {
    s = Shape();
    s.type = c.shape_type;
    s.radius = c.radius;
    s.v1 = V3(c.vertex1);
    s.v2 = V3(c.vertex2);
    s.v3 = V3(c.vertex3);
    s.hasE1 = s.hasE2 = s.hasE3 = false;
    if (c.shape_type == BX_SHAPE_HULL) {
        if (hulls) {
            if (c.shape_index >= hulls->size()) return false;
            s.hull = &(*hulls)[c.shape_index];
            s.hullIndex = c.shape_index;
        } else {  // a library hull's vertices as the proxy (gjk.cpp:63-70)
            if (!lib || c.shape_index >= lib->hull_count) return false;
            const bx_hull_desc& d = lib->hulls[c.shape_index];
            proxy.vertices = reinterpret_cast<const Vec3*>(d.vertices);
            proxy.count = d.vertex_count;
            proxy.radius = c.radius;
            return true;
        }
    } else if (c.shape_type > BX_SHAPE_HULL || c.shape_type == BX_SHAPE_MESH) {
        return false;
    }
    proxy.Set(s);
    return true;
}

static Transform XF(const float* p, const float* q) {
    return Transform(V3(p), Quat(Vec3(q[0], q[1], q[2]), q[3]));
}

static Sweep SW(const bx_sweep& s) {
    Sweep w;
    w.localCenter = V3(s.local_center);
    w.worldCenter0 = V3(s.center0);
    w.worldCenter = V3(s.center);
    w.orientation0 = Quat(Vec3(s.orientation0[0], s.orientation0[1], s.orientation0[2]), s.orientation0[3]);
    w.orientation = Quat(Vec3(s.orientation[0], s.orientation[1], s.orientation[2]), s.orientation[3]);
    w.t0 = s.t0;
    return w;
}

// The three b3TimeOfImpact forms (bx_toi_query, bxstep.h) over the hulls of `lib`.
int orc_time_of_impact(const bx_shape_library* lib, const bx_toi_query* q, uint32_t count, bx_toi_result* out) {
    for (uint32_t i = 0; i < count; ++i) {
        const bx_toi_query& x = q[i];
        TOIOutput o;
        if (x.kind == BX_TOI_AABB) {
            AABB a, b;
            a.lowerBound = V3(x.lower1);
            a.upperBound = V3(x.upper1);
            b.lowerBound = V3(x.lower2);
            b.upperBound = V3(x.upper2);
            o = TimeOfImpact(a, V3(x.displacement1), b, V3(x.displacement2));
        } else {
            Shape s1, s2;
            GJKProxy p1, p2;
            if (!CastShape(x.shape1, nullptr, lib, s1, p1) || !CastShape(x.shape2, nullptr, lib, s2, p2)) return -1;
            if (x.kind == BX_TOI_SWEEP) {
                o = TimeOfImpact(p1, SW(x.sweep1), p2, SW(x.sweep2), x.t_max);
            } else if (x.kind == BX_TOI_TRANSLATION) {
                o = TimeOfImpact(XF(x.position1, x.orientation1), p1, V3(x.displacement1),
                                 XF(x.position2, x.orientation2), p2, V3(x.displacement2), x.max_iterations);
            } else {
                return -1;
            }
        }
        out[i].state = o.state;
        out[i].t = o.t;
        out[i].iterations = o.iterations;
    }
    return 0;
}

// b3World::ShapeCast with a listener returning 1: every hit, in increasing (fraction, fixture, triangle).
int orc_shape_cast(void* h, const bx_shape_cast_query* q, const uint8_t* mask, bx_shape_hit* out, uint32_t cap,
                   uint32_t* count) {
    World* w = (World*)h;
    Shape s;
    GJKProxy unused;
    if (!CastShape(q->shape, &w->hulls, nullptr, s, unused)) return -1;
    std::vector<ShapeHit> hits;
    ShapeCastAll(*w, s, XF(q->position, q->orientation), V3(q->displacement), mask, hits);
    *count = u32(hits.size());
    for (u32 i = 0; i < cap && i < hits.size(); ++i) {
        out[i].fixture = hits[i].fixture;
        out[i].child = hits[i].child;
        out[i].fraction = hits[i].fraction;
        for (int a = 0; a < 3; ++a) {
            out[i].point[a] = hits[i].point[a];
            out[i].normal[a] = hits[i].normal[a];
        }
    }
    return 0;
}

// b3World::ShapeCastSingle: the first hit of that order (fixture BX_NULL_INDEX for none).
int orc_shape_cast_single(void* h, const bx_shape_cast_query* q, const uint8_t* mask, bx_shape_hit* out) {
    uint32_t n = 0;
    std::memset(out, 0, sizeof(*out));
    int rc = orc_shape_cast(h, q, mask, out, 1, &n);
    if (rc) return rc;
    if (n == 0) out->fixture = BX_NULL_INDEX;
    return 0;
}

// CPU baseline with the restatement: same contract as the reference harness's bench entry.
double orc_bench(const bx_shape_library* lib, const bx_world_desc* worlds, uint32_t world_count, float dt,
                 uint32_t vit, uint32_t pit, uint32_t warmup, uint32_t steps, uint32_t threads) {
    if (threads == 0) threads = 1;
    std::atomic<uint32_t> next{0};
    std::vector<double> busy(threads, 0.0);
    auto worker = [&](uint32_t t) {
        for (;;) {
            uint32_t i = next.fetch_add(1);
            if (i >= world_count) break;
            World* w = Create(lib, worlds + i);
            if (!w) continue;
            for (uint32_t s = 0; s < warmup; ++s) w->Step(dt, vit, pit);
            auto t0 = std::chrono::steady_clock::now();
            for (uint32_t s = 0; s < steps; ++s) w->Step(dt, vit, pit);
            auto t1 = std::chrono::steady_clock::now();
            busy[t] += std::chrono::duration<double>(t1 - t0).count();
            delete w;
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < threads; ++t) pool.emplace_back(worker, t);
    for (auto& th : pool) th.join();
    double wall = 0.0;
    for (double b : busy) wall = std::max(wall, b);
    return wall;
}

}  // extern "C"
