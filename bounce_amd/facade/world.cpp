// world.cpp -- the drop-in b3World over the C ABI of libbxstep (include/bxstep.h).
//
// The device world is the state of record. Each call that changes the world (creating or destroying
// bodies, fixtures and joints; the body setters, forces and impulses; the world's settings) is queued in
// call order with the creation indices it names at that moment. Before a step, a query, or a getter that
// reads state a queued call changes, the queue is applied to the resident world: runs of topology
// changes through bx_batch_edit_topology, runs of body and world edits through bx_batch_edit_bodies, so
// contacts, impulses, proxies and sleep state carry across them as they do in the reference. The first
// time, the leading run of creations and world settings becomes the world's descriptor
// (bx_batch_create). Getters read back what the device computed, once per change.
#include <algorithm>
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
namespace {
void Copy3(float* out, const b3Vec3& v) { out[0] = v.x; out[1] = v.y; out[2] = v.z; }
void Copy4(float* out, const b3Quat& q) { out[0] = q.v.x; out[1] = q.v.y; out[2] = q.v.z; out[3] = q.s; }
b3Vec3 V3(const float* p) { return b3Vec3(p[0], p[1], p[2]); }
b3Mat33 M33(const float* m) {  // column-major
    return b3Mat33(b3Vec3(m[0], m[1], m[2]), b3Vec3(m[3], m[4], m[5]), b3Vec3(m[6], m[7], m[8]));
}

// The library keeps its own copy of every hull and mesh it has sent to the device: the caller's arrays
// only need to live until the fixture is created, and a later, grown library must begin with the same
// bytes.
struct OwnedHull {
    const b3Hull* source;
    b3Vec3 centroid;
    std::vector<b3Vec3> vertices;
    std::vector<b3HalfEdge> edges;
    std::vector<b3Face> faces;
    std::vector<b3Plane> planes;
    bool Same(const b3Hull* h) const {
        return h->vertexCount == vertices.size() && h->edgeCount == edges.size() && h->faceCount == faces.size() &&
               std::memcmp(&h->centroid, &centroid, sizeof(b3Vec3)) == 0 &&
               std::memcmp(h->vertices, vertices.data(), vertices.size() * sizeof(b3Vec3)) == 0 &&
               std::memcmp(h->edges, edges.data(), edges.size() * sizeof(b3HalfEdge)) == 0 &&
               std::memcmp(h->faces, faces.data(), faces.size() * sizeof(b3Face)) == 0 &&
               std::memcmp(h->planes, planes.data(), planes.size() * sizeof(b3Plane)) == 0;
    }
};

struct OwnedMesh {
    const b3Mesh* source;
    std::vector<b3Vec3> vertices;
    std::vector<uint32> triangles;  // v1 v2 v3 per triangle
    bool Same(const b3Mesh* m) const {
        if (m->vertexCount != vertices.size() || 3 * size_t(m->triangleCount) != triangles.size()) return false;
        if (std::memcmp(m->vertices, vertices.data(), vertices.size() * sizeof(b3Vec3)) != 0) return false;
        for (uint32 k = 0; k < m->triangleCount; ++k)
            if (m->triangles[k].v1 != triangles[3 * k] || m->triangles[k].v2 != triangles[3 * k + 1] ||
                m->triangles[k].v3 != triangles[3 * k + 2])
                return false;
        return true;
    }
};

struct QueuedCall {
    bool topology;
    bx_topology_op topo;  // index into the descriptor arrays below (creations) or a creation index
    bx_body_edit edit;
};
}  // namespace

struct b3WorldDevice {
    bx_context* ctx = nullptr;
    bx_batch* batch = nullptr;
    std::string error;
    // shape library
    std::vector<OwnedHull> hulls;
    std::vector<OwnedMesh> meshes;
    std::unordered_map<const b3Hull*, uint32> hullOf;
    std::unordered_map<const b3Mesh*, uint32> meshOf;
    uint32 hullsOnDevice = 0, meshesOnDevice = 0;
    std::vector<bx_hull_desc> hullDescs;
    std::vector<bx_mesh_desc> meshDescs;
    // calls not yet applied to the device world, and the descriptors their creations name
    std::vector<QueuedCall> queue;
    std::vector<bx_body_desc> newBodies;
    std::vector<bx_fixture_desc> newFixtures;
    std::vector<bx_joint_desc> newJoints;
    // what the host copies of device state miss
    bool stateStale = true, massStale = true, contactsStale = true;
    std::vector<b3Contact*> contacts;  // the world's contact list, newest first
    std::vector<bx_body_state> bodyStates;
    std::vector<bx_body_mass> masses;
    std::vector<bx_contact_state> contactStates;

    const bx_shape_library& Library() {
        hullDescs.resize(hulls.size());
        for (size_t i = 0; i < hulls.size(); ++i) {
            const OwnedHull& h = hulls[i];
            bx_hull_desc& d = hullDescs[i];
            d.vertex_count = uint32(h.vertices.size());
            d.edge_count = uint32(h.edges.size());
            d.face_count = uint32(h.faces.size());
            d.vertices = &h.vertices[0].x;
            d.edges = &h.edges[0].origin;
            d.faces = &h.faces[0].edge;
            d.planes = &h.planes[0].normal.x;
            Copy3(d.centroid, h.centroid);
        }
        meshDescs.resize(meshes.size());
        for (size_t i = 0; i < meshes.size(); ++i) {
            const OwnedMesh& m = meshes[i];
            bx_mesh_desc& d = meshDescs[i];
            d.vertex_count = uint32(m.vertices.size());
            d.triangle_count = uint32(m.triangles.size() / 3);
            d.vertices = m.vertices.empty() ? nullptr : &m.vertices[0].x;
            d.triangles = m.triangles.data();
        }
        std::memset(&lib, 0, sizeof(lib));
        lib.hull_count = uint32(hullDescs.size());
        lib.hulls = hullDescs.data();
        lib.mesh_count = uint32(meshDescs.size());
        lib.meshes = meshDescs.data();
        return lib;
    }
    void ClearContacts() {
        for (b3Contact* c : contacts) delete c;
        contacts.clear();
    }
    bx_shape_library lib;
};

uint32 b3World::HullIndex(const b3Hull* h) {
    b3WorldDevice& D = *m_device;
    auto it = D.hullOf.find(h);
    if (it != D.hullOf.end() && D.hulls[it->second].Same(h)) return it->second;
    OwnedHull o;
    o.source = h;
    o.centroid = h->centroid;
    o.vertices.assign(h->vertices, h->vertices + h->vertexCount);
    o.edges.assign(h->edges, h->edges + h->edgeCount);
    o.faces.assign(h->faces, h->faces + h->faceCount);
    o.planes.assign(h->planes, h->planes + h->faceCount);
    D.hulls.push_back(std::move(o));
    return D.hullOf[h] = uint32(D.hulls.size() - 1);
}

uint32 b3World::MeshIndex(const b3Mesh* m) {
    b3WorldDevice& D = *m_device;
    auto it = D.meshOf.find(m);
    if (it != D.meshOf.end() && D.meshes[it->second].Same(m)) return it->second;
    OwnedMesh o;
    o.source = m;
    o.vertices.assign(m->vertices, m->vertices + m->vertexCount);
    o.triangles.resize(3 * size_t(m->triangleCount));
    for (uint32 k = 0; k < m->triangleCount; ++k) {
        o.triangles[3 * k] = m->triangles[k].v1;
        o.triangles[3 * k + 1] = m->triangles[k].v2;
        o.triangles[3 * k + 2] = m->triangles[k].v3;
    }
    D.meshes.push_back(std::move(o));
    return D.meshOf[m] = uint32(D.meshes.size() - 1);
}

void b3World::QueueTopology(uint32 op, uint32 index) {
    QueuedCall c;
    std::memset(&c, 0, sizeof(c));
    c.topology = true;
    c.topo.op = op;
    c.topo.index = index;
    m_device->queue.push_back(c);
    m_device->stateStale = m_device->massStale = m_device->contactsStale = true;
}

void b3World::QueueEdit(const void* edit, bool changesState) {
    QueuedCall c;
    std::memset(&c, 0, sizeof(c));
    c.topology = false;
    std::memcpy(&c.edit, edit, sizeof(bx_body_edit));
    m_device->queue.push_back(c);
    if (changesState) m_device->stateStale = m_device->massStale = m_device->contactsStale = true;
}

namespace {
bx_body_edit WorldEdit(uint32 op, bool flag) {
    bx_body_edit e;
    std::memset(&e, 0, sizeof(e));
    e.op = op;
    e.wake = flag ? 1u : 0u;
    return e;
}
}  // namespace

int b3World::Flush() {
    b3WorldDevice& D = *m_device;
    if (D.queue.empty() && D.batch) return BX_OK;
    int rc = BX_OK;
    if (!D.ctx && (rc = bx_context_create(0, &D.ctx)) != BX_OK) return rc;
    size_t k = 0;
    if (!D.batch) {
        // the leading creations and world settings are the world's descriptor
        std::vector<bx_body_desc> bodies;
        std::vector<bx_fixture_desc> fixtures;
        std::vector<bx_joint_desc> joints;
        bx_world_desc wd;
        std::memset(&wd, 0, sizeof(wd));
        wd.gravity[1] = -9.8f;  // world.cpp:41-55 defaults: sleeping off, warm starting and clearing on
        bool leading = true;
        for (; k < D.queue.size() && leading; ++k) {
            const QueuedCall& c = D.queue[k];
            if (c.topology) {
                switch (c.topo.op) {
                case BX_TOPO_CREATE_BODY: bodies.push_back(D.newBodies[c.topo.index]); break;
                case BX_TOPO_CREATE_FIXTURE: fixtures.push_back(D.newFixtures[c.topo.index]); break;
                case BX_TOPO_CREATE_JOINT: joints.push_back(D.newJoints[c.topo.index]); break;
                default: leading = false; break;
                }
                continue;
            }
            const bx_body_edit& e = c.edit;
            switch (e.op) {
            case BX_EDIT_WORLD_GRAVITY: std::memcpy(wd.gravity, e.v, sizeof(wd.gravity)); break;
            case BX_EDIT_WORLD_SLEEPING:
                if (e.wake) {
                    wd.flags |= BX_WORLD_SLEEPING;
                } else {
                    wd.flags &= ~BX_WORLD_SLEEPING;
                    for (bx_body_desc& b : bodies) b.flags |= BX_BODY_AWAKE;  // b3Body::SetAwake(true)
                }
                break;
            case BX_EDIT_WORLD_WARM_STARTING:
                wd.flags = e.wake ? (wd.flags & ~BX_WORLD_NO_WARM_START) : (wd.flags | BX_WORLD_NO_WARM_START);
                break;
            case BX_EDIT_WORLD_AUTO_CLEAR_FORCES:
                wd.flags = e.wake ? (wd.flags & ~BX_WORLD_NO_CLEAR_FORCES) : (wd.flags | BX_WORLD_NO_CLEAR_FORCES);
                break;
            case BX_EDIT_WORLD_CLEAR_FORCES: break;  // nothing has a force yet
            default: leading = false; break;
            }
        }
        if (!leading) --k;  // the call that ended the run is applied below
        if (bodies.empty()) return BX_OK;  // nothing to simulate yet: keep the queue
        wd.body_count = uint32(bodies.size());
        wd.bodies = bodies.data();
        wd.fixture_count = uint32(fixtures.size());
        wd.fixtures = fixtures.data();
        wd.joint_count = uint32(joints.size());
        wd.joints = joints.data();
        rc = bx_batch_create(D.ctx, &D.Library(), &wd, 1, nullptr, &D.batch);
        if (rc != BX_OK) return rc;
        D.hullsOnDevice = uint32(D.hulls.size());
        D.meshesOnDevice = uint32(D.meshes.size());
    }
    // the rest in runs of one kind
    while (k < D.queue.size() && rc == BX_OK) {
        size_t e = k;
        if (D.queue[k].topology) {
            std::vector<bx_topology_op> ops;
            while (e < D.queue.size() && D.queue[e].topology) ops.push_back(D.queue[e++].topo);
            bx_world_desc adds;
            std::memset(&adds, 0, sizeof(adds));
            adds.body_count = uint32(D.newBodies.size());
            adds.bodies = D.newBodies.data();
            adds.fixture_count = uint32(D.newFixtures.size());
            adds.fixtures = D.newFixtures.data();
            adds.joint_count = uint32(D.newJoints.size());
            adds.joints = D.newJoints.data();
            const bool grown = D.hulls.size() > D.hullsOnDevice || D.meshes.size() > D.meshesOnDevice;
            rc = bx_batch_edit_topology(D.batch, 0, grown ? &D.Library() : nullptr, &adds, ops.data(),
                                        uint32(ops.size()));
            if (rc == BX_OK && grown) {
                D.hullsOnDevice = uint32(D.hulls.size());
                D.meshesOnDevice = uint32(D.meshes.size());
            }
        } else {
            std::vector<bx_body_edit> edits;
            while (e < D.queue.size() && !D.queue[e].topology) edits.push_back(D.queue[e++].edit);
            rc = bx_batch_edit_bodies(D.batch, 0, edits.data(), uint32(edits.size()));
        }
        if (rc == BX_OK) k = e;
    }
    D.queue.erase(D.queue.begin(), D.queue.begin() + k);
    if (D.queue.empty()) {
        D.newBodies.clear();
        D.newFixtures.clear();
        D.newJoints.clear();
    }
    return rc;
}

void b3World::ReadBodies() {
    b3WorldDevice& D = *m_device;
    D.bodyStates.resize(m_bodies.size());
    int rc = bx_batch_get_bodies(D.batch, 0, D.bodyStates.data(), uint32(m_bodies.size()));
    if (rc < 0) {
        D.error = bx_last_error();
        return;
    }
    for (size_t i = 0; i < m_bodies.size(); ++i) {
        b3Body* b = m_bodies[i];
        const bx_body_state& x = D.bodyStates[i];
        b->m_xf.translation = V3(x.position);
        b->m_xf.rotation = b3Quat(b3Vec3(x.orientation[0], x.orientation[1], x.orientation[2]), x.orientation[3]);
        b->m_worldCenter = V3(x.world_center);
        b->m_linearVelocity = V3(x.linear_velocity);
        b->m_angularVelocity = V3(x.angular_velocity);
        b->m_awake = (x.flags & BX_BODY_AWAKE) != 0;
        b->m_sleepTime = x.sleep_time;
    }
    D.stateStale = false;
}

void b3World::ReadMass() {
    b3WorldDevice& D = *m_device;
    D.masses.resize(m_bodies.size());
    int rc = bx_batch_get_mass(D.batch, 0, D.masses.data(), uint32(m_bodies.size()));
    if (rc < 0) {
        D.error = bx_last_error();
        return;
    }
    for (size_t i = 0; i < m_bodies.size(); ++i) {
        b3Body* b = m_bodies[i];
        const bx_body_mass& m = D.masses[i];
        b->m_mass = m.mass;
        b->m_invMass = m.inv_mass;
        b->m_localCenter = V3(m.local_center);
        b->m_I = M33(m.inertia);
        b->m_worldInvI = M33(m.world_inv_inertia);
    }
    D.massStale = false;
}

void b3World::Renumber() {
    for (uint32 i = 0; i < m_bodies.size(); ++i) m_bodies[i]->m_index = i;
    for (uint32 i = 0; i < m_fixtures.size(); ++i) m_fixtures[i]->m_index = i;
    for (uint32 i = 0; i < m_joints.size(); ++i) m_joints[i]->m_index = i;
}

// ------------------------------------------------------------------ b3Fixture / b3Body
b3Fixture::~b3Fixture() { delete m_shape; }

b3Body::b3Body(const b3BodyDef& def, b3World* world, uint32 index)
    : m_def(def), m_type(def.type), m_world(world), m_index(index), m_xf(def.position, def.orientation),
      m_worldCenter(def.position), m_localCenter(scalar(0), scalar(0), scalar(0)),
      m_linearVelocity(def.linearVelocity), m_angularVelocity(def.angularVelocity), m_awake(def.awake),
      m_sleepTime(0), m_mass(0), m_invMass(0), m_linearDamping(def.linearDamping),
      m_angularDamping(def.angularDamping), m_gravityScale(def.gravityScale),
      m_linearSleepTolerance(def.linearSleepTolerance), m_angularSleepTolerance(def.angularSleepTolerance),
      m_allowSleep(def.allowSleep), m_fixedX(def.fixedRotationX), m_fixedY(def.fixedRotationY),
      m_fixedZ(def.fixedRotationZ), m_fixtureList(nullptr), m_fixtureCount(0), m_jointList(nullptr),
      m_next(nullptr), m_userData(def.userData) {
    m_I.SetZero();
    m_worldInvI.SetZero();
}

b3Body::~b3Body() {
    b3Fixture* f = m_fixtureList;
    while (f) {
        b3Fixture* n = f->m_next;
        delete f;
        f = n;
    }
}

const b3Body& b3Body::State() const {
    b3World* w = m_world;
    if (w->m_device->stateStale) {
        int rc = w->Flush();
        if (rc == BX_OK && w->m_device->batch) w->ReadBodies();
        else if (rc < 0) w->m_device->error = bx_last_error();
    }
    return *this;
}

const b3Body& b3Body::Mass() const {
    State();
    b3World* w = m_world;
    if (w->m_device->massStale) {
        int rc = w->Flush();
        if (rc == BX_OK && w->m_device->batch) w->ReadMass();
        else if (rc < 0) w->m_device->error = bx_last_error();
    }
    return *this;
}

scalar b3Body::GetAngularEnergy() const {  // body.h:713-718
    const b3Body& b = Mass();
    b3Mat33 I = b3RotateToFrame(b.m_I, b.m_xf.rotation);
    b3Vec3 L = I * b.m_angularVelocity;
    return b3Dot(L, b.m_angularVelocity);
}

void b3Body::QueueEdit(uint32 op, const b3Vec3& v, const float* p, bool flag, bool changesState) {
    bx_body_edit e;
    std::memset(&e, 0, sizeof(e));
    e.body = m_index;
    e.op = op;
    e.wake = flag ? 1u : 0u;
    Copy3(e.v, v);
    if (p) std::memcpy(e.p, p, sizeof(e.p));
    m_world->QueueEdit(&e, changesState);
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

    bx_fixture_desc d;
    std::memset(&d, 0, sizeof(d));
    d.body = m_index;
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
        d.shape_index = m_world->MeshIndex(m->m_mesh);
        Copy3(d.mesh_scale, m->m_scale);
    } else {
        d.shape_index = m_world->HullIndex(static_cast<const b3HullShape*>(f->m_shape)->m_hull);
    }
    b3WorldDevice& D = *m_world->m_device;
    D.newFixtures.push_back(d);
    m_world->QueueTopology(BX_TOPO_CREATE_FIXTURE, uint32(D.newFixtures.size() - 1));

    f->m_index = uint32(m_world->m_fixtures.size());
    f->m_next = m_fixtureList;  // the reference prepends (body.cpp:120-122)
    m_fixtureList = f;
    ++m_fixtureCount;
    m_world->m_fixtures.push_back(f);
    return f;
}

void b3Body::DestroyFixture(b3Fixture* f) {
    if (!f || f->m_body != this) return;
    m_world->QueueTopology(BX_TOPO_DESTROY_FIXTURE, f->m_index);
    b3Fixture** link = &m_fixtureList;
    while (*link && *link != f) link = &(*link)->m_next;
    if (*link) *link = f->m_next;
    --m_fixtureCount;
    std::vector<b3Fixture*>& all = m_world->m_fixtures;
    all.erase(std::find(all.begin(), all.end(), f));
    delete f;
    m_world->Renumber();
}

void b3Body::SetTransform(const b3Transform& xf) {
    float q[4];
    Copy4(q, xf.rotation);
    QueueEdit(BX_EDIT_TRANSFORM, xf.translation, q, false, true);
}

void b3Body::SetLinearVelocity(const b3Vec3& v) {
    if (m_type == e_staticBody) return;
    QueueEdit(BX_EDIT_LINEAR_VELOCITY, v, nullptr, false, true);
}

void b3Body::SetAngularVelocity(const b3Vec3& w) {
    if (m_type == e_staticBody) return;
    QueueEdit(BX_EDIT_ANGULAR_VELOCITY, w, nullptr, false, true);
}

// A force or torque changes nothing a getter reads unless it wakes the body.
void b3Body::ApplyForce(const b3Vec3& force, const b3Vec3& point, bool wake) {
    if (m_type != e_dynamicBody) return;
    float p[4] = {point.x, point.y, point.z, 0.0f};
    QueueEdit(BX_EDIT_FORCE, force, p, wake, wake && !State().m_awake);
}

void b3Body::ApplyForceToCenter(const b3Vec3& force, bool wake) {
    if (m_type != e_dynamicBody) return;
    QueueEdit(BX_EDIT_FORCE_TO_CENTER, force, nullptr, wake, wake && !State().m_awake);
}

void b3Body::ApplyTorque(const b3Vec3& torque, bool wake) {
    if (m_type != e_dynamicBody) return;
    QueueEdit(BX_EDIT_TORQUE, torque, nullptr, wake, wake && !State().m_awake);
}

void b3Body::ApplyLinearImpulse(const b3Vec3& impulse, const b3Vec3& point, bool wake) {
    if (m_type != e_dynamicBody) return;
    float p[4] = {point.x, point.y, point.z, 0.0f};
    QueueEdit(BX_EDIT_LINEAR_IMPULSE, impulse, p, wake, true);
}

void b3Body::ApplyAngularImpulse(const b3Vec3& impulse, bool wake) {
    if (m_type != e_dynamicBody) return;
    QueueEdit(BX_EDIT_ANGULAR_IMPULSE, impulse, nullptr, wake, true);
}

void b3Body::SetLinearDamping(const b3Vec3& damping) {
    m_linearDamping = damping;
    QueueEdit(BX_EDIT_LINEAR_DAMPING, damping, nullptr, false, false);
}

void b3Body::SetAngularDamping(const b3Vec3& damping) {
    m_angularDamping = damping;
    QueueEdit(BX_EDIT_ANGULAR_DAMPING, damping, nullptr, false, false);
}

void b3Body::SetGravityScale(const b3Vec3& scale) {
    if (m_type == e_staticBody) return;
    m_gravityScale = scale;
    QueueEdit(BX_EDIT_GRAVITY_SCALE, scale, nullptr, false, false);
}

void b3Body::SetLinearSleepTolerance(scalar tolerance) {
    m_linearSleepTolerance = tolerance;
    QueueEdit(BX_EDIT_LINEAR_SLEEP_TOLERANCE, b3Vec3(tolerance, 0, 0), nullptr, false, false);
}

void b3Body::SetAngularSleepTolerance(scalar tolerance) {
    m_angularSleepTolerance = tolerance;
    QueueEdit(BX_EDIT_ANGULAR_SLEEP_TOLERANCE, b3Vec3(tolerance, 0, 0), nullptr, false, false);
}

void b3Body::SetFixedRotation(bool flagX, bool flagY, bool flagZ) {
    m_fixedX = flagX;
    m_fixedY = flagY;
    m_fixedZ = flagZ;
    QueueEdit(BX_EDIT_FIXED_ROTATION, b3Vec3(flagX ? 1.0f : 0.0f, flagY ? 1.0f : 0.0f, flagZ ? 1.0f : 0.0f), nullptr,
              false, true);
}

void b3Body::SetAwake(bool flag) { QueueEdit(BX_EDIT_AWAKE, b3Vec3(0, 0, 0), nullptr, flag, true); }

void b3Body::SetSleepingAllowed(bool flag) {
    m_allowSleep = flag;
    QueueEdit(BX_EDIT_SLEEPING_ALLOWED, b3Vec3(0, 0, 0), nullptr, flag, !flag);
}

// ------------------------------------------------------------------ b3Contact
void b3Contact::GetWorldManifold(b3WorldManifold* out, uint32 index) const {  // contact.cpp:165-179
    const b3Fixture* fA = m_fixtureA;
    const b3Fixture* fB = m_fixtureB;
    out->Initialize(m_manifolds + index, fA->GetShape()->m_radius, fA->GetBody()->GetTransform(),
                    fB->GetShape()->m_radius, fB->GetBody()->GetTransform());
}

// ------------------------------------------------------------------ b3World
b3World::b3World()
    : m_gravity(scalar(0), scalar(-9.8), scalar(0)), m_sleeping(false), m_warmStarting(true), m_autoClearForces(true),
      m_bodyList(nullptr), m_jointList(nullptr), m_device(new b3WorldDevice()) {}

b3World::~b3World() {
    m_device->ClearContacts();
    if (m_device->batch) bx_batch_destroy(m_device->batch);
    if (m_device->ctx) bx_context_destroy(m_device->ctx);
    delete m_device;
    for (b3Joint* j : m_joints) delete j;
    for (b3Body* b : m_bodies) delete b;
}

void b3World::SetSleeping(bool flag) {
    m_sleeping = flag;
    bx_body_edit e = WorldEdit(BX_EDIT_WORLD_SLEEPING, flag);
    QueueEdit(&e, !flag);
}

void b3World::SetWarmStarting(bool flag) {
    m_warmStarting = flag;
    bx_body_edit e = WorldEdit(BX_EDIT_WORLD_WARM_STARTING, flag);
    QueueEdit(&e, false);
}

void b3World::SetGravity(const b3Vec3& g) {
    m_gravity = g;
    bx_body_edit e = WorldEdit(BX_EDIT_WORLD_GRAVITY, false);
    Copy3(e.v, g);
    QueueEdit(&e, false);
}

void b3World::SetAutoClearForces(bool flag) {
    m_autoClearForces = flag;
    bx_body_edit e = WorldEdit(BX_EDIT_WORLD_AUTO_CLEAR_FORCES, flag);
    QueueEdit(&e, false);
}

void b3World::ClearForces() {
    bx_body_edit e = WorldEdit(BX_EDIT_WORLD_CLEAR_FORCES, false);
    QueueEdit(&e, false);
}

const char* b3World::GetLastError() const { return m_device->error.c_str(); }

b3Body* b3World::CreateBody(const b3BodyDef& def) {
    bx_body_desc d;
    std::memset(&d, 0, sizeof(d));
    d.type = uint32(def.type);
    d.flags = (def.awake ? BX_BODY_AWAKE : 0u) | (def.allowSleep ? BX_BODY_AUTO_SLEEP : 0u) |
              (def.fixedRotationX ? BX_BODY_FIXED_ROT_X : 0u) | (def.fixedRotationY ? BX_BODY_FIXED_ROT_Y : 0u) |
              (def.fixedRotationZ ? BX_BODY_FIXED_ROT_Z : 0u);
    Copy3(d.position, def.position);
    Copy4(d.orientation, def.orientation);
    Copy3(d.linear_velocity, def.linearVelocity);
    Copy3(d.angular_velocity, def.angularVelocity);
    Copy3(d.linear_damping, def.linearDamping);
    Copy3(d.angular_damping, def.angularDamping);
    Copy3(d.gravity_scale, def.gravityScale);
    d.linear_sleep_tolerance = def.linearSleepTolerance;
    d.angular_sleep_tolerance = def.angularSleepTolerance;
    m_device->newBodies.push_back(d);
    QueueTopology(BX_TOPO_CREATE_BODY, uint32(m_device->newBodies.size() - 1));

    b3Body* b = new b3Body(def, this, uint32(m_bodies.size()));
    b->m_next = m_bodyList;  // world.cpp:94-110 prepends
    m_bodyList = b;
    m_bodies.push_back(b);
    return b;
}

void b3World::RemoveJoint(b3Joint* j) {
    b3Joint** link = &m_jointList;
    while (*link && *link != j) link = &(*link)->m_next;
    if (*link) *link = j->m_next;
    b3JointEdge* edges[2] = {&j->m_edgeA, &j->m_edgeB};
    b3Body* owners[2] = {j->m_def.bodyA, j->m_def.bodyB};
    for (int k = 0; k < 2; ++k) {  // joint_manager.cpp:75-141 unlinks both edges
        b3JointEdge* e = edges[k];
        if (e->prev) e->prev->next = e->next;
        if (e->next) e->next->prev = e->prev;
        if (owners[k]->m_jointList == e) owners[k]->m_jointList = e->next;
    }
    m_joints.erase(std::find(m_joints.begin(), m_joints.end(), j));
    delete j;
}

void b3World::DestroyBody(b3Body* body) {
    if (!body || body->m_world != this) return;
    QueueTopology(BX_TOPO_DESTROY_BODY, body->m_index);
    // the device destroys the body's joints and fixtures with it (world.cpp:112-175)
    while (body->m_jointList) RemoveJoint(body->m_jointList->joint);
    for (b3Fixture* f = body->m_fixtureList; f; f = f->m_next)
        m_fixtures.erase(std::find(m_fixtures.begin(), m_fixtures.end(), f));
    b3Body** link = &m_bodyList;
    while (*link && *link != body) link = &(*link)->m_next;
    if (*link) *link = body->m_next;
    m_bodies.erase(std::find(m_bodies.begin(), m_bodies.end(), body));
    delete body;
    Renumber();
}

b3Joint* b3World::CreateJoint(const b3JointDef& def) {
    if (!def.bodyA || !def.bodyB || def.bodyA->m_world != this || def.bodyB->m_world != this) return nullptr;
    bx_joint_desc d;
    std::memset(&d, 0, sizeof(d));
    d.body_a = def.bodyA->m_index;
    d.body_b = def.bodyB->m_index;
    d.collide_connected = def.collideConnected ? 1u : 0u;
    switch (def.type) {
        case e_revoluteJoint: {
            const b3RevoluteJointDef& r = static_cast<const b3RevoluteJointDef&>(def);
            d.type = BX_JOINT_REVOLUTE;
            d.use_initialize = r.m_initialize ? 1u : 0u;
            Copy3(d.axis, r.m_axis);
            Copy3(d.anchor, r.m_anchor);
            Copy3(d.local_anchor_a, r.localAnchorA);
            Copy4(d.local_rotation_a, r.localRotationA);
            Copy3(d.local_anchor_b, r.localAnchorB);
            Copy4(d.local_rotation_b, r.localRotationB);
            Copy4(d.reference_rotation, r.referenceRotation);
            d.enable_limit = r.enableLimit ? 1u : 0u;
            d.lower_angle = r.lowerAngle;
            d.upper_angle = r.upperAngle;
            d.enable_motor = r.enableMotor ? 1u : 0u;
            d.motor_speed = r.motorSpeed;
            d.max_motor_torque = r.maxMotorTorque;
            break;
        }
        case e_sphereJoint: {
            const b3SphereJointDef& r = static_cast<const b3SphereJointDef&>(def);
            d.type = BX_JOINT_SPHERE;
            d.use_initialize = r.m_initialize ? 1u : 0u;
            Copy3(d.anchor, r.m_anchor);
            Copy3(d.local_anchor_a, r.localAnchorA);
            Copy3(d.local_anchor_b, r.localAnchorB);
            break;
        }
        case e_weldJoint: {
            const b3WeldJointDef& r = static_cast<const b3WeldJointDef&>(def);
            d.type = BX_JOINT_WELD;
            d.use_initialize = r.m_initialize ? 1u : 0u;
            Copy3(d.anchor, r.m_anchor);
            Copy3(d.local_anchor_a, r.localAnchorA);
            Copy3(d.local_anchor_b, r.localAnchorB);
            Copy4(d.reference_rotation, r.referenceRotation);
            d.frequency_hz = r.frequencyHz;
            d.damping_ratio = r.dampingRatio;
            d.correction_factor = r.correctionFactor;
            break;
        }
        case e_springJoint: {
            const b3SpringJointDef& r = static_cast<const b3SpringJointDef&>(def);
            d.type = BX_JOINT_SPRING;
            d.use_initialize = r.m_initialize ? 1u : 0u;
            Copy3(d.anchor, r.m_anchorA);
            Copy3(d.anchor_b, r.m_anchorB);
            Copy3(d.local_anchor_a, r.localAnchorA);
            Copy3(d.local_anchor_b, r.localAnchorB);
            d.length = r.length;
            d.frequency_hz = r.frequencyHz;
            d.damping_ratio = r.dampingRatio;
            break;
        }
        case e_mouseJoint: {
            const b3MouseJointDef& r = static_cast<const b3MouseJointDef&>(def);
            d.type = BX_JOINT_MOUSE;
            Copy3(d.target, r.target);
            d.max_force = r.maxForce;
            d.frequency_hz = r.frequencyHz;
            d.damping_ratio = r.dampingRatio;
            break;
        }
        case e_frictionJoint: {
            const b3FrictionJointDef& r = static_cast<const b3FrictionJointDef&>(def);
            d.type = BX_JOINT_FRICTION;
            d.use_initialize = r.m_initialize ? 1u : 0u;
            Copy3(d.anchor, r.m_anchor);
            Copy3(d.local_anchor_a, r.localAnchorA);
            Copy3(d.local_anchor_b, r.localAnchorB);
            d.max_force = r.maxForce;
            d.max_torque = r.maxTorque;
            break;
        }
        case e_motorJoint: {
            const b3MotorJointDef& r = static_cast<const b3MotorJointDef&>(def);
            d.type = BX_JOINT_MOTOR;
            d.use_initialize = r.m_initialize ? 1u : 0u;
            Copy3(d.linear_offset, r.linearOffset);
            Copy4(d.angular_offset, r.angularOffset);
            d.max_force = r.maxForce;
            d.max_torque = r.maxTorque;
            d.correction_factor = r.correctionFactor;
            break;
        }
        case e_coneJoint: {
            const b3ConeJointDef& r = static_cast<const b3ConeJointDef&>(def);
            d.type = BX_JOINT_CONE;
            d.use_initialize = r.m_initialize ? 1u : 0u;
            Copy3(d.axis, r.m_axis);
            Copy3(d.anchor, r.m_anchor);
            Copy3(d.local_anchor_a, r.localAnchorA);
            Copy4(d.local_rotation_a, r.localRotationA);
            Copy3(d.local_anchor_b, r.localAnchorB);
            Copy4(d.local_rotation_b, r.localRotationB);
            Copy4(d.reference_rotation, r.referenceRotation);
            d.enable_limit = r.enableConeLimit ? 1u : 0u;
            d.cone_angle = r.coneAngle;
            break;
        }
        case e_prismaticJoint: {
            const b3PrismaticJointDef& r = static_cast<const b3PrismaticJointDef&>(def);
            d.type = BX_JOINT_PRISMATIC;
            d.use_initialize = r.m_initialize ? 1u : 0u;
            Copy3(d.anchor, r.m_anchor);
            Copy3(d.axis, r.m_axis);
            Copy3(d.local_anchor_a, r.localAnchorA);
            Copy3(d.local_anchor_b, r.localAnchorB);
            Copy3(d.local_axis_a, r.localAxisA);
            Copy4(d.reference_rotation, r.referenceRotation);
            d.enable_limit = r.enableLimit ? 1u : 0u;
            d.lower_translation = r.lowerTranslation;
            d.upper_translation = r.upperTranslation;
            d.enable_motor = r.enableMotor ? 1u : 0u;
            d.max_force = r.maxMotorForce;
            d.motor_speed = r.motorSpeed;
            d.correction_factor = r.correctionFactor;
            break;
        }
        case e_wheelJoint: {
            const b3WheelJointDef& r = static_cast<const b3WheelJointDef&>(def);
            d.type = BX_JOINT_WHEEL;
            d.use_initialize = r.m_initialize ? 1u : 0u;
            Copy3(d.anchor, r.m_anchor);
            Copy3(d.axis, r.m_axisA);
            Copy3(d.axis_b, r.m_axisB);
            Copy3(d.local_anchor_a, r.localAnchorA);
            Copy3(d.local_anchor_b, r.localAnchorB);
            Copy3(d.local_axis_a, r.localAxisA);
            Copy3(d.local_axis_b, r.localAxisB);
            d.enable_motor = r.enableMotor ? 1u : 0u;
            d.max_motor_torque = r.maxMotorTorque;
            d.motor_speed = r.motorSpeed;
            d.frequency_hz = r.frequencyHz;
            d.damping_ratio = r.dampingRatio;
            break;
        }
        default:
            return nullptr;
    }
    m_device->newJoints.push_back(d);
    QueueTopology(BX_TOPO_CREATE_JOINT, uint32(m_device->newJoints.size() - 1));

    b3Joint* j = new b3Joint();
    j->m_def = def;
    j->m_userData = def.userData;
    j->m_index = uint32(m_joints.size());
    j->m_next = m_jointList;  // joint_manager.cpp:30-73: world list and both bodies' lists prepend
    m_jointList = j;
    b3Body* bodies[2] = {def.bodyA, def.bodyB};
    b3JointEdge* edges[2] = {&j->m_edgeA, &j->m_edgeB};
    for (int k = 0; k < 2; ++k) {
        b3JointEdge* e = edges[k];
        e->joint = j;
        e->other = bodies[1 - k];
        e->prev = nullptr;
        e->next = bodies[k]->m_jointList;
        if (bodies[k]->m_jointList) bodies[k]->m_jointList->prev = e;
        bodies[k]->m_jointList = e;
    }
    m_joints.push_back(j);
    return j;
}

void b3World::DestroyJoint(b3Joint* joint) {
    if (!joint) return;
    QueueTopology(BX_TOPO_DESTROY_JOINT, joint->m_index);
    RemoveJoint(joint);
    Renumber();
}

void b3World::Step(scalar dt, uint32 velocityIterations, uint32 positionIterations) {
    b3WorldDevice& D = *m_device;
    D.error.clear();
    int rc = Flush();
    if (rc == BX_OK && D.batch) {
        bx_step_cfg cfg;
        cfg.dt = dt;
        cfg.velocity_iterations = velocityIterations;
        cfg.position_iterations = positionIterations;
        rc = bx_batch_step(D.batch, &cfg, 1);
        if (rc == BX_OK) rc = bx_batch_sync(D.batch);
        D.stateStale = D.massStale = D.contactsStale = true;
    }
    if (rc < 0) D.error = bx_last_error();
}

b3Contact* b3World::GetContactList() {
    b3WorldDevice& D = *m_device;
    if (D.contactsStale) {
        int rc = Flush();
        if (rc < 0) {
            D.error = bx_last_error();
        } else if (D.batch) {
            bx_world_counts counts;
            rc = bx_batch_get_counts(D.batch, 0, &counts);
            if (rc == BX_OK) {
                D.contactStates.resize(counts.contact_count);
                rc = bx_batch_get_contacts(D.batch, 0, D.contactStates.data(), counts.contact_count);
            }
            if (rc < 0) {
                D.error = bx_last_error();
            } else {
                // keep the objects of pairs that are still in contact
                std::unordered_map<uint64, b3Contact*> old;
                for (b3Contact* c : D.contacts)
                    old[(uint64(c->m_fixtureA->m_index) << 32) | c->m_fixtureB->m_index] = c;
                std::vector<b3Contact*> list;
                for (const bx_contact_state& s : D.contactStates) {
                    b3Fixture* fA = m_fixtures[s.fixture_a];
                    b3Fixture* fB = m_fixtures[s.fixture_b];
                    const uint64 key = (uint64(s.fixture_a) << 32) | s.fixture_b;
                    auto it = old.find(key);
                    b3Contact* c = nullptr;
                    if (it != old.end() && it->second->m_fixtureA == fA && it->second->m_fixtureB == fB) {
                        c = it->second;
                        old.erase(it);
                    } else {
                        c = new b3Contact();
                    }
                    c->m_fixtureA = fA;
                    c->m_fixtureB = fB;
                    c->m_touching = s.touching != 0;
                    c->m_manifoldCount = std::min<uint32>(s.manifold_count, BX_MAX_MANIFOLDS);
                    for (uint32 m = 0; m < c->m_manifoldCount; ++m) {
                        const bx_manifold_state& ms = s.manifolds[m];
                        b3Manifold& out = c->m_manifolds[m];
                        std::memset(static_cast<void*>(&out), 0, sizeof(out));
                        out.pointCount = ms.point_count;
                        out.tangentImpulse.Set(ms.tangent_impulse[0], ms.tangent_impulse[1]);
                        out.motorImpulse = ms.motor_impulse;
                        for (uint32 p = 0; p < ms.point_count && p < B3_MAX_MANIFOLD_POINTS; ++p) {
                            const bx_manifold_point_state& ps = ms.points[p];
                            b3ManifoldPoint& op = out.points[p];
                            op.localNormal1 = V3(ps.local_normal1);
                            op.localPoint1 = V3(ps.local_point1);
                            op.localPoint2 = V3(ps.local_point2);
                            op.normalImpulse = ps.normal_impulse;
                            op.id.triangleKey = ps.triangle_key;
                            op.id.key.pair.inEdge1 = ps.key[0];
                            op.id.key.pair.outEdge1 = ps.key[1];
                            op.id.key.pair.inEdge2 = ps.key[2];
                            op.id.key.pair.outEdge2 = ps.key[3];
                            op.persisting = ps.persisting != 0;
                            op.edgeContact = ps.edge_contact != 0;
                        }
                    }
                    list.push_back(c);
                }
                for (auto& kv : old) delete kv.second;
                for (size_t i = 0; i < list.size(); ++i) list[i]->m_next = i + 1 < list.size() ? list[i + 1] : nullptr;
                D.contacts.swap(list);
                D.contactsStale = false;
            }
        } else {
            D.ClearContacts();
            D.contactsStale = false;
        }
    }
    return D.contacts.empty() ? nullptr : D.contacts[0];
}

uint32 b3World::GetContactCount() const {
    const_cast<b3World*>(this)->GetContactList();
    return uint32(m_device->contacts.size());
}

// ------------------------------------------------------------------ world queries
void b3World::RayCast(b3RayCastListener* listener, b3RayCastFilter* filter, const b3Vec3& p1,
                      const b3Vec3& p2) const {
    b3World* self = const_cast<b3World*>(this);
    b3WorldDevice& D = *m_device;
    int rc = self->Flush();
    const uint32 nf = uint32(m_fixtures.size());
    if (rc == BX_OK && D.batch && nf) {
        bx_ray_query q;
        q.world = 0;
        q.mask = BX_NULL_INDEX;
        Copy3(q.p1, p1);
        Copy3(q.p2, p2);
        std::vector<uint8_t> mask;
        if (filter) {
            mask.resize(nf);
            for (uint32 i = 0; i < nf; ++i) mask[i] = filter->ShouldRayCast(m_fixtures[i]) ? 1 : 0;
            q.mask = 0;
        }
        std::vector<bx_ray_hit> hits(nf);
        uint32 count = 0;
        rc = bx_batch_raycast(D.batch, &q, 1, mask.empty() ? nullptr : mask.data(), mask.size(), hits.data(), nf,
                              &count);
        if (rc == BX_OK) {
            // a listener's return value clips the ray (positive), ends the cast (zero) or ignores the hit
            // (negative), as b3DynamicTree::RayCast treats it (dynamic_tree.h:312-329)
            scalar maxFraction = scalar(1);
            for (uint32 i = 0; i < std::min(count, nf); ++i) {
                const bx_ray_hit& h = hits[i];
                if (h.fraction > maxFraction) continue;
                scalar r = listener->ReportFixture(m_fixtures[h.fixture], V3(h.point), V3(h.normal), h.fraction);
                if (r == scalar(0)) break;
                if (r > scalar(0)) maxFraction = r;
            }
        }
    }
    if (rc < 0) D.error = bx_last_error();
}

bool b3World::RayCastSingle(b3RayCastSingleOutput* output, b3RayCastFilter* filter, const b3Vec3& p1,
                            const b3Vec3& p2) const {
    b3World* self = const_cast<b3World*>(this);
    b3WorldDevice& D = *m_device;
    int rc = self->Flush();
    const uint32 nf = uint32(m_fixtures.size());
    bool found = false;
    if (rc == BX_OK && D.batch && nf) {
        bx_ray_query q;
        q.world = 0;
        q.mask = BX_NULL_INDEX;
        Copy3(q.p1, p1);
        Copy3(q.p2, p2);
        std::vector<uint8_t> mask;
        if (filter) {
            mask.resize(nf);
            for (uint32 i = 0; i < nf; ++i) mask[i] = filter->ShouldRayCast(m_fixtures[i]) ? 1 : 0;
            q.mask = 0;
        }
        bx_ray_hit h;
        rc = bx_batch_raycast_single(D.batch, &q, 1, mask.empty() ? nullptr : mask.data(), mask.size(), &h);
        if (rc == BX_OK && h.fixture != BX_NULL_INDEX) {
            output->fixture = m_fixtures[h.fixture];
            output->point = V3(h.point);
            output->normal = V3(h.normal);
            output->fraction = h.fraction;
            found = true;
        }
    }
    if (rc < 0) D.error = bx_last_error();
    return found;
}

void b3World::QueryAABB(b3QueryListener* listener, b3QueryFilter* filter, const b3AABB& aabb) const {
    b3World* self = const_cast<b3World*>(this);
    b3WorldDevice& D = *m_device;
    int rc = self->Flush();
    const uint32 nf = uint32(m_fixtures.size());
    if (rc == BX_OK && D.batch && nf) {
        bx_aabb_query q;
        q.world = 0;
        q.mask = BX_NULL_INDEX;
        Copy3(q.lower, aabb.lowerBound);
        Copy3(q.upper, aabb.upperBound);
        std::vector<uint8_t> mask;
        if (filter) {
            mask.resize(nf);
            for (uint32 i = 0; i < nf; ++i) mask[i] = filter->ShouldReport(m_fixtures[i]) ? 1 : 0;
            q.mask = 0;
        }
        std::vector<uint32_t> found(nf);
        uint32 count = 0;
        rc = bx_batch_query_aabb(D.batch, &q, 1, mask.empty() ? nullptr : mask.data(), mask.size(), found.data(), nf,
                                 &count);
        if (rc == BX_OK)
            for (uint32 i = 0; i < std::min(count, nf); ++i)
                if (!listener->ReportFixture(m_fixtures[found[i]])) break;
    }
    if (rc < 0) D.error = bx_last_error();
}

// ------------------------------------------------------------------ shape casts
bool b3World::CastShapeOf(const b3Shape* shape, void* out) {
    bx_cast_shape& c = *static_cast<bx_cast_shape*>(out);
    std::memset(&c, 0, sizeof(c));
    if (!shape) return false;
    c.shape_type = uint32(shape->GetType());
    c.radius = shape->m_radius;
    switch (shape->GetType()) {
    case b3Shape::e_sphere: Copy3(c.vertex1, static_cast<const b3SphereShape*>(shape)->m_center); return true;
    case b3Shape::e_capsule:
        Copy3(c.vertex1, static_cast<const b3CapsuleShape*>(shape)->m_vertex1);
        Copy3(c.vertex2, static_cast<const b3CapsuleShape*>(shape)->m_vertex2);
        return true;
    case b3Shape::e_triangle: {
        const b3TriangleShape* t = static_cast<const b3TriangleShape*>(shape);
        Copy3(c.vertex1, t->m_vertex1);
        Copy3(c.vertex2, t->m_vertex2);
        Copy3(c.vertex3, t->m_vertex3);
        return true;
    }
    case b3Shape::e_hull:
        if (!static_cast<const b3HullShape*>(shape)->m_hull) return false;
        c.shape_index = HullIndex(static_cast<const b3HullShape*>(shape)->m_hull);
        return true;
    default: return false;  // the reference casts convex shapes only (world.cpp:755-760)
    }
}

int b3World::FlushLibrary() {
    b3WorldDevice& D = *m_device;
    int rc = Flush();
    if (rc != BX_OK || !D.batch || (D.hulls.size() <= D.hullsOnDevice && D.meshes.size() <= D.meshesOnDevice))
        return rc;
    bx_world_desc none;
    std::memset(&none, 0, sizeof(none));
    rc = bx_batch_edit_topology(D.batch, 0, &D.Library(), &none, nullptr, 0);
    if (rc == BX_OK) {
        D.hullsOnDevice = uint32(D.hulls.size());
        D.meshesOnDevice = uint32(D.meshes.size());
    }
    return rc;
}

void b3World::ShapeCast(b3ShapeCastListener* listener, b3ShapeCastFilter* filter, const b3Shape* shape,
                        const b3Transform& xf, const b3Vec3& displacement) const {
    b3World* self = const_cast<b3World*>(this);
    b3WorldDevice& D = *m_device;
    bx_shape_cast_query q;
    std::memset(&q, 0, sizeof(q));
    if (!self->CastShapeOf(shape, &q.shape)) return;
    int rc = self->FlushLibrary();
    const uint32 nf = uint32(m_fixtures.size());
    if (rc == BX_OK && D.batch && nf) {
        q.world = 0;
        q.mask = BX_NULL_INDEX;
        Copy3(q.position, xf.translation);
        Copy4(q.orientation, xf.rotation);
        Copy3(q.displacement, displacement);
        std::vector<uint8_t> mask;
        if (filter) {
            mask.resize(nf);
            for (uint32 i = 0; i < nf; ++i) mask[i] = filter->ShouldShapeCast(m_fixtures[i]) ? 1 : 0;
            q.mask = 0;
        }
        // a mesh fixture is hit once per triangle: a second call when the first capacity was short
        uint32 cap = 4 * nf + 16, count = 0;
        std::vector<bx_shape_hit> hits(cap);
        rc = bx_batch_shape_cast(D.batch, &q, 1, mask.empty() ? nullptr : mask.data(), mask.size(), hits.data(), cap,
                                 &count);
        if (rc == BX_OK && count > cap) {
            cap = count;
            hits.resize(cap);
            rc = bx_batch_shape_cast(D.batch, &q, 1, mask.empty() ? nullptr : mask.data(), mask.size(), hits.data(),
                                     cap, &count);
        }
        if (rc == BX_OK) {
            scalar maxFraction = scalar(1);  // as the world's cast wrapper (world.cpp:660-695)
            for (uint32 i = 0; i < std::min(count, cap); ++i) {
                const bx_shape_hit& h = hits[i];
                if (h.fraction > maxFraction) continue;
                maxFraction = listener->ReportFixture(m_fixtures[h.fixture], V3(h.point), V3(h.normal), h.fraction);
                if (maxFraction == scalar(0)) break;
            }
        }
    }
    if (rc < 0) D.error = bx_last_error();
}

bool b3World::ShapeCastSingle(b3ShapeCastSingleOutput* output, b3ShapeCastFilter* filter, const b3Shape* shape,
                              const b3Transform& xf, const b3Vec3& displacement) const {
    b3World* self = const_cast<b3World*>(this);
    b3WorldDevice& D = *m_device;
    bx_shape_cast_query q;
    std::memset(&q, 0, sizeof(q));
    if (!self->CastShapeOf(shape, &q.shape)) return false;
    int rc = self->FlushLibrary();
    const uint32 nf = uint32(m_fixtures.size());
    bool found = false;
    if (rc == BX_OK && D.batch && nf) {
        q.world = 0;
        q.mask = BX_NULL_INDEX;
        Copy3(q.position, xf.translation);
        Copy4(q.orientation, xf.rotation);
        Copy3(q.displacement, displacement);
        std::vector<uint8_t> mask;
        if (filter) {
            mask.resize(nf);
            for (uint32 i = 0; i < nf; ++i) mask[i] = filter->ShouldShapeCast(m_fixtures[i]) ? 1 : 0;
            q.mask = 0;
        }
        bx_shape_hit h;
        rc = bx_batch_shape_cast_single(D.batch, &q, 1, mask.empty() ? nullptr : mask.data(), mask.size(), &h);
        if (rc == BX_OK && h.fixture != BX_NULL_INDEX) {
            output->fixture = m_fixtures[h.fixture];
            output->point = V3(h.point);
            output->normal = V3(h.normal);
            output->fraction = h.fraction;
            found = true;
        }
    }
    if (rc < 0) D.error = bx_last_error();
    return found;
}
