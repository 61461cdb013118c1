// bx_api.hip -- host implementation of the C ABI declared in include/bxstep.h.
//
// World construction repeats, on the host and in binary32, what the reference does when a
// program builds a world: b3World::CreateBody (world.cpp:94-110, body.cpp:25-103),
// b3Body::CreateFixture (body.cpp:105-136) with b3Body::ResetMass (body.cpp:265-376) and
// b3BroadPhase::CreateProxy (broad_phase.cpp:78-84, dynamic_tree.cpp:108-128). Proxy ids are the
// node ids the reference's dynamic tree would hand out: its node pool is a LIFO free list over
// sequentially numbered nodes, a proxy takes one node and every insertion into a non-empty tree
// takes one more for the new parent; moving a proxy frees and retakes the same parent, so ids
// are fixed at creation (dynamic_tree.cpp:44-106).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/bxstep.h"
#include "bx_shapes.h"
#include "bx_joint.h"

namespace bx {
struct StepParams { scalar dt, inv_dt; u32 velocityIterations, positionIterations; };
void LaunchStep(const BatchView& view, u32 firstWorld, u32 worldCount, const StepParams& sp, bool findFirst,
                u32 maxIslands, hipStream_t stream, hipEvent_t* ev);
}

using namespace bx;

static thread_local std::string g_error;

static int Fail(int code, const std::string& msg) {
    g_error = msg;
    return code;
}

#define HIP_CHECK(expr)                                                                                  \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) return Fail(BX_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

// Steps of one bx_batch_step call whose kernels are timed with HIP events (12 events per step,
// see LaunchStep); later steps of a longer call run untimed.
static const u32 kMaxTimedSteps = 1024;

struct bx_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev;  // 12 per timed step of the last bx_batch_step call
    u32 timedSteps = 0;
};

static int EnsureEvents(bx_context* ctx, u32 steps) {
    size_t need = size_t(12) * steps;
    while (ctx->ev.size() < need) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        ctx->ev.push_back(e);
    }
    return BX_OK;
}

template <class T>
struct DevArray {
    T* ptr = nullptr;
    size_t count = 0;
};

struct bx_batch {
    bx_context* ctx = nullptr;
    u32 worldCount = 0;
    BatchView view{};
    std::vector<void*> allocations;
    std::vector<WorldInfo> worlds;
    std::vector<std::vector<u32>> proxyIds;  // per world: proxy id of each fixture
    std::vector<u32> hullRadius;
    bool newContacts = true;
    float timings[5] = {};
    size_t deviceBytes = 0;
};

template <class T>
static int DevAlloc(bx_batch* b, T** out, size_t count, const T* init = nullptr) {
    size_t bytes = std::max<size_t>(1, count) * sizeof(T);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return Fail(BX_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    b->allocations.push_back(p);
    b->deviceBytes += bytes;
    if (init && count) {
        e = hipMemcpy(p, init, count * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return Fail(BX_E_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
    } else {
        e = hipMemset(p, 0, bytes);
        if (e != hipSuccess) return Fail(BX_E_DEVICE, std::string("hipMemset: ") + hipGetErrorString(e));
    }
    *out = (T*)p;
    return BX_OK;
}

static Vec3 V3(const float* p) { return Vec3(p[0], p[1], p[2]); }
static Quat Q4(const float* p) { return Quat(Vec3(p[0], p[1], p[2]), p[3]); }

// Host copy of the shape library.
struct HostLibrary {
    std::vector<HullInfo> info;
    std::vector<Vec3> vertices;
    std::vector<HalfEdge> edges;
    std::vector<u32> faces;
    std::vector<Plane> planes;
    HullView View(u32 i) const {
        const HullInfo& hi = info[i];
        HullView h;
        h.vertices = vertices.data() + hi.vBase;
        h.edges = edges.data() + hi.eBase;
        h.faces = faces.data() + hi.fBase;
        h.planes = planes.data() + hi.fBase;
        h.centroid = Vec3(hi.cx, hi.cy, hi.cz);
        h.vertexCount = hi.vCount;
        h.edgeCount = hi.eCount;
        h.faceCount = hi.fCount;
        return h;
    }
};

static int BuildLibrary(const bx_shape_library* lib, HostLibrary& L) {
    if (!lib) return BX_OK;
    for (u32 i = 0; i < lib->hull_count; ++i) {
        const bx_hull_desc& d = lib->hulls[i];
        if (d.vertex_count < 4 || d.edge_count < 6 || d.face_count < 4 || (d.edge_count & 1u))
            return Fail(BX_E_INVALID, "hull " + std::to_string(i) + ": invalid counts");
        HullInfo hi;
        hi.vBase = u32(L.vertices.size());
        hi.vCount = d.vertex_count;
        hi.eBase = u32(L.edges.size());
        hi.eCount = d.edge_count;
        hi.fBase = u32(L.faces.size());
        hi.fCount = d.face_count;
        hi.cx = d.centroid[0];
        hi.cy = d.centroid[1];
        hi.cz = d.centroid[2];
        for (u32 k = 0; k < d.vertex_count; ++k) L.vertices.push_back(V3(d.vertices + 3 * k));
        for (u32 k = 0; k < d.edge_count; ++k) {
            const uint32_t* e = d.edges + 5 * k;
            for (int c = 0; c < 5; ++c)
                if ((c == 0 && e[c] >= d.vertex_count) || (c == 2 && e[c] >= d.face_count) ||
                    ((c == 1 || c >= 3) && e[c] >= d.edge_count))
                    return Fail(BX_E_INVALID, "hull " + std::to_string(i) + ": edge index out of range");
            L.edges.push_back(HalfEdge{e[0], e[1], e[2], e[3], e[4]});
        }
        for (u32 k = 0; k < d.face_count; ++k) {
            if (d.faces[k] >= d.edge_count) return Fail(BX_E_INVALID, "hull face edge out of range");
            L.faces.push_back(d.faces[k]);
            L.planes.push_back(Plane(V3(d.planes + 4 * k), d.planes[4 * k + 3]));
        }
        // Every loop over a face walks `next` until it returns to the face's first edge, and
        // clipping stores at most 32 vertices: reject hulls whose face cycles do not close
        // within the edge count, whose twins are not adjacent, or whose faces are too large.
        for (u32 k = 0; k < d.face_count; ++k) {
            u32 begin = d.faces[k], e = begin, steps = 0;
            do {
                if (d.edges[5 * e + 2] != k) return Fail(BX_E_INVALID, "hull face cycle leaves its face");
                e = d.edges[5 * e + 4];
                if (++steps > d.edge_count || steps > 16) return Fail(BX_E_INVALID, "hull face cycle does not close");
            } while (e != begin);
        }
        for (u32 k = 0; k < d.edge_count; ++k)
            if (d.edges[5 * k + 1] != (k ^ 1u)) return Fail(BX_E_INVALID, "hull twins must be adjacent half-edges");
        L.info.push_back(hi);
    }
    if (lib->mesh_count > 0) return Fail(BX_E_UNSUPPORTED, "meshes are not supported by the device path yet");
    return BX_OK;
}

// b3Body::ResetMass (body.cpp:265-376)
static void ResetMass(Body& b, const std::vector<Fixture>& fixtures, const std::vector<u32>& list, u32 listBase,
                      const HostLibrary& L) {
    b.mass = 0;
    b.invMass = 0;
    b.I.SetZero();
    b.invI.SetZero();
    b.worldInvI.SetZero();
    b.localCenter.SetZero();
    if (b.type == kStatic || b.type == kKinematic) {
        b.x0 = b.xfT;
        b.x = b.xfT;
        b.q0 = b.q;
        return;
    }
    Vec3 localCenter;
    localCenter.SetZero();
    for (u32 k = 0; k < b.fixtureCount; ++k) {
        const Fixture& f = fixtures[list[listBase + b.fixtureBegin + k]];
        if (f.density == scalar(0)) continue;
        MassData md;
        if (f.shapeType == BX_SHAPE_HULL) HullMass(L.View(f.shapeIndex), f.density, &md);
        else if (f.shapeType == BX_SHAPE_SPHERE) SphereMass(f.v1, f.radius, f.density, &md);
        else if (f.shapeType == BX_SHAPE_CAPSULE) CapsuleMass(f.v1, f.v2, f.radius, f.density, &md);
        else { md.mass = 0; md.center.SetZero(); md.I.SetZero(); }
        localCenter += md.mass * md.center;
        b.mass += md.mass;
        b.I += md.I;
    }
    if (b.mass > scalar(0)) {
        b.invMass = scalar(1) / b.mass;
        localCenter *= b.invMass;
        b.I = b.I - b.mass * Steiner(localCenter);
        b.invI = Inverse(b.I);
        b.worldInvI = RotateToFrame(b.invI, b.q);
        if (b.flags & kFixedX) {
            b.invI.y.y = 0; b.invI.z.y = 0; b.invI.y.z = 0; b.invI.z.z = 0;
            b.worldInvI.y.y = 0; b.worldInvI.z.y = 0; b.worldInvI.y.z = 0; b.worldInvI.z.z = 0;
        }
        if (b.flags & kFixedY) {
            b.invI.x.x = 0; b.invI.x.z = 0; b.invI.z.x = 0; b.invI.z.z = 0;
            b.worldInvI.x.x = 0; b.worldInvI.x.z = 0; b.worldInvI.z.x = 0; b.worldInvI.z.z = 0;
        }
        if (b.flags & kFixedZ) {
            b.invI.x.x = 0; b.invI.x.y = 0; b.invI.y.x = 0; b.invI.y.y = 0;
            b.worldInvI.x.x = 0; b.worldInvI.x.y = 0; b.worldInvI.y.x = 0; b.worldInvI.y.y = 0;
        }
    } else {
        b.mass = 1;
        b.invMass = 1;
    }
    Vec3 oldCenter = b.x;
    b.localCenter = localCenter;
    b.x = Mul(Transform(b.xfT, b.q), b.localCenter);
    b.x0 = b.x;
    b.v += Cross(b.w, b.x - oldCenter);
}

static void HostAABB(const Fixture& f, const Transform& xf, const HostLibrary& L, AABB& out) {
    if (f.shapeType == BX_SHAPE_HULL) {
        HullView h = L.View(f.shapeIndex);
        out.lowerBound = out.upperBound = Mul(xf, h.vertices[0]);
        for (u32 i = 1; i < h.vertexCount; ++i) {
            Vec3 v = Mul(xf, h.vertices[i]);
            out.lowerBound = Min(out.lowerBound, v);
            out.upperBound = Max(out.upperBound, v);
        }
        out.Extend(f.radius);
    } else if (f.shapeType == BX_SHAPE_SPHERE) {
        Vec3 c = Mul(xf, f.v1);
        Vec3 r(f.radius, f.radius, f.radius);
        out.lowerBound = c - r;
        out.upperBound = c + r;
    } else if (f.shapeType == BX_SHAPE_CAPSULE) {
        Vec3 c1 = Mul(xf, f.v1), c2 = Mul(xf, f.v2);
        Vec3 r(f.radius, f.radius, f.radius);
        out.lowerBound = Min(c1, c2) - r;
        out.upperBound = Max(c1, c2) + r;
    } else {
        out.lowerBound.SetZero();
        out.upperBound.SetZero();
    }
}

template <class T>
static int Download(bx_batch* b, std::vector<T>& out, const T* src, size_t count) {
    out.resize(count);
    if (count == 0) return BX_OK;
    HIP_CHECK(hipMemcpyAsync(out.data(), src, count * sizeof(T), hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    return BX_OK;
}

static int GetState(bx_batch* b, u32 world, WorldState& ws) {
    if (!b || world >= b->worldCount) return Fail(BX_E_INVALID, "bad batch or world");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    HIP_CHECK(hipMemcpyAsync(&ws, b->view.wstate + world, sizeof(ws), hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    return BX_OK;
}

static int StepRange(bx_batch* b, const bx_step_cfg* cfg, u32 first, u32 count, u32 steps, bool timed) {
    StepParams sp;
    sp.dt = cfg->dt;
    sp.inv_dt = cfg->dt > scalar(0) ? scalar(1) / cfg->dt : scalar(0);
    sp.velocityIterations = cfg->velocity_iterations;
    sp.positionIterations = cfg->position_iterations;
    u32 maxIslands = 0;
    for (u32 i = 0; i < count; ++i) maxIslands += b->worlds[first + i].bodyCount;
    u32 timedSteps = timed ? std::min(steps, kMaxTimedSteps) : 0u;
    if (timedSteps) {
        int rc = EnsureEvents(b->ctx, timedSteps);
        if (rc) return rc;
        b->ctx->timedSteps = timedSteps;
    }
    for (u32 s = 0; s < steps; ++s) {
        hipEvent_t* ev = s < timedSteps ? b->ctx->ev.data() + size_t(12) * s : nullptr;
        LaunchStep(b->view, first, count, sp, b->newContacts, maxIslands, b->ctx->stream, ev);
        b->newContacts = false;
    }
    HIP_CHECK(hipGetLastError());
    return BX_OK;
}

extern "C" {

const char* bx_version(void) { return "bxstep 0.1 gfx950"; }
const char* bx_last_error(void) { return g_error.c_str(); }

int bx_context_create(int device, bx_context** out) {
    if (!out) return Fail(BX_E_INVALID, "null out");
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return Fail(BX_E_INVALID, "no such HIP device");
    HIP_CHECK(hipSetDevice(device));
    bx_context* ctx = new bx_context();
    ctx->device = device;
    HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    int rc = EnsureEvents(ctx, 1);
    if (rc) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return BX_OK;
}

void bx_context_destroy(bx_context* ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
    hipStreamDestroy(ctx->stream);
    delete ctx;
}

void bx_batch_destroy(bx_batch* b) {
    if (!b) return;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->ctx->stream);
    for (void* p : b->allocations) hipFree(p);
    delete b;
}

int bx_batch_create(bx_context* ctx, const bx_shape_library* library, const bx_world_desc* worlds,
                    uint32_t world_count, const bx_batch_opts* opts, bx_batch** out) {
    if (!ctx || !out || (!worlds && world_count)) return Fail(BX_E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(ctx->device));
    HostLibrary L;
    int rc = BuildLibrary(library, L);
    if (rc) return rc;
    bx_batch* b = new bx_batch();
    b->ctx = ctx;
    b->worldCount = world_count;
    std::vector<WorldState> wstate(world_count);
    std::vector<Body> bodies;
    std::vector<Fixture> fixtures;
    std::vector<u32> fixtureList, proxyOrder;
    std::vector<AABB> fat;
    std::vector<Joint> joints;
    std::vector<JointEdge> jointEdges;
    std::vector<u32> freeSlots;
    u32 contactTotal = 0, pairTotal = 0, islandBodyTotal = 0;
    b->worlds.resize(world_count);
    b->proxyIds.resize(world_count);
    for (u32 w = 0; w < world_count; ++w) {
        const bx_world_desc& d = worlds[w];
        WorldInfo& wi = b->worlds[w];
        wi.bodyBase = u32(bodies.size());
        wi.bodyCount = d.body_count;
        wi.fixtureBase = u32(fixtures.size());
        wi.fixtureCount = d.fixture_count;
        wi.fixtureListBase = u32(fixtureList.size());
        wi.edgeBase = wi.fixtureBase + w;
        wi.jointBase = u32(joints.size());
        wi.jointCount = d.joint_count;
        wi.jointEdgeBase = u32(jointEdges.size());
        wi.flags = d.flags;
        wi.gravity = V3(d.gravity);
        wi.serialIslands = 0;
        // bodies (b3Body::b3Body, body.cpp:25-103)
        for (u32 i = 0; i < d.body_count; ++i) {
            const bx_body_desc& bd = d.bodies[i];
            if (bd.type > 2) { delete b; return Fail(BX_E_INVALID, "bad body type"); }
            Body x;
            std::memset(&x, 0, sizeof(x));
            x.type = bd.type;
            x.flags = 0;
            if (bd.flags & BX_BODY_AWAKE) x.flags |= kAwake;
            if (bd.flags & BX_BODY_AUTO_SLEEP) x.flags |= kAutoSleep;
            if (x.type == kDynamic) {
                x.mass = 1;
                x.invMass = 1;
                x.flags |= bd.flags & (kFixedX | kFixedY | kFixedZ);
            }
            x.v = V3(bd.linear_velocity);
            x.w = V3(bd.angular_velocity);
            x.localCenter.SetZero();
            x.x = V3(bd.position);
            x.q = Q4(bd.orientation);
            x.x0 = x.x;
            x.q0 = x.q;
            x.xfT = x.x;
            x.linearDamping = V3(bd.linear_damping);
            x.angularDamping = V3(bd.angular_damping);
            x.gravityScale = V3(bd.gravity_scale);
            x.linTol = bd.linear_sleep_tolerance;
            x.angTol = bd.angular_sleep_tolerance;
            x.islandIndex = 0;
            bodies.push_back(x);
        }
        // fixtures, grouped per body in list order (newest first)
        std::vector<std::vector<u32>> perBody(d.body_count);
        for (u32 i = 0; i < d.fixture_count; ++i) {
            const bx_fixture_desc& fd = d.fixtures[i];
            if (fd.body >= d.body_count) { delete b; return Fail(BX_E_INVALID, "fixture body out of range"); }
            Fixture f;
            std::memset(&f, 0, sizeof(f));
            f.body = fd.body;
            f.shapeType = fd.shape_type;
            f.shapeIndex = fd.shape_index;
            f.isSensor = fd.is_sensor ? 1u : 0u;
            f.density = fd.density;
            f.friction = fd.friction;
            f.restitution = fd.restitution;
            f.v1 = V3(fd.vertex1);
            f.v2 = V3(fd.vertex2);
            f.v3 = V3(fd.vertex3);
            switch (fd.shape_type) {
            case BX_SHAPE_HULL:
                if (fd.shape_index >= L.info.size()) { delete b; return Fail(BX_E_INVALID, "hull index out of range"); }
                f.radius = kHullRadius;
                break;
            case BX_SHAPE_SPHERE:
            case BX_SHAPE_CAPSULE:
                f.radius = fd.radius;
                break;
            default:
                delete b;
                return Fail(BX_E_UNSUPPORTED, "shape type not supported by the device path yet");
            }
            fixtures.push_back(f);
            perBody[fd.body].insert(perBody[fd.body].begin(), i);
        }
        u32 cursor = 0;
        for (u32 i = 0; i < d.body_count; ++i) {
            Body& x = bodies[wi.bodyBase + i];
            x.fixtureBegin = cursor;
            x.fixtureCount = u32(perBody[i].size());
            for (u32 f : perBody[i]) fixtureList.push_back(f);
            cursor += x.fixtureCount;
        }
        // proxies, fat AABBs and mass, in fixture creation order (b3Body::CreateFixture)
        std::vector<u32>& pid = b->proxyIds[w];
        pid.resize(d.fixture_count);
        u32 nextNode = 0, leaves = 0;
        std::vector<Fixture> local(fixtures.begin() + wi.fixtureBase, fixtures.end());
        std::vector<u32> localList(fixtureList.begin() + wi.fixtureListBase, fixtureList.end());
        for (u32 i = 0; i < d.fixture_count; ++i) {
            Fixture& f = fixtures[wi.fixtureBase + i];
            Body& x = bodies[wi.bodyBase + f.body];
            AABB aabb;
            HostAABB(f, Transform(x.xfT, x.q), L, aabb);
            Vec3 r(kAabbExtension, kAabbExtension, kAabbExtension);
            AABB fa;
            fa.lowerBound = aabb.lowerBound - r;
            fa.upperBound = aabb.upperBound + r;
            fat.push_back(fa);
            pid[i] = nextNode++;
            if (leaves > 0) ++nextNode;  // InsertLeaf allocates the new parent
            ++leaves;
            f.proxyId = pid[i];
            if (f.density > scalar(0)) {
                // ResetMass sees the fixtures created so far: restrict the list to them.
                Body tmp = x;
                std::vector<u32> visible;
                for (u32 k = 0; k < x.fixtureCount; ++k) {
                    u32 fi = localList[x.fixtureBegin + k];
                    if (fi <= i) visible.push_back(fi);
                }
                tmp.fixtureBegin = 0;
                tmp.fixtureCount = u32(visible.size());
                ResetMass(tmp, local, visible, 0, L);
                u32 fb = x.fixtureBegin, fc = x.fixtureCount;
                x = tmp;
                x.fixtureBegin = fb;
                x.fixtureCount = fc;
            }
        }
        // revolute joints (b3World::CreateJoint after the bodies' fixtures; joint_manager.cpp:30-72
        // prepends each joint's edges to both bodies' joint lists)
        std::vector<std::vector<JointEdge>> edgesOf(d.body_count);
        for (u32 i = 0; i < d.joint_count; ++i) {
            const bx_joint_desc& jd = d.joints[i];
            if (jd.type != BX_JOINT_REVOLUTE || jd.body_a >= d.body_count || jd.body_b >= d.body_count ||
                jd.body_a == jd.body_b) {
                delete b;
                return Fail(BX_E_INVALID, "bad joint descriptor");
            }
            if (jd.enable_limit) {
                delete b;
                return Fail(BX_E_UNSUPPORTED, "revolute joint limits are not supported by the device path");
            }
            Joint j;
            std::memset(&j, 0, sizeof(j));
            j.bodyA = jd.body_a;
            j.bodyB = jd.body_b;
            j.collideConnected = jd.collide_connected ? 1u : 0u;
            if (jd.use_initialize) {  // b3RevoluteJointDef::Initialize (revolute_joint.cpp:22-48)
                Mat33 rotation;
                rotation.z = V3(jd.axis);
                ComputeBasis(rotation.z, rotation.x, rotation.y);
                Quat q = Mat33Quat(rotation);
                const Body& A = bodies[wi.bodyBase + jd.body_a];
                const Body& Bd = bodies[wi.bodyBase + jd.body_b];
                Vec3 anchor = V3(jd.anchor);
                j.localAnchorA = MulT(Transform(A.xfT, A.q), anchor);
                j.localRotationA = MulC(A.q, q);
                j.localAnchorB = MulT(Transform(Bd.xfT, Bd.q), anchor);
                j.localRotationB = MulC(Bd.q, q);
                j.referenceRotation = Quat(Vec3(0, 0, 0), scalar(1));
            } else {
                j.localAnchorA = V3(jd.local_anchor_a);
                j.localRotationA = Q4(jd.local_rotation_a);
                j.localAnchorB = V3(jd.local_anchor_b);
                j.localRotationB = Q4(jd.local_rotation_b);
                j.referenceRotation = Q4(jd.reference_rotation);
            }
            j.enableMotor = jd.enable_motor ? 1u : 0u;
            j.enableLimit = 0;
            j.lowerAngle = jd.lower_angle;
            j.upperAngle = jd.upper_angle;
            j.motorSpeed = jd.motor_speed;
            j.maxMotorTorque = jd.max_motor_torque;
            joints.push_back(j);
            edgesOf[jd.body_a].insert(edgesOf[jd.body_a].begin(), JointEdge{i, jd.body_b});
            edgesOf[jd.body_b].insert(edgesOf[jd.body_b].begin(), JointEdge{i, jd.body_a});
        }
        u32 ecursor = 0;
        for (u32 i = 0; i < d.body_count; ++i) {
            Body& x = bodies[wi.bodyBase + i];
            x.jointBegin = ecursor;
            x.jointCount = u32(edgesOf[i].size());
            for (const JointEdge& e : edgesOf[i]) jointEdges.push_back(e);
            ecursor += x.jointCount;
        }
        std::vector<u32> po(d.fixture_count);
        for (u32 i = 0; i < d.fixture_count; ++i) po[i] = i;
        std::sort(po.begin(), po.end(), [&](u32 a, u32 c) { return pid[a] < pid[c]; });
        proxyOrder.insert(proxyOrder.end(), po.begin(), po.end());
        // Static bodies join every island they touch; the position solver renormalizes their
        // orientation once per contact point. If that is not a fixed point of the binary32
        // normalization the outcome depends on island order, so such worlds solve islands in order.
        for (u32 i = 0; i < d.body_count; ++i) {
            const Body& x = bodies[wi.bodyBase + i];
            if (x.type != kStatic) continue;
            Quat q = x.q;
            Quat n1 = q;
            n1.Normalize();
            if (std::memcmp(&n1, &q, sizeof(Quat)) != 0) wi.serialIslands = 1;
        }
        // capacities
        u32 cap = opts && opts->max_contacts_per_world ? opts->max_contacts_per_world : 12 * d.fixture_count + 64;
        u32 pcap = opts && opts->max_pairs_per_world ? opts->max_pairs_per_world : 2 * cap;
        wi.contactBase = contactTotal;
        wi.contactCap = cap;
        wi.pairBase = pairTotal;
        wi.pairCap = pcap;
        wi.islandBodyBase = islandBodyTotal;
        wi.islandBodyCap = d.body_count + cap;
        contactTotal += cap;
        pairTotal += pcap;
        islandBodyTotal += wi.islandBodyCap;
        for (u32 k = 0; k < cap; ++k) freeSlots.push_back(cap - 1 - k);
        WorldState& ws = wstate[w];
        std::memset(&ws, 0, sizeof(ws));
        ws.freeCount = cap;
        ws.newContacts = 1;
    }
    const u32 nb = u32(bodies.size()), nf = u32(fixtures.size());
    std::vector<u32> moved(nf, 1u);  // CreateProxy buffers every new proxy (broad_phase.cpp:78-84)
    BatchView& V = b->view;
    V.worldCount = world_count;
#define ALLOC(field, count, init)                                   \
    do {                                                            \
        int _r = DevAlloc(b, &field, (count), (init));              \
        if (_r) { bx_batch_destroy(b); return _r; }                 \
    } while (0)
    WorldInfo* dWorlds;
    ALLOC(dWorlds, world_count, b->worlds.data());
    V.worlds = dWorlds;
    ALLOC(V.wstate, world_count, wstate.data());
    HullInfo* dh;
    ALLOC(dh, L.info.size(), L.info.data());
    V.hulls = dh;
    Vec3* dv;
    ALLOC(dv, L.vertices.size(), L.vertices.data());
    V.hullVertices = dv;
    HalfEdge* de;
    ALLOC(de, L.edges.size(), L.edges.data());
    V.hullEdges = de;
    u32* dfc;
    ALLOC(dfc, L.faces.size(), L.faces.data());
    V.hullFaces = dfc;
    Plane* dp;
    ALLOC(dp, L.planes.size(), L.planes.data());
    V.hullPlanes = dp;
    ALLOC(V.bodies, nb, bodies.data());
    ALLOC(V.fixtures, nf, fixtures.data());
    u32* dfl;
    ALLOC(dfl, fixtureList.size(), fixtureList.data());
    V.fixtureList = dfl;
    u32* dpo;
    ALLOC(dpo, proxyOrder.size(), proxyOrder.data());
    V.proxyOrder = dpo;
    ALLOC(V.fatAABB, nf, fat.data());
    ALLOC(V.moved, nf, moved.data());
    ALLOC(V.joints, joints.size(), joints.data());
    JointEdge* dje;
    ALLOC(dje, jointEdges.size(), jointEdges.data());
    V.jointEdges = dje;
    ALLOC(V.contacts, contactTotal, (Contact*)nullptr);
    ALLOC(V.contactOrder, contactTotal, (u32*)nullptr);
    ALLOC(V.freeSlots, contactTotal, freeSlots.data());
    ALLOC(V.contactAlive, contactTotal, (u32*)nullptr);
    ALLOC(V.edgeOffset, nf + world_count, (u32*)nullptr);
    ALLOC(V.edgeSlots, 2 * size_t(contactTotal), (u32*)nullptr);
    ALLOC(V.pairs, pairTotal, (uint2*)nullptr);
    ALLOC(V.pairCounts, nf + world_count, (u32*)nullptr);
    ALLOC(V.islands, nb, (IslandInfo*)nullptr);
    ALLOC(V.islandBodies, islandBodyTotal, (u32*)nullptr);
    ALLOC(V.islandContacts, contactTotal, (uint4*)nullptr);
    ALLOC(V.islandJoints, joints.size(), (uint4*)nullptr);
    ALLOC(V.jointCons, joints.size(), (JointCon*)nullptr);
    ALLOC(V.dfsStack, nb, (u32*)nullptr);
    ALLOC(V.velBodies, islandBodyTotal, (VelBody*)nullptr);
    ALLOC(V.posBodies, islandBodyTotal, (PosBody*)nullptr);
    ALLOC(V.velCons, contactTotal, (VelCon*)nullptr);
    ALLOC(V.posCons, contactTotal, (PosCon*)nullptr);
    ALLOC(V.islandList, nb, (uint2*)nullptr);
    ALLOC(V.islandListCount, 1, (u32*)nullptr);
    {
        // island builder: LDS when a world's scratch fits, otherwise a global area
        const u32 kLdsWords = 64 * 1024 / 4;
        u32 need = 0, scratch = 0;
        for (u32 w = 0; w < world_count; ++w) {
            WorldInfo& wi = b->worlds[w];
            u32 words = IslandScratchWords(wi.bodyCount, wi.fixtureCount, wi.contactCap);
            wi.islandScratchWords = words;
            if (words <= kLdsWords) {
                wi.islandScratchBase = kNull;
                need = std::max(need, words);
            } else {
                wi.islandScratchBase = scratch;
                scratch += words;
            }
        }
        V.islandLdsWords = need;
        ALLOC(V.islandScratch, scratch, (u32*)nullptr);
        HIP_CHECK(hipMemcpy((void*)V.worlds, b->worlds.data(), world_count * sizeof(WorldInfo), hipMemcpyHostToDevice));
    }
#undef ALLOC
    *out = b;
    return BX_OK;
}


int bx_batch_step(bx_batch* b, const bx_step_cfg* cfg, uint32_t step_count) {
    if (!b || !cfg) return Fail(BX_E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    if (b->worldCount == 0 || step_count == 0) return BX_OK;
    return StepRange(b, cfg, 0, b->worldCount, step_count, true);
}

int bx_batch_step_subset(bx_batch* b, const bx_step_cfg* cfg, const uint32_t* worlds, const uint32_t* step_counts,
                         uint32_t count) {
    if (!b || !cfg || (!worlds && count)) return Fail(BX_E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    if (b->newContacts) {
        // the first step of a batch runs the creation-time FindNewContacts on every world
        int rc = StepRange(b, cfg, 0, b->worldCount, 0, false);
        if (rc) return rc;
    }
    for (u32 i = 0; i < count; ++i) {
        if (worlds[i] >= b->worldCount) return Fail(BX_E_INVALID, "world index out of range");
        u32 n = step_counts ? step_counts[i] : 1u;
        bool saved = b->newContacts;
        int rc = StepRange(b, cfg, worlds[i], 1, n, false);
        b->newContacts = saved;
        if (rc) return rc;
    }
    b->newContacts = false;
    return BX_OK;
}

int bx_batch_sync(bx_batch* b) {
    if (!b) return Fail(BX_E_INVALID, "null batch");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    return BX_OK;
}



int bx_batch_get_counts(bx_batch* b, uint32_t world, bx_world_counts* out) {
    WorldState ws;
    int rc = GetState(b, world, ws);
    if (rc) return rc;
    const WorldInfo& wi = b->worlds[world];
    std::memset(out, 0, sizeof(*out));
    out->body_count = wi.bodyCount;
    out->fixture_count = wi.fixtureCount;
    out->joint_count = wi.jointCount;
    out->contact_count = ws.contactCount;
    out->pair_count = ws.pairCount;
    out->island_count = ws.islandCount;
    out->status = ws.status;
    std::vector<Body> bodies;
    rc = Download(b, bodies, b->view.bodies + wi.bodyBase, wi.bodyCount);
    if (rc) return rc;
    for (const Body& x : bodies) out->awake_count += (x.flags & kAwake) ? 1u : 0u;
    std::vector<u32> order;
    rc = Download(b, order, b->view.contactOrder + wi.contactBase, ws.contactCount);
    if (rc) return rc;
    std::vector<Contact> contacts;
    rc = Download(b, contacts, b->view.contacts + wi.contactBase, wi.contactCap);
    if (rc) return rc;
    for (u32 s : order) out->touching_count += (contacts[s].flags & 2u) ? 1u : 0u;
    return BX_OK;
}

int bx_batch_get_bodies(bx_batch* b, uint32_t world, bx_body_state* out, uint32_t capacity) {
    if (!b || world >= b->worldCount || !out) return Fail(BX_E_INVALID, "bad argument");
    const WorldInfo& wi = b->worlds[world];
    std::vector<Body> bodies;
    int rc = Download(b, bodies, b->view.bodies + wi.bodyBase, wi.bodyCount);
    if (rc) return rc;
    u32 n = std::min<u32>(capacity, wi.bodyCount);
    for (u32 i = 0; i < n; ++i) {
        const Body& x = bodies[i];
        bx_body_state& s = out[i];
        std::memset(&s, 0, sizeof(s));
        s.position[0] = x.xfT.x; s.position[1] = x.xfT.y; s.position[2] = x.xfT.z;
        s.orientation[0] = x.q.v.x; s.orientation[1] = x.q.v.y; s.orientation[2] = x.q.v.z; s.orientation[3] = x.q.s;
        s.world_center[0] = x.x.x; s.world_center[1] = x.x.y; s.world_center[2] = x.x.z;
        s.linear_velocity[0] = x.v.x; s.linear_velocity[1] = x.v.y; s.linear_velocity[2] = x.v.z;
        s.angular_velocity[0] = x.w.x; s.angular_velocity[1] = x.w.y; s.angular_velocity[2] = x.w.z;
        s.sleep_time = x.sleepTime;
        s.flags = x.flags;
        s.island_index = x.islandIndex;
    }
    return int(n);
}

int bx_batch_get_contacts(bx_batch* b, uint32_t world, bx_contact_state* out, uint32_t capacity) {
    WorldState ws;
    int rc = GetState(b, world, ws);
    if (rc) return rc;
    const WorldInfo& wi = b->worlds[world];
    std::vector<u32> order;
    rc = Download(b, order, b->view.contactOrder + wi.contactBase, ws.contactCount);
    if (rc) return rc;
    std::vector<Contact> contacts;
    rc = Download(b, contacts, b->view.contacts + wi.contactBase, wi.contactCap);
    if (rc) return rc;
    std::vector<Fixture> fixtures;
    rc = Download(b, fixtures, b->view.fixtures + wi.fixtureBase, wi.fixtureCount);
    if (rc) return rc;
    u32 n = std::min<u32>(capacity, ws.contactCount);
    for (u32 i = 0; i < n; ++i) {
        const Contact& c = contacts[order[ws.contactCount - 1 - i]];  // world list: newest first
        bx_contact_state& s = out[i];
        std::memset(&s, 0, sizeof(s));
        s.fixture_a = c.fixtureA;
        s.fixture_b = c.fixtureB;
        s.touching = (c.flags & 2u) ? 1u : 0u;
        s.manifold_count = c.manifoldCount;
        if (fixtures[c.fixtureA].shapeType == BX_SHAPE_HULL && fixtures[c.fixtureB].shapeType == BX_SHAPE_HULL) {
            s.cache[0] = c.cache.cacheType;
            s.cache[1] = c.cache.featureType;
            s.cache[2] = c.cache.index1;
            s.cache[3] = c.cache.index2;
        }
        if (c.manifoldCount > 0) {
            const Manifold& m = c.manifold;
            bx_manifold_state& ms = s.manifolds[0];
            ms.point_count = m.pointCount;
            ms.tangent_impulse[0] = m.tangentImpulse.x;
            ms.tangent_impulse[1] = m.tangentImpulse.y;
            ms.motor_impulse = m.motorImpulse;
            for (u32 k = 0; k < m.pointCount; ++k) {
                const ManifoldPoint& p = m.points[k];
                bx_manifold_point_state& ps = ms.points[k];
                ps.local_normal1[0] = p.localNormal1.x; ps.local_normal1[1] = p.localNormal1.y; ps.local_normal1[2] = p.localNormal1.z;
                ps.local_point1[0] = p.localPoint1.x; ps.local_point1[1] = p.localPoint1.y; ps.local_point1[2] = p.localPoint1.z;
                ps.local_point2[0] = p.localPoint2.x; ps.local_point2[1] = p.localPoint2.y; ps.local_point2[2] = p.localPoint2.z;
                ps.normal_impulse = p.normalImpulse;
                ps.triangle_key = p.triangleKey;
                for (int j = 0; j < 4; ++j) ps.key[j] = p.key[j];
                ps.persisting = p.persisting;
                ps.edge_contact = p.edgeContact;
            }
        }
    }
    return int(n);
}

int bx_batch_get_pairs(bx_batch* b, uint32_t world, bx_pair* out, uint32_t capacity) {
    WorldState ws;
    int rc = GetState(b, world, ws);
    if (rc) return rc;
    const WorldInfo& wi = b->worlds[world];
    std::vector<uint2> pairs;
    rc = Download(b, pairs, b->view.pairs + wi.pairBase, std::min(ws.pairCount, wi.pairCap));
    if (rc) return rc;
    const std::vector<u32>& pid = b->proxyIds[world];
    u32 n = std::min<u32>(capacity, u32(pairs.size()));
    for (u32 i = 0; i < n; ++i) {
        out[i].proxy1 = pid[pairs[i].x];
        out[i].proxy2 = pid[pairs[i].y];
    }
    return int(n);
}

int bx_batch_get_island_contacts(bx_batch* b, uint32_t world, uint32_t* out, uint32_t capacity, uint32_t* count) {
    WorldState ws;
    int rc = GetState(b, world, ws);
    if (rc) return rc;
    const WorldInfo& wi = b->worlds[world];
    std::vector<u32> order;
    rc = Download(b, order, b->view.contactOrder + wi.contactBase, ws.contactCount);
    if (rc) return rc;
    std::vector<uint4> ic;
    rc = Download(b, ic, b->view.islandContacts + wi.contactBase, ws.islandContactCount);
    if (rc) return rc;
    std::vector<u32> listIndex(wi.contactCap, BX_NULL_INDEX);
    for (u32 k = 0; k < ws.contactCount; ++k) listIndex[order[k]] = ws.contactCount - 1 - k;
    u32 n = std::min<u32>(capacity, u32(ic.size()));
    for (u32 i = 0; i < n; ++i) out[i] = listIndex[ic[i].x];
    if (count) *count = u32(ic.size());
    return BX_OK;
}

int bx_batch_set_bodies(bx_batch* b, uint32_t world, const bx_body_state* in, uint32_t count) {
    if (!b || world >= b->worldCount || !in) return Fail(BX_E_INVALID, "bad argument");
    const WorldInfo& wi = b->worlds[world];
    if (count != wi.bodyCount) return Fail(BX_E_INVALID, "body count mismatch");
    std::vector<Body> bodies;
    int rc = Download(b, bodies, b->view.bodies + wi.bodyBase, wi.bodyCount);
    if (rc) return rc;
    for (u32 i = 0; i < count; ++i) {
        Body& x = bodies[i];
        x.v = V3(in[i].linear_velocity);
        x.w = V3(in[i].angular_velocity);
    }
    HIP_CHECK(hipMemcpyAsync(b->view.bodies + wi.bodyBase, bodies.data(), count * sizeof(Body), hipMemcpyHostToDevice,
                             b->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    return BX_OK;
}

int bx_batch_get_timings(bx_batch* b, float* ms, uint32_t count) {
    if (!b || !ms) return Fail(BX_E_INVALID, "bad argument");
    bx_context* ctx = b->ctx;
    if (ctx->timedSteps == 0) return Fail(BX_E_INVALID, "no timed step yet");
    HIP_CHECK(hipSetDevice(ctx->device));
    HIP_CHECK(hipEventSynchronize(ctx->ev[size_t(12) * (ctx->timedSteps - 1) + 11]));
    // phase boundaries (event indices): update contacts (+ initial pair search) 0-2, islands 2-3,
    // island solve kernels 3-9, fixture sync + pair search 9-11, whole step 0-11
    static const int kFrom[5] = {0, 2, 3, 9, 0}, kTo[5] = {2, 3, 9, 11, 11};
    double sum[5] = {};
    for (u32 s = 0; s < ctx->timedSteps; ++s) {
        hipEvent_t* e = ctx->ev.data() + size_t(12) * s;
        for (int i = 0; i < 5; ++i) {
            float t = 0.f;
            HIP_CHECK(hipEventElapsedTime(&t, e[kFrom[i]], e[kTo[i]]));
            sum[i] += t;
        }
    }
    for (u32 i = 0; i < count && i < 5; ++i) ms[i] = float(sum[i] / ctx->timedSteps);
    return BX_OK;
}

static const char* kKernelNames[11] = {"k_find_pairs(initial)", "k_update_contacts", "k_islands", "k_integrate_vel",
                                       "k_init_contacts", "k_solve_velocity", "k_mid", "k_solve_position",
                                       "k_finish", "k_sync", "k_find_pairs"};

const char* bx_kernel_name(uint32_t index) { return index < 11 ? kKernelNames[index] : ""; }

int bx_batch_get_kernel_timings(bx_batch* b, float* ms, uint32_t count) {
    if (!b || !ms) return Fail(BX_E_INVALID, "bad argument");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    bx_context* ctx = b->ctx;
    if (ctx->timedSteps == 0) return Fail(BX_E_INVALID, "no timed step yet");
    HIP_CHECK(hipEventSynchronize(ctx->ev[size_t(12) * (ctx->timedSteps - 1) + 11]));
    u32 n = std::min<u32>(count, 11);
    for (u32 i = 0; i < n; ++i) {
        double sum = 0.0;
        for (u32 s = 0; s < ctx->timedSteps; ++s) {
            hipEvent_t* e = ctx->ev.data() + size_t(12) * s;
            float t = 0.f;
            HIP_CHECK(hipEventElapsedTime(&t, e[i], e[i + 1]));
            sum += t;
        }
        ms[i] = float(sum / ctx->timedSteps);
    }
    return int(n);
}

// Size in bytes of each ABI structure, for layout checks by bindings.
uint32_t bx_struct_size(const char* name) {
    std::string n(name ? name : "");
#define S(T) if (n == #T) return uint32_t(sizeof(T));
    S(bx_hull_desc) S(bx_mesh_desc) S(bx_shape_library) S(bx_body_desc) S(bx_fixture_desc) S(bx_joint_desc)
    S(bx_world_desc) S(bx_body_state) S(bx_manifold_point_state) S(bx_manifold_state) S(bx_contact_state)
    S(bx_pair) S(bx_world_counts) S(bx_step_cfg) S(bx_batch_opts)
#undef S
    return 0;
}

uint64_t bx_batch_device_bytes(bx_batch* b) { return b ? uint64_t(b->deviceBytes) : 0; }

}  // extern "C"
