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
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <map>
#include <cstring>
#include <functional>
#include <string>
#include <vector>
#include "../../include/bxstep.h"
#include "bx_shapes.h"
#include "bx_joint.h"


namespace bx {
struct StepParams { scalar dt, inv_dt; u32 velocityIterations, positionIterations; };
void LaunchStep(const BatchView& view, const u32* worldList, u32 firstWorld, u32 worldCount, const StepParams& sp,
                bool findFirst, u32 maxIslands, u32 maxWorldBodies, hipStream_t stream, hipEvent_t* ev);
void LaunchInitialPairs(const BatchView& view, const u32* worldList, u32 worldCount, hipStream_t stream);
void LaunchBuildBvh(const BatchView& view, const u32* worldList, u32 worldCount, hipStream_t stream);
void LaunchQueries(const BatchView& view, const QueryParams& qp, hipStream_t stream);
void LaunchTimeOfImpact(const BatchView& view, const ToiQueryD* queries, u32 count, ToiResultD* out,
                        hipStream_t stream);
void LaunchShapeCasts(const BatchView& view, const CastParams& cp, hipStream_t stream);
}

using namespace bx;
typedef uint8_t u8;

static thread_local std::string g_error;
static thread_local const char* g_errorStatic = nullptr;  // set instead of g_error where nothing may allocate

static int Fail(int code, const std::string& msg) {
    g_error = msg;
    g_errorStatic = nullptr;
    return code;
}

// Failure with a message of static storage (no allocation on the error path).
static int FailStatic(int code, const char* msg) {
    g_errorStatic = msg;
    return code;
}

#define HIP_CHECK(expr)                                                                                  \
    do {                                                                                                 \
        hipError_t _e = (expr);                                                                          \
        if (_e != hipSuccess) return Fail(BX_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(_e)); \
    } while (0)

// HIP_CHECK for paths that allocate nothing, even when they fail: the message is HIP's own.
#define HIP_CHECK_S(expr)                                          \
    do {                                                           \
        hipError_t _e = (expr);                                    \
        if (_e != hipSuccess) return FailStatic(BX_E_DEVICE, hipGetErrorString(_e)); \
    } while (0)

// Steps of one bx_batch_step call whose kernels are timed with HIP events (12 events per step,
// see LaunchStep); later steps of a longer call run untimed.
static const u32 kMaxTimedSteps = 1024;

struct bx_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev;  // 12 per timed step of the last bx_batch_step call
    u32 timedSteps = 0;
    unsigned char* scratch = nullptr;  // this device's copy of `scratch` (bx_batch_step_noalloc)
};

// bx_batch_step_noalloc keeps the launch's island work list (one {world, island} entry per island, at
// most one per body) and its counter in this file-scope device array instead of in memory the batch
// allocated. Every device the library runs on holds one copy until the process ends; a context looks
// up its device's copy once, at creation. Calls that overlap on one device share it (and race).
static const u32 kNoAllocIslands = 1u << 21;
static const size_t kNoAllocListBytes = size_t(kNoAllocIslands) * sizeof(uint2);
static __device__ __align__(256) unsigned char scratch[kNoAllocListBytes + 256];

static int EnsureEvents(bx_context* ctx, u32 steps) {
    size_t need = size_t(12) * steps;
    while (ctx->ev.size() < need) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        ctx->ev.push_back(e);
    }
    return BX_OK;
}

// The per-world arrays of a batch live in seven index spaces. A world owns one contiguous range of
// units in each space, and every array of a space is indexed from the world's base in that space,
// so a world's whole state is relocatable: every index stored inside the arrays is world-local.
enum Space : u32 {
    kSpaceBody,        // bodyBase: bodies, islands, DFS stack
    kSpaceFixture,     // fixtureBase (= fixtureListBase = edgeBase): fixtureCount + 1 units
    kSpaceJoint,       // jointBase (jointEdgeBase = 2 x jointBase)
    kSpaceContact,     // contactBase: contactCap units
    kSpacePair,        // pairBase: pairCap units
    kSpaceIslandBody,  // islandBodyBase: islandBodyCap units
    kSpaceScratch,     // islandScratchBase: island-builder words at full capacity
    kSpaceCon,         // conBase: island constraint entries (3 x contactCap with meshes, else contactCap)
    kSpaceMesh,        // meshBase: mesh contact records, contactCap units in worlds with a mesh fixture
    kSpaceMeshTri,     // meshTriBase: contactCap x meshTriCap triangle indices
    kSpaceMeshScratch, // meshScratchBase: clustering stripes of the world's kBlock lanes
    kSpaces
};

static void WorldUnits(const WorldInfo& wi, u32 u[kSpaces]) {
    u[kSpaceBody] = wi.bodyCount;
    u[kSpaceFixture] = wi.fixtureCount + 1;  // the CSR offsets hold fixtureCount + 1 entries
    u[kSpaceJoint] = wi.jointCount;
    u[kSpaceContact] = wi.contactCap;
    u[kSpacePair] = wi.pairCap;
    u[kSpaceIslandBody] = wi.islandBodyCap;
    u[kSpaceScratch] = wi.islandScratchWords;
    u[kSpaceCon] = wi.conCap;
    u[kSpaceMesh] = wi.meshTriCap ? wi.contactCap : 0u;
    u[kSpaceMeshTri] = wi.meshTriCap ? wi.contactCap * wi.meshTriCap : 0u;
    u[kSpaceMeshScratch] = wi.meshTriCap ? kMeshLanes * MeshLaneStride(wi.meshTriCap) : 0u;
}

static void SetBases(WorldInfo& wi, const u32 base[kSpaces]) {
    wi.bodyBase = base[kSpaceBody];
    wi.fixtureBase = wi.fixtureListBase = wi.edgeBase = base[kSpaceFixture];
    wi.jointBase = base[kSpaceJoint];
    wi.jointEdgeBase = 2 * base[kSpaceJoint];
    wi.contactBase = base[kSpaceContact];
    wi.pairBase = base[kSpacePair];
    wi.islandBodyBase = base[kSpaceIslandBody];
    wi.islandScratchBase = base[kSpaceScratch];
    wi.conBase = base[kSpaceCon];
    wi.meshBase = base[kSpaceMesh];
    wi.meshTriBase = base[kSpaceMeshTri];
    wi.meshScratchBase = base[kSpaceMeshScratch];
}

static void GetBases(const WorldInfo& wi, u32 base[kSpaces]) {
    base[kSpaceBody] = wi.bodyBase;
    base[kSpaceFixture] = wi.fixtureBase;
    base[kSpaceJoint] = wi.jointBase;
    base[kSpaceContact] = wi.contactBase;
    base[kSpacePair] = wi.pairBase;
    base[kSpaceIslandBody] = wi.islandBodyBase;
    base[kSpaceScratch] = wi.islandScratchBase;
    base[kSpaceCon] = wi.conBase;
    base[kSpaceMesh] = wi.meshBase;
    base[kSpaceMeshTri] = wi.meshTriBase;
    base[kSpaceMeshScratch] = wi.meshScratchBase;
}

// Every per-world device array: its BatchView member, element size, space and elements per unit.
struct ArraySpec {
    size_t member;
    u32 elemSize, space, perUnit;
};
#define BX_ARRAY(m, space, per) {offsetof(BatchView, m), u32(sizeof(*static_cast<BatchView*>(nullptr)->m)), space, per}
static const ArraySpec kArrays[] = {
    BX_ARRAY(bodies, kSpaceBody, 1), BX_ARRAY(islands, kSpaceBody, 1), BX_ARRAY(dfsStack, kSpaceBody, 1),
    BX_ARRAY(fixtures, kSpaceFixture, 1), BX_ARRAY(fixtureList, kSpaceFixture, 1),
    BX_ARRAY(proxyOrder, kSpaceFixture, 1), BX_ARRAY(fatAABB, kSpaceFixture, 1), BX_ARRAY(moved, kSpaceFixture, 1),
    BX_ARRAY(edgeOffset, kSpaceFixture, 1), BX_ARRAY(pairCounts, kSpaceFixture, 1),
    BX_ARRAY(bvhLeaves, kSpaceFixture, 1), BX_ARRAY(bvhNodes, kSpaceFixture, 1),
    BX_ARRAY(joints, kSpaceJoint, 1), BX_ARRAY(islandJoints, kSpaceJoint, 1), BX_ARRAY(jointCons, kSpaceJoint, 1),
    BX_ARRAY(jointEdges, kSpaceJoint, 2),
    BX_ARRAY(contacts, kSpaceContact, 1), BX_ARRAY(contactOrder, kSpaceContact, 1),
    BX_ARRAY(freeSlots, kSpaceContact, 1), BX_ARRAY(contactAlive, kSpaceContact, 1),
#if BX_DEVICE_CALLBACKS
    BX_ARRAY(events, kSpaceContact, 1),
#endif
    BX_ARRAY(edgeSlots, kSpaceContact, 2),
    BX_ARRAY(islandContacts, kSpaceCon, 1), BX_ARRAY(velCons, kSpaceCon, 1), BX_ARRAY(posCons, kSpaceCon, 1),
    BX_ARRAY(meshContacts, kSpaceMesh, 1), BX_ARRAY(meshTriLists, kSpaceMeshTri, 1),
    BX_ARRAY(meshScratch, kSpaceMeshScratch, 1),
    BX_ARRAY(pairs, kSpacePair, 1),
    BX_ARRAY(islandBodies, kSpaceIslandBody, 1), BX_ARRAY(velBodies, kSpaceIslandBody, 1),
    BX_ARRAY(posBodies, kSpaceIslandBody, 1),
    BX_ARRAY(islandScratch, kSpaceScratch, 1),
};
#undef BX_ARRAY
static const u32 kArrayCount = u32(sizeof(kArrays) / sizeof(kArrays[0]));

static char* ArrayPtr(const BatchView& V, const ArraySpec& a) {
    char* p;
    std::memcpy(&p, reinterpret_cast<const char*>(&V) + a.member, sizeof(p));
    return p;
}
static void SetArrayPtr(BatchView& V, const ArraySpec& a, void* p) {
    std::memcpy(reinterpret_cast<char*>(&V) + a.member, &p, sizeof(p));
}

// Device bytes one unit of each space costs (the island work list is one entry per body unit).
static void UnitBytes(u64 out[kSpaces]) {
    for (u32 s = 0; s < kSpaces; ++s) out[s] = 0;
    for (const ArraySpec& a : kArrays) out[a.space] += u64(a.elemSize) * a.perUnit;
    out[kSpaceBody] += sizeof(uint2);
}

// First-fit allocator of unit ranges in one space, coalescing on free.
struct RangeAllocator {
    u64 capacity = 0;
    std::map<u64, u64> free;  // start -> length
    void Reset(u64 cap) {
        capacity = cap;
        free.clear();
        if (cap) free[0] = cap;
    }
    bool Alloc(u64 n, u64& start) {
        if (n == 0) { start = 0; return true; }
        for (auto it = free.begin(); it != free.end(); ++it) {
            if (it->second < n) continue;
            start = it->first;
            u64 rest = it->second - n;
            free.erase(it);
            if (rest) free[start + n] = rest;
            return true;
        }
        return false;
    }
    void Free(u64 start, u64 n) {
        if (n == 0) return;
        auto next = free.lower_bound(start);
        if (next != free.end() && start + n == next->first) {
            n += next->second;
            next = free.erase(next);
        }
        if (next != free.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                prev->second += n;
                return;
            }
        }
        free[start] = n;
    }
    u64 FreeUnits() const {
        u64 t = 0;
        for (auto& r : free) t += r.second;
        return t;
    }
};

// Device-side world lists of list launches: a ring of regions reused once the launches that read
// them have finished (one event per region).
static const u32 kListRing = 4;

// Tuning of the velocity-sweep layout (solver_mode 0) on a batch's first steps: kTuneSkip steps
// with the first layout (contacts and their points are still appearing), then every layout twice
// in the palindrome kTuneOrder; a layout's time is the smaller of its two samples. The first steps
// after an upload run slower (the device is coming out of idle), so a layout timed only early would
// lose to one timed later; with the first layout sampled first and last, such a start cannot move
// the choice away from it, while a layout that is faster by more than the early slowdown still wins.
static const u32 kTuneSkip = 3;
static const u32 kTuneOrder[2 * kSolveLayouts] = {0, 1, 2, 3, 4, 4, 3, 2, 1, 0};
static const u32 kTuneSteps = kTuneSkip + 2 * kSolveLayouts;
// Layouts within a few percent of each other are a tie in one or two samples; the first layout is
// kept unless another is faster by more than this fraction.
static const float kTuneMargin = 0.05f;

// b3DynamicTree's node pool as far as it fixes proxy ids (dynamic_tree.cpp:44-136): a LIFO free list
// over a sequential pool. CreateProxy takes the proxy's node and, when the tree already has a leaf,
// InsertLeaf takes a parent node; DestroyProxy's RemoveLeaf frees the leaf's parent (when the leaf is
// not the root) and then the proxy's node. Moving a proxy frees and takes back the same parent, so
// it leaves the pool unchanged. The ids of freed parents depend on the tree's shape and are not
// tracked (kNull entries); a proxy always takes a node freed by a proxy or a fresh one.
struct NodePool {
    std::vector<u32> freeStack;  // top last
    u32 nextNode = 0;
    u32 leafCount = 0;
};

struct bx_batch {
    bx_context* ctx = nullptr;
    u32 worldCount = 0;  // world slots
    BatchView view{};
    std::vector<void*> allocations;
    std::vector<WorldInfo> worlds;
    std::vector<std::vector<u32>> proxyIds;  // per world: proxy id of each fixture
    std::vector<NodePool> nodePools;  // per world: the dynamic tree's node pool (proxy ids of new fixtures)
    std::vector<u8> resident;                // per slot: holds a world
    std::vector<u8> needsInitialPairs;       // per slot: creation-time pair search still to run
    bool newContacts = true;                 // some resident world still needs its initial pair search
    size_t deviceBytes = 0;
    // pools
    bool pool = false;
    RangeAllocator spaces[kSpaces];
    std::vector<u32> freeWorldSlots;
    bx_batch_opts worldOpts{};
    struct HostLibrary* library = nullptr;
    // world lists
    u32* listHost = nullptr;  // pinned, kListRing regions of 2 x worldCount
    u32* listDev = nullptr;
    hipEvent_t listEv[kListRing] = {};
    bool listUsed[kListRing] = {};
    u32 listNext = 0;
    // one captured step of every world, replayed by bx_batch_step_graph (captured with graphParams)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graphExec = nullptr;
    StepParams graphParams{};
    u32 graphLayout = 0;
    // velocity-sweep layout tuning (solver_mode 0): steps of the tuning sequence done, and the
    // time of each layout (the smaller of its two samples)
    u32 tuneStep = kTuneSteps;
    float tuneMs[kSolveLayouts] = {};
    // world queries: the state generation (advanced by every call that moves fat AABBs), the
    // generation each world's query hierarchy was built at, and the device buffer of the queries
    u64 stateGen = 1;
    std::vector<u64> bvhGen;
    char* queryDev = nullptr;
    size_t queryBytes = 0;
};

static bool IsResident(const bx_batch* b, u32 world) { return world < b->worldCount && b->resident[world]; }

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
    std::vector<MeshInfo> meshes;
    std::vector<Vec3> meshVertices;
    std::vector<MeshTri> meshTris;
    std::vector<TreeNode> meshNodes;
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

// b3StaticTree::Build (static_tree.cpp:40-170) over the triangle AABBs (mesh.h:80-96) of one mesh,
// appended to L.meshNodes: nodes in pre-order (a node, its first child's subtree, then its second
// child's), each interior node splitting its triangles, sorted by AABB centre along the longest
// axis of its AABB with std::sort, at the first centre beyond the AABB centre (at the middle when
// that leaves one side empty). Returns the depth of the deepest node.
static u32 BuildMeshTree(const std::vector<AABB>& boxes, std::vector<TreeNode>& nodes) {
    const u32 n = u32(boxes.size());
    std::vector<u32> idx(n);
    for (u32 i = 0; i < n; ++i) idx[i] = i;
    struct Job { u32 first, count, parent, side, depth; };
    std::vector<Job> jobs{{0u, n, kNull, 0u, 0u}};
    const u32 base = u32(nodes.size());
    u32 maxDepth = 0;
    while (!jobs.empty()) {
        Job j = jobs.back();
        jobs.pop_back();
        const u32 id = u32(nodes.size()) - base;
        nodes.push_back(TreeNode{});
        if (j.parent != kNull) {
            TreeNode& p = nodes[base + j.parent];
            (j.side == 1u ? p.child1 : p.child2) = id;
        }
        maxDepth = std::max(maxDepth, j.depth);
        u32* r = idx.data() + j.first;
        TreeNode& node = nodes.back();
        if (j.count <= 1) {
            node.aabb = boxes[r[0]];
            node.child1 = node.child2 = kNull;
            node.index = r[0];
            continue;
        }
        AABB a = boxes[r[0]];
        for (u32 i = 1; i < j.count; ++i) a = Combine(a, boxes[r[i]]);
        node.aabb = a;
        node.index = kNull;
        Vec3 d = a.upperBound - a.lowerBound;  // b3AABB longest axis (aabb.h:164-184)
        u32 axis = 0;
        scalar longest = d.x;
        if (d.y > longest) { longest = d.y; axis = 1; }
        if (d.z > longest) axis = 2;
        auto coord = [&](u32 t) { Vec3 c = boxes[t].GetCenter(); return axis == 0 ? c.x : (axis == 1 ? c.y : c.z); };
        std::sort(r, r + j.count, [&](u32 t1, u32 t2) { return coord(t1) < coord(t2); });
        Vec3 ac = a.GetCenter();
        scalar split = axis == 0 ? ac.x : (axis == 1 ? ac.y : ac.z);
        u32 k = 0;
        while (k < j.count && !(coord(r[k]) > split)) ++k;
        if (k == 0 || k == j.count) k = j.count / 2;
        jobs.push_back(Job{j.first + k, j.count - k, id, 2u, j.depth + 1});  // built after the first child
        jobs.push_back(Job{j.first, k, id, 1u, j.depth + 1});
    }
    return maxDepth;
}

// b3Mesh::BuildAdjacency (mesh.cpp:35-112): for each triangle in order and each of its edges whose
// wing is unset, the first later triangle (and its first edge) sharing the edge in either
// direction sets both wings: this edge's to the other triangle's vertex off the edge, and the
// other edge's (overwriting it) to this triangle's. Found through a sorted edge list instead of a
// scan over the later triangles.
static void BuildMeshWings(MeshTri* t, u32 n) {
    struct EdgeRef { u64 key; u32 tri, edge; };
    auto vert = [&](const MeshTri& x, u32 i) { return i == 0 ? x.v1 : (i == 1 ? x.v2 : x.v3); };
    auto wing = [&](MeshTri& x, u32 i) -> u32& { return i == 0 ? x.u1 : (i == 1 ? x.u2 : x.u3); };
    std::vector<EdgeRef> refs;
    refs.reserve(size_t(3) * n);
    for (u32 i = 0; i < n; ++i) {
        t[i].u1 = t[i].u2 = t[i].u3 = kNull;
        for (u32 e = 0; e < 3; ++e) {
            u32 a = vert(t[i], e), b = vert(t[i], e + 1 < 3 ? e + 1 : 0);
            refs.push_back(EdgeRef{(u64(std::min(a, b)) << 32) | std::max(a, b), i, e});
        }
    }
    auto less = [](const EdgeRef& x, const EdgeRef& y) {
        return x.key != y.key ? x.key < y.key : (x.tri != y.tri ? x.tri < y.tri : x.edge < y.edge);
    };
    std::sort(refs.begin(), refs.end(), less);
    for (u32 i = 0; i < n; ++i) {
        for (u32 e = 0; e < 3; ++e) {
            if (wing(t[i], e) != kNull) continue;
            u32 a = vert(t[i], e), b = vert(t[i], e + 1 < 3 ? e + 1 : 0);
            EdgeRef probe{(u64(std::min(a, b)) << 32) | std::max(a, b), i + 1, 0};
            auto it = std::lower_bound(refs.begin(), refs.end(), probe, less);
            if (it == refs.end() || it->key != probe.key) continue;
            u32 e3 = e + 2 < 3 ? e + 2 : e - 1;            // vertex of triangle i off edge e
            u32 f3 = it->edge + 2 < 3 ? it->edge + 2 : it->edge - 1;
            wing(t[it->tri], it->edge) = vert(t[i], e3);
            wing(t[i], e) = vert(t[it->tri], f3);
        }
    }
}

static int BuildLibrary(const bx_shape_library* lib, HostLibrary& L) {
    if (!lib) return BX_OK;
    for (u32 i = 0; i < lib->hull_count; ++i) {
        const bx_hull_desc& d = lib->hulls[i];
        if (d.vertex_count < 4 || d.edge_count < 6 || d.face_count < 4 || (d.edge_count & 1u))
            return Fail(BX_E_INVALID, "hull " + std::to_string(i) + ": invalid counts");
        if (d.edge_count >= 0xFFFFu)  // feature ids are clipped in 16 bits
            return Fail(BX_E_INVALID, "hull " + std::to_string(i) + ": more than 65,534 half-edges");
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
    if (lib->mesh_count && !lib->meshes) return Fail(BX_E_INVALID, "null mesh array");
    for (u32 i = 0; i < lib->mesh_count; ++i) {
        const bx_mesh_desc& d = lib->meshes[i];
        if (d.vertex_count == 0 || d.triangle_count == 0 || !d.vertices || !d.triangles)
            return Fail(BX_E_INVALID, "mesh " + std::to_string(i) + ": empty");
        if (d.triangle_count > (1u << 26)) return Fail(BX_E_INVALID, "mesh " + std::to_string(i) + ": too many triangles");
        MeshInfo mi;
        std::memset(&mi, 0, sizeof(mi));
        mi.vBase = u32(L.meshVertices.size());
        mi.vCount = d.vertex_count;
        mi.tBase = u32(L.meshTris.size());
        mi.tCount = d.triangle_count;
        mi.nBase = u32(L.meshNodes.size());
        mi.root = 0;  // pre-order
        for (u32 k = 0; k < d.vertex_count; ++k) L.meshVertices.push_back(V3(d.vertices + 3 * k));
        std::vector<AABB> boxes(d.triangle_count);
        for (u32 k = 0; k < d.triangle_count; ++k) {
            MeshTri t;
            t.v1 = d.triangles[3 * k];
            t.v2 = d.triangles[3 * k + 1];
            t.v3 = d.triangles[3 * k + 2];
            if (t.v1 >= d.vertex_count || t.v2 >= d.vertex_count || t.v3 >= d.vertex_count)
                return Fail(BX_E_INVALID, "mesh " + std::to_string(i) + ": vertex index out of range");
            Vec3 a = V3(d.vertices + 3 * t.v1), b = V3(d.vertices + 3 * t.v2), c = V3(d.vertices + 3 * t.v3);
            boxes[k].lowerBound = Min(a, Min(b, c));  // b3Mesh::GetTriangleAABB (mesh.h:80-96)
            boxes[k].upperBound = Max(a, Max(b, c));
            boxes[k].Extend(kLinearSlop);
            L.meshTris.push_back(t);
        }
        BuildMeshWings(L.meshTris.data() + mi.tBase, d.triangle_count);
        u32 depth = BuildMeshTree(boxes, L.meshNodes);
        if (depth + 2 > kTreeStack) return Fail(BX_E_INVALID, "mesh " + std::to_string(i) + ": tree too deep");
        L.meshes.push_back(mi);
    }
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
        else if (f.shapeType == BX_SHAPE_TRIANGLE) {  // triangle_shape.cpp:67-73
            md.center = (f.v1 + f.v2 + f.v3) / scalar(3);
            md.mass = 0;
            md.I.SetZero();
        }
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
    } else if (f.shapeType == BX_SHAPE_TRIANGLE) {  // triangle_shape.cpp:75-83 (local bounds)
        Vec3 r(f.radius, f.radius, f.radius);
        out.lowerBound = Min(f.v1, Min(f.v2, f.v3)) - r;
        out.upperBound = Max(f.v1, Max(f.v2, f.v3)) + r;
    } else if (f.shapeType == BX_SHAPE_MESH) {  // mesh_shape.cpp:44-50
        const MeshInfo& mi = L.meshes[f.shapeIndex];
        const Vec3* v = L.meshVertices.data() + mi.vBase;
        out.lowerBound = out.upperBound = Mul(xf, Mul(f.v1, v[0]));
        for (u32 i = 1; i < mi.vCount; ++i) {
            Vec3 p = Mul(xf, Mul(f.v1, v[i]));
            out.lowerBound = Min(out.lowerBound, p);
            out.upperBound = Max(out.upperBound, p);
        }
        out.Extend(f.radius);
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
    if (!b || !IsResident(b, world)) return Fail(BX_E_INVALID, "bad batch or world not resident");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    HIP_CHECK(hipMemcpyAsync(&ws, b->view.wstate + world, sizeof(ws), hipMemcpyDeviceToHost, b->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    return BX_OK;
}

// Default contact-slot capacity of a world: the reference grows its contact list without bound;
// here a world holds a fixed number of slots, enough for the densest piles of the mixed-shapes
// workload (up to about 12.5 fat-AABB overlaps per fixture while bodies settle). A step that needs
// more sets status bit 0x2 and the world no longer follows the reference.
static const u32 kContactsPerFixture = 16, kContactSlack = 64;
static u64 DefaultContactCap(u32 fixtures) { return u64(kContactsPerFixture) * fixtures + kContactSlack; }

// Triangles a mesh contact holds by default: the fat AABB of a shape of about a metre over a
// heightfield of 1 m cells, moving a few metres per step, overlaps well under this many.
static const u32 kMeshTrianglesPerContact = 64;

// The sizes of a world that follow from its counts and the options: contact slots, pairs, island
// entries, island-builder words, and for a world with a mesh fixture the mesh contact records and
// three constraint entries per contact slot (a mesh contact has up to three manifolds).
static int SetCapacities(WorldInfo& wi, bool hasMesh, const bx_batch_opts* o) {
    u64 cap = o && o->max_contacts_per_world ? o->max_contacts_per_world : DefaultContactCap(wi.fixtureCount);
    u64 pcap = o && o->max_pairs_per_world ? o->max_pairs_per_world : 2 * cap;
    u64 tcap = hasMesh ? (o && o->mesh_triangles_per_contact ? o->mesh_triangles_per_contact : kMeshTrianglesPerContact) : 0;
    if (cap > (1u << 24) || pcap > (1u << 25)) return Fail(BX_E_INVALID, "contact capacity too large");
    if (tcap > (1u << 20) || cap * tcap > 0x7FFFFFFFull || u64(kMeshLanes) * MeshLaneStride(u32(tcap)) > 0x7FFFFFFFull)
        return Fail(BX_E_INVALID, "mesh triangle capacity too large");
    wi.contactCap = u32(cap);
    wi.pairCap = u32(pcap);
    wi.islandBodyCap = wi.bodyCount + wi.contactCap;
    wi.islandScratchWords = IslandScratchWords(wi.bodyCount, wi.fixtureCount, wi.contactCap);
    wi.conCap = hasMesh ? kMaxManifolds * wi.contactCap : wi.contactCap;
    wi.meshTriCap = u32(tcap);
    return BX_OK;
}

static bool HasMesh(const bx_world_desc& d) {
    for (u32 i = 0; i < d.fixture_count; ++i)
        if (d.fixtures[i].shape_type == BX_SHAPE_MESH) return true;
    return false;
}

static u32 TakeNode(NodePool& p) {
    if (p.freeStack.empty()) return p.nextNode++;
    const u32 n = p.freeStack.back();
    p.freeStack.pop_back();
    return n;
}

// The proxy id b3DynamicTree::CreateProxy returns; false if it would be an untracked node.
static bool PoolCreateProxy(NodePool& p, u32* proxy) {
    *proxy = TakeNode(p);
    if (*proxy == kNull) return false;
    if (p.leafCount > 0) (void)TakeNode(p);  // InsertLeaf's new parent
    ++p.leafCount;
    return true;
}

static void PoolDestroyProxy(NodePool& p, u32 proxy) {
    if (p.leafCount > 1) p.freeStack.push_back(kNull);  // RemoveLeaf frees the parent first
    p.freeStack.push_back(proxy);
    --p.leafCount;
}

// A pair-list entry that holds a proxy id rather than a fixture index: the fixture was destroyed
// after the pair search that reported it (only the host reads such entries; the next search
// rewrites the list).
static const u32 kPairProxyBit = 0x80000000u;

// One world built on the host from its descriptor, or taken from a resident world for a topology
// edit. Every index in these arrays is world-local; `info` holds sizes, capacities and flags (its
// bases are set when the world is placed).
struct HostWorld {
    WorldInfo info;
    WorldState state;
    std::vector<Body> bodies;
    std::vector<Fixture> fixtures;
    std::vector<u32> fixtureList, proxyOrder, proxyIds, freeSlots;
    std::vector<AABB> fat;
    std::vector<Joint> joints;
    std::vector<JointEdge> jointEdges;
    NodePool nodes;
    // State carried over from a resident world: the fixtures' moved flags, the live contacts in list
    // order with their mesh records and triangle lists, the last pair search's pairs, and whether the
    // next step begins with the new-contact search.
    bool carried = false;
    std::vector<u32> moved;
    std::vector<Contact> contacts;
    std::vector<MeshContact> meshContacts;
    std::vector<u32> meshTris;  // contacts.size() x info.meshTriCap
    std::vector<uint2> pairs;
    bool needsInitialPairs = true;
};

// b3Body::b3Body (body.cpp:25-103) from a descriptor; mass data follows from the fixtures.
static int MakeBody(const bx_body_desc& bd, Body& x) {
    if (bd.type > 2) return Fail(BX_E_INVALID, "bad body type");
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
    return BX_OK;
}

// b3Fixture::Create (fixture.cpp:43-56) from a descriptor (the owning body is checked by the caller).
static int MakeFixture(const bx_fixture_desc& fd, const HostLibrary& L, Fixture& f) {
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
        if (fd.shape_index >= L.info.size()) return Fail(BX_E_INVALID, "hull index out of range");
        f.radius = kHullRadius;
        break;
    case BX_SHAPE_SPHERE:
    case BX_SHAPE_CAPSULE:
        f.radius = fd.radius;
        break;
    case BX_SHAPE_TRIANGLE:  // b3TriangleShape::Set: hull skin radius, no edge wings
        f.radius = kHullRadius;
        break;
    case BX_SHAPE_MESH:  // b3MeshShape: the mesh by index, its scale in v1, hull skin radius
        if (fd.shape_index >= L.meshes.size()) return Fail(BX_E_INVALID, "mesh index out of range");
        if (fd.mesh_scale[0] == 0.0f || fd.mesh_scale[1] == 0.0f || fd.mesh_scale[2] == 0.0f)
            return Fail(BX_E_INVALID, "mesh scale must be nonzero");
        f.radius = kHullRadius;
        f.v1 = V3(fd.mesh_scale);
        f.v2.SetZero();
        f.v3.SetZero();
        break;
    default:
        return Fail(BX_E_UNSUPPORTED, "shape type not supported by the device path yet");
    }
    return BX_OK;
}

// b3Joint::Create (joint.cpp:33-100) between two of `bodies`: the joint's constructor, with the frames
// of its def's Initialize taken from the bodies' current transforms when asked.
static int MakeJoint(const bx_joint_desc& jd, const std::vector<Body>& bodies, Joint& j) {
    if (jd.type < BX_JOINT_MOUSE || jd.type > BX_JOINT_WHEEL || jd.body_a >= bodies.size() ||
        jd.body_b >= bodies.size() || jd.body_a == jd.body_b)
        return Fail(BX_E_INVALID, "bad joint descriptor");
    std::memset(&j, 0, sizeof(j));
    j.type = jd.type;
    j.bodyA = jd.body_a;
    j.bodyB = jd.body_b;
    j.collideConnected = jd.collide_connected ? 1u : 0u;
    j.frequencyHz = jd.frequency_hz;
    j.dampingRatio = jd.damping_ratio;
    j.length = jd.length;
    j.maxForce = jd.max_force;
    j.maxTorque = jd.max_torque;
    j.correctionFactor = jd.correction_factor;
    j.target = V3(jd.target);
    j.linearOffset = V3(jd.linear_offset);
    j.angularOffset = Q4(jd.angular_offset);
    const Body& A0 = bodies[jd.body_a];
    const Body& B0 = bodies[jd.body_b];
    const Transform xfA(A0.xfT, A0.q), xfB(B0.xfT, B0.q);
    if (jd.type == BX_JOINT_MOUSE) {  // mouse_joint.cpp:22-31: the target is B's anchor at creation
        j.localAnchorB = MulT(xfB, j.target);
        j.localRotationA = j.localRotationB = j.referenceRotation = Quat(Vec3(0, 0, 0), scalar(1));
        return BX_OK;
    }
    if (jd.type == BX_JOINT_PRISMATIC || jd.type == BX_JOINT_WHEEL) {
        j.referenceRotation = Q4(jd.reference_rotation);
        j.localRotationA = j.localRotationB = Quat(Vec3(0, 0, 0), scalar(1));
        Vec3 axisA = V3(jd.local_axis_a), axisB = V3(jd.local_axis_b);
        if (jd.use_initialize) {  // prismatic_joint.cpp:23-35, wheel_joint.cpp:22-31
            const Vec3 anchor = V3(jd.anchor);
            j.localAnchorA = MulT(xfA, anchor);
            j.localAnchorB = MulT(xfB, anchor);
            axisA = MulC(A0.q, V3(jd.axis));
            if (jd.type == BX_JOINT_PRISMATIC) j.referenceRotation = Conjugate(A0.q) * B0.q;
            else axisB = MulC(B0.q, V3(jd.axis_b));
        } else {
            j.localAnchorA = V3(jd.local_anchor_a);
            j.localAnchorB = V3(jd.local_anchor_b);
        }
        // the constructors (prismatic_joint.cpp:37-61, wheel_joint.cpp:33-66)
        j.localXAxisA = axisA;
        j.localXAxisA.Normalize();
        ComputeBasis(j.localXAxisA, j.localYAxisA, j.localZAxisA);
        j.enableMotor = jd.enable_motor ? 1u : 0u;
        j.motorSpeed = jd.motor_speed;
        if (jd.type == BX_JOINT_PRISMATIC) {
            j.enableLimit = jd.enable_limit ? 1u : 0u;
            j.lowerTranslation = jd.lower_translation;
            j.upperTranslation = jd.upper_translation;
        } else {
            j.maxMotorTorque = jd.max_motor_torque;
            j.localXAxisB = axisB;
            j.localXAxisB.Normalize();
            const Vec3 wA = Mul(A0.q, j.localXAxisA), wB = Mul(B0.q, j.localXAxisB);
            const Vec3 u = Cross(wA, wB);
            const scalar x = Dot(wA, wB), y = Length(u);
            j.referenceCosine = x;
            j.referenceAngle = scalar(::atan2(double(y), double(x)));
        }
        return BX_OK;
    }
    if (jd.type == BX_JOINT_CONE) {
        j.coneAngle = jd.cone_angle;
        j.enableLimit = jd.enable_limit ? 1u : 0u;
    }
    if (jd.type != BX_JOINT_REVOLUTE && jd.type != BX_JOINT_CONE) {
        j.localRotationA = Q4(jd.local_rotation_a);
        j.localRotationB = Q4(jd.local_rotation_b);
        j.referenceRotation = Q4(jd.reference_rotation);
        if (!jd.use_initialize) {
            j.localAnchorA = V3(jd.local_anchor_a);
            j.localAnchorB = V3(jd.local_anchor_b);
        } else if (jd.type == BX_JOINT_MOTOR) {  // motor_joint.cpp:23-31
            j.linearOffset = MulT(xfA, B0.xfT);
            j.angularOffset = MulC(A0.q, B0.q);
        } else {  // sphere_joint.cpp:23-29, weld_joint.cpp:23-33, spring_joint.cpp:23-30, friction_joint.cpp:23-29
            const Vec3 anchor = V3(jd.anchor);
            const Vec3 anchorB = jd.type == BX_JOINT_SPRING ? V3(jd.anchor_b) : anchor;
            j.localAnchorA = MulT(xfA, anchor);
            j.localAnchorB = MulT(xfB, anchorB);
            if (jd.type == BX_JOINT_WELD) j.referenceRotation = Conjugate(A0.q) * B0.q;
            if (jd.type == BX_JOINT_SPRING) j.length = Distance(anchor, anchorB);
        }
        return BX_OK;
    }
    if (jd.use_initialize) {  // b3RevoluteJointDef::Initialize (revolute_joint.cpp:22-48),
                              // b3ConeJointDef::Initialize (cone_joint.cpp:35-55) about x
        Mat33 rotation;
        if (jd.type == BX_JOINT_CONE) {
            rotation.x = V3(jd.axis);
            ComputeBasis(rotation.x, rotation.y, rotation.z);
        } else {
            rotation.z = V3(jd.axis);
            ComputeBasis(rotation.z, rotation.x, rotation.y);
        }
        Quat q = Mat33Quat(rotation);
        const Body& A = bodies[jd.body_a];
        const Body& Bd = bodies[jd.body_b];
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
    if (jd.type == BX_JOINT_CONE) return BX_OK;
    j.enableMotor = jd.enable_motor ? 1u : 0u;
    j.enableLimit = jd.enable_limit ? 1u : 0u;
    j.lowerAngle = jd.lower_angle;
    j.upperAngle = jd.upper_angle;
    j.motorSpeed = jd.motor_speed;
    j.maxMotorTorque = jd.max_motor_torque;
    return BX_OK;
}

// b3Body::ResetMass (body.cpp:265-376) of a body whose fixture list (list order) is `list`.
static void ResetMassOf(Body& x, const std::vector<Fixture>& fixtures, const std::vector<u32>& list, const HostLibrary& L) {
    const u32 fb = x.fixtureBegin, fc = x.fixtureCount;
    x.fixtureBegin = 0;
    x.fixtureCount = u32(list.size());
    ResetMass(x, fixtures, list, 0, L);
    x.fixtureBegin = fb;
    x.fixtureCount = fc;
}

// The fat AABB b3DynamicTree::CreateProxy stores for a fixture at transform xf.
static AABB CreationFatAABB(const Fixture& f, const Transform& xf, const HostLibrary& L) {
    AABB aabb;
    HostAABB(f, xf, L, aabb);
    Vec3 r(kAabbExtension, kAabbExtension, kAabbExtension);
    AABB fa;
    fa.lowerBound = aabb.lowerBound - r;
    fa.upperBound = aabb.upperBound + r;
    return fa;
}

// Sorted proxy order, the serial-island flag, capacities, the contact slot stack and the counters
// of a world whose bodies, fixtures, joints (and carried contacts) are final.
static int FinishWorld(HostWorld& hw, bool hasMesh, const bx_batch_opts* opts) {
    WorldInfo& wi = hw.info;
    const std::vector<u32>& pid = hw.proxyIds;
    hw.proxyOrder.resize(wi.fixtureCount);
    for (u32 i = 0; i < wi.fixtureCount; ++i) hw.proxyOrder[i] = i;
    std::sort(hw.proxyOrder.begin(), hw.proxyOrder.end(), [&](u32 a, u32 c) { return pid[a] < pid[c]; });
    // Static bodies join every island they touch; the position solver renormalizes their
    // orientation once per contact point. If that is not a fixed point of the binary32
    // normalization the outcome depends on island order, so such worlds solve islands in order.
    wi.serialIslands = 0;
    for (const Body& x : hw.bodies) {
        if (x.type != kStatic) continue;
        Quat q = x.q;
        Quat n1 = q;
        n1.Normalize();
        if (std::memcmp(&n1, &q, sizeof(Quat)) != 0) wi.serialIslands = 1;
    }
    const u32 oldTriCap = wi.meshTriCap;
    int rc = SetCapacities(wi, hasMesh, opts);
    if (rc) return rc;
    const u32 cap = wi.contactCap;
    const u32 live = u32(hw.contacts.size());
    if (live > cap) return Fail(BX_E_CAPACITY, "the world's live contacts exceed its contact capacity");
    hw.freeSlots.resize(cap - live);  // slots [live, cap) free, slot `live` on top
    for (u32 k = 0; k < cap - live; ++k) hw.freeSlots[k] = cap - 1 - k;
    if (!hw.carried) {
        std::memset(&hw.state, 0, sizeof(hw.state));
        hw.state.newContacts = 1;
        hw.moved.assign(wi.fixtureCount, 1u);  // CreateProxy buffers every new proxy
    } else {
        // the contacts keep their list order in slots [0, live); the last step's islands are not
        // carried (the next step builds its own)
        if (wi.meshTriCap != oldTriCap) {
            hw.meshContacts.assign(wi.meshTriCap ? live : 0, MeshContact{});
            hw.meshTris.assign(size_t(live) * wi.meshTriCap, 0u);
            if (oldTriCap) return Fail(BX_E_UNSUPPORTED, "mesh triangle capacity changed under live mesh contacts");
        }
        if (hw.pairs.size() > wi.pairCap) hw.pairs.resize(wi.pairCap);
        hw.state.contactCount = live;
        hw.state.freeCount = cap - live;
        hw.state.pairCount = u32(hw.pairs.size());
        hw.state.islandCount = hw.state.islandBodyCount = hw.state.islandContactCount = hw.state.islandJointCount = 0;
        hw.state.newContacts = hw.needsInitialPairs ? 1u : 0u;
    }
    hw.state.freeCount = cap - live;
    return BX_OK;
}

static int BuildWorld(const bx_world_desc& d, const HostLibrary& L, const bx_batch_opts* opts, HostWorld& hw) {
    WorldInfo& wi = hw.info;
    std::memset(&wi, 0, sizeof(wi));
    wi.bodyCount = d.body_count;
    wi.fixtureCount = d.fixture_count;
    wi.jointCount = d.joint_count;
    wi.flags = d.flags;
#if !BX_DEVICE_CALLBACKS
    if (d.flags & BX_WORLD_EVENTS)
        return Fail(BX_E_UNSUPPORTED, "contact event recording (BX_WORLD_EVENTS) is not enabled in this build");
#endif
    wi.gravity = V3(d.gravity);
    wi.serialIslands = 0;
    if ((d.body_count && !d.bodies) || (d.fixture_count && !d.fixtures) || (d.joint_count && !d.joints))
        return Fail(BX_E_INVALID, "null descriptor array");
    // the island builder packs a body's first fixture-list entry and first joint edge (two per
    // joint) with their counts into 16-bit halves
    if (d.fixture_count > 0xFFFFu || d.body_count > 0xFFFFFFu || d.joint_count > 0x7FFFu)
        return Fail(BX_E_INVALID, "world too large");
    // bodies (b3Body::b3Body, body.cpp:25-103)
    std::vector<Body>& bodies = hw.bodies;
    bodies.resize(d.body_count);
    for (u32 i = 0; i < d.body_count; ++i) {
        int rc = MakeBody(d.bodies[i], bodies[i]);
        if (rc) return rc;
    }
    // fixtures, grouped per body in list order (newest first)
    std::vector<Fixture>& fixtures = hw.fixtures;
    fixtures.resize(d.fixture_count);
    std::vector<std::vector<u32>> perBody(d.body_count);
    for (u32 i = 0; i < d.fixture_count; ++i) {
        if (d.fixtures[i].body >= d.body_count) return Fail(BX_E_INVALID, "fixture body out of range");
        int rc = MakeFixture(d.fixtures[i], L, fixtures[i]);
        if (rc) return rc;
        perBody[d.fixtures[i].body].insert(perBody[d.fixtures[i].body].begin(), i);
    }
    u32 cursor = 0;
    for (u32 i = 0; i < d.body_count; ++i) {
        Body& x = bodies[i];
        x.fixtureBegin = cursor;
        x.fixtureCount = u32(perBody[i].size());
        for (u32 f : perBody[i]) hw.fixtureList.push_back(f);
        cursor += x.fixtureCount;
    }
    // proxies, fat AABBs and mass, in fixture creation order (b3Body::CreateFixture)
    std::vector<u32>& pid = hw.proxyIds;
    pid.resize(d.fixture_count);
    const std::vector<Fixture> local(fixtures);
    for (u32 i = 0; i < d.fixture_count; ++i) {
        Fixture& f = fixtures[i];
        Body& x = bodies[f.body];
        hw.fat.push_back(CreationFatAABB(f, Transform(x.xfT, x.q), L));
        (void)PoolCreateProxy(hw.nodes, &pid[i]);  // a fresh pool: every node is tracked
        f.proxyId = pid[i];
        if (f.density > scalar(0)) {
            // ResetMass sees the fixtures created so far: restrict the list to them.
            std::vector<u32> visible;
            for (u32 k = 0; k < x.fixtureCount; ++k) {
                u32 fi = hw.fixtureList[x.fixtureBegin + k];
                if (fi <= i) visible.push_back(fi);
            }
            ResetMassOf(x, local, visible, L);
        }
    }
    // revolute joints (b3World::CreateJoint after the bodies' fixtures; joint_manager.cpp:30-72
    // prepends each joint's edges to both bodies' joint lists)
    std::vector<std::vector<JointEdge>> edgesOf(d.body_count);
    hw.joints.resize(d.joint_count);
    for (u32 i = 0; i < d.joint_count; ++i) {
        const bx_joint_desc& jd = d.joints[i];
        int rc = MakeJoint(jd, bodies, hw.joints[i]);
        if (rc) return rc;
        edgesOf[jd.body_a].insert(edgesOf[jd.body_a].begin(), JointEdge{i, jd.body_b});
        edgesOf[jd.body_b].insert(edgesOf[jd.body_b].begin(), JointEdge{i, jd.body_a});
    }
    u32 ecursor = 0;
    for (u32 i = 0; i < d.body_count; ++i) {
        Body& x = bodies[i];
        x.jointBegin = ecursor;
        x.jointCount = u32(edgesOf[i].size());
        for (const JointEdge& e : edgesOf[i]) hw.jointEdges.push_back(e);
        ecursor += x.jointCount;
    }
    return FinishWorld(hw, HasMesh(d), opts);
}

// Island-builder LDS words for a launch over worlds with the given full-capacity needs: worlds
// whose working arrays fit this budget at the current contact count use LDS, the rest their
// global area; the budget is small enough to keep several worlds per CU in flight.
static const u32 kIslandLdsBudgetWords = 6144;

// Allocates every per-world array for the given units per space, zeroed, plus the world tables
// and the world-list ring.
static int AllocArrays(bx_batch* b, const u64 cap[kSpaces], u32 worldSlots) {
    BatchView& V = b->view;
    for (const ArraySpec& a : kArrays) {
        u64 count = cap[a.space] * a.perUnit;
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, std::max<u64>(1, count) * a.elemSize);
        if (e != hipSuccess) return Fail(BX_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
        b->allocations.push_back(p);
        b->deviceBytes += std::max<u64>(1, count) * a.elemSize;
        HIP_CHECK(hipMemset(p, 0, std::max<u64>(1, count) * a.elemSize));
        SetArrayPtr(V, a, p);
    }
    int rc;
    WorldInfo* dw;
    if ((rc = DevAlloc(b, &dw, worldSlots))) return rc;
    V.worlds = dw;
    if ((rc = DevAlloc(b, &V.wstate, worldSlots))) return rc;
    if ((rc = DevAlloc(b, &V.islandList, cap[kSpaceBody]))) return rc;
    if ((rc = DevAlloc(b, &V.islandListCount, 1))) return rc;
    if ((rc = DevAlloc(b, &b->listDev, size_t(kListRing) * 2 * std::max(1u, worldSlots)))) return rc;
    HIP_CHECK(hipHostMalloc((void**)&b->listHost, sizeof(u32) * kListRing * 2 * std::max(1u, worldSlots)));
    for (u32 r = 0; r < kListRing; ++r) HIP_CHECK(hipEventCreateWithFlags(&b->listEv[r], hipEventDisableTiming));
    V.worldCount = worldSlots;
    return BX_OK;
}

static int UploadLibrary(bx_batch* b, const HostLibrary& L) {
    BatchView& V = b->view;
    int rc;
    HullInfo* dh;
    if ((rc = DevAlloc(b, &dh, L.info.size(), L.info.data()))) return rc;
    V.hulls = dh;
    Vec3* dv;
    if ((rc = DevAlloc(b, &dv, L.vertices.size(), L.vertices.data()))) return rc;
    V.hullVertices = dv;
    HalfEdge* de;
    if ((rc = DevAlloc(b, &de, L.edges.size(), L.edges.data()))) return rc;
    V.hullEdges = de;
    u32* dfc;
    if ((rc = DevAlloc(b, &dfc, L.faces.size(), L.faces.data()))) return rc;
    V.hullFaces = dfc;
    Plane* dp;
    if ((rc = DevAlloc(b, &dp, L.planes.size(), L.planes.data()))) return rc;
    V.hullPlanes = dp;
    MeshInfo* dm;
    if ((rc = DevAlloc(b, &dm, L.meshes.size(), L.meshes.data()))) return rc;
    V.meshes = dm;
    Vec3* dmv;
    if ((rc = DevAlloc(b, &dmv, L.meshVertices.size(), L.meshVertices.data()))) return rc;
    V.meshVertices = dmv;
    MeshTri* dmt;
    if ((rc = DevAlloc(b, &dmt, L.meshTris.size(), L.meshTris.data()))) return rc;
    V.meshTris = dmt;
    TreeNode* dn;
    if ((rc = DevAlloc(b, &dn, L.meshNodes.size(), L.meshNodes.data()))) return rc;
    V.meshNodes = dn;
    return BX_OK;
}

// Copies a slice of host elements to element `at` of a device array, on the batch stream.
template <class T>
static int PutAsync(bx_batch* b, T* dev, u64 at, const std::vector<T>& host) {
    if (host.empty()) return BX_OK;
    HIP_CHECK(hipMemcpyAsync(dev + at, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, b->ctx->stream));
    return BX_OK;
}

// Writes a built world into its ranges (bases already in hw.info) and its slot; every other
// per-world array of the world is cleared first, since a pool reuses ranges.
static int PlaceWorld(bx_batch* b, u32 slot, const HostWorld& hw) {
    ++b->stateGen;  // the slot's fat AABBs change
    const WorldInfo& wi = hw.info;
    BatchView& V = b->view;
    hipStream_t st = b->ctx->stream;
    u32 units[kSpaces], base[kSpaces];
    WorldUnits(wi, units);
    GetBases(wi, base);
    for (const ArraySpec& a : kArrays) {
        u64 bytes = u64(units[a.space]) * a.perUnit * a.elemSize;
        if (!bytes) continue;
        HIP_CHECK(hipMemsetAsync(ArrayPtr(V, a) + u64(base[a.space]) * a.perUnit * a.elemSize, 0, bytes, st));
    }
    int rc;
    if ((rc = PutAsync(b, V.bodies, wi.bodyBase, hw.bodies))) return rc;
    if ((rc = PutAsync(b, V.fixtures, wi.fixtureBase, hw.fixtures))) return rc;
    if ((rc = PutAsync(b, const_cast<u32*>(V.fixtureList), wi.fixtureListBase, hw.fixtureList))) return rc;
    if ((rc = PutAsync(b, const_cast<u32*>(V.proxyOrder), wi.fixtureBase, hw.proxyOrder))) return rc;
    if ((rc = PutAsync(b, V.fatAABB, wi.fixtureBase, hw.fat))) return rc;
    std::vector<u32> moved(wi.fixtureCount, 1u);  // CreateProxy buffers every new proxy (broad_phase.cpp:78-84)
    if ((rc = PutAsync(b, V.moved, wi.fixtureBase, moved))) return rc;
    if ((rc = PutAsync(b, V.joints, wi.jointBase, hw.joints))) return rc;
    if ((rc = PutAsync(b, const_cast<JointEdge*>(V.jointEdges), wi.jointEdgeBase, hw.jointEdges))) return rc;
    if ((rc = PutAsync(b, V.freeSlots, wi.contactBase, hw.freeSlots))) return rc;
    HIP_CHECK(hipMemcpyAsync((void*)(V.worlds + slot), &wi, sizeof(wi), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(V.wstate + slot, &hw.state, sizeof(hw.state), hipMemcpyHostToDevice, st));
    // the host sources are temporaries of the caller
    HIP_CHECK(hipStreamSynchronize(st));
    b->worlds[slot] = wi;
    b->proxyIds[slot] = hw.proxyIds;
    b->resident[slot] = 1;
    b->needsInitialPairs[slot] = 1;
    b->newContacts = true;
    return BX_OK;
}

static void InitSlots(bx_batch* b, u32 slots) {
    b->worldCount = slots;
    b->worlds.assign(slots, WorldInfo{});
    b->proxyIds.assign(slots, {});
    b->nodePools.assign(slots, NodePool{});
    b->resident.assign(slots, 0);
    b->needsInitialPairs.assign(slots, 0);
    b->bvhGen.assign(slots, 0);
}


static StepParams MakeParams(const bx_step_cfg* cfg) {
    StepParams sp;
    sp.dt = cfg->dt;
    sp.inv_dt = cfg->dt > scalar(0) ? scalar(1) / cfg->dt : scalar(0);
    sp.velocityIterations = cfg->velocity_iterations;
    sp.positionIterations = cfg->position_iterations;
    return sp;
}

// Takes the next region of the world-list ring (waiting for the launches that last read it).
static int NextListRegion(bx_batch* b, u32** host, u32** dev, u32* r) {
    *r = b->listNext;
    b->listNext = (b->listNext + 1) % kListRing;
    if (b->listUsed[*r]) HIP_CHECK_S(hipEventSynchronize(b->listEv[*r]));
    size_t off = size_t(*r) * 2 * std::max(1u, b->worldCount);
    *host = b->listHost + off;
    *dev = b->listDev + off;
    return BX_OK;
}

// The creation-time pair search of every listed world (every world when `worlds` is null) that has
// not had it yet.
static int RunInitialPairs(bx_batch* b, const u32* worlds, u32 count) {
    if (!b->newContacts) return BX_OK;
    u32 *h, *d, r;
    int rc = NextListRegion(b, &h, &d, &r);
    if (rc) return rc;
    u32 n = 0;
    if (!worlds) count = b->worldCount;
    for (u32 i = 0; i < count; ++i) {
        const u32 w = worlds ? worlds[i] : i;
        if (b->needsInitialPairs[w]) {
            h[n++] = w;
            b->needsInitialPairs[w] = 0;
        }
    }
    if (n) {
        HIP_CHECK_S(hipMemcpyAsync(d, h, n * sizeof(u32), hipMemcpyHostToDevice, b->ctx->stream));
        LaunchInitialPairs(b->view, d, n, b->ctx->stream);
        HIP_CHECK_S(hipGetLastError());
    }
    HIP_CHECK_S(hipEventRecord(b->listEv[r], b->ctx->stream));
    b->listUsed[r] = true;
    bool any = false;
    for (u32 w = 0; w < b->worldCount && !any; ++w) any = b->needsInitialPairs[w] != 0;
    b->newContacts = any;
    return BX_OK;
}

// One step of the listed worlds in one launch sequence (continuous batching: any subset of the
// resident worlds, each block of the one-workgroup-per-world kernels stepping one listed world).
static int StepList(bx_batch* b, const bx_step_cfg* cfg, const u32* worlds, u32 count) {
    ++b->stateGen;
    if (count == 0) return BX_OK;
    if (count > b->worldCount) return Fail(BX_E_INVALID, "more worlds listed than the batch holds");
    std::vector<u8> seen(b->worldCount, 0);
    for (u32 i = 0; i < count; ++i) {
        if (!IsResident(b, worlds[i])) return Fail(BX_E_INVALID, "listed world is not resident");
        if (seen[worlds[i]]) return Fail(BX_E_INVALID, "world listed twice");
        seen[worlds[i]] = 1;
    }
    int rc = RunInitialPairs(b, worlds, count);
    if (rc) return rc;
    u32 *h, *d, r;
    if ((rc = NextListRegion(b, &h, &d, &r))) return rc;
    u32 maxIslands = 0, maxBodies = 0, lds = 0;
    for (u32 i = 0; i < count; ++i) {
        const WorldInfo& wi = b->worlds[worlds[i]];
        h[i] = worlds[i];
        maxIslands += wi.bodyCount;
        maxBodies = std::max(maxBodies, wi.bodyCount);
        lds = std::max(lds, std::min(wi.islandScratchWords, kIslandLdsBudgetWords));
    }
    HIP_CHECK(hipMemcpyAsync(d, h, count * sizeof(u32), hipMemcpyHostToDevice, b->ctx->stream));
    BatchView v = b->view;
    v.islandLdsWords = lds;
    LaunchStep(v, d, 0, count, MakeParams(cfg), false, maxIslands, maxBodies, b->ctx->stream, nullptr);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(b->listEv[r], b->ctx->stream));
    b->listUsed[r] = true;
    return BX_OK;
}

// `steps` steps of every world of a (non-pool) batch, the first kMaxTimedSteps timed per kernel.
static int StepAll(bx_batch* b, const bx_step_cfg* cfg, u32 steps) {
    ++b->stateGen;
    const StepParams sp = MakeParams(cfg);
    u32 maxIslands = 0, maxBodies = 0;
    bool allNew = true;
    for (u32 i = 0; i < b->worldCount; ++i) {
        maxIslands += b->worlds[i].bodyCount;
        maxBodies = std::max(maxBodies, b->worlds[i].bodyCount);
        allNew = allNew && b->needsInitialPairs[i];
    }
    bool findFirst = false;
    if (b->newContacts) {
        if (allNew) {  // the usual case: the initial search runs inside the first (timed) step
            findFirst = true;
            std::fill(b->needsInitialPairs.begin(), b->needsInitialPairs.end(), 0);
            b->newContacts = false;
        } else {
            int rc = RunInitialPairs(b, nullptr, 0);
            if (rc) return rc;
        }
    }
    u32 timedSteps = std::min(steps, kMaxTimedSteps);
    int rc = EnsureEvents(b->ctx, timedSteps);
    if (rc) return rc;
    b->ctx->timedSteps = timedSteps;
    for (u32 s = 0; s < steps; ++s) {
        hipEvent_t* ev = s < timedSteps ? b->ctx->ev.data() + size_t(12) * s : nullptr;
        if (b->tuneStep >= kTuneSkip && b->tuneStep < kTuneSteps && ev && sp.dt > scalar(0)) {
            // solver_mode 0: this step runs and times one velocity-sweep layout of the tuning
            // sequence (events 5 and 6 bracket k_solve_velocity); after the last, the fastest is kept
            const u32 layout = kTuneOrder[b->tuneStep - kTuneSkip];
            BatchView v = b->view;
            v.solveLayout = layout;
            LaunchStep(v, nullptr, 0, b->worldCount, sp, findFirst, maxIslands, maxBodies, b->ctx->stream, ev);
            HIP_CHECK(hipGetLastError());
            HIP_CHECK(hipEventSynchronize(ev[6]));
            float ms = 0.0f;
            HIP_CHECK(hipEventElapsedTime(&ms, ev[5], ev[6]));
            const bool first = b->tuneStep - kTuneSkip < kSolveLayouts;
            b->tuneMs[layout] = first ? ms : std::min(b->tuneMs[layout], ms);
            if (++b->tuneStep == kTuneSteps) {
                // another layout replaces the first only when faster by more than kTuneMargin
                u32 best = 0;
                for (u32 l = 1; l < kSolveLayouts; ++l)
                    if (b->tuneMs[l] < b->tuneMs[best] && b->tuneMs[l] < b->tuneMs[0] * (1.0f - kTuneMargin)) best = l;
                b->view.solveLayout = best;
            }
        } else {
            if (b->tuneStep < kTuneSkip && sp.dt > scalar(0)) ++b->tuneStep;
            LaunchStep(b->view, nullptr, 0, b->worldCount, sp, findFirst, maxIslands, maxBodies, b->ctx->stream, ev);
        }
        findFirst = false;
    }
    HIP_CHECK(hipGetLastError());
    return BX_OK;
}

static void DropGraph(bx_batch* b) {
    if (b->graphExec) hipGraphExecDestroy(b->graphExec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->graphExec = nullptr;
    b->graph = nullptr;
}

// `steps` steps of every world of a (non-pool) batch as replays of one captured step (after the
// directly launched steps that still time velocity-sweep layouts under solver_mode 0). The graph
// holds the step's kernel launches with their arguments by value; the batch's device arrays are
// fixed at creation, so it stays valid for the batch's lifetime and is captured again only when
// the step configuration changes. The creation-time pair search runs before it, as a launch of its
// own. The step's kernels write no memory through copies or fills, so the graph has kernel nodes
// only.
static int StepGraph(bx_batch* b, const bx_step_cfg* cfg, u32 steps) {
    ++b->stateGen;
    const StepParams sp = MakeParams(cfg);
    if (b->tuneStep < kTuneSteps && sp.dt > scalar(0)) {  // solver_mode 0: time the layouts first
        const u32 k = std::min(steps, kTuneSteps - b->tuneStep);
        int rc = StepAll(b, cfg, k);
        if (rc || k == steps) return rc;
        steps -= k;
    }
    u32 maxIslands = 0, maxBodies = 0;
    for (u32 i = 0; i < b->worldCount; ++i) {
        maxIslands += b->worlds[i].bodyCount;
        maxBodies = std::max(maxBodies, b->worlds[i].bodyCount);
    }
    if (b->newContacts) {
        int rc = RunInitialPairs(b, nullptr, 0);
        if (rc) return rc;
    }
    const StepParams& g = b->graphParams;
    const bool same = b->graphExec && g.dt == sp.dt && g.inv_dt == sp.inv_dt &&
                      g.velocityIterations == sp.velocityIterations && g.positionIterations == sp.positionIterations &&
                      b->graphLayout == b->view.solveLayout;
    if (!same) {
        DropGraph(b);
        hipStream_t st = b->ctx->stream;
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        LaunchStep(b->view, nullptr, 0, b->worldCount, sp, false, maxIslands, maxBodies, st, nullptr);
        hipError_t launch = hipGetLastError();
        hipError_t e = hipStreamEndCapture(st, &b->graph);
        if (launch != hipSuccess || e != hipSuccess) {
            DropGraph(b);
            return Fail(BX_E_DEVICE, std::string("step capture: ") +
                                         hipGetErrorString(launch != hipSuccess ? launch : e));
        }
        e = hipGraphInstantiate(&b->graphExec, b->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            DropGraph(b);
            return Fail(BX_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        b->graphParams = sp;
        b->graphLayout = b->view.solveLayout;
    }
    for (u32 s = 0; s < steps; ++s) HIP_CHECK(hipGraphLaunch(b->graphExec, b->ctx->stream));
    b->ctx->timedSteps = 0;  // replayed steps carry no per-kernel events
    return BX_OK;
}

extern "C" {

const char* bx_version(void) { return "bxstep 0.1 gfx950"; }
const char* bx_last_error(void) { return g_errorStatic ? g_errorStatic : g_error.c_str(); }

int bx_context_create(int device, bx_context** out) {
    if (!out) return Fail(BX_E_INVALID, "null out");
    int n = 0;
    HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return Fail(BX_E_INVALID, "no such HIP device");
    HIP_CHECK(hipSetDevice(device));
    bx_context* ctx = new bx_context();
    ctx->device = device;
    HIP_CHECK(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    // this device's copy of the no-allocation scratch array (the array's name is a host-side stand-in)
    void* sym = nullptr;
    hipError_t se = hipGetSymbolAddress(&sym, HIP_SYMBOL(scratch));
    if (se != hipSuccess || !sym) {
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return Fail(BX_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se));
    }
    ctx->scratch = static_cast<unsigned char*>(sym);
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
    for (hipEvent_t e : b->listEv)
        if (e) hipEventDestroy(e);
    DropGraph(b);
    if (b->listHost) hipHostFree(b->listHost);
    if (b->queryDev) hipFree(b->queryDev);
    delete b->library;
    delete b;
}

// Uploads the carried state of a placed world (a topology edit): its live contacts in slots
// [0, live) in list order, their mesh records and triangle lists, and the last pair search's pairs.
static int PutCarried(bx_batch* b, const HostWorld& hw) {
    const WorldInfo& wi = hw.info;
    const u32 live = u32(hw.contacts.size());
    int rc;
    if ((rc = PutAsync(b, b->view.contacts, wi.contactBase, hw.contacts))) return rc;
    std::vector<u32> order(live);
    for (u32 i = 0; i < live; ++i) order[i] = i;
    if ((rc = PutAsync(b, b->view.contactOrder, wi.contactBase, order))) return rc;
    if (wi.meshTriCap) {
        if ((rc = PutAsync(b, b->view.meshContacts, wi.meshBase, hw.meshContacts))) return rc;
        if ((rc = PutAsync(b, b->view.meshTriLists, wi.meshTriBase, hw.meshTris))) return rc;
    }
    if ((rc = PutAsync(b, b->view.pairs, wi.pairBase, hw.pairs))) return rc;
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    return BX_OK;
}

// A batch of `world_count` worlds, world w produced by produce(w, hw) (built from a descriptor or
// taken from a resident world), on the shape library L (owned by the batch from here on).
static int CreateBatchFrom(bx_context* ctx, HostLibrary* L, const bx_batch_opts* opts, u32 world_count,
                           const std::function<int(u32, HostWorld&)>& produce, bx_batch** out) {
    HIP_CHECK(hipSetDevice(ctx->device));
    bx_batch* b = new bx_batch();
    b->ctx = ctx;
    b->library = L;
    if (opts) b->worldOpts = *opts;
    InitSlots(b, world_count);
    // the initialized arrays, concatenated in world order (one unit of padding per world in the
    // fixture space, whose ranges hold fixtureCount + 1 units)
    std::vector<Body> bodies;
    std::vector<Fixture> fixtures;
    std::vector<u32> fixtureList, proxyOrder, moved, freeSlots;
    std::vector<AABB> fat;
    std::vector<Joint> joints;
    std::vector<JointEdge> jointEdges;
    std::vector<WorldState> wstate(world_count);
    std::vector<HostWorld> carried;  // worlds with state to upload after the arrays exist
    u64 next[kSpaces] = {};
    u32 lds = 0;
    int rc;
    for (u32 w = 0; w < world_count; ++w) {
        HostWorld hw;
        rc = produce(w, hw);
        if (rc) { bx_batch_destroy(b); return rc; }
        u32 units[kSpaces], base[kSpaces];
        WorldUnits(hw.info, units);
        for (u32 s = 0; s < kSpaces; ++s) {
            base[s] = u32(next[s]);
            next[s] += units[s];
            if (next[s] > 0x7FFFFFFFull) { bx_batch_destroy(b); return Fail(BX_E_CAPACITY, "batch too large for 32-bit indices"); }
        }
        SetBases(hw.info, base);
        bodies.insert(bodies.end(), hw.bodies.begin(), hw.bodies.end());
        fixtures.insert(fixtures.end(), hw.fixtures.begin(), hw.fixtures.end());
        fixtures.push_back(Fixture{});
        fixtureList.insert(fixtureList.end(), hw.fixtureList.begin(), hw.fixtureList.end());
        fixtureList.push_back(0);
        proxyOrder.insert(proxyOrder.end(), hw.proxyOrder.begin(), hw.proxyOrder.end());
        proxyOrder.push_back(0);
        fat.insert(fat.end(), hw.fat.begin(), hw.fat.end());
        fat.push_back(AABB{});
        moved.insert(moved.end(), hw.moved.begin(), hw.moved.end());
        moved.push_back(0);
        joints.insert(joints.end(), hw.joints.begin(), hw.joints.end());
        jointEdges.insert(jointEdges.end(), hw.jointEdges.begin(), hw.jointEdges.end());
        freeSlots.insert(freeSlots.end(), hw.freeSlots.begin(), hw.freeSlots.end());
        freeSlots.insert(freeSlots.end(), hw.info.contactCap - hw.freeSlots.size(), 0u);  // live slots' units
        wstate[w] = hw.state;
        b->worlds[w] = hw.info;
        b->proxyIds[w] = std::move(hw.proxyIds);
        b->nodePools[w] = std::move(hw.nodes);
        b->resident[w] = 1;
        b->needsInitialPairs[w] = hw.needsInitialPairs ? 1 : 0;
        lds = std::max(lds, std::min(hw.info.islandScratchWords, kIslandLdsBudgetWords));
        if (hw.carried && (!hw.contacts.empty() || !hw.pairs.empty())) carried.push_back(std::move(hw));
    }
    for (u32 s = 0; s < kSpaces; ++s) b->spaces[s].Reset(0);
    BatchView& V = b->view;
    rc = AllocArrays(b, next, world_count);
    if (!rc) rc = UploadLibrary(b, *b->library);
    if (rc) { bx_batch_destroy(b); return rc; }
    V.islandLdsWords = lds;
    V.meshWorlds = 0;
    for (u32 w = 0; w < world_count; ++w) V.meshWorlds += b->worlds[w].meshTriCap ? 1u : 0u;
    // velocity-sweep layout: fixed by solver_mode 1..5, else timed on the first steps (StepAll)
    const u32 mode = opts ? opts->solver_mode : 0u;
    if (mode > kSolveLayouts) {
        bx_batch_destroy(b);
        return Fail(BX_E_INVALID, "solver_mode must be 0 (tuned) or 1..5");
    }
    V.solveLayout = mode ? mode - 1 : 0u;
    b->tuneStep = mode ? kTuneSteps : 0u;
#define PUT(dev, host)                                                                                      \
    if (!host.empty()) {                                                                                    \
        hipError_t _e = hipMemcpy((void*)(dev), host.data(), host.size() * sizeof(host[0]), hipMemcpyHostToDevice); \
        if (_e != hipSuccess) { bx_batch_destroy(b); return Fail(BX_E_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(_e)); } \
    }
    PUT(V.bodies, bodies);
    PUT(V.fixtures, fixtures);
    PUT(V.fixtureList, fixtureList);
    PUT(V.proxyOrder, proxyOrder);
    PUT(V.fatAABB, fat);
    PUT(V.moved, moved);
    PUT(V.joints, joints);
    PUT(V.jointEdges, jointEdges);
    PUT(V.freeSlots, freeSlots);
    PUT(V.worlds, b->worlds);
    PUT(V.wstate, wstate);
#undef PUT
    for (const HostWorld& hw : carried) {  // hw.info holds the world's bases
        rc = PutCarried(b, hw);
        if (rc) { bx_batch_destroy(b); return rc; }
    }
    b->newContacts = false;
    for (u32 w = 0; w < world_count; ++w) b->newContacts = b->newContacts || b->needsInitialPairs[w];
    *out = b;
    return BX_OK;
}

int bx_batch_create(bx_context* ctx, const bx_shape_library* library, const bx_world_desc* worlds,
                    uint32_t world_count, const bx_batch_opts* opts, bx_batch** out) {
    if (!ctx || !out || (!worlds && world_count)) return Fail(BX_E_INVALID, "null argument");
    HostLibrary* L = new HostLibrary();
    int rc = BuildLibrary(library, *L);
    if (rc) { delete L; return rc; }
    return CreateBatchFrom(ctx, L, opts, world_count,
                           [&](u32 w, HostWorld& hw) { return BuildWorld(worlds[w], *L, opts, hw); }, out);
}

int bx_pool_create(bx_context* ctx, const bx_shape_library* library, const bx_pool_opts* opts, bx_batch** out) {
    if (!ctx || !out || !opts) return Fail(BX_E_INVALID, "null argument");
    if (opts->hbm_budget_bytes == 0) return Fail(BX_E_INVALID, "zero memory budget");
    HIP_CHECK(hipSetDevice(ctx->device));
    const u32 slots = opts->max_worlds ? opts->max_worlds : 4096u;
    // Split the budget over the spaces in the proportions of worlds of one fixture per body, about
    // 200 bodies, the default contact capacity and the given joint density.
    const double fixturesPerBody = 1.0 + 1.0 / 200.0;
    const double jointsPerBody = (opts->joints_per_1000_bodies ? opts->joints_per_1000_bodies : 50u) / 1000.0;
    const double contactsPerBody = opts->max_contacts_per_world ? opts->max_contacts_per_world / 200.0
                                                                : double(kContactsPerFixture) * fixturesPerBody + double(kContactSlack) / 200.0;
    const double pairsPerBody = opts->max_pairs_per_world ? opts->max_pairs_per_world / 200.0 : 2.0 * contactsPerBody;
    double per[kSpaces];
    per[kSpaceBody] = 1.0;
    per[kSpaceFixture] = fixturesPerBody;
    per[kSpaceJoint] = jointsPerBody;
    per[kSpaceContact] = contactsPerBody;
    per[kSpacePair] = pairsPerBody;
    per[kSpaceIslandBody] = 1.0 + contactsPerBody;
    per[kSpaceScratch] = 5.0 + 4.0 * fixturesPerBody + 5.0 * contactsPerBody;
    per[kSpaceCon] = contactsPerBody;
    per[kSpaceMesh] = per[kSpaceMeshTri] = per[kSpaceMeshScratch] = 0.0;  // pools hold no mesh worlds
    u64 ub[kSpaces];
    UnitBytes(ub);
    double bytesPerBody = 0;
    for (u32 s = 0; s < kSpaces; ++s) bytesPerBody += per[s] * double(ub[s]);
    double fixed = double(slots) * (sizeof(WorldInfo) + sizeof(WorldState) + 8.0 * kListRing);
    double bodiesCap = (double(opts->hbm_budget_bytes) - fixed) / bytesPerBody;
    // every index of every array stays below 2^31
    for (u32 s = 0; s < kSpaces; ++s) bodiesCap = std::min(bodiesCap, double(0x7FFFFFFF / 2) / per[s]);
    if (bodiesCap < 64.0) return Fail(BX_E_INVALID, "memory budget too small");
    u64 cap[kSpaces];
    for (u32 s = 0; s < kSpaces; ++s) cap[s] = u64(bodiesCap * per[s]) + 1;
    bx_batch* b = new bx_batch();
    b->ctx = ctx;
    b->pool = true;
    b->library = new HostLibrary();
    int rc = BuildLibrary(library, *b->library);
    if (rc) { bx_batch_destroy(b); return rc; }
    b->worldOpts.max_contacts_per_world = opts->max_contacts_per_world;
    b->worldOpts.max_pairs_per_world = opts->max_pairs_per_world;
    InitSlots(b, slots);
    rc = AllocArrays(b, cap, slots);
    if (!rc) rc = UploadLibrary(b, *b->library);
    if (rc) { bx_batch_destroy(b); return rc; }
    for (u32 s = 0; s < kSpaces; ++s) b->spaces[s].Reset(cap[s]);
    b->freeWorldSlots.resize(slots);
    for (u32 i = 0; i < slots; ++i) b->freeWorldSlots[i] = slots - 1 - i;  // lowest slot first
    b->view.islandLdsWords = kIslandLdsBudgetWords;
    b->newContacts = false;
    if (opts->solver_mode > kSolveLayouts) {
        bx_batch_destroy(b);
        return Fail(BX_E_INVALID, "solver_mode must be 0 (16 islands per wavefront) or 1..5");
    }
    b->view.solveLayout = opts->solver_mode ? opts->solver_mode - 1 : 0u;
    *out = b;
    return BX_OK;
}

// Units a world descriptor needs in each space (the sizes BuildWorld gives it).
static void DescUnits(const bx_world_desc& d, const bx_batch_opts* o, u32 units[kSpaces]) {
    WorldInfo wi;
    std::memset(&wi, 0, sizeof(wi));
    wi.bodyCount = d.body_count;
    wi.fixtureCount = d.fixture_count;
    wi.jointCount = d.joint_count;
    SetCapacities(wi, d.fixtures && HasMesh(d), o);
    WorldUnits(wi, units);
}

// Takes a free slot and ranges for a world of the given units; BX_E_CAPACITY when they are not free.
static int Reserve(bx_batch* b, const u32 units[kSpaces], u32 base[kSpaces], u32* slot) {
    if (b->freeWorldSlots.empty()) return Fail(BX_E_CAPACITY, "no free world slot in the pool");
    u64 got[kSpaces];
    for (u32 s = 0; s < kSpaces; ++s) {
        if (!b->spaces[s].Alloc(units[s], got[s])) {
            for (u32 t = 0; t < s; ++t) b->spaces[t].Free(got[t], units[t]);
            return Fail(BX_E_CAPACITY, "pool memory budget exhausted");
        }
        base[s] = u32(got[s]);
    }
    *slot = b->freeWorldSlots.back();
    b->freeWorldSlots.pop_back();
    return BX_OK;
}

int bx_pool_admit(bx_batch* b, const bx_world_desc* world, uint32_t* slot) {
    if (!b || !world || !slot) return Fail(BX_E_INVALID, "null argument");
    if (!b->pool) return Fail(BX_E_INVALID, "not a pool");
    if (world->fixture_count && !world->fixtures) return Fail(BX_E_INVALID, "null fixture array");
    if (HasMesh(*world)) return Fail(BX_E_UNSUPPORTED, "pools do not hold worlds with mesh fixtures");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    // reserve first (from the counts alone), so a world that does not fit costs no host build
    u32 units[kSpaces], base[kSpaces], s;
    DescUnits(*world, &b->worldOpts, units);
    int rc = Reserve(b, units, base, &s);
    if (rc) return rc;
    HostWorld hw;
    rc = BuildWorld(*world, *b->library, &b->worldOpts, hw);
    if (!rc) {
        u32 built[kSpaces];
        WorldUnits(hw.info, built);
        if (std::memcmp(built, units, sizeof(units)) != 0) rc = Fail(BX_E_INVALID, "world sizes changed while building");
    }
    if (!rc) {
        SetBases(hw.info, base);
        rc = PlaceWorld(b, s, hw);
    }
    if (rc) {
        for (u32 k = 0; k < kSpaces; ++k) b->spaces[k].Free(base[k], units[k]);
        b->freeWorldSlots.push_back(s);
        b->resident[s] = 0;
        return rc;
    }
    *slot = s;
    return BX_OK;
}

int bx_pool_release(bx_batch* b, uint32_t slot) {
    if (!b || !b->pool) return Fail(BX_E_INVALID, "not a pool");
    if (!IsResident(b, slot)) return Fail(BX_E_INVALID, "slot holds no world");
    u32 units[kSpaces], base[kSpaces];
    WorldUnits(b->worlds[slot], units);
    GetBases(b->worlds[slot], base);
    for (u32 s = 0; s < kSpaces; ++s) b->spaces[s].Free(base[s], units[s]);
    b->resident[slot] = 0;
    b->needsInitialPairs[slot] = 0;
    b->proxyIds[slot].clear();
    b->freeWorldSlots.push_back(slot);
    return BX_OK;
}

uint64_t bx_pool_world_bytes(bx_batch* b, const bx_world_desc* d) {
    if (!d) return 0;
    u32 units[kSpaces];
    DescUnits(*d, b ? &b->worldOpts : nullptr, units);
    u64 ub[kSpaces];
    UnitBytes(ub);
    u64 bytes = 0;
    for (u32 s = 0; s < kSpaces; ++s) bytes += units[s] * ub[s];
    return bytes;
}

int bx_pool_get_stats(bx_batch* b, bx_pool_stats* out) {
    if (!b || !out || !b->pool) return Fail(BX_E_INVALID, "not a pool");
    std::memset(out, 0, sizeof(*out));
    u64 ub[kSpaces];
    UnitBytes(ub);
    for (u32 s = 0; s < kSpaces; ++s) {
        out->capacity_bytes += b->spaces[s].capacity * ub[s];
        out->used_bytes += (b->spaces[s].capacity - b->spaces[s].FreeUnits()) * ub[s];
    }
    for (u32 w = 0; w < b->worldCount; ++w) out->resident_worlds += b->resident[w];
    out->free_slots = u32(b->freeWorldSlots.size());
    out->world_slots = b->worldCount;
    return BX_OK;
}

// Snapshot layout: header, proxy ids, then every per-world array's range in kArrays order.
struct SnapshotHeader {
    u32 magic, version, needsInitialPairs, proxyCount;
    WorldInfo info;
    WorldState state;
};
static const u32 kSnapshotMagic = 0x53584231u;  // "1BXS"

static u64 SnapshotBytes(const WorldInfo& wi) {
    u32 units[kSpaces];
    WorldUnits(wi, units);
    u64 bytes = sizeof(SnapshotHeader) + u64(wi.fixtureCount) * sizeof(u32);
    for (const ArraySpec& a : kArrays) bytes += u64(units[a.space]) * a.perUnit * a.elemSize;
    return bytes;
}

int bx_batch_snapshot(bx_batch* b, uint32_t world, void* out, uint64_t capacity, uint64_t* size) {
    if (!b || !size) return Fail(BX_E_INVALID, "null argument");
    if (!IsResident(b, world)) return Fail(BX_E_INVALID, "world not resident");
    const WorldInfo& wi = b->worlds[world];
    *size = SnapshotBytes(wi);
    if (!out) return BX_OK;
    if (capacity < *size) return Fail(BX_E_INVALID, "snapshot buffer too small");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    SnapshotHeader h;
    std::memset(&h, 0, sizeof(h));
    h.magic = kSnapshotMagic;
    h.version = 2;
    h.needsInitialPairs = b->needsInitialPairs[world];
    h.proxyCount = wi.fixtureCount;
    h.info = wi;
    HIP_CHECK(hipMemcpyAsync(&h.state, b->view.wstate + world, sizeof(WorldState), hipMemcpyDeviceToHost, st));
    char* p = static_cast<char*>(out) + sizeof(h) + u64(wi.fixtureCount) * sizeof(u32);
    u32 units[kSpaces], base[kSpaces];
    WorldUnits(wi, units);
    GetBases(wi, base);
    for (const ArraySpec& a : kArrays) {
        u64 bytes = u64(units[a.space]) * a.perUnit * a.elemSize;
        if (bytes)
            HIP_CHECK(hipMemcpyAsync(p, ArrayPtr(b->view, a) + u64(base[a.space]) * a.perUnit * a.elemSize, bytes,
                                     hipMemcpyDeviceToHost, st));
        p += bytes;
    }
    HIP_CHECK(hipStreamSynchronize(st));
    std::memcpy(out, &h, sizeof(h));
    if (wi.fixtureCount)
        std::memcpy(static_cast<char*>(out) + sizeof(h), b->proxyIds[world].data(), wi.fixtureCount * sizeof(u32));
    return BX_OK;
}

int bx_pool_restore(bx_batch* b, const void* snapshot, uint64_t size, uint32_t* slot) {
    if (b) ++b->stateGen;  // a restored world's fat AABBs replace the slot's
    if (!b || !snapshot || !slot) return Fail(BX_E_INVALID, "null argument");
    if (!b->pool) return Fail(BX_E_INVALID, "not a pool");
    SnapshotHeader h;
    if (size < sizeof(h)) return Fail(BX_E_INVALID, "snapshot too short");
    std::memcpy(&h, snapshot, sizeof(h));
    if (h.magic != kSnapshotMagic || h.version != 2 || h.proxyCount != h.info.fixtureCount ||
        size != SnapshotBytes(h.info))
        return Fail(BX_E_INVALID, "not a world snapshot of this library version");
    if (h.info.meshTriCap) return Fail(BX_E_UNSUPPORTED, "pools do not hold worlds with mesh fixtures");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    hipStream_t st = b->ctx->stream;
    WorldInfo wi = h.info;
#if BX_DEVICE_CALLBACKS
    wi.filter = nullptr;  // contact filter tables belong to the batch that set them (pools hold none)
#endif
    u32 units[kSpaces], base[kSpaces], s;
    WorldUnits(wi, units);
    int rc = Reserve(b, units, base, &s);
    if (rc) return rc;
    SetBases(wi, base);
    const char* p = static_cast<const char*>(snapshot) + sizeof(h) + u64(wi.fixtureCount) * sizeof(u32);
    for (const ArraySpec& a : kArrays) {
        u64 bytes = u64(units[a.space]) * a.perUnit * a.elemSize;
        hipError_t e = bytes ? hipMemcpyAsync(ArrayPtr(b->view, a) + u64(base[a.space]) * a.perUnit * a.elemSize, p,
                                              bytes, hipMemcpyHostToDevice, st)
                             : hipSuccess;
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) {
            for (u32 k = 0; k < kSpaces; ++k) b->spaces[k].Free(base[k], units[k]);
            b->freeWorldSlots.push_back(s);
            return Fail(BX_E_DEVICE, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
        }
        p += bytes;
    }
    HIP_CHECK(hipMemcpyAsync((void*)(b->view.worlds + s), &wi, sizeof(wi), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(b->view.wstate + s, &h.state, sizeof(h.state), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    b->worlds[s] = wi;
    const u32* ids = reinterpret_cast<const u32*>(static_cast<const char*>(snapshot) + sizeof(h));
    b->proxyIds[s].assign(ids, ids + wi.fixtureCount);
    b->resident[s] = 1;
    b->needsInitialPairs[s] = h.needsInitialPairs ? 1 : 0;
    if (h.needsInitialPairs) b->newContacts = true;
    *slot = s;
    return BX_OK;
}

int bx_batch_step(bx_batch* b, const bx_step_cfg* cfg, uint32_t step_count) {
    if (!b || !cfg) return Fail(BX_E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    if (step_count == 0) return BX_OK;
    if (!b->pool) return b->worldCount ? StepAll(b, cfg, step_count) : BX_OK;
    std::vector<u32> list;
    for (u32 w = 0; w < b->worldCount; ++w)
        if (b->resident[w]) list.push_back(w);
    for (u32 s = 0; s < step_count; ++s) {
        int rc = StepList(b, cfg, list.data(), u32(list.size()));
        if (rc) return rc;
    }
    return BX_OK;
}

int bx_batch_step_noalloc(bx_batch* b, const bx_step_cfg* cfg, uint32_t step_count) {
    if (!b || !cfg) return FailStatic(BX_E_INVALID, "null argument");
    if (b->pool) return FailStatic(BX_E_UNSUPPORTED, "bx_batch_step_noalloc steps a batch, not a pool");
    if (!b->ctx->scratch) return FailStatic(BX_E_DEVICE, "bx_batch_step_noalloc: no scratch array on this device");
    HIP_CHECK_S(hipSetDevice(b->ctx->device));
    if (step_count == 0 || b->worldCount == 0) return BX_OK;
    u64 bodies = 0;
    u32 maxBodies = 0;
    bool allNew = true;
    for (u32 w = 0; w < b->worldCount; ++w) {
        bodies += b->worlds[w].bodyCount;
        maxBodies = std::max(maxBodies, b->worlds[w].bodyCount);
        allNew = allNew && b->needsInitialPairs[w];
    }
    if (bodies > kNoAllocIslands)
        return FailStatic(BX_E_CAPACITY, "bx_batch_step_noalloc: the batch holds more than 2,097,152 bodies, the "
                                         "island work list of the no-allocation scratch array");
    bool findFirst = false;
    if (b->newContacts) {
        if (allNew) {  // the creation-time pair search runs as the first step's first kernel
            findFirst = true;
            for (u32 w = 0; w < b->worldCount; ++w) b->needsInitialPairs[w] = 0;
            b->newContacts = false;
        } else {
            int rc = RunInitialPairs(b, nullptr, 0);
            if (rc) return rc;
        }
    }
    ++b->stateGen;
    BatchView v = b->view;
    v.islandList = reinterpret_cast<uint2*>(b->ctx->scratch);
    v.islandListCount = reinterpret_cast<u32*>(b->ctx->scratch + kNoAllocListBytes);  // reset by the contact update
    const StepParams sp = MakeParams(cfg);
    for (u32 s = 0; s < step_count; ++s) {
        LaunchStep(v, nullptr, 0, b->worldCount, sp, findFirst, u32(bodies), maxBodies, b->ctx->stream, nullptr);
        findFirst = false;
    }
    HIP_CHECK_S(hipGetLastError());
    b->ctx->timedSteps = 0;
    return BX_OK;
}

int bx_batch_step_graph(bx_batch* b, const bx_step_cfg* cfg, uint32_t step_count) {
    if (!b || !cfg) return Fail(BX_E_INVALID, "null argument");
    if (b->pool) return Fail(BX_E_UNSUPPORTED, "graph replay steps a batch, not a pool");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    if (step_count == 0 || b->worldCount == 0) return BX_OK;
    return StepGraph(b, cfg, step_count);
}

int bx_batch_get_solver_layout(bx_batch* b, uint32_t* islands_per_wavefront, float* tuned_ms, uint32_t count) {
    if (!b || !islands_per_wavefront || (!tuned_ms && count)) return Fail(BX_E_INVALID, "null argument");
    *islands_per_wavefront = kSolveLayoutLanes[std::min(b->view.solveLayout, kSolveLayouts - 1)];
    for (u32 i = 0; i < count; ++i) tuned_ms[i] = i < kSolveLayouts ? b->tuneMs[i] : 0.0f;
    return int(kSolveLayouts);
}

int bx_batch_step_list(bx_batch* b, const bx_step_cfg* cfg, const uint32_t* worlds, uint32_t count) {
    if (!b || !cfg || (!worlds && count)) return Fail(BX_E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    return StepList(b, cfg, worlds, count);
}

int bx_batch_step_subset(bx_batch* b, const bx_step_cfg* cfg, const uint32_t* worlds, const uint32_t* step_counts,
                         uint32_t count) {
    if (!b || !cfg || (!worlds && count)) return Fail(BX_E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    u32 most = 0;
    for (u32 i = 0; i < count; ++i) most = std::max(most, step_counts ? step_counts[i] : 1u);
    // launch s steps every world with more than s steps to go
    std::vector<u32> list;
    for (u32 s = 0; s < most; ++s) {
        list.clear();
        for (u32 i = 0; i < count; ++i)
            if ((step_counts ? step_counts[i] : 1u) > s) list.push_back(worlds[i]);
        int rc = StepList(b, cfg, list.data(), u32(list.size()));
        if (rc) return rc;
    }
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

int bx_batch_get_totals(bx_batch* b, bx_batch_totals* out) {
    if (!b || !out) return Fail(BX_E_INVALID, "null argument");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    std::memset(out, 0, sizeof(*out));
    std::vector<WorldState> ws;
    int rc = Download(b, ws, b->view.wstate, b->worldCount);
    if (rc) return rc;
    // the body ranges of the resident worlds, read in one copy
    u64 lo = ~u64(0), hi = 0;
    for (u32 w = 0; w < b->worldCount; ++w) {
        if (!b->resident[w]) continue;
        lo = std::min<u64>(lo, b->worlds[w].bodyBase);
        hi = std::max<u64>(hi, u64(b->worlds[w].bodyBase) + b->worlds[w].bodyCount);
    }
    std::vector<Body> bodies;
    if (hi > lo && (rc = Download(b, bodies, b->view.bodies + lo, size_t(hi - lo)))) return rc;
    for (u32 w = 0; w < b->worldCount; ++w) {
        if (!b->resident[w]) continue;
        const WorldInfo& wi = b->worlds[w];
        const WorldState& s = ws[w];
        ++out->worlds;
        out->bodies += wi.bodyCount;
        out->fixtures += wi.fixtureCount;
        out->joints += wi.jointCount;
        out->contacts += s.contactCount;
        if (wi.meshTriCap) out->mesh_world_contacts += s.contactCount;
        out->pairs += s.pairCount;
        out->islands += s.islandCount;
        out->island_contacts += std::min(s.islandContactCount, wi.conCap);
        out->status_or |= s.status;
        out->status_worlds += s.status ? 1u : 0u;
        for (u32 i = 0; i < wi.bodyCount; ++i) {
            const Body& x = bodies[wi.bodyBase - lo + i];
            out->awake_bodies += (x.type != kStatic && (x.flags & kAwake)) ? 1u : 0u;
        }
    }
    return BX_OK;
}

int bx_batch_set_contact_filter(bx_batch* b, uint32_t world, const uint8_t* collide, const uint8_t* respond) {
    if (!b || !IsResident(b, world)) return Fail(BX_E_INVALID, "bad batch or world not resident");
    if (b->pool) return Fail(BX_E_UNSUPPORTED, "contact filter tables are set on batches, not pools");
    if ((collide == nullptr) != (respond == nullptr)) return Fail(BX_E_INVALID, "set both filter tables or neither");
#if !BX_DEVICE_CALLBACKS
    return Fail(BX_E_UNSUPPORTED, "contact filter tables are not enabled in this build");
#else
    HIP_CHECK(hipSetDevice(b->ctx->device));
    WorldInfo& wi = b->worlds[world];
    if (!collide) {
        wi.filter = nullptr;
    } else {
        if (wi.fixtureCount > 16384) return Fail(BX_E_CAPACITY, "contact filter tables hold at most 16,384 fixtures");
        const u32 nf = wi.fixtureCount, words = FilterWords(nf);
        std::vector<u32> bits(size_t(2) * std::max(1u, words), 0u);
        for (u64 k = 0; k < u64(nf) * nf; ++k) {
            if (collide[k >> 3] & (1u << (k & 7))) bits[k >> 5] |= 1u << (k & 31);
            if (respond[k >> 3] & (1u << (k & 7))) bits[words + (k >> 5)] |= 1u << (k & 31);
        }
        u32* dev = nullptr;
        int rc = DevAlloc(b, &dev, bits.size(), bits.data());
        if (rc) return rc;
        wi.filter = dev;
    }
    HIP_CHECK(hipMemcpy((void*)(b->view.worlds + world), &wi, sizeof(wi), hipMemcpyHostToDevice));
    return BX_OK;
#endif
}

int bx_batch_get_events(bx_batch* b, uint32_t world, bx_contact_event* out, uint32_t capacity, uint32_t* count) {
#if !BX_DEVICE_CALLBACKS
    (void)out, (void)capacity, (void)count;
    if (!b || !IsResident(b, world)) return Fail(BX_E_INVALID, "bad batch or world not resident");
    return Fail(BX_E_UNSUPPORTED, "contact event recording is not enabled in this build");
#else
    WorldState ws;
    int rc = GetState(b, world, ws);
    if (rc) return rc;
    const WorldInfo& wi = b->worlds[world];
    if ((wi.flags & kWorldEvents) == 0) return Fail(BX_E_INVALID, "the world does not record contact events");
    std::vector<u32> order;
    std::vector<uint4> visits, entries;
    if ((rc = Download(b, order, b->view.contactOrder + wi.contactBase, ws.contactCount))) return rc;
    if ((rc = Download(b, visits, b->view.events + wi.contactBase, std::min(ws.eventCount, wi.contactCap)))) return rc;
    if ((rc = Download(b, entries, b->view.islandContacts + wi.conBase, std::min(ws.islandContactCount, wi.conCap))))
        return rc;
    std::vector<u32> listIndex(wi.contactCap, BX_NULL_INDEX);  // slot -> position in the contact list
    for (u32 k = 0; k < ws.contactCount; ++k) listIndex[order[k]] = ws.contactCount - 1 - k;
    u32 total = 0;
    auto emit = [&](u32 type, u32 fa, u32 fb, u32 contact) {
        if (total < capacity) out[total] = bx_contact_event{type, fa, fb, contact};
        ++total;
    };
    // b3ContactManager::UpdateContacts visits the world list, newest contact first
    for (u32 k = u32(visits.size()); k-- > 0;) {
        const uint4 v = visits[k];
        const u32 contact = (v.w & kEvDestroyed) ? BX_NULL_INDEX : listIndex[v.x];
        if (v.w & kEvBegin) emit(BX_EVENT_BEGIN_CONTACT, v.y, v.z, contact);
        if (v.w & kEvEnd) emit(BX_EVENT_END_CONTACT, v.y, v.z, contact);
        if (v.w & kEvPreSolve) emit(BX_EVENT_PRE_SOLVE, v.y, v.z, contact);
    }
    if (ws.islandContactCount) {  // b3Island::Report: every solved contact, island by island
        std::vector<Contact> contacts;
        if ((rc = Download(b, contacts, b->view.contacts + wi.contactBase, wi.contactCap))) return rc;
        for (const uint4& e : entries) {
            if (e.x >> 30) continue;  // one PostSolve per contact (a mesh contact has one entry per manifold)
            const Contact& c = contacts[e.x & 0x3FFFFFFFu];
            emit(BX_EVENT_POST_SOLVE, c.fixtureA, c.fixtureB, listIndex[e.x & 0x3FFFFFFFu]);
        }
    }
    if (count) *count = total;
    return BX_OK;
#endif
}

int bx_batch_get_bodies(bx_batch* b, uint32_t world, bx_body_state* out, uint32_t capacity) {
    if (!b || !IsResident(b, world) || !out) return Fail(BX_E_INVALID, "bad argument or world not resident");
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
    std::vector<MeshContact> mesh;
    if (wi.meshTriCap) {
        rc = Download(b, mesh, b->view.meshContacts + wi.meshBase, wi.contactCap);
        if (rc) return rc;
    }
    u32 n = std::min<u32>(capacity, ws.contactCount);
    for (u32 i = 0; i < n; ++i) {
        const u32 slot = order[ws.contactCount - 1 - i];  // world list: newest first
        const Contact& c = contacts[slot];
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
        for (u32 j = 0; j < std::min(c.manifoldCount, kMaxManifolds); ++j) {
            const Manifold& m = j == 0 ? c.manifold : mesh[slot].extra[j - 1];
            bx_manifold_state& ms = s.manifolds[j];
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
    // entries of fixtures destroyed since the pair search hold the proxy id itself (kPairProxyBit)
    auto proxy = [&](u32 x) { return (x & kPairProxyBit) ? (x & ~kPairProxyBit) : pid[x]; };
    for (u32 i = 0; i < n; ++i) {
        out[i].proxy1 = proxy(pairs[i].x);
        out[i].proxy2 = proxy(pairs[i].y);
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
    std::vector<uint4> entries;
    rc = Download(b, entries, b->view.islandContacts + wi.conBase, std::min(ws.islandContactCount, wi.conCap));
    if (rc) return rc;
    std::vector<u32> listIndex(wi.contactCap, BX_NULL_INDEX);
    for (u32 k = 0; k < ws.contactCount; ++k) listIndex[order[k]] = ws.contactCount - 1 - k;
    u32 total = 0;
    for (const uint4& e : entries) {  // one entry per manifold: the contact is listed at its first
        if (e.x >> 30) continue;
        if (total < capacity) out[total] = listIndex[e.x];
        ++total;
    }
    if (count) *count = total;
    return BX_OK;
}

// b3Body::SetAwake (body.h:588-607) on a body record.
static void HostSetAwake(Body& x, bool flag) {
    if (flag) {
        if ((x.flags & kAwake) == 0) {
            x.flags |= kAwake;
            x.sleepTime = 0;
        }
    } else {
        x.flags &= ~kAwake;
        x.sleepTime = 0;
        x.force.SetZero();
        x.torque.SetZero();
        x.v.SetZero();
        x.w.SetZero();
    }
}

// b3Body::SetLinearVelocity / SetAngularVelocity (body.h:652-685): ignored for a static body; a
// nonzero velocity wakes the body.
static void HostSetVelocity(Body& x, Vec3& dst, const Vec3& v) {
    if (x.type == kStatic) return;
    if (Dot(v, v) > scalar(0)) HostSetAwake(x, true);
    dst = v;
}

static int UploadBodies(bx_batch* b, const WorldInfo& wi, const std::vector<Body>& bodies) {
    HIP_CHECK(hipMemcpyAsync(b->view.bodies + wi.bodyBase, bodies.data(), bodies.size() * sizeof(Body),
                             hipMemcpyHostToDevice, b->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    return BX_OK;
}

int bx_batch_get_mass(bx_batch* b, uint32_t world, bx_body_mass* out, uint32_t capacity) {
    if (!b || !IsResident(b, world) || (!out && capacity)) return Fail(BX_E_INVALID, "bad argument or world not resident");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    const WorldInfo& wi = b->worlds[world];
    std::vector<Body> bodies;
    int rc = Download(b, bodies, b->view.bodies + wi.bodyBase, wi.bodyCount);
    if (rc) return rc;
    const u32 n = std::min<u32>(capacity, wi.bodyCount);
    auto put = [](float* o, const Mat33& m) {
        const Vec3 c[3] = {m.x, m.y, m.z};
        for (int k = 0; k < 3; ++k) {
            o[3 * k + 0] = c[k].x;
            o[3 * k + 1] = c[k].y;
            o[3 * k + 2] = c[k].z;
        }
    };
    for (u32 i = 0; i < n; ++i) {
        const Body& x = bodies[i];
        out[i].mass = x.mass;
        out[i].inv_mass = x.invMass;
        out[i].local_center[0] = x.localCenter.x;
        out[i].local_center[1] = x.localCenter.y;
        out[i].local_center[2] = x.localCenter.z;
        put(out[i].inertia, x.I);
        put(out[i].inv_inertia, x.invI);
        put(out[i].world_inv_inertia, x.worldInvI);
    }
    return int(n);
}

int bx_batch_set_bodies(bx_batch* b, uint32_t world, const bx_body_state* in, uint32_t count) {
    if (!b || !IsResident(b, world) || !in) return Fail(BX_E_INVALID, "bad argument or world not resident");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    const WorldInfo& wi = b->worlds[world];
    if (count != wi.bodyCount) return Fail(BX_E_INVALID, "body count mismatch");
    std::vector<Body> bodies;
    int rc = Download(b, bodies, b->view.bodies + wi.bodyBase, wi.bodyCount);
    if (rc) return rc;
    for (u32 i = 0; i < count; ++i) {
        Body& x = bodies[i];
        HostSetVelocity(x, x.v, V3(in[i].linear_velocity));
        HostSetVelocity(x, x.w, V3(in[i].angular_velocity));
    }
    return UploadBodies(b, wi, bodies);
}

// ------------------------------------------------------------------ world queries
// Builds the query hierarchy (k_build_bvh) of every listed world whose hierarchy predates the
// batch's current state.
static int EnsureBvh(bx_batch* b, const std::vector<u32>& worlds) {
    std::vector<u32> need;
    for (u32 w : worlds) {
        if (b->bvhGen[w] == b->stateGen) continue;
        b->bvhGen[w] = b->stateGen;  // also marks a world listed twice
        need.push_back(w);
    }
    if (need.empty()) return BX_OK;
    u32 *h, *d, r;
    int rc = NextListRegion(b, &h, &d, &r);
    if (rc) return rc;
    std::memcpy(h, need.data(), need.size() * sizeof(u32));
    HIP_CHECK(hipMemcpyAsync(d, h, need.size() * sizeof(u32), hipMemcpyHostToDevice, b->ctx->stream));
    LaunchBuildBvh(b->view, d, u32(need.size()), b->ctx->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(b->listEv[r], b->ctx->stream));
    b->listUsed[r] = true;
    return BX_OK;
}

static size_t Align256(size_t x) { return (x + 255) & ~size_t(255); }

// Runs `count` queries of one kind (0: single ray, 1: every ray hit, 2: boxes) given as world, mask
// offset and six coordinates each; outputs go to hits (kinds 0, 1) or fixtures (kind 2) and counts.
// The batch's query buffer grown to at least `total` bytes (contents not kept).
static int QueryScratch(bx_batch* b, size_t total) {
    if (total <= b->queryBytes) return BX_OK;
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    if (b->queryDev) HIP_CHECK(hipFree(b->queryDev));
    b->queryDev = nullptr;
    b->queryBytes = 0;
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, total));
    b->queryDev = static_cast<char*>(p);
    b->queryBytes = total;
    return BX_OK;
}

static int RunQueries(bx_batch* b, u32 kind, const u32* world, const u32* maskOff, const float* coords, u32 count,
                      const uint8_t* maskBytes, u64 maskSize, u32 perQuery, bx_ray_hit* hits, uint32_t* fixtures,
                      uint32_t* counts) {
    if (count == 0) return BX_OK;
    if ((kind != 0 && !counts) || (kind == 0 && !hits) || (kind == 1 && !hits && perQuery) ||
        (kind == 2 && !fixtures && perQuery))
        return Fail(BX_E_INVALID, "null output");
    std::vector<u32> worlds(count);
    for (u32 i = 0; i < count; ++i) {
        const u32 w = world[i];
        if (!IsResident(b, w)) return Fail(BX_E_INVALID, "query names a world that is not resident");
        if (maskOff[i] != kNull) {
            if (!maskBytes || u64(maskOff[i]) + b->worlds[w].fixtureCount > maskSize)
                return Fail(BX_E_INVALID, "query mask lies outside the mask bytes");
        }
        for (u32 k = 0; k < 6; ++k)
            if (!std::isfinite(coords[6 * u64(i) + k])) return Fail(BX_E_INVALID, "query coordinates are not finite");
        worlds[i] = w;
    }
    HIP_CHECK(hipSetDevice(b->ctx->device));
    int rc = EnsureBvh(b, worlds);
    if (rc) return rc;
    // device layout: query headers, coordinates, mask bytes, outputs, counts
    const size_t perOut = kind == 2 ? sizeof(u32) : sizeof(RayHitOut);
    const size_t outCount = kind == 0 ? size_t(count) : size_t(count) * perQuery;
    const size_t oQ = 0, oC = Align256(oQ + size_t(count) * sizeof(uint4));
    const size_t oM = Align256(oC + size_t(count) * 6 * sizeof(float));
    const size_t oO = Align256(oM + (maskBytes ? size_t(maskSize) : 0));
    const size_t oN = Align256(oO + std::max<size_t>(1, outCount) * perOut);
    const size_t total = Align256(oN + size_t(count) * sizeof(u32));
    if ((rc = QueryScratch(b, total))) return rc;
    std::vector<uint4> heads(count);
    for (u32 i = 0; i < count; ++i) heads[i] = make_uint4(world[i], maskOff[i], 0u, 0u);
    hipStream_t st = b->ctx->stream;
    HIP_CHECK(hipMemcpyAsync(b->queryDev + oQ, heads.data(), heads.size() * sizeof(uint4), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(b->queryDev + oC, coords, size_t(count) * 6 * sizeof(float), hipMemcpyHostToDevice, st));
    if (maskBytes && maskSize)
        HIP_CHECK(hipMemcpyAsync(b->queryDev + oM, maskBytes, size_t(maskSize), hipMemcpyHostToDevice, st));
    QueryParams qp;
    qp.queries = reinterpret_cast<const uint4*>(b->queryDev + oQ);
    qp.coords = reinterpret_cast<const float*>(b->queryDev + oC);
    qp.mask = reinterpret_cast<const uint8_t*>(b->queryDev + oM);
    qp.count = count;
    qp.kind = kind;
    qp.perQuery = perQuery;
    qp.hits = reinterpret_cast<RayHitOut*>(b->queryDev + oO);
    qp.fixtures = reinterpret_cast<u32*>(b->queryDev + oO);
    qp.counts = reinterpret_cast<u32*>(b->queryDev + oN);
    LaunchQueries(b->view, qp, st);
    HIP_CHECK(hipGetLastError());
    if (kind == 2) {
        if (outCount)
            HIP_CHECK(hipMemcpyAsync(fixtures, b->queryDev + oO, outCount * sizeof(u32), hipMemcpyDeviceToHost, st));
    } else if (outCount) {
        std::vector<RayHitOut> out(outCount);
        HIP_CHECK(hipMemcpyAsync(out.data(), b->queryDev + oO, outCount * sizeof(RayHitOut), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        for (size_t i = 0; i < outCount; ++i) {
            hits[i].fixture = out[i].fixture;
            hits[i].fraction = out[i].fraction;
            hits[i].point[0] = out[i].point.x;
            hits[i].point[1] = out[i].point.y;
            hits[i].point[2] = out[i].point.z;
            hits[i].normal[0] = out[i].normal.x;
            hits[i].normal[1] = out[i].normal.y;
            hits[i].normal[2] = out[i].normal.z;
        }
    }
    if (kind != 0)
        HIP_CHECK(hipMemcpyAsync(counts, b->queryDev + oN, size_t(count) * sizeof(u32), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return BX_OK;
}

// World, mask offset and the six coordinates (p1 p2, or lower upper) of each query record.
static_assert(offsetof(bx_ray_query, p1) == 8 && offsetof(bx_ray_query, p2) == 20 &&
                  offsetof(bx_aabb_query, lower) == 8 && offsetof(bx_aabb_query, upper) == 20,
              "query records: world, mask, then six floats");
static_assert(sizeof(bx_ray_query) == 32 && sizeof(bx_aabb_query) == 32, "query records are 32 bytes");
static void SplitQueries(const void* q, u32 count, std::vector<u32>& world, std::vector<u32>& mask,
                         std::vector<float>& coords) {
    world.resize(count);
    mask.resize(count);
    coords.resize(size_t(count) * 6);
    const char* p = static_cast<const char*>(q);
    for (u32 i = 0; i < count; ++i, p += 32) {
        std::memcpy(&world[i], p, sizeof(u32));
        std::memcpy(&mask[i], p + 4, sizeof(u32));
        std::memcpy(&coords[6 * size_t(i)], p + 8, 6 * sizeof(float));
    }
}

int bx_batch_raycast_single(bx_batch* b, const bx_ray_query* rays, uint32_t count, const uint8_t* mask_bytes,
                            uint64_t mask_size, bx_ray_hit* out) {
    if (!b || (!rays && count) || (!out && count)) return Fail(BX_E_INVALID, "null argument");
    std::vector<u32> world, mask;
    std::vector<float> coords;
    SplitQueries(rays, count, world, mask, coords);
    return RunQueries(b, 0, world.data(), mask.data(), coords.data(), count, mask_bytes, mask_size, 1, out, nullptr,
                      nullptr);
}

int bx_batch_raycast(bx_batch* b, const bx_ray_query* rays, uint32_t count, const uint8_t* mask_bytes,
                     uint64_t mask_size, bx_ray_hit* hits, uint32_t hits_per_ray, uint32_t* hit_counts) {
    if (!b || (!rays && count)) return Fail(BX_E_INVALID, "null argument");
    std::vector<u32> world, mask;
    std::vector<float> coords;
    SplitQueries(rays, count, world, mask, coords);
    return RunQueries(b, 1, world.data(), mask.data(), coords.data(), count, mask_bytes, mask_size, hits_per_ray, hits,
                      nullptr, hit_counts);
}

int bx_batch_query_aabb(bx_batch* b, const bx_aabb_query* boxes, uint32_t count, const uint8_t* mask_bytes,
                        uint64_t mask_size, uint32_t* fixtures, uint32_t per_query, uint32_t* counts) {
    if (!b || (!boxes && count)) return Fail(BX_E_INVALID, "null argument");
    std::vector<u32> world, mask;
    std::vector<float> coords;
    SplitQueries(boxes, count, world, mask, coords);
    return RunQueries(b, 2, world.data(), mask.data(), coords.data(), count, mask_bytes, mask_size, per_query, nullptr,
                      fixtures, counts);
}

// ------------------------------------------------------------------ time of impact and shape casts
static bool Finite(const float* v, u32 n) {
    for (u32 i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

// A cast shape checked against the library and converted; false with the reason in `why`.
static bool ToCastShape(const bx_batch* b, const bx_cast_shape& c, CastShapeD& out, const char** why) {
    if (c.shape_type > BX_SHAPE_HULL || c.shape_type == BX_SHAPE_MESH) {
        *why = "a cast shape must be a sphere, capsule, triangle or hull";
        return false;
    }
    if (c.shape_type == BX_SHAPE_HULL && c.shape_index >= b->library->info.size()) {
        *why = "a cast hull index is outside the shape library";
        return false;
    }
    if (!std::isfinite(c.radius) || c.radius < 0.0f || !Finite(c.vertex1, 3) || !Finite(c.vertex2, 3) ||
        !Finite(c.vertex3, 3)) {
        *why = "a cast shape has a negative or non-finite radius or vertex";
        return false;
    }
    out.type = c.shape_type;
    out.index = c.shape_type == BX_SHAPE_HULL ? c.shape_index : 0u;
    out.radius = c.radius;
    out.v1 = V3(c.vertex1);
    out.v2 = V3(c.vertex2);
    out.v3 = V3(c.vertex3);
    return true;
}

static Transform XF(const float* p, const float* q) { return Transform(V3(p), Quat(V3(q), q[3])); }

static SweepD ToSweep(const bx_sweep& s) {
    SweepD w;
    w.localCenter = V3(s.local_center);
    w.c0 = V3(s.center0);
    w.c = V3(s.center);
    w.q0 = Quat(V3(s.orientation0), s.orientation0[3]);
    w.q = Quat(V3(s.orientation), s.orientation[3]);
    return w;
}

int bx_batch_time_of_impact(bx_batch* b, const bx_toi_query* queries, uint32_t count, bx_toi_result* out) {
    if (!b || ((!queries || !out) && count)) return Fail(BX_E_INVALID, "null argument");
    if (count == 0) return BX_OK;
    std::vector<ToiQueryD> qs(count);
    for (u32 i = 0; i < count; ++i) {
        const bx_toi_query& x = queries[i];
        ToiQueryD& q = qs[i];
        std::memset(static_cast<void*>(&q), 0, sizeof(q));
        q.kind = x.kind;
        q.maxIterations = x.max_iterations;
        q.tMax = x.t_max;
        const char* why = nullptr;
        if (x.kind == BX_TOI_AABB) {
            if (!Finite(x.lower1, 3) || !Finite(x.upper1, 3) || !Finite(x.lower2, 3) || !Finite(x.upper2, 3) ||
                !Finite(x.displacement1, 3) || !Finite(x.displacement2, 3))
                return Fail(BX_E_INVALID, "time of impact: non-finite box or displacement");
            q.a1.lowerBound = V3(x.lower1);
            q.a1.upperBound = V3(x.upper1);
            q.a2.lowerBound = V3(x.lower2);
            q.a2.upperBound = V3(x.upper2);
            q.d1 = V3(x.displacement1);
            q.d2 = V3(x.displacement2);
            continue;
        }
        if (x.kind != BX_TOI_SWEEP && x.kind != BX_TOI_TRANSLATION) return Fail(BX_E_INVALID, "time of impact: unknown kind");
        if (!ToCastShape(b, x.shape1, q.s1, &why) || !ToCastShape(b, x.shape2, q.s2, &why)) return Fail(BX_E_INVALID, why);
        if (x.kind == BX_TOI_SWEEP) {
            const bx_sweep* sw[2] = {&x.sweep1, &x.sweep2};
            for (const bx_sweep* w : sw)
                if (!Finite(w->local_center, 3) || !Finite(w->center0, 3) || !Finite(w->center, 3) ||
                    !Finite(w->orientation0, 4) || !Finite(w->orientation, 4))
                    return Fail(BX_E_INVALID, "time of impact: non-finite sweep");
            if (!std::isfinite(x.t_max)) return Fail(BX_E_INVALID, "time of impact: non-finite tMax");
            q.w1 = ToSweep(x.sweep1);
            q.w2 = ToSweep(x.sweep2);
        } else {
            // conservative advancement stops only at its iteration cap in the worst case
            if (x.max_iterations == 0 || x.max_iterations > 1000)
                return Fail(BX_E_INVALID, "time of impact: max_iterations must be 1..1000");
            if (!Finite(x.position1, 3) || !Finite(x.orientation1, 4) || !Finite(x.position2, 3) ||
                !Finite(x.orientation2, 4) || !Finite(x.displacement1, 3) || !Finite(x.displacement2, 3))
                return Fail(BX_E_INVALID, "time of impact: non-finite transform or displacement");
            q.xf1 = XF(x.position1, x.orientation1);
            q.xf2 = XF(x.position2, x.orientation2);
            q.d1 = V3(x.displacement1);
            q.d2 = V3(x.displacement2);
        }
    }
    HIP_CHECK(hipSetDevice(b->ctx->device));
    const size_t oR = Align256(size_t(count) * sizeof(ToiQueryD));
    int rc = QueryScratch(b, oR + size_t(count) * sizeof(ToiResultD));
    if (rc) return rc;
    hipStream_t st = b->ctx->stream;
    HIP_CHECK(hipMemcpyAsync(b->queryDev, qs.data(), qs.size() * sizeof(ToiQueryD), hipMemcpyHostToDevice, st));
    LaunchTimeOfImpact(b->view, reinterpret_cast<const ToiQueryD*>(b->queryDev), count,
                       reinterpret_cast<ToiResultD*>(b->queryDev + oR), st);
    HIP_CHECK(hipGetLastError());
    std::vector<ToiResultD> res(count);
    HIP_CHECK(hipMemcpyAsync(res.data(), b->queryDev + oR, res.size() * sizeof(ToiResultD), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    for (u32 i = 0; i < count; ++i) {
        out[i].state = res[i].state;
        out[i].t = res[i].t;
        out[i].iterations = res[i].iterations;
    }
    return BX_OK;
}

static int RunShapeCasts(bx_batch* b, const bx_shape_cast_query* casts, u32 count, const uint8_t* maskBytes,
                         u64 maskSize, u32 single, u32 perQuery, bx_shape_hit* hits, uint32_t* counts) {
    if (!b || (!casts && count)) return Fail(BX_E_INVALID, "null argument");
    if (count == 0) return BX_OK;
    if ((single && !hits) || (!single && (!counts || (!hits && perQuery)))) return Fail(BX_E_INVALID, "null output");
    std::vector<ShapeCastD> cs(count);
    std::vector<u32> worlds(count);
    for (u32 i = 0; i < count; ++i) {
        const bx_shape_cast_query& x = casts[i];
        if (!IsResident(b, x.world)) return Fail(BX_E_INVALID, "shape cast names a world that is not resident");
        if (x.mask != kNull && (!maskBytes || u64(x.mask) + b->worlds[x.world].fixtureCount > maskSize))
            return Fail(BX_E_INVALID, "shape cast mask lies outside the mask bytes");
        if (!Finite(x.position, 3) || !Finite(x.orientation, 4) || !Finite(x.displacement, 3))
            return Fail(BX_E_INVALID, "shape cast transform or displacement is not finite");
        ShapeCastD& c = cs[i];
        std::memset(static_cast<void*>(&c), 0, sizeof(c));
        const char* why = nullptr;
        if (!ToCastShape(b, x.shape, c.shape, &why)) return Fail(BX_E_INVALID, why);
        c.world = x.world;
        c.mask = x.mask;
        c.xf = XF(x.position, x.orientation);
        c.d = V3(x.displacement);
        worlds[i] = x.world;
    }
    HIP_CHECK(hipSetDevice(b->ctx->device));
    int rc = EnsureBvh(b, worlds);
    if (rc) return rc;
    const size_t outCount = single ? size_t(count) : size_t(count) * perQuery;
    const size_t oM = Align256(size_t(count) * sizeof(ShapeCastD));
    const size_t oO = Align256(oM + (maskBytes ? size_t(maskSize) : 0));
    const size_t oN = Align256(oO + std::max<size_t>(1, outCount) * sizeof(ShapeHitOut));
    if ((rc = QueryScratch(b, Align256(oN + size_t(count) * sizeof(u32))))) return rc;
    hipStream_t st = b->ctx->stream;
    HIP_CHECK(hipMemcpyAsync(b->queryDev, cs.data(), cs.size() * sizeof(ShapeCastD), hipMemcpyHostToDevice, st));
    if (maskBytes && maskSize)
        HIP_CHECK(hipMemcpyAsync(b->queryDev + oM, maskBytes, size_t(maskSize), hipMemcpyHostToDevice, st));
    CastParams cp;
    cp.casts = reinterpret_cast<const ShapeCastD*>(b->queryDev);
    cp.mask = reinterpret_cast<const uint8_t*>(b->queryDev + oM);
    cp.count = count;
    cp.single = single;
    cp.perQuery = perQuery;
    cp.hits = reinterpret_cast<ShapeHitOut*>(b->queryDev + oO);
    cp.counts = reinterpret_cast<u32*>(b->queryDev + oN);
    LaunchShapeCasts(b->view, cp, st);
    HIP_CHECK(hipGetLastError());
    if (outCount) {
        std::vector<ShapeHitOut> out(outCount);
        HIP_CHECK(hipMemcpyAsync(out.data(), b->queryDev + oO, outCount * sizeof(ShapeHitOut), hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        for (size_t i = 0; i < outCount; ++i) {
            hits[i].fixture = out[i].fixture;
            hits[i].child = out[i].child;
            hits[i].fraction = out[i].fraction;
            hits[i].point[0] = out[i].point.x;
            hits[i].point[1] = out[i].point.y;
            hits[i].point[2] = out[i].point.z;
            hits[i].normal[0] = out[i].normal.x;
            hits[i].normal[1] = out[i].normal.y;
            hits[i].normal[2] = out[i].normal.z;
        }
    }
    if (!single)
        HIP_CHECK(hipMemcpyAsync(counts, b->queryDev + oN, size_t(count) * sizeof(u32), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return BX_OK;
}

int bx_batch_shape_cast_single(bx_batch* b, const bx_shape_cast_query* casts, uint32_t count,
                               const uint8_t* mask_bytes, uint64_t mask_size, bx_shape_hit* out) {
    return RunShapeCasts(b, casts, count, mask_bytes, mask_size, 1u, 1u, out, nullptr);
}

int bx_batch_shape_cast(bx_batch* b, const bx_shape_cast_query* casts, uint32_t count, const uint8_t* mask_bytes,
                        uint64_t mask_size, bx_shape_hit* hits, uint32_t hits_per_cast, uint32_t* hit_counts) {
    return RunShapeCasts(b, casts, count, mask_bytes, mask_size, 0u, hits_per_cast, hits, hit_counts);
}

// ------------------------------------------------------------------ topology edits
// A resident world's whole state on the host with per-body fixture and joint lists (list order,
// newest first), as b3World::CreateBody / DestroyBody / CreateJoint / DestroyJoint and
// b3Body::CreateFixture / DestroyFixture change it.
struct EditWorld {
    HostWorld hw;
    std::vector<std::vector<u32>> fixturesOf;
    std::vector<std::vector<JointEdge>> jointsOf;
};

// The state of world w of batch b (synchronous): bodies, fixtures, joints, proxies, the live
// contacts in list order with their mesh records, and the last pairs.
static int TakeWorld(bx_batch* b, u32 w, EditWorld& e) {
    HostWorld& hw = e.hw;
    const WorldInfo wi = b->worlds[w];
    const BatchView& V = b->view;
    hw.info = wi;
    hw.carried = true;
    int rc = GetState(b, w, hw.state);
    if (rc) return rc;
    if ((rc = Download(b, hw.bodies, V.bodies + wi.bodyBase, wi.bodyCount))) return rc;
    if ((rc = Download(b, hw.fixtures, V.fixtures + wi.fixtureBase, wi.fixtureCount))) return rc;
    std::vector<u32> flist;
    if ((rc = Download(b, flist, V.fixtureList + wi.fixtureListBase, wi.fixtureCount))) return rc;
    if ((rc = Download(b, hw.fat, V.fatAABB + wi.fixtureBase, wi.fixtureCount))) return rc;
    if ((rc = Download(b, hw.moved, V.moved + wi.fixtureBase, wi.fixtureCount))) return rc;
    if ((rc = Download(b, hw.joints, V.joints + wi.jointBase, wi.jointCount))) return rc;
    std::vector<JointEdge> edges;
    if ((rc = Download(b, edges, V.jointEdges + wi.jointEdgeBase, 2 * size_t(wi.jointCount)))) return rc;
    std::vector<u32> order;
    if ((rc = Download(b, order, V.contactOrder + wi.contactBase, hw.state.contactCount))) return rc;
    std::vector<Contact> slots;
    if ((rc = Download(b, slots, V.contacts + wi.contactBase, wi.contactCap))) return rc;
    std::vector<MeshContact> meshSlots;
    std::vector<u32> triSlots;
    if (wi.meshTriCap) {
        if ((rc = Download(b, meshSlots, V.meshContacts + wi.meshBase, wi.contactCap))) return rc;
        if ((rc = Download(b, triSlots, V.meshTriLists + wi.meshTriBase, size_t(wi.contactCap) * wi.meshTriCap)))
            return rc;
    }
    if ((rc = Download(b, hw.pairs, V.pairs + wi.pairBase, std::min(hw.state.pairCount, wi.pairCap)))) return rc;
    for (u32 slot : order) {
        hw.contacts.push_back(slots[slot]);
        if (wi.meshTriCap) {
            hw.meshContacts.push_back(meshSlots[slot]);
            hw.meshTris.insert(hw.meshTris.end(), triSlots.begin() + size_t(slot) * wi.meshTriCap,
                               triSlots.begin() + size_t(slot + 1) * wi.meshTriCap);
        }
    }
    hw.proxyIds = b->proxyIds[w];
    hw.nodes = b->nodePools[w];
    hw.needsInitialPairs = b->needsInitialPairs[w] != 0;
    e.fixturesOf.assign(wi.bodyCount, {});
    e.jointsOf.assign(wi.bodyCount, {});
    for (u32 i = 0; i < wi.bodyCount; ++i) {
        const Body& x = hw.bodies[i];
        e.fixturesOf[i].assign(flist.begin() + x.fixtureBegin, flist.begin() + x.fixtureBegin + x.fixtureCount);
        e.jointsOf[i].assign(edges.begin() + x.jointBegin, edges.begin() + x.jointBegin + x.jointCount);
    }
    return BX_OK;
}

// Removes the live contacts of fixture f, keeping the others in list order (b3ContactManager::
// Destroy, contact_manager.cpp:229-290). b3Contact::Destroy (contact.cpp:108-131) wakes both bodies
// of a contact with a nonempty manifold unless a fixture is a sensor.
static void DropContactsOf(HostWorld& hw, u32 f) {
    const u32 tri = hw.info.meshTriCap;
    u32 kept = 0;
    for (u32 i = 0; i < hw.contacts.size(); ++i) {
        const Contact& c = hw.contacts[i];
        if (c.fixtureA == f || c.fixtureB == f) {
            const Fixture& fA = hw.fixtures[c.fixtureA];
            const Fixture& fB = hw.fixtures[c.fixtureB];
            bool points = false;
            for (u32 m = 0; m < c.manifoldCount && m < kMaxManifolds; ++m)
                points = points || (m == 0 ? c.manifold.pointCount : hw.meshContacts[i].extra[m - 1].pointCount) > 0;
            if (points && !fA.isSensor && !fB.isSensor) {
                HostSetAwake(hw.bodies[fA.body], true);
                HostSetAwake(hw.bodies[fB.body], true);
            }
            continue;
        }
        hw.contacts[kept] = hw.contacts[i];
        if (tri) {
            hw.meshContacts[kept] = hw.meshContacts[i];
            std::copy(hw.meshTris.begin() + size_t(i) * tri, hw.meshTris.begin() + size_t(i + 1) * tri,
                      hw.meshTris.begin() + size_t(kept) * tri);
        }
        ++kept;
    }
    hw.contacts.resize(kept);
    if (tri) {
        hw.meshContacts.resize(kept);
        hw.meshTris.resize(size_t(kept) * tri);
    }
}

// b3JointManager::Destroy (joint_manager.cpp:75-141): wakes both bodies, unlinks the joint's edges.
static void EditDestroyJoint(EditWorld& e, u32 j) {
    HostWorld& hw = e.hw;
    const Joint J = hw.joints[j];
    HostSetAwake(hw.bodies[J.bodyA], true);
    HostSetAwake(hw.bodies[J.bodyB], true);
    for (auto& list : e.jointsOf) {
        list.erase(std::remove_if(list.begin(), list.end(), [&](const JointEdge& x) { return x.joint == j; }), list.end());
        for (JointEdge& x : list)
            if (x.joint > j) --x.joint;
    }
    hw.joints.erase(hw.joints.begin() + j);
}

// The contacts and the broadphase proxy of fixture f go; fixture indices above f move down.
static void EditRemoveFixture(EditWorld& e, u32 f) {
    HostWorld& hw = e.hw;
    DropContactsOf(hw, f);
    for (Contact& c : hw.contacts) {
        if (c.fixtureA > f) --c.fixtureA;
        if (c.fixtureB > f) --c.fixtureB;
    }
    // The pairs of the last search stay as the broadphase's pair buffer keeps them; an entry of the
    // destroyed fixture keeps its proxy id (kPairProxyBit), the others follow the new indices.
    const u32 gone = kPairProxyBit | hw.proxyIds[f];
    for (uint2& p : hw.pairs) {
        for (u32* x : {&p.x, &p.y}) {
            if (*x & kPairProxyBit) continue;
            if (*x == f) *x = gone;
            else if (*x > f) --*x;
        }
    }
    PoolDestroyProxy(hw.nodes, hw.proxyIds[f]);  // b3BroadPhase::DestroyProxy (also unbuffers its move)
    const u32 body = hw.fixtures[f].body;
    auto& own = e.fixturesOf[body];
    own.erase(std::remove(own.begin(), own.end(), f), own.end());
    for (auto& list : e.fixturesOf)
        for (u32& x : list)
            if (x > f) --x;
    hw.fixtures.erase(hw.fixtures.begin() + f);
    hw.proxyIds.erase(hw.proxyIds.begin() + f);
    hw.fat.erase(hw.fat.begin() + f);
    hw.moved.erase(hw.moved.begin() + f);
}

static int EditOp(EditWorld& e, const HostLibrary& L, const bx_world_desc* adds, const bx_topology_op& op) {
    HostWorld& hw = e.hw;
    const u32 nb = u32(hw.bodies.size()), nf = u32(hw.fixtures.size()), nj = u32(hw.joints.size());
    switch (op.op) {
    case BX_TOPO_CREATE_BODY: {  // b3World::CreateBody (world.cpp:94-110)
        if (!adds || op.index >= adds->body_count) return Fail(BX_E_INVALID, "body descriptor index out of range");
        Body x;
        int rc = MakeBody(adds->bodies[op.index], x);
        if (rc) return rc;
        hw.bodies.push_back(x);
        e.fixturesOf.emplace_back();
        e.jointsOf.emplace_back();
        return BX_OK;
    }
    case BX_TOPO_CREATE_FIXTURE: {  // b3Body::CreateFixture (body.cpp:105-136)
        if (!adds || op.index >= adds->fixture_count) return Fail(BX_E_INVALID, "fixture descriptor index out of range");
        const bx_fixture_desc& fd = adds->fixtures[op.index];
        if (fd.body >= nb) return Fail(BX_E_INVALID, "fixture body out of range");
        if (nf >= 0xFFFFu) return Fail(BX_E_INVALID, "world too large");
        Fixture f;
        int rc = MakeFixture(fd, L, f);
        if (rc) return rc;
        Body& x = hw.bodies[fd.body];
        u32 proxy;
        if (!PoolCreateProxy(hw.nodes, &proxy))
            return Fail(BX_E_UNSUPPORTED, "the new proxy's node id depends on the dynamic tree's shape");
        f.proxyId = proxy;
        hw.fat.push_back(CreationFatAABB(f, Transform(x.xfT, x.q), L));
        hw.moved.push_back(1u);  // BufferMove
        hw.proxyIds.push_back(proxy);
        hw.fixtures.push_back(f);
        e.fixturesOf[fd.body].insert(e.fixturesOf[fd.body].begin(), nf);
        if (f.density > scalar(0)) ResetMassOf(x, hw.fixtures, e.fixturesOf[fd.body], L);
        hw.needsInitialPairs = true;  // m_newContacts
        return BX_OK;
    }
    case BX_TOPO_CREATE_JOINT: {  // b3JointManager::Create (joint_manager.cpp:30-73)
        if (!adds || op.index >= adds->joint_count) return Fail(BX_E_INVALID, "joint descriptor index out of range");
        if (nj >= 0x7FFFu) return Fail(BX_E_INVALID, "world too large");
        const bx_joint_desc& jd = adds->joints[op.index];
        Joint j;
        int rc = MakeJoint(jd, hw.bodies, j);
        if (rc) return rc;
        hw.joints.push_back(j);
        e.jointsOf[jd.body_a].insert(e.jointsOf[jd.body_a].begin(), JointEdge{nj, jd.body_b});
        e.jointsOf[jd.body_b].insert(e.jointsOf[jd.body_b].begin(), JointEdge{nj, jd.body_a});
        return BX_OK;
    }
    case BX_TOPO_DESTROY_JOINT:
        if (op.index >= nj) return Fail(BX_E_INVALID, "joint index out of range");
        EditDestroyJoint(e, op.index);
        return BX_OK;
    case BX_TOPO_DESTROY_FIXTURE: {  // b3Body::DestroyFixture (body.cpp:138-191)
        if (op.index >= nf) return Fail(BX_E_INVALID, "fixture index out of range");
        const u32 body = hw.fixtures[op.index].body;
        EditRemoveFixture(e, op.index);
        ResetMassOf(hw.bodies[body], hw.fixtures, e.fixturesOf[body], L);
        return BX_OK;
    }
    case BX_TOPO_DESTROY_BODY: {  // b3World::DestroyBody (world.cpp:112-175)
        const u32 b = op.index;
        if (b >= nb) return Fail(BX_E_INVALID, "body index out of range");
        while (!e.jointsOf[b].empty()) EditDestroyJoint(e, e.jointsOf[b].front().joint);
        while (!e.fixturesOf[b].empty()) EditRemoveFixture(e, e.fixturesOf[b].front());
        hw.bodies.erase(hw.bodies.begin() + b);
        e.fixturesOf.erase(e.fixturesOf.begin() + b);
        e.jointsOf.erase(e.jointsOf.begin() + b);
        for (Fixture& f : hw.fixtures)
            if (f.body > b) --f.body;
        for (Joint& j : hw.joints) {
            if (j.bodyA > b) --j.bodyA;
            if (j.bodyB > b) --j.bodyB;
        }
        for (auto& list : e.jointsOf)
            for (JointEdge& x : list)
                if (x.other > b) --x.other;
        return BX_OK;
    }
    }
    return Fail(BX_E_INVALID, "unknown topology operation");
}

// Back to the CSR lists of the device layout, with counts, capacities and the slot stack.
static int FinishEdit(EditWorld& e, const bx_batch_opts* opts) {
    HostWorld& hw = e.hw;
    WorldInfo& wi = hw.info;
    wi.bodyCount = u32(hw.bodies.size());
    wi.fixtureCount = u32(hw.fixtures.size());
    wi.jointCount = u32(hw.joints.size());
    hw.fixtureList.clear();
    hw.jointEdges.clear();
    for (u32 i = 0; i < wi.bodyCount; ++i) {
        Body& x = hw.bodies[i];
        x.fixtureBegin = u32(hw.fixtureList.size());
        x.fixtureCount = u32(e.fixturesOf[i].size());
        hw.fixtureList.insert(hw.fixtureList.end(), e.fixturesOf[i].begin(), e.fixturesOf[i].end());
        x.jointBegin = u32(hw.jointEdges.size());
        x.jointCount = u32(e.jointsOf[i].size());
        hw.jointEdges.insert(hw.jointEdges.end(), e.jointsOf[i].begin(), e.jointsOf[i].end());
    }
    for (u32 i = 0; i < wi.fixtureCount; ++i) hw.fixtures[i].proxyId = hw.proxyIds[i];
    bool hasMesh = false;
    for (const Fixture& f : hw.fixtures) hasMesh = hasMesh || f.shapeType == BX_SHAPE_MESH;
    return FinishWorld(hw, hasMesh, opts);
}

// Whether the first hulls and meshes of `next` are those of `prev` (indices stay valid).
static bool PrefixBytes(const void* a, size_t na, const void* b, size_t nb) {
    return na <= nb && (na == 0 || std::memcmp(a, b, na) == 0);
}
#define BX_PREFIX(m) PrefixBytes(prev.m.data(), prev.m.size() * sizeof(prev.m[0]), next.m.data(), next.m.size() * sizeof(next.m[0]))
static bool LibraryExtends(const HostLibrary& prev, const HostLibrary& next) {
    return BX_PREFIX(info) && BX_PREFIX(vertices) && BX_PREFIX(edges) && BX_PREFIX(faces) && BX_PREFIX(planes) &&
           BX_PREFIX(meshes) && BX_PREFIX(meshVertices) && BX_PREFIX(meshTris) && BX_PREFIX(meshNodes);
}
#undef BX_PREFIX

int bx_batch_edit_topology(bx_batch* b, uint32_t world, const bx_shape_library* library, const bx_world_desc* adds,
                           const bx_topology_op* ops, uint32_t op_count) {
    if (!b || (!ops && op_count)) return Fail(BX_E_INVALID, "null argument");
    if (b->pool) return Fail(BX_E_UNSUPPORTED, "topology edits apply to a batch, not a pool");
    if (!IsResident(b, world)) return Fail(BX_E_INVALID, "world not resident");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    HIP_CHECK(hipStreamSynchronize(b->ctx->stream));
    HostLibrary* L = new HostLibrary();
    int rc = library ? BuildLibrary(library, *L) : BX_OK;
    if (!library) *L = *b->library;
    if (!rc && !LibraryExtends(*b->library, *L))
        rc = Fail(BX_E_INVALID, "the new shape library must begin with the batch's hulls and meshes");
    if (rc) { delete L; return rc; }
    // the edited world first, so that a failing operation leaves the batch unchanged
    EditWorld edited;
    rc = TakeWorld(b, world, edited);
    for (u32 k = 0; !rc && k < op_count; ++k) rc = EditOp(edited, *L, adds, ops[k]);
    if (!rc) rc = FinishEdit(edited, &b->worldOpts);
    if (rc) { delete L; return rc; }
    bx_batch_opts opts = b->worldOpts;
    opts.solver_mode = b->view.solveLayout + 1;  // the batch's current layout (tuned or fixed)
    bx_batch* nb = nullptr;
    rc = CreateBatchFrom(b->ctx, L, &opts, b->worldCount, [&](u32 w, HostWorld& hw) -> int {
        if (w == world) {
            hw = std::move(edited.hw);
            return BX_OK;
        }
        EditWorld other;
        int r = TakeWorld(b, w, other);
        if (!r) r = FinishEdit(other, &b->worldOpts);
        if (!r) hw = std::move(other.hw);
        return r;
    }, &nb);
    if (rc) return rc;
    // the new arrays replace the old ones under the same handle; the velocity-sweep tuning state and
    // the options carry over
    nb->worldOpts = b->worldOpts;
    nb->tuneStep = b->tuneStep;
    std::memcpy(nb->tuneMs, b->tuneMs, sizeof(b->tuneMs));
    std::swap(*b, *nb);
    bx_batch_destroy(nb);
    ++b->stateGen;
    return BX_OK;
}

int bx_batch_edit_bodies(bx_batch* b, uint32_t world, const bx_body_edit* edits, uint32_t count) {
    if (!b || !IsResident(b, world) || (!edits && count)) return Fail(BX_E_INVALID, "bad argument or world not resident");
    HIP_CHECK(hipSetDevice(b->ctx->device));
    WorldInfo& wi = b->worlds[world];
    bool transform = false, fixed = false, worldEdit = false;
    for (u32 k = 0; k < count; ++k) {
        const u32 op = edits[k].op;
        const bool worldOp = op >= BX_EDIT_WORLD_GRAVITY && op <= BX_EDIT_WORLD_CLEAR_FORCES;
        if (!worldOp && (op < BX_EDIT_LINEAR_VELOCITY || op > BX_EDIT_FIXED_ROTATION))
            return Fail(BX_E_INVALID, "unknown body edit");
        if (!worldOp && edits[k].body >= wi.bodyCount) return Fail(BX_E_INVALID, "edit names a body out of range");
        transform = transform || op == BX_EDIT_TRANSFORM;
        fixed = fixed || op == BX_EDIT_FIXED_ROTATION;
        worldEdit = worldEdit || worldOp;
    }
    if (count == 0) return BX_OK;
    if (transform) ++b->stateGen;  // SetTransform moves fat AABBs
    std::vector<Body> bodies;
    int rc = Download(b, bodies, b->view.bodies + wi.bodyBase, wi.bodyCount);
    if (rc) return rc;
    // SetTransform moves the body's fixture proxies: their fixtures, fat AABBs and moved flags;
    // SetFixedRotation recomputes the mass from the body's fixtures
    std::vector<u32> flist, moved;
    std::vector<Fixture> fixtures;
    std::vector<AABB> fat;
    if (transform || fixed) {
        if ((rc = Download(b, flist, b->view.fixtureList + wi.fixtureListBase, wi.fixtureCount))) return rc;
        if ((rc = Download(b, fixtures, b->view.fixtures + wi.fixtureBase, wi.fixtureCount))) return rc;
    }
    if (transform) {
        if ((rc = Download(b, fat, b->view.fatAABB + wi.fixtureBase, wi.fixtureCount))) return rc;
        if ((rc = Download(b, moved, b->view.moved + wi.fixtureBase, wi.fixtureCount))) return rc;
    }
    u32 worldFlags = wi.flags;
    Vec3 gravity = wi.gravity;
    for (u32 k = 0; k < count; ++k) {
        const bx_body_edit& e = edits[k];
        switch (e.op) {  // world edits
        case BX_EDIT_WORLD_GRAVITY: gravity = V3(e.v); continue;
        case BX_EDIT_WORLD_SLEEPING:  // world.cpp:82-92
            if (e.wake) {
                worldFlags |= BX_WORLD_SLEEPING;
            } else {
                worldFlags &= ~BX_WORLD_SLEEPING;
                for (Body& x : bodies) HostSetAwake(x, true);
            }
            continue;
        case BX_EDIT_WORLD_WARM_STARTING:
            worldFlags = e.wake ? (worldFlags & ~BX_WORLD_NO_WARM_START) : (worldFlags | BX_WORLD_NO_WARM_START);
            continue;
        case BX_EDIT_WORLD_AUTO_CLEAR_FORCES:
            worldFlags = e.wake ? (worldFlags & ~BX_WORLD_NO_CLEAR_FORCES) : (worldFlags | BX_WORLD_NO_CLEAR_FORCES);
            continue;
        case BX_EDIT_WORLD_CLEAR_FORCES:  // world.cpp:244-251
            for (Body& x : bodies) {
                x.force.SetZero();
                x.torque.SetZero();
            }
            continue;
        default: break;
        }
        Body& x = bodies[e.body];
        const Vec3 v = V3(e.v), p(e.p[0], e.p[1], e.p[2]);
        const bool dynamic = x.type == kDynamic;
        // the Apply* methods: dynamic bodies only; wake first if asked; act only on an awake body
        auto wakeFirst = [&]() {
            if (e.wake && (x.flags & kAwake) == 0) HostSetAwake(x, true);
            return (x.flags & kAwake) != 0;
        };
        switch (e.op) {
        case BX_EDIT_LINEAR_VELOCITY: HostSetVelocity(x, x.v, v); break;
        case BX_EDIT_ANGULAR_VELOCITY: HostSetVelocity(x, x.w, v); break;
        case BX_EDIT_FORCE:  // body.h:727-744
            if (dynamic && wakeFirst()) {
                x.force += v;
                x.torque += Cross(p - x.x, v);
            }
            break;
        case BX_EDIT_FORCE_TO_CENTER:
            if (dynamic && wakeFirst()) x.force += v;
            break;
        case BX_EDIT_TORQUE:
            if (dynamic && wakeFirst()) x.torque += v;
            break;
        case BX_EDIT_LINEAR_IMPULSE:  // body.h:782-799
            if (dynamic && wakeFirst()) {
                x.v += x.invMass * v;
                x.w += x.worldInvI * Cross(p - x.x, v);
            }
            break;
        case BX_EDIT_ANGULAR_IMPULSE:
            if (dynamic && wakeFirst()) x.w += x.worldInvI * v;
            break;
        case BX_EDIT_TRANSFORM: {  // body.cpp:194-214
            const Transform xf(v, Quat(Vec3(e.p[0], e.p[1], e.p[2]), e.p[3]));
            x.xfT = xf.translation;
            x.q = xf.rotation;
            x.x = Mul(xf, x.localCenter);
            x.x0 = x.x;
            x.q0 = x.q;
            x.worldInvI = RotateToFrame(x.invI, x.q);
            for (u32 i = 0; i < x.fixtureCount; ++i) {  // b3Fixture::Synchronize(broadPhase, xf, xf)
                const u32 f = flist[x.fixtureBegin + i];
                AABB a1, a2;
                HostAABB(fixtures[f], xf, *b->library, a1);
                HostAABB(fixtures[f], xf, *b->library, a2);
                if (MoveFatAABB(fat[f], Combine(a1, a2), a2.GetCenter() - a1.GetCenter())) moved[f] = 1u;
            }
            b->needsInitialPairs[world] = 1;  // m_newContacts: the next step begins with FindNewContacts
            b->newContacts = true;
            break;
        }
        case BX_EDIT_AWAKE: HostSetAwake(x, e.wake != 0); break;
        case BX_EDIT_SLEEPING_ALLOWED:  // body.h:829-840
            if (e.wake) {
                x.flags |= kAutoSleep;
            } else {
                x.flags &= ~kAutoSleep;
                HostSetAwake(x, true);
            }
            break;
        case BX_EDIT_GRAVITY_SCALE:
            if (x.type != kStatic) x.gravityScale = v;
            break;
        case BX_EDIT_LINEAR_DAMPING: x.linearDamping = v; break;
        case BX_EDIT_ANGULAR_DAMPING: x.angularDamping = v; break;
        case BX_EDIT_LINEAR_SLEEP_TOLERANCE: x.linTol = e.v[0]; break;
        case BX_EDIT_ANGULAR_SLEEP_TOLERANCE: x.angTol = e.v[0]; break;
        case BX_EDIT_FIXED_ROTATION: {  // body.cpp:507-548
            const u32 want = (e.v[0] != 0.0f ? kFixedX : 0u) | (e.v[1] != 0.0f ? kFixedY : 0u) |
                             (e.v[2] != 0.0f ? kFixedZ : 0u);
            if ((x.flags & (kFixedX | kFixedY | kFixedZ)) == want) break;
            x.flags = (x.flags & ~(kFixedX | kFixedY | kFixedZ)) | want;
            x.w.SetZero();
            ResetMass(x, fixtures, flist, 0, *b->library);
            break;
        }
        }
    }
    if (worldEdit) {
        wi.flags = worldFlags;
        wi.gravity = gravity;
        HIP_CHECK(hipMemcpyAsync((void*)(b->view.worlds + world), &wi, sizeof(wi), hipMemcpyHostToDevice,
                                 b->ctx->stream));
    }
    if (transform) {
        HIP_CHECK(hipMemcpyAsync(b->view.fatAABB + wi.fixtureBase, fat.data(), fat.size() * sizeof(AABB),
                                 hipMemcpyHostToDevice, b->ctx->stream));
        HIP_CHECK(hipMemcpyAsync(b->view.moved + wi.fixtureBase, moved.data(), moved.size() * sizeof(u32),
                                 hipMemcpyHostToDevice, b->ctx->stream));
    }
    return UploadBodies(b, wi, bodies);
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
    S(bx_pair) S(bx_world_counts) S(bx_step_cfg) S(bx_batch_opts) S(bx_pool_opts) S(bx_pool_stats) S(bx_body_edit) S(bx_batch_totals) S(bx_contact_event) S(bx_ray_query)
    S(bx_ray_hit) S(bx_aabb_query) S(bx_topology_op) S(bx_body_mass) S(bx_cast_shape) S(bx_sweep) S(bx_toi_query)
    S(bx_toi_result) S(bx_shape_cast_query) S(bx_shape_hit)
#undef S
    return 0;
}

uint64_t bx_batch_device_bytes(bx_batch* b) { return b ? uint64_t(b->deviceBytes) : 0; }

// Host-side mesh queries over the structures the device path builds at upload (no device needed).
static int BuildOneMesh(const bx_mesh_desc* mesh, HostLibrary& L) {
    if (!mesh) return Fail(BX_E_INVALID, "null mesh");
    bx_shape_library lib;
    std::memset(&lib, 0, sizeof(lib));
    lib.mesh_count = 1;
    lib.meshes = mesh;
    return BuildLibrary(&lib, L);
}

int bx_mesh_query_aabbs(const bx_mesh_desc* mesh, const float* aabbs, uint32_t n, uint32_t* counts, uint32_t* out,
                        uint32_t cap_per) {
    if (n && (!aabbs || !counts || (cap_per && !out))) return Fail(BX_E_INVALID, "null argument");
    HostLibrary L;
    int rc = BuildOneMesh(mesh, L);
    if (rc) return rc;
    const TreeNode* nodes = L.meshNodes.data();
    std::vector<u32> stack;
    for (u32 k = 0; k < n; ++k) {
        AABB box;
        box.lowerBound = V3(aabbs + 6 * k);
        box.upperBound = V3(aabbs + 6 * k + 3);
        u32 found = 0;
        stack.assign(1, L.meshes[0].root);
        while (!stack.empty()) {  // the device's traversal (MeshContactFindPairs): stack order
            u32 id = stack.back();
            stack.pop_back();
            if (id == kNull) continue;
            const TreeNode& node = nodes[id];
            if (!TestOverlap(node.aabb, box)) continue;
            if (node.child1 == kNull) {
                if (found < cap_per) out[size_t(k) * cap_per + found] = node.index;
                ++found;
            } else {
                stack.push_back(node.child1);
                stack.push_back(node.child2);
            }
        }
        counts[k] = found;
    }
    return BX_OK;
}

int bx_mesh_wings(const bx_mesh_desc* mesh, uint32_t* out) {
    if (!out) return Fail(BX_E_INVALID, "null argument");
    HostLibrary L;
    int rc = BuildOneMesh(mesh, L);
    if (rc) return rc;
    for (size_t i = 0; i < L.meshTris.size(); ++i) {
        out[3 * i] = L.meshTris[i].u1;
        out[3 * i + 1] = L.meshTris[i].u2;
        out[3 * i + 2] = L.meshTris[i].u3;
    }
    return BX_OK;
}

}  // extern "C"
