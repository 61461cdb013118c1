// bx_types.h -- device data layout of a batch of worlds.
//
// All worlds of a batch live in flat arrays; each world owns contiguous ranges (WorldInfo).
// Indices stored inside a world (bodies, fixtures, contact slots) are world-local.
//
// Contacts: each world has a pool of contact slots and an order list holding the live slots in
// creation order. The reference keeps contacts in linked lists with prepend insertion
// (contact_manager.cpp:105-139), so its world list and every fixture's edge list run from the
// newest contact to the oldest: the reverse of the order list. Creation order inside one
// FindNewContacts call is the sorted proxy-pair order (broad_phase.h:184-206).
#pragma once
#include "bx_math.h"

// Contact listener events (BX_WORLD_EVENTS) and contact filter tables in the kernels and their host
// side: compiled in when BX_DEVICE_CALLBACKS is nonzero (the host build of these kernels sets it). The
// GPU library leaves them out: a device test of event recording ended in a GPU memory fault whose
// cause has not been found, and without them its kernels are those that run clean on the GPU.
#ifndef BX_DEVICE_CALLBACKS
#define BX_DEVICE_CALLBACKS 0
#endif

namespace bx {

struct HalfEdge { u32 origin, twin, face, prev, next; };

struct HullInfo {
    u32 vBase, vCount;
    u32 eBase, eCount;
    u32 fBase, fCount;
    scalar cx, cy, cz;  // centroid
};

// Read-only view of one hull in the shape library (b3Hull, hull.h:37-152).
struct HullView {
    const Vec3* vertices;
    const HalfEdge* edges;
    const u32* faces;
    const Plane* planes;
    Vec3 centroid;
    u32 vertexCount, edgeCount, faceCount;

    BX_HD u32 GetSupportVertex(const Vec3& d) const {
        u32 maxIndex = 0;
        scalar maxProjection = Dot(d, vertices[0]);
        for (u32 i = 1; i < vertexCount; ++i) {
            scalar p = Dot(d, vertices[i]);
            if (p > maxProjection) { maxIndex = i; maxProjection = p; }
        }
        return maxIndex;
    }
    BX_HD u32 GetSupportFace(const Vec3& d) const {
        u32 maxIndex = 0;
        scalar maxProjection = Dot(d, planes[0].normal);
        for (u32 i = 1; i < faceCount; ++i) {
            scalar p = Dot(d, planes[i].normal);
            if (p > maxProjection) { maxIndex = i; maxProjection = p; }
        }
        return maxIndex;
    }
    BX_HD Plane GetEdgeSidePlane(u32 index) const {
        const HalfEdge e = edges[index];
        const HalfEdge t = edges[e.twin];
        Vec3 P = vertices[e.origin];
        Vec3 Q = vertices[t.origin];
        Vec3 D = Cross(Q - P, planes[e.face].normal);
        Plane p;
        p.normal = Normalize(D);
        p.offset = Dot(p.normal, P);
        return p;
    }
};

// Triangle meshes of the shape library (b3Mesh, mesh.h:30-96; b3StaticTree, static_tree.h:30-110).
struct MeshInfo {
    u32 vBase, vCount;      // vertices
    u32 tBase, tCount;      // triangles
    u32 nBase, root;        // tree nodes, root node (mesh-local)
    u32 pad0, pad1;
};

// b3MeshTriangle with its edge wings (b3MeshTriangleWings, mesh.cpp:35-112): the vertex across
// edge (v1, v2), (v2, v3), (v3, v1) in the adjacent triangle, kNull on a boundary edge.
struct MeshTri { u32 v1, v2, v3, u1, u2, u3; };

// b3StaticTree node: a leaf has child1 == kNull and index = its triangle.
struct TreeNode {
    AABB aabb;
    u32 child1, child2, index, pad;
};
constexpr u32 kTreeStack = 64;  // traversal stack of a tree query; the host rejects deeper trees

// b3DynamicTree::MoveProxy (dynamic_tree.cpp:139-207) of a proxy whose fat AABB is `tree`: the shape
// now covers `aabb` and is predicted to move by `displacement`. Returns whether the proxy moved (its fat
// AABB is replaced by the new, extended and displaced one, and it is buffered for the pair search).
BX_HD bool MoveFatAABB(AABB& tree, const AABB& aabb, const Vec3& displacement) {
    const Vec3 d = kAabbMultiplier * displacement;
    AABB fat = aabb;
    fat.Extend(kAabbExtension);
    if (d.x < scalar(0)) fat.lowerBound.x += d.x; else fat.upperBound.x += d.x;
    if (d.y < scalar(0)) fat.lowerBound.y += d.y; else fat.upperBound.y += d.y;
    if (d.z < scalar(0)) fat.lowerBound.z += d.z; else fat.upperBound.z += d.z;
    if (tree.Contains(aabb)) {
        AABB huge = fat;
        huge.Extend(scalar(4) * kAabbExtension);
        if (huge.Contains(tree)) return false;
    }
    tree = fat;
    return true;
}

// Body flags: b3Body::e_* (body.h:379-384).
constexpr u32 kAwake = 0x1, kIsland = 0x2, kAutoSleep = 0x4;
constexpr u32 kFixedX = 0x8, kFixedY = 0x10, kFixedZ = 0x20;
constexpr u32 kStatic = 0, kKinematic = 1, kDynamic = 2;

struct Body {
    Vec3 x;            // sweep.worldCenter
    Quat q;            // sweep.orientation == xf.rotation
    Vec3 x0;           // sweep.worldCenter0
    Quat q0;           // sweep.orientation0
    Vec3 localCenter;
    Vec3 xfT;          // xf.translation
    Vec3 v, w;
    Vec3 force, torque;
    scalar invMass, mass;
    Mat33 I, invI, worldInvI;
    Vec3 linearDamping, angularDamping, gravityScale;
    scalar linTol, angTol, sleepTime;
    u32 type, flags, islandIndex;
    u32 fixtureBegin, fixtureCount;   // into WorldInfo fixture-list range, list order
    u32 jointBegin, jointCount;       // into the joint-edge range, list order
};

struct Fixture {
    u32 body;
    u32 shapeType;
    u32 shapeIndex;
    u32 proxyId;
    scalar radius;
    Vec3 v1, v2, v3;
    scalar density, friction, restitution;
    u32 isSensor;
};

struct ManifoldPoint {
    Vec3 localNormal1, localPoint1, localPoint2;
    scalar normalImpulse;
    u32 triangleKey;
    u32 key[4];        // b3FeaturePair: inEdge1 outEdge1 inEdge2 outEdge2
    u32 persisting, edgeContact;
};

struct Manifold {
    ManifoldPoint points[kMaxManifoldPoints];
    u32 pointCount;
    Vec2 tangentImpulse;
    scalar motorImpulse, motorSpeed, tangentSpeed1, tangentSpeed2;
};

struct FeatureCache { u32 cacheType, featureType, index1, index2; };

struct Contact {
    u32 fixtureA, fixtureB;
    u32 flags;          // 1 island, 2 touching (contact.h)
    u32 manifoldCount;
    FeatureCache cache;
    Manifold manifold;  // convex contacts carry one manifold (convex_contact.cpp:23-30)
};

// State of a mesh contact besides its first manifold (b3MeshContact, mesh_contact.h:50-80): the
// fat AABB of the other shape in the mesh's unscaled frame, whether it moved since the triangle
// query, the queried triangle count (the triangles themselves are in BatchView::meshTriLists) and
// manifolds 2 and 3 of the up to kMaxManifolds cluster manifolds.
struct MeshContact {
    AABB meshAABB;
    u32 moved, triangleCount;
    Manifold extra[kMaxManifolds - 1];
};

// One contact point of a mesh contact's per-triangle manifolds during clustering (cluster.cpp:
// 471-584): world normal, world point (moved onto the cluster plane during reduction), separation,
// and (triangle list position << 2 | point).
struct MeshObs {
    Vec3 n, p;
    scalar sep;
    u32 tag;
};

// Lanes of the per-world contact update and the stripe of MeshObs each lane owns there: its three
// cluster manifolds, up to four observations per triangle and the member list of one cluster.
constexpr u32 kMeshLanes = 256;
constexpr u32 kMeshManifoldObs = u32((kMaxManifolds * sizeof(Manifold) + sizeof(MeshObs) - 1) / sizeof(MeshObs));
BX_HD u32 MeshLaneStride(u32 triCap) { return kMeshManifoldObs + 4u * triCap + (4u * triCap + 7u) / 8u; }

struct JointEdge { u32 joint, other; };

// b3JointType values of the joints the device solves (joint.h:33-46).
constexpr u32 kJointMouse = 1, kJointSpring = 2, kJointWeld = 3, kJointRevolute = 4, kJointSphere = 5,
              kJointCone = 6, kJointFriction = 7, kJointMotor = 8, kJointPrismatic = 9, kJointWheel = 10;

// A joint (joints/*_joint.h): definition and the warm-start state that persists between steps. The
// revolute members come first; the other types use the anchors, the reference rotation, the linear
// impulse and the block after `type`.
struct Joint {
    u32 bodyA, bodyB;
    u32 collideConnected, islandFlag;
    Vec3 localAnchorA, localAnchorB;
    Quat localRotationA, localRotationB, referenceRotation;
    u32 enableLimit, enableMotor;
    scalar lowerAngle, upperAngle, motorSpeed, maxMotorTorque;
    scalar motorImpulse, limitImpulse;
    Vec3 linearImpulse;                // revolute, sphere, weld, mouse, friction, motor
    Vec2 angularImpulse;
    u32 limitState;
    u32 type;                          // kJoint*
    scalar frequencyHz, dampingRatio, length, maxForce, maxTorque, correctionFactor;
    Vec3 target;                       // mouse: world target
    Vec3 linearOffset;                 // motor
    Quat angularOffset;                // motor
    Vec3 angularImpulse3;              // weld, friction, motor, prismatic
    scalar impulse1;                   // spring; wheel: angular impulse
    // cone, prismatic and wheel joints
    Vec3 localXAxisA, localYAxisA, localZAxisA, localXAxisB;
    scalar coneAngle, lowerTranslation, upperTranslation, referenceCosine, referenceAngle;
    Vec2 linearImpulse2;               // prismatic, wheel: the two perpendicular rows
    scalar springImpulse;              // wheel
};

// Solver constants of one island joint (b3RevoluteJoint's per-solve members), built in parallel
// before the island sweeps; the sweeps update its impulses.
struct JointCon {
    u32 indexA, indexB;
    u32 enableMotor, enableLimit, limitState;
    scalar mA, mB;
    Mat33 iA, iB;
    Vec3 motorAxis;
    scalar motorMass;
    Vec3 rA, rB;
    Mat33 linearMass;
    Vec3 a1, a2;
    Mat22 angularMass;
    scalar motorSpeed, maxMotorTorque;
    scalar motorImpulse, limitImpulse;
    Vec3 linearImpulse;
    Vec2 angularImpulse;
    // position solver inputs
    Vec3 localAnchorA, localAnchorB, localCenterA, localCenterB;
    Mat33 localInvIA, localInvIB;
    Quat localRotationA, localRotationB;
    scalar lowerAngle, upperAngle;
    // the other joint types (linearMass, rA, rB and linearImpulse are shared with the revolute joint)
    u32 type;
    scalar frequencyHz, length, maxForce, maxTorque, correctionFactor, invDt;
    Mat33 gamma3, angularMass3;
    Vec3 bias3, n, linearError, angularVelocity, angularImpulse3;
    scalar mass1, gamma1, bias1, impulse1;
    // cone, prismatic and wheel joints: motorAxis holds the cone axis, the prismatic axis or the
    // wheel's axis on A; mass1 the cone or spring mass; angularMass3 and bias3 the prismatic
    // angular row
    Vec3 perp1, perp2, s1, s2, s3, s4, axisB, u;
    Mat22 mass22;
    scalar angularMass1, coneAngle, referenceCosine, referenceAngle, springImpulse;
    Vec3 localXAxisA, localYAxisA, localZAxisA, localXAxisB;
    Vec2 linearImpulse2;
};


struct WorldInfo {
    u32 bodyBase, bodyCount;
    u32 fixtureBase, fixtureCount;     // fixtures in creation order (world-local index)
    u32 fixtureListBase;               // per-body fixture lists (list order), fixtureCount entries
    u32 jointBase, jointCount;
    u32 jointEdgeBase;                 // 2 * jointCount entries
    u32 contactBase, contactCap;       // slot pool, order list, free stack, island contact list
    u32 pairBase, pairCap;
    u32 islandBodyBase, islandBodyCap; // island body entries (statics may repeat across islands)
    u32 flags;                         // BX_WORLD_* flags
    u32 serialIslands;                 // islands must be solved in order (see bx_step.hip)
    Vec3 gravity;
    u32 islandScratchBase;             // words in BatchView::islandScratch (used when LDS is too small)
    u32 islandScratchWords;            // words the island builder needs for this world at full capacity
    u32 edgeBase;                      // CSR offsets of the contact edges (fixtureCount + 1 entries)
    u32 conBase, conCap;               // island constraint entries (one per manifold of an island contact)
    u32 meshBase;                      // mesh contact records (contactCap when the world has a mesh fixture)
    u32 meshTriBase;                   // mesh contact triangle lists (contactCap x meshTriCap)
    u32 meshTriCap;                    // triangles a mesh contact can hold (0: no mesh fixture)
    u32 meshScratchBase;               // clustering stripes (kMeshLanes x MeshLaneStride(meshTriCap))
#if BX_DEVICE_CALLBACKS
    u32 pad0;
    // contact filter tables (b3ContactFilter) or null: bit i * fixtureCount + j of the first table is
    // ShouldCollide(fixture i, fixture j), of the second (FilterWords(fixtureCount) words later)
    // ShouldRespond(i, j)
    const u32* filter;
#endif
};

BX_HD u32 FilterWords(u32 nf) { return u32((u64(nf) * nf + 31) / 32); }
BX_HD bool FilterBit(const u32* table, u32 nf, u32 i, u32 j) {
    const u32 bit = i * nf + j;
    return ((table[bit >> 5] >> (bit & 31u)) & 1u) != 0;
}

// World flag: record contact events (bx_batch_get_events).
constexpr u32 kWorldEvents = 0x8;
// Contact event bits of one visited contact in b3ContactManager::UpdateContacts: destroyed (b3Contact
// removed; EndContact when it was touching), BeginContact, EndContact, PreSolve (contact.cpp:268-284,
// contact_manager.cpp:230-237).
constexpr u32 kEvDestroyed = 1, kEvBegin = 2, kEvEnd = 4, kEvPreSolve = 8;

// Per-world mutable counters.
struct WorldState {
    u32 contactCount;     // live contacts (order list length)
    u32 freeCount;        // free slot stack length
    u32 pairCount;        // unique pairs of the last FindPairs
    u32 islandCount;
    u32 islandBodyCount;
    u32 islandContactCount;
    u32 islandJointCount;
    u32 newContacts;
    u32 status;
    scalar inv_dt0;
    u32 touchingCount;
    u32 anyAsleep;
#if BX_DEVICE_CALLBACKS
    u32 eventCount;       // contacts visited by the last contact update (entries of BatchView::events)
    u32 pad;
#endif
};

// Words of island-builder scratch a world needs: per body flags, fixture range, joint range,
// island index and DFS stack; per fixture its list entry and owner, CSR offsets and cursors; per
// contact slot its two fixtures, flags and two edge entries.
BX_HD u32 IslandScratchWords(u32 nb, u32 nf, u32 cap) { return 5 * nb + 4 * nf + 2 + 5 * cap; }

struct IslandInfo {
    u32 bodyStart, bodyCount;
    u32 contactStart, contactCount;
    u32 jointStart, jointCount;
    u32 positionsSolved, pad;
};

// Solver state of one island body (b3Velocity / b3Position, time_step.h), split so that the
// velocity iterations touch only 24 bytes per body.
struct VelBody { Vec3 v, w; };
struct PosBody { Vec3 x; Quat q; Mat33 I; };

// Velocity constraint of one convex contact (b3ContactVelocityConstraint with its single
// manifold, contact_solver.h:57-111). Impulses are the only fields the iterations write.
struct VCPoint { Vec3 rA, rB, normal; scalar normalMass, normalImpulse, velocityBias; };
struct VelCon {
    u32 indexA, indexB, pointCount, pad;
    scalar invMassA, invMassB, friction, restitution;
    Mat33 invIA, invIB;
    Vec3 rA, rB, normal, tangent1, tangent2;
    scalar tangentSpeed1, tangentSpeed2, motorSpeed, motorMass;
    Mat22 tangentMass;
    Vec2 tangentImpulse;
    scalar motorImpulse;
    VCPoint points[kMaxManifoldPoints];
};

// Position constraint of one convex contact (b3ContactPositionConstraint, contact_solver.h:31-55).
struct PosCon {
    u32 indexA, indexB, pointCount, pad;
    scalar invMassA, invMassB, radiusA, radiusB;
    Vec3 localCenterA, localCenterB;
    Mat33 localInvIA, localInvIB;
    Vec3 localNormalA[kMaxManifoldPoints];
    Vec3 localPointA[kMaxManifoldPoints];
    Vec3 localPointB[kMaxManifoldPoints];
};

// Layouts of the velocity sweeps: islands per 64-wide workgroup and the largest island whose
// velocities live in LDS. Layouts 0-2 (k_solve_velocity) give each island one lane and request
// 96 KB of LDS, so one workgroup runs per CU; layouts 3-4 (k_solve_velocity_quad) give each island
// eight lanes that share each contact's vector arithmetic. The layout changes only which lanes
// solve which island, never the order of operations inside an island, so all layouts give
// identical results. A wavefront's sweep is a serial chain whose instruction count does not depend
// on how many of its lanes are active: few islands per wavefront spread a small island count over
// more CUs, many islands per wavefront need fewer wavefronts when islands are many.
constexpr u32 kSolveLayouts = 5;
constexpr u32 kSolveLayoutLanes[kSolveLayouts] = {16, 32, 64, 4, 8};
constexpr u32 kSolveLayoutSlab[kSolveLayouts] = {256, 128, 64, 320, 224};
constexpr bool SolveLayoutQuad(u32 layout) { return layout >= 3; }
// The eight-lane layouts solve positions the same way (k_solve_position_quad), with a 32-byte pose
// per body in LDS for islands of at most this many bodies.
constexpr u32 kPosQuadSlab = 224;

// Everything a step kernel needs; all pointers are device pointers.
struct BatchView {
    u32 worldCount;
    const u32* worldList;     // launched worlds (block i steps world worldList[i]); null: block i steps world i
    const WorldInfo* worlds;
    WorldState* wstate;
    // shape library
    const HullInfo* hulls;
    const Vec3* hullVertices;
    const HalfEdge* hullEdges;
    const u32* hullFaces;
    const Plane* hullPlanes;
    const MeshInfo* meshes;
    const Vec3* meshVertices;
    const MeshTri* meshTris;
    const TreeNode* meshNodes;
    // bodies / fixtures / joints
    Body* bodies;
    Fixture* fixtures;
    const u32* fixtureList;   // per world: bodies' fixture lists in list order (world-local fixture ids)
    const u32* proxyOrder;    // per world: fixtures sorted by proxy id
    AABB* fatAABB;            // per fixture
    u32* moved;               // per fixture: proxy is in the move buffer
    Joint* joints;
    const JointEdge* jointEdges;
    // contacts
    Contact* contacts;        // slot pool
    u32* contactOrder;        // per world: live slots, creation order
    u32* freeSlots;           // per world: free slot stack
    u32* contactAlive;        // per slot: scratch flag for the update pass
#if BX_DEVICE_CALLBACKS
    uint4* events;            // per world: the last contact update's visits {slot, fixtureA, fixtureB, kEv* bits}
#endif
    // adjacency (rebuilt each step)
    u32* edgeOffset;          // per fixture + 1 per world (fixtureCount + 1 entries from WorldInfo::edgeBase)
    u32* edgeSlots;           // per world: 2 * contactCap
    // pairs
    uint2* pairs;             // per world: pairCap
    u32* pairCounts;          // per fixture (world-local proxy order) scratch
    // the query hierarchy (bx_query.h), rebuilt after a step before the next query
    u32* bvhLeaves;           // per fixture: fixture indices sorted by the Morton code of their fat-AABB centre
    AABB* bvhNodes;           // per fixture: the node levels above them, bottom-up
    // islands
    IslandInfo* islands;      // per world: up to bodyCount islands (base = bodyBase)
    u32* islandBodies;        // per world: islandBodyCap
    uint4* islandContacts;    // per world: conCap entries {slot | manifold << 30, localA, localB, island bodyStart}
    uint4* islandJoints;      // per world: jointCount entries {joint, localA, localB, island bodyStart}
    JointCon* jointCons;      // per world: jointCount entries, island order
    u32* dfsStack;            // per world: bodyCount
    VelBody* velBodies;       // per world: islandBodyCap
    PosBody* posBodies;       // per world: islandBodyCap
    VelCon* velCons;          // per world: conCap
    PosCon* posCons;          // per world: conCap
    // mesh contacts (worlds with a mesh fixture)
    MeshContact* meshContacts; // per contact slot
    u32* meshTriLists;        // per contact slot: meshTriCap triangle indices, tree query order
    MeshObs* meshScratch;     // per world: kBlock x 4 x meshTriCap
    uint2* islandList;        // islands of the launched worlds: {world, island} (island = kNull: all, in order)
    u32* islandListCount;
    u32* islandScratch;       // island-builder working arrays of worlds whose arrays do not fit LDS
    u32 islandLdsWords;       // LDS words available to the island builder per world
    u32 meshWorlds;           // nonzero when some world holds a mesh fixture (mesh kernel variants)
    u32 solveLayout;          // layout of the velocity sweeps (index into kSolveLayoutLanes)
};

// Output of one ray of RayCastSingle / RayCast or one box of QueryAABB, as the C ABI lays it out.
struct RayHitOut { u32 fixture; scalar fraction; Vec3 point, normal; };

struct QueryParams {
    const uint4* queries;   // per query: {world, mask offset, -, -} then 6 floats (p1 p2 or lower upper)
    const float* coords;    // 6 per query
    const uint8_t* mask;
    u32 count;
    u32 kind;               // 0 single ray, 1 all ray hits, 2 boxes
    u32 perQuery;           // output slots per query (kinds 1 and 2)
    RayHitOut* hits;        // kinds 0 and 1
    u32* fixtures;          // kind 2
    u32* counts;            // kinds 1 and 2
};

// A convex shape as b3GJKProxy::Set takes it (gjk.cpp:37-95): sphere (v1 the centre), capsule (v1, v2),
// triangle (v1..v3) or hull `index` of the library.
struct CastShapeD { u32 type, index; scalar radius; Vec3 v1, v2, v3; };
// b3Sweep (sweep.h:27-63) without t0, which the time of impact does not read.
struct SweepD { Vec3 localCenter, c0, c; Quat q0, q; };

// One b3TimeOfImpact problem: kind 0 two sweeps up to tMax, 1 two translations (conservative
// advancement, at most maxIterations), 2 two moving boxes.
struct ToiQueryD {
    u32 kind, maxIterations;
    scalar tMax;
    CastShapeD s1, s2;
    SweepD w1, w2;
    Transform xf1, xf2;
    Vec3 d1, d2;
    AABB a1, a2;
};
// b3TOIOutput (time_of_impact.h:37-51).
struct ToiResultD { u32 state; scalar t; u32 iterations; };

// One b3World::ShapeCast of `shape` at `xf` moved by `d` in world `world`; `mask` is kNull or an offset
// into the call's mask bytes.
struct ShapeCastD { u32 world, mask; CastShapeD shape; Transform xf; Vec3 d; };
struct ShapeHitOut { u32 fixture, child; scalar fraction; Vec3 point, normal; };

struct CastParams {
    const ShapeCastD* casts;
    const uint8_t* mask;
    u32 count;
    u32 single;             // 1: the first hit only (ShapeCastSingle); 0: every hit, perQuery kept
    u32 perQuery;
    ShapeHitOut* hits;
    u32* counts;            // single == 0
};

// BX_WORLD_HOOK(view, w) marks where a workgroup or lane starts work on world w; it compiles to
// nothing in the device library (the CPU test build uses it to check that a world's kernels touch
// only that world's ranges).
#ifndef BX_WORLD_HOOK
#define BX_WORLD_HOOK(view, w)
#endif

// World stepped by this workgroup in the one-workgroup-per-world kernels.
__device__ __forceinline__ u32 WorldOfBlock(const BatchView& B) {
    const u32 w = B.worldList ? B.worldList[blockIdx.x] : blockIdx.x;
    BX_WORLD_HOOK(B, w);
    return w;
}

}  // namespace bx
