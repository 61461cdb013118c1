// bx_step.hip -- the batched b3World::Step on the GPU (gfx950).
//
// One step is a fixed sequence of kernels, each launched over all worlds of the batch
// (blockIdx.x = world, one workgroup per world):
//
//   k_find_pairs (only when new fixtures were added)   b3ContactManager::FindNewContacts
//   k_update_contacts   narrowphase + stable compaction   b3ContactManager::UpdateContacts
//   k_islands           fixture adjacency + island DFS    b3World::Solve (world.cpp:253-422)
//   k_solve             per-island integration + solver   b3Island::Solve (island.cpp:81-307)
//   k_sync              fixture AABBs / fat-AABB update   b3Body::SynchronizeFixtures, MoveProxy
//   k_find_pairs        pairs + contact creation          b3BroadPhase::FindPairs + AddPair
//
// Exactness: every floating-point formula is the reference's (see bx_math.h); every ordering
// decision the reference takes implicitly through list order is reproduced explicitly:
// contact order lists, fixture edge lists rebuilt in reference order, DFS over bodies in body-
// list order, Gauss-Seidel in island contact order. Work is parallel only where the reference's
// result is order-independent (per contact narrowphase, per fixture AABBs, per pair filtering,
// independent islands).
#include <algorithm>
#include "bx_collide.h"
#include "bx_convex.h"
#include "bx_triangle.h"
#include "bx_joint.h"
#include "bx_solver.h"

namespace bx {

constexpr int kBlock = 256;

__device__ __forceinline__ HullView GetHull(const BatchView& B, u32 index) {
    HullInfo hi = B.hulls[index];
    HullView h;
    h.vertices = B.hullVertices + hi.vBase;
    h.edges = B.hullEdges + hi.eBase;
    h.faces = B.hullFaces + hi.fBase;
    h.planes = B.hullPlanes + hi.fBase;
    h.centroid = Vec3(hi.cx, hi.cy, hi.cz);
    h.vertexCount = hi.vCount;
    h.edgeCount = hi.eCount;
    h.faceCount = hi.fCount;
    return h;
}

// Block-wide exclusive scan of one u32 per thread; returns the prefix and writes the total.
__device__ u32 BlockExclusiveScan(u32 value, u32* total) {
    __shared__ u32 warpSums[kBlock / 64 + 1];
    const u32 lane = threadIdx.x & 63u;
    const u32 warp = threadIdx.x >> 6;
    u32 x = value;
    for (u32 d = 1; d < 64; d <<= 1) {
        u32 y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) warpSums[warp] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 run = 0;
        for (u32 i = 0; i < blockDim.x / 64; ++i) {
            u32 s = warpSums[i];
            warpSums[i] = run;
            run += s;
        }
        warpSums[kBlock / 64] = run;
    }
    __syncthreads();
    u32 result = x - value + warpSums[warp];
    *total = warpSums[kBlock / 64];
    __syncthreads();
    return result;
}

__device__ __forceinline__ bool BodiesShouldCollide(const BatchView& B, const WorldInfo& wi, u32 a, u32 b) {
    // b3Body::ShouldCollide (body.cpp:240-263)
    const Body& A = B.bodies[wi.bodyBase + a];
    const Body& Bb = B.bodies[wi.bodyBase + b];
    if (A.type != kDynamic && Bb.type != kDynamic) return false;
    for (u32 k = 0; k < A.jointCount; ++k) {
        JointEdge je = B.jointEdges[wi.jointEdgeBase + A.jointBegin + k];
        if (je.other == b && B.joints[wi.jointBase + je.joint].collideConnected == 0) return false;
    }
    return true;
}

__device__ __forceinline__ void WakeBody(Body& b) {  // b3Body::SetAwake(true)
    if ((b.flags & kAwake) == 0) {
        atomicOr(&b.flags, kAwake);
        b.sleepTime = 0;
    }
}

// ================================================================== shape AABBs
__device__ void ComputeAABB(const BatchView& B, const Fixture& f, const Transform& xf, AABB& out) {
    if (f.shapeType == 3u) {  // hull_shape.cpp:160-164, aabb.h:53-62
        HullView h = GetHull(B, f.shapeIndex);
        out.lowerBound = out.upperBound = Mul(xf, h.vertices[0]);
        for (u32 i = 1; i < h.vertexCount; ++i) {
            Vec3 v = Mul(xf, h.vertices[i]);
            out.lowerBound = Min(out.lowerBound, v);
            out.upperBound = Max(out.upperBound, v);
        }
        out.Extend(f.radius);
    } else if (f.shapeType == 0u) {  // sphere_shape.cpp:51-57
        Vec3 c = Mul(xf, f.v1);
        Vec3 r(f.radius, f.radius, f.radius);
        out.lowerBound = c - r;
        out.upperBound = c + r;
    } else if (f.shapeType == 1u) {  // capsule_shape.cpp:124-131
        Vec3 c1 = Mul(xf, f.v1), c2 = Mul(xf, f.v2);
        Vec3 r(f.radius, f.radius, f.radius);
        out.lowerBound = Min(c1, c2) - r;
        out.upperBound = Max(c1, c2) + r;
    } else if (f.shapeType == 2u) {  // triangle_shape.cpp:75-83: local bounds, the transform is unused
        Vec3 r(f.radius, f.radius, f.radius);
        out.lowerBound = Min(f.v1, Min(f.v2, f.v3)) - r;
        out.upperBound = Max(f.v1, Max(f.v2, f.v3)) + r;
    } else if (f.shapeType == 4u) {  // mesh_shape.cpp:44-50, aabb.h:65-74
        const MeshInfo mi = B.meshes[f.shapeIndex];
        const Vec3* v = B.meshVertices + mi.vBase;
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

#include "bx_mesh.h"

// ================================================================== update contacts
// b3Manifold::Initialize(const b3Manifold&) (collision.cpp:40-61): impulses of `old` carried over to
// the points of `m` with the same contact id.
__device__ __forceinline__ void MatchManifold(Manifold& m, const Manifold& old) {
    m.tangentImpulse = old.tangentImpulse;
    m.motorImpulse = old.motorImpulse;
    for (u32 i = 0; i < m.pointCount; ++i) {
        ManifoldPoint& p1 = m.points[i];
        for (u32 j = 0; j < old.pointCount; ++j) {
            const ManifoldPoint& p2 = old.points[j];
            if (p2.triangleKey == p1.triangleKey && p2.key[0] == p1.key[0] && p2.key[1] == p1.key[1] &&
                p2.key[2] == p1.key[2] && p2.key[3] == p1.key[3]) {
                p1.normalImpulse = p2.normalImpulse;
                p1.persisting = 1;
                break;
            }
        }
    }
}

// b3Contact::Update (contact.cpp:218-262) of a mesh contact: b3Manifold::Initialize of its
// manifolds, b3MeshContact::Collide, then every new manifold matched against every old one in
// order (a point takes the impulse of its match in the last old manifold that has one; tangent and
// motor impulses come from the last old manifold). Returns touching.
__device__ __noinline__ bool UpdateMeshManifolds(const BatchView& B, const WorldInfo& wi, WorldState& ws, u32 slot,
                                                 Contact& c, const Fixture& fA, const Fixture& fB,
                                                 const Transform& xfA, const Transform& xfB, FaceScratch& scratch) {
    MeshObs* lane = B.meshScratch + wi.meshScratchBase + threadIdx.x * MeshLaneStride(wi.meshTriCap);
    Manifold* nm = reinterpret_cast<Manifold*>(lane);
    for (u32 i = 0; i < kMaxManifolds; ++i) {
        Manifold& m = nm[i];
        m.pointCount = 0;
        m.tangentImpulse.SetZero();
        m.motorImpulse = m.motorSpeed = m.tangentSpeed1 = m.tangentSpeed2 = 0;
        for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
            m.points[k].normalImpulse = 0;
            m.points[k].persisting = 0;
            m.points[k].edgeContact = 0;
        }
    }
    const u32 oldCount = c.manifoldCount;
    u32 overflow = 0;
    const u32 count = MeshCollide(B, wi, slot, fA, fB, xfA, xfB, lane, nm, scratch, overflow);
    if (overflow) atomicOr(&ws.status, 0x2u);
    if ((wi.flags & 0x2u) == 0)
        for (u32 i = 0; i < count; ++i)
            for (u32 j = 0; j < oldCount; ++j) MatchManifold(nm[i], ContactManifold(B, wi, c, slot, j));
    bool touching = false;
    for (u32 i = 0; i < count; ++i) {
        if (nm[i].pointCount > 0) touching = true;
        ContactManifold(B, wi, c, slot, i) = nm[i];
    }
    c.manifoldCount = count;
    return touching;
}

// b3ContactManager::UpdateContacts (contact_manager.cpp:165-227) + b3Contact::Update
// (contact.cpp:183-285) for one contact. Returns false if the contact is destroyed. kMesh compiles
// the mesh contact path in (launches over batches that hold a mesh fixture).
template <bool kMesh>
__device__ bool UpdateOneContact(const BatchView& B, const WorldInfo& wi, WorldState& ws, u32 slot, bool applyWakes
#if BX_DEVICE_CALLBACKS
                                 , u32& events  // kEv* bits of this visit
#endif
) {
    Contact& c = B.contacts[wi.contactBase + slot];
    const Fixture fA = B.fixtures[wi.fixtureBase + c.fixtureA];
    const Fixture fB = B.fixtures[wi.fixtureBase + c.fixtureB];
    Body& bA = B.bodies[wi.bodyBase + fA.body];
    Body& bB = B.bodies[wi.bodyBase + fB.body];
    bool sensor = fA.isSensor || fB.isSensor;
    const bool isMesh = kMesh && (fA.shapeType == kMeshShape || fB.shapeType == kMeshShape);
    FaceScratch scratch;  // hull clipping buffers of every narrowphase below
#if BX_DEVICE_CALLBACKS
    events = 0;
#endif
    auto destroy = [&]() {
#if BX_DEVICE_CALLBACKS
        events = kEvDestroyed | ((c.flags & 2u) ? kEvEnd : 0u);  // contact_manager.cpp:230-237
#endif
        if (applyWakes && !sensor) {
            for (u32 j = 0; j < c.manifoldCount; ++j) {
                if (ContactManifold(B, wi, c, slot, j).pointCount > 0) {
                    WakeBody(bA);
                    WakeBody(bB);
                    break;
                }
            }
        }
        return false;
    };
    if (!BodiesShouldCollide(B, wi, fA.body, fB.body)) return destroy();
#if BX_DEVICE_CALLBACKS
    // b3ContactFilter::ShouldCollide (contact_manager.cpp:189-199)
    if (wi.filter && !FilterBit(wi.filter, wi.fixtureCount, c.fixtureA, c.fixtureB)) return destroy();
#endif
    bool activeA = (bA.flags & kAwake) && bA.type != kStatic;
    bool activeB = (bB.flags & kAwake) && bB.type != kStatic;
    if (!activeA && !activeB) return true;
    if (!TestOverlap(B.fatAABB[wi.fixtureBase + c.fixtureA], B.fatAABB[wi.fixtureBase + c.fixtureB])) return destroy();

    bool wasTouching = (c.flags & 2u) != 0;
    bool touching = false;
    if (sensor && isMesh) {
        touching = MeshContactTestOverlap(B, wi, slot, c, Transform(bA.xfT, bA.q), Transform(bB.xfT, bB.q));
        c.manifoldCount = 0;
    } else if (isMesh) {
        touching = UpdateMeshManifolds(B, wi, ws, slot, c, fA, fB, Transform(bA.xfT, bA.q), Transform(bB.xfT, bB.q),
                                       scratch);
        if (touching != wasTouching && applyWakes) {
            WakeBody(bA);
            WakeBody(bB);
        }
    } else if (sensor) {
        // b3ConvexContact::TestOverlap (convex_contact.cpp:34-43)
        Transform xfA(bA.xfT, bA.q), xfB(bB.xfT, bB.q);
        HullView hA, hB;
        if (fA.shapeType == 3u) hA = GetHull(B, fA.shapeIndex);
        if (fB.shapeType == 3u) hB = GetHull(B, fB.shapeIndex);
        touching = ConvexTestOverlap(xfA, MakeProxy(fA, &hA), xfB, MakeProxy(fB, &hB));
        c.manifoldCount = 0;
    } else {
        const u32 oldCount = c.manifoldCount;
        // the last step's manifold, read in place: the contact record is written only at the end
        const Manifold& old = c.manifold;
        // b3Manifold::Initialize (collision.cpp:23-38). Every narrowphase writes the geometry and id
        // of each point it reports, so the new manifold starts from the initialized fields only;
        // the record's unused point slots keep their old contents, as the reference's reused
        // manifold storage does.
        Manifold m;
        m.pointCount = 0;
        m.tangentImpulse.SetZero();
        m.motorImpulse = 0;
        m.motorSpeed = 0;
        m.tangentSpeed1 = 0;
        m.tangentSpeed2 = 0;
        for (u32 i = 0; i < kMaxManifoldPoints; ++i) {
            m.points[i].normalImpulse = 0;
            m.points[i].persisting = 0;
            m.points[i].edgeContact = 0;
        }
        Transform xfA(bA.xfT, bA.q), xfB(bB.xfT, bB.q);
        if (fA.shapeType == 3u && fB.shapeType == 3u) {
            // b3HullContact::Evaluate (hull_contact.cpp:41-47)
            Transform xf0A = SweepTransform(bA.x0, bA.q0, bA.x, bA.q, bA.localCenter, scalar(0));
            Transform xf0B = SweepTransform(bB.x0, bB.q0, bB.x, bB.q, bB.localCenter, scalar(0));
            HullView hA = GetHull(B, fA.shapeIndex), hB = GetHull(B, fB.shapeIndex);
            u32 overflow = 0;
            FeatureCache cache = c.cache;
            CollideHulls(m, xfA, hA, fA.radius, xfB, hB, fB.radius, cache, xf0A, xf0B, scratch, overflow);
            c.cache = cache;
            if (overflow) atomicOr(&ws.status, 0x2u);
        } else if (fA.shapeType == 0u && fB.shapeType == 0u) {  // sphere_contact.cpp:42-45
            CollideSpheres(m, xfA, fA, xfB, fB);
        } else if (fA.shapeType == 1u && fB.shapeType == 0u) {  // capsule_sphere_contact.cpp:42-45
            CollideCapsuleAndSphere(m, xfA, fA, xfB, fB);
        } else if (fA.shapeType == 1u && fB.shapeType == 1u) {  // capsule_contact.cpp:41-44
            CollideCapsules(m, xfA, fA, xfB, fB);
        } else if (fA.shapeType == 3u && fB.shapeType == 0u) {  // hull_sphere_contact.cpp:42-45
            HullView hA = GetHull(B, fA.shapeIndex);
            CollideHullAndSphere(m, xfA, hA, fA, xfB, fB);
        } else if (fA.shapeType == 3u && fB.shapeType == 1u) {  // hull_capsule_contact.cpp:42-45
            HullView hA = GetHull(B, fA.shapeIndex);
            CollideHullAndCapsule(m, xfA, hA, fA, xfB, fB);
        } else if (fA.shapeType == 2u && fB.shapeType == 0u) {  // triangle_sphere_contact.cpp
            CollideTriangleAndSphere(m, xfA, fA, xfB, fB);
        } else if (fA.shapeType == 2u && fB.shapeType == 1u) {  // triangle_capsule_contact.cpp
            CollideTriangleAndCapsule(m, xfA, fA, xfB, fB);
        } else if (fA.shapeType == 3u && fB.shapeType == 2u) {  // hull_triangle_contact.cpp
            HullView hA = GetHull(B, fA.shapeIndex);
            u32 overflow = 0;
            CollideHullAndTriangle(m, xfA, hA, fA.radius, xfB, fB, scratch, overflow);
            if (overflow) atomicOr(&ws.status, 0x2u);
        } else {
            atomicOr(&ws.status, 0x200u);  // shape pair without a device narrowphase
        }
        c.manifoldCount = 1;
        if ((wi.flags & 0x2u) == 0 && oldCount > 0) MatchManifold(m, old);  // no matching without warm starting
        Manifold& dst = c.manifold;
        dst.pointCount = m.pointCount;
        dst.tangentImpulse = m.tangentImpulse;
        dst.motorImpulse = m.motorImpulse;
        dst.motorSpeed = m.motorSpeed;
        dst.tangentSpeed1 = m.tangentSpeed1;
        dst.tangentSpeed2 = m.tangentSpeed2;
        for (u32 i = 0; i < m.pointCount; ++i) dst.points[i] = m.points[i];
        touching = m.pointCount > 0;
        if (touching != wasTouching && applyWakes) {
            WakeBody(bA);
            WakeBody(bB);
        }
    }
    if (touching) c.flags |= 2u;
    else c.flags &= ~2u;
#if BX_DEVICE_CALLBACKS
    // the listener calls of b3Contact::Update (contact.cpp:268-284)
    if (!wasTouching && touching) events |= kEvBegin;
    if (wasTouching && !touching) events |= kEvEnd;
    if (!sensor && touching) events |= kEvPreSolve;
#endif
    return true;
}

#ifndef BX_UC_BLOCKS_PER_CU
#define BX_UC_BLOCKS_PER_CU 4
#endif
template <bool kMesh>
__global__ void __launch_bounds__(kBlock, BX_UC_BLOCKS_PER_CU) k_update_contacts(BatchView B) {
    const u32 w = WorldOfBlock(B);
    const WorldInfo wi = B.worlds[w];
    WorldState& ws = B.wstate[w];
    const u32 n = ws.contactCount;
    u32* order = B.contactOrder + wi.contactBase;
    u32* alive = B.contactAlive + wi.contactBase;
    if (blockIdx.x == 0 && threadIdx.x == 0) *B.islandListCount = 0;  // the island builder refills it
    // The reference visits contacts newest-first and a contact can wake a sleeping body that a
    // later contact tests for activity; with every non-static body awake the visits commute.
    int asleep = 0;
    for (u32 i = threadIdx.x; i < wi.bodyCount; i += blockDim.x) {
        const Body& b = B.bodies[wi.bodyBase + i];
        if (b.type != kStatic && (b.flags & kAwake) == 0) asleep = 1;
    }
    asleep = __syncthreads_or(asleep);
#if BX_DEVICE_CALLBACKS
    const bool record = (wi.flags & kWorldEvents) != 0;
    auto update = [&](u32 i) {
        u32 ev = 0;
        const u32 slot = order[i];
        alive[i] = UpdateOneContact<kMesh>(B, wi, ws, slot, true, ev) ? 1u : 0u;
        if (record) {
            const Contact& c = B.contacts[wi.contactBase + slot];
            B.events[wi.contactBase + i] = make_uint4(slot, c.fixtureA, c.fixtureB, ev);
        }
    };
    if (!asleep) {
        for (u32 i = threadIdx.x; i < n; i += blockDim.x) update(i);
    } else if (threadIdx.x == 0) {
        for (u32 k = n; k-- > 0;) update(k);
    }
    if (record && threadIdx.x == 0) ws.eventCount = n;
#else
    if (!asleep) {
        for (u32 i = threadIdx.x; i < n; i += blockDim.x)
            alive[i] = UpdateOneContact<kMesh>(B, wi, ws, order[i], true) ? 1u : 0u;
    } else if (threadIdx.x == 0) {
        for (u32 k = n; k-- > 0;) alive[k] = UpdateOneContact<kMesh>(B, wi, ws, order[k], true) ? 1u : 0u;
    }
#endif
    __syncthreads();
    // stable compaction of the order list; destroyed slots go back on the free stack
    u32 kept = 0, freed = 0;
    u32 freeBase = ws.freeCount;
    __syncthreads();
    for (u32 base = 0; base < n; base += blockDim.x) {
        u32 i = base + threadIdx.x;
        u32 a = i < n ? alive[i] : 0u;
        u32 d = (i < n && !a) ? 1u : 0u;
        u32 slot = i < n ? order[i] : 0u;
        u32 totalA, totalD;
        u32 pa = BlockExclusiveScan(a, &totalA);
        u32 pd = BlockExclusiveScan(d, &totalD);
        __syncthreads();
        if (a) order[kept + pa] = slot;
        if (d) B.freeSlots[wi.contactBase + freeBase + freed + pd] = slot;
        kept += totalA;
        freed += totalD;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        ws.contactCount = kept;
        ws.freeCount = freeBase + freed;
    }
}

#include "bx_islands.h"
#include "bx_solve_kernels.h"
#include "bx_solve_quad.h"
#include "bx_query.h"
#include "bx_toi.h"

// ================================================================== fixture sync + clear forces
__global__ void __launch_bounds__(kBlock) k_sync(BatchView B, StepParams sp) {
    const u32 w = WorldOfBlock(B);
    const WorldInfo wi = B.worlds[w];
    const bool solved = sp.dt > scalar(0);
    if (threadIdx.x == 0 && solved) B.wstate[w].inv_dt0 = sp.inv_dt;  // world.cpp:233-236
    for (u32 i = threadIdx.x; i < wi.bodyCount; i += blockDim.x) {
        Body& b = B.bodies[wi.bodyBase + i];
        if (solved && (b.flags & kIsland) && b.type != kStatic) {
            // b3Body::SynchronizeFixtures (body.cpp:216-238)
            Transform xf2(b.xfT, b.q);
            Transform xf1 = xf2;
            if (b.flags & kAwake) xf1 = Transform(b.x0 - Mul(b.q0, b.localCenter), b.q0);
            for (u32 k = 0; k < b.fixtureCount; ++k) {
                u32 f = B.fixtureList[wi.fixtureListBase + b.fixtureBegin + k];
                const Fixture& fx = B.fixtures[wi.fixtureBase + f];
                AABB a1, a2;
                ComputeAABB(B, fx, xf1, a1);
                ComputeAABB(B, fx, xf2, a2);
                AABB aabb = Combine(a1, a2);
                if (MoveFatAABB(B.fatAABB[wi.fixtureBase + f], aabb, a2.GetCenter() - a1.GetCenter()))
                    B.moved[wi.fixtureBase + f] = 1u;
            }
        }
        if ((wi.flags & 0x4u) == 0) {  // b3World::ClearForces
            b.force.SetZero();
            b.torque.SetZero();
        }
    }
    if (solved && wi.meshTriCap) {
        // every mesh contact follows its shapes (world.cpp: after the fixtures are synchronized)
        const u32 n = B.wstate[w].contactCount;
        const u32* order = B.contactOrder + wi.contactBase;
        for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
            const u32 slot = order[i];
            const Contact& c = B.contacts[wi.contactBase + slot];
            if (B.fixtures[wi.fixtureBase + c.fixtureA].shapeType == kMeshShape ||
                B.fixtures[wi.fixtureBase + c.fixtureB].shapeType == kMeshShape)
                MeshContactSynchronize(B, wi, slot, c);
        }
    }
}

// ================================================================== pairs + contact creation
// contact.cpp:58-72: registered shape pairs; +1 primary order, -1 swapped, 0 none.
__device__ __forceinline__ int RegisterLookup(u32 t1, u32 t2) {
    // bit (4*t1 + t2) set when (t1, t2) is a primary registered pair
    const u32 primary = (1u << (0 * 5 + 0)) | (1u << (1 * 5 + 0)) | (1u << (1 * 5 + 1)) | (1u << (2 * 5 + 0)) |
                        (1u << (2 * 5 + 1)) | (1u << (3 * 5 + 0)) | (1u << (3 * 5 + 1)) | (1u << (3 * 5 + 2)) |
                        (1u << (3 * 5 + 3)) | (1u << (3 * 5 + 4)) | (1u << (4 * 5 + 0)) | (1u << (4 * 5 + 1));
    if (primary & (1u << (t1 * 5 + t2))) return 1;
    if (primary & (1u << (t2 * 5 + t1))) return -1;
    return 0;
}

__global__ void __launch_bounds__(kBlock) k_find_pairs(BatchView B, u32 rebuildAdjacency) {
    const u32 w = WorldOfBlock(B);
    const WorldInfo wi = B.worlds[w];
    WorldState& ws = B.wstate[w];
    if (rebuildAdjacency) {
        BuildAdjacency(B, wi, ws, B.edgeOffset + wi.edgeBase, B.pairCounts + wi.edgeBase);
    }
    const u32 nf = wi.fixtureCount;
    const u32* porder = B.proxyOrder + wi.fixtureBase;
    const u32* moved = B.moved + wi.fixtureBase;
    const AABB* fat = B.fatAABB + wi.fixtureBase;
    uint2* pairs = B.pairs + wi.pairBase;
    // b3BroadPhase::FindPairs: every overlap of a moved proxy with any other proxy, as sorted
    // unique (min, max) proxy-id pairs. Enumerated directly in sorted order: thread k owns the
    // pairs whose smaller proxy is the k-th proxy.
    u32 carry = 0;
    for (u32 base = 0; base < nf; base += blockDim.x) {
        u32 k = base + threadIdx.x;
        u32 cnt = 0;
        if (k < nf) {
            u32 p = porder[k];
            AABB ap = fat[p];
            u32 mp = moved[p];
            for (u32 l = k + 1; l < nf; ++l) {
                u32 q = porder[l];
                if ((mp | moved[q]) && TestOverlap(ap, fat[q])) ++cnt;
            }
        }
        u32 total;
        u32 pos = BlockExclusiveScan(cnt, &total);
        if (k < nf && cnt) {
            u32 p = porder[k];
            AABB ap = fat[p];
            u32 mp = moved[p];
            u32 o = carry + pos;
            for (u32 l = k + 1; l < nf; ++l) {
                u32 q = porder[l];
                if ((mp | moved[q]) && TestOverlap(ap, fat[q])) {
                    if (o < wi.pairCap) pairs[o] = make_uint2(p, q);
                    ++o;
                }
            }
        }
        carry += total;
    }
    __syncthreads();
    u32 npairs = carry;
    if (npairs > wi.pairCap) {
        if (threadIdx.x == 0) ws.status |= 0x2u;
        npairs = wi.pairCap;
    }
    for (u32 f = threadIdx.x; f < nf; f += blockDim.x) B.moved[wi.fixtureBase + f] = 0u;
    // b3ContactManager::AddPair filters (contact_manager.cpp:38-103), then creation in pair order.
    const u32* off = B.edgeOffset + wi.edgeBase;
    const u32* edges = B.edgeSlots + 2 * wi.contactBase;
    const u32* order = B.contactOrder + wi.contactBase;
    u32 created = 0;
    const u32 n0 = ws.contactCount;
    const u32 free0 = ws.freeCount;
    __syncthreads();
    for (u32 base = 0; base < npairs; base += blockDim.x) {
        u32 i = base + threadIdx.x;
        u32 keep = 0;
        u32 fa = 0, fb = 0;
        if (i < npairs) {
            uint2 pr = pairs[i];
            fa = pr.x;
            fb = pr.y;
            const Fixture& FA = B.fixtures[wi.fixtureBase + fa];
            const Fixture& FB = B.fixtures[wi.fixtureBase + fb];
            keep = 1;
            if (FA.body == FB.body) keep = 0;
            if (keep) {
                for (u32 e = off[fb]; e < off[fb + 1]; ++e) {
                    const Contact& c = B.contacts[wi.contactBase + order[edges[e]]];
                    if ((c.fixtureA == fa && c.fixtureB == fb) || (c.fixtureA == fb && c.fixtureB == fa)) { keep = 0; break; }
                }
            }
            if (keep && !BodiesShouldCollide(B, wi, FA.body, FB.body)) keep = 0;
#if BX_DEVICE_CALLBACKS
            // b3ContactFilter::ShouldCollide (contact_manager.cpp:86-92), fixtures in pair order
            if (keep && wi.filter && !FilterBit(wi.filter, nf, fa, fb)) keep = 0;
#endif
            if (keep) {
                int reg = RegisterLookup(FA.shapeType, FB.shapeType);
                if (reg == 0) keep = 0;
                else if (reg < 0) { u32 t = fa; fa = fb; fb = t; }
            }
        }
        u32 total;
        u32 rank = BlockExclusiveScan(keep, &total);
        if (keep) {
            u32 r = created + rank;
            if (r < free0) {
                u32 slot = B.freeSlots[wi.contactBase + free0 - 1 - r];
                Contact& c = B.contacts[wi.contactBase + slot];
                c.fixtureA = fa;
                c.fixtureB = fb;
                c.flags = 0;
                c.manifoldCount = 0;
                c.cache.cacheType = 0;
                c.cache.featureType = 0;
                c.cache.index1 = 0;
                c.cache.index2 = 0;
                B.contactOrder[wi.contactBase + n0 + r] = slot;
                if (B.fixtures[wi.fixtureBase + fa].shapeType == kMeshShape ||
                    B.fixtures[wi.fixtureBase + fb].shapeType == kMeshShape)
                    MeshContactCreate(B, wi, slot, c);
            }
        }
        created += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (created > free0) atomicOr(&ws.status, 0x2u);
        ws.contactCount = n0 + Min(created, free0);
        ws.freeCount = free0 - Min(created, free0);
        ws.pairCount = npairs;
    }
    if (wi.meshTriCap) {
        // b3ContactManager::FindNewContacts: every mesh contact refreshes its triangles
        // (contact_manager.cpp:153-163, mesh_contact.cpp:228-258)
        const u32 n = n0 + Min(created, free0);
        bool ok = true;
        for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
            const u32 slot = order[i];
            const Contact& c = B.contacts[wi.contactBase + slot];
            if (B.fixtures[wi.fixtureBase + c.fixtureA].shapeType == kMeshShape ||
                B.fixtures[wi.fixtureBase + c.fixtureB].shapeType == kMeshShape)
                ok = MeshContactFindPairs(B, wi, slot, c) && ok;
        }
        if (!ok) atomicOr(&ws.status, 0x2u);
    }
}

// ================================================================== host launch
// Launches one step of worldCount worlds: worlds firstWorld.. when worldList is null, else the
// worlds listed in the device array worldList (any subset of the batch, in any order).
void LaunchStep(const BatchView& view, const u32* worldList, u32 firstWorld, u32 worldCount, const StepParams& sp,
                bool findFirst, u32 maxIslands, u32 maxWorldBodies, hipStream_t stream, hipEvent_t* ev) {
    BatchView B = view;
    B.worldList = worldList;
    if (!worldList) {
        B.worlds += firstWorld;
        B.wstate += firstWorld;
    }
    B.worldCount = worldCount;
    dim3 grid(worldCount);
    // ev (when given) holds kStepEvents events: one before the first kernel and one after each
    // kernel slot, so the host can report per-kernel times measured on this stream.
    auto mark = [&](int i) {
        if (ev) hipEventRecord(ev[i], stream);
    };
    mark(0);
    if (findFirst) hipLaunchKernelGGL(k_find_pairs, grid, dim3(kBlock), 0, stream, B, 1u);
    mark(1);
    if (B.meshWorlds) hipLaunchKernelGGL(k_update_contacts<true>, grid, dim3(kBlock), 0, stream, B);
    else hipLaunchKernelGGL(k_update_contacts<false>, grid, dim3(kBlock), 0, stream, B);
    mark(2);
    if (sp.dt > scalar(0)) {
        if (B.meshWorlds) hipLaunchKernelGGL(k_islands<true>, grid, dim3(kBlock), B.islandLdsWords * sizeof(u32), stream, B);
        else hipLaunchKernelGGL(k_islands<false>, grid, dim3(kBlock), B.islandLdsWords * sizeof(u32), stream, B);
        mark(3);
        // islands <= bodies of the launched worlds
        static_assert(kSolveLayoutLanes[0] == 16 && kSolveLayoutSlab[0] == 256 && kSolveLayoutLanes[1] == 32 &&
                          kSolveLayoutSlab[1] == 128 && kSolveLayoutLanes[2] == 64 && kSolveLayoutSlab[2] == 64 &&
                          kSolveLayoutLanes[3] == 4 && kSolveLayoutSlab[3] == 320 && kSolveLayoutLanes[4] == 8 &&
                          kSolveLayoutSlab[4] == 224,
                      "launched layouts");
        const u32 layout = std::min(B.solveLayout, kSolveLayouts - 1);
        const u32 lanes = kSolveLayoutLanes[layout];
        u32 vblocks = std::max(1u, std::min((maxIslands + lanes - 1) / lanes, 16384u));
        u32 pblocks = std::max(1u, std::min((maxIslands + kPosLanes - 1) / kPosLanes, 16384u));
        hipLaunchKernelGGL(k_integrate_vel, grid, dim3(kBlock), 0, stream, B, sp);
        mark(4);
        hipLaunchKernelGGL(k_init_contacts, grid, dim3(kBlock), 0, stream, B, sp);
        mark(5);
        if (layout == 0) hipLaunchKernelGGL((k_solve_velocity<16, 256>), dim3(vblocks), dim3(64), 0, stream, B, sp, 0u);
        else if (layout == 1) hipLaunchKernelGGL((k_solve_velocity<32, 128>), dim3(vblocks), dim3(64), 0, stream, B, sp, 0u);
        else if (layout == 2) hipLaunchKernelGGL((k_solve_velocity<64, 64>), dim3(vblocks), dim3(64), 0, stream, B, sp, 0u);
        else if (layout == 3) hipLaunchKernelGGL((k_solve_velocity_quad<4, 320>), dim3(vblocks), dim3(64), 0, stream, B, sp);
        else hipLaunchKernelGGL((k_solve_velocity_quad<8, 224>), dim3(vblocks), dim3(64), 0, stream, B, sp);
        // islands above the eight-lane layouts' slab, when a launched world is large enough to hold one
        if (SolveLayoutQuad(layout) && maxWorldBodies > kSolveLayoutSlab[layout]) {
            const u32 big = std::max(1u, std::min((maxIslands + 15) / 16, 16384u));
            hipLaunchKernelGGL((k_solve_velocity<16, 256>), dim3(big), dim3(64), 0, stream, B, sp, kSolveLayoutSlab[layout]);
        }
        mark(6);
        hipLaunchKernelGGL(k_mid, grid, dim3(kBlock), 0, stream, B, sp);
        mark(7);
        if (layout == 3) {
            hipLaunchKernelGGL((k_solve_position_quad<4, kPosQuadSlab>), dim3(vblocks), dim3(64), 0, stream, B, sp);
        } else if (layout == 4) {
            hipLaunchKernelGGL((k_solve_position_quad<8, kPosQuadSlab>), dim3(vblocks), dim3(64), 0, stream, B, sp);
        }
        if (!SolveLayoutQuad(layout))
            hipLaunchKernelGGL(k_solve_position, dim3(pblocks), dim3(64), kPosSpreadLds, stream, B, sp, 0u);
        else if (maxWorldBodies > kPosQuadSlab)  // islands above the eight-lane layouts' pose slab
            hipLaunchKernelGGL(k_solve_position, dim3(pblocks), dim3(64), kPosSpreadLds, stream, B, sp, kPosQuadSlab);
        mark(8);
        hipLaunchKernelGGL(k_finish, grid, dim3(kBlock), 0, stream, B, sp);
        mark(9);
        hipLaunchKernelGGL(k_sync, grid, dim3(kBlock), 0, stream, B, sp);
        mark(10);
        hipLaunchKernelGGL(k_find_pairs, grid, dim3(kBlock), 0, stream, B, 0u);
        mark(11);
    } else {
        for (int i = 3; i <= 9; ++i) mark(i);
        hipLaunchKernelGGL(k_sync, grid, dim3(kBlock), 0, stream, B, sp);
        mark(10);
        mark(11);
    }
}

// The query hierarchy of the worlds listed in the device array worldList (k_build_bvh).
void LaunchBuildBvh(const BatchView& view, const u32* worldList, u32 worldCount, hipStream_t stream) {
    BatchView B = view;
    B.worldList = worldList;
    B.worldCount = worldCount;
    hipLaunchKernelGGL(k_build_bvh, dim3(worldCount), dim3(kBlock), 0, stream, B);
}

// One lane per query (k_query).
void LaunchQueries(const BatchView& view, const QueryParams& qp, hipStream_t stream) {
    const u32 blocks = (qp.count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_query, dim3(blocks), dim3(kBlock), 0, stream, view, qp);
}

// One lane per time-of-impact problem (k_toi) and per shape cast (k_shape_cast).
void LaunchTimeOfImpact(const BatchView& view, const ToiQueryD* queries, u32 count, ToiResultD* out,
                        hipStream_t stream) {
    const u32 blocks = (count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_toi, dim3(blocks), dim3(kBlock), 0, stream, view, queries, count, out);
}

void LaunchShapeCasts(const BatchView& view, const CastParams& cp, hipStream_t stream) {
    const u32 blocks = (cp.count + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_shape_cast, dim3(blocks), dim3(kBlock), 0, stream, view, cp);
}

// The creation-time pair search (b3World::Step's FindNewContacts after shapes were added) of the
// worlds listed in the device array worldList.
void LaunchInitialPairs(const BatchView& view, const u32* worldList, u32 worldCount, hipStream_t stream) {
    BatchView B = view;
    B.worldList = worldList;
    B.worldCount = worldCount;
    hipLaunchKernelGGL(k_find_pairs, dim3(worldCount), dim3(kBlock), 0, stream, B, 1u);
}

}  // namespace bx
