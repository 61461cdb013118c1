// orc_world.cpp -- oracle (test infrastructure): b3World::Step restated on CPU.
//   world.cpp:187-450            Step, Solve (island DFS), ClearForces
//   island.cpp:59-307            integration, solver calls, sleep
//   contact_solver.cpp:28-624    PGS velocity / NGS position contact solver
//   contact_manager.cpp:38-289   AddPair, UpdateContacts, Destroy
//   contact.cpp:183-285          b3Contact::Update (warm-start matching)
//   body.cpp:216-376             SynchronizeFixtures, ShouldCollide, ResetMass
//   broad_phase.h:159-207, broad_phase.cpp, dynamic_tree.cpp:44-207  pairs, proxies, MoveProxy
// The pair *set* of FindPairs does not depend on the dynamic tree's shape (every fat AABB that
// overlaps a moved proxy's fat AABB is reported, internal nodes are exact unions), so this
// oracle enumerates overlaps directly and only reproduces the tree's node-id allocation, which
// fixes the proxy ids and therefore the sort order of the pairs.
#include <algorithm>
#include <cstring>
#include "orc_world.h"

namespace orc {

void SolveJointsInit(World& w, const TimeStep& step, std::vector<u32>& joints, Position* pos, Velocity* vel);
void SolveJointsWarmStart(World& w, std::vector<u32>& joints, Velocity* vel);
void SolveJointsVelocity(World& w, const TimeStep& step, std::vector<u32>& joints, Velocity* vel);
bool SolveJointsPosition(World& w, std::vector<u32>& joints, Position* pos);
void MeshContactFindPairs(World& w, Contact* c);
void MeshContactSynchronize(World& w, Contact* c);
void MeshContactCollide(World& w, Contact* c);
void MeshContactCreate(World& w, Contact* c);
bool MeshContactTestOverlap(World& w, Contact* c);
void MeshAABB(const Shape& s, const Transform& xf, AABB* out);

World::~World() {
    for (Contact* c : contactList) delete c;
}

// ------------------------------------------------------------------ broadphase
u32 World::AllocateNode() {  // dynamic_tree.cpp:44-79: LIFO free list over a sequential pool
    if (!freeStack.empty()) { u32 n = freeStack.back(); freeStack.pop_back(); return n; }
    return nextNode++;
}

u32 World::CreateProxy(const AABB& aabb, u32 fixture) {  // dynamic_tree.cpp:108-128, broad_phase.cpp:78-84
    u32 proxy = AllocateNode();
    if (proxy == kNullU32) return kNullU32;  // an untracked node: the id depends on the tree's shape
    if (leafCount > 0) AllocateNode();  // InsertLeaf allocates the new parent
    ++leafCount;
    if (fatAABB.size() <= proxy) { fatAABB.resize(proxy + 1); proxyFixture.resize(proxy + 1, kNullU32); }
    Vec3 r(kAabbExtension, kAabbExtension, kAabbExtension);
    fatAABB[proxy].lowerBound = aabb.lowerBound - r;
    fatAABB[proxy].upperBound = aabb.upperBound + r;
    proxyFixture[proxy] = fixture;
    moveBuffer.push_back(proxy);
    return proxy;
}

// b3BroadPhase::DestroyProxy (broad_phase.cpp:87-92): unbuffer the move, then RemoveLeaf frees the
// leaf's parent when the leaf is not the root (its id depends on the tree's shape and is not
// tracked) and FreeNode frees the proxy's node (dynamic_tree.cpp:130-137, 333-373).
void World::DestroyProxy(u32 proxy) {
    for (u32& m : moveBuffer)
        if (m == proxy) m = kNullU32;
    if (leafCount > 1) freeStack.push_back(kNullU32);
    freeStack.push_back(proxy);
    --leafCount;
    proxyFixture[proxy] = kNullU32;
}

void World::MoveProxy(u32 proxy, const AABB& aabb, const Vec3& displacement) {  // dynamic_tree.cpp:139-207
    AABB fat = aabb;
    fat.Extend(kAabbExtension);
    Vec3 d = kAabbMultiplier * displacement;
    if (d.x < scalar(0)) fat.lowerBound.x += d.x; else fat.upperBound.x += d.x;
    if (d.y < scalar(0)) fat.lowerBound.y += d.y; else fat.upperBound.y += d.y;
    if (d.z < scalar(0)) fat.lowerBound.z += d.z; else fat.upperBound.z += d.z;
    const AABB& tree = fatAABB[proxy];
    if (tree.Contains(aabb)) {
        AABB huge = fat;
        huge.Extend(scalar(4) * kAabbExtension);
        if (huge.Contains(tree)) return;
    }
    // RemoveLeaf frees the parent and InsertLeaf takes it straight back: ids are unchanged.
    fatAABB[proxy] = fat;
    moveBuffer.push_back(proxy);
}

void World::SynchronizeFixture(u32 f, const Transform& xf1, const Transform& xf2) {  // fixture.cpp:120-132
    Fixture& fx = fixtures[f];
    AABB a1, a2;
    ComputeAABB(fx.shape, xf1, &a1);
    ComputeAABB(fx.shape, xf2, &a2);
    AABB a = Combine(a1, a2);
    Vec3 disp = a2.GetCenter() - a1.GetCenter();
    MoveProxy(fx.proxyId, a, disp);
}

void World::SynchronizeFixtures(u32 bi) {  // body.cpp:216-238
    Body& b = bodies[bi];
    if (b.flags & e_awakeFlag) {
        Transform xf1;
        xf1.rotation = b.sweep.orientation0;
        xf1.translation = b.sweep.worldCenter0 - Mul(xf1.rotation, b.sweep.localCenter);
        for (u32 f : b.fixtures) SynchronizeFixture(f, xf1, b.xf);
    } else {
        for (u32 f : b.fixtures) SynchronizeFixture(f, b.xf, b.xf);
    }
}

bool World::ShouldCollide(u32 ai, u32 bi) const {  // body.cpp:240-263
    const Body& a = bodies[ai];
    const Body& b = bodies[bi];
    if (a.type != e_dynamicBody && b.type != e_dynamicBody) return false;
    for (const JointEdge& je : a.joints)
        if (je.other == bi && !joints[je.joint].collideConnected) return false;
    return true;
}

void World::ResetMass(u32 bi) {  // body.cpp:265-376
    Body& b = bodies[bi];
    b.mass = 0; b.invMass = 0;
    b.I.SetZero(); b.invI.SetZero(); b.worldInvI.SetZero();
    b.sweep.localCenter.SetZero();
    if (b.type == e_staticBody || b.type == e_kinematicBody) {
        b.sweep.worldCenter0 = b.xf.translation;
        b.sweep.worldCenter = b.xf.translation;
        b.sweep.orientation0 = b.sweep.orientation;
        return;
    }
    Vec3 localCenter; localCenter.SetZero();
    for (u32 f : b.fixtures) {
        const Fixture& fx = fixtures[f];
        if (fx.density == scalar(0)) continue;
        MassData md;
        ComputeMass(fx.shape, fx.density, &md);
        localCenter += md.mass * md.center;
        b.mass += md.mass;
        b.I += md.I;
    }
    if (b.mass > scalar(0)) {
        b.invMass = scalar(1) / b.mass;
        localCenter *= b.invMass;
        b.I = b.I - b.mass * Steiner(localCenter);
        b.invI = Inverse(b.I);
        b.worldInvI = RotateToFrame(b.invI, b.xf.rotation);
        if (b.flags & e_fixedRotationX) {
            b.invI.y.y = 0; b.invI.z.y = 0; b.invI.y.z = 0; b.invI.z.z = 0;
            b.worldInvI.y.y = 0; b.worldInvI.z.y = 0; b.worldInvI.y.z = 0; b.worldInvI.z.z = 0;
        }
        if (b.flags & e_fixedRotationY) {
            b.invI.x.x = 0; b.invI.x.z = 0; b.invI.z.x = 0; b.invI.z.z = 0;
            b.worldInvI.x.x = 0; b.worldInvI.x.z = 0; b.worldInvI.z.x = 0; b.worldInvI.z.z = 0;
        }
        if (b.flags & e_fixedRotationZ) {
            b.invI.x.x = 0; b.invI.x.y = 0; b.invI.y.x = 0; b.invI.y.y = 0;
            b.worldInvI.x.x = 0; b.worldInvI.x.y = 0; b.worldInvI.y.x = 0; b.worldInvI.y.y = 0;
        }
    } else {
        b.mass = 1;
        b.invMass = 1;
    }
    Vec3 oldCenter = b.sweep.worldCenter;
    b.sweep.localCenter = localCenter;
    b.sweep.worldCenter = Mul(b.xf, b.sweep.localCenter);
    b.sweep.worldCenter0 = b.sweep.worldCenter;
    b.linearVelocity += Cross(b.angularVelocity, b.sweep.worldCenter - oldCenter);
}

// contact.cpp:58-72 type table: the primary (first) type of each registered pair.
static int RegisterLookup(u32 t1, u32 t2) {
    static const u32 pairs[12][2] = {
        {e_sphere, e_sphere}, {e_capsule, e_sphere}, {e_capsule, e_capsule}, {e_triangle, e_sphere},
        {e_triangle, e_capsule}, {e_hull, e_sphere}, {e_hull, e_capsule}, {e_hull, e_triangle},
        {e_hull, e_hull}, {e_hull, e_mesh}, {e_mesh, e_sphere}, {e_mesh, e_capsule}};
    for (int i = 0; i < 12; ++i) {
        if (pairs[i][0] == t1 && pairs[i][1] == t2) return 1;   // primary order
        if (pairs[i][0] == t2 && pairs[i][1] == t1) return -1;  // swapped
    }
    return 0;
}

void World::AddPair(u32 fA, u32 fB) {  // contact_manager.cpp:38-141
    u32 bA = fixtures[fA].body, bB = fixtures[fB].body;
    if (bA == bB) return;
    for (Contact* c : fixtures[fB].contacts) {
        u32 other = c->fixtureA == fB ? c->fixtureB : c->fixtureA;
        if (other == fA) {
            if ((c->fixtureA == fA && c->fixtureB == fB) || (c->fixtureA == fB && c->fixtureB == fA)) return;
        }
    }
    if (!ShouldCollide(bA, bB)) return;
    if (!FilterAllows(filterCollide, fA, fB)) return;  // b3ContactFilter::ShouldCollide (contact_manager.cpp:86-92)
    int reg = RegisterLookup(fixtures[fA].shape.type, fixtures[fB].shape.type);
    if (reg == 0) return;
    if (reg < 0) std::swap(fA, fB);
    Contact* c = new Contact();
    c->fixtureA = fA;
    c->fixtureB = fB;
    c->isMesh = fixtures[fA].shape.type == e_mesh || fixtures[fB].shape.type == e_mesh;
    c->manifoldCapacity = c->isMesh ? kMaxManifolds : 1;
    contactList.insert(contactList.begin(), c);
    fixtures[fA].contacts.insert(fixtures[fA].contacts.begin(), c);
    fixtures[fB].contacts.insert(fixtures[fB].contacts.begin(), c);
    if (c->isMesh) MeshContactCreate(*this, c);
}

void World::FindNewContacts() {  // contact_manager.cpp:153-163, broad_phase.h:159-207
    std::vector<std::pair<u32, u32>> pairs;
    u32 nProxies = u32(fatAABB.size());
    for (u32 q : moveBuffer) {
        if (q == kNullU32) continue;
        for (u32 p = 0; p < nProxies; ++p) {
            if (p == q || proxyFixture[p] == kNullU32) continue;
            if (TestOverlap(fatAABB[p], fatAABB[q])) pairs.push_back(std::make_pair(std::min(p, q), std::max(p, q)));
        }
    }
    moveBuffer.clear();
    std::sort(pairs.begin(), pairs.end());
    pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
    lastPairs = pairs;
    for (auto& pr : pairs) AddPair(proxyFixture[pr.first], proxyFixture[pr.second]);
    for (Contact* c : contactList)
        if (c->isMesh) MeshContactFindPairs(*this, c);
}

void World::DestroyContact(Contact* c) {  // contact_manager.cpp:229-289, contact.cpp:108-131
    if (recordEvents && (c->flags & 2u)) events.push_back(Event{2u, c->fixtureA, c->fixtureB, nullptr});
    contactList.erase(std::find(contactList.begin(), contactList.end(), c));
    auto& la = fixtures[c->fixtureA].contacts;
    la.erase(std::find(la.begin(), la.end(), c));
    auto& lb = fixtures[c->fixtureB].contacts;
    lb.erase(std::find(lb.begin(), lb.end(), c));
    for (u32 i = 0; i < c->manifoldCount; ++i) {
        if (c->manifolds[i].pointCount > 0) {
            if (!fixtures[c->fixtureA].isSensor && !fixtures[c->fixtureB].isSensor) {
                bodies[fixtures[c->fixtureA].body].SetAwake(true);
                bodies[fixtures[c->fixtureB].body].SetAwake(true);
                break;
            }
        }
    }
    delete c;
}

void World::Collide(Contact* c) {
    const Fixture& fA = fixtures[c->fixtureA];
    const Fixture& fB = fixtures[c->fixtureB];
    if (c->isMesh) { MeshContactCollide(*this, c); return; }
    const Body& bA = bodies[fA.body];
    const Body& bB = bodies[fB.body];
    Transform xfA = bA.xf, xfB = bB.xf;
    Manifold& m = c->manifolds[0];
    u32 tA = fA.shape.type, tB = fB.shape.type;
    if (tA == e_hull && tB == e_hull) {  // hull_contact.cpp:41-47
        Transform xf0A = bA.sweep.GetTransform(scalar(0));
        Transform xf0B = bB.sweep.GetTransform(scalar(0));
        CollideHulls(m, xfA, fA.shape, xfB, fB.shape, c->cache, xf0A, xf0B);
    } else if (tA == e_sphere && tB == e_sphere) {
        CollideSpheres(m, xfA, fA.shape, xfB, fB.shape);
    } else if (tA == e_capsule && tB == e_sphere) {
        CollideCapsuleAndSphere(m, xfA, fA.shape, xfB, fB.shape);
    } else if (tA == e_capsule && tB == e_capsule) {
        CollideCapsules(m, xfA, fA.shape, xfB, fB.shape);
    } else if (tA == e_hull && tB == e_sphere) {
        CollideHullAndSphere(m, xfA, fA.shape, xfB, fB.shape);
    } else if (tA == e_hull && tB == e_capsule) {
        CollideHullAndCapsule(m, xfA, fA.shape, xfB, fB.shape);
    } else if (tA == e_triangle && tB == e_sphere) {  // triangle_sphere_contact.cpp
        CollideTriangleAndSphere(m, xfA, fA.shape, xfB, fB.shape);
    } else if (tA == e_triangle && tB == e_capsule) {  // triangle_capsule_contact.cpp
        CollideTriangleAndCapsule(m, xfA, fA.shape, xfB, fB.shape);
    } else if (tA == e_hull && tB == e_triangle) {  // hull_triangle_contact.cpp
        CollideHullAndTriangle(m, xfA, fA.shape, xfB, fB.shape);
    }
    c->manifoldCount = 1;  // convex_contact.cpp:44-52
}

void World::UpdateContact(Contact* c) {  // contact.cpp:183-285
    bool touching = false;
    bool wasTouching = (c->flags & 2u) != 0;
    const Fixture& fA = fixtures[c->fixtureA];
    const Fixture& fB = fixtures[c->fixtureB];
    bool sensor = fA.isSensor || fB.isSensor;
    Body& bA = bodies[fA.body];
    Body& bB = bodies[fB.body];
    if (sensor) {
        touching = c->isMesh ? MeshContactTestOverlap(*this, c) : ConvexTestOverlap(bA.xf, fA.shape, bB.xf, fB.shape);
        c->manifoldCount = 0;
    } else {
        u32 oldCount = c->manifoldCount;
        Manifold old[kMaxManifolds];
        for (u32 i = 0; i < oldCount; ++i) old[i] = c->manifolds[i];
        c->manifoldCount = 0;
        for (u32 i = 0; i < c->manifoldCapacity; ++i) c->manifolds[i].Initialize();
        Collide(c);
        if (warmStarting) {
            for (u32 i = 0; i < c->manifoldCount; ++i)
                for (u32 j = 0; j < oldCount; ++j) c->manifolds[i].Initialize(old[j]);
        }
        for (u32 i = 0; i < c->manifoldCount; ++i)
            if (c->manifolds[i].pointCount > 0) { touching = true; break; }
        if (touching != wasTouching) { bA.SetAwake(true); bB.SetAwake(true); }
    }
    if (touching) c->flags |= 2u; else c->flags &= ~2u;
    if (recordEvents) {  // contact.cpp:268-284
        if (!wasTouching && touching) events.push_back(Event{1u, c->fixtureA, c->fixtureB, c});
        if (wasTouching && !touching) events.push_back(Event{2u, c->fixtureA, c->fixtureB, c});
        if (!sensor && touching) events.push_back(Event{3u, c->fixtureA, c->fixtureB, c});
    }
}

void World::UpdateContacts() {  // contact_manager.cpp:165-227
    std::vector<Contact*> list = contactList;  // iteration snapshot; destroyed ones are skipped below
    for (Contact* c : list) {
        u32 bA = fixtures[c->fixtureA].body, bB = fixtures[c->fixtureB].body;
        if (!ShouldCollide(bA, bB)) { DestroyContact(c); continue; }
        if (!FilterAllows(filterCollide, c->fixtureA, c->fixtureB)) { DestroyContact(c); continue; }
        bool activeA = bodies[bA].IsAwake() && bodies[bA].type != e_staticBody;
        bool activeB = bodies[bB].IsAwake() && bodies[bB].type != e_staticBody;
        if (!activeA && !activeB) continue;
        bool overlap = TestOverlap(fatAABB[fixtures[c->fixtureA].proxyId], fatAABB[fixtures[c->fixtureB].proxyId]);
        if (!overlap) { DestroyContact(c); continue; }
        UpdateContact(c);
    }
}

// ------------------------------------------------------------------ contact solver
namespace {

struct VCPoint { Vec3 rA, rB, normal; scalar normalMass, normalImpulse, velocityBias; };
struct VCManifold {
    Vec3 rA, rB, normal, tangent1, tangent2;
    scalar tangentSpeed1, tangentSpeed2;
    Mat22 tangentMass;
    Vec2 tangentImpulse;
    scalar motorImpulse, motorMass, motorSpeed;
    u32 pointCount;
    VCPoint points[kMaxManifoldPoints];
};
struct VC {
    u32 indexA, indexB;
    scalar invMassA, invMassB;
    Mat33 invIA, invIB;
    scalar friction, restitution;
    u32 manifoldCount;
    VCManifold manifolds[kMaxManifolds];
};
struct PCPoint { Vec3 localNormalA, localPointA, localPointB; };
struct PC {
    u32 indexA, indexB;
    scalar invMassA, invMassB, radiusA, radiusB;
    Mat33 localInvIA, localInvIB;
    Vec3 localCenterA, localCenterB;
    u32 manifoldCount;
    u32 pointCount[kMaxManifolds];
    PCPoint points[kMaxManifolds][kMaxManifoldPoints];
};

}  // namespace

static void InitContactConstraints(World& w, const TimeStep& step, std::vector<Contact*>& contacts, Position* pos,
                                   std::vector<VC>& vcs, std::vector<PC>& pcs) {
    // contact_solver.cpp:28-150 (constructor)
    u32 n = u32(contacts.size());
    vcs.resize(n);
    pcs.resize(n);
    for (u32 i = 0; i < n; ++i) {
        Contact* c = contacts[i];
        const Fixture& fA = w.fixtures[c->fixtureA];
        const Fixture& fB = w.fixtures[c->fixtureB];
        const Body& bA = w.bodies[fA.body];
        const Body& bB = w.bodies[fB.body];
        PC& pc = pcs[i];
        VC& vc = vcs[i];
        pc.indexA = bA.islandIndex; pc.invMassA = bA.invMass; pc.localInvIA = bA.invI;
        pc.localCenterA = bA.sweep.localCenter; pc.radiusA = fA.shape.radius;
        pc.indexB = bB.islandIndex; pc.invMassB = bB.invMass; pc.localInvIB = bB.invI;
        pc.localCenterB = bB.sweep.localCenter; pc.radiusB = fB.shape.radius;
        pc.manifoldCount = c->manifoldCount;
        vc.indexA = bA.islandIndex; vc.invMassA = bA.invMass; vc.invIA = pos[vc.indexA].I;
        vc.indexB = bB.islandIndex; vc.invMassB = bB.invMass; vc.invIB = pos[vc.indexB].I;
        vc.friction = Sqrt(fA.friction * fB.friction);         // b3MixFriction
        vc.restitution = Max(fA.restitution, fB.restitution);  // b3MixRestitution
        vc.manifoldCount = c->manifoldCount;
        for (u32 j = 0; j < c->manifoldCount; ++j) {
            const Manifold& m = c->manifolds[j];
            VCManifold& vcm = vc.manifolds[j];
            pc.pointCount[j] = m.pointCount;
            vcm.pointCount = m.pointCount;
            vcm.rA.SetZero(); vcm.rB.SetZero(); vcm.normal.SetZero();
            vcm.tangent1.SetZero(); vcm.tangent2.SetZero();
            vcm.motorSpeed = m.motorSpeed;
            vcm.tangentSpeed1 = m.tangentSpeed1;
            vcm.tangentSpeed2 = m.tangentSpeed2;
            if (step.warmStarting) {
                vcm.tangentImpulse = step.dtRatio * m.tangentImpulse;
                vcm.motorImpulse = step.dtRatio * m.motorImpulse;
            } else {
                vcm.tangentImpulse.SetZero();
                vcm.motorImpulse = 0;
            }
            vcm.tangentMass.SetZero();
            vcm.motorMass = 0;
            for (u32 k = 0; k < m.pointCount; ++k) {
                const ManifoldPoint& cp = m.points[k];
                pc.points[j][k].localNormalA = cp.localNormal1;
                pc.points[j][k].localPointA = cp.localPoint1;
                pc.points[j][k].localPointB = cp.localPoint2;
                VCPoint& vcp = vcm.points[k];
                vcp.rA.SetZero(); vcp.rB.SetZero(); vcp.normal.SetZero();
                vcp.normalMass = 0; vcp.velocityBias = 0;
                vcp.normalImpulse = step.warmStarting ? step.dtRatio * cp.normalImpulse : scalar(0);
            }
        }
    }
}

static void InitializeVelocityConstraints(std::vector<Contact*>& contacts, Position* pos, Velocity* vel,
                                          std::vector<VC>& vcs, std::vector<PC>& pcs) {
    // contact_solver.cpp:152-270
    for (u32 i = 0; i < contacts.size(); ++i) {
        Contact* c = contacts[i];
        PC& pc = pcs[i];
        VC& vc = vcs[i];
        scalar radiusA = pc.radiusA, radiusB = pc.radiusB;
        u32 iAi = vc.indexA, iBi = vc.indexB;
        scalar mA = vc.invMassA, mB = vc.invMassB;
        Mat33 iA = vc.invIA, iB = vc.invIB;
        Vec3 xA = pos[iAi].x; Quat qA = pos[iAi].q;
        Vec3 xB = pos[iBi].x; Quat qB = pos[iBi].q;
        Vec3 vA = vel[iAi].v, wA = vel[iAi].w, vB = vel[iBi].v, wB = vel[iBi].w;
        Transform xfA, xfB;
        xfA.rotation = qA; xfA.translation = xA - Mul(qA, pc.localCenterA);
        xfB.rotation = qB; xfB.translation = xB - Mul(qB, pc.localCenterB);
        for (u32 j = 0; j < c->manifoldCount; ++j) {
            const Manifold* m = c->manifolds + j;
            VCManifold& vcm = vc.manifolds[j];
            WorldManifold wm;
            wm.Initialize(m, radiusA, xfA, radiusB, xfB);
            vcm.normal = wm.normal;
            vcm.tangent1 = wm.tangent1;
            vcm.tangent2 = wm.tangent2;
            for (u32 k = 0; k < wm.pointCount; ++k) {
                WorldManifoldPoint* mp = wm.points + k;
                VCPoint& vcp = vcm.points[k];
                Vec3 normal = mp->normal;
                Vec3 point = mp->point;
                Vec3 rA = point - xA;
                Vec3 rB = point - xB;
                vcp.rA = rA;
                vcp.rB = rB;
                vcp.normal = normal;
                Vec3 rnA = Cross(rA, normal);
                Vec3 rnB = Cross(rB, normal);
                scalar K = mA + mB + Dot(iA * rnA, rnA) + Dot(iB * rnB, rnB);
                vcp.normalMass = K > scalar(0) ? scalar(1) / K : scalar(0);
                Vec3 dv = vB + Cross(wB, rB) - vA - Cross(wA, rA);
                scalar vn = Dot(normal, dv);
                vcp.velocityBias = 0;
                if (vn < -kVelocityThreshold) vcp.velocityBias = -vc.restitution * vn;
            }
            if (wm.pointCount > 0) {
                Vec3 rA = wm.center - xA;
                Vec3 rB = wm.center - xB;
                vcm.rA = rA;
                vcm.rB = rB;
                Vec3 t1 = vcm.tangent1, t2 = vcm.tangent2;
                Vec3 rn1A = Cross(rA, t1), rn1B = Cross(rB, t1);
                Vec3 rn2A = Cross(rA, t2), rn2B = Cross(rB, t2);
                scalar k11 = mA + mB + Dot(iA * rn1A, rn1A) + Dot(iB * rn1B, rn1B);
                scalar k12 = Dot(iA * rn1A, rn2A) + Dot(iB * rn1B, rn2B);
                scalar k22 = mA + mB + Dot(iA * rn2A, rn2A) + Dot(iB * rn2B, rn2B);
                Mat22 K;
                K.x.Set(k11, k12);
                K.y.Set(k12, k22);
                vcm.tangentMass = Inverse(K);
                scalar mass = Dot(vcm.normal, (iA + iB) * vcm.normal);
                vcm.motorMass = mass > scalar(0) ? scalar(1) / mass : scalar(0);
            }
        }
    }
}

static void WarmStartContacts(std::vector<VC>& vcs, Velocity* vel) {  // contact_solver.cpp:272-326
    for (VC& vc : vcs) {
        scalar mA = vc.invMassA, mB = vc.invMassB;
        Mat33 iA = vc.invIA, iB = vc.invIB;
        Vec3 vA = vel[vc.indexA].v, wA = vel[vc.indexA].w, vB = vel[vc.indexB].v, wB = vel[vc.indexB].w;
        for (u32 j = 0; j < vc.manifoldCount; ++j) {
            VCManifold& vcm = vc.manifolds[j];
            for (u32 k = 0; k < vcm.pointCount; ++k) {
                VCPoint& vcp = vcm.points[k];
                Vec3 P = vcp.normalImpulse * vcp.normal;
                vA -= mA * P;
                wA -= iA * Cross(vcp.rA, P);
                vB += mB * P;
                wB += iB * Cross(vcp.rB, P);
            }
            if (vcm.pointCount > 0) {
                Vec3 P1 = vcm.tangentImpulse.x * vcm.tangent1;
                Vec3 P2 = vcm.tangentImpulse.y * vcm.tangent2;
                Vec3 P3 = vcm.motorImpulse * vcm.normal;
                vA -= mA * (P1 + P2);
                wA -= iA * (Cross(vcm.rA, P1 + P2) + P3);
                vB += mB * (P1 + P2);
                wB += iB * (Cross(vcm.rB, P1 + P2) + P3);
            }
        }
        vel[vc.indexA].v = vA; vel[vc.indexA].w = wA;
        vel[vc.indexB].v = vB; vel[vc.indexB].w = wB;
    }
}

static void SolveContactVelocities(std::vector<VC>& vcs, Velocity* vel) {  // contact_solver.cpp:328-486
    for (VC& vc : vcs) {
        scalar mA = vc.invMassA, mB = vc.invMassB;
        Mat33 iA = vc.invIA, iB = vc.invIB;
        Vec3 vA = vel[vc.indexA].v, wA = vel[vc.indexA].w, vB = vel[vc.indexB].v, wB = vel[vc.indexB].w;
        for (u32 j = 0; j < vc.manifoldCount; ++j) {
            VCManifold& vcm = vc.manifolds[j];
            scalar normalImpulse = 0;
            for (u32 k = 0; k < vcm.pointCount; ++k) {
                VCPoint& vcp = vcm.points[k];
                Vec3 dv = vB + Cross(wB, vcp.rB) - vA - Cross(wA, vcp.rA);
                scalar Cdot = Dot(vcp.normal, dv);
                scalar impulse = -vcp.normalMass * (Cdot - vcp.velocityBias);
                scalar oldImpulse = vcp.normalImpulse;
                vcp.normalImpulse = Max(vcp.normalImpulse + impulse, scalar(0));
                impulse = vcp.normalImpulse - oldImpulse;
                Vec3 P = impulse * vcp.normal;
                vA -= mA * P;
                wA -= iA * Cross(vcp.rA, P);
                vB += mB * P;
                wB += iB * Cross(vcp.rB, P);
                normalImpulse += vcp.normalImpulse;
            }
            if (vcm.pointCount > 0) {
                {
                    Vec3 dv = vB + Cross(wB, vcm.rB) - vA - Cross(wA, vcm.rA);
                    Vec2 Cdot;
                    Cdot.x = Dot(dv, vcm.tangent1) - vcm.tangentSpeed1;
                    Cdot.y = Dot(dv, vcm.tangent2) - vcm.tangentSpeed2;
                    Vec2 impulse = vcm.tangentMass * -Cdot;
                    Vec2 oldImpulse = vcm.tangentImpulse;
                    vcm.tangentImpulse += impulse;
                    scalar maxImpulse = vc.friction * normalImpulse;
                    if (Dot(vcm.tangentImpulse, vcm.tangentImpulse) > maxImpulse * maxImpulse) {
                        vcm.tangentImpulse.Normalize();
                        vcm.tangentImpulse *= maxImpulse;
                    }
                    impulse = vcm.tangentImpulse - oldImpulse;
                    Vec3 P1 = impulse.x * vcm.tangent1;
                    Vec3 P2 = impulse.y * vcm.tangent2;
                    Vec3 P = P1 + P2;
                    vA -= mA * P;
                    wA -= iA * Cross(vcm.rA, P);
                    vB += mB * P;
                    wB += iB * Cross(vcm.rB, P);
                }
                {
                    scalar Cdot = Dot(vcm.normal, wB - wA) - vcm.motorSpeed;
                    scalar impulse = -vcm.motorMass * Cdot;
                    scalar oldImpulse = vcm.motorImpulse;
                    scalar maxImpulse = vc.friction * normalImpulse;
                    vcm.motorImpulse = Clamp(vcm.motorImpulse + impulse, -maxImpulse, maxImpulse);
                    impulse = vcm.motorImpulse - oldImpulse;
                    Vec3 P = impulse * vcm.normal;
                    wA -= iA * P;
                    wB += iB * P;
                }
            }
        }
        vel[vc.indexA].v = vA; vel[vc.indexA].w = wA;
        vel[vc.indexB].v = vB; vel[vc.indexB].w = wB;
    }
}

static void StoreImpulses(std::vector<Contact*>& contacts, std::vector<VC>& vcs) {  // contact_solver.cpp:488-515
    for (u32 i = 0; i < contacts.size(); ++i) {
        Contact* c = contacts[i];
        for (u32 j = 0; j < c->manifoldCount; ++j) {
            Manifold& m = c->manifolds[j];
            VCManifold& vcm = vcs[i].manifolds[j];
            m.tangentImpulse = vcm.tangentImpulse;
            m.motorImpulse = vcm.motorImpulse;
            for (u32 k = 0; k < m.pointCount; ++k) m.points[k].normalImpulse = vcm.points[k].normalImpulse;
        }
    }
}

static bool SolveContactPositions(std::vector<PC>& pcs, Position* pos) {  // contact_solver.cpp:517-624
    scalar minSeparation = 0;
    for (PC& pc : pcs) {
        scalar mA = pc.invMassA, mB = pc.invMassB;
        Vec3 cA = pos[pc.indexA].x; Quat qA = pos[pc.indexA].q; Mat33 iA = pos[pc.indexA].I;
        Vec3 cB = pos[pc.indexB].x; Quat qB = pos[pc.indexB].q; Mat33 iB = pos[pc.indexB].I;
        for (u32 j = 0; j < pc.manifoldCount; ++j) {
            for (u32 k = 0; k < pc.pointCount[j]; ++k) {
                const PCPoint& pcp = pc.points[j][k];
                Transform xfA, xfB;
                xfA.rotation = qA; xfA.translation = cA - Mul(qA, pc.localCenterA);
                xfB.rotation = qB; xfB.translation = cB - Mul(qB, pc.localCenterB);
                Vec3 normal = Mul(xfA.rotation, pcp.localNormalA);
                Vec3 c1 = Mul(xfA, pcp.localPointA);
                Vec3 c2 = Mul(xfB, pcp.localPointB);
                Vec3 point = c2;
                scalar separation = Dot(c2 - c1, normal) - pc.radiusA - pc.radiusB;
                minSeparation = Min(minSeparation, separation);
                scalar C = Clamp(kBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, scalar(0));
                Vec3 rA = point - cA;
                Vec3 rB = point - cB;
                Vec3 rnA = Cross(rA, normal);
                Vec3 rnB = Cross(rB, normal);
                scalar K = mA + mB + Dot(rnA, iA * rnA) + Dot(rnB, iB * rnB);
                scalar impulse = K > scalar(0) ? -C / K : scalar(0);
                Vec3 P = impulse * normal;
                cA -= mA * P;
                qA -= Derivative(qA, iA * Cross(rA, P));
                qA.Normalize();
                iA = RotateToFrame(pc.localInvIA, qA);
                cB += mB * P;
                qB += Derivative(qB, iB * Cross(rB, P));
                qB.Normalize();
                iB = RotateToFrame(pc.localInvIB, qB);
            }
        }
        pos[pc.indexA].x = cA; pos[pc.indexA].q = qA; pos[pc.indexA].I = iA;
        pos[pc.indexB].x = cB; pos[pc.indexB].q = qB; pos[pc.indexB].I = iB;
    }
    return minSeparation >= scalar(-3) * kLinearSlop;
}

// ------------------------------------------------------------------ island (island.cpp)
static Vec3 SolveGyro(const Quat& q, const Mat33& Ib, const Vec3& w1, scalar h) {  // island.cpp:60-79
    Vec3 w1b = MulC(q, w1);
    Mat33 J = Ib + h * (Skew(w1b) * Ib - Skew(Ib * w1b));
    Vec3 f = h * Cross(w1b, Ib * w1b);
    w1b -= J.Solve(f);
    return Mul(q, w1b);
}

void World::SolveIsland(const TimeStep& step, std::vector<u32>& ib, std::vector<Contact*>& ic, std::vector<u32>& ij) {
    scalar h = step.dt;
    u32 nb = u32(ib.size());
    std::vector<Position> pos(nb);
    std::vector<Velocity> vel(nb);
    for (u32 i = 0; i < nb; ++i) {  // 1. integrate velocities
        Body& b = bodies[ib[i]];
        Vec3 v = b.linearVelocity, w = b.angularVelocity;
        Vec3 x = b.sweep.worldCenter;
        Quat q = b.sweep.orientation;
        b.sweep.worldCenter0 = b.sweep.worldCenter;
        b.sweep.orientation0 = b.sweep.orientation;
        if (b.type == e_dynamicBody) {
            v += h * (Mul(b.gravityScale, gravity) + b.invMass * b.force);
            Vec3 dw1 = h * b.worldInvI * b.torque;
            Vec3 w2 = SolveGyro(q, b.I, w, h);
            Vec3 dw2 = w2 - w;
            w += dw1 + dw2;
            v.x *= scalar(1) / (scalar(1) + h * b.linearDamping.x);
            v.y *= scalar(1) / (scalar(1) + h * b.linearDamping.y);
            v.z *= scalar(1) / (scalar(1) + h * b.linearDamping.z);
            w.x *= scalar(1) / (scalar(1) + h * b.angularDamping.x);
            w.y *= scalar(1) / (scalar(1) + h * b.angularDamping.y);
            w.z *= scalar(1) / (scalar(1) + h * b.angularDamping.z);
        }
        vel[i].v = v; vel[i].w = w;
        pos[i].x = x; pos[i].q = q; pos[i].I = b.worldInvI;
    }
    std::vector<VC> vcs;
    std::vector<PC> pcs;
    InitContactConstraints(*this, step, ic, pos.data(), vcs, pcs);
    InitializeVelocityConstraints(ic, pos.data(), vel.data(), vcs, pcs);  // 2.
    if (step.warmStarting) WarmStartContacts(vcs, vel.data());
    SolveJointsInit(*this, step, ij, pos.data(), vel.data());
    if (step.warmStarting) SolveJointsWarmStart(*this, ij, vel.data());
    for (u32 i = 0; i < step.velocityIterations; ++i) {  // 3.
        SolveJointsVelocity(*this, step, ij, vel.data());
        SolveContactVelocities(vcs, vel.data());
    }
    if (step.warmStarting) StoreImpulses(ic, vcs);
    for (u32 i = 0; i < nb; ++i) {  // 4. integrate positions
        Body& b = bodies[ib[i]];
        Vec3 x = pos[i].x; Quat q = pos[i].q; Mat33 I = pos[i].I;
        Vec3 v = vel[i].v, w = vel[i].w;
        if (b.type != e_staticBody) {
            Vec3 translation = h * v;
            if (Dot(translation, translation) > kMaxTranslationSq) {
                scalar ratio = kMaxTranslation / Length(translation);
                v *= ratio;
            }
            Vec3 rotation = h * w;
            if (Dot(rotation, rotation) > kMaxRotationSq) {
                scalar ratio = kMaxRotation / Length(rotation);
                w *= ratio;
            }
            x += h * v;
            q = Integrate(q, w, h);
            I = RotateToFrame(b.invI, q);
        }
        pos[i].x = x; pos[i].q = q; pos[i].I = I;
        vel[i].v = v; vel[i].w = w;
    }
    bool positionsSolved = false;  // 5.
    for (u32 i = 0; i < step.positionIterations; ++i) {
        bool contactsSolved = SolveContactPositions(pcs, pos.data());
        bool jointsSolved = SolveJointsPosition(*this, ij, pos.data());
        if (contactsSolved && jointsSolved) { positionsSolved = true; break; }
    }
    for (u32 i = 0; i < nb; ++i) {  // 6. copy back
        Body& b = bodies[ib[i]];
        b.sweep.worldCenter = pos[i].x;
        b.sweep.orientation = pos[i].q;
        b.worldInvI = pos[i].I;
        b.linearVelocity = vel[i].v;
        b.angularVelocity = vel[i].w;
        b.SynchronizeTransform();
    }
    for (Contact* c : ic) lastIslandOrder.push_back(c);
    if (recordEvents)  // b3Island::Report (island.cpp:309-320)
        for (Contact* c : ic) events.push_back(Event{4u, c->fixtureA, c->fixtureB, c});
    if (sleeping) {  // 7. sleep
        scalar minSleepTime = kMaxScalar;
        for (u32 i = 0; i < nb; ++i) {
            Body& b = bodies[ib[i]];
            if (b.type == e_staticBody) continue;
            scalar sqrLin = Dot(b.linearVelocity, b.linearVelocity);
            scalar sqrAng = Dot(b.angularVelocity, b.angularVelocity);
            if ((b.flags & e_autoSleepFlag) == 0 || sqrLin > b.linearSleepTolerance * b.linearSleepTolerance ||
                sqrAng > b.angularSleepTolerance * b.angularSleepTolerance) {
                b.sleepTime = 0;
            } else {
                b.sleepTime += h;
            }
            minSleepTime = Min(minSleepTime, b.sleepTime);
        }
        if (minSleepTime >= kTimeToSleep && positionsSolved)
            for (u32 i = 0; i < nb; ++i) bodies[ib[i]].SetAwake(false);
    }
}

void World::Solve(const TimeStep& step) {  // world.cpp:253-450
    u32 nb = u32(bodies.size());
    for (Body& b : bodies) b.flags &= ~u32(e_islandFlag);
    for (Contact* c : contactList) c->flags &= ~1u;
    for (Joint& j : joints) j.islandFlag = false;
    lastIslandOrder.clear();
    lastIslandCount = 0;
    std::vector<u32> stack;
    stack.reserve(nb);
    std::vector<u32> ib;
    std::vector<Contact*> ic;
    std::vector<u32> ij;
    for (u32 s = nb; s-- > 0;) {  // body list order = reverse creation
        Body& seed = bodies[s];
        if (seed.flags & e_islandFlag) continue;
        if (!seed.IsAwake()) continue;
        if (seed.type == e_staticBody) continue;
        ib.clear(); ic.clear(); ij.clear();
        stack.clear();
        stack.push_back(s);
        seed.flags |= e_islandFlag;
        while (!stack.empty()) {
            u32 bi = stack.back();
            stack.pop_back();
            Body& b = bodies[bi];
            b.islandIndex = u32(ib.size());
            ib.push_back(bi);
            if (b.type == e_staticBody) continue;
            b.flags |= e_awakeFlag;
            for (u32 f : b.fixtures) {
                for (Contact* c : fixtures[f].contacts) {
                    if (c->flags & 1u) continue;
                    if ((c->flags & 2u) == 0) continue;
                    if (fixtures[c->fixtureA].isSensor || fixtures[c->fixtureB].isSensor) continue;
                    if (!FilterAllows(filterRespond, c->fixtureA, c->fixtureB)) continue;  // world.cpp:351-357
                    ic.push_back(c);
                    c->flags |= 1u;
                    u32 otherF = c->fixtureA == f ? c->fixtureB : c->fixtureA;
                    u32 other = fixtures[otherF].body;
                    if (bodies[other].flags & e_islandFlag) continue;
                    stack.push_back(other);
                    bodies[other].flags |= e_islandFlag;
                }
            }
            for (const JointEdge& je : b.joints) {
                Joint& j = joints[je.joint];
                if (j.islandFlag) continue;
                ij.push_back(je.joint);
                j.islandFlag = true;
                if (bodies[je.other].flags & e_islandFlag) continue;
                stack.push_back(je.other);
                bodies[je.other].flags |= e_islandFlag;
            }
        }
        SolveIsland(step, ib, ic, ij);
        ++lastIslandCount;
        for (u32 bi : ib)
            if (bodies[bi].type == e_staticBody) bodies[bi].flags &= ~u32(e_islandFlag);
    }
    for (u32 s = nb; s-- > 0;) {  // "Find New Pairs"
        Body& b = bodies[s];
        if ((b.flags & e_islandFlag) == 0) continue;
        if (b.type == e_staticBody) continue;
        SynchronizeFixtures(s);
    }
    for (Contact* c : contactList)
        if (c->isMesh) MeshContactSynchronize(*this, c);
    FindNewContacts();
}

void World::Step(scalar dt, u32 vit, u32 pit) {  // world.cpp:187-242
    events.clear();
    if (newContacts) { FindNewContacts(); newContacts = false; }
    TimeStep step;
    step.dt = dt;
    step.velocityIterations = vit;
    step.positionIterations = pit;
    step.inv_dt = dt > scalar(0) ? scalar(1) / dt : scalar(0);
    step.dtRatio = inv_dt0 * dt;
    step.warmStarting = warmStarting;
    UpdateContacts();
    if (step.dt > scalar(0)) Solve(step);
    if (step.dt > scalar(0)) inv_dt0 = step.inv_dt;
    if (clearForces)
        for (Body& b : bodies) { b.force.SetZero(); b.torque.SetZero(); }
}

}  // namespace orc
