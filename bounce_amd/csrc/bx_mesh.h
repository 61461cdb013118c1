// bx_mesh.h -- mesh contacts in the device step (included by bx_step.hip after the shape helpers).
//
// Same binary32 expression trees and visiting orders as the reference:
//   mesh_shape.cpp:44-50,142-175      b3MeshShape AABB and child triangles (scaled, with wings)
//   static_tree.h:215-255             b3StaticTree::QueryAABB: leaves in stack order
//   mesh_contact.cpp:27-90            fat AABB of the other shape in the mesh's unscaled frame
//   mesh_contact.cpp:97-226           SynchronizeFixture / MoveAABB
//   mesh_contact.cpp:228-258,293-371  FindPairs, TestOverlap, Collide
//   cluster.cpp:202-584               b3ClusterSolver: k-means over contact normals (at most three
//                                     clusters, ten iterations) and b3ReducePolygon per cluster
//
// A mesh contact is updated by one lane: the triangle list of the contact is walked in query
// order, each triangle's manifold is reduced to observations (world normal, point, separation)
// in the lane's stripe of BatchView::meshScratch, the observations are clustered, and each
// cluster's reduced points are copied from the triangle manifolds they came from, recomputed
// (the triangle narrowphase is a pure function of the two transforms and the triangle).
#pragma once
// (included inside namespace bx by bx_step.hip)

constexpr u32 kMeshShape = 4u;
static_assert(kMeshLanes == u32(kBlock), "one clustering stripe per lane of the contact update");

__device__ __forceinline__ Manifold& ContactManifold(const BatchView& B, const WorldInfo& wi, Contact& c, u32 slot,
                                                     u32 j) {
    return j == 0u ? c.manifold : B.meshContacts[wi.meshBase + slot].extra[j - 1u];
}

// b3AABB scaled by the inverse of the mesh scale (mesh_contact.cpp:45-50): bounds re-sorted per axis
// for negative scales.
__device__ __forceinline__ AABB ScaledInv(const AABB& a, const Vec3& scale) {
    Vec3 inv(scalar(1) / scale.x, scalar(1) / scale.y, scalar(1) / scale.z);
    Vec3 lo = Mul(inv, a.lowerBound), hi = Mul(inv, a.upperBound);
    AABB out;
    out.lowerBound = Min(lo, hi);
    out.upperBound = Max(lo, hi);
    return out;
}

// b3MeshShape::GetChildTriangle (mesh_shape.cpp:142-175): triangle `index` scaled, as a triangle
// fixture with the mesh's radius, and its wings.
__device__ __forceinline__ void MeshChild(const BatchView& B, const Fixture& mf, u32 index, Fixture& t, TriWings& w) {
    const MeshInfo mi = B.meshes[mf.shapeIndex];
    const MeshTri tri = B.meshTris[mi.tBase + index];
    const Vec3* v = B.meshVertices + mi.vBase;
    const Vec3 scale = mf.v1;
    t = mf;
    t.shapeType = 2u;
    t.v1 = Mul(scale, v[tri.v1]);
    t.v2 = Mul(scale, v[tri.v2]);
    t.v3 = Mul(scale, v[tri.v3]);
    w.mask = (tri.u1 != kNull ? 1u : 0u) | (tri.u2 != kNull ? 2u : 0u) | (tri.u3 != kNull ? 4u : 0u);
    w.e1 = tri.u1 != kNull ? Mul(scale, v[tri.u1]) : Vec3(scalar(0), scalar(0), scalar(0));
    w.e2 = tri.u2 != kNull ? Mul(scale, v[tri.u2]) : Vec3(scalar(0), scalar(0), scalar(0));
    w.e3 = tri.u3 != kNull ? Mul(scale, v[tri.u3]) : Vec3(scalar(0), scalar(0), scalar(0));
}

__device__ __forceinline__ u32* MeshTriangles(const BatchView& B, const WorldInfo& wi, u32 slot) {
    return B.meshTriLists + wi.meshTriBase + slot * wi.meshTriCap;
}

// b3MeshContact constructor (mesh_contact.cpp:27-90), at contact creation.
__device__ void MeshContactCreate(const BatchView& B, const WorldInfo& wi, u32 slot, const Contact& c) {
    const Fixture fA = B.fixtures[wi.fixtureBase + c.fixtureA];
    const Fixture fB = B.fixtures[wi.fixtureBase + c.fixtureB];
    const bool meshIsA = fA.shapeType == kMeshShape;
    const Fixture& fm = meshIsA ? fA : fB;
    const Fixture& fo = meshIsA ? fB : fA;
    const Body& bm = B.bodies[wi.bodyBase + fm.body];
    const Body& bo = B.bodies[wi.bodyBase + fo.body];
    Transform xf = MulT(Transform(bm.xfT, bm.q), Transform(bo.xfT, bo.q));
    AABB fat;
    ComputeAABB(B, fo, xf, fat);
    fat = ScaledInv(fat, fm.v1);
    fat.Extend(kAabbExtension);
    MeshContact& rec = B.meshContacts[wi.meshBase + slot];
    rec.meshAABB = fat;
    rec.moved = 1u;
    rec.triangleCount = 0u;
}

// b3MeshContact::SynchronizeFixture + MoveAABB (mesh_contact.cpp:97-226), after the island solve.
__device__ void MeshContactSynchronize(const BatchView& B, const WorldInfo& wi, u32 slot, const Contact& c) {
    const Fixture fA = B.fixtures[wi.fixtureBase + c.fixtureA];
    const Fixture fB = B.fixtures[wi.fixtureBase + c.fixtureB];
    const bool meshIsA = fA.shapeType == kMeshShape;
    const Fixture& fm = meshIsA ? fA : fB;
    const Fixture& fo = meshIsA ? fB : fA;
    const Body& bm = B.bodies[wi.bodyBase + fm.body];
    const Body& bo = B.bodies[wi.bodyBase + fo.body];
    Transform xfM(bm.xfT, bm.q);
    Transform xfO0(bo.x0 - Mul(bo.q0, bo.localCenter), bo.q0);
    Transform xf1 = MulT(xfM, xfO0);
    Transform xf2 = MulT(xfM, Transform(bo.xfT, bo.q));
    AABB aabb1, aabb2;
    ComputeAABB(B, fo, xf1, aabb1);
    ComputeAABB(B, fo, xf2, aabb2);
    aabb1 = ScaledInv(aabb1, fm.v1);
    aabb2 = ScaledInv(aabb2, fm.v1);
    AABB aabb = Combine(aabb1, aabb2);
    Vec3 displacement = aabb2.GetCenter() - aabb1.GetCenter();
    MeshContact& rec = B.meshContacts[wi.meshBase + slot];
    if (rec.meshAABB.Contains(aabb)) {
        rec.moved = 0u;
        return;
    }
    AABB fat = aabb;
    fat.Extend(kAabbExtension);
    if (displacement.x < scalar(0)) fat.lowerBound.x += kAabbMultiplier * displacement.x;
    else fat.upperBound.x += kAabbMultiplier * displacement.x;
    if (displacement.y < scalar(0)) fat.lowerBound.y += kAabbMultiplier * displacement.y;
    else fat.upperBound.y += kAabbMultiplier * displacement.y;
    if (displacement.z < scalar(0)) fat.lowerBound.z += kAabbMultiplier * displacement.z;
    else fat.upperBound.z += kAabbMultiplier * displacement.z;
    rec.meshAABB = fat;
    rec.moved = 1u;
}

// b3MeshContact::FindPairs (mesh_contact.cpp:228-258): when the fat AABB moved, the triangles whose
// tree leaves overlap it, in b3StaticTree::QueryAABB's stack order. Returns false on overflow of
// the contact's triangle list or of the traversal stack (the list then holds what fitted).
__device__ bool MeshContactFindPairs(const BatchView& B, const WorldInfo& wi, u32 slot, const Contact& c) {
    MeshContact& rec = B.meshContacts[wi.meshBase + slot];
    if (!rec.moved) return true;
    const Fixture fA = B.fixtures[wi.fixtureBase + c.fixtureA];
    const Fixture& fm = fA.shapeType == kMeshShape ? fA : B.fixtures[wi.fixtureBase + c.fixtureB];
    const MeshInfo mi = B.meshes[fm.shapeIndex];
    const TreeNode* nodes = B.meshNodes + mi.nBase;
    const AABB box = rec.meshAABB;
    u32* out = MeshTriangles(B, wi, slot);
    u32 stack[kTreeStack];
    u32 sp = 0, n = 0;
    bool ok = true;
    stack[sp++] = mi.root;
    while (sp > 0) {
        u32 id = stack[--sp];
        if (id == kNull) continue;
        const TreeNode node = nodes[id];
        if (!TestOverlap(node.aabb, box)) continue;
        if (node.child1 == kNull) {
            if (n < wi.meshTriCap) out[n] = node.index;
            else ok = false;
            ++n;
        } else if (sp + 2u <= kTreeStack) {
            stack[sp++] = node.child1;
            stack[sp++] = node.child2;
        } else {
            ok = false;
        }
    }
    rec.triangleCount = Min(n, wi.meshTriCap);
    return ok;
}

// b3MeshContact::TestOverlap (mesh_contact.cpp:293-326) for sensor contacts.
__device__ bool MeshContactTestOverlap(const BatchView& B, const WorldInfo& wi, u32 slot, const Contact& c,
                                       const Transform& xfA, const Transform& xfB) {
    const Fixture fA = B.fixtures[wi.fixtureBase + c.fixtureA];
    const Fixture fB = B.fixtures[wi.fixtureBase + c.fixtureB];
    const bool meshIsA = fA.shapeType == kMeshShape;
    const Fixture& fo = meshIsA ? fB : fA;
    HullView ho;
    if (fo.shapeType == 3u) ho = GetHull(B, fo.shapeIndex);
    const ProxyD po = MakeProxy(fo, &ho);
    const u32 nt = B.meshContacts[wi.meshBase + slot].triangleCount;
    const u32* tris = MeshTriangles(B, wi, slot);
    for (u32 i = 0; i < nt; ++i) {
        Fixture t;
        TriWings w;
        MeshChild(B, meshIsA ? fA : fB, tris[i], t, w);
        const ProxyD pt = MakeProxy(t, nullptr);
        GJKOut q = meshIsA ? GJK(xfA, pt, xfB, po, true) : GJK(xfA, po, xfB, pt, true);
        const scalar kTol = scalar(10) * kEpsilon;
        if (q.distance <= kTol) return true;
    }
    return false;
}

// The manifold of one child triangle (hull_mesh_contact.cpp, mesh_sphere_contact.cpp,
// mesh_capsule_contact.cpp), from a fresh manifold, its points tagged with the triangle.
__device__ __noinline__ void MeshTriangleManifold(const BatchView& B, const Fixture& fA, const Fixture& fB,
                                                  const Transform& xfA, const Transform& xfB, u32 triangle, Manifold& m,
                                                  FaceScratch& scratch, u32& overflow) {
    for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
        ManifoldPoint& p = m.points[k];
        p.localNormal1 = p.localPoint1 = p.localPoint2 = Vec3(scalar(0), scalar(0), scalar(0));
        p.normalImpulse = 0;
        p.triangleKey = 0u;
        MakeKey(p.key, 0u, 0u, 0u, 0u);
        p.persisting = 0u;
        p.edgeContact = 0u;
    }
    m.pointCount = 0;
    m.tangentImpulse.SetZero();
    m.motorImpulse = m.motorSpeed = m.tangentSpeed1 = m.tangentSpeed2 = 0;
    Fixture t;
    TriWings w;
    if (fA.shapeType == kMeshShape) {
        MeshChild(B, fA, triangle, t, w);
        if (fB.shapeType == 0u) CollideTriangleAndSphere(m, xfA, t, xfB, fB, &w);
        else if (fB.shapeType == 1u) CollideTriangleAndCapsule(m, xfA, t, xfB, fB, &w);
    } else {
        MeshChild(B, fB, triangle, t, w);
        HullView hA = GetHull(B, fA.shapeIndex);
        CollideHullAndTriangle(m, xfA, hA, fA.radius, xfB, t, scratch, overflow, &w);
    }
    for (u32 k = 0; k < m.pointCount; ++k) m.points[k].triangleKey = triangle;
}

// Cluster centroids of b3ClusterSolver (at most three), kept in registers.
struct Clusters {
    Vec3 c0, c1, c2;
    u32 n;
    __device__ __forceinline__ Vec3 Get(u32 i) const { return i == 0u ? c0 : (i == 1u ? c1 : c2); }
    __device__ __forceinline__ void Set(u32 i, const Vec3& v) {
        if (i == 0u) c0 = v;
        else if (i == 1u) c1 = v;
        else c2 = v;
    }
    __device__ __forceinline__ u32 Find(const Vec3& p) const {  // cluster.cpp:247-263
        u32 bestIndex = 0;
        scalar bestValue = kMaxScalar;
        for (u32 i = 0; i < n; ++i) {
            Vec3 d = p - Get(i);
            scalar metric = Dot(d, d);
            if (metric < bestValue) { bestValue = metric; bestIndex = i; }
        }
        return bestIndex;
    }
    __device__ __forceinline__ void Add(const Vec3& centroid) {  // cluster.cpp:265-290
        if (n == 0u) { c0 = centroid; n = 1; return; }
        u32 best = Find(centroid);
        const scalar kTol = scalar(0.05);
        Vec3 d = centroid - Get(best);
        if (Dot(d, d) <= kTol * kTol) {
            Vec3 c = Get(best);
            c += centroid;
            c.Normalize();
            Set(best, c);
            return;
        }
        Set(n, centroid);
        ++n;
    }
};

__device__ __forceinline__ u32 ObsCluster(const MeshObs& o) { return o.tag >> 28; }
__device__ __forceinline__ void SetObsCluster(MeshObs& o, u32 k) { o.tag = (o.tag & 0x0FFFFFFFu) | (k << 28); }

// b3ClusterSolver::InitializeClusters + Solve (cluster.cpp:292-469) over obs[0, n).
__device__ void SolveClusters(MeshObs* obs, u32 n, Clusters& cs) {
    cs.n = 0;
    const u32 kMaxClusters = 3;
    if (n <= kMaxClusters) {
        for (u32 i = 0; i < n; ++i) cs.Add(obs[i].n);
    } else {
        cs.Add(obs[0].n);
        u32 chosen1;
        {
            Vec3 A = cs.c0;
            u32 index = 0;
            scalar max = -kMaxScalar;
            for (u32 i = 1; i < n; ++i) {
                Vec3 d = A - obs[i].n;
                scalar dd = Dot(d, d);
                if (dd > max) { max = dd; index = i; }
            }
            cs.Add(obs[index].n);
            chosen1 = index;
        }
        if (cs.n == 1u) {
            Vec3 A = cs.c0;
            u32 index = 0;
            scalar max = -kMaxScalar;
            for (u32 i = 1; i < n; ++i) {
                if (i == chosen1) continue;
                Vec3 d = A - obs[i].n;
                scalar dd = Dot(d, d);
                if (dd > max) { max = dd; index = i; }
            }
            cs.Add(obs[index].n);
        } else {
            Vec3 A = cs.c0, Bv = cs.c1;
            u32 index = 0;
            scalar max = -kMaxScalar;
            for (u32 i = 1; i < n; ++i) {
                if (i == chosen1) continue;
                Vec3 C = obs[i].n;
                Vec3 Q = ClosestPointOnSegment(C, A, Bv);
                Vec3 d = C - Q;
                scalar dd = Dot(d, d);
                if (dd > max) { max = dd; index = i; }
            }
            cs.Add(obs[index].n);
        }
    }
    const u32 kMaxIters = 10;
    for (u32 iter = 0; iter < kMaxIters; ++iter) {
        Vec3 sum0(scalar(0), scalar(0), scalar(0)), sum1 = sum0, sum2 = sum0;
        u32 cnt0 = 0, cnt1 = 0, cnt2 = 0;
        for (u32 i = 0; i < n; ++i) {
            MeshObs& o = obs[i];
            u32 k = cs.Find(o.n);
            SetObsCluster(o, k);
            if (k == 0u) { sum0 += o.n; ++cnt0; }
            else if (k == 1u) { sum1 += o.n; ++cnt1; }
            else { sum2 += o.n; ++cnt2; }
        }
        if (cnt0 > 0u) { sum0 /= scalar(cnt0); cs.c0 = sum0; }
        if (cs.n > 1u && cnt1 > 0u) { sum1 /= scalar(cnt1); cs.c1 = sum1; }
        if (cs.n > 2u && cnt2 > 0u) { sum2 /= scalar(cnt2); cs.c2 = sum2; }
    }
    // drop empty clusters, renumbering the observations in order
    u32 count[3] = {0u, 0u, 0u};
    for (u32 i = 0; i < n; ++i) ++count[ObsCluster(obs[i])];
    u32 remap[3];
    Clusters used;
    used.n = 0;
    for (u32 k = 0; k < cs.n; ++k) {
        remap[k] = used.n;
        if (count[k] > 0u) {
            used.Set(used.n, cs.Get(k));
            ++used.n;
        }
    }
    for (u32 i = 0; i < n; ++i) SetObsCluster(obs[i], remap[ObsCluster(obs[i])]);
    cs = used;
}

// b3ReducePolygon (cluster.cpp:100-200) over n polygon vertices at pos(i), choices tracked by index.
template <class Pos>
__device__ u32 ReducePolygonIndexed(ClusterVertex out[kMaxManifoldPoints], Pos pos, u32 n, const Vec3& N,
                                    u32 initialPoint) {
    if (n <= kMaxManifoldPoints) {
        for (u32 i = 0; i < n; ++i) {
            out[i].position = pos(i);
            out[i].clipIndex = i;
        }
        SortPolygon(out, n, N);
        return n;
    }
    u32 count = 0;
    out[count].position = pos(initialPoint);
    out[count].clipIndex = initialPoint;
    ++count;
    u32 chosen0 = initialPoint, chosen1 = kNull, chosen2 = kNull;
    {
        Vec3 A = out[0].position;
        u32 index = 0;
        scalar max = -kMaxScalar;
        for (u32 i = 0; i < n; ++i) {
            if (i == chosen0) continue;
            Vec3 d = pos(i) - A;
            scalar dd = Dot(d, d);
            if (dd > max) { max = dd; index = i; }
        }
        if (max < kEpsilon * kEpsilon) return count;
        out[count].position = pos(index);
        out[count].clipIndex = index;
        ++count;
        chosen1 = index;
    }
    {
        Vec3 A = out[0].position, Bv = out[1].position;
        u32 index = 0;
        scalar max = -kMaxScalar;
        for (u32 i = 0; i < n; ++i) {
            if (i == chosen0 || i == chosen1) continue;
            Vec3 C = pos(i);
            scalar a2 = Abs(Dot(Cross(Bv - A, C - A), N));
            if (a2 > max) { max = a2; index = i; }
        }
        const scalar kAreaTol = scalar(0.01);
        if (max < scalar(2) * kAreaTol) return count;
        Vec3 C = pos(index);
        if (!IsCCW(A, Bv, C, N)) Swap(out[0], out[1]);
        out[count].position = C;
        out[count].clipIndex = index;
        ++count;
        chosen2 = index;
    }
    {
        Vec3 A = out[0].position, Bv = out[1].position;
        u32 index = 0;
        scalar min = kMaxScalar;
        for (u32 i = 0; i < n; ++i) {
            if (i == chosen0 || i == chosen1 || i == chosen2) continue;
            Vec3 D = pos(i);
            scalar sa2 = Dot(Cross(Bv - A, D - A), N);
            if (sa2 < min) { min = sa2; index = i; }
        }
        const scalar kAreaTol = scalar(0.01);
        if (Abs(min) < scalar(2) * kAreaTol) return count;
        out[count].position = pos(index);
        out[count].clipIndex = index;
        ++count;
    }
    SortPolygon(out, count, N);
    return count;
}

// b3MeshContact::Collide (mesh_contact.cpp:328-371) + b3ClusterSolver::Run (cluster.cpp:471-584):
// the cluster manifolds of a mesh contact into nm[0, returned count), which arrive initialized
// (b3Manifold::Initialize). `lane` is the lane's stripe of the world's clustering scratch and
// `scratch` the caller's hull clipping buffers.
__device__ __noinline__ u32 MeshCollide(const BatchView& B, const WorldInfo& wi, u32 slot, const Fixture& fA,
                                        const Fixture& fB, const Transform& xfA, const Transform& xfB, MeshObs* lane,
                                        Manifold* nm, FaceScratch& scratch, u32& overflow) {
    const u32 nt = B.meshContacts[wi.meshBase + slot].triangleCount;
    const u32* tris = MeshTriangles(B, wi, slot);
    MeshObs* obs = lane + kMeshManifoldObs;
    u32* members = reinterpret_cast<u32*>(obs + 4u * wi.meshTriCap);
    const scalar rA = fA.radius, rB = fB.radius;
    // 1. observations: the points of every triangle manifold, in triangle then point order
    u32 no = 0;
    for (u32 i = 0; i < nt; ++i) {
        Manifold m;
        MeshTriangleManifold(B, fA, fB, xfA, xfB, tris[i], m, scratch, overflow);
        for (u32 j = 0; j < m.pointCount; ++j) {
            const ManifoldPoint& p = m.points[j];
            // b3WorldManifold point (collision.cpp:63-75)
            Vec3 nA = Mul(xfA.rotation, p.localNormal1);
            Vec3 cA = Mul(xfA, p.localPoint1);
            Vec3 cB = Mul(xfB, p.localPoint2);
            Vec3 pA = cA + rA * nA;
            Vec3 pB = cB - rB * nA;
            MeshObs o;
            o.n = nA;
            o.p = scalar(0.5) * (pA + pB);
            o.sep = Dot(cB - cA, nA) - rA - rB;
            o.tag = (i << 2) | j;
            obs[no++] = o;
        }
    }
    if (no == 0u) return 0u;
    // 2. clusters over the normals
    Clusters cs;
    SolveClusters(obs, no, cs);
    // 3. one manifold per cluster: its points moved onto the cluster plane, reduced to at most four
    u32 outCount = 0;
    for (u32 k = 0; k < cs.n; ++k) {
        Vec3 normal(scalar(0), scalar(0), scalar(0));
        u32 np = 0;
        for (u32 i = 0; i < no; ++i) {
            if (ObsCluster(obs[i]) != k) continue;
            members[np++] = i;
            normal += obs[i].n;
        }
        if (np == 0u) continue;
        normal.Normalize();
        u32 minIndex = 0;
        scalar minSeparation = kMaxScalar;
        for (u32 j = 0; j < np; ++j) {
            MeshObs& o = obs[members[j]];
            if (o.sep < minSeparation) { minIndex = j; minSeparation = o.sep; }
            o.p = o.p + o.sep * normal;
        }
        ClusterVertex quad[kMaxManifoldPoints];
        u32 nq = ReducePolygonIndexed(quad, [&](u32 j) { return obs[members[j]].p; }, np, normal, minIndex);
        Manifold& out = nm[outCount++];
        for (u32 j = 0; j < nq; ++j) {
            const u32 tag = obs[members[quad[j].clipIndex]].tag & 0x0FFFFFFFu;
            Manifold m;
            MeshTriangleManifold(B, fA, fB, xfA, xfB, tris[tag >> 2], m, scratch, overflow);
            const u32 pt = tag & 3u;
            out.points[j] = pt == 0u ? m.points[0] : (pt == 1u ? m.points[1] : (pt == 2u ? m.points[2] : m.points[3]));
        }
        out.pointCount = nq;
    }
    return outCount;
}
