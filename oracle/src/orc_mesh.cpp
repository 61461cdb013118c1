// orc_mesh.cpp -- oracle (test infrastructure): triangle meshes and mesh contacts.
//
// Restates, in the reference's binary32 expression order and visiting order:
//   mesh.h:80-96, mesh.cpp:21-124    b3Mesh: triangle AABBs, BuildTree, BuildAdjacency (edge wings)
//   static_tree.cpp:40-170           b3StaticTree::Build (median-split on the longest axis, std::sort)
//   static_tree.h:215-255            b3StaticTree::QueryAABB (stack order)
//   mesh_shape.cpp:37-60,142-175     b3MeshShape mass, AABB, GetChildTriangle
//   gjk.cpp:76-92                    b3GJKProxy of a mesh triangle
//   mesh_contact.cpp:27-371          b3MeshContact: fat AABB in the mesh frame, triangle query, collide
//   cluster.cpp:202-584              b3ClusterSolver: k-means over contact normals, reduction
//   hull_mesh_contact.cpp, mesh_sphere_contact.cpp, mesh_capsule_contact.cpp (Evaluate)
#include <algorithm>
#include "orc_world.h"
#include "../../include/bxstep.h"

namespace orc {

// ------------------------------------------------------------------ static tree
static u32 LongestAxis(const AABB& a) {  // aabb.h:164-184
    Vec3 d = a.upperBound - a.lowerBound;
    scalar maxValue = d.x;
    u32 maxIndex = 0;
    if (d.y > maxValue) { maxValue = d.y; maxIndex = 1; }
    if (d.z > maxValue) { maxValue = d.z; maxIndex = 2; }
    return maxIndex;
}

struct AxisLess {  // b3SortPredicate (static_tree.cpp:40-52)
    const std::vector<AABB>* aabbs;
    u32 axis;
    bool operator()(u32 i1, u32 i2) const {
        Vec3 c1 = (*aabbs)[i1].GetCenter();
        Vec3 c2 = (*aabbs)[i2].GetCenter();
        return c1[axis] < c2[axis];
    }
};

static u32 Partition(const AABB& aabb, const std::vector<AABB>& aabbs, u32* indices, u32 count) {
    u32 axis = LongestAxis(aabb);
    AxisLess less;
    less.aabbs = &aabbs;
    less.axis = axis;
    std::sort(indices, indices + count, less);
    scalar splitPos = aabb.GetCenter()[axis];
    u32 k = 0;
    while (k < count) {
        Vec3 center = aabbs[indices[k]].GetCenter();
        if (center[axis] > splitPos) break;
        ++k;
    }
    if (k == 0 || k == count) k = count / 2;
    return k;
}

static u32 BuildNode(Mesh& m, const std::vector<AABB>& aabbs, u32* indices, u32 count) {
    u32 id = u32(m.nodes.size());
    m.nodes.push_back(StaticNode());
    if (count <= 1) {
        m.nodes[id].aabb = aabbs[indices[0]];
        m.nodes[id].child1 = kNullU32;
        m.nodes[id].child2 = kNullU32;
        m.nodes[id].index = indices[0];
        return id;
    }
    AABB aabb = aabbs[indices[0]];
    for (u32 i = 1; i < count; ++i) aabb = Combine(aabb, aabbs[indices[i]]);
    m.nodes[id].aabb = aabb;
    m.nodes[id].index = kNullU32;
    u32 k = Partition(aabb, aabbs, indices, count);
    u32 c1 = BuildNode(m, aabbs, indices, k);
    u32 c2 = BuildNode(m, aabbs, indices + k, count - k);
    m.nodes[id].child1 = c1;
    m.nodes[id].child2 = c2;
    return id;
}

// ------------------------------------------------------------------ mesh
static AABB TriangleAABB(const Mesh& m, u32 i) {  // mesh.h:80-96
    const MeshTriangle& t = m.triangles[i];
    Vec3 v1 = m.vertices[t.v1], v2 = m.vertices[t.v2], v3 = m.vertices[t.v3];
    AABB a;
    a.lowerBound = Min(v1, Min(v2, v3));
    a.upperBound = Max(v1, Max(v2, v3));
    a.Extend(kLinearSlop);
    return a;
}

static u32& Wing(MeshTriangle& t, u32 i) { return i == 0 ? t.u1 : (i == 1 ? t.u2 : t.u3); }
static u32 Vert(const MeshTriangle& t, u32 i) { return i == 0 ? t.v1 : (i == 1 ? t.v2 : t.v3); }

static void BuildAdjacency(Mesh& m) {  // mesh.cpp:35-112
    u32 n = u32(m.triangles.size());
    for (u32 i = 0; i < n; ++i) m.triangles[i].u1 = m.triangles[i].u2 = m.triangles[i].u3 = kNullU32;
    for (u32 i = 0; i < n; ++i) {
        MeshTriangle& t1 = m.triangles[i];
        for (u32 j1 = 0; j1 < 3; ++j1) {
            u32 k1 = j1 + 1 < 3 ? j1 + 1 : 0;
            u32 t1v1 = Vert(t1, j1), t1v2 = Vert(t1, k1);
            u32& u1 = Wing(t1, j1);
            if (u1 != kNullU32) continue;
            for (u32 j = i + 1; j < n; ++j) {
                MeshTriangle& t2 = m.triangles[j];
                for (u32 j2 = 0; j2 < 3; ++j2) {
                    u32 k2 = j2 + 1 < 3 ? j2 + 1 : 0;
                    u32 t2v1 = Vert(t2, j2), t2v2 = Vert(t2, k2);
                    u32& u2 = Wing(t2, j2);
                    if ((t1v1 == t2v2 && t1v2 == t2v1) || (t1v1 == t2v1 && t1v2 == t2v2)) {
                        u32 k3 = k1 + 1 < 3 ? k1 + 1 : 0;
                        u2 = Vert(t1, k3);
                        u32 k4 = k2 + 1 < 3 ? k2 + 1 : 0;
                        u1 = Vert(t2, k4);
                        break;
                    }
                }
                if (u1 != kNullU32) break;
            }
        }
    }
}

bool BuildMesh(Mesh& m, const bx_mesh_desc& d) {
    if (d.vertex_count == 0 || d.triangle_count == 0) return false;
    m.vertices.resize(d.vertex_count);
    for (u32 i = 0; i < d.vertex_count; ++i)
        m.vertices[i] = Vec3(d.vertices[3 * i], d.vertices[3 * i + 1], d.vertices[3 * i + 2]);
    m.triangles.resize(d.triangle_count);
    for (u32 i = 0; i < d.triangle_count; ++i) {
        MeshTriangle& t = m.triangles[i];
        t.v1 = d.triangles[3 * i];
        t.v2 = d.triangles[3 * i + 1];
        t.v3 = d.triangles[3 * i + 2];
        if (t.v1 >= d.vertex_count || t.v2 >= d.vertex_count || t.v3 >= d.vertex_count) return false;
    }
    // b3Mesh::BuildTree (mesh.cpp:21-33), then BuildAdjacency, as the reference harness builds it
    std::vector<AABB> aabbs(d.triangle_count);
    for (u32 i = 0; i < d.triangle_count; ++i) aabbs[i] = TriangleAABB(m, i);
    std::vector<u32> indices(d.triangle_count);
    for (u32 i = 0; i < d.triangle_count; ++i) indices[i] = i;
    m.nodes.clear();
    m.nodes.reserve(2 * d.triangle_count - 1);
    m.root = BuildNode(m, aabbs, indices.data(), d.triangle_count);
    BuildAdjacency(m);
    return true;
}

// b3StaticTree::QueryAABB (static_tree.h:215-255): leaves reported in stack order
static void QueryTree(const Mesh& m, const AABB& aabb, std::vector<u32>& out);
void QueryMeshTree(const Mesh& m, const AABB& aabb, std::vector<u32>& out) { QueryTree(m, aabb, out); }
static void QueryTree(const Mesh& m, const AABB& aabb, std::vector<u32>& out) {
    std::vector<u32> stack;
    stack.push_back(m.root);
    while (!stack.empty()) {
        u32 id = stack.back();
        stack.pop_back();
        if (id == kNullU32) continue;
        const StaticNode& n = m.nodes[id];
        if (!TestOverlap(n.aabb, aabb)) continue;
        if (n.child1 == kNullU32) {
            out.push_back(n.index);
        } else {
            stack.push_back(n.child1);
            stack.push_back(n.child2);
        }
    }
}

// ------------------------------------------------------------------ mesh shape
void MeshAABB(const Shape& s, const Transform& xf, AABB* out) {  // mesh_shape.cpp:44-50, aabb.h:65-74
    const Mesh& m = *s.mesh;
    out->lowerBound = out->upperBound = Mul(xf, Mul(s.scale, m.vertices[0]));
    for (u32 i = 1; i < m.vertices.size(); ++i) {
        Vec3 v = Mul(xf, Mul(s.scale, m.vertices[i]));
        out->lowerBound = Min(out->lowerBound, v);
        out->upperBound = Max(out->upperBound, v);
    }
    out->Extend(s.radius);
}

static void ChildTriangle(Shape& t, const Shape& s, u32 index) {  // mesh_shape.cpp:142-175
    const Mesh& m = *s.mesh;
    const MeshTriangle& tri = m.triangles[index];
    t.type = e_triangle;
    t.v1 = Mul(s.scale, m.vertices[tri.v1]);
    t.v2 = Mul(s.scale, m.vertices[tri.v2]);
    t.v3 = Mul(s.scale, m.vertices[tri.v3]);
    t.radius = s.radius;
    t.hasE1 = tri.u1 != kNullU32;
    t.hasE2 = tri.u2 != kNullU32;
    t.hasE3 = tri.u3 != kNullU32;
    if (t.hasE1) t.e1 = Mul(s.scale, m.vertices[tri.u1]);
    if (t.hasE2) t.e2 = Mul(s.scale, m.vertices[tri.u2]);
    if (t.hasE3) t.e3 = Mul(s.scale, m.vertices[tri.u3]);
}

// ------------------------------------------------------------------ mesh contacts
static AABB ScaledInv(AABB a, const Vec3& scale) {  // b3AABB::Scale by the inverse mesh scale
    Vec3 inv(scalar(1) / scale.x, scalar(1) / scale.y, scalar(1) / scale.z);
    Vec3 lo = Mul(inv, a.lowerBound), hi = Mul(inv, a.upperBound);
    a.lowerBound = Min(lo, hi);
    a.upperBound = Max(lo, hi);
    return a;
}

static bool MeshIsA(const World& w, const Contact* c) { return w.fixtures[c->fixtureA].shape.type == e_mesh; }

void MeshContactCreate(World& w, Contact* c) {  // mesh_contact.cpp:27-90
    bool meshIsA = MeshIsA(w, c);
    const Fixture& fm = w.fixtures[meshIsA ? c->fixtureA : c->fixtureB];
    const Fixture& fo = w.fixtures[meshIsA ? c->fixtureB : c->fixtureA];
    Transform xfM = w.bodies[fm.body].xf, xfO = w.bodies[fo.body].xf;
    Transform xf = MulT(xfM, xfO);
    AABB fat;
    ComputeAABB(fo.shape, xf, &fat);
    fat = ScaledInv(fat, fm.shape.scale);
    fat.Extend(kAabbExtension);
    c->meshAABB = fat;
    c->meshAABBMoved = true;
    c->triangles.clear();
}

void MeshContactSynchronize(World& w, Contact* c) {  // mesh_contact.cpp:97-174 (SynchronizeFixture)
    bool meshIsA = MeshIsA(w, c);
    const Fixture& fm = w.fixtures[meshIsA ? c->fixtureA : c->fixtureB];
    const Fixture& fo = w.fixtures[meshIsA ? c->fixtureB : c->fixtureA];
    const Body& bm = w.bodies[fm.body];
    const Body& bo = w.bodies[fo.body];
    Transform xfO0;
    xfO0.rotation = bo.sweep.orientation0;
    xfO0.translation = bo.sweep.worldCenter0 - Mul(xfO0.rotation, bo.sweep.localCenter);
    Transform xf1 = MulT(bm.xf, xfO0);
    Transform xf2 = MulT(bm.xf, bo.xf);
    AABB aabb1, aabb2;
    ComputeAABB(fo.shape, xf1, &aabb1);
    ComputeAABB(fo.shape, xf2, &aabb2);
    aabb1 = ScaledInv(aabb1, fm.shape.scale);
    aabb2 = ScaledInv(aabb2, fm.shape.scale);
    AABB aabb = Combine(aabb1, aabb2);
    Vec3 displacement = aabb2.GetCenter() - aabb1.GetCenter();
    // MoveAABB (mesh_contact.cpp:176-226)
    if (c->meshAABB.Contains(aabb)) { c->meshAABBMoved = false; return; }
    AABB fat = aabb;
    fat.Extend(kAabbExtension);
    if (displacement.x < scalar(0)) fat.lowerBound.x += kAabbMultiplier * displacement.x;
    else fat.upperBound.x += kAabbMultiplier * displacement.x;
    if (displacement.y < scalar(0)) fat.lowerBound.y += kAabbMultiplier * displacement.y;
    else fat.upperBound.y += kAabbMultiplier * displacement.y;
    if (displacement.z < scalar(0)) fat.lowerBound.z += kAabbMultiplier * displacement.z;
    else fat.upperBound.z += kAabbMultiplier * displacement.z;
    c->meshAABB = fat;
    c->meshAABBMoved = true;
}

void MeshContactFindPairs(World& w, Contact* c) {  // mesh_contact.cpp:228-258
    if (!c->meshAABBMoved) return;
    c->triangles.clear();
    const Fixture& fm = w.fixtures[MeshIsA(w, c) ? c->fixtureA : c->fixtureB];
    QueryTree(*fm.shape.mesh, c->meshAABB, c->triangles);
}

// b3GJKProxy of one triangle of a mesh (gjk.cpp:76-92)
static void MeshProxy(GJKProxy& p, const Shape& s, u32 index) {
    const Mesh& m = *s.mesh;
    const MeshTriangle& t = m.triangles[index];
    p.buffer[0] = Mul(s.scale, m.vertices[t.v1]);
    p.buffer[1] = Mul(s.scale, m.vertices[t.v2]);
    p.buffer[2] = Mul(s.scale, m.vertices[t.v3]);
    p.count = 3;
    p.vertices = p.buffer;
    p.radius = s.radius;
}

bool MeshContactTestOverlap(World& w, Contact* c) {  // mesh_contact.cpp:293-326, collision.cpp:100-112
    bool meshIsA = MeshIsA(w, c);
    const Fixture& fA = w.fixtures[c->fixtureA];
    const Fixture& fB = w.fixtures[c->fixtureB];
    Transform xfA = w.bodies[fA.body].xf, xfB = w.bodies[fB.body].xf;
    for (u32 index : c->triangles) {
        GJKProxy pA, pB;
        if (meshIsA) { MeshProxy(pA, fA.shape, index); pB.Set(fB.shape); }
        else { pA.Set(fA.shape); MeshProxy(pB, fB.shape, index); }
        GJKOutput q = GJK(xfA, pA, xfB, pB, true);
        const scalar kTol = scalar(10) * kEpsilon;
        if (q.distance <= kTol) return true;
    }
    return false;
}

// ------------------------------------------------------------------ cluster solver (cluster.cpp:202-584)
namespace {
struct Observation { Vec3 point; u32 cluster; u32 manifold; u32 manifoldPoint; };

struct ClusterSolver {
    std::vector<Observation> obs;
    std::vector<Vec3> clusters;  // centroids

    u32 FindCluster(const Vec3& p) const {
        u32 bestIndex = 0;
        scalar bestValue = kMaxScalar;
        for (u32 i = 0; i < clusters.size(); ++i) {
            Vec3 d = p - clusters[i];
            scalar metric = Dot(d, d);
            if (metric < bestValue) { bestValue = metric; bestIndex = i; }
        }
        return bestIndex;
    }
    void AddCluster(const Vec3& centroid) {
        if (clusters.empty()) { clusters.push_back(centroid); return; }
        u32 best = FindCluster(centroid);
        const scalar kTol = scalar(0.05);
        Vec3 d = centroid - clusters[best];
        if (Dot(d, d) <= kTol * kTol) {
            clusters[best] += centroid;
            clusters[best].Normalize();
            return;
        }
        clusters.push_back(centroid);
    }
    void InitializeClusters() {
        const u32 kMaxClusters = 3;
        u32 n = u32(obs.size());
        if (n <= kMaxClusters) {
            for (u32 i = 0; i < n; ++i) AddCluster(obs[i].point);
            return;
        }
        std::vector<bool> chosen(n, false);
        AddCluster(obs[0].point);
        chosen[0] = true;
        {
            Vec3 A = clusters[0];
            u32 index = 0;
            scalar max = -kMaxScalar;
            for (u32 i = 0; i < n; ++i) {
                if (chosen[i]) continue;
                Vec3 d = A - obs[i].point;
                scalar dd = Dot(d, d);
                if (dd > max) { max = dd; index = i; }
            }
            AddCluster(obs[index].point);
            chosen[index] = true;
        }
        if (clusters.size() == 1) {
            Vec3 A = clusters[0];
            u32 index = 0;
            scalar max = -kMaxScalar;
            for (u32 i = 0; i < n; ++i) {
                if (chosen[i]) continue;
                Vec3 d = A - obs[i].point;
                scalar dd = Dot(d, d);
                if (dd > max) { max = dd; index = i; }
            }
            AddCluster(obs[index].point);
            chosen[index] = true;
            return;
        }
        {
            Vec3 A = clusters[0], B = clusters[1];
            u32 index = 0;
            scalar max = -kMaxScalar;
            for (u32 i = 0; i < n; ++i) {
                if (chosen[i]) continue;
                Vec3 C = obs[i].point;
                Vec3 Q = ClosestPointOnSegment(C, A, B);
                Vec3 d = C - Q;
                scalar dd = Dot(d, d);
                if (dd > max) { max = dd; index = i; }
            }
            AddCluster(obs[index].point);
            chosen[index] = true;
        }
    }
    void Solve() {
        InitializeClusters();
        const u32 kMaxIters = 10;
        for (u32 iter = 0; iter < kMaxIters; ++iter) {
            for (Observation& o : obs) o.cluster = FindCluster(o.point);
            for (u32 i = 0; i < clusters.size(); ++i) {
                Vec3 centroid;
                centroid.SetZero();
                u32 pointCount = 0;
                for (const Observation& o : obs)
                    if (o.cluster == i) { centroid += o.point; ++pointCount; }
                if (pointCount > 0) {
                    centroid /= scalar(pointCount);
                    clusters[i] = centroid;
                }
            }
        }
        std::vector<Vec3> used;  // drop empty clusters, renumbering the observations
        for (u32 i = 0; i < clusters.size(); ++i) {
            u32 count = 0;
            for (const Observation& o : obs)
                if (o.cluster == i) ++count;
            if (count > 0) {
                u32 to = u32(used.size());
                for (Observation& o : obs)
                    if (o.cluster == i) o.cluster = to;
                used.push_back(clusters[i]);
            }
        }
        clusters = used;
    }
};
}  // namespace

static void WorldPoint(const ManifoldPoint* p, scalar rA, const Transform& xfA, scalar rB, const Transform& xfB,
                       Vec3& point, Vec3& normal, scalar& separation) {  // collision.cpp:63-75
    Vec3 nA = Mul(xfA.rotation, p->localNormal1);
    Vec3 cA = xfA * p->localPoint1;
    Vec3 cB = xfB * p->localPoint2;
    Vec3 pA = cA + rA * nA;
    Vec3 pB = cB - rB * nA;
    point = scalar(0.5) * (pA + pB);
    normal = nA;
    separation = Dot(cB - cA, nA) - rA - rB;
}

static void ClusterRun(Manifold* out, u32& numOut, const std::vector<Manifold>& in, const Transform& xfA,
                       scalar rA, const Transform& xfB, scalar rB) {  // cluster.cpp:471-584
    ClusterSolver cs;
    for (u32 i = 0; i < in.size(); ++i) {
        WorldManifold wm;
        wm.Initialize(&in[i], rA, xfA, rB, xfB);
        for (u32 j = 0; j < wm.pointCount; ++j) {
            Observation o;
            o.point = wm.points[j].normal;
            o.cluster = kNullU32;
            o.manifold = i;
            o.manifoldPoint = j;
            cs.obs.push_back(o);
        }
    }
    cs.Solve();
    for (u32 i = 0; i < cs.clusters.size(); ++i) {
        Vec3 center, normal;
        center.SetZero();
        normal.SetZero();
        ClusterPolygon polygonB;
        for (u32 j = 0; j < cs.obs.size(); ++j) {
            const Observation& o = cs.obs[j];
            if (o.cluster != i) continue;
            const ManifoldPoint* mp = in[o.manifold].points + o.manifoldPoint;
            Vec3 p, n;
            scalar sep;
            WorldPoint(mp, rA, xfA, rB, xfB, p, n, sep);
            ClusterVertex cv;
            cv.position = p;
            cv.clipIndex = j;
            polygonB.push_back(cv);
            center += p;
            normal += o.point;
        }
        if (polygonB.empty()) continue;
        center /= scalar(polygonB.size());
        normal.Normalize();
        Manifold& m = out[numOut];
        ++numOut;
        u32 minIndex = 0;
        scalar minSeparation = kMaxScalar;
        for (u32 j = 0; j < polygonB.size(); ++j) {
            const Observation& o = cs.obs[polygonB[j].clipIndex];
            const ManifoldPoint* ip = in[o.manifold].points + o.manifoldPoint;
            Vec3 p, n;
            scalar separation;
            WorldPoint(ip, rA, xfA, rB, xfB, p, n, separation);
            if (separation < minSeparation) { minIndex = j; minSeparation = separation; }
            polygonB[j].position = polygonB[j].position + separation * normal;
        }
        ClusterPolygon quadB;
        ReducePolygon(quadB, polygonB, normal, minIndex);
        for (u32 j = 0; j < quadB.size(); ++j) {
            const Observation& o = cs.obs[quadB[j].clipIndex];
            m.points[j] = in[o.manifold].points[o.manifoldPoint];
        }
        m.pointCount = u32(quadB.size());
    }
}

void MeshContactCollide(World& w, Contact* c) {  // mesh_contact.cpp:328-371
    bool meshIsA = MeshIsA(w, c);
    const Fixture& fA = w.fixtures[c->fixtureA];
    const Fixture& fB = w.fixtures[c->fixtureB];
    Transform xfA = w.bodies[fA.body].xf, xfB = w.bodies[fB.body].xf;
    std::vector<Manifold> manifolds(c->triangles.size());
    for (u32 i = 0; i < c->triangles.size(); ++i) {
        Manifold& m = manifolds[i];
        m.Initialize();
        Shape tri;
        if (meshIsA) {
            ChildTriangle(tri, fA.shape, c->triangles[i]);
            if (fB.shape.type == e_sphere) CollideTriangleAndSphere(m, xfA, tri, xfB, fB.shape);
            else if (fB.shape.type == e_capsule) CollideTriangleAndCapsule(m, xfA, tri, xfB, fB.shape);
        } else {  // hull-mesh (hull_mesh_contact.cpp)
            ChildTriangle(tri, fB.shape, c->triangles[i]);
            CollideHullAndTriangle(m, xfA, fA.shape, xfB, tri);
        }
        for (u32 j = 0; j < m.pointCount; ++j) m.points[j].id.triangleKey = c->triangles[i];
    }
    ClusterRun(c->manifolds, c->manifoldCount, manifolds, xfA, fA.shape.radius, xfB, fB.shape.radius);
}

}  // namespace orc

extern "C" {

// b3StaticTree::QueryAABB of mesh `d` (built as b3Mesh::BuildTree builds it) for n boxes
// (aabbs[6 k .. 6 k + 5]: lower xyz, upper xyz): counts[k] triangles, the first cap_per of them in
// out[k * cap_per ..]. Returns 0, or -1 for an invalid mesh.
int orc_mesh_query(const bx_mesh_desc* d, const float* aabbs, uint32_t n, uint32_t* counts, uint32_t* out,
                   uint32_t cap_per) {
    orc::Mesh m;
    if (!orc::BuildMesh(m, *d)) return -1;
    std::vector<orc::u32> tris;
    for (uint32_t k = 0; k < n; ++k) {
        orc::AABB box;
        box.lowerBound = orc::Vec3(aabbs[6 * k], aabbs[6 * k + 1], aabbs[6 * k + 2]);
        box.upperBound = orc::Vec3(aabbs[6 * k + 3], aabbs[6 * k + 4], aabbs[6 * k + 5]);
        tris.clear();
        orc::QueryMeshTree(m, box, tris);
        counts[k] = uint32_t(tris.size());
        for (uint32_t i = 0; i < tris.size() && i < cap_per; ++i) out[size_t(k) * cap_per + i] = tris[i];
    }
    return 0;
}

// The wing vertices (u1, u2, u3) of every triangle after b3Mesh::BuildAdjacency.
int orc_mesh_wings(const bx_mesh_desc* d, uint32_t* out) {
    orc::Mesh m;
    if (!orc::BuildMesh(m, *d)) return -1;
    for (size_t i = 0; i < m.triangles.size(); ++i) {
        out[3 * i] = m.triangles[i].u1;
        out[3 * i + 1] = m.triangles[i].u2;
        out[3 * i + 2] = m.triangles[i].u3;
    }
    return 0;
}

}  // extern "C"
