// orc_geom.h -- oracle (test infrastructure): shapes, hulls, manifolds and narrowphase entry
// points, restating include/bounce/collision/{collision,clip,cluster}.h,
// include/bounce/collision/geometry/{hull,capsule,mesh}.h and include/bounce/collision/sat/sat.h.
#pragma once
#include <vector>
#include "orc_math.h"

namespace orc {

struct HalfEdge { u32 origin, twin, face, prev, next; };

// b3Hull (hull.h:37-152)
struct Hull {
    Vec3 centroid;
    std::vector<Vec3> vertices;
    std::vector<HalfEdge> edges;
    std::vector<u32> faces;      // first half-edge of each face
    std::vector<Plane> planes;
    u32 vertexCount() const { return u32(vertices.size()); }
    u32 edgeCount() const { return u32(edges.size()); }
    u32 faceCount() const { return u32(faces.size()); }
    u32 GetSupportVertex(const Vec3& d) const {  // first index of the maximum (hull.h:103-118)
        u32 maxIndex = 0;
        scalar maxProjection = Dot(d, vertices[0]);
        for (u32 i = 1; i < vertexCount(); ++i) {
            scalar p = Dot(d, vertices[i]);
            if (p > maxProjection) { maxIndex = i; maxProjection = p; }
        }
        return maxIndex;
    }
    u32 GetSupportFace(const Vec3& d) const {  // hull.h:120-135
        u32 maxIndex = 0;
        scalar maxProjection = Dot(d, planes[0].normal);
        for (u32 i = 1; i < faceCount(); ++i) {
            scalar p = Dot(d, planes[i].normal);
            if (p > maxProjection) { maxIndex = i; maxProjection = p; }
        }
        return maxIndex;
    }
    Plane GetEdgeSidePlane(u32 index) const {  // hull.h:137-150
        const HalfEdge& e = edges[index];
        const HalfEdge& t = edges[e.twin];
        Vec3 P = vertices[e.origin];
        Vec3 Q = vertices[t.origin];
        Vec3 E = Q - P;
        Vec3 D = Cross(E, planes[e.face].normal);
        Plane p;
        p.normal = Normalize(D);
        p.offset = Dot(p.normal, P);
        return p;
    }
};

// b3MeshTriangle + b3Mesh (mesh.h)
struct MeshTriangle { u32 u1, u2, u3; u32 v1, v2, v3; };

struct StaticNode { AABB aabb; u32 child1, child2, index; };

struct Mesh {
    std::vector<Vec3> vertices;
    std::vector<MeshTriangle> triangles;
    std::vector<StaticNode> nodes;   // b3StaticTree node array (root = 0)
    u32 root = kNullU32;
};

// b3StaticTree::QueryAABB (static_tree.h:215-255): the triangles whose leaf box overlaps `aabb`, in stack order
void QueryMeshTree(const Mesh& m, const AABB& aabb, std::vector<u32>& out);

enum ShapeType : u32 { e_sphere = 0, e_capsule = 1, e_triangle = 2, e_hull = 3, e_mesh = 4 };

struct Shape {
    u32 type;
    scalar radius;
    Vec3 v1, v2, v3;                 // sphere centre, capsule and triangle vertices
    bool hasE1, hasE2, hasE3;        // triangle wing vertices (triangle_shape.h)
    Vec3 e1, e2, e3;
    const Hull* hull = nullptr;
    const Mesh* mesh = nullptr;
    u32 hullIndex = 0, meshIndex = 0;  // library indices of hull and mesh
    Vec3 scale;                      // mesh scale
};

// b3FeaturePair / b3ContactID (clip.h:27-43, collision.h:52-81)
struct FeaturePair { u32 inEdge1, outEdge1, inEdge2, outEdge2; };
inline FeaturePair MakePair(u32 a, u32 b, u32 c, u32 d) { FeaturePair p; p.inEdge1 = a; p.outEdge1 = b; p.inEdge2 = c; p.outEdge2 = d; return p; }
struct ContactID {
    u32 triangleKey;
    FeaturePair pair;
    bool operator==(const ContactID& o) const {
        return triangleKey == o.triangleKey && pair.inEdge1 == o.pair.inEdge1 && pair.outEdge1 == o.pair.outEdge1 &&
               pair.inEdge2 == o.pair.inEdge2 && pair.outEdge2 == o.pair.outEdge2;
    }
};
inline ContactID MakeID(const FeaturePair& p) { ContactID id; id.triangleKey = kNullU32; id.pair = p; return id; }
inline ContactID MakeID(u32 k1, u32 k2) {  // b3MakeID(uint64 key1, uint64 key2) on little-endian
    ContactID id; id.triangleKey = kNullU32; id.pair = MakePair(k1, 0, k2, 0); return id;
}

// b3ClipVertex / b3ClipPlane (clip.h:47-61)
struct ClipVertex { Vec3 position; FeaturePair pair; };
struct ClipPlane { Plane plane; u32 edge; };

struct ManifoldPoint {
    Vec3 localNormal1, localPoint1, localPoint2;
    scalar normalImpulse;
    ContactID id;
    bool persisting, edgeContact;
};

struct Manifold {
    ManifoldPoint points[kMaxManifoldPoints];
    u32 pointCount;
    Vec2 tangentImpulse;
    scalar motorImpulse, motorSpeed, tangentSpeed1, tangentSpeed2;
    void Initialize() {  // collision.cpp:23-38
        pointCount = 0;
        tangentImpulse.SetZero();
        motorImpulse = 0; motorSpeed = 0; tangentSpeed1 = 0; tangentSpeed2 = 0;
        for (u32 i = 0; i < kMaxManifoldPoints; ++i) {
            points[i].normalImpulse = 0;
            points[i].persisting = false;
            points[i].edgeContact = false;
        }
    }
    void Initialize(const Manifold& old) {  // collision.cpp:40-61
        tangentImpulse = old.tangentImpulse;
        motorImpulse = old.motorImpulse;
        for (u32 i = 0; i < pointCount; ++i) {
            ManifoldPoint* p1 = points + i;
            for (u32 j = 0; j < old.pointCount; ++j) {
                const ManifoldPoint* p2 = old.points + j;
                if (p2->id == p1->id) { p1->normalImpulse = p2->normalImpulse; p1->persisting = true; break; }
            }
        }
    }
};

struct WorldManifoldPoint { Vec3 point, normal; scalar separation; };
struct WorldManifold {
    u32 pointCount;
    WorldManifoldPoint points[kMaxManifoldPoints];
    Vec3 center, normal, tangent1, tangent2;
    void Initialize(const Manifold* m, scalar rA, const Transform& xfA, scalar rB, const Transform& xfB);
};

// b3FeatureCache (sat.h:67-96)
enum CacheType : u32 { e_empty = 0, e_separation = 1, e_overlap = 2 };
enum FeatureType : u32 { e_edges = 0, e_face1 = 1, e_face2 = 2 };
struct FeatureCache {
    u32 cacheType = e_empty, featureType = 0, index1 = 0, index2 = 0;
};

// Shape operations (shapes/*_shape.cpp)
void ComputeAABB(const Shape& s, const Transform& xf, AABB* out);
struct MassData { Vec3 center; scalar mass; Mat33 I; };
void ComputeMass(const Shape& s, scalar density, MassData* out);
void SphereMass(const Vec3& center, scalar radius, scalar density, MassData* out);
void CapsuleMass(const Vec3& A, const Vec3& B, scalar radius, scalar density, MassData* out);

// GJK (gjk.cpp:37-925, gjk.h): closest points between two convex vertex clouds with radii
struct GJKProxy {
    const Vec3* vertices = nullptr;
    u32 count = 0;
    scalar radius = 0;
    Vec3 buffer[3];
    void Set(const Shape& s);
    u32 GetSupportIndex(const Vec3& d) const;
};
// simplexCount / simplexIndex1: the final simplex (the cache b3GJK writes, gjk.cpp:781, 868-876)
struct GJKOutput { Vec3 point1, point2; scalar distance; u32 iterations; u32 simplexCount; u32 simplexIndex1[4]; u32 simplexIndex2[4]; };
GJKOutput GJK(const Transform& xf1, const GJKProxy& p1, const Transform& xf2, const GJKProxy& p2, bool applyRadius);
// b3SimplexCache (gjk.h): the final simplex of one call, read back by the next
struct SimplexCache { scalar metric; u32 count; u32 index1[4]; u32 index2[4]; };
GJKOutput GJK(const Transform& xf1, const GJKProxy& p1, const Transform& xf2, const GJKProxy& p2, bool applyRadius,
              SimplexCache* cache);

// Time of impact (time_of_impact.cpp). States are b3TOIOutput::State's (time_of_impact.h:37-51).
enum TOIState : u32 { e_toiUnknown = 0, e_toiFailed = 1, e_toiOverlapped = 2, e_toiTouching = 3, e_toiSeparated = 4 };
struct TOIOutput { u32 state; scalar t; u32 iterations; };
TOIOutput TimeOfImpact(const GJKProxy& proxyA, const Sweep& sweepA, const GJKProxy& proxyB, const Sweep& sweepB,
                       scalar tMax);
TOIOutput TimeOfImpact(const Transform& xf1, const GJKProxy& proxy1, const Vec3& d1, const Transform& xf2,
                       const GJKProxy& proxy2, const Vec3& d2, u32 maxIterations);
TOIOutput TimeOfImpact(const AABB& aabb1, const Vec3& d1, const AABB& aabb2, const Vec3& d2);

// Narrowphase (collide/*.cpp)
void CollideHulls(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2, const Shape& s2,
                  FeatureCache& cache, const Transform& xf01, const Transform& xf02);
void CollideSpheres(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2, const Shape& s2);
void CollideCapsuleAndSphere(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2, const Shape& s2);
void CollideCapsules(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2, const Shape& s2);
void CollideHullAndSphere(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2, const Shape& s2);
void CollideHullAndCapsule(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2, const Shape& s2);
bool ConvexTestOverlap(const Transform& xfA, const Shape& sA, const Transform& xfB, const Shape& sB);
// b3ClusterVertex and b3ReducePolygon (cluster.h, cluster.cpp:23-200)
struct ClusterVertex { Vec3 position; u32 clipIndex; };
typedef std::vector<ClusterVertex> ClusterPolygon;
void ReducePolygon(ClusterPolygon& out, const ClusterPolygon& in, const Vec3& N, u32 initialPoint);
void CollideTriangleAndSphere(Manifold& m, const Transform& xf1, const Shape& t1, const Transform& xf2, const Shape& s2);
void CollideTriangleAndCapsule(Manifold& m, const Transform& xf1, const Shape& t1, const Transform& xf2, const Shape& s2);
void CollideHullAndTriangle(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2, const Shape& t2);

void ClosestPointsOnSegments(Vec3& C1, Vec3& C2, const Vec3& P1, const Vec3& Q1, const Vec3& P2, const Vec3& Q2);
Vec3 ClosestPointOnSegment(const Vec3& Q, const Vec3& A, const Vec3& B);

}  // namespace orc
