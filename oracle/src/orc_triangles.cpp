// orc_triangles.cpp -- oracle (test infrastructure): triangle shapes and their narrowphase.
//
// Restates, in the reference's binary32 expression order:
//   triangle_shape.cpp:25-73        b3TriangleShape (skin radius, Set, mass)
//   triangle_hull.h:24-72           b3TriangleHull: the triangle as a two-face hull
//   geometry.h:27-94                b3BarycentricCoordinates (segment and triangle forms)
//   collide_triangle_sphere.cpp:24-296
//   collide_triangle_capsule.cpp:107-278 (the hull-capsule sequence on the triangle hull, with the
//                                   edge-wing step in CollideConvexAndCapsule)
//   collide_hull_triangle.cpp:24-124
#include <cstdio>
#include <cstdlib>
#include "orc_world.h"

namespace orc {

void CollideHullsUncached(Manifold& m, const Transform& xf1, const Hull& h1, scalar r1, const Transform& xf2,
                          const Hull& h2, scalar r2);
void CollideConvexAndCapsule(Manifold& m, const Transform& xf1, const Shape& s1, const Hull& h1, const Transform& xf2,
                             const Shape& s2);

void InitTriangleShape(Shape& s, const Vec3& a, const Vec3& b, const Vec3& c) {  // triangle_shape.cpp:25-45
    s.type = e_triangle;
    s.radius = kHullRadius;
    s.v1 = a;
    s.v2 = b;
    s.v3 = c;
    s.hasE1 = s.hasE2 = s.hasE3 = false;
}

void MakeTriangleHull(Hull& h, const Vec3& A, const Vec3& B, const Vec3& C) {  // triangle_hull.h:36-70
    h.centroid = (A + B + C) / scalar(3);
    h.vertices = {A, B, C};
    // origin, twin, face, prev, next; twins adjacent
    h.edges = {{0, 1, 0, 4, 2}, {1, 0, 1, 3, 5}, {1, 3, 0, 0, 4}, {2, 2, 1, 5, 1}, {2, 5, 0, 2, 0}, {0, 4, 1, 1, 3}};
    h.faces = {0, 1};
    Vec3 N = Normalize(Cross(B - A, C - A));
    h.planes = {Plane(N, Dot(N, h.centroid)), Plane(-N, Dot(-N, h.centroid))};
}

// geometry.h:27-43: segment AB, the divisor is |AB|^2
static inline void Barycentric(scalar out[3], const Vec3& A, const Vec3& B, const Vec3& Q) {
    Vec3 AB = B - A;
    Vec3 QA = A - Q;
    Vec3 QB = B - Q;
    scalar divisor = Dot(AB, AB);
    out[0] = Dot(QB, AB);
    out[1] = -Dot(QA, AB);
    out[2] = divisor;
}

// geometry.h:45-70: triangle ABC, the divisor is |AB x AC|^2
static inline void Barycentric(scalar out[4], const Vec3& A, const Vec3& B, const Vec3& C, const Vec3& Q) {
    Vec3 AB = B - A;
    Vec3 AC = C - A;
    Vec3 QA = A - Q;
    Vec3 QB = B - Q;
    Vec3 QC = C - Q;
    Vec3 QB_x_QC = Cross(QB, QC);
    Vec3 QC_x_QA = Cross(QC, QA);
    Vec3 QA_x_QB = Cross(QA, QB);
    Vec3 AB_x_AC = Cross(AB, AC);
    scalar divisor = Dot(AB_x_AC, AB_x_AC);
    out[0] = Dot(QB_x_QC, AB_x_AC);
    out[1] = Dot(QC_x_QA, AB_x_AC);
    out[2] = Dot(QA_x_QB, AB_x_AC);
    out[3] = divisor;
}

// One point at P on the triangle towards the sphere centre Q (normal (0,1,0) when they coincide).
static void SpherePoint(Manifold& m, const Vec3& P, const Vec3& Q, const Vec3& center2, scalar radius) {
    Vec3 d = Q - P;
    scalar dd = Dot(d, d);
    if (dd > radius * radius) return;
    Vec3 n(0, 1, 0);
    scalar len = Length(d);
    if (len > kEpsilon) n = d / len;
    m.pointCount = 1;
    m.points[0].localNormal1 = n;
    m.points[0].localPoint1 = P;
    m.points[0].localPoint2 = center2;
    m.points[0].id = MakeID(0, 0);
}

// Is Q inside the triangle (A1, B1, C1) adjacent across an edge?
static bool InAdjacentFace(const Vec3& A1, const Vec3& B1, const Vec3& C1, const Vec3& Q) {
    scalar w[4];
    Barycentric(w, A1, B1, C1, Q);
    return w[0] > scalar(0) && w[1] > scalar(0) && w[2] > scalar(0);
}

void CollideTriangleAndSphere(Manifold& m, const Transform& xf1, const Shape& t1, const Transform& xf2,
                              const Shape& s2) {  // collide_triangle_sphere.cpp:24-296
    Vec3 Q = MulT(xf1, Mul(xf2, s2.v1));
    Vec3 A = t1.v1, B = t1.v2, C = t1.v3;
    scalar radius = t1.radius + s2.radius;
    scalar wAB[3], wBC[3], wCA[3];
    Barycentric(wAB, A, B, Q);
    Barycentric(wBC, B, C, Q);
    Barycentric(wCA, C, A, Q);
    if (wAB[1] <= scalar(0) && wCA[0] <= scalar(0)) { SpherePoint(m, A, Q, s2.v1, radius); return; }
    if (wAB[0] <= scalar(0) && wBC[1] <= scalar(0)) { SpherePoint(m, B, Q, s2.v1, radius); return; }
    if (wBC[0] <= scalar(0) && wCA[1] <= scalar(0)) { SpherePoint(m, C, Q, s2.v1, radius); return; }
    scalar wABC[4];
    Barycentric(wABC, A, B, C, Q);
    if (wAB[0] > scalar(0) && wAB[1] > scalar(0) && wABC[3] * wABC[2] <= scalar(0)) {  // edge AB
        scalar den = scalar(1) / wAB[2];
        Vec3 P = den * (wAB[0] * A + wAB[1] * B);
        Vec3 d = Q - P;
        if (Dot(d, d) > radius * radius) return;
        if (t1.hasE1 && InAdjacentFace(t1.e1, B, A, Q)) return;
        SpherePoint(m, P, Q, s2.v1, radius);
        return;
    }
    if (wBC[0] > scalar(0) && wBC[1] > scalar(0) && wABC[3] * wABC[0] <= scalar(0)) {  // edge BC
        scalar den = scalar(1) / wBC[2];
        Vec3 P = den * (wBC[0] * B + wBC[1] * C);
        Vec3 d = Q - P;
        if (Dot(d, d) > radius * radius) return;
        if (t1.hasE2 && InAdjacentFace(t1.e2, C, B, Q)) return;
        SpherePoint(m, P, Q, s2.v1, radius);
        return;
    }
    if (wCA[0] > scalar(0) && wCA[1] > scalar(0) && wABC[3] * wABC[1] <= scalar(0)) {  // edge CA
        scalar den = scalar(1) / wCA[2];
        Vec3 P = den * (wCA[0] * C + wCA[1] * A);
        Vec3 d = Q - P;
        if (Dot(d, d) > radius * radius) return;
        if (t1.hasE3 && InAdjacentFace(t1.e3, A, C, Q)) return;
        SpherePoint(m, P, Q, s2.v1, radius);
        return;
    }
    if (wABC[3] == scalar(0)) return;  // face region
    scalar den = scalar(1) / wABC[3];
    Vec3 P = den * (wABC[0] * A + wABC[1] * B + wABC[2] * C);
    SpherePoint(m, P, Q, s2.v1, radius);
}

void CollideTriangleAndCapsule(Manifold& m, const Transform& xf1, const Shape& t1, const Transform& xf2,
                               const Shape& s2) {  // collide_triangle_capsule.cpp:107-278
    Hull h;
    MakeTriangleHull(h, t1.v1, t1.v2, t1.v3);
    CollideConvexAndCapsule(m, xf1, t1, h, xf2, s2);
}

void CollideHullAndTriangle(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2,
                            const Shape& t2) {  // collide_hull_triangle.cpp:24-124
    const Hull& h1 = *s1.hull;
    Hull h2;
    MakeTriangleHull(h2, t2.v1, t2.v2, t2.v3);
    CollideHullsUncached(m, xf1, h1, s1.radius, xf2, h2, t2.radius);
    // normals of points on an edge shared with a coplanar wing triangle snap to the triangle plane
    const Vec3 vertices[3] = {t2.v1, t2.v2, t2.v3};
    const bool hasWing[3] = {t2.hasE1, t2.hasE2, t2.hasE3};
    const Vec3 wings[3] = {t2.e1, t2.e2, t2.e3};
    static const u32 kHalfEdgeEdge[6] = {0, 0, 1, 1, 2, 2};
    Plane plane2 = Mul(xf2, h2.planes[0]);
    Vec3 centroid1 = xf1 * h1.centroid;
    Vec3 centroid2 = xf2 * h2.centroid;
    for (u32 i = 0; i < m.pointCount; ++i) {
        ManifoldPoint& mp = m.points[i];
        FeaturePair pair = mp.id.pair;
        u32 e2;
        if (mp.edgeContact) {
            e2 = pair.inEdge2;
        } else {
            e2 = pair.inEdge2 == kNullU32 ? pair.outEdge2 : pair.inEdge2;
            if (e2 == kNullU32) continue;  // B3_NULL_EDGE
        }
        u32 edgeIndex = kHalfEdgeEdge[e2];
        if (!hasWing[edgeIndex]) continue;
        u32 ev1 = edgeIndex;
        u32 ev2 = edgeIndex + 1 < 3 ? edgeIndex + 1 : 0;
        Vec3 A = wings[edgeIndex], B = vertices[ev2], C = vertices[ev1];
        Vec3 center = (A + B + C) / scalar(3);
        scalar distance = Distance(center, h2.planes[0]);
        const scalar kCoplanarTol = scalar(0.005);
        if (distance > -kCoplanarTol && distance < kCoplanarTol) {
            Vec3 n1 = Mul(xf1.rotation, mp.localNormal1);
            Vec3 c1 = xf1 * mp.localPoint1;
            Vec3 c2 = xf2 * mp.localPoint2;
            scalar s = Dot(c2 - c1, n1);
            Vec3 n = plane2.normal;
            if (Dot(n, centroid2 - centroid1) < scalar(0)) n = -n;
            Vec3 nc1 = c2 - s * n;
            mp.localNormal1 = MulC(xf1.rotation, n);
            mp.localPoint1 = MulT(xf1, nc1);
        }
    }
}

}  // namespace orc
