// bx_triangle.h -- triangle shapes in the device narrowphase (included by bx_step.hip).
//
// Same binary32 expression trees as the reference:
//   triangle_hull.h:36-70              b3TriangleHull, the triangle as a two-face hull
//   geometry.h:27-70                   b3BarycentricCoordinates (segment and triangle forms)
//   collide_triangle_sphere.cpp:24-296 Voronoi regions of the triangle
//   collide_triangle_capsule.cpp:107-278 (triangles without edge wings: the hull-capsule sequence on
//                                      the triangle hull)
//   collide_hull_triangle.cpp:24-124 and collide_hulls.cpp:189-264 (b3CollideHulls without the
//                                      feature cache on the triangle hull)
// Triangle fixtures have no edge wings (b3TriangleShape::Set); the child triangles of a mesh
// (b3MeshShape::GetChildTriangle, mesh_shape.cpp:142-175) carry them, and the functions below take
// them as an optional TriWings.
#pragma once

namespace bx {

// The two-face hull of a triangle in private memory, and a HullView of it.
struct TriangleHull {
    Vec3 v[3];
    HalfEdge e[6];
    u32 f[2];
    Plane p[2];
    Vec3 centroid;

    BX_HD void Set(const Vec3& A, const Vec3& B, const Vec3& C) {
        centroid = (A + B + C) / scalar(3);
        v[0] = A;
        v[1] = B;
        v[2] = C;
        // origin, twin, face, prev, next (each edge followed by its twin)
        e[0] = HalfEdge{0, 1, 0, 4, 2};
        e[1] = HalfEdge{1, 0, 1, 3, 5};
        e[2] = HalfEdge{1, 3, 0, 0, 4};
        e[3] = HalfEdge{2, 2, 1, 5, 1};
        e[4] = HalfEdge{2, 5, 0, 2, 0};
        e[5] = HalfEdge{0, 4, 1, 1, 3};
        f[0] = 0;
        f[1] = 1;
        Vec3 N = Normalize(Cross(B - A, C - A));
        p[0] = Plane(N, Dot(N, centroid));
        Vec3 M = -N;
        p[1] = Plane(M, Dot(M, centroid));
    }
    BX_HD HullView View() const {
        HullView h;
        h.vertices = v;
        h.edges = e;
        h.faces = f;
        h.planes = p;
        h.centroid = centroid;
        h.vertexCount = 3;
        h.edgeCount = 6;
        h.faceCount = 2;
        return h;
    }
};

// One point at P on the triangle towards the sphere centre Q.
BX_HD void TriangleSpherePoint(Manifold& m, const Vec3& P, const Vec3& Q, const Vec3& center2, scalar radius) {
    Vec3 d = Q - P;
    scalar dd = Dot(d, d);
    if (dd > radius * radius) return;
    Vec3 n(scalar(0), scalar(1), scalar(0));
    scalar len = Length(d);
    if (len > kEpsilon) n = d / len;
    OnePoint(m, n, P, center2);
}

// Is Q inside the triangle (A1, B1, C1) adjacent across an edge?
BX_HD bool InAdjacentFace(const Vec3& A1, const Vec3& B1, const Vec3& C1, const Vec3& Q) {
    scalar w[4];
    BaryGeom3(w, A1, B1, C1, Q);
    return w[0] > scalar(0) && w[1] > scalar(0) && w[2] > scalar(0);
}

// collide_triangle_sphere.cpp:24-296. On an edge with a wing, a sphere centre over the adjacent
// triangle makes no point (that triangle's face region has it).
BX_HD void CollideTriangleAndSphere(Manifold& m, const Transform& xf1, const Fixture& t1, const Transform& xf2,
                                    const Fixture& s2, const TriWings* wings = nullptr) {
    Vec3 Q = MulT(xf1, Mul(xf2, s2.v1));
    Vec3 A = t1.v1, B = t1.v2, C = t1.v3;
    scalar radius = t1.radius + s2.radius;
    scalar wAB[3], wBC[3], wCA[3];
    BaryGeom2(wAB, A, B, Q);
    BaryGeom2(wBC, B, C, Q);
    BaryGeom2(wCA, C, A, Q);
    if (wAB[1] <= scalar(0) && wCA[0] <= scalar(0)) { TriangleSpherePoint(m, A, Q, s2.v1, radius); return; }
    if (wAB[0] <= scalar(0) && wBC[1] <= scalar(0)) { TriangleSpherePoint(m, B, Q, s2.v1, radius); return; }
    if (wBC[0] <= scalar(0) && wCA[1] <= scalar(0)) { TriangleSpherePoint(m, C, Q, s2.v1, radius); return; }
    scalar wABC[4];
    BaryGeom3(wABC, A, B, C, Q);
    if (wAB[0] > scalar(0) && wAB[1] > scalar(0) && wABC[3] * wABC[2] <= scalar(0)) {  // edge AB
        scalar den = scalar(1) / wAB[2];
        if (wings && wings->Has(0u) && InAdjacentFace(wings->e1, B, A, Q)) return;
        TriangleSpherePoint(m, den * (wAB[0] * A + wAB[1] * B), Q, s2.v1, radius);
        return;
    }
    if (wBC[0] > scalar(0) && wBC[1] > scalar(0) && wABC[3] * wABC[0] <= scalar(0)) {  // edge BC
        scalar den = scalar(1) / wBC[2];
        if (wings && wings->Has(1u) && InAdjacentFace(wings->e2, C, B, Q)) return;
        TriangleSpherePoint(m, den * (wBC[0] * B + wBC[1] * C), Q, s2.v1, radius);
        return;
    }
    if (wCA[0] > scalar(0) && wCA[1] > scalar(0) && wABC[3] * wABC[1] <= scalar(0)) {  // edge CA
        scalar den = scalar(1) / wCA[2];
        if (wings && wings->Has(2u) && InAdjacentFace(wings->e3, A, C, Q)) return;
        TriangleSpherePoint(m, den * (wCA[0] * C + wCA[1] * A), Q, s2.v1, radius);
        return;
    }
    if (wABC[3] == scalar(0)) return;  // face region
    scalar den = scalar(1) / wABC[3];
    TriangleSpherePoint(m, den * (wABC[0] * A + wABC[1] * B + wABC[2] * C), Q, s2.v1, radius);
}

BX_HD void CollideTriangleAndCapsule(Manifold& m, const Transform& xf1, const Fixture& t1, const Transform& xf2,
                                     const Fixture& s2, const TriWings* wings = nullptr) {
    TriangleHull th;
    th.Set(t1.v1, t1.v2, t1.v3);
    CollideHullAndCapsule(m, xf1, th.View(), t1, xf2, s2, wings);
}

// b3CollideHulls without the feature cache (collide_hulls.cpp:189-264).
BX_HD void CollideHullsUncached(Manifold& m, const Transform& xf1, const HullView& h1, scalar r1, const Transform& xf2,
                                const HullView& h2, scalar r2, FaceScratch& s, u32& overflow) {
    scalar totalRadius = r1 + r2;
    FaceQuery fq1 = QueryFaceSeparation(xf1, h1, xf2, h2);
    if (fq1.separation > totalRadius) return;
    FaceQuery fq2 = QueryFaceSeparation(xf2, h2, xf1, h1);
    if (fq2.separation > totalRadius) return;
    EdgeQuery eq = QueryEdgeSeparation(xf1, h1, xf2, h2);
    if (eq.separation > totalRadius) return;
    const scalar kRelEdgeTol = scalar(0.90);
    const scalar kRelFaceTol = scalar(0.98);
    const scalar kAbsTol = scalar(0.5) * kLinearSlop;
    if (eq.separation > kRelEdgeTol * Max(fq1.separation, fq2.separation) + kAbsTol) {
        BuildEdgeContact(m, xf1, eq.index1, h1, xf2, eq.index2, h2);
    } else if (fq2.separation > kRelFaceTol * fq1.separation + kAbsTol) {
        BuildFaceContact(m, xf2, fq2.index, h2, r2, xf1, h1, r1, true, s, overflow);
    } else {
        BuildFaceContact(m, xf1, fq1.index, h1, r1, xf2, h2, r2, false, s, overflow);
    }
    if (m.pointCount > 0) return;  // the heuristic succeeded
    if (eq.separation > Max(fq1.separation, fq2.separation)) {
        BuildEdgeContact(m, xf1, eq.index1, h1, xf2, eq.index2, h2);
    } else if (fq2.separation > fq1.separation) {
        BuildFaceContact(m, xf2, fq2.index, h2, r2, xf1, h1, r1, true, s, overflow);
    } else {
        BuildFaceContact(m, xf1, fq1.index, h1, r1, xf2, h2, r2, false, s, overflow);
    }
}

// collide_hull_triangle.cpp:24-124: with wings, the normal of a point on an edge shared with a
// coplanar wing triangle is snapped to the triangle plane.
BX_HD void CollideHullAndTriangle(Manifold& m, const Transform& xf1, const HullView& h1, scalar r1,
                                  const Transform& xf2, const Fixture& t2, FaceScratch& s, u32& overflow,
                                  const TriWings* wings = nullptr) {
    TriangleHull th;
    th.Set(t2.v1, t2.v2, t2.v3);
    CollideHullsUncached(m, xf1, h1, r1, xf2, th.View(), t2.radius, s, overflow);
    if (!wings || !wings->mask) return;
    Plane plane2 = Mul(xf2, th.p[0]);
    Vec3 centroid1 = Mul(xf1, h1.centroid);
    Vec3 centroid2 = Mul(xf2, th.centroid);
    for (u32 i = 0; i < m.pointCount; ++i) {
        ManifoldPoint& mp = m.points[i];
        u32 e2;
        if (mp.edgeContact) {
            e2 = mp.key[2];
        } else {
            e2 = mp.key[2] == kNull ? mp.key[3] : mp.key[2];
        }
        if (e2 >= 6u) continue;  // no triangle edge (B3_NULL_EDGE)
        u32 edgeIndex = e2 >> 1;  // half-edges 2k, 2k + 1 lie on edge k
        if (!wings->Has(edgeIndex)) continue;
        u32 ev1 = edgeIndex;
        u32 ev2 = edgeIndex + 1u < 3u ? edgeIndex + 1u : 0u;
        Vec3 A = wings->Get(edgeIndex);
        Vec3 B = ev2 == 0u ? t2.v1 : (ev2 == 1u ? t2.v2 : t2.v3);
        Vec3 C = ev1 == 0u ? t2.v1 : (ev1 == 1u ? t2.v2 : t2.v3);
        Vec3 center = (A + B + C) / scalar(3);
        scalar distance = Distance(center, th.p[0]);
        const scalar kCoplanarTol = scalar(0.005);
        if (distance > -kCoplanarTol && distance < kCoplanarTol) {
            Vec3 n1 = Mul(xf1.rotation, mp.localNormal1);
            Vec3 c1 = Mul(xf1, mp.localPoint1);
            Vec3 c2 = Mul(xf2, mp.localPoint2);
            scalar sep = Dot(c2 - c1, n1);
            Vec3 n = plane2.normal;
            if (Dot(n, centroid2 - centroid1) < scalar(0)) n = -n;
            Vec3 nc1 = c2 - sep * n;
            mp.localNormal1 = MulC(xf1.rotation, n);
            mp.localPoint1 = MulT(xf1, nc1);
        }
    }
}

}  // namespace bx
