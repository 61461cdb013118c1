// bx_convex.h -- convex narrowphase other than hull-hull, and GJK, for the device step.
//
// Same binary32 expression trees and tie-breaking as the reference:
//   collide/collide_spheres.cpp:22-52, collide/collide_capsule_sphere.cpp:23-120,
//   collide/collide_capsules.cpp:25-123, collide/collide_hull_sphere.cpp:26-97,
//   collide/collide_hull_capsule.cpp:28-181, sat/sat_hull_edge.cpp:22-146,
//   clip.cpp:22-30,57-100,153-241, gjk/gjk.cpp:37-925, collision.cpp:100-112.
// Written for registers: the simplex, the proxies of spheres and capsules and the two-point clip
// segments are named scalars rather than indexed arrays, so nothing spills to scratch memory.
#pragma once
#include "bx_collide.h"

namespace bx {

// GJK proxy (gjk_proxy.h): hull vertices in memory, or up to three inline vertices.
struct ProxyD {
    const Vec3* g;
    Vec3 a, b, c;
    u32 count;
    scalar radius;
    BX_HD Vec3 Get(u32 i) const { return g ? g[i] : (i == 0u ? a : (i == 1u ? b : c)); }
    BX_HD u32 Support(const Vec3& d) const {  // gjk_proxy.h:72-86: first index of the maximum
        u32 maxIndex = 0;
        scalar maxProjection = Dot(d, Get(0));
        for (u32 i = 1; i < count; ++i) {
            scalar p = Dot(d, Get(i));
            if (p > maxProjection) { maxIndex = i; maxProjection = p; }
        }
        return maxIndex;
    }
};

BX_HD ProxyD MakeProxy(const Fixture& f, const HullView* h) {  // gjk.cpp:37-95
    ProxyD p;
    p.g = nullptr;
    p.a = f.v1;
    p.b = f.v2;
    p.c = f.v3;
    p.radius = f.radius;
    if (f.shapeType == 3u) {
        p.g = h->vertices;
        p.count = h->vertexCount;
    } else {  // sphere 1, capsule 2, triangle 3 vertices (gjk.cpp:37-95)
        p.count = f.shapeType == 1u ? 2u : (f.shapeType == 2u ? 3u : 1u);
    }
    return p;
}

struct SVert {
    Vec3 p1, p2, p;
    scalar w;
    u32 i1, i2;
};

struct GJKOut {
    Vec3 point1, point2;
    scalar distance;
    // the final simplex: vertex count and the proxy vertex indices of its first three vertices
    u32 simplexCount;
    u32 a1, b1, c1, a2, b2, c2;
};

// Edge wings of a mesh triangle (b3TriangleShape::m_hasE1..3 / m_e1..3, triangle_shape.h:60-70):
// bit i of mask set when edge i (v1v2, v2v3, v3v1) has a wing vertex e(i+1).
struct TriWings {
    Vec3 e1, e2, e3;
    u32 mask;
    BX_HD bool Has(u32 i) const { return (mask >> i) & 1u; }
    BX_HD Vec3 Get(u32 i) const { return i == 0u ? e1 : (i == 1u ? e2 : e3); }
};

namespace gjk {

BX_HD void Bary2(scalar out[3], const Vec3& A, const Vec3& B, const Vec3& Q) {  // gjk.cpp:104-118
    Vec3 AB = B - A;
    Vec3 QA = A - Q;
    Vec3 QB = B - Q;
    out[0] = Dot(QB, AB);
    out[1] = -Dot(QA, AB);
    out[2] = out[0] + out[1];
}

BX_HD void Bary3(scalar out[4], const Vec3& A, const Vec3& B, const Vec3& C, const Vec3& Q) {  // gjk.cpp:123-145
    Vec3 AB = B - A;
    Vec3 AC = C - A;
    Vec3 QA = A - Q;
    Vec3 QB = B - Q;
    Vec3 QC = C - Q;
    Vec3 QB_x_QC = Cross(QB, QC);
    Vec3 QC_x_QA = Cross(QC, QA);
    Vec3 QA_x_QB = Cross(QA, QB);
    Vec3 AB_x_AC = Cross(AB, AC);
    out[0] = Dot(QB_x_QC, AB_x_AC);
    out[1] = Dot(QC_x_QA, AB_x_AC);
    out[2] = Dot(QA_x_QB, AB_x_AC);
    out[3] = out[0] + out[1] + out[2];
}

BX_HD void Bary4(scalar out[5], const Vec3& A, const Vec3& B, const Vec3& C, const Vec3& D, const Vec3& Q) {
    // gjk.cpp:150-173
    Vec3 AB = B - A;
    Vec3 AC = C - A;
    Vec3 AD = D - A;
    Vec3 QA = A - Q;
    Vec3 QB = B - Q;
    Vec3 QC = C - Q;
    Vec3 QD = D - Q;
    scalar divisor = Det(AB, AC, AD);
    scalar sign = Sign(divisor);
    out[0] = sign * Det(QB, QC, QD);
    out[1] = sign * Det(QA, QD, QC);
    out[2] = sign * Det(QA, QB, QD);
    out[3] = sign * Det(QA, QC, QB);
    out[4] = sign * divisor;
}

}  // namespace gjk

struct SimplexD {
    SVert v0, v1, v2, v3;
    u32 count;

    BX_HD void Keep1(const SVert& A) {
        count = 1;
        v0 = A;
        v0.w = scalar(1);
    }
    BX_HD void Keep2(const SVert& A, const SVert& B, scalar wa, scalar wb, scalar divisor) {
        scalar s = scalar(1) / divisor;
        count = 2;
        v0 = A;
        v0.w = s * wa;
        v1 = B;
        v1.w = s * wb;
    }
    BX_HD void Keep3(const SVert& A, const SVert& B, const SVert& C, scalar wa, scalar wb, scalar wc,
                     scalar divisor) {
        scalar s = scalar(1) / divisor;
        count = 3;
        v0 = A;
        v0.w = s * wa;
        v1 = B;
        v1.w = s * wb;
        v2 = C;
        v2.w = s * wc;
    }

    BX_HD Vec3 SearchDirection(const Vec3& Q) const {  // gjk.cpp:175-216
        if (count == 1) return Q - v0.p;
        if (count == 2) {
            Vec3 AB = v1.p - v0.p;
            Vec3 AQ = Q - v0.p;
            return Cross(Cross(AB, AQ), AB);
        }
        Vec3 AB = v1.p - v0.p;
        Vec3 AC = v2.p - v0.p;
        Vec3 AQ = Q - v0.p;
        Vec3 AB_x_AC = Cross(AB, AC);
        scalar sign = Dot(AB_x_AC, AQ);
        if (sign > scalar(0)) return AB_x_AC;
        return -AB_x_AC;
    }

    BX_HD void ClosestPoints(Vec3* p1, Vec3* p2) const {  // gjk.cpp:238-268
        if (count == 1) {
            *p1 = v0.p1;
            *p2 = v0.p2;
        } else if (count == 2) {
            *p1 = v0.w * v0.p1 + v1.w * v1.p1;
            *p2 = v0.w * v0.p2 + v1.w * v1.p2;
        } else if (count == 3) {
            *p1 = v0.w * v0.p1 + v1.w * v1.p1 + v2.w * v2.p1;
            *p2 = v0.w * v0.p2 + v1.w * v1.p2 + v2.w * v2.p2;
        } else {
            *p1 = v0.w * v0.p1 + v1.w * v1.p1 + v2.w * v2.p1 + v3.w * v3.p1;
            *p2 = *p1;
        }
    }

    BX_HD void Solve2(const Vec3& Q) {  // gjk.cpp:271-301
        SVert A = v0, B = v1;
        scalar wAB[3];
        gjk::Bary2(wAB, A.p, B.p, Q);
        if (wAB[1] <= scalar(0)) { Keep1(A); return; }
        if (wAB[0] <= scalar(0)) { Keep1(B); return; }
        Keep2(A, B, wAB[0], wAB[1], wAB[2]);
    }

    BX_HD void Solve3(const Vec3& Q) {  // gjk.cpp:305-414
        SVert A = v0, B = v1, C = v2;
        scalar wAB[3], wBC[3], wCA[3];
        gjk::Bary2(wAB, A.p, B.p, Q);
        gjk::Bary2(wBC, B.p, C.p, Q);
        gjk::Bary2(wCA, C.p, A.p, Q);
        if (wAB[1] <= scalar(0) && wCA[0] <= scalar(0)) { Keep1(A); return; }
        if (wAB[0] <= scalar(0) && wBC[1] <= scalar(0)) { Keep1(B); return; }
        if (wBC[0] <= scalar(0) && wCA[1] <= scalar(0)) { Keep1(C); return; }
        scalar wABC[4];
        gjk::Bary3(wABC, A.p, B.p, C.p, Q);
        scalar area = wABC[3];
        if (wAB[0] > scalar(0) && wAB[1] > scalar(0) && area * wABC[2] <= scalar(0)) {
            Keep2(A, B, wAB[0], wAB[1], wAB[2]);
            return;
        }
        if (wBC[0] > scalar(0) && wBC[1] > scalar(0) && area * wABC[0] <= scalar(0)) {
            Keep2(B, C, wBC[0], wBC[1], wBC[2]);
            return;
        }
        if (wCA[0] > scalar(0) && wCA[1] > scalar(0) && area * wABC[1] <= scalar(0)) {
            Keep2(C, A, wCA[0], wCA[1], wCA[2]);
            return;
        }
        scalar divisor = wABC[3];
        if (divisor <= scalar(0)) return;  // degenerate triangle: the simplex stays as it is
        Keep3(A, B, C, wABC[0], wABC[1], wABC[2], divisor);
    }

    BX_HD void Solve4(const Vec3& Q) {  // gjk.cpp:417-644
        SVert A = v0, B = v1, C = v2, D = v3;
        scalar wAB[3], wAC[3], wAD[3], wBC[3], wCD[3], wDB[3];
        gjk::Bary2(wAB, A.p, B.p, Q);
        gjk::Bary2(wBC, B.p, C.p, Q);
        gjk::Bary2(wAC, A.p, C.p, Q);
        gjk::Bary2(wAD, A.p, D.p, Q);
        gjk::Bary2(wCD, C.p, D.p, Q);
        gjk::Bary2(wDB, D.p, B.p, Q);
        if (wAB[1] <= scalar(0) && wAC[1] <= scalar(0) && wAD[1] <= scalar(0)) { Keep1(A); return; }
        if (wAB[0] <= scalar(0) && wDB[0] <= scalar(0) && wBC[1] <= scalar(0)) { Keep1(B); return; }
        if (wAC[0] <= scalar(0) && wBC[0] <= scalar(0) && wCD[1] <= scalar(0)) { Keep1(C); return; }
        if (wAD[0] <= scalar(0) && wCD[0] <= scalar(0) && wDB[1] <= scalar(0)) { Keep1(D); return; }
        scalar wACB[4], wABD[4], wADC[4], wBCD[4];
        gjk::Bary3(wACB, A.p, C.p, B.p, Q);
        gjk::Bary3(wABD, A.p, B.p, D.p, Q);
        gjk::Bary3(wADC, A.p, D.p, C.p, Q);
        gjk::Bary3(wBCD, B.p, C.p, D.p, Q);
        if (wABD[2] <= scalar(0) && wACB[1] <= scalar(0) && wAB[0] > scalar(0) && wAB[1] > scalar(0)) {
            Keep2(A, B, wAB[0], wAB[1], wAB[2]);
            return;
        }
        if (wACB[2] <= scalar(0) && wADC[1] <= scalar(0) && wAC[0] > scalar(0) && wAC[1] > scalar(0)) {
            Keep2(A, C, wAC[0], wAC[1], wAC[2]);
            return;
        }
        if (wADC[2] <= scalar(0) && wABD[1] <= scalar(0) && wAD[0] > scalar(0) && wAD[1] > scalar(0)) {
            Keep2(A, D, wAD[0], wAD[1], wAD[2]);
            return;
        }
        if (wACB[0] <= scalar(0) && wBCD[2] <= scalar(0) && wBC[0] > scalar(0) && wBC[1] > scalar(0)) {
            Keep2(B, C, wBC[0], wBC[1], wBC[2]);
            return;
        }
        if (wADC[0] <= scalar(0) && wBCD[0] <= scalar(0) && wCD[0] > scalar(0) && wCD[1] > scalar(0)) {
            Keep2(C, D, wCD[0], wCD[1], wCD[2]);
            return;
        }
        if (wABD[0] <= scalar(0) && wBCD[1] <= scalar(0) && wDB[0] > scalar(0) && wDB[1] > scalar(0)) {
            Keep2(D, B, wDB[0], wDB[1], wDB[2]);
            return;
        }
        scalar wABCD[5];
        gjk::Bary4(wABCD, A.p, B.p, C.p, D.p, Q);
        if (wABCD[3] <= scalar(0) && wACB[0] > scalar(0) && wACB[1] > scalar(0) && wACB[2] > scalar(0)) {
            Keep3(A, C, B, wACB[0], wACB[1], wACB[2], wACB[0] + wACB[1] + wACB[2]);
            return;
        }
        if (wABCD[2] <= scalar(0) && wABD[0] > scalar(0) && wABD[1] > scalar(0) && wABD[2] > scalar(0)) {
            Keep3(A, B, D, wABD[0], wABD[1], wABD[2], wABD[0] + wABD[1] + wABD[2]);
            return;
        }
        if (wABCD[1] <= scalar(0) && wADC[0] > scalar(0) && wADC[1] > scalar(0) && wADC[2] > scalar(0)) {
            Keep3(A, D, C, wADC[0], wADC[1], wADC[2], wADC[0] + wADC[1] + wADC[2]);
            return;
        }
        if (wABCD[0] <= scalar(0) && wBCD[0] > scalar(0) && wBCD[1] > scalar(0) && wBCD[2] > scalar(0)) {
            Keep3(B, C, D, wBCD[0], wBCD[1], wBCD[2], wBCD[0] + wBCD[1] + wBCD[2]);
            return;
        }
        scalar divisor = wABCD[0] + wABCD[1] + wABCD[2] + wABCD[3];
        if (divisor <= scalar(0)) return;
        scalar s = scalar(1) / divisor;
        count = 4;
        v0.w = s * wABCD[0];
        v1.w = s * wABCD[1];
        v2.w = s * wABCD[2];
        v3.w = s * wABCD[3];
    }
};

// b3GJK with an empty simplex cache (gjk.cpp:651-785, 911-917).
BX_HD GJKOut GJK(const Transform& xf1, const ProxyD& proxy1, const Transform& xf2, const ProxyD& proxy2,
                 bool applyRadius) {
    SimplexD s;
    s.count = 1;
    s.v0.p1 = Mul(xf1, proxy1.Get(0));
    s.v0.p2 = Mul(xf2, proxy2.Get(0));
    s.v0.p = s.v0.p2 - s.v0.p1;
    s.v0.w = scalar(1);
    s.v0.i1 = 0;
    s.v0.i2 = 0;
    s.v1 = s.v0;
    s.v2 = s.v0;
    s.v3 = s.v0;
    const Vec3 origin(scalar(0), scalar(0), scalar(0));
    const u32 kMaxIters = 20;
    u32 iter = 0;
    while (iter < kMaxIters) {
        // vertices of the simplex before solving: a new support point equal to one of them ends
        // the search (gjk.cpp:682-688, 744-757)
        u32 saveCount = s.count;
        u32 sa1 = s.v0.i1, sa2 = s.v0.i2, sb1 = s.v1.i1, sb2 = s.v1.i2;
        u32 sc1 = s.v2.i1, sc2 = s.v2.i2, sd1 = s.v3.i1, sd2 = s.v3.i2;
        if (s.count == 2) s.Solve2(origin);
        else if (s.count == 3) s.Solve3(origin);
        else if (s.count == 4) s.Solve4(origin);
        if (s.count == 4) break;
        Vec3 d = s.SearchDirection(origin);
        if (Dot(d, d) < kEpsilon * kEpsilon) break;
        SVert nv;
        nv.i1 = proxy1.Support(MulC(xf1.rotation, -d));
        nv.p1 = Mul(xf1, proxy1.Get(nv.i1));
        nv.i2 = proxy2.Support(MulC(xf2.rotation, d));
        nv.p2 = Mul(xf2, proxy2.Get(nv.i2));
        nv.p = nv.p2 - nv.p1;
        nv.w = scalar(0);
        ++iter;
        bool duplicate = (saveCount > 0u && nv.i1 == sa1 && nv.i2 == sa2) ||
                         (saveCount > 1u && nv.i1 == sb1 && nv.i2 == sb2) ||
                         (saveCount > 2u && nv.i1 == sc1 && nv.i2 == sc2) ||
                         (saveCount > 3u && nv.i1 == sd1 && nv.i2 == sd2);
        if (duplicate) break;
        if (s.count == 1) s.v1 = nv;
        else if (s.count == 2) s.v2 = nv;
        else s.v3 = nv;
        ++s.count;
    }
    GJKOut out;
    out.simplexCount = s.count;
    out.a1 = s.v0.i1;
    out.b1 = s.v1.i1;
    out.c1 = s.v2.i1;
    out.a2 = s.v0.i2;
    out.b2 = s.v1.i2;
    out.c2 = s.v2.i2;
    s.ClosestPoints(&out.point1, &out.point2);
    out.distance = Distance(out.point1, out.point2);
    if (applyRadius) {
        scalar r1 = proxy1.radius, r2 = proxy2.radius;
        if (out.distance > r1 + r2 && out.distance > kEpsilon) {
            out.distance -= r1 + r2;
            Vec3 n = Normalize(out.point2 - out.point1);
            out.point1 += r1 * n;
            out.point2 -= r2 * n;
        } else {
            Vec3 p = scalar(0.5) * (out.point1 + out.point2);
            out.point1 = p;
            out.point2 = p;
            out.distance = scalar(0);
        }
    }
    return out;
}

// b3TestOverlap (collision.cpp:100-112) as used by b3ConvexContact::TestOverlap for sensors.
BX_HD bool ConvexTestOverlap(const Transform& xfA, const ProxyD& pA, const Transform& xfB, const ProxyD& pB) {
    GJKOut q = GJK(xfA, pA, xfB, pB, true);
    const scalar kTol = scalar(10) * kEpsilon;
    return q.distance <= kTol;
}

// ------------------------------------------------------------------ helpers
BX_HD void OnePoint(Manifold& m, const Vec3& n1, const Vec3& p1, const Vec3& p2) {
    m.pointCount = 1;
    m.points[0].localNormal1 = n1;
    m.points[0].localPoint1 = p1;
    m.points[0].localPoint2 = p2;
    m.points[0].triangleKey = kNull;
    MakeKey(m.points[0].key, 0u, 0u, 0u, 0u);  // b3MakeID(0, 0)
}

struct SegClip {  // two b3ClipVertex of a clipped segment
    ClipVertex c0, c1;
};

BX_HD SegClip BuildSegment(const Vec3& v1, const Vec3& v2) {  // clip.cpp:22-30
    SegClip s;
    s.c0.position = v1;
    MakeClipKey(s.c0.pair, kNull, kNull, 0u, 0u);
    s.c1.position = v2;
    MakeClipKey(s.c1.pair, kNull, kNull, 1u, 0u);
    return s;
}

// b3ClipSegmentToPlane (clip.cpp:57-100). It keeps both points, or one point and the
// intersection, or nothing, so the result has 0 or 2 points; `io` keeps its old contents when
// nothing is kept, as the reference's output buffer does.
BX_HD u32 ClipSegmentToPlane(SegClip& io, const ClipPlane& plane) {
    scalar d1 = Distance(io.c0.position, plane.plane);
    scalar d2 = Distance(io.c1.position, plane.plane);
    if (d1 <= scalar(0) && d2 <= scalar(0)) return 2;
    if (d1 <= scalar(0) && d2 > scalar(0)) {
        scalar fraction = d1 / (d1 - d2);
        ClipVertex x;
        x.position = io.c0.position + fraction * (io.c1.position - io.c0.position);
        MakeClipKey(x.pair, plane.edge, kNull, Id32(io.c0.pair[3]), kNull);
        io.c1 = x;
        return 2;
    }
    if (d1 > scalar(0) && d2 <= scalar(0)) {
        scalar fraction = d1 / (d1 - d2);
        ClipVertex x;
        x.position = io.c0.position + fraction * (io.c1.position - io.c0.position);
        MakeClipKey(x.pair, kNull, plane.edge, kNull, Id32(io.c1.pair[2]));
        io.c0 = io.c1;
        io.c1 = x;
        return 2;
    }
    return 0;
}

BX_HD u32 ClipSegmentToSegmentSidePlanes(SegClip& io, const Vec3& P1, const Vec3& Q1) {  // clip.cpp:153-199
    Vec3 E1 = Q1 - P1;
    ClipPlane plane1;
    plane1.plane.normal = Normalize(E1);
    plane1.plane.offset = Dot(plane1.plane.normal, Q1);
    plane1.edge = 0;
    u32 n = ClipSegmentToPlane(io, plane1);
    if (n < 2) return n;
    ClipPlane plane2;
    plane2.plane.normal = -plane1.plane.normal;
    plane2.plane.offset = Dot(plane2.plane.normal, P1);
    plane2.edge = 1;
    return ClipSegmentToPlane(io, plane2);
}

BX_HD u32 ClipSegmentToFaceSidePlanes(SegClip& io, const Transform& xf, scalar r, u32 index,
                                      const HullView& h) {  // clip.cpp:202-241
    u32 n = 0;
    u32 begin = h.faces[index];
    u32 e = begin;
    do {
        HalfEdge edge = h.edges[e];
        u32 edgeId = h.edges[edge.twin].twin;
        Plane plane = h.GetEdgeSidePlane(edgeId);
        plane.offset += r;
        ClipPlane cp;
        cp.edge = edgeId;
        cp.plane = Mul(xf, plane);
        n = ClipSegmentToPlane(io, cp);
        if (n == 0) return n;
        e = edge.next;
    } while (e != begin);
    return n;
}

// ------------------------------------------------------------------ colliders
BX_HD void CollideSpheres(Manifold& m, const Transform& xf1, const Fixture& s1, const Transform& xf2,
                          const Fixture& s2) {  // collide_spheres.cpp:22-52
    Vec3 c1 = Mul(xf1, s1.v1);
    scalar r1 = s1.radius;
    Vec3 c2 = Mul(xf2, s2.v1);
    scalar r2 = s2.radius;
    Vec3 d = c2 - c1;
    scalar dd = Dot(d, d);
    scalar totalRadius = r1 + r2;
    if (dd > totalRadius * totalRadius) return;
    scalar distance = Length(d);
    Vec3 normal(scalar(0), scalar(1), scalar(0));
    if (distance > kEpsilon) normal = d / distance;
    OnePoint(m, MulC(xf1.rotation, normal), s1.v1, s2.v1);
}

BX_HD void CollideCapsuleAndSphere(Manifold& m, const Transform& xf1, const Fixture& s1, const Transform& xf2,
                                   const Fixture& s2) {  // collide_capsule_sphere.cpp:23-120
    Vec3 Q = MulT(xf1, Mul(xf2, s2.v1));
    Vec3 A = s1.v1, B = s1.v2;
    Vec3 AB = B - A;
    scalar u = Dot(B - Q, AB);
    scalar v = Dot(Q - A, AB);
    scalar radius = s1.radius + s2.radius;
    if (v <= scalar(0) || u <= scalar(0)) {  // vertex region A, else B
        Vec3 P = v <= scalar(0) ? A : B;
        Vec3 d = Q - P;
        scalar dd = Dot(d, d);
        if (dd > radius * radius) return;
        Vec3 n(scalar(0), scalar(1), scalar(0));
        scalar len = Length(d);
        if (len > kEpsilon) n = d / len;
        OnePoint(m, n, P, s2.v1);
        return;
    }
    scalar s = Dot(AB, AB);
    Vec3 P = (u * A + v * B) / s;
    Vec3 d = Q - P;
    scalar dd = Dot(d, d);
    if (dd > radius * radius) return;
    Vec3 AQ = Q - A;
    Vec3 AB_x_AQ = Cross(AB, AQ);
    Vec3 n = Cross(AB_x_AQ, AB);
    if (Dot(n, AQ) < scalar(0)) n = -n;
    n.Normalize();
    OnePoint(m, n, P, s2.v1);
}

BX_HD bool SegmentsParallel(const Vec3& P1, const Vec3& Q1, const Vec3& P2, const Vec3& Q2) {  // collide_capsules.cpp:25-44
    Vec3 E1 = Q1 - P1;
    scalar L1 = Length(E1);
    if (L1 < kLinearSlop) return false;
    Vec3 E2 = Q2 - P2;
    scalar L2 = Length(E2);
    if (L2 < kLinearSlop) return false;
    const scalar kTol = scalar(0.005f);
    Vec3 N = Cross(E1, E2);
    return Length(N) < kTol * L1 * L2;
}

BX_HD void SetPoint(ManifoldPoint& p, const Vec3& n1, const Vec3& p1, const Vec3& p2, const u32 key[4]) {
    p.localNormal1 = n1;
    p.localPoint1 = p1;
    p.localPoint2 = p2;
    p.triangleKey = kNull;
    MakeKey(p.key, key[0], key[1], key[2], key[3]);
}

BX_HD void SetPoint(ManifoldPoint& p, const Vec3& n1, const Vec3& p1, const Vec3& p2, const uint16_t key[4]) {
    u32 wide[4];
    WidenKey(wide, key);
    SetPoint(p, n1, p1, p2, wide);
}

BX_HD void CollideCapsules(Manifold& m, const Transform& xf1, const Fixture& s1, const Transform& xf2,
                           const Fixture& s2) {  // collide_capsules.cpp:46-123
    Vec3 P1 = Mul(xf1, s1.v1), Q1 = Mul(xf1, s1.v2);
    Vec3 P2 = Mul(xf2, s2.v1), Q2 = Mul(xf2, s2.v2);
    Vec3 point1, point2;
    ClosestPointsOnSegments(point1, point2, P1, Q1, P2, Q2);
    scalar distance = Distance(point1, point2);
    scalar r1 = s1.radius, r2 = s2.radius;
    scalar totalRadius = r1 + r2;
    if (distance > totalRadius) return;
    if (distance > scalar(0) && SegmentsParallel(P1, Q1, P2, Q2)) {
        SegClip seg = BuildSegment(P1, Q1);
        u32 clipCount = ClipSegmentToSegmentSidePlanes(seg, P2, Q2);
        if (clipCount == 2) {
            Vec3 cp1 = ClosestPointOnSegment(seg.c0.position, P2, Q2);
            Vec3 cp2 = ClosestPointOnSegment(seg.c1.position, P2, Q2);
            scalar d1 = Distance(seg.c0.position, cp1);
            scalar d2 = Distance(seg.c1.position, cp2);
            if (d1 > kEpsilon && d1 <= totalRadius && d2 > kEpsilon && d2 <= totalRadius) {
                Vec3 n1 = (cp1 - seg.c0.position) / d1;
                Vec3 n2 = (cp2 - seg.c1.position) / d2;
                m.pointCount = 2;
                SetPoint(m.points[0], MulC(xf1.rotation, n1), MulT(xf1, seg.c0.position), MulT(xf2, cp1), seg.c0.pair);
                SetPoint(m.points[1], MulC(xf1.rotation, n2), MulT(xf1, seg.c1.position), MulT(xf2, cp2), seg.c1.pair);
                return;
            }
        }
    }
    if (distance > scalar(0)) {
        Vec3 normal = (point2 - point1) / distance;
        OnePoint(m, MulC(xf1.rotation, normal), MulT(xf1, point1), MulT(xf2, point2));
    }
}

BX_HD void CollideHullAndSphere(Manifold& m, const Transform& xf1, const HullView& h1, const Fixture& s1,
                                const Transform& xf2, const Fixture& s2) {  // collide_hull_sphere.cpp:26-97
    scalar radius = s1.radius + s2.radius;
    Vec3 cLocal = MulT(xf1, Mul(xf2, s2.v1));
    u32 faceIndex = 0;
    scalar separation = -kMaxScalar;
    for (u32 i = 0; i < h1.faceCount; ++i) {
        scalar s = Distance(cLocal, h1.planes[i]);
        if (s > radius) return;
        if (s > separation) { faceIndex = i; separation = s; }
    }
    if (separation < scalar(0)) {
        Plane localPlane1 = h1.planes[faceIndex];
        OnePoint(m, localPlane1.normal, ClosestPointOnPlane(cLocal, localPlane1), s2.v1);
        return;
    }
    ProxyD p1 = MakeProxy(s1, &h1), p2 = MakeProxy(s2, nullptr);
    GJKOut q = GJK(xf1, p1, xf2, p2, false);
    if (q.distance > radius) return;
    if (q.distance > scalar(0)) {
        Vec3 normal = (q.point2 - q.point1) / q.distance;
        OnePoint(m, MulC(xf1.rotation, normal), MulT(xf1, q.point1), s2.v1);
    }
}

BX_HD void CapsuleFaceContact(Manifold& m, const Transform& xf1, const HullView& h1, u32 index1, scalar r1,
                              const Transform& xf2, const Fixture& s2, scalar r2) {  // collide_hull_capsule.cpp:65-105
    SegClip seg = BuildSegment(Mul(xf2, s2.v1), Mul(xf2, s2.v2));
    u32 clipCount = ClipSegmentToFaceSidePlanes(seg, xf1, r1, index1, h1);
    Plane localPlane1 = h1.planes[index1];
    Plane plane1 = Mul(xf1, localPlane1);
    scalar totalRadius = r1 + r2;
    u32 pointCount = 0;
#pragma unroll
    for (u32 i = 0; i < 2; ++i) {
        if (i >= clipCount) break;
        const ClipVertex& cv = i == 0 ? seg.c0 : seg.c1;
        Vec3 c2 = cv.position;
        scalar s = Distance(c2, plane1);
        if (s <= totalRadius) {
            Vec3 c1 = ClosestPointOnPlane(c2, plane1);
            SetPoint(m.points[pointCount], localPlane1.normal, MulT(xf1, c1), MulT(xf2, c2), cv.pair);
            ++pointCount;
        }
    }
    m.pointCount = pointCount;
}

BX_HD void CapsuleEdgeContact(Manifold& m, const Transform& xf1, const HullView& h1, u32 index1,
                              const Transform& xf2, const Fixture& s2) {  // collide_hull_capsule.cpp:28-63
    HalfEdge e1 = h1.edges[index1];
    HalfEdge t1 = h1.edges[index1 + 1];
    Vec3 C1 = Mul(xf1, h1.centroid);
    Vec3 P1 = Mul(xf1, h1.vertices[e1.origin]);
    Vec3 Q1 = Mul(xf1, h1.vertices[t1.origin]);
    Vec3 E1 = Q1 - P1;
    Vec3 P2 = Mul(xf2, s2.v1);
    Vec3 Q2 = Mul(xf2, s2.v2);
    Vec3 E2 = Q2 - P2;
    Vec3 N = Cross(E1, E2);
    N.Normalize();
    if (Dot(N, P1 - C1) < scalar(0)) N = -N;
    Vec3 point1, point2;
    ClosestPointsOnSegments(point1, point2, P1, Q1, P2, Q2);
    u32 key[4] = {index1, index1 + 1, 0u, 1u};
    m.pointCount = 1;
    SetPoint(m.points[0], MulC(xf1.rotation, N), MulT(xf1, point1), MulT(xf2, point2), key);
}

// geometry.h:27-43 (segment AB; the divisor is |AB|^2)
BX_HD void BaryGeom2(scalar out[3], const Vec3& A, const Vec3& B, const Vec3& Q) {
    Vec3 AB = B - A;
    Vec3 QA = A - Q;
    Vec3 QB = B - Q;
    scalar divisor = Dot(AB, AB);
    out[0] = Dot(QB, AB);
    out[1] = -Dot(QA, AB);
    out[2] = divisor;
}

// geometry.h:45-70 (triangle ABC; the divisor is |AB x AC|^2)
BX_HD void BaryGeom3(scalar out[4], const Vec3& A, const Vec3& B, const Vec3& C, const Vec3& Q) {
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

// b3UniqueCount (math.h:84-104) of the first n <= 3 of (a, b, c).
BX_HD u32 UniqueCount3(u32 n, u32 a, u32 b, u32 c) {
    if (n == 1u) return 1u;
    if (n == 2u) return a == b ? 1u : 2u;
    u32 count = 1u;
    if (b != a) ++count;
    if (c != a && c != b) ++count;
    return count;
}

// The proxy-1 side of b3GetFeaturePair (gjk_feature_pair.cpp:21-208) for the final simplex of q:
// the number of proxy-1 vertices of the closest feature and, when it is an edge, its two vertices.
BX_HD u32 FeatureCount1(const GJKOut& q, u32& e0, u32& e1) {
    const u32 n = q.simplexCount;
    if (n == 1u) return 1u;
    if (n == 2u) {
        u32 u1 = UniqueCount3(2u, q.a1, q.b1, 0u), u2 = UniqueCount3(2u, q.a2, q.b2, 0u);
        if ((u1 == 2u && u2 == 2u) || (u1 != 1u && u2 == 1u)) { e0 = q.a1; e1 = q.b1; return 2u; }
        if (u1 == 1u) return 1u;
        return 0u;
    }
    if (n == 3u) {
        u32 u1 = UniqueCount3(3u, q.a1, q.b1, q.c1), u2 = UniqueCount3(3u, q.a2, q.b2, q.c2);
        if (u1 == 1u && u2 == 3u) return 1u;
        if (u1 == 3u && u2 == 1u) return 3u;
        if (u1 == 2u && (u2 == 2u || u2 == 3u)) {
            e0 = q.a1;
            e1 = q.a1 == q.b1 ? q.c1 : q.b1;
            return 2u;
        }
        if (u1 == 3u) return 3u;
    }
    return 0u;
}

// collide_hull_capsule.cpp:107-181; with `wings`, the triangle-capsule sequence of a mesh triangle
// (collide_triangle_capsule.cpp:107-278), whose closest-edge step consults the edge wings.
BX_HD void CollideHullAndCapsule(Manifold& m, const Transform& xf1, const HullView& h1, const Fixture& s1,
                                 const Transform& xf2, const Fixture& s2, const TriWings* wings = nullptr) {
    scalar r1 = s1.radius, r2 = s2.radius;
    scalar totalRadius = r1 + r2;
    ProxyD p1 = MakeProxy(s1, &h1), p2 = MakeProxy(s2, nullptr);
    GJKOut q = GJK(xf1, p1, xf2, p2, false);
    if (q.distance > totalRadius) return;
    if (q.distance > scalar(0)) {
        Vec3 c1 = q.point1, c2 = q.point2;
        Vec3 N1 = (c2 - c1) / q.distance;
        Vec3 localN1 = MulC(xf1.rotation, N1);
        u32 index1 = h1.GetSupportFace(localN1);
        Vec3 localFaceN1 = h1.planes[index1].normal;
        const scalar kTol = scalar(0.005);
        Vec3 N = Cross(localN1, localFaceN1);
        scalar L = Dot(N, N);
        if (L < kTol * kTol) {
            CapsuleFaceContact(m, xf1, h1, index1, r1, xf2, s2, r2);
            if (m.pointCount == 2) return;
        }
        u32 v1 = 0, v2 = 0;
        if (wings && wings->mask && FeatureCount1(q, v1, v2) == 2u) {
            // collide_triangle_capsule.cpp:162-241: the closest feature is a triangle edge
            u32 edgeIndex = 0;
            for (u32 i = 0; i < 3u; ++i) {
                u32 j = i + 1u < 3u ? i + 1u : 0u;
                if ((v1 == i && v2 == j) || (v2 == i && v1 == j)) { edgeIndex = i; break; }
            }
            if (wings->Has(edgeIndex)) {
                u32 ev1 = edgeIndex;
                u32 ev2 = edgeIndex + 1u < 3u ? edgeIndex + 1u : 0u;
                Vec3 Q = MulT(xf1, c2);
                Vec3 A = wings->Get(edgeIndex);
                Vec3 B = ev2 == 0u ? s1.v1 : (ev2 == 1u ? s1.v2 : s1.v3);
                Vec3 C = ev1 == 0u ? s1.v1 : (ev1 == 1u ? s1.v2 : s1.v3);
                scalar wABC[4];
                BaryGeom3(wABC, A, B, C, Q);
                if (wABC[0] > scalar(0) && wABC[1] > scalar(0) && wABC[2] > scalar(0)) return;
                Vec3 center = (A + B + C) / scalar(3);
                Plane plane = h1.planes[0];
                scalar distance = Distance(center, plane);
                const scalar kCoplanarTol = scalar(0.005);
                if (distance > -kCoplanarTol && distance < kCoplanarTol) {
                    Vec3 n = plane.normal;
                    if (Dot(n, localN1) < scalar(0)) n = -n;
                    Vec3 nc2 = c1 + q.distance * Mul(xf1.rotation, n);
                    localN1 = n;
                    c2 = nc2;
                }
            }
        }
        OnePoint(m, localN1, MulT(xf1, c1), MulT(xf2, c2));
        return;
    }
    // face and edge separation of the hull and the capsule segment (sat_hull_edge.cpp:28-146),
    // in the hull's frame
    Transform xf = MulT(xf1, xf2);
    Vec3 a = Mul(xf, s2.v1), b = Mul(xf, s2.v2);
    u32 faceIndex = 0;
    scalar faceSeparation = -kMaxScalar;
    for (u32 i = 0; i < h1.faceCount; ++i) {
        Plane plane = h1.planes[i];
        Vec3 nd = -plane.normal;
        Vec3 support = Dot(nd, a) > Dot(nd, b) ? a : b;  // b3Capsule::GetSupportVertex
        scalar separation = Distance(support, plane);
        if (separation > faceSeparation) { faceIndex = i; faceSeparation = separation; }
    }
    if (faceSeparation > totalRadius) return;
    u32 edgeIndex = 0;
    scalar edgeSeparation = -kMaxScalar;
    {
        Vec3 C1 = h1.centroid;
        Vec3 E2 = b - a;
        for (u32 i = 0; i < h1.edgeCount; i += 2) {
            HalfEdge e1 = h1.edges[i];
            HalfEdge t1 = h1.edges[i + 1];
            Vec3 P1 = h1.vertices[e1.origin];
            Vec3 Q1 = h1.vertices[t1.origin];
            Vec3 E1 = Q1 - P1;
            Vec3 U1 = h1.planes[e1.face].normal;
            Vec3 V1 = h1.planes[t1.face].normal;
            if (Dot(E2, U1) * Dot(E2, V1) < scalar(0)) {  // b3IsMinkowskiFace
                scalar separation = -kMaxScalar;          // b3ProjectEdge (sat_hull_edge.cpp:63-96)
                scalar L1 = Length(E1);
                scalar L2 = Length(E2);
                if (L1 >= kLinearSlop && L2 >= kLinearSlop) {
                    const scalar kTolE = 0.005f;
                    Vec3 E1_x_E2 = Cross(E1, E2);
                    scalar L = Length(E1_x_E2);
                    if (L >= kTolE * L1 * L2) {
                        Vec3 N = (scalar(1) / L) * E1_x_E2;
                        if (Dot(N, P1 - C1) < scalar(0)) N = -N;
                        separation = Dot(N, a - P1);
                    }
                }
                if (separation > edgeSeparation) { edgeSeparation = separation; edgeIndex = i; }
            }
        }
    }
    if (edgeSeparation > totalRadius) return;
    const scalar kRelEdgeTol = scalar(0.90);
    const scalar kAbsTol = scalar(0.5) * kLinearSlop;
    if (edgeSeparation > kRelEdgeTol * faceSeparation + kAbsTol)
        CapsuleEdgeContact(m, xf1, h1, edgeIndex, xf2, s2);
    else
        CapsuleFaceContact(m, xf1, h1, faceIndex, r1, xf2, s2, r2);
}

}  // namespace bx
