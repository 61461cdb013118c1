// bx_collide.h -- narrowphase contact generation, host/device.
//
// Hull-hull contacts with the SAT feature cache, restating
//   src/bounce/collision/sat/sat.cpp:26-282            face / edge separation queries, cache reads
//   src/bounce/collision/clip.cpp:33-150, 202-281      polygon building and side-plane clipping
//   src/bounce/collision/cluster.cpp:23-200            polygon sort and reduction to <= 4 points
//   src/bounce/collision/collide/collide_hulls.cpp:32-533
//   src/bounce/collision/collision.cpp:145-215         closest points between segments
// Polygons live in fixed-capacity private arrays (kMaxClip, the capacity of the reference's
// b3StackArray<b3ClipVertex, 32>); an overflow is reported, never silently truncated.
#pragma once
#include "bx_types.h"

namespace bx {

constexpr u32 kMaxClip = 32;

// b3ClipVertex with its b3FeaturePair held in 16 bits per id while clipping (the ids are hull
// half-edge indices, below 0xFFFF by the upload checks, or the null edge, held as 0xFFFF); they
// are widened back to the reference's 32-bit ids when a manifold point is written.
struct ClipVertex { Vec3 position; uint16_t pair[4]; };
constexpr uint16_t kNullId16 = 0xFFFFu;
BX_HD uint16_t Id16(u32 x) { return x == kNull ? kNullId16 : uint16_t(x); }
BX_HD u32 Id32(uint16_t x) { return x == kNullId16 ? kNull : u32(x); }
BX_HD void MakeClipKey(uint16_t out[4], u32 a, u32 b, u32 c, u32 d) {
    out[0] = Id16(a); out[1] = Id16(b); out[2] = Id16(c); out[3] = Id16(d);
}
BX_HD void WidenKey(u32 out[4], const uint16_t in[4]) {
    out[0] = Id32(in[0]); out[1] = Id32(in[1]); out[2] = Id32(in[2]); out[3] = Id32(in[3]);
}
struct ClipPlane { Plane plane; u32 edge; };
struct ClusterVertex { Vec3 position; u32 clipIndex; };

struct ClipPolygon {
    ClipVertex v[kMaxClip];
    u32 count;
};

BX_HD void MakeKey(u32 out[4], u32 a, u32 b, u32 c, u32 d) { out[0] = a; out[1] = b; out[2] = c; out[3] = d; }

BX_HD Vec3 ClosestPointOnSegment(const Vec3& Q, const Vec3& A, const Vec3& B) {
    Vec3 AB = B - A;
    scalar u = Dot(B - Q, AB);
    scalar v = Dot(Q - A, AB);
    if (v <= scalar(0)) return A;
    if (u <= scalar(0)) return B;
    scalar den = Dot(AB, AB);
    if (den < kLinearSlop * kLinearSlop) return A;
    return (u * A + v * B) / den;
}

BX_HD void ClosestPointsOnSegments(Vec3& C1, Vec3& C2, const Vec3& P1, const Vec3& Q1, const Vec3& P2, const Vec3& Q2) {
    Vec3 E1 = Q1 - P1;
    scalar L1 = Length(E1);
    Vec3 E2 = Q2 - P2;
    scalar L2 = Length(E2);
    if (L1 < kLinearSlop && L2 < kLinearSlop) { C1 = P1; C2 = P2; return; }
    if (L1 < kLinearSlop) { C1 = P1; C2 = ClosestPointOnSegment(P1, P2, Q2); return; }
    if (L2 < kLinearSlop) { C1 = ClosestPointOnSegment(P2, P1, Q1); C2 = P2; return; }
    Vec3 N1 = E1 / L1;
    Vec3 N2 = E2 / L2;
    scalar b = Dot(N1, N2);
    scalar den = scalar(1) - b * b;
    if (den != scalar(0)) {
        scalar inv_den = scalar(1) / den;
        Vec3 E3 = P1 - P2;
        scalar d = Dot(N1, E3);
        scalar e = Dot(N2, E3);
        scalar s = inv_den * (b * e - d);
        scalar t = inv_den * (e - b * d);
        C1 = P1 + s * N1;
        C2 = P2 + t * N2;
    } else {
        C1 = P1;
        C2 = P2;
    }
    C1 = ClosestPointOnSegment(C1, P1, Q1);
    C2 = ClosestPointOnSegment(C1, P2, Q2);
    C1 = ClosestPointOnSegment(C2, P1, Q1);
}

// ------------------------------------------------------------------ SAT
struct FaceQuery { u32 index; scalar separation; };
struct EdgeQuery { u32 index1, index2; scalar separation; };

BX_HD scalar ProjectHull(const HullView& h, const Plane& plane) {
    Vec3 support = h.vertices[h.GetSupportVertex(-plane.normal)];
    return Distance(support, plane);
}

BX_HD FaceQuery QueryFaceSeparation(const Transform& xf1, const HullView& h1, const Transform& xf2, const HullView& h2) {
    Transform xf = MulT(xf2, xf1);
    u32 maxIndex = 0;
    scalar maxSeparation = -kMaxScalar;
    for (u32 i = 0; i < h1.faceCount; ++i) {
        Plane plane = Mul(xf, h1.planes[i]);
        scalar separation = ProjectHull(h2, plane);
        if (separation > maxSeparation) { maxIndex = i; maxSeparation = separation; }
    }
    FaceQuery out;
    out.index = maxIndex;
    out.separation = maxSeparation;
    return out;
}

BX_HD bool IsMinkowskiFace(const Vec3& A, const Vec3& B, const Vec3& B_x_A, const Vec3& C, const Vec3& D,
                           const Vec3& D_x_C) {
    scalar ADC = Dot(A, D_x_C);
    scalar BDC = Dot(B, D_x_C);
    scalar CBA = Dot(C, B_x_A);
    scalar DBA = Dot(D, B_x_A);
    return CBA * DBA < scalar(0) && ADC * BDC < scalar(0) && CBA * BDC > scalar(0);
}

BX_HD scalar ProjectEdges(const Vec3& P1, const Vec3& E1, const Vec3& P2, const Vec3& E2, const Vec3& C1) {
    scalar L1 = Length(E1);
    if (L1 < kLinearSlop) return -kMaxScalar;
    scalar L2 = Length(E2);
    if (L2 < kLinearSlop) return -kMaxScalar;
    const scalar kTol = scalar(0.005);
    Vec3 E1_x_E2 = Cross(E1, E2);
    scalar L = Length(E1_x_E2);
    if (L < kTol * L1 * L2) return -kMaxScalar;
    Vec3 N = (scalar(1) / L) * E1_x_E2;
    if (Dot(N, P1 - C1) < scalar(0)) N = -N;
    return Dot(N, P2 - P1);
}

BX_HD EdgeQuery QueryEdgeSeparation(const Transform& xf1, const HullView& h1, const Transform& xf2, const HullView& h2) {
    Transform xf = MulT(xf2, xf1);
    Vec3 C1 = Mul(xf, h1.centroid);
    u32 maxIndex1 = 0, maxIndex2 = 0;
    scalar maxSeparation = -kMaxScalar;
    for (u32 i = 0; i < h1.edgeCount; i += 2) {
        const HalfEdge edge1 = h1.edges[i];
        const HalfEdge twin1 = h1.edges[i + 1];
        Vec3 P1 = Mul(xf, h1.vertices[edge1.origin]);
        Vec3 Q1 = Mul(xf, h1.vertices[twin1.origin]);
        Vec3 E1 = Q1 - P1;
        Vec3 U1 = Mul(xf.rotation, h1.planes[edge1.face].normal);
        Vec3 V1 = Mul(xf.rotation, h1.planes[twin1.face].normal);
        for (u32 j = 0; j < h2.edgeCount; j += 2) {
            const HalfEdge edge2 = h2.edges[j];
            const HalfEdge twin2 = h2.edges[j + 1];
            Vec3 P2 = h2.vertices[edge2.origin];
            Vec3 Q2 = h2.vertices[twin2.origin];
            Vec3 E2 = Q2 - P2;
            Vec3 U2 = h2.planes[edge2.face].normal;
            Vec3 V2 = h2.planes[twin2.face].normal;
            if (IsMinkowskiFace(U1, V1, -E1, -U2, -V2, -E2)) {
                scalar separation = ProjectEdges(P1, E1, P2, E2, C1);
                if (separation > maxSeparation) { maxSeparation = separation; maxIndex1 = i; maxIndex2 = j; }
            }
        }
    }
    EdgeQuery out;
    out.index1 = maxIndex1;
    out.index2 = maxIndex2;
    out.separation = maxSeparation;
    return out;
}

constexpr u32 kCacheEmpty = 0, kCacheSeparation = 1, kCacheOverlap = 2;
constexpr u32 kFeatEdges = 0, kFeatFace1 = 1, kFeatFace2 = 2;

BX_HD u32 ReadFace(u32 index1, const Transform& xf1, const HullView& h1, const Transform& xf2, const HullView& h2,
                   scalar totalRadius) {
    Transform xf = MulT(xf2, xf1);
    Plane plane = Mul(xf, h1.planes[index1]);
    return ProjectHull(h2, plane) > totalRadius ? kCacheSeparation : kCacheOverlap;
}

BX_HD u32 ReadEdges(u32 i, u32 j, const Transform& xf1, const HullView& h1, const Transform& xf2, const HullView& h2,
                    scalar totalRadius) {
    Transform xf = MulT(xf2, xf1);
    Vec3 C1 = Mul(xf, h1.centroid);
    const HalfEdge edge1 = h1.edges[i];
    const HalfEdge twin1 = h1.edges[i + 1];
    Vec3 P1 = Mul(xf, h1.vertices[edge1.origin]);
    Vec3 Q1 = Mul(xf, h1.vertices[twin1.origin]);
    Vec3 E1 = Q1 - P1;
    Vec3 U1 = Mul(xf.rotation, h1.planes[edge1.face].normal);
    Vec3 V1 = Mul(xf.rotation, h1.planes[twin1.face].normal);
    const HalfEdge edge2 = h2.edges[j];
    const HalfEdge twin2 = h2.edges[j + 1];
    Vec3 P2 = h2.vertices[edge2.origin];
    Vec3 Q2 = h2.vertices[twin2.origin];
    Vec3 E2 = Q2 - P2;
    Vec3 U2 = h2.planes[edge2.face].normal;
    Vec3 V2 = h2.planes[twin2.face].normal;
    if (IsMinkowskiFace(U1, V1, -E1, -U2, -V2, -E2))
        return ProjectEdges(P1, E1, P2, E2, C1) > totalRadius ? kCacheSeparation : kCacheOverlap;
    return kCacheEmpty;
}

BX_HD u32 ReadState(FeatureCache& c, const Transform& xf1, const HullView& h1, const Transform& xf2,
                    const HullView& h2, scalar totalRadius) {
    if (c.cacheType == kCacheEmpty) {
        c.cacheType = kCacheSeparation;
        c.featureType = kFeatFace1;
        c.index1 = 0;
        c.index2 = 0;
    }
    switch (c.featureType) {
    case kFeatEdges: return ReadEdges(c.index1, c.index2, xf1, h1, xf2, h2, totalRadius);
    case kFeatFace1: return ReadFace(c.index1, xf1, h1, xf2, h2, totalRadius);
    case kFeatFace2: return ReadFace(c.index1, xf2, h2, xf1, h1, totalRadius);
    default: return kCacheEmpty;
    }
}

// ------------------------------------------------------------------ clipping
BX_HD void BuildPolygon(ClipPolygon& out, const Transform& xf, u32 index, const HullView& h, u32& overflow) {
    out.count = 0;
    u32 begin = h.faces[index];
    u32 e = begin;
    do {
        const HalfEdge edge = h.edges[e];
        const HalfEdge twin = h.edges[edge.twin];
        if (out.count < kMaxClip) {
            ClipVertex& v = out.v[out.count++];
            v.position = Mul(xf, h.vertices[edge.origin]);
            MakeClipKey(v.pair, kNull, kNull, edge.prev, twin.twin);
        } else {
            overflow = 1;
        }
        e = edge.next;
    } while (e != begin);
}

BX_HD void PushClip(ClipPolygon& p, const ClipVertex& v, u32& overflow) {
    if (p.count < kMaxClip) p.v[p.count++] = v;
    else overflow = 1;
}

BX_HD void ClipPolygonToPlane(ClipPolygon& out, const ClipPolygon& in, const ClipPlane& plane, u32& overflow) {
    out.count = 0;
    ClipVertex v1 = in.v[in.count - 1];
    scalar d1 = Distance(v1.position, plane.plane);
    for (u32 i = 0; i < in.count; ++i) {
        ClipVertex v2 = in.v[i];
        scalar d2 = Distance(v2.position, plane.plane);
        if (d1 <= scalar(0) && d2 <= scalar(0)) {
            PushClip(out, v2, overflow);
        } else if (d1 <= scalar(0) && d2 > scalar(0)) {
            scalar fraction = d1 / (d1 - d2);
            ClipVertex v;
            v.position = v1.position + fraction * (v2.position - v1.position);
            MakeClipKey(v.pair, plane.edge, kNull, Id32(v1.pair[3]), kNull);
            PushClip(out, v, overflow);
        } else if (d1 > scalar(0) && d2 <= scalar(0)) {
            scalar fraction = d1 / (d1 - d2);
            ClipVertex v;
            v.position = v1.position + fraction * (v2.position - v1.position);
            MakeClipKey(v.pair, kNull, plane.edge, kNull, Id32(v2.pair[2]));
            PushClip(out, v, overflow);
            PushClip(out, v2, overflow);
        }
        v1 = v2;
        d1 = d2;
    }
}

// Clip `poly` against the side planes of face `index`, alternating between `poly` and `tmp`;
// returns the buffer holding the result (its count is 0 once a plane leaves nothing).
BX_HD const ClipPolygon* ClipPolygonToFaceSidePlanes(ClipPolygon& poly, ClipPolygon& tmp, const Transform& xf,
                                                     scalar r, u32 index, const HullView& h, u32& overflow) {
    ClipPolygon* in = &poly;
    ClipPolygon* out = &tmp;
    u32 begin = h.faces[index];
    u32 e = begin;
    do {
        const HalfEdge edge = h.edges[e];
        const HalfEdge twin = h.edges[edge.twin];
        u32 edgeId = twin.twin;
        Plane plane = h.GetEdgeSidePlane(edgeId);
        plane.offset += r;
        ClipPlane cp;
        cp.edge = edgeId;
        cp.plane = Mul(xf, plane);
        ClipPolygonToPlane(*out, *in, cp, overflow);
        ClipPolygon* t = in;
        in = out;
        out = t;
        if (in->count == 0) return in;
        e = edge.next;
    } while (e != begin);
    return in;
}

// ------------------------------------------------------------------ reduction (cluster.cpp:23-200)
BX_HD bool IsCCW(const Vec3& A, const Vec3& B, const Vec3& C, const Vec3& N) {
    return Dot(Cross(B - A, C - A), N) > scalar(0);
}

BX_HD void SortPolygon(ClusterVertex* p, u32 n, const Vec3& N) {
    Vec3 A = p[0].position;
    for (u32 i = 1; i < n; ++i) {
        Vec3 B = p[i].position;  // captured once per i, as in the reference
        for (u32 j = i + 1; j < n; ++j) {
            Vec3 C = p[j].position;
            if (!IsCCW(A, B, C, N)) Swap(p[i], p[j]);
        }
    }
}

// Reduce `in` (count n) to out[<=4]; returns the output count.
BX_HD u32 ReducePolygon(ClusterVertex* out, const ClusterVertex* in, u32 n, const Vec3& N, u32 initialPoint) {
    if (n <= kMaxManifoldPoints) {
        for (u32 i = 0; i < n; ++i) out[i] = in[i];
        SortPolygon(out, n, N);
        return n;
    }
    u32 chosen = 0;  // bit mask over the <= 32 inputs
    u32 count = 0;
    out[count++] = in[initialPoint];
    chosen |= 1u << initialPoint;
    {
        Vec3 A = out[0].position;
        u32 index = 0;
        scalar max = -kMaxScalar;
        for (u32 i = 0; i < n; ++i) {
            if (chosen & (1u << i)) continue;
            Vec3 d = in[i].position - A;
            scalar dd = Dot(d, d);
            if (dd > max) { max = dd; index = i; }
        }
        if (max < kEpsilon * kEpsilon) return count;
        out[count++] = in[index];
        chosen |= 1u << index;
    }
    {
        Vec3 A = out[0].position, B = out[1].position;
        u32 index = 0;
        scalar max = -kMaxScalar;
        for (u32 i = 0; i < n; ++i) {
            if (chosen & (1u << i)) continue;
            Vec3 C = in[i].position;
            scalar a2 = Abs(Dot(Cross(B - A, C - A), N));
            if (a2 > max) { max = a2; index = i; }
        }
        const scalar kAreaTol = scalar(0.01);
        if (max < scalar(2) * kAreaTol) return count;
        Vec3 C = in[index].position;
        if (!IsCCW(A, B, C, N)) Swap(out[0], out[1]);
        out[count++] = in[index];
        chosen |= 1u << index;
    }
    {
        Vec3 A = out[0].position, B = out[1].position;
        u32 index = 0;
        scalar min = kMaxScalar;
        for (u32 i = 0; i < n; ++i) {
            if (chosen & (1u << i)) continue;
            Vec3 D = in[i].position;
            scalar sa2 = Dot(Cross(B - A, D - A), N);
            if (sa2 < min) { min = sa2; index = i; }
        }
        const scalar kAreaTol = scalar(0.01);
        if (Abs(min) < scalar(2) * kAreaTol) return count;
        out[count++] = in[index];
        chosen |= 1u << index;
    }
    SortPolygon(out, count, N);
    return count;
}

// ------------------------------------------------------------------ contact builders
BX_HD void SetPointId(ManifoldPoint& p, const u32 key[4]) {
    p.triangleKey = kNull;
    p.key[0] = key[0]; p.key[1] = key[1]; p.key[2] = key[2]; p.key[3] = key[3];
}

BX_HD void EdgeContactPoint(Manifold& m, const Transform& xf1, u32 index1, const Transform& xf2, u32 index2,
                            const Vec3& N, const Vec3& point1, const Vec3& point2) {
    m.pointCount = 1;
    ManifoldPoint& p = m.points[0];
    p.localNormal1 = MulC(xf1.rotation, N);
    p.localPoint1 = MulT(xf1, point1);
    p.localPoint2 = MulT(xf2, point2);
    u32 key[4];
    MakeKey(key, index1, index1 + 1, index2, index2 + 1);
    SetPointId(p, key);
    p.edgeContact = 1;
}

BX_HD void BuildEdgeContact(Manifold& m, const Transform& xf1, u32 index1, const HullView& h1, const Transform& xf2,
                            u32 index2, const HullView& h2) {
    const HalfEdge edge1 = h1.edges[index1];
    const HalfEdge twin1 = h1.edges[index1 + 1];
    Vec3 C1 = Mul(xf1, h1.centroid);
    Vec3 P1 = Mul(xf1, h1.vertices[edge1.origin]);
    Vec3 Q1 = Mul(xf1, h1.vertices[twin1.origin]);
    Vec3 E1 = Q1 - P1;
    const HalfEdge edge2 = h2.edges[index2];
    const HalfEdge twin2 = h2.edges[index2 + 1];
    Vec3 P2 = Mul(xf2, h2.vertices[edge2.origin]);
    Vec3 Q2 = Mul(xf2, h2.vertices[twin2.origin]);
    Vec3 E2 = Q2 - P2;
    Vec3 N = Cross(E1, E2);
    N.Normalize();
    if (Dot(N, P1 - C1) < scalar(0)) N = -N;
    Vec3 point1, point2;
    ClosestPointsOnSegments(point1, point2, P1, Q1, P2, Q2);
    EdgeContactPoint(m, xf1, index1, xf2, index2, N, point1, point2);
}

struct FaceScratch {
    ClipPolygon poly, tmp;
    ClusterVertex polygon1[kMaxClip];
    ClusterVertex reduced[kMaxManifoldPoints];
};

BX_HD void BuildFaceContact(Manifold& m, const Transform& xf1, u32 index1, const HullView& h1, scalar r1,
                            const Transform& xf2, const HullView& h2, scalar r2, bool flipNormal, FaceScratch& s,
                            u32& overflow) {
    scalar totalRadius = r1 + r2;
    Plane plane1 = Mul(xf1, h1.planes[index1]);
    Vec3 normal1 = MulC(xf2.rotation, plane1.normal);
    u32 index2 = h2.GetSupportFace(-normal1);
    BuildPolygon(s.poly, xf2, index2, h2, overflow);
    const ClipPolygon& poly = *ClipPolygonToFaceSidePlanes(s.poly, s.tmp, xf1, totalRadius, index1, h1, overflow);
    if (poly.count == 0) return;
    u32 n1 = 0;
    u32 minIndex = 0;
    scalar minSeparation = kMaxScalar;
    for (u32 i = 0; i < poly.count; ++i) {
        Vec3 p2 = poly.v[i].position;
        scalar separation = Distance(p2, plane1);
        if (separation <= totalRadius) {
            if (separation < minSeparation) { minIndex = n1; minSeparation = separation; }
            s.polygon1[n1].position = ClosestPointOnPlane(p2, plane1);
            s.polygon1[n1].clipIndex = i;
            ++n1;
        }
    }
    if (n1 == 0) return;
    Vec3 normal = plane1.normal;
    Vec3 orientedNormal = flipNormal ? -normal : normal;
    u32 pointCount = ReducePolygon(s.reduced, s.polygon1, n1, orientedNormal, minIndex);
    for (u32 i = 0; i < pointCount; ++i) {
        const ClipVertex& v2 = poly.v[s.reduced[i].clipIndex];
        Vec3 v1 = ClosestPointOnPlane(v2.position, plane1);
        ManifoldPoint& mp = m.points[i];
        if (flipNormal) {
            u32 key[4];
            MakeKey(key, Id32(v2.pair[2]), Id32(v2.pair[3]), Id32(v2.pair[0]), Id32(v2.pair[1]));
            mp.localNormal1 = MulC(xf2.rotation, orientedNormal);
            mp.localPoint1 = MulT(xf2, v2.position);
            mp.localPoint2 = MulT(xf1, v1);
            SetPointId(mp, key);
        } else {
            mp.localNormal1 = MulC(xf1.rotation, orientedNormal);
            mp.localPoint1 = MulT(xf1, v1);
            mp.localPoint2 = MulT(xf2, v2.position);
            u32 key[4];
            WidenKey(key, v2.pair);
            SetPointId(mp, key);
        }
        mp.edgeContact = 0;
    }
    m.pointCount = pointCount;
}

BX_HD void RebuildEdgeContact(Manifold& m, const Transform& xf1, u32 index1, const HullView& h1,
                              const Transform& xf2, u32 index2, const HullView& h2) {
    const HalfEdge edge1 = h1.edges[index1];
    const HalfEdge twin1 = h1.edges[index1 + 1];
    Vec3 C1 = Mul(xf1, h1.centroid);
    Vec3 P1 = Mul(xf1, h1.vertices[edge1.origin]);
    Vec3 Q1 = Mul(xf1, h1.vertices[twin1.origin]);
    Vec3 E1 = Q1 - P1;
    const HalfEdge edge2 = h2.edges[index2];
    const HalfEdge twin2 = h2.edges[index2 + 1];
    Vec3 P2 = Mul(xf2, h2.vertices[edge2.origin]);
    Vec3 Q2 = Mul(xf2, h2.vertices[twin2.origin]);
    Vec3 E2 = Q2 - P2;
    Vec3 point1, point2;
    ClosestPointsOnSegments(point1, point2, P1, Q1, P2, Q2);
    scalar w1[3], w2[3];
    Barycentric2(w1, P2, Q2, point1);
    Barycentric2(w2, P1, Q1, point2);
    if (w1[1] > scalar(0) && w1[1] <= w1[2] && w2[1] > scalar(0) && w2[1] <= w2[2]) {
        Vec3 N = Cross(E1, E2);
        N.Normalize();
        if (Dot(N, P1 - C1) < scalar(0)) N = -N;
        EdgeContactPoint(m, xf1, index1, xf2, index2, N, point1, point2);
    }
}

BX_HD void RebuildFaceContact(Manifold& m, const Transform& xf1, u32 index1, const HullView& h1, scalar r1,
                              const Transform& xf2, const HullView& h2, scalar r2, bool flipNormal,
                              const Transform& xf01, const Transform& xf02, FaceScratch& s, u32& overflow) {
    Quat dq0 = Conjugate(xf01.rotation) * xf02.rotation;
    Quat dq = Conjugate(xf1.rotation) * xf2.rotation;
    Quat q = Conjugate(dq0) * dq;
    const scalar kCosTol = scalar(0.995);
    if (Abs(q.s) > kCosTol) BuildFaceContact(m, xf1, index1, h1, r1, xf2, h2, r2, flipNormal, s, overflow);
}

BX_HD void SetCache(FeatureCache& c, u32 ct, u32 ft, u32 i1, u32 i2) {
    c.cacheType = ct; c.featureType = ft; c.index1 = i1; c.index2 = i2;
}

// b3CollideHullsCache (collide_hulls.cpp:355-515).
BX_HD void CollideHulls(Manifold& m, const Transform& xf1, const HullView& h1, scalar r1, const Transform& xf2,
                        const HullView& h2, scalar r2, FeatureCache& cache, const Transform& xf01,
                        const Transform& xf02, FaceScratch& s, u32& overflow) {
    scalar totalRadius = r1 + r2;
    u32 state0 = cache.cacheType;
    u32 state1 = ReadState(cache, xf1, h1, xf2, h2, totalRadius);
    if (state0 == kCacheSeparation && state1 == kCacheSeparation) return;
    if (state0 == kCacheOverlap && state1 == kCacheOverlap) {
        switch (cache.featureType) {
        case kFeatEdges: RebuildEdgeContact(m, xf1, cache.index1, h1, xf2, cache.index2, h2); break;
        case kFeatFace1:
            RebuildFaceContact(m, xf1, cache.index1, h1, r1, xf2, h2, r2, false, xf01, xf02, s, overflow);
            break;
        case kFeatFace2:
            RebuildFaceContact(m, xf2, cache.index1, h2, r2, xf1, h1, r1, true, xf02, xf01, s, overflow);
            break;
        default: break;
        }
        if (m.pointCount > 0) return;
    }
    cache.cacheType = kCacheEmpty;
    FaceQuery fq1 = QueryFaceSeparation(xf1, h1, xf2, h2);
    if (fq1.separation > totalRadius) { SetCache(cache, kCacheSeparation, kFeatFace1, fq1.index, fq1.index); return; }
    FaceQuery fq2 = QueryFaceSeparation(xf2, h2, xf1, h1);
    if (fq2.separation > totalRadius) { SetCache(cache, kCacheSeparation, kFeatFace2, fq2.index, fq2.index); return; }
    EdgeQuery eq = QueryEdgeSeparation(xf1, h1, xf2, h2);
    if (eq.separation > totalRadius) { SetCache(cache, kCacheSeparation, kFeatEdges, eq.index1, eq.index2); return; }
    const scalar kRelEdgeTol = scalar(0.90);
    const scalar kRelFaceTol = scalar(0.98);
    const scalar kAbsTol = scalar(0.5) * kLinearSlop;
    if (eq.separation > kRelEdgeTol * Max(fq1.separation, fq2.separation) + kAbsTol) {
        BuildEdgeContact(m, xf1, eq.index1, h1, xf2, eq.index2, h2);
        SetCache(cache, kCacheOverlap, kFeatEdges, eq.index1, eq.index2);
        return;
    } else if (fq2.separation > kRelFaceTol * fq1.separation + kAbsTol) {
        BuildFaceContact(m, xf2, fq2.index, h2, r2, xf1, h1, r1, true, s, overflow);
        if (m.pointCount > 0) { SetCache(cache, kCacheOverlap, kFeatFace2, fq2.index, fq2.index); return; }
    } else {
        BuildFaceContact(m, xf1, fq1.index, h1, r1, xf2, h2, r2, false, s, overflow);
        if (m.pointCount > 0) { SetCache(cache, kCacheOverlap, kFeatFace1, fq1.index, fq1.index); return; }
    }
    if (eq.separation > Max(fq1.separation, fq2.separation)) {
        BuildEdgeContact(m, xf1, eq.index1, h1, xf2, eq.index2, h2);
        SetCache(cache, kCacheOverlap, kFeatEdges, eq.index1, eq.index2);
    } else if (fq2.separation > fq1.separation) {
        BuildFaceContact(m, xf2, fq2.index, h2, r2, xf1, h1, r1, true, s, overflow);
        if (m.pointCount > 0) SetCache(cache, kCacheOverlap, kFeatFace2, fq2.index, fq2.index);
    } else {
        BuildFaceContact(m, xf1, fq1.index, h1, r1, xf2, h2, r2, false, s, overflow);
        if (m.pointCount > 0) SetCache(cache, kCacheOverlap, kFeatFace1, fq1.index, fq1.index);
    }
}

}  // namespace bx
