// orc_hulls.cpp -- oracle (test infrastructure): hull shape operations and the cached hull-hull
// narrowphase, restating
//   src/bounce/collision/shapes/hull_shape.cpp:40-164   (Kallay mass properties, AABB)
//   src/bounce/collision/sat/sat.cpp:26-282              (face/edge SAT queries, feature cache)
//   src/bounce/collision/clip.cpp:23-281                 (polygon building and clipping)
//   src/bounce/collision/cluster.cpp:23-200              (polygon sort and reduction)
//   src/bounce/collision/collide/collide_hulls.cpp:32-533
//   src/bounce/collision/collision.cpp:23-215            (manifolds, closest points)
#include <vector>
#include "orc_geom.h"

namespace orc {

// ------------------------------------------------------------------ manifolds (collision.cpp)
void WorldManifold::Initialize(const Manifold* m, scalar rA, const Transform& xfA, scalar rB, const Transform& xfB) {
    center.SetZero();
    normal.SetZero();
    pointCount = m->pointCount;
    for (u32 i = 0; i < pointCount; ++i) {
        const ManifoldPoint* p = m->points + i;
        WorldManifoldPoint* wp = points + i;
        Vec3 nA = Mul(xfA.rotation, p->localNormal1);
        Vec3 cA = xfA * p->localPoint1;
        Vec3 cB = xfB * p->localPoint2;
        Vec3 pA = cA + rA * nA;
        Vec3 pB = cB - rB * nA;
        wp->point = scalar(0.5) * (pA + pB);
        wp->normal = nA;
        wp->separation = Dot(cB - cA, nA) - rA - rB;
        center += wp->point;
        normal += wp->normal;
    }
    if (pointCount > 0) {
        center /= scalar(pointCount);
        normal.Normalize();
        tangent1 = Perp(normal);
        tangent2 = Cross(tangent1, normal);
    }
}

Vec3 ClosestPointOnSegment(const Vec3& Q, const Vec3& A, const Vec3& B) {
    Vec3 AB = B - A;
    scalar u = Dot(B - Q, AB);
    scalar v = Dot(Q - A, AB);
    if (v <= scalar(0)) return A;
    if (u <= scalar(0)) return B;
    scalar den = Dot(AB, AB);
    if (den < kLinearSlop * kLinearSlop) return A;
    return (u * A + v * B) / den;
}

void ClosestPointsOnSegments(Vec3& C1, Vec3& C2, const Vec3& P1, const Vec3& Q1, const Vec3& P2, const Vec3& Q2) {
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

// ------------------------------------------------------------------ hull shape (hull_shape.cpp)
static void HullMass(const Hull& h, scalar density, MassData* md) {
    Vec3 s; s.SetZero();
    for (u32 i = 0; i < h.vertexCount(); ++i) s += h.vertices[i];
    s /= scalar(h.vertexCount());
    scalar volume = 0;
    Vec3 center; center.SetZero();
    scalar xx = 0, xy = 0, yy = 0, xz = 0, zz = 0, yz = 0;
    for (u32 i = 0; i < h.faceCount(); ++i) {
        const HalfEdge* begin = &h.edges[h.faces[i]];
        const HalfEdge* edge = &h.edges[begin->next];
        do {
            const HalfEdge* next = &h.edges[edge->next];
            Vec3 v1 = h.vertices[begin->origin] - s;
            Vec3 v2 = h.vertices[edge->origin] - s;
            Vec3 v3 = h.vertices[next->origin] - s;
            scalar D = Det(v1, v2, v3);
            volume += D;
            Vec3 v4 = v1 + v2 + v3;
            center += D * v4;
            xx += D * (v1.x * v1.x + v2.x * v2.x + v3.x * v3.x + v4.x * v4.x);
            yy += D * (v1.y * v1.y + v2.y * v2.y + v3.y * v3.y + v4.y * v4.y);
            zz += D * (v1.z * v1.z + v2.z * v2.z + v3.z * v3.z + v4.z * v4.z);
            xy += D * (v1.x * v1.y + v2.x * v2.y + v3.x * v3.y + v4.x * v4.y);
            xz += D * (v1.x * v1.z + v2.x * v2.z + v3.x * v3.z + v4.x * v4.z);
            yz += D * (v1.y * v1.z + v2.y * v2.z + v3.y * v3.z + v4.y * v4.z);
            edge = next;
        } while (&h.edges[edge->next] != begin);
    }
    Mat33 I;
    I.x.x = yy + zz; I.x.y = -xy;    I.x.z = -xz;
    I.y.x = -xy;    I.y.y = xx + zz; I.y.z = -yz;
    I.z.x = -xz;    I.z.y = -yz;    I.z.z = xx + yy;
    md->mass = density * volume / 6.0f;
    center /= scalar(4) * volume;
    md->center = center + s;
    md->I = (density / scalar(120)) * I;
    md->I += md->mass * (Steiner(md->center) - Steiner(center));
}

// ------------------------------------------------------------------ SAT (sat.cpp)
struct FaceQuery { u32 index; scalar separation; };
struct EdgeQuery { u32 index1, index2; scalar separation; };

static inline scalar ProjectHull(const Hull& h, const Plane& plane) {
    Vec3 support = h.vertices[h.GetSupportVertex(-plane.normal)];
    return Distance(support, plane);
}

static FaceQuery QueryFaceSeparation(const Transform& xf1, const Hull& h1, const Transform& xf2, const Hull& h2) {
    Transform xf = MulT(xf2, xf1);
    u32 maxIndex = 0;
    scalar maxSeparation = -kMaxScalar;
    for (u32 i = 0; i < h1.faceCount(); ++i) {
        Plane plane = Mul(xf, h1.planes[i]);
        scalar separation = ProjectHull(h2, plane);
        if (separation > maxSeparation) { maxIndex = i; maxSeparation = separation; }
    }
    FaceQuery out; out.index = maxIndex; out.separation = maxSeparation;
    return out;
}

static inline bool IsMinkowskiFace(const Vec3& A, const Vec3& B, const Vec3& B_x_A, const Vec3& C, const Vec3& D,
                                   const Vec3& D_x_C) {
    scalar ADC = Dot(A, D_x_C);
    scalar BDC = Dot(B, D_x_C);
    scalar CBA = Dot(C, B_x_A);
    scalar DBA = Dot(D, B_x_A);
    return CBA * DBA < scalar(0) && ADC * BDC < scalar(0) && CBA * BDC > scalar(0);
}

static scalar ProjectEdges(const Vec3& P1, const Vec3& Q1, const Vec3& E1, const Vec3& P2, const Vec3& E2, const Vec3& C1) {
    (void)Q1;
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

static EdgeQuery QueryEdgeSeparation(const Transform& xf1, const Hull& h1, const Transform& xf2, const Hull& h2) {
    Transform xf = MulT(xf2, xf1);
    Vec3 C1 = xf * h1.centroid;
    u32 maxIndex1 = 0, maxIndex2 = 0;
    scalar maxSeparation = -kMaxScalar;
    for (u32 i = 0; i < h1.edgeCount(); i += 2) {
        const HalfEdge& edge1 = h1.edges[i];
        const HalfEdge& twin1 = h1.edges[i + 1];
        Vec3 P1 = xf * h1.vertices[edge1.origin];
        Vec3 Q1 = xf * h1.vertices[twin1.origin];
        Vec3 E1 = Q1 - P1;
        Vec3 U1 = Mul(xf.rotation, h1.planes[edge1.face].normal);
        Vec3 V1 = Mul(xf.rotation, h1.planes[twin1.face].normal);
        for (u32 j = 0; j < h2.edgeCount(); j += 2) {
            const HalfEdge& edge2 = h2.edges[j];
            const HalfEdge& twin2 = h2.edges[j + 1];
            Vec3 P2 = h2.vertices[edge2.origin];
            Vec3 Q2 = h2.vertices[twin2.origin];
            Vec3 E2 = Q2 - P2;
            Vec3 U2 = h2.planes[edge2.face].normal;
            Vec3 V2 = h2.planes[twin2.face].normal;
            if (IsMinkowskiFace(U1, V1, -E1, -U2, -V2, -E2)) {
                scalar separation = ProjectEdges(P1, Q1, E1, P2, E2, C1);
                if (separation > maxSeparation) { maxSeparation = separation; maxIndex1 = i; maxIndex2 = j; }
            }
        }
    }
    EdgeQuery out; out.index1 = maxIndex1; out.index2 = maxIndex2; out.separation = maxSeparation;
    return out;
}

static u32 ReadFace(const FeatureCache& c, const Transform& xf1, const Hull& h1, const Transform& xf2, const Hull& h2,
                    scalar totalRadius) {
    Transform xf = MulT(xf2, xf1);
    Plane plane = Mul(xf, h1.planes[c.index1]);
    scalar separation = ProjectHull(h2, plane);
    return separation > totalRadius ? e_separation : e_overlap;
}

static u32 ReadEdges(const FeatureCache& c, const Transform& xf1, const Hull& h1, const Transform& xf2, const Hull& h2,
                     scalar totalRadius) {
    u32 i = c.index1, j = c.index2;
    Transform xf = MulT(xf2, xf1);
    Vec3 C1 = xf * h1.centroid;
    const HalfEdge& edge1 = h1.edges[i];
    const HalfEdge& twin1 = h1.edges[i + 1];
    Vec3 P1 = xf * h1.vertices[edge1.origin];
    Vec3 Q1 = xf * h1.vertices[twin1.origin];
    Vec3 E1 = Q1 - P1;
    Vec3 U1 = Mul(xf.rotation, h1.planes[edge1.face].normal);
    Vec3 V1 = Mul(xf.rotation, h1.planes[twin1.face].normal);
    const HalfEdge& edge2 = h2.edges[j];
    const HalfEdge& twin2 = h2.edges[j + 1];
    Vec3 P2 = h2.vertices[edge2.origin];
    Vec3 Q2 = h2.vertices[twin2.origin];
    Vec3 E2 = Q2 - P2;
    Vec3 U2 = h2.planes[edge2.face].normal;
    Vec3 V2 = h2.planes[twin2.face].normal;
    if (IsMinkowskiFace(U1, V1, -E1, -U2, -V2, -E2)) {
        scalar separation = ProjectEdges(P1, Q1, E1, P2, E2, C1);
        return separation > totalRadius ? e_separation : e_overlap;
    }
    return e_empty;
}

static u32 ReadState(FeatureCache& c, const Transform& xf1, const Hull& h1, const Transform& xf2, const Hull& h2,
                     scalar totalRadius) {
    if (c.cacheType == e_empty) {  // sat.cpp:178-190: pick an arbitrary feature
        c.cacheType = e_separation;
        c.featureType = e_face1;
        c.index1 = 0;
        c.index2 = 0;
    }
    switch (c.featureType) {
    case e_edges: return ReadEdges(c, xf1, h1, xf2, h2, totalRadius);
    case e_face1: return ReadFace(c, xf1, h1, xf2, h2, totalRadius);
    case e_face2: return ReadFace(c, xf2, h2, xf1, h1, totalRadius);
    default: return e_empty;
    }
}

// ------------------------------------------------------------------ clipping (clip.cpp)
typedef std::vector<ClipVertex> ClipPolygon;

static void BuildPolygon(ClipPolygon& out, const Transform& xf, u32 index, const Hull& h) {
    const HalfEdge* begin = &h.edges[h.faces[index]];
    const HalfEdge* edge = begin;
    do {
        const HalfEdge* twin = &h.edges[edge->twin];
        u32 inEdge = edge->prev;
        u32 outEdge = twin->twin;
        ClipVertex v;
        v.position = Mul(xf, h.vertices[edge->origin]);
        v.pair = MakePair(kNullU32, kNullU32, inEdge, outEdge);
        out.push_back(v);
        edge = &h.edges[edge->next];
    } while (edge != begin);
}

static void ClipPolygonToPlane(ClipPolygon& out, const ClipPolygon& in, const ClipPlane& plane) {
    ClipVertex v1 = in.back();
    scalar d1 = Distance(v1.position, plane.plane);
    for (u32 i = 0; i < in.size(); ++i) {
        ClipVertex v2 = in[i];
        scalar d2 = Distance(v2.position, plane.plane);
        if (d1 <= scalar(0) && d2 <= scalar(0)) {
            out.push_back(v2);
        } else if (d1 <= scalar(0) && d2 > scalar(0)) {
            scalar fraction = d1 / (d1 - d2);
            ClipVertex v;
            v.position = v1.position + fraction * (v2.position - v1.position);
            v.pair = MakePair(plane.edge, kNullU32, v1.pair.outEdge2, kNullU32);
            out.push_back(v);
        } else if (d1 > scalar(0) && d2 <= scalar(0)) {
            scalar fraction = d1 / (d1 - d2);
            ClipVertex v;
            v.position = v1.position + fraction * (v2.position - v1.position);
            v.pair = MakePair(kNullU32, plane.edge, kNullU32, v2.pair.inEdge2);
            out.push_back(v);
            out.push_back(v2);
        }
        v1 = v2;
        d1 = d2;
    }
}

static void ClipPolygonToFaceSidePlanes(ClipPolygon& out, const ClipPolygon& in, const Transform& xf, scalar r,
                                        u32 index, const Hull& h) {
    out = in;
    const HalfEdge* begin = &h.edges[h.faces[index]];
    const HalfEdge* edge = begin;
    do {
        const HalfEdge* twin = &h.edges[edge->twin];
        u32 edgeId = twin->twin;
        Plane plane = h.GetEdgeSidePlane(edgeId);
        plane.offset += r;
        ClipPlane cp;
        cp.edge = edgeId;
        cp.plane = Mul(xf, plane);
        ClipPolygon clipped;
        ClipPolygonToPlane(clipped, out, cp);
        out = clipped;
        if (out.empty()) return;
        edge = &h.edges[edge->next];
    } while (edge != begin);
}

// ------------------------------------------------------------------ reduction (cluster.cpp:23-200)

static inline bool IsCCW(const Vec3& A, const Vec3& B, const Vec3& C, const Vec3& N) {
    Vec3 n = Cross(B - A, C - A);
    return Dot(n, N) > scalar(0);
}

static void SortPolygon(ClusterPolygon& out, const ClusterPolygon& in, const Vec3& N) {
    out = in;
    Vec3 A = out[0].position;
    for (u32 i = 1; i < out.size(); ++i) {
        Vec3 B = out[i].position;  // captured once per i, as in the reference
        for (u32 j = i + 1; j < out.size(); ++j) {
            Vec3 C = out[j].position;
            if (!IsCCW(A, B, C, N)) Swap(out[i], out[j]);
        }
    }
}

void ReducePolygon(ClusterPolygon& out, const ClusterPolygon& in, const Vec3& N, u32 initialPoint) {
    if (in.size() <= kMaxManifoldPoints) { SortPolygon(out, in, N); return; }
    std::vector<char> chosen(in.size(), 0);
    out.push_back(in[initialPoint]);
    chosen[initialPoint] = 1;
    {
        Vec3 A = out[0].position;
        u32 index = 0;
        scalar max = -kMaxScalar;
        for (u32 i = 0; i < in.size(); ++i) {
            if (chosen[i]) continue;
            Vec3 d = in[i].position - A;
            scalar dd = Dot(d, d);
            if (dd > max) { max = dd; index = i; }
        }
        if (max < kEpsilon * kEpsilon) return;
        out.push_back(in[index]);
        chosen[index] = 1;
    }
    {
        Vec3 A = out[0].position, B = out[1].position;
        u32 index = 0;
        scalar max = -kMaxScalar;
        for (u32 i = 0; i < in.size(); ++i) {
            if (chosen[i]) continue;
            Vec3 C = in[i].position;
            Vec3 n = Cross(B - A, C - A);
            scalar a2 = Abs(Dot(n, N));
            if (a2 > max) { max = a2; index = i; }
        }
        const scalar kAreaTol = scalar(0.01);
        if (max < scalar(2) * kAreaTol) return;
        Vec3 C = in[index].position;
        if (!IsCCW(A, B, C, N)) Swap(out[0], out[1]);
        out.push_back(in[index]);
        chosen[index] = 1;
    }
    {
        Vec3 A = out[0].position, B = out[1].position;
        u32 index = 0;
        scalar min = kMaxScalar;
        for (u32 i = 0; i < in.size(); ++i) {
            if (chosen[i]) continue;
            Vec3 D = in[i].position;
            Vec3 n = Cross(B - A, D - A);
            scalar sa2 = Dot(n, N);
            if (sa2 < min) { min = sa2; index = i; }
        }
        const scalar kAreaTol = scalar(0.01);
        if (Abs(min) < scalar(2) * kAreaTol) return;
        out.push_back(in[index]);
        chosen[index] = 1;
    }
    ClusterPolygon quad;
    SortPolygon(quad, out, N);
    out = quad;
}

// ------------------------------------------------------------------ hull contacts (collide_hulls.cpp)
static void BuildEdgeContact(Manifold& m, const Transform& xf1, u32 index1, const Hull& h1, const Transform& xf2,
                             u32 index2, const Hull& h2) {
    const HalfEdge& edge1 = h1.edges[index1];
    const HalfEdge& twin1 = h1.edges[index1 + 1];
    Vec3 C1 = xf1 * h1.centroid;
    Vec3 P1 = xf1 * h1.vertices[edge1.origin];
    Vec3 Q1 = xf1 * h1.vertices[twin1.origin];
    Vec3 E1 = Q1 - P1;
    const HalfEdge& edge2 = h2.edges[index2];
    const HalfEdge& twin2 = h2.edges[index2 + 1];
    Vec3 P2 = xf2 * h2.vertices[edge2.origin];
    Vec3 Q2 = xf2 * h2.vertices[twin2.origin];
    Vec3 E2 = Q2 - P2;
    Vec3 N = Cross(E1, E2);
    N.Normalize();
    if (Dot(N, P1 - C1) < scalar(0)) N = -N;
    Vec3 point1, point2;
    ClosestPointsOnSegments(point1, point2, P1, Q1, P2, Q2);
    m.pointCount = 1;
    m.points[0].localNormal1 = MulC(xf1.rotation, N);
    m.points[0].localPoint1 = MulT(xf1, point1);
    m.points[0].localPoint2 = MulT(xf2, point2);
    m.points[0].id = MakeID(MakePair(index1, index1 + 1, index2, index2 + 1));
    m.points[0].edgeContact = true;
}

static void BuildFaceContact(Manifold& m, const Transform& xf1, u32 index1, const Hull& h1, scalar r1,
                             const Transform& xf2, const Hull& h2, scalar r2, bool flipNormal) {
    scalar totalRadius = r1 + r2;
    Plane localPlane1 = h1.planes[index1];
    Plane plane1 = Mul(xf1, localPlane1);
    Vec3 normal1 = MulC(xf2.rotation, plane1.normal);
    ClipPolygon polygon2;
    u32 index2 = h2.GetSupportFace(-normal1);
    BuildPolygon(polygon2, xf2, index2, h2);
    ClipPolygon clipPolygon2;
    ClipPolygonToFaceSidePlanes(clipPolygon2, polygon2, xf1, totalRadius, index1, h1);
    if (clipPolygon2.empty()) return;
    ClusterPolygon polygon1;
    u32 minIndex = 0;
    scalar minSeparation = kMaxScalar;
    for (u32 i = 0; i < clipPolygon2.size(); ++i) {
        ClipVertex v2 = clipPolygon2[i];
        scalar separation = Distance(v2.position, plane1);
        if (separation <= totalRadius) {
            if (separation < minSeparation) { minIndex = u32(polygon1.size()); minSeparation = separation; }
            ClusterVertex v1;
            v1.position = ClosestPointOnPlane(v2.position, plane1);
            v1.clipIndex = i;
            polygon1.push_back(v1);
        }
    }
    if (polygon1.empty()) return;
    Vec3 normal = plane1.normal;
    Vec3 orientedNormal = flipNormal ? -normal : normal;
    ClusterPolygon reduced;
    ReducePolygon(reduced, polygon1, orientedNormal, minIndex);
    u32 pointCount = u32(reduced.size());
    for (u32 i = 0; i < pointCount; ++i) {
        ClipVertex v2 = clipPolygon2[reduced[i].clipIndex];
        Vec3 v1 = ClosestPointOnPlane(v2.position, plane1);
        ManifoldPoint* mp = m.points + i;
        if (flipNormal) {
            FeaturePair pair = MakePair(v2.pair.inEdge2, v2.pair.outEdge2, v2.pair.inEdge1, v2.pair.outEdge1);
            mp->localNormal1 = MulC(xf2.rotation, orientedNormal);
            mp->localPoint1 = MulT(xf2, v2.position);
            mp->localPoint2 = MulT(xf1, v1);
            mp->id = MakeID(pair);
            mp->edgeContact = false;
        } else {
            mp->localNormal1 = MulC(xf1.rotation, orientedNormal);
            mp->localPoint1 = MulT(xf1, v1);
            mp->localPoint2 = MulT(xf2, v2.position);
            mp->id = MakeID(v2.pair);
            mp->edgeContact = false;
        }
    }
    m.pointCount = pointCount;
}

static void RebuildEdgeContact(Manifold& m, const Transform& xf1, u32 index1, const Hull& h1, const Transform& xf2,
                               u32 index2, const Hull& h2) {
    const HalfEdge& edge1 = h1.edges[index1];
    const HalfEdge& twin1 = h1.edges[index1 + 1];
    Vec3 C1 = xf1 * h1.centroid;
    Vec3 P1 = xf1 * h1.vertices[edge1.origin];
    Vec3 Q1 = xf1 * h1.vertices[twin1.origin];
    Vec3 E1 = Q1 - P1;
    const HalfEdge& edge2 = h2.edges[index2];
    const HalfEdge& twin2 = h2.edges[index2 + 1];
    Vec3 P2 = xf2 * h2.vertices[edge2.origin];
    Vec3 Q2 = xf2 * h2.vertices[twin2.origin];
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
        m.pointCount = 1;
        m.points[0].localNormal1 = MulC(xf1.rotation, N);
        m.points[0].localPoint1 = MulT(xf1, point1);
        m.points[0].localPoint2 = MulT(xf2, point2);
        m.points[0].id = MakeID(MakePair(index1, index1 + 1, index2, index2 + 1));
        m.points[0].edgeContact = true;
    }
}

static void RebuildFaceContact(Manifold& m, const Transform& xf1, u32 index1, const Hull& h1, scalar r1,
                               const Transform& xf2, const Hull& h2, scalar r2, bool flipNormal,
                               const Transform& xf01, const Transform& xf02) {
    Quat q01 = xf01.rotation, q1 = xf1.rotation;
    Quat q02 = xf02.rotation, q2 = xf2.rotation;
    Quat dq0 = Conjugate(q01) * q02;
    Quat dq = Conjugate(q1) * q2;
    Quat q = Conjugate(dq0) * dq;
    const scalar kCosTol = scalar(0.995);
    if (Abs(q.s) > kCosTol) BuildFaceContact(m, xf1, index1, h1, r1, xf2, h2, r2, flipNormal);
}

static inline FeatureCache MakeCache(u32 ct, u32 ft, u32 i1, u32 i2) {
    FeatureCache c; c.cacheType = ct; c.featureType = ft; c.index1 = i1; c.index2 = i2; return c;
}

void CollideHulls(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2, const Shape& s2,
                  FeatureCache& cache, const Transform& xf01, const Transform& xf02) {
    const Hull& h1 = *s1.hull;
    const Hull& h2 = *s2.hull;
    scalar r1 = s1.radius, r2 = s2.radius;
    scalar totalRadius = r1 + r2;
    u32 state0 = cache.cacheType;
    u32 state1 = ReadState(cache, xf1, h1, xf2, h2, totalRadius);
    if (state0 == e_separation && state1 == e_separation) return;
    if (state0 == e_overlap && state1 == e_overlap) {
        switch (cache.featureType) {
        case e_edges: RebuildEdgeContact(m, xf1, cache.index1, h1, xf2, cache.index2, h2); break;
        case e_face1: RebuildFaceContact(m, xf1, cache.index1, h1, r1, xf2, h2, r2, false, xf01, xf02); break;
        case e_face2: RebuildFaceContact(m, xf2, cache.index1, h2, r2, xf1, h1, r1, true, xf02, xf01); break;
        default: break;
        }
        if (m.pointCount > 0) return;
    }
    cache.cacheType = e_empty;
    FaceQuery fq1 = QueryFaceSeparation(xf1, h1, xf2, h2);
    if (fq1.separation > totalRadius) { cache = MakeCache(e_separation, e_face1, fq1.index, fq1.index); return; }
    FaceQuery fq2 = QueryFaceSeparation(xf2, h2, xf1, h1);
    if (fq2.separation > totalRadius) { cache = MakeCache(e_separation, e_face2, fq2.index, fq2.index); return; }
    EdgeQuery eq = QueryEdgeSeparation(xf1, h1, xf2, h2);
    if (eq.separation > totalRadius) { cache = MakeCache(e_separation, e_edges, eq.index1, eq.index2); return; }
    const scalar kRelEdgeTol = scalar(0.90);
    const scalar kRelFaceTol = scalar(0.98);
    const scalar kAbsTol = scalar(0.5) * kLinearSlop;
    if (eq.separation > kRelEdgeTol * Max(fq1.separation, fq2.separation) + kAbsTol) {
        BuildEdgeContact(m, xf1, eq.index1, h1, xf2, eq.index2, h2);
        cache = MakeCache(e_overlap, e_edges, eq.index1, eq.index2);
        return;
    } else {
        if (fq2.separation > kRelFaceTol * fq1.separation + kAbsTol) {
            BuildFaceContact(m, xf2, fq2.index, h2, r2, xf1, h1, r1, true);
            if (m.pointCount > 0) { cache = MakeCache(e_overlap, e_face2, fq2.index, fq2.index); return; }
        } else {
            BuildFaceContact(m, xf1, fq1.index, h1, r1, xf2, h2, r2, false);
            if (m.pointCount > 0) { cache = MakeCache(e_overlap, e_face1, fq1.index, fq1.index); return; }
        }
    }
    if (eq.separation > Max(fq1.separation, fq2.separation)) {
        BuildEdgeContact(m, xf1, eq.index1, h1, xf2, eq.index2, h2);
        cache = MakeCache(e_overlap, e_edges, eq.index1, eq.index2);
        return;
    } else {
        if (fq2.separation > fq1.separation) {
            BuildFaceContact(m, xf2, fq2.index, h2, r2, xf1, h1, r1, true);
            if (m.pointCount > 0) { cache = MakeCache(e_overlap, e_face2, fq2.index, fq2.index); return; }
        } else {
            BuildFaceContact(m, xf1, fq1.index, h1, r1, xf2, h2, r2, false);
            if (m.pointCount > 0) { cache = MakeCache(e_overlap, e_face1, fq1.index, fq1.index); return; }
        }
    }
}

// b3CollideHulls without the feature cache (collide_hulls.cpp:189-264), as b3CollideHullAndTriangle
// calls it on the triangle's two-face hull.
void CollideHullsUncached(Manifold& m, const Transform& xf1, const Hull& h1, scalar r1, const Transform& xf2,
                          const Hull& h2, scalar r2) {
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
    } else {
        if (fq2.separation > kRelFaceTol * fq1.separation + kAbsTol)
            BuildFaceContact(m, xf2, fq2.index, h2, r2, xf1, h1, r1, true);
        else
            BuildFaceContact(m, xf1, fq1.index, h1, r1, xf2, h2, r2, false);
    }
    if (m.pointCount > 0) return;  // the heuristic succeeded
    if (eq.separation > Max(fq1.separation, fq2.separation)) {
        BuildEdgeContact(m, xf1, eq.index1, h1, xf2, eq.index2, h2);
    } else {
        if (fq2.separation > fq1.separation)
            BuildFaceContact(m, xf2, fq2.index, h2, r2, xf1, h1, r1, true);
        else
            BuildFaceContact(m, xf1, fq1.index, h1, r1, xf2, h2, r2, false);
    }
}

// ------------------------------------------------------------------ shape dispatch
void ComputeAABB(const Shape& s, const Transform& xf, AABB* out) {
    switch (s.type) {
    case e_hull: {  // hull_shape.cpp:160-164 via aabb.h:53-62
        const Hull& h = *s.hull;
        out->lowerBound = out->upperBound = Mul(xf, h.vertices[0]);
        for (u32 i = 1; i < h.vertexCount(); ++i) {
            Vec3 v = Mul(xf, h.vertices[i]);
            out->lowerBound = Min(out->lowerBound, v);
            out->upperBound = Max(out->upperBound, v);
        }
        out->Extend(s.radius);
        break;
    }
    case e_sphere: {  // sphere_shape.cpp:51-57
        Vec3 center = Mul(xf, s.v1);
        Vec3 r(s.radius, s.radius, s.radius);
        out->lowerBound = center - r;
        out->upperBound = center + r;
        break;
    }
    case e_capsule: {  // capsule_shape.cpp:124-131
        Vec3 c1 = Mul(xf, s.v1);
        Vec3 c2 = Mul(xf, s.v2);
        Vec3 r(s.radius, s.radius, s.radius);
        out->lowerBound = Min(c1, c2) - r;
        out->upperBound = Max(c1, c2) + r;
        break;
    }
    case e_mesh: {  // mesh_shape.cpp:44-50
        void MeshAABB(const Shape& s, const Transform& xf, AABB* out);
        MeshAABB(s, xf, out);
        break;
    }
    case e_triangle: {  // triangle_shape.cpp:75-83: the local bounds, the transform is not applied
        Vec3 lower = Min(s.v1, Min(s.v2, s.v3));
        Vec3 upper = Max(s.v1, Max(s.v2, s.v3));
        Vec3 r(s.radius, s.radius, s.radius);
        out->lowerBound = lower - r;
        out->upperBound = upper + r;
        break;
    }
    default:
        out->lowerBound.SetZero();
        out->upperBound.SetZero();
        break;
    }
}

void ComputeMass(const Shape& s, scalar density, MassData* out) {
    switch (s.type) {
    case e_hull: HullMass(*s.hull, density, out); break;
    case e_sphere: SphereMass(s.v1, s.radius, density, out); break;
    case e_capsule: CapsuleMass(s.v1, s.v2, s.radius, density, out); break;
    case e_triangle:  // triangle_shape.cpp:67-73
        out->center = (s.v1 + s.v2 + s.v3) / scalar(3);
        out->mass = 0;
        out->I.SetZero();
        break;
    default:
        out->mass = 0; out->center.SetZero(); out->I.SetZero();
        break;
    }
}

}  // namespace orc
