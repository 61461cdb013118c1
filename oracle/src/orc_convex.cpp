// orc_convex.cpp -- oracle (test infrastructure): convex narrowphase other than hull-hull, restated
// on the CPU in the reference's binary32 expression order.
//   collide/collide_spheres.cpp:22-52         CollideSpheres
//   collide/collide_capsule_sphere.cpp:23-120 CollideCapsuleAndSphere
//   collide/collide_capsules.cpp:25-123       CollideCapsules (+ parallel test, clipped 2-point case)
//   collide/collide_hull_sphere.cpp:26-97     CollideHullAndSphere
//   collide/collide_hull_capsule.cpp:28-181   CollideHullAndCapsule
//   sat/sat_hull_edge.cpp:22-146              face / edge separation of a hull and a segment
//   clip.cpp:22-30,57-100,153-241             segment building and clipping
//   gjk/gjk.cpp:37-925                        GJK with Voronoi-region simplex solver
//   collision.cpp:100-112                     TestOverlap (sensors)
//   shapes/sphere_shape.cpp:37-49, shapes/capsule_shape.cpp:36-122  mass data
#include "orc_geom.h"

namespace orc {

// ------------------------------------------------------------------ mass data
void SphereMass(const Vec3& c, scalar r, scalar density, MassData* md) {  // sphere_shape.cpp:37-49
    scalar volume = (scalar(4) / scalar(3)) * kPi * r * r * r;
    scalar mass = density * volume;
    scalar I = mass * (scalar(0.4) * r * r + Dot(c, c));
    md->center = c;
    md->mass = mass;
    md->I = Diagonal(I, I, I);
}

static Mat33 RotateToFrameM(const Mat33& I, const Mat33& R) { return R * I * Transpose(R); }  // mat33.h:362-365

void CapsuleMass(const Vec3& A, const Vec3& B, scalar r, scalar density, MassData* md) {  // capsule_shape.cpp:36-122
    Vec3 d = B - A;
    scalar h = Length(d);
    scalar h2 = h * h;
    scalar r2 = r * r;
    scalar r3 = r2 * r;
    Vec3 center = scalar(0.5) * (A + B);
    scalar mass = 0;
    Mat33 I;
    I.SetZero();
    Mat33 rotation;
    rotation.y = (scalar(1) / h) * d;
    rotation.x = Perp(rotation.y);
    rotation.z = Cross(rotation.y, rotation.x);
    {
        scalar cylinderVolume = kPi * r2 * h;
        scalar cylinderMass = density * cylinderVolume;
        scalar Ixx = (scalar(1) / scalar(12)) * cylinderMass * (scalar(3) * r2 + h2);
        scalar Iyy = scalar(0.5) * cylinderMass * r2;
        Mat33 cylinderI = Diagonal(Ixx, Iyy, Ixx);
        cylinderI = RotateToFrameM(cylinderI, rotation);
        cylinderI += cylinderMass * Steiner(center);
        mass += cylinderMass;
        I += cylinderI;
    }
    {
        scalar hemiVolume = (scalar(2) / scalar(3)) * kPi * r3;
        scalar hemiMass = density * hemiVolume;
        scalar Io = (scalar(2) / scalar(5)) * hemiMass * r2;
        scalar coy = (scalar(3) / scalar(8)) * r;
        scalar Iyy = Io - hemiMass * coy * coy;
        scalar ccy = coy + scalar(0.5) * h;
        scalar Ixx = Io + hemiMass * ccy * ccy;
        Mat33 hemiI = Diagonal(Ixx, Iyy, Ixx);
        hemiI = RotateToFrameM(hemiI, rotation);
        hemiI += hemiMass * Steiner(center);
        mass += scalar(2) * hemiMass;
        I += scalar(2) * hemiI;
    }
    md->center = center;
    md->mass = mass;
    md->I = I;
}

// ------------------------------------------------------------------ GJK
void GJKProxy::Set(const Shape& s) {  // gjk.cpp:37-95
    switch (s.type) {
    case e_sphere: buffer[0] = s.v1; count = 1; break;
    case e_capsule: buffer[0] = s.v1; buffer[1] = s.v2; count = 2; break;
    case e_triangle: buffer[0] = s.v1; buffer[1] = s.v2; buffer[2] = s.v3; count = 3; break;
    case e_hull: vertices = s.hull->vertices.data(); count = s.hull->vertexCount(); radius = s.radius; return;
    default: count = 0; break;
    }
    vertices = buffer;
    radius = s.radius;
}

u32 GJKProxy::GetSupportIndex(const Vec3& d) const {  // gjk_proxy.h:72-86: first index of the maximum
    u32 maxIndex = 0;
    scalar maxProjection = Dot(d, vertices[maxIndex]);
    for (u32 i = 1; i < count; ++i) {
        scalar projection = Dot(d, vertices[i]);
        if (projection > maxProjection) { maxIndex = i; maxProjection = projection; }
    }
    return maxIndex;
}

namespace {

struct SimplexVertex {
    Vec3 point1, point2, point;
    scalar weight;
    u32 index1, index2;
};

// Barycentric coordinates of Q on a segment, triangle and tetrahedron (gjk.cpp:104-173); the last
// entry is the divisor.
inline void Bary2(scalar out[3], const Vec3& A, const Vec3& B, const Vec3& Q) {
    Vec3 AB = B - A;
    Vec3 QA = A - Q;
    Vec3 QB = B - Q;
    out[0] = Dot(QB, AB);
    out[1] = -Dot(QA, AB);
    out[2] = out[0] + out[1];
}

inline void Bary3(scalar out[4], const Vec3& A, const Vec3& B, const Vec3& C, const Vec3& Q) {
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

inline void Bary4(scalar out[5], const Vec3& A, const Vec3& B, const Vec3& C, const Vec3& D, const Vec3& Q) {
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

struct Simplex {
    SimplexVertex v[4];
    u32 count;

    void Keep1(const SimplexVertex& a) {
        count = 1;
        v[0] = a;
        v[0].weight = scalar(1);
    }
    void Keep2(const SimplexVertex& a, const SimplexVertex& b, scalar wa, scalar wb, scalar divisor) {
        scalar s = scalar(1) / divisor;
        count = 2;
        v[0] = a;
        v[0].weight = s * wa;
        v[1] = b;
        v[1].weight = s * wb;
    }
    void Keep3(const SimplexVertex& a, const SimplexVertex& b, const SimplexVertex& c, scalar wa, scalar wb,
               scalar wc, scalar divisor) {
        scalar s = scalar(1) / divisor;
        count = 3;
        v[0] = a;
        v[0].weight = s * wa;
        v[1] = b;
        v[1].weight = s * wb;
        v[2] = c;
        v[2].weight = s * wc;
    }

    Vec3 SearchDirection(const Vec3& Q) const {  // gjk.cpp:175-216
        switch (count) {
        case 1: return Q - v[0].point;
        case 2: {
            Vec3 AB = v[1].point - v[0].point;
            Vec3 AQ = Q - v[0].point;
            return Cross(Cross(AB, AQ), AB);
        }
        case 3: {
            Vec3 AB = v[1].point - v[0].point;
            Vec3 AC = v[2].point - v[0].point;
            Vec3 AQ = Q - v[0].point;
            Vec3 AB_x_AC = Cross(AB, AC);
            scalar sign = Dot(AB_x_AC, AQ);
            if (sign > scalar(0)) return AB_x_AC;
            return -AB_x_AC;
        }
        default: return Vec3(0, 0, 0);
        }
    }

    void ClosestPoints(Vec3* p1, Vec3* p2) const {  // gjk.cpp:238-268
        switch (count) {
        case 1:
            *p1 = v[0].point1;
            *p2 = v[0].point2;
            break;
        case 2:
            *p1 = v[0].weight * v[0].point1 + v[1].weight * v[1].point1;
            *p2 = v[0].weight * v[0].point2 + v[1].weight * v[1].point2;
            break;
        case 3:
            *p1 = v[0].weight * v[0].point1 + v[1].weight * v[1].point1 + v[2].weight * v[2].point1;
            *p2 = v[0].weight * v[0].point2 + v[1].weight * v[1].point2 + v[2].weight * v[2].point2;
            break;
        case 4:
            *p1 = v[0].weight * v[0].point1 + v[1].weight * v[1].point1 + v[2].weight * v[2].point1 +
                  v[3].weight * v[3].point1;
            *p2 = *p1;
            break;
        default: break;
        }
    }

    void Solve2(const Vec3& Q) {  // gjk.cpp:271-301
        SimplexVertex A = v[0], B = v[1];
        scalar wAB[3];
        Bary2(wAB, A.point, B.point, Q);
        if (wAB[1] <= scalar(0)) { Keep1(A); return; }
        if (wAB[0] <= scalar(0)) { Keep1(B); return; }
        Keep2(A, B, wAB[0], wAB[1], wAB[2]);
    }

    void Solve3(const Vec3& Q) {  // gjk.cpp:305-414
        SimplexVertex A = v[0], B = v[1], C = v[2];
        scalar wAB[3], wBC[3], wCA[3];
        Bary2(wAB, A.point, B.point, Q);
        Bary2(wBC, B.point, C.point, Q);
        Bary2(wCA, C.point, A.point, Q);
        if (wAB[1] <= scalar(0) && wCA[0] <= scalar(0)) { Keep1(A); return; }
        if (wAB[0] <= scalar(0) && wBC[1] <= scalar(0)) { Keep1(B); return; }
        if (wBC[0] <= scalar(0) && wCA[1] <= scalar(0)) { Keep1(C); return; }
        scalar wABC[4];
        Bary3(wABC, A.point, B.point, C.point, Q);
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
        if (divisor <= scalar(0)) return;  // degenerate: keep the simplex as it is
        Keep3(A, B, C, wABC[0], wABC[1], wABC[2], divisor);
    }

    void Solve4(const Vec3& Q) {  // gjk.cpp:417-644
        SimplexVertex A = v[0], B = v[1], C = v[2], D = v[3];
        scalar wAB[3], wAC[3], wAD[3], wBC[3], wCD[3], wDB[3];
        Bary2(wAB, A.point, B.point, Q);
        Bary2(wBC, B.point, C.point, Q);
        Bary2(wAC, A.point, C.point, Q);
        Bary2(wAD, A.point, D.point, Q);
        Bary2(wCD, C.point, D.point, Q);
        Bary2(wDB, D.point, B.point, Q);
        if (wAB[1] <= scalar(0) && wAC[1] <= scalar(0) && wAD[1] <= scalar(0)) { Keep1(A); return; }
        if (wAB[0] <= scalar(0) && wDB[0] <= scalar(0) && wBC[1] <= scalar(0)) { Keep1(B); return; }
        if (wAC[0] <= scalar(0) && wBC[0] <= scalar(0) && wCD[1] <= scalar(0)) { Keep1(C); return; }
        if (wAD[0] <= scalar(0) && wCD[0] <= scalar(0) && wDB[1] <= scalar(0)) { Keep1(D); return; }
        scalar wACB[4], wABD[4], wADC[4], wBCD[4];
        Bary3(wACB, A.point, C.point, B.point, Q);
        Bary3(wABD, A.point, B.point, D.point, Q);
        Bary3(wADC, A.point, D.point, C.point, Q);
        Bary3(wBCD, B.point, C.point, D.point, Q);
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
        Bary4(wABCD, A.point, B.point, C.point, D.point, Q);
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
        v[0].weight = s * wABCD[0];
        v[1].weight = s * wABCD[1];
        v[2].weight = s * wABCD[2];
        v[3].weight = s * wABCD[3];
    }
};

}  // namespace

// b3Simplex::GetMetric (gjk.cpp:881-917): length, area or volume of the simplex.
static scalar SimplexMetric(const Simplex& s) {
    switch (s.count) {
    case 2: return Distance(s.v[0].point, s.v[1].point);
    case 3: return Length(Cross(s.v[1].point - s.v[0].point, s.v[2].point - s.v[0].point));
    case 4: {
        scalar det = Det(s.v[1].point - s.v[0].point, s.v[2].point - s.v[0].point, s.v[3].point - s.v[0].point);
        return Sign(det) * det;
    }
    default: return scalar(0);
    }
}

GJKOutput GJK(const Transform& xf1, const GJKProxy& proxy1, const Transform& xf2, const GJKProxy& proxy2,
              bool applyRadius) {  // an empty simplex cache (gjk.cpp:911-917)
    SimplexCache cache;
    cache.count = 0;
    return GJK(xf1, proxy1, xf2, proxy2, applyRadius, &cache);
}

GJKOutput GJK(const Transform& xf1, const GJKProxy& proxy1, const Transform& xf2, const GJKProxy& proxy2,
              bool applyRadius, SimplexCache* cache) {  // gjk.cpp:651-812
    Simplex simplex;
    // b3Simplex::ReadCache (gjk.cpp:817-868): the cached vertices at the new transforms, dropped when
    // their metric changed by more than a factor of two; an empty simplex starts from the first vertex
    // of each proxy
    simplex.count = cache->count;
    for (u32 i = 0; i < simplex.count; ++i) {
        SimplexVertex& v = simplex.v[i];
        v.index1 = cache->index1[i];
        v.index2 = cache->index2[i];
        v.point1 = xf1 * proxy1.vertices[v.index1];
        v.point2 = xf2 * proxy2.vertices[v.index2];
        v.point = v.point2 - v.point1;
        v.weight = scalar(0);
    }
    if (simplex.count > 1) {
        scalar metric1 = cache->metric;
        scalar metric2 = SimplexMetric(simplex);
        if (metric2 < scalar(0.5) * metric1 || scalar(2) * metric1 < metric2 || metric2 < kEpsilon) simplex.count = 0;
    }
    if (simplex.count == 0) {
        simplex.count = 1;
        simplex.v[0].point1 = Mul(xf1, proxy1.vertices[0]);
        simplex.v[0].point2 = Mul(xf2, proxy2.vertices[0]);
        simplex.v[0].point = simplex.v[0].point2 - simplex.v[0].point1;
        simplex.v[0].weight = scalar(1);
        simplex.v[0].index1 = 0;
        simplex.v[0].index2 = 0;
    }
    u32 save1[4], save2[4];
    const Vec3 origin(scalar(0), scalar(0), scalar(0));
    const u32 kMaxIters = 20;
    u32 iter = 0;
    while (iter < kMaxIters) {
        u32 saveCount = simplex.count;
        for (u32 i = 0; i < saveCount; ++i) {
            save1[i] = simplex.v[i].index1;
            save2[i] = simplex.v[i].index2;
        }
        switch (simplex.count) {
        case 2: simplex.Solve2(origin); break;
        case 3: simplex.Solve3(origin); break;
        case 4: simplex.Solve4(origin); break;
        default: break;
        }
        if (simplex.count == 4) break;
        Vec3 d = simplex.SearchDirection(origin);
        if (Dot(d, d) < kEpsilon * kEpsilon) break;
        SimplexVertex* vertex = simplex.v + simplex.count;
        vertex->index1 = proxy1.GetSupportIndex(MulC(xf1.rotation, -d));
        vertex->point1 = Mul(xf1, proxy1.vertices[vertex->index1]);
        vertex->index2 = proxy2.GetSupportIndex(MulC(xf2.rotation, d));
        vertex->point2 = Mul(xf2, proxy2.vertices[vertex->index2]);
        vertex->point = vertex->point2 - vertex->point1;
        ++iter;
        bool duplicate = false;
        for (u32 i = 0; i < saveCount; ++i) {
            if (vertex->index1 == save1[i] && vertex->index2 == save2[i]) { duplicate = true; break; }
        }
        if (duplicate) break;
        ++simplex.count;
    }
    GJKOutput out;
    simplex.ClosestPoints(&out.point1, &out.point2);
    out.distance = Distance(out.point1, out.point2);
    out.iterations = iter;
    out.simplexCount = simplex.count;
    for (u32 i = 0; i < simplex.count; ++i) {
        out.simplexIndex1[i] = simplex.v[i].index1;
        out.simplexIndex2[i] = simplex.v[i].index2;
    }
    // b3Simplex::WriteCache (gjk.cpp:870-879)
    cache->metric = SimplexMetric(simplex);
    cache->count = simplex.count;
    for (u32 i = 0; i < simplex.count; ++i) {
        cache->index1[i] = simplex.v[i].index1;
        cache->index2[i] = simplex.v[i].index2;
    }
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

bool ConvexTestOverlap(const Transform& xfA, const Shape& sA, const Transform& xfB, const Shape& sB) {
    // collision.cpp:100-112 via convex_contact.cpp:34-43
    GJKProxy pA, pB;
    pA.Set(sA);
    pB.Set(sB);
    GJKOutput q = GJK(xfA, pA, xfB, pB, true);
    const scalar kTol = scalar(10) * kEpsilon;
    return q.distance <= kTol;
}

// ------------------------------------------------------------------ segment clipping (clip.cpp)
static void BuildSegment(ClipVertex out[2], const Vec3& v1, const Vec3& v2) {  // clip.cpp:22-30
    out[0].position = v1;
    out[0].pair = MakePair(kNullU32, kNullU32, 0, 0);
    out[1].position = v2;
    out[1].pair = MakePair(kNullU32, kNullU32, 1, 0);
}

static u32 ClipSegmentToPlane(ClipVertex out[2], const ClipVertex in[2], const ClipPlane& plane) {  // clip.cpp:57-100
    u32 n = 0;
    scalar d1 = Distance(in[0].position, plane.plane);
    scalar d2 = Distance(in[1].position, plane.plane);
    if (d1 <= scalar(0)) out[n++] = in[0];
    if (d2 <= scalar(0)) out[n++] = in[1];
    if (d1 <= scalar(0) && d2 > scalar(0)) {
        scalar fraction = d1 / (d1 - d2);
        out[n].position = in[0].position + fraction * (in[1].position - in[0].position);
        out[n].pair = MakePair(plane.edge, kNullU32, in[0].pair.outEdge2, kNullU32);
        ++n;
    } else if (d1 > scalar(0) && d2 <= scalar(0)) {
        scalar fraction = d1 / (d1 - d2);
        out[n].position = in[0].position + fraction * (in[1].position - in[0].position);
        out[n].pair = MakePair(kNullU32, plane.edge, kNullU32, in[1].pair.inEdge2);
        ++n;
    }
    return n;
}

static u32 ClipSegmentToSegmentSidePlanes(ClipVertex out[2], const ClipVertex in[2], const Vec3& P1,
                                          const Vec3& Q1) {  // clip.cpp:153-199
    out[0] = in[0];
    out[1] = in[1];
    Vec3 E1 = Q1 - P1;
    ClipPlane plane1;
    plane1.plane.normal = Normalize(E1);
    plane1.plane.offset = Dot(plane1.plane.normal, Q1);
    plane1.edge = 0;
    ClipVertex tmp[2];
    u32 n = ClipSegmentToPlane(tmp, out, plane1);
    out[0] = tmp[0];
    out[1] = tmp[1];
    if (n < 2) return n;
    ClipPlane plane2;
    plane2.plane.normal = -plane1.plane.normal;
    plane2.plane.offset = Dot(plane2.plane.normal, P1);
    plane2.edge = 1;
    n = ClipSegmentToPlane(tmp, out, plane2);
    out[0] = tmp[0];
    out[1] = tmp[1];
    return n;
}

static u32 ClipSegmentToFaceSidePlanes(ClipVertex out[2], const ClipVertex in[2], const Transform& xf, scalar r,
                                       u32 index, const Hull& h) {  // clip.cpp:202-241
    out[0] = in[0];
    out[1] = in[1];
    u32 n = 0;
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
        ClipVertex tmp[2];
        n = ClipSegmentToPlane(tmp, out, cp);
        out[0] = tmp[0];
        out[1] = tmp[1];
        if (n == 0) return n;
        edge = &h.edges[edge->next];
    } while (edge != begin);
    return n;
}

// ------------------------------------------------------------------ sphere / capsule colliders
static inline void OnePoint(Manifold& m, const Vec3& n1, const Vec3& p1, const Vec3& p2) {
    m.pointCount = 1;
    m.points[0].localNormal1 = n1;
    m.points[0].localPoint1 = p1;
    m.points[0].localPoint2 = p2;
    m.points[0].id = MakeID(0, 0);
}

void CollideSpheres(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2, const Shape& s2) {
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

void CollideCapsuleAndSphere(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2,
                             const Shape& s2) {
    Vec3 Q = MulT(xf1, Mul(xf2, s2.v1));
    Vec3 A = s1.v1, B = s1.v2;
    Vec3 AB = B - A;
    scalar u = Dot(B - Q, AB);
    scalar v = Dot(Q - A, AB);
    scalar radius = s1.radius + s2.radius;
    if (v <= scalar(0) || u <= scalar(0)) {  // vertex regions A, then B
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

static bool AreParallel(const Vec3& P1, const Vec3& Q1, const Vec3& P2, const Vec3& Q2) {  // collide_capsules.cpp:25-44
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

void CollideCapsules(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2, const Shape& s2) {
    Vec3 P1 = Mul(xf1, s1.v1), Q1 = Mul(xf1, s1.v2);
    Vec3 P2 = Mul(xf2, s2.v1), Q2 = Mul(xf2, s2.v2);
    Vec3 point1, point2;
    ClosestPointsOnSegments(point1, point2, P1, Q1, P2, Q2);
    scalar distance = Distance(point1, point2);
    scalar r1 = s1.radius, r2 = s2.radius;
    scalar totalRadius = r1 + r2;
    if (distance > totalRadius) return;
    if (distance > scalar(0) && AreParallel(P1, Q1, P2, Q2)) {
        ClipVertex in1[2];
        BuildSegment(in1, P1, Q1);
        ClipVertex clip1[2];
        u32 clipCount = ClipSegmentToSegmentSidePlanes(clip1, in1, P2, Q2);
        if (clipCount == 2) {
            Vec3 cp1 = ClosestPointOnSegment(clip1[0].position, P2, Q2);
            Vec3 cp2 = ClosestPointOnSegment(clip1[1].position, P2, Q2);
            scalar d1 = Distance(clip1[0].position, cp1);
            scalar d2 = Distance(clip1[1].position, cp2);
            if (d1 > kEpsilon && d1 <= totalRadius && d2 > kEpsilon && d2 <= totalRadius) {
                Vec3 n1 = (cp1 - clip1[0].position) / d1;
                Vec3 n2 = (cp2 - clip1[1].position) / d2;
                m.pointCount = 2;
                m.points[0].localNormal1 = MulC(xf1.rotation, n1);
                m.points[0].localPoint1 = MulT(xf1, clip1[0].position);
                m.points[0].localPoint2 = MulT(xf2, cp1);
                m.points[0].id = MakeID(clip1[0].pair);
                m.points[1].localNormal1 = MulC(xf1.rotation, n2);
                m.points[1].localPoint1 = MulT(xf1, clip1[1].position);
                m.points[1].localPoint2 = MulT(xf2, cp2);
                m.points[1].id = MakeID(clip1[1].pair);
                return;
            }
        }
    }
    if (distance > scalar(0)) {
        Vec3 normal = (point2 - point1) / distance;
        OnePoint(m, MulC(xf1.rotation, normal), MulT(xf1, point1), MulT(xf2, point2));
    }
}

// ------------------------------------------------------------------ hull vs sphere / capsule
void CollideHullAndSphere(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2,
                          const Shape& s2) {
    scalar radius = s1.radius + s2.radius;
    const Hull& h1 = *s1.hull;
    Vec3 cLocal = MulT(xf1, Mul(xf2, s2.v1));
    u32 faceIndex = 0;
    scalar separation = -kMaxScalar;
    for (u32 i = 0; i < h1.faceCount(); ++i) {
        scalar s = Distance(cLocal, h1.planes[i]);
        if (s > radius) return;
        if (s > separation) { faceIndex = i; separation = s; }
    }
    if (separation < scalar(0)) {
        Plane localPlane1 = h1.planes[faceIndex];
        OnePoint(m, localPlane1.normal, ClosestPointOnPlane(cLocal, localPlane1), s2.v1);
        return;
    }
    GJKProxy p1, p2;
    p1.Set(s1);
    p2.Set(s2);
    GJKOutput q = GJK(xf1, p1, xf2, p2, false);
    if (q.distance > radius) return;
    if (q.distance > scalar(0)) {
        Vec3 normal = (q.point2 - q.point1) / q.distance;
        OnePoint(m, MulC(xf1.rotation, normal), MulT(xf1, q.point1), s2.v1);
    }
}

namespace {
struct SegFaceQuery { u32 index; scalar separation; };
struct SegEdgeQuery { u32 index1; scalar separation; };
}  // namespace

static SegFaceQuery QueryFaceSeparationSeg(const Transform& xf1, const Hull& h1, const Transform& xf2,
                                           const Vec3& v1, const Vec3& v2) {  // sat_hull_edge.cpp:28-55
    Transform xf = MulT(xf1, xf2);
    Vec3 a = Mul(xf, v1), b = Mul(xf, v2);
    u32 maxIndex = 0;
    scalar maxSeparation = -kMaxScalar;
    for (u32 i = 0; i < h1.faceCount(); ++i) {
        const Plane& plane = h1.planes[i];
        Vec3 nd = -plane.normal;
        Vec3 support = Dot(nd, a) > Dot(nd, b) ? a : b;  // b3Capsule::GetSupportVertex (capsule.h:47-54)
        scalar separation = Distance(support, plane);
        if (separation > maxSeparation) { maxIndex = i; maxSeparation = separation; }
    }
    return SegFaceQuery{maxIndex, maxSeparation};
}

static scalar ProjectEdgeSeg(const Vec3& P1, const Vec3& E1, const Vec3& C1, const Vec3& P2, const Vec3& E2) {
    // sat_hull_edge.cpp:63-96
    scalar L1 = Length(E1);
    if (L1 < kLinearSlop) return -kMaxScalar;
    scalar L2 = Length(E2);
    if (L2 < kLinearSlop) return -kMaxScalar;
    const scalar kTol = 0.005f;
    Vec3 E1_x_E2 = Cross(E1, E2);
    scalar L = Length(E1_x_E2);
    if (L < kTol * L1 * L2) return -kMaxScalar;
    Vec3 N = (scalar(1) / L) * E1_x_E2;
    if (Dot(N, P1 - C1) < scalar(0)) N = -N;
    return Dot(N, P2 - P1);
}

static SegEdgeQuery QueryEdgeSeparationSeg(const Transform& xf1, const Hull& h1, const Transform& xf2,
                                           const Vec3& v1, const Vec3& v2) {  // sat_hull_edge.cpp:98-146
    u32 maxIndex = 0;
    scalar maxSeparation = -kMaxScalar;
    Transform xf = MulT(xf1, xf2);
    Vec3 C1 = h1.centroid;
    Vec3 P2 = Mul(xf, v1);
    Vec3 Q2 = Mul(xf, v2);
    Vec3 E2 = Q2 - P2;
    for (u32 i = 0; i < h1.edgeCount(); i += 2) {
        const HalfEdge& e1 = h1.edges[i];
        const HalfEdge& t1 = h1.edges[i + 1];
        Vec3 P1 = h1.vertices[e1.origin];
        Vec3 Q1 = h1.vertices[t1.origin];
        Vec3 E1 = Q1 - P1;
        Vec3 U1 = h1.planes[e1.face].normal;
        Vec3 V1 = h1.planes[t1.face].normal;
        if (Dot(E2, U1) * Dot(E2, V1) < scalar(0)) {  // b3IsMinkowskiFace
            scalar separation = ProjectEdgeSeg(P1, E1, C1, P2, E2);
            if (separation > maxSeparation) { maxSeparation = separation; maxIndex = i; }
        }
    }
    return SegEdgeQuery{maxIndex, maxSeparation};
}

static void BuildCapsuleEdgeContact(Manifold& m, const Transform& xf1, const Hull& h1, u32 index1,
                                    const Transform& xf2, const Vec3& v1, const Vec3& v2) {  // collide_hull_capsule.cpp:28-63
    const HalfEdge& e1 = h1.edges[index1];
    const HalfEdge& t1 = h1.edges[index1 + 1];
    Vec3 C1 = Mul(xf1, h1.centroid);
    Vec3 P1 = Mul(xf1, h1.vertices[e1.origin]);
    Vec3 Q1 = Mul(xf1, h1.vertices[t1.origin]);
    Vec3 E1 = Q1 - P1;
    Vec3 P2 = Mul(xf2, v1);
    Vec3 Q2 = Mul(xf2, v2);
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
    m.points[0].id = MakeID(MakePair(index1, index1 + 1, 0, 1));
}

static void BuildCapsuleFaceContact(Manifold& m, const Transform& xf1, const Hull& h1, u32 index1, scalar r1,
                                    const Transform& xf2, const Vec3& v1, const Vec3& v2,
                                    scalar r2) {  // collide_hull_capsule.cpp:65-105
    ClipVertex seg2[2];
    BuildSegment(seg2, Mul(xf2, v1), Mul(xf2, v2));
    ClipVertex clip2[2];
    u32 clipCount = ClipSegmentToFaceSidePlanes(clip2, seg2, xf1, r1, index1, h1);
    Plane localPlane1 = h1.planes[index1];
    Plane plane1 = Mul(xf1, localPlane1);
    scalar totalRadius = r1 + r2;
    u32 pointCount = 0;
    for (u32 i = 0; i < clipCount; ++i) {
        Vec3 c2 = clip2[i].position;
        scalar s = Distance(c2, plane1);
        if (s <= totalRadius) {
            Vec3 c1 = ClosestPointOnPlane(c2, plane1);
            ManifoldPoint* mp = m.points + pointCount;
            mp->localNormal1 = localPlane1.normal;
            mp->localPoint1 = MulT(xf1, c1);
            mp->localPoint2 = MulT(xf2, c2);
            mp->id = MakeID(clip2[i].pair);
            ++pointCount;
        }
    }
    m.pointCount = pointCount;
}

// b3CollideHullAndCapsule (collide_hull_capsule.cpp:107-181) for a convex shape s1 whose polyhedron is
// h1. b3CollideTriangleAndCapsule (collide_triangle_capsule.cpp:107-278) is the same sequence on the
// triangle's two-face hull (b3TriangleHull) when the triangle has no edge wings.
// b3UniqueCount (math.h:84-104)
static u32 UniqueCount(const u32* v, u32 n) {
    u32 count = 0;
    for (u32 i = 0; i < n; ++i) {
        u32 j;
        for (j = 0; j < n; ++j)
            if (v[i] == v[j]) break;
        if (i == j) ++count;
    }
    return count;
}

// The proxy-1 side of b3GetFeaturePair (gjk_feature_pair.cpp:21-208) for the final simplex of q:
// returns count1 and, when it is 2, the two vertex indices of the edge on proxy 1.
static u32 FeatureCount1(const GJKOutput& q, u32& e0, u32& e1) {
    u32 n = q.simplexCount;
    const u32* i1 = q.simplexIndex1;
    const u32* i2 = q.simplexIndex2;
    u32 u1 = UniqueCount(i1, n), u2 = UniqueCount(i2, n);
    if (n == 1) return 1;
    if (n == 2) {
        if ((u1 == 2 && u2 == 2) || (u1 != 1 && u2 == 1)) { e0 = i1[0]; e1 = i1[1]; return 2; }
        if (u1 == 1) return 1;
        return 0;
    }
    if (n == 3) {
        if (u1 == 1 && u2 == 3) return 1;
        if (u1 == 3 && u2 == 1) return 3;
        if (u1 == 2 && (u2 == 2 || u2 == 3)) {
            e0 = i1[0];
            e1 = i1[0] == i1[1] ? i1[2] : i1[1];
            return 2;
        }
        if (u1 == 3) return 3;
    }
    return 0;
}

void CollideConvexAndCapsule(Manifold& m, const Transform& xf1, const Shape& s1, const Hull& h1, const Transform& xf2,
                             const Shape& s2) {
    scalar r1 = s1.radius, r2 = s2.radius;
    scalar totalRadius = r1 + r2;
    GJKProxy p1, p2;
    p1.Set(s1);
    p2.Set(s2);
    GJKOutput q = GJK(xf1, p1, xf2, p2, false);
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
            BuildCapsuleFaceContact(m, xf1, h1, index1, r1, xf2, s2.v1, s2.v2, r2);
            if (m.pointCount == 2) return;
        }
        if (s1.type == e_triangle && (s1.hasE1 || s1.hasE2 || s1.hasE3)) {
            // collide_triangle_capsule.cpp:162-241: the closest feature is a triangle edge with a wing
            u32 v1 = 0, v2 = 0;
            if (FeatureCount1(q, v1, v2) == 2) {
                const Vec3 vertices[3] = {s1.v1, s1.v2, s1.v3};
                const bool hasWing[3] = {s1.hasE1, s1.hasE2, s1.hasE3};
                const Vec3 wings[3] = {s1.e1, s1.e2, s1.e3};
                u32 edgeIndex = 0;
                for (u32 i = 0; i < 3; ++i) {
                    u32 j = i + 1 < 3 ? i + 1 : 0;
                    if ((v1 == i && v2 == j) || (v2 == i && v1 == j)) { edgeIndex = i; break; }
                }
                if (hasWing[edgeIndex]) {
                    u32 ev1 = edgeIndex;
                    u32 ev2 = edgeIndex + 1 < 3 ? edgeIndex + 1 : 0;
                    Vec3 Q = MulT(xf1, c2);
                    Vec3 A = wings[edgeIndex], B = vertices[ev2], C = vertices[ev1];
                    scalar wABC[4];
                    Barycentric3(wABC, A, B, C, Q);
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
        }
        m.pointCount = 1;
        m.points[0].localNormal1 = localN1;
        m.points[0].localPoint1 = MulT(xf1, c1);
        m.points[0].localPoint2 = MulT(xf2, c2);
        m.points[0].id = MakeID(0, 0);
        return;
    }
    SegFaceQuery fq = QueryFaceSeparationSeg(xf1, h1, xf2, s2.v1, s2.v2);
    if (fq.separation > totalRadius) return;
    SegEdgeQuery eq = QueryEdgeSeparationSeg(xf1, h1, xf2, s2.v1, s2.v2);
    if (eq.separation > totalRadius) return;
    const scalar kRelEdgeTol = scalar(0.90);
    const scalar kAbsTol = scalar(0.5) * kLinearSlop;
    if (eq.separation > kRelEdgeTol * fq.separation + kAbsTol)
        BuildCapsuleEdgeContact(m, xf1, h1, eq.index1, xf2, s2.v1, s2.v2);
    else
        BuildCapsuleFaceContact(m, xf1, h1, fq.index, r1, xf2, s2.v1, s2.v2, r2);
}

void CollideHullAndCapsule(Manifold& m, const Transform& xf1, const Shape& s1, const Transform& xf2,
                           const Shape& s2) {  // collide_hull_capsule.cpp:107-181
    CollideConvexAndCapsule(m, xf1, s1, *s1.hull, xf2, s2);
}

}  // namespace orc
