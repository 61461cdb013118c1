// Time of impact (src/bounce/collision/time_of_impact.cpp), restated for the oracle: the local
// separating-axis method over two sweeps, conservative advancement of two translating shapes, and the
// interval test of two moving boxes. Test infrastructure: the device's bx_toi.h is checked against
// these, and these against the reference build (oracle/_ref).
#include "orc_geom.h"

namespace orc {

namespace {

// b3GJKFeaturePair (gjk_feature_pair.cpp:21-236): the distinct vertices of each proxy in a simplex cache.
struct GJKFeatures {
    u32 count1, count2;
    u32 index1[3], index2[3];
};

u32 Unique(const u32* v, u32 n) {  // b3UniqueCount (math.h:84-104)
    u32 count = 0;
    for (u32 i = 0; i < n; ++i) {
        u32 j;
        for (j = 0; j < n; ++j)
            if (v[i] == v[j]) break;
        if (i == j) ++count;
    }
    return count;
}

// The second distinct index of a three-vertex simplex side holding two distinct indices.
u32 SecondOf3(const u32* v) { return v[0] == v[1] ? v[2] : v[1]; }

GJKFeatures GetFeaturePair(const SimplexCache& c) {
    GJKFeatures p;
    p.count1 = p.count2 = 0;
    const u32 n = c.count;
    const u32 u1 = Unique(c.index1, n), u2 = Unique(c.index2, n);
    auto take = [](u32* out, const u32* in, u32 k) {
        for (u32 i = 0; i < k; ++i) out[i] = in[i];
    };
    if (n == 1) {
        p.count1 = p.count2 = 1;
        p.index1[0] = c.index1[0];
        p.index2[0] = c.index2[0];
    } else if (n == 2) {
        if (u1 == 2 && u2 == 2) {
            p.count1 = p.count2 = 2;
            take(p.index1, c.index1, 2);
            take(p.index2, c.index2, 2);
        } else if (u1 == 1) {
            p.count1 = 1;
            p.count2 = 2;
            p.index1[0] = c.index1[0];
            take(p.index2, c.index2, 2);
        } else if (u2 == 1) {
            p.count1 = 2;
            p.count2 = 1;
            take(p.index1, c.index1, 2);
            p.index2[0] = c.index2[0];
        }
    } else if (n == 3) {
        if (u1 == 1 && u2 == 3) {
            p.count1 = 1;
            p.count2 = 3;
            p.index1[0] = c.index1[0];
            take(p.index2, c.index2, 3);
        } else if (u1 == 3 && u2 == 1) {
            p.count1 = 3;
            p.count2 = 1;
            take(p.index1, c.index1, 3);
            p.index2[0] = c.index2[0];
        } else if (u1 == 2 && u2 == 2) {
            p.count1 = p.count2 = 2;
            p.index1[0] = c.index1[0];
            p.index1[1] = SecondOf3(c.index1);
            p.index2[0] = c.index2[0];
            p.index2[1] = SecondOf3(c.index2);
        } else if (u1 == 2 && u2 == 3) {
            p.count1 = 2;
            p.count2 = 3;
            p.index1[0] = c.index1[0];
            p.index1[1] = SecondOf3(c.index1);
            take(p.index2, c.index2, 3);
        } else if (u1 == 3 && u2 == 2) {
            p.count1 = 3;
            p.count2 = 2;
            take(p.index1, c.index1, 3);
            p.index2[0] = c.index2[0];
            p.index2[1] = SecondOf3(c.index2);
        } else if (u1 == 3 && u2 == 3) {
            p.count1 = p.count2 = 3;
            take(p.index1, c.index1, 3);
            take(p.index2, c.index2, 3);
        }
    }
    return p;
}

// b3SeparationFunction (time_of_impact.cpp:27-458): the separating axis found at t1 from the closest
// features, and the separation along it at any t.
struct SeparationFunction {
    enum Type { e_points, e_edgeA, e_faceA, e_edgeB, e_faceB, e_edges, e_none };
    const GJKProxy* proxyA;
    const GJKProxy* proxyB;
    Sweep sweepA, sweepB;
    Type type;
    Vec3 localPoint, axis;
    bool edgesParallel;

    // The axis of an edge of proxy P (vertices a, b in P's frame) towards point c of the other proxy,
    // in the world frame of P at `xf`.
    static Vec3 EdgeAxis(const Vec3& a, const Vec3& b, const Vec3& c, const Transform& xf) {
        Vec3 AB = b - a;
        Vec3 AC = c - a;
        Vec3 PC = Cross(Cross(AB, AC), AB);
        Vec3 n = Mul(xf.rotation, PC);
        n.Normalize();
        return n;
    }
    // The normal of face (a, b, c) of proxy P facing point d of the other proxy, in P's frame.
    static Vec3 FaceAxis(const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& d) {
        Vec3 N = Cross(b - a, c - a);
        N = -N;
        if (Dot(d - a, N) < scalar(0)) N = -N;
        N.Normalize();
        return N;
    }

    void Initialize(const SimplexCache& cache, const GJKProxy& pA, const Sweep& sA, const GJKProxy& pB,
                    const Sweep& sB, scalar t1) {
        proxyA = &pA;
        proxyB = &pB;
        sweepA = sA;
        sweepB = sB;
        edgesParallel = false;
        localPoint.SetZero();
        axis.SetZero();
        const Transform xfA = sweepA.GetTransform(t1);
        const Transform xfB = sweepB.GetTransform(t1);
        const GJKFeatures fp = GetFeaturePair(cache);
        const Vec3* vA = pA.vertices;
        const Vec3* vB = pB.vertices;
        if (fp.count1 == 1 && fp.count2 == 1) {
            type = e_points;
            Vec3 A = xfA * vA[fp.index1[0]];
            Vec3 B = xfB * vB[fp.index2[0]];
            axis = B - A;
            axis.Normalize();
        } else if (fp.count1 == 2 && fp.count2 == 1) {
            type = e_edgeA;
            Vec3 C = MulT(xfA, xfB * vB[fp.index2[0]]);
            axis = EdgeAxis(vA[fp.index1[0]], vA[fp.index1[1]], C, xfA);
        } else if (fp.count1 == 3) {
            type = e_faceA;
            Vec3 D = MulT(xfA, xfB * vB[fp.index2[0]]);
            const Vec3 A = vA[fp.index1[0]];
            axis = FaceAxis(A, vA[fp.index1[1]], vA[fp.index1[2]], D);
            localPoint = A;
        } else if (fp.count2 == 2 && fp.count1 == 1) {
            type = e_edgeB;
            Vec3 C = MulT(xfB, xfA * vA[fp.index1[0]]);
            axis = EdgeAxis(vB[fp.index2[0]], vB[fp.index2[1]], C, xfB);
        } else if (fp.count2 == 3) {
            type = e_faceB;
            Vec3 D = MulT(xfB, xfA * vA[fp.index1[0]]);
            const Vec3 A = vB[fp.index2[0]];
            axis = FaceAxis(A, vB[fp.index2[1]], vB[fp.index2[2]], D);
            localPoint = A;
        } else if (fp.count1 == 2 && fp.count2 == 2) {
            type = e_edges;
            Vec3 A = vA[fp.index1[0]];
            Vec3 B = vA[fp.index1[1]];
            Vec3 C = MulT(xfA, xfB * vB[fp.index2[0]]);
            Vec3 D = MulT(xfA, xfB * vB[fp.index2[1]]);
            Vec3 E1 = B - A;
            Vec3 E2 = D - C;
            scalar L1 = Length(E1);
            scalar L2 = Length(E2);
            const scalar kTol = scalar(0.005);
            Vec3 E1_x_E2 = Cross(E1, E2);
            scalar L = Length(E1_x_E2);
            if (L < kTol * L1 * L2) {  // nearly parallel: the closest points give the axis
                Vec3 C1, C2;
                ClosestPointsOnSegments(C1, C2, A, B, C, D);
                C1 = xfA * C1;
                C2 = xfA * C2;
                axis = C2 - C1;
                axis.Normalize();
                edgesParallel = true;
                return;
            }
            Vec3 C1 = scalar(0.5) * (A + B);
            Vec3 C2 = scalar(0.5) * (C + D);
            Vec3 N = (scalar(1) / L) * E1_x_E2;
            if (Dot(N, C2 - C1) < scalar(0)) N = -N;
            axis = N;
            localPoint = C1;
        } else {
            type = e_none;  // not reached for a simplex GJK leaves
        }
    }

    // Separation of the deepest points along the axis at t, and their indices (kNullU32 on a face side).
    scalar FindMinSeparation(u32* indexA, u32* indexB, scalar t) const {
        const Transform xfA = sweepA.GetTransform(t);
        const Transform xfB = sweepB.GetTransform(t);
        const Quat qA = xfA.rotation, qB = xfB.rotation;
        switch (type) {
        case e_points:
        case e_edgeA: {
            *indexA = proxyA->GetSupportIndex(MulC(qA, axis));
            *indexB = proxyB->GetSupportIndex(MulC(qB, -axis));
            Vec3 A = xfA * proxyA->vertices[*indexA];
            Vec3 B = xfB * proxyB->vertices[*indexB];
            return Dot(B - A, axis);
        }
        case e_faceA: {
            Vec3 N = Mul(qA, axis);
            Vec3 A = xfA * localPoint;
            *indexA = kNullU32;
            *indexB = proxyB->GetSupportIndex(MulC(qB, -N));
            Vec3 B = xfB * proxyB->vertices[*indexB];
            return Dot(B - A, N);
        }
        case e_edgeB: {
            *indexA = proxyA->GetSupportIndex(MulC(qA, -axis));
            *indexB = proxyB->GetSupportIndex(MulC(qB, axis));
            Vec3 A = xfA * proxyA->vertices[*indexA];
            Vec3 B = xfB * proxyB->vertices[*indexB];
            return Dot(A - B, axis);
        }
        case e_faceB: {
            Vec3 N = Mul(qB, axis);
            Vec3 B = xfB * localPoint;
            *indexA = proxyA->GetSupportIndex(MulC(qA, -N));
            *indexB = kNullU32;
            Vec3 A = xfA * proxyA->vertices[*indexA];
            return Dot(A - B, N);
        }
        case e_edges: {
            if (edgesParallel) {
                *indexA = proxyA->GetSupportIndex(MulC(qA, axis));
                *indexB = proxyB->GetSupportIndex(MulC(qB, -axis));
                Vec3 A = xfA * proxyA->vertices[*indexA];
                Vec3 B = xfB * proxyB->vertices[*indexB];
                return Dot(B - A, axis);
            }
            Vec3 N = Mul(qA, axis);
            Vec3 A = xfA * localPoint;
            *indexA = kNullU32;
            *indexB = proxyB->GetSupportIndex(MulC(qB, -N));
            Vec3 B = xfB * proxyB->vertices[*indexB];
            return Dot(B - A, N);
        }
        default:
            *indexA = kNullU32;
            *indexB = kNullU32;
            return -kMaxScalar;
        }
    }

    // Separation of the given points along the axis at t.
    scalar Evaluate(u32 indexA, u32 indexB, scalar t) const {
        const Transform xfA = sweepA.GetTransform(t);
        const Transform xfB = sweepB.GetTransform(t);
        const Quat qA = xfA.rotation, qB = xfB.rotation;
        switch (type) {
        case e_points:
        case e_edgeA: {
            Vec3 A = xfA * proxyA->vertices[indexA];
            Vec3 B = xfB * proxyB->vertices[indexB];
            return Dot(B - A, axis);
        }
        case e_faceA: {
            Vec3 N = Mul(qA, axis);
            Vec3 A = xfA * localPoint;
            Vec3 B = xfB * proxyB->vertices[indexB];
            return Dot(B - A, N);
        }
        case e_edgeB: {
            Vec3 A = xfA * proxyA->vertices[indexA];
            Vec3 B = xfB * proxyB->vertices[indexB];
            return Dot(A - B, axis);
        }
        case e_faceB: {
            Vec3 N = Mul(qB, axis);
            Vec3 B = xfB * localPoint;
            Vec3 A = xfA * proxyA->vertices[indexA];
            return Dot(A - B, N);
        }
        case e_edges: {
            if (edgesParallel) {
                Vec3 A = xfA * proxyA->vertices[indexA];
                Vec3 B = xfB * proxyB->vertices[indexB];
                return Dot(B - A, axis);
            }
            Vec3 N = Mul(qA, axis);
            Vec3 A = xfA * localPoint;
            Vec3 B = xfB * proxyB->vertices[indexB];
            return Dot(B - A, N);
        }
        default: return scalar(0);
        }
    }
};

TOIOutput Out(u32 state, scalar t, u32 iterations) {
    TOIOutput o;
    o.state = state;
    o.t = t;
    o.iterations = iterations;
    return o;
}

}  // namespace

// b3TimeOfImpact(const b3TOIInput&) (time_of_impact.cpp:465-643): advance t1 along the sweeps by
// resolving the deepest points of a separating axis, a root finder of bisection and secant steps.
TOIOutput TimeOfImpact(const GJKProxy& proxyA, const Sweep& sweepA, const GJKProxy& proxyB, const Sweep& sweepB,
                       scalar tMax) {
    TOIOutput output = Out(e_toiUnknown, tMax, 0);
    const scalar totalRadius = proxyA.radius + proxyB.radius;
    const scalar target = Max(kLinearSlop, totalRadius - scalar(3) * kLinearSlop);
    const scalar tolerance = scalar(0.25) * kLinearSlop;
    scalar t1 = scalar(0);
    const u32 kMaxIterations = 20;
    u32 iteration = 0;
    SimplexCache cache;
    cache.count = 0;
    for (;;) {
        const Transform xfA = sweepA.GetTransform(t1);
        const Transform xfB = sweepB.GetTransform(t1);
        GJKOutput gjk = GJK(xfA, proxyA, xfB, proxyB, false, &cache);
        if (gjk.distance == scalar(0)) {
            output = Out(e_toiOverlapped, scalar(0), iteration);
            break;
        }
        if (gjk.distance < target + tolerance) {
            output = Out(e_toiTouching, t1, iteration);
            break;
        }
        SeparationFunction fcn;
        fcn.Initialize(cache, proxyA, sweepA, proxyB, sweepB, t1);
        bool done = false;
        scalar t2 = tMax;
        u32 pushBackIter = 0;
        for (;;) {
            u32 indexA, indexB;
            scalar s2 = fcn.FindMinSeparation(&indexA, &indexB, t2);
            if (s2 > target + tolerance) {
                output = Out(e_toiSeparated, tMax, iteration);
                done = true;
                break;
            }
            if (s2 > target - tolerance) {
                t1 = t2;
                break;
            }
            scalar s1 = fcn.Evaluate(indexA, indexB, t1);
            if (s1 < target - tolerance) {
                output = Out(e_toiFailed, t1, iteration);
                done = true;
                break;
            }
            if (s1 <= target + tolerance) {
                output = Out(e_toiTouching, t1, iteration);
                done = true;
                break;
            }
            u32 rootIteration = 0;
            scalar a1 = t1, a2 = t2;
            for (;;) {
                scalar t;
                if (rootIteration & 1) t = a1 + (target - s1) * (a2 - a1) / (s2 - s1);
                else t = scalar(0.5) * (a1 + a2);
                ++rootIteration;
                scalar s = fcn.Evaluate(indexA, indexB, t);
                if (Abs(s - target) < tolerance) {
                    t2 = t;
                    break;
                }
                if (s > target) {
                    a1 = t;
                    s1 = s;
                } else {
                    a2 = t;
                    s2 = s;
                }
                if (rootIteration == 50) break;
            }
            ++pushBackIter;
            if (pushBackIter == 64) break;
        }
        ++iteration;
        if (done) break;
        if (iteration == kMaxIterations) {
            output = Out(e_toiFailed, t1, iteration);
            break;
        }
    }
    output.iterations = iteration;
    return output;
}

// b3TimeOfImpact of two translating shapes (time_of_impact.cpp:647-758): conservative advancement along
// the closest points' direction.
TOIOutput TimeOfImpact(const Transform& xf1, const GJKProxy& proxy1, const Vec3& d1, const Transform& xf2,
                       const GJKProxy& proxy2, const Vec3& d2, u32 maxIterations) {
    const scalar totalRadius = proxy1.radius + proxy2.radius;
    const scalar target = Max(kLinearSlop, totalRadius - scalar(3) * kLinearSlop);
    const scalar tolerance = scalar(0.25) * kLinearSlop;
    SimplexCache cache;
    cache.count = 0;
    scalar t = scalar(0);
    Transform xf1t = xf1, xf2t = xf2;
    u32 iteration = 0;
    for (;;) {
        GJKOutput query = GJK(xf1t, proxy1, xf2t, proxy2, false, &cache);
        if (query.distance == scalar(0)) return Out(e_toiOverlapped, scalar(0), iteration);
        if (query.distance < target + tolerance) return Out(e_toiTouching, t, iteration);
        Vec3 p1 = query.point1, p2 = query.point2;
        scalar d = query.distance;
        Vec3 n = (p2 - p1) / d;
        scalar denominator = Dot(d2 - d1, n);
        if (denominator >= scalar(0)) return Out(e_toiSeparated, scalar(1), iteration);
        scalar dt = (target - d) / denominator;
        t += dt;
        if (t >= scalar(1)) return Out(e_toiSeparated, scalar(1), iteration);
        xf1t.translation = xf1.translation + t * d1;
        xf2t.translation = xf2.translation + t * d2;
        ++iteration;
        if (iteration == maxIterations) break;
    }
    return Out(e_toiUnknown, t, iteration);
}

// b3TimeOfImpact of two moving boxes (time_of_impact.cpp:764-901): per axis, the interval of times the
// projections overlap.
TOIOutput TimeOfImpact(const AABB& A, const Vec3& d1, const AABB& B, const Vec3& d2) {
    if (TestOverlap(A, B)) return Out(e_toiOverlapped, scalar(0), 0);
    const Vec3 d = d2 - d1;
    Vec3 t1(scalar(0), scalar(0), scalar(0));
    Vec3 t2(scalar(1), scalar(1), scalar(1));
    for (u32 i = 0; i < 3; ++i) {
        const scalar di = d[i];
        const scalar Al = A.lowerBound[i], Au = A.upperBound[i], Bl = B.lowerBound[i], Bu = B.upperBound[i];
        if (di == scalar(0) && (Al > Bu || Au < Bl)) return Out(e_toiSeparated, scalar(1), 0);
        if (di < scalar(0)) {
            if (Bu < Al) return Out(e_toiSeparated, scalar(1), 0);
            t2[i] = (Al - Bu) / di;
            if (Bl > Au) {
                t1[i] = (Au - Bl) / di;
                if (t1[i] >= scalar(1)) return Out(e_toiSeparated, scalar(1), 0);
            }
        }
        if (di > scalar(0)) {
            if (Bl > Au) return Out(e_toiSeparated, scalar(1), 0);
            t2[i] = (Au - Bl) / di;
            if (Bu < Al) {
                t1[i] = (Al - Bu) / di;
                if (t1[i] >= scalar(1)) return Out(e_toiSeparated, scalar(1), 0);
            }
        }
    }
    scalar t1Max = Max(t1.x, Max(t1.y, t1.z));
    scalar t2Min = Min(t2.x, Min(t2.y, t2.z));
    if (t1Max > t2Min) return Out(e_toiSeparated, scalar(0), 0);
    return Out(e_toiTouching, t1Max, 0);
}

}  // namespace orc
