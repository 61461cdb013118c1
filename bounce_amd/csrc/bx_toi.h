// bx_toi.h -- time of impact and world shape casts on the device.
//   gjk.cpp:651-879              b3GJK with a simplex cache (ReadCache, WriteCache, GetMetric)
//   gjk_feature_pair.cpp:21-236  b3GetFeaturePair
//   time_of_impact.cpp:27-458    b3SeparationFunction
//   time_of_impact.cpp:465-643   b3TimeOfImpact(const b3TOIInput&): the local separating-axis method
//   time_of_impact.cpp:647-758   b3TimeOfImpact of two translating shapes (conservative advancement)
//   time_of_impact.cpp:764-901   b3TimeOfImpact of two moving boxes
//   world.cpp:568-850            b3World::ShapeCast / ShapeCastSingle
// A lane per problem or per cast. Every loop is bounded: GJK by 20 support points, the sweep method by
// 20 x 64 x 50 steps, conservative advancement by the caller's iteration cap (the host checks it), the
// mesh tree walk by its fixed stack.
#pragma once

// ================================================================== GJK with a simplex cache
struct SimplexCacheD {
    scalar metric;
    u32 count;
    u32 i1[4], i2[4];
};

BX_HD SVert& SimplexAt(SimplexD& s, u32 i) { return i == 0 ? s.v0 : (i == 1 ? s.v1 : (i == 2 ? s.v2 : s.v3)); }

// b3Simplex::GetMetric (gjk.cpp:881-917): length, area or volume.
BX_HD scalar SimplexMetric(const SimplexD& s) {
    if (s.count == 2) return Distance(s.v0.p, s.v1.p);
    if (s.count == 3) return Length(Cross(s.v1.p - s.v0.p, s.v2.p - s.v0.p));
    if (s.count == 4) {
        scalar det = Det(s.v1.p - s.v0.p, s.v2.p - s.v0.p, s.v3.p - s.v0.p);
        return Sign(det) * det;
    }
    return scalar(0);
}

// b3GJK(xf1, proxy1, xf2, proxy2, false, &cache): the simplex starts from the cached vertices at the new
// transforms unless their metric changed by more than a factor of two, and the final simplex is written
// back. The search loop is the one of GJK (bx_convex.h).
BX_HD GJKOut GJKCached(const Transform& xf1, const ProxyD& proxy1, const Transform& xf2, const ProxyD& proxy2,
                       SimplexCacheD& cache) {
    SimplexD s;
    s.count = cache.count;
    for (u32 i = 0; i < 4; ++i) {
        if (i >= s.count) break;
        SVert& v = SimplexAt(s, i);
        v.i1 = cache.i1[i];
        v.i2 = cache.i2[i];
        v.p1 = Mul(xf1, proxy1.Get(v.i1));
        v.p2 = Mul(xf2, proxy2.Get(v.i2));
        v.p = v.p2 - v.p1;
        v.w = scalar(0);
    }
    if (s.count > 1) {
        scalar metric1 = cache.metric;
        scalar metric2 = SimplexMetric(s);
        if (metric2 < scalar(0.5) * metric1 || scalar(2) * metric1 < metric2 || metric2 < kEpsilon) s.count = 0;
    }
    if (s.count == 0) {
        s.count = 1;
        s.v0.p1 = Mul(xf1, proxy1.Get(0));
        s.v0.p2 = Mul(xf2, proxy2.Get(0));
        s.v0.p = s.v0.p2 - s.v0.p1;
        s.v0.w = scalar(1);
        s.v0.i1 = 0;
        s.v0.i2 = 0;
    }
    for (u32 i = s.count; i < 4; ++i) SimplexAt(s, i) = s.v0;  // defined contents in unused slots
    const Vec3 origin(scalar(0), scalar(0), scalar(0));
    const u32 kMaxIters = 20;
    u32 iter = 0;
    while (iter < kMaxIters) {
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
    cache.metric = SimplexMetric(s);
    cache.count = s.count;
    for (u32 i = 0; i < 4; ++i) {
        if (i >= s.count) break;
        const SVert& v = SimplexAt(s, i);
        cache.i1[i] = v.i1;
        cache.i2[i] = v.i2;
    }
    return out;
}

// ================================================================== separation function
// b3GetFeaturePair: the distinct vertices of each proxy in a simplex cache of one to three vertices.
struct GJKFeaturesD {
    u32 count1, count2;
    u32 a[3], b[3];
};

BX_HD u32 UniqueOf(const u32* v, u32 n) {  // b3UniqueCount (math.h:84-104)
    u32 count = 0;
    for (u32 i = 0; i < n; ++i) {
        u32 j;
        for (j = 0; j < n; ++j)
            if (v[i] == v[j]) break;
        if (i == j) ++count;
    }
    return count;
}

BX_HD GJKFeaturesD FeaturesOf(const SimplexCacheD& c) {
    GJKFeaturesD p;
    p.count1 = p.count2 = 0;
    for (u32 i = 0; i < 3; ++i) p.a[i] = p.b[i] = 0;
    const u32 n = c.count;
    if (n < 1 || n > 3) return p;
    const u32 u1 = UniqueOf(c.i1, n), u2 = UniqueOf(c.i2, n);
    if (n == 1) {
        p.count1 = p.count2 = 1;
        p.a[0] = c.i1[0];
        p.b[0] = c.i2[0];
    } else if (n == 2) {
        if (u1 == 2 && u2 == 2) {
            p.count1 = p.count2 = 2;
        } else if (u1 == 1) {
            p.count1 = 1;
            p.count2 = 2;
        } else if (u2 == 1) {
            p.count1 = 2;
            p.count2 = 1;
        }
        p.a[0] = c.i1[0];
        p.a[1] = c.i1[1];
        p.b[0] = c.i2[0];
        p.b[1] = c.i2[1];
    } else {
        if (u1 == 1 && u2 == 3) {
            p.count1 = 1;
            p.count2 = 3;
        } else if (u1 == 3 && u2 == 1) {
            p.count1 = 3;
            p.count2 = 1;
        } else if ((u1 == 2 || u1 == 3) && (u2 == 2 || u2 == 3)) {
            p.count1 = u1;
            p.count2 = u2;
        }
        // a side with two distinct of its three vertices keeps the first and the other one
        p.a[0] = c.i1[0];
        p.a[1] = u1 == 2 && c.i1[0] == c.i1[1] ? c.i1[2] : c.i1[1];
        p.a[2] = c.i1[2];
        p.b[0] = c.i2[0];
        p.b[1] = u2 == 2 && c.i2[0] == c.i2[1] ? c.i2[2] : c.i2[1];
        p.b[2] = c.i2[2];
    }
    return p;
}

// b3SeparationFunction: an axis found at t1 from the closest features, and the separation along it at
// any t.
struct SeparationD {
    enum Type : u32 { kPoints, kEdgeA, kFaceA, kEdgeB, kFaceB, kEdges, kNone };
    Type type;
    bool edgesParallel;
    Vec3 localPoint, axis;

    BX_HD static Vec3 EdgeAxis(const Vec3& a, const Vec3& b, const Vec3& c, const Transform& xf) {
        Vec3 AB = b - a;
        Vec3 AC = c - a;
        Vec3 PC = Cross(Cross(AB, AC), AB);
        Vec3 n = Mul(xf.rotation, PC);
        n.Normalize();
        return n;
    }
    BX_HD static Vec3 FaceAxis(const Vec3& a, const Vec3& b, const Vec3& c, const Vec3& d) {
        Vec3 N = Cross(b - a, c - a);
        N = -N;
        if (Dot(d - a, N) < scalar(0)) N = -N;
        N.Normalize();
        return N;
    }

    BX_HD void Initialize(const SimplexCacheD& cache, const ProxyD& pA, const Transform& xfA, const ProxyD& pB,
                          const Transform& xfB) {
        edgesParallel = false;
        localPoint.SetZero();
        axis.SetZero();
        const GJKFeaturesD fp = FeaturesOf(cache);
        if (fp.count1 == 1 && fp.count2 == 1) {
            type = kPoints;
            axis = Mul(xfB, pB.Get(fp.b[0])) - Mul(xfA, pA.Get(fp.a[0]));
            axis.Normalize();
        } else if (fp.count1 == 2 && fp.count2 == 1) {
            type = kEdgeA;
            Vec3 C = MulT(xfA, Mul(xfB, pB.Get(fp.b[0])));
            axis = EdgeAxis(pA.Get(fp.a[0]), pA.Get(fp.a[1]), C, xfA);
        } else if (fp.count1 == 3) {
            type = kFaceA;
            Vec3 D = MulT(xfA, Mul(xfB, pB.Get(fp.b[0])));
            localPoint = pA.Get(fp.a[0]);
            axis = FaceAxis(localPoint, pA.Get(fp.a[1]), pA.Get(fp.a[2]), D);
        } else if (fp.count2 == 2 && fp.count1 == 1) {
            type = kEdgeB;
            Vec3 C = MulT(xfB, Mul(xfA, pA.Get(fp.a[0])));
            axis = EdgeAxis(pB.Get(fp.b[0]), pB.Get(fp.b[1]), C, xfB);
        } else if (fp.count2 == 3) {
            type = kFaceB;
            Vec3 D = MulT(xfB, Mul(xfA, pA.Get(fp.a[0])));
            localPoint = pB.Get(fp.b[0]);
            axis = FaceAxis(localPoint, pB.Get(fp.b[1]), pB.Get(fp.b[2]), D);
        } else if (fp.count1 == 2 && fp.count2 == 2) {
            type = kEdges;
            Vec3 A = pA.Get(fp.a[0]);
            Vec3 B = pA.Get(fp.a[1]);
            Vec3 C = MulT(xfA, Mul(xfB, pB.Get(fp.b[0])));
            Vec3 D = MulT(xfA, Mul(xfB, pB.Get(fp.b[1])));
            Vec3 E1 = B - A;
            Vec3 E2 = D - C;
            scalar L1 = Length(E1);
            scalar L2 = Length(E2);
            const scalar kTol = scalar(0.005);
            Vec3 E1_x_E2 = Cross(E1, E2);
            scalar L = Length(E1_x_E2);
            if (L < kTol * L1 * L2) {  // nearly parallel edges: the closest points give the axis
                Vec3 C1, C2;
                ClosestPointsOnSegments(C1, C2, A, B, C, D);
                C1 = Mul(xfA, C1);
                C2 = Mul(xfA, C2);
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
            type = kNone;  // not reached for a simplex GJK leaves
        }
    }

    // find: the deepest points along the axis (their indices out, kNull on a face side); otherwise the
    // separation of the given points.
    BX_HD scalar Separation(const ProxyD& pA, const Transform& xfA, const ProxyD& pB, const Transform& xfB, bool find,
                            u32* indexA, u32* indexB) const {
        const Quat qA = xfA.rotation, qB = xfB.rotation;
        switch (type) {
        case kPoints:
        case kEdgeA: {
            if (find) {
                *indexA = pA.Support(MulC(qA, axis));
                *indexB = pB.Support(MulC(qB, -axis));
            }
            return Dot(Mul(xfB, pB.Get(*indexB)) - Mul(xfA, pA.Get(*indexA)), axis);
        }
        case kFaceA: {
            Vec3 N = Mul(qA, axis);
            Vec3 A = Mul(xfA, localPoint);
            if (find) {
                *indexA = kNull;
                *indexB = pB.Support(MulC(qB, -N));
            }
            return Dot(Mul(xfB, pB.Get(*indexB)) - A, N);
        }
        case kEdgeB: {
            if (find) {
                *indexA = pA.Support(MulC(qA, -axis));
                *indexB = pB.Support(MulC(qB, axis));
            }
            return Dot(Mul(xfA, pA.Get(*indexA)) - Mul(xfB, pB.Get(*indexB)), axis);
        }
        case kFaceB: {
            Vec3 N = Mul(qB, axis);
            Vec3 B = Mul(xfB, localPoint);
            if (find) {
                *indexA = pA.Support(MulC(qA, -N));
                *indexB = kNull;
            }
            return Dot(Mul(xfA, pA.Get(*indexA)) - B, N);
        }
        case kEdges: {
            if (edgesParallel) {
                if (find) {
                    *indexA = pA.Support(MulC(qA, axis));
                    *indexB = pB.Support(MulC(qB, -axis));
                }
                return Dot(Mul(xfB, pB.Get(*indexB)) - Mul(xfA, pA.Get(*indexA)), axis);
            }
            Vec3 N = Mul(qA, axis);
            Vec3 A = Mul(xfA, localPoint);
            if (find) {
                *indexA = kNull;
                *indexB = pB.Support(MulC(qB, -N));
            }
            return Dot(Mul(xfB, pB.Get(*indexB)) - A, N);
        }
        default:
            if (find) {
                *indexA = kNull;
                *indexB = kNull;
                return -kMaxScalar;
            }
            return scalar(0);
        }
    }
};

BX_HD Transform SweepAt(const SweepD& s, scalar t) { return SweepTransform(s.c0, s.q0, s.c, s.q, s.localCenter, t); }

BX_HD ToiResultD ToiOut(u32 state, scalar t, u32 iterations) {
    ToiResultD o;
    o.state = state;
    o.t = t;
    o.iterations = iterations;
    return o;
}

constexpr u32 kToiUnknown = 0, kToiFailed = 1, kToiOverlapped = 2, kToiTouching = 3, kToiSeparated = 4;

// b3TimeOfImpact(const b3TOIInput&): advance t1 along the sweeps by resolving the deepest points of a
// separating axis with a root finder of bisection and secant steps.
BX_HD ToiResultD TimeOfImpactSweeps(const ProxyD& pA, const SweepD& sA, const ProxyD& pB, const SweepD& sB,
                                    scalar tMax) {
    ToiResultD output = ToiOut(kToiUnknown, tMax, 0);
    const scalar totalRadius = pA.radius + pB.radius;
    const scalar target = Max(kLinearSlop, totalRadius - scalar(3) * kLinearSlop);
    const scalar tolerance = scalar(0.25) * kLinearSlop;
    scalar t1 = scalar(0);
    const u32 kMaxIterations = 20;
    u32 iteration = 0;
    SimplexCacheD cache;
    cache.count = 0;
    cache.metric = scalar(0);
    for (;;) {
        const Transform xfA = SweepAt(sA, t1), xfB = SweepAt(sB, t1);
        GJKOut gjk = GJKCached(xfA, pA, xfB, pB, cache);
        if (gjk.distance == scalar(0)) {
            output = ToiOut(kToiOverlapped, scalar(0), iteration);
            break;
        }
        if (gjk.distance < target + tolerance) {
            output = ToiOut(kToiTouching, t1, iteration);
            break;
        }
        SeparationD fcn;
        fcn.Initialize(cache, pA, xfA, pB, xfB);
        bool done = false;
        scalar t2 = tMax;
        u32 pushBackIter = 0;
        for (;;) {
            u32 indexA = 0, indexB = 0;
            scalar s2 = fcn.Separation(pA, SweepAt(sA, t2), pB, SweepAt(sB, t2), true, &indexA, &indexB);
            if (s2 > target + tolerance) {
                output = ToiOut(kToiSeparated, tMax, iteration);
                done = true;
                break;
            }
            if (s2 > target - tolerance) {
                t1 = t2;
                break;
            }
            scalar s1 = fcn.Separation(pA, SweepAt(sA, t1), pB, SweepAt(sB, t1), false, &indexA, &indexB);
            if (s1 < target - tolerance) {
                output = ToiOut(kToiFailed, t1, iteration);
                done = true;
                break;
            }
            if (s1 <= target + tolerance) {
                output = ToiOut(kToiTouching, t1, iteration);
                done = true;
                break;
            }
            u32 rootIteration = 0;
            scalar a1 = t1, a2 = t2;
            for (;;) {
                scalar t;
                if (rootIteration & 1u) t = a1 + (target - s1) * (a2 - a1) / (s2 - s1);
                else t = scalar(0.5) * (a1 + a2);
                ++rootIteration;
                scalar s = fcn.Separation(pA, SweepAt(sA, t), pB, SweepAt(sB, t), false, &indexA, &indexB);
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
            output = ToiOut(kToiFailed, t1, iteration);
            break;
        }
    }
    output.iterations = iteration;
    return output;
}

// b3TimeOfImpact of two translating shapes: conservative advancement along the closest points'
// direction, at most maxIterations steps (maxIterations >= 1).
BX_HD ToiResultD TimeOfImpactTranslation(const Transform& xf1, const ProxyD& p1, const Vec3& d1,
                                         const Transform& xf2, const ProxyD& p2, const Vec3& d2, u32 maxIterations) {
    const scalar totalRadius = p1.radius + p2.radius;
    const scalar target = Max(kLinearSlop, totalRadius - scalar(3) * kLinearSlop);
    const scalar tolerance = scalar(0.25) * kLinearSlop;
    SimplexCacheD cache;
    cache.count = 0;
    cache.metric = scalar(0);
    scalar t = scalar(0);
    Transform xf1t = xf1, xf2t = xf2;
    u32 iteration = 0;
    for (;;) {
        GJKOut query = GJKCached(xf1t, p1, xf2t, p2, cache);
        if (query.distance == scalar(0)) return ToiOut(kToiOverlapped, scalar(0), iteration);
        if (query.distance < target + tolerance) return ToiOut(kToiTouching, t, iteration);
        scalar d = query.distance;
        Vec3 n = (query.point2 - query.point1) / d;
        scalar denominator = Dot(d2 - d1, n);
        if (denominator >= scalar(0)) return ToiOut(kToiSeparated, scalar(1), iteration);
        t += (target - d) / denominator;
        if (t >= scalar(1)) return ToiOut(kToiSeparated, scalar(1), iteration);
        xf1t.translation = xf1.translation + t * d1;
        xf2t.translation = xf2.translation + t * d2;
        ++iteration;
        if (iteration == maxIterations) break;
    }
    return ToiOut(kToiUnknown, t, iteration);
}

// b3TimeOfImpact of two moving boxes: per axis, the interval of times their projections overlap.
BX_HD ToiResultD TimeOfImpactBoxes(const AABB& A, const Vec3& d1, const AABB& B, const Vec3& d2) {
    if (TestOverlap(A, B)) return ToiOut(kToiOverlapped, scalar(0), 0);
    const Vec3 dv = d2 - d1;
    const scalar d[3] = {dv.x, dv.y, dv.z};
    const scalar Al[3] = {A.lowerBound.x, A.lowerBound.y, A.lowerBound.z};
    const scalar Au[3] = {A.upperBound.x, A.upperBound.y, A.upperBound.z};
    const scalar Bl[3] = {B.lowerBound.x, B.lowerBound.y, B.lowerBound.z};
    const scalar Bu[3] = {B.upperBound.x, B.upperBound.y, B.upperBound.z};
    scalar t1[3] = {scalar(0), scalar(0), scalar(0)};
    scalar t2[3] = {scalar(1), scalar(1), scalar(1)};
    for (u32 i = 0; i < 3; ++i) {
        if (d[i] == scalar(0) && (Al[i] > Bu[i] || Au[i] < Bl[i])) return ToiOut(kToiSeparated, scalar(1), 0);
        if (d[i] < scalar(0)) {
            if (Bu[i] < Al[i]) return ToiOut(kToiSeparated, scalar(1), 0);
            t2[i] = (Al[i] - Bu[i]) / d[i];
            if (Bl[i] > Au[i]) {
                t1[i] = (Au[i] - Bl[i]) / d[i];
                if (t1[i] >= scalar(1)) return ToiOut(kToiSeparated, scalar(1), 0);
            }
        }
        if (d[i] > scalar(0)) {
            if (Bl[i] > Au[i]) return ToiOut(kToiSeparated, scalar(1), 0);
            t2[i] = (Au[i] - Bl[i]) / d[i];
            if (Bu[i] < Al[i]) {
                t1[i] = (Al[i] - Bu[i]) / d[i];
                if (t1[i] >= scalar(1)) return ToiOut(kToiSeparated, scalar(1), 0);
            }
        }
    }
    scalar t1Max = Max(t1[0], Max(t1[1], t1[2]));
    scalar t2Min = Min(t2[0], Min(t2[1], t2[2]));
    if (t1Max > t2Min) return ToiOut(kToiSeparated, scalar(0), 0);
    return ToiOut(kToiTouching, t1Max, 0);
}

// b3GJKProxy of a cast shape.
__device__ __forceinline__ ProxyD CastProxy(const BatchView& B, const CastShapeD& c) {
    ProxyD p;
    p.g = nullptr;
    p.a = c.v1;
    p.b = c.v2;
    p.c = c.v3;
    p.radius = c.radius;
    p.count = c.type == 1u ? 2u : (c.type == 2u ? 3u : 1u);
    if (c.type == 3u) {
        const HullView h = GetHull(B, c.index);
        p.g = h.vertices;
        p.count = h.vertexCount;
    }
    return p;
}

__global__ void __launch_bounds__(kBlock) k_toi(BatchView B, const ToiQueryD* queries, u32 count, ToiResultD* out) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const ToiQueryD& q = queries[i];
    ToiResultD r;
    if (q.kind == 2u) {
        r = TimeOfImpactBoxes(q.a1, q.d1, q.a2, q.d2);
    } else {
        const ProxyD p1 = CastProxy(B, q.s1), p2 = CastProxy(B, q.s2);
        if (q.kind == 0u) r = TimeOfImpactSweeps(p1, q.w1, p2, q.w2, q.tMax);
        else r = TimeOfImpactTranslation(q.xf1, p1, q.d1, q.xf2, p2, q.d2, q.maxIterations);
    }
    out[i] = r;
}

// ================================================================== world shape casts
// (fraction, fixture, triangle) order of shape-cast hits.
__device__ __forceinline__ bool CastBefore(const ShapeHitOut& a, const ShapeHitOut& b) {
    return a.fraction < b.fraction || (a.fraction == b.fraction && (a.fixture < b.fixture ||
                                                                    (a.fixture == b.fixture && a.child < b.child)));
}

// One lane per cast: candidates from the world's hierarchy (fat AABB overlapping the swept AABB), each
// convex fixture or reported mesh triangle cast by conservative advancement; a touching cast is a hit,
// its point and normal from GJK at the hit (the query wrapper's Evaluate, world.cpp:703-716).
__global__ void __launch_bounds__(kBlock) k_shape_cast(BatchView B, CastParams cp) {
    const u32 qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= cp.count) return;
    const ShapeCastD c = cp.casts[qi];
    BX_WORLD_HOOK(B, c.world);
    const WorldInfo wi = B.worlds[c.world];
    const uint8_t* mask = c.mask == kNull ? nullptr : cp.mask + c.mask;
    Fixture fa;  // the cast shape in the form ComputeAABB takes
    fa.shapeType = c.shape.type;
    fa.shapeIndex = c.shape.index;
    fa.radius = c.shape.radius;
    fa.v1 = c.shape.v1;
    fa.v2 = c.shape.v2;
    fa.v3 = c.shape.v3;
    const ProxyD proxyA = CastProxy(B, c.shape);
    Transform xf2 = c.xf;
    xf2.translation += c.d;
    AABB a1, a2;
    ComputeAABB(B, fa, c.xf, a1);
    ComputeAABB(B, fa, xf2, a2);
    const AABB box = Combine(a1, a2);
    const Vec3 zero(scalar(0), scalar(0), scalar(0));
    ShapeHitOut* out = cp.hits + (cp.single ? u64(qi) : u64(qi) * cp.perQuery);
    ShapeHitOut best;
    best.fixture = kNull;
    best.child = 0;
    best.fraction = scalar(0);
    best.point.SetZero();
    best.normal.SetZero();
    u32 found = 0;
    auto cast = [&](u32 f, u32 child, const Transform& xfB, const ProxyD& proxyB) {
        ToiResultD toi = TimeOfImpactTranslation(c.xf, proxyA, c.d, xfB, proxyB, zero, 20u);
        if (toi.state != kToiTouching) return;
        ShapeHitOut h;
        h.fixture = f;
        h.child = child;
        h.fraction = toi.t;
        Transform xt(c.xf.translation + toi.t * c.d, c.xf.rotation);
        GJKOut q = GJK(xt, proxyA, xfB, proxyB, false);
        Vec3 n = Normalize(q.point1 - q.point2);
        h.point = q.point2 + proxyB.radius * n;
        h.normal = n;
        if (cp.single) {
            if (best.fixture == kNull || CastBefore(h, best)) best = h;
            return;
        }
        u32 k = Min(found, cp.perQuery);
        while (k > 0 && CastBefore(h, out[k - 1])) {
            if (k < cp.perQuery) out[k] = out[k - 1];
            --k;
        }
        if (k < cp.perQuery) out[k] = h;
        ++found;
    };
    WalkBvh(B, wi, [&](const AABB& node) { return TestOverlap(node, box); }, [&](u32 f) {
        if (!TestOverlap(B.fatAABB[wi.fixtureBase + f], box) || (mask && !mask[f])) return;
        const Fixture fx = B.fixtures[wi.fixtureBase + f];
        const Body& b = B.bodies[wi.bodyBase + fx.body];
        const Transform xfB(b.xfT, b.q);
        if (fx.shapeType != 4u) {
            HullView h;
            if (fx.shapeType == 3u) h = GetHull(B, fx.shapeIndex);
            cast(f, 0u, xfB, MakeProxy(fx, fx.shapeType == 3u ? &h : nullptr));
            return;
        }
        // a mesh: its static tree in stack order over the swept AABB in the unscaled mesh frame
        const MeshInfo mi = B.meshes[fx.shapeIndex];
        if (mi.tCount == 0) return;
        AABB m1, m2;
        ComputeAABB(B, fa, MulT(xfB, c.xf), m1);
        ComputeAABB(B, fa, MulT(xfB, xf2), m2);
        const AABB mb = Combine(ScaledInv(m1, fx.v1), ScaledInv(m2, fx.v1));
        const TreeNode* nodes = B.meshNodes + mi.nBase;
        const Vec3* verts = B.meshVertices + mi.vBase;
        const MeshTri* tris = B.meshTris + mi.tBase;
        u32 stack[kTreeStack];
        u32 sp = 0;
        stack[sp++] = mi.root;
        while (sp > 0) {
            const u32 id = stack[--sp];
            if (id == kNull) continue;
            const TreeNode node = nodes[id];
            if (!TestOverlap(node.aabb, mb)) continue;
            if (node.child1 == kNull) {
                const MeshTri t = tris[node.index];
                ProxyD pB;
                pB.g = nullptr;
                pB.a = Mul(fx.v1, verts[t.v1]);
                pB.b = Mul(fx.v1, verts[t.v2]);
                pB.c = Mul(fx.v1, verts[t.v3]);
                pB.count = 3;
                pB.radius = fx.radius;
                cast(f, node.index, xfB, pB);
            } else if (sp + 2u <= kTreeStack) {
                stack[sp++] = node.child1;
                stack[sp++] = node.child2;
            }
        }
    });
    if (cp.single) *out = best;
    else cp.counts[qi] = found;
}
