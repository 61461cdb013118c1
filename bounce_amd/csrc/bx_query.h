// bx_query.h -- world queries on the device (included inside namespace bx by bx_step.hip):
// b3World::RayCast, RayCastSingle and QueryAABB (world.cpp:452-566, 852-870) over many worlds.
//
// Candidates come from a bounding-volume hierarchy over each world's fat AABBs, built on the device
// after a step (k_build_bvh) and walked without a stack: the fixtures are sorted by the Morton code
// of their fat-AABB centres, and an implicit tree of arity 8 sits over them (node j of level l
// covers children 8j..8j+7 of level l-1; level 0 is the sorted fixtures). A walk moves to the first
// child on a hit and otherwise to the next sibling, climbing while a node is the last of its group,
// so its whole state is (level, index). Internal nodes are unions of fat AABBs and are tested with
// comparisons only (no rounding), so they never cut off a fixture that passes its own test; a fixture
// is then tested exactly as b3DynamicTree::RayCast tests a leaf (dynamic_tree.h:201-339) and its
// shape cast as b3Fixture::RayCast does:
//   sphere_shape.cpp:69-121, capsule_shape.cpp:180-324, hull_shape.cpp:189-281,
//   triangle_shape.cpp:95-176, mesh_shape.cpp:81-146 (the mesh's static tree in stack order,
//   static_tree.h:150-290).
#pragma once
// (included inside namespace bx by bx_step.hip)

constexpr u32 kBvhArity = 8;
constexpr u32 kBvhLdsKeys = 2048;  // worlds up to this many fixtures sort their keys in LDS

// Nodes in level l >= 1 of a world with n fixtures (level 0: the n sorted fixtures).
BX_HD u32 BvhLevelCount(u32 n, u32 l) {
    u32 c = n;
    for (u32 k = 0; k < l; ++k) c = (c + kBvhArity - 1) / kBvhArity;
    return c;
}
// Levels above the fixtures (0 when n <= 1).
BX_HD u32 BvhLevels(u32 n) {
    u32 L = 0;
    for (u32 c = n; c > 1; c = (c + kBvhArity - 1) / kBvhArity) ++L;
    return L;
}
// Offset of level l >= 1 in the world's node range (levels stored bottom-up).
BX_HD u32 BvhLevelOffset(u32 n, u32 l) {
    u32 off = 0, c = n;
    for (u32 k = 1; k < l; ++k) {
        c = (c + kBvhArity - 1) / kBvhArity;
        off += c;
    }
    return off;
}

struct RayIn { Vec3 p1, p2; scalar maxFraction; };
struct RayOut { scalar fraction; Vec3 normal; };

BX_HD Vec3 AbsV(const Vec3& v) { return Vec3(Abs(v.x), Abs(v.y), Abs(v.z)); }

// b3Vec3 r = p2 - p1; r.Normalize() (vec3.h:93-104).
BX_HD Vec3 RayDirection(const Vec3& p1, const Vec3& p2) {
    Vec3 r = p2 - p1;
    r.Normalize();
    return r;
}

BX_HD bool SegmentBoxOverlap(const Vec3& p1, const Vec3& q2, const AABB& box) {
    AABB seg;
    seg.lowerBound = Min(p1, q2);
    seg.upperBound = Max(p1, q2);
    return TestOverlap(seg, box);
}

// The node test of b3DynamicTree::RayCast for segment p1..q2, direction r: the segment's AABB, then
// the three box axes and the three cross axes (Gino, p80).
BX_HD bool SegmentOverlapsBox(const Vec3& p1, const Vec3& q2, const Vec3& r, const AABB& box) {
    if (!SegmentBoxOverlap(p1, q2, box)) return false;
    const Vec3 c = scalar(0.5) * (box.lowerBound + box.upperBound);
    const Vec3 h = scalar(0.5) * (box.upperBound - box.lowerBound);
    const Vec3 s = p1 - c, t = q2 - c;
    if (Abs(s.x + t.x) - (Abs(s.x - t.x) + scalar(2) * h.x) > scalar(0)) return false;
    if (Abs(s.y + t.y) - (Abs(s.y - t.y) + scalar(2) * h.y) > scalar(0)) return false;
    if (Abs(s.z + t.z) - (Abs(s.z - t.z) + scalar(2) * h.z) > scalar(0)) return false;
    const Vec3 v1 = Cross(Vec3(1, 0, 0), r);
    if (Abs(Dot(v1, s)) - Dot(AbsV(v1), h) > scalar(0)) return false;
    const Vec3 v2 = Cross(Vec3(0, 1, 0), r);
    if (Abs(Dot(v2, s)) - Dot(AbsV(v2), h) > scalar(0)) return false;
    const Vec3 v3 = Cross(Vec3(0, 0, 1), r);
    if (Abs(Dot(v3, s)) - Dot(AbsV(v3), h) > scalar(0)) return false;
    return true;
}

// Quadratic of a ray against the sphere of radius^2 rr at offset m = p1 - centre, direction d
// (a = |d|^2); the smaller root when it lies in (0, maxFraction].
BX_HD bool RaySphereRoot(const Vec3& m, const Vec3& d, scalar a, scalar rr, scalar maxFraction, RayOut* out) {
    const scalar b = Dot(m, d);
    const scalar c = Dot(m, m) - rr;
    const scalar disc = b * b - a * c;
    if (disc < scalar(0)) return false;
    scalar t = -b - Sqrt(disc);
    if (t > scalar(0) && t <= maxFraction * a) {
        t /= a;
        out->fraction = t;
        out->normal = Normalize(m + t * d);
        return true;
    }
    return false;
}

BX_HD bool SphereRay(const Fixture& f, const Transform& xf, const RayIn& in, RayOut* out) {
    const Vec3 d = in.p2 - in.p1;
    const scalar a = Dot(d, d);
    if (a < kEpsilon * kEpsilon) return false;
    return RaySphereRoot(in.p1 - Mul(xf, f.v1), d, a, f.radius * f.radius, in.maxFraction, out);
}

BX_HD bool CapsuleRay(const Fixture& f, const Transform& xf, const RayIn& in, RayOut* out) {
    const Vec3 A = in.p1, B = in.p2;
    const Vec3 n = B - A;
    const scalar nn = Dot(n, n);
    if (nn < kEpsilon * kEpsilon) return false;
    const Vec3 P = Mul(xf, f.v1), Q = Mul(xf, f.v2);
    const Vec3 d = Q - P;
    const scalar dd = Dot(d, d);
    const scalar rr = f.radius * f.radius;
    if (dd < kEpsilon * kEpsilon) return RaySphereRoot(A - P, n, nn, rr, in.maxFraction, out);
    const Vec3 m = A - P;
    const scalar nd = Dot(n, d), mm = Dot(m, m), md = Dot(m, d), mn = Dot(m, n);
    const scalar a = dd * nn - nd * nd;
    if (Abs(a) < scalar(2) * (kEpsilon * kEpsilon)) return false;
    const scalar b = dd * mn - nd * md;
    const scalar c = dd * (mm - rr) - md * md;
    const scalar disc = b * b - a * c;
    if (disc < scalar(0)) return false;
    scalar t = -b - Sqrt(disc);
    t /= a;
    if (t < scalar(0) || t > scalar(1)) return false;
    const scalar tp = md + t * nd;  // where along PQ the line meets the capsule's cylinder
    Vec3 C;                         // the end whose hemisphere is hit, unless the cylinder is
    if (tp < scalar(0)) {
        C = P;
    } else if (tp > dd) {
        C = Q;
    } else {
        const Vec3 X = (scalar(1) - t) * A + t * B;
        const Vec3 e2 = Cross(d, X - P);
        Vec3 e3 = Cross(e2, d);
        e3.Normalize();
        out->fraction = t;
        out->normal = e3;
        return true;
    }
    return RaySphereRoot(A - C, n, nn, rr, in.maxFraction, out);
}

BX_HD bool HullRay(const HullView& h, const Transform& xf, const RayIn& in, RayOut* out) {
    const Vec3 p1 = MulC(xf.rotation, in.p1 - xf.translation);
    const Vec3 p2 = MulC(xf.rotation, in.p2 - xf.translation);
    const Vec3 d = p2 - p1;
    scalar lower = scalar(0), upper = in.maxFraction;
    u32 index = kNull;
    for (u32 i = 0; i < h.faceCount; ++i) {
        const Plane pl = h.planes[i];
        const scalar numerator = pl.offset - Dot(pl.normal, p1);
        const scalar denominator = Dot(pl.normal, d);
        if (denominator == scalar(0)) {
            if (numerator < scalar(0)) return false;
            continue;
        }
        if (denominator < scalar(0)) {
            if (numerator < lower * denominator) {
                lower = numerator / denominator;
                index = i;
            }
        } else if (numerator < upper * denominator) {
            upper = numerator / denominator;
        }
        if (upper < lower) return false;
    }
    if (index == kNull) return false;
    out->fraction = lower;
    out->normal = Mul(xf.rotation, h.planes[index].normal);
    return true;
}

BX_HD bool TriangleRay(const Vec3& v1, const Vec3& v2, const Vec3& v3, const Transform& xf, const RayIn& in,
                       RayOut* out) {
    const Vec3 p1 = MulT(xf, in.p1), p2 = MulT(xf, in.p2);
    const Vec3 d = p2 - p1;
    if (LengthSquared(d) < kEpsilon * kEpsilon) return false;
    Vec3 n = Cross(v2 - v1, v3 - v1);
    const scalar len = Length(n);
    if (len == scalar(0)) return false;
    n /= len;
    const scalar num = Dot(n, v1 - p1);
    const scalar den = Dot(n, d);
    if (den == scalar(0)) return false;
    const scalar fraction = num / den;
    if (fraction < scalar(0) || in.maxFraction < fraction) return false;
    const Vec3 Q = p1 + fraction * d;
    const Vec3 QA = v1 - Q, QB = v2 - Q, QC = v3 - Q;
    const Vec3 AB_x_AC = Cross(v2 - v1, v3 - v1);
    const scalar u = Dot(Cross(QB, QC), AB_x_AC);
    const scalar v = Dot(Cross(QC, QA), AB_x_AC);
    const scalar w = Dot(Cross(QA, QB), AB_x_AC);
    if (u >= scalar(0) && v >= scalar(0) && w >= scalar(0)) {
        out->fraction = fraction;
        out->normal = num > scalar(0) ? Mul(xf.rotation, -n) : Mul(xf.rotation, n);
        return true;
    }
    return false;
}

// b3MeshShape::RayCast: the static tree in stack order with the ray in the unscaled mesh frame (never
// clipped), each candidate child triangle cast in the world frame, the first strictly smallest
// fraction kept. The mesh scale is the fixture's v1.
__device__ bool MeshRay(const BatchView& B, const Fixture& f, const Transform& xf, const RayIn& in, RayOut* out) {
    const MeshInfo mi = B.meshes[f.shapeIndex];
    if (mi.tCount == 0) return false;
    const Vec3 inv(scalar(1) / f.v1.x, scalar(1) / f.v1.y, scalar(1) / f.v1.z);
    const Vec3 tp1 = Mul(inv, MulT(xf, in.p1)), tp2 = Mul(inv, MulT(xf, in.p2));
    const Vec3 r = RayDirection(tp1, tp2);
    const Vec3 q2 = tp1 + in.maxFraction * (tp2 - tp1);
    const TreeNode* nodes = B.meshNodes + mi.nBase;
    const Vec3* verts = B.meshVertices + mi.vBase;
    const MeshTri* tris = B.meshTris + mi.tBase;
    bool hit = false;
    RayOut best;
    best.fraction = kMaxScalar;
    best.normal.SetZero();
    u32 stack[kTreeStack];
    u32 sp = 0;
    stack[sp++] = mi.root;
    while (sp > 0) {
        const u32 id = stack[--sp];
        if (id == kNull) continue;
        const TreeNode node = nodes[id];
        if (!SegmentOverlapsBox(tp1, q2, r, node.aabb)) continue;
        if (node.child1 == kNull) {
            const MeshTri t = tris[node.index];
            RayOut o;
            if (TriangleRay(Mul(f.v1, verts[t.v1]), Mul(f.v1, verts[t.v2]), Mul(f.v1, verts[t.v3]), xf, in, &o) &&
                o.fraction < best.fraction) {
                hit = true;
                best = o;
            }
        } else if (sp + 2u <= kTreeStack) {
            stack[sp++] = node.child1;
            stack[sp++] = node.child2;
        }
    }
    *out = best;
    return hit;
}

// b3Fixture::RayCast of fixture `fi` of world `wi`.
__device__ bool FixtureRay(const BatchView& B, const WorldInfo& wi, u32 fi, const RayIn& in, RayOut* out) {
    const Fixture f = B.fixtures[wi.fixtureBase + fi];
    const Body& b = B.bodies[wi.bodyBase + f.body];
    const Transform xf(b.xfT, b.q);
    switch (f.shapeType) {
    case 0u: return SphereRay(f, xf, in, out);
    case 1u: return CapsuleRay(f, xf, in, out);
    case 2u: return TriangleRay(f.v1, f.v2, f.v3, xf, in, out);
    case 3u: return HullRay(GetHull(B, f.shapeIndex), xf, in, out);
    case 4u: return MeshRay(B, f, xf, in, out);
    }
    return false;
}

// ================================================================== hierarchy build
// Morton code of a point quantized to 10 bits per axis within [lo, lo + 1/scale].
__device__ __forceinline__ u32 Spread10(u32 x) {
    x &= 0x3FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}
__device__ __forceinline__ u32 Quantize10(scalar v) {
    const scalar q = v * scalar(1023);
    return q >= scalar(1023) ? 1023u : (q > scalar(0) ? u32(q) : 0u);
}

// Sorts keys[0..P) ascending (P a power of two) with the block's threads; `keys` is LDS or the
// world's global scratch.
__device__ void BlockBitonicSort(u64* keys, u32 P) {
    for (u32 k = 2; k <= P; k <<= 1) {
        for (u32 j = k >> 1; j > 0; j >>= 1) {
            for (u32 i = threadIdx.x; i < P; i += blockDim.x) {
                const u32 l = i ^ j;
                if (l > i) {
                    const u64 a = keys[i], b = keys[l];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        keys[i] = b;
                        keys[l] = a;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// One workgroup per listed world: the sorted fixture list and the node levels of its hierarchy.
__global__ void __launch_bounds__(kBlock) k_build_bvh(BatchView B) {
    __shared__ u64 ldsKeys[kBvhLdsKeys];
    __shared__ scalar red[6][kBlock];
    const u32 w = WorldOfBlock(B);
    const WorldInfo wi = B.worlds[w];
    const u32 n = wi.fixtureCount;
    if (n == 0) return;
    const AABB* fat = B.fatAABB + wi.fixtureBase;
    u32* leaves = B.bvhLeaves + wi.fixtureBase;
    AABB* nodes = B.bvhNodes + wi.fixtureBase;
    // bounds of the fat-AABB centres
    Vec3 lo(kMaxScalar, kMaxScalar, kMaxScalar), hi(-kMaxScalar, -kMaxScalar, -kMaxScalar);
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
        const AABB a = fat[i];
        const Vec3 c = scalar(0.5) * (a.lowerBound + a.upperBound);
        lo = Min(lo, c);
        hi = Max(hi, c);
    }
    red[0][threadIdx.x] = lo.x; red[1][threadIdx.x] = lo.y; red[2][threadIdx.x] = lo.z;
    red[3][threadIdx.x] = hi.x; red[4][threadIdx.x] = hi.y; red[5][threadIdx.x] = hi.z;
    __syncthreads();
    for (u32 s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (u32 k = 0; k < 3; ++k) red[k][threadIdx.x] = Min(red[k][threadIdx.x], red[k][threadIdx.x + s]);
            for (u32 k = 3; k < 6; ++k) red[k][threadIdx.x] = Max(red[k][threadIdx.x], red[k][threadIdx.x + s]);
        }
        __syncthreads();
    }
    const Vec3 bmin(red[0][0], red[1][0], red[2][0]);
    const Vec3 ext(red[3][0] - red[0][0], red[4][0] - red[1][0], red[5][0] - red[2][0]);
    const Vec3 inv(ext.x > scalar(0) ? scalar(1) / ext.x : scalar(0), ext.y > scalar(0) ? scalar(1) / ext.y : scalar(0),
                   ext.z > scalar(0) ? scalar(1) / ext.z : scalar(0));
    // keys (Morton code, fixture), padded to a power of two with the largest key
    u32 P = 1;
    while (P < n) P <<= 1;
    u64* keys = P <= kBvhLdsKeys ? ldsKeys : reinterpret_cast<u64*>(nodes);
    for (u32 i = threadIdx.x; i < P; i += blockDim.x) {
        u64 key = ~u64(0);
        if (i < n) {
            const AABB a = fat[i];
            const Vec3 c = scalar(0.5) * (a.lowerBound + a.upperBound) - bmin;
            const u32 code = (Spread10(Quantize10(c.x * inv.x)) << 2) | (Spread10(Quantize10(c.y * inv.y)) << 1) |
                             Spread10(Quantize10(c.z * inv.z));
            key = (u64(code) << 32) | i;
        }
        keys[i] = key;
    }
    __syncthreads();
    BlockBitonicSort(keys, P);
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) leaves[i] = u32(keys[i]);
    __syncthreads();
    // node levels, bottom-up (the first level reads the sorted fixtures' fat AABBs)
    const u32 L = BvhLevels(n);
    u32 childCount = n, off = 0, prevOff = 0;
    for (u32 l = 1; l <= L; ++l) {
        const u32 count = (childCount + kBvhArity - 1) / kBvhArity;
        for (u32 j = threadIdx.x; j < count; j += blockDim.x) {
            const u32 c0 = j * kBvhArity, c1 = Min(c0 + kBvhArity, childCount);
            AABB u = l == 1 ? fat[leaves[c0]] : nodes[prevOff + c0];
            for (u32 c = c0 + 1; c < c1; ++c) u = Combine(u, l == 1 ? fat[leaves[c]] : nodes[prevOff + c]);
            nodes[off + j] = u;
        }
        __syncthreads();
        prevOff = off;
        off += count;
        childCount = count;
    }
}

// The stackless walk of a world's hierarchy: visits(box) decides whether to enter an internal node;
// leaf(fixture) handles a fixture that reached level 0.
template <class Visit, class Leaf>
__device__ __forceinline__ void WalkBvh(const BatchView& B, const WorldInfo& wi, Visit visits, Leaf leaf) {
    const u32 n = wi.fixtureCount;
    if (n == 0) return;
    const u32* leaves = B.bvhLeaves + wi.fixtureBase;
    const AABB* nodes = B.bvhNodes + wi.fixtureBase;
    const u32 L = BvhLevels(n);
    u32 l = L, i = 0;
    u32 count = L == 0 ? n : 1u;  // nodes in the current level
    for (;;) {
        bool descend = false;
        if (l == 0) {
            leaf(leaves[i]);
        } else {
            descend = visits(nodes[BvhLevelOffset(n, l) + i]);
        }
        if (descend) {
            --l;
            i *= kBvhArity;
            count = BvhLevelCount(n, l);
            continue;
        }
        // next sibling, climbing past the last child of each group
        for (;;) {
            if (l == L) {
                if (L == 0 && i + 1 < count) { ++i; break; }  // a world of one level: its fixtures in turn
                return;
            }
            if ((i + 1) % kBvhArity != 0 && i + 1 < count) { ++i; break; }
            i /= kBvhArity;
            ++l;
            count = BvhLevelCount(n, l);
        }
    }
}

// (fraction, fixture) order of hits: fraction first, then the lower fixture index.
__device__ __forceinline__ bool HitBefore(scalar fa, u32 a, scalar fb, u32 b) { return fa < fb || (fa == fb && a < b); }

// One lane per query.
__global__ void __launch_bounds__(kBlock) k_query(BatchView B, QueryParams qp) {
    const u32 qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= qp.count) return;
    const uint4 q = qp.queries[qi];
    BX_WORLD_HOOK(B, q.x);
    const WorldInfo wi = B.worlds[q.x];
    const uint8_t* mask = q.y == kNull ? nullptr : qp.mask + q.y;
    const float* cq = qp.coords + 6 * u64(qi);
    const Vec3 a(cq[0], cq[1], cq[2]), b(cq[3], cq[4], cq[5]);
    if (qp.kind == 2u) {  // QueryAABB: fixtures in creation order, the first perQuery kept
        AABB box;
        box.lowerBound = a;
        box.upperBound = b;
        u32* out = qp.fixtures + u64(qi) * qp.perQuery;
        u32 found = 0;
        WalkBvh(B, wi, [&](const AABB& node) { return TestOverlap(node, box); }, [&](u32 f) {
            if (!TestOverlap(B.fatAABB[wi.fixtureBase + f], box) || (mask && !mask[f])) return;
            u32 k = Min(found, qp.perQuery);
            while (k > 0 && out[k - 1] > f) {
                if (k < qp.perQuery) out[k] = out[k - 1];
                --k;
            }
            if (k < qp.perQuery) out[k] = f;
            ++found;
        });
        qp.counts[qi] = found;
        return;
    }
    const Vec3 r = RayDirection(a, b);
    const scalar maxFraction = scalar(1);
    const Vec3 q2 = a + maxFraction * (b - a);
    RayIn in;
    in.p1 = a;
    in.p2 = b;
    in.maxFraction = maxFraction;
    RayHitOut* out = qp.hits + (qp.kind == 0u ? u64(qi) : u64(qi) * qp.perQuery);
    u32 found = 0;
    RayHitOut best;
    best.fixture = kNull;
    best.fraction = scalar(0);
    best.point.SetZero();
    best.normal.SetZero();
    WalkBvh(B, wi, [&](const AABB& node) { return SegmentBoxOverlap(a, q2, node); }, [&](u32 f) {
        if (!SegmentOverlapsBox(a, q2, r, B.fatAABB[wi.fixtureBase + f]) || (mask && !mask[f])) return;
        RayOut o;
        if (!FixtureRay(B, wi, f, in, &o)) return;
        RayHitOut h;
        h.fixture = f;
        h.fraction = o.fraction;
        h.point = (scalar(1) - o.fraction) * a + o.fraction * b;
        h.normal = o.normal;
        if (qp.kind == 0u) {
            if (best.fixture == kNull || HitBefore(h.fraction, f, best.fraction, best.fixture)) best = h;
            return;
        }
        u32 k = Min(found, qp.perQuery);
        while (k > 0 && HitBefore(h.fraction, f, out[k - 1].fraction, out[k - 1].fixture)) {
            if (k < qp.perQuery) out[k] = out[k - 1];
            --k;
        }
        if (k < qp.perQuery) out[k] = h;
        ++found;
    });
    if (qp.kind == 0u) *out = best;
    else qp.counts[qi] = found;
}
