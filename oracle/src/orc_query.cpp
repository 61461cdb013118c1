// orc_query.cpp -- oracle (test infrastructure): the world queries restated on CPU.
//   world.cpp:452-566          b3World::RayCast, RayCastSingle (the broadphase ray wrappers)
//   world.cpp:852-870          b3World::QueryAABB
//   dynamic_tree.h:201-339     b3DynamicTree::RayCast: segment AABB + separating-axis test per node
//   static_tree.h:150-290      b3StaticTree::RayCast (the same test; mesh child triangles)
//   sphere_shape.cpp:69-121, capsule_shape.cpp:180-324, hull_shape.cpp:189-281,
//   triangle_shape.cpp:95-176, mesh_shape.cpp:81-146   b3Shape::RayCast of each shape type
//   world.cpp:568-850          b3World::ShapeCast, ShapeCastSingle (b3TimeOfImpact of the translating
//                              shape against each candidate fixture or mesh triangle)
// The dynamic tree's node shapes are not reproduced: a leaf (fixture proxy) is a candidate exactly
// when its fat AABB passes the segment test with the ray's full extent, which is what the tree
// reports for a listener that does not clip the ray (RayCastSingle's wrapper, or a listener
// returning 1). Hits are returned in increasing fraction, ties by fixture creation index.
#include <algorithm>
#include "orc_world.h"

namespace orc {

struct RayIn { Vec3 p1, p2; scalar maxFraction; };
struct RayOut { scalar fraction; Vec3 normal; };

// The per-node test of b3DynamicTree::RayCast / b3StaticTree::RayCast for the segment p1..q2 with
// direction r (normalized p2 - p1): segment AABB overlap, then the six separating axes.
static bool SegmentOverlapsBox(const Vec3& p1, const Vec3& q2, const Vec3& r, const AABB& box) {
    AABB seg;
    seg.lowerBound = Min(p1, q2);
    seg.upperBound = Max(p1, q2);
    if (!TestOverlap(seg, box)) return false;
    const Vec3 c = scalar(0.5) * (box.lowerBound + box.upperBound);
    const Vec3 h = scalar(0.5) * (box.upperBound - box.lowerBound);
    const Vec3 s = p1 - c, t = q2 - c;
    if (Abs(s.x + t.x) - (Abs(s.x - t.x) + scalar(2) * h.x) > scalar(0)) return false;
    if (Abs(s.y + t.y) - (Abs(s.y - t.y) + scalar(2) * h.y) > scalar(0)) return false;
    if (Abs(s.z + t.z) - (Abs(s.z - t.z) + scalar(2) * h.z) > scalar(0)) return false;
    const Vec3 axes[3] = {Vec3(1, 0, 0), Vec3(0, 1, 0), Vec3(0, 0, 1)};
    for (int k = 0; k < 3; ++k) {
        const Vec3 v = Cross(axes[k], r);
        if (Abs(Dot(v, s)) - Dot(Abs(v), h) > scalar(0)) return false;
    }
    return true;
}

static Vec3 RayDirection(const Vec3& p1, const Vec3& p2) {  // r = p2 - p1; r.Normalize() (vec3.h:93-104)
    Vec3 r = p2 - p1;
    const scalar len = Sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
    if (len > kEpsilon) {
        const scalar s = scalar(1) / len;
        r.x *= s;
        r.y *= s;
        r.z *= s;
    }
    return r;
}

static bool SphereRay(const Shape& sh, const Transform& xf, const RayIn& in, RayOut* out) {
    const Vec3 d = in.p2 - in.p1;
    const scalar a = Dot(d, d);
    if (a < kEpsilon * kEpsilon) return false;
    const Vec3 m = in.p1 - Mul(xf, sh.v1);
    const scalar b = Dot(m, d);
    const scalar c = Dot(m, m) - sh.radius * sh.radius;
    const scalar disc = b * b - a * c;
    if (disc < scalar(0)) return false;
    scalar t = -b - Sqrt(disc);
    if (t > scalar(0) && t <= in.maxFraction * a) {
        t /= a;
        out->fraction = t;
        out->normal = Normalize(m + t * d);
        return true;
    }
    return false;
}

static bool CapsuleRay(const Shape& sh, const Transform& xf, const RayIn& in, RayOut* out) {
    const Vec3 A = in.p1, B = in.p2;
    Vec3 n = B - A;
    const scalar nn = Dot(n, n);
    if (nn < kEpsilon * kEpsilon) return false;
    const Vec3 P = Mul(xf, sh.v1), Q = Mul(xf, sh.v2);
    Vec3 d = Q - P;
    const scalar dd = Dot(d, d);
    const scalar rr = sh.radius * sh.radius;
    if (dd < kEpsilon * kEpsilon) {  // a point capsule: the sphere at P
        const scalar a = nn;
        const Vec3 m = A - P;
        const scalar b = Dot(m, n);
        const scalar c = Dot(m, m) - rr;
        const scalar disc = b * b - a * c;
        if (disc < scalar(0)) return false;
        scalar t = -b - Sqrt(disc);
        if (t > scalar(0) && t <= in.maxFraction * a) {
            t /= a;
            out->fraction = t;
            out->normal = Normalize(m + t * n);
            return true;
        }
        return false;
    }
    Vec3 m = A - P;
    const scalar nd = Dot(n, d), mm = Dot(m, m), md = Dot(m, d), mn = Dot(m, n);
    scalar a = dd * nn - nd * nd;
    if (Abs(a) < scalar(2) * (kEpsilon * kEpsilon)) return false;
    scalar b = dd * mn - nd * md;
    scalar c = dd * (mm - rr) - md * md;
    scalar disc = b * b - a * c;
    if (disc < scalar(0)) return false;
    scalar t = -b - Sqrt(disc);
    t /= a;
    if (t < scalar(0) || t > scalar(1)) return false;
    const scalar tp = md + t * nd;  // where along PQ the line meets the infinite cylinder
    Vec3 C;
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
    // the hemisphere at C
    d = n;
    a = nn;
    m = A - C;
    b = Dot(m, d);
    c = Dot(m, m) - rr;
    disc = b * b - a * c;
    if (disc < scalar(0)) return false;
    t = -b - Sqrt(disc);
    if (t > scalar(0) && t <= in.maxFraction * a) {
        t /= a;
        out->fraction = t;
        out->normal = Normalize(m + t * d);
        return true;
    }
    return false;
}

static bool HullRay(const Shape& sh, const Transform& xf, const RayIn& in, RayOut* out) {
    const Hull& h = *sh.hull;
    const Vec3 p1 = MulC(xf.rotation, in.p1 - xf.translation);
    const Vec3 p2 = MulC(xf.rotation, in.p2 - xf.translation);
    const Vec3 d = p2 - p1;
    scalar lower = scalar(0), upper = in.maxFraction;
    u32 index = kNullU32;
    for (u32 i = 0; i < h.faceCount(); ++i) {
        const scalar numerator = h.planes[i].offset - Dot(h.planes[i].normal, p1);
        const scalar denominator = Dot(h.planes[i].normal, d);
        if (denominator == scalar(0)) {
            if (numerator < scalar(0)) return false;
        } else {
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
    }
    if (index == kNullU32) return false;
    out->fraction = lower;
    out->normal = Mul(xf.rotation, h.planes[index].normal);
    return true;
}

static bool TriangleRay(const Vec3& v1, const Vec3& v2, const Vec3& v3, const Transform& xf, const RayIn& in,
                        RayOut* out) {
    const Vec3 p1 = MulT(xf, in.p1), p2 = MulT(xf, in.p2);
    const scalar maxFraction = in.maxFraction;
    const Vec3 d = p2 - p1;
    if (LengthSquared(d) < kEpsilon * kEpsilon) return false;
    Vec3 n = Cross(v2 - v1, v3 - v1);
    const scalar len = Length(n);
    if (len == scalar(0)) return false;
    n /= len;  // b3Vec3::operator/= multiplies by the reciprocal
    const scalar num = Dot(n, v1 - p1);
    const scalar den = Dot(n, d);
    if (den == scalar(0)) return false;
    const scalar fraction = num / den;
    if (fraction < scalar(0) || maxFraction < fraction) return false;
    const Vec3 Q = p1 + fraction * d;
    const Vec3 A = v1, B = v2, C = v3;
    const Vec3 AB = B - A, AC = C - A;
    const Vec3 QA = A - Q, QB = B - Q, QC = C - Q;
    const Vec3 QB_x_QC = Cross(QB, QC), QC_x_QA = Cross(QC, QA), QA_x_QB = Cross(QA, QB);
    const Vec3 AB_x_AC = Cross(AB, AC);
    const scalar u = Dot(QB_x_QC, AB_x_AC), v = Dot(QC_x_QA, AB_x_AC), w = Dot(QA_x_QB, AB_x_AC);
    if (u >= scalar(0) && v >= scalar(0) && w >= scalar(0)) {
        out->fraction = fraction;
        out->normal = num > scalar(0) ? Mul(xf.rotation, -n) : Mul(xf.rotation, n);
        return true;
    }
    return false;
}

// b3MeshShape::RayCast: the static tree walked in stack order with the ray in the unscaled mesh
// frame (never clipped), every candidate triangle cast in the world frame, the first strictly
// smallest fraction kept.
static bool MeshRay(const Shape& sh, const Transform& xf, const RayIn& in, RayOut* out) {
    const Mesh& m = *sh.mesh;
    if (m.nodes.empty()) return false;
    const Vec3 inv(scalar(1) / sh.scale.x, scalar(1) / sh.scale.y, scalar(1) / sh.scale.z);
    const Vec3 tp1 = Mul(inv, MulT(xf, in.p1)), tp2 = Mul(inv, MulT(xf, in.p2));
    const Vec3 r = RayDirection(tp1, tp2);
    const Vec3 q2 = tp1 + in.maxFraction * (tp2 - tp1);
    bool hit = false;
    RayOut best;
    best.fraction = kMaxScalar;
    std::vector<u32> stack;
    stack.push_back(m.root);
    while (!stack.empty()) {
        const u32 id = stack.back();
        stack.pop_back();
        if (id == kNullU32) continue;
        const StaticNode& node = m.nodes[id];
        if (!SegmentOverlapsBox(tp1, q2, r, node.aabb)) continue;
        if (node.child1 == kNullU32) {
            const MeshTriangle& t = m.triangles[node.index];
            RayOut o;
            if (TriangleRay(Mul(sh.scale, m.vertices[t.v1]), Mul(sh.scale, m.vertices[t.v2]),
                            Mul(sh.scale, m.vertices[t.v3]), xf, in, &o) &&
                o.fraction < best.fraction) {
                hit = true;
                best = o;
            }
        } else {
            stack.push_back(node.child1);
            stack.push_back(node.child2);
        }
    }
    *out = best;
    return hit;
}

static bool FixtureRay(const World& w, u32 f, const RayIn& in, RayOut* out) {  // b3Fixture::RayCast
    const Fixture& fx = w.fixtures[f];
    const Transform& xf = w.bodies[fx.body].xf;
    switch (fx.shape.type) {
    case e_sphere: return SphereRay(fx.shape, xf, in, out);
    case e_capsule: return CapsuleRay(fx.shape, xf, in, out);
    case e_triangle: return TriangleRay(fx.shape.v1, fx.shape.v2, fx.shape.v3, xf, in, out);
    case e_hull: return HullRay(fx.shape, xf, in, out);
    case e_mesh: return MeshRay(fx.shape, xf, in, out);
    }
    return false;
}

struct RayHit { u32 fixture; scalar fraction; Vec3 point, normal; };

// Every fixture the ray p1..p2 hits (ShouldRayCast(fixture) when mask is given), as a listener that
// never clips sees them, in increasing fraction (ties by fixture index).
void RayCastAll(const World& w, const Vec3& p1, const Vec3& p2, const uint8_t* mask, std::vector<RayHit>& hits) {
    hits.clear();
    const Vec3 r = RayDirection(p1, p2);
    const scalar maxFraction = scalar(1);
    const Vec3 q2 = p1 + maxFraction * (p2 - p1);
    for (u32 f = 0; f < w.fixtures.size(); ++f) {
        if (!SegmentOverlapsBox(p1, q2, r, w.fatAABB[w.fixtures[f].proxyId])) continue;
        if (mask && !mask[f]) continue;
        RayIn in{p1, p2, maxFraction};
        RayOut o;
        if (!FixtureRay(w, f, in, &o)) continue;
        RayHit h;
        h.fixture = f;
        h.fraction = o.fraction;
        h.point = (scalar(1) - o.fraction) * p1 + o.fraction * p2;
        h.normal = o.normal;
        hits.push_back(h);
    }
    std::stable_sort(hits.begin(), hits.end(), [](const RayHit& a, const RayHit& b) { return a.fraction < b.fraction; });
}

// b3World::QueryAABB: fixtures whose fat AABB overlaps `aabb` (b3TestOverlap), in creation order.
void QueryAABB(const World& w, const AABB& aabb, const uint8_t* mask, std::vector<u32>& out) {
    out.clear();
    for (u32 f = 0; f < w.fixtures.size(); ++f)
        if (TestOverlap(w.fatAABB[w.fixtures[f].proxyId], aabb) && (!mask || mask[f])) out.push_back(f);
}

// ------------------------------------------------------------------ shape casts
struct ShapeHit { u32 fixture, child; scalar fraction; Vec3 point, normal; };

// The world query wrapper's Evaluate (world.cpp:703-716): the closest points at the hit; the point on
// the fixture's surface and the normal from the fixture towards the cast shape.
static void EvaluateHit(Vec3* point, Vec3* normal, const Transform& xA, const GJKProxy& pA, const Transform& xB,
                        const GJKProxy& pB) {
    GJKOutput q = GJK(xA, pA, xB, pB, false);
    Vec3 n = Normalize(q.point1 - q.point2);
    *point = q.point2 + pB.radius * n;
    *normal = n;
}

// b3GJKProxy of child triangle `tri` of a mesh shape (gjk.cpp:71-86): its scaled vertices.
static void MeshTriangleProxy(GJKProxy& p, const Shape& mesh, u32 tri) {
    const Mesh& m = *mesh.mesh;
    const MeshTriangle& t = m.triangles[tri];
    p.buffer[0] = Mul(mesh.scale, m.vertices[t.v1]);
    p.buffer[1] = Mul(mesh.scale, m.vertices[t.v2]);
    p.buffer[2] = Mul(mesh.scale, m.vertices[t.v3]);
    p.vertices = p.buffer;
    p.count = 3;
    p.radius = mesh.radius;
}

static AABB ScaleAABB(const AABB& a, const Vec3& s) {  // b3AABB::Scale (aabb.h:287-294)
    Vec3 lo = Mul(s, a.lowerBound), hi = Mul(s, a.upperBound);
    AABB r;
    r.lowerBound = Min(lo, hi);
    r.upperBound = Max(lo, hi);
    return r;
}

// Every touching cast of `shape` at `xf` moved by `d` against the fixtures (ShouldShapeCast(fixture) when
// mask is given), as a listener that never clips sees them: one per convex fixture, one per reported
// mesh triangle; in increasing (fraction, fixture, triangle).
void ShapeCastAll(const World& w, const Shape& shape, const Transform& xf, const Vec3& d, const uint8_t* mask,
                  std::vector<ShapeHit>& hits) {
    hits.clear();
    GJKProxy proxyA;
    proxyA.Set(shape);
    Transform xf2 = xf;
    xf2.translation += d;
    AABB a1, a2;
    ComputeAABB(shape, xf, &a1);
    ComputeAABB(shape, xf2, &a2);
    const AABB aabb = Combine(a1, a2);
    const Vec3 zero(scalar(0), scalar(0), scalar(0));
    auto cast = [&](u32 f, u32 child, const Transform& xfB, const GJKProxy& proxyB) {
        TOIOutput toi = TimeOfImpact(xf, proxyA, d, xfB, proxyB, zero, 20);
        if (toi.state != e_toiTouching) return;
        ShapeHit h;
        h.fixture = f;
        h.child = child;
        h.fraction = toi.t;
        Transform xt;
        xt.rotation = xf.rotation;
        xt.translation = xf.translation + toi.t * d;
        EvaluateHit(&h.point, &h.normal, xt, proxyA, xfB, proxyB);
        hits.push_back(h);
    };
    for (u32 f = 0; f < w.fixtures.size(); ++f) {
        if (!TestOverlap(w.fatAABB[w.fixtures[f].proxyId], aabb) || (mask && !mask[f])) continue;
        const Fixture& fx = w.fixtures[f];
        const Transform& xfB = w.bodies[fx.body].xf;
        if (fx.shape.type == e_mesh) {
            const Vec3 s = fx.shape.scale;
            const Vec3 inv(scalar(1) / s.x, scalar(1) / s.y, scalar(1) / s.z);
            AABB b1, b2;
            ComputeAABB(shape, MulT(xfB, xf), &b1);
            ComputeAABB(shape, MulT(xfB, xf2), &b2);
            std::vector<u32> tris;
            QueryMeshTree(*fx.shape.mesh, Combine(ScaleAABB(b1, inv), ScaleAABB(b2, inv)), tris);
            for (u32 t : tris) {
                GJKProxy proxyB;
                MeshTriangleProxy(proxyB, fx.shape, t);
                cast(f, t, xfB, proxyB);
            }
        } else {
            GJKProxy proxyB;
            proxyB.Set(fx.shape);
            cast(f, 0, xfB, proxyB);
        }
    }
    std::stable_sort(hits.begin(), hits.end(), [](const ShapeHit& a, const ShapeHit& b) {
        if (a.fraction != b.fraction) return a.fraction < b.fraction;
        if (a.fixture != b.fixture) return a.fixture < b.fixture;
        return a.child < b.child;
    });
}

}  // namespace orc
