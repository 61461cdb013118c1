// CPU build of the device library (test tool): adversarial fuzz of the device narrowphase routines
// (bx_collide.h, bx_convex.h) under AddressSanitizer and UndefinedBehaviorSanitizer.
//
// Every convex shape pair the device path collides (sphere, capsule, hull in each registered order),
// plus GJK and the sensor overlap test, over shapes and poses drawn to be hard: coincident and nearly
// coincident centres, parallel and zero-length capsules, zero and denormal radii, huge coordinates,
// non-unit, zero and NaN rotations, NaN and infinite positions, and hull pairs collided repeatedly
// under small perturbations so that the feature cache is read back in every state. After every call
// it asserts that the manifold has at most four points, that every feature id is a half-edge of the
// hull it names (or null) and that the feature cache indexes the hulls' faces and edges; a call that
// does not return within a few seconds (a face loop that never ends) aborts the run.
//
//   tests/cpu_build/_out/fuzz_narrowphase scene.bin iterations seed
#include "hip/hip_runtime.h"
#include "../../bounce_amd/csrc/bx_step.hip"

#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <unistd.h>
#include <vector>

namespace bx {
thread_local u32 lds[16384];
}
extern "C" void* bxcpu_lds_buffer(size_t* bytes) {
    *bytes = sizeof(bx::lds);
    return bx::lds;
}

using namespace bx;

namespace {

struct HullData {
    std::vector<Vec3> v;
    std::vector<HalfEdge> e;
    std::vector<u32> f;
    std::vector<Plane> p;
    Vec3 c;
    HullView View() const {
        HullView h;
        h.vertices = v.data();
        h.edges = e.data();
        h.faces = f.data();
        h.planes = p.data();
        h.centroid = c;
        h.vertexCount = u32(v.size());
        h.edgeCount = u32(e.size());
        h.faceCount = u32(f.size());
        return h;
    }
};

std::vector<HullData> ReadHulls(const char* path) {
    FILE* fp = std::fopen(path, "rb");
    if (!fp) std::exit(2);
    auto u = [&]() {
        uint32_t x;
        if (std::fread(&x, 4, 1, fp) != 1) std::exit(2);
        return x;
    };
    char magic[4];
    if (std::fread(magic, 1, 4, fp) != 4 || u() != 1) std::exit(2);
    std::vector<HullData> hulls(u());
    for (HullData& h : hulls) {
        u32 nv = u(), ne = u(), nf = u();
        h.v.resize(nv);
        h.e.resize(ne);
        h.f.resize(nf);
        h.p.resize(nf);
        float c[3];
        if (std::fread(h.v.data(), sizeof(Vec3), nv, fp) != nv || std::fread(h.e.data(), sizeof(HalfEdge), ne, fp) != ne ||
            std::fread(h.f.data(), 4, nf, fp) != nf || std::fread(h.p.data(), sizeof(Plane), nf, fp) != nf ||
            std::fread(c, 4, 3, fp) != 3)
            std::exit(2);
        h.c = Vec3(c[0], c[1], c[2]);
    }
    std::fclose(fp);
    return hulls;
}

std::mt19937_64 rng;
float U(float a, float b) { return std::uniform_real_distribution<float>(a, b)(rng); }
u32 Pick(u32 n) { return u32(std::uniform_int_distribution<uint64_t>(0, n - 1)(rng)); }

float Special() {
    static const float v[] = {0.0f, -0.0f, 1e-40f, -1e-40f, 1e-30f, 1e-7f, 1e30f, -1e30f, 3.4e38f, INFINITY, -INFINITY,
                              NAN};
    return v[Pick(sizeof(v) / sizeof(v[0]))];
}

Quat RandomRotation() {
    Quat q;
    switch (Pick(8)) {
    case 0: q = Quat(Vec3(0, 0, 0), 1); break;
    case 1: q = Quat(Vec3(0, 0, 0), 0); break;                       // zero
    case 2: q = Quat(Vec3(U(-1, 1), U(-1, 1), U(-1, 1)), U(-1, 1)); break;  // not normalized
    case 3: q = Quat(Vec3(Special(), 0, 0), 1); break;
    default: {
        q = Quat(Vec3(U(-1, 1), U(-1, 1), U(-1, 1)), U(-1, 1));
        q.Normalize();
    }
    }
    return q;
}

Vec3 RandomOffset(const Vec3& base) {
    switch (Pick(10)) {
    case 0: return base;                                                  // coincident
    case 1: return base + Vec3(1e-7f, 0, 0);                              // nearly
    case 2: return base + Vec3(1e-40f, -1e-40f, 1e-40f);                   // denormal offset
    case 3: return Vec3(Special(), U(-2, 2), U(-2, 2));
    case 4: return base + Vec3(U(-0.01f, 0.01f), U(-0.01f, 0.01f), U(-0.01f, 0.01f));
    default: return base + Vec3(U(-2.5f, 2.5f), U(-2.5f, 2.5f), U(-2.5f, 2.5f));
    }
}

Fixture RandomShape(u32 type, u32 hullCount) {
    Fixture f;
    std::memset(&f, 0, sizeof(f));
    f.shapeType = type;
    if (type == 0u) {  // sphere
        f.radius = Pick(6) == 0 ? std::fabs(Special()) : U(0.05f, 1.0f);
        f.v1 = Pick(4) == 0 ? Vec3(U(-1, 1), U(-1, 1), U(-1, 1)) : Vec3(0, 0, 0);
    } else if (type == 1u) {  // capsule
        float half = Pick(5) == 0 ? (Pick(2) ? 0.0f : 1e-6f) : U(0.1f, 1.5f);
        f.v1 = Vec3(-half, 0, 0);
        f.v2 = Vec3(half, 0, 0);
        f.radius = Pick(8) == 0 ? std::fabs(Special()) : U(0.05f, 0.6f);
    } else {  // hull
        f.shapeIndex = Pick(hullCount);
        f.radius = 0;
    }
    return f;
}

void Fail(const char* what, u32 iter) {
    std::fprintf(stderr, "fuzz_narrowphase: iteration %u: %s\n", iter, what);
    std::abort();
}

void CheckManifold(const Manifold& m, const HullView* h1, const HullView* h2, u32 iter) {
    if (m.pointCount > kMaxManifoldPoints) Fail("more than four manifold points", iter);
    const u32 lim = std::max(h1 ? h1->edgeCount : 2u, h2 ? h2->edgeCount : 2u);
    for (u32 i = 0; i < m.pointCount; ++i)
        for (u32 k = 0; k < 4; ++k)
            if (m.points[i].key[k] != kNull && m.points[i].key[k] >= std::max(lim, 2u)) Fail("feature id out of range", iter);
}

void CheckCache(const FeatureCache& c, const HullView& h1, const HullView& h2, u32 iter) {
    if (c.cacheType == kCacheEmpty) return;
    if (c.featureType == kFeatFace1 && c.index1 >= h1.faceCount) Fail("cached face of hull 1 out of range", iter);
    if (c.featureType == kFeatFace2 && c.index1 >= h2.faceCount) Fail("cached face of hull 2 out of range", iter);
    if (c.featureType == kFeatEdges && (c.index1 + 1 >= h1.edgeCount || c.index2 + 1 >= h2.edgeCount))
        Fail("cached edge pair out of range", iter);
    if (c.featureType > kFeatFace2) Fail("cached feature type out of range", iter);
}

void OnAlarm(int) {
    static const char msg[] = "fuzz_narrowphase: a call did not return (loop over a face that never ends?)\n";
    (void)!write(2, msg, sizeof(msg) - 1);
    std::abort();
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s scene.bin iterations seed\n", argv[0]);
        return 2;
    }
    std::vector<HullData> hulls = ReadHulls(argv[1]);
    const u32 iterations = u32(std::strtoul(argv[2], nullptr, 10));
    rng.seed(std::strtoull(argv[3], nullptr, 10));
    std::signal(SIGALRM, OnAlarm);
    uint64_t points = 0, calls = 0;
    FaceScratch scratch;
    for (u32 it = 0; it < iterations; ++it) {
        alarm(10);
        // registered convex pairs in primary order (contact.cpp:60-71): A's type >= B's type
        static const u32 pairs[][2] = {{0, 0}, {1, 0}, {1, 1}, {3, 0}, {3, 1}, {3, 3}};
        const u32* pt = pairs[Pick(6)];
        Fixture fA = RandomShape(pt[0], u32(hulls.size())), fB = RandomShape(pt[1], u32(hulls.size()));
        HullView hA, hB;
        if (fA.shapeType == 3u) hA = hulls[fA.shapeIndex].View();
        if (fB.shapeType == 3u) hB = hulls[fB.shapeIndex].View();
        Vec3 pA(U(-3, 3), U(-3, 3), U(-3, 3));
        Transform xfA(pA, RandomRotation()), xfB(RandomOffset(pA), RandomRotation());
        // the sensor overlap test and GJK with and without radii, for every pair
        GJKOut q = GJK(xfA, MakeProxy(fA, &hA), xfB, MakeProxy(fB, &hB), Pick(2) != 0);
        if (q.simplexCount > 4) Fail("GJK simplex of more than four vertices", it);
        (void)ConvexTestOverlap(xfA, MakeProxy(fA, &hA), xfB, MakeProxy(fB, &hB));
        Manifold m;
        std::memset(&m, 0, sizeof(m));
        ++calls;
        if (pt[0] == 3u && pt[1] == 3u) {
            FeatureCache cache{0, 0, 0, 0};
            // the same pair over a short drift: every cache state is read back
            for (u32 k = 0; k < 6; ++k) {
                Transform xf0A = xfA, xf0B = xfB;
                if (k) {
                    xfB.translation = xfB.translation + Vec3(U(-0.05f, 0.05f), U(-0.05f, 0.05f), U(-0.05f, 0.05f));
                    if (Pick(3) == 0) xfB.rotation = RandomRotation();
                }
                std::memset(&m, 0, sizeof(m));
                u32 overflow = 0;
                CollideHulls(m, xfA, hA, fA.radius, xfB, hB, fB.radius, cache, xf0A, xf0B, scratch, overflow);
                CheckManifold(m, &hA, &hB, it);
                CheckCache(cache, hA, hB, it);
                points += m.pointCount;
            }
        } else if (pt[0] == 0u) {
            CollideSpheres(m, xfA, fA, xfB, fB);
        } else if (pt[0] == 1u && pt[1] == 0u) {
            CollideCapsuleAndSphere(m, xfA, fA, xfB, fB);
        } else if (pt[0] == 1u) {
            CollideCapsules(m, xfA, fA, xfB, fB);
        } else if (pt[1] == 0u) {
            CollideHullAndSphere(m, xfA, hA, fA, xfB, fB);
        } else {
            CollideHullAndCapsule(m, xfA, hA, fA, xfB, fB);
        }
        CheckManifold(m, fA.shapeType == 3u ? &hA : nullptr, fB.shapeType == 3u ? &hB : nullptr, it);
        points += m.pointCount;
    }
    alarm(0);
    std::printf("fuzz_narrowphase: %u iterations, %llu collide calls, %llu manifold points, all in range\n",
                iterations, (unsigned long long)calls, (unsigned long long)points);
    return 0;
}
