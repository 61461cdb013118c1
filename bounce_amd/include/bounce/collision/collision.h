// bounce/collision/collision.h -- the collision value types the drop-in API hands back: b3AABB,
// the manifold types a b3Contact exposes (b3FeaturePair, b3ContactID, b3ManifoldPoint, b3Manifold)
// and the ray-cast input and output. Names and members follow the reference's
// include/bounce/collision/{collision,clip}.h and geometry/aabb.h.
#pragma once
#include "../common/math.h"

#define B3_MAX_MANIFOLD_POINTS 4
#define B3_NULL_TRIANGLE 0xFFFFFFFF

struct b3AABB {
    b3AABB() {}
    b3Vec3 GetCenter() const { return scalar(0.5) * (lowerBound + upperBound); }
    b3Vec3 GetExtents() const { return scalar(0.5) * (upperBound - lowerBound); }
    bool Contains(const b3Vec3& p) const {
        return lowerBound.x <= p.x && p.x <= upperBound.x && lowerBound.y <= p.y && p.y <= upperBound.y &&
               lowerBound.z <= p.z && p.z <= upperBound.z;
    }
    b3Vec3 lowerBound;
    b3Vec3 upperBound;
};

// Inclusive overlap of two boxes (aabb.h:347-367).
inline bool b3TestOverlap(const b3AABB& a, const b3AABB& b) {
    if (a.upperBound.x < b.lowerBound.x || a.lowerBound.x > b.upperBound.x) return false;
    if (a.upperBound.y < b.lowerBound.y || a.lowerBound.y > b.upperBound.y) return false;
    if (a.upperBound.z < b.lowerBound.z || a.lowerBound.z > b.upperBound.z) return false;
    return true;
}

struct b3RayCastInput {
    b3Vec3 p1;
    b3Vec3 p2;
    scalar maxFraction;
};

struct b3RayCastOutput {
    scalar fraction;
    b3Vec3 normal;
};

struct b3FeaturePair {
    uint32 inEdge1;   // incoming side plane edge on hull 1
    uint32 outEdge1;  // outgoing side plane edge on hull 1
    uint32 inEdge2;   // incoming edge on hull 2
    uint32 outEdge2;  // outgoing edge on hull 2
};

struct b3Key128 {
    uint64 key1;
    uint64 key2;
};

union b3PairKey {
    b3FeaturePair pair;
    b3Key128 key;
};

struct b3ContactID {
    bool operator==(const b3ContactID& o) const {
        return triangleKey == o.triangleKey && key.key.key1 == o.key.key.key1 && key.key.key2 == o.key.key.key2;
    }
    uint32 triangleKey;
    b3PairKey key;
};

struct b3ManifoldPoint {
    b3Vec3 localNormal1;  // normal on the first shape, in its frame
    b3Vec3 localPoint1;   // point on the first shape without its radius, in its frame
    b3Vec3 localPoint2;   // point on the second shape without its radius, in its frame
    scalar normalImpulse;
    b3ContactID id;
    bool persisting;
    bool edgeContact;
};

struct b3Manifold {
    b3ManifoldPoint points[B3_MAX_MANIFOLD_POINTS];
    uint32 pointCount;
    b3Vec2 tangentImpulse;
    scalar motorImpulse;
    scalar motorSpeed;
    scalar tangentSpeed1;
    scalar tangentSpeed2;
};

// A manifold point in world space (collision.cpp:63-75): the midpoint of the two surface points, the
// normal of the first shape and the separation.
struct b3WorldManifoldPoint {
    void Initialize(const b3ManifoldPoint* p, scalar rA, const b3Transform& xfA, scalar rB, const b3Transform& xfB) {
        b3Vec3 nA = b3Mul(xfA.rotation, p->localNormal1);
        b3Vec3 cA = xfA * p->localPoint1;
        b3Vec3 cB = xfB * p->localPoint2;
        b3Vec3 pA = cA + rA * nA;
        b3Vec3 pB = cB - rB * nA;
        point = scalar(0.5) * (pA + pB);
        normal = nA;
        separation = b3Dot(cB - cA, nA) - rA - rB;
    }
    b3Vec3 point;
    b3Vec3 normal;
    scalar separation;
};

// A manifold in world space (collision.cpp:77-100): its points, their mean, the normalized mean normal
// and a tangent basis.
struct b3WorldManifold {
    void Initialize(const b3Manifold* m, scalar rA, const b3Transform& xfA, scalar rB, const b3Transform& xfB) {
        center.SetZero();
        normal.SetZero();
        pointCount = m->pointCount;
        for (uint32 i = 0; i < pointCount; ++i) {
            points[i].Initialize(m->points + i, rA, xfA, rB, xfB);
            center += points[i].point;
            normal += points[i].normal;
        }
        if (pointCount > 0) {
            center /= scalar(pointCount);
            normal.Normalize();
            tangent1 = b3Perp(normal);
            tangent2 = b3Cross(tangent1, normal);
        }
    }
    uint32 pointCount;
    b3WorldManifoldPoint points[B3_MAX_MANIFOLD_POINTS];
    b3Vec3 center;
    b3Vec3 normal;
    b3Vec3 tangent1;
    b3Vec3 tangent2;
};
