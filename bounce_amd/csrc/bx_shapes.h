// bx_shapes.h -- shape mass properties and body mass reset (host side of world construction).
//   hull:    src/bounce/collision/shapes/hull_shape.cpp:40-158 (Kallay's polyhedral integrals)
//   sphere:  src/bounce/collision/shapes/sphere_shape.cpp
//   capsule: src/bounce/collision/shapes/capsule_shape.cpp
//   b3Body::ResetMass: src/bounce/dynamics/body.cpp:265-376
#pragma once
#include "bx_types.h"

namespace bx {

struct MassData { Vec3 center; scalar mass; Mat33 I; };

BX_HD void HullMass(const HullView& h, scalar density, MassData* md) {
    Vec3 s;
    s.SetZero();
    for (u32 i = 0; i < h.vertexCount; ++i) s += h.vertices[i];
    s /= scalar(h.vertexCount);
    scalar volume = 0;
    Vec3 center;
    center.SetZero();
    scalar xx = 0, xy = 0, yy = 0, xz = 0, zz = 0, yz = 0;
    for (u32 i = 0; i < h.faceCount; ++i) {
        u32 begin = h.faces[i];
        u32 edge = h.edges[begin].next;
        do {
            u32 next = h.edges[edge].next;
            Vec3 v1 = h.vertices[h.edges[begin].origin] - s;
            Vec3 v2 = h.vertices[h.edges[edge].origin] - s;
            Vec3 v3 = h.vertices[h.edges[next].origin] - s;
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
        } while (h.edges[edge].next != begin);
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

BX_HD void SphereMass(const Vec3& c, scalar r, scalar density, MassData* md) {
    scalar volume = (scalar(4) / scalar(3)) * kPi * r * r * r;
    scalar mass = density * volume;
    scalar I = mass * (scalar(0.4) * r * r + Dot(c, c));
    md->center = c;
    md->mass = mass;
    md->I = Diagonal(I, I, I);
}

BX_HD Mat33 RotateToFrameM(const Mat33& I, const Mat33& R) { return R * I * Transpose(R); }

BX_HD void CapsuleMass(const Vec3& A, const Vec3& B, scalar r, scalar density, MassData* md) {
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

}  // namespace bx
