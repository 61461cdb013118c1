// bx_math.h -- binary32 vector / quaternion / matrix algebra shared by host and device code.
//
// Operation order follows the reference's math headers exactly
// (include/bounce/common/math/{math,vec2,vec3,mat22,mat33,quat,transform,sweep}.h and
// src/bounce/common/math/math.cpp), because the step is required to be bit-exact against the
// reference's FP32 build: sums are evaluated left to right, b3Vec3::operator/= multiplies by the
// reciprocal while b3Vec2::Normalize and the quaternion scalar part divide, min/max are
// ternaries. Every translation unit that includes this header is compiled with
// -ffp-contract=off and IEEE division/sqrt (no fast-math), on host and device alike.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

#define BX_HD __host__ __device__ __forceinline__

namespace bx {

typedef float scalar;
typedef uint32_t u32;
typedef uint64_t u64;

// include/bounce/common/common.h:73-123, same literals evaluated in binary32.
constexpr scalar kEpsilon = 1.19209290e-07f;             // FLT_EPSILON
constexpr scalar kMaxScalar = 3.40282347e+38f;           // FLT_MAX
constexpr scalar kPi = scalar(3.14159265358979323846);
constexpr scalar kAabbExtension = scalar(0.2);
constexpr scalar kAabbMultiplier = scalar(2);
constexpr scalar kLinearSlop = scalar(0.005);
constexpr scalar kAngularSlop = scalar(2.0) / scalar(180) * kPi;
constexpr scalar kHullRadius = scalar(0.0) * kLinearSlop;
constexpr scalar kMaxTranslation = scalar(2.0);
constexpr scalar kMaxTranslationSq = kMaxTranslation * kMaxTranslation;
constexpr scalar kMaxRotation = scalar(0.5) * kPi;
constexpr scalar kMaxRotationSq = kMaxRotation * kMaxRotation;
constexpr scalar kMaxLinearCorrection = scalar(0.2);
constexpr scalar kMaxAngularCorrection = scalar(8.0) / scalar(180) * kPi;
constexpr scalar kBaumgarte = scalar(0.1);
constexpr scalar kVelocityThreshold = scalar(1.0);
constexpr scalar kTimeToSleep = scalar(0.2);
constexpr u32 kMaxManifoldPoints = 4;
constexpr u32 kMaxManifolds = 3;
constexpr u32 kNull = 0xFFFFFFFFu;
// b3Perp / b3ComputeBasis keep sqrt(scalar(1) / scalar(3)) in a function-static; its binary32
// value is 0x1.279a74p-1.
constexpr scalar kSqrtInv3 = 0x1.279a74p-1f;

template <class T> BX_HD T Min(T a, T b) { return a < b ? a : b; }
template <class T> BX_HD T Max(T a, T b) { return a > b ? a : b; }
template <class T> BX_HD T Clamp(T a, T lo, T hi) { return Max(lo, Min(a, hi)); }
BX_HD scalar Abs(scalar x) { return fabsf(x); }
BX_HD scalar Sqrt(scalar x) { return sqrtf(x); }
template <class T> BX_HD void Swap(T& a, T& b) { T t = a; a = b; b = t; }
BX_HD scalar Sign(scalar x) { return scalar(scalar(0) < x) - scalar(x < scalar(0)); }

struct Vec2 {
    scalar x, y;
    Vec2() = default;
    BX_HD Vec2(scalar a, scalar b) : x(a), y(b) {}
    BX_HD void SetZero() { x = y = 0; }
    BX_HD void Set(scalar a, scalar b) { x = a; y = b; }
    BX_HD void operator+=(const Vec2& b) { x += b.x; y += b.y; }
    BX_HD void operator-=(const Vec2& b) { x -= b.x; y -= b.y; }
    BX_HD void operator*=(scalar s) { x *= s; y *= s; }
    BX_HD scalar Normalize() {
        scalar len = Sqrt(x * x + y * y);
        if (len > kEpsilon) { x /= len; y /= len; }
        return len;
    }
};
BX_HD Vec2 operator+(const Vec2& a, const Vec2& b) { return Vec2(a.x + b.x, a.y + b.y); }
BX_HD Vec2 operator-(const Vec2& a, const Vec2& b) { return Vec2(a.x - b.x, a.y - b.y); }
BX_HD Vec2 operator-(const Vec2& a) { return Vec2(-a.x, -a.y); }
BX_HD Vec2 operator*(scalar s, const Vec2& v) { return Vec2(s * v.x, s * v.y); }
BX_HD scalar Dot(const Vec2& a, const Vec2& b) { return a.x * b.x + a.y * b.y; }
BX_HD scalar Det(const Vec2& a, const Vec2& b) { return a.x * b.y - a.y * b.x; }

struct Vec3 {
    scalar x, y, z;
    Vec3() = default;
    BX_HD Vec3(scalar a, scalar b, scalar c) : x(a), y(b), z(c) {}
    BX_HD scalar operator[](u32 i) const { return i == 0 ? x : (i == 1 ? y : z); }
    BX_HD void SetZero() { x = y = z = 0; }
    BX_HD void Set(scalar a, scalar b, scalar c) { x = a; y = b; z = c; }
    BX_HD void operator+=(const Vec3& b) { x += b.x; y += b.y; z += b.z; }
    BX_HD void operator-=(const Vec3& b) { x -= b.x; y -= b.y; z -= b.z; }
    BX_HD void operator*=(scalar s) { x *= s; y *= s; z *= s; }
    BX_HD void operator/=(scalar a) { scalar s = scalar(1) / a; x *= s; y *= s; z *= s; }
    BX_HD scalar Normalize() {
        scalar len = Sqrt(x * x + y * y + z * z);
        if (len > kEpsilon) { scalar s = scalar(1) / len; x *= s; y *= s; z *= s; }
        return len;
    }
};
BX_HD Vec3 operator-(const Vec3& v) { return Vec3(-v.x, -v.y, -v.z); }
BX_HD Vec3 operator+(const Vec3& a, const Vec3& b) { return Vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
BX_HD Vec3 operator-(const Vec3& a, const Vec3& b) { return Vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
BX_HD Vec3 operator*(scalar s, const Vec3& v) { return Vec3(s * v.x, s * v.y, s * v.z); }
BX_HD Vec3 operator*(const Vec3& v, scalar s) { return s * v; }
BX_HD Vec3 operator/(const Vec3& v, scalar s) { return Vec3(v.x / s, v.y / s, v.z / s); }
BX_HD scalar Dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
BX_HD Vec3 Cross(const Vec3& a, const Vec3& b) {
    return Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
BX_HD scalar Det(const Vec3& a, const Vec3& b, const Vec3& c) { return Dot(a, Cross(b, c)); }
BX_HD scalar Length(const Vec3& v) { return Sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
BX_HD scalar LengthSquared(const Vec3& v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
BX_HD Vec3 Normalize(const Vec3& v) {
    scalar len = Length(v);
    if (len > kEpsilon) { scalar s = scalar(1) / len; return s * v; }
    return v;
}
BX_HD scalar DistanceSquared(const Vec3& a, const Vec3& b) { return LengthSquared(a - b); }
BX_HD scalar Distance(const Vec3& a, const Vec3& b) { return Length(a - b); }
BX_HD Vec3 Min(const Vec3& a, const Vec3& b) { return Vec3(Min(a.x, b.x), Min(a.y, b.y), Min(a.z, b.z)); }
BX_HD Vec3 Max(const Vec3& a, const Vec3& b) { return Vec3(Max(a.x, b.x), Max(a.y, b.y), Max(a.z, b.z)); }
BX_HD Vec3 Mul(const Vec3& a, const Vec3& b) { return Vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
BX_HD Vec3 Perp(const Vec3& v) {  // vec3.h:243-259
    Vec3 u;
    if (Abs(v.x) >= kSqrtInv3) u.Set(v.y, -v.x, scalar(0));
    else u.Set(scalar(0), v.z, -v.y);
    u.Normalize();
    return u;
}

struct Mat22 {
    Vec2 x, y;
    Mat22() = default;
    BX_HD Mat22(const Vec2& a, const Vec2& b) : x(a), y(b) {}
    BX_HD void SetZero() { x.SetZero(); y.SetZero(); }
};
BX_HD Vec2 operator*(const Mat22& A, const Vec2& v) { return v.x * A.x + v.y * A.y; }
BX_HD Mat22 operator*(scalar s, const Mat22& A) { return Mat22(s * A.x, s * A.y); }
BX_HD Mat22 Inverse(const Mat22& A) {  // math.cpp:76-84
    scalar det = Det(A.x, A.y);
    if (det != scalar(0)) det = scalar(1) / det;
    return det * Mat22(Vec2(A.y.y, -A.x.y), Vec2(-A.y.x, A.x.x));
}
BX_HD Vec2 Solve22(const Mat22& A, const Vec2& b) {  // math.cpp:62-74
    scalar a11 = A.x.x, a12 = A.y.x, a21 = A.x.y, a22 = A.y.y;
    scalar det = a11 * a22 - a12 * a21;
    if (det != scalar(0)) det = scalar(1) / det;
    Vec2 r;
    r.x = det * (a22 * b.x - a12 * b.y);
    r.y = det * (a11 * b.y - a21 * b.x);
    return r;
}

struct Mat33 {
    Vec3 x, y, z;
    Mat33() = default;
    BX_HD Mat33(const Vec3& a, const Vec3& b, const Vec3& c) : x(a), y(b), z(c) {}
    BX_HD void SetZero() { x.SetZero(); y.SetZero(); z.SetZero(); }
    BX_HD void SetIdentity() { x.Set(1, 0, 0); y.Set(0, 1, 0); z.Set(0, 0, 1); }
    BX_HD void operator+=(const Mat33& B) { x += B.x; y += B.y; z += B.z; }
    BX_HD Vec3 Solve(const Vec3& b) const {  // math.cpp:88-101
        scalar det = Det(x, y, z);
        if (det != scalar(0)) det = scalar(1) / det;
        Vec3 r;
        r.x = det * Det(b, y, z);
        r.y = det * Det(x, b, z);
        r.z = det * Det(x, y, b);
        return r;
    }
};
BX_HD Mat33 operator+(const Mat33& A, const Mat33& B) { return Mat33(A.x + B.x, A.y + B.y, A.z + B.z); }
BX_HD Mat33 operator-(const Mat33& A, const Mat33& B) { return Mat33(A.x - B.x, A.y - B.y, A.z - B.z); }
BX_HD Mat33 operator*(scalar s, const Mat33& A) { return Mat33(s * A.x, s * A.y, s * A.z); }
BX_HD Mat33 operator-(const Mat33& A) { return scalar(-1) * A; }  // mat33.h:123-126
BX_HD Vec3 operator*(const Mat33& A, const Vec3& v) { return v.x * A.x + v.y * A.y + v.z * A.z; }
BX_HD Mat33 operator*(const Mat33& A, const Mat33& B) { return Mat33(A * B.x, A * B.y, A * B.z); }
BX_HD Vec3 MulT(const Mat33& A, const Vec3& v) { return Vec3(Dot(A.x, v), Dot(A.y, v), Dot(A.z, v)); }
BX_HD Mat33 Transpose(const Mat33& A) {
    return Mat33(Vec3(A.x.x, A.y.x, A.z.x), Vec3(A.x.y, A.y.y, A.z.y), Vec3(A.x.z, A.y.z, A.z.z));
}
BX_HD Mat33 Diagonal(scalar a, scalar b, scalar c) { return Mat33(Vec3(a, 0, 0), Vec3(0, b, 0), Vec3(0, 0, c)); }
BX_HD Mat33 Skew(const Vec3& v) { return Mat33(Vec3(0, v.z, -v.y), Vec3(-v.z, 0, v.x), Vec3(v.y, -v.x, 0)); }
BX_HD Mat33 Inverse(const Mat33& A) {  // math.cpp:103-113, cofactor method
    scalar det = Det(A.x, A.y, A.z);
    if (det != scalar(0)) det = scalar(1) / det;
    return det * Transpose(Mat33(Cross(A.y, A.z), Cross(A.z, A.x), Cross(A.x, A.y)));
}
BX_HD Mat33 Steiner(const Vec3& v) {  // mat33.h:337-356
    scalar xx = v.x * v.x, yy = v.y * v.y, zz = v.z * v.z;
    Mat33 S;
    S.x.x = yy + zz; S.x.y = -v.x * v.y; S.x.z = -v.x * v.z;
    S.y.x = S.x.y;   S.y.y = xx + zz;    S.y.z = -v.y * v.z;
    S.z.x = S.x.z;   S.z.y = S.y.z;      S.z.z = xx + yy;
    return S;
}

struct Quat {
    Vec3 v;
    scalar s;
    Quat() = default;
    BX_HD Quat(const Vec3& a, scalar b) : v(a), s(b) {}
    BX_HD void SetIdentity() { v.SetZero(); s = 1; }
    BX_HD void operator+=(const Quat& q) { v += q.v; s += q.s; }
    BX_HD void operator-=(const Quat& q) { v -= q.v; s -= q.s; }
    BX_HD scalar Normalize() {  // quat.h:63-72
        scalar len = Sqrt(Dot(v, v) + s * s);
        if (len > kEpsilon) { v /= len; s /= len; }
        return len;
    }
    BX_HD Mat33 GetRotationMatrix() const {
        scalar x = v.x, y = v.y, z = v.z, w = s;
        scalar x2 = x + x, y2 = y + y, z2 = z + z;
        scalar xx = x * x2, xy = x * y2, xz = x * z2;
        scalar yy = y * y2, yz = y * z2, zz = z * z2;
        scalar wx = w * x2, wy = w * y2, wz = w * z2;
        return Mat33(Vec3(scalar(1) - (yy + zz), xy + wz, xz - wy),
                     Vec3(xy - wz, scalar(1) - (xx + zz), yz + wx),
                     Vec3(xz + wy, yz - wx, scalar(1) - (xx + yy)));
    }
};
BX_HD Quat operator+(const Quat& a, const Quat& b) { return Quat(a.v + b.v, a.s + b.s); }
BX_HD Quat operator-(const Quat& a, const Quat& b) { return Quat(a.v - b.v, a.s - b.s); }
BX_HD Quat operator*(scalar s, const Quat& q) { return Quat(s * q.v, s * q.s); }
BX_HD Quat operator-(const Quat& q) { return Quat(-q.v, -q.s); }
BX_HD Quat Mul(const Quat& a, const Quat& b) {
    return Quat(Cross(a.v, b.v) + a.s * b.v + b.s * a.v, a.s * b.s - Dot(a.v, b.v));
}
BX_HD Quat operator*(const Quat& a, const Quat& b) { return Mul(a, b); }
BX_HD scalar Dot(const Quat& a, const Quat& b) { return Dot(a.v, b.v) + a.s * b.s; }
BX_HD Quat Conjugate(const Quat& q) { return Quat(-q.v, q.s); }
BX_HD Quat MulC(const Quat& a, const Quat& b) { return Mul(Conjugate(a), b); }
BX_HD Vec3 Mul(const Quat& q, const Vec3& v) {  // quat.h:270-274
    Vec3 t = scalar(2) * Cross(q.v, v);
    return v + q.s * t + Cross(q.v, t);
}
BX_HD Vec3 MulC(const Quat& q, const Vec3& v) { return Mul(Conjugate(q), v); }
BX_HD Mat33 RotateToFrame(const Mat33& I, const Quat& q) {
    Mat33 R = q.GetRotationMatrix();
    return R * I * Transpose(R);
}
BX_HD Quat Derivative(const Quat& q, const Vec3& omega) { return scalar(0.5) * Quat(omega, scalar(0)) * q; }
BX_HD Quat Integrate(const Quat& q, const Vec3& omega, scalar dt) {
    Quat qdot = Derivative(q, omega);
    Quat q2 = q + dt * qdot;
    q2.Normalize();
    return q2;
}

struct Transform {
    Vec3 translation;
    Quat rotation;
    Transform() = default;
    BX_HD Transform(const Vec3& t, const Quat& q) : translation(t), rotation(q) {}
};
BX_HD Vec3 Mul(const Transform& T, const Vec3& v) { return Mul(T.rotation, v) + T.translation; }
BX_HD Transform MulT(const Transform& A, const Transform& B) {
    return Transform(MulC(A.rotation, B.translation - A.translation), MulC(A.rotation, B.rotation));
}
BX_HD Vec3 MulT(const Transform& A, const Vec3& v) { return MulC(A.rotation, v - A.translation); }

// b3Sweep::GetTransform (sweep.h:43-63) from its stored pieces.
BX_HD Transform SweepTransform(const Vec3& c0, const Quat& q0, const Vec3& c, const Quat& qn, const Vec3& lc,
                               scalar t) {
    Vec3 cc = (scalar(1) - t) * c0 + t * c;
    Quat q1 = q0;
    Quat q2 = qn;
    if (Dot(q1, q2) < scalar(0)) q1 = -q1;
    Quat q = (scalar(1) - t) * q1 + t * q2;
    q.Normalize();
    return Transform(cc - Mul(q, lc), q);
}

struct Plane {
    Vec3 normal;
    scalar offset;
    Plane() = default;
    BX_HD Plane(const Vec3& n, scalar o) : normal(n), offset(o) {}
};
BX_HD Plane Mul(const Transform& T, const Plane& p) {
    Vec3 n = Mul(T.rotation, p.normal);
    return Plane(n, p.offset + Dot(n, T.translation));
}
BX_HD scalar Distance(const Vec3& P, const Plane& p) { return Dot(p.normal, P) - p.offset; }
BX_HD Vec3 ClosestPointOnPlane(const Vec3& P, const Plane& p) { return P - Distance(P, p) * p.normal; }

struct AABB {
    Vec3 lowerBound, upperBound;
    BX_HD void Extend(scalar s) { Vec3 r(s, s, s); lowerBound -= r; upperBound += r; }
    BX_HD Vec3 GetCenter() const { return scalar(0.5) * (lowerBound + upperBound); }
    BX_HD bool Contains(const AABB& a) const {
        return lowerBound.x <= a.lowerBound.x && lowerBound.y <= a.lowerBound.y && lowerBound.z <= a.lowerBound.z &&
               a.upperBound.x <= upperBound.x && a.upperBound.y <= upperBound.y && a.upperBound.z <= upperBound.z;
    }
};
BX_HD AABB Combine(const AABB& a, const AABB& b) {
    AABB r;
    r.lowerBound = Min(a.lowerBound, b.lowerBound);
    r.upperBound = Max(a.upperBound, b.upperBound);
    return r;
}
BX_HD bool TestOverlap(const AABB& a, const AABB& b) {  // aabb.h:347-367, inclusive
    if (a.upperBound.x < b.lowerBound.x || a.lowerBound.x > b.upperBound.x) return false;
    if (a.upperBound.y < b.lowerBound.y || a.lowerBound.y > b.upperBound.y) return false;
    if (a.upperBound.z < b.lowerBound.z || a.lowerBound.z > b.upperBound.z) return false;
    return true;
}

BX_HD void Barycentric2(scalar out[3], const Vec3& A, const Vec3& B, const Vec3& Q) {  // geometry.h:27-39
    Vec3 AB = B - A, QA = A - Q, QB = B - Q;
    scalar divisor = Dot(AB, AB);
    out[0] = Dot(QB, AB);
    out[1] = -Dot(QA, AB);
    out[2] = divisor;
}

}  // namespace bx
