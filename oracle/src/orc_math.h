// orc_math.h -- oracle (test infrastructure): CPU restatement of the reference's math core.
//
// Follows include/bounce/common/math/{math,vec2,vec3,mat22,mat33,quat,transform,sweep}.h and
// src/bounce/common/math/math.cpp. Every function evaluates the same IEEE binary32 expression
// tree as the reference (left-to-right sums, the reference's choice of reciprocal-multiply vs
// division), because the parity bar for this path is bit-exactness.
// Parity of each function here is pinned by the golden trajectories in tests/golden (any change
// of rounding shows up as a bit difference in body states within a few steps) and by the unit
// checks in tests/test_oracle_math.py (closed forms: rotation by quaternions, inverses).
#pragma once
#include <cmath>
#include <cfloat>
#include <cstdint>

namespace orc {

typedef float scalar;
typedef uint32_t u32;

static const scalar kEpsilon = FLT_EPSILON;    // common.h B3_EPSILON
static const scalar kMaxScalar = FLT_MAX;      // common.h B3_MAX_SCALAR
static const scalar kPi = scalar(3.14159265358979323846);
// common.h:73-123 tunables (same literals, evaluated in binary32 like the reference macros)
static const scalar kAabbExtension = scalar(0.2);
static const scalar kAabbMultiplier = scalar(2);
static const scalar kLinearSlop = scalar(0.005);
static const scalar kAngularSlop = scalar(2.0) / scalar(180) * kPi;
static const scalar kHullRadius = scalar(0.0) * kLinearSlop;
static const scalar kMaxTranslation = scalar(2.0);
static const scalar kMaxTranslationSq = kMaxTranslation * kMaxTranslation;
static const scalar kMaxRotation = scalar(0.5) * kPi;
static const scalar kMaxRotationSq = kMaxRotation * kMaxRotation;
static const scalar kMaxLinearCorrection = scalar(0.2);
static const scalar kMaxAngularCorrection = scalar(8.0) / scalar(180) * kPi;
static const scalar kBaumgarte = scalar(0.1);
static const scalar kVelocityThreshold = scalar(1.0);
static const scalar kTimeToSleep = scalar(0.2);
static const u32 kMaxManifoldPoints = 4;
static const u32 kMaxManifolds = 3;
static const u32 kNullU32 = 0xFFFFFFFFu;

// math.h:53-68 -- ternaries, not fmin/fmax (NaN and signed-zero behaviour matter).
template <class T> inline T Min(T a, T b) { return a < b ? a : b; }
template <class T> inline T Max(T a, T b) { return a > b ? a : b; }
template <class T> inline T Clamp(T a, T lo, T hi) { return Max(lo, Min(a, hi)); }
inline scalar Abs(scalar x) { return std::fabs(x); }
inline scalar Sqrt(scalar x) { return std::sqrt(x); }
template <class T> inline void Swap(T& a, T& b) { T t = a; a = b; b = t; }
inline scalar Sign(scalar x) { return scalar(scalar(0) < x) - scalar(x < scalar(0)); }

// ---------------------------------------------------------------- Vec2 (vec2.h)
struct Vec2 {
    scalar x, y;
    Vec2() {}
    Vec2(scalar a, scalar b) : x(a), y(b) {}
    void SetZero() { x = y = 0; }
    void Set(scalar a, scalar b) { x = a; y = b; }
    void operator+=(const Vec2& b) { x += b.x; y += b.y; }
    void operator-=(const Vec2& b) { x -= b.x; y -= b.y; }
    void operator*=(scalar s) { x *= s; y *= s; }
    scalar Normalize() {  // vec2.h: divides (vec3 multiplies by the reciprocal)
        scalar len = Sqrt(x * x + y * y);
        if (len > kEpsilon) { x /= len; y /= len; }
        return len;
    }
};
inline Vec2 operator+(const Vec2& a, const Vec2& b) { return Vec2(a.x + b.x, a.y + b.y); }
inline Vec2 operator-(const Vec2& a, const Vec2& b) { return Vec2(a.x - b.x, a.y - b.y); }
inline Vec2 operator-(const Vec2& a) { return Vec2(-a.x, -a.y); }
inline Vec2 operator*(scalar s, const Vec2& v) { return Vec2(s * v.x, s * v.y); }
inline scalar Dot(const Vec2& a, const Vec2& b) { return a.x * b.x + a.y * b.y; }
inline scalar Det(const Vec2& a, const Vec2& b) { return a.x * b.y - a.y * b.x; }

// ---------------------------------------------------------------- Vec3 (vec3.h)
struct Vec3 {
    scalar x, y, z;
    Vec3() {}
    Vec3(scalar a, scalar b, scalar c) : x(a), y(b), z(c) {}
    scalar operator[](u32 i) const { return (&x)[i]; }
    scalar& operator[](u32 i) { return (&x)[i]; }
    void SetZero() { x = y = z = 0; }
    void Set(scalar a, scalar b, scalar c) { x = a; y = b; z = c; }
    void operator+=(const Vec3& b) { x += b.x; y += b.y; z += b.z; }
    void operator-=(const Vec3& b) { x -= b.x; y -= b.y; z -= b.z; }
    void operator*=(scalar s) { x *= s; y *= s; z *= s; }
    void operator/=(scalar a) { scalar s = scalar(1) / a; x *= s; y *= s; z *= s; }
    scalar Normalize() {
        scalar len = Sqrt(x * x + y * y + z * z);
        if (len > kEpsilon) { scalar s = scalar(1) / len; x *= s; y *= s; z *= s; }
        return len;
    }
};
inline Vec3 operator-(const Vec3& v) { return Vec3(-v.x, -v.y, -v.z); }
inline Vec3 operator+(const Vec3& a, const Vec3& b) { return Vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline Vec3 operator-(const Vec3& a, const Vec3& b) { return Vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline Vec3 operator*(scalar s, const Vec3& v) { return Vec3(s * v.x, s * v.y, s * v.z); }
inline Vec3 operator*(const Vec3& v, scalar s) { return s * v; }
inline Vec3 operator/(const Vec3& v, scalar s) { return Vec3(v.x / s, v.y / s, v.z / s); }
inline scalar Dot(const Vec3& a, const Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline Vec3 Cross(const Vec3& a, const Vec3& b) {
    return Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline scalar Det(const Vec3& a, const Vec3& b, const Vec3& c) { return Dot(a, Cross(b, c)); }
inline scalar Length(const Vec3& v) { return Sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline scalar LengthSquared(const Vec3& v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline Vec3 Normalize(const Vec3& v) {
    scalar len = Length(v);
    if (len > kEpsilon) { scalar s = scalar(1) / len; return s * v; }
    return v;
}
inline scalar Distance(const Vec3& a, const Vec3& b) { return Length(a - b); }
inline scalar DistanceSquared(const Vec3& a, const Vec3& b) { return LengthSquared(a - b); }
inline Vec3 Min(const Vec3& a, const Vec3& b) { return Vec3(Min(a.x, b.x), Min(a.y, b.y), Min(a.z, b.z)); }
inline Vec3 Max(const Vec3& a, const Vec3& b) { return Vec3(Max(a.x, b.x), Max(a.y, b.y), Max(a.z, b.z)); }
inline Vec3 Abs(const Vec3& v) { return Vec3(Abs(v.x), Abs(v.y), Abs(v.z)); }
inline Vec3 Mul(const Vec3& a, const Vec3& b) { return Vec3(a.x * b.x, a.y * b.y, a.z * b.z); }
// vec3.h:243-259: one function-static sqrt(1/3) in binary32.
inline Vec3 Perp(const Vec3& v) {
    static const scalar sqrt_inv3 = Sqrt(scalar(1) / scalar(3));
    Vec3 u;
    if (Abs(v.x) >= sqrt_inv3) u.Set(v.y, -v.x, scalar(0));
    else u.Set(scalar(0), v.z, -v.y);
    u.Normalize();
    return u;
}

// ---------------------------------------------------------------- Mat22 (mat22.h, math.cpp)
struct Mat22 {
    Vec2 x, y;
    Mat22() {}
    Mat22(const Vec2& a, const Vec2& b) : x(a), y(b) {}
    void SetZero() { x.SetZero(); y.SetZero(); }
};
inline Vec2 operator*(const Mat22& A, const Vec2& v) { return v.x * A.x + v.y * A.y; }
inline Vec2 Solve(const Mat22& A, const Vec2& b) {  // b3Mat22::Solve, Cramer (math.cpp:70-86)
    scalar a11 = A.x.x, a12 = A.y.x;
    scalar a21 = A.x.y, a22 = A.y.y;
    scalar det = a11 * a22 - a12 * a21;
    if (det != scalar(0)) det = scalar(1) / det;
    Vec2 r;
    r.x = det * (a22 * b.x - a12 * b.y);
    r.y = det * (a11 * b.y - a21 * b.x);
    return r;
}
inline Mat22 operator*(scalar s, const Mat22& A) { return Mat22(s * A.x, s * A.y); }
inline Mat22 Adjugate(const Mat22& A) { return Mat22(Vec2(A.y.y, -A.x.y), Vec2(-A.y.x, A.x.x)); }
inline Mat22 Inverse(const Mat22& A) {
    scalar det = Det(A.x, A.y);
    if (det != scalar(0)) det = scalar(1) / det;
    return det * Adjugate(A);
}

// ---------------------------------------------------------------- Mat33 (mat33.h, math.cpp)
struct Mat33 {
    Vec3 x, y, z;
    Mat33() {}
    Mat33(const Vec3& a, const Vec3& b, const Vec3& c) : x(a), y(b), z(c) {}
    const Vec3& operator[](u32 i) const { return (&x)[i]; }
    Vec3& operator[](u32 i) { return (&x)[i]; }
    void SetZero() { x.SetZero(); y.SetZero(); z.SetZero(); }
    void SetIdentity() { x.Set(1, 0, 0); y.Set(0, 1, 0); z.Set(0, 0, 1); }
    void operator+=(const Mat33& B) { x += B.x; y += B.y; z += B.z; }
    Vec3 Solve(const Vec3& b) const {  // Cramer (math.cpp:88-101)
        scalar det = Det(x, y, z);
        if (det != scalar(0)) det = scalar(1) / det;
        Vec3 r;
        r.x = det * Det(b, y, z);
        r.y = det * Det(x, b, z);
        r.z = det * Det(x, y, b);
        return r;
    }
};
inline Mat33 operator+(const Mat33& A, const Mat33& B) { return Mat33(A.x + B.x, A.y + B.y, A.z + B.z); }
inline Mat33 operator-(const Mat33& A, const Mat33& B) { return Mat33(A.x - B.x, A.y - B.y, A.z - B.z); }
inline Mat33 operator*(scalar s, const Mat33& A) { return Mat33(s * A.x, s * A.y, s * A.z); }
inline Vec3 operator*(const Mat33& A, const Vec3& v) { return v.x * A.x + v.y * A.y + v.z * A.z; }
inline Mat33 operator*(const Mat33& A, const Mat33& B) { return Mat33(A * B.x, A * B.y, A * B.z); }
inline Vec3 MulT(const Mat33& A, const Vec3& v) { return Vec3(Dot(A.x, v), Dot(A.y, v), Dot(A.z, v)); }
inline Mat33 Transpose(const Mat33& A) {
    return Mat33(Vec3(A.x.x, A.y.x, A.z.x), Vec3(A.x.y, A.y.y, A.z.y), Vec3(A.x.z, A.y.z, A.z.z));
}
inline Mat33 operator-(const Mat33& A) { return scalar(-1) * A; }  // mat33.h:123-126
inline Mat33 Diagonal(scalar s) { return Mat33(Vec3(s, 0, 0), Vec3(0, s, 0), Vec3(0, 0, s)); }
inline Mat33 Diagonal(scalar a, scalar b, scalar c) { return Mat33(Vec3(a, 0, 0), Vec3(0, b, 0), Vec3(0, 0, c)); }
inline Mat33 Skew(const Vec3& v) {
    return Mat33(Vec3(0, v.z, -v.y), Vec3(-v.z, 0, v.x), Vec3(v.y, -v.x, 0));
}
inline Mat33 Inverse(const Mat33& A) {  // cofactor method (math.cpp:103-113)
    scalar det = Det(A.x, A.y, A.z);
    if (det != scalar(0)) det = scalar(1) / det;
    Mat33 C(Cross(A.y, A.z), Cross(A.z, A.x), Cross(A.x, A.y));
    return det * Transpose(C);
}
inline Mat33 Steiner(const Vec3& v) {  // mat33.h:337-356
    scalar xx = v.x * v.x, yy = v.y * v.y, zz = v.z * v.z;
    Mat33 S;
    S.x.x = yy + zz; S.x.y = -v.x * v.y; S.x.z = -v.x * v.z;
    S.y.x = S.x.y;   S.y.y = xx + zz;    S.y.z = -v.y * v.z;
    S.z.x = S.x.z;   S.z.y = S.y.z;      S.z.z = xx + yy;
    return S;
}

// ---------------------------------------------------------------- Quat (quat.h)
struct Quat {
    Vec3 v;
    scalar s;
    Quat() {}
    Quat(const Vec3& a, scalar b) : v(a), s(b) {}
    void SetIdentity() { v.SetZero(); s = 1; }
    void operator+=(const Quat& q) { v += q.v; s += q.s; }
    void operator-=(const Quat& q) { v -= q.v; s -= q.s; }
    scalar Normalize() {  // quat.h:63-72: vector part by reciprocal, scalar part by division
        scalar len = Sqrt(Dot(v, v) + s * s);
        if (len > kEpsilon) { v /= len; s /= len; }
        return len;
    }
    Mat33 GetRotationMatrix() const {
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
inline Quat QuatIdentity() { return Quat(Vec3(0, 0, 0), 1); }
inline Quat operator+(const Quat& a, const Quat& b) { return Quat(a.v + b.v, a.s + b.s); }
inline Quat operator-(const Quat& a, const Quat& b) { return Quat(a.v - b.v, a.s - b.s); }
inline Quat operator*(scalar s, const Quat& q) { return Quat(s * q.v, s * q.s); }
inline Quat operator/(const Quat& q, scalar s) { return Quat(q.v / s, q.s / s); }
inline Quat operator-(const Quat& q) { return Quat(-q.v, -q.s); }
inline Quat Mul(const Quat& a, const Quat& b) {
    return Quat(Cross(a.v, b.v) + a.s * b.v + b.s * a.v, a.s * b.s - Dot(a.v, b.v));
}
inline Quat operator*(const Quat& a, const Quat& b) { return Mul(a, b); }
inline scalar Dot(const Quat& a, const Quat& b) { return Dot(a.v, b.v) + a.s * b.s; }
inline Quat Conjugate(const Quat& q) { return Quat(-q.v, q.s); }
inline Quat MulC(const Quat& a, const Quat& b) { return Mul(Conjugate(a), b); }
inline Vec3 Mul(const Quat& q, const Vec3& v) {  // quat.h:270-274
    Vec3 t = scalar(2) * Cross(q.v, v);
    return v + q.s * t + Cross(q.v, t);
}
inline Vec3 MulC(const Quat& q, const Vec3& v) { return Mul(Conjugate(q), v); }
inline Mat33 RotateToFrame(const Mat33& I, const Quat& q) {
    Mat33 R = q.GetRotationMatrix();
    return R * I * Transpose(R);
}
inline Quat Derivative(const Quat& q, const Vec3& omega) {
    Quat w(omega, scalar(0));
    return scalar(0.5) * w * q;
}
inline Quat Integrate(const Quat& q, const Vec3& omega, scalar dt) {
    Quat qdot = Derivative(q, omega);
    Quat q2 = q + dt * qdot;
    q2.Normalize();
    return q2;
}

// ---------------------------------------------------------------- Transform (transform.h)
struct Transform {
    Vec3 translation;
    Quat rotation;
    Transform() {}
    Transform(const Vec3& t, const Quat& q) : translation(t), rotation(q) {}
    void SetIdentity() { translation.SetZero(); rotation.SetIdentity(); }
};
inline Vec3 Mul(const Transform& T, const Vec3& v) { return Mul(T.rotation, v) + T.translation; }
inline Vec3 operator*(const Transform& T, const Vec3& v) { return Mul(T, v); }
inline Transform MulT(const Transform& A, const Transform& B) {
    Transform C;
    C.rotation = MulC(A.rotation, B.rotation);
    C.translation = MulC(A.rotation, B.translation - A.translation);
    return C;
}
inline Vec3 MulT(const Transform& A, const Vec3& v) { return MulC(A.rotation, v - A.translation); }

// ---------------------------------------------------------------- Sweep (sweep.h:43-63)
struct Sweep {
    Vec3 localCenter;
    Quat orientation0;
    Vec3 worldCenter0;
    scalar t0;
    Quat orientation;
    Vec3 worldCenter;
    Transform GetTransform(scalar t) const {
        Vec3 c = (scalar(1) - t) * worldCenter0 + t * worldCenter;
        Quat q1 = orientation0;
        Quat q2 = orientation;
        if (Dot(q1, q2) < scalar(0)) q1 = -q1;
        Quat q = (scalar(1) - t) * q1 + t * q2;
        q.Normalize();
        Transform xf;
        xf.translation = c - Mul(q, localCenter);
        xf.rotation = q;
        return xf;
    }
};

// ---------------------------------------------------------------- Plane (plane.h)
struct Plane {
    Vec3 normal;
    scalar offset;
    Plane() {}
    Plane(const Vec3& n, scalar o) : normal(n), offset(o) {}
};
inline Plane Mul(const Transform& T, const Plane& p) {
    Vec3 n = Mul(T.rotation, p.normal);
    return Plane(n, p.offset + Dot(n, T.translation));
}
inline scalar Distance(const Vec3& P, const Plane& p) { return Dot(p.normal, P) - p.offset; }
inline Vec3 ClosestPointOnPlane(const Vec3& P, const Plane& p) {
    scalar f = Distance(P, p);
    return P - f * p.normal;
}

// ---------------------------------------------------------------- AABB (aabb.h)
struct AABB {
    Vec3 lowerBound, upperBound;
    void Extend(scalar s) { Vec3 r(s, s, s); lowerBound -= r; upperBound += r; }
    Vec3 GetCenter() const { return scalar(0.5) * (lowerBound + upperBound); }
    bool Contains(const AABB& a) const {
        bool r = true;
        r = r && lowerBound.x <= a.lowerBound.x;
        r = r && lowerBound.y <= a.lowerBound.y;
        r = r && lowerBound.z <= a.lowerBound.z;
        r = r && a.upperBound.x <= upperBound.x;
        r = r && a.upperBound.y <= upperBound.y;
        r = r && a.upperBound.z <= upperBound.z;
        return r;
    }
};
inline AABB Combine(const AABB& a, const AABB& b) {
    AABB r;
    r.lowerBound = Min(a.lowerBound, b.lowerBound);
    r.upperBound = Max(a.upperBound, b.upperBound);
    return r;
}
inline bool TestOverlap(const AABB& a, const AABB& b) {  // inclusive (aabb.h:347-367)
    if (a.upperBound.x < b.lowerBound.x || a.lowerBound.x > b.upperBound.x) return false;
    if (a.upperBound.y < b.lowerBound.y || a.lowerBound.y > b.upperBound.y) return false;
    if (a.upperBound.z < b.lowerBound.z || a.lowerBound.z > b.upperBound.z) return false;
    return true;
}

// geometry.h:27-39
inline void Barycentric2(scalar out[3], const Vec3& A, const Vec3& B, const Vec3& Q) {
    Vec3 AB = B - A, QA = A - Q, QB = B - Q;
    scalar divisor = Dot(AB, AB);
    out[0] = Dot(QB, AB);
    out[1] = -Dot(QA, AB);
    out[2] = divisor;
}
inline void Barycentric3(scalar out[4], const Vec3& A, const Vec3& B, const Vec3& C, const Vec3& Q) {
    Vec3 AB = B - A, AC = C - A, QA = A - Q, QB = B - Q, QC = C - Q;
    Vec3 QB_x_QC = Cross(QB, QC), QC_x_QA = Cross(QC, QA), QA_x_QB = Cross(QA, QB);
    Vec3 AB_x_AC = Cross(AB, AC);
    scalar divisor = Dot(AB_x_AC, AB_x_AC);
    out[0] = Dot(QB_x_QC, AB_x_AC);
    out[1] = Dot(QC_x_QA, AB_x_AC);
    out[2] = Dot(QA_x_QB, AB_x_AC);
    out[3] = divisor;
}

}  // namespace orc
