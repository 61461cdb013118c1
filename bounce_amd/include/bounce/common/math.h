// bounce/common/math.h -- the value types of the drop-in API (b3Vec3, b3Quat, b3Mat33, b3Transform).
//
// Names, members and argument meanings follow the reference's public math headers
// (include/bounce/common/math/{math,vec3,quat,mat33,transform}.h), so user code written against the
// reference compiles unchanged. The arithmetic of the step itself lives in the device library; these
// types only carry data across the API and offer the conveniences a caller uses.
#pragma once
#include <cmath>
#include <cstdint>

typedef float scalar;
typedef uint8_t uint8;
typedef uint16_t uint16;
typedef uint32_t uint32;
typedef int32_t int32;
typedef uint64_t uint64;

#define B3_PI scalar(3.14159265358979323846)
#define B3_EPSILON scalar(1.19209290e-07)
#define B3_MAX_SCALAR scalar(3.40282347e+38)
#define B3_LINEAR_SLOP scalar(0.005)
#define B3_ANGULAR_SLOP (scalar(2.0) / scalar(180) * B3_PI)

template <class T> inline T b3Min(T a, T b) { return a < b ? a : b; }
template <class T> inline T b3Max(T a, T b) { return a > b ? a : b; }
template <class T> inline T b3Clamp(T a, T low, T high) { return b3Max(low, b3Min(a, high)); }
template <class T> inline T b3Abs(T a) { return a >= T(0) ? a : -a; }
inline scalar b3Sqrt(scalar x) { return std::sqrt(x); }

struct b3Vec2 {
    scalar x, y;
    b3Vec2() {}
    b3Vec2(scalar _x, scalar _y) : x(_x), y(_y) {}
    void SetZero() { x = y = scalar(0); }
    void Set(scalar _x, scalar _y) { x = _x; y = _y; }
};

struct b3Vec3 {
    scalar x, y, z;
    b3Vec3() {}
    b3Vec3(scalar _x, scalar _y, scalar _z) : x(_x), y(_y), z(_z) {}
    scalar& operator[](uint32 i) { return (&x)[i]; }
    scalar operator[](uint32 i) const { return (&x)[i]; }
    void SetZero() { x = y = z = scalar(0); }
    void Set(scalar _x, scalar _y, scalar _z) { x = _x; y = _y; z = _z; }
    void operator+=(const b3Vec3& b) { x += b.x; y += b.y; z += b.z; }
    void operator-=(const b3Vec3& b) { x -= b.x; y -= b.y; z -= b.z; }
    void operator*=(scalar s) { x *= s; y *= s; z *= s; }
    void operator/=(scalar a) { scalar s = scalar(1) / a; x *= s; y *= s; z *= s; }
    b3Vec3 operator-() const { return b3Vec3(-x, -y, -z); }
    scalar Normalize() {
        scalar len = std::sqrt(x * x + y * y + z * z);
        if (len > B3_EPSILON) { scalar s = scalar(1) / len; x *= s; y *= s; z *= s; }
        return len;
    }
};

inline b3Vec3 operator+(const b3Vec3& a, const b3Vec3& b) { return b3Vec3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline b3Vec3 operator-(const b3Vec3& a, const b3Vec3& b) { return b3Vec3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline b3Vec3 operator*(scalar s, const b3Vec3& v) { return b3Vec3(s * v.x, s * v.y, s * v.z); }
inline b3Vec3 operator*(const b3Vec3& v, scalar s) { return s * v; }
inline b3Vec3 operator/(const b3Vec3& v, scalar s) { return b3Vec3(v.x / s, v.y / s, v.z / s); }
inline scalar b3Dot(const b3Vec3& a, const b3Vec3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline b3Vec3 b3Cross(const b3Vec3& a, const b3Vec3& b) {
    return b3Vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline scalar b3Length(const b3Vec3& v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
inline scalar b3LengthSquared(const b3Vec3& v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
inline scalar b3Distance(const b3Vec3& a, const b3Vec3& b) { return b3Length(a - b); }
inline b3Vec3 b3Normalize(const b3Vec3& v) {
    scalar len = b3Length(v);
    if (len > B3_EPSILON) return (scalar(1) / len) * v;
    return v;
}

struct b3Mat33 {
    b3Vec3 x, y, z;
    b3Mat33() {}
    b3Mat33(const b3Vec3& a, const b3Vec3& b, const b3Vec3& c) : x(a), y(b), z(c) {}
    b3Vec3& operator[](uint32 i) { return (&x)[i]; }
    const b3Vec3& operator[](uint32 i) const { return (&x)[i]; }
    void SetZero() { x.SetZero(); y.SetZero(); z.SetZero(); }
    void SetIdentity() { x.Set(1, 0, 0); y.Set(0, 1, 0); z.Set(0, 0, 1); }
};
inline b3Vec3 operator*(const b3Mat33& A, const b3Vec3& v) { return v.x * A.x + v.y * A.y + v.z * A.z; }
inline b3Vec3 b3Mul(const b3Mat33& A, const b3Vec3& v) { return A * v; }
inline b3Mat33 operator*(const b3Mat33& A, const b3Mat33& B) { return b3Mat33(A * B.x, A * B.y, A * B.z); }
inline b3Mat33 b3Transpose(const b3Mat33& A) {
    return b3Mat33(b3Vec3(A.x.x, A.y.x, A.z.x), b3Vec3(A.x.y, A.y.y, A.z.y), b3Vec3(A.x.z, A.y.z, A.z.z));
}

// A unit vector orthogonal to v (vec3.h b3Perp).
inline b3Vec3 b3Perp(const b3Vec3& v) {
    static const scalar kSqrtInv3 = std::sqrt(scalar(1) / scalar(3));
    b3Vec3 u;
    if (b3Abs(v.x) >= kSqrtInv3) u.Set(v.y, -v.x, scalar(0));
    else u.Set(scalar(0), v.z, -v.y);
    u.Normalize();
    return u;
}

// Quaternion (x, y, z) = v, w = s.
struct b3Quat {
    b3Vec3 v;
    scalar s;
    b3Quat() {}
    b3Quat(const b3Vec3& _v, scalar _s) : v(_v), s(_s) {}
    b3Quat(scalar x, scalar y, scalar z, scalar w) : v(x, y, z), s(w) {}
    void SetIdentity() { v.SetZero(); s = scalar(1); }
    void Set(const b3Vec3& _v, scalar _s) { v = _v; s = _s; }
    void SetAxisAngle(const b3Vec3& axis, scalar angle) {  // axis must be normalized
        scalar theta = scalar(0.5) * angle;
        v = std::sin(theta) * axis;
        s = std::cos(theta);
    }
    scalar Normalize() {
        scalar len = std::sqrt(b3Dot(v, v) + s * s);
        if (len > B3_EPSILON) { v *= scalar(1) / len; s = s / len; }
        return len;
    }
    // Axis and angle of the rotation (quat.h GetAxisAngle).
    void GetAxisAngle(b3Vec3* axis, scalar* angle) const {
        scalar sine = b3Length(v);
        axis->SetZero();
        if (sine > B3_EPSILON) *axis = (scalar(1) / sine) * v;
        scalar cosine = b3Clamp(s, scalar(-1), scalar(1));
        *angle = scalar(2) * std::acos(cosine);
    }
    b3Mat33 GetRotationMatrix() const {
        scalar x = v.x, y = v.y, z = v.z, w = s;
        scalar x2 = x + x, y2 = y + y, z2 = z + z;
        scalar xx = x * x2, xy = x * y2, xz = x * z2;
        scalar yy = y * y2, yz = y * z2, zz = z * z2;
        scalar wx = w * x2, wy = w * y2, wz = w * z2;
        return b3Mat33(b3Vec3(scalar(1) - (yy + zz), xy + wz, xz - wy),
                       b3Vec3(xy - wz, scalar(1) - (xx + zz), yz + wx),
                       b3Vec3(xz + wy, yz - wx, scalar(1) - (xx + yy)));
    }
};

inline b3Quat b3QuatIdentity() { return b3Quat(b3Vec3(0, 0, 0), scalar(1)); }
inline b3Quat b3Conjugate(const b3Quat& q) { return b3Quat(-q.v, q.s); }
inline b3Quat b3Mul(const b3Quat& a, const b3Quat& b) {
    return b3Quat(b3Cross(a.v, b.v) + a.s * b.v + b.s * a.v, a.s * b.s - b3Dot(a.v, b.v));
}
inline b3Quat operator*(const b3Quat& a, const b3Quat& b) { return b3Mul(a, b); }
inline b3Vec3 b3Mul(const b3Quat& q, const b3Vec3& v) {  // v + 2 s (q x v) + 2 q x (q x v)
    b3Vec3 t = scalar(2) * b3Cross(q.v, v);
    return v + q.s * t + b3Cross(q.v, t);
}
inline b3Vec3 b3MulC(const b3Quat& q, const b3Vec3& v) { return b3Mul(b3Conjugate(q), v); }
// R I R^T with R the rotation matrix of q (quat.h b3RotateToFrame).
inline b3Mat33 b3RotateToFrame(const b3Mat33& I, const b3Quat& q) {
    b3Mat33 R = q.GetRotationMatrix();
    return R * I * b3Transpose(R);
}

struct b3Transform {
    b3Vec3 translation;
    b3Quat rotation;
    b3Transform() {}
    b3Transform(const b3Vec3& t, const b3Quat& q) : translation(t), rotation(q) {}
    void SetIdentity() { translation.SetZero(); rotation.SetIdentity(); }
};
inline b3Vec3 b3Mul(const b3Transform& T, const b3Vec3& v) { return b3Mul(T.rotation, v) + T.translation; }
inline b3Vec3 b3MulT(const b3Transform& T, const b3Vec3& v) { return b3MulC(T.rotation, v - T.translation); }
inline b3Vec3 operator*(const b3Transform& T, const b3Vec3& v) { return b3Mul(T, v); }
