// bx_joint.h -- the revolute joint's solver rows for the device island sweeps.
//
// Same binary32 expression trees as revolute_joint.cpp:78-531 (InitializeVelocityConstraints,
// WarmStart, SolveVelocityConstraints, SolvePositionConstraints). The joint angle of the limit rows
// (revolute_joint.cpp:135, 434) is the reference's atan2 of the two binary32 cosines promoted to
// double, rounded to binary32: JointAngle evaluates it in double as well, so the binary32 angle is
// the reference's unless the two double results fall on opposite sides of a binary32 rounding
// boundary (their last-bit difference makes that a chance of about 2^-26 per evaluation).
#pragma once
#include "bx_types.h"

namespace bx {

// b3ComputeBasis (mat33.h:310-328)
BX_HD void ComputeBasis(const Vec3& a, Vec3& b, Vec3& c) {
    if (Abs(a.x) >= kSqrtInv3) b = Vec3(a.y, -a.x, scalar(0));
    else b = Vec3(scalar(0), a.z, -a.y);
    b.Normalize();
    c = Cross(a, b);
}

// b3Mat33Quat (quat.h:283-345); columns x, y, z are m[0], m[1], m[2].
BX_HD Quat Mat33Quat(const Mat33& m) {
    auto at = [&](u32 c, u32 r) -> scalar {
        const Vec3& col = c == 0 ? m.x : (c == 1 ? m.y : m.z);
        return r == 0 ? col.x : (r == 1 ? col.y : col.z);
    };
    scalar trace = at(0, 0) + at(1, 1) + at(2, 2);
    if (trace > scalar(0)) {
        Quat r;
        scalar s = Sqrt(trace + scalar(1));
        r.s = scalar(0.5) * s;
        scalar t = scalar(0.5) / s;
        r.v.x = t * (at(1, 2) - at(2, 1));
        r.v.y = t * (at(2, 0) - at(0, 2));
        r.v.z = t * (at(0, 1) - at(1, 0));
        return r;
    }
    u32 i = 0;
    if (at(1, 1) > at(0, 0)) i = 1;
    if (at(2, 2) > at(i, i)) i = 2;
    u32 j = i == 0 ? 1u : (i == 1 ? 2u : 0u);
    u32 k = j == 0 ? 1u : (j == 1 ? 2u : 0u);
    scalar s = Sqrt((at(i, i) - (at(j, j) + at(k, k))) + scalar(1));
    scalar q[4];
    q[i] = s * scalar(0.5);
    scalar t = s != scalar(0) ? scalar(0.5) / s : s;
    q[3] = t * (at(j, k) - at(k, j));
    q[j] = t * (at(i, j) + at(j, i));
    q[k] = t * (at(i, k) + at(k, i));
    Quat r;
    r.v.x = q[0];
    r.v.y = q[1];
    r.v.z = q[2];
    r.s = q[3];
    return r;
}

// b3RevoluteJoint's limit states (revolute_joint.h)
constexpr u32 kInactiveLimit = 0, kAtLowerLimit = 1, kAtUpperLimit = 2, kEqualLimits = 3;

// The relative angle of the joint frames: atan2(s, c) on the binary32 arguments promoted to double.
BX_HD scalar JointAngle(scalar s, scalar c) { return scalar(::atan2(double(s), double(c))); }

BX_HD bool IsAxisJoint(u32 type) { return type == kJointCone || type == kJointPrismatic || type == kJointWheel; }

// revolute_joint.cpp:78-231. `pos` is the island's position array, `indexA/B` the joint bodies'
// island-local indices in this island.
BX_HD void InitRevoluteCon(JointCon& jc, Joint& j, const Body& A, const Body& B, u32 indexA, u32 indexB,
                           const PosBody* pos, scalar dtRatio, bool warm) {
    jc.angularImpulse3.SetZero();
    jc.impulse1 = scalar(0);
    jc.linearImpulse2 = j.linearImpulse2;
    jc.springImpulse = j.springImpulse;
    jc.indexA = indexA;
    jc.indexB = indexB;
    jc.mA = A.invMass;
    jc.mB = B.invMass;
    jc.localCenterA = A.localCenter;
    jc.localCenterB = B.localCenter;
    jc.localInvIA = A.invI;
    jc.localInvIB = B.invI;
    jc.iA = pos[indexA].I;
    jc.iB = pos[indexB].I;
    jc.localAnchorA = j.localAnchorA;
    jc.localAnchorB = j.localAnchorB;
    jc.localRotationA = j.localRotationA;
    jc.localRotationB = j.localRotationB;
    jc.lowerAngle = j.lowerAngle;
    jc.upperAngle = j.upperAngle;
    jc.enableMotor = j.enableMotor;
    jc.enableLimit = j.enableLimit;
    jc.motorSpeed = j.motorSpeed;
    jc.maxMotorTorque = j.maxMotorTorque;
    Quat qA = pos[indexA].q;
    Quat qB = pos[indexB].q;
    scalar mA = jc.mA;
    Mat33 iA = jc.iA;
    scalar mB = jc.mB;
    Mat33 iB = jc.iB;
    Quat fA = qA * j.localRotationA;
    Quat fB = qB * j.localRotationB;
    Mat33 RfA = fA.GetRotationMatrix();
    Mat33 RfB = fB.GetRotationMatrix();
    Vec3 u1 = RfA.x, v1 = RfA.y, w1 = RfA.z;
    Vec3 u2 = RfB.x, v2 = RfB.y;
    scalar motorImpulse = j.motorImpulse;
    if (j.enableMotor || j.enableLimit) {
        jc.motorAxis = w1;
        jc.motorAxis.Normalize();
        scalar K = Dot((iA + iB) * jc.motorAxis, jc.motorAxis);
        jc.motorMass = K > scalar(0) ? scalar(1) / K : scalar(0);
    } else {
        jc.motorAxis.SetZero();
        jc.motorMass = scalar(0);
    }
    if (!j.enableMotor) motorImpulse = scalar(0);
    // the limit state carries over from the last step; entering a state clears the limit impulse
    // (revolute_joint.cpp:122-168)
    u32 limitState = j.limitState;
    scalar limitImpulse = j.limitImpulse;
    if (j.enableLimit) {
        scalar c = Dot(u2, u1);
        scalar s = Dot(u2, v1);
        scalar angle = JointAngle(s, c);
        if (Abs(j.upperAngle - j.lowerAngle) < scalar(2) * kAngularSlop) {
            if (limitState != kEqualLimits) { limitState = kEqualLimits; limitImpulse = scalar(0); }
        } else if (angle <= j.lowerAngle) {
            if (limitState != kAtLowerLimit) { limitState = kAtLowerLimit; limitImpulse = scalar(0); }
        } else if (angle >= j.upperAngle) {
            if (limitState != kAtUpperLimit) { limitState = kAtUpperLimit; limitImpulse = scalar(0); }
        } else {
            limitState = kInactiveLimit;
            limitImpulse = scalar(0);
        }
    } else {
        limitState = kInactiveLimit;
        limitImpulse = scalar(0);
    }
    j.limitState = limitState;
    jc.limitState = limitState;
    {
        jc.rA = Mul(qA, j.localAnchorA - jc.localCenterA);
        jc.rB = Mul(qB, j.localAnchorB - jc.localCenterB);
        Mat33 RA = Skew(jc.rA);
        Mat33 RAT = Transpose(RA);
        Mat33 RB = Skew(jc.rB);
        Mat33 RBT = Transpose(RB);
        Mat33 M = Diagonal(mA + mB, mA + mB, mA + mB);
        jc.linearMass = M + RA * iA * RAT + RB * iB * RBT;
    }
    {
        Vec3 a1 = Cross(u2, w1);
        Vec3 a2 = Cross(v2, w1);
        scalar k11 = Dot(a1, (iA + iB) * a1);
        scalar k12 = Dot(a1, iA * a2) + Dot(a1, iB * a2);
        scalar k22 = Dot(a2, (iA + iB) * a2);
        jc.a1 = a1;
        jc.a2 = a2;
        jc.angularMass.x.x = k11;
        jc.angularMass.x.y = k12;
        jc.angularMass.y.x = k12;
        jc.angularMass.y.y = k22;
    }
    Vec3 linearImpulse = j.linearImpulse;
    Vec2 angularImpulse = j.angularImpulse;
    if (warm) {
        motorImpulse *= dtRatio;
        limitImpulse *= dtRatio;
        linearImpulse *= dtRatio;
        angularImpulse *= dtRatio;
    } else {
        motorImpulse = scalar(0);
        limitImpulse = scalar(0);
        linearImpulse.SetZero();
        angularImpulse.SetZero();
    }
    jc.motorImpulse = motorImpulse;
    jc.limitImpulse = limitImpulse;
    jc.linearImpulse = linearImpulse;
    jc.angularImpulse = angularImpulse;
}

// revolute_joint.cpp:233-277
template <class V>
BX_HD void WarmStartRevolute(const JointCon& jc, V* vel) {
    Vec3 vA = vel[jc.indexA].v, wA = vel[jc.indexA].w;
    Vec3 vB = vel[jc.indexB].v, wB = vel[jc.indexB].w;
    if (jc.enableMotor && jc.limitState != kEqualLimits) {
        Vec3 P = jc.motorAxis * jc.motorImpulse;
        wA -= jc.iA * P;
        wB += jc.iB * P;
    }
    if (jc.enableLimit && jc.limitState != kInactiveLimit) {
        Vec3 P = jc.motorAxis * jc.limitImpulse;
        wA -= jc.iA * P;
        wB += jc.iB * P;
    }
    {
        vA -= jc.mA * jc.linearImpulse;
        wA -= jc.iA * Cross(jc.rA, jc.linearImpulse);
        vB += jc.mB * jc.linearImpulse;
        wB += jc.iB * Cross(jc.rB, jc.linearImpulse);
    }
    {
        Vec3 L = jc.angularImpulse.x * jc.a1 + jc.angularImpulse.y * jc.a2;
        wA -= jc.iA * L;
        wB += jc.iB * L;
    }
    vel[jc.indexA].v = vA;
    vel[jc.indexA].w = wA;
    vel[jc.indexB].v = vB;
    vel[jc.indexB].w = wB;
}

// revolute_joint.cpp:279-367
template <class V>
BX_HD void SolveRevoluteVelocity(JointCon& jc, V* vel, scalar dt) {
    Vec3 vA = vel[jc.indexA].v, wA = vel[jc.indexA].w;
    Vec3 vB = vel[jc.indexB].v, wB = vel[jc.indexB].w;
    Mat33 iA = jc.iA, iB = jc.iB;
    if (jc.enableMotor && jc.limitState != kEqualLimits) {
        scalar dw = Dot(wB - wA, jc.motorAxis);
        scalar Cdot = dw - jc.motorSpeed;
        scalar impulse = -jc.motorMass * Cdot;
        scalar oldImpulse = jc.motorImpulse;
        scalar maxImpulse = dt * jc.maxMotorTorque;
        jc.motorImpulse = Clamp(jc.motorImpulse + impulse, -maxImpulse, maxImpulse);
        impulse = jc.motorImpulse - oldImpulse;
        Vec3 P = jc.motorAxis * impulse;
        wA -= iA * P;
        wB += iB * P;
    }
    if (jc.enableLimit && jc.limitState != kInactiveLimit) {
        scalar Cdot = Dot(wB - wA, jc.motorAxis);
        scalar impulse = -jc.motorMass * Cdot;
        if (jc.limitState == kEqualLimits) {
            jc.limitImpulse += impulse;
        } else if (jc.limitState == kAtLowerLimit) {
            scalar oldImpulse = jc.limitImpulse;
            jc.limitImpulse = Max(jc.limitImpulse + impulse, scalar(0));
            impulse = jc.limitImpulse - oldImpulse;
        } else if (jc.limitState == kAtUpperLimit) {
            scalar oldImpulse = jc.limitImpulse;
            jc.limitImpulse = Min(jc.limitImpulse + impulse, scalar(0));
            impulse = jc.limitImpulse - oldImpulse;
        }
        Vec3 P = impulse * jc.motorAxis;
        wA -= iA * P;
        wB += iB * P;
    }
    {
        Vec3 Cdot = vB + Cross(wB, jc.rB) - vA - Cross(wA, jc.rA);
        Vec3 impulse = jc.linearMass.Solve(-Cdot);
        jc.linearImpulse += impulse;
        vA -= jc.mA * impulse;
        wA -= jc.iA * Cross(jc.rA, impulse);
        vB += jc.mB * impulse;
        wB += jc.iB * Cross(jc.rB, impulse);
    }
    {
        Vec3 dw = wB - wA;
        Vec2 Cdot;
        Cdot.x = Dot(dw, jc.a1);
        Cdot.y = Dot(dw, jc.a2);
        Vec2 impulse = Solve22(jc.angularMass, -Cdot);
        jc.angularImpulse += impulse;
        Vec3 L = impulse.x * jc.a1 + impulse.y * jc.a2;
        wA -= jc.iA * L;
        wB += jc.iB * L;
    }
    vel[jc.indexA].v = vA;
    vel[jc.indexA].w = wA;
    vel[jc.indexB].v = vB;
    vel[jc.indexB].w = wB;
}

// revolute_joint.cpp:369-531. `P` is an array of PosBody.
template <class P>
BX_HD bool SolveRevolutePosition(const JointCon& jc, P* pos) {
    Vec3 xA = pos[jc.indexA].x;
    Quat qA = pos[jc.indexA].q;
    Vec3 xB = pos[jc.indexB].x;
    Quat qB = pos[jc.indexB].q;
    Mat33 iA = pos[jc.indexA].I;
    Mat33 iB = pos[jc.indexB].I;
    scalar mA = jc.mA, mB = jc.mB;
    scalar linearError = scalar(0);
    {
        Vec3 rA = Mul(qA, jc.localAnchorA - jc.localCenterA);
        Vec3 rB = Mul(qB, jc.localAnchorB - jc.localCenterB);
        Vec3 C = xB + rB - xA - rA;
        linearError += Length(C);
        Mat33 M = Diagonal(mA + mB, mA + mB, mA + mB);
        Mat33 RA = Skew(rA);
        Mat33 RAT = Transpose(RA);
        Mat33 RB = Skew(rB);
        Mat33 RBT = Transpose(RB);
        Mat33 mass = M + RA * iA * RAT + RB * iB * RBT;
        Vec3 impulse = mass.Solve(-C);
        xA -= mA * impulse;
        qA -= Derivative(qA, iA * Cross(rA, impulse));
        qA.Normalize();
        iA = RotateToFrame(jc.localInvIA, qA);
        xB += mB * impulse;
        qB += Derivative(qB, iB * Cross(rB, impulse));
        qB.Normalize();
        iB = RotateToFrame(jc.localInvIB, qB);
    }
    scalar limitError = scalar(0);
    if (jc.enableLimit) {
        Quat fA = qA * jc.localRotationA;
        Quat fB = qB * jc.localRotationB;
        Mat33 RfA = fA.GetRotationMatrix();
        Mat33 RfB = fB.GetRotationMatrix();
        Vec3 u1 = RfA.x, v1 = RfA.y, w1 = RfA.z;
        Vec3 u2 = RfB.x;
        Vec3 limitAxis = w1;
        scalar K = Dot((iA + iB) * limitAxis, limitAxis);
        scalar limitMass = K > scalar(0) ? scalar(1) / K : scalar(0);
        scalar c = Dot(u2, u1);
        scalar s = Dot(u2, v1);
        scalar angle = JointAngle(s, c);
        scalar C = scalar(0);
        if (Abs(jc.upperAngle - jc.lowerAngle) < scalar(2) * kAngularSlop)
            C = Clamp(angle - jc.lowerAngle, -kMaxAngularCorrection, kMaxAngularCorrection);
        else if (angle <= jc.lowerAngle)
            C = Clamp(angle - jc.lowerAngle + kAngularSlop, -kMaxAngularCorrection, scalar(0));
        else if (angle >= jc.upperAngle)
            C = Clamp(angle - jc.upperAngle - kAngularSlop, scalar(0), kMaxAngularCorrection);
        limitError = Abs(C);
        scalar limitImpulse = -C * limitMass;
        Vec3 Pl = limitImpulse * limitAxis;
        qA -= Derivative(qA, iA * Pl);
        qA.Normalize();
        iA = RotateToFrame(jc.localInvIA, qA);
        qB += Derivative(qB, iB * Pl);
        qB.Normalize();
        iB = RotateToFrame(jc.localInvIB, qB);
    }
    scalar angularError = scalar(0);
    const scalar kCosSlop = scalar(0.002);
    {
        Quat fA = qA * jc.localRotationA;
        Quat fB = qB * jc.localRotationB;
        Mat33 RfA = fA.GetRotationMatrix();
        Mat33 RfB = fB.GetRotationMatrix();
        Vec3 w1 = RfA.z;
        Vec3 u2 = RfB.x, v2 = RfB.y;
        Vec3 a1 = Cross(u2, w1);
        Vec3 a2 = Cross(v2, w1);
        scalar k11 = Dot(a1, (iA + iB) * a1);
        scalar k12 = Dot(a1, iA * a2) + Dot(a1, iB * a2);
        scalar k22 = Dot(a2, (iA + iB) * a2);
        Mat22 K;
        K.x.x = k11;
        K.x.y = k12;
        K.y.x = k12;
        K.y.y = k22;
        Vec2 C;
        C.x = Dot(u2, w1);
        C.y = Dot(v2, w1);
        Vec2 impulse = Solve22(K, -C);
        Vec3 Pa = impulse.x * a1 + impulse.y * a2;
        qA -= Derivative(qA, iA * Pa);
        qA.Normalize();
        iA = RotateToFrame(jc.localInvIA, qA);
        qB += Derivative(qB, iB * Pa);
        qB.Normalize();
        iB = RotateToFrame(jc.localInvIB, qB);
        angularError = Abs(C.x) + Abs(C.y);
    }
    pos[jc.indexA].x = xA;
    pos[jc.indexA].q = qA;
    pos[jc.indexB].x = xB;
    pos[jc.indexB].q = qB;
    pos[jc.indexA].I = iA;
    pos[jc.indexB].I = iB;
    return linearError <= kLinearSlop && limitError <= kAngularSlop && angularError <= kCosSlop;
}

// ================================================================== the other joint types
// b3ComputeSoftConstraintCoefficients (weld_joint.cpp:64-81, mouse_joint.cpp:35-52): gamma and bias of
// a soft constraint of stiffness and damping from `frequencyHz` and `dampingRatio` on mass m.
BX_HD void SoftCoefficients(scalar& gamma, scalar& bias, scalar frequencyHz, scalar dampingRatio, scalar m, scalar C,
                            scalar h) {
    scalar omega = scalar(2) * kPi * frequencyHz;
    scalar k = omega * omega * m;
    scalar d = scalar(2) * dampingRatio * omega * m;
    gamma = h * (d + h * k);
    gamma = gamma != scalar(0) ? scalar(1) / gamma : scalar(0);
    bias = gamma * h * k * C;
}

// M + [rA]x iA [rA]x^T + [rB]x iB [rB]x^T: the point constraint's effective mass.
BX_HD Mat33 PointMass(scalar mA, scalar mB, const Mat33& iA, const Mat33& iB, const Vec3& rA, const Vec3& rB) {
    Mat33 RA = Skew(rA);
    Mat33 RAT = Transpose(RA);
    Mat33 RB = Skew(rB);
    Mat33 RBT = Transpose(RB);
    Mat33 M = Diagonal(mA + mB, mA + mB, mA + mB);
    return M + RA * iA * RAT + RB * iB * RBT;
}

// InitializeVelocityConstraints of sphere_joint.cpp:37-79, weld_joint.cpp:83-177,
// spring_joint.cpp:105-173, mouse_joint.cpp:54-100, friction_joint.cpp:39-86, motor_joint.cpp:41-113.
BX_HD void InitOtherCon(JointCon& jc, const Joint& j, const Body& A, const Body& B, u32 indexA, u32 indexB,
                        const PosBody* pos, scalar dt, scalar invDt, scalar dtRatio, bool warm) {
    jc.indexA = indexA;
    jc.indexB = indexB;
    jc.mA = A.invMass;
    jc.mB = B.invMass;
    jc.localCenterA = A.localCenter;
    jc.localCenterB = B.localCenter;
    jc.localInvIA = A.invI;
    jc.localInvIB = B.invI;
    jc.iA = pos[indexA].I;
    jc.iB = pos[indexB].I;
    jc.localAnchorA = j.localAnchorA;
    jc.localAnchorB = j.localAnchorB;
    jc.frequencyHz = j.frequencyHz;
    jc.length = j.length;
    jc.maxForce = j.maxForce;
    jc.maxTorque = j.maxTorque;
    jc.correctionFactor = j.correctionFactor;
    jc.invDt = invDt;
    // the members of other joint types the impulse store-back reads
    jc.motorImpulse = j.motorImpulse;
    jc.limitImpulse = j.limitImpulse;
    jc.angularImpulse = j.angularImpulse;
    jc.linearImpulse2 = j.linearImpulse2;
    jc.springImpulse = j.springImpulse;
    const Vec3 xA = pos[indexA].x, xB = pos[indexB].x;
    const Quat qA = pos[indexA].q, qB = pos[indexB].q;
    const scalar h = dt;
    switch (j.type) {
    case kJointSphere:
    case kJointFriction:
        jc.rA = Mul(qA, j.localAnchorA - jc.localCenterA);
        jc.rB = Mul(qB, j.localAnchorB - jc.localCenterB);
        jc.linearMass = PointMass(jc.mA, jc.mB, jc.iA, jc.iB, jc.rA, jc.rB);
        jc.angularMass3 = jc.iA + jc.iB;
        break;
    case kJointWeld: {
        jc.rA = Mul(qA, j.localAnchorA - jc.localCenterA);
        jc.rB = Mul(qB, j.localAnchorB - jc.localCenterB);
        jc.linearMass = PointMass(jc.mA, jc.mB, jc.iA, jc.iB, jc.rA, jc.rB);
        Quat q1 = Conjugate(qA) * qB;  // weld_joint.cpp:132-148
        if (Dot(q1, j.referenceRotation) < scalar(0)) q1 = -q1;
        Quat q = scalar(2) * (j.referenceRotation - q1) * Conjugate(q1);
        Vec3 C = Mul(qA, -q.v);
        if (j.frequencyHz > scalar(0)) {
            Mat33 invM = jc.iA + jc.iB;
            Mat33 m = Inverse(invM);
            Vec3 gamma, bias;
            SoftCoefficients(gamma.x, bias.x, j.frequencyHz, j.dampingRatio, m.x.x, C.x, h);
            SoftCoefficients(gamma.y, bias.y, j.frequencyHz, j.dampingRatio, m.y.y, C.y, h);
            SoftCoefficients(gamma.z, bias.z, j.frequencyHz, j.dampingRatio, m.z.z, C.z, h);
            jc.gamma3 = Diagonal(gamma.x, gamma.y, gamma.z);
            jc.bias3 = bias;
            invM = invM + jc.gamma3;
            jc.angularMass3 = Inverse(invM);
        } else {
            jc.gamma3 = Diagonal(scalar(0), scalar(0), scalar(0));
            jc.angularMass3 = Inverse(jc.iA + jc.iB);
            jc.bias3 = j.correctionFactor * invDt * C;
        }
        break;
    }
    case kJointSpring: {
        jc.rA = Mul(qA, j.localAnchorA - jc.localCenterA);
        jc.rB = Mul(qB, j.localAnchorB - jc.localCenterB);
        jc.n = xB + jc.rB - xA - jc.rA;
        scalar length = Length(jc.n);
        if (length > kLinearSlop) jc.n /= length;
        else jc.n.SetZero();
        Vec3 rnA = Cross(jc.rA, jc.n);
        Vec3 rnB = Cross(jc.rB, jc.n);
        scalar mass = jc.mA + jc.mB + Dot(jc.iA * rnA, rnA) + Dot(jc.iB * rnB, rnB);
        jc.mass1 = mass > scalar(0) ? scalar(1) / mass : scalar(0);
        if (j.frequencyHz > scalar(0)) {
            scalar C = length - j.length;
            scalar omega = scalar(2) * kPi * j.frequencyHz;
            scalar d = scalar(2) * jc.mass1 * j.dampingRatio * omega;
            scalar k = jc.mass1 * omega * omega;
            jc.gamma1 = h * (d + h * k);
            jc.gamma1 = jc.gamma1 != scalar(0) ? scalar(1) / jc.gamma1 : scalar(0);
            jc.bias1 = jc.gamma1 * h * k * C;
            mass += jc.gamma1;
            jc.mass1 = mass != scalar(0) ? scalar(1) / mass : scalar(0);
        } else {
            jc.bias1 = scalar(0);
            jc.gamma1 = scalar(0);
        }
        break;
    }
    case kJointMouse: {
        jc.rB = Mul(qB, j.localAnchorB - jc.localCenterB);
        Mat33 M = Diagonal(jc.mB, jc.mB, jc.mB);
        Mat33 RB = Skew(jc.rB);
        Mat33 RBT = Transpose(RB);
        Mat33 invM = M + RB * jc.iB * RBT;
        Mat33 m = Inverse(invM);
        Vec3 C = xB + jc.rB - j.target;
        Vec3 gamma, bias;
        SoftCoefficients(gamma.x, bias.x, j.frequencyHz, j.dampingRatio, m.x.x, C.x, h);
        SoftCoefficients(gamma.y, bias.y, j.frequencyHz, j.dampingRatio, m.y.y, C.y, h);
        SoftCoefficients(gamma.z, bias.z, j.frequencyHz, j.dampingRatio, m.z.z, C.z, h);
        jc.gamma3 = Diagonal(gamma.x, gamma.y, gamma.z);
        jc.bias3 = bias;
        invM = invM + jc.gamma3;
        jc.linearMass = Inverse(invM);
        break;
    }
    case kJointMotor: {
        jc.rA = Mul(qA, j.linearOffset - jc.localCenterA);
        jc.rB = Mul(qB, -jc.localCenterB);
        jc.linearMass = PointMass(jc.mA, jc.mB, jc.iA, jc.iB, jc.rA, jc.rB);
        jc.linearError = xB + jc.rB - xA - jc.rA;
        Quat q1 = Conjugate(qA) * qB;
        if (Dot(q1, j.angularOffset) < scalar(0)) q1 = -q1;
        Quat qw = invDt * scalar(2) * (j.angularOffset - q1) * Conjugate(q1);
        jc.angularVelocity = Mul(qA, qw.v);
        jc.angularVelocity *= j.correctionFactor;
        jc.angularMass3 = jc.iA + jc.iB;
        break;
    }
    default: break;
    }
    Vec3 linearImpulse = j.linearImpulse, angularImpulse3 = j.angularImpulse3;
    scalar impulse1 = j.impulse1;
    if (warm) {
        linearImpulse *= dtRatio;
        angularImpulse3 *= dtRatio;
        impulse1 *= dtRatio;
    } else {
        linearImpulse.SetZero();
        angularImpulse3.SetZero();
        impulse1 = scalar(0);
    }
    jc.linearImpulse = linearImpulse;
    jc.angularImpulse3 = angularImpulse3;
    jc.impulse1 = impulse1;
}

// A point impulse P at rA / rB on the two bodies.
BX_HD void ApplyPoint(const JointCon& jc, const Vec3& P, Vec3& vA, Vec3& wA, Vec3& vB, Vec3& wB) {
    vA -= jc.mA * P;
    wA -= jc.iA * Cross(jc.rA, P);
    vB += jc.mB * P;
    wB += jc.iB * Cross(jc.rB, P);
}

// An accumulated impulse clamped to |impulse| <= maxImpulse; returns the applied change
// (friction_joint.cpp:118-133, motor_joint.cpp:160-175).
BX_HD Vec3 ClampedDelta(Vec3& accumulated, const Vec3& impulse, scalar maxImpulse) {
    Vec3 old = accumulated;
    accumulated += impulse;
    if (LengthSquared(accumulated) > maxImpulse * maxImpulse) {
        accumulated.Normalize();
        accumulated *= maxImpulse;
    }
    return accumulated - old;
}

template <class V>
BX_HD void WarmStartOther(const JointCon& jc, V* vel) {
    if (jc.type == kJointMouse) {  // mouse_joint.cpp:102-106
        vel[jc.indexB].v += jc.mB * jc.linearImpulse;
        vel[jc.indexB].w += jc.iB * Cross(jc.rB, jc.linearImpulse);
        return;
    }
    if (jc.type == kJointSpring) {  // spring_joint.cpp:175-183, in place
        Vec3 P = jc.impulse1 * jc.n;
        vel[jc.indexA].v -= jc.mA * P;
        vel[jc.indexA].w -= jc.iA * Cross(jc.rA, P);
        vel[jc.indexB].v += jc.mB * P;
        vel[jc.indexB].w += jc.iB * Cross(jc.rB, P);
        return;
    }
    Vec3 vA = vel[jc.indexA].v, wA = vel[jc.indexA].w, vB = vel[jc.indexB].v, wB = vel[jc.indexB].w;
    if (jc.type == kJointFriction) {  // friction_joint.cpp:88-107
        Vec3 P = jc.linearImpulse;
        vA -= jc.mA * P;
        wA -= jc.iA * (Cross(jc.rA, P) + jc.angularImpulse3);
        vB += jc.mB * P;
        wB += jc.iB * (Cross(jc.rB, P) + jc.angularImpulse3);
    } else {  // sphere_joint.cpp:81-97; the weld and motor joints add their angular impulse
        ApplyPoint(jc, jc.linearImpulse, vA, wA, vB, wB);
        if (jc.type == kJointWeld || jc.type == kJointMotor) {
            wA -= jc.iA * jc.angularImpulse3;
            wB += jc.iB * jc.angularImpulse3;
        }
    }
    vel[jc.indexA].v = vA;
    vel[jc.indexA].w = wA;
    vel[jc.indexB].v = vB;
    vel[jc.indexB].w = wB;
}

template <class V>
BX_HD void SolveOtherVelocity(JointCon& jc, V* vel, scalar dt) {
    const scalar h = dt;
    if (jc.type == kJointMouse) {  // mouse_joint.cpp:108-131
        Vec3 vB = vel[jc.indexB].v, wB = vel[jc.indexB].w;
        Vec3 Cdot = vB + Cross(wB, jc.rB);
        Vec3 impulse = -jc.linearMass * (Cdot + jc.bias3 + jc.gamma3 * jc.linearImpulse);
        Vec3 oldImpulse = jc.linearImpulse;
        jc.linearImpulse += impulse;
        scalar maxImpulse = h * jc.maxForce;
        if (Dot(jc.linearImpulse, jc.linearImpulse) > maxImpulse * maxImpulse)
            jc.linearImpulse *= maxImpulse / Length(jc.linearImpulse);
        impulse = jc.linearImpulse - oldImpulse;
        vB += jc.mB * impulse;
        wB += jc.iB * Cross(jc.rB, impulse);
        vel[jc.indexB].v = vB;
        vel[jc.indexB].w = wB;
        return;
    }
    Vec3 vA = vel[jc.indexA].v, wA = vel[jc.indexA].w, vB = vel[jc.indexB].v, wB = vel[jc.indexB].w;
    if (jc.type == kJointSphere || jc.type == kJointWeld) {  // sphere_joint.cpp:99-122, weld_joint.cpp:207-246
        Vec3 Cdot = vB + Cross(wB, jc.rB) - vA - Cross(wA, jc.rA);
        Vec3 impulse = jc.linearMass.Solve(-Cdot);
        jc.linearImpulse += impulse;
        ApplyPoint(jc, impulse, vA, wA, vB, wB);
        if (jc.type == kJointWeld) {
            Vec3 dw = wB - wA;
            Vec3 ai = -jc.angularMass3 * (dw + jc.bias3 + jc.gamma3 * jc.angularImpulse3);
            jc.angularImpulse3 += ai;
            wA -= jc.iA * ai;
            wB += jc.iB * ai;
        }
    } else if (jc.type == kJointSpring) {  // spring_joint.cpp:185-212
        Vec3 dv = vB + Cross(wB, jc.rB) - vA - Cross(wA, jc.rA);
        scalar Cdot = Dot(jc.n, dv);
        scalar impulse = -jc.mass1 * (Cdot + jc.bias1 + jc.gamma1 * jc.impulse1);
        jc.impulse1 += impulse;
        ApplyPoint(jc, impulse * jc.n, vA, wA, vB, wB);
    } else if (jc.type == kJointFriction) {  // friction_joint.cpp:109-158: angular, then linear
        Vec3 ai = ClampedDelta(jc.angularImpulse3, jc.angularMass3.Solve(-(wB - wA)), h * jc.maxTorque);
        wA -= jc.iA * ai;
        wB += jc.iB * ai;
        Vec3 Cdot = vB + Cross(wB, jc.rB) - vA - Cross(wA, jc.rA);
        Vec3 li = ClampedDelta(jc.linearImpulse, jc.linearMass.Solve(-Cdot), h * jc.maxForce);
        ApplyPoint(jc, li, vA, wA, vB, wB);
    } else if (jc.type == kJointMotor) {  // motor_joint.cpp:143-200: linear, then angular
        Vec3 Cdot = vB + Cross(wB, jc.rB) - vA - Cross(wA, jc.rA) + jc.invDt * jc.correctionFactor * jc.linearError;
        Vec3 li = ClampedDelta(jc.linearImpulse, jc.linearMass.Solve(-Cdot), h * jc.maxForce);
        ApplyPoint(jc, li, vA, wA, vB, wB);
        Vec3 ai = ClampedDelta(jc.angularImpulse3, jc.angularMass3.Solve(-((wB - wA) - jc.angularVelocity)),
                               h * jc.maxTorque);
        wA -= jc.iA * ai;
        wB += jc.iB * ai;
    }
    vel[jc.indexA].v = vA;
    vel[jc.indexA].w = wA;
    vel[jc.indexB].v = vB;
    vel[jc.indexB].w = wB;
}

// SolvePositionConstraints of the sphere and weld joints (sphere_joint.cpp:124-165,
// weld_joint.cpp:248-296) and the rigid spring (spring_joint.cpp:214-271, orientations not
// renormalized); the mouse, friction and motor joints and a soft spring have none.
template <class P>
BX_HD bool SolveOtherPosition(const JointCon& jc, P* pos) {
    if (jc.type == kJointMouse || jc.type == kJointFriction || jc.type == kJointMotor ||
        (jc.type == kJointSpring && jc.frequencyHz > scalar(0)))
        return true;
    Vec3 xA = pos[jc.indexA].x, xB = pos[jc.indexB].x;
    Quat qA = pos[jc.indexA].q, qB = pos[jc.indexB].q;
    Mat33 iA = pos[jc.indexA].I, iB = pos[jc.indexB].I;
    const scalar mA = jc.mA, mB = jc.mB;
    bool solved;
    Vec3 rA = Mul(qA, jc.localAnchorA - jc.localCenterA);
    Vec3 rB = Mul(qB, jc.localAnchorB - jc.localCenterB);
    if (jc.type == kJointSpring) {
        Vec3 n = xB + rB - xA - rA;
        scalar length = Length(n);
        scalar C = length - jc.length;
        C = Clamp(C, -kMaxLinearCorrection, kMaxLinearCorrection);
        n.Normalize();
        Vec3 rnA = Cross(rA, n);
        Vec3 rnB = Cross(rB, n);
        scalar kMass = mA + mB + Dot(iA * rnA, rnA) + Dot(iB * rnB, rnB);
        scalar mass = kMass > scalar(0) ? scalar(1) / kMass : scalar(0);
        scalar lambda = -mass * C;
        Vec3 impulse = lambda * n;
        xA -= mA * impulse;
        qA -= Derivative(qA, iA * Cross(rA, impulse));
        iA = RotateToFrame(jc.localInvIA, qA);
        xB += mB * impulse;
        qB += Derivative(qB, iB * Cross(rB, impulse));
        iB = RotateToFrame(jc.localInvIB, qB);
        solved = Abs(C) < kLinearSlop;
    } else {
        Mat33 mass = PointMass(mA, mB, iA, iB, rA, rB);
        Vec3 C = xB + rB - xA - rA;
        Vec3 impulse = mass.Solve(-C);
        xA -= mA * impulse;
        qA -= Derivative(qA, iA * Cross(rA, impulse));
        qA.Normalize();
        iA = RotateToFrame(jc.localInvIA, qA);
        xB += mB * impulse;
        qB += Derivative(qB, iB * Cross(rB, impulse));
        qB.Normalize();
        iB = RotateToFrame(jc.localInvIB, qB);
        solved = Length(C) <= kLinearSlop;
    }
    pos[jc.indexA].x = xA;
    pos[jc.indexA].q = qA;
    pos[jc.indexB].x = xB;
    pos[jc.indexB].q = qB;
    pos[jc.indexA].I = iA;
    pos[jc.indexB].I = iB;
    return solved;
}


// ================================================================== cone, prismatic and wheel joints
// The prismatic and wheel joints' two rows perpendicular to the axis (prismatic_joint.cpp:113-136,
// wheel_joint.cpp:103-126 and the position solvers): K = [k11 k12; k12 k22] over the perp1/perp2
// Jacobians, inverted.
BX_HD Mat22 PerpMass(scalar mA, scalar mB, const Mat33& iA, const Mat33& iB, const Vec3& s1, const Vec3& s2,
                     const Vec3& s3, const Vec3& s4) {
    scalar k11 = mA + mB + Dot(s1, iA * s1) + Dot(s2, iB * s2);
    scalar k12 = Dot(s1, iA * s3) + Dot(s2, iB * s4);
    scalar k22 = mA + mB + Dot(s3, iA * s3) + Dot(s4, iB * s4);
    Mat22 K;
    K.x.x = k11;
    K.x.y = k12;
    K.y.x = k12;
    K.y.y = k22;
    return Inverse(K);
}

// Linear and angular impulses P, LA, LB on the two bodies: vA -= mA P, wA -= iA LA, vB += mB P,
// wB += iB LB.
BX_HD void ApplyRows(const JointCon& jc, const Vec3& P, const Vec3& LA, const Vec3& LB, Vec3& vA, Vec3& wA, Vec3& vB,
                     Vec3& wB) {
    vA -= jc.mA * P;
    wA -= jc.iA * LA;
    vB += jc.mB * P;
    wB += jc.iB * LB;
}

// InitializeVelocityConstraints of cone_joint.cpp:70-168, prismatic_joint.cpp:63-200 and
// wheel_joint.cpp:62-190. The cone and prismatic limit states are written back to the joint.
BX_HD void InitAxisCon(JointCon& jc, Joint& j, const Body& A, const Body& B, u32 indexA, u32 indexB,
                       const PosBody* pos, scalar dt, scalar invDt, scalar dtRatio, bool warm) {
    jc.indexA = indexA;
    jc.indexB = indexB;
    jc.mA = A.invMass;
    jc.mB = B.invMass;
    jc.localCenterA = A.localCenter;
    jc.localCenterB = B.localCenter;
    jc.localInvIA = A.invI;
    jc.localInvIB = B.invI;
    jc.iA = pos[indexA].I;
    jc.iB = pos[indexB].I;
    jc.localAnchorA = j.localAnchorA;
    jc.localAnchorB = j.localAnchorB;
    jc.localRotationA = j.localRotationA;
    jc.localRotationB = j.localRotationB;
    jc.localXAxisA = j.localXAxisA;
    jc.localYAxisA = j.localYAxisA;
    jc.localZAxisA = j.localZAxisA;
    jc.localXAxisB = j.localXAxisB;
    jc.enableLimit = j.enableLimit;
    jc.enableMotor = j.enableMotor;
    jc.motorSpeed = j.motorSpeed;
    jc.maxMotorTorque = j.maxMotorTorque;
    jc.maxForce = j.maxForce;
    jc.lowerAngle = j.lowerTranslation;
    jc.upperAngle = j.upperTranslation;
    jc.coneAngle = j.coneAngle;
    jc.referenceCosine = j.referenceCosine;
    jc.referenceAngle = j.referenceAngle;
    jc.frequencyHz = j.frequencyHz;
    const scalar mA = jc.mA, mB = jc.mB;
    const Mat33 iA = jc.iA, iB = jc.iB;
    const Vec3 xA = pos[indexA].x, xB = pos[indexB].x;
    const Quat qA = pos[indexA].q, qB = pos[indexB].q;
    u32 limitState = j.limitState;
    scalar limitImpulse = j.limitImpulse, motorImpulse = j.motorImpulse, springImpulse = j.springImpulse;
    Vec3 linearImpulse = j.linearImpulse, angularImpulse3 = j.angularImpulse3;
    Vec2 linearImpulse2 = j.linearImpulse2;
    scalar impulse1 = j.impulse1;
    if (j.type == kJointCone) {
        Quat fA = qA * j.localRotationA;
        Quat fB = qB * j.localRotationB;
        Mat33 RfA = fA.GetRotationMatrix();
        Mat33 RfB = fB.GetRotationMatrix();
        jc.rA = Mul(qA, j.localAnchorA - jc.localCenterA);
        jc.rB = Mul(qB, j.localAnchorB - jc.localCenterB);
        jc.linearMass = PointMass(mA, mB, iA, iB, jc.rA, jc.rB);
        if (j.enableLimit) {
            Vec3 u1 = RfA.x;
            Vec3 u2 = RfB.x;
            jc.motorAxis = Cross(u2, u1);
            scalar mass = Dot((iA + iB) * jc.motorAxis, jc.motorAxis);
            jc.mass1 = mass > scalar(0) ? scalar(1) / mass : scalar(0);
            scalar cosine = Dot(u2, u1);
            scalar sine = Length(jc.motorAxis);
            scalar angle = JointAngle(sine, cosine);
            if (scalar(0.5) * j.coneAngle < angle) {
                if (limitState != kAtLowerLimit) {
                    limitState = kAtLowerLimit;
                    limitImpulse = scalar(0);
                }
            } else {
                limitState = kInactiveLimit;
                limitImpulse = scalar(0);
            }
        } else {
            jc.motorAxis.SetZero();
            jc.mass1 = scalar(0);
            limitState = kInactiveLimit;
        }
        if (warm) {
            linearImpulse *= dtRatio;
            limitImpulse *= dtRatio;
        } else {
            linearImpulse.SetZero();
            limitImpulse = scalar(0);
        }
    } else {
        Vec3 rA = Mul(qA, j.localAnchorA - jc.localCenterA);
        Vec3 rB = Mul(qB, j.localAnchorB - jc.localCenterB);
        Vec3 d = xB + rB - xA - rA;
        jc.perp1 = Mul(qA, j.localYAxisA);
        jc.s1 = Cross(d + rA, jc.perp1);
        jc.s2 = Cross(rB, jc.perp1);
        jc.perp2 = Mul(qA, j.localZAxisA);
        jc.s3 = Cross(d + rA, jc.perp2);
        jc.s4 = Cross(rB, jc.perp2);
        if (j.type == kJointPrismatic) {
            jc.motorAxis = Mul(qA, j.localXAxisA);
            jc.a1 = Cross(d + rA, jc.motorAxis);
            jc.a2 = Cross(rB, jc.motorAxis);
            jc.motorMass = mA + mB + Dot(iA * jc.a1, jc.a1) + Dot(iB * jc.a2, jc.a2);
            if (jc.motorMass > scalar(0)) jc.motorMass = scalar(1) / jc.motorMass;
            jc.mass22 = PerpMass(mA, mB, iA, iB, jc.s1, jc.s2, jc.s3, jc.s4);
            jc.angularMass3 = Inverse(iA + iB);
            Quat q1 = Conjugate(qA) * qB;
            Quat q2 = j.referenceRotation;
            if (Dot(q1, q2) < scalar(0)) q1 = -q1;
            Quat q = scalar(2) * (q2 - q1) * Conjugate(q1);
            Vec3 C = Mul(qA, -q.v);
            jc.bias3 = j.correctionFactor * invDt * C;
            if (j.enableLimit) {
                scalar translation = Dot(jc.motorAxis, d);
                if (Abs(j.upperTranslation - j.lowerTranslation) < scalar(2) * kLinearSlop) {
                    limitState = kEqualLimits;
                } else if (translation <= j.lowerTranslation) {
                    if (limitState != kAtLowerLimit) {
                        limitState = kAtLowerLimit;
                        limitImpulse = scalar(0);
                    }
                } else if (translation >= j.upperTranslation) {
                    if (limitState != kAtUpperLimit) {
                        limitState = kAtUpperLimit;
                        limitImpulse = scalar(0);
                    }
                } else {
                    limitState = kInactiveLimit;
                    limitImpulse = scalar(0);
                }
            } else {
                limitState = kInactiveLimit;
                limitImpulse = scalar(0);
            }
            if (!j.enableMotor) motorImpulse = scalar(0);
            if (warm) {
                linearImpulse2 *= dtRatio;
                angularImpulse3 *= dtRatio;
                motorImpulse *= dtRatio;
                limitImpulse *= dtRatio;
            } else {
                linearImpulse2.SetZero();
                angularImpulse3.SetZero();
                motorImpulse = scalar(0);
                limitImpulse = scalar(0);
            }
        } else {  // wheel
            jc.mass22 = PerpMass(mA, mB, iA, iB, jc.s1, jc.s2, jc.s3, jc.s4);
            jc.mass1 = scalar(0);
            jc.bias1 = scalar(0);
            jc.gamma1 = scalar(0);
            // with the spring off (or the motor off) the reference leaves the spring's axis and
            // arms (the motor's axis) unset and multiplies them by a zero impulse; they are zero here
            jc.motorAxis.SetZero();
            jc.a1.SetZero();
            jc.a2.SetZero();
            jc.axisB.SetZero();
            if (j.frequencyHz > scalar(0)) {
                jc.motorAxis = Mul(qA, j.localXAxisA);
                jc.a1 = Cross(d + rA, jc.motorAxis);
                jc.a2 = Cross(rB, jc.motorAxis);
                scalar invMass = mA + mB + Dot(iA * jc.a1, jc.a1) + Dot(iB * jc.a2, jc.a2);
                if (invMass > scalar(0)) {
                    jc.mass1 = scalar(1) / invMass;
                    scalar C = Dot(d, jc.motorAxis);
                    scalar omega = scalar(2) * kPi * j.frequencyHz;
                    scalar dd = scalar(2) * jc.mass1 * j.dampingRatio * omega;
                    scalar k = jc.mass1 * omega * omega;
                    scalar h = dt;
                    jc.gamma1 = h * (dd + h * k);
                    if (jc.gamma1 > scalar(0)) jc.gamma1 = scalar(1) / jc.gamma1;
                    jc.bias1 = C * h * k * jc.gamma1;
                    jc.mass1 = invMass + jc.gamma1;
                    if (jc.mass1 > scalar(0)) jc.mass1 = scalar(1) / jc.mass1;
                }
            } else {
                springImpulse = scalar(0);
            }
            if (j.enableMotor) {
                jc.axisB = Mul(qB, j.localXAxisB);
                jc.motorMass = Dot((iA + iB) * jc.axisB, jc.axisB);
                if (jc.motorMass > scalar(0)) jc.motorMass = scalar(1) / jc.motorMass;
            } else {
                motorImpulse = scalar(0);
            }
            {
                Vec3 axisA = Mul(qA, j.localXAxisA);
                Vec3 axisB = Mul(qB, j.localXAxisB);
                jc.u = Cross(axisA, axisB);
                jc.angularMass1 = Dot((iA + iB) * jc.u, jc.u);
                if (jc.angularMass1 > scalar(0)) jc.angularMass1 = scalar(1) / jc.angularMass1;
            }
            if (warm) {
                linearImpulse2 *= dtRatio;
                springImpulse *= dtRatio;
                motorImpulse *= dtRatio;
                impulse1 *= dtRatio;
            } else {
                linearImpulse2.SetZero();
                springImpulse = scalar(0);
                motorImpulse = scalar(0);
                impulse1 = scalar(0);
            }
        }
    }
    j.limitState = limitState;
    jc.limitState = limitState;
    jc.limitImpulse = limitImpulse;
    jc.motorImpulse = motorImpulse;
    jc.springImpulse = springImpulse;
    jc.linearImpulse = linearImpulse;
    jc.angularImpulse3 = angularImpulse3;
    jc.linearImpulse2 = linearImpulse2;
    jc.impulse1 = impulse1;
    jc.angularImpulse = j.angularImpulse;
}

template <class V>
BX_HD void WarmStartAxis(const JointCon& jc, V* vel) {
    Vec3 vA = vel[jc.indexA].v, wA = vel[jc.indexA].w, vB = vel[jc.indexB].v, wB = vel[jc.indexB].w;
    if (jc.type == kJointCone) {  // cone_joint.cpp:170-196
        ApplyPoint(jc, jc.linearImpulse, vA, wA, vB, wB);
        if (jc.enableLimit && jc.limitState != kInactiveLimit) {
            Vec3 L = jc.limitImpulse * jc.motorAxis;
            wA -= jc.iA * L;
            wB += jc.iB * L;
        }
    } else if (jc.type == kJointPrismatic) {  // prismatic_joint.cpp:202-222
        const Vec2 li = jc.linearImpulse2;
        const scalar axial = jc.motorImpulse + jc.limitImpulse;
        Vec3 P = li.x * jc.perp1 + li.y * jc.perp2 + axial * jc.motorAxis;
        Vec3 LA = li.x * jc.s1 + li.y * jc.s3 + jc.angularImpulse3 + axial * jc.a1;
        Vec3 LB = li.x * jc.s2 + li.y * jc.s4 + jc.angularImpulse3 + axial * jc.a2;
        ApplyRows(jc, P, LA, LB, vA, wA, vB, wB);
    } else {  // wheel_joint.cpp:192-222
        const Vec2 li = jc.linearImpulse2;
        Vec3 P = li.x * jc.perp1 + li.y * jc.perp2 + jc.springImpulse * jc.motorAxis;
        Vec3 LA = li.x * jc.s1 + li.y * jc.s3 + jc.springImpulse * jc.a1 + jc.motorImpulse * jc.axisB;
        Vec3 LB = li.x * jc.s2 + li.y * jc.s4 + jc.springImpulse * jc.a2 + jc.motorImpulse * jc.axisB;
        ApplyRows(jc, P, LA, LB, vA, wA, vB, wB);
        Vec3 L = jc.impulse1 * jc.u;
        wA += jc.iA * L;
        wB -= jc.iB * L;
    }
    vel[jc.indexA].v = vA;
    vel[jc.indexA].w = wA;
    vel[jc.indexB].v = vB;
    vel[jc.indexB].w = wB;
}

// The two rows perpendicular to the axis (prismatic_joint.cpp:249-268, wheel_joint.cpp:263-282).
BX_HD void SolvePerpRows(JointCon& jc, Vec3& vA, Vec3& wA, Vec3& vB, Vec3& wB) {
    Vec2 Cdot;
    Cdot.x = Dot(jc.perp1, vB - vA) + Dot(jc.s2, wB) - Dot(jc.s1, wA);
    Cdot.y = Dot(jc.perp2, vB - vA) + Dot(jc.s4, wB) - Dot(jc.s3, wA);
    Vec2 impulse = jc.mass22 * (-Cdot);
    jc.linearImpulse2 += impulse;
    Vec3 P = impulse.x * jc.perp1 + impulse.y * jc.perp2;
    Vec3 LA = impulse.x * jc.s1 + impulse.y * jc.s3;
    Vec3 LB = impulse.x * jc.s2 + impulse.y * jc.s4;
    ApplyRows(jc, P, LA, LB, vA, wA, vB, wB);
}

template <class V>
BX_HD void SolveAxisVelocity(JointCon& jc, V* vel, scalar dt) {
    Vec3 vA = vel[jc.indexA].v, wA = vel[jc.indexA].w, vB = vel[jc.indexB].v, wB = vel[jc.indexB].w;
    if (jc.type == kJointCone) {  // cone_joint.cpp:198-240
        Vec3 Cdot = vB + Cross(wB, jc.rB) - vA - Cross(wA, jc.rA);
        Vec3 P = jc.linearMass.Solve(-Cdot);
        jc.linearImpulse += P;
        ApplyPoint(jc, P, vA, wA, vB, wB);
        if (jc.enableLimit && jc.limitState != kInactiveLimit) {
            scalar cdot = Dot(jc.motorAxis, wB - wA);
            scalar impulse = -jc.mass1 * cdot;
            scalar oldImpulse = jc.limitImpulse;
            jc.limitImpulse = Max(jc.limitImpulse + impulse, scalar(0));
            impulse = jc.limitImpulse - oldImpulse;
            Vec3 L = impulse * jc.motorAxis;
            wA -= jc.iA * L;
            wB += jc.iB * L;
        }
    } else if (jc.type == kJointPrismatic) {  // prismatic_joint.cpp:224-315
        if (jc.enableMotor && jc.limitState != kEqualLimits) {
            scalar Cdot = Dot(jc.motorAxis, vB - vA) + Dot(jc.a2, wB) - Dot(jc.a1, wA);
            scalar impulse = jc.motorMass * (jc.motorSpeed - Cdot);
            scalar oldImpulse = jc.motorImpulse;
            scalar maxImpulse = dt * jc.maxForce;
            jc.motorImpulse = Clamp(jc.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = jc.motorImpulse - oldImpulse;
            ApplyRows(jc, impulse * jc.motorAxis, impulse * jc.a1, impulse * jc.a2, vA, wA, vB, wB);
        }
        SolvePerpRows(jc, vA, wA, vB, wB);
        {
            Vec3 Cdot = wB - wA;
            Vec3 impulse = -jc.angularMass3 * (Cdot + jc.bias3);
            jc.angularImpulse3 += impulse;
            wA -= jc.iA * impulse;
            wB += jc.iB * impulse;
        }
        if (jc.enableLimit && jc.limitState != kInactiveLimit) {
            scalar Cdot = Dot(jc.motorAxis, vB - vA) + Dot(jc.a2, wB) - Dot(jc.a1, wA);
            scalar impulse = jc.motorMass * (-Cdot);
            scalar oldImpulse = jc.limitImpulse;
            if (jc.limitState == kAtLowerLimit) jc.limitImpulse = Max(jc.limitImpulse + impulse, scalar(0));
            else if (jc.limitState == kAtUpperLimit) jc.limitImpulse = Min(jc.limitImpulse + impulse, scalar(0));
            impulse = jc.limitImpulse - oldImpulse;
            ApplyRows(jc, impulse * jc.motorAxis, impulse * jc.a1, impulse * jc.a2, vA, wA, vB, wB);
        }
    } else {  // wheel_joint.cpp:224-299
        {
            scalar Cdot = Dot(jc.motorAxis, vB - vA) + Dot(jc.a2, wB) - Dot(jc.a1, wA);
            scalar impulse = -jc.mass1 * (Cdot + jc.bias1 + jc.gamma1 * jc.springImpulse);
            jc.springImpulse += impulse;
            ApplyRows(jc, impulse * jc.motorAxis, impulse * jc.a1, impulse * jc.a2, vA, wA, vB, wB);
        }
        if (jc.enableMotor) {
            scalar dw = Dot(wB - wA, jc.axisB);
            scalar Cdot = dw - jc.motorSpeed;
            scalar impulse = -jc.motorMass * Cdot;
            scalar oldImpulse = jc.motorImpulse;
            scalar maxImpulse = dt * jc.maxMotorTorque;
            jc.motorImpulse = Clamp(jc.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = jc.motorImpulse - oldImpulse;
            Vec3 P = jc.axisB * impulse;
            wA -= jc.iA * P;
            wB += jc.iB * P;
        }
        SolvePerpRows(jc, vA, wA, vB, wB);
        {
            scalar Cdot = Dot(wA - wB, jc.u);
            scalar impulse = -jc.angularMass1 * Cdot;
            jc.impulse1 += impulse;
            Vec3 L = impulse * jc.u;
            wA += jc.iA * L;
            wB -= jc.iB * L;
        }
    }
    vel[jc.indexA].v = vA;
    vel[jc.indexA].w = wA;
    vel[jc.indexB].v = vB;
    vel[jc.indexB].w = wB;
}

// One position correction: xA -= mA P, qA -= dq(iA LA) renormalized, the inverse inertia rotated
// again; and the opposite on B.
BX_HD void MoveBodies(const JointCon& jc, const Vec3& P, const Vec3& LA, const Vec3& LB, Vec3& xA, Quat& qA,
                      Mat33& iA, Vec3& xB, Quat& qB, Mat33& iB) {
    xA -= jc.mA * P;
    qA -= Derivative(qA, iA * LA);
    qA.Normalize();
    iA = RotateToFrame(jc.localInvIA, qA);
    xB += jc.mB * P;
    qB += Derivative(qB, iB * LB);
    qB.Normalize();
    iB = RotateToFrame(jc.localInvIB, qB);
}

// SolvePositionConstraints of cone_joint.cpp:242-322, prismatic_joint.cpp:317-441 and
// wheel_joint.cpp:301-423.
template <class PB>
BX_HD bool SolveAxisPosition(const JointCon& jc, PB* pos) {
    Vec3 xA = pos[jc.indexA].x, xB = pos[jc.indexB].x;
    Quat qA = pos[jc.indexA].q, qB = pos[jc.indexB].q;
    Mat33 iA = pos[jc.indexA].I, iB = pos[jc.indexB].I;
    const scalar mA = jc.mA, mB = jc.mB;
    bool solved;
    if (jc.type == kJointCone) {
        scalar limitError = scalar(0);
        if (jc.enableLimit) {
            Quat fA = qA * jc.localRotationA;
            Quat fB = qB * jc.localRotationB;
            Mat33 RfA = fA.GetRotationMatrix();
            Mat33 RfB = fB.GetRotationMatrix();
            Vec3 u1 = RfA.x;
            Vec3 u2 = RfB.x;
            Vec3 coneAxis = Cross(u2, u1);
            scalar mass = Dot((iA + iB) * coneAxis, coneAxis);
            scalar coneMass = mass > scalar(0) ? scalar(1) / mass : scalar(0);
            scalar cosine = Dot(u2, u1);
            scalar sine = Length(coneAxis);
            scalar angle = JointAngle(sine, cosine);
            if (scalar(0.5) * jc.coneAngle < angle) {
                scalar C = scalar(0.5) * jc.coneAngle - angle;
                limitError = -C;
                C = Clamp(C + kAngularSlop, -kMaxAngularCorrection, scalar(0));
                scalar coneImpulse = -C * coneMass;
                Vec3 L = coneImpulse * coneAxis;
                qA -= Derivative(qA, iA * L);
                qA.Normalize();
                iA = RotateToFrame(jc.localInvIA, qA);
                qB += Derivative(qB, iB * L);
                qB.Normalize();
                iB = RotateToFrame(jc.localInvIB, qB);
            }
        }
        scalar linearError;
        {
            Vec3 rA = Mul(qA, jc.localAnchorA - jc.localCenterA);
            Vec3 rB = Mul(qB, jc.localAnchorB - jc.localCenterB);
            Vec3 C = xB + rB - xA - rA;
            linearError = Length(C);
            Mat33 mass = PointMass(mA, mB, iA, iB, rA, rB);
            Vec3 impulse = mass.Solve(-C);
            MoveBodies(jc, impulse, Cross(rA, impulse), Cross(rB, impulse), xA, qA, iA, xB, qB, iB);
        }
        solved = linearError <= kLinearSlop && limitError <= kAngularSlop;
    } else {
        Vec3 rA = Mul(qA, jc.localAnchorA - jc.localCenterA);
        Vec3 rB = Mul(qB, jc.localAnchorB - jc.localCenterB);
        Vec3 d = xB + rB - xA - rA;
        Vec3 axis = Mul(qA, jc.localXAxisA);
        Vec3 a1 = Cross(d + rA, axis);
        Vec3 a2 = Cross(rB, axis);
        Vec3 perp1 = Mul(qA, jc.localYAxisA);
        Vec3 s1 = Cross(d + rA, perp1);
        Vec3 s2 = Cross(rB, perp1);
        Vec3 perp2 = Mul(qA, jc.localZAxisA);
        Vec3 s3 = Cross(d + rA, perp2);
        Vec3 s4 = Cross(rB, perp2);
        Vec3 axisB = jc.type == kJointWheel ? Mul(qB, jc.localXAxisB) : Vec3(0, 0, 0);
        scalar linearError = scalar(0);
        {
            Vec2 C;
            C.x = Dot(perp1, d);
            C.y = Dot(perp2, d);
            linearError += Abs(C.x) + Abs(C.y);
            Mat22 linearMass = PerpMass(mA, mB, iA, iB, s1, s2, s3, s4);
            Vec2 impulse = linearMass * (-C);
            Vec3 P = impulse.x * perp1 + impulse.y * perp2;
            Vec3 LA = impulse.x * s1 + impulse.y * s3;
            Vec3 LB = impulse.x * s2 + impulse.y * s4;
            MoveBodies(jc, P, LA, LB, xA, qA, iA, xB, qB, iB);
        }
        if (jc.type == kJointPrismatic) {
            if (jc.enableLimit) {
                scalar C = scalar(0);
                scalar translation = Dot(axis, d);
                if (Abs(jc.upperAngle - jc.lowerAngle) < scalar(2) * kLinearSlop) {
                    C = Clamp(translation, -kMaxLinearCorrection, kMaxLinearCorrection);
                    linearError = Max(linearError, Abs(translation));
                } else if (translation <= jc.lowerAngle) {
                    C = Clamp(translation - jc.lowerAngle + kLinearSlop, -kMaxLinearCorrection, scalar(0));
                    linearError = Max(linearError, jc.lowerAngle - translation);
                } else if (translation >= jc.upperAngle) {
                    C = Clamp(translation - jc.upperAngle - kLinearSlop, scalar(0), kMaxLinearCorrection);
                    linearError = Max(linearError, translation - jc.upperAngle);
                }
                scalar limitMass = mA + mB + Dot(iA * a1, a1) + Dot(iB * a2, a2);
                limitMass = limitMass > scalar(0) ? scalar(1) / limitMass : scalar(0);
                scalar impulse = limitMass * -C;
                MoveBodies(jc, impulse * axis, impulse * a1, impulse * a2, xA, qA, iA, xB, qB, iB);
            }
            solved = linearError <= kLinearSlop;
        } else {
            scalar angularError = scalar(0);
            {
                Vec3 u = Cross(axis, axisB);
                scalar cosine = Dot(axis, axisB);
                scalar C = cosine - jc.referenceCosine;
                scalar sine = Length(u);
                scalar angle = JointAngle(sine, cosine);
                scalar angleC = angle - jc.referenceAngle;
                angularError += Abs(angleC);
                scalar mass = Dot((iA + iB) * u, u);
                if (mass > scalar(0)) mass = scalar(1) / mass;
                scalar impulse = -mass * C;
                Vec3 L = impulse * u;
                qA += Derivative(qA, iA * L);
                qA.Normalize();
                iA = RotateToFrame(jc.localInvIA, qA);
                qB -= Derivative(qB, iB * L);
                qB.Normalize();
                iB = RotateToFrame(jc.localInvIB, qB);
            }
            solved = linearError <= kLinearSlop && angularError <= kAngularSlop;
        }
    }
    pos[jc.indexA].x = xA;
    pos[jc.indexA].q = qA;
    pos[jc.indexB].x = xB;
    pos[jc.indexB].q = qB;
    pos[jc.indexA].I = iA;
    pos[jc.indexB].I = iB;
    return solved;
}

// ================================================================== dispatch by joint type
BX_HD void InitJointCon(JointCon& jc, Joint& j, const Body& A, const Body& B, u32 indexA, u32 indexB,
                        const PosBody* pos, scalar dt, scalar invDt, scalar dtRatio, bool warm) {
    jc.type = j.type;
    if (j.type == kJointRevolute) InitRevoluteCon(jc, j, A, B, indexA, indexB, pos, dtRatio, warm);
    else if (IsAxisJoint(j.type)) InitAxisCon(jc, j, A, B, indexA, indexB, pos, dt, invDt, dtRatio, warm);
    else InitOtherCon(jc, j, A, B, indexA, indexB, pos, dt, invDt, dtRatio, warm);
}

template <class V>
BX_HD void WarmStartJoint(const JointCon& jc, V* vel) {
    if (jc.type == kJointRevolute) WarmStartRevolute(jc, vel);
    else if (IsAxisJoint(jc.type)) WarmStartAxis(jc, vel);
    else WarmStartOther(jc, vel);
}

template <class V>
BX_HD void SolveJointVelocity(JointCon& jc, V* vel, scalar dt) {
    if (jc.type == kJointRevolute) SolveRevoluteVelocity(jc, vel, dt);
    else if (IsAxisJoint(jc.type)) SolveAxisVelocity(jc, vel, dt);
    else SolveOtherVelocity(jc, vel, dt);
}

template <class P>
BX_HD bool SolveJointPosition(const JointCon& jc, P* pos) {
    if (jc.type == kJointRevolute) return SolveRevolutePosition(jc, pos);
    if (IsAxisJoint(jc.type)) return SolveAxisPosition(jc, pos);
    return SolveOtherPosition(jc, pos);
}

}  // namespace bx
