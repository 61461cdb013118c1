// bx_solver.h -- island integration and the sequential-impulse contact solver, host/device.
//
// Restates src/bounce/dynamics/island.cpp:59-307 and
// src/bounce/dynamics/contacts/contact_solver.cpp:28-624 for one island whose bodies are
// addressed by island-local index. Constraint order is the island's contact order, which the
// island builder reproduces exactly (reference DFS order), so the Gauss-Seidel sweep is the
// reference's sweep.
#pragma once
#include "bx_types.h"

namespace bx {

BX_HD Vec3 SolveGyro(const Quat& q, const Mat33& Ib, const Vec3& w1, scalar h) {  // island.cpp:60-79
    Vec3 w1b = MulC(q, w1);
    Mat33 J = Ib + h * (Skew(w1b) * Ib - Skew(Ib * w1b));
    Vec3 f = h * Cross(w1b, Ib * w1b);
    w1b -= J.Solve(f);
    return Mul(q, w1b);
}

// island.cpp:86-135: velocity integration of one body; returns the solver body.
BX_HD SolverBody IntegrateVelocity(const Body& b, const Vec3& gravity, scalar h) {
    Vec3 v = b.v, w = b.w;
    if (b.type == kDynamic) {
        v += h * (Mul(b.gravityScale, gravity) + b.invMass * b.force);
        Vec3 dw1 = h * b.worldInvI * b.torque;
        Vec3 w2 = SolveGyro(b.q, b.I, w, h);
        Vec3 dw2 = w2 - w;
        w += dw1 + dw2;
        v.x *= scalar(1) / (scalar(1) + h * b.linearDamping.x);
        v.y *= scalar(1) / (scalar(1) + h * b.linearDamping.y);
        v.z *= scalar(1) / (scalar(1) + h * b.linearDamping.z);
        w.x *= scalar(1) / (scalar(1) + h * b.angularDamping.x);
        w.y *= scalar(1) / (scalar(1) + h * b.angularDamping.y);
        w.z *= scalar(1) / (scalar(1) + h * b.angularDamping.z);
    }
    SolverBody s;
    s.x = b.x;
    s.q = b.q;
    s.I = b.worldInvI;
    s.v = v;
    s.w = w;
    return s;
}

// island.cpp:190-229: position integration with translation / rotation clamps.
BX_HD void IntegratePosition(SolverBody& s, const Body& b, scalar h) {
    if (b.type == kStatic) return;
    Vec3 translation = h * s.v;
    if (Dot(translation, translation) > kMaxTranslationSq) {
        scalar ratio = kMaxTranslation / Length(translation);
        s.v *= ratio;
    }
    Vec3 rotation = h * s.w;
    if (Dot(rotation, rotation) > kMaxRotationSq) {
        scalar ratio = kMaxRotation / Length(rotation);
        s.w *= ratio;
    }
    s.x += h * s.v;
    s.q = Integrate(s.q, s.w, h);
    s.I = RotateToFrame(b.invI, s.q);
}

// contact_solver.cpp:28-150 (constructor) + 152-270 (InitializeVelocityConstraints) for one
// convex contact. `pos` is read at the island's post-velocity-integration state.
BX_HD void InitContactConstraint(ContactConstraint& cc, const Contact& c, const Fixture& fA, const Fixture& fB,
                                 const Body& bA, const Body& bB, u32 indexA, u32 indexB, const SolverBody* sb,
                                 scalar dtRatio, bool warmStarting) {
    cc.indexA = indexA;
    cc.indexB = indexB;
    cc.invMassA = bA.invMass;
    cc.invMassB = bB.invMass;
    cc.localInvIA = bA.invI;
    cc.localInvIB = bB.invI;
    cc.localCenterA = bA.localCenter;
    cc.localCenterB = bB.localCenter;
    cc.radiusA = fA.radius;
    cc.radiusB = fB.radius;
    cc.invIA = sb[indexA].I;
    cc.invIB = sb[indexB].I;
    cc.friction = Sqrt(fA.friction * fB.friction);         // b3MixFriction
    cc.restitution = Max(fA.restitution, fB.restitution);  // b3MixRestitution
    const Manifold& m = c.manifold;
    u32 pointCount = c.manifoldCount > 0 ? m.pointCount : 0;
    cc.pointCount = pointCount;
    cc.motorSpeed = m.motorSpeed;
    cc.tangentSpeed1 = m.tangentSpeed1;
    cc.tangentSpeed2 = m.tangentSpeed2;
    if (warmStarting) {
        cc.tangentImpulse = dtRatio * m.tangentImpulse;
        cc.motorImpulse = dtRatio * m.motorImpulse;
    } else {
        cc.tangentImpulse.SetZero();
        cc.motorImpulse = 0;
    }
    cc.tangentMass.SetZero();
    cc.motorMass = 0;
    cc.rA.SetZero(); cc.rB.SetZero(); cc.normal.SetZero(); cc.tangent1.SetZero(); cc.tangent2.SetZero();
    for (u32 k = 0; k < pointCount; ++k)
        cc.points[k].normalImpulse = warmStarting ? dtRatio * m.points[k].normalImpulse : scalar(0);

    if (c.manifoldCount == 0) return;
    scalar mA = cc.invMassA, mB = cc.invMassB;
    Mat33 iA = cc.invIA, iB = cc.invIB;
    Vec3 xA = sb[indexA].x, xB = sb[indexB].x;
    Quat qA = sb[indexA].q, qB = sb[indexB].q;
    Vec3 vA = sb[indexA].v, wA = sb[indexA].w, vB = sb[indexB].v, wB = sb[indexB].w;
    Transform xfA(xA - Mul(qA, cc.localCenterA), qA);
    Transform xfB(xB - Mul(qB, cc.localCenterB), qB);
    // b3WorldManifold::Initialize (collision.cpp:76-101)
    Vec3 center, normal;
    center.SetZero();
    normal.SetZero();
    for (u32 k = 0; k < pointCount; ++k) {
        const ManifoldPoint& p = m.points[k];
        Vec3 nA = Mul(xfA.rotation, p.localNormal1);
        Vec3 cA = Mul(xfA, p.localPoint1);
        Vec3 cB = Mul(xfB, p.localPoint2);
        Vec3 pA = cA + cc.radiusA * nA;
        Vec3 pB = cB - cc.radiusB * nA;
        Vec3 point = scalar(0.5) * (pA + pB);
        center += point;
        normal += nA;
        VCPoint& vcp = cc.points[k];
        Vec3 rA = point - xA;
        Vec3 rB = point - xB;
        vcp.rA = rA;
        vcp.rB = rB;
        vcp.normal = nA;
        Vec3 rnA = Cross(rA, nA);
        Vec3 rnB = Cross(rB, nA);
        scalar K = mA + mB + Dot(iA * rnA, rnA) + Dot(iB * rnB, rnB);
        vcp.normalMass = K > scalar(0) ? scalar(1) / K : scalar(0);
        Vec3 dv = vB + Cross(wB, rB) - vA - Cross(wA, rA);
        scalar vn = Dot(nA, dv);
        vcp.velocityBias = 0;
        if (vn < -kVelocityThreshold) vcp.velocityBias = -cc.restitution * vn;
    }
    Vec3 t1, t2;
    if (pointCount > 0) {
        center /= scalar(pointCount);
        normal.Normalize();
        t1 = Perp(normal);
        t2 = Cross(t1, normal);
        cc.normal = normal;
        cc.tangent1 = t1;
        cc.tangent2 = t2;
        Vec3 rA = center - xA;
        Vec3 rB = center - xB;
        cc.rA = rA;
        cc.rB = rB;
        Vec3 rn1A = Cross(rA, t1), rn1B = Cross(rB, t1);
        Vec3 rn2A = Cross(rA, t2), rn2B = Cross(rB, t2);
        scalar k11 = mA + mB + Dot(iA * rn1A, rn1A) + Dot(iB * rn1B, rn1B);
        scalar k12 = Dot(iA * rn1A, rn2A) + Dot(iB * rn1B, rn2B);
        scalar k22 = mA + mB + Dot(iA * rn2A, rn2A) + Dot(iB * rn2B, rn2B);
        Mat22 K(Vec2(k11, k12), Vec2(k12, k22));
        cc.tangentMass = Inverse(K);
        scalar mass = Dot(cc.normal, (iA + iB) * cc.normal);
        cc.motorMass = mass > scalar(0) ? scalar(1) / mass : scalar(0);
    }
}

// contact_solver.cpp:272-326
BX_HD void WarmStartContact(const ContactConstraint& cc, SolverBody* sb) {
    scalar mA = cc.invMassA, mB = cc.invMassB;
    Mat33 iA = cc.invIA, iB = cc.invIB;
    Vec3 vA = sb[cc.indexA].v, wA = sb[cc.indexA].w, vB = sb[cc.indexB].v, wB = sb[cc.indexB].w;
    for (u32 k = 0; k < cc.pointCount; ++k) {
        const VCPoint& vcp = cc.points[k];
        Vec3 P = vcp.normalImpulse * vcp.normal;
        vA -= mA * P;
        wA -= iA * Cross(vcp.rA, P);
        vB += mB * P;
        wB += iB * Cross(vcp.rB, P);
    }
    if (cc.pointCount > 0) {
        Vec3 P1 = cc.tangentImpulse.x * cc.tangent1;
        Vec3 P2 = cc.tangentImpulse.y * cc.tangent2;
        Vec3 P3 = cc.motorImpulse * cc.normal;
        vA -= mA * (P1 + P2);
        wA -= iA * (Cross(cc.rA, P1 + P2) + P3);
        vB += mB * (P1 + P2);
        wB += iB * (Cross(cc.rB, P1 + P2) + P3);
    }
    sb[cc.indexA].v = vA; sb[cc.indexA].w = wA;
    sb[cc.indexB].v = vB; sb[cc.indexB].w = wB;
}

// contact_solver.cpp:328-486: one contact of one PGS sweep.
BX_HD void SolveContactVelocity(ContactConstraint& cc, Vec3& vA, Vec3& wA, Vec3& vB, Vec3& wB) {
    scalar mA = cc.invMassA, mB = cc.invMassB;
    const Mat33 iA = cc.invIA, iB = cc.invIB;
    scalar normalImpulse = 0;
    for (u32 k = 0; k < cc.pointCount; ++k) {
        VCPoint& vcp = cc.points[k];
        Vec3 dv = vB + Cross(wB, vcp.rB) - vA - Cross(wA, vcp.rA);
        scalar Cdot = Dot(vcp.normal, dv);
        scalar impulse = -vcp.normalMass * (Cdot - vcp.velocityBias);
        scalar oldImpulse = vcp.normalImpulse;
        vcp.normalImpulse = Max(vcp.normalImpulse + impulse, scalar(0));
        impulse = vcp.normalImpulse - oldImpulse;
        Vec3 P = impulse * vcp.normal;
        vA -= mA * P;
        wA -= iA * Cross(vcp.rA, P);
        vB += mB * P;
        wB += iB * Cross(vcp.rB, P);
        normalImpulse += vcp.normalImpulse;
    }
    if (cc.pointCount > 0) {
        {
            Vec3 dv = vB + Cross(wB, cc.rB) - vA - Cross(wA, cc.rA);
            Vec2 Cdot;
            Cdot.x = Dot(dv, cc.tangent1) - cc.tangentSpeed1;
            Cdot.y = Dot(dv, cc.tangent2) - cc.tangentSpeed2;
            Vec2 impulse = cc.tangentMass * -Cdot;
            Vec2 oldImpulse = cc.tangentImpulse;
            cc.tangentImpulse += impulse;
            scalar maxImpulse = cc.friction * normalImpulse;
            if (Dot(cc.tangentImpulse, cc.tangentImpulse) > maxImpulse * maxImpulse) {
                cc.tangentImpulse.Normalize();
                cc.tangentImpulse *= maxImpulse;
            }
            impulse = cc.tangentImpulse - oldImpulse;
            Vec3 P1 = impulse.x * cc.tangent1;
            Vec3 P2 = impulse.y * cc.tangent2;
            Vec3 P = P1 + P2;
            vA -= mA * P;
            wA -= iA * Cross(cc.rA, P);
            vB += mB * P;
            wB += iB * Cross(cc.rB, P);
        }
        {
            scalar Cdot = Dot(cc.normal, wB - wA) - cc.motorSpeed;
            scalar impulse = -cc.motorMass * Cdot;
            scalar oldImpulse = cc.motorImpulse;
            scalar maxImpulse = cc.friction * normalImpulse;
            cc.motorImpulse = Clamp(cc.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = cc.motorImpulse - oldImpulse;
            Vec3 P = impulse * cc.normal;
            wA -= iA * P;
            wB += iB * P;
        }
    }
}

// contact_solver.cpp:517-624: one contact of one NGS sweep; returns its minimum separation (<= 0 start).
BX_HD scalar SolveContactPosition(const ContactConstraint& cc, const Manifold& m, Vec3& cA, Quat& qA, Mat33& iA,
                                  Vec3& cB, Quat& qB, Mat33& iB, scalar minSeparation) {
    scalar mA = cc.invMassA, mB = cc.invMassB;
    for (u32 k = 0; k < cc.pointCount; ++k) {
        const ManifoldPoint& p = m.points[k];
        Transform xfA(cA - Mul(qA, cc.localCenterA), qA);
        Transform xfB(cB - Mul(qB, cc.localCenterB), qB);
        Vec3 normal = Mul(xfA.rotation, p.localNormal1);
        Vec3 c1 = Mul(xfA, p.localPoint1);
        Vec3 c2 = Mul(xfB, p.localPoint2);
        Vec3 point = c2;
        scalar separation = Dot(c2 - c1, normal) - cc.radiusA - cc.radiusB;
        minSeparation = Min(minSeparation, separation);
        scalar C = Clamp(kBaumgarte * (separation + kLinearSlop), -kMaxLinearCorrection, scalar(0));
        Vec3 rA = point - cA;
        Vec3 rB = point - cB;
        Vec3 rnA = Cross(rA, normal);
        Vec3 rnB = Cross(rB, normal);
        scalar K = mA + mB + Dot(rnA, iA * rnA) + Dot(rnB, iB * rnB);
        scalar impulse = K > scalar(0) ? -C / K : scalar(0);
        Vec3 P = impulse * normal;
        cA -= mA * P;
        qA -= Derivative(qA, iA * Cross(rA, P));
        qA.Normalize();
        iA = RotateToFrame(cc.localInvIA, qA);
        cB += mB * P;
        qB += Derivative(qB, iB * Cross(rB, P));
        qB.Normalize();
        iB = RotateToFrame(cc.localInvIB, qB);
    }
    return minSeparation;
}

}  // namespace bx
