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

// island.cpp:86-135: velocity integration of one body.
BX_HD void IntegrateVelocity(const Body& b, const Vec3& gravity, scalar h, VelBody& vb, PosBody& pb) {
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
    vb.v = v;
    vb.w = w;
    pb.x = b.x;
    pb.q = b.q;
    pb.I = b.worldInvI;
}

// island.cpp:190-229: position integration with translation / rotation clamps.
BX_HD void IntegratePosition(VelBody& vb, PosBody& pb, const Body& b, scalar h) {
    if (b.type == kStatic) return;
    Vec3 translation = h * vb.v;
    if (Dot(translation, translation) > kMaxTranslationSq) {
        scalar ratio = kMaxTranslation / Length(translation);
        vb.v *= ratio;
    }
    Vec3 rotation = h * vb.w;
    if (Dot(rotation, rotation) > kMaxRotationSq) {
        scalar ratio = kMaxRotation / Length(rotation);
        vb.w *= ratio;
    }
    pb.x += h * vb.v;
    pb.q = Integrate(pb.q, vb.w, h);
    pb.I = RotateToFrame(b.invI, pb.q);
}

// contact_solver.cpp:28-150 (constructor) and 152-270 (InitializeVelocityConstraints) for one
// manifold of a contact, reading the island state right after velocity integration. A convex
// contact has one manifold; the manifolds of a mesh contact become consecutive constraints on the
// same two bodies, which the sweeps then visit in b3ContactSolver's manifold order.
BX_HD void InitContact(VelCon& vc, PosCon& pc, const Manifold& m, u32 manifoldCount, const Fixture& fA,
                       const Fixture& fB, const Body& bA, const Body& bB, u32 indexA, u32 indexB, const VelBody* vel,
                       const PosBody* pos, scalar dtRatio, bool warmStarting) {
    const u32 pointCount = manifoldCount > 0 ? m.pointCount : 0;
    pc.indexA = indexA;
    pc.indexB = indexB;
    pc.pointCount = pointCount;
    pc.invMassA = bA.invMass;
    pc.invMassB = bB.invMass;
    pc.radiusA = fA.radius;
    pc.radiusB = fB.radius;
    pc.localCenterA = bA.localCenter;
    pc.localCenterB = bB.localCenter;
    pc.localInvIA = bA.invI;
    pc.localInvIB = bB.invI;
#pragma unroll
    for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
        if (k >= pointCount) break;
        pc.localNormalA[k] = m.points[k].localNormal1;
        pc.localPointA[k] = m.points[k].localPoint1;
        pc.localPointB[k] = m.points[k].localPoint2;
    }
    vc.indexA = indexA;
    vc.indexB = indexB;
    vc.pointCount = pointCount;
    vc.invMassA = bA.invMass;
    vc.invMassB = bB.invMass;
    vc.invIA = pos[indexA].I;
    vc.invIB = pos[indexB].I;
    vc.friction = Sqrt(fA.friction * fB.friction);         // b3MixFriction
    vc.restitution = Max(fA.restitution, fB.restitution);  // b3MixRestitution
    vc.motorSpeed = m.motorSpeed;
    vc.tangentSpeed1 = m.tangentSpeed1;
    vc.tangentSpeed2 = m.tangentSpeed2;
    if (warmStarting) {
        vc.tangentImpulse = dtRatio * m.tangentImpulse;
        vc.motorImpulse = dtRatio * m.motorImpulse;
    } else {
        vc.tangentImpulse.SetZero();
        vc.motorImpulse = 0;
    }
    vc.tangentMass.SetZero();
    vc.motorMass = 0;
    vc.rA.SetZero(); vc.rB.SetZero(); vc.normal.SetZero(); vc.tangent1.SetZero(); vc.tangent2.SetZero();
    if (pointCount == 0) return;
    const scalar mA = vc.invMassA, mB = vc.invMassB;
    const Mat33 iA = vc.invIA, iB = vc.invIB;
    const Vec3 xA = pos[indexA].x, xB = pos[indexB].x;
    const Quat qA = pos[indexA].q, qB = pos[indexB].q;
    const Vec3 vA = vel[indexA].v, wA = vel[indexA].w, vB = vel[indexB].v, wB = vel[indexB].w;
    Transform xfA(xA - Mul(qA, pc.localCenterA), qA);
    Transform xfB(xB - Mul(qB, pc.localCenterB), qB);
    // b3WorldManifold::Initialize (collision.cpp:76-101) fused with the per-point constraint setup
    Vec3 center, normal;
    center.SetZero();
    normal.SetZero();
#pragma unroll
    for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
        if (k >= pointCount) break;
        Vec3 nA = Mul(xfA.rotation, pc.localNormalA[k]);
        Vec3 cA = Mul(xfA, pc.localPointA[k]);
        Vec3 cB = Mul(xfB, pc.localPointB[k]);
        Vec3 pA = cA + pc.radiusA * nA;
        Vec3 pB = cB - pc.radiusB * nA;
        Vec3 point = scalar(0.5) * (pA + pB);
        center += point;
        normal += nA;
        VCPoint& vcp = vc.points[k];
        Vec3 rA = point - xA;
        Vec3 rB = point - xB;
        vcp.rA = rA;
        vcp.rB = rB;
        vcp.normal = nA;
        Vec3 rnA = Cross(rA, nA);
        Vec3 rnB = Cross(rB, nA);
        scalar K = mA + mB + Dot(iA * rnA, rnA) + Dot(iB * rnB, rnB);
        vcp.normalMass = K > scalar(0) ? scalar(1) / K : scalar(0);
        vcp.normalImpulse = warmStarting ? dtRatio * m.points[k].normalImpulse : scalar(0);
        Vec3 dv = vB + Cross(wB, rB) - vA - Cross(wA, rA);
        scalar vn = Dot(nA, dv);
        vcp.velocityBias = 0;
        if (vn < -kVelocityThreshold) vcp.velocityBias = -vc.restitution * vn;
    }
    center /= scalar(pointCount);
    normal.Normalize();
    Vec3 t1 = Perp(normal);
    Vec3 t2 = Cross(t1, normal);
    vc.normal = normal;
    vc.tangent1 = t1;
    vc.tangent2 = t2;
    Vec3 rA = center - xA;
    Vec3 rB = center - xB;
    vc.rA = rA;
    vc.rB = rB;
    Vec3 rn1A = Cross(rA, t1), rn1B = Cross(rB, t1);
    Vec3 rn2A = Cross(rA, t2), rn2B = Cross(rB, t2);
    scalar k11 = mA + mB + Dot(iA * rn1A, rn1A) + Dot(iB * rn1B, rn1B);
    scalar k12 = Dot(iA * rn1A, rn2A) + Dot(iB * rn1B, rn2B);
    scalar k22 = mA + mB + Dot(iA * rn2A, rn2A) + Dot(iB * rn2B, rn2B);
    vc.tangentMass = Inverse(Mat22(Vec2(k11, k12), Vec2(k12, k22)));
    scalar mass = Dot(vc.normal, (iA + iB) * vc.normal);
    vc.motorMass = mass > scalar(0) ? scalar(1) / mass : scalar(0);
}

// contact_solver.cpp:272-326
BX_HD void WarmStartContact(const VelCon& vc, Vec3& vA, Vec3& wA, Vec3& vB, Vec3& wB) {
    const scalar mA = vc.invMassA, mB = vc.invMassB;
    const Mat33 iA = vc.invIA, iB = vc.invIB;
#pragma unroll
    for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
        if (k >= vc.pointCount) break;
        const VCPoint& vcp = vc.points[k];
        Vec3 P = vcp.normalImpulse * vcp.normal;
        vA -= mA * P;
        wA -= iA * Cross(vcp.rA, P);
        vB += mB * P;
        wB += iB * Cross(vcp.rB, P);
    }
    if (vc.pointCount > 0) {
        Vec3 P1 = vc.tangentImpulse.x * vc.tangent1;
        Vec3 P2 = vc.tangentImpulse.y * vc.tangent2;
        Vec3 P3 = vc.motorImpulse * vc.normal;
        vA -= mA * (P1 + P2);
        wA -= iA * (Cross(vc.rA, P1 + P2) + P3);
        vB += mB * (P1 + P2);
        wB += iB * (Cross(vc.rB, P1 + P2) + P3);
    }
}

// contact_solver.cpp:328-486: one contact of one PGS sweep.
BX_HD void SolveContactVelocity(VelCon& vc, Vec3& vA, Vec3& wA, Vec3& vB, Vec3& wB) {
    const scalar mA = vc.invMassA, mB = vc.invMassB;
    const Mat33 iA = vc.invIA, iB = vc.invIB;
    scalar normalImpulse = 0;
#pragma unroll
    for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
        if (k >= vc.pointCount) break;
        VCPoint& vcp = vc.points[k];
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
    if (vc.pointCount > 0) {
        {
            Vec3 dv = vB + Cross(wB, vc.rB) - vA - Cross(wA, vc.rA);
            Vec2 Cdot;
            Cdot.x = Dot(dv, vc.tangent1) - vc.tangentSpeed1;
            Cdot.y = Dot(dv, vc.tangent2) - vc.tangentSpeed2;
            Vec2 impulse = vc.tangentMass * -Cdot;
            Vec2 oldImpulse = vc.tangentImpulse;
            vc.tangentImpulse += impulse;
            scalar maxImpulse = vc.friction * normalImpulse;
            if (Dot(vc.tangentImpulse, vc.tangentImpulse) > maxImpulse * maxImpulse) {
                vc.tangentImpulse.Normalize();
                vc.tangentImpulse *= maxImpulse;
            }
            impulse = vc.tangentImpulse - oldImpulse;
            Vec3 P1 = impulse.x * vc.tangent1;
            Vec3 P2 = impulse.y * vc.tangent2;
            Vec3 P = P1 + P2;
            vA -= mA * P;
            wA -= iA * Cross(vc.rA, P);
            vB += mB * P;
            wB += iB * Cross(vc.rB, P);
        }
        {
            scalar Cdot = Dot(vc.normal, wB - wA) - vc.motorSpeed;
            scalar impulse = -vc.motorMass * Cdot;
            scalar oldImpulse = vc.motorImpulse;
            scalar maxImpulse = vc.friction * normalImpulse;
            vc.motorImpulse = Clamp(vc.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = vc.motorImpulse - oldImpulse;
            Vec3 P = impulse * vc.normal;
            wA -= iA * P;
            wB += iB * P;
        }
    }
}

// contact_solver.cpp:517-624: one contact of one NGS sweep; returns the running minimum separation.
BX_HD scalar SolveContactPosition(const PosCon& pc, Vec3& cA, Quat& qA, Mat33& iA, Vec3& cB, Quat& qB, Mat33& iB,
                                  scalar minSeparation) {
    const scalar mA = pc.invMassA, mB = pc.invMassB;
#pragma unroll
    for (u32 k = 0; k < kMaxManifoldPoints; ++k) {
        if (k >= pc.pointCount) break;
        Transform xfA(cA - Mul(qA, pc.localCenterA), qA);
        Transform xfB(cB - Mul(qB, pc.localCenterB), qB);
        Vec3 normal = Mul(xfA.rotation, pc.localNormalA[k]);
        Vec3 c1 = Mul(xfA, pc.localPointA[k]);
        Vec3 c2 = Mul(xfB, pc.localPointB[k]);
        Vec3 point = c2;
        scalar separation = Dot(c2 - c1, normal) - pc.radiusA - pc.radiusB;
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
        iA = RotateToFrame(pc.localInvIA, qA);
        cB += mB * P;
        qB += Derivative(qB, iB * Cross(rB, P));
        qB.Normalize();
        iB = RotateToFrame(pc.localInvIB, qB);
    }
    return minSeparation;
}

}  // namespace bx
