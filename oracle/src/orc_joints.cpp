// orc_joints.cpp -- oracle (test infrastructure): the revolute joint and the joint solver, restated.
//   joints/revolute_joint.cpp:22-48     b3RevoluteJointDef::Initialize
//   joints/revolute_joint.cpp:50-76     b3RevoluteJoint constructor (warm-start state)
//   joints/revolute_joint.cpp:78-231    InitializeVelocityConstraints
//   joints/revolute_joint.cpp:233-277   WarmStart
//   joints/revolute_joint.cpp:279-367   SolveVelocityConstraints
//   joints/revolute_joint.cpp:369-531   SolvePositionConstraints
//   joints/joint_solver.cpp:31-64       the joint solver loops (island order)
//   joint_manager.cpp:30-72             joint edges prepended to both bodies' joint lists
//   joints/cone_joint.cpp, prismatic_joint.cpp, wheel_joint.cpp: Initialize, constructors and the
//   solver functions of the cone, prismatic and wheel joints (the axis joints below)
#include <cmath>
#include "orc_world.h"
#include "../../include/bxstep.h"

namespace orc {

static void ComputeBasis(const Vec3& a, Vec3& b, Vec3& c) {  // mat33.h:310-328
    static const scalar sqrt_inv3 = Sqrt(scalar(1) / scalar(3));
    if (Abs(a.x) >= sqrt_inv3) b.Set(a.y, -a.x, scalar(0));
    else b.Set(scalar(0), a.z, -a.y);
    b.Normalize();
    c = Cross(a, b);
}

static Quat Mat33Quat(const Mat33& m) {  // quat.h:283-345
    scalar trace = m[0][0] + m[1][1] + m[2][2];
    if (trace > scalar(0)) {
        Quat r;
        scalar s = Sqrt(trace + scalar(1));
        r.s = scalar(0.5) * s;
        scalar t = scalar(0.5) / s;
        r.v.x = t * (m[1][2] - m[2][1]);
        r.v.y = t * (m[2][0] - m[0][2]);
        r.v.z = t * (m[0][1] - m[1][0]);
        return r;
    }
    const u32 next[3] = {1, 2, 0};
    u32 i = 0;
    if (m[1][1] > m[0][0]) i = 1;
    if (m[2][2] > m[i][i]) i = 2;
    u32 j = next[i];
    u32 k = next[j];
    scalar s = Sqrt((m[i][i] - (m[j][j] + m[k][k])) + scalar(1));
    scalar q[4];
    q[i] = s * scalar(0.5);
    scalar t = s != scalar(0) ? scalar(0.5) / s : s;
    q[3] = t * (m[j][k] - m[k][j]);
    q[j] = t * (m[i][j] + m[j][i]);
    q[k] = t * (m[i][k] + m[k][i]);
    Quat r;
    r.v.x = q[0];
    r.v.y = q[1];
    r.v.z = q[2];
    r.s = q[3];
    return r;
}

static Vec3 V3(const float* p) { return Vec3(p[0], p[1], p[2]); }
static Quat Q4(const float* p) { return Quat(Vec3(p[0], p[1], p[2]), p[3]); }

void CreateJoint(World& w, const bx_joint_desc& d) {
    Joint j;
    j.type = d.type;
    j.bodyA = d.body_a;
    j.bodyB = d.body_b;
    j.collideConnected = d.collide_connected != 0;
    j.frequencyHz = d.frequency_hz;
    j.dampingRatio = d.damping_ratio;
    j.length = d.length;
    j.maxForce = d.max_force;
    j.maxTorque = d.max_torque;
    j.correctionFactor = d.correction_factor;
    j.target = V3(d.target);
    j.linearOffset = V3(d.linear_offset);
    j.angularOffset = Q4(d.angular_offset);
    j.angularImpulse3.SetZero();
    j.impulse1 = 0;
    j.linearImpulse2.SetZero();
    j.springImpulse = 0;
    for (Vec3* v : {&j.localXAxisA, &j.localYAxisA, &j.localZAxisA, &j.localXAxisB, &j.motorAxis, &j.a1, &j.a2,
                    &j.perp1, &j.perp2, &j.s1, &j.s2, &j.s3, &j.s4, &j.axisB, &j.u})
        v->SetZero();
    const Body& A0 = w.bodies[j.bodyA];
    const Body& B0 = w.bodies[j.bodyB];
    if (d.type == BX_JOINT_PRISMATIC || d.type == BX_JOINT_WHEEL) {
        j.referenceRotation = Q4(d.reference_rotation);
        j.localRotationA = j.localRotationB = Quat(Vec3(0, 0, 0), scalar(1));
        Vec3 axisA = V3(d.local_axis_a), axisB = V3(d.local_axis_b);
        if (d.use_initialize) {  // prismatic_joint.cpp:23-35, wheel_joint.cpp:22-31
            Vec3 anchor = V3(d.anchor);
            j.localAnchorA = MulT(A0.xf, anchor);
            j.localAnchorB = MulT(B0.xf, anchor);
            axisA = MulC(A0.xf.rotation, V3(d.axis));
            if (d.type == BX_JOINT_PRISMATIC) j.referenceRotation = Conjugate(A0.sweep.orientation) * B0.sweep.orientation;
            else axisB = MulC(B0.xf.rotation, V3(d.axis_b));
        } else {
            j.localAnchorA = V3(d.local_anchor_a);
            j.localAnchorB = V3(d.local_anchor_b);
        }
        j.localXAxisA = axisA;  // the constructors (prismatic_joint.cpp:37-61, wheel_joint.cpp:33-66)
        j.localXAxisA.Normalize();
        ComputeBasis(j.localXAxisA, j.localYAxisA, j.localZAxisA);
        if (d.type == BX_JOINT_WHEEL) {
            j.localXAxisB = axisB;
            j.localXAxisB.Normalize();
            Vec3 wA = Mul(A0.xf.rotation, j.localXAxisA);
            Vec3 wB = Mul(B0.xf.rotation, j.localXAxisB);
            Vec3 u = Cross(wA, wB);
            scalar x = Dot(wA, wB);
            scalar y = Length(u);
            j.referenceCosine = x;
            j.referenceAngle = scalar(::atan2(double(y), double(x)));  // the reference calls ::atan2(double, double)
        }
        j.lowerTranslation = d.lower_translation;
        j.upperTranslation = d.upper_translation;
    } else if (d.type == BX_JOINT_MOUSE) {  // mouse_joint.cpp:22-31: the target is B's anchor at creation
        j.localAnchorB = MulT(B0.xf, j.target);
    } else if (d.type != BX_JOINT_REVOLUTE && d.type != BX_JOINT_CONE) {
        if (d.use_initialize) {
            const Vec3 anchor = V3(d.anchor);
            if (d.type == BX_JOINT_MOTOR) {  // motor_joint.cpp:23-31
                j.linearOffset = MulT(A0.xf, B0.xf.translation);
                j.angularOffset = MulC(A0.sweep.orientation, B0.xf.rotation);
            } else {  // sphere, weld, spring, friction Initialize
                j.localAnchorA = MulT(A0.xf, anchor);
                j.localAnchorB = MulT(B0.xf, d.type == BX_JOINT_SPRING ? V3(d.anchor_b) : anchor);
                if (d.type == BX_JOINT_WELD) j.referenceRotation = Conjugate(A0.xf.rotation) * B0.xf.rotation;
                if (d.type == BX_JOINT_SPRING) j.length = Distance(anchor, V3(d.anchor_b));
            }
            if (d.type != BX_JOINT_WELD) j.referenceRotation = Q4(d.reference_rotation);
        } else {
            j.localAnchorA = V3(d.local_anchor_a);
            j.localAnchorB = V3(d.local_anchor_b);
            j.referenceRotation = Q4(d.reference_rotation);
        }
        j.localRotationA = Q4(d.local_rotation_a);
        j.localRotationB = Q4(d.local_rotation_b);
    } else if (d.use_initialize) {  // revolute_joint.cpp:22-48; cone_joint.cpp:35-55 builds the frame about x
        Mat33 rotation;
        if (d.type == BX_JOINT_CONE) {
            rotation.x = V3(d.axis);
            ComputeBasis(rotation.x, rotation.y, rotation.z);
        } else {
            rotation.z = V3(d.axis);
            ComputeBasis(rotation.z, rotation.x, rotation.y);
        }
        Quat q = Mat33Quat(rotation);
        const Body& A = w.bodies[j.bodyA];
        const Body& B = w.bodies[j.bodyB];
        Vec3 anchor = V3(d.anchor);
        j.localAnchorA = MulT(A.xf, anchor);
        j.localRotationA = MulC(A.sweep.orientation, q);
        j.localAnchorB = MulT(B.xf, anchor);
        j.localRotationB = MulC(B.sweep.orientation, q);
        j.referenceRotation = Quat(Vec3(0, 0, 0), scalar(1));
    } else {
        j.localAnchorA = V3(d.local_anchor_a);
        j.localRotationA = Q4(d.local_rotation_a);
        j.localAnchorB = V3(d.local_anchor_b);
        j.localRotationB = Q4(d.local_rotation_b);
        j.referenceRotation = Q4(d.reference_rotation);
    }
    j.coneAngle = d.cone_angle;
    j.lowerAngle = d.lower_angle;
    j.upperAngle = d.upper_angle;
    j.enableLimit = d.enable_limit != 0;
    j.enableMotor = d.enable_motor != 0;
    j.motorSpeed = d.motor_speed;
    j.maxMotorTorque = d.max_motor_torque;
    j.motorImpulse = 0;
    j.limitState = e_inactiveLimit;
    j.limitImpulse = 0;
    j.linearImpulse.SetZero();
    j.angularImpulse.SetZero();
    u32 index = u32(w.joints.size());
    w.joints.push_back(j);
    Body& A = w.bodies[j.bodyA];
    Body& B = w.bodies[j.bodyB];
    A.joints.insert(A.joints.begin(), JointEdge{index, j.bodyB});
    B.joints.insert(B.joints.begin(), JointEdge{index, j.bodyA});
}

static void InitRevolute(World& w, const TimeStep& step, Joint& j, Position* pos) {  // revolute_joint.cpp:78-231
    const Body& A = w.bodies[j.bodyA];
    const Body& B = w.bodies[j.bodyB];
    j.indexA = A.islandIndex;
    j.indexB = B.islandIndex;
    j.mA = A.invMass;
    j.mB = B.invMass;
    j.localCenterA = A.sweep.localCenter;
    j.localCenterB = B.sweep.localCenter;
    j.localInvIA = A.invI;
    j.localInvIB = B.invI;
    j.iA = pos[j.indexA].I;
    j.iB = pos[j.indexB].I;
    Quat qA = pos[j.indexA].q;
    Quat qB = pos[j.indexB].q;
    scalar mA = j.mA;
    Mat33 iA = j.iA;
    scalar mB = j.mB;
    Mat33 iB = j.iB;
    Quat fA = qA * j.localRotationA;
    Quat fB = qB * j.localRotationB;
    Mat33 RfA = fA.GetRotationMatrix();
    Mat33 RfB = fB.GetRotationMatrix();
    Vec3 u1 = RfA.x, v1 = RfA.y, w1 = RfA.z;
    Vec3 u2 = RfB.x, v2 = RfB.y;
    if (j.enableMotor || j.enableLimit) {
        j.motorAxis = w1;
        j.motorAxis.Normalize();
        scalar K = Dot((iA + iB) * j.motorAxis, j.motorAxis);
        j.motorMass = K > scalar(0) ? scalar(1) / K : scalar(0);
    }
    if (!j.enableMotor) j.motorImpulse = scalar(0);
    if (j.enableLimit) {
        scalar c = Dot(u2, u1);
        scalar s = Dot(u2, v1);
        scalar angle = scalar(::atan2(double(s), double(c)));  // the reference calls ::atan2(double, double)
        if (Abs(j.upperAngle - j.lowerAngle) < scalar(2) * kAngularSlop) {
            if (j.limitState != e_equalLimits) { j.limitState = e_equalLimits; j.limitImpulse = scalar(0); }
        } else if (angle <= j.lowerAngle) {
            if (j.limitState != e_atLowerLimit) { j.limitState = e_atLowerLimit; j.limitImpulse = scalar(0); }
        } else if (angle >= j.upperAngle) {
            if (j.limitState != e_atUpperLimit) { j.limitState = e_atUpperLimit; j.limitImpulse = scalar(0); }
        } else {
            j.limitState = e_inactiveLimit;
            j.limitImpulse = scalar(0);
        }
    } else {
        j.limitState = e_inactiveLimit;
        j.limitImpulse = scalar(0);
    }
    {
        j.rA = Mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = Mul(qB, j.localAnchorB - j.localCenterB);
        Mat33 RA = Skew(j.rA);
        Mat33 RAT = Transpose(RA);
        Mat33 RB = Skew(j.rB);
        Mat33 RBT = Transpose(RB);
        Mat33 M = Diagonal(mA + mB);
        j.linearMass = M + RA * iA * RAT + RB * iB * RBT;
    }
    {
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
        j.a1 = a1;
        j.a2 = a2;
        j.angularMass = K;
    }
    if (step.warmStarting) {
        j.motorImpulse *= step.dtRatio;
        j.limitImpulse *= step.dtRatio;
        j.linearImpulse *= step.dtRatio;
        j.angularImpulse *= step.dtRatio;
    } else {
        j.motorImpulse = scalar(0);
        j.limitImpulse = scalar(0);
        j.linearImpulse.SetZero();
        j.angularImpulse.SetZero();
    }
}

static void WarmStartRevolute(Joint& j, Velocity* vel) {  // revolute_joint.cpp:233-277
    Vec3 vA = vel[j.indexA].v, wA = vel[j.indexA].w;
    Vec3 vB = vel[j.indexB].v, wB = vel[j.indexB].w;
    if (j.enableMotor && j.limitState != e_equalLimits) {
        Vec3 P = j.motorAxis * j.motorImpulse;
        wA -= j.iA * P;
        wB += j.iB * P;
    }
    if (j.enableLimit && j.limitState != e_inactiveLimit) {
        Vec3 P = j.motorAxis * j.limitImpulse;
        wA -= j.iA * P;
        wB += j.iB * P;
    }
    {
        vA -= j.mA * j.linearImpulse;
        wA -= j.iA * Cross(j.rA, j.linearImpulse);
        vB += j.mB * j.linearImpulse;
        wB += j.iB * Cross(j.rB, j.linearImpulse);
    }
    {
        Vec3 L = j.angularImpulse.x * j.a1 + j.angularImpulse.y * j.a2;
        wA -= j.iA * L;
        wB += j.iB * L;
    }
    vel[j.indexA].v = vA;
    vel[j.indexA].w = wA;
    vel[j.indexB].v = vB;
    vel[j.indexB].w = wB;
}

static void SolveRevoluteVelocity(const TimeStep& step, Joint& j, Velocity* vel) {  // revolute_joint.cpp:279-367
    Vec3 vA = vel[j.indexA].v, wA = vel[j.indexA].w;
    Vec3 vB = vel[j.indexB].v, wB = vel[j.indexB].w;
    Mat33 iA = j.iA, iB = j.iB;
    if (j.enableMotor && j.limitState != e_equalLimits) {
        scalar dw = Dot(wB - wA, j.motorAxis);
        scalar Cdot = dw - j.motorSpeed;
        scalar impulse = -j.motorMass * Cdot;
        scalar oldImpulse = j.motorImpulse;
        scalar maxImpulse = step.dt * j.maxMotorTorque;
        j.motorImpulse = Clamp(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
        impulse = j.motorImpulse - oldImpulse;
        Vec3 P = j.motorAxis * impulse;
        wA -= iA * P;
        wB += iB * P;
    }
    if (j.enableLimit && j.limitState != e_inactiveLimit) {
        scalar Cdot = Dot(wB - wA, j.motorAxis);
        scalar impulse = -j.motorMass * Cdot;
        if (j.limitState == e_equalLimits) {
            j.limitImpulse += impulse;
        } else if (j.limitState == e_atLowerLimit) {
            scalar oldImpulse = j.limitImpulse;
            j.limitImpulse = Max(j.limitImpulse + impulse, scalar(0));
            impulse = j.limitImpulse - oldImpulse;
        } else if (j.limitState == e_atUpperLimit) {
            scalar oldImpulse = j.limitImpulse;
            j.limitImpulse = Min(j.limitImpulse + impulse, scalar(0));
            impulse = j.limitImpulse - oldImpulse;
        }
        Vec3 P = impulse * j.motorAxis;
        wA -= iA * P;
        wB += iB * P;
    }
    {
        Vec3 Cdot = vB + Cross(wB, j.rB) - vA - Cross(wA, j.rA);
        Vec3 impulse = j.linearMass.Solve(-Cdot);
        j.linearImpulse += impulse;
        vA -= j.mA * impulse;
        wA -= j.iA * Cross(j.rA, impulse);
        vB += j.mB * impulse;
        wB += j.iB * Cross(j.rB, impulse);
    }
    {
        Vec3 dw = wB - wA;
        Vec2 Cdot;
        Cdot.x = Dot(dw, j.a1);
        Cdot.y = Dot(dw, j.a2);
        Vec2 impulse = Solve(j.angularMass, -Cdot);
        j.angularImpulse += impulse;
        Vec3 L = impulse.x * j.a1 + impulse.y * j.a2;
        wA -= j.iA * L;
        wB += j.iB * L;
    }
    vel[j.indexA].v = vA;
    vel[j.indexA].w = wA;
    vel[j.indexB].v = vB;
    vel[j.indexB].w = wB;
}

static bool SolveRevolutePosition(Joint& j, Position* pos) {  // revolute_joint.cpp:369-531
    Vec3 xA = pos[j.indexA].x;
    Quat qA = pos[j.indexA].q;
    Vec3 xB = pos[j.indexB].x;
    Quat qB = pos[j.indexB].q;
    Mat33 iA = pos[j.indexA].I;
    Mat33 iB = pos[j.indexB].I;
    scalar mA = j.mA, mB = j.mB;
    scalar linearError = scalar(0);
    {
        Vec3 rA = Mul(qA, j.localAnchorA - j.localCenterA);
        Vec3 rB = Mul(qB, j.localAnchorB - j.localCenterB);
        Vec3 C = xB + rB - xA - rA;
        linearError += Length(C);
        Mat33 M = Diagonal(mA + mB);
        Mat33 RA = Skew(rA);
        Mat33 RAT = Transpose(RA);
        Mat33 RB = Skew(rB);
        Mat33 RBT = Transpose(RB);
        Mat33 mass = M + RA * iA * RAT + RB * iB * RBT;
        Vec3 impulse = mass.Solve(-C);
        xA -= mA * impulse;
        qA -= Derivative(qA, iA * Cross(rA, impulse));
        qA.Normalize();
        iA = RotateToFrame(j.localInvIA, qA);
        xB += mB * impulse;
        qB += Derivative(qB, iB * Cross(rB, impulse));
        qB.Normalize();
        iB = RotateToFrame(j.localInvIB, qB);
    }
    scalar limitError = scalar(0);
    if (j.enableLimit) {
        Quat fA = qA * j.localRotationA;
        Quat fB = qB * j.localRotationB;
        Mat33 RfA = fA.GetRotationMatrix();
        Mat33 RfB = fB.GetRotationMatrix();
        Vec3 u1 = RfA.x, v1 = RfA.y, w1 = RfA.z;
        Vec3 u2 = RfB.x;
        Vec3 limitAxis = w1;
        scalar K = Dot((iA + iB) * limitAxis, limitAxis);
        scalar limitMass = K > scalar(0) ? scalar(1) / K : scalar(0);
        scalar c = Dot(u2, u1);
        scalar s = Dot(u2, v1);
        scalar angle = scalar(::atan2(double(s), double(c)));  // the reference calls ::atan2(double, double)
        scalar C = scalar(0);
        if (Abs(j.upperAngle - j.lowerAngle) < scalar(2) * kAngularSlop)
            C = Clamp(angle - j.lowerAngle, -kMaxAngularCorrection, kMaxAngularCorrection);
        else if (angle <= j.lowerAngle)
            C = Clamp(angle - j.lowerAngle + kAngularSlop, -kMaxAngularCorrection, scalar(0));
        else if (angle >= j.upperAngle)
            C = Clamp(angle - j.upperAngle - kAngularSlop, scalar(0), kMaxAngularCorrection);
        limitError = Abs(C);
        scalar limitImpulse = -C * limitMass;
        Vec3 P = limitImpulse * limitAxis;
        qA -= Derivative(qA, iA * P);
        qA.Normalize();
        iA = RotateToFrame(j.localInvIA, qA);
        qB += Derivative(qB, iB * P);
        qB.Normalize();
        iB = RotateToFrame(j.localInvIB, qB);
    }
    scalar angularError = scalar(0);
    const scalar kCosSlop = scalar(0.002);
    {
        Quat fA = qA * j.localRotationA;
        Quat fB = qB * j.localRotationB;
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
        Vec2 impulse = Solve(K, -C);
        Vec3 P = impulse.x * a1 + impulse.y * a2;
        qA -= Derivative(qA, iA * P);
        qA.Normalize();
        iA = RotateToFrame(j.localInvIA, qA);
        qB += Derivative(qB, iB * P);
        qB.Normalize();
        iB = RotateToFrame(j.localInvIB, qB);
        angularError = Abs(C.x) + Abs(C.y);
    }
    pos[j.indexA].x = xA;
    pos[j.indexA].q = qA;
    pos[j.indexB].x = xB;
    pos[j.indexB].q = qB;
    pos[j.indexA].I = iA;
    pos[j.indexB].I = iB;
    return linearError <= kLinearSlop && limitError <= kAngularSlop && angularError <= kCosSlop;
}

// ------------------------------------------------------------------ the other joint types
// b3ComputeSoftConstraintCoefficients (weld_joint.cpp:64-81, mouse_joint.cpp:35-52): the soft
// constraint's gamma and bias of a mass-spring-damper at `frequencyHz`.
static void SoftCoefficients(scalar& gamma, scalar& bias, scalar frequencyHz, scalar dampingRatio, scalar m,
                             scalar C, scalar h) {
    scalar omega = scalar(2) * kPi * frequencyHz;
    scalar k = omega * omega * m;
    scalar d = scalar(2) * dampingRatio * omega * m;
    gamma = h * (d + h * k);
    gamma = gamma != scalar(0) ? scalar(1) / gamma : scalar(0);
    bias = gamma * h * k * C;
}

// The point-to-point effective mass shared by the sphere, weld and friction joints and the position
// solvers: M + [rA]x iA [rA]x^T + [rB]x iB [rB]x^T.
static Mat33 PointMass(scalar mA, scalar mB, const Mat33& iA, const Mat33& iB, const Vec3& rA, const Vec3& rB) {
    Mat33 RA = Skew(rA);
    Mat33 RAT = Transpose(RA);
    Mat33 RB = Skew(rB);
    Mat33 RBT = Transpose(RB);
    Mat33 M = Diagonal(mA + mB);
    return M + RA * iA * RAT + RB * iB * RBT;
}

// The relative rotation error of weld_joint.cpp:132-148 / motor_joint.cpp:86-99: 2 (q2 - q1) q1*, with
// q1 = qA* qB flipped to q2's hemisphere.
static Quat RotationError(const Quat& qA, const Quat& qB, const Quat& q2) {
    Quat q1 = Conjugate(qA) * qB;
    if (Dot(q1, q2) < scalar(0)) q1 = -q1;
    return scalar(2) * (q2 - q1) * Conjugate(q1);
}

// InitializeVelocityConstraints of sphere_joint.cpp:37-79, weld_joint.cpp:83-177,
// spring_joint.cpp:105-173, mouse_joint.cpp:54-100, friction_joint.cpp:39-86, motor_joint.cpp:41-113.
static void InitOther(World& w, const TimeStep& step, Joint& j, Position* pos) {
    const Body& A = w.bodies[j.bodyA];
    const Body& B = w.bodies[j.bodyB];
    j.indexA = A.islandIndex;
    j.indexB = B.islandIndex;
    j.mA = A.invMass;
    j.mB = B.invMass;
    j.localCenterA = A.sweep.localCenter;
    j.localCenterB = B.sweep.localCenter;
    j.localInvIA = A.invI;
    j.localInvIB = B.invI;
    j.iA = pos[j.indexA].I;
    j.iB = pos[j.indexB].I;
    const Vec3 xA = pos[j.indexA].x, xB = pos[j.indexB].x;
    const Quat qA = pos[j.indexA].q, qB = pos[j.indexB].q;
    const scalar h = step.dt;
    switch (j.type) {
    case BX_JOINT_SPHERE:
        j.rA = Mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = Mul(qB, j.localAnchorB - j.localCenterB);
        j.linearMass = PointMass(j.mA, j.mB, j.iA, j.iB, j.rA, j.rB);
        break;
    case BX_JOINT_WELD: {
        j.rA = Mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = Mul(qB, j.localAnchorB - j.localCenterB);
        j.linearMass = PointMass(j.mA, j.mB, j.iA, j.iB, j.rA, j.rB);
        Quat q = RotationError(qA, qB, j.referenceRotation);
        Vec3 C = Mul(qA, -q.v);
        if (j.frequencyHz > scalar(0)) {
            Mat33 invM = j.iA + j.iB;
            Mat33 m = Inverse(invM);
            Vec3 gamma, bias;
            SoftCoefficients(gamma.x, bias.x, j.frequencyHz, j.dampingRatio, m.x.x, C.x, h);
            SoftCoefficients(gamma.y, bias.y, j.frequencyHz, j.dampingRatio, m.y.y, C.y, h);
            SoftCoefficients(gamma.z, bias.z, j.frequencyHz, j.dampingRatio, m.z.z, C.z, h);
            j.gamma3 = Diagonal(gamma.x, gamma.y, gamma.z);
            j.bias3 = bias;
            invM = invM + j.gamma3;
            j.angularMass3 = Inverse(invM);
        } else {
            j.gamma3 = Diagonal(scalar(0));
            j.angularMass3 = Inverse(j.iA + j.iB);
            j.bias3 = j.correctionFactor * step.inv_dt * C;
        }
        break;
    }
    case BX_JOINT_SPRING: {
        j.rA = Mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = Mul(qB, j.localAnchorB - j.localCenterB);
        j.n = xB + j.rB - xA - j.rA;
        scalar length = Length(j.n);
        if (length > kLinearSlop) j.n /= length;
        else j.n.SetZero();
        Vec3 rnA = Cross(j.rA, j.n);
        Vec3 rnB = Cross(j.rB, j.n);
        scalar mass = j.mA + j.mB + Dot(j.iA * rnA, rnA) + Dot(j.iB * rnB, rnB);
        j.mass1 = mass > scalar(0) ? scalar(1) / mass : scalar(0);
        if (j.frequencyHz > scalar(0)) {
            scalar C = length - j.length;
            scalar omega = scalar(2) * kPi * j.frequencyHz;
            scalar d = scalar(2) * j.mass1 * j.dampingRatio * omega;
            scalar k = j.mass1 * omega * omega;
            j.gamma1 = h * (d + h * k);
            j.gamma1 = j.gamma1 != scalar(0) ? scalar(1) / j.gamma1 : scalar(0);
            j.bias1 = j.gamma1 * h * k * C;
            mass += j.gamma1;
            j.mass1 = mass != scalar(0) ? scalar(1) / mass : scalar(0);
        } else {
            j.bias1 = scalar(0);
            j.gamma1 = scalar(0);
        }
        break;
    }
    case BX_JOINT_MOUSE: {
        j.rB = Mul(qB, j.localAnchorB - j.localCenterB);
        Mat33 M = Diagonal(j.mB);
        Mat33 RB = Skew(j.rB);
        Mat33 RBT = Transpose(RB);
        Mat33 invM = M + RB * j.iB * RBT;
        Mat33 m = Inverse(invM);
        Vec3 C = xB + j.rB - j.target;
        Vec3 gamma, bias;
        SoftCoefficients(gamma.x, bias.x, j.frequencyHz, j.dampingRatio, m.x.x, C.x, h);
        SoftCoefficients(gamma.y, bias.y, j.frequencyHz, j.dampingRatio, m.y.y, C.y, h);
        SoftCoefficients(gamma.z, bias.z, j.frequencyHz, j.dampingRatio, m.z.z, C.z, h);
        j.gamma3 = Diagonal(gamma.x, gamma.y, gamma.z);
        j.bias3 = bias;
        invM = invM + j.gamma3;
        j.linearMass = Inverse(invM);
        break;
    }
    case BX_JOINT_FRICTION:
        j.rA = Mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = Mul(qB, j.localAnchorB - j.localCenterB);
        j.linearMass = PointMass(j.mA, j.mB, j.iA, j.iB, j.rA, j.rB);
        j.angularMass3 = j.iA + j.iB;
        break;
    case BX_JOINT_MOTOR: {
        j.rA = Mul(qA, j.linearOffset - j.localCenterA);
        j.rB = Mul(qB, -j.localCenterB);
        j.linearMass = PointMass(j.mA, j.mB, j.iA, j.iB, j.rA, j.rB);
        j.linearError = xB + j.rB - xA - j.rA;
        Quat q1 = Conjugate(qA) * qB;
        if (Dot(q1, j.angularOffset) < scalar(0)) q1 = -q1;
        Quat qw = step.inv_dt * scalar(2) * (j.angularOffset - q1) * Conjugate(q1);
        j.angularVelocity = Mul(qA, qw.v);
        j.angularVelocity *= j.correctionFactor;
        j.angularMass3 = j.iA + j.iB;
        break;
    }
    default: break;
    }
    if (step.warmStarting) {
        j.linearImpulse *= step.dtRatio;
        j.angularImpulse3 *= step.dtRatio;
        j.impulse1 *= step.dtRatio;
    } else {
        j.linearImpulse.SetZero();
        j.angularImpulse3.SetZero();
        j.impulse1 = scalar(0);
    }
}

// Point impulse P at rA / rB: vA -= mA P, wA -= iA (rA x P), and the opposite on B.
static void ApplyPoint(const Joint& j, const Vec3& P, Vec3& vA, Vec3& wA, Vec3& vB, Vec3& wB) {
    vA -= j.mA * P;
    wA -= j.iA * Cross(j.rA, P);
    vB += j.mB * P;
    wB += j.iB * Cross(j.rB, P);
}

// A point or angular impulse clamped to |impulse| <= maxImpulse (friction_joint.cpp:118-133).
static Vec3 ClampedDelta(Vec3& accumulated, const Vec3& impulse, scalar maxImpulse) {
    Vec3 old = accumulated;
    accumulated += impulse;
    if (LengthSquared(accumulated) > maxImpulse * maxImpulse) {
        accumulated.Normalize();
        accumulated *= maxImpulse;
    }
    return accumulated - old;
}

static void WarmStartOther(Joint& j, Velocity* vel) {
    if (j.type == BX_JOINT_MOUSE) {  // mouse_joint.cpp:102-106
        vel[j.indexB].v += j.mB * j.linearImpulse;
        vel[j.indexB].w += j.iB * Cross(j.rB, j.linearImpulse);
        return;
    }
    if (j.type == BX_JOINT_SPRING) {  // spring_joint.cpp:175-183: in place
        Vec3 P = j.impulse1 * j.n;
        vel[j.indexA].v -= j.mA * P;
        vel[j.indexA].w -= j.iA * Cross(j.rA, P);
        vel[j.indexB].v += j.mB * P;
        vel[j.indexB].w += j.iB * Cross(j.rB, P);
        return;
    }
    Vec3 vA = vel[j.indexA].v, wA = vel[j.indexA].w, vB = vel[j.indexB].v, wB = vel[j.indexB].w;
    if (j.type == BX_JOINT_FRICTION) {  // friction_joint.cpp:88-107
        Vec3 P = j.linearImpulse;
        vA -= j.mA * P;
        wA -= j.iA * (Cross(j.rA, P) + j.angularImpulse3);
        vB += j.mB * P;
        wB += j.iB * (Cross(j.rB, P) + j.angularImpulse3);
    } else {  // sphere_joint.cpp:81-97; weld_joint.cpp:179-205 and motor_joint.cpp:115-141 add the angular part
        ApplyPoint(j, j.linearImpulse, vA, wA, vB, wB);
        if (j.type == BX_JOINT_WELD || j.type == BX_JOINT_MOTOR) {
            wA -= j.iA * j.angularImpulse3;
            wB += j.iB * j.angularImpulse3;
        }
    }
    vel[j.indexA].v = vA;
    vel[j.indexA].w = wA;
    vel[j.indexB].v = vB;
    vel[j.indexB].w = wB;
}

static void SolveOtherVelocity(const TimeStep& step, Joint& j, Velocity* vel) {
    const scalar h = step.dt;
    if (j.type == BX_JOINT_MOUSE) {  // mouse_joint.cpp:108-131
        Vec3 vB = vel[j.indexB].v, wB = vel[j.indexB].w;
        Vec3 Cdot = vB + Cross(wB, j.rB);
        Vec3 impulse = -j.linearMass * (Cdot + j.bias3 + j.gamma3 * j.linearImpulse);
        Vec3 oldImpulse = j.linearImpulse;
        j.linearImpulse += impulse;
        scalar maxImpulse = h * j.maxForce;
        if (Dot(j.linearImpulse, j.linearImpulse) > maxImpulse * maxImpulse)
            j.linearImpulse *= maxImpulse / Length(j.linearImpulse);
        impulse = j.linearImpulse - oldImpulse;
        vB += j.mB * impulse;
        wB += j.iB * Cross(j.rB, impulse);
        vel[j.indexB].v = vB;
        vel[j.indexB].w = wB;
        return;
    }
    Vec3 vA = vel[j.indexA].v, wA = vel[j.indexA].w, vB = vel[j.indexB].v, wB = vel[j.indexB].w;
    switch (j.type) {
    case BX_JOINT_SPHERE: {  // sphere_joint.cpp:99-122
        Vec3 Cdot = vB + Cross(wB, j.rB) - vA - Cross(wA, j.rA);
        Vec3 impulse = j.linearMass.Solve(-Cdot);
        j.linearImpulse += impulse;
        ApplyPoint(j, impulse, vA, wA, vB, wB);
        break;
    }
    case BX_JOINT_WELD: {  // weld_joint.cpp:207-246
        {
            Vec3 Cdot = vB + Cross(wB, j.rB) - vA - Cross(wA, j.rA);
            Vec3 impulse = j.linearMass.Solve(-Cdot);
            j.linearImpulse += impulse;
            ApplyPoint(j, impulse, vA, wA, vB, wB);
        }
        {
            Vec3 Cdot = wB - wA;
            Vec3 impulse = -j.angularMass3 * (Cdot + j.bias3 + j.gamma3 * j.angularImpulse3);
            j.angularImpulse3 += impulse;
            wA -= j.iA * impulse;
            wB += j.iB * impulse;
        }
        break;
    }
    case BX_JOINT_SPRING: {  // spring_joint.cpp:185-212
        Vec3 dv = vB + Cross(wB, j.rB) - vA - Cross(wA, j.rA);
        scalar Cdot = Dot(j.n, dv);
        scalar impulse = -j.mass1 * (Cdot + j.bias1 + j.gamma1 * j.impulse1);
        j.impulse1 += impulse;
        ApplyPoint(j, impulse * j.n, vA, wA, vB, wB);
        break;
    }
    case BX_JOINT_FRICTION: {  // friction_joint.cpp:109-158: angular, then linear
        {
            Vec3 Cdot = wB - wA;
            Vec3 impulse = ClampedDelta(j.angularImpulse3, j.angularMass3.Solve(-Cdot), h * j.maxTorque);
            wA -= j.iA * impulse;
            wB += j.iB * impulse;
        }
        {
            Vec3 Cdot = vB + Cross(wB, j.rB) - vA - Cross(wA, j.rA);
            Vec3 impulse = ClampedDelta(j.linearImpulse, j.linearMass.Solve(-Cdot), h * j.maxForce);
            ApplyPoint(j, impulse, vA, wA, vB, wB);
        }
        break;
    }
    case BX_JOINT_MOTOR: {  // motor_joint.cpp:143-200: linear, then angular
        {
            Vec3 Cdot = vB + Cross(wB, j.rB) - vA - Cross(wA, j.rA) + step.inv_dt * j.correctionFactor * j.linearError;
            Vec3 impulse = ClampedDelta(j.linearImpulse, j.linearMass.Solve(-Cdot), h * j.maxForce);
            ApplyPoint(j, impulse, vA, wA, vB, wB);
        }
        {
            Vec3 Cdot = (wB - wA) - j.angularVelocity;
            Vec3 impulse = ClampedDelta(j.angularImpulse3, j.angularMass3.Solve(-Cdot), h * j.maxTorque);
            wA -= j.iA * impulse;
            wB += j.iB * impulse;
        }
        break;
    }
    default: break;
    }
    vel[j.indexA].v = vA;
    vel[j.indexA].w = wA;
    vel[j.indexB].v = vB;
    vel[j.indexB].w = wB;
}

// SolvePositionConstraints: the point constraint of the sphere and weld joints
// (sphere_joint.cpp:124-165, weld_joint.cpp:248-296), the rigid spring (spring_joint.cpp:214-271,
// whose orientations are not renormalized); the mouse, friction and motor joints and a soft spring
// have none.
static bool SolveOtherPosition(Joint& j, Position* pos) {
    if (j.type == BX_JOINT_MOUSE || j.type == BX_JOINT_FRICTION || j.type == BX_JOINT_MOTOR ||
        (j.type == BX_JOINT_SPRING && j.frequencyHz > scalar(0)))
        return true;
    Vec3 xA = pos[j.indexA].x, xB = pos[j.indexB].x;
    Quat qA = pos[j.indexA].q, qB = pos[j.indexB].q;
    Mat33 iA = pos[j.indexA].I, iB = pos[j.indexB].I;
    bool solved;
    if (j.type == BX_JOINT_SPRING) {
        scalar mA = j.mA, mB = j.mB;
        Vec3 rA = Mul(qA, j.localAnchorA - j.localCenterA);
        Vec3 rB = Mul(qB, j.localAnchorB - j.localCenterB);
        Vec3 n = xB + rB - xA - rA;
        scalar length = Length(n);
        scalar C = length - j.length;
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
        iA = RotateToFrame(j.localInvIA, qA);
        xB += mB * impulse;
        qB += Derivative(qB, iB * Cross(rB, impulse));
        iB = RotateToFrame(j.localInvIB, qB);
        solved = Abs(C) < kLinearSlop;
    } else {  // sphere, weld
        Vec3 rA = Mul(qA, j.localAnchorA - j.localCenterA);
        Vec3 rB = Mul(qB, j.localAnchorB - j.localCenterB);
        Mat33 mass = PointMass(j.mA, j.mB, iA, iB, rA, rB);
        Vec3 C = xB + rB - xA - rA;
        Vec3 impulse = mass.Solve(-C);
        xA -= j.mA * impulse;
        qA -= Derivative(qA, iA * Cross(rA, impulse));
        qA.Normalize();
        iA = RotateToFrame(j.localInvIA, qA);
        xB += j.mB * impulse;
        qB += Derivative(qB, iB * Cross(rB, impulse));
        qB.Normalize();
        iB = RotateToFrame(j.localInvIB, qB);
        solved = Length(C) <= kLinearSlop;
    }
    pos[j.indexA].x = xA;
    pos[j.indexA].q = qA;
    pos[j.indexB].x = xB;
    pos[j.indexB].q = qB;
    pos[j.indexA].I = iA;
    pos[j.indexB].I = iB;
    return solved;
}


// ------------------------------------------------------------------ cone, prismatic and wheel joints
// k = [k11 k12; k12 k22] of the two rows perpendicular to the axis, inverted (prismatic_joint.cpp:113-136,
// wheel_joint.cpp:103-126, and again in both position solvers).
static Mat22 PerpMass(scalar mA, scalar mB, const Mat33& iA, const Mat33& iB, const Vec3& s1, const Vec3& s2,
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

static void InitAxis(World& w, const TimeStep& step, Joint& j, Position* pos) {
    const Body& A = w.bodies[j.bodyA];
    const Body& B = w.bodies[j.bodyB];
    j.indexA = A.islandIndex;
    j.indexB = B.islandIndex;
    j.mA = A.invMass;
    j.mB = B.invMass;
    j.localCenterA = A.sweep.localCenter;
    j.localCenterB = B.sweep.localCenter;
    j.localInvIA = A.invI;
    j.localInvIB = B.invI;
    j.iA = pos[j.indexA].I;
    j.iB = pos[j.indexB].I;
    Vec3 xA = pos[j.indexA].x, xB = pos[j.indexB].x;
    Quat qA = pos[j.indexA].q, qB = pos[j.indexB].q;
    scalar mA = j.mA, mB = j.mB;
    Mat33 iA = j.iA, iB = j.iB;
    if (j.type == BX_JOINT_CONE) {  // cone_joint.cpp:70-168
        Quat fA = qA * j.localRotationA;
        Quat fB = qB * j.localRotationB;
        Mat33 RfA = fA.GetRotationMatrix();
        Mat33 RfB = fB.GetRotationMatrix();
        j.rA = Mul(qA, j.localAnchorA - j.localCenterA);
        j.rB = Mul(qB, j.localAnchorB - j.localCenterB);
        j.linearMass = PointMass(mA, mB, iA, iB, j.rA, j.rB);
        if (j.enableLimit) {
            Vec3 u1 = RfA.x;
            Vec3 u2 = RfB.x;
            j.motorAxis = Cross(u2, u1);
            scalar mass = Dot((iA + iB) * j.motorAxis, j.motorAxis);
            j.mass1 = mass > scalar(0) ? scalar(1) / mass : scalar(0);
            scalar cosine = Dot(u2, u1);
            scalar sine = Length(j.motorAxis);
            scalar angle = scalar(::atan2(double(sine), double(cosine)));
            if (scalar(0.5) * j.coneAngle < angle) {
                if (j.limitState != e_atLowerLimit) {
                    j.limitState = e_atLowerLimit;
                    j.limitImpulse = 0;
                }
            } else {
                j.limitState = e_inactiveLimit;
                j.limitImpulse = 0;
            }
        } else {
            j.limitState = e_inactiveLimit;
        }
        if (step.warmStarting) {
            j.linearImpulse *= step.dtRatio;
            j.limitImpulse *= step.dtRatio;
        } else {
            j.linearImpulse.SetZero();
            j.limitImpulse = 0;
        }
        return;
    }
    Vec3 rA = Mul(qA, j.localAnchorA - j.localCenterA);
    Vec3 rB = Mul(qB, j.localAnchorB - j.localCenterB);
    Vec3 d = xB + rB - xA - rA;
    if (j.type == BX_JOINT_PRISMATIC) {  // prismatic_joint.cpp:63-200
        j.motorAxis = Mul(qA, j.localXAxisA);
        j.a1 = Cross(d + rA, j.motorAxis);
        j.a2 = Cross(rB, j.motorAxis);
        j.motorMass = mA + mB + Dot(iA * j.a1, j.a1) + Dot(iB * j.a2, j.a2);
        if (j.motorMass > scalar(0)) j.motorMass = scalar(1) / j.motorMass;
    }
    j.perp1 = Mul(qA, j.localYAxisA);
    j.s1 = Cross(d + rA, j.perp1);
    j.s2 = Cross(rB, j.perp1);
    j.perp2 = Mul(qA, j.localZAxisA);
    j.s3 = Cross(d + rA, j.perp2);
    j.s4 = Cross(rB, j.perp2);
    j.mass22 = PerpMass(mA, mB, iA, iB, j.s1, j.s2, j.s3, j.s4);
    if (j.type == BX_JOINT_PRISMATIC) {
        j.angularMass3 = Inverse(iA + iB);
        Quat q1 = Conjugate(qA) * qB;
        Quat q2 = j.referenceRotation;
        if (Dot(q1, q2) < scalar(0)) q1 = -q1;
        Quat q = scalar(2) * (q2 - q1) * Conjugate(q1);
        Vec3 C = Mul(qA, -q.v);
        j.bias3 = j.correctionFactor * step.inv_dt * C;
        if (j.enableLimit) {
            scalar translation = Dot(j.motorAxis, d);
            if (Abs(j.upperTranslation - j.lowerTranslation) < scalar(2) * kLinearSlop) {
                j.limitState = e_equalLimits;
            } else if (translation <= j.lowerTranslation) {
                if (j.limitState != e_atLowerLimit) {
                    j.limitState = e_atLowerLimit;
                    j.limitImpulse = 0;
                }
            } else if (translation >= j.upperTranslation) {
                if (j.limitState != e_atUpperLimit) {
                    j.limitState = e_atUpperLimit;
                    j.limitImpulse = 0;
                }
            } else {
                j.limitState = e_inactiveLimit;
                j.limitImpulse = 0;
            }
        } else {
            j.limitState = e_inactiveLimit;
            j.limitImpulse = 0;
        }
        if (!j.enableMotor) j.motorImpulse = 0;
        if (step.warmStarting) {
            j.linearImpulse2 *= step.dtRatio;
            j.angularImpulse3 *= step.dtRatio;
            j.motorImpulse *= step.dtRatio;
            j.limitImpulse *= step.dtRatio;
        } else {
            j.linearImpulse2.SetZero();
            j.angularImpulse3.SetZero();
            j.motorImpulse = 0;
            j.limitImpulse = 0;
        }
        return;
    }
    // wheel_joint.cpp:62-190; the spring's axis and arms keep their values while the spring is off
    j.mass1 = 0;
    j.bias1 = 0;
    j.gamma1 = 0;
    if (j.frequencyHz > scalar(0)) {
        j.motorAxis = Mul(qA, j.localXAxisA);
        j.a1 = Cross(d + rA, j.motorAxis);
        j.a2 = Cross(rB, j.motorAxis);
        scalar invMass = mA + mB + Dot(iA * j.a1, j.a1) + Dot(iB * j.a2, j.a2);
        if (invMass > scalar(0)) {
            j.mass1 = scalar(1) / invMass;
            scalar C = Dot(d, j.motorAxis);
            scalar omega = scalar(2) * kPi * j.frequencyHz;
            scalar damp = scalar(2) * j.mass1 * j.dampingRatio * omega;
            scalar k = j.mass1 * omega * omega;
            scalar h = step.dt;
            j.gamma1 = h * (damp + h * k);
            if (j.gamma1 > scalar(0)) j.gamma1 = scalar(1) / j.gamma1;
            j.bias1 = C * h * k * j.gamma1;
            j.mass1 = invMass + j.gamma1;
            if (j.mass1 > scalar(0)) j.mass1 = scalar(1) / j.mass1;
        }
    } else {
        j.springImpulse = 0;
    }
    if (j.enableMotor) {
        j.axisB = Mul(qB, j.localXAxisB);
        j.motorMass = Dot((iA + iB) * j.axisB, j.axisB);
        if (j.motorMass > scalar(0)) j.motorMass = scalar(1) / j.motorMass;
    } else {
        j.motorImpulse = 0;
    }
    {
        Vec3 axisA = Mul(qA, j.localXAxisA);
        Vec3 axisB = Mul(qB, j.localXAxisB);
        j.u = Cross(axisA, axisB);
        j.angularMass1 = Dot((iA + iB) * j.u, j.u);
        if (j.angularMass1 > scalar(0)) j.angularMass1 = scalar(1) / j.angularMass1;
    }
    if (step.warmStarting) {
        j.linearImpulse2 *= step.dtRatio;
        j.springImpulse *= step.dtRatio;
        j.motorImpulse *= step.dtRatio;
        j.impulse1 *= step.dtRatio;
    } else {
        j.linearImpulse2.SetZero();
        j.springImpulse = 0;
        j.motorImpulse = 0;
        j.impulse1 = 0;
    }
}

// vA -= mA P, wA -= iA LA, vB += mB P, wB += iB LB.
static void ApplyRows(const Joint& j, const Vec3& P, const Vec3& LA, const Vec3& LB, Vec3& vA, Vec3& wA, Vec3& vB,
                      Vec3& wB) {
    vA -= j.mA * P;
    wA -= j.iA * LA;
    vB += j.mB * P;
    wB += j.iB * LB;
}

static void WarmStartAxis(Joint& j, Velocity* vel) {
    Vec3 vA = vel[j.indexA].v, wA = vel[j.indexA].w, vB = vel[j.indexB].v, wB = vel[j.indexB].w;
    if (j.type == BX_JOINT_CONE) {  // cone_joint.cpp:170-196
        ApplyPoint(j, j.linearImpulse, vA, wA, vB, wB);
        if (j.enableLimit && j.limitState != e_inactiveLimit) {
            Vec3 L = j.limitImpulse * j.motorAxis;
            wA -= j.iA * L;
            wB += j.iB * L;
        }
    } else if (j.type == BX_JOINT_PRISMATIC) {  // prismatic_joint.cpp:202-222
        Vec3 P = j.linearImpulse2.x * j.perp1 + j.linearImpulse2.y * j.perp2 +
                 (j.motorImpulse + j.limitImpulse) * j.motorAxis;
        Vec3 LA = j.linearImpulse2.x * j.s1 + j.linearImpulse2.y * j.s3 + j.angularImpulse3 +
                  (j.motorImpulse + j.limitImpulse) * j.a1;
        Vec3 LB = j.linearImpulse2.x * j.s2 + j.linearImpulse2.y * j.s4 + j.angularImpulse3 +
                  (j.motorImpulse + j.limitImpulse) * j.a2;
        ApplyRows(j, P, LA, LB, vA, wA, vB, wB);
    } else {  // wheel_joint.cpp:192-222
        Vec3 P = j.linearImpulse2.x * j.perp1 + j.linearImpulse2.y * j.perp2 + j.springImpulse * j.motorAxis;
        Vec3 LA = j.linearImpulse2.x * j.s1 + j.linearImpulse2.y * j.s3 + j.springImpulse * j.a1 +
                  j.motorImpulse * j.axisB;
        Vec3 LB = j.linearImpulse2.x * j.s2 + j.linearImpulse2.y * j.s4 + j.springImpulse * j.a2 +
                  j.motorImpulse * j.axisB;
        ApplyRows(j, P, LA, LB, vA, wA, vB, wB);
        Vec3 L = j.impulse1 * j.u;
        wA += j.iA * L;
        wB -= j.iB * L;
    }
    vel[j.indexA].v = vA;
    vel[j.indexA].w = wA;
    vel[j.indexB].v = vB;
    vel[j.indexB].w = wB;
}

static void SolvePerp(Joint& j, Vec3& vA, Vec3& wA, Vec3& vB, Vec3& wB) {
    Vec2 Cdot;
    Cdot.x = Dot(j.perp1, vB - vA) + Dot(j.s2, wB) - Dot(j.s1, wA);
    Cdot.y = Dot(j.perp2, vB - vA) + Dot(j.s4, wB) - Dot(j.s3, wA);
    Vec2 impulse = j.mass22 * (-Cdot);
    j.linearImpulse2 += impulse;
    Vec3 P = impulse.x * j.perp1 + impulse.y * j.perp2;
    Vec3 LA = impulse.x * j.s1 + impulse.y * j.s3;
    Vec3 LB = impulse.x * j.s2 + impulse.y * j.s4;
    ApplyRows(j, P, LA, LB, vA, wA, vB, wB);
}

static void SolveAxisVelocity(const TimeStep& step, Joint& j, Velocity* vel) {
    Vec3 vA = vel[j.indexA].v, wA = vel[j.indexA].w, vB = vel[j.indexB].v, wB = vel[j.indexB].w;
    if (j.type == BX_JOINT_CONE) {  // cone_joint.cpp:198-240
        Vec3 Cdot = vB + Cross(wB, j.rB) - vA - Cross(wA, j.rA);
        Vec3 P = j.linearMass.Solve(-Cdot);
        j.linearImpulse += P;
        ApplyPoint(j, P, vA, wA, vB, wB);
        if (j.enableLimit && j.limitState != e_inactiveLimit) {
            scalar c = Dot(j.motorAxis, wB - wA);
            scalar impulse = -j.mass1 * c;
            scalar oldImpulse = j.limitImpulse;
            j.limitImpulse = Max(j.limitImpulse + impulse, scalar(0));
            impulse = j.limitImpulse - oldImpulse;
            Vec3 L = impulse * j.motorAxis;
            wA -= j.iA * L;
            wB += j.iB * L;
        }
    } else if (j.type == BX_JOINT_PRISMATIC) {  // prismatic_joint.cpp:224-315
        if (j.enableMotor && j.limitState != e_equalLimits) {
            scalar Cdot = Dot(j.motorAxis, vB - vA) + Dot(j.a2, wB) - Dot(j.a1, wA);
            scalar impulse = j.motorMass * (j.motorSpeed - Cdot);
            scalar oldImpulse = j.motorImpulse;
            scalar maxImpulse = step.dt * j.maxForce;
            j.motorImpulse = Clamp(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            ApplyRows(j, impulse * j.motorAxis, impulse * j.a1, impulse * j.a2, vA, wA, vB, wB);
        }
        SolvePerp(j, vA, wA, vB, wB);
        {
            Vec3 Cdot = wB - wA;
            Vec3 impulse = -j.angularMass3 * (Cdot + j.bias3);
            j.angularImpulse3 += impulse;
            wA -= j.iA * impulse;
            wB += j.iB * impulse;
        }
        if (j.enableLimit && j.limitState != e_inactiveLimit) {
            scalar Cdot = Dot(j.motorAxis, vB - vA) + Dot(j.a2, wB) - Dot(j.a1, wA);
            scalar impulse = j.motorMass * (-Cdot);
            scalar oldImpulse = j.limitImpulse;
            if (j.limitState == e_atLowerLimit) j.limitImpulse = Max(j.limitImpulse + impulse, scalar(0));
            else if (j.limitState == e_atUpperLimit) j.limitImpulse = Min(j.limitImpulse + impulse, scalar(0));
            impulse = j.limitImpulse - oldImpulse;
            ApplyRows(j, impulse * j.motorAxis, impulse * j.a1, impulse * j.a2, vA, wA, vB, wB);
        }
    } else {  // wheel_joint.cpp:224-299
        {
            scalar Cdot = Dot(j.motorAxis, vB - vA) + Dot(j.a2, wB) - Dot(j.a1, wA);
            scalar impulse = -j.mass1 * (Cdot + j.bias1 + j.gamma1 * j.springImpulse);
            j.springImpulse += impulse;
            ApplyRows(j, impulse * j.motorAxis, impulse * j.a1, impulse * j.a2, vA, wA, vB, wB);
        }
        if (j.enableMotor) {
            scalar dw = Dot(wB - wA, j.axisB);
            scalar Cdot = dw - j.motorSpeed;
            scalar impulse = -j.motorMass * Cdot;
            scalar oldImpulse = j.motorImpulse;
            scalar maxImpulse = step.dt * j.maxMotorTorque;
            j.motorImpulse = Clamp(j.motorImpulse + impulse, -maxImpulse, maxImpulse);
            impulse = j.motorImpulse - oldImpulse;
            Vec3 P = j.axisB * impulse;
            wA -= j.iA * P;
            wB += j.iB * P;
        }
        SolvePerp(j, vA, wA, vB, wB);
        {
            scalar Cdot = Dot(wA - wB, j.u);
            scalar impulse = -j.angularMass1 * Cdot;
            j.impulse1 += impulse;
            Vec3 L = impulse * j.u;
            wA += j.iA * L;
            wB -= j.iB * L;
        }
    }
    vel[j.indexA].v = vA;
    vel[j.indexA].w = wA;
    vel[j.indexB].v = vB;
    vel[j.indexB].w = wB;
}

// Moves both bodies by P (linear) and LA / LB (angular), renormalizing and rotating the inverse inertia.
static void Move(const Joint& j, const Vec3& P, const Vec3& LA, const Vec3& LB, Vec3& xA, Quat& qA, Mat33& iA,
                 Vec3& xB, Quat& qB, Mat33& iB) {
    xA -= j.mA * P;
    qA -= Derivative(qA, iA * LA);
    qA.Normalize();
    iA = RotateToFrame(j.localInvIA, qA);
    xB += j.mB * P;
    qB += Derivative(qB, iB * LB);
    qB.Normalize();
    iB = RotateToFrame(j.localInvIB, qB);
}

static bool SolveAxisPosition(Joint& j, Position* pos) {
    Vec3 xA = pos[j.indexA].x, xB = pos[j.indexB].x;
    Quat qA = pos[j.indexA].q, qB = pos[j.indexB].q;
    Mat33 iA = pos[j.indexA].I, iB = pos[j.indexB].I;
    scalar mA = j.mA, mB = j.mB;
    bool solved;
    if (j.type == BX_JOINT_CONE) {  // cone_joint.cpp:242-322
        scalar limitError = 0;
        if (j.enableLimit) {
            Quat fA = qA * j.localRotationA;
            Quat fB = qB * j.localRotationB;
            Mat33 RfA = fA.GetRotationMatrix();
            Mat33 RfB = fB.GetRotationMatrix();
            Vec3 u1 = RfA.x;
            Vec3 u2 = RfB.x;
            Vec3 coneAxis = Cross(u2, u1);
            scalar mass = Dot((iA + iB) * coneAxis, coneAxis);
            scalar coneMass = mass > scalar(0) ? scalar(1) / mass : scalar(0);
            scalar cosine = Dot(u2, u1);
            scalar sine = Length(coneAxis);
            scalar angle = scalar(::atan2(double(sine), double(cosine)));
            if (scalar(0.5) * j.coneAngle < angle) {
                scalar C = scalar(0.5) * j.coneAngle - angle;
                limitError = -C;
                C = Clamp(C + kAngularSlop, -kMaxAngularCorrection, scalar(0));
                scalar coneImpulse = -C * coneMass;
                Vec3 P = coneImpulse * coneAxis;
                qA -= Derivative(qA, iA * P);
                qA.Normalize();
                iA = RotateToFrame(j.localInvIA, qA);
                qB += Derivative(qB, iB * P);
                qB.Normalize();
                iB = RotateToFrame(j.localInvIB, qB);
            }
        }
        Vec3 rA = Mul(qA, j.localAnchorA - j.localCenterA);
        Vec3 rB = Mul(qB, j.localAnchorB - j.localCenterB);
        Vec3 C = xB + rB - xA - rA;
        scalar linearError = Length(C);
        Mat33 mass = PointMass(mA, mB, iA, iB, rA, rB);
        Vec3 P = mass.Solve(-C);
        Move(j, P, Cross(rA, P), Cross(rB, P), xA, qA, iA, xB, qB, iB);
        solved = linearError <= kLinearSlop && limitError <= kAngularSlop;
    } else {  // prismatic_joint.cpp:317-441, wheel_joint.cpp:301-423
        Vec3 rA = Mul(qA, j.localAnchorA - j.localCenterA);
        Vec3 rB = Mul(qB, j.localAnchorB - j.localCenterB);
        Vec3 d = xB + rB - xA - rA;
        Vec3 axis = Mul(qA, j.localXAxisA);
        Vec3 a1 = Cross(d + rA, axis);
        Vec3 a2 = Cross(rB, axis);
        Vec3 perp1 = Mul(qA, j.localYAxisA);
        Vec3 s1 = Cross(d + rA, perp1);
        Vec3 s2 = Cross(rB, perp1);
        Vec3 perp2 = Mul(qA, j.localZAxisA);
        Vec3 s3 = Cross(d + rA, perp2);
        Vec3 s4 = Cross(rB, perp2);
        Vec3 axisB = Mul(qB, j.localXAxisB);
        scalar linearError = 0;
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
            Move(j, P, LA, LB, xA, qA, iA, xB, qB, iB);
        }
        if (j.type == BX_JOINT_PRISMATIC) {
            if (j.enableLimit) {
                scalar C = 0;
                scalar translation = Dot(axis, d);
                if (Abs(j.upperTranslation - j.lowerTranslation) < scalar(2) * kLinearSlop) {
                    C = Clamp(translation, -kMaxLinearCorrection, kMaxLinearCorrection);
                    linearError = Max(linearError, Abs(translation));
                } else if (translation <= j.lowerTranslation) {
                    C = Clamp(translation - j.lowerTranslation + kLinearSlop, -kMaxLinearCorrection, scalar(0));
                    linearError = Max(linearError, j.lowerTranslation - translation);
                } else if (translation >= j.upperTranslation) {
                    C = Clamp(translation - j.upperTranslation - kLinearSlop, scalar(0), kMaxLinearCorrection);
                    linearError = Max(linearError, translation - j.upperTranslation);
                }
                scalar limitMass = mA + mB + Dot(iA * a1, a1) + Dot(iB * a2, a2);
                limitMass = limitMass > scalar(0) ? scalar(1) / limitMass : scalar(0);
                scalar impulse = limitMass * -C;
                Move(j, impulse * axis, impulse * a1, impulse * a2, xA, qA, iA, xB, qB, iB);
            }
            solved = linearError <= kLinearSlop;
        } else {
            scalar angularError = 0;
            Vec3 u = Cross(axis, axisB);
            scalar cosine = Dot(axis, axisB);
            scalar C = cosine - j.referenceCosine;
            scalar sine = Length(u);
            scalar angle = scalar(::atan2(double(sine), double(cosine)));
            scalar angleC = angle - j.referenceAngle;
            angularError += Abs(angleC);
            scalar mass = Dot((iA + iB) * u, u);
            if (mass > scalar(0)) mass = scalar(1) / mass;
            scalar impulse = -mass * C;
            Vec3 P = impulse * u;
            qA += Derivative(qA, iA * P);
            qA.Normalize();
            iA = RotateToFrame(j.localInvIA, qA);
            qB -= Derivative(qB, iB * P);
            qB.Normalize();
            iB = RotateToFrame(j.localInvIB, qB);
            solved = linearError <= kLinearSlop && angularError <= kAngularSlop;
        }
    }
    pos[j.indexA].x = xA;
    pos[j.indexA].q = qA;
    pos[j.indexB].x = xB;
    pos[j.indexB].q = qB;
    pos[j.indexA].I = iA;
    pos[j.indexB].I = iB;
    return solved;
}

static bool IsAxis(u32 type) { return type == BX_JOINT_CONE || type == BX_JOINT_PRISMATIC || type == BX_JOINT_WHEEL; }

static void InitJoint(World& w, const TimeStep& step, Joint& j, Position* pos) {
    if (j.type == BX_JOINT_REVOLUTE) InitRevolute(w, step, j, pos);
    else if (IsAxis(j.type)) InitAxis(w, step, j, pos);
    else InitOther(w, step, j, pos);
}
static void WarmStartJoint(Joint& j, Velocity* vel) {
    if (j.type == BX_JOINT_REVOLUTE) WarmStartRevolute(j, vel);
    else if (IsAxis(j.type)) WarmStartAxis(j, vel);
    else WarmStartOther(j, vel);
}
static void SolveJointVelocity(const TimeStep& step, Joint& j, Velocity* vel) {
    if (j.type == BX_JOINT_REVOLUTE) SolveRevoluteVelocity(step, j, vel);
    else if (IsAxis(j.type)) SolveAxisVelocity(step, j, vel);
    else SolveOtherVelocity(step, j, vel);
}
static bool SolveJointPosition(Joint& j, Position* pos) {
    if (j.type == BX_JOINT_REVOLUTE) return SolveRevolutePosition(j, pos);
    if (IsAxis(j.type)) return SolveAxisPosition(j, pos);
    return SolveOtherPosition(j, pos);
}

// b3JointSolver (joint_solver.cpp:31-64): the island's joints in island order.
void SolveJointsInit(World& w, const TimeStep& step, std::vector<u32>& joints, Position* pos, Velocity*) {
    for (u32 i : joints) InitJoint(w, step, w.joints[i], pos);
}
void SolveJointsWarmStart(World& w, std::vector<u32>& joints, Velocity* vel) {
    for (u32 i : joints) WarmStartJoint(w.joints[i], vel);
}
void SolveJointsVelocity(World& w, const TimeStep& step, std::vector<u32>& joints, Velocity* vel) {
    for (u32 i : joints) SolveJointVelocity(step, w.joints[i], vel);
}
bool SolveJointsPosition(World& w, std::vector<u32>& joints, Position* pos) {
    bool solved = true;
    for (u32 i : joints) {
        bool s = SolveJointPosition(w.joints[i], pos);
        solved = solved && s;
    }
    return solved;
}

}  // namespace orc
