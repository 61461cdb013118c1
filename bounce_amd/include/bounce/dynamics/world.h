// bounce/dynamics/world.h -- b3World, b3Body, b3Fixture, b3Contact and the revolute joint of the drop-in
// API.
//
// The classes keep the reference's public names, members and call semantics
// (include/bounce/dynamics/{world,body,fixture,world_callbacks}.h, contacts/contact.h,
// joints/revolute_joint.h) so that a program written against the reference builds unchanged. Underneath,
// b3World::Step runs on the GPU through the C ABI of libbxstep (include/bxstep.h), and the device world
// is the state of record: creating or destroying bodies, fixtures and joints, the body setters and the
// force and impulse methods are queued in call order and applied to the resident world before the next
// step or the next getter that reads state they change; contacts, manifolds, impulses, proxies and
// sleep state carry across them as in the reference. Getters read back what the device computed.
#pragma once
#include <vector>
#include "../collision/collision.h"
#include "../collision/shapes.h"

class b3World;
class b3Body;
class b3Fixture;
class b3Joint;
class b3Contact;

enum b3BodyType { e_staticBody = 0, e_kinematicBody = 1, e_dynamicBody = 2 };

struct b3BodyDef {
    b3BodyDef() {
        type = e_staticBody;
        awake = true;
        allowSleep = true;
        fixedRotationX = false;
        fixedRotationY = false;
        fixedRotationZ = false;
        userData = nullptr;
        position.SetZero();
        orientation.SetIdentity();
        linearVelocity.SetZero();
        angularVelocity.SetZero();
        gravityScale.Set(scalar(1), scalar(1), scalar(1));
        linearDamping.SetZero();
        angularDamping.SetZero();
        linearSleepTolerance = scalar(0.05);
        angularSleepTolerance = scalar(2) / scalar(180) * B3_PI;
    }
    b3BodyType type;
    bool awake;
    bool allowSleep;
    bool fixedRotationX, fixedRotationY, fixedRotationZ;
    void* userData;
    b3Vec3 position;
    b3Quat orientation;
    b3Vec3 linearVelocity;
    b3Vec3 angularVelocity;
    b3Vec3 linearDamping;
    b3Vec3 angularDamping;
    b3Vec3 gravityScale;
    scalar linearSleepTolerance;
    scalar angularSleepTolerance;
};

// Mass properties of a body (shape.h b3MassData): centre of mass in the body frame, mass, and the
// rotational inertia about the centre of mass.
struct b3MassData {
    b3Vec3 center;
    scalar mass;
    b3Mat33 I;
};

struct b3FixtureDef {
    b3FixtureDef() : shape(nullptr), userData(nullptr), isSensor(false), density(0), restitution(0), friction(0.3f) {}
    const b3Shape* shape;  // cloned by CreateFixture
    void* userData;
    bool isSensor;
    scalar density;
    scalar restitution;
    scalar friction;
};

class b3Fixture {
public:
    b3Shape::Type GetType() const { return m_type; }
    b3Shape* GetShape() { return m_shape; }
    const b3Shape* GetShape() const { return m_shape; }
    b3Body* GetBody() { return m_body; }
    const b3Body* GetBody() const { return m_body; }
    bool IsSensor() const { return m_isSensor; }
    scalar GetDensity() const { return m_density; }
    scalar GetFriction() const { return m_friction; }
    scalar GetRestitution() const { return m_restitution; }
    void* GetUserData() const { return m_userData; }
    void SetUserData(void* d) { m_userData = d; }
    b3Fixture* GetNext() { return m_next; }
    const b3Fixture* GetNext() const { return m_next; }

private:
    friend class b3Body;
    friend class b3World;
    b3Fixture() {}
    ~b3Fixture();
    b3Shape::Type m_type;
    b3Shape* m_shape;  // owned clone
    b3Body* m_body;
    b3Fixture* m_next;
    uint32 m_index;    // creation index in the world (position among the live fixtures)
    bool m_isSensor;
    scalar m_density, m_friction, m_restitution;
    void* m_userData;
};

// A joint in a body's joint list (joint.h b3JointEdge).
struct b3JointEdge {
    b3Body* other;
    b3Joint* joint;
    b3JointEdge* prev;
    b3JointEdge* next;
};

class b3Body {
public:
    b3BodyType GetType() const { return m_type; }
    b3World* GetWorld() { return m_world; }
    const b3World* GetWorld() const { return m_world; }

    // b3Body::CreateFixture / DestroyFixture (body.cpp:105-191): the mass is reset on the device.
    b3Fixture* CreateFixture(const b3FixtureDef& def);
    void DestroyFixture(b3Fixture* fixture);
    b3Fixture* GetFixtureList() { return m_fixtureList; }
    const b3Fixture* GetFixtureList() const { return m_fixtureList; }
    uint32 GetFixtureCount() const { return m_fixtureCount; }
    b3JointEdge* GetJointList() { return m_jointList; }
    const b3JointEdge* GetJointList() const { return m_jointList; }

    // b3Body::SetTransform (body.cpp:194-214): moves the body and its proxies.
    void SetTransform(const b3Transform& xf);
    const b3Transform& GetTransform() const { return State().m_xf; }
    b3Vec3 GetPosition() const { return State().m_xf.translation; }
    b3Quat GetOrientation() const { return State().m_xf.rotation; }
    b3Vec3 GetWorldCenter() const { return State().m_worldCenter; }
    b3Vec3 GetLocalCenter() const { return Mass().m_localCenter; }
    b3Vec3 GetPointVelocity(const b3Vec3& point) const {
        const b3Body& b = State();
        return b.m_linearVelocity + b3Cross(b.m_angularVelocity, point - b.m_worldCenter);
    }

    b3Vec3 GetLinearVelocity() const { return State().m_linearVelocity; }
    void SetLinearVelocity(const b3Vec3& v);
    b3Vec3 GetAngularVelocity() const { return State().m_angularVelocity; }
    void SetAngularVelocity(const b3Vec3& w);

    // Forces, torques and impulses (body.h:727-817); `wake` wakes a sleeping body first, and a body
    // that stays asleep ignores the call.
    void ApplyForce(const b3Vec3& force, const b3Vec3& point, bool wake);
    void ApplyForceToCenter(const b3Vec3& force, bool wake);
    void ApplyTorque(const b3Vec3& torque, bool wake);
    void ApplyLinearImpulse(const b3Vec3& impulse, const b3Vec3& point, bool wake);
    void ApplyAngularImpulse(const b3Vec3& impulse, bool wake);

    scalar GetMass() const { return Mass().m_mass; }
    scalar GetInverseMass() const { return Mass().m_invMass; }
    const b3Mat33& GetInertia() const { return Mass().m_I; }  // about the centre of mass, body frame
    const b3Mat33& GetWorldInverseInertia() const { return Mass().m_worldInvI; }
    void GetMassData(b3MassData* data) const {
        const b3Body& b = Mass();
        data->mass = b.m_mass;
        data->I = b.m_I;
        data->center = b.m_localCenter;
    }
    scalar GetLinearEnergy() const {
        const b3Body& b = Mass();
        b3Vec3 P = b.m_mass * b.m_linearVelocity;
        return b3Dot(P, b.m_linearVelocity);
    }
    scalar GetAngularEnergy() const;
    scalar GetEnergy() const { return scalar(0.5) * (GetLinearEnergy() + GetAngularEnergy()); }

    b3Vec3 GetLocalVector(const b3Vec3& v) const { return b3MulC(State().m_xf.rotation, v); }
    b3Vec3 GetWorldVector(const b3Vec3& v) const { return b3Mul(State().m_xf.rotation, v); }
    b3Vec3 GetLocalPoint(const b3Vec3& p) const { return b3MulT(State().m_xf, p); }
    b3Vec3 GetWorldPoint(const b3Vec3& p) const { return b3Mul(State().m_xf, p); }
    b3Quat GetLocalFrame(const b3Quat& frame) const { return b3Conjugate(State().m_xf.rotation) * frame; }
    b3Quat GetWorldFrame(const b3Quat& frame) const { return State().m_xf.rotation * frame; }

    const b3Vec3& GetLinearDamping() const { return m_linearDamping; }
    void SetLinearDamping(const b3Vec3& damping);
    const b3Vec3& GetAngularDamping() const { return m_angularDamping; }
    void SetAngularDamping(const b3Vec3& damping);
    const b3Vec3& GetGravityScale() const { return m_gravityScale; }
    void SetGravityScale(const b3Vec3& scale);
    void SetLinearSleepTolerance(scalar tolerance);
    scalar GetLinearSleepTolerance() const { return m_linearSleepTolerance; }
    void SetAngularSleepTolerance(scalar tolerance);
    scalar GetAngularSleepTolerance() const { return m_angularSleepTolerance; }
    // b3Body::SetFixedRotation (body.cpp:507-548): zeroes the angular velocity and resets the mass.
    void SetFixedRotation(bool flagX, bool flagY, bool flagZ);

    bool IsAwake() const { return State().m_awake; }
    void SetAwake(bool flag);
    scalar GetSleepTime() const { return State().m_sleepTime; }
    void SetSleepingAllowed(bool flag);
    bool IsSleepingAllowed() const { return m_allowSleep; }

    b3Body* GetNext() { return m_next; }
    const b3Body* GetNext() const { return m_next; }
    void* GetUserData() const { return m_userData; }
    void SetUserData(void* d) { m_userData = d; }

private:
    friend class b3World;
    friend class b3Fixture;
    b3Body(const b3BodyDef& def, b3World* world, uint32 index);
    ~b3Body();
    const b3Body& State() const;  // this body after the world's queued operations, read back
    const b3Body& Mass() const;   // the same, with the mass properties read back
    void QueueEdit(uint32 op, const b3Vec3& v, const float* p, bool flag, bool changesState);

    b3BodyDef m_def;  // creation parameters
    b3BodyType m_type;
    b3World* m_world;
    uint32 m_index;   // creation index in the world (position among the live bodies)
    // read back from the device
    b3Transform m_xf;
    b3Vec3 m_worldCenter, m_localCenter;
    b3Vec3 m_linearVelocity, m_angularVelocity;
    bool m_awake;
    scalar m_sleepTime;
    scalar m_mass, m_invMass;
    b3Mat33 m_I, m_worldInvI;
    // kept on the host as well (set only through this API)
    b3Vec3 m_linearDamping, m_angularDamping, m_gravityScale;
    scalar m_linearSleepTolerance, m_angularSleepTolerance;
    bool m_allowSleep;
    bool m_fixedX, m_fixedY, m_fixedZ;
    b3Fixture* m_fixtureList;  // newest first, as the reference's list
    uint32 m_fixtureCount;
    b3JointEdge* m_jointList;  // newest first
    b3Body* m_next;            // world body list, newest first
    void* m_userData;
};

enum b3JointType {
    e_unknownJoint, e_mouseJoint, e_springJoint, e_weldJoint, e_revoluteJoint, e_sphereJoint, e_coneJoint,
    e_frictionJoint, e_motorJoint, e_prismaticJoint, e_wheelJoint
};

struct b3JointDef {
    b3JointDef() : type(e_unknownJoint), userData(nullptr), bodyA(nullptr), bodyB(nullptr), collideConnected(false) {}
    b3JointType type;
    void* userData;
    b3Body* bodyA;
    b3Body* bodyB;
    bool collideConnected;
};

struct b3RevoluteJointDef : public b3JointDef {
    b3RevoluteJointDef() {
        type = e_revoluteJoint;
        localAnchorA.SetZero();
        localRotationA.SetIdentity();
        localAnchorB.SetZero();
        localRotationB.SetIdentity();
        referenceRotation.SetIdentity();
        enableLimit = false;
        lowerAngle = scalar(0);
        upperAngle = scalar(0);
        enableMotor = false;
        motorSpeed = scalar(0);
        maxMotorTorque = scalar(0);
        m_initialize = false;
        m_axis.SetZero();
        m_anchor.SetZero();
    }
    // Frames from a world hinge axis and anchor and the bodies' current transforms
    // (b3RevoluteJointDef::Initialize); evaluated on the device when the joint is created there.
    void Initialize(b3Body* bA, b3Body* bB, const b3Vec3& axis, const b3Vec3& anchor, scalar lower, scalar upper) {
        bodyA = bA;
        bodyB = bB;
        m_initialize = true;
        m_axis = axis;
        m_anchor = anchor;
        lowerAngle = lower;
        upperAngle = upper;
    }
    b3Vec3 localAnchorA;
    b3Quat localRotationA;
    b3Vec3 localAnchorB;
    b3Quat localRotationB;
    b3Quat referenceRotation;
    bool enableLimit;
    scalar lowerAngle, upperAngle;
    bool enableMotor;
    scalar motorSpeed;
    scalar maxMotorTorque;
    bool m_initialize;
    b3Vec3 m_axis, m_anchor;
};

// The other joint definitions (sphere_joint.h, weld_joint.h, spring_joint.h, mouse_joint.h,
// friction_joint.h, motor_joint.h): the same members and defaults. Initialize records its
// arguments; the local frames are computed from the bodies' transforms on the device when the
// joint is created there, as the reference's Initialize computes them from the current transforms.
struct b3SphereJointDef : public b3JointDef {
    b3SphereJointDef() : m_initialize(false) {
        type = e_sphereJoint;
        localAnchorA.SetZero();
        localAnchorB.SetZero();
        m_anchor.SetZero();
    }
    void Initialize(b3Body* bA, b3Body* bB, const b3Vec3& anchor) {
        bodyA = bA;
        bodyB = bB;
        m_initialize = true;
        m_anchor = anchor;
    }
    b3Vec3 localAnchorA;
    b3Vec3 localAnchorB;
    bool m_initialize;
    b3Vec3 m_anchor;
};

struct b3WeldJointDef : public b3JointDef {
    b3WeldJointDef() : m_initialize(false) {
        type = e_weldJoint;
        localAnchorA.SetZero();
        localAnchorB.SetZero();
        referenceRotation.SetIdentity();
        frequencyHz = scalar(0);
        dampingRatio = scalar(0);
        correctionFactor = scalar(0.2);
        m_anchor.SetZero();
    }
    void Initialize(b3Body* bA, b3Body* bB, const b3Vec3& anchor) {
        bodyA = bA;
        bodyB = bB;
        m_initialize = true;
        m_anchor = anchor;
    }
    b3Vec3 localAnchorA;
    b3Vec3 localAnchorB;
    b3Quat referenceRotation;
    scalar frequencyHz;
    scalar dampingRatio;
    scalar correctionFactor;
    bool m_initialize;
    b3Vec3 m_anchor;
};

struct b3SpringJointDef : public b3JointDef {
    b3SpringJointDef() : m_initialize(false) {
        type = e_springJoint;
        localAnchorA.SetZero();
        localAnchorB.SetZero();
        length = scalar(0);
        frequencyHz = scalar(0);
        dampingRatio = scalar(0);
        m_anchorA.SetZero();
        m_anchorB.SetZero();
    }
    void Initialize(b3Body* bA, b3Body* bB, const b3Vec3& anchorA, const b3Vec3& anchorB) {
        bodyA = bA;
        bodyB = bB;
        m_initialize = true;
        m_anchorA = anchorA;
        m_anchorB = anchorB;
    }
    b3Vec3 localAnchorA;
    b3Vec3 localAnchorB;
    scalar length;
    scalar frequencyHz;
    scalar dampingRatio;
    bool m_initialize;
    b3Vec3 m_anchorA, m_anchorB;
};

struct b3MouseJointDef : public b3JointDef {
    b3MouseJointDef() {
        type = e_mouseJoint;
        target.SetZero();
        maxForce = scalar(0);
        frequencyHz = scalar(6);
        dampingRatio = scalar(0.8);
    }
    b3Vec3 target;
    scalar maxForce;
    scalar frequencyHz;
    scalar dampingRatio;
};

struct b3FrictionJointDef : public b3JointDef {
    b3FrictionJointDef() : m_initialize(false) {
        type = e_frictionJoint;
        localAnchorA.SetZero();
        localAnchorB.SetZero();
        maxForce = scalar(0);
        maxTorque = scalar(0);
        m_anchor.SetZero();
    }
    void Initialize(b3Body* bA, b3Body* bB, const b3Vec3& anchor) {
        bodyA = bA;
        bodyB = bB;
        m_initialize = true;
        m_anchor = anchor;
    }
    b3Vec3 localAnchorA;
    b3Vec3 localAnchorB;
    scalar maxForce;
    scalar maxTorque;
    bool m_initialize;
    b3Vec3 m_anchor;
};

struct b3MotorJointDef : public b3JointDef {
    b3MotorJointDef() : m_initialize(false) {
        type = e_motorJoint;
        linearOffset.SetZero();
        angularOffset.SetIdentity();
        maxForce = scalar(0);
        maxTorque = scalar(0);
        correctionFactor = scalar(0.2);
    }
    void Initialize(b3Body* bA, b3Body* bB) {
        bodyA = bA;
        bodyB = bB;
        m_initialize = true;
    }
    b3Vec3 linearOffset;
    b3Quat angularOffset;
    scalar maxForce;
    scalar maxTorque;
    scalar correctionFactor;
    bool m_initialize;
};

struct b3ConeJointDef : public b3JointDef {
    b3ConeJointDef() : m_initialize(false) {
        type = e_coneJoint;
        localAnchorA.SetZero();
        localRotationA.SetIdentity();
        localAnchorB.SetZero();
        localRotationB.SetIdentity();
        referenceRotation.SetIdentity();
        enableConeLimit = false;
        coneAngle = scalar(0);
        m_axis.SetZero();
        m_anchor.SetZero();
    }
    void Initialize(b3Body* bA, b3Body* bB, const b3Vec3& axis, const b3Vec3& anchor, scalar angle) {
        bodyA = bA;
        bodyB = bB;
        m_initialize = true;
        m_axis = axis;
        m_anchor = anchor;
        coneAngle = angle;
    }
    b3Vec3 localAnchorA;
    b3Quat localRotationA;
    b3Vec3 localAnchorB;
    b3Quat localRotationB;
    b3Quat referenceRotation;
    bool enableConeLimit;
    scalar coneAngle;
    bool m_initialize;
    b3Vec3 m_axis, m_anchor;
};

struct b3PrismaticJointDef : public b3JointDef {
    b3PrismaticJointDef() : m_initialize(false) {
        type = e_prismaticJoint;
        localAnchorA.SetZero();
        localAnchorB.SetZero();
        localAxisA.Set(scalar(1), scalar(0), scalar(0));
        referenceRotation.SetIdentity();
        enableLimit = false;
        lowerTranslation = scalar(0);
        upperTranslation = scalar(0);
        enableMotor = false;
        maxMotorForce = scalar(0);
        motorSpeed = scalar(0);
        correctionFactor = scalar(0.2);
        m_anchor.SetZero();
        m_axis.SetZero();
    }
    void Initialize(b3Body* bA, b3Body* bB, const b3Vec3& anchor, const b3Vec3& axis) {
        bodyA = bA;
        bodyB = bB;
        m_initialize = true;
        m_anchor = anchor;
        m_axis = axis;
    }
    b3Vec3 localAnchorA;
    b3Vec3 localAnchorB;
    b3Vec3 localAxisA;
    b3Quat referenceRotation;
    bool enableLimit;
    scalar lowerTranslation;
    scalar upperTranslation;
    bool enableMotor;
    scalar maxMotorForce;
    scalar motorSpeed;
    scalar correctionFactor;
    bool m_initialize;
    b3Vec3 m_anchor, m_axis;
};

struct b3WheelJointDef : public b3JointDef {
    b3WheelJointDef() : m_initialize(false) {
        type = e_wheelJoint;
        localAnchorA.SetZero();
        localAnchorB.SetZero();
        localAxisA.Set(scalar(0), scalar(1), scalar(0));
        localAxisB.Set(scalar(1), scalar(0), scalar(0));
        enableMotor = false;
        maxMotorTorque = scalar(0);
        motorSpeed = scalar(0);
        frequencyHz = scalar(4);
        dampingRatio = scalar(0.7);
        m_anchor.SetZero();
        m_axisA.SetZero();
        m_axisB.SetZero();
    }
    void Initialize(b3Body* bA, b3Body* bB, const b3Vec3& anchor, const b3Vec3& axisA, const b3Vec3& axisB) {
        bodyA = bA;
        bodyB = bB;
        m_initialize = true;
        m_anchor = anchor;
        m_axisA = axisA;
        m_axisB = axisB;
    }
    b3Vec3 localAnchorA;
    b3Vec3 localAnchorB;
    b3Vec3 localAxisA;
    b3Vec3 localAxisB;
    bool enableMotor;
    scalar maxMotorTorque;
    scalar motorSpeed;
    scalar frequencyHz;
    scalar dampingRatio;
    bool m_initialize;
    b3Vec3 m_anchor, m_axisA, m_axisB;
};

class b3Joint {
public:
    b3JointType GetType() const { return m_def.type; }
    b3Body* GetBodyA() { return m_def.bodyA; }
    const b3Body* GetBodyA() const { return m_def.bodyA; }
    b3Body* GetBodyB() { return m_def.bodyB; }
    const b3Body* GetBodyB() const { return m_def.bodyB; }
    bool GetCollideConnected() const { return m_def.collideConnected; }
    void* GetUserData() const { return m_userData; }
    void SetUserData(void* d) { m_userData = d; }
    b3Joint* GetNext() { return m_next; }
    const b3Joint* GetNext() const { return m_next; }

private:
    friend class b3World;
    b3Joint() {}
    b3JointDef m_def;  // the definition's common members
    uint32 m_index;    // creation index in the world
    b3JointEdge m_edgeA, m_edgeB;
    b3Joint* m_next;
    void* m_userData;
};

typedef b3Joint b3RevoluteJoint;
typedef b3Joint b3SphereJoint;
typedef b3Joint b3WeldJoint;
typedef b3Joint b3SpringJoint;
typedef b3Joint b3MouseJoint;
typedef b3Joint b3FrictionJoint;
typedef b3Joint b3MotorJoint;
typedef b3Joint b3ConeJoint;
typedef b3Joint b3PrismaticJoint;
typedef b3Joint b3WheelJoint;

// A touching or merely overlapping pair of fixtures (contact.h); the world's list holds the pairs the
// last step (or topology change) left, newest first. Valid until the next step or world change.
class b3Contact {
public:
    b3Fixture* GetFixtureA() { return m_fixtureA; }
    const b3Fixture* GetFixtureA() const { return m_fixtureA; }
    b3Fixture* GetFixtureB() { return m_fixtureB; }
    const b3Fixture* GetFixtureB() const { return m_fixtureB; }
    uint32 GetManifoldCount() const { return m_manifoldCount; }
    b3Manifold* GetManifold(uint32 index) { return m_manifolds + index; }
    const b3Manifold* GetManifold(uint32 index) const { return m_manifolds + index; }
    // Contact points in world space (b3WorldManifold::Initialize, collision.cpp:63-100).
    void GetWorldManifold(b3WorldManifold* out, uint32 index) const;
    bool IsTouching() const { return m_touching; }
    b3Contact* GetNext() { return m_next; }
    const b3Contact* GetNext() const { return m_next; }

private:
    friend class b3World;
    b3Fixture* m_fixtureA;
    b3Fixture* m_fixtureB;
    uint32 m_manifoldCount;
    b3Manifold m_manifolds[3];
    bool m_touching;
    b3Contact* m_next;
};

// Query callbacks (world_callbacks.h).
class b3QueryListener {
public:
    virtual ~b3QueryListener() {}
    // A fixture whose fat AABB overlaps the box; return false to end the query.
    virtual bool ReportFixture(b3Fixture* fixture) = 0;
};

class b3QueryFilter {
public:
    virtual ~b3QueryFilter() {}
    virtual bool ShouldReport(b3Fixture* fixture) = 0;
};

class b3RayCastListener {
public:
    virtual ~b3RayCastListener() {}
    // A hit; return 0 to end the cast, a positive fraction to clip the ray there, -1 to ignore the hit.
    virtual scalar ReportFixture(b3Fixture* fixture, const b3Vec3& point, const b3Vec3& normal, scalar fraction) = 0;
};

class b3RayCastFilter {
public:
    virtual ~b3RayCastFilter() {}
    virtual bool ShouldRayCast(b3Fixture* fixture) = 0;
};

class b3ShapeCastListener {
public:
    virtual ~b3ShapeCastListener() {}
    // A touching cast; the return value becomes the cast's maximum fraction (world.cpp:684-690): 0 ends
    // the cast, and later casts beyond it are not reported.
    virtual scalar ReportFixture(b3Fixture* fixture, const b3Vec3& point, const b3Vec3& normal, scalar fraction) = 0;
};

class b3ShapeCastFilter {
public:
    virtual ~b3ShapeCastFilter() {}
    virtual bool ShouldShapeCast(b3Fixture* fixture) = 0;
};

// Output of b3World::ShapeCastSingle.
struct b3ShapeCastSingleOutput {
    b3Fixture* fixture;
    b3Vec3 point;
    b3Vec3 normal;
    scalar fraction;
};

// Output of b3World::RayCastSingle.
struct b3RayCastSingleOutput {
    b3Fixture* fixture;
    b3Vec3 point;
    b3Vec3 normal;
    scalar fraction;
};

struct b3WorldDevice;  // the device batch behind a world

class b3World {
public:
    b3World();
    ~b3World();

    void SetSleeping(bool flag);  // clearing it wakes every body (world.cpp:82-92)
    bool GetSleeping() const { return m_sleeping; }
    void SetWarmStarting(bool flag);
    bool GetWarmStarting() const { return m_warmStarting; }
    void SetGravity(const b3Vec3& gravity);
    const b3Vec3& GetGravity() const { return m_gravity; }
    void SetAutoClearForces(bool flag);
    bool GetAutoClearForces() const { return m_autoClearForces; }
    void ClearForces();

    b3Body* CreateBody(const b3BodyDef& def);
    // b3World::DestroyBody (world.cpp:112-175): its joints (waking their bodies), its fixtures' contacts
    // (waking the other bodies of touching ones) and proxies go with it.
    void DestroyBody(b3Body* body);
    b3Joint* CreateJoint(const b3JointDef& def);
    void DestroyJoint(b3Joint* joint);  // wakes both bodies

    // b3World::Step: one step of this world on the GPU.
    void Step(scalar dt, uint32 velocityIterations, uint32 positionIterations);

    // World queries on the device (world.cpp:452-566, 852-870). RayCast reports hits in increasing
    // fraction (the reference reports them in its tree's order; a listener that returns the fraction
    // it is given ends with the same closest hit, one that returns 1 sees the same hits); QueryAABB
    // reports fixtures in creation order. A null filter admits every fixture.
    void RayCast(b3RayCastListener* listener, b3RayCastFilter* filter, const b3Vec3& point1,
                 const b3Vec3& point2) const;
    bool RayCastSingle(b3RayCastSingleOutput* output, b3RayCastFilter* filter, const b3Vec3& point1,
                       const b3Vec3& point2) const;
    void QueryAABB(b3QueryListener* listener, b3QueryFilter* filter, const b3AABB& aabb) const;
    // Shape casts (world.cpp:568-850) of a sphere, capsule, triangle or hull at `xf` moved by
    // `displacement`: every touching cast against a convex fixture or a mesh triangle, reported in
    // increasing fraction as RayCast reports hits.
    void ShapeCast(b3ShapeCastListener* listener, b3ShapeCastFilter* filter, const b3Shape* shape,
                   const b3Transform& xf, const b3Vec3& displacement) const;
    bool ShapeCastSingle(b3ShapeCastSingleOutput* output, b3ShapeCastFilter* filter, const b3Shape* shape,
                         const b3Transform& xf, const b3Vec3& displacement) const;

    b3Body* GetBodyList() { return m_bodyList; }
    const b3Body* GetBodyList() const { return m_bodyList; }
    uint32 GetBodyCount() const { return uint32(m_bodies.size()); }
    b3Joint* GetJointList() { return m_jointList; }
    const b3Joint* GetJointList() const { return m_jointList; }
    uint32 GetJointCount() const { return uint32(m_joints.size()); }
    b3Contact* GetContactList();
    const b3Contact* GetContactList() const { return const_cast<b3World*>(this)->GetContactList(); }
    uint32 GetContactCount() const;

    // Text of the last device error ("" if none). A failed call leaves the device world as it was and
    // keeps the queued operations.
    const char* GetLastError() const;

private:
    friend class b3Body;
    friend class b3Fixture;
    // queue of operations not yet applied to the device world
    void QueueTopology(uint32 op, uint32 index);
    void QueueEdit(const void* edit, bool changesState);
    int Flush();         // apply the queue (creating the device world on first use)
    void ReadBodies();   // transforms, velocities, sleep state
    void ReadMass();     // mass properties
    void Renumber();
    void RemoveJoint(b3Joint* joint);
    uint32 HullIndex(const b3Hull* hull);
    bool CastShapeOf(const b3Shape* shape, void* cast);  // a bx_cast_shape
    int FlushLibrary();  // Flush, then send hulls the device world has not seen (cast shapes)
    uint32 MeshIndex(const b3Mesh* mesh);

    b3Vec3 m_gravity;
    bool m_sleeping, m_warmStarting, m_autoClearForces;
    std::vector<b3Body*> m_bodies;       // creation order
    std::vector<b3Fixture*> m_fixtures;  // creation order
    std::vector<b3Joint*> m_joints;      // creation order
    b3Body* m_bodyList;
    b3Joint* m_jointList;
    b3WorldDevice* m_device;
};
