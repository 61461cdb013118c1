// bounce/dynamics/world.h -- b3World, b3Body, b3Fixture and the revolute joint of the drop-in API.
//
// The classes keep the reference's public names, members and call semantics
// (include/bounce/dynamics/{world,body,fixture}.h, joints/revolute_joint.h) so that a program written
// against the reference builds unchanged. Underneath, b3World::Step runs on the GPU through the C ABI
// of libbxstep (include/bxstep.h): the world mirrors its bodies, fixtures and joints into descriptors
// when they change, steps on the device and reads back the state the getters return.
#pragma once
#include <vector>
#include "../collision/shapes.h"

class b3World;
class b3Body;

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
    ~b3Fixture();

private:
    friend class b3Body;
    friend class b3World;
    b3Fixture() {}
    b3Shape::Type m_type;
    b3Shape* m_shape;  // owned clone
    b3Body* m_body;
    b3Fixture* m_next;
    uint32 m_index;    // creation index in the world
    bool m_isSensor;
    scalar m_density, m_friction, m_restitution;
    void* m_userData;
};

class b3Body {
public:
    b3BodyType GetType() const { return m_type; }
    b3Fixture* CreateFixture(const b3FixtureDef& def);

    const b3Transform& GetTransform() const { return m_xf; }
    const b3Vec3& GetPosition() const { return m_xf.translation; }
    const b3Quat& GetOrientation() const { return m_xf.rotation; }
    const b3Vec3& GetWorldCenter() const { return m_worldCenter; }
    b3Vec3 GetWorldPoint(const b3Vec3& p) const { return b3Mul(m_xf, p); }
    b3Vec3 GetLocalPoint(const b3Vec3& p) const { return b3MulT(m_xf, p); }

    const b3Vec3& GetLinearVelocity() const { return m_linearVelocity; }
    void SetLinearVelocity(const b3Vec3& v);
    const b3Vec3& GetAngularVelocity() const { return m_angularVelocity; }
    void SetAngularVelocity(const b3Vec3& w);

    bool IsAwake() const { return m_awake; }
    scalar GetSleepTime() const { return m_sleepTime; }

    b3Fixture* GetFixtureList() { return m_fixtureList; }
    const b3Fixture* GetFixtureList() const { return m_fixtureList; }
    b3Body* GetNext() { return m_next; }
    const b3Body* GetNext() const { return m_next; }
    b3World* GetWorld() { return m_world; }
    void* GetUserData() const { return m_userData; }
    void SetUserData(void* d) { m_userData = d; }

private:
    friend class b3World;
    b3Body(const b3BodyDef& def, b3World* world, uint32 index);
    ~b3Body();
    b3BodyDef m_def;  // creation parameters (damping, gravity scale, tolerances, flags)
    b3BodyType m_type;
    b3World* m_world;
    uint32 m_index;   // creation index in the world
    b3Transform m_xf;
    b3Vec3 m_worldCenter;
    b3Vec3 m_linearVelocity, m_angularVelocity;
    bool m_awake;
    scalar m_sleepTime;
    b3Fixture* m_fixtureList;  // newest first, as the reference's list
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
    // (b3RevoluteJointDef::Initialize); evaluated when the world uploads the joint.
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

class b3Joint {
public:
    b3JointType GetType() const { return m_def.type; }
    b3Body* GetBodyA() { return m_def.bodyA; }
    b3Body* GetBodyB() { return m_def.bodyB; }
    b3Joint* GetNext() { return m_next; }

private:
    friend class b3World;
    b3Joint() {}
    b3RevoluteJointDef m_def;
    b3Joint* m_next;
};

typedef b3Joint b3RevoluteJoint;

struct b3WorldDevice;  // the device batch behind a world

class b3World {
public:
    b3World();
    ~b3World();

    void SetSleeping(bool flag);
    bool GetSleeping() const { return m_sleeping; }
    void SetWarmStarting(bool flag);
    bool GetWarmStarting() const { return m_warmStarting; }
    void SetGravity(const b3Vec3& gravity);
    const b3Vec3& GetGravity() const { return m_gravity; }
    void SetAutoClearForces(bool flag);
    bool GetAutoClearForces() const { return m_autoClearForces; }

    b3Body* CreateBody(const b3BodyDef& def);
    void DestroyBody(b3Body* body);
    b3Joint* CreateJoint(const b3JointDef& def);
    void DestroyJoint(b3Joint* joint);

    // b3World::Step: one step of every body of this world on the GPU.
    void Step(scalar dt, uint32 velocityIterations, uint32 positionIterations);

    b3Body* GetBodyList() { return m_bodyList; }
    const b3Body* GetBodyList() const { return m_bodyList; }
    uint32 GetBodyCount() const { return uint32(m_bodies.size()); }
    b3Joint* GetJointList() { return m_jointList; }
    uint32 GetJointCount() const { return uint32(m_joints.size()); }
    uint32 GetContactCount() const { return m_contactCount; }

    // Text of the last device error ("" if none); Step leaves the world unchanged on an error.
    const char* GetLastError() const;

private:
    friend class b3Body;
    void MarkTopologyChanged() { m_topologyChanged = true; }
    void MarkStateChanged() { m_stateChanged = true; }
    int Upload();
    int PushState();
    int Pull();

    b3Vec3 m_gravity;
    bool m_sleeping, m_warmStarting, m_autoClearForces;
    std::vector<b3Body*> m_bodies;       // creation order
    std::vector<b3Fixture*> m_fixtures;  // creation order
    std::vector<b3Joint*> m_joints;      // creation order
    b3Body* m_bodyList;
    b3Joint* m_jointList;
    uint32 m_contactCount;
    bool m_topologyChanged, m_stateChanged;
    b3WorldDevice* m_device;
};
