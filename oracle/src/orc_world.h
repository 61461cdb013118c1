// orc_world.h -- oracle (test infrastructure): world state, restating the reference's dynamics
// data model (include/bounce/dynamics/{world,body,fixture}.h, contacts/contact.h,
// joints/joint.h, joints/revolute_joint.h) with the same list-order semantics.
#pragma once
#include <vector>
#include <cstdint>
#include "orc_geom.h"

namespace orc {

enum BodyType : u32 { e_staticBody = 0, e_kinematicBody = 1, e_dynamicBody = 2 };
enum BodyFlags : u32 {
    e_awakeFlag = 0x0001, e_islandFlag = 0x0002, e_autoSleepFlag = 0x0004,
    e_fixedRotationX = 0x0008, e_fixedRotationY = 0x0010, e_fixedRotationZ = 0x0020
};

struct Contact;
struct Joint;

struct JointEdge { u32 joint; u32 other; };

struct Body {
    u32 type = e_staticBody;
    u32 flags = 0;
    u32 islandIndex = 0;
    scalar mass = 0, invMass = 0;
    Mat33 I, invI, worldInvI;
    Vec3 force, torque;
    Vec3 linearVelocity, angularVelocity;
    Sweep sweep;
    Transform xf;
    Vec3 linearDamping, angularDamping, gravityScale;
    scalar linearSleepTolerance = 0, angularSleepTolerance = 0, sleepTime = 0;
    std::vector<u32> fixtures;       // list order: front = last created (body.cpp:120-122)
    std::vector<JointEdge> joints;   // list order: front = last created (joint_manager.cpp)
    bool IsAwake() const { return (flags & e_awakeFlag) != 0; }
    void SetAwake(bool flag) {       // body.h:588-607
        if (flag) {
            if (!IsAwake()) { flags |= e_awakeFlag; sleepTime = 0; }
        } else {
            flags &= ~u32(e_awakeFlag);
            sleepTime = 0;
            force.SetZero(); torque.SetZero();
            linearVelocity.SetZero(); angularVelocity.SetZero();
        }
    }
    void SynchronizeTransform() {    // body.h:847-851
        xf.rotation = sweep.orientation;
        xf.translation = sweep.worldCenter - Mul(xf.rotation, sweep.localCenter);
    }
};

struct Fixture {
    u32 body = 0;
    Shape shape;
    scalar density = 0, friction = 0, restitution = 0;
    bool isSensor = false;
    u32 proxyId = 0;
    std::vector<Contact*> contacts;  // contact edge list, front = last created (contact_manager.cpp:118-139)
};

struct Contact {
    u32 fixtureA = 0, fixtureB = 0;
    u32 flags = 0;                   // e_islandFlag 1, e_touchingFlag 2 (contact.h)
    u32 manifoldCapacity = 1;
    u32 manifoldCount = 0;
    Manifold manifolds[kMaxManifolds];
    FeatureCache cache;
    bool isMesh = false;
    // mesh contact state (mesh_contact.h)
    AABB meshAABB;
    bool meshAABBMoved = false;
    std::vector<u32> triangles;
};

// b3RevoluteJoint (revolute_joint.h/.cpp): definition, warm-start state and solver temporaries.
enum LimitState : u32 { e_inactiveLimit = 0, e_atLowerLimit = 1, e_atUpperLimit = 2, e_equalLimits = 3 };
struct Joint {
    u32 type = 0;
    u32 bodyA = 0, bodyB = 0;
    bool collideConnected = false;
    bool islandFlag = false;
    Vec3 localAnchorA, localAnchorB;
    Quat localRotationA, localRotationB, referenceRotation;
    bool enableLimit = false, enableMotor = false;
    scalar lowerAngle = 0, upperAngle = 0, motorSpeed = 0, maxMotorTorque = 0;
    // warm-start state
    scalar motorImpulse = 0, limitImpulse = 0;
    u32 limitState = e_inactiveLimit;
    Vec3 linearImpulse;
    Vec2 angularImpulse;
    // per-solve temporaries
    u32 indexA = 0, indexB = 0;
    scalar mA = 0, mB = 0;
    Vec3 localCenterA, localCenterB;
    Mat33 localInvIA, localInvIB, iA, iB;
    Vec3 motorAxis;
    scalar motorMass = 0;
    Vec3 rA, rB;
    Mat33 linearMass;
    Vec3 a1, a2;
    Mat22 angularMass;
    // the other joint types (sphere, weld, spring, mouse, friction, motor): definitions, carried
    // impulses and per-solve values
    scalar frequencyHz = 0, dampingRatio = 0, length = 0, maxForce = 0, maxTorque = 0, correctionFactor = 0;
    Vec3 target, linearOffset;
    Quat angularOffset;
    Vec3 angularImpulse3;       // weld, friction, motor
    scalar impulse1 = 0;        // spring
    Mat33 gamma3, angularMass3;
    Vec3 bias3, n, linearError, angularVelocity;
    scalar mass1 = 0, gamma1 = 0, bias1 = 0;
    // cone, prismatic and wheel joints (cone_joint.h, prismatic_joint.h, wheel_joint.h): motorAxis
    // holds the cone axis, the prismatic axis or the wheel's axis on A; mass1 the cone or spring
    // mass; limitImpulse/limitState the cone and prismatic limits; impulse1 the wheel's angular
    // impulse; angularImpulse3, angularMass3 and bias3 the prismatic angular row
    Vec3 localXAxisA, localYAxisA, localZAxisA, localXAxisB;
    scalar coneAngle = 0, lowerTranslation = 0, upperTranslation = 0, referenceCosine = 0, referenceAngle = 0;
    Vec2 linearImpulse2;
    scalar springImpulse = 0, angularMass1 = 0;
    Vec3 perp1, perp2, s1, s2, s3, s4, axisB, u;
    Mat22 mass22;
};

struct Position { Vec3 x; Quat q; Mat33 I; };
struct Velocity { Vec3 v, w; };

struct TimeStep {
    scalar dt, inv_dt, dtRatio;
    u32 velocityIterations, positionIterations;
    bool warmStarting;
};

struct World {
    Vec3 gravity;
    bool sleeping = false, warmStarting = true, clearForces = true, newContacts = false;
    scalar inv_dt0 = 0;
    std::vector<Hull> hulls;
    std::vector<Mesh> meshes;
    std::vector<Body> bodies;        // creation order; body list order is reverse creation
    std::vector<Fixture> fixtures;   // creation order
    std::vector<Joint> joints;       // creation order; joint list order is reverse creation
    // broadphase (broad_phase.cpp, dynamic_tree.cpp)
    std::vector<AABB> fatAABB;       // indexed by proxy id
    std::vector<u32> proxyFixture;   // proxy id -> fixture, kNullU32 for internal nodes
    std::vector<u32> moveBuffer;
    std::vector<u32> freeStack;      // freed node ids (LIFO; kNullU32: an internal node whose id is not
                                     // tracked), then sequential ids from nextNode
    u32 nextNode = 0;
    u32 leafCount = 0;
    std::vector<std::pair<u32, u32>> lastPairs;  // unique sorted pairs of the last FindPairs
    // contacts: world contact list, front = last created (contact_manager.cpp:105-113)
    std::vector<Contact*> contactList;
    std::vector<Contact*> lastIslandOrder;
    u32 lastIslandCount = 0;
    // contact listener events of the last step (b3ContactListener, world_callbacks.h); `contact` is
    // null for a contact the step destroyed
    struct Event { u32 type, fixtureA, fixtureB; Contact* contact; };
    bool recordEvents = false;
    std::vector<Event> events;
    // contact filter tables (b3ContactFilter): bit i * n + j over fixtures in creation order; empty = none
    std::vector<uint8_t> filterCollide, filterRespond;
    bool FilterAllows(const std::vector<uint8_t>& t, u32 i, u32 j) const {
        if (t.empty()) return true;
        const size_t k = size_t(i) * fixtures.size() + j;
        return (t[k >> 3] >> (k & 7)) & 1u;
    }

    ~World();
    u32 AllocateNode();
    u32 CreateProxy(const AABB& aabb, u32 fixture);
    void DestroyProxy(u32 proxy);
    void MoveProxy(u32 proxy, const AABB& aabb, const Vec3& displacement);
    void SynchronizeFixture(u32 f, const Transform& xf1, const Transform& xf2);
    void SynchronizeFixtures(u32 b);
    void ResetMass(u32 b);
    bool ShouldCollide(u32 a, u32 b) const;
    void FindNewContacts();
    void AddPair(u32 fA, u32 fB);
    void UpdateContacts();
    void DestroyContact(Contact* c);
    void UpdateContact(Contact* c);
    void Collide(Contact* c);
    void Solve(const TimeStep& step);
    void SolveIsland(const TimeStep& step, std::vector<u32>& bodies, std::vector<Contact*>& contacts,
                     std::vector<u32>& joints);
    void Step(scalar dt, u32 vit, u32 pit);
};

}  // namespace orc
