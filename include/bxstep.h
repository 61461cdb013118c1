/*
 * bxstep.h -- C ABI of the MI355X batched rigid-body step (libbxstep.so).
 *
 * This is the one boundary between host code and the HIP step. It replaces the
 * body of the reference's
 *     void b3World::Step(scalar dt, uint32 velocityIterations, uint32 positionIterations)
 * (include/bounce/dynamics/world.h:127, src/bounce/dynamics/world.cpp:187-242)
 * for a batch of independent worlds. The C++ drop-in API in bounce_amd/include/bounce/
 * (b3World, b3Body, b3Fixture, b3Shape, b3Hull, b3RevoluteJoint ...) builds these
 * descriptors and calls bx_batch_step() from b3World::Step.
 *
 * Conventions
 *  - Plain C: POD structs, pointers and sizes, no exceptions, no C++ types.
 *  - Every entry point returns BX_OK (0) or a negative BX_E_* code; the text of the last
 *    error on the calling thread is available from bx_last_error().
 *  - The caller owns every array passed in; the library copies what it keeps during the call.
 *  - A context owns one device, one HIP stream and all device memory of its batches.
 *    A context is used by one host thread at a time.
 *  - Orientation quaternions are stored (x, y, z, w) = (v.x, v.y, v.z, s).
 *  - Bodies, fixtures and joints are listed in creation order; the reference semantics of
 *    list order (bodies and fixtures are prepended to their lists, world.cpp:94-110,
 *    body.cpp:105-136) are applied by the library.
 */
#ifndef BXSTEP_H
#define BXSTEP_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------------
 * Status codes
 * ------------------------------------------------------------------------------------------- */
#define BX_OK 0
#define BX_E_INVALID (-1)      /* bad argument or descriptor */
#define BX_E_DEVICE (-2)       /* HIP runtime error */
#define BX_E_CAPACITY (-3)     /* a fixed capacity was exceeded (see bx_batch_opts) */
#define BX_E_UNSUPPORTED (-4)  /* feature not supported by the device path */
#define BX_E_NOMEM (-5)        /* device or host allocation failed */

/* ---------------------------------------------------------------------------------------------
 * Enumerations (values equal the reference's enums)
 * ------------------------------------------------------------------------------------------- */
/* b3Shape::Type (include/bounce/collision/shapes/shape.h) */
#define BX_SHAPE_SPHERE 0u
#define BX_SHAPE_CAPSULE 1u
#define BX_SHAPE_TRIANGLE 2u
#define BX_SHAPE_HULL 3u
#define BX_SHAPE_MESH 4u

/* b3BodyType (include/bounce/dynamics/body.h) */
#define BX_BODY_STATIC 0u
#define BX_BODY_KINEMATIC 1u
#define BX_BODY_DYNAMIC 2u

/* body flags: b3Body::e_* values (include/bounce/dynamics/body.h:379-384) */
#define BX_BODY_AWAKE 0x0001u
#define BX_BODY_ISLAND 0x0002u
#define BX_BODY_AUTO_SLEEP 0x0004u
#define BX_BODY_FIXED_ROT_X 0x0008u
#define BX_BODY_FIXED_ROT_Y 0x0010u
#define BX_BODY_FIXED_ROT_Z 0x0020u

/* world flags */
#define BX_WORLD_SLEEPING 0x0001u      /* b3World::SetSleeping(true); default off (world.cpp:46) */
#define BX_WORLD_NO_WARM_START 0x0002u /* b3World::SetWarmStart(false); default on */
#define BX_WORLD_NO_CLEAR_FORCES 0x0004u
#define BX_WORLD_EVENTS 0x0008u        /* record the contact listener's events of every step (bx_batch_get_events);
                                          compiled in only with -DBX_DEVICE_CALLBACKS=1: the GPU library as
                                          built refuses such a world with BX_E_UNSUPPORTED */

/* b3JointType (include/bounce/dynamics/joints/joint.h:33-46); cone, prismatic and wheel joints are
   not supported by the device path (BX_E_UNSUPPORTED) */
#define BX_JOINT_MOUSE 1u
#define BX_JOINT_SPRING 2u
#define BX_JOINT_WELD 3u
#define BX_JOINT_REVOLUTE 4u
#define BX_JOINT_SPHERE 5u
#define BX_JOINT_CONE 6u
#define BX_JOINT_FRICTION 7u
#define BX_JOINT_MOTOR 8u
#define BX_JOINT_PRISMATIC 9u
#define BX_JOINT_WHEEL 10u

#define BX_NULL_INDEX 0xFFFFFFFFu
#define BX_MAX_MANIFOLD_POINTS 4
#define BX_MAX_MANIFOLDS 3

/* ---------------------------------------------------------------------------------------------
 * Shape library (shared by all worlds of a batch; geometry is borrowed by index, as b3HullShape
 * and b3MeshShape borrow their b3Hull / b3Mesh, hull_shape.h:42-43, mesh_shape.h:49-50)
 * ------------------------------------------------------------------------------------------- */
typedef struct bx_hull_desc {
    uint32_t vertex_count;
    uint32_t edge_count;      /* half-edges; twins are adjacent (2k, 2k+1) */
    uint32_t face_count;
    const float* vertices;    /* [vertex_count][3] */
    const uint32_t* edges;    /* [edge_count][5]: origin, twin, face, prev, next (b3HalfEdge) */
    const uint32_t* faces;    /* [face_count]: first half-edge of the face (b3Face) */
    const float* planes;      /* [face_count][4]: normal x y z, offset (b3Plane) */
    float centroid[3];
} bx_hull_desc;

typedef struct bx_mesh_desc {
    uint32_t vertex_count;
    uint32_t triangle_count;
    const float* vertices;      /* [vertex_count][3] */
    const uint32_t* triangles;  /* [triangle_count][3] vertex indices (b3MeshTriangle) */
} bx_mesh_desc;

typedef struct bx_shape_library {
    uint32_t hull_count;
    const bx_hull_desc* hulls;
    uint32_t mesh_count;
    const bx_mesh_desc* meshes;
} bx_shape_library;

/* ---------------------------------------------------------------------------------------------
 * World descriptors (b3BodyDef body.h:49-123, b3FixtureDef fixture.h:37-70,
 * b3RevoluteJointDef revolute_joint.h:25-78)
 * ------------------------------------------------------------------------------------------- */
typedef struct bx_body_desc {
    uint32_t type;                 /* BX_BODY_* */
    uint32_t flags;                /* BX_BODY_AWAKE | BX_BODY_AUTO_SLEEP | BX_BODY_FIXED_ROT_* */
    float position[3];
    float orientation[4];          /* x y z w */
    float linear_velocity[3];
    float angular_velocity[3];
    float linear_damping[3];
    float angular_damping[3];
    float gravity_scale[3];
    float linear_sleep_tolerance;
    float angular_sleep_tolerance;
} bx_body_desc;

typedef struct bx_fixture_desc {
    uint32_t body;                 /* creation index of the owning body */
    uint32_t shape_type;           /* BX_SHAPE_* */
    uint32_t shape_index;          /* hull or mesh index in the library */
    uint32_t is_sensor;
    float radius;                  /* sphere / capsule radius; hulls use the hull skin (0) */
    float vertex1[3];              /* sphere centre, capsule / triangle vertex 1 */
    float vertex2[3];
    float vertex3[3];
    float mesh_scale[3];           /* b3MeshShape::m_scale */
    float density;
    float friction;
    float restitution;
} bx_fixture_desc;

typedef struct bx_joint_desc {
    uint32_t type;                 /* BX_JOINT_* */
    uint32_t body_a;               /* creation indices */
    uint32_t body_b;
    uint32_t collide_connected;
    /* If use_initialize != 0 the local frames are computed from the bodies' transforms at creation,
       as the def's Initialize does:
         revolute  b3RevoluteJointDef::Initialize(A, B, axis, anchor, lower_angle, upper_angle)
         sphere    b3SphereJointDef::Initialize(A, B, anchor)         (sphere_joint.cpp:23-29)
         weld      b3WeldJointDef::Initialize(A, B, anchor)           (weld_joint.cpp:23-33)
         spring    b3SpringJointDef::Initialize(A, B, anchor, anchor_b) (spring_joint.cpp:23-30)
         friction  b3FrictionJointDef::Initialize(A, B, anchor)      (friction_joint.cpp:23-29)
         motor     b3MotorJointDef::Initialize(A, B)                  (motor_joint.cpp:23-31)
         cone      b3ConeJointDef::Initialize(A, B, axis, anchor, cone_angle) (cone_joint.cpp:35-55)
         prismatic b3PrismaticJointDef::Initialize(A, B, anchor, axis) (prismatic_joint.cpp:23-35)
         wheel     b3WheelJointDef::Initialize(A, B, anchor, axis, axis_b) (wheel_joint.cpp:22-31)
       otherwise the local_*, reference_rotation, length and offsets are used as given. A mouse
       joint's body-B anchor is always the target in body B's frame at creation
       (mouse_joint.cpp:22-31). */
    uint32_t use_initialize;
    float axis[3];
    float anchor[3];
    float local_anchor_a[3];
    float local_rotation_a[4];
    float local_anchor_b[3];
    float local_rotation_b[4];
    float reference_rotation[4];
    uint32_t enable_limit;
    float lower_angle;
    float upper_angle;
    uint32_t enable_motor;
    float motor_speed;
    float max_motor_torque;
    /* the other joint types' definitions (b3*JointDef members) */
    float anchor_b[3];             /* spring Initialize: the anchor on body B */
    float frequency_hz;            /* weld, spring, mouse */
    float damping_ratio;           /* weld, spring, mouse */
    float length;                  /* spring rest length */
    float max_force;               /* mouse, friction, motor */
    float max_torque;              /* friction, motor */
    float correction_factor;       /* weld, motor */
    float target[3];               /* mouse: the world target */
    float linear_offset[3];        /* motor: position of B in A's frame */
    float angular_offset[4];       /* motor: orientation of B in A's frame (x y z w) */
    /* cone, prismatic and wheel joints (cone_joint.h, prismatic_joint.h, wheel_joint.h) */
    float cone_angle;              /* cone: the cone's full angle (enable_limit: enableConeLimit) */
    float local_axis_a[3];         /* prismatic, wheel: the axis in A's frame */
    float local_axis_b[3];         /* wheel: the axis in B's frame */
    float lower_translation;       /* prismatic limit (enable_limit); max_force is maxMotorForce */
    float upper_translation;
    float axis_b[3];               /* wheel Initialize: the world axis on B (axis: on A) */
} bx_joint_desc;

typedef struct bx_world_desc {
    float gravity[3];              /* default (0, -9.8, 0), world.cpp:51 */
    uint32_t flags;                /* BX_WORLD_* */
    uint32_t body_count;
    const bx_body_desc* bodies;
    uint32_t fixture_count;
    const bx_fixture_desc* fixtures;
    uint32_t joint_count;
    const bx_joint_desc* joints;
} bx_world_desc;

/* ---------------------------------------------------------------------------------------------
 * State read back after stepping
 * ------------------------------------------------------------------------------------------- */
typedef struct bx_body_state {
    float position[3];             /* b3Body::GetTransform().translation */
    float orientation[4];          /* b3Body::GetTransform().rotation, x y z w */
    float world_center[3];         /* b3Sweep::worldCenter */
    float linear_velocity[3];
    float angular_velocity[3];
    float sleep_time;
    uint32_t flags;                /* BX_BODY_AWAKE | BX_BODY_ISLAND ... after the step */
    uint32_t island_index;         /* b3Body::m_islandIndex (position in its last island) */
} bx_body_state;

typedef struct bx_manifold_point_state {
    float local_normal1[3];
    float local_point1[3];
    float local_point2[3];
    float normal_impulse;
    uint32_t triangle_key;         /* b3ContactID::triangleKey */
    uint32_t key[4];               /* b3FeaturePair: inEdge1 outEdge1 inEdge2 outEdge2 */
    uint32_t persisting;
    uint32_t edge_contact;
} bx_manifold_point_state;

typedef struct bx_manifold_state {
    uint32_t point_count;
    bx_manifold_point_state points[BX_MAX_MANIFOLD_POINTS];
    float tangent_impulse[2];
    float motor_impulse;
} bx_manifold_state;

typedef struct bx_contact_state {
    uint32_t fixture_a;            /* creation index of b3Contact::GetFixtureA() */
    uint32_t fixture_b;
    uint32_t touching;
    uint32_t manifold_count;
    uint32_t cache[4];             /* b3FeatureCache: cacheType, featureType, index1, index2
                                      (hull-hull contacts; zero otherwise) */
    bx_manifold_state manifolds[BX_MAX_MANIFOLDS];
} bx_contact_state;

typedef struct bx_pair {
    uint32_t proxy1;               /* b3Pair after the last FindPairs: proxy1 <= proxy2 */
    uint32_t proxy2;
} bx_pair;

typedef struct bx_world_counts {
    uint32_t body_count;
    uint32_t fixture_count;
    uint32_t joint_count;
    uint32_t contact_count;        /* b3World::GetContactCount */
    uint32_t touching_count;
    uint32_t pair_count;           /* unique pairs of the last FindPairs */
    uint32_t island_count;         /* islands solved in the last step */
    uint32_t awake_count;
    uint32_t status;               /* BX_WS_* bits */
} bx_world_counts;

/* per-world status bits */
#define BX_WS_NONFINITE 0x1u       /* a non-finite body state was produced */
#define BX_WS_CAPACITY 0x2u        /* a per-world capacity overflowed */

/* ---------------------------------------------------------------------------------------------
 * Device path
 * ------------------------------------------------------------------------------------------- */
typedef struct bx_context bx_context;
typedef struct bx_batch bx_batch;

typedef struct bx_step_cfg {
    float dt;                      /* b3World::Step(dt, velocityIterations, positionIterations) */
    uint32_t velocity_iterations;
    uint32_t position_iterations;
} bx_step_cfg;

typedef struct bx_batch_opts {
    uint32_t max_contacts_per_world;   /* 0: 16 x fixtures + 64 */
    uint32_t max_pairs_per_world;      /* 0: 2 x max_contacts_per_world */
    uint32_t solver_mode;              /* velocity-sweep layout: 1, 2, 3 = 16, 32, 64 islands per
                                          wavefront, one lane each; 4, 5 = 4, 8 islands per
                                          wavefront, eight lanes each (the contact's vector
                                          arithmetic split over components and bodies); 0: tuned
                                          on the batch's steps 4-13 (each layout timed twice, in
                                          the order 1 2 3 4 5 5 4 3 2 1), the one with the smallest
                                          sample is kept; mode 1 unless another is faster by more
                                          than 5% */
    uint32_t mesh_triangles_per_contact; /* triangles a mesh contact can overlap (0: 64) */
    uint32_t reserved[4];
} bx_batch_opts;

typedef struct bx_pool_opts {
    uint64_t hbm_budget_bytes;         /* device memory for the resident worlds' state, allocated once */
    uint32_t max_worlds;               /* resident world slots (0: 4096) */
    uint32_t max_contacts_per_world;   /* 0: 16 x fixtures + 64 */
    uint32_t max_pairs_per_world;      /* 0: 2 x max_contacts_per_world */
    uint32_t joints_per_1000_bodies;   /* share of the budget kept for joints (0: 50) */
    uint32_t solver_mode;              /* velocity-sweep layout as bx_batch_opts.solver_mode 1..5;
                                          0: mode 1 (pools are not tuned: their world mix changes
                                          from tick to tick) */
    uint32_t reserved[3];
} bx_pool_opts;

typedef struct bx_pool_stats {
    uint64_t capacity_bytes;           /* per-world state the pool can hold */
    uint64_t used_bytes;               /* held by resident worlds */
    uint32_t resident_worlds;
    uint32_t free_slots;
    uint32_t world_slots;
    uint32_t reserved[5];
} bx_pool_stats;

/* Totals over every resident world of a batch or pool after the last step: the counts the step's
   compulsory-traffic model is written in, and every world's status. */
typedef struct bx_batch_totals {
    uint64_t bodies;                   /* all bodies */
    uint64_t awake_bodies;             /* awake non-static bodies */
    uint64_t fixtures;                 /* fixtures = broadphase proxies */
    uint64_t joints;
    uint64_t contacts;                 /* live contacts (b3World::GetContactCount summed) */
    uint64_t mesh_world_contacts;      /* of those, contacts of worlds that hold a mesh fixture */
    uint64_t pairs;                    /* unique pairs of the last pair search */
    uint64_t islands;                  /* islands solved in the last step */
    uint64_t island_contacts;          /* contact constraints solved in the last step */
    uint32_t worlds;                   /* resident worlds */
    uint32_t status_worlds;            /* worlds whose status is nonzero */
    uint32_t status_or;                /* OR of every world's status */
    uint32_t reserved;
} bx_batch_totals;

/* Library version and build information ("bxstep gfx950 ..."). */
const char* bx_version(void);

/* Text of the last error on the calling thread (empty string if none). */
const char* bx_last_error(void);

/* Create a context on HIP device `device` with its own stream. */
int bx_context_create(int device, bx_context** out);
void bx_context_destroy(bx_context* ctx);

/* Upload `world_count` worlds built from the descriptors into device memory.
   The shape library is copied; descriptors may be freed after the call. */
int bx_batch_create(bx_context* ctx, const bx_shape_library* library,
                    const bx_world_desc* worlds, uint32_t world_count,
                    const bx_batch_opts* opts, bx_batch** out);
void bx_batch_destroy(bx_batch* batch);

/* Advance every world of the batch by `step_count` steps of b3World::Step(cfg).
   Asynchronous on the context stream; call bx_batch_sync before reading state. */
int bx_batch_step(bx_batch* batch, const bx_step_cfg* cfg, uint32_t step_count);

/* The same steps as bx_batch_step, for callers that cannot allocate (code under HIP graph capture,
   tight serving loops, harnesses that forbid allocation). No counterpart in the reference. It calls no
   hipMalloc, hipFree or other allocating routine, creates no events and times nothing: the launch's
   island work list lives in a device array of the library (16 MiB per device, held until the process
   ends), so a batch may hold at most 2,097,152 bodies (BX_E_CAPACITY beyond). Calls that overlap on one
   device share that array and race. Batches only (a pool returns BX_E_UNSUPPORTED); the velocity-sweep
   layout is the batch's current one (no tuning). */
int bx_batch_step_noalloc(bx_batch* batch, const bx_step_cfg* cfg, uint32_t step_count);

/* The same steps as bx_batch_step (b3World::Step, world.cpp:187-242, step_count times) with the same
   results, each step a replay of one step captured as a HIP graph on the first call (captured
   again when cfg changes), so a step costs one graph launch on the host. The steps are not timed
   per kernel (bx_batch_get_timings and bx_batch_get_kernel_timings then report no timed step).
   Batches only: a pool returns BX_E_UNSUPPORTED. */
int bx_batch_step_graph(bx_batch* batch, const bx_step_cfg* cfg, uint32_t step_count);

/* The velocity-sweep layout in use: islands per wavefront (16 until solver_mode 0 has timed every
   layout), and the time of k_solve_velocity in ms measured for each layout by solver_mode 0 (the
   smaller of its two samples; 0 before the tuning steps) in tuned_ms[0..count). Returns the number of layouts (5:
   16, 32, 64 islands per wavefront with one lane each, 4 and 8 with eight lanes each). No counterpart
   in the reference, whose solver runs one island at a time on the CPU (island.cpp:154-181). */
int bx_batch_get_solver_layout(bx_batch* batch, uint32_t* islands_per_wavefront, float* tuned_ms, uint32_t count);

/* Advance world worlds[i] by step_counts[i] steps (1 each when step_counts is NULL). Step s launches
   once for every listed world with more than s steps to go, so the worlds share each launch. */
int bx_batch_step_subset(bx_batch* batch, const bx_step_cfg* cfg, const uint32_t* worlds,
                         const uint32_t* step_counts, uint32_t count);

/* One step of each listed world (distinct, resident) in one launch sequence: the unit of
   continuous batching. A world's result does not depend on which other worlds share the launch. */
int bx_batch_step_list(bx_batch* batch, const bx_step_cfg* cfg, const uint32_t* worlds, uint32_t count);

/* ---------------------------------------------------------------------------------------------
 * Pools (serving): a batch with world slots and a fixed device-memory budget, allocated once.
 * Worlds are admitted into free slots while their state fits the budget, stepped in any mix with
 * bx_batch_step_list, read with the bx_batch_get_* calls (world = slot) and released. The pool's
 * memory is split into index spaces (bodies, fixtures, joints, contact slots, pairs, island
 * entries, island-builder words); a world holds one range in each.
 * ------------------------------------------------------------------------------------------- */
int bx_pool_create(bx_context* ctx, const bx_shape_library* library, const bx_pool_opts* opts,
                   bx_batch** out);
/* Admit a world (built as bx_batch_create would build it) into a free slot. BX_E_CAPACITY when no
   slot or not enough of the budget is free now; the pool is unchanged then. */
int bx_pool_admit(bx_batch* pool, const bx_world_desc* world, uint32_t* slot);
/* Free a slot and its memory. Launches already enqueued that list the slot still run correctly. */
int bx_pool_release(bx_batch* pool, uint32_t slot);
/* Device bytes a world would hold while resident in this pool (NULL pool: default capacities). */
uint64_t bx_pool_world_bytes(bx_batch* pool, const bx_world_desc* world);
int bx_pool_get_stats(bx_batch* pool, bx_pool_stats* out);

/* Whole device state of a resident world as a relocatable byte image (eviction and resume): with
   out == NULL only *size is written. bx_pool_restore admits an image into a free slot of a pool
   built with the same shape library; stepping continues exactly as if the world had stayed. */
int bx_batch_snapshot(bx_batch* batch, uint32_t world, void* out, uint64_t capacity, uint64_t* size);
int bx_pool_restore(bx_batch* pool, const void* snapshot, uint64_t size, uint32_t* slot);

int bx_batch_sync(bx_batch* batch);

/* Read back one world. Arrays must hold the counts given by bx_batch_get_counts. */
int bx_batch_get_counts(bx_batch* batch, uint32_t world, bx_world_counts* out);
/* Contact listener events of the last step (b3ContactListener, world_callbacks.h), in the order the
   reference calls them: BeginContact / EndContact / PreSolve as b3ContactManager::UpdateContacts visits
   the contacts (contact_manager.cpp:165-237, contact.cpp:268-284; EndContact also for a touching contact
   it destroys), then PostSolve for every solved contact, island by island (island.cpp:309-320). The
   replay runs after the step, so a listener reading a contact in BeginContact or PreSolve sees its
   state after the step, and changes it makes there do not reach the solver. */
#define BX_EVENT_BEGIN_CONTACT 1
#define BX_EVENT_END_CONTACT 2
#define BX_EVENT_PRE_SOLVE 3
#define BX_EVENT_POST_SOLVE 4
typedef struct bx_contact_event {
    uint32_t type;                 /* BX_EVENT_* */
    uint32_t fixture_a;            /* creation indices of b3Contact::GetFixtureA() / GetFixtureB() */
    uint32_t fixture_b;
    uint32_t contact;              /* position in bx_batch_get_contacts' list after the step, BX_NULL_INDEX
                                      for a contact the step destroyed */
} bx_contact_event;
/* Events of the last step of a world created with BX_WORLD_EVENTS; `count` receives the total.
   Returns BX_E_UNSUPPORTED in a library built without -DBX_DEVICE_CALLBACKS=1 (the GPU library). */
int bx_batch_get_events(bx_batch* batch, uint32_t world, bx_contact_event* out, uint32_t capacity,
                        uint32_t* count);

/* A contact filter (b3ContactFilter, world_callbacks.h; b3World::SetContactFilter, world.h:77-81) as two
   tables over the world's fixtures in creation order: bit i * n + j (n = fixture count, bit k at
   byte k / 8, bit k % 8) of `collide` is ShouldCollide(fixture i, fixture j), of `respond`
   ShouldRespond(fixture i, fixture j). The device consults them where the reference calls the filter:
   when a pair becomes a contact (contact_manager.cpp:86-92), when a contact is updated
   (contact_manager.cpp:189-199) and when the island search meets a touching contact (world.cpp:351-357),
   so a filter whose answers depend only on the two fixtures is reproduced exactly. Null tables remove
   the filter. Batches only; at most 16,384 fixtures. Returns BX_E_UNSUPPORTED in a library built
   without -DBX_DEVICE_CALLBACKS=1 (the GPU library). */
int bx_batch_set_contact_filter(bx_batch* batch, uint32_t world, const uint8_t* collide, const uint8_t* respond);

/* bx_batch_totals of the whole batch or pool, from one read of the world states and bodies. */
int bx_batch_get_totals(bx_batch* batch, bx_batch_totals* out);
int bx_batch_get_bodies(bx_batch* batch, uint32_t world, bx_body_state* out, uint32_t capacity);
int bx_batch_get_contacts(bx_batch* batch, uint32_t world, bx_contact_state* out, uint32_t capacity);
int bx_batch_get_pairs(bx_batch* batch, uint32_t world, bx_pair* out, uint32_t capacity);
/* Contacts in island solve order (indices into the bx_batch_get_contacts list) for the last step. */
int bx_batch_get_island_contacts(bx_batch* batch, uint32_t world, uint32_t* out, uint32_t capacity,
                                 uint32_t* count);

/* Mass data of every body in creation order (b3Body::GetMass, GetInverseMass, GetLocalCenter,
   GetInertia, GetWorldInverseInertia; body.h:237-246, 193): what ResetMass computed from the
   fixtures (body.cpp:265-376). Matrices are b3Mat33 columns x, y, z. */
typedef struct bx_body_mass {
    float mass;
    float inv_mass;
    float local_center[3];         /* b3Sweep::localCenter */
    float inertia[9];              /* about the centre of mass, body frame */
    float inv_inertia[9];
    float world_inv_inertia[9];
} bx_body_mass;
int bx_batch_get_mass(bx_batch* batch, uint32_t world, bx_body_mass* out, uint32_t capacity);

/* Set every body's linear and angular velocity between steps, as b3Body::SetLinearVelocity and
   SetAngularVelocity (body.h:652-684) would for each body in creation order: static bodies are left
   alone, and a nonzero velocity wakes the body (b3Body::SetAwake(true), body.h:588-607). `bodies` is in
   creation order with `count` == body_count of the world; only the velocities are read. */
int bx_batch_set_bodies(bx_batch* batch, uint32_t world, const bx_body_state* bodies, uint32_t count);

/* Body edits between steps (b3Body's setters and force/impulse methods), applied in order. */
#define BX_EDIT_LINEAR_VELOCITY 1   /* b3Body::SetLinearVelocity(v)              body.h:652-665 */
#define BX_EDIT_ANGULAR_VELOCITY 2  /* b3Body::SetAngularVelocity(v)             body.h:672-685 */
#define BX_EDIT_FORCE 3             /* b3Body::ApplyForce(v, p, wake)            body.h:727-744 */
#define BX_EDIT_FORCE_TO_CENTER 4   /* b3Body::ApplyForceToCenter(v, wake)       body.h:746-762 */
#define BX_EDIT_TORQUE 5            /* b3Body::ApplyTorque(v, wake)              body.h:764-780 */
#define BX_EDIT_LINEAR_IMPULSE 6    /* b3Body::ApplyLinearImpulse(v, p, wake)    body.h:782-799 */
#define BX_EDIT_ANGULAR_IMPULSE 7   /* b3Body::ApplyAngularImpulse(v, wake)      body.h:801-817 */
#define BX_EDIT_TRANSFORM 8         /* b3Body::SetTransform({v, p})              body.cpp:194-214 */
#define BX_EDIT_AWAKE 9             /* b3Body::SetAwake(wake)                    body.h:588-607 */
#define BX_EDIT_SLEEPING_ALLOWED 10 /* b3Body::SetSleepingAllowed(wake)          body.h:829-840 */
#define BX_EDIT_GRAVITY_SCALE 11    /* b3Body::SetGravityScale(v)                body.h:634-640 */
#define BX_EDIT_LINEAR_DAMPING 12   /* b3Body::SetLinearDamping(v)               body.h:614-617 */
#define BX_EDIT_ANGULAR_DAMPING 13  /* b3Body::SetAngularDamping(v)              body.h:624-627 */
#define BX_EDIT_LINEAR_SLEEP_TOLERANCE 14  /* b3Body::SetLinearSleepTolerance(v[0])  body.cpp:550-554 */
#define BX_EDIT_ANGULAR_SLEEP_TOLERANCE 15 /* b3Body::SetAngularSleepTolerance(v[0]) body.cpp:556-560 */
#define BX_EDIT_FIXED_ROTATION 16   /* b3Body::SetFixedRotation(v[0] != 0, v[1] != 0, v[2] != 0)
                                                                                 body.cpp:507-548 */
/* World edits in the same list (`body` is not read), so that they keep their order among the body
   edits: */
#define BX_EDIT_WORLD_GRAVITY 32           /* b3World::SetGravity(v)             world.h:255-258 */
#define BX_EDIT_WORLD_SLEEPING 33          /* b3World::SetSleeping(wake): clearing it wakes every
                                              body                               world.cpp:82-92 */
#define BX_EDIT_WORLD_WARM_STARTING 34     /* b3World::SetWarmStarting(wake)     world.h:95 */
#define BX_EDIT_WORLD_AUTO_CLEAR_FORCES 35 /* b3World::SetAutoClearForces(wake)  world.h:265-268 */
#define BX_EDIT_WORLD_CLEAR_FORCES 36      /* b3World::ClearForces()             world.cpp:244-251 */

typedef struct bx_body_edit {
    uint32_t body;                 /* creation index */
    uint32_t op;                   /* BX_EDIT_* */
    uint32_t wake;                 /* `wake` of the Apply* methods; the flag of SetAwake, SetSleepingAllowed
                                      and the world's Set* edits */
    uint32_t reserved;
    float v[3];                    /* velocity, force, torque, impulse, translation, scale or damping */
    float p[4];                    /* point of application (x y z), or rotation (x y z w) of SetTransform */
} bx_body_edit;

/* Apply `count` edits to bodies of `world`, in order, as the b3Body calls named above would between two
   b3World::Step calls. SetTransform also moves the body's fixture proxies (b3Fixture::Synchronize with
   the new transform at both ends) and makes the next step begin with the world's new-contact search,
   as the reference's m_newContacts does. Synchronous. */
int bx_batch_edit_bodies(bx_batch* batch, uint32_t world, const bx_body_edit* edits, uint32_t count);

/* ---------------------------------------------------------------------------------------------
 * Topology changes between steps, keeping the world's state: contacts with their manifolds,
 * impulses and feature caches, joint impulses, proxies and their fat AABBs, sleep state and the
 * previous step's inverse time step all carry over, as they do in the reference.
 *   BX_TOPO_CREATE_BODY     b3World::CreateBody (world.cpp:94-110): adds->bodies[index] becomes the
 *                           newest body (creation index = the world's body count at that point)
 *   BX_TOPO_CREATE_FIXTURE  b3Body::CreateFixture (body.cpp:105-136): adds->fixtures[index], whose
 *                           `body` is a creation index at that point; new proxy, fat AABB from the
 *                           body's current transform, ResetMass when the density is positive, and
 *                           the next step begins with the new-contact search
 *   BX_TOPO_CREATE_JOINT    b3World::CreateJoint (joint_manager.cpp:30-73): adds->joints[index]
 *                           (Initialize frames from the bodies' current transforms); no body wakes
 *   BX_TOPO_DESTROY_JOINT   b3World::DestroyJoint (joint_manager.cpp:75-141): wakes both bodies
 *   BX_TOPO_DESTROY_FIXTURE b3Body::DestroyFixture (body.cpp:138-191): its contacts and proxy go,
 *                           then ResetMass
 *   BX_TOPO_DESTROY_BODY    b3World::DestroyBody (world.cpp:112-175): its joints (waking both
 *                           bodies of each), its fixtures' contacts and proxies, then the body
 * For a destroy, `index` is the creation index at that point; later creation indices move down
 * by one, as they do in a list that loses an element. Operations apply in order. Proxy ids follow
 * b3DynamicTree's node free list (dynamic_tree.cpp:44-136), so pair order and contact creation
 * order match the reference; BX_E_UNSUPPORTED when a new proxy would take a node whose id depends
 * on the dynamic tree's shape (an internal node freed by a destroy).
 * ------------------------------------------------------------------------------------------- */
#define BX_TOPO_CREATE_BODY 1
#define BX_TOPO_CREATE_FIXTURE 2
#define BX_TOPO_CREATE_JOINT 3
#define BX_TOPO_DESTROY_BODY 4
#define BX_TOPO_DESTROY_FIXTURE 5
#define BX_TOPO_DESTROY_JOINT 6

typedef struct bx_topology_op {
    uint32_t op;                   /* BX_TOPO_* */
    uint32_t index;                /* descriptor index in `adds` (create) or creation index (destroy) */
} bx_topology_op;

/* Apply `op_count` topology operations to world `world` of a batch (not a pool). The batch keeps its
   handle; its device arrays are rebuilt with the world's new capacities and every other world
   carried over unchanged. `library` (or NULL to keep the batch's) must begin with the batch's
   hulls and meshes, so existing indices stay valid. Synchronous; on an error the batch is
   unchanged. */
int bx_batch_edit_topology(bx_batch* batch, uint32_t world, const bx_shape_library* library,
                           const bx_world_desc* adds, const bx_topology_op* ops, uint32_t op_count);

/* ---------------------------------------------------------------------------------------------
 * World queries between steps, many at once over any worlds of a batch or pool: the rays and
 * boxes of b3World::RayCast, RayCastSingle and QueryAABB (world.h:141-172, world.cpp:452-566,
 * 852-870) against every fixture's broadphase proxy (its fat AABB) and shape. A ray is a candidate
 * for a fixture when the fixture's fat AABB passes b3DynamicTree::RayCast's segment test with the
 * ray's full extent (dynamic_tree.h:201-339); the fixture's shape is then cast as b3Fixture::RayCast
 * does (sphere_shape.cpp:69-121, capsule_shape.cpp:180-324, hull_shape.cpp:189-281,
 * triangle_shape.cpp:95-176, mesh_shape.cpp:81-146, the mesh's static tree in stack order). The
 * device finds candidates with a bounding-volume hierarchy of its own over the fat AABBs, rebuilt
 * after each step; the reference's dynamic-tree shape does not enter the results, except through
 * the order of hits at exactly equal fractions (here: lowest fixture index first).
 * A fixture mask stands for b3RayCastFilter::ShouldRayCast / b3QueryFilter::ShouldReport
 * (world_callbacks.h): a query whose `mask` is not BX_NULL_INDEX considers fixture f of its world only
 * when mask_bytes[mask + f] != 0.
 * ------------------------------------------------------------------------------------------- */
typedef struct bx_ray_query {
    uint32_t world;                /* world (pool slot) the ray is cast in */
    uint32_t mask;                 /* BX_NULL_INDEX or an offset into the call's mask bytes */
    float p1[3];                   /* b3RayCastInput::p1, p2 (p1 != p2) */
    float p2[3];
} bx_ray_query;

typedef struct bx_ray_hit {
    uint32_t fixture;              /* creation index of the fixture hit, BX_NULL_INDEX for none */
    float fraction;                /* b3RayCastOutput::fraction */
    float point[3];                /* (1 - fraction) * p1 + fraction * p2 */
    float normal[3];               /* b3RayCastOutput::normal */
} bx_ray_hit;

typedef struct bx_aabb_query {
    uint32_t world;
    uint32_t mask;                 /* BX_NULL_INDEX or an offset into the call's mask bytes */
    float lower[3];                /* b3AABB lowerBound, upperBound */
    float upper[3];
} bx_aabb_query;

/* b3World::RayCastSingle for each ray: the hit with the smallest fraction, in out[i] (fixture ==
   BX_NULL_INDEX when the ray hits nothing). Synchronous. */
int bx_batch_raycast_single(bx_batch* batch, const bx_ray_query* rays, uint32_t count, const uint8_t* mask_bytes,
                            uint64_t mask_size, bx_ray_hit* out);
/* b3World::RayCast with a listener that never clips the ray (ReportFixture returning 1): every hit
   of ray i in increasing fraction, the first hits_per_ray of them in hits[i * hits_per_ray ..];
   hit_counts[i] receives the number of hits. Synchronous. */
int bx_batch_raycast(bx_batch* batch, const bx_ray_query* rays, uint32_t count, const uint8_t* mask_bytes,
                     uint64_t mask_size, bx_ray_hit* hits, uint32_t hits_per_ray, uint32_t* hit_counts);
/* b3World::QueryAABB: the fixtures whose fat AABB overlaps box i (b3TestOverlap, aabb.h:347-367),
   in creation order, the first per_query of them in fixtures[i * per_query ..]; counts[i]
   receives their number. Synchronous. */
int bx_batch_query_aabb(bx_batch* batch, const bx_aabb_query* boxes, uint32_t count, const uint8_t* mask_bytes,
                        uint64_t mask_size, uint32_t* fixtures, uint32_t per_query, uint32_t* counts);

/* ---------------------------------------------------------------------------------------------
 * Time of impact and shape casts (time_of_impact.cpp, world.cpp:568-850).
 * A cast shape is a convex shape as b3GJKProxy::Set takes it (gjk.cpp:37-95): a sphere (vertex1 the
 * centre), a capsule (vertex1, vertex2), a triangle (vertex1..3) or a hull of the batch's shape library,
 * with its radius. GJK inside the time of impact carries its simplex cache between iterations
 * (gjk.cpp:651-879), as the reference's does.
 * ------------------------------------------------------------------------------------------- */
typedef struct bx_cast_shape {
    uint32_t shape_type;           /* BX_SHAPE_SPHERE, BX_SHAPE_CAPSULE, BX_SHAPE_TRIANGLE or BX_SHAPE_HULL */
    uint32_t shape_index;          /* BX_SHAPE_HULL: index into the batch's shape library */
    float radius;                  /* b3Shape::m_radius */
    float vertex1[3];
    float vertex2[3];
    float vertex3[3];
} bx_cast_shape;

typedef struct bx_sweep {          /* b3Sweep (sweep.h:27-63) */
    float local_center[3];
    float center0[3];              /* worldCenter0 */
    float center[3];               /* worldCenter */
    float orientation0[4];         /* x y z w */
    float orientation[4];
    float t0;                      /* not read by the time of impact, as in the reference */
} bx_sweep;

#define BX_TOI_SWEEP 0u            /* b3TimeOfImpact(const b3TOIInput&)           time_of_impact.cpp:465-643 */
#define BX_TOI_TRANSLATION 1u      /* b3TimeOfImpact(xf1, proxy1, d1, xf2, proxy2, d2, maxIterations)
                                                                                  time_of_impact.cpp:647-758 */
#define BX_TOI_AABB 2u             /* b3TimeOfImpact(aabb1, d1, aabb2, d2)        time_of_impact.cpp:764-901 */

typedef struct bx_toi_query {
    uint32_t kind;                 /* BX_TOI_* */
    uint32_t max_iterations;       /* BX_TOI_TRANSLATION: maxIterations (the reference's default is 20) */
    float t_max;                   /* BX_TOI_SWEEP: b3TOIInput::tMax */
    uint32_t reserved;
    bx_cast_shape shape1;          /* BX_TOI_SWEEP, BX_TOI_TRANSLATION */
    bx_cast_shape shape2;
    bx_sweep sweep1;               /* BX_TOI_SWEEP */
    bx_sweep sweep2;
    float position1[3];            /* BX_TOI_TRANSLATION: xf1, xf2 (orientation x y z w) */
    float orientation1[4];
    float position2[3];
    float orientation2[4];
    float displacement1[3];        /* BX_TOI_TRANSLATION and BX_TOI_AABB: d1, d2 */
    float displacement2[3];
    float lower1[3];               /* BX_TOI_AABB: aabb1, aabb2 */
    float upper1[3];
    float lower2[3];
    float upper2[3];
} bx_toi_query;

#define BX_TOI_UNKNOWN 0u          /* b3TOIOutput::State (time_of_impact.h:37-51) */
#define BX_TOI_FAILED 1u
#define BX_TOI_OVERLAPPED 2u
#define BX_TOI_TOUCHING 3u
#define BX_TOI_SEPARATED 4u

typedef struct bx_toi_result {
    uint32_t state;                /* BX_TOI_UNKNOWN .. BX_TOI_SEPARATED */
    float t;
    uint32_t iterations;
} bx_toi_result;

/* `count` time-of-impact problems, one per lane, hull shapes from the batch's library. Synchronous. */
int bx_batch_time_of_impact(bx_batch* batch, const bx_toi_query* queries, uint32_t count, bx_toi_result* out);

/* b3World::ShapeCast / ShapeCastSingle (world.cpp:568-850) in world `world`: `shape` at `position`,
   `orientation` moved by `displacement`. Candidates are the fixtures whose fat AABB overlaps the swept
   AABB (b3TestOverlap with the shape's AABBs at both ends combined); a convex fixture is cast as one
   b3TimeOfImpact of the translating shape against the still fixture, a mesh fixture once per triangle
   its static tree reports for the swept AABB in the mesh frame. A hit is a cast that ends touching;
   its point and normal come from GJK at the hit (point on the fixture's surface, normal towards the
   shape). */
typedef struct bx_shape_cast_query {
    uint32_t world;
    uint32_t mask;                 /* BX_NULL_INDEX or an offset into the call's mask bytes (b3ShapeCastFilter) */
    bx_cast_shape shape;
    float position[3];
    float orientation[4];
    float displacement[3];
} bx_shape_cast_query;

typedef struct bx_shape_hit {
    uint32_t fixture;              /* creation index, BX_NULL_INDEX for none */
    uint32_t child;                /* mesh triangle index; 0 for a convex fixture */
    float fraction;
    float point[3];
    float normal[3];
} bx_shape_hit;

/* b3World::ShapeCastSingle for each cast: the touching hit of smallest fraction (ties: lowest fixture,
   then triangle). Synchronous. */
int bx_batch_shape_cast_single(bx_batch* batch, const bx_shape_cast_query* casts, uint32_t count,
                               const uint8_t* mask_bytes, uint64_t mask_size, bx_shape_hit* out);
/* b3World::ShapeCast with a listener that never clips (ReportFixture returning 1): every hit of cast i
   in increasing (fraction, fixture, triangle), the first hits_per_cast of them in
   hits[i * hits_per_cast ..]; hit_counts[i] receives their number. Synchronous. */
int bx_batch_shape_cast(bx_batch* batch, const bx_shape_cast_query* casts, uint32_t count, const uint8_t* mask_bytes,
                        uint64_t mask_size, bx_shape_hit* hits, uint32_t hits_per_cast, uint32_t* hit_counts);

/* Per-phase GPU time per step, in milliseconds, averaged over the steps of the last bx_batch_step
   call on the batch's context (its first 1024 steps), from HIP events recorded on the context
   stream between the kernels: [0] update contacts, [1] islands, [2] island solve, [3] fixture
   sync and new pairs, [4] whole step. Waits for those steps to finish. */
int bx_batch_get_timings(bx_batch* batch, float* ms, uint32_t count);

/* Per-kernel GPU time (ms) per step over the same steps, in launch order; names from
   bx_kernel_name(i). Returns the number of entries written (at most 11). */
int bx_batch_get_kernel_timings(bx_batch* batch, float* ms, uint32_t count);
const char* bx_kernel_name(uint32_t index);

/* Size in bytes of an ABI structure by name (layout checks for bindings); 0 if unknown. */
uint32_t bx_struct_size(const char* name);

/* Device memory held by a batch, in bytes. */
uint64_t bx_batch_device_bytes(bx_batch* batch);

/* ---------------------------------------------------------------------------------------------
 * Mesh structures as the device path builds them at upload (host only, no device needed)
 * ------------------------------------------------------------------------------------------- */
/* b3StaticTree::QueryAABB (static_tree.h:215-255) on the mesh's tree as b3Mesh::BuildTree builds it
   (mesh.cpp:21-33, static_tree.cpp:40-170), for n boxes aabbs[6k..6k+5] (lower xyz, upper xyz):
   counts[k] overlapping triangles, the first cap_per of them in out[k * cap_per ..] in the order a
   mesh contact visits them. */
int bx_mesh_query_aabbs(const bx_mesh_desc* mesh, const float* aabbs, uint32_t n, uint32_t* counts, uint32_t* out,
                        uint32_t cap_per);
/* b3Mesh::BuildAdjacency (mesh.cpp:35-112): the wing vertices u1, u2, u3 of every triangle
   (BX_NULL_INDEX on a boundary edge) into out[3 * triangle_count]. */
int bx_mesh_wings(const bx_mesh_desc* mesh, uint32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* BXSTEP_H */
