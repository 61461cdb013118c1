"""ctypes mirror of include/bxstep.h (the C ABI of libbxstep.so).

The structures here must stay byte-identical to the C declarations; tests/test_capi_layout.py
checks every size and offset against the compiled library's own report.
"""
import ctypes as C
import os

u32 = C.c_uint32
f32 = C.c_float

BX_OK = 0
BX_E_INVALID, BX_E_DEVICE, BX_E_CAPACITY, BX_E_UNSUPPORTED, BX_E_NOMEM = -1, -2, -3, -4, -5
BX_SHAPE_SPHERE, BX_SHAPE_CAPSULE, BX_SHAPE_TRIANGLE, BX_SHAPE_HULL, BX_SHAPE_MESH = 0, 1, 2, 3, 4
BX_BODY_STATIC, BX_BODY_KINEMATIC, BX_BODY_DYNAMIC = 0, 1, 2
BX_BODY_AWAKE = 0x1
BX_BODY_ISLAND = 0x2
BX_BODY_AUTO_SLEEP = 0x4
BX_BODY_FIXED_ROT_X, BX_BODY_FIXED_ROT_Y, BX_BODY_FIXED_ROT_Z = 0x8, 0x10, 0x20
BX_WORLD_SLEEPING = 0x1
BX_WORLD_EVENTS = 0x8
BX_EVENT_BEGIN_CONTACT, BX_EVENT_END_CONTACT, BX_EVENT_PRE_SOLVE, BX_EVENT_POST_SOLVE = 1, 2, 3, 4
(BX_EDIT_LINEAR_VELOCITY, BX_EDIT_ANGULAR_VELOCITY, BX_EDIT_FORCE, BX_EDIT_FORCE_TO_CENTER, BX_EDIT_TORQUE,
 BX_EDIT_LINEAR_IMPULSE, BX_EDIT_ANGULAR_IMPULSE, BX_EDIT_TRANSFORM, BX_EDIT_AWAKE, BX_EDIT_SLEEPING_ALLOWED,
 BX_EDIT_GRAVITY_SCALE, BX_EDIT_LINEAR_DAMPING, BX_EDIT_ANGULAR_DAMPING, BX_EDIT_LINEAR_SLEEP_TOLERANCE,
 BX_EDIT_ANGULAR_SLEEP_TOLERANCE, BX_EDIT_FIXED_ROTATION) = range(1, 17)
(BX_EDIT_WORLD_GRAVITY, BX_EDIT_WORLD_SLEEPING, BX_EDIT_WORLD_WARM_STARTING, BX_EDIT_WORLD_AUTO_CLEAR_FORCES,
 BX_EDIT_WORLD_CLEAR_FORCES) = range(32, 37)
BX_WORLD_NO_WARM_START = 0x2
BX_WORLD_NO_CLEAR_FORCES = 0x4
(BX_JOINT_MOUSE, BX_JOINT_SPRING, BX_JOINT_WELD, BX_JOINT_REVOLUTE, BX_JOINT_SPHERE, BX_JOINT_CONE,
 BX_JOINT_FRICTION, BX_JOINT_MOTOR, BX_JOINT_PRISMATIC, BX_JOINT_WHEEL) = range(1, 11)
BX_NULL_INDEX = 0xFFFFFFFF
(BX_TOPO_CREATE_BODY, BX_TOPO_CREATE_FIXTURE, BX_TOPO_CREATE_JOINT, BX_TOPO_DESTROY_BODY, BX_TOPO_DESTROY_FIXTURE,
 BX_TOPO_DESTROY_JOINT) = range(1, 7)


def _arr(t, n):
    return t * n


class bx_hull_desc(C.Structure):
    _fields_ = [("vertex_count", u32), ("edge_count", u32), ("face_count", u32),
                ("vertices", C.POINTER(f32)), ("edges", C.POINTER(u32)), ("faces", C.POINTER(u32)),
                ("planes", C.POINTER(f32)), ("centroid", _arr(f32, 3))]


class bx_mesh_desc(C.Structure):
    _fields_ = [("vertex_count", u32), ("triangle_count", u32),
                ("vertices", C.POINTER(f32)), ("triangles", C.POINTER(u32))]


class bx_shape_library(C.Structure):
    _fields_ = [("hull_count", u32), ("hulls", C.POINTER(bx_hull_desc)),
                ("mesh_count", u32), ("meshes", C.POINTER(bx_mesh_desc))]


class bx_body_desc(C.Structure):
    _fields_ = [("type", u32), ("flags", u32), ("position", _arr(f32, 3)), ("orientation", _arr(f32, 4)),
                ("linear_velocity", _arr(f32, 3)), ("angular_velocity", _arr(f32, 3)),
                ("linear_damping", _arr(f32, 3)), ("angular_damping", _arr(f32, 3)),
                ("gravity_scale", _arr(f32, 3)), ("linear_sleep_tolerance", f32),
                ("angular_sleep_tolerance", f32)]


class bx_fixture_desc(C.Structure):
    _fields_ = [("body", u32), ("shape_type", u32), ("shape_index", u32), ("is_sensor", u32),
                ("radius", f32), ("vertex1", _arr(f32, 3)), ("vertex2", _arr(f32, 3)),
                ("vertex3", _arr(f32, 3)), ("mesh_scale", _arr(f32, 3)),
                ("density", f32), ("friction", f32), ("restitution", f32)]


class bx_joint_desc(C.Structure):
    _fields_ = [("type", u32), ("body_a", u32), ("body_b", u32), ("collide_connected", u32),
                ("use_initialize", u32), ("axis", _arr(f32, 3)), ("anchor", _arr(f32, 3)),
                ("local_anchor_a", _arr(f32, 3)), ("local_rotation_a", _arr(f32, 4)),
                ("local_anchor_b", _arr(f32, 3)), ("local_rotation_b", _arr(f32, 4)),
                ("reference_rotation", _arr(f32, 4)), ("enable_limit", u32), ("lower_angle", f32),
                ("upper_angle", f32), ("enable_motor", u32), ("motor_speed", f32),
                ("max_motor_torque", f32), ("anchor_b", _arr(f32, 3)), ("frequency_hz", f32),
                ("damping_ratio", f32), ("length", f32), ("max_force", f32), ("max_torque", f32),
                ("correction_factor", f32), ("target", _arr(f32, 3)), ("linear_offset", _arr(f32, 3)),
                ("angular_offset", _arr(f32, 4)), ("cone_angle", f32), ("local_axis_a", _arr(f32, 3)),
                ("local_axis_b", _arr(f32, 3)), ("lower_translation", f32), ("upper_translation", f32),
                ("axis_b", _arr(f32, 3))]


class bx_world_desc(C.Structure):
    _fields_ = [("gravity", _arr(f32, 3)), ("flags", u32),
                ("body_count", u32), ("bodies", C.POINTER(bx_body_desc)),
                ("fixture_count", u32), ("fixtures", C.POINTER(bx_fixture_desc)),
                ("joint_count", u32), ("joints", C.POINTER(bx_joint_desc))]


class bx_body_state(C.Structure):
    _fields_ = [("position", _arr(f32, 3)), ("orientation", _arr(f32, 4)), ("world_center", _arr(f32, 3)),
                ("linear_velocity", _arr(f32, 3)), ("angular_velocity", _arr(f32, 3)),
                ("sleep_time", f32), ("flags", u32), ("island_index", u32)]


class bx_batch_totals(C.Structure):
    _fields_ = [("bodies", C.c_uint64), ("awake_bodies", C.c_uint64), ("fixtures", C.c_uint64), ("joints", C.c_uint64),
                ("contacts", C.c_uint64), ("mesh_world_contacts", C.c_uint64), ("pairs", C.c_uint64),
                ("islands", C.c_uint64), ("island_contacts", C.c_uint64), ("worlds", u32), ("status_worlds", u32),
                ("status_or", u32), ("reserved", u32)]


class bx_contact_event(C.Structure):
    _fields_ = [("type", u32), ("fixture_a", u32), ("fixture_b", u32), ("contact", u32)]


class bx_body_edit(C.Structure):
    _fields_ = [("body", u32), ("op", u32), ("wake", u32), ("reserved", u32), ("v", _arr(f32, 3)), ("p", _arr(f32, 4))]


def body_edits(rows):
    """bx_body_edit records from (body, op, v, p, wake) tuples (p and wake optional)."""
    out = (bx_body_edit * max(1, len(rows)))()
    for i, r in enumerate(rows):
        body, op, v = r[0], r[1], r[2]
        p = r[3] if len(r) > 3 and r[3] is not None else (0.0, 0.0, 0.0, 0.0)
        out[i].body, out[i].op, out[i].wake = body, op, int(r[4]) if len(r) > 4 else 1
        out[i].v[:] = [float(x) for x in v]
        out[i].p[:] = [float(x) for x in (list(p) + [0.0] * 4)[:4]]
    return out


class bx_body_mass(C.Structure):
    _fields_ = [("mass", f32), ("inv_mass", f32), ("local_center", _arr(f32, 3)), ("inertia", _arr(f32, 9)),
                ("inv_inertia", _arr(f32, 9)), ("world_inv_inertia", _arr(f32, 9))]


class bx_topology_op(C.Structure):
    _fields_ = [("op", u32), ("index", u32)]


def topology_ops(rows):
    """bx_topology_op records from (op, index) pairs."""
    out = (bx_topology_op * max(1, len(rows)))()
    for i, (op, index) in enumerate(rows):
        out[i].op, out[i].index = op, index
    return out


class bx_cast_shape(C.Structure):
    _fields_ = [("shape_type", u32), ("shape_index", u32), ("radius", f32), ("vertex1", _arr(f32, 3)),
                ("vertex2", _arr(f32, 3)), ("vertex3", _arr(f32, 3))]


class bx_sweep(C.Structure):
    _fields_ = [("local_center", _arr(f32, 3)), ("center0", _arr(f32, 3)), ("center", _arr(f32, 3)),
                ("orientation0", _arr(f32, 4)), ("orientation", _arr(f32, 4)), ("t0", f32)]


BX_TOI_SWEEP, BX_TOI_TRANSLATION, BX_TOI_AABB = 0, 1, 2
BX_TOI_UNKNOWN, BX_TOI_FAILED, BX_TOI_OVERLAPPED, BX_TOI_TOUCHING, BX_TOI_SEPARATED = 0, 1, 2, 3, 4


class bx_toi_query(C.Structure):
    _fields_ = [("kind", u32), ("max_iterations", u32), ("t_max", f32), ("reserved", u32),
                ("shape1", bx_cast_shape), ("shape2", bx_cast_shape), ("sweep1", bx_sweep), ("sweep2", bx_sweep),
                ("position1", _arr(f32, 3)), ("orientation1", _arr(f32, 4)), ("position2", _arr(f32, 3)),
                ("orientation2", _arr(f32, 4)), ("displacement1", _arr(f32, 3)), ("displacement2", _arr(f32, 3)),
                ("lower1", _arr(f32, 3)), ("upper1", _arr(f32, 3)), ("lower2", _arr(f32, 3)), ("upper2", _arr(f32, 3))]


class bx_toi_result(C.Structure):
    _fields_ = [("state", u32), ("t", f32), ("iterations", u32)]


class bx_shape_cast_query(C.Structure):
    _fields_ = [("world", u32), ("mask", u32), ("shape", bx_cast_shape), ("position", _arr(f32, 3)),
                ("orientation", _arr(f32, 4)), ("displacement", _arr(f32, 3))]


class bx_shape_hit(C.Structure):
    _fields_ = [("fixture", u32), ("child", u32), ("fraction", f32), ("point", _arr(f32, 3)),
                ("normal", _arr(f32, 3))]


class bx_ray_query(C.Structure):
    _fields_ = [("world", u32), ("mask", u32), ("p1", _arr(f32, 3)), ("p2", _arr(f32, 3))]


class bx_ray_hit(C.Structure):
    _fields_ = [("fixture", u32), ("fraction", f32), ("point", _arr(f32, 3)), ("normal", _arr(f32, 3))]


class bx_aabb_query(C.Structure):
    _fields_ = [("world", u32), ("mask", u32), ("lower", _arr(f32, 3)), ("upper", _arr(f32, 3))]


class bx_manifold_point_state(C.Structure):
    _fields_ = [("local_normal1", _arr(f32, 3)), ("local_point1", _arr(f32, 3)),
                ("local_point2", _arr(f32, 3)), ("normal_impulse", f32), ("triangle_key", u32),
                ("key", _arr(u32, 4)), ("persisting", u32), ("edge_contact", u32)]


class bx_manifold_state(C.Structure):
    _fields_ = [("point_count", u32), ("points", _arr(bx_manifold_point_state, 4)),
                ("tangent_impulse", _arr(f32, 2)), ("motor_impulse", f32)]


class bx_contact_state(C.Structure):
    _fields_ = [("fixture_a", u32), ("fixture_b", u32), ("touching", u32), ("manifold_count", u32),
                ("cache", _arr(u32, 4)), ("manifolds", _arr(bx_manifold_state, 3))]


class bx_pair(C.Structure):
    _fields_ = [("proxy1", u32), ("proxy2", u32)]


class bx_world_counts(C.Structure):
    _fields_ = [("body_count", u32), ("fixture_count", u32), ("joint_count", u32), ("contact_count", u32),
                ("touching_count", u32), ("pair_count", u32), ("island_count", u32), ("awake_count", u32),
                ("status", u32)]


class bx_step_cfg(C.Structure):
    _fields_ = [("dt", f32), ("velocity_iterations", u32), ("position_iterations", u32)]


class bx_batch_opts(C.Structure):
    _fields_ = [("max_contacts_per_world", u32), ("max_pairs_per_world", u32), ("solver_mode", u32),
                ("mesh_triangles_per_contact", u32), ("reserved", _arr(u32, 4))]


class bx_pool_opts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_worlds", u32), ("max_contacts_per_world", u32),
                ("max_pairs_per_world", u32), ("joints_per_1000_bodies", u32), ("solver_mode", u32),
                ("reserved", _arr(u32, 3))]


class bx_pool_stats(C.Structure):
    _fields_ = [("capacity_bytes", C.c_uint64), ("used_bytes", C.c_uint64), ("resident_worlds", u32),
                ("free_slots", u32), ("world_slots", u32), ("reserved", _arr(u32, 5))]


STRUCTS = [bx_hull_desc, bx_mesh_desc, bx_shape_library, bx_body_desc, bx_fixture_desc, bx_joint_desc,
           bx_world_desc, bx_body_state, bx_manifold_point_state, bx_manifold_state, bx_contact_state,
           bx_pair, bx_world_counts, bx_step_cfg, bx_batch_opts, bx_pool_opts, bx_pool_stats]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "lib", "libbxstep.so")
