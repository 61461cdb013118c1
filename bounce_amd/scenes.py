"""Scene recipes and descriptor building for the batched step.

A SceneSet holds a shape library and a list of worlds as numpy structured arrays laid out exactly
like the C descriptors of include/bxstep.h, and exposes them as ctypes pointers for
bx_batch_create (and, in tests, for the oracles, which take the same descriptors).

Workloads (sizes as named by the benchmark configurations):
  * drop_boxes    -- 1,000 unit boxes (half-extent 0.5, density 1) at seeded uniform positions in a
                     20 x 50 x 20 m volume over a static 100 x 1 x 100 m ground box.
  * pyramid       -- a 20-level 2-D box pyramid (row r has 20 - r boxes, 210 boxes) of unit boxes,
                     friction 0.6, seeded +-1 mm jitter per axis, on the same ground box.
  * hello_world   -- examples/hello_world/main.cpp:24-112 of the reference.
Defaults follow b3BodyDef (body.h:49-123), b3FixtureDef (fixture.h:37-70) and b3World
(world.cpp:41-69: gravity (0, -9.8, 0), warm starting on, sleeping off).
"""
import ctypes as C
import numpy as np

from . import capi
from .capi import (bx_body_desc, bx_fixture_desc, bx_joint_desc, bx_world_desc, bx_hull_desc,
                   bx_mesh_desc, bx_shape_library)
from . import hulls as hull_lib

F32 = np.float32
BODY_DT = np.dtype(bx_body_desc)
FIX_DT = np.dtype(bx_fixture_desc)
JOINT_DT = np.dtype(bx_joint_desc)

# b3BodyDef::angularSleepTolerance = scalar(2) / scalar(180) * B3_PI, evaluated in binary32
ANGULAR_SLEEP_TOL = F32(F32(2.0) / F32(180.0)) * F32(np.pi)


def body_array(n):
    a = np.zeros(n, dtype=BODY_DT)
    a["type"] = capi.BX_BODY_STATIC
    a["flags"] = capi.BX_BODY_AWAKE | capi.BX_BODY_AUTO_SLEEP
    a["orientation"][:, 3] = 1.0
    a["gravity_scale"] = 1.0
    a["linear_sleep_tolerance"] = F32(0.05)
    a["angular_sleep_tolerance"] = ANGULAR_SLEEP_TOL
    return a


def fixture_array(n):
    a = np.zeros(n, dtype=FIX_DT)
    a["friction"] = F32(0.3)
    a["mesh_scale"] = 1.0
    return a


def joint_array(n):
    """Joint descriptors with the defs' defaults (revolute type; weld and motor correction 0.2)."""
    a = np.zeros(n, dtype=JOINT_DT)
    a["type"] = capi.BX_JOINT_REVOLUTE
    for k in ("local_rotation_a", "local_rotation_b", "reference_rotation", "angular_offset"):
        a[k][:, 3] = 1.0
    a["correction_factor"] = F32(0.2)
    return a


class World:
    """One world's descriptor arrays (creation order)."""

    def __init__(self, bodies, fixtures, joints=None, gravity=(0.0, -9.8, 0.0), flags=0):
        self.bodies = bodies
        self.fixtures = fixtures
        self.joints = joints if joints is not None else joint_array(0)
        self.gravity = gravity
        self.flags = flags

    @property
    def body_count(self):
        return len(self.bodies)

    def desc(self):
        d = bx_world_desc()
        d.gravity[:] = [float(g) for g in self.gravity]
        d.flags = self.flags
        d.body_count = len(self.bodies)
        d.bodies = self.bodies.ctypes.data_as(C.POINTER(bx_body_desc))
        d.fixture_count = len(self.fixtures)
        d.fixtures = self.fixtures.ctypes.data_as(C.POINTER(bx_fixture_desc))
        d.joint_count = len(self.joints)
        d.joints = self.joints.ctypes.data_as(C.POINTER(bx_joint_desc))
        return d


class SceneSet:
    """A shape library plus worlds; keeps every array referenced by the descriptors alive."""

    def __init__(self, name):
        self.name = name
        self.hulls = []      # list of hull_lib.HullData
        self.meshes = []     # list of (vertices f32[V,3], triangles u32[T,3])
        self.worlds = []
        self._lib = None
        self._descs = None

    def add_hull(self, h):
        self.hulls.append(h)
        return len(self.hulls) - 1

    def library(self):
        hd = (bx_hull_desc * max(1, len(self.hulls)))()
        for i, h in enumerate(self.hulls):
            hd[i].vertex_count = len(h.vertices)
            hd[i].edge_count = len(h.edges)
            hd[i].face_count = len(h.faces)
            hd[i].vertices = h.vertices.ctypes.data_as(C.POINTER(C.c_float))
            hd[i].edges = h.edges.ctypes.data_as(C.POINTER(C.c_uint32))
            hd[i].faces = h.faces.ctypes.data_as(C.POINTER(C.c_uint32))
            hd[i].planes = h.planes.ctypes.data_as(C.POINTER(C.c_float))
            hd[i].centroid[:] = [float(x) for x in h.centroid]
        md = (bx_mesh_desc * max(1, len(self.meshes)))()
        for i, (v, t) in enumerate(self.meshes):
            md[i].vertex_count = len(v)
            md[i].triangle_count = len(t)
            md[i].vertices = v.ctypes.data_as(C.POINTER(C.c_float))
            md[i].triangles = t.ctypes.data_as(C.POINTER(C.c_uint32))
        lib = bx_shape_library()
        lib.hull_count = len(self.hulls)
        lib.hulls = hd
        lib.mesh_count = len(self.meshes)
        lib.meshes = md
        self._lib = (lib, hd, md)
        return lib

    def world_descs(self):
        arr = (bx_world_desc * len(self.worlds))()
        for i, w in enumerate(self.worlds):
            arr[i] = w.desc()
        self._descs = arr
        return arr

    @property
    def body_count(self):
        return sum(w.body_count for w in self.worlds)


def _ground(bodies, fixtures, hull_index, friction=0.3, y=-0.5):
    bodies[0]["type"] = capi.BX_BODY_STATIC
    bodies[0]["position"] = (0.0, y, 0.0)
    fixtures[0]["body"] = 0
    fixtures[0]["shape_type"] = capi.BX_SHAPE_HULL
    fixtures[0]["shape_index"] = hull_index
    fixtures[0]["friction"] = F32(friction)


def hello_world():
    """examples/hello_world/main.cpp: a 2 m box spinning at pi rad/s dropped from y = 10."""
    s = SceneSet("hello_world")
    g = s.add_hull(hull_lib.box_hull(10.0, 1.0, 10.0))
    b = s.add_hull(hull_lib.box_hull(1.0, 1.0, 1.0))
    bodies = body_array(2)
    fixtures = fixture_array(2)
    fixtures[0]["shape_type"] = capi.BX_SHAPE_HULL
    fixtures[0]["shape_index"] = g
    bodies[1]["type"] = capi.BX_BODY_DYNAMIC
    bodies[1]["position"] = (0.0, 10.0, 0.0)
    bodies[1]["angular_velocity"] = (0.0, F32(np.pi), 0.0)
    fixtures[1]["body"] = 1
    fixtures[1]["shape_type"] = capi.BX_SHAPE_HULL
    fixtures[1]["shape_index"] = b
    fixtures[1]["density"] = 1.0
    s.worlds.append(World(bodies, fixtures, gravity=(0.0, F32(-9.8), 0.0)))
    return s


def mesh_ground_program():
    """tests/drop_in/mesh_ground.cpp: a sphere (r 0.5), a capsule (1 m, r 0.3) and a 1 m box dropped
    on a static 6 x 6-cell heightfield mesh of 1 m cells, heights ((7i + 3j) mod 5) / 10 - 0.2."""
    s = SceneSet("mesh_ground_program")
    box = s.add_hull(hull_lib.box_hull(0.5, 0.5, 0.5))
    n = 6
    v = np.zeros(((n + 1) * (n + 1), 3), dtype=F32)
    for i in range(n + 1):
        for j in range(n + 1):
            v[i * (n + 1) + j] = (F32(-3.0) + F32(i), F32((7 * i + 3 * j) % 5) * F32(0.1) - F32(0.2), F32(-3.0) + F32(j))
    t = []
    for i in range(n):
        for j in range(n):
            a, b = i * (n + 1) + j, (i + 1) * (n + 1) + j
            c, d = (i + 1) * (n + 1) + j + 1, i * (n + 1) + j + 1
            t.append((a, c, b))
            t.append((a, d, c))
    s.meshes.append((v, np.array(t, dtype=np.uint32)))
    bodies = body_array(4)
    fixtures = fixture_array(4)
    fixtures[0]["shape_type"] = capi.BX_SHAPE_MESH
    fixtures[0]["shape_index"] = 0
    for k, pos in enumerate([(0.3, 2.5, -0.2), (-1.2, 3.0, 0.8), (1.1, 3.5, 1.3)], start=1):
        bodies[k]["type"] = capi.BX_BODY_DYNAMIC
        bodies[k]["position"] = pos
        fixtures[k]["body"] = k
        fixtures[k]["density"] = 1.0
    fixtures[1]["shape_type"] = capi.BX_SHAPE_SPHERE
    fixtures[1]["radius"] = F32(0.5)
    fixtures[2]["shape_type"] = capi.BX_SHAPE_CAPSULE
    fixtures[2]["vertex1"] = (-0.5, 0.0, 0.0)
    fixtures[2]["vertex2"] = (0.5, 0.0, 0.0)
    fixtures[2]["radius"] = F32(0.3)
    fixtures[3]["shape_type"] = capi.BX_SHAPE_HULL
    fixtures[3]["shape_index"] = box
    s.worlds.append(World(bodies, fixtures, gravity=(0.0, F32(-9.8), 0.0)))
    return s


def drop_boxes_world(rng, n_boxes, ground_hull, box_hull, extent=(20.0, 50.0, 20.0)):
    bodies = body_array(n_boxes + 1)
    fixtures = fixture_array(n_boxes + 1)
    _ground(bodies, fixtures, ground_hull)
    ex, ey, ez = extent
    pos = np.empty((n_boxes, 3), dtype=F32)
    pos[:, 0] = rng.uniform(-ex / 2, ex / 2, n_boxes)
    pos[:, 1] = rng.uniform(1.0, 1.0 + ey, n_boxes)
    pos[:, 2] = rng.uniform(-ez / 2, ez / 2, n_boxes)
    bodies["type"][1:] = capi.BX_BODY_DYNAMIC
    bodies["position"][1:] = pos
    fixtures["body"][1:] = np.arange(1, n_boxes + 1)
    fixtures["shape_type"][1:] = capi.BX_SHAPE_HULL
    fixtures["shape_index"][1:] = box_hull
    fixtures["density"][1:] = 1.0
    return World(bodies, fixtures)


def drop_boxes(n_boxes=1000, n_worlds=1, seed=1, extent=(20.0, 50.0, 20.0)):
    s = SceneSet("drop_boxes")
    g = s.add_hull(hull_lib.box_hull(50.0, 0.5, 50.0))
    b = s.add_hull(hull_lib.box_hull(0.5, 0.5, 0.5))
    rng = np.random.default_rng(seed)
    for _ in range(n_worlds):
        s.worlds.append(drop_boxes_world(rng, n_boxes, g, b, extent))
    return s


def box_worlds(n_worlds, seed=8, median=200, sigma=0.8, clip=(50, 5000)):
    """Worlds of unit boxes dropped on a ground box, per-world counts round(LogNormal(ln median,
    sigma)) clipped to `clip`, each in a column whose footprint grows with the count (about one box
    per 2 m^2 of ground, 20 m tall)."""
    s = SceneSet("box_worlds")
    g = s.add_hull(hull_lib.box_hull(50.0, 0.5, 50.0))
    b = s.add_hull(hull_lib.box_hull(0.5, 0.5, 0.5))
    rng = np.random.default_rng(seed)
    sizes = np.clip(np.round(rng.lognormal(np.log(median), sigma, n_worlds)), clip[0], clip[1]).astype(int)
    for n in sizes:
        side = float(min(90.0, max(6.0, np.sqrt(2.0 * n / 4.0))))
        s.worlds.append(drop_boxes_world(rng, int(n), g, b, extent=(side, 20.0, side)))
    return s


def pyramid_world(rng, levels, ground_hull, box_hull, friction=0.6, jitter=1e-3):
    n = levels * (levels + 1) // 2
    bodies = body_array(n + 1)
    fixtures = fixture_array(n + 1)
    _ground(bodies, fixtures, ground_hull, friction=friction)
    pos = []
    for r in range(levels):
        count = levels - r
        for k in range(count):
            pos.append((-0.5 * (count - 1) + k, 0.5 + r, 0.0))
    pos = np.asarray(pos, dtype=np.float64)
    if jitter > 0:
        pos += rng.uniform(-jitter, jitter, pos.shape)
    bodies["type"][1:] = capi.BX_BODY_DYNAMIC
    bodies["position"][1:] = pos.astype(F32)
    fixtures["body"][1:] = np.arange(1, n + 1)
    fixtures["shape_type"][1:] = capi.BX_SHAPE_HULL
    fixtures["shape_index"][1:] = box_hull
    fixtures["density"][1:] = 1.0
    fixtures["friction"][1:] = F32(friction)
    return World(bodies, fixtures)


def pyramids(n_worlds=4096, levels=20, seed=2, jitter=1e-3):
    s = SceneSet("pyramids")
    g = s.add_hull(hull_lib.box_hull(50.0, 0.5, 50.0))
    b = s.add_hull(hull_lib.box_hull(0.5, 0.5, 0.5))
    rng = np.random.default_rng(seed)
    for _ in range(n_worlds):
        s.worlds.append(pyramid_world(rng, levels, g, b, jitter=jitter))
    return s


def box_stack(n_boxes=5, seed=3):
    """A vertical stack of unit boxes with small seeded offsets (testbed box_stack.h recipe)."""
    s = SceneSet("box_stack")
    g = s.add_hull(hull_lib.box_hull(50.0, 0.5, 50.0))
    b = s.add_hull(hull_lib.box_hull(0.5, 0.5, 0.5))
    rng = np.random.default_rng(seed)
    bodies = body_array(n_boxes + 1)
    fixtures = fixture_array(n_boxes + 1)
    _ground(bodies, fixtures, g)
    for i in range(n_boxes):
        bodies[i + 1]["type"] = capi.BX_BODY_DYNAMIC
        bodies[i + 1]["position"] = (rng.uniform(-0.05, 0.05), 0.5 + 1.0 * i + 0.01 * i, rng.uniform(-0.05, 0.05))
        fixtures[i + 1]["body"] = i + 1
        fixtures[i + 1]["shape_type"] = capi.BX_SHAPE_HULL
        fixtures[i + 1]["shape_index"] = b
        fixtures[i + 1]["density"] = 1.0
    s.worlds.append(World(bodies, fixtures))
    return s


def with_flags(scene, flags):
    """Set the BX_WORLD_* flags (e.g. BX_WORLD_SLEEPING) of every world of a scene set."""
    for w in scene.worlds:
        w.flags = flags
    return scene


def concat(sets, name="mix"):
    """Merge scene sets into one batch (hull indices are remapped)."""
    out = SceneSet(name)
    for s in sets:
        base = len(out.hulls)
        out.hulls.extend(s.hulls)
        mbase = len(out.meshes)
        out.meshes.extend(s.meshes)
        for w in s.worlds:
            f = w.fixtures.copy()
            hull = f["shape_type"] == capi.BX_SHAPE_HULL
            f["shape_index"][hull] += base
            mesh = f["shape_type"] == capi.BX_SHAPE_MESH
            f["shape_index"][mesh] += mbase
            out.worlds.append(World(w.bodies.copy(), f, w.joints.copy(), w.gravity, w.flags))
    return out


def random_orientations(rng, n):
    """Uniform random unit quaternions (x, y, z, w), rounded to binary32."""
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q.astype(F32)


def mixed_drop_world(rng, n_bodies, ground_hull, box_hull=None, kinds=("sphere", "capsule", "box"),
                     extent=(8.0, 16.0, 8.0), sensors=0.0, hull_choices=None):
    """A static ground box plus `n_bodies` dynamic bodies of the given kinds at seeded positions and
    orientations: spheres r ~ U[0.2, 0.6], capsules with segment length ~ U[0.5, 2] along local x and
    radius 0.25, unit boxes, or hulls drawn from `hull_choices` (library indices). A fraction
    `sensors` of the fixtures are sensors."""
    bodies = body_array(n_bodies + 1)
    fixtures = fixture_array(n_bodies + 1)
    _ground(bodies, fixtures, ground_hull)
    ex, ey, ez = extent
    bodies["type"][1:] = capi.BX_BODY_DYNAMIC
    bodies["position"][1:, 0] = rng.uniform(-ex / 2, ex / 2, n_bodies)
    bodies["position"][1:, 1] = rng.uniform(1.0, 1.0 + ey, n_bodies)
    bodies["position"][1:, 2] = rng.uniform(-ez / 2, ez / 2, n_bodies)
    bodies["orientation"][1:] = random_orientations(rng, n_bodies)
    fixtures["body"][1:] = np.arange(1, n_bodies + 1)
    fixtures["density"][1:] = 1.0
    kind = rng.integers(0, len(kinds), n_bodies)
    for i in range(n_bodies):
        f = fixtures[i + 1]
        k = kinds[kind[i]]
        if k == "sphere":
            f["shape_type"] = capi.BX_SHAPE_SPHERE
            f["radius"] = F32(rng.uniform(0.2, 0.6))
        elif k == "capsule":
            half = F32(0.5 * rng.uniform(0.5, 2.0))
            f["shape_type"] = capi.BX_SHAPE_CAPSULE
            f["vertex1"] = (-half, 0.0, 0.0)
            f["vertex2"] = (half, 0.0, 0.0)
            f["radius"] = F32(0.25)
        elif k == "box":
            f["shape_type"] = capi.BX_SHAPE_HULL
            f["shape_index"] = box_hull
        else:
            f["shape_type"] = capi.BX_SHAPE_HULL
            f["shape_index"] = hull_choices[rng.integers(0, len(hull_choices))]
    if sensors > 0:
        fixtures["is_sensor"][1:] = (rng.uniform(size=n_bodies) < sensors).astype(np.uint32)
    return World(bodies, fixtures)


def mixed_drop(n_bodies=60, n_worlds=1, seed=4, kinds=("sphere", "capsule", "box"), extent=(8.0, 16.0, 8.0),
               sensors=0.0):
    """Spheres, capsules and boxes dropped on a ground box (the convex narrowphase mix)."""
    s = SceneSet("mixed_drop")
    g = s.add_hull(hull_lib.box_hull(50.0, 0.5, 50.0))
    b = s.add_hull(hull_lib.box_hull(0.5, 0.5, 0.5))
    rng = np.random.default_rng(seed)
    for _ in range(n_worlds):
        s.worlds.append(mixed_drop_world(rng, n_bodies, g, b, kinds, extent, sensors))
    return s


def triangle_ground_world(rng, n_bodies, box_hull, grid=6, size=12.0, bumps=0.3, kinds=("sphere", "capsule", "box"),
                          extent=(8.0, 12.0, 8.0), hull_choices=None):
    """Bodies of the given kinds (as in mixed_drop_world) dropped on a static ground made of isolated
    triangle fixtures (b3TriangleShape, no edge wings): a grid x grid lattice over a size x size
    square at y = 0 whose vertex heights are U[-bumps, bumps], two triangles per cell. The ground
    body sits at the origin with the identity orientation (triangle AABBs are computed in the
    shape's frame)."""
    nt = 2 * grid * grid
    ground = mixed_drop_world(rng, n_bodies, 0, box_hull, kinds, extent, 0.0, hull_choices)
    bodies = ground.bodies
    bodies[0]["position"] = (0.0, 0.0, 0.0)
    xs = np.linspace(-size / 2, size / 2, grid + 1)
    h = rng.uniform(-bumps, bumps, (grid + 1, grid + 1))
    v = lambda i, j: (F32(xs[i]), F32(h[i, j]), F32(xs[j]))
    tris = fixture_array(nt)
    k = 0
    for i in range(grid):
        for j in range(grid):
            a, b, c, d = v(i, j), v(i + 1, j), v(i + 1, j + 1), v(i, j + 1)
            for t in ((a, c, b), (a, d, c)):  # counter-clockwise seen from above: normals point up
                tris[k]["body"] = 0
                tris[k]["shape_type"] = capi.BX_SHAPE_TRIANGLE
                tris[k]["vertex1"], tris[k]["vertex2"], tris[k]["vertex3"] = t
                k += 1
    fixtures = np.concatenate([tris, ground.fixtures[1:]])
    return World(bodies, fixtures)


def triangle_drop(n_bodies=40, n_worlds=1, seed=14, grid=6, kinds=("sphere", "capsule", "box"), bumps=0.3,
                  extent=(8.0, 12.0, 8.0), hull_count=0):
    """Spheres, capsules and boxes (and random hulls when hull_count > 0) dropped on a bumpy ground
    of isolated triangles (the triangle narrowphase mix)."""
    s = SceneSet("triangle_drop")
    b = s.add_hull(hull_lib.box_hull(0.5, 0.5, 0.5))
    choices = None
    if hull_count:
        base = len(s.hulls)
        s.hulls.extend(hull_lib.random_hull_library(hull_count, seed=seed + 1))
        choices = np.arange(base, base + hull_count)
        kinds = tuple(kinds) + ("hull",)
    rng = np.random.default_rng(seed)
    for _ in range(n_worlds):
        s.worlds.append(triangle_ground_world(rng, n_bodies, b, grid, 12.0, bumps, kinds, extent, choices))
    return s


def grid_mesh(rng, cells=8, size=12.0, bumps=0.3):
    """A heightfield triangle mesh: (cells + 1)^2 vertices over a size x size square at heights
    U[-bumps, bumps], two counter-clockwise (upward-facing) triangles per cell."""
    xs = np.linspace(-size / 2, size / 2, cells + 1)
    h = rng.uniform(-bumps, bumps, (cells + 1, cells + 1))
    v = np.zeros(((cells + 1) * (cells + 1), 3), dtype=F32)
    for i in range(cells + 1):
        for j in range(cells + 1):
            v[i * (cells + 1) + j] = (xs[i], h[i, j], xs[j])
    t = []
    for i in range(cells):
        for j in range(cells):
            a, b = i * (cells + 1) + j, (i + 1) * (cells + 1) + j
            c, d = (i + 1) * (cells + 1) + j + 1, i * (cells + 1) + j + 1
            t.append((a, c, b))
            t.append((a, d, c))
    return v, np.array(t, dtype=np.uint32)


def mesh_ground_world(rng, n_bodies, box_hull, mesh_index, kinds=("sphere", "capsule", "box"), extent=(8.0, 12.0, 8.0),
                      hull_choices=None, scale=(1.0, 1.0, 1.0), sensors=0.0):
    """Bodies of the given kinds (as in mixed_drop_world) dropped on a static mesh fixture
    (b3MeshShape with the given scale) at the origin."""
    w = mixed_drop_world(rng, n_bodies, 0, box_hull, kinds, extent, sensors, hull_choices)
    w.bodies[0]["position"] = (0.0, 0.0, 0.0)
    f = w.fixtures[0]
    f["shape_type"] = capi.BX_SHAPE_MESH
    f["shape_index"] = mesh_index
    f["mesh_scale"] = [F32(x) for x in scale]
    f["is_sensor"] = 0
    w.fixtures[0] = f
    return w


def mesh_drop(n_bodies=40, n_worlds=1, seed=24, cells=8, kinds=("sphere", "capsule", "box"), bumps=0.3,
              extent=(8.0, 12.0, 8.0), hull_count=0, scale=(1.0, 1.0, 1.0), sensors=0.0):
    """Spheres, capsules and boxes (and random hulls when hull_count > 0) dropped on a bumpy
    heightfield mesh (the mesh-contact mix: static tree query, per-triangle narrowphase with edge
    wings, clustering into at most three manifolds)."""
    s = SceneSet("mesh_drop")
    b = s.add_hull(hull_lib.box_hull(0.5, 0.5, 0.5))
    choices = None
    if hull_count:
        base = len(s.hulls)
        s.hulls.extend(hull_lib.random_hull_library(hull_count, seed=seed + 1))
        choices = np.arange(base, base + hull_count)
        kinds = tuple(kinds) + ("hull",)
    rng = np.random.default_rng(seed)
    s.meshes.append(grid_mesh(rng, cells, 12.0, bumps))
    for _ in range(n_worlds):
        s.worlds.append(mesh_ground_world(rng, n_bodies, b, 0, kinds, extent, choices, scale, sensors))
    return s


def chain_world(rng, links=10, ground_hull=0, link_hull=1, motor=True, limit=False, extra=0, origin=(0.0, 12.0, 0.0),
                motor_speed=0.5 * np.pi, max_torque=50.0, lower=-0.25 * np.pi, upper=0.25 * np.pi):
    """A chain of `links` boxes hanging from the static ground body by revolute joints about z,
    built with b3RevoluteJointDef::Initialize (anchors between consecutive links), plus `extra`
    dynamic boxes dropped next to it."""
    nb = 1 + links + extra
    bodies = body_array(nb)
    fixtures = fixture_array(nb)
    _ground(bodies, fixtures, ground_hull)
    ox, oy, oz = origin
    for i in range(links):
        b = bodies[1 + i]
        b["type"] = capi.BX_BODY_DYNAMIC
        b["position"] = (ox + 1.0 * i + 0.5, oy, oz)
        f = fixtures[1 + i]
        f["body"] = 1 + i
        f["shape_type"] = capi.BX_SHAPE_HULL
        f["shape_index"] = link_hull
        f["density"] = 1.0
    for k in range(extra):
        i = 1 + links + k
        bodies[i]["type"] = capi.BX_BODY_DYNAMIC
        bodies[i]["position"] = (ox + rng.uniform(-3, 3), oy - 6.0 + rng.uniform(-1, 1), oz + rng.uniform(-3, 3))
        fixtures[i]["body"] = i
        fixtures[i]["shape_type"] = capi.BX_SHAPE_HULL
        fixtures[i]["shape_index"] = link_hull
        fixtures[i]["density"] = 1.0
    joints = joint_array(links)
    for i in range(links):
        j = joints[i]
        j["body_a"] = 0 if i == 0 else i
        j["body_b"] = 1 + i
        j["use_initialize"] = 1
        j["axis"] = (0.0, 0.0, 1.0)
        j["anchor"] = (ox + 1.0 * i, oy, oz)
        j["enable_motor"] = 1 if motor else 0
        j["motor_speed"] = F32(motor_speed)
        j["max_motor_torque"] = F32(max_torque)
        j["enable_limit"] = 1 if limit else 0
        j["lower_angle"] = F32(lower)
        j["upper_angle"] = F32(upper)
    return World(bodies, fixtures, joints)


def chains(n_worlds=1, links=10, seed=6, motor=True, limit=False, extra=0):
    """Revolute chains of boxes (half-extents 0.5 x 0.125 x 0.125) with motors, one per world."""
    s = SceneSet("chains")
    g = s.add_hull(hull_lib.box_hull(50.0, 0.5, 50.0))
    b = s.add_hull(hull_lib.box_hull(0.5, 0.125, 0.125))
    rng = np.random.default_rng(seed)
    for _ in range(n_worlds):
        s.worlds.append(chain_world(rng, links, g, b, motor, limit, extra))
    return s


def mixed_world(rng, n_bodies, ground_hull, link_hull, hull_indices, links=10, sleeping=False):
    """One world of the mixed-shapes workload: a static ground box; `n_bodies` dynamic bodies,
    each with equal probability a sphere (r ~ U[0.2, 0.6]), a capsule (segment length ~ U[0.5, 2]
    along local x, radius 0.25) or a hull drawn from `hull_indices`, at seeded uniform positions
    and orientations in a drop box of side L = 1.5 n^(1/3) m (height 2L); and a `links`-link
    revolute chain of boxes with motors (speed pi/2 rad/s, max torque 50 N m, no limits) hanging
    from the ground body beside the drop box."""
    nb = 1 + n_bodies + links
    bodies = body_array(nb)
    fixtures = fixture_array(nb)
    _ground(bodies, fixtures, ground_hull)
    L = 1.5 * n_bodies ** (1.0 / 3.0)
    sl = slice(1, 1 + n_bodies)
    bodies["type"][sl] = capi.BX_BODY_DYNAMIC
    bodies["position"][sl, 0] = rng.uniform(-L / 2, L / 2, n_bodies)
    bodies["position"][sl, 1] = rng.uniform(1.0, 1.0 + 2 * L, n_bodies)
    bodies["position"][sl, 2] = rng.uniform(-L / 2, L / 2, n_bodies)
    bodies["orientation"][sl] = random_orientations(rng, n_bodies)
    fixtures["body"][sl] = np.arange(1, 1 + n_bodies)
    fixtures["density"][sl] = 1.0
    kind = rng.integers(0, 3, n_bodies)
    radius = rng.uniform(0.2, 0.6, n_bodies).astype(F32)
    half = (0.5 * rng.uniform(0.5, 2.0, n_bodies)).astype(F32)
    hull = np.asarray(hull_indices, dtype=np.uint32)[rng.integers(0, len(hull_indices), n_bodies)]
    f = fixtures[sl]
    sph, cap, hul = kind == 0, kind == 1, kind == 2
    f["shape_type"][sph] = capi.BX_SHAPE_SPHERE
    f["radius"][sph] = radius[sph]
    f["shape_type"][cap] = capi.BX_SHAPE_CAPSULE
    f["radius"][cap] = F32(0.25)
    f["vertex1"][cap, 0] = -half[cap]
    f["vertex2"][cap, 0] = half[cap]
    f["shape_type"][hul] = capi.BX_SHAPE_HULL
    f["shape_index"][hul] = hull[hul]
    fixtures[sl] = f
    # the chain
    ox, oy = L / 2 + 2.0, 2 * L + 2.0
    cl = slice(1 + n_bodies, nb)
    bodies["type"][cl] = capi.BX_BODY_DYNAMIC
    bodies["position"][cl, 0] = ox + np.arange(links) + 0.5
    bodies["position"][cl, 1] = oy
    fixtures["body"][cl] = np.arange(1 + n_bodies, nb)
    fixtures["shape_type"][cl] = capi.BX_SHAPE_HULL
    fixtures["shape_index"][cl] = link_hull
    fixtures["density"][cl] = 1.0
    joints = joint_array(links)
    joints["body_a"][0] = 0
    joints["body_a"][1:] = np.arange(1 + n_bodies, nb - 1)
    joints["body_b"] = np.arange(1 + n_bodies, nb)
    joints["use_initialize"] = 1
    joints["axis"] = (0.0, 0.0, 1.0)
    joints["anchor"][:, 0] = ox + np.arange(links)
    joints["anchor"][:, 1] = oy
    joints["enable_motor"] = 1
    joints["motor_speed"] = F32(0.5 * np.pi)
    joints["max_motor_torque"] = F32(50.0)
    return World(bodies, fixtures, joints, flags=capi.BX_WORLD_SLEEPING if sleeping else 0)


def mixed_worlds(n_worlds=16384, seed=8, median=200, sigma=0.8, clip=(20, 2000), links=10, hull_count=1024,
                 sleeping=False, sizes=None):
    """The batched mixed-shapes workload: per-world body counts round(LogNormal(ln median, sigma))
    clipped to `clip` (or the given `sizes`), shapes from a seeded library of `hull_count` hulls of
    8-16 points in the unit ball, one motor chain per world."""
    s = SceneSet("mixed_worlds")
    g = s.add_hull(hull_lib.box_hull(50.0, 0.5, 50.0))
    link = s.add_hull(hull_lib.box_hull(0.5, 0.125, 0.125))
    base = len(s.hulls)
    s.hulls.extend(hull_lib.random_hull_library(hull_count, seed=seed + 1))
    hull_indices = np.arange(base, base + hull_count)
    rng = np.random.default_rng(seed)
    if sizes is None:
        sizes = np.clip(np.round(rng.lognormal(np.log(median), sigma, n_worlds)), clip[0], clip[1]).astype(int)
    for n in sizes:
        s.worlds.append(mixed_world(rng, int(n), g, link, hull_indices, links, sleeping))
    return s


def joint_zoo(n_worlds=1, seed=21, sleeping=False):
    """Every joint type of the device path in one world, on a ground box: boxes hung from the ground
    and from each other by a sphere joint, a rigid and a soft weld, a rigid and a soft spring, a motor
    joint pulling a box to an offset, a friction joint holding a box sliding on the ground, a mouse
    joint dragging a box to a target, a revolute joint, a swinging box held inside a cone, a box driven
    along a rail by a prismatic motor against its limits with another box on an equal-limit prismatic
    below it, and a chassis on a sprung, motorized wheel; some frames from Initialize, some given.
    Each world is jittered by its seed."""
    s = SceneSet("joint_zoo")
    g = s.add_hull(hull_lib.box_hull(30.0, 0.5, 30.0))
    b = s.add_hull(hull_lib.box_hull(0.4, 0.3, 0.35))
    rng = np.random.default_rng(seed)
    J = capi
    for _ in range(n_worlds):
        n = 18
        bodies = body_array(n)
        fixtures = fixture_array(n)
        _ground(bodies, fixtures, g)
        pos = np.array([[-6, 6, 0], [-6, 4.8, 0.2], [-3, 6, 0], [-3, 4.9, 0.1], [0, 6, 0], [0, 4.6, 0],
                        [3, 6, 0], [3, 4.4, 0.3], [6, 0.81, 0], [-8, 0.81, 3], [8, 3, -3], [9, 6, 0],
                        [-6, 8, 4], [-3, 9, 4], [3, 2, 4], [3, 1.2, 4], [-3, 7.9, 4]],
                       dtype=F32) + rng.uniform(-0.05, 0.05, (n - 1, 3)).astype(F32)
        sl = slice(1, n)
        bodies["type"][sl] = J.BX_BODY_DYNAMIC
        bodies["position"][sl] = pos
        bodies["orientation"][sl] = random_orientations(rng, n - 1) if rng.random() < 0.0 else (0, 0, 0, 1)
        bodies["linear_velocity"][9] = (3.0, 0.0, 1.0)      # slides on the ground against its friction joint
        bodies["angular_velocity"][9] = (0.0, 4.0, 0.0)
        bodies["linear_velocity"][13] = (4.0, 0.0, 1.0)     # swings out against its cone
        fixtures["body"][sl] = np.arange(1, n)
        fixtures["shape_type"][sl] = J.BX_SHAPE_HULL
        fixtures["shape_index"][sl] = b
        fixtures["density"][sl] = 1.0
        joints = joint_array(16)
        spec = [
            # (type, a, b, use_initialize, anchor)
            (J.BX_JOINT_SPHERE, 0, 1, 1, (-6, 6.6, 0)),
            (J.BX_JOINT_SPHERE, 1, 2, 0, None),
            (J.BX_JOINT_WELD, 0, 3, 1, (-3, 6.6, 0)),
            (J.BX_JOINT_WELD, 3, 4, 1, (-3, 5.45, 0.05)),
            (J.BX_JOINT_SPRING, 0, 5, 1, (0, 7.5, 0)),
            (J.BX_JOINT_SPRING, 5, 6, 1, (0, 5.7, 0)),
            (J.BX_JOINT_MOTOR, 0, 7, 1, None),
            (J.BX_JOINT_MOTOR, 7, 8, 0, None),
            (J.BX_JOINT_FRICTION, 0, 10, 1, (-8, 0.5, 3)),
            (J.BX_JOINT_MOUSE, 0, 11, 0, None),
            (J.BX_JOINT_REVOLUTE, 11, 12, 1, (8.5, 3, -3)),
            (J.BX_JOINT_CONE, 0, 13, 1, (-6, 9, 4)),
            (J.BX_JOINT_PRISMATIC, 0, 14, 1, (-3, 9, 4)),
            (J.BX_JOINT_WHEEL, 15, 16, 1, (3, 1.2, 4)),
            (J.BX_JOINT_PRISMATIC, 14, 17, 0, None),
            (J.BX_JOINT_CONE, 13, 0, 0, None),
        ]
        for k, (t, a, bb, init, anchor) in enumerate(spec):
            j = joints[k]
            j["type"], j["body_a"], j["body_b"], j["use_initialize"] = t, a, bb, init
            if anchor is not None:
                j["anchor"] = anchor
        joints[1]["local_anchor_a"] = (0.0, -0.5, 0.0)
        joints[1]["local_anchor_b"] = (0.1, 0.5, 0.0)
        joints[3]["frequency_hz"], joints[3]["damping_ratio"] = F32(3.0), F32(0.4)
        joints[4]["anchor_b"] = (0.05, 5.0, 0.02)
        joints[5]["anchor_b"] = (0.0, 5.0, 0.0)
        joints[5]["frequency_hz"], joints[5]["damping_ratio"] = F32(2.0), F32(0.3)
        joints[6]["max_force"], joints[6]["max_torque"] = F32(200.0), F32(50.0)
        joints[7]["linear_offset"] = (0.5, -1.5, 0.0)
        joints[7]["angular_offset"] = (0.0, 0.38268343, 0.0, 0.9238795)
        joints[7]["max_force"], joints[7]["max_torque"], joints[7]["correction_factor"] = F32(80.0), F32(20.0), F32(0.3)
        joints[8]["max_force"], joints[8]["max_torque"] = F32(5.0), F32(2.0)
        joints[9]["target"] = (7.0, 5.0, -2.0)
        joints[9]["max_force"] = F32(300.0)
        joints[9]["frequency_hz"], joints[9]["damping_ratio"] = F32(6.0), F32(0.8)
        joints[10]["axis"] = (0.0, 0.0, 1.0)
        joints[10]["collide_connected"] = 1
        joints[11]["axis"], joints[11]["cone_angle"], joints[11]["enable_limit"] = (0.0, -1.0, 0.0), F32(0.6), 1
        joints[12]["axis"] = (1.0, 0.0, 0.3)
        joints[12]["enable_limit"], joints[12]["lower_translation"], joints[12]["upper_translation"] = 1, -1.0, 1.5
        joints[12]["enable_motor"], joints[12]["motor_speed"], joints[12]["max_force"] = 1, F32(2.0), F32(50.0)
        joints[13]["axis"], joints[13]["axis_b"] = (0.0, 1.0, 0.0), (0.0, 0.0, 1.0)
        joints[13]["enable_motor"], joints[13]["motor_speed"], joints[13]["max_motor_torque"] = 1, F32(5.0), F32(10.0)
        joints[13]["frequency_hz"], joints[13]["damping_ratio"] = F32(4.0), F32(0.7)
        joints[14]["local_anchor_a"], joints[14]["local_anchor_b"] = (0.0, -0.55, 0.0), (0.0, 0.55, 0.0)
        joints[14]["local_axis_a"] = (0.0, 1.0, 0.0)
        joints[14]["enable_limit"] = 1          # lower = upper = 0: equal limits
        joints[15]["local_anchor_b"] = (-6.0, 9.0, 4.0)    # the cone 11's anchor, on body 13 too
        joints[15]["local_anchor_a"] = np.array([-6.0, 9.0, 4.0], dtype=F32) - pos[12]
        joints[15]["cone_angle"] = F32(1.2)     # given frames, limit off
        s.worlds.append(World(bodies, fixtures, joints, flags=capi.BX_WORLD_SLEEPING if sleeping else 0))
    return s
