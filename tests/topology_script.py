"""Scripted topology changes between steps (bodies, fixtures and joints created and destroyed in a
running world), shared by the oracle test against the reference build and the device test against the
oracle.

Each scenario is a world and, for some steps, a list of operations (capi.BX_TOPO_*, index) with the
descriptors they name in an `adds` world. Creation indices in the descriptors and in destroy operations
are those of the world at that point, as the reference's object lists would be counted.
"""
import numpy as np

from bounce_amd import capi, hulls, scenes

T = capi


class Scenario:
    def __init__(self, name, scene, steps, edits, library=None):
        self.name = name
        self.scene = scene          # scenes.SceneSet with one world
        self.steps = steps
        self.edits = edits          # step -> (ops, adds World)
        self.library = library      # SceneSet whose library replaces the scene's from the first edit


def _body(pos, dynamic=True, q=(0.0, 0.0, 0.0, 1.0), v=(0.0, 0.0, 0.0)):
    b = scenes.body_array(1)
    b["type"] = capi.BX_BODY_DYNAMIC if dynamic else capi.BX_BODY_STATIC
    b["position"] = pos
    b["orientation"] = q
    b["linear_velocity"] = v
    return b


def _fixture(body, kind, index=0, radius=0.4, half=0.6, density=1.0, friction=0.3, offset=(0.0, 0.0, 0.0)):
    f = scenes.fixture_array(1)
    f["body"] = body
    f["density"] = density
    f["friction"] = friction
    if kind == "box":
        f["shape_type"] = capi.BX_SHAPE_HULL
        f["shape_index"] = index
    elif kind == "sphere":
        f["shape_type"] = capi.BX_SHAPE_SPHERE
        f["radius"] = radius
        f["vertex1"] = offset
    else:
        f["shape_type"] = capi.BX_SHAPE_CAPSULE
        f["radius"] = 0.25
        f["vertex1"] = (-half + offset[0], offset[1], offset[2])
        f["vertex2"] = (half + offset[0], offset[1], offset[2])
    return f


def _adds(bodies=(), fixtures=(), joints=()):
    b = np.concatenate(list(bodies)) if bodies else scenes.body_array(0)
    f = np.concatenate(list(fixtures)) if fixtures else scenes.fixture_array(0)
    j = np.concatenate(list(joints)) if joints else scenes.joint_array(0)
    return scenes.World(b, f, j)


def spawn():
    """A 6-level pyramid; boxes, a sphere, a capsule and a compound body (box + sphere, so ResetMass
    moves its centre) appear above it while it settles; then a sphere fixture is added to a resting
    pyramid box."""
    s = scenes.pyramids(1, levels=6, seed=3)
    nb = s.worlds[0].body_count  # 22: ground + 21 boxes
    box = 1  # the unit box hull of the pyramid scene
    e = {}
    b0 = [_body((-0.5 + i, 9.0 + 0.7 * i, 0.3)) for i in range(3)]
    f0 = [_fixture(nb + i, "box", box) for i in range(3)]
    e[20] = ([(T.BX_TOPO_CREATE_BODY, 0), (T.BX_TOPO_CREATE_FIXTURE, 0), (T.BX_TOPO_CREATE_BODY, 1),
              (T.BX_TOPO_CREATE_BODY, 2), (T.BX_TOPO_CREATE_FIXTURE, 1), (T.BX_TOPO_CREATE_FIXTURE, 2)],
             _adds(b0, f0))
    nb += 3
    e[40] = ([(T.BX_TOPO_CREATE_BODY, 0), (T.BX_TOPO_CREATE_FIXTURE, 0), (T.BX_TOPO_CREATE_BODY, 1),
              (T.BX_TOPO_CREATE_FIXTURE, 1)],
             _adds([_body((1.2, 11.0, -0.2)), _body((-1.0, 12.0, 0.1), q=(0.0, 0.0, 0.38268343, 0.9238795))],
                   [_fixture(nb, "sphere"), _fixture(nb + 1, "capsule")]))
    nb += 2
    e[60] = ([(T.BX_TOPO_CREATE_BODY, 0), (T.BX_TOPO_CREATE_FIXTURE, 0), (T.BX_TOPO_CREATE_FIXTURE, 1)],
             _adds([_body((0.2, 14.0, 0.0))],
                   [_fixture(nb, "box", box), _fixture(nb, "sphere", radius=0.3, offset=(0.0, 0.8, 0.0), density=2.0)]))
    nb += 1
    e[80] = ([(T.BX_TOPO_CREATE_FIXTURE, 0)], _adds(fixtures=[_fixture(4, "sphere", radius=0.2, offset=(0.0, 0.6, 0.0))]))
    return Scenario("spawn", s, 120, e)


def destroy():
    """Thirty boxes settle with sleeping on; resting bodies are destroyed (their contacts go and wake
    their neighbours), proxies are freed and taken again by new boxes, a fixture is destroyed, and a
    joint between two boxes is created and destroyed."""
    s = scenes.with_flags(scenes.drop_boxes(30, extent=(4.0, 8.0, 4.0), seed=12), capi.BX_WORLD_SLEEPING)
    box = 1
    e = {}
    e[100] = ([(T.BX_TOPO_DESTROY_BODY, 5)], _adds())
    e[110] = ([(T.BX_TOPO_DESTROY_BODY, 10), (T.BX_TOPO_DESTROY_BODY, 11)], _adds())
    nb = 31 - 3
    e[120] = ([(T.BX_TOPO_CREATE_BODY, 0), (T.BX_TOPO_CREATE_FIXTURE, 0), (T.BX_TOPO_CREATE_BODY, 1),
               (T.BX_TOPO_CREATE_FIXTURE, 1)],
              _adds([_body((0.3, 7.0, 0.2)), _body((-0.4, 9.0, -0.1))], [_fixture(nb, "box", box), _fixture(nb + 1, "box", box)]))
    nb += 2
    j = scenes.joint_array(1)
    j["body_a"], j["body_b"] = nb - 2, nb - 1
    j["use_initialize"] = 1
    j["axis"] = (0.0, 0.0, 1.0)
    j["anchor"] = (0.0, 8.0, 0.0)
    e[125] = ([(T.BX_TOPO_CREATE_JOINT, 0)], _adds(joints=[j]))
    e[140] = ([(T.BX_TOPO_DESTROY_FIXTURE, 7)], _adds())  # a resting box loses its only fixture
    e[150] = ([(T.BX_TOPO_DESTROY_JOINT, 0), (T.BX_TOPO_DESTROY_BODY, 2)], _adds())
    e[155] = ([(T.BX_TOPO_CREATE_BODY, 0), (T.BX_TOPO_CREATE_FIXTURE, 0)],
              _adds([_body((0.0, 10.0, 0.0))], [_fixture(nb - 1, "sphere", radius=0.5)]))
    return Scenario("destroy", s, 190, e)


def joints():
    """A motor chain: a joint in the middle is destroyed (the chain breaks, both links wake), a new
    joint ties the loose end to the ground, and a link with two joints is destroyed."""
    s = scenes.chains(1, 10)
    e = {}
    e[30] = ([(T.BX_TOPO_DESTROY_JOINT, 4)], _adds())
    j = scenes.joint_array(1)
    j["body_a"], j["body_b"] = 0, 10
    j["use_initialize"] = 1
    j["axis"] = (0.0, 0.0, 1.0)
    j["anchor"] = (9.5, 12.0, 0.0)
    j["enable_motor"] = 1
    j["motor_speed"] = np.float32(-0.5)
    j["max_motor_torque"] = np.float32(20.0)
    e[50] = ([(T.BX_TOPO_CREATE_JOINT, 0)], _adds(joints=[j]))
    e[70] = ([(T.BX_TOPO_DESTROY_BODY, 3)], _adds())
    return Scenario("joints", s, 120, e)


def library():
    """Spheres, capsules and boxes; a body whose shape is a hull that was not in the library joins
    mid-run (the library grows, earlier indices unchanged)."""
    s = scenes.mixed_drop(20, seed=9, extent=(5.0, 8.0, 5.0))
    grown = scenes.SceneSet("grown")
    grown.hulls = list(s.hulls)
    rng = np.random.default_rng(4)
    pts = rng.normal(size=(12, 3))
    pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    h = grown.add_hull(hulls.convex_hull((0.8 * pts).astype(np.float32)))
    nb = s.worlds[0].body_count
    e = {30: ([(T.BX_TOPO_CREATE_BODY, 0), (T.BX_TOPO_CREATE_FIXTURE, 0)],
              _adds([_body((0.0, 6.0, 0.0))], [_fixture(nb, "box", h)]))}
    return Scenario("library", s, 90, e, library=grown)


SCENARIOS = {f.__name__: f for f in (spawn, destroy, joints, library)}
