"""The calls of tests/drop_in/spawn_apply.cpp replayed on the oracle, and the lines that program must print.

The replica keeps the drop-in world's rules: calls are queued with the creation indices they name when
they are made; the first time the queue is applied, its leading creations and world settings become the
world's descriptor, and the rest go through the oracle's topology and body edits in call order. Values the
program computes on the host from state it read back (a point on a body, a joint anchor) are computed
here in binary32 with the same operation order.
"""
import numpy as np

from bounce_amd import capi, hulls as hull_lib, scenes
from oracle.oracle import CpuWorld

F = np.float32
T = capi


def _cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], dtype=F)


def world_point(state, local):
    """b3Mul(xf, p) = b3Mul(q, p) + translation, b3Mul(q, v) = v + s t + q.v x t with t = 2 q.v x v."""
    q = np.asarray(state["orientation"], dtype=F)
    v = np.asarray(local, dtype=F)
    t = F(2) * _cross(q[:3], v)
    r = (v + q[3] * t) + _cross(q[:3], t)
    return r + np.asarray(state["position"], dtype=F)


class Replica:
    def __init__(self, kind="oracle"):
        self.kind = kind
        self.scene = scenes.SceneSet("drop_in")
        self.hull_of = {}
        self.hulls_on_world = 0
        self.queue = []
        self.new_bodies, self.new_fixtures, self.new_joints = [], [], []
        self.world = None
        self.bodies = []    # body ids, creation order
        self.fixtures = []  # (fixture id, body id), creation order
        self.joints = []    # (body a id, body b id), creation order

    # ---------------------------------------------------------------- calls
    def create_body(self, bid, pos, dynamic=True, q=(0.0, 0.0, 0.0, 1.0)):
        d = scenes.body_array(1)
        d["type"] = T.BX_BODY_DYNAMIC if dynamic else T.BX_BODY_STATIC
        d["position"] = pos
        d["orientation"] = q
        self.new_bodies.append(d)
        self.queue.append(("t", T.BX_TOPO_CREATE_BODY, len(self.new_bodies) - 1))
        self.bodies.append(bid)

    def _hull(self, ext):
        if ext not in self.hull_of:
            self.hull_of[ext] = self.scene.add_hull(hull_lib.box_hull(*ext))
        return self.hull_of[ext]

    def create_fixture(self, bid, fid, kind, density=0.0, friction=0.3, ext=None, radius=0.0, v1=(0, 0, 0),
                       v2=(0, 0, 0)):
        d = scenes.fixture_array(1)
        d["body"] = self.bodies.index(bid)
        d["density"] = F(density)
        d["friction"] = F(friction)
        if kind == "hull":
            d["shape_type"] = T.BX_SHAPE_HULL
            d["shape_index"] = self._hull(ext)
        elif kind == "sphere":
            d["shape_type"] = T.BX_SHAPE_SPHERE
            d["radius"] = F(radius)
            d["vertex1"] = v1
        else:
            d["shape_type"] = T.BX_SHAPE_CAPSULE
            d["radius"] = F(radius)
            d["vertex1"] = v1
            d["vertex2"] = v2
        self.new_fixtures.append(d)
        self.queue.append(("t", T.BX_TOPO_CREATE_FIXTURE, len(self.new_fixtures) - 1))
        self.fixtures.append((fid, bid))

    def create_joint(self, a, b, kind=T.BX_JOINT_REVOLUTE, **fields):
        j = scenes.joint_array(1)
        j["type"] = kind
        j["body_a"], j["body_b"] = self.bodies.index(a), self.bodies.index(b)
        for k, v in fields.items():
            j[k] = v
        self.new_joints.append(j)
        self.queue.append(("t", T.BX_TOPO_CREATE_JOINT, len(self.new_joints) - 1))
        self.joints.append((a, b))

    def destroy_body(self, bid):
        self.queue.append(("t", T.BX_TOPO_DESTROY_BODY, self.bodies.index(bid)))
        self.joints = [j for j in self.joints if bid not in j]
        self.fixtures = [f for f in self.fixtures if f[1] != bid]
        self.bodies.remove(bid)

    def destroy_joint(self, k):
        self.queue.append(("t", T.BX_TOPO_DESTROY_JOINT, k))
        del self.joints[k]

    def destroy_fixture(self, fid):
        k = [f[0] for f in self.fixtures].index(fid)
        self.queue.append(("t", T.BX_TOPO_DESTROY_FIXTURE, k))
        del self.fixtures[k]

    def edit(self, bid, op, v=(0.0, 0.0, 0.0), p=None, flag=False):
        body = 0 if bid is None else self.bodies.index(bid)
        self.queue.append(("e", (body, op, v, p, int(flag))))

    # ---------------------------------------------------------------- applying the queue
    def flush(self):
        k = 0
        if self.world is None:
            bodies, fixtures, joints = [], [], []
            gravity, flags = [F(0), F(-9.8), F(0)], 0
            while k < len(self.queue):
                c = self.queue[k]
                if c[0] == "t":
                    src = {T.BX_TOPO_CREATE_BODY: (bodies, self.new_bodies),
                           T.BX_TOPO_CREATE_FIXTURE: (fixtures, self.new_fixtures),
                           T.BX_TOPO_CREATE_JOINT: (joints, self.new_joints)}.get(c[1])
                    if src is None:
                        break
                    src[0].append(src[1][c[2]])
                else:
                    body, op, v, p, flag = c[1]
                    if op == T.BX_EDIT_WORLD_GRAVITY:
                        gravity = list(v)
                    elif op == T.BX_EDIT_WORLD_SLEEPING:
                        flags = flags | T.BX_WORLD_SLEEPING if flag else flags & ~T.BX_WORLD_SLEEPING
                        if not flag:
                            for b in bodies:
                                b["flags"] |= T.BX_BODY_AWAKE
                    elif op != T.BX_EDIT_WORLD_CLEAR_FORCES:
                        break
                k += 1
            if not bodies:
                return
            self.scene.worlds = [scenes.World(np.concatenate(bodies), np.concatenate(fixtures)
                                              if fixtures else scenes.fixture_array(0),
                                              np.concatenate(joints) if joints else scenes.joint_array(0),
                                              gravity=gravity, flags=flags)]
            self.world = CpuWorld(self.kind, self.scene, 0)
            self.hulls_on_world = len(self.scene.hulls)
        while k < len(self.queue):
            e = k
            if self.queue[k][0] == "t":
                ops = []
                while e < len(self.queue) and self.queue[e][0] == "t":
                    ops.append(self.queue[e][1:])
                    e += 1
                adds = scenes.World(np.concatenate(self.new_bodies) if self.new_bodies else scenes.body_array(0),
                                    np.concatenate(self.new_fixtures) if self.new_fixtures else scenes.fixture_array(0),
                                    np.concatenate(self.new_joints) if self.new_joints else scenes.joint_array(0))
                grown = len(self.scene.hulls) > self.hulls_on_world
                self.world.edit_topology(ops, adds, self.scene if grown else None)
                self.hulls_on_world = len(self.scene.hulls)
            else:
                rows = []
                while e < len(self.queue) and self.queue[e][0] == "e":
                    rows.append(self.queue[e][1])
                    e += 1
                self.world.edit_bodies(rows)
            k = e
        self.queue = []
        self.new_bodies, self.new_fixtures, self.new_joints = [], [], []

    def state(self, bid):
        self.flush()
        return self.world.bodies()[self.bodies.index(bid)]

    def step(self):
        self.flush()
        self.world.step(1.0 / 60.0, 8, 2)


def _g(x):
    return "%.9g" % float(x)


def spawn_apply_lines(kind="oracle"):
    """What tests/drop_in/spawn_apply.cpp prints, from the oracle (or the reference build, `kind` "ref32")."""
    out = []
    R = Replica(kind)
    R.edit(None, T.BX_EDIT_WORLD_SLEEPING, flag=True)
    R.create_body(0, (0.0, 0.0, 0.0), dynamic=False)
    R.create_fixture(0, 100, "hull", ext=(12.0, 1.0, 12.0))

    def add_box(bid, pos, density, friction):
        R.create_body(bid, pos)
        R.create_fixture(bid, 100 + bid, "hull", density=density, friction=friction, ext=(0.5, 0.5, 0.5))

    def print_mass(when, bid):
        R.flush()
        m = R.world.mass()[R.bodies.index(bid)]
        vals = [m[0], m[1], m[2], m[3], m[4], m[5], m[9], m[13], m[6], m[7], m[10]]
        out.append(("mass %s body %d = %s %s center %s %s %s inertia %s %s %s %s %s %s" % ((when, bid) + tuple(
            _g(x) for x in vals))))

    stack = [1, 2, 3, 4, 5]
    for i in range(5):
        add_box(1 + i, (0.0, F(1.5) + F(1.0) * F(i), 0.0), 1.0, 0.4)
    print_mass("start", 1)
    R.create_body(6, (3.0, 1.5, 0.0))
    R.create_fixture(6, 106, "sphere", density=2.0, radius=0.5)
    joint = False
    for step in range(180):
        R.edit(6, T.BX_EDIT_FORCE_TO_CENTER, (4.0, 0.0, 0.5), flag=True)
        if stack[4]:
            p = world_point(R.state(5), (0.5, 0.5, 0.0))
            R.edit(5, T.BX_EDIT_FORCE, (0.0, 0.0, 2.0), tuple(p) + (0.0,), flag=False)
        if step == 30:
            add_box(7, (0.3, 8.0, 0.2), 1.0, 0.3)
            R.create_body(8, (-1.0, 9.0, 0.1), q=(0.0, 0.0, 0.38268343, 0.9238795))
            R.create_fixture(8, 108, "capsule", density=1.0, radius=0.25, v1=(-0.5, 0.0, 0.0), v2=(0.5, 0.0, 0.0))
        if step == 50:
            c = np.asarray(R.state(7)["world_center"], dtype=F) + np.array([0.0, 0.2, 0.0], dtype=F)
            R.edit(7, T.BX_EDIT_LINEAR_IMPULSE, (1.0, 2.0, 0.0), tuple(c) + (0.0,), flag=True)
            R.edit(8, T.BX_EDIT_ANGULAR_VELOCITY, (0.0, 3.0, 0.0))
        if step == 70:
            a = np.asarray(R.state(7)["position"], dtype=F)
            b = np.asarray(R.state(8)["position"], dtype=F)
            R.create_joint(7, 8, use_initialize=1, axis=(0.0, 0.0, 1.0), anchor=F(0.5) * (a + b))
            joint = True
        if step == 90:
            R.destroy_body(3)
            stack[2] = None
        if step == 100:
            R.edit(6, T.BX_EDIT_TRANSFORM, (-3.0, 3.0, 0.0), (0.0, 0.0, 0.0, 1.0))
        if step == 110:
            R.destroy_joint(0)
            joint = False
        if step == 120:
            add_box(9, (2.0, 6.0, -2.0), 1.0, 0.3)
            R.create_fixture(9, 110, "sphere", density=3.0, radius=0.3, v1=(0.0, 0.7, 0.0))
            print_mass("compound", 9)
        if step == 130:
            R.destroy_fixture(110)
            print_mass("one fixture", 9)
        if step == 140:
            R.edit(None, T.BX_EDIT_WORLD_GRAVITY, (0.5, -6.0, 0.0))
            R.edit(8, T.BX_EDIT_FIXED_ROTATION, (1.0, 0.0, 1.0))
            R.edit(7, T.BX_EDIT_LINEAR_DAMPING, (0.2, 0.2, 0.2))
        if step == 150:
            if stack[4]:
                R.edit(5, T.BX_EDIT_AWAKE, flag=False)
            R.edit(6, T.BX_EDIT_SLEEPING_ALLOWED, flag=False)
            target = np.asarray(R.state(8)["position"], dtype=F) + np.array([0.0, 2.0, 0.0], dtype=F)
            R.create_joint(0, 8, T.BX_JOINT_MOUSE, target=target, max_force=200.0, frequency_hz=5.0,
                           damping_ratio=0.7)
            R.create_joint(7, 9, T.BX_JOINT_SPRING, use_initialize=1, anchor=R.state(7)["position"],
                           anchor_b=R.state(9)["world_center"], frequency_hz=2.0, damping_ratio=0.3)
            mid = F(0.5) * (np.asarray(R.state(1)["position"], dtype=F) + np.asarray(R.state(2)["position"], dtype=F))
            R.create_joint(1, 2, T.BX_JOINT_WELD, use_initialize=1, anchor=mid, correction_factor=0.2)
            R.create_joint(0, 6, T.BX_JOINT_PRISMATIC, use_initialize=1, anchor=R.state(6)["position"],
                           axis=(1.0, 0.0, 0.0), enable_limit=1, lower_translation=-1.0, upper_translation=2.0,
                           correction_factor=0.2)
        if step == 160:
            R.edit(None, T.BX_EDIT_WORLD_SLEEPING, flag=False)
        if step == 170:
            R.destroy_joint(0)
        R.step()
        st = R.world.bodies()
        for bid in reversed(R.bodies):
            s = st[R.bodies.index(bid)]
            vals = (list(s["position"]) + list(s["orientation"]) + list(s["linear_velocity"])
                    + list(s["angular_velocity"]))
            out.append("step %d body %d p %s %s %s q %s %s %s %s v %s %s %s w %s %s %s awake %d sleep %s" % (
                (step + 1, bid) + tuple(_g(x) for x in vals) + (1 if s["flags"] & T.BX_BODY_AWAKE else 0,
                                                                _g(s["sleep_time"]))))
        if step % 10 == 9:
            cs = R.world.contacts()
            out.append("contacts %d count %d" % (step + 1, len(cs)))
            for c in cs:
                line = "contact %d %d touching %d manifolds %d" % (
                    R.fixtures[c["fixture_a"]][0], R.fixtures[c["fixture_b"]][0], 1 if c["touching"] else 0,
                    min(int(c["manifold_count"]), 3))
                for m in range(min(int(c["manifold_count"]), 3)):
                    mf = c["manifolds"][m]
                    line += " [%d" % mf["point_count"]
                    for k in range(int(mf["point_count"])):
                        line += " %s %d" % (_g(mf["points"][k]["normal_impulse"]), mf["points"][k]["key"][0])
                    line += " t %s %s]" % (_g(mf["tangent_impulse"][0]), _g(mf["tangent_impulse"][1]))
                out.append(line)

    fid = [f[0] for f in R.fixtures]
    not_ground = np.array([f != 100 for f in fid], dtype=np.uint8)

    def hit_text(h):
        return "%d %s" % (fid[h["fixture"]], " ".join(_g(x) for x in list(h["point"]) + list(h["normal"])
                                                      + [h["fraction"]]))

    down1, down2 = (0.1, 20.0, 0.05), (0.1, -5.0, 0.05)
    h = R.world.raycast_single(down1, down2)
    out.append("single none" if h["fixture"] == T.BX_NULL_INDEX else "single " + hit_text(h))
    h = R.world.raycast_single((-8.0, 5.0, 7.0), (-8.0, -5.0, 7.0), mask=not_ground)
    out.append("single filtered none" if h["fixture"] == T.BX_NULL_INDEX else "single filtered %d" % fid[h["fixture"]])
    slant1, slant2 = (-6.0, 4.0, 0.3), (6.0, 0.5, -0.2)
    out.append("ray down")
    out += ["hit " + hit_text(h) for h in R.world.raycast(down1, down2)]
    out.append("ray slanted")
    out += ["hit " + hit_text(h) for h in R.world.raycast(slant1, slant2)]
    out.append("ray slanted without ground")
    out += ["hit " + hit_text(h) for h in R.world.raycast(slant1, slant2, mask=not_ground)]
    best, frac, reports, maxf = -1, F(1), 0, F(1)
    for h in R.world.raycast(slant1, slant2):
        if h["fraction"] > maxf:
            continue
        best, frac, reports, maxf = fid[h["fixture"]], h["fraction"], reports + 1, h["fraction"]
    out.append("closest %d %s reports %d" % (best, _g(frac), reports))
    out.append("query" + "".join(" %d" % fid[i] for i in R.world.query_aabb((-2.0, 0.0, -2.0), (2.0, 4.0, 2.0))))

    def cast(shape_type, pos, q, disp, radius=0.0, v1=(0, 0, 0), v2=(0, 0, 0), index=0):
        c = T.bx_shape_cast_query()
        c.world, c.mask = 0, T.BX_NULL_INDEX
        c.shape.shape_type, c.shape.shape_index, c.shape.radius = shape_type, index, F(radius)
        c.shape.vertex1[:], c.shape.vertex2[:] = list(v1), list(v2)
        c.position[:], c.orientation[:], c.displacement[:] = list(pos), list(q), list(disp)
        return c

    def cast_text(h):
        return "%d %s" % (fid[h["fixture"]], " ".join(_g(x) for x in list(h["point"]) + list(h["normal"])
                                                      + [h["fraction"]]))

    h = R.world.shape_cast_single(cast(T.BX_SHAPE_SPHERE, (0.2, 12.0, 0.1), (0, 0, 0, 1), (0.0, -14.0, 0.0), 0.4))
    out.append("cast single none" if h["fixture"] == T.BX_NULL_INDEX else "cast single " + cast_text(h))
    grown = R._hull((0.3, 0.2, 0.4))  # the facade sends the new hull before casting it
    R.world.edit_topology([], scenes.World(scenes.body_array(0), scenes.fixture_array(0)), R.scene)
    out.append("cast box")
    out += ["cast hit " + cast_text(h) for h in R.world.shape_cast(
        cast(T.BX_SHAPE_HULL, (-6.0, 2.5, 0.2), (0.0, 0.0, 0.38268343, 0.9238795), (12.0, -1.0, 0.0), index=grown))]
    out.append("cast capsule without ground")
    out += ["cast hit " + cast_text(h) for h in R.world.shape_cast(
        cast(T.BX_SHAPE_CAPSULE, (1.0, 9.0, -0.5), (0, 0, 0, 1), (-1.5, -10.0, 0.8), 0.2, (-0.4, 0, 0), (0.4, 0, 0)),
        mask=not_ground)]
    R.step()
    st = R.world.bodies()
    for bid in reversed(R.bodies):
        s = st[R.bodies.index(bid)]
        out.append("last body %d p %s %s %s q %s %s %s %s" % ((bid,) + tuple(
            _g(x) for x in list(s["position"]) + list(s["orientation"]))))
    out.append("bodies %d joints %d contacts %d" % (len(R.bodies), len(R.joints), R.world.counts().contact_count))
    return out
