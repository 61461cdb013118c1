"""Seeded time-of-impact problems and shape casts, shared by the oracle tests (against the reference build)
and the device tests (against the oracle).

Problems pair spheres, capsules, triangles, boxes and random hulls; one shape moves towards the other so
that most problems end touching, and the rest cover the other outcomes (separated, overlapped at the
start, a root finder that gives up).
"""
import numpy as np

from bounce_amd import capi, hulls as hull_lib, scenes

F = np.float32


def library():
    """A scene whose hull library the problems index (a unit box, a flat box and four random hulls), with
    one world of a static box so that a batch can be made of it."""
    s = scenes.SceneSet("toi")
    s.add_hull(hull_lib.box_hull(0.5, 0.5, 0.5))
    s.add_hull(hull_lib.box_hull(1.5, 0.2, 0.8))
    rng = np.random.default_rng(77)
    for k in range(4):
        pts = rng.normal(size=(10 + 4 * k, 3))
        pts /= np.linalg.norm(pts, axis=1, keepdims=True)
        pts *= rng.uniform(0.4, 1.0, size=(len(pts), 1))
        s.add_hull(hull_lib.convex_hull((0.8 * pts).astype(F)))
    f = scenes.fixture_array(1)
    f["shape_type"] = capi.BX_SHAPE_HULL
    s.worlds.append(scenes.World(scenes.body_array(1), f))
    return s


def _quat(rng, angle=None):
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    a = rng.uniform(0, 2 * np.pi) if angle is None else angle
    return np.concatenate([np.sin(a / 2) * axis, [np.cos(a / 2)]]).astype(F)


def _qmul(a, b):
    av, aw, bv, bw = a[:3].astype(np.float64), float(a[3]), b[:3].astype(np.float64), float(b[3])
    v = np.cross(av, bv) + aw * bv + bw * av
    q = np.concatenate([v, [aw * bw - av @ bv]])
    return (q / np.linalg.norm(q)).astype(F)


def random_shape(rng, n_hulls, allow=("sphere", "capsule", "triangle", "hull")):
    s = capi.bx_cast_shape()
    kind = allow[rng.integers(len(allow))]
    if kind == "sphere":
        s.shape_type = capi.BX_SHAPE_SPHERE
        s.radius = F(rng.uniform(0.2, 0.8))
        s.vertex1[:] = list(rng.uniform(-0.1, 0.1, 3).astype(F))
    elif kind == "capsule":
        s.shape_type = capi.BX_SHAPE_CAPSULE
        s.radius = F(rng.uniform(0.1, 0.4))
        h = rng.uniform(0.2, 0.8)
        s.vertex1[:] = [F(-h), F(0), F(0)]
        s.vertex2[:] = [F(h), F(0), F(0)]
    elif kind == "triangle":
        s.shape_type = capi.BX_SHAPE_TRIANGLE
        s.radius = F(0)
        v = rng.uniform(-1.0, 1.0, size=(3, 3)).astype(F)
        s.vertex1[:], s.vertex2[:], s.vertex3[:] = list(v[0]), list(v[1]), list(v[2])
    else:
        s.shape_type = capi.BX_SHAPE_HULL
        s.shape_index = int(rng.integers(n_hulls))
        s.radius = F(0)
    return s


def toi_queries(n, seed, n_hulls=6):
    """n problems: one in three of each kind (sweeps with rotation, translations, boxes)."""
    rng = np.random.default_rng(seed)
    qs = (capi.bx_toi_query * n)()
    for i in range(n):
        q = qs[i]
        kind = i % 3
        q.kind = kind
        q.t_max = F(1.0)
        q.max_iterations = 20 if rng.random() < 0.9 else int(rng.integers(1, 4))
        q.shape1 = random_shape(rng, n_hulls)
        q.shape2 = random_shape(rng, n_hulls)
        start = rng.uniform(-0.3, 0.3, 3).astype(F) + F(3.0) * rng.normal(size=3).astype(F) / F(1.7)
        motion = (-start * F(rng.uniform(0.6, 1.6)) + rng.uniform(-0.5, 0.5, 3).astype(F)).astype(F)
        if rng.random() < 0.1:
            motion = -motion  # moving apart
        if rng.random() < 0.05:
            start = (F(0.05) * start).astype(F)  # overlapping at the start
        if kind == capi.BX_TOI_SWEEP:
            for sw, c0, c1, spin in ((q.sweep1, start, start + motion, 1.5), (q.sweep2, np.zeros(3, F),
                                                                            rng.uniform(-0.2, 0.2, 3), 0.5)):
                q0 = _quat(rng)
                sw.local_center[:] = list(rng.uniform(-0.2, 0.2, 3).astype(F))
                sw.center0[:] = list(np.asarray(c0, F))
                sw.center[:] = list(np.asarray(c1, F))
                sw.orientation0[:] = list(q0)
                sw.orientation[:] = list(_qmul(_quat(rng, rng.uniform(0, spin)), q0))
                sw.t0 = F(0)
            if rng.random() < 0.2:
                q.t_max = F(rng.uniform(0.3, 1.0))
        elif kind == capi.BX_TOI_TRANSLATION:
            q.position1[:] = list(start)
            q.orientation1[:] = list(_quat(rng))
            q.position2[:] = list(rng.uniform(-0.2, 0.2, 3).astype(F))
            q.orientation2[:] = list(_quat(rng))
            q.displacement1[:] = list(motion)
            q.displacement2[:] = list((rng.uniform(-0.3, 0.3, 3) if rng.random() < 0.3 else np.zeros(3)).astype(F))
        else:
            e1, e2 = rng.uniform(0.2, 1.0, 3).astype(F), rng.uniform(0.2, 1.0, 3).astype(F)
            q.lower1[:], q.upper1[:] = list(start - e1), list(start + e1)
            q.lower2[:], q.upper2[:] = list(-e2), list(e2)
            q.displacement1[:] = list(motion)
            if rng.random() < 0.3:
                q.displacement1[int(rng.integers(3))] = F(0)
            q.displacement2[:] = list((rng.uniform(-0.3, 0.3, 3) if rng.random() < 0.3 else np.zeros(3)).astype(F))
    return qs


def cast_queries(bodies, n, seed, n_hulls, world=0, mask=capi.BX_NULL_INDEX):
    """n shape casts aimed at the bodies of a world (its body states), from outside towards a body."""
    rng = np.random.default_rng(seed)
    qs = (capi.bx_shape_cast_query * n)()
    pos = np.asarray(bodies["position"], dtype=F)
    for i in range(n):
        q = qs[i]
        q.world = world
        q.mask = mask
        q.shape = random_shape(rng, n_hulls)
        target = pos[rng.integers(len(pos))] + rng.uniform(-0.5, 0.5, 3).astype(F)
        direction = rng.normal(size=3)
        direction /= np.linalg.norm(direction)
        if rng.random() < 0.5:
            direction = np.array([0.0, 1.0, 0.0]) + 0.3 * rng.normal(size=3)  # from above
        start = (target + rng.uniform(3.0, 6.0) * direction).astype(F)
        q.position[:] = list(start)
        q.orientation[:] = list(_quat(rng))
        q.displacement[:] = list(((target - start) * F(rng.uniform(1.0, 1.8))).astype(F))
    return qs
