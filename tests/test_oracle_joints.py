"""The oracle's sphere, weld, spring, mouse, friction, motor, cone, prismatic and wheel joints
(joints/*_joint.cpp) against the
reference build (oracle/_ref), bit for bit on every step of a world holding each of them (rigid and soft
welds and springs, frames from Initialize and given), with sleeping off and on; and what each joint
holds in the oracle's settled world: the sphere joint's anchors meet, the rigid weld keeps its
reference rotation, the rigid spring keeps its length, the mouse joint's anchor sits on its target
up to the soft constraint's sag, the cone holds its body inside the cone, the prismatic joint moves
its body along the axis inside the limits, and the wheel stays on the chassis's suspension axis.
The cone, prismatic and wheel joints are in the same world and compared with the reference build
the same way."""
import numpy as np
import pytest

from bounce_amd import capi, scenes
from oracle import oracle as O
from tests.parity import compare_worlds

needs_ref = pytest.mark.skipif(not O.available("ref32"), reason="reference build (make -C oracle ref) not present")


@needs_ref
@pytest.mark.parametrize("sleeping", [False, True])
def test_joint_zoo_oracle_equals_reference(sleeping):
    scene = scenes.joint_zoo(2, seed=21, sleeping=sleeping)
    for w in range(2):
        o = O.CpuWorld("oracle", scene, w)
        r = O.CpuWorld("ref32", scene, w)
        for step in range(300):
            o.step()
            r.step()
            d = compare_worlds(o, r, mask_cache=True)
            assert not d, f"world {w} step {step}: {d}"


def _qmul(a, b):
    av, aw, bv, bw = a[:3], a[3], b[:3], b[3]
    return np.concatenate([np.cross(av, bv) + aw * bv + bw * av, [aw * bw - av @ bv]])


def _rot(q, v):
    t = 2.0 * np.cross(q[:3], v)
    return v + q[3] * t + np.cross(q[:3], t)


def _world_point(b, local):
    return _rot(b["orientation"].astype(np.float64), np.asarray(local, np.float64)) + b["position"]


def test_joints_hold_what_they_constrain():
    scene = scenes.joint_zoo(1, seed=21)
    o = O.CpuWorld("oracle", scene)
    o.step(n=240)
    b = o.bodies()
    j = scene.worlds[0].joints
    q = lambda i: b[i]["orientation"].astype(np.float64)
    # sphere joint 1 (bodies 1 and 2, anchors given): the anchors meet
    assert np.linalg.norm(_world_point(b[1], j[1]["local_anchor_a"]) - _world_point(b[2], j[1]["local_anchor_b"])) < 0.01
    # rigid weld 2 (ground and body 3, from Initialize): body 3 keeps its initial orientation
    assert np.linalg.norm(np.abs(q(3)) - np.array([0, 0, 0, 1.0])) < 0.01
    # rigid spring 4 (ground and body 5): the anchors stay `length` apart
    pa = np.asarray(j[4]["anchor"], np.float64)
    pb = _world_point(b[5], _rot(np.array([0, 0, 0, 1.0]), np.asarray(j[4]["anchor_b"], np.float64)
                                 - scene.worlds[0].bodies[5]["position"]))
    length = np.linalg.norm(np.asarray(j[4]["anchor_b"], np.float64) - pa)
    assert abs(np.linalg.norm(pb - pa) - length) < 0.02
    # mouse 9 (body 11, with body 12 hanging from it): its anchor on the target up to the sag of a soft
    # constraint tuned to one body (g / omega^2) carrying two, doubled for the offset anchor
    target = np.asarray(j[9]["target"], np.float64)
    anchor0 = target - scene.worlds[0].bodies[11]["position"]
    omega = 2 * np.pi * float(j[9]["frequency_hz"])
    assert np.linalg.norm(_world_point(b[11], anchor0) - target) < 4 * 9.8 / omega ** 2
    # cone 11 (ground and body 13, from Initialize about -y, cone 0.6): the anchors meet, and body 13's
    # frame axis stays within half the cone angle of the ground's, up to the angular slop and the
    # solver's residual
    w0 = scene.worlds[0].bodies
    anchor = np.asarray(j[11]["anchor"], np.float64)
    assert np.linalg.norm(_world_point(b[13], anchor - w0[13]["position"]) - anchor) < 0.01
    axis = np.asarray(j[11]["axis"], np.float64)
    cosang = np.clip(_rot(q(13), axis) @ axis, -1.0, 1.0)
    assert np.arccos(cosang) <= 0.5 * float(j[11]["cone_angle"]) + 0.05
    # prismatic 12 (ground and body 14 along (1, 0, 0.3), limits [-1, 1.5]): the anchor moves along
    # the axis only, inside the limits
    anchor = np.asarray(j[12]["anchor"], np.float64)
    axis = np.asarray(j[12]["axis"], np.float64)
    axis /= np.linalg.norm(axis)
    d = _world_point(b[14], anchor - w0[14]["position"]) - anchor
    t = d @ axis
    assert -1.0 - 0.02 <= t <= 1.5 + 0.02
    assert np.linalg.norm(d - t * axis) < 0.05
    # wheel 13 (chassis 15, wheel 16): the wheel's anchor stays on the chassis's suspension axis
    anchor = np.asarray(j[13]["anchor"], np.float64)
    pa = _world_point(b[15], anchor - w0[15]["position"])
    pb = _world_point(b[16], anchor - w0[16]["position"])
    up = _rot(q(15), np.asarray(j[13]["axis"], np.float64))
    d = pb - pa
    assert np.linalg.norm(d - (d @ up) * up) < 0.02
