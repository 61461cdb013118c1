"""The CPU oracle against what the mathematics fixes, independently of the reference's fixtures.

Each check is chosen so that a plausible slip in the restatement fails it:
  * free fall: symplectic Euler gives v_n = n h g and y_n = y_0 + h^2 g n(n+1)/2 (a position update
    that used the old velocity would give n(n-1)/2);
  * box mass data: m = 8 rho ex ey ez and I_xx = m (ey^2 + ez^2) / 3 about the centre (hull_shape.cpp
    ComputeMass; a swapped axis or a missing factor shows);
  * broadphase: every reported pair's fat AABBs overlap, pairs are sorted and unique, and every
    overlapping pair of fixtures whose bodies may collide owns a contact (brute force over all
    fixture pairs, aabb.h:347-367, body.cpp:240-263);
  * islands: with sleeping off the island count equals the number of connected components of
    dynamic bodies under touching contacts (union-find), and the solved contacts are exactly the
    touching contacts (world.cpp:303-406).
"""
import numpy as np
import pytest

from bounce_amd import capi, scenes
from oracle.oracle import CpuWorld

H = np.float32(1.0 / 60.0)


def single_box(y0=10.0):
    s = scenes.SceneSet("free_fall")
    hb = s.add_hull(scenes.hull_lib.box_hull(0.5, 0.5, 0.5))
    b = scenes.body_array(1)
    b["type"] = capi.BX_BODY_DYNAMIC
    b["position"][0] = (0.0, y0, 0.0)
    f = scenes.fixture_array(1)
    f["body"] = 0
    f["shape_type"] = capi.BX_SHAPE_HULL
    f["shape_index"] = hb
    f["density"] = 1.0
    s.worlds.append(scenes.World(b, f))
    return s


def test_free_fall_closed_form():
    w = CpuWorld("oracle", single_box())
    g = -9.8
    h = float(H)
    for n in range(1, 31):
        w.step()
        b = w.bodies()[0]
        v = float(b["linear_velocity"][1])
        y = float(b["position"][1])
        assert abs(v - n * h * g) <= 1e-5 * n, (n, v)
        assert abs(y - (10.0 + h * h * g * n * (n + 1) / 2)) <= 1e-5 * n, (n, y)
        assert b["angular_velocity"].tolist() == [0.0, 0.0, 0.0]


@pytest.mark.parametrize("ext", [(0.5, 0.5, 0.5), (0.3, 1.7, 2.9), (2.0, 0.25, 1.0)])
@pytest.mark.parametrize("density", [1.0, 2.5])
def test_box_mass_closed_form(ext, density):
    s = scenes.SceneSet("mass")
    hb = s.add_hull(scenes.hull_lib.box_hull(*ext))
    b = scenes.body_array(1)
    b["type"] = capi.BX_BODY_DYNAMIC
    b["position"][0] = (1.0, 2.0, 3.0)
    f = scenes.fixture_array(1)
    f["shape_type"] = capi.BX_SHAPE_HULL
    f["shape_index"] = hb
    f["density"] = density
    s.worlds.append(scenes.World(b, f))
    m = CpuWorld("oracle", s).mass()[0].astype(np.float64)
    ex, ey, ez = ext
    mass = 8.0 * density * ex * ey * ez
    assert abs(m[0] - mass) <= 1e-5 * mass
    assert abs(m[1] - 1.0 / mass) <= 1e-5 / mass
    assert np.all(np.abs(m[2:5]) <= 1e-5)  # centre of mass at the body origin
    inertia = m[5:14].reshape(3, 3)
    expect = np.diag([mass * (ey * ey + ez * ez) / 3, mass * (ex * ex + ez * ez) / 3, mass * (ex * ex + ey * ey) / 3])
    assert np.allclose(inertia, expect, rtol=1e-5, atol=1e-5 * mass)
    assert np.allclose(m[14:23].reshape(3, 3) @ inertia, np.eye(3), atol=1e-5)


def _fixture_bodies(scene):
    return scene.worlds[0].fixtures["body"].astype(np.int64)


def _overlap(a, b):
    return bool(np.all(a[:3] <= b[3:]) and np.all(b[:3] <= a[3:]))


@pytest.mark.parametrize("recipe", ["drop", "pyramid"])
def test_pairs_and_contacts_against_brute_force(recipe):
    scene = scenes.drop_boxes(60, seed=11, extent=(6.0, 12.0, 6.0)) if recipe == "drop" else scenes.pyramids(1, levels=8)
    w = CpuWorld("oracle", scene)
    fb = _fixture_bodies(scene)
    types = scene.worlds[0].bodies["type"]
    for step in range(1, 91):
        w.step()
        if step % 15:
            continue
        fat = w.fat_aabbs()
        pid = w.proxy_ids()
        fix_of_proxy = {int(p): i for i, p in enumerate(pid)}
        pairs = w.pairs()
        keys = [(int(p["proxy1"]), int(p["proxy2"])) for p in pairs]
        assert keys == sorted(set(keys))
        for p1, p2 in keys:
            assert p1 < p2
            assert _overlap(fat[fix_of_proxy[p1]], fat[fix_of_proxy[p2]])
        have = {tuple(sorted((int(c["fixture_a"]), int(c["fixture_b"])))) for c in w.contacts()}
        n = len(fat)
        for i in range(n):
            for j in range(i + 1, n):
                if fb[i] == fb[j] or not _overlap(fat[i], fat[j]):
                    continue
                if types[fb[i]] != capi.BX_BODY_DYNAMIC and types[fb[j]] != capi.BX_BODY_DYNAMIC:
                    continue
                assert (i, j) in have, f"step {step}: overlapping fixtures {i},{j} have no contact"


def _find(parent, x):
    while parent[x] != x:
        parent[x] = parent[parent[x]]
        x = parent[x]
    return x


@pytest.mark.parametrize("recipe", ["drop", "pyramid", "stack"])
def test_islands_against_union_find(recipe):
    if recipe == "drop":
        scene = scenes.drop_boxes(80, seed=5, extent=(5.0, 10.0, 5.0))
    elif recipe == "pyramid":
        scene = scenes.pyramids(1, levels=6)
    else:
        scene = scenes.box_stack(5)
    w = CpuWorld("oracle", scene)
    fb = _fixture_bodies(scene)
    types = scene.worlds[0].bodies["type"]
    nb = len(types)
    for step in range(1, 121):
        w.step()
        if step % 20:
            continue
        contacts = w.contacts()
        parent = list(range(nb))
        touching = set()
        for k, c in enumerate(contacts):
            if not c["touching"]:
                continue
            a, b = int(fb[c["fixture_a"]]), int(fb[c["fixture_b"]])
            touching.add(k)
            if types[a] == capi.BX_BODY_STATIC or types[b] == capi.BX_BODY_STATIC:
                continue
            parent[_find(parent, a)] = _find(parent, b)
        roots = {_find(parent, i) for i in range(nb) if types[i] == capi.BX_BODY_DYNAMIC}
        assert w.counts().island_count == len(roots), f"step {step}"
        solved = [int(i) for i in w.island_contacts()]
        assert len(solved) == len(set(solved))
        assert set(solved) == touching, f"step {step}"
