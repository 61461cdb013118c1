"""The oracle's world queries (oracle/src/orc_query.cpp) against the reference.

* tests/golden/queries.npz holds the reference build's own b3World::RayCastSingle, RayCast and
  QueryAABB results on five seeded worlds (spheres, capsules, boxes, random hulls, triangles, a
  scaled mesh) after stepping; the oracle must reproduce every hit bit for bit (fixture, fraction,
  point, normal) and every box result.
* Where the reference build is present, the same comparison runs on fresh random rays.
* Closed forms pin the shape casts independently on a hand-built world: vertical rays meet a
  sphere's, a box's, a capsule's and the ground's tops at the analytic fractions with normal +y, a
  horizontal ray enters the sphere at its radius, a ray above everything reports nothing, and the
  mask removes a fixture; QueryAABB equals the brute-force fat-AABB overlap set.
"""
import os

import numpy as np
import pytest

from bounce_amd import capi, scenes
from tests.golden.make_query_golden import QUERY_SCENES, query_set, run_queries

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "queries.npz")


@pytest.mark.parametrize("name", list(QUERY_SCENES))
def test_oracle_queries_match_reference_fixture(name):
    from oracle.oracle import CpuWorld
    g = np.load(GOLDEN)
    factory, steps = QUERY_SCENES[name]
    w = CpuWorld("oracle", factory())
    w.step(n=steps)
    single, hits, rc, found, bc = run_queries(w, g[f"{name}_rays"], g[f"{name}_boxes"], g[f"{name}_mask"],
                                              g[f"{name}_use_mask"])
    assert single.tobytes() == g[f"{name}_single"].tobytes()
    assert np.array_equal(rc, g[f"{name}_ray_counts"])
    assert hits.tobytes() == g[f"{name}_hits"].tobytes()
    assert np.array_equal(bc, g[f"{name}_box_counts"])
    assert np.array_equal(found, g[f"{name}_found"])


@pytest.mark.parametrize("name", ["mixed60", "mesh30_scaled", "mixed_world200"])
def test_oracle_queries_match_reference_build(name):
    from oracle.oracle import CpuWorld, available
    if not available("ref32"):
        pytest.skip("reference build not present")
    factory, steps = QUERY_SCENES[name]
    o, r = CpuWorld("oracle", factory()), CpuWorld("ref32", factory())
    o.step(n=steps // 2)
    r.step(n=steps // 2)
    rays, boxes, mask, use_mask = query_set(r.bodies(), r.counts().fixture_count, seed=7, n=240)
    a = run_queries(o, rays, boxes, mask, use_mask)
    b = run_queries(r, rays, boxes, mask, use_mask)
    for x, y in zip(a, b):
        assert x.tobytes() == y.tobytes()


def _probe_scene():
    """A ground box, a sphere (r 0.5) at (0, 3, 0), an axis-aligned unit box at (4, 3, 0) and a capsule
    (segment along x, half-length 1, r 0.25) at (-4, 3, 0), queried before any step."""
    from bounce_amd import hulls
    s = scenes.SceneSet("probe")
    g = s.add_hull(hulls.box_hull(50.0, 0.5, 50.0))
    box = s.add_hull(hulls.box_hull(0.5, 0.5, 0.5))
    bodies = scenes.body_array(4)
    fixtures = scenes.fixture_array(4)
    scenes._ground(bodies, fixtures, g)
    bodies["type"][1:] = capi.BX_BODY_DYNAMIC
    bodies["position"][1:] = [(0.0, 3.0, 0.0), (4.0, 3.0, 0.0), (-4.0, 3.0, 0.0)]
    fixtures["body"][1:] = [1, 2, 3]
    fixtures["density"][1:] = 1.0
    fixtures[1]["shape_type"], fixtures[1]["radius"] = capi.BX_SHAPE_SPHERE, 0.5
    fixtures[2]["shape_type"], fixtures[2]["shape_index"] = capi.BX_SHAPE_HULL, box
    fixtures[3]["shape_type"], fixtures[3]["radius"] = capi.BX_SHAPE_CAPSULE, 0.25
    fixtures[3]["vertex1"], fixtures[3]["vertex2"] = (-1.0, 0.0, 0.0), (1.0, 0.0, 0.0)
    s.worlds.append(scenes.World(bodies, fixtures))
    return s


def test_ray_closed_forms():
    """Vertical rays from y = 13 to y = -7 (length 20) hit the sphere's top at y = 3.5, the box's top
    face at y = 3.5, the capsule's cylinder at y = 3.25 and the ground's top face at y = 0, with
    upward normals; a ray above everything hits nothing."""
    from oracle.oracle import CpuWorld
    w = CpuWorld("oracle", _probe_scene())
    for x, fixture, top in [(0.0, 1, 3.5), (4.0, 2, 3.5), (-4.0, 3, 3.25), (10.0, 0, 0.0)]:
        p1 = np.array([x, 13.0, 0.0], np.float32)
        p2 = np.array([x, -7.0, 0.0], np.float32)
        hit = w.raycast_single(p1, p2)
        assert hit["fixture"] == fixture
        assert abs(hit["fraction"] - (13.0 - top) / 20.0) < 1e-6
        assert np.allclose(hit["normal"], [0.0, 1.0, 0.0], atol=1e-6)
        assert np.allclose(hit["point"], [x, top, 0.0], atol=1e-5)
        every = w.raycast(p1, p2)
        assert every[0].tobytes() == hit.tobytes()
        if fixture != 0:  # the ray goes on to the ground
            assert len(every) == 2 and every[1]["fixture"] == 0
    # a horizontal ray along z through the sphere's centre: entry at z = -0.5
    hit = w.raycast_single(np.array([0, 3, -5], np.float32), np.array([0, 3, 5], np.float32))
    assert hit["fixture"] == 1 and abs(hit["fraction"] - 0.45) < 1e-6
    assert np.allclose(hit["normal"], [0.0, 0.0, -1.0], atol=1e-6)
    miss = w.raycast_single(np.array([0, 500, 0], np.float32), np.array([1, 501, 0], np.float32))
    assert miss["fixture"] == capi.BX_NULL_INDEX
    assert len(w.raycast(np.array([0, 500, 0], np.float32), np.array([1, 501, 0], np.float32))) == 0
    # the mask stands for the filter: without the sphere, the vertical ray at x = 0 hits the ground
    mask = np.array([1, 0, 1, 1], np.uint8)
    assert w.raycast_single(np.array([0, 13, 0], np.float32), np.array([0, -7, 0], np.float32), mask)["fixture"] == 0


def test_query_aabb_is_the_fat_aabb_overlap_set():
    """QueryAABB reports exactly the fixtures whose fat AABB overlaps the box (brute force)."""
    from oracle.oracle import CpuWorld
    w = CpuWorld("oracle", scenes.mixed_drop(60, seed=0))
    w.step(n=60)
    fat = w.fat_aabbs()
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = rng.uniform(-8, 8, (2, 3)).astype(np.float32)
        lo, hi = p.min(0), p.max(0)
        expect = np.nonzero(np.all(fat[:, 3:] >= lo, axis=1) & np.all(fat[:, :3] <= hi, axis=1))[0]
        assert np.array_equal(w.query_aabb(lo, hi), expect.astype(np.uint32))
