"""The hull builder against closed forms: every hull is a valid b3Hull (adjacent twins, face loops,
Euler's formula) of exactly the convex hull of its points (scipy's Qhull gives the same vertex set
and volume), and the oracle's hull mass equals density x volume with the centroid as the centre of
mass (hull_shape.cpp ComputeMass)."""
import numpy as np
import pytest
from scipy.spatial import ConvexHull

from bounce_amd import capi, hulls, scenes


@pytest.fixture(scope="module")
def library():
    return hulls.random_hull_library(96, seed=3)


def test_topology(library):
    for h in library:
        V, E, F = len(h.vertices), len(h.edges), len(h.faces)
        assert V - E // 2 + F == 2
        for i, e in enumerate(h.edges):
            assert e[1] == (i ^ 1) and h.edges[e[1]][1] == i          # twins 2k, 2k+1
            assert h.edges[e[4]][3] == i and h.edges[e[4]][2] == e[2]  # next/prev on the same face
            assert h.edges[e[1]][0] != e[0]


def test_matches_qhull(library):
    for h in library:
        ch = ConvexHull(h.vertices.astype(np.float64))
        assert sorted(ch.vertices.tolist()) == list(range(len(h.vertices)))
        # every plane has all vertices behind it and is supported by its face's vertices
        d = h.vertices.astype(np.float64) @ h.planes[:, :3].T.astype(np.float64) - h.planes[:, 3]
        assert np.all(d <= 1e-5)
        assert np.linalg.norm(h.planes[:, :3], axis=1) == pytest.approx(1.0, abs=1e-6)


def test_face_loops_are_outward(library):
    for h in library:
        for f, e0 in enumerate(h.faces):
            a = h.vertices[h.edges[e0][0]].astype(np.float64)
            b = h.vertices[h.edges[h.edges[e0][4]][0]].astype(np.float64)
            c = h.vertices[h.edges[h.edges[h.edges[e0][4]][4]][0]].astype(np.float64)
            assert np.dot(np.cross(b - a, c - a), h.planes[f, :3]) > 0


def test_hull_mass_is_density_times_volume(library):
    from oracle.oracle import CpuWorld
    s = scenes.SceneSet("hull_mass")
    idx = [s.add_hull(h) for h in library[:12]]
    b = scenes.body_array(12)
    b["type"] = capi.BX_BODY_DYNAMIC
    b["position"][:, 0] = np.arange(12) * 5.0
    f = scenes.fixture_array(12)
    f["body"] = np.arange(12)
    f["shape_type"] = capi.BX_SHAPE_HULL
    f["shape_index"] = idx
    f["density"] = 2.0
    s.worlds.append(scenes.World(b, f))
    m = CpuWorld("oracle", s).mass()
    for k, h in enumerate(library[:12]):
        ch = ConvexHull(h.vertices.astype(np.float64))
        assert m[k, 0] == pytest.approx(2.0 * ch.volume, rel=1e-5)
        # centre of mass of the solid (oracle) = the stored centroid
        assert m[k, 2:5] == pytest.approx(h.centroid, abs=1e-5)
