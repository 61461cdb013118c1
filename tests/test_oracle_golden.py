"""The CPU oracle against the reference's golden fixtures (tests/golden/*.npz), bit-exact.

Each fixture holds, for every step, a SHA-256 of the reference FP32 build's full world state
(bodies, contacts with manifolds and feature ids, feature caches, broadphase pairs, island
order). Equality of every hash pins the oracle's integration, narrowphase, contact lifecycle,
broadphase, island builder, solver and sleep logic on these scenes.
"""
import hashlib
import os

import numpy as np
import pytest

from bounce_amd import hulls, scenes
from tests.golden.make_golden import SCENES, state_hash
from tests.parity import first_diff, masked_bodies, masked_contacts

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


@pytest.mark.parametrize("name", list(SCENES))
def test_oracle_matches_reference_every_step(name):
    from oracle.oracle import CpuWorld
    g = load(name)
    factory, steps, checkpoints = SCENES[name]
    w = CpuWorld("oracle", factory())
    assert np.array_equal(w.proxy_ids(), g["proxy_ids"])
    assert w.mass().tobytes() == g["mass"].tobytes()
    for s in range(1, steps + 1):
        w.step()
        if s in checkpoints:
            d = first_diff(masked_bodies(w.bodies()), g[f"bodies_{s}"], "body")
            assert d is None, f"{name} step {s}: {d}"
            d = first_diff(masked_contacts(w.contacts()), g[f"contacts_{s}"], "contact")
            assert d is None, f"{name} step {s}: {d}"
        h = np.frombuffer(state_hash(w), dtype=np.uint8)
        assert np.array_equal(h, g["hashes"][s - 1]), f"{name}: state diverges at step {s}"


def test_fp32_vs_fp64_error_reported():
    """FP32 poses stay close to the FP64 reference on a resting stack (reported, loose bound)."""
    from oracle.oracle import CpuWorld
    g = load("box_stack5")
    w = CpuWorld("oracle", scenes.box_stack(5))
    w.step(n=int(g["steps"][0]))
    b = w.bodies()
    err = np.abs(b["position"].astype(np.float64) - g["f64_final"][:, 0:3]).max()
    assert err < 1e-3, err


@pytest.mark.parametrize("ext", [(0.5, 0.5, 0.5), (50.0, 0.5, 50.0), (10.0, 1.0, 10.0), (1.0, 1.0, 1.0),
                                 (0.3, 1.7, 2.9)])
def test_box_hull_matches_reference(ext):
    g = np.load(os.path.join(GOLDEN, "box_hulls.npz"))
    key = "_".join(str(x) for x in ext)
    h = hulls.box_hull(*ext)
    assert h.vertices.ravel().tobytes() == g[f"v_{key}"].tobytes()
    assert h.edges.ravel().tobytes() == g[f"e_{key}"].tobytes()
    assert h.faces.tobytes() == g[f"f_{key}"].tobytes()
    assert h.planes.ravel().tobytes() == g[f"p_{key}"].tobytes()
    assert h.centroid.tobytes() == g[f"c_{key}"].tobytes()


def test_hello_world_printed_pose():
    """examples/hello_world prints 'position = 0.00 5.02 0.00' and angle 3.14 after 60 steps."""
    from oracle.oracle import CpuWorld
    w = CpuWorld("oracle", scenes.hello_world())
    w.step(n=60)
    b = w.bodies()[1]
    assert "%.2f %.2f %.2f" % tuple(b["position"]) == "0.00 5.02 0.00"
    q = b["orientation"].astype(np.float64)
    angle = 2.0 * np.arccos(np.clip(q[3], -1, 1))
    assert "%.2f" % angle == "3.14"
