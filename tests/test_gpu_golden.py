"""The device step against the reference's golden fixtures (tests/golden/*.npz), bit-exact.

Each fixture holds, for every step, a SHA-256 of the reference FP32 build's full world state
(bodies, contacts with every manifold and feature id, broadphase pairs, island order). Here the
device library steps the same scene and its state hash must equal the reference's at every step:
the product is pinned to the reference directly, not only through the oracle.

One fixture is not replayed: mixed150_dense_sleep (kept off the GPU; see tests/test_gpu_parity.py).
"""
import os

import numpy as np
import pytest

from tests.golden.make_golden import SCENES, state_hash

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NOT_REPLAYED = {"mixed150_dense_sleep"}


class DeviceWorld:
    def __init__(self, batch):
        self.batch = batch

    def bodies(self):
        return self.batch.bodies(0)

    def contacts(self):
        return self.batch.contacts(0)

    def pairs(self):
        return self.batch.pairs(0)

    def island_contacts(self):
        return self.batch.island_contacts(0)


@pytest.mark.parametrize("name", [n for n in SCENES if n not in NOT_REPLAYED])
def test_device_matches_reference_every_step(gpu_ctx, name):
    from bounce_amd import device
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    factory, steps, _ = SCENES[name]
    print(f"[golden] {name}: {steps} steps", flush=True)
    batch = device.Batch(gpu_ctx, factory())
    dw = DeviceWorld(batch)
    for s in range(1, steps + 1):
        batch.step()
        h = np.frombuffer(state_hash(dw), dtype=np.uint8)
        assert np.array_equal(h, g["hashes"][s - 1]), f"{name}: device state differs from the reference at step {s}"
    assert batch.counts(0).status == 0
    batch.close()
