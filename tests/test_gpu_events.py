"""Contact listener events and the contact filter on the device (BX_WORLD_EVENTS, bx_batch_get_events,
bx_batch_set_contact_filter) against the oracle: the event sequence and the whole world state equal at
every step, without and with the filter tables (tests/event_script.py)."""
import pytest

from bounce_amd import device
from oracle.oracle import CpuWorld
from tests import event_script
from tests.parity import compare_worlds

pytestmark = pytest.mark.gpu


class _W:
    def __init__(self, b, w):
        self.b, self.w = b, w

    def bodies(self):
        return self.b.bodies(self.w)

    def contacts(self):
        return self.b.contacts(self.w)

    def pairs(self):
        return self.b.pairs(self.w)

    def island_contacts(self):
        return self.b.island_contacts(self.w)


@pytest.mark.parametrize("filtered", [False, True])
def test_events_device_equal_oracle(gpu_ctx, filtered):
    scene = event_script.scene()
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    refs = [CpuWorld("oracle", scene, w) for w in range(len(scene.worlds))]
    if filtered:
        for w, r in enumerate(refs):
            tables = event_script.filter_tables(r.counts().fixture_count)
            r.set_contact_filter(*tables)
            batch.set_contact_filter(w, *tables)
    for step in range(event_script.STEPS):
        batch.step()
        for w, r in enumerate(refs):
            r.step()
            ed, eo = batch.events(w), r.events()
            assert ed.tobytes() == eo.tobytes(), f"step {step} world {w}: {len(ed)} vs {len(eo)} events"
            d = compare_worlds(_W(batch, w), r, mask_cache=True)
            assert not d, f"step {step} world {w}: {d}"
    batch.close()
