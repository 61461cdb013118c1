"""Contact listener events and the contact filter on the device (BX_WORLD_EVENTS, bx_batch_get_events,
bx_batch_set_contact_filter) against the oracle.

Both are compiled out of the GPU library (BX_DEVICE_CALLBACKS): the device test of event recording, the
sleeping mixed pile of tests/event_script.py with events on and no filter, ended in a GPU memory fault
whose cause has not been found. The host build of the same kernels compiles them in and runs every test
of this file clean under AddressSanitizer, equal to the oracle at every step (tests/test_cpu_build.py).
On the GPU the library must refuse both, and the equality tests stay skipped."""
import numpy as np
import pytest

from bounce_amd import capi, device, scenes
from oracle.oracle import CpuWorld
from tests import event_script
from tests.parity import compare_worlds

pytestmark = pytest.mark.gpu

FAULT_REASON = ("contact events and filters are compiled out of the GPU library: the event test ended in "
                "a GPU memory fault whose cause has not been found (run on the host build instead)")


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


def test_gpu_library_refuses_event_worlds(gpu_ctx):
    """Batch creation fails in the host-side world build, before any kernel runs."""
    with pytest.raises(device.BxError, match=f"error {-4}:.*BX_WORLD_EVENTS"):
        device.Batch(gpu_ctx, event_script.scene(), solver_mode=1)


def test_gpu_library_refuses_filter_tables(gpu_ctx):
    scene = scenes.mixed_drop(8, seed=0)
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    tables = event_script.filter_tables(9)
    with pytest.raises(device.BxError, match=f"error {-4}:.*filter"):
        batch.set_contact_filter(0, *tables)
    batch.close()


@pytest.mark.skip(reason=FAULT_REASON)
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


@pytest.mark.skip(reason=FAULT_REASON)
def test_contact_filter_device_equal_oracle(gpu_ctx):
    """Filter tables on two of three worlds (the third unfiltered): every world's bodies, contacts,
    pairs and island order equal the oracle's every 8 steps and at the end, and the filtered worlds
    differ from the unfiltered run (the tables take effect)."""
    scene = scenes.mixed_drop(60, n_worlds=3, seed=0)
    assert not any(w.flags & capi.BX_WORLD_EVENTS for w in scene.worlds)
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    refs = [CpuWorld("oracle", scene, w) for w in range(3)]
    for w in (0, 2):
        tables = event_script.filter_tables(refs[w].counts().fixture_count)
        refs[w].set_contact_filter(*tables)
        batch.set_contact_filter(w, *tables)
    plain = CpuWorld("oracle", scene, 0)
    steps = 160
    for step in range(steps):
        batch.step()
        plain.step()
        for r in refs:
            r.step()
        if step % 8 == 7 or step == steps - 1:
            for w, r in enumerate(refs):
                d = compare_worlds(_W(batch, w), r, mask_cache=True)
                assert not d, f"step {step} world {w}: {d}"
    moved = np.abs(refs[0].bodies()["position"] - plain.bodies()["position"]).max()
    assert moved > 1e-3, "the filter tables changed nothing"
    batch.close()
