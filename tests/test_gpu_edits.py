"""Body edits between steps on the device (bx_batch_edit_bodies: b3Body's setters, force and impulse
methods, SetTransform) against the oracle, bit for bit on every step of a scripted sequence, with all
worlds of the script in one batch; and bx_batch_set_bodies with the velocity setters' semantics (static
bodies untouched, a nonzero velocity wakes a sleeping body)."""
import numpy as np
import pytest

from bounce_amd import capi, device
from oracle.oracle import CpuWorld
from tests import edit_script
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


def test_edits_device_equals_oracle(gpu_ctx):
    scene = edit_script.scene()
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    refs = [CpuWorld("oracle", scene, w) for w in range(len(scene.worlds))]
    for step in range(edit_script.STEPS):
        rows = edit_script.edits(step)
        for w, r in enumerate(refs):
            batch.edit_bodies(w, rows)
            r.edit_bodies(rows)
            r.step()
        batch.step()
        for w, r in enumerate(refs):
            assert batch.counts(w).status == 0
            d = compare_worlds(_W(batch, w), r, mask_cache=True)
            assert not d, f"step {step} world {w}: {d}"
    batch.close()


def test_set_bodies_wakes_and_skips_static(gpu_ctx):
    scene = edit_script.scene()
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    ref = CpuWorld("oracle", scene, 2)
    for _ in range(150):
        batch.step()
        ref.step()
    st = batch.bodies(2)
    asleep = np.where((st["flags"] & capi.BX_BODY_AWAKE) == 0)[0]
    asleep = asleep[asleep > 0]
    assert len(asleep) > 0, "some box should be asleep by step 150"
    want = st.copy()
    want["linear_velocity"][asleep[0]] = (0.0, 2.0, 0.0)
    want["linear_velocity"][0] = (5.0, 0.0, 0.0)  # the static ground ignores it
    rc = device.load().bx_batch_set_bodies(batch.h, 2, want.ctypes.data_as(device.C.POINTER(capi.bx_body_state)),
                                           len(want))
    assert rc == 0
    rows = [(int(i), capi.BX_EDIT_LINEAR_VELOCITY, tuple(want["linear_velocity"][i])) for i in range(len(want))]
    ref.edit_bodies(rows)
    got = batch.bodies(2)
    assert got["flags"][asleep[0]] & capi.BX_BODY_AWAKE
    assert tuple(got["linear_velocity"][0]) == (0.0, 0.0, 0.0)
    for _ in range(30):
        batch.step()
        ref.step()
        d = compare_worlds(_W(batch, 2), ref, mask_cache=True)
        assert not d, d
    batch.close()
