"""Topology changes between steps on the device (bx_batch_edit_topology) against the oracle, bit for bit
after every edit and every step of the scenarios of tests/topology_script.py: contacts, impulses,
feature caches, joint impulses, sleep state and proxy ids carry over; new proxies take the ids the
reference's dynamic tree would give them. A batch of several worlds keeps the unedited ones stepping
exactly as before. The same functions run on the host build under AddressSanitizer
(tests/test_cpu_build.py)."""
import pytest

from bounce_amd import device, scenes
from oracle.oracle import CpuWorld
from tests import topology_script as TS
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


@pytest.mark.parametrize("name", list(TS.SCENARIOS))
def test_topology_device_equals_oracle(gpu_ctx, name):
    sc = TS.SCENARIOS[name]()
    print(f"[topology] {name}: {sc.steps} steps", flush=True)
    batch = device.Batch(gpu_ctx, sc.scene, solver_mode=1)
    ref = CpuWorld("oracle", sc.scene)
    dw = _W(batch, 0)
    for step in range(sc.steps):
        if step in sc.edits:
            ops, adds = sc.edits[step]
            batch.edit_topology(0, ops, adds, sc.library)
            ref.edit_topology(ops, adds, sc.library)
            d = compare_worlds(dw, ref, mask_cache=True)
            assert not d, f"after the edit before step {step}: {d}"
        batch.step()
        ref.step()
        d = compare_worlds(dw, ref, mask_cache=True)
        assert not d, f"step {step}: {d}"
    assert batch.counts(0).status == 0
    batch.close()


def test_topology_edit_in_a_batch_keeps_other_worlds(gpu_ctx):
    """World 1 of three is edited (the spawn scenario); worlds 0 and 2 carry over and stay equal to
    their oracle worlds."""
    sc = TS.spawn()
    scene = scenes.concat([scenes.pyramids(1, levels=5, seed=8), sc.scene, scenes.drop_boxes(20, seed=3,
                                                                                             extent=(5.0, 8.0, 5.0))])
    print("[topology] batch of three, world 1 edited", flush=True)
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    refs = [CpuWorld("oracle", scene, w) for w in range(3)]
    for step in range(90):
        if step in sc.edits:
            ops, adds = sc.edits[step]
            batch.edit_topology(1, ops, adds, sc.library)
            refs[1].edit_topology(ops, adds, sc.library)
        batch.step()
        for r in refs:
            r.step()
        if step % 10 == 9:
            for w, r in enumerate(refs):
                d = compare_worlds(_W(batch, w), r, mask_cache=True)
                assert not d, f"world {w} step {step}: {d}"
    batch.close()


def test_topology_errors_leave_the_batch_unchanged(gpu_ctx):
    sc = TS.spawn()
    batch = device.Batch(gpu_ctx, sc.scene, solver_mode=1)
    ref = CpuWorld("oracle", sc.scene)
    for _ in range(5):
        batch.step()
        ref.step()
    from bounce_amd import capi
    with pytest.raises(device.BxError):
        batch.edit_topology(0, [(capi.BX_TOPO_DESTROY_BODY, 500)], TS._adds())
    with pytest.raises(device.BxError):  # the second operation fails: the first must not apply either
        batch.edit_topology(0, [(capi.BX_TOPO_DESTROY_BODY, 3), (capi.BX_TOPO_CREATE_FIXTURE, 9)], TS._adds())
    for _ in range(5):
        batch.step()
        ref.step()
    assert not compare_worlds(_W(batch, 0), ref, mask_cache=True)
    batch.close()
