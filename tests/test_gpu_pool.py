"""Pools of resident worlds on the device (bx_pool_*, bx_batch_step_list, snapshots).

Worlds are admitted, stepped in changing mixes, released (their ranges then reused by other
worlds), snapshotted and restored into other slots; after every phase each resident world must be
bit-exact equal to the oracle world stepped alone the same number of times, and a world stepped
alone in a pool must equal, byte for byte, the same world stepped in a mix.
"""
import numpy as np
import pytest

from bounce_amd import scenes
from tests.parity import compare_worlds


def _mix():
    return scenes.concat([scenes.pyramids(1, levels=8), scenes.drop_boxes(60, extent=(6.0, 14.0, 6.0), seed=21),
                          scenes.chains(1, 10), scenes.mixed_drop(40, seed=0), scenes.hello_world(),
                          scenes.with_flags(scenes.drop_boxes(50, extent=(6.0, 12.0, 6.0), seed=22), 1)],
                         name="pool_mix")


class SlotView:
    def __init__(self, pool, slot):
        self.pool, self.slot = pool, slot

    def bodies(self):
        return self.pool.bodies(self.slot)

    def contacts(self):
        return self.pool.contacts(self.slot)

    def pairs(self):
        return self.pool.pairs(self.slot)

    def island_contacts(self):
        return self.pool.island_contacts(self.slot)


@pytest.mark.gpu
def test_pool_mixes_release_reuse_and_restore_match_oracle():
    from bounce_amd import device
    from oracle.oracle import CpuWorld
    mix = _mix()
    ctx = device.Context(0)
    pool = device.Pool(ctx, mix, budget_bytes=256 << 20, max_worlds=8)
    refs = {}

    def admit(i):
        s = pool.admit(mix.worlds[i])
        refs[s] = (i, CpuWorld("oracle", mix, i))
        return s

    def ticks(slots, n):
        for _ in range(n):
            pool.step_list(slots)
            for s in slots:
                refs[s][1].step()

    def check(tag):
        for s, (i, ref) in refs.items():
            d = compare_worlds(SlotView(pool, s), ref)
            assert not d, (tag, s, i, d)
        print(f"[pool] {tag} ok: slots {sorted(refs)}", flush=True)

    a, b = admit(0), admit(1)
    ticks([a, b], 20)
    check("two worlds")
    c = admit(2)
    ticks([c, a, b], 15)
    ticks([c], 5)
    ticks([b, c], 10)
    check("three worlds, uneven")
    pool.release(b)
    del refs[b]
    d = admit(3)  # takes the released ranges
    e = admit(4)
    ticks([a, c, d, e], 25)
    check("released ranges reused")
    image = pool.snapshot(a)
    ia, ra = refs.pop(a)
    pool.release(a)
    f = admit(5)
    a2 = pool.restore(mix.worlds[ia], image)
    refs[a2] = (ia, ra)
    ticks([a2, c, d, e, f], 30)
    check("snapshot restored into another slot")
    st = pool.stats()
    assert st.resident_worlds == len(refs) and 0 < st.used_bytes <= st.capacity_bytes
    ctx.close()


@pytest.mark.gpu
def test_world_alone_equals_world_in_mix():
    """The same world stepped alone and in a changing mix of other worlds: identical bytes."""
    from bounce_amd import device
    mix = _mix()
    ctx = device.Context(0)
    alone = device.Pool(ctx, mix, budget_bytes=128 << 20, max_worlds=2)
    crowd = device.Pool(ctx, mix, budget_bytes=256 << 20, max_worlds=8)
    x = alone.admit(mix.worlds[3])
    for _ in range(60):
        alone.step_list([x])
    others = [crowd.admit(mix.worlds[i]) for i in (0, 1)]
    y = crowd.admit(mix.worlds[3])
    rng = np.random.default_rng(4)
    for k in range(60):
        if k == 20:
            others.append(crowd.admit(mix.worlds[2]))
        if k == 40:
            crowd.release(others.pop(0))
        sel = [o for o in others if rng.random() < 0.7]
        order = sel + [y]
        rng.shuffle(order)
        crowd.step_list(order)
    for name in ("bodies", "contacts", "pairs"):
        a, b = getattr(alone, name)(x), getattr(crowd, name)(y)
        assert a.tobytes() == b.tobytes(), name
    ctx.close()


@pytest.mark.gpu
def test_pool_admission_by_budget():
    from bounce_amd import device
    mix = _mix()
    ctx = device.Context(0)
    probe = device.Pool(ctx, mix, budget_bytes=64 << 20, max_worlds=2)
    need = [probe.world_bytes(w) for w in mix.worlds]
    probe.close()
    pool = device.Pool(ctx, mix, budget_bytes=int(need[0] * 2.5), max_worlds=8)
    got = []
    with pytest.raises(device.PoolFull):
        for _ in range(8):
            got.append(pool.admit(mix.worlds[0]))
    assert 1 <= len(got) < 8
    pool.release(got.pop())
    got.append(pool.admit(mix.worlds[0]))  # fits again once memory is free
    with pytest.raises(device.BxError):
        pool.release(99)
    with pytest.raises(device.BxError):
        pool.step_list([got[0], got[0]])  # a world listed twice
    ctx.close()
