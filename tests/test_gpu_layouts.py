"""Velocity-sweep layouts (bx_batch_opts.solver_mode) give identical results, bit-exact.

solver_mode 1, 2, 3 put 16, 32 or 64 islands on one wavefront (one lane each) with LDS slabs of
256, 128 or 64 bodies; solver_mode 4, 5 put 4 or 8 islands on one wavefront with eight lanes each
(the contact's vector arithmetic split over components and bodies) and slabs of 320 or 224 bodies.
Larger islands keep their velocities in global memory. solver_mode 0 runs each layout on two of the
first timed steps and keeps the fastest. A layout only decides which lanes solve which island, so
every mode must give the same state hash for every world at every checkpoint, on islands that fit
every slab, islands that fit only the larger slabs or none, joints, meshes and many small islands.
"""
import pytest

from bounce_amd import scenes
from tests.golden.make_golden import state_hash

pytestmark = pytest.mark.gpu

ISLANDS_PER_WAVEFRONT = (16, 32, 64, 4, 8)  # solver_mode 1..5
LAYOUTS = len(ISLANDS_PER_WAVEFRONT)


class DeviceWorld:
    def __init__(self, batch, w):
        self.batch, self.w = batch, w

    def bodies(self):
        return self.batch.bodies(self.w)

    def contacts(self):
        return self.batch.contacts(self.w)

    def pairs(self):
        return self.batch.pairs(self.w)

    def island_contacts(self):
        return self.batch.island_contacts(self.w)


CASES = {
    # 10-, 12- and 16-level pyramids: islands of 56, 79 and 137 bodies
    "pyramids": lambda: scenes.concat([scenes.pyramids(2, levels=10, seed=3), scenes.pyramids(2, levels=12, seed=4),
                                       scenes.pyramids(2, levels=16, seed=5)]),
    "drop200": lambda: scenes.drop_boxes(200, extent=(10.0, 25.0, 10.0), seed=9),
    # a 30-level pyramid: one island of 466 bodies, above every eight-lane slab
    "pyramid30": lambda: scenes.pyramids(1, levels=30, seed=12),
    "limited_chains_sleep": lambda: scenes.with_flags(scenes.chains(2, 10, limit=True, extra=20), 1),
    "mesh_grounds": lambda: scenes.mesh_drop(20, n_worlds=2, seed=37, hull_count=4),
}


@pytest.mark.parametrize("name", list(CASES))
def test_layouts_match(gpu_ctx, name):
    from bounce_amd import device
    scene = CASES[name]()
    worlds = len(scene.worlds)
    print(f"[layouts] {name}: {worlds} worlds, modes 0-5, 1+2+57 steps", flush=True)
    batches = {m: device.Batch(gpu_ctx, scene, solver_mode=m) for m in range(LAYOUTS + 1)}
    for k in (1, 2, 57):
        hashes = {}
        for m, b in batches.items():
            b.step(n=k)
            b.sync()
            hashes[m] = [state_hash(DeviceWorld(b, w)) for w in range(worlds)]
        for m in (0, 2, 3, 4, 5):
            for w in range(worlds):
                assert hashes[m][w] == hashes[1][w], f"{name}: mode {m} differs from mode 1 in world {w}"
    for m in range(1, LAYOUTS + 1):
        lanes, times = batches[m].solver_layout()
        assert lanes == ISLANDS_PER_WAVEFRONT[m - 1] and times == [0.0] * LAYOUTS
    lanes, times = batches[0].solver_layout()
    assert lanes in ISLANDS_PER_WAVEFRONT and all(t > 0.0 for t in times)
    best = min(range(LAYOUTS), key=lambda i: times[i])
    expect = best if times[best] < times[0] * (1.0 - 0.05) else 0  # 16 unless faster by more than 5%
    assert lanes == ISLANDS_PER_WAVEFRONT[expect]
    for b in batches.values():
        for w in range(worlds):
            assert b.counts(w).status == 0
        b.close()


def test_graph_replay_uses_the_tuned_layout(gpu_ctx):
    from bounce_amd import device
    scene = scenes.pyramids(4, levels=12, seed=6)
    print("[layouts] graph replay after tuning: 2 + 20 steps", flush=True)
    direct = device.Batch(gpu_ctx, scene, solver_mode=0)
    replay = device.Batch(gpu_ctx, scene, solver_mode=0)
    direct.step(n=2)
    replay.step_graph(n=2)  # the first steps, before the tuner times anything, launched directly
    direct.step(n=20)
    # the tuner's steps (4-13) launch directly; the steps after them are replays of the graph
    # captured with the layout it kept (all layouts give the same bits, so the hashes check the
    # replays' results, and the tuner's recorded times show that tuning finished first)
    replay.step_graph(n=20)
    direct.sync()
    replay.sync()
    assert all(t > 0.0 for t in replay.solver_layout()[1])
    for w in range(4):
        assert state_hash(DeviceWorld(direct, w)) == state_hash(DeviceWorld(replay, w))
    direct.close()
    replay.close()


def test_invalid_solver_mode_is_rejected(gpu_ctx):
    from bounce_amd import device
    with pytest.raises(device.BxError):
        device.Batch(gpu_ctx, scenes.pyramids(1, levels=4), solver_mode=LAYOUTS + 1)


def test_pool_layouts_match(gpu_ctx):
    """Pools take a fixed layout (bx_pool_opts.solver_mode); 16 and 64 islands per wavefront, one lane
    each, and 8 islands per wavefront, eight lanes each, give the same state for worlds stepped in a
    changing mix."""
    from bounce_amd import device
    scene = scenes.drop_boxes(30, n_worlds=6, seed=21, extent=(12.0, 6.0, 12.0))
    print("[layouts] pools at solver modes 1, 3 and 5: 40 ticks", flush=True)
    pools = [device.Pool(gpu_ctx, scene, budget_bytes=64 << 20, max_worlds=6, solver_mode=m) for m in (1, 3, 5)]
    slots = [[p.admit(w) for w in scene.worlds] for p in pools]
    for t in range(40):
        pick = [i for i in range(6) if (t + i) % 3 != 0]
        for p, s in zip(pools, slots):
            p.step_list([s[i] for i in pick])
    for i in range(6):
        a = state_hash(DeviceWorld(pools[0], slots[0][i]))
        for k in (1, 2):
            assert state_hash(DeviceWorld(pools[k], slots[k][i])) == a, f"world {i}, pool {k}"
    for p in pools:
        p.close()
    with pytest.raises(device.BxError):
        device.Pool(gpu_ctx, scene, budget_bytes=64 << 20, max_worlds=2, solver_mode=LAYOUTS + 1)
