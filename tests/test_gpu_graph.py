"""Graph replay (bx_batch_step_graph) against direct launches (bx_batch_step), bit-exact.

Two copies of a batch step side by side: one with the step's kernels launched directly, one with
replays of the captured step. After every call (several replays each, with reads of the state in
between) every world's state hash must be equal, also across a change of the step configuration
(which captures the step again) and on the mesh and joint kernel paths.
"""
import numpy as np
import pytest

from bounce_amd import scenes
from tests.golden.make_golden import state_hash

pytestmark = pytest.mark.gpu


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


def _run_pair(ctx, scene, calls):
    from bounce_amd import device
    worlds = len(scene.worlds)
    direct = device.Batch(ctx, scene)
    replay = device.Batch(ctx, scene)
    done = 0
    for k, cfg in calls:
        direct.step(n=k, **cfg)
        replay.step_graph(n=k, **cfg)
        done += k
        direct.sync()
        replay.sync()
        for w in range(worlds):
            a = state_hash(DeviceWorld(direct, w))
            b = state_hash(DeviceWorld(replay, w))
            assert a == b, f"world {w}: replayed state differs after {done} steps ({cfg or 'default cfg'})"
        assert np.array_equal(direct.bodies(0), replay.bodies(0))
    for w in range(worlds):
        assert replay.counts(w).status == 0
    direct.close()
    replay.close()


SLOW = {"dt": 1.0 / 120.0, "vit": 4, "pit": 3}


def test_graph_replay_pyramids(gpu_ctx):
    scene = scenes.pyramids(16, levels=10, seed=4)
    print("[graph] pyramids x16: 1+5+3+12 steps, 6 at another cfg, 5 more", flush=True)
    _run_pair(gpu_ctx, scene, [(1, {}), (5, {}), (3, {}), (12, {}), (6, SLOW), (5, {})])


def test_graph_replay_mesh_grounds(gpu_ctx):
    scene = scenes.mesh_drop(20, n_worlds=3, seed=35, hull_count=4)
    print("[graph] mesh grounds x3: 1+7+30 steps", flush=True)
    _run_pair(gpu_ctx, scene, [(1, {}), (7, {}), (30, {})])


def test_graph_replay_limited_chains_with_sleep(gpu_ctx):
    scene = scenes.with_flags(scenes.chains(1, 10, limit=True, extra=20), 1)
    print("[graph] limited chain with sleep: 1+20+60+4 steps", flush=True)
    _run_pair(gpu_ctx, scene, [(1, {}), (20, {}), (60, {}), (4, SLOW)])


def test_graph_replay_reports_no_kernel_times_and_refuses_pools(gpu_ctx):
    from bounce_amd import device
    # a fixed layout: under solver_mode 0 the first nine steps run directly (timed) to tune it
    batch = device.Batch(gpu_ctx, scenes.pyramids(2, levels=5, seed=1), solver_mode=1)
    batch.step(n=2)
    assert batch.kernel_timings()
    batch.step_graph(n=2)
    with pytest.raises(device.BxError):
        batch.kernel_timings()
    batch.close()
    pool = device.Pool(gpu_ctx, scenes.pyramids(1, levels=5), budget_bytes=64 << 20, max_worlds=2)
    with pytest.raises(device.BxError):
        pool.step_graph(n=1)
    pool.close()
