"""The continuous-batching server (bounce_amd/serving.py).

CPU: the scheduler over a model pool that keeps the device pool's contract (slots, a byte budget,
BX_E_CAPACITY as PoolFull, snapshots that carry a world's state) and a simulated clock: every
request completes with exactly its steps, a world never has two requests running, requests of a
world finish in arrival order, ticks list each world once, and admission respects the budget with
eviction and restore of idle worlds.

GPU: the server on a real pool, with a budget that forces evictions, against the oracle: every
world of the catalog ends bit-exact equal to the oracle world stepped alone by the same total
number of steps (each result is the same alone and in the mix).
"""
import collections

import numpy as np
import pytest

from bounce_amd import scenes, serving
from bounce_amd.device import PoolFull


class ModelWorld:
    def __init__(self, body_count):
        self.body_count = body_count


class ModelPool:
    """Slots and a byte budget (100 bytes per body); a world's state is its step count."""

    def __init__(self, budget, slots):
        self.budget, self.free_slots = budget, list(range(slots))[::-1]
        self.resident = {}   # slot -> [world object, steps taken]
        self.ticks = []

    def _used(self):
        return sum(100 * w.body_count for w, _ in self.resident.values())

    def _take(self, world, steps):
        if not self.free_slots or self._used() + 100 * world.body_count > self.budget:
            raise PoolFull("full")
        s = self.free_slots.pop()
        self.resident[s] = [world, steps]
        return s

    def admit(self, world):
        return self._take(world, 0)

    def restore(self, world, image):
        assert image[0] is world
        return self._take(world, image[1])

    def snapshot(self, slot):
        w, n = self.resident[slot]
        return (w, n)

    def release(self, slot):
        del self.resident[slot]
        self.free_slots.append(slot)

    def step_list(self, slots, dt, vit, pit):
        assert len(set(slots)) == len(slots)
        assert self._used() <= self.budget
        for s in slots:
            self.resident[s][1] += 1
        self.ticks.append(list(slots))

    def sync(self):
        pass


class SimClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 1e-4  # each look at the clock costs a little simulated time
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-5)


def _serve(budget, n_worlds=12, rate=400.0, duration=0.5, seed=3, max_active=None, slots=6):
    rng = np.random.default_rng(seed)
    catalog = [ModelWorld(int(n)) for n in rng.integers(5, 50, n_worlds)]
    pool = ModelPool(budget, slots=slots)
    srv = serving.Server(pool, catalog, max_active=max_active)
    reqs = serving.poisson_requests(rate, duration, n_worlds, seed=seed)
    clock = SimClock()
    wall = srv.run(reqs, clock=clock, sleep=clock.sleep)
    srv.flush()
    return catalog, pool, srv, reqs, wall


@pytest.mark.parametrize("budget", [10 ** 9, 12000])
def test_every_request_completes_with_its_steps(budget):
    catalog, pool, srv, reqs, wall = _serve(budget)
    assert len(srv.finished) == len(reqs) > 50
    want = collections.Counter()
    for r in reqs:
        want[r.world] += r.steps
        assert r.done is not None and r.done >= r.start >= r.arrival
    got = {}
    for w, image in srv.images.items():
        got[w] = image[1]
    assert got == dict(want)
    # a world runs one request at a time, in arrival order
    by_world = collections.defaultdict(list)
    for r in srv.finished:
        by_world[r.world].append(r)
    for rs in by_world.values():
        rs.sort(key=lambda r: r.arrival)
        for a, b in zip(rs, rs[1:]):
            assert b.start >= a.done
    rep = srv.report(wall)
    assert rep["requests"] == len(reqs)
    assert rep["p50_request_ms"] <= rep["p99_request_ms"]


def test_small_budget_evicts_and_restores():
    _, _, srv, _, _ = _serve(12000)
    assert srv.evictions > 0 and srv.restores > 0
    _, _, few_slots, _, _ = _serve(10 ** 9, slots=4)
    assert few_slots.evictions > 0
    _, _, big, _, _ = _serve(10 ** 9, slots=12)
    assert big.evictions == 0


def test_ticks_share_worlds():
    """Requests that overlap in time share ticks (continuous batching, not a queue)."""
    _, pool, srv, _, _ = _serve(10 ** 9, rate=2000.0, duration=0.2)
    assert max(len(t) for t in pool.ticks) > 3
    assert srv.report(1.0)["mean_worlds_per_tick"] > 1.5


def test_max_active_bounds_tick_size():
    _, pool, _, _, _ = _serve(10 ** 9, rate=2000.0, duration=0.2, max_active=2)
    assert max(len(t) for t in pool.ticks) <= 2


def test_world_larger_than_budget_is_an_error():
    pool = ModelPool(1000, slots=2)
    srv = serving.Server(pool, [ModelWorld(50)])
    with pytest.raises(PoolFull):
        srv.run([serving.Request(0, 0, 3, 0.0)])


# ------------------------------------------------------------------------------------ GPU
def _catalog():
    return scenes.concat([scenes.pyramids(1, levels=6), scenes.drop_boxes(40, extent=(6.0, 12.0, 6.0), seed=11),
                          scenes.chains(1, 6), scenes.mixed_drop(30, seed=0), scenes.box_stack(4, seed=9),
                          scenes.hello_world(), scenes.drop_boxes(25, extent=(5.0, 10.0, 5.0), seed=12),
                          scenes.mixed_drop(20, seed=2, kinds=("sphere", "capsule"))], name="serve_catalog")


@pytest.mark.gpu
def test_served_worlds_match_oracle_alone():
    from bounce_amd import device
    from oracle.oracle import CpuWorld
    from tests.parity import compare_worlds

    cat = _catalog()
    ctx = device.Context(0)
    probe = device.Pool(ctx, cat, budget_bytes=1 << 26, max_worlds=4)
    sizes = [probe.world_bytes(w) for w in cat.worlds]
    probe.close()
    budget = int(sum(sorted(sizes)[-3:]) * 1.6)  # about three worlds resident: forces evictions
    pool = device.Pool(ctx, cat, budget_bytes=budget, max_worlds=8)
    srv = serving.Server(pool, cat.worlds)
    reqs = serving.poisson_requests(300.0, 0.4, len(cat.worlds), max_steps=16, seed=5)
    print(f"[serving] {len(reqs)} requests over {len(cat.worlds)} worlds, budget {budget}", flush=True)
    wall = srv.run(reqs)
    rep = srv.report(wall)
    print("[serving]", rep, flush=True)
    assert rep["requests"] == len(reqs)
    assert srv.evictions > 0
    steps = collections.Counter()
    for r in reqs:
        steps[r.world] += r.steps
    srv.flush()
    for w in range(len(cat.worlds)):
        ref = CpuWorld("oracle", cat, w)
        if steps[w]:
            ref.step(n=steps[w])
        if w not in srv.images:
            continue
        slot = pool.restore(cat.worlds[w], srv.images[w])

        class View:
            def bodies(self):
                return pool.bodies(slot)

            def contacts(self):
                return pool.contacts(slot)

            def pairs(self):
                return pool.pairs(slot)

            def island_contacts(self):
                return pool.island_contacts(slot)

        d = compare_worlds(View(), ref)
        assert not d, (w, steps[w], d)
        pool.release(slot)
    ctx.close()
