"""Continuous batching of step requests over a device pool of resident worlds.

A request asks to advance one world of a catalog by k steps (b3World::Step(dt, vIt, pIt) k times).
The server keeps worlds resident in a bx_pool (bounce_amd.device.Pool) whose device memory is a
fixed budget, and runs ticks: one tick is one launch sequence (bx_batch_step_list) that advances
every world with an active request by one step. Between ticks it

  * admits arrived requests in arrival order: a request becomes active when its world has no other
    active request and the world is resident, or can be made resident in a free slot within the
    budget (a world seen before is restored from its snapshot; idle resident worlds are evicted to
    host snapshots, least recently used first, when memory is needed);
  * retires requests whose last step ran in the tick.

So requests that arrive while others run join the next tick and share its launches, and a world's
result is the same whichever other worlds share its ticks (each world is one workgroup of its own
in the per-world kernels and its islands are solved by lanes of their own).

Latency of a request = completion time - arrival time, on the host clock after the device finished
the tick that completed it. Tick latency = wall time of one tick (launch to device completion).
"""
import collections
import time

import numpy as np

from .device import PoolFull


class Request:
    __slots__ = ("rid", "world", "steps", "arrival", "start", "done", "remaining")

    def __init__(self, rid, world, steps, arrival):
        self.rid, self.world, self.steps, self.arrival = rid, world, int(steps), float(arrival)
        self.start = self.done = None
        self.remaining = int(steps)


def poisson_requests(rate, duration, n_worlds, max_steps=16, seed=0, weights=None):
    """Requests with exponential inter-arrival times (rate per second) over `duration` seconds,
    each for a uniformly drawn world (or drawn by `weights`) and U{1..max_steps} steps."""
    rng = np.random.default_rng(seed)
    out, t, rid = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        w = int(rng.choice(n_worlds, p=weights)) if weights is not None else int(rng.integers(0, n_worlds))
        out.append(Request(rid, w, int(rng.integers(1, max_steps + 1)), t))
        rid += 1


class Server:
    """Continuous-batching step server over `pool` for the worlds of `catalog` (a list of
    scenes.World sharing the pool's shape library)."""

    def __init__(self, pool, catalog, dt=1.0 / 60.0, velocity_iterations=8, position_iterations=2,
                 max_active=None):
        self.pool, self.catalog = pool, catalog
        self.cfg = (dt, velocity_iterations, position_iterations)
        self.max_active = max_active
        self.slot_of = {}                        # world -> slot (resident worlds)
        self.images = {}                         # world -> host snapshot of an evicted world
        self.last_used = {}                      # world -> tick of last use (eviction order)
        self.waiting = collections.deque()       # arrived, not yet active
        self.active = collections.OrderedDict()  # world -> active request
        self.finished = []
        self.tick_times = []
        self.tick_sizes = []
        self.body_steps = 0
        self.ticks = 0
        self.evictions = 0
        self.restores = 0

    # -- residency ------------------------------------------------------------------------------
    def _make_resident(self, world):
        if world in self.slot_of:
            return True
        while True:
            try:
                if world in self.images:
                    slot = self.pool.restore(self.catalog[world], self.images[world])
                    del self.images[world]
                    self.restores += 1
                else:
                    slot = self.pool.admit(self.catalog[world])
                self.slot_of[world] = slot
                return True
            except PoolFull:
                if not self._evict_one(keep=world):
                    return False

    def _evict_one(self, keep):
        busy = set(self.active)
        busy.update(r.world for r in self.waiting)
        idle = [w for w in self.slot_of if w != keep and w not in busy]
        if not idle:
            idle = [w for w in self.slot_of if w != keep and w not in self.active]
        if not idle:
            return False
        victim = min(idle, key=lambda w: self.last_used.get(w, -1))
        slot = self.slot_of.pop(victim)
        self.images[victim] = self.pool.snapshot(slot)
        self.pool.release(slot)
        self.evictions += 1
        return True

    # -- scheduling -----------------------------------------------------------------------------
    def submit(self, req):
        self.waiting.append(req)

    def _admit(self, now):
        kept = collections.deque()
        while self.waiting:
            r = self.waiting.popleft()
            full = self.max_active is not None and len(self.active) >= self.max_active
            if full or r.world in self.active or not self._make_resident(r.world):
                kept.append(r)
                continue
            r.start = now
            self.active[r.world] = r
        self.waiting = kept

    def tick(self, clock):
        """Admit, then run one step of every active world; returns the number of worlds stepped."""
        self._admit(clock())
        if not self.active:
            return 0
        worlds = list(self.active)
        slots = [self.slot_of[w] for w in worlds]
        t0 = time.perf_counter()
        self.pool.step_list(slots, *self.cfg)
        self.pool.sync()
        self.tick_times.append(time.perf_counter() - t0)
        self.tick_sizes.append(len(slots))
        now = clock()
        self.ticks += 1
        for w in worlds:
            r = self.active[w]
            r.remaining -= 1
            self.body_steps += self.catalog[w].body_count
            self.last_used[w] = self.ticks
            if r.remaining == 0:
                r.done = now
                self.finished.append(r)
                del self.active[w]
        return len(slots)

    def run(self, requests, clock=None, sleep=time.sleep, idle_sleep=2e-4, progress=None, progress_every=2.0):
        """Serve `requests` (arrival in seconds from the start) in real time (or on `clock`).
        `progress(server)` is called about every `progress_every` seconds of wall time."""
        t0 = time.perf_counter()
        clock = clock or (lambda: time.perf_counter() - t0)
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        last = time.perf_counter()
        while pending or self.waiting or self.active:
            if progress is not None and time.perf_counter() - last >= progress_every:
                last = time.perf_counter()
                progress(self)
            now = clock()
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            if self.tick(clock) == 0:
                if self.waiting:
                    raise PoolFull(f"world {self.waiting[0].world} does not fit the pool even when empty")
                if pending:
                    sleep(min(idle_sleep, max(0.0, pending[0].arrival - clock())))
        return clock()

    def flush(self):
        """Snapshot every resident world back to the host (the catalog's state after serving)."""
        for w in list(self.slot_of):
            slot = self.slot_of.pop(w)
            self.images[w] = self.pool.snapshot(slot)
            self.pool.release(slot)

    def report(self, wall):
        lat = np.array([r.done - r.arrival for r in self.finished]) if self.finished else np.zeros(1)
        tick = np.array(self.tick_times) if self.tick_times else np.zeros(1)
        steps = sum(r.steps for r in self.finished)
        return {
            "requests": len(self.finished),
            "steps": steps,
            "ticks": self.ticks,
            "mean_worlds_per_tick": float(np.mean(self.tick_sizes)) if self.tick_sizes else 0.0,
            "p50_request_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_request_ms": float(np.percentile(lat, 99) * 1e3),
            "p50_tick_ms": float(np.percentile(tick, 50) * 1e3),
            "p99_tick_ms": float(np.percentile(tick, 99) * 1e3),
            "body_steps_per_s": self.body_steps / wall if wall > 0 else 0.0,
            "evictions": self.evictions,
            "restores": self.restores,
            "wall_s": wall,
        }
