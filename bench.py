#!/usr/bin/env python3
"""Benchmark of the batched b3World::Step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--worlds 4096] [--levels 20]

Workload (default): the batched-stacking configuration, 4,096 independent worlds, each a 20-level
box pyramid (210 unit boxes, friction 0.6, seeded +-1 mm jitter) on a static ground box, stepped
together at dt = 1/60 s with 8 velocity / 2 position iterations in FP32 (sleeping off). A "step"
is one b3World::Step of every world of the batch; inputs are resident in HBM before timing.
`--workload mixed` runs the mixed-shapes batch instead: per-world body counts lognormal (median
200, clipped to 20..2000), spheres, capsules and random hulls of 8-16 points dropped on a ground box,
and a 10-link revolute chain with a motor per world (scenes.mixed_worlds, seed 8). `--workload drops`
runs worlds of 30 unit boxes dropped on a ground box (many small islands), and `--solver-mode`
fixes the velocity-sweep layout (default 0: tuned on the batch's first steps).

Printed (one JSON line): body-steps/s of the batched path (value; the step's kernels launched
directly, with HIP events between them), the same steps of a fresh copy of the batch as replays
of one captured HIP graph (graph_replay), the plain per-world GPU path
(one world per launch sequence, measured on a sample of worlds), their ratio, p50/p99 step
latency, the roofline of the dominant kernel (the island solver) from HIP events on the
library's stream, the reference CPU Step timed on this host's cores (one world per thread)
on a bounded sample of the same worlds, and the serving configuration at one arrival rate
(request and tick latency p50/p99, body-steps/s; bounce_amd/serve_bench.py sweeps rates).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
# Steps a batch spends timing the velocity-sweep layouts when solver_mode is 0 (three steps first, then
# each of five layouts twice): the warm-up covers them, so no tuning step lands in the timed region.
TUNE_STEPS = 13
# Velocity-sweep layouts (solver_mode 1..5): islands per wavefront and lanes per island.
LAYOUTS = [(16, 1), (32, 1), (64, 1), (4, 8), (8, 8)]
# Compulsory bytes of one step (the unit of work is one body-step), from the counts of the step: per
# awake non-static body 432 B (state read 256 B, written 176 B), per proxy 52 B (fat AABB read and
# written, moved flag), per broadphase pair 16 B (sorted pair written and read), per convex contact
# 580 B (manifold read and written, feature cache, header), per mesh contact 1,668 B (three manifolds),
# per revolute joint 164 B.
STEP_BYTES = {"awake_body": 432, "proxy": 52, "pair": 16, "contact": 580, "mesh_contact": 1668, "joint": 164}
# Algorithmic (compulsory) bytes of one launch of each heavy kernel, per unit of work, from the
# data it must read and write once (bounce_amd/csrc/bx_types.h layouts):
#  * k_solve_velocity: per island contact the velocity constraint (400 B) read and its impulses
#    (28 B) written; per island body its velocity (24 B) read and written;
#  * k_solve_position: per island contact the position constraint (272 B) read; per island body
#    its pose and world inverse inertia (64 B) read and written;
#  * k_update_contacts: per contact its record with manifold and feature cache (332 B) read and
#    written, and for each of its two fixtures the fixture (72 B), fat AABB (24 B) and the body's
#    pose, sweep start, local centre, type and flags (88 B) read.
KERNEL_BYTES = {
    "k_solve_velocity": lambda ic, ib, c: 428 * ic + 48 * ib,
    "k_solve_position": lambda ic, ib, c: 272 * ic + 128 * ib,
    "k_update_contacts": lambda ic, ib, c: (2 * 332 + 2 * (72 + 24 + 88)) * c,
}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=60)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--workload", choices=("stacking", "mixed", "drops"), default="stacking")
    p.add_argument("--solver-mode", type=int, default=0,
                   help="velocity-sweep layout: 1, 2, 3 = 16, 32, 64 islands per wavefront; 0 = tuned")
    p.add_argument("--worlds", type=int, default=4096)
    p.add_argument("--levels", type=int, default=20)
    p.add_argument("--plain-worlds", type=int, default=16)
    p.add_argument("--no-graph", dest="graph", action="store_false",
                   help="skip the graph-replay leg (the K steps again as replays of one captured step)")
    p.add_argument("--cpu-worlds", type=int, default=256)
    p.add_argument("--cpu-steps", type=int, default=60)
    p.add_argument("--latency-steps", type=int, default=20)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--parity-worlds", type=int, default=8,
                   help="worlds of the benchmarked batch compared with the oracle after timing (0: none)")
    p.add_argument("--serve-worlds", type=int, default=512)
    p.add_argument("--serve-rate", type=float, default=500.0)
    p.add_argument("--serve-seconds", type=float, default=4.0,
                   help="seconds of the serving leg after the batched timing (0: none)")
    p.add_argument("--serve-budget-gb", type=float, default=16.0)
    p.add_argument("--serve-catalog", choices=("boxes", "mixed"), default="boxes")
    return p.parse_args()


def cpu_baseline(args, scene_factory):
    """Reference Step (oracle/_ref build) on the host cores this process may run on, one world per
    thread, bounded sample."""
    try:
        from oracle import oracle as O
    except Exception as e:  # pragma: no cover
        return {"value": None, "error": str(e)}
    kind = "reference" if O.available("ref32") else "port"
    lib = "ref32" if kind == "reference" else "oracle"
    threads = len(os.sched_getaffinity(0))
    sample = scene_factory(args.cpu_worlds)
    secs = O.bench(lib, sample, warmup=0, steps=args.cpu_steps, threads=threads)
    value = sample.body_count * args.cpu_steps / secs if secs > 0 else None
    return {"value": value, "unit": "body-steps/s", "cores": threads, "kind": kind,
            "sample": f"{args.cpu_worlds} worlds x {args.cpu_steps} steps from step 0, one world per thread on "
                      f"all {threads} cores this process may run on"}


def step_bytes(t):
    """Compulsory bytes of one step from the batch totals (bx_batch_totals)."""
    convex = t.contacts - t.mesh_world_contacts
    return (STEP_BYTES["awake_body"] * t.awake_bodies + STEP_BYTES["proxy"] * t.fixtures + STEP_BYTES["pair"] * t.pairs
            + STEP_BYTES["contact"] * convex + STEP_BYTES["mesh_contact"] * t.mesh_world_contacts
            + STEP_BYTES["joint"] * t.joints)


class _DeviceWorld:
    def __init__(self, batch, w):
        self.b, self.w = batch, w

    def bodies(self):
        return self.b.bodies(self.w)

    def contacts(self):
        return self.b.contacts(self.w)

    def pairs(self):
        return self.b.pairs(self.w)

    def island_contacts(self):
        return self.b.island_contacts(self.w)


def parity(batch, scene, worlds, steps):
    """The sampled worlds of the benchmarked batch against the oracle stepped as often from the same
    descriptors: every world must be bit-identical (bodies, contacts, pairs, island order)."""
    from oracle.oracle import CpuWorld
    from tests.parity import compare_worlds
    diffs = {}
    for w in worlds:
        ref = CpuWorld("oracle", scene, w)
        ref.step(n=steps)
        d = compare_worlds(_DeviceWorld(batch, w), ref, mask_cache=True)
        if d:
            diffs[w] = d[:2]
        ref.close()
    return {"worlds_compared": len(worlds), "steps": steps, "bit_exact": not diffs,
            "differing_worlds": {str(k): v for k, v in diffs.items()}}


def serve(args):
    """The serving configuration at one arrival rate, in a child process of its own (its JSON line
    is copied into this benchmark's line)."""
    import subprocess
    cmd = [sys.executable, "-m", "bounce_amd.serve_bench", "--worlds", str(args.serve_worlds),
           "--budget-gb", str(args.serve_budget_gb), "--rates", str(args.serve_rate),
           "--seconds", str(args.serve_seconds), "--catalog", args.serve_catalog]
    try:
        r = subprocess.run(cmd, cwd=REPO, capture_output=True, text=True, timeout=600)
    except subprocess.TimeoutExpired:
        return {"error": "timed out"}
    sys.stderr.write(r.stderr[-2000:])
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        return {"error": f"exit {r.returncode}"}
    rep = json.loads(lines[-1])
    rep["workload"] = (f"Poisson requests at {args.serve_rate}/s for {args.serve_seconds} s, each 1-16 steps of "
                       f"one of {args.serve_worlds} {args.serve_catalog} worlds (50-5000 bodies), "
                       f"pool budget {args.serve_budget_gb} GiB")
    return rep


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    from bounce_amd import device, gates, scenes
    if args.workload == "mixed":
        gates.require_mixed_gpu("bench.py --workload mixed")
    if args.serve_seconds > 0 and args.serve_catalog == "mixed":
        gates.require_mixed_gpu("bench.py --serve-catalog mixed")

    def factory(n):
        if args.workload == "drops":
            return scenes.drop_boxes(30, n_worlds=n, seed=21, extent=(12.0, 6.0, 12.0))
        if args.workload == "mixed":
            return scenes.mixed_worlds(n)
        return scenes.pyramids(n, levels=args.levels, seed=2)

    scene = factory(args.worlds)
    bodies_per_world = scene.worlds[0].body_count
    total_bodies = scene.body_count
    ctx = device.Context(0)
    batch = device.Batch(ctx, scene, solver_mode=args.solver_mode)
    # warm-up: at least the layout-tuning steps, so that the timed region runs the chosen layout only
    warmup = max(args.warmup, TUNE_STEPS) if args.solver_mode == 0 else args.warmup
    for _ in range(warmup):
        batch.step()
    batch.sync()
    lanes, tuned_ms = batch.solver_layout()
    tuned_before_timing = args.solver_mode != 0 or all(t > 0 for t in tuned_ms)
    # timed region: K steps, synchronised on both sides
    t0 = time.perf_counter()
    batch.step(n=args.steps)
    batch.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    value = total_bodies * args.steps / wall
    # per-kernel and per-phase times averaged over the K timed steps (HIP events recorded between
    # the kernels on the library's stream)
    kernel_ms = batch.kernel_timings()
    phase_ms = batch.timings()
    dominant = max(KERNEL_BYTES, key=lambda k: kernel_ms.get(k, 0.0))
    # counts of the last timed step over every world (device counters): the compulsory bytes of a step
    tot = batch.totals()
    counts = {"awake_bodies": tot.awake_bodies, "proxies": tot.fixtures, "pairs": tot.pairs,
              "contacts": tot.contacts, "mesh_world_contacts": tot.mesh_world_contacts, "joints": tot.joints,
              "islands": tot.islands, "island_contacts": tot.island_contacts, "bodies": tot.bodies}
    bytes_step = step_bytes(tot)
    status = int(tot.status_or)
    device_bytes = batch.device_bytes()
    model_bytes = KERNEL_BYTES[dominant](tot.island_contacts, tot.awake_bodies, tot.contacts)
    achieved = bytes_step / (kernel_ms[dominant] * 1e-3) / 1e9
    # step latency: single steps, each bracketed by events on the stream
    lat = []
    for _ in range(args.latency_steps):
        batch.step()
        lat.append(batch.timings()[4])
    batch.sync()
    par = None
    if args.parity_worlds > 0:
        sample = sorted(set(np.linspace(0, args.worlds - 1, args.parity_worlds).astype(int).tolist()))
        par = parity(batch, scene, sample, warmup + args.steps + args.latency_steps)
    # HBM bytes per launch of the dominant kernel, from the committed FETCH_SIZE / WRITE_SIZE profile
    # of this workload (bounce_amd/profile_summary.py), when one exists for this configuration
    traffic = None
    traffic_file = os.path.join(REPO, "profiles", "traffic.json")
    if os.path.exists(traffic_file):
        with open(traffic_file) as f:
            tj = json.load(f)
        if args.workload == "stacking" and tj.get("worlds") == args.worlds and tj.get("levels") == args.levels:
            traffic = tj["bytes_per_launch"].get(dominant)
    batch.close()
    # the same W + K steps of a fresh copy of the batch, each a replay of one captured step
    # (bx_batch_step_graph; the capture happens in the first warm-up step)
    graph = None
    if args.graph:
        gb = device.Batch(ctx, scene, solver_mode=args.solver_mode)
        for _ in range(warmup):
            gb.step_graph()
        gb.sync()
        g0 = time.perf_counter()
        gb.step_graph(n=args.steps)
        gb.sync()
        g1 = time.perf_counter()
        graph = {"value": total_bodies * args.steps / (g1 - g0), "unit": "body-steps/s",
                 "ms_per_step": (g1 - g0) / args.steps * 1e3, "steps": args.steps, "warmup": warmup}
        gb.close()
    # plain per-world path: the same kernels, one world per launch sequence
    plain = None
    if args.plain_worlds > 0:
        plain_scene = factory(args.plain_worlds)
        pb = device.Batch(ctx, plain_scene)
        for _ in range(2):
            for w in range(args.plain_worlds):
                pb.step_list([w])
        pb.sync()
        p0 = time.perf_counter()
        plain_steps = 5
        for _ in range(plain_steps):
            for w in range(args.plain_worlds):
                pb.step_list([w])
        pb.sync()
        p1 = time.perf_counter()
        plain = {"value": plain_scene.body_count * plain_steps / (p1 - p0), "unit": "body-steps/s",
                 "sample": f"{args.plain_worlds} worlds x {plain_steps} steps, one world per launch"}
        pb.close()
    ctx.close()
    cpu = None if args.no_cpu else cpu_baseline(args, factory)
    serving = None if args.serve_seconds <= 0 else serve(args)
    out = {
        "metric": "body-steps/sec on 1 MI355X, batched vs plain per-world; p99 step latency",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": {"stacking": "synthetic (seeded pyramids)", "mixed": "synthetic (seeded mixed worlds)",
                 "drops": "synthetic (seeded dropped boxes)"}[args.workload],
        "config": {"workload": (f"batched stacking: {args.worlds} worlds x {args.levels}-level box pyramid "
                                f"({bodies_per_world} bodies incl. ground), dt=1/60, 8/2 iterations, sleep off")
                   if args.workload == "stacking" else
                   (f"batched drops: {args.worlds} worlds x 30 unit boxes dropped from up to 6 m on a ground box "
                    f"(many small islands), dt=1/60, 8/2 iterations, sleep off")
                   if args.workload == "drops" else
                   (f"batched mixed shapes: {args.worlds} worlds, {total_bodies} bodies (lognormal median 200, "
                    f"20..2000 per world; spheres, capsules, 8-16 point hulls, a 10-link motor chain), "
                    f"dt=1/60, 8/2 iterations, sleep off"),
                   "worlds": args.worlds, "bodies": total_bodies, "parallelism": "worlds x islands"},
        "solver_layout": {"islands_per_wavefront": lanes,
                          "lanes_per_island": dict(LAYOUTS).get(lanes),
                          "layouts": [{"islands_per_wavefront": i, "lanes_per_island": n} for i, n in LAYOUTS],
                          "tuned_k_solve_velocity_ms": [round(t, 3) for t in tuned_ms],
                          "tuned_before_timing": tuned_before_timing,
                          "note": "solver_mode 0: each layout timed twice on warm-up steps 4-13 (order "
                                  "1 2 3 4 5 5 4 3 2 1), the one with the smallest sample kept (layout 1 "
                                  "unless another is faster by more than 5%)"},
        "graph_replay": graph,
        "plain_per_world": plain,
        "batched_over_plain": value / plain["value"] if plain else None,
        "step_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                            "samples": len(lat)},
        "phase_ms": {"update_contacts": phase_ms[0], "islands": phase_ms[1], "solve": phase_ms[2],
                     "sync_and_pairs": phase_ms[3], "step": phase_ms[4]},
        "kernel_ms": kernel_ms,
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": bytes_step, "kernel_ms": kernel_ms[dominant], "counts": counts,
                     "bytes_per_body_step": bytes_step / total_bodies,
                     "note": "algorithmic bytes = compulsory bytes of one whole step from the device counts "
                             "(432 per awake body, 52 per proxy, 16 per pair, 580 per convex contact, 1668 per "
                             "mesh contact, 164 per joint), over the dominant kernel's mean time",
                     "kernel_model": {"bytes": model_bytes, "achieved": model_bytes / (kernel_ms[dominant] * 1e-3) / 1e9,
                                      "note": "the kernel's own compulsory bytes: 428 per island contact, 48 per "
                                              "island body"},
                     "whole_step": {"achieved": bytes_step / (wall / args.steps) / 1e9,
                                    "frac": bytes_step / (wall / args.steps) / 1e9 / HBM_PEAK_GBS}},
        "cpu_baseline": cpu,
        "serving": serving,
        "device_bytes": device_bytes,
        "status": status,
        "status_worlds": int(tot.status_worlds),
        "valid": status == 0 and (par is None or par["bit_exact"]) and tuned_before_timing,
        "parity": par,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
