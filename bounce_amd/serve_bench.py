"""Serving benchmark: Poisson step requests over resident worlds on one GPU.

A catalog of worlds (`--catalog mixed`: the mixed-shapes workload, spheres, capsules, random hulls
of 8-16 points and a 10-link motor chain per world; `--catalog boxes`: unit boxes dropped on a
ground box; body counts lognormal, median 200, clipped to 50..5000) is served by
bounce_amd.serving.Server from a device pool with a fixed memory budget. Requests arrive
as a Poisson process of the given rates, each asking to advance a uniformly drawn world by 1-16
steps of 1/60 s with 8 velocity and 2 position iterations. One JSON line per rate: request latency
p50/p99, tick latency p50/p99, body-steps/s, mean worlds per tick, evictions.

    python -m bounce_amd.serve_bench --worlds 1024 --budget-gb 2 --rates 500,1000,2000 --seconds 8
"""
import argparse
import json
import sys
import time

import numpy as np

from bounce_amd import device, gates, scenes, serving


def catalog(n_worlds, seed=8, clip=(50, 5000), kind="mixed"):
    if kind == "boxes":
        return scenes.box_worlds(n_worlds, seed=seed, clip=clip)
    return scenes.mixed_worlds(n_worlds=n_worlds, seed=seed, clip=clip, hull_count=256)


def run(n_worlds=1024, budget_gb=2.0, rates=(1000.0,), seconds=8.0, seed=1, max_slots=4096, log=sys.stderr,
        kind="boxes", solver_mode=0):
    """Yields one report per rate as soon as that rate has been served."""
    cat = catalog(n_worlds, kind=kind)
    ctx = device.Context(0)

    def progress(srv):
        print(f"[serve] {len(srv.finished)} requests done, {srv.ticks} ticks, {srv.evictions} evictions",
              file=log, flush=True)

    for rate in rates:
        pool = device.Pool(ctx, cat, budget_bytes=int(budget_gb * (1 << 30)), max_worlds=max_slots,
                           solver_mode=solver_mode)
        srv = serving.Server(pool, cat.worlds)
        reqs = serving.poisson_requests(rate, seconds, n_worlds, max_steps=16, seed=seed)
        print(f"[serve] rate {rate}/s: {len(reqs)} requests over {n_worlds} worlds", file=log, flush=True)
        t = time.perf_counter()
        wall = srv.run(reqs, progress=progress)
        rep = srv.report(wall)
        rep.update({"rate_per_s": rate, "worlds": n_worlds, "budget_gb": budget_gb, "catalog": kind,
                    "solver_mode": solver_mode,
                    "catalog_bodies": int(cat.body_count), "serve_s": time.perf_counter() - t})
        st = pool.stats()
        rep["pool_capacity_gb"] = st.capacity_bytes / 2 ** 30
        print(f"[serve] rate {rate}/s done: p99 request {rep['p99_request_ms']:.2f} ms", file=log, flush=True)
        pool.close()
        yield rep
    ctx.close()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--worlds", type=int, default=1024)
    ap.add_argument("--budget-gb", type=float, default=2.0)
    ap.add_argument("--rates", default="500,1000,2000")
    ap.add_argument("--seconds", type=float, default=8.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--catalog", choices=("mixed", "boxes"), default="boxes")
    ap.add_argument("--solver-mode", type=int, default=0, help="1, 2, 3 = 16, 32, 64 islands per wavefront")
    a = ap.parse_args(argv)
    rates = [float(x) for x in a.rates.split(",")]
    if a.catalog == "mixed":
        gates.require_mixed_gpu("serve_bench --catalog mixed")
    for rep in run(a.worlds, a.budget_gb, rates, a.seconds, a.seed, kind=a.catalog, solver_mode=a.solver_mode):
        print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
