"""Profiling driver for the HBM-traffic counters of the batched stacking workload.

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT -- python3 -m bounce_amd.profile_step
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT -- python3 -m bounce_amd.profile_step

Steps the same workload and the same steps as bench.py's timed region (worlds x levels pyramids,
dt = 1/60, 8/2 iterations; `--warmup` untimed steps, then `--steps` steps), then runs a plain
device copy of a known size (k_copy in libbxcal.so, which reads and writes each byte once): the
copy's FETCH_SIZE and WRITE_SIZE readings give the factor that turns the counters into bytes.
Prints one JSON line with the units of work of the last step and the per-kernel HIP-event times.
"""
import argparse
import ctypes as C
import json
import os
import sys

import numpy as np


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--worlds", type=int, default=4096)
    p.add_argument("--levels", type=int, default=20)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--steps", type=int, default=60)
    p.add_argument("--copy-mib", type=int, default=256)
    p.add_argument("--solver-mode", type=int, default=0, help="bx_batch_opts.solver_mode (fix it to profile one layout)")
    a = p.parse_args()
    from bounce_amd import device, scenes

    scene = scenes.pyramids(a.worlds, levels=a.levels, seed=2)
    ctx = device.Context(0)
    batch = device.Batch(ctx, scene, solver_mode=a.solver_mode)
    for _ in range(a.warmup):
        batch.step()
    batch.sync()
    print(f"warmup done ({a.warmup} steps)", flush=True)
    batch.step(n=a.steps)
    batch.sync()
    kernel_ms = batch.kernel_timings()
    sample = list(range(0, a.worlds, max(1, a.worlds // 16)))
    counts = [batch.counts(w) for w in sample]
    units = {
        "island_bodies": float(np.mean([c.awake_count for c in counts]) * a.worlds),
        "island_contacts": float(np.mean([len(batch.island_contacts(w)) for w in sample]) * a.worlds),
        "contacts": float(np.mean([c.contact_count for c in counts]) * a.worlds),
    }
    batch.close()
    ctx.close()
    cal = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libbxcal.so"))
    cal.bxcal_copy.restype = C.c_double
    cal.bxcal_copy.argtypes = [C.c_uint64, C.c_int]
    n = a.copy_mib << 20
    copy_us = cal.bxcal_copy(n, 3)
    if copy_us < 0:
        raise RuntimeError(f"calibration copy failed ({copy_us})")
    print(json.dumps({"warmup": a.warmup, "steps": a.steps, "worlds": a.worlds, "levels": a.levels,
                      "solver_mode": a.solver_mode,
                      "units": units, "kernel_ms_events": kernel_ms, "copy_bytes_each_way": n,
                      "copy_us": copy_us}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
