"""Runs workloads through the CPU build of the device library and compares every world with the oracle.

    python tests/cpu_build/check.py --lib cpu  --workload mixed --first 0 --worlds 2048 --steps 25
    LD_PRELOAD=$(/opt/rocm/llvm/bin/clang++ -print-file-name=libclang_rt.asan-x86_64.so) \\
    ASAN_OPTIONS=detect_leaks=0 python tests/cpu_build/check.py --lib asan --workload dense --steps 300

The library is the product's own kernels and host code compiled for the host (Makefile in this
directory); with --lib asan every access a world's kernel code makes outside that world's ranges of the
per-world arrays, or outside any allocation, stops the run with a report. Worlds are stepped together
as one batch, as on the device; at every comparison point each world's bodies, contacts (with ids,
impulses and feature cache), pairs and island contact order must equal the oracle's bit for bit, and
every world's status must be 0.
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from bounce_amd import device, scenes  # noqa: E402
from oracle.oracle import CpuWorld  # noqa: E402
from tests.parity import compare_worlds  # noqa: E402

LIBS = {"cpu": "libbxstep_cpu.so", "asan": "libbxstep_asan.so"}


def subset(s, first, count):
    out = scenes.SceneSet(f"{s.name}[{first}:{first + count}]")
    out.hulls = list(s.hulls)
    out.meshes = list(getattr(s, "meshes", []))
    out.worlds = s.worlds[first:first + count]
    return out


def workload(name, first, count, seed):
    if name == "pyramids":
        return scenes.pyramids(count, levels=20, seed=2)
    if name == "dense":  # 150 spheres, capsules and boxes in a 4 x 20 x 4 m column, sleeping on
        return scenes.with_flags(scenes.mixed_drop(150, seed=2, extent=(4.0, 20.0, 4.0)), 1)
    if name == "mixed":  # the batched mixed-shapes workload: slice [first, first + count) of 16,384 worlds
        return subset(scenes.mixed_worlds(16384, seed=seed), first, count)
    if name == "serving":  # the serving catalog (50-5,000 bodies, 256 hulls)
        return subset(scenes.mixed_worlds(n_worlds=first + count, seed=seed, clip=(50, 5000), hull_count=256),
                      first, count)
    raise SystemExit(f"unknown workload {name}")


def dump(scene, path):
    """Writes the shape library and world descriptors as little-endian words for the MemorySanitizer
    driver (msan_driver.cpp): hull count, each hull's counts and arrays, mesh count, each mesh, world
    count, each world's gravity, flags and its body, fixture and joint descriptors as raw records."""
    u = lambda *x: np.asarray(x, dtype=np.uint32).tobytes()
    with open(path, "wb") as f:
        f.write(b"BXSC" + u(1, len(scene.hulls)))
        for h in scene.hulls:
            f.write(u(len(h.vertices), len(h.edges), len(h.faces)))
            for arr, dt in ((h.vertices, np.float32), (h.edges, np.uint32), (h.faces, np.uint32),
                            (h.planes, np.float32), (np.asarray(h.centroid), np.float32)):
                f.write(np.ascontiguousarray(arr, dtype=dt).tobytes())
        f.write(u(len(scene.meshes)))
        for v, t in scene.meshes:
            f.write(u(len(v), len(t)))
            f.write(np.ascontiguousarray(v, dtype=np.float32).tobytes())
            f.write(np.ascontiguousarray(t, dtype=np.uint32).tobytes())
        f.write(u(len(scene.worlds)))
        for w in scene.worlds:
            f.write(np.asarray(w.gravity, dtype=np.float32).tobytes() + u(w.flags))
            for arr in (w.bodies, w.fixtures, w.joints):
                f.write(u(len(arr)) + np.ascontiguousarray(arr).tobytes())


class DeviceWorld:
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", choices=sorted(LIBS), default="cpu")
    ap.add_argument("--workload", default="mixed")
    ap.add_argument("--first", type=int, default=0)
    ap.add_argument("--worlds", type=int, default=64)
    ap.add_argument("--seed", type=int, default=8)
    ap.add_argument("--steps", type=int, default=25)
    ap.add_argument("--compare-every", type=int, default=0, help="0: compare after the last step only")
    ap.add_argument("--no-oracle", action="store_true", help="only run (address checks and status)")
    ap.add_argument("--solver-mode", type=int, default=1)
    ap.add_argument("--dump", help="write the workload for msan_driver.cpp to this file and stop")
    a = ap.parse_args(argv)
    if a.dump:
        dump(workload(a.workload, a.first, a.worlds, a.seed), a.dump)
        return 0

    device.load(os.path.join(HERE, "_out", LIBS[a.lib]))
    t0 = time.time()
    scene = workload(a.workload, a.first, a.worlds, a.seed)
    n = len(scene.worlds)
    sizes = [w.body_count for w in scene.worlds]
    print(f"[check] {a.workload} worlds {a.first}..{a.first + n - 1}: {sum(sizes)} bodies "
          f"(max {max(sizes)}), built in {time.time() - t0:.1f} s", flush=True)
    ctx = device.Context(0)
    batch = device.Batch(ctx, scene, solver_mode=a.solver_mode)
    refs = None if a.no_oracle else [CpuWorld("oracle", scene, w) for w in range(n)]
    every = a.compare_every or a.steps
    failures = 0
    for step in range(1, a.steps + 1):
        t = time.time()
        batch.step()
        msg = f"[check] step {step}: {time.time() - t:.2f} s"
        if refs is not None:
            for r in refs:
                r.step()
        if step % every == 0 or step == a.steps:
            status = [batch.counts(w).status for w in range(n)]
            bad = [w for w in range(n) if status[w]]
            if bad:
                print(f"[check] step {step}: status nonzero in worlds {bad[:10]} "
                      f"({[hex(status[w]) for w in bad[:10]]})", flush=True)
                failures += 1
            if refs is not None:
                ndiff = 0
                for w in range(n):
                    d = compare_worlds(DeviceWorld(batch, w), refs[w], mask_cache=True)
                    if d:
                        ndiff += 1
                        if ndiff <= 5:
                            print(f"[check] step {step} world {a.first + w}: {d}", flush=True)
                msg += f", {n - ndiff}/{n} worlds equal to the oracle"
                failures += ndiff > 0
        print(msg, flush=True)
    batch.close()
    ctx.close()
    print(f"[check] {'FAIL' if failures else 'OK'} in {time.time() - t0:.0f} s", flush=True)
    return 1 if failures else 0


if __name__ == "__main__":
    sys.exit(main())
