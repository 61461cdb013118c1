"""Writes the golden fixtures under tests/golden/ from the reference build (oracle/_ref).

Run in a checkout that has the reference build (make -C oracle ref):
    python tests/golden/make_golden.py

For every scene it records, per step of b3World::Step(1/60, 8, 2):
  * a SHA-256 of the FP32 reference's body states, contact states (world-list order, manifolds,
    feature ids, feature cache), unique broadphase pairs and island solve order, with the two
    fields the reference leaves uninitialised masked (see tests/parity.py);
  * the full body states and contact states at a few checkpoints;
  * the FP64 reference's body poses at the final step (error of FP32 against FP64).
The fixtures are data only (inputs are the seeded scene recipes in bounce_amd/scenes.py).
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

from bounce_amd import scenes  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tests.parity import masked_bodies, masked_contacts  # noqa: E402

# name -> (scene factory, steps, checkpoints)
SCENES = {
    "hello_world": (lambda: scenes.hello_world(), 90, (30, 90)),
    "box_stack5": (lambda: scenes.box_stack(5), 120, (60, 120)),
    "pyramid10": (lambda: scenes.pyramids(1, levels=10), 120, (60, 120)),
    "pyramid20": (lambda: scenes.pyramids(1, levels=20), 60, (60,)),
    "drop100": (lambda: scenes.drop_boxes(100, extent=(8.0, 20.0, 8.0)), 150, (75, 150)),
    "drop1000": (lambda: scenes.drop_boxes(1000), 60, (60,)),
    "pyramid10_sleep": (lambda: scenes.with_flags(scenes.pyramids(1, levels=10), 1), 240, (120, 240)),
    "drop100_sleep": (lambda: scenes.with_flags(scenes.drop_boxes(100, extent=(8.0, 20.0, 8.0)), 1), 360, (180, 360)),
    "mixed60": (lambda: scenes.mixed_drop(60, seed=0), 240, (120, 240)),
    "mixed60_sensors": (lambda: scenes.mixed_drop(60, seed=1, sensors=0.3), 240, (240,)),
    "mixed150_dense_sleep": (lambda: scenes.with_flags(scenes.mixed_drop(150, seed=2, extent=(4.0, 20.0, 4.0)), 1),
                             300, (150, 300)),
    "chain10_motor": (lambda: scenes.chains(1, 10), 240, (120, 240)),
    "chain10_limit": (lambda: scenes.chains(1, 10, motor=False, limit=True), 240, (240,)),
    "chain10_extra_sleep": (lambda: scenes.with_flags(scenes.chains(1, 10, extra=20), 1), 300, (300,)),
    "mixed_world200": (lambda: scenes.mixed_worlds(1, sizes=[200], hull_count=64), 180, (90, 180)),
    "triangles40_hulls": (lambda: scenes.triangle_drop(40, hull_count=8), 240, (120, 240)),
    "triangles60_round": (lambda: scenes.triangle_drop(60, seed=15, kinds=("sphere", "capsule")), 240, (120, 240)),
    "triangles50_boxes_sleep": (lambda: scenes.with_flags(scenes.triangle_drop(50, seed=16, kinds=("box",)), 1),
                                300, (150, 300)),
    "mesh40_round": (lambda: scenes.mesh_drop(40, kinds=("sphere", "capsule")), 240, (120, 240)),
    "mesh40_boxes": (lambda: scenes.mesh_drop(40, seed=25, kinds=("box",)), 240, (120, 240)),
    "mesh40_hulls": (lambda: scenes.mesh_drop(40, seed=26, hull_count=8), 240, (120, 240)),
    "mesh30_scaled": (lambda: scenes.mesh_drop(30, seed=27, scale=(1.5, 0.5, 1.2)), 240, (240,)),
    "mesh40_sensors": (lambda: scenes.mesh_drop(40, seed=28, sensors=0.3), 240, (240,)),
    "mesh50_sleep": (lambda: scenes.with_flags(scenes.mesh_drop(50, seed=29, hull_count=4), 1), 300, (150, 300)),
}


def state_hash(w):
    h = hashlib.sha256()
    h.update(masked_bodies(w.bodies()).tobytes())
    h.update(masked_contacts(w.contacts()).tobytes())
    h.update(w.pairs().tobytes())
    h.update(w.island_contacts().tobytes())
    return h.digest()


def make(name):
    factory, steps, checkpoints = SCENES[name]
    scene = factory()
    ref = O.CpuWorld("ref32", scene)
    ref64 = O.CpuWorld("ref64", scene)
    hashes = []
    out = {}
    for s in range(1, steps + 1):
        ref.step()
        hashes.append(np.frombuffer(state_hash(ref), dtype=np.uint8))
        if s in checkpoints:
            out[f"bodies_{s}"] = masked_bodies(ref.bodies())
            out[f"contacts_{s}"] = masked_contacts(ref.contacts())
    ref64.step(n=steps)
    out["hashes"] = np.stack(hashes)
    out["f64_final"] = ref64.bodies_f64()
    out["mass"] = ref.mass()
    out["proxy_ids"] = ref.proxy_ids()
    out["steps"] = np.array([steps])
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
    return out


def box_hull_fixture():
    api = O.load("ref32")
    import ctypes as C
    rows = {}
    for ext in [(0.5, 0.5, 0.5), (50.0, 0.5, 50.0), (10.0, 1.0, 10.0), (1.0, 1.0, 1.0), (0.3, 1.7, 2.9)]:
        v = np.zeros(24, np.float32)
        e = np.zeros(120, np.uint32)
        f = np.zeros(6, np.uint32)
        p = np.zeros(24, np.float32)
        c = np.zeros(3, np.float32)
        api["box_hull"](*[C.c_float(x) for x in ext], v.ctypes.data_as(C.POINTER(C.c_float)),
                        e.ctypes.data_as(C.POINTER(C.c_uint32)), f.ctypes.data_as(C.POINTER(C.c_uint32)),
                        p.ctypes.data_as(C.POINTER(C.c_float)), c.ctypes.data_as(C.POINTER(C.c_float)))
        key = "_".join(str(x) for x in ext)
        rows[f"v_{key}"] = v
        rows[f"e_{key}"] = e
        rows[f"f_{key}"] = f
        rows[f"p_{key}"] = p
        rows[f"c_{key}"] = c
    np.savez_compressed(os.path.join(HERE, "box_hulls.npz"), **rows)


if __name__ == "__main__":
    names = sys.argv[1:] or list(SCENES)
    box_hull_fixture()
    for n in names:
        make(n)
        print("wrote", n, flush=True)
