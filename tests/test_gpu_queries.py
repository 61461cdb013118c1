"""World queries on the device (bx_batch_raycast_single, bx_batch_raycast, bx_batch_query_aabb)
against the oracle, bit for bit: every hit's fixture, fraction, point and normal, and every box's
fixtures, with and without fixture masks, over many worlds in one call.

Worlds: a batch of pyramids, spheres/capsules/boxes, random hulls and triangle grounds; a scaled mesh
heightfield; a 2,500-box world, whose query hierarchy sorts its keys in global memory rather than
LDS; and pool slots. After further steps the hierarchy is rebuilt and the results follow the new
state. The same functions run on the host build of the kernels under AddressSanitizer with the
per-world range check (tests/test_cpu_build.py).
"""
import numpy as np
import pytest

from bounce_amd import capi, scenes
from tests.golden.make_query_golden import QUERY_SCENES, query_set, run_queries

pytestmark = pytest.mark.gpu


PER_RAY, PER_BOX = 24, 256


def _truncate(flat, counts, cap):
    """A flat result list of run_queries cut to the first `cap` entries of each query."""
    out, q = [], 0
    for c in counts:
        out.append(flat[q:q + min(int(c), cap)])
        q += int(c)
    return np.concatenate(out) if out else flat[:0]


def _device_queries(batch, worlds, sets):
    """Runs every world's query set in one call per query kind; returns per-world tuples laid out as
    make_query_golden.run_queries returns them, each query's list cut to the call's capacity
    (PER_RAY hits, PER_BOX fixtures; the counts are the full numbers)."""
    rays, boxes, masks, offsets = [], [], [], []
    off = 0
    for w, (r, b, m, use) in zip(worlds, sets):
        masks.append(m)
        for k in range(len(r)):
            mo = off if use[k] else None
            rays.append((w, r[k, 0], r[k, 1], mo))
            boxes.append((w, b[k, 0], b[k, 1], mo))
        offsets.append(off)
        off += len(m)
    mask = np.concatenate(masks)
    single = batch.raycast_single(rays, mask=mask)
    hits, rc = batch.raycast(rays, per_ray=PER_RAY, mask=mask)
    found, bc = batch.query_aabb(boxes, per_query=PER_BOX, mask=mask)
    out, q = [], 0
    for (r, b, m, use) in sets:
        n = len(r)
        sl = slice(q, q + n)
        out.append((single[sl], np.concatenate([hits[q + k, :min(rc[q + k], PER_RAY)] for k in range(n)]), rc[sl],
                    np.concatenate([found[q + k, :min(bc[q + k], PER_BOX)] for k in range(n)]).astype(np.uint32),
                    bc[sl]))
        q += n
    return out


def _compare(batch, worlds, refs, seed):
    sets = [query_set(r.bodies(), r.counts().fixture_count, seed=seed + i) for i, r in enumerate(refs)]
    got = _device_queries(batch, worlds, sets)
    hits = 0
    for w, (ref, qs, dev) in enumerate(zip(refs, sets, got)):
        exp = list(run_queries(ref, *qs))
        exp[1] = _truncate(exp[1], exp[2], PER_RAY)
        exp[3] = _truncate(exp[3], exp[4], PER_BOX)
        for name, a, b in zip(("single", "hits", "ray counts", "box fixtures", "box counts"), dev, exp):
            assert a.tobytes() == b.tobytes(), f"world {worlds[w]}: {name} differ from the oracle"
        hits += int(exp[2].sum())
    return hits


def test_queries_equal_oracle_batched(gpu_ctx):
    from bounce_amd import device
    from oracle.oracle import CpuWorld
    names = ["pyramid10", "mixed60", "triangles40_hulls", "mixed_world200"]
    scene = scenes.concat([QUERY_SCENES[n][0]() for n in names])
    print("[queries] batch of 4 worlds, 60 steps, queried twice", flush=True)
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    refs = [CpuWorld("oracle", scene, w) for w in range(len(names))]
    for _ in range(40):
        batch.step()
        for r in refs:
            r.step()
    hits = _compare(batch, list(range(len(names))), refs, seed=11)
    assert hits > 100
    for _ in range(20):  # the hierarchy is rebuilt after these steps
        batch.step()
        for r in refs:
            r.step()
    _compare(batch, list(range(len(names))), refs, seed=31)
    batch.close()


def test_queries_equal_oracle_mesh(gpu_ctx):
    from bounce_amd import device
    from oracle.oracle import CpuWorld
    factory, steps = QUERY_SCENES["mesh30_scaled"]
    scene = factory()
    print("[queries] scaled mesh heightfield", flush=True)
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    ref = CpuWorld("oracle", scene)
    for _ in range(60):
        batch.step()
        ref.step()
    _compare(batch, [0], [ref], seed=5)
    batch.close()


def test_queries_equal_oracle_large_world(gpu_ctx):
    """A world of 2,501 fixtures (more than the LDS key sort holds) beside two small worlds."""
    from bounce_amd import device
    from oracle.oracle import CpuWorld
    scene = scenes.concat([scenes.drop_boxes(2500, extent=(20.0, 40.0, 20.0), seed=3), scenes.pyramids(1, levels=5),
                           scenes.drop_boxes(1, seed=4)])
    print("[queries] 2,500-box world", flush=True)
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    refs = [CpuWorld("oracle", scene, w) for w in range(3)]
    for _ in range(20):
        batch.step()
        for r in refs:
            r.step()
    _compare(batch, [0, 1, 2], refs, seed=17)  # boxes over the pile find more fixtures than PER_BOX
    batch.close()


def test_queries_on_pool_slots(gpu_ctx):
    from bounce_amd import device
    from oracle.oracle import CpuWorld
    parts = [scenes.pyramids(1, levels=6, seed=3), scenes.drop_boxes(40, extent=(6.0, 12.0, 6.0), seed=9)]
    scene = scenes.concat(parts)
    print("[queries] pool slots", flush=True)
    pool = device.Pool(gpu_ctx, scene, budget_bytes=64 << 20, solver_mode=1)
    slots = [pool.admit(scene.worlds[w]) for w in range(2)]
    refs = [CpuWorld("oracle", scene, w) for w in range(2)]
    for _ in range(30):
        pool.step_list(slots)
        for r in refs:
            r.step()
    _compare(pool, slots, refs, seed=23)
    pool.close()


def test_query_arguments_are_checked(gpu_ctx):
    from bounce_amd import device
    batch = device.Batch(gpu_ctx, scenes.pyramids(1, levels=3), solver_mode=1)
    with pytest.raises(device.BxError):
        batch.raycast_single([(5, (0, 1, 0), (0, -1, 0))])  # no such world
    with pytest.raises(device.BxError):
        batch.raycast_single([(0, (0, 1, 0), (0, -1, 0), 3)], mask=np.ones(4, np.uint8))  # mask too short
    with pytest.raises(device.BxError):
        batch.query_aabb([(0, (0, np.nan, 0), (1, 1, 1))])
    hit = batch.raycast_single([(0, (0.0, 20.0, 0.0), (0.0, -1.0, 0.0))])[0]
    assert hit["fixture"] != capi.BX_NULL_INDEX
    batch.close()
