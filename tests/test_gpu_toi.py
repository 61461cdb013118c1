"""Time of impact and shape casts on the device (bx_batch_time_of_impact, bx_batch_shape_cast,
bx_batch_shape_cast_single) against the oracle, bit for bit.

- 3,000 seeded problems of the three b3TimeOfImpact forms over spheres, capsules, triangles, boxes and
  random hulls (tests/toi_cases.py): every state, time and iteration count.
- Shape casts of every cast shape kind over a batch of stepped worlds (pyramid, mixed shapes, triangles
  with hulls, a mixed world of 200 bodies) and a scaled mesh heightfield, with and without fixture
  masks: every hit's fixture, mesh triangle, fraction, point and normal in order, and the single cast.
- Arguments the library refuses. The same tests run on the host build of the kernels under
  AddressSanitizer with the per-world range check (tests/test_cpu_build.py).
"""
import numpy as np
import pytest

from bounce_amd import capi, scenes
from tests import toi_cases as T
from tests.golden.make_query_golden import QUERY_SCENES

pytestmark = pytest.mark.gpu

PER_CAST = 32


def test_time_of_impact_equals_oracle(gpu_ctx):
    from bounce_amd import device
    from oracle import oracle as O
    lib = T.library()
    batch = device.Batch(gpu_ctx, lib, solver_mode=1)
    print("[toi] 3,000 problems", flush=True)
    for seed in (1, 2):
        qs = T.toi_queries(1500, seed=seed)
        got = batch.time_of_impact(qs)
        want = O.time_of_impact("oracle", lib, qs)
        assert got.tobytes() == want.tobytes(), np.where(got != want)[0][:10]
    batch.close()


def _compare_casts(batch, refs, worlds, seed, n_hulls):
    hits_total = 0
    for w, ref in zip(worlds, refs):
        nf = ref.counts().fixture_count
        mask = (np.random.default_rng(seed + w).random(nf) < 0.6).astype(np.uint8)
        plain = T.cast_queries(ref.bodies(), 40, seed=seed + 7 * w, n_hulls=n_hulls, world=w)
        masked = T.cast_queries(ref.bodies(), 20, seed=seed + 7 * w + 1, n_hulls=n_hulls, world=w, mask=0)
        for casts, m in ((plain, None), (masked, mask)):
            single = batch.shape_cast_single(casts, mask=m)
            hits, counts = batch.shape_cast(casts, per_cast=PER_CAST, mask=m)
            for i in range(len(casts)):
                want = ref.shape_cast(casts[i], m)
                assert counts[i] == len(want), f"world {w} cast {i}: {counts[i]} hits, oracle {len(want)}"
                k = min(len(want), PER_CAST)
                assert hits[i, :k].tobytes() == want[:k].tobytes(), f"world {w} cast {i}"
                assert single[i].tobytes() == ref.shape_cast_single(casts[i], m).tobytes(), f"world {w} cast {i}"
                hits_total += len(want)
    return hits_total


def test_shape_casts_equal_oracle(gpu_ctx):
    from bounce_amd import device
    from oracle.oracle import CpuWorld
    names = ["pyramid10", "mixed60", "triangles40_hulls", "mixed_world200"]
    scene = scenes.concat([QUERY_SCENES[n][0]() for n in names])
    print("[casts] batch of 4 worlds", flush=True)
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    refs = [CpuWorld("oracle", scene, w) for w in range(len(names))]
    for _ in range(40):
        batch.step()
        for r in refs:
            r.step()
    assert _compare_casts(batch, refs, list(range(len(names))), 11, len(scene.hulls)) > 200
    batch.close()


def test_shape_casts_on_a_mesh_equal_oracle(gpu_ctx):
    from bounce_amd import device
    from oracle.oracle import CpuWorld
    factory, _ = QUERY_SCENES["mesh30_scaled"]
    scene = factory()
    print("[casts] scaled mesh heightfield", flush=True)
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    ref = CpuWorld("oracle", scene)
    for _ in range(60):
        batch.step()
        ref.step()
    assert _compare_casts(batch, [ref], [0], 5, len(scene.hulls)) > 50
    batch.close()


def test_toi_and_cast_arguments_are_checked(gpu_ctx):
    from bounce_amd import device
    lib = T.library()
    batch = device.Batch(gpu_ctx, lib, solver_mode=1)
    qs = T.toi_queries(3, seed=9)
    bad = (capi.bx_toi_query * 1)()
    bad[0] = qs[1]  # a translation problem
    bad[0].max_iterations = 0
    with pytest.raises(device.BxError):
        batch.time_of_impact(bad)
    bad[0] = qs[0]
    bad[0].shape1.shape_type = capi.BX_SHAPE_MESH
    with pytest.raises(device.BxError):
        batch.time_of_impact(bad)
    bad[0] = qs[0]
    bad[0].shape1.shape_type = capi.BX_SHAPE_HULL
    bad[0].shape1.shape_index = 99
    with pytest.raises(device.BxError):
        batch.time_of_impact(bad)
    casts = T.cast_queries(np.zeros(1, dtype=[("position", np.float32, 3)]), 1, seed=1, n_hulls=1)
    casts[0].world = 3  # no such world
    with pytest.raises(device.BxError):
        batch.shape_cast_single(casts)
    casts[0].world = 0
    casts[0].displacement[1] = float("nan")
    with pytest.raises(device.BxError):
        batch.shape_cast_single(casts)
    casts[0].displacement[1] = 0.0
    assert batch.shape_cast_single(casts)[0]["fixture"] in (0, capi.BX_NULL_INDEX)
    batch.close()
