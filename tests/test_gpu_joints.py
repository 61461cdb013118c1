"""The sphere, weld, spring, mouse, friction, motor, cone, prismatic and wheel joints on the device
against the oracle, bit for bit on every step: four worlds of every joint type (sleeping off and on) in
one batch, under the one-lane and the eight-lane solver layouts; joints of these types created and
destroyed between steps; and descriptors of no joint type refused. The same tests run on the host
build of the kernels under AddressSanitizer (tests/test_cpu_build.py)."""
import numpy as np
import pytest

from bounce_amd import capi, scenes
from tests.parity import compare_worlds

pytestmark = pytest.mark.gpu


class _W:
    def __init__(self, b, w):
        self.b, self.w = b, w

    def bodies(self):
        return self.b.bodies(self.w)

    def contacts(self):
        return self.b.contacts(self.w)

    def pairs(self):
        return self.b.pairs(self.w)

    def island_contacts(self):
        return self.b.island_contacts(self.w)


@pytest.mark.parametrize("solver_mode", [1, 5])
def test_joint_zoo_device_equals_oracle(gpu_ctx, solver_mode):
    from bounce_amd import device
    from oracle.oracle import CpuWorld
    scene = scenes.concat([scenes.joint_zoo(2, seed=21), scenes.joint_zoo(2, seed=22, sleeping=True)])
    print(f"[joints] joint zoo, solver mode {solver_mode}", flush=True)
    batch = device.Batch(gpu_ctx, scene, solver_mode=solver_mode)
    refs = [CpuWorld("oracle", scene, w) for w in range(4)]
    for step in range(240):
        batch.step()
        for w, r in enumerate(refs):
            r.step()
            d = compare_worlds(_W(batch, w), r, mask_cache=True)
            assert not d, f"step {step} world {w}: {d}"
    batch.close()


def test_joints_created_and_destroyed_between_steps(gpu_ctx):
    from bounce_amd import device
    from oracle.oracle import CpuWorld
    scene = scenes.joint_zoo(1, seed=23)
    print("[joints] created and destroyed between steps", flush=True)
    batch = device.Batch(gpu_ctx, scene, solver_mode=1)
    ref = CpuWorld("oracle", scene)
    T = capi
    adds = scenes.joint_array(2)
    adds[0]["type"], adds[0]["body_a"], adds[0]["body_b"], adds[0]["use_initialize"] = T.BX_JOINT_WELD, 2, 5, 1
    adds[0]["anchor"] = (-1.5, 5.5, 0.0)
    adds[0]["frequency_hz"], adds[0]["damping_ratio"] = np.float32(1.5), np.float32(0.2)
    adds[1]["type"], adds[1]["body_a"], adds[1]["body_b"], adds[1]["use_initialize"] = T.BX_JOINT_SPRING, 4, 6, 1
    adds[1]["anchor"], adds[1]["anchor_b"] = (-3.0, 4.9, 0.1), (0.0, 4.6, 0.0)
    edits = {40: [(T.BX_TOPO_CREATE_JOINT, 0)], 70: [(T.BX_TOPO_DESTROY_JOINT, 4), (T.BX_TOPO_CREATE_JOINT, 1)],
             100: [(T.BX_TOPO_DESTROY_JOINT, 0), (T.BX_TOPO_DESTROY_BODY, 8)]}
    world = scenes.World(scenes.body_array(0), scenes.fixture_array(0), adds)
    for step in range(160):
        if step in edits:
            batch.edit_topology(0, edits[step], world)
            ref.edit_topology(edits[step], world)
        batch.step()
        ref.step()
        d = compare_worlds(_W(batch, 0), ref, mask_cache=True)
        assert not d, f"step {step}: {d}"
    batch.close()


def test_unsupported_joint_types_are_refused(gpu_ctx):
    from bounce_amd import device
    scene = scenes.joint_zoo(1, seed=21)
    for t in (0, capi.BX_JOINT_WHEEL + 1):
        scene.worlds[0].joints[0]["type"] = t
        with pytest.raises(device.BxError):
            device.Batch(gpu_ctx, scene, solver_mode=1)


def test_joint_worlds_in_a_pool_equal_oracle(gpu_ctx):
    """Worlds of every joint type admitted into a pool and stepped in changing mixes (one launch
    sequence per tick over the listed slots) equal the oracle world stepped as often, at the end."""
    from bounce_amd import device
    from oracle.oracle import CpuWorld
    scene = scenes.concat([scenes.joint_zoo(2, seed=31), scenes.joint_zoo(1, seed=32, sleeping=True)])
    print("[joints] pool of joint worlds: 60 ticks", flush=True)
    pool = device.Pool(gpu_ctx, scene, budget_bytes=64 << 20, max_worlds=3, solver_mode=5)
    slots = [pool.admit(w) for w in scene.worlds]
    steps = [0, 0, 0]
    for t in range(60):
        pick = [i for i in range(3) if (t + i) % 4 != 0]
        pool.step_list([slots[i] for i in pick])
        for i in pick:
            steps[i] += 1
    for i in range(3):
        ref = CpuWorld("oracle", scene, i)
        ref.step(n=steps[i])
        d = compare_worlds(_W(pool, slots[i]), ref)
        assert not d, f"world {i} after {steps[i]} steps: {d}"
    pool.close()
