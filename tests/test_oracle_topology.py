"""The oracle restatement of topology changes between steps (b3World::CreateBody / DestroyBody /
CreateJoint / DestroyJoint, b3Body::CreateFixture / DestroyFixture) against the reference build
(oracle/_ref), bit for bit after every edit and every step of the scripted scenarios of
tests/topology_script.py: bodies, contacts with manifolds and feature ids, broadphase pairs with their
proxy ids (the dynamic tree's node free list), and island order."""
import pytest

from oracle import oracle as O
from tests import topology_script as TS
from tests.parity import compare_worlds

pytestmark = pytest.mark.skipif(not O.available("ref32"), reason="reference build (make -C oracle ref) not present")


@pytest.mark.parametrize("name", list(TS.SCENARIOS))
def test_topology_oracle_equals_reference(name):
    sc = TS.SCENARIOS[name]()
    orc = O.CpuWorld("oracle", sc.scene)
    ref = O.CpuWorld("ref32", sc.scene)
    for step in range(sc.steps):
        if step in sc.edits:
            ops, adds = sc.edits[step]
            orc.edit_topology(ops, adds, sc.library)
            ref.edit_topology(ops, adds, sc.library)
            d = compare_worlds(orc, ref, mask_cache=True)
            assert not d, f"after the edit before step {step}: {d}"
        orc.step()
        ref.step()
        d = compare_worlds(orc, ref, mask_cache=True)
        assert not d, f"step {step}: {d}"
