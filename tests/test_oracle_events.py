"""The oracle's contact listener events and contact filter against the reference build: the sequence of
BeginContact / EndContact / PreSolve / PostSolve calls (with the fixtures and the contact's position in
the world list) and the whole world state equal at every step, with a filter installed through the
reference's own b3World::SetContactFilter (tests/event_script.py)."""
import pytest

from oracle import oracle as O
from tests import event_script
from tests.parity import compare_worlds

pytestmark = pytest.mark.skipif(not O.available("ref32"), reason="reference build (make -C oracle ref) not present")


@pytest.mark.parametrize("world", [0, 1])
@pytest.mark.parametrize("filtered", [False, True])
def test_events_oracle_equal_reference(world, filtered):
    scene = event_script.scene()
    orc = O.CpuWorld("oracle", scene, world)
    ref = O.CpuWorld("ref32", scene, world)
    if filtered:
        n = orc.counts().fixture_count
        collide, respond = event_script.filter_tables(n)
        orc.set_contact_filter(collide, respond)
        ref.set_contact_filter(collide, respond)
    kinds = set()
    for step in range(event_script.STEPS):
        orc.step()
        ref.step()
        eo, er = orc.events(), ref.events()
        assert eo.tobytes() == er.tobytes(), f"step {step}: {len(eo)} vs {len(er)} events"
        kinds |= set(eo["type"].tolist())
        d = compare_worlds(orc, ref, mask_cache=True)
        assert not d, f"step {step}: {d}"
    assert kinds == {1, 2, 3, 4}, kinds
