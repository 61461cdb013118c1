"""The oracle restatement of b3Body's setters, force and impulse methods and SetTransform against the
reference build (oracle/_ref), bit for bit on every step of a scripted edit sequence (tests/edit_script.py)."""
import pytest

from oracle import oracle as O
from tests import edit_script
from tests.parity import compare_worlds

pytestmark = pytest.mark.skipif(not O.available("ref32"), reason="reference build (make -C oracle ref) not present")


@pytest.mark.parametrize("world", [0, 1, 2])
def test_edits_oracle_equals_reference(world):
    scene = edit_script.scene()
    orc = O.CpuWorld("oracle", scene, world)
    ref = O.CpuWorld("ref32", scene, world)
    asleep_seen = False
    for step in range(edit_script.STEPS):
        rows = edit_script.edits(step)
        orc.edit_bodies(rows)
        ref.edit_bodies(rows)
        orc.step()
        ref.step()
        d = compare_worlds(orc, ref, mask_cache=True)
        assert not d, f"step {step}: {d}"
        asleep_seen = asleep_seen or (orc.bodies()["flags"][1:] & 1 == 0).any()
    assert asleep_seen, "the script should reach sleeping bodies"
