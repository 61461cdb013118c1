"""The device library's own kernels and host code, compiled for the host (tests/cpu_build), against the
oracle; the AddressSanitizer build also checks that a world's kernel code touches only that world's
ranges of the per-world arrays, and the MemorySanitizer build that no branch, index or address depends on
memory never written.

These run the product's code paths on the CPU: they find out-of-range and uninitialized accesses that
a GPU would turn into memory faults or arbitrary values. Each run is a subprocess (the CPU library
takes the place of libbxstep.so in it).
"""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "cpu_build")
CLANG = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not os.path.exists(CLANG) or not shutil.which("make"),
                                reason="clang++ of the ROCm toolchain is not present")


def _make(target):
    subprocess.run(["make", "-C", BUILD, target], check=True, capture_output=True, text=True)


def _asan_runtime():
    out = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    return out.stdout.strip()


def _check(lib, *args, timeout=900):
    env = dict(os.environ)
    if lib == "asan":
        env["LD_PRELOAD"] = _asan_runtime()
        env["ASAN_OPTIONS"] = "detect_leaks=0:detect_stack_use_after_return=0"
    r = subprocess.run([sys.executable, os.path.join(BUILD, "check.py"), "--lib", lib, *args], env=env,
                       capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "[check] OK" in r.stdout
    return r.stdout


def test_cpu_build_mixed_worlds_equal_oracle():
    _make("all")
    out = _check("cpu", "--workload", "mixed", "--first", "0", "--worlds", "24", "--steps", "8", "--compare-every", "4")
    assert "24/24 worlds equal to the oracle" in out


def test_cpu_build_address_check_mixed_and_dense_pile():
    _make("asan")
    # the largest worlds of the mixed-shapes batch's first slice, and the dense sleeping pile
    _check("asan", "--workload", "mixed", "--first", "0", "--worlds", "12", "--steps", "4")
    _check("asan", "--workload", "dense", "--worlds", "1", "--steps", "40", "--compare-every", "20")


def test_cpu_build_address_check_catches_cross_world_access(tmp_path):
    """The per-world check itself: a read one element past world 0's range of a per-world array (the
    next world's first element of the same flat array) is reported; the last element of the range is
    not."""
    _make("asan")
    script = tmp_path / "probe.py"
    script.write_text(f"""
import ctypes as C, sys
sys.path.insert(0, {os.path.dirname(HERE)!r})
from bounce_amd import device, scenes
lib = device.load({os.path.join(BUILD, "_out", "libbxstep_asan.so")!r})
ctx = device.Context(0)
batch = device.Batch(ctx, scenes.pyramids(2, levels=3, seed=1))
batch.step()
lib.bxcpu_probe_world_range(batch.h, 0, int(sys.argv[1]))
print("probe read done", flush=True)
""")
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(), ASAN_OPTIONS="detect_leaks=0")
    inside = subprocess.run([sys.executable, str(script), "0"], env=env, capture_output=True, text=True)
    assert inside.returncode == 0 and "probe read done" in inside.stdout, inside.stderr[-2000:]
    past = subprocess.run([sys.executable, str(script), "1"], env=env, capture_output=True, text=True)
    assert past.returncode != 0 and "use-after-poison" in past.stderr, past.stderr[-2000:]


def test_cpu_build_events_and_filter(tmp_path):
    """The device tests of contact events and the filter (tests/test_gpu_events.py) on the host build,
    which compiles event recording in, under AddressSanitizer with the per-world check: the event
    sequence and the world state equal the oracle's at every step, without and with the filter tables,
    and the filter alone over the mixed drop."""
    _make("asan")
    script = tmp_path / "events.py"
    script.write_text(f"""
import sys
sys.path.insert(0, {os.path.dirname(HERE)!r})
from bounce_amd import device
device.load({os.path.join(BUILD, "_out", "libbxstep_asan.so")!r})
from tests import test_gpu_events as t
ctx = device.Context(0)
for filtered in (False, True):
    t.test_events_device_equal_oracle(ctx, filtered)
t.test_contact_filter_device_equal_oracle(ctx)
print("events and filter equal to the oracle", flush=True)
""")
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(), ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0")
    r = subprocess.run([sys.executable, str(script)], env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and "equal to the oracle" in r.stdout, r.stdout[-2000:] + r.stderr[-3000:]


def test_cpu_build_world_queries(tmp_path):
    """The device tests of the world queries (tests/test_gpu_queries.py) on the host build under
    AddressSanitizer with the per-world check: the hierarchy build (LDS and global-memory key sorts),
    the stackless walk and every shape's ray cast touch only the queried world's ranges, and every
    result equals the oracle's."""
    _make("asan")
    script = tmp_path / "queries.py"
    script.write_text(f"""
import sys
sys.path.insert(0, {os.path.dirname(HERE)!r})
from bounce_amd import device
device.load({os.path.join(BUILD, "_out", "libbxstep_asan.so")!r})
from tests import test_gpu_queries as t
ctx = device.Context(0)
for name in dir(t):
    if name.startswith("test_"):
        getattr(t, name)(ctx)
print("queries equal to the oracle", flush=True)
""")
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(), ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0")
    r = subprocess.run([sys.executable, str(script)], env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and "equal to the oracle" in r.stdout, r.stdout[-2000:] + r.stderr[-3000:]


def test_cpu_build_topology_edits(tmp_path):
    """The device tests of topology changes (tests/test_gpu_topology.py) on the host build under
    AddressSanitizer with the per-world check: the rebuilt batches carry every world's state and the
    edited world equals the oracle after every edit and step."""
    _make("asan")
    script = tmp_path / "topology.py"
    script.write_text(f"""
import sys
sys.path.insert(0, {os.path.dirname(HERE)!r})
from bounce_amd import device
device.load({os.path.join(BUILD, "_out", "libbxstep_asan.so")!r})
from tests import test_gpu_topology as t, topology_script as TS
ctx = device.Context(0)
for name in TS.SCENARIOS:
    t.test_topology_device_equals_oracle(ctx, name)
t.test_topology_edit_in_a_batch_keeps_other_worlds(ctx)
t.test_topology_errors_leave_the_batch_unchanged(ctx)
print("topology equal to the oracle", flush=True)
""")
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(), ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0")
    r = subprocess.run([sys.executable, str(script)], env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and "equal to the oracle" in r.stdout, r.stdout[-2000:] + r.stderr[-3000:]


def test_cpu_build_time_of_impact_and_shape_casts(tmp_path):
    """The device tests of the time of impact and shape casts (tests/test_gpu_toi.py) on the host build
    under AddressSanitizer with the per-world check: every problem and cast equals the oracle, and each
    cast touches only its world's ranges."""
    _make("asan")
    script = tmp_path / "toi.py"
    script.write_text(f"""
import sys
sys.path.insert(0, {os.path.dirname(HERE)!r})
from bounce_amd import device
device.load({os.path.join(BUILD, "_out", "libbxstep_asan.so")!r})
from tests import test_gpu_toi as t
ctx = device.Context(0)
for name in dir(t):
    if name.startswith("test_"):
        getattr(t, name)(ctx)
print("time of impact and casts equal to the oracle", flush=True)
""")
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(), ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0")
    r = subprocess.run([sys.executable, str(script)], env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and "equal to the oracle" in r.stdout, r.stdout[-2000:] + r.stderr[-3000:]


def test_cpu_build_joints(tmp_path):
    """The device tests of the sphere, weld, spring, mouse, friction and motor joints
    (tests/test_gpu_joints.py) on the host build under AddressSanitizer with the per-world check."""
    _make("asan")
    script = tmp_path / "joints.py"
    script.write_text(f"""
import sys
sys.path.insert(0, {os.path.dirname(HERE)!r})
from bounce_amd import device
device.load({os.path.join(BUILD, "_out", "libbxstep_asan.so")!r})
from tests import test_gpu_joints as t
ctx = device.Context(0)
for mode in (1, 5):
    t.test_joint_zoo_device_equals_oracle(ctx, mode)
t.test_joints_created_and_destroyed_between_steps(ctx)
t.test_unsupported_joint_types_are_refused(ctx)
print("joints equal to the oracle", flush=True)
""")
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(), ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0")
    r = subprocess.run([sys.executable, str(script)], env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and "equal to the oracle" in r.stdout, r.stdout[-2000:] + r.stderr[-3000:]


def test_cpu_build_body_and_world_edits(tmp_path):
    """The device test of body and world edits (tests/test_gpu_edits.py) on the host build under
    AddressSanitizer with the per-world check: every step of the scripted setters, forces, impulses,
    transforms, fixed rotations and world settings equals the oracle."""
    _make("asan")
    script = tmp_path / "edits.py"
    script.write_text(f"""
import sys
sys.path.insert(0, {os.path.dirname(HERE)!r})
from bounce_amd import device
device.load({os.path.join(BUILD, "_out", "libbxstep_asan.so")!r})
from tests import test_gpu_edits as t
ctx = device.Context(0)
t.test_edits_device_equals_oracle(ctx)
print("edits equal to the oracle", flush=True)
""")
    env = dict(os.environ, LD_PRELOAD=_asan_runtime(), ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0")
    r = subprocess.run([sys.executable, str(script)], env=env, capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and "equal to the oracle" in r.stdout, r.stdout[-2000:] + r.stderr[-3000:]


@pytest.mark.slow
def test_cpu_build_memory_sanitizer_mixed_worlds(tmp_path):
    _make("msan")
    scene = tmp_path / "mixed.bin"
    subprocess.run([sys.executable, os.path.join(BUILD, "check.py"), "--workload", "mixed", "--first", "0",
                    "--worlds", "8", "--dump", str(scene)], check=True)
    r = subprocess.run([os.path.join(BUILD, "_out", "bxcpu_msan"), str(scene), "3"],
                       env=dict(os.environ, MSAN_OPTIONS="halt_on_error=1"), capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0 and "0 with nonzero status" in r.stdout, r.stderr[-3000:]


def test_cpu_build_narrowphase_fuzz(tmp_path):
    """Adversarial shapes and poses (coincident, degenerate, denormal, huge, NaN, infinite) through every
    convex narrowphase routine of the device path, GJK and the sensor test, under AddressSanitizer and
    UndefinedBehaviorSanitizer, over the mixed-shapes workload's hull library: every point count,
    feature id and cached feature stays in range and every call returns."""
    _make("fuzz")
    scene = tmp_path / "hulls.bin"
    subprocess.run([sys.executable, os.path.join(BUILD, "check.py"), "--workload", "mixed", "--first", "0",
                    "--worlds", "1", "--dump", str(scene)], check=True)
    r = subprocess.run([os.path.join(BUILD, "_out", "fuzz_narrowphase"), str(scene), "40000", "7"],
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=0"), capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0 and "all in range" in r.stdout, r.stderr[-3000:]
