"""The drop-in C++ API (bounce_amd/include/bounce/bounce.h over libbxstep).

CPU: programs written against the reference's API compile and link against the drop-in headers
(the reference's own examples/hello_world/main.cpp when the reference tree is present), and the
drop-in b3BoxHull builds the reference's box table bit for bit. The programs, linked with the host build
of the device library (tests/cpu_build), print exactly the oracle's lines; the oracle's replay of the
program that changes its world while it runs equals the reference build's. GPU: the same programs step
on the device and print the oracle's lines: a hello world, a sphere, a capsule and a box on a b3Mesh
heightfield, and a world whose bodies, fixtures and joints are created and destroyed between steps while
forces push its bodies every step (tests/drop_in/spawn_apply.cpp), with its contact lists and queries.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from bounce_amd import hulls, scenes

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(REPO, "bounce_amd", "include")
LIB = os.path.join(REPO, "bounce_amd", "lib")
REF_HELLO = "/root/reference/examples/hello_world/main.cpp"
HELLO_BIN = os.path.join(LIB, "bx_hello_world")
MESH_BIN = os.path.join(LIB, "bx_mesh_ground")
SPAWN_BIN = os.path.join(LIB, "bx_spawn_apply")
CPU_BUILD = os.path.join(REPO, "tests", "cpu_build")
CPU_LIB = os.path.join(CPU_BUILD, "_out", "libbxstep_cpu.so")


def _have_libs():
    return os.path.exists(os.path.join(LIB, "libbounce_amd.so")) and os.path.exists(os.path.join(LIB, "libbxstep.so"))


def _compile(src, out, extra=()):
    cmd = ["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-I", INC, "-I", os.path.join(REPO, "include"), src,
           "-o", out, "-L", LIB, "-lbounce_amd", "-lbxstep", "-Wl,-rpath," + LIB] + list(extra)
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.skipif(not os.path.exists(REF_HELLO), reason="reference tree not present")
def test_reference_hello_world_builds_unchanged(tmp_path):
    if not _have_libs():
        pytest.skip("libraries not built")
    _compile(REF_HELLO, str(tmp_path / "hello"))


def test_box_hull_table_matches(tmp_path):
    if not _have_libs():
        pytest.skip("libraries not built")
    src = tmp_path / "box.cpp"
    src.write_text("""
#include <bounce/bounce.h>
#include <stdio.h>
int main() {
    float ext[5][3] = {{0.5f, 0.5f, 0.5f}, {50.0f, 0.5f, 50.0f}, {10.0f, 1.0f, 10.0f}, {1.0f, 1.0f, 1.0f}, {0.3f, 1.7f, 2.9f}};
    for (int k = 0; k < 5; ++k) {
        b3BoxHull h(ext[k][0], ext[k][1], ext[k][2]);
        fwrite(h.vertices, sizeof(b3Vec3), h.vertexCount, stdout);
        fwrite(h.edges, sizeof(b3HalfEdge), h.edgeCount, stdout);
        fwrite(h.faces, sizeof(b3Face), h.faceCount, stdout);
        fwrite(h.planes, sizeof(b3Plane), h.faceCount, stdout);
    }
    return 0;
}
""")
    exe = str(tmp_path / "box")
    _compile(str(src), exe)
    raw = subprocess.run([exe], capture_output=True, check=True).stdout
    per = 8 * 12 + 24 * 20 + 6 * 4 + 6 * 16
    assert len(raw) == 5 * per
    for k, ext in enumerate([(0.5, 0.5, 0.5), (50.0, 0.5, 50.0), (10.0, 1.0, 10.0), (1.0, 1.0, 1.0), (0.3, 1.7, 2.9)]):
        chunk = raw[k * per:(k + 1) * per]
        h = hulls.box_hull(*ext)
        expect = h.vertices.tobytes() + h.edges.tobytes() + h.faces.tobytes() + h.planes.tobytes()
        assert chunk == expect, ext


@pytest.mark.gpu
def test_hello_world_program_on_device():
    """The drop-in hello world prints, on every step, exactly the oracle's pose of the box."""
    from oracle.oracle import CpuWorld
    assert os.path.exists(HELLO_BIN), "bx_hello_world is built by bounce_amd/build.py"
    print("[drop-in] bx_hello_world", flush=True)
    out = subprocess.run([HELLO_BIN], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    lines = [l for l in out.stdout.splitlines() if l.startswith("step ")]
    assert len(lines) == 60
    w = CpuWorld("oracle", scenes.hello_world())
    for i in range(60):
        w.step()
        b = w.bodies()[1]
        vals = tuple(float(x) for x in b["position"]) + tuple(float(x) for x in b["orientation"])
        expect = "step %d position = %.9g %.9g %.9g orientation = %.9g %.9g %.9g %.9g" % ((i + 1,) + vals)
        assert lines[i] == expect
    printed = [l for l in out.stdout.splitlines() if l.startswith("position = ")]
    assert printed[-1] == "position = 0.00 5.02 0.00"  # what the reference's hello_world prints at step 60


def mesh_program_lines(exe):
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return [l for l in out.stdout.splitlines() if l.startswith("step ")]


def expected_mesh_lines():
    from oracle.oracle import CpuWorld
    w = CpuWorld("oracle", scenes.mesh_ground_program())
    lines = []
    for i in range(120):
        w.step()
        b = w.bodies()
        for k in (1, 2, 3):
            vals = tuple(float(x) for x in b[k]["position"]) + tuple(float(x) for x in b[k]["orientation"])
            lines.append("step %d body %d position = %.9g %.9g %.9g orientation = %.9g %.9g %.9g %.9g"
                         % ((i + 1, k) + vals))
    return lines


@pytest.mark.gpu
def test_mesh_program_on_device():
    """The drop-in mesh program prints, on every step, exactly the oracle's poses of its bodies."""
    assert os.path.exists(MESH_BIN), "bx_mesh_ground is built by bounce_amd/build.py"
    print("[drop-in] bx_mesh_ground", flush=True)
    got = mesh_program_lines(MESH_BIN)
    want = expected_mesh_lines()
    assert len(got) == len(want) == 360
    for g, w in zip(got, want):
        assert g == w


def _program_lines(exe, prefix="step "):
    out = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return out.stdout.splitlines()


def expected_hello_lines():
    from oracle.oracle import CpuWorld
    w = CpuWorld("oracle", scenes.hello_world())
    lines = []
    for i in range(60):
        w.step()
        b = w.bodies()[1]
        vals = tuple(float(x) for x in b["position"]) + tuple(float(x) for x in b["orientation"])
        lines.append("step %d position = %.9g %.9g %.9g orientation = %.9g %.9g %.9g %.9g" % ((i + 1,) + vals))
    return lines


def _compare(got, want):
    assert len(got) == len(want), (len(got), len(want))
    for i, (g, w) in enumerate(zip(got, want)):
        assert g == w, f"line {i}:\n got  {g}\n want {w}"


@pytest.mark.skipif(not shutil.which("make") or not os.path.exists("/opt/rocm/llvm/bin/clang++"),
                    reason="the host build of the device library needs the ROCm clang++")
@pytest.mark.parametrize("program", ["hello_world", "mesh_ground", "spawn_apply"])
def test_drop_in_programs_on_the_host_build(tmp_path, program):
    """The drop-in programs and the facade, linked with the host build of the kernels and the C ABI."""
    subprocess.run(["make", "-C", CPU_BUILD, "all"], check=True, capture_output=True)
    exe = str(tmp_path / program)
    cmd = ["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-fno-fast-math", "-I", INC, "-I",
           os.path.join(REPO, "include"), os.path.join(REPO, "tests", "drop_in", program + ".cpp"),
           os.path.join(REPO, "bounce_amd", "facade", "world.cpp"), "-o", exe, CPU_LIB,
           "-Wl,-rpath," + os.path.dirname(CPU_LIB)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = _program_lines(exe)
    if program == "hello_world":
        _compare([l for l in lines if l.startswith("step ")], expected_hello_lines())
    elif program == "mesh_ground":
        _compare([l for l in lines if l.startswith("step ")], expected_mesh_lines())
    else:
        from tests.drop_in_replica import spawn_apply_lines
        _compare(lines, spawn_apply_lines())


def test_spawn_apply_replay_oracle_equals_reference():
    """The oracle's replay of the program that changes its world while it runs (creations and destroys
    between steps, forces every step, transforms, fixed rotation, gravity, sleeping, contact lists and
    queries) prints what the reference build prints."""
    from oracle import oracle as O
    if not O.available("ref32"):
        pytest.skip("reference build (make -C oracle ref) not present")
    from tests.drop_in_replica import spawn_apply_lines
    _compare(spawn_apply_lines("oracle"), spawn_apply_lines("ref32"))


@pytest.mark.gpu
def test_spawn_apply_program_on_device():
    """A world changed while it runs, through the drop-in API on the device: every body on every step, the
    contact list every ten steps and the final queries equal the oracle's replay of the same calls."""
    from tests.drop_in_replica import spawn_apply_lines
    assert os.path.exists(SPAWN_BIN), "bx_spawn_apply is built by bounce_amd/build.py"
    print("[drop-in] bx_spawn_apply", flush=True)
    _compare(_program_lines(SPAWN_BIN), spawn_apply_lines())
