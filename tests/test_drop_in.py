"""The drop-in C++ API (bounce_amd/include/bounce/bounce.h over libbxstep).

CPU: programs written against the reference's API compile and link against the drop-in headers
(the reference's own examples/hello_world/main.cpp when the reference tree is present), and the
drop-in b3BoxHull builds the reference's box table bit for bit. GPU: a hello-world program and a
program dropping a sphere, a capsule and a box on a b3Mesh heightfield, built against the drop-in
API, step on the device and print the oracle's poses on every step.
"""
import os
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
