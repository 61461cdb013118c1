"""The revolute joint's limit angle on the device against the host's atan2.

The reference computes the angle as atan2 of two binary32 values promoted to double (glibc's
atan2) and stores it as binary32 (revolute_joint.cpp:135, 434). The device evaluates the same
expression with its own double atan2 (bx::JointAngle, run here through libbxcal.so); the binary32
results agree unless the two double results straddle a binary32 rounding boundary. This checks
4M seeded (sin, cos) pairs, covering every angle, the axes and the cosines of joints near their
limits, for equality.
"""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "bounce_amd", "lib", "libbxcal.so")


def test_joint_angle_matches_host_atan2():
    cal = C.CDLL(LIB)
    fp = C.POINTER(C.c_float)
    cal.bxcal_joint_angles.argtypes = [fp, fp, fp, C.c_uint64]
    cal.bxcal_joint_angles.restype = C.c_int
    rng = np.random.default_rng(17)
    n = 1 << 22
    theta = rng.uniform(-np.pi, np.pi, n)
    r = rng.uniform(0.999, 1.001, n)  # binary32 frame columns are unit only to rounding
    s = (r * np.sin(theta)).astype(np.float32)
    c = (r * np.cos(theta)).astype(np.float32)
    s[:8] = [0, 0, 1, -1, 0, -0.0, 1e-30, -1e-30]
    c[:8] = [1, -1, 0, 0, -1, -1, -1, -1]
    out = np.empty(n, dtype=np.float32)
    rc = cal.bxcal_joint_angles(s.ctypes.data_as(fp), c.ctypes.data_as(fp), out.ctypes.data_as(fp), n)
    assert rc == 0
    want = np.arctan2(s.astype(np.float64), c.astype(np.float64)).astype(np.float32)
    bad = np.flatnonzero(out.view(np.uint32) != want.view(np.uint32))
    print(f"[joint angle] {len(bad)} of {n} differ", flush=True)
    assert len(bad) == 0, (bad[:5], s[bad[:5]], c[bad[:5]], out[bad[:5]], want[bad[:5]])
