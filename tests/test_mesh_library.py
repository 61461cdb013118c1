"""The mesh structures the device path builds at upload, against the oracle (CPU only).

libbxstep builds every mesh's static tree and edge wings on the host when a batch is created
(bx_mesh_query_aabbs and bx_mesh_wings expose them without a device). The oracle restates
b3Mesh::BuildTree / b3StaticTree::QueryAABB and the reference's O(n^2) b3Mesh::BuildAdjacency;
the library must give the same triangle order for every query box (the order a mesh contact
visits triangles in) and the same wings, also on meshes far larger than the GPU parity scenes and
on tie-heavy flat grids (where std::sort's handling of equal keys fixes the tree). The candidate set
of each query is also checked by brute force: exactly the triangles whose slop-extended AABB
overlaps the box.
"""
import ctypes as C

import numpy as np
import pytest

from bounce_amd import capi, scenes

F32 = np.float32
u32p = C.POINTER(C.c_uint32)
f32p = C.POINTER(C.c_float)


def _libs():
    from oracle import oracle as O
    bx = C.CDLL(capi.LIB_PATH)
    bx.bx_mesh_query_aabbs.argtypes = [C.POINTER(capi.bx_mesh_desc), f32p, C.c_uint32, u32p, u32p, C.c_uint32]
    bx.bx_mesh_query_aabbs.restype = C.c_int
    bx.bx_mesh_wings.argtypes = [C.POINTER(capi.bx_mesh_desc), u32p]
    bx.bx_mesh_wings.restype = C.c_int
    orc = C.CDLL(O.LIBS["oracle"][0])
    orc.orc_mesh_query.argtypes = [C.POINTER(capi.bx_mesh_desc), f32p, C.c_uint32, u32p, u32p, C.c_uint32]
    orc.orc_mesh_query.restype = C.c_int
    orc.orc_mesh_wings.argtypes = [C.POINTER(capi.bx_mesh_desc), u32p]
    orc.orc_mesh_wings.restype = C.c_int
    return bx, orc


def _desc(v, t):
    v = np.ascontiguousarray(v, dtype=F32)
    t = np.ascontiguousarray(t, dtype=np.uint32)
    d = capi.bx_mesh_desc(len(v), len(t), v.ctypes.data_as(f32p), t.ctypes.data_as(u32p))
    return d, (v, t)


def _meshes():
    rng = np.random.default_rng(5)
    out = {"bumpy64": scenes.grid_mesh(rng, cells=64, size=64.0, bumps=0.4),
           "flat64": scenes.grid_mesh(rng, cells=64, size=64.0, bumps=0.0),
           "flat_odd37": scenes.grid_mesh(rng, cells=37, size=20.0, bumps=0.0)}
    # a non-manifold fan: three triangles on one edge, plus a duplicated triangle
    v = np.array([(0, 0, 0), (1, 0, 0), (0.5, 1, 0), (0.5, -1, 0), (0.5, 0, 1), (2, 0, 0)], dtype=F32)
    t = np.array([(0, 1, 2), (1, 0, 3), (0, 1, 4), (1, 5, 2), (0, 1, 2)], dtype=np.uint32)
    out["nonmanifold"] = (v, t)
    return out


def _boxes(v, rng, n):
    lo, hi = v.min(axis=0), v.max(axis=0)
    c = rng.uniform(lo, hi, (n, 3)).astype(F32)
    h = rng.uniform(0.05, 3.0, (n, 3)).astype(F32)
    return np.concatenate([c - h, c + h], axis=1).astype(F32)


@pytest.mark.parametrize("name", list(_meshes()))
def test_tree_query_and_wings_match_oracle(name):
    bx, orc = _libs()
    v, t = _meshes()[name]
    d, keep = _desc(v, t)
    rng = np.random.default_rng(11)
    boxes = np.ascontiguousarray(_boxes(v, rng, 400))
    cap = 4096
    res = {}
    for lib, fn in ((bx, bx.bx_mesh_query_aabbs), (orc, orc.orc_mesh_query)):
        counts = np.zeros(len(boxes), dtype=np.uint32)
        out = np.zeros(len(boxes) * cap, dtype=np.uint32)
        rc = fn(C.byref(d), boxes.ctypes.data_as(f32p), len(boxes), counts.ctypes.data_as(u32p),
                out.ctypes.data_as(u32p), cap)
        assert rc == 0
        res[fn.__name__] = (counts, out.reshape(len(boxes), cap))
    (cb, ob), (co, oo) = res["bx_mesh_query_aabbs"], res["orc_mesh_query"]
    assert np.array_equal(cb, co)
    assert cb.max() < cap
    for k in range(len(boxes)):
        assert np.array_equal(ob[k, :cb[k]], oo[k, :co[k]]), k
    # brute force: the candidate set is every triangle whose slop-extended AABB overlaps the box
    tri = v[t]
    tlo = tri.min(axis=1) - F32(0.005)
    thi = tri.max(axis=1) + F32(0.005)
    for k in range(0, len(boxes), 7):
        blo, bhi = boxes[k, :3], boxes[k, 3:]
        want = np.flatnonzero(np.all((tlo <= bhi) & (blo <= thi), axis=1))
        assert np.array_equal(np.sort(ob[k, :cb[k]]), want), k
    wb = np.zeros(3 * len(t), dtype=np.uint32)
    wo = np.zeros(3 * len(t), dtype=np.uint32)
    assert bx.bx_mesh_wings(C.byref(d), wb.ctypes.data_as(u32p)) == 0
    assert orc.orc_mesh_wings(C.byref(d), wo.ctypes.data_as(u32p)) == 0
    assert np.array_equal(wb, wo)
