"""Convex hull data for the shape library (b3Hull layout: half-edges with adjacent twins).

box_hull(ex, ey, ez) builds the same data as the reference's b3BoxHull(ex, ey, ez)
(include/bounce/collision/geometry/box_hull.h:41-121 then b3Hull::Scale, hull.cpp:120-150):
the box topology table has to be identical because contact feature ids are half-edge indices and
SAT tie-breaking follows face order. The arithmetic of Scale is repeated in binary32, operation by
operation.
"""
import numpy as np

F32 = np.float32


class HullData:
    def __init__(self, vertices, edges, faces, planes, centroid):
        self.vertices = np.ascontiguousarray(vertices, dtype=F32)
        self.edges = np.ascontiguousarray(edges, dtype=np.uint32)
        self.faces = np.ascontiguousarray(faces, dtype=np.uint32)
        self.planes = np.ascontiguousarray(planes, dtype=F32)
        self.centroid = np.ascontiguousarray(centroid, dtype=F32)


# b3BoxHull::SetIdentity topology: 8 vertices, 24 half-edges (origin, twin, face, prev, next), 6 faces.
_BOX_VERTICES = [(-1, -1, 1), (1, -1, -1), (1, 1, 1), (-1, 1, -1), (1, 1, -1), (-1, 1, 1), (1, -1, 1), (-1, -1, -1)]
_BOX_EDGES = [
    (0, 1, 0, 6, 2), (5, 0, 1, 12, 8), (5, 3, 0, 0, 4), (3, 2, 5, 19, 13), (3, 5, 0, 2, 6), (7, 4, 3, 15, 18),
    (7, 7, 0, 4, 0), (0, 6, 2, 9, 14), (0, 9, 1, 1, 10), (6, 8, 2, 16, 7), (6, 11, 1, 8, 12), (2, 10, 4, 22, 17),
    (2, 13, 1, 10, 1), (5, 12, 5, 3, 23), (7, 15, 2, 7, 16), (1, 14, 3, 20, 5), (1, 17, 2, 14, 9), (6, 16, 4, 11, 21),
    (3, 19, 3, 5, 20), (4, 18, 5, 23, 3), (4, 21, 3, 18, 15), (1, 20, 4, 17, 22), (4, 23, 4, 21, 11), (2, 22, 5, 13, 19),
]
_BOX_FACES = [0, 8, 7, 15, 21, 23]
_BOX_PLANES = [(-1, 0, 0, 1), (0, 0, 1, 1), (0, -1, 0, 1), (0, 0, -1, 1), (1, 0, 0, 1), (0, 1, 0, 1)]


def _normalize3(n):
    """b3Vec3::Normalize in binary32: length, then multiply by the reciprocal."""
    x, y, z = F32(n[0]), F32(n[1]), F32(n[2])
    length = np.sqrt(F32(F32(F32(x * x) + F32(y * y)) + F32(z * z)), dtype=F32)
    if length > np.finfo(F32).eps:
        s = F32(F32(1.0) / length)
        x, y, z = F32(x * s), F32(y * s), F32(z * s)
    return (x, y, z), F32(length)


def box_hull(ex, ey, ez):
    scale = (F32(ex), F32(ey), F32(ez))
    vertices = np.array(_BOX_VERTICES, dtype=F32) * np.array(scale, dtype=F32)
    inv = tuple(F32(F32(1.0) / s) for s in scale)
    planes = np.zeros((6, 4), dtype=F32)
    for i, (nx, ny, nz, off) in enumerate(_BOX_PLANES):
        n = (F32(inv[0] * F32(nx)), F32(inv[1] * F32(ny)), F32(inv[2] * F32(nz)))
        n, length = _normalize3(n)
        planes[i, :3] = n
        planes[i, 3] = F32(F32(off) / length)
    return HullData(vertices, np.array(_BOX_EDGES, dtype=np.uint32), np.array(_BOX_FACES, dtype=np.uint32),
                    planes, np.zeros(3, dtype=F32))
