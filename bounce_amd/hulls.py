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


def convex_hull(points, eps=1e-9):
    """A b3Hull for the convex hull of a small point cloud (the role of the testbed's quickhull).

    Faces are the hull's triangles, found by testing every triple of points against all others
    (exact for the <= 16-point clouds of the mixed-shape workloads; points in general position).
    Every face loop runs counter-clockwise seen from outside, so a face's normal is
    cross(e1, e2) of its first two edges; the two half-edges of an edge take indices 2k and 2k+1;
    planes are (unit outward normal, offset) and the centroid is the solid's centroid, all computed
    in binary64 and stored in binary32. Only the vertices on the hull are kept.
    """
    p32 = np.ascontiguousarray(points, dtype=F32)
    p = p32.astype(np.float64)
    n = len(p)
    tri = np.array([(i, j, k) for i in range(n) for j in range(i + 1, n) for k in range(j + 1, n)])
    nrm = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]])
    d = np.einsum("tk,tmk->tm", nrm, p[None, :, :] - p[tri[:, 0]][:, None, :])
    rows = np.arange(len(tri))[:, None]
    d[rows, tri] = 0.0
    ok = np.linalg.norm(nrm, axis=1) >= eps
    below = ok & np.all(d <= eps, axis=1)
    above = ok & ~below & np.all(d >= -eps, axis=1)
    faces = [tuple(t) for t in tri[below]] + [(t[0], t[2], t[1]) for t in tri[above]]
    faces.sort()
    if len(faces) < 4:
        raise ValueError("degenerate point cloud")
    used = sorted({v for f in faces for v in f})
    remap = {v: r for r, v in enumerate(used)}
    verts = p32[used]
    tris = [tuple(remap[v] for v in f) for f in faces]
    # half-edges: twins adjacent
    directed = {}
    edges = []
    for (a, b, c) in tris:
        for (u, v) in ((a, b), (b, c), (c, a)):
            if (u, v) in directed:
                continue
            base = len(edges)
            directed[(u, v)] = base
            directed[(v, u)] = base + 1
            edges.append([u, base + 1, 0, 0, 0])
            edges.append([v, base, 0, 0, 0])
    face_first = []
    for fi, (a, b, c) in enumerate(tris):
        loop = [directed[(a, b)], directed[(b, c)], directed[(c, a)]]
        for m in range(3):
            e = loop[m]
            edges[e][2] = fi
            edges[e][3] = loop[(m + 2) % 3]
            edges[e][4] = loop[(m + 1) % 3]
        face_first.append(loop[0])
    pv = verts.astype(np.float64)
    planes = np.zeros((len(tris), 4), dtype=F32)
    volume = 0.0
    centroid = np.zeros(3)
    ref = pv.mean(axis=0)
    for fi, (a, b, c) in enumerate(tris):
        nrm = np.cross(pv[b] - pv[a], pv[c] - pv[a])
        nrm /= np.linalg.norm(nrm)
        planes[fi, :3] = nrm
        planes[fi, 3] = np.dot(nrm, pv[a])
        vol = np.dot(np.cross(pv[a] - ref, pv[b] - ref), pv[c] - ref) / 6.0
        volume += vol
        centroid += vol * (ref + pv[a] + pv[b] + pv[c]) / 4.0
    centroid /= volume
    return HullData(verts, np.array(edges, dtype=np.uint32), np.array(face_first, dtype=np.uint32), planes,
                    centroid.astype(F32))


def random_hull_library(count=1024, seed=7, min_points=8, max_points=16):
    """Hulls of 8-16 points drawn uniformly in the unit ball (seeded)."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < count:
        m = int(rng.integers(min_points, max_points + 1))
        d = rng.normal(size=(m, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        r = rng.uniform(size=(m, 1)) ** (1.0 / 3.0)
        try:
            out.append(convex_hull(d * r))
        except ValueError:
            continue
    return out
