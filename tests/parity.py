"""Comparison helpers shared by the parity tests: bit-exact state equality with readable diffs."""
import numpy as np


def masked_bodies(b):
    """Body states with the one field the reference leaves undefined masked.

    b3Body::m_islandIndex is never initialised (body.cpp:25-103) and is only meaningful for a
    body that was placed in an island this step; static bodies keep the index of the last island
    they joined, so it is compared only for non-static bodies flagged as in an island.
    """
    b = b.copy()
    keep = (b["flags"] & 2) != 0
    b["island_index"][~keep] = 0
    return b


def masked_contacts(c):
    """Contact states with the feature-cache fields the reference leaves undefined masked.

    A fresh b3HullContact sets only cacheType = e_empty (sat.h:69); featureType and the indices
    are written on the first b3FeatureCache::ReadState, so they are compared only when the
    cache is not empty.
    """
    c = c.copy()
    empty = c["cache"][:, 0] == 0
    c["cache"][empty, 1:] = 0
    return c


def first_diff(a, b, name):
    if a.shape != b.shape:
        return f"{name}: count {a.shape} != {b.shape}"
    if a.tobytes() == b.tobytes():
        return None
    av = a.view(np.uint8).reshape(len(a), -1)
    bv = b.view(np.uint8).reshape(len(b), -1)
    rows, cols = np.where(av != bv)
    i = int(rows[0])
    field = None
    for fname in a.dtype.names:
        off = a.dtype.fields[fname][1]
        size = a.dtype.fields[fname][0].itemsize
        if off <= cols[0] < off + size:
            field = fname
            break
    return f"{name}[{i}].{field}: {a[i][field]} != {b[i][field]} ({len(set(rows.tolist()))} rows differ)"


def compare_worlds(x, y, mask_cache=False):
    """List of differences between two stepped worlds (objects exposing bodies/contacts/pairs)."""
    diffs = []
    d = first_diff(masked_bodies(x.bodies()), masked_bodies(y.bodies()), "body")
    if d:
        diffs.append(d)
    cx, cy = x.contacts(), y.contacts()
    if mask_cache:
        cx, cy = masked_contacts(cx), masked_contacts(cy)
    d = first_diff(cx, cy, "contact")
    if d:
        diffs.append(d)
    d = first_diff(x.pairs(), y.pairs(), "pair")
    if d:
        diffs.append(d)
    ix, iy = x.island_contacts(), y.island_contacts()
    if ix.tobytes() != iy.tobytes():
        diffs.append(f"island order: {len(ix)} vs {len(iy)} entries differ")
    return diffs
