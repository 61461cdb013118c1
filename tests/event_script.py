"""Worlds for the contact listener and filter tests: spheres, capsules and boxes dropped on a ground box
with sleeping and event recording on, and a contact filter whose answers depend only on the two fixtures
(b3ContactFilter): fixtures i and j with i % 3 == 0 and j % 3 == 1 (either order) never collide, and
touching contacts between two odd fixtures get no response (they touch, are reported, and are not
solved)."""
import numpy as np

from bounce_amd import capi, scenes

STEPS = 150


def scene():
    return scenes.with_flags(scenes.mixed_drop(36, n_worlds=2, seed=41, extent=(5.0, 8.0, 5.0)),
                             capi.BX_WORLD_SLEEPING | capi.BX_WORLD_EVENTS)


def filter_tables(n):
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    collide = ~(((i % 3 == 0) & (j % 3 == 1)) | ((i % 3 == 1) & (j % 3 == 0)))
    respond = ~((i % 2 == 1) & (j % 2 == 1))
    pack = lambda t: np.packbits(t.reshape(-1).astype(np.uint8), bitorder="little")
    return pack(collide), pack(respond)
