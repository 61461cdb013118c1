"""Python binding of libbxstep.so (the HIP step), used by tests, the bench and the smoke check.

The library is loaded from bounce_amd/lib; if it is missing, loading raises: there is no
fallback path.
"""
import ctypes as C
import os
import weakref

import numpy as np

from . import capi

_lib = None


class BxError(RuntimeError):
    pass


def load(path=capi.LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise BxError(f"libbxstep.so not built at {path}; run bounce_amd/build.py")
    lib = C.CDLL(path)
    P = C.c_void_p
    u = C.c_uint32
    sig = {
        "bx_version": (C.c_char_p, []),
        "bx_last_error": (C.c_char_p, []),
        "bx_context_create": (C.c_int, [C.c_int, C.POINTER(P)]),
        "bx_context_destroy": (None, [P]),
        "bx_batch_create": (C.c_int, [P, C.POINTER(capi.bx_shape_library), C.POINTER(capi.bx_world_desc), u,
                                      C.POINTER(capi.bx_batch_opts), C.POINTER(P)]),
        "bx_batch_destroy": (None, [P]),
        "bx_batch_step": (C.c_int, [P, C.POINTER(capi.bx_step_cfg), u]),
        "bx_batch_step_graph": (C.c_int, [P, C.POINTER(capi.bx_step_cfg), u]),
        "bx_batch_step_noalloc": (C.c_int, [P, C.POINTER(capi.bx_step_cfg), u]),
        "bx_batch_get_solver_layout": (C.c_int, [P, C.POINTER(u), C.POINTER(C.c_float), u]),
        "bx_batch_step_subset": (C.c_int, [P, C.POINTER(capi.bx_step_cfg), C.POINTER(u), C.POINTER(u), u]),
        "bx_batch_sync": (C.c_int, [P]),
        "bx_batch_get_counts": (C.c_int, [P, u, C.POINTER(capi.bx_world_counts)]),
        "bx_batch_get_totals": (C.c_int, [P, C.POINTER(capi.bx_batch_totals)]),
        "bx_batch_get_events": (C.c_int, [P, u, C.POINTER(capi.bx_contact_event), u, C.POINTER(u)]),
        "bx_batch_set_contact_filter": (C.c_int, [P, u, C.c_void_p, C.c_void_p]),
        "bx_batch_get_bodies": (C.c_int, [P, u, C.POINTER(capi.bx_body_state), u]),
        "bx_batch_get_contacts": (C.c_int, [P, u, C.POINTER(capi.bx_contact_state), u]),
        "bx_batch_get_pairs": (C.c_int, [P, u, C.POINTER(capi.bx_pair), u]),
        "bx_batch_get_island_contacts": (C.c_int, [P, u, C.POINTER(u), u, C.POINTER(u)]),
        "bx_batch_set_bodies": (C.c_int, [P, u, C.POINTER(capi.bx_body_state), u]),
        "bx_batch_edit_bodies": (C.c_int, [P, u, C.POINTER(capi.bx_body_edit), u]),
        "bx_batch_get_timings": (C.c_int, [P, C.POINTER(C.c_float), u]),
        "bx_batch_get_kernel_timings": (C.c_int, [P, C.POINTER(C.c_float), u]),
        "bx_kernel_name": (C.c_char_p, [u]),
        "bx_struct_size": (u, [C.c_char_p]),
        "bx_batch_device_bytes": (C.c_uint64, [P]),
        "bx_batch_step_list": (C.c_int, [P, C.POINTER(capi.bx_step_cfg), C.POINTER(u), u]),
        "bx_pool_create": (C.c_int, [P, C.POINTER(capi.bx_shape_library), C.POINTER(capi.bx_pool_opts), C.POINTER(P)]),
        "bx_pool_admit": (C.c_int, [P, C.POINTER(capi.bx_world_desc), C.POINTER(u)]),
        "bx_pool_release": (C.c_int, [P, u]),
        "bx_pool_world_bytes": (C.c_uint64, [P, C.POINTER(capi.bx_world_desc)]),
        "bx_pool_get_stats": (C.c_int, [P, C.POINTER(capi.bx_pool_stats)]),
        "bx_batch_snapshot": (C.c_int, [P, u, P, C.c_uint64, C.POINTER(C.c_uint64)]),
        "bx_pool_restore": (C.c_int, [P, P, C.c_uint64, C.POINTER(u)]),
        "bx_batch_edit_topology": (C.c_int, [P, u, C.POINTER(capi.bx_shape_library), C.POINTER(capi.bx_world_desc),
                                             C.POINTER(capi.bx_topology_op), u]),
        "bx_batch_raycast_single": (C.c_int, [P, C.POINTER(capi.bx_ray_query), u, C.c_void_p, C.c_uint64,
                                              C.POINTER(capi.bx_ray_hit)]),
        "bx_batch_raycast": (C.c_int, [P, C.POINTER(capi.bx_ray_query), u, C.c_void_p, C.c_uint64,
                                       C.POINTER(capi.bx_ray_hit), u, C.POINTER(u)]),
        "bx_batch_query_aabb": (C.c_int, [P, C.POINTER(capi.bx_aabb_query), u, C.c_void_p, C.c_uint64,
                                          C.POINTER(u), u, C.POINTER(u)]),
        "bx_batch_time_of_impact": (C.c_int, [P, C.POINTER(capi.bx_toi_query), u, C.POINTER(capi.bx_toi_result)]),
        "bx_batch_shape_cast_single": (C.c_int, [P, C.POINTER(capi.bx_shape_cast_query), u, C.c_void_p, C.c_uint64,
                                                 C.POINTER(capi.bx_shape_hit)]),
        "bx_batch_shape_cast": (C.c_int, [P, C.POINTER(capi.bx_shape_cast_query), u, C.c_void_p, C.c_uint64,
                                          C.POINTER(capi.bx_shape_hit), u, C.POINTER(u)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc < 0:
        raise BxError(f"bxstep error {rc}: {_lib.bx_last_error().decode()}")
    return rc


class Context:
    def __init__(self, device=0):
        lib = load()
        h = C.c_void_p()
        _check(lib.bx_context_create(device, C.byref(h)))
        self.h = h
        self._batches = weakref.WeakSet()

    def close(self):
        """Destroy the context; batches still alive on it are destroyed first."""
        if self.h:
            for b in list(self._batches):
                b.close()
            _lib.bx_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BODY_DT = np.dtype(capi.bx_body_state)
CONTACT_DT = np.dtype(capi.bx_contact_state)
PAIR_DT = np.dtype(capi.bx_pair)


class Batch:
    """A batch of worlds resident on the device."""

    def __init__(self, ctx, scene, max_contacts=0, max_pairs=0, solver_mode=0, mesh_triangles=0):
        self.ctx = ctx
        self.scene = scene
        self._lib = scene.library()
        self._descs = scene.world_descs()
        opts = capi.bx_batch_opts()
        opts.max_contacts_per_world = max_contacts
        opts.max_pairs_per_world = max_pairs
        opts.solver_mode = solver_mode
        opts.mesh_triangles_per_contact = mesh_triangles
        h = C.c_void_p()
        _check(_lib.bx_batch_create(ctx.h, C.byref(self._lib), self._descs, len(scene.worlds), C.byref(opts),
                                    C.byref(h)))
        self.h = h
        self.world_count = len(scene.worlds)
        ctx._batches.add(self)

    def close(self):
        if self.h and self.ctx.h:
            _lib.bx_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def status(self, world=0):
        return self.counts(world).status

    def step(self, dt=1.0 / 60.0, vit=8, pit=2, n=1):
        cfg = capi.bx_step_cfg(C.c_float(dt), vit, pit)
        _check(_lib.bx_batch_step(self.h, C.byref(cfg), n))

    def step_graph(self, dt=1.0 / 60.0, vit=8, pit=2, n=1):
        """n steps, each a replay of one captured step (same results as step; untimed)."""
        cfg = capi.bx_step_cfg(C.c_float(dt), vit, pit)
        _check(_lib.bx_batch_step_graph(self.h, C.byref(cfg), n))

    def edit_bodies(self, world, rows):
        """Body edits between steps: (body, op, v, p, wake) tuples, op a capi.BX_EDIT_* value."""
        e = capi.body_edits(rows)
        _check(_lib.bx_batch_edit_bodies(self.h, world, e, len(rows)))

    def step_noalloc(self, dt=1.0 / 60.0, vit=8, pit=2, n=1):
        """n steps through bx_batch_step_noalloc (same results as step; allocates nothing, untimed)."""
        cfg = capi.bx_step_cfg(C.c_float(dt), vit, pit)
        _check(_lib.bx_batch_step_noalloc(self.h, C.byref(cfg), n))

    def step_list(self, worlds, dt=1.0 / 60.0, vit=8, pit=2):
        """One step of each listed world, all in one launch sequence."""
        cfg = capi.bx_step_cfg(C.c_float(dt), vit, pit)
        w = np.ascontiguousarray(worlds, dtype=np.uint32)
        _check(_lib.bx_batch_step_list(self.h, C.byref(cfg), w.ctypes.data_as(C.POINTER(C.c_uint32)), len(w)))

    def snapshot(self, world):
        """The world's whole device state as bytes (bx_pool_restore resumes it in a pool)."""
        size = C.c_uint64(0)
        _check(_lib.bx_batch_snapshot(self.h, world, None, 0, C.byref(size)))
        buf = C.create_string_buffer(size.value)
        _check(_lib.bx_batch_snapshot(self.h, world, buf, size.value, C.byref(size)))
        return buf.raw

    def step_subset(self, worlds, counts, dt=1.0 / 60.0, vit=8, pit=2):
        cfg = capi.bx_step_cfg(C.c_float(dt), vit, pit)
        w = np.ascontiguousarray(worlds, dtype=np.uint32)
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        _check(_lib.bx_batch_step_subset(self.h, C.byref(cfg), w.ctypes.data_as(C.POINTER(C.c_uint32)),
                                         c.ctypes.data_as(C.POINTER(C.c_uint32)), len(w)))

    def sync(self):
        _check(_lib.bx_batch_sync(self.h))

    def counts(self, world=0):
        c = capi.bx_world_counts()
        _check(_lib.bx_batch_get_counts(self.h, world, C.byref(c)))
        return c

    def events(self, world=0):
        """Listener events of the world's last step (BX_WORLD_EVENTS): (type, fixture_a, fixture_b, contact)."""
        cnt = C.c_uint32(0)
        _check(_lib.bx_batch_get_events(self.h, world, None, 0, C.byref(cnt)))
        out = np.zeros(cnt.value, dtype=np.dtype(capi.bx_contact_event))
        _check(_lib.bx_batch_get_events(self.h, world, out.ctypes.data_as(C.POINTER(capi.bx_contact_event)),
                                        cnt.value, C.byref(cnt)))
        return out

    def set_contact_filter(self, world, collide, respond):
        """Filter tables (uint8 arrays of n*n bits, bx_batch_set_contact_filter) or None for no filter."""
        _check(_lib.bx_batch_set_contact_filter(self.h, world, None if collide is None else collide.ctypes.data,
                                                None if respond is None else respond.ctypes.data))

    def totals(self):
        """bx_batch_totals over every resident world (one read of world states and bodies)."""
        t = capi.bx_batch_totals()
        _check(_lib.bx_batch_get_totals(self.h, C.byref(t)))
        return t

    def _body_count(self, world):
        return self.counts(world).body_count

    def edit_topology(self, world, ops, adds, library_scene=None):
        """Topology operations ((op, index) pairs, capi.BX_TOPO_*) on one world, descriptors from the
        scenes.World `adds`, an optionally grown shape library from `library_scene`."""
        lib = library_scene.library() if library_scene is not None else None
        desc = adds.desc()
        _check(_lib.bx_batch_edit_topology(self.h, world, None if lib is None else C.byref(lib), C.byref(desc),
                                           capi.topology_ops(ops), len(ops)))
        if library_scene is not None:
            self._lib_keep = (lib, library_scene)

    def bodies(self, world=0):
        n = self._body_count(world)
        out = np.zeros(n, dtype=BODY_DT)
        _check(_lib.bx_batch_get_bodies(self.h, world, out.ctypes.data_as(C.POINTER(capi.bx_body_state)), n))
        return out

    def contacts(self, world=0):
        n = self.counts(world).contact_count
        out = np.zeros(n, dtype=CONTACT_DT)
        _check(_lib.bx_batch_get_contacts(self.h, world, out.ctypes.data_as(C.POINTER(capi.bx_contact_state)), n))
        return out

    def pairs(self, world=0):
        n = self.counts(world).pair_count
        out = np.zeros(n, dtype=PAIR_DT)
        _check(_lib.bx_batch_get_pairs(self.h, world, out.ctypes.data_as(C.POINTER(capi.bx_pair)), n))
        return out

    def island_contacts(self, world=0):
        cap = max(1, self.counts(world).contact_count)
        out = np.zeros(cap, dtype=np.uint32)
        cnt = C.c_uint32(0)
        _check(_lib.bx_batch_get_island_contacts(self.h, world, out.ctypes.data_as(C.POINTER(C.c_uint32)), cap,
                                                 C.byref(cnt)))
        return out[: cnt.value]

    def timings(self):
        t = (C.c_float * 5)()
        _check(_lib.bx_batch_get_timings(self.h, t, 5))
        return list(t)

    def kernel_timings(self):
        """{kernel name: ms} of the last step of the last step() call (HIP events on the stream)."""
        t = (C.c_float * 11)()
        n = _check(_lib.bx_batch_get_kernel_timings(self.h, t, 11))
        return {_lib.bx_kernel_name(i).decode(): float(t[i]) for i in range(n)}

    def solver_layout(self):
        """(islands per wavefront of the velocity sweeps, [k_solve_velocity ms timed per layout])."""
        lanes = C.c_uint32(0)
        t = (C.c_float * 8)()
        n = _check(_lib.bx_batch_get_solver_layout(self.h, C.byref(lanes), t, 8))
        return lanes.value, [float(t[i]) for i in range(n)]

    def device_bytes(self):
        return int(_lib.bx_batch_device_bytes(self.h))

    @staticmethod
    def _queries(kind, rows):
        """bx_ray_query / bx_aabb_query records from (world, a[3], b[3], mask_offset) rows."""
        out = np.zeros(len(rows), dtype=np.dtype(kind))
        a_name, b_name = out.dtype.names[2], out.dtype.names[3]
        for i, r in enumerate(rows):
            out["world"][i] = r[0]
            out["mask"][i] = capi.BX_NULL_INDEX if len(r) < 4 or r[3] is None else r[3]
            out[a_name][i] = np.asarray(r[1], dtype=np.float32)
            out[b_name][i] = np.asarray(r[2], dtype=np.float32)
        return out

    @staticmethod
    def _mask_args(mask):
        if mask is None:
            return None, 0, None
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        return m.ctypes.data, m.size, m

    def raycast_single(self, rays, mask=None):
        """b3World::RayCastSingle for each (world, p1, p2[, mask offset]) row: a bx_ray_hit array."""
        q = self._queries(capi.bx_ray_query, rays)
        out = np.zeros(len(rays), dtype=np.dtype(capi.bx_ray_hit))
        mp, ms, _keep = self._mask_args(mask)
        _check(_lib.bx_batch_raycast_single(self.h, q.ctypes.data_as(C.POINTER(capi.bx_ray_query)), len(rays), mp, ms,
                                            out.ctypes.data_as(C.POINTER(capi.bx_ray_hit))))
        return out

    def raycast(self, rays, per_ray=16, mask=None):
        """b3World::RayCast (a listener returning 1) for each row: (hits [n, per_ray], counts [n])."""
        q = self._queries(capi.bx_ray_query, rays)
        hits = np.zeros((len(rays), max(1, per_ray)), dtype=np.dtype(capi.bx_ray_hit))
        counts = np.zeros(len(rays), dtype=np.uint32)
        mp, ms, _keep = self._mask_args(mask)
        _check(_lib.bx_batch_raycast(self.h, q.ctypes.data_as(C.POINTER(capi.bx_ray_query)), len(rays), mp, ms,
                                     hits.ctypes.data_as(C.POINTER(capi.bx_ray_hit)), per_ray,
                                     counts.ctypes.data_as(C.POINTER(C.c_uint32))))
        return hits[:, :per_ray], counts

    def query_aabb(self, boxes, per_query=64, mask=None):
        """b3World::QueryAABB for each (world, lower, upper[, mask offset]) row: (fixtures, counts)."""
        q = self._queries(capi.bx_aabb_query, boxes)
        fx = np.zeros((len(boxes), max(1, per_query)), dtype=np.uint32)
        counts = np.zeros(len(boxes), dtype=np.uint32)
        mp, ms, _keep = self._mask_args(mask)
        _check(_lib.bx_batch_query_aabb(self.h, q.ctypes.data_as(C.POINTER(capi.bx_aabb_query)), len(boxes), mp, ms,
                                        fx.ctypes.data_as(C.POINTER(C.c_uint32)), per_query,
                                        counts.ctypes.data_as(C.POINTER(C.c_uint32))))
        return fx[:, :per_query], counts

    def time_of_impact(self, queries):
        """b3TimeOfImpact for each bx_toi_query of a ctypes array (hulls from this batch's library): a
        structured bx_toi_result array."""
        out = (capi.bx_toi_result * max(1, len(queries)))()
        _check(_lib.bx_batch_time_of_impact(self.h, queries, len(queries), out))
        return np.ctypeslib.as_array(out)[: len(queries)].copy()

    def shape_cast_single(self, casts, mask=None):
        """b3World::ShapeCastSingle for each bx_shape_cast_query of a ctypes array: a bx_shape_hit array."""
        out = np.zeros(len(casts), dtype=np.dtype(capi.bx_shape_hit))
        mp, ms, _keep = self._mask_args(mask)
        _check(_lib.bx_batch_shape_cast_single(self.h, casts, len(casts), mp, ms,
                                               out.ctypes.data_as(C.POINTER(capi.bx_shape_hit))))
        return out

    def shape_cast(self, casts, per_cast=16, mask=None):
        """b3World::ShapeCast (a listener returning 1) for each cast: (hits [n, per_cast], counts [n])."""
        hits = np.zeros((len(casts), max(1, per_cast)), dtype=np.dtype(capi.bx_shape_hit))
        counts = np.zeros(len(casts), dtype=np.uint32)
        mp, ms, _keep = self._mask_args(mask)
        _check(_lib.bx_batch_shape_cast(self.h, casts, len(casts), mp, ms,
                                        hits.ctypes.data_as(C.POINTER(capi.bx_shape_hit)), per_cast,
                                        counts.ctypes.data_as(C.POINTER(C.c_uint32))))
        return hits[:, :per_cast], counts


class PoolFull(BxError):
    """The pool has no free slot or not enough free memory for the world right now."""


class Pool(Batch):
    """Resident worlds in a fixed device-memory budget (bx_pool_*): worlds are admitted into slots
    while they fit, stepped in any mix with step_list, read like a batch's worlds and released."""

    def __init__(self, ctx, scene, budget_bytes, max_worlds=0, max_contacts=0, max_pairs=0, solver_mode=0):
        self.ctx = ctx
        self.scene = scene  # shape library; its worlds are not admitted by construction
        self._lib = scene.library()
        opts = capi.bx_pool_opts()
        opts.hbm_budget_bytes = int(budget_bytes)
        opts.max_worlds = max_worlds
        opts.max_contacts_per_world = max_contacts
        opts.max_pairs_per_world = max_pairs
        opts.solver_mode = solver_mode
        h = C.c_void_p()
        _check(_lib.bx_pool_create(ctx.h, C.byref(self._lib), C.byref(opts), C.byref(h)))
        self.h = h
        self.slots = {}  # slot -> World
        ctx._batches.add(self)

    def _body_count(self, slot):
        return self.slots[slot].body_count

    def world_bytes(self, world):
        d = world.desc()
        return int(_lib.bx_pool_world_bytes(self.h, C.byref(d)))

    def admit(self, world):
        """Slot of the admitted world; PoolFull when it does not fit now."""
        d = world.desc()
        slot = C.c_uint32(0)
        rc = _lib.bx_pool_admit(self.h, C.byref(d), C.byref(slot))
        if rc == capi.BX_E_CAPACITY:
            raise PoolFull(_lib.bx_last_error().decode())
        _check(rc)
        self.slots[slot.value] = world
        return slot.value

    def release(self, slot):
        _check(_lib.bx_pool_release(self.h, slot))
        del self.slots[slot]

    def restore(self, world, image):
        """Admit a snapshot taken of `world`; PoolFull when it does not fit now."""
        slot = C.c_uint32(0)
        buf = C.create_string_buffer(image, len(image))
        rc = _lib.bx_pool_restore(self.h, buf, len(image), C.byref(slot))
        if rc == capi.BX_E_CAPACITY:
            raise PoolFull(_lib.bx_last_error().decode())
        _check(rc)
        self.slots[slot.value] = world
        return slot.value

    def stats(self):
        st = capi.bx_pool_stats()
        _check(_lib.bx_pool_get_stats(self.h, C.byref(st)))
        return st
