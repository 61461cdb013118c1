"""A scripted sequence of body edits between steps (bx_batch_edit_bodies / b3Body's setters and force and
impulse methods), shared by the oracle test against the reference and the device test against the
oracle. Worlds are spheres, capsules and boxes dropped on a ground box with sleeping on, so that edits
land on awake, sleeping, static and freshly woken bodies."""
import numpy as np

from bounce_amd import capi, scenes

STEPS = 160


def scene():
    return scenes.with_flags(scenes.concat([scenes.mixed_drop(40, n_worlds=2, seed=31, extent=(6.0, 8.0, 6.0)),
                                            scenes.drop_boxes(30, n_worlds=1, seed=32, extent=(6.0, 6.0, 6.0))]),
                             capi.BX_WORLD_SLEEPING)


def edits(step, rng_seed=5):
    """Edits to apply before `step` (0-based): (body, op, v, p, wake) tuples; body 0 is the static ground."""
    rng = np.random.default_rng(rng_seed * 1000 + step)
    E = capi
    rows = [(3, E.BX_EDIT_FORCE, (4.0, 30.0, -2.0), (0.1, 2.0, 0.3), 0)]  # every step, only while awake
    if step == 20:
        rows += [(0, E.BX_EDIT_LINEAR_VELOCITY, (1.0, 0.0, 0.0)),      # static: ignored
                 (5, E.BX_EDIT_LINEAR_VELOCITY, (0.0, 0.0, 0.0)),
                 (6, E.BX_EDIT_FORCE_TO_CENTER, (0.0, 50.0, 0.0), None, 1)]
    if step == 40:
        rows += [(6, E.BX_EDIT_TORQUE, (0.0, 3.0, 1.0), None, 1),
                 (7, E.BX_EDIT_LINEAR_IMPULSE, (1.5, 2.0, 0.0), (0.2, 3.0, 0.1), 1),
                 (8, E.BX_EDIT_ANGULAR_IMPULSE, (0.3, 0.0, 0.2), None, 1)]
    if step == 60:
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        rows += [(9, E.BX_EDIT_TRANSFORM, (0.5, 6.0, -0.5), tuple(np.float32(q)))]
    if step == 80:
        rows += [(10, E.BX_EDIT_AWAKE, (0, 0, 0), None, 0),
                 (11, E.BX_EDIT_GRAVITY_SCALE, (0.5, 0.25, 0.5)),
                 (12, E.BX_EDIT_LINEAR_DAMPING, (0.3, 0.3, 0.3)),
                 (12, E.BX_EDIT_ANGULAR_DAMPING, (0.2, 0.2, 0.2))]
    if step == 100:
        rows += [(13, E.BX_EDIT_SLEEPING_ALLOWED, (0, 0, 0), None, 0),
                 (14, E.BX_EDIT_ANGULAR_VELOCITY, (0.0, 2.0, 0.0))]
    if step in (120, 140, 150):
        # velocities and forces on whatever is asleep by now: the velocity wakes the body, a force with
        # wake false does nothing to a sleeping body
        rows += [(15, E.BX_EDIT_LINEAR_VELOCITY, (0.0, 1.0, 0.0)),
                 (16, E.BX_EDIT_FORCE, (0.0, 100.0, 0.0), (0.0, 0.0, 0.0), 0),
                 (17, E.BX_EDIT_LINEAR_IMPULSE, (0.0, 3.0, 0.0), (0.0, 1.0, 0.0), 0)]
    if step == 50:
        rows += [(19, E.BX_EDIT_LINEAR_SLEEP_TOLERANCE, (0.2, 0.0, 0.0)),
                 (20, E.BX_EDIT_ANGULAR_SLEEP_TOLERANCE, (0.1, 0.0, 0.0)),
                 (21, E.BX_EDIT_FIXED_ROTATION, (1.0, 0.0, 1.0))]
    # world edits (the body field is not read)
    if step == 70:
        rows += [(0, E.BX_EDIT_WORLD_GRAVITY, (0.5, -9.0, 0.2))]
    if step == 90:
        rows += [(0, E.BX_EDIT_WORLD_AUTO_CLEAR_FORCES, (0, 0, 0), None, 0)]  # body 3's force accumulates
    if step == 95:
        rows += [(0, E.BX_EDIT_WORLD_CLEAR_FORCES, (0, 0, 0)), (0, E.BX_EDIT_WORLD_AUTO_CLEAR_FORCES, (0, 0, 0), None, 1)]
    if step == 110:
        rows += [(0, E.BX_EDIT_WORLD_WARM_STARTING, (0, 0, 0), None, 0)]
    if step == 115:
        rows += [(0, E.BX_EDIT_WORLD_WARM_STARTING, (0, 0, 0), None, 1)]
    if step == 125:
        rows += [(21, E.BX_EDIT_FIXED_ROTATION, (0.0, 0.0, 0.0))]
    if step == 152:
        rows += [(0, E.BX_EDIT_WORLD_SLEEPING, (0, 0, 0), None, 0)]  # wakes every body
    if step == 156:
        rows += [(0, E.BX_EDIT_WORLD_SLEEPING, (0, 0, 0), None, 1)]
    if step == 130:
        rows += [(0, E.BX_EDIT_TRANSFORM, (0.0, -0.55, 0.1), (0.0, 0.0, 0.0, 1.0)),  # move the static ground
                 (18, E.BX_EDIT_AWAKE, (0, 0, 0), None, 1)]
    return rows
