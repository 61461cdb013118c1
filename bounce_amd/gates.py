"""Workloads that are held off the GPU.

The mixed-shapes workload (spheres, capsules and random hulls of 8-16 points in lognormal-sized
worlds of up to 2,000 bodies, and the serving catalog of 50-5,000-body worlds drawn from it) ended
in GPU memory faults at batch scale before any cause was shown. Until a cause has been found and
fixed, the entry points that run it at that scale refuse to start.
"""

MIXED_GPU_ALLOWED = False


class WorkloadBlocked(RuntimeError):
    pass


def require_mixed_gpu(what):
    if not MIXED_GPU_ALLOWED:
        raise WorkloadBlocked(
            f"{what}: the mixed-shapes workload at batch scale is disabled on the GPU (it ended in GPU "
            "memory faults whose cause is not fixed); use the stacking, drops or boxes workloads")
