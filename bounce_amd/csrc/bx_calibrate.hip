// bx_calibrate.hip -- measurement and check support around the step library:
//   * a plain device copy of a known size, for calibrating HBM traffic counters: a profiler's
//     FETCH_SIZE and WRITE_SIZE readings for k_copy (which reads and writes exactly `bytes` once
//     each) give the factor that turns the same counters read for the step's kernels into bytes;
//   * the revolute joint's limit angle (bx::JointAngle, the step's own function) evaluated on the
//     device over given inputs, so tests can compare it with the host's atan2.
// libbxcal.so is not part of the step.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "bx_joint.h"

namespace {

__global__ void k_fill(uint4* p, size_t n, uint32_t v) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = make_uint4(v, v + 1, v + 2, v + 3);
}

__global__ void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_joint_angles(const float* s, const float* c, float* out, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        out[i] = bx::JointAngle(s[i], c[i]);
}

}  // namespace

extern "C" {

// Copies `bytes` (a multiple of 16) between two fresh device buffers `reps` times and returns the
// average kernel time in microseconds, or a negative value on a HIP error.
double bxcal_copy(uint64_t bytes, int reps) {
    if (bytes == 0 || bytes % 16 != 0 || reps <= 0) return -1.0;
    size_t n = bytes / 16;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return -2.0;
    if (hipMalloc(&b, bytes) != hipSuccess) { (void)hipFree(a); return -2.0; }
    const unsigned blocks = 4096, threads = 256;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    k_fill<<<blocks, threads>>>(a, n, 7u);
    k_fill<<<blocks, threads>>>(b, n, 0u);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; ++r) k_copy<<<blocks, threads>>>(a, b, n);
    (void)hipEventRecord(e1, 0);
    hipError_t err = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (err != hipSuccess) return -3.0;
    return double(ms) * 1e3 / reps;
}

// out[i] = bx::JointAngle(s[i], c[i]) computed on the device (host arrays of n floats). Returns 0,
// or a negative value on a HIP error.
int bxcal_joint_angles(const float* s, const float* c, float* out, uint64_t n) {
    if (n == 0) return 0;
    if (!s || !c || !out) return -1;
    float *ds = nullptr, *dc = nullptr, *dout = nullptr;
    const size_t bytes = size_t(n) * sizeof(float);
    int rc = 0;
    if (hipMalloc(&ds, bytes) != hipSuccess || hipMalloc(&dc, bytes) != hipSuccess ||
        hipMalloc(&dout, bytes) != hipSuccess) {
        rc = -2;
    } else if (hipMemcpy(ds, s, bytes, hipMemcpyHostToDevice) != hipSuccess ||
               hipMemcpy(dc, c, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        rc = -3;
    } else {
        k_joint_angles<<<1024, 256>>>(ds, dc, dout, size_t(n));
        if (hipDeviceSynchronize() != hipSuccess || hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = -4;
    }
    if (ds) (void)hipFree(ds);
    if (dc) (void)hipFree(dc);
    if (dout) (void)hipFree(dout);
    return rc;
}

}  // extern "C"
