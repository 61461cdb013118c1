// bx_calibrate.hip -- plain device copy of a known size, for calibrating HBM traffic counters.
//
// libbxcal.so is measurement support, not part of the step: a profiler's FETCH_SIZE and
// WRITE_SIZE readings for k_copy (which reads and writes exactly `bytes` once each) give the
// factor that turns the same counters read for the step's kernels into bytes.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__global__ void k_fill(uint4* p, size_t n, uint32_t v) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        p[i] = make_uint4(v, v + 1, v + 2, v + 3);
}

__global__ void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
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

}  // extern "C"
