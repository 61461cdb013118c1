// CPU build of the device library (test tool): the HIP surface that bounce_amd/csrc uses, emulated on
// the host so that the product's own kernels run under AddressSanitizer, MemorySanitizer and
// UndefinedBehaviorSanitizer.
//
// Execution model: a kernel launch runs its workgroups one after another on the calling host
// thread; the threads of a workgroup are fibers (ucontext) switched only at barriers, so a
// workgroup's threads interleave exactly as far as the kernel's own synchronisation allows.
// __shfl_up, __syncthreads_or and the atomics keep their HIP semantics. Device memory is host
// memory; copies and fills are synchronous, so stream order is program order. Stream capture
// records the launches and a graph launch replays them.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <utility>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (reinterpret_cast<const void*>(&(x)))
// Block-scope thread_local is static per host thread; one workgroup runs at a time per host
// thread, so these are the workgroup's shared memory. `extern __shared__ T x[]` names the dynamic
// LDS buffer of the namespace (bx::lds), defined by the CPU build.
#define __shared__ thread_local

struct dim3 {
    unsigned x, y, z;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint2 { unsigned x, y; };
struct uint4 { unsigned x, y, z, w; };
inline uint2 make_uint2(unsigned x, unsigned y) { return uint2{x, y}; }
inline uint4 make_uint4(unsigned x, unsigned y, unsigned z, unsigned w) { return uint4{x, y, z, w}; }

typedef enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 } hipError_t;
typedef enum { hipMemcpyHostToHost, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
               hipMemcpyDefault } hipMemcpyKind;
typedef enum { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed }
    hipStreamCaptureMode;
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2 };

struct bxcpu_stream;
struct bxcpu_event;
struct bxcpu_graph;
typedef bxcpu_stream* hipStream_t;
typedef bxcpu_event* hipEvent_t;
typedef bxcpu_graph* hipGraph_t;
typedef bxcpu_graph* hipGraphExec_t;

namespace bxcpu {
struct ThreadCtx { dim3 threadIdx, blockIdx, blockDim, gridDim; };
extern thread_local ThreadCtx tl_ctx;
void Barrier();
int BarrierOr(int v);
unsigned ShflUp(unsigned v, unsigned d);
void Launch(std::function<void()> body, dim3 grid, dim3 block, size_t dynLds, hipStream_t s);
// Called by a kernel when its workgroup (or lane) starts work on world w of the batch view at `view`;
// the address-checking build allows that world's ranges of every per-world array and no other's.
void WorldHook(const void* view, uint32_t w);
// Cross-lane moves inside a group of eight consecutive threads (the velocity sweeps' lane groups):
// the caller's group must reach the same call together, as a wavefront's lanes execute a DPP move
// together. `ctrl` is a DPP control word (quad_perm, row_shl:n, row_shr:n); lanes outside `bankMask`
// keep `old`.
unsigned GroupDpp(unsigned old, unsigned src, int ctrl, int bankMask);
// Waits until the caller's group of eight threads has reached this call.
void GroupSync();
}  // namespace bxcpu

#define threadIdx (bxcpu::tl_ctx.threadIdx)
#define blockIdx (bxcpu::tl_ctx.blockIdx)
#define blockDim (bxcpu::tl_ctx.blockDim)
#define gridDim (bxcpu::tl_ctx.gridDim)
#define BX_WORLD_HOOK(view, w) bxcpu::WorldHook(&(view), (w))
#define BX_DPP_MOV(src, ctrl) int(bxcpu::GroupDpp(0u, unsigned(src), (ctrl), 0xF))
#define BX_DPP_UPDATE(old, src, ctrl, bankMask) int(bxcpu::GroupDpp(unsigned(old), unsigned(src), (ctrl), (bankMask)))
#define BX_GROUP_SYNC() bxcpu::GroupSync()

inline void __syncthreads() { bxcpu::Barrier(); }
inline int __syncthreads_or(int v) { return bxcpu::BarrierOr(v); }
template <class T> inline T __shfl_up(T v, unsigned d, int /*width*/) {
    static_assert(sizeof(T) == 4, "32-bit shuffles only");
    unsigned u;
    std::memcpy(&u, &v, 4);
    u = bxcpu::ShflUp(u, d);
    T r;
    std::memcpy(&r, &u, 4);
    return r;
}
template <class T> inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_RELAXED); }
template <class T> inline T atomicOr(T* p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_RELAXED); }
template <class T> inline T atomicAnd(T* p, T v) { return __atomic_fetch_and(p, v, __ATOMIC_RELAXED); }

#define hipLaunchKernelGGL(kernel, grid, block, lds, stream, ...) \
    bxcpu::Launch([=]() { kernel(__VA_ARGS__); }, dim3(grid), dim3(block), size_t(lds), stream)

hipError_t hipMalloc(void** p, size_t bytes);
template <class T> inline hipError_t hipMalloc(T** p, size_t bytes) { return hipMalloc(reinterpret_cast<void**>(p), bytes); }
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags = 0);
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind k, hipStream_t s = nullptr);
hipError_t hipMemset(void* dst, int v, size_t bytes);
hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t s = nullptr);
hipError_t hipGetDeviceCount(int* n);
hipError_t hipSetDevice(int d);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipGraphDestroy(hipGraph_t g);
// A __device__ variable is an ordinary host object here: its address is the symbol's.
inline hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
    *p = const_cast<void*>(symbol);
    return hipSuccess;
}
