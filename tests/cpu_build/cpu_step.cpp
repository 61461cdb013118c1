// CPU build of the device library (test tool): the step kernels of bounce_amd/csrc/bx_step.hip.
#include "hip/hip_runtime.h"
#include "../../bounce_amd/csrc/bx_step.hip"

namespace bx {
// The dynamic LDS of a workgroup (`extern __shared__ u32 lds[]`); 64 KiB, the per-workgroup limit.
thread_local u32 lds[16384];
}  // namespace bx

extern "C" void* bxcpu_lds_buffer(size_t* bytes) {
    *bytes = sizeof(bx::lds);
    return bx::lds;
}
