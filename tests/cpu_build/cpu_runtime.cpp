// CPU build of the device library (test tool): the emulated HIP runtime behind hip/hip_runtime.h.
#include "hip/hip_runtime.h"
// the runtime itself names the members, not the kernels' per-thread values
#undef threadIdx
#undef blockIdx
#undef blockDim
#undef gridDim

#include <ucontext.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <mutex>
#include <vector>

#if defined(__has_feature)
#if __has_feature(address_sanitizer)
#define BXCPU_ASAN 1
#endif
#endif
#if defined(__SANITIZE_ADDRESS__)
#define BXCPU_ASAN 1
#endif
#ifdef BXCPU_ASAN
#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>
#endif

struct bxcpu_stream { int unused; };
struct bxcpu_event { std::chrono::steady_clock::time_point t; };
struct LaunchRec {
    std::function<void()> body;
    dim3 grid, block;
    size_t lds;
};
struct bxcpu_graph { std::vector<LaunchRec> launches; };

extern "C" void* bxcpu_lds_buffer(size_t* bytes);  // the dynamic LDS buffer (defined with the kernels)

namespace bxcpu {

thread_local ThreadCtx tl_ctx;

// Hooks of the address-checking build (set by the translation unit that knows the array layout).
void (*g_worldHook)(const void* view, uint32_t w) = nullptr;
void (*g_blockEnd)() = nullptr;
void (*g_hostAccessEnd)(void* p, size_t n) = nullptr;
void (*g_onFree)(void* p) = nullptr;

namespace {

const size_t kStackBytes = size_t(128) << 10;  // per workgroup thread; the largest kernel frame is a few KiB
enum { kReady, kWaiting, kDone, kWaitGroup };
const unsigned kGroup = 8;  // threads of a lane group (GroupDpp, GroupSync)

struct Fiber {
    ucontext_t ctx;
    char* stack = nullptr;
    int state = kDone;
    void* fake = nullptr;
    // the world this thread works on (BX_WORLD_HOOK): lanes of one workgroup may work on different
    // worlds (an island per lane group), so the range check follows the thread that runs
    const void* view = nullptr;
    uint32_t world = 0;
};

struct Block {
    std::vector<Fiber> fibers;
    ucontext_t sched;
    int cur = -1;
    unsigned alive = 0;
    int orAcc = 0, orResult = 0;
    std::vector<unsigned> xch;
    // group exchanges: two value buffers alternate by each thread's exchange count, so a thread
    // that has gone on to the next exchange never overwrites a value its group has yet to read
    std::vector<unsigned> gx[2], garrive[2], gcount;
    const std::function<void()>* body = nullptr;
    void* schedFake = nullptr;
    const void* schedBottom = nullptr;
    size_t schedSize = 0;
};

thread_local Block* tl_block = nullptr;

}  // namespace

void WorldHook(const void* view, uint32_t w) {
    if (!g_worldHook) return;
    if (tl_block && tl_block->cur >= 0) {
        Fiber& f = tl_block->fibers[size_t(tl_block->cur)];
        f.view = view;
        f.world = w;
    }
    g_worldHook(view, w);
}

namespace {

void ToScheduler(bool dying) {
    Block* b = tl_block;
    Fiber& f = b->fibers[b->cur];
#ifdef BXCPU_ASAN
    __sanitizer_start_switch_fiber(dying ? nullptr : &f.fake, b->schedBottom, b->schedSize);
#else
    (void)dying;
#endif
    swapcontext(&f.ctx, &b->sched);
#ifdef BXCPU_ASAN
    __sanitizer_finish_switch_fiber(f.fake, &b->schedBottom, &b->schedSize);
#endif
}

void FiberEntry() {
    Block* b = tl_block;
#ifdef BXCPU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &b->schedBottom, &b->schedSize);
#endif
    (*b->body)();
    b->fibers[b->cur].state = kDone;
    --b->alive;
    ToScheduler(true);
    std::abort();  // a finished fiber is never resumed
}

void RunBlock(Block& b, const std::function<void()>& body, unsigned n) {
    if (b.fibers.size() < n) b.fibers.resize(n);
    b.xch.assign(n, 0u);
    for (int p = 0; p < 2; ++p) {
        b.gx[p].assign(n, 0u);
        b.garrive[p].assign((n + kGroup - 1) / kGroup, 0u);
    }
    b.gcount.assign(n, 0u);
    b.body = &body;
    b.alive = n;
    b.orAcc = b.orResult = 0;
    for (unsigned i = 0; i < n; ++i) {
        Fiber& f = b.fibers[i];
        if (!f.stack) f.stack = static_cast<char*>(std::malloc(kStackBytes));
        getcontext(&f.ctx);
        f.ctx.uc_stack.ss_sp = f.stack;
        f.ctx.uc_stack.ss_size = kStackBytes;
        f.ctx.uc_link = nullptr;
        makecontext(&f.ctx, FiberEntry, 0);
        f.state = kReady;
        f.fake = nullptr;
        f.view = nullptr;
    }
    Block* prev = tl_block;
    tl_block = &b;
    while (b.alive > 0) {
        bool ran = false;
        for (unsigned i = 0; i < n; ++i) {
            if (b.fibers[i].state != kReady) continue;
            ran = true;
            b.cur = int(i);
            tl_ctx.threadIdx = dim3(i, 0, 0);
            if (g_worldHook && b.fibers[i].view) g_worldHook(b.fibers[i].view, b.fibers[i].world);
#ifdef BXCPU_ASAN
            __sanitizer_start_switch_fiber(&b.schedFake, b.fibers[i].stack, kStackBytes);
#endif
            swapcontext(&b.sched, &b.fibers[i].ctx);
#ifdef BXCPU_ASAN
            __sanitizer_finish_switch_fiber(b.schedFake, nullptr, nullptr);
#endif
        }
        if (b.alive == 0) break;
        bool ready = false, inGroup = false;
        for (unsigned i = 0; i < n; ++i) {
            ready = ready || b.fibers[i].state == kReady;
            inGroup = inGroup || b.fibers[i].state == kWaitGroup;
        }
        if (ready) continue;  // a group exchange woke threads this pass had already passed
        if (inGroup || !ran) {
            std::fprintf(stderr, "bxcpu: a lane group waits for threads that never reach its exchange\n");
            std::abort();
        }
        // every live thread waits at the barrier: release them (finished threads do not take part,
        // as a wavefront that has ended does not hold a workgroup barrier)
        b.orResult = b.orAcc;
        b.orAcc = 0;
        for (unsigned i = 0; i < n; ++i)
            if (b.fibers[i].state == kWaiting) b.fibers[i].state = kReady;
    }
    tl_block = prev;
    b.cur = -1;
}

std::mutex g_memMutex;
// device allocations, sorted by start (a vector rather than a std::map: the map's rebalancing code
// lives in the uninstrumented libstdc++, which MemorySanitizer would report on)
std::vector<std::pair<char*, size_t>> g_allocs;

std::vector<std::pair<char*, size_t>>::iterator AllocAfter(const void* p) {
    return std::upper_bound(g_allocs.begin(), g_allocs.end(), static_cast<const char*>(p),
                            [](const char* x, const std::pair<char*, size_t>& a) { return x < a.first; });
}
bxcpu_graph* g_capture = nullptr;

void RunLaunch(const LaunchRec& L) {
    static thread_local Block block;
    size_t ldsBytes = 0;
    char* lds = static_cast<char*>(bxcpu_lds_buffer(&ldsBytes));
    if (L.lds > ldsBytes) {
        std::fprintf(stderr, "bxcpu: launch asks %zu bytes of dynamic LDS, emulation holds %zu\n", L.lds, ldsBytes);
        std::abort();
    }
    if (L.block.y != 1 || L.block.z != 1) {
        std::fprintf(stderr, "bxcpu: only one-dimensional workgroups are emulated\n");
        std::abort();
    }
#ifdef BXCPU_ASAN
    __asan_unpoison_memory_region(lds, ldsBytes);
    __asan_poison_memory_region(lds + L.lds, ldsBytes - L.lds);
#endif
    ThreadCtx saved = tl_ctx;
    for (unsigned z = 0; z < L.grid.z; ++z)
        for (unsigned y = 0; y < L.grid.y; ++y)
            for (unsigned x = 0; x < L.grid.x; ++x) {
                tl_ctx.blockIdx = dim3(x, y, z);
                tl_ctx.blockDim = L.block;
                tl_ctx.gridDim = L.grid;
                RunBlock(block, L.body, L.block.x);
                if (g_blockEnd) g_blockEnd();
            }
    tl_ctx = saved;
#ifdef BXCPU_ASAN
    __asan_unpoison_memory_region(lds, ldsBytes);
#endif
}

void HostBegin(void* p, size_t n) {
#ifdef BXCPU_ASAN
    if (p && n) __asan_unpoison_memory_region(p, n);
#else
    (void)p;
    (void)n;
#endif
}
void HostEnd(void* p, size_t n) {
    if (p && n && g_hostAccessEnd) g_hostAccessEnd(p, n);
}

}  // namespace

void Barrier() {
    Block* b = tl_block;
    b->fibers[b->cur].state = kWaiting;
    ToScheduler(false);
}

int BarrierOr(int v) {
    tl_block->orAcc |= v ? 1 : 0;
    Barrier();
    return tl_block->orResult;
}

unsigned ShflUp(unsigned v, unsigned d) {
    Block* b = tl_block;
    const unsigned t = unsigned(b->cur);
    b->xch[t] = v;
    Barrier();
    const unsigned r = (t & 63u) >= d ? b->xch[t - d] : v;
    Barrier();
    return r;
}

namespace {
// Deposits v and returns the eight values of the caller's group once all eight have deposited.
const unsigned* GroupExchange(unsigned v) {
    Block* b = tl_block;
    const unsigned t = unsigned(b->cur), g = t / kGroup;
    const unsigned par = b->gcount[t]++ & 1u;
    b->gx[par][t] = v;
    if (++b->garrive[par][g] == kGroup) {
        b->garrive[par][g] = 0;
        for (unsigned i = g * kGroup; i < g * kGroup + kGroup; ++i)
            if (i != t && b->fibers[i].state == kWaitGroup) b->fibers[i].state = kReady;
    } else {
        b->fibers[t].state = kWaitGroup;
        ToScheduler(false);
    }
    return &b->gx[par][g * kGroup];
}
}  // namespace

unsigned GroupDpp(unsigned old, unsigned src, int ctrl, int bankMask) {
    const unsigned* vals = GroupExchange(src);
    const unsigned t = unsigned(tl_block->cur);
    const unsigned r = t & 15u;  // lane within its DPP row of 16
    if (((unsigned(bankMask) >> (r >> 2)) & 1u) == 0) return old;
    unsigned s;
    if (ctrl >= 0 && ctrl < 0x100) {
        s = (t & ~3u) | ((unsigned(ctrl) >> (2 * (t & 3u))) & 3u);
    } else if (ctrl > 0x100 && ctrl < 0x110) {  // row_shl:n
        const unsigned n = unsigned(ctrl) - 0x100u;
        if (r + n >= 16) return old;
        s = t + n;
    } else if (ctrl > 0x110 && ctrl < 0x120) {  // row_shr:n
        const unsigned n = unsigned(ctrl) - 0x110u;
        if (r < n) return old;
        s = t - n;
    } else {
        std::fprintf(stderr, "bxcpu: DPP control 0x%x is not emulated\n", unsigned(ctrl));
        std::abort();
    }
    if (s / kGroup != t / kGroup) {
        std::fprintf(stderr, "bxcpu: DPP move of thread %u reads thread %u outside its group\n", t, s);
        std::abort();
    }
    return vals[s % kGroup];
}

void GroupSync() { (void)GroupExchange(0u); }

void Launch(std::function<void()> body, dim3 grid, dim3 block, size_t dynLds, hipStream_t) {
    if (tl_block) {
        std::fprintf(stderr, "bxcpu: kernel launched from device code\n");
        std::abort();
    }
    LaunchRec L{std::move(body), grid, block, dynLds};
    if (g_capture) {
        g_capture->launches.push_back(std::move(L));
        return;
    }
    RunLaunch(L);
}

// The device allocation holding p: its start and size.
bool FindAlloc(const void* p, char** base, size_t* n) {
    std::lock_guard<std::mutex> lk(g_memMutex);
    auto it = AllocAfter(p);
    if (it == g_allocs.begin()) return false;
    --it;
    char* b = it->first;
    if (static_cast<const char*>(p) >= b + it->second) return false;
    *base = b;
    *n = it->second;
    return true;
}

}  // namespace bxcpu

using namespace bxcpu;

hipError_t hipMalloc(void** p, size_t bytes) {
    void* q = std::malloc(bytes ? bytes : 1);
    if (!q) return hipErrorOutOfMemory;
    {
        std::lock_guard<std::mutex> lk(g_memMutex);
        g_allocs.insert(AllocAfter(q), std::make_pair(static_cast<char*>(q), bytes));
    }
    *p = q;
    return hipSuccess;
}
hipError_t hipFree(void* p) {
    if (!p) return hipSuccess;
    size_t n = 0;
    {
        std::lock_guard<std::mutex> lk(g_memMutex);
        auto it = AllocAfter(p);
        if (it == g_allocs.begin() || (--it)->first != p) return hipErrorInvalidValue;
        n = it->second;
        g_allocs.erase(it);
    }
if (g_onFree) g_onFree(p);
#ifdef BXCPU_ASAN
    __asan_unpoison_memory_region(p, n);
#endif
    std::free(p);
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
    *p = std::malloc(bytes ? bytes : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
    if (!bytes) return hipSuccess;
    void* s = const_cast<void*>(src);
    HostBegin(dst, bytes);
    HostBegin(s, bytes);
    std::memcpy(dst, src, bytes);
    HostEnd(dst, bytes);
    HostEnd(s, bytes);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind k, hipStream_t) {
    return hipMemcpy(dst, src, bytes, k);
}
hipError_t hipMemset(void* dst, int v, size_t bytes) {
    if (!bytes) return hipSuccess;
    HostBegin(dst, bytes);
    std::memset(dst, v, bytes);
    HostEnd(dst, bytes);
    return hipSuccess;
}
hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t) { return hipMemset(dst, v, bytes); }
hipError_t hipGetDeviceCount(int* n) {
    *n = 1;
    return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "hipSuccess" : "emulated HIP error"; }
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    *s = new bxcpu_stream();
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
    *e = new bxcpu_event();
    (*e)->t = std::chrono::steady_clock::now();
    return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
    if (!g_capture) e->t = std::chrono::steady_clock::now();
    return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    if (g_capture) return hipErrorInvalidValue;
    g_capture = new bxcpu_graph();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    if (!g_capture) return hipErrorInvalidValue;
    *g = g_capture;
    g_capture = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, void*, size_t) {
    *e = new bxcpu_graph(*g);
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (const LaunchRec& L : e->launches) RunLaunch(L);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
