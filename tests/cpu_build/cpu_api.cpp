// CPU build of the device library (test tool): the host side (bounce_amd/csrc/bx_api.hip) and, in the
// AddressSanitizer build, the per-world range check.
//
// Every per-world device array (the kArrays table of bx_api.hip, plus the world table and world
// states) is poisoned as a whole; when a workgroup or a lane starts work on world w (BX_WORLD_HOOK),
// exactly w's ranges of every array are unpoisoned, and they are poisoned again when it moves to
// another world or the workgroup ends. Any access by a world's kernel code outside its own ranges,
// including into a neighbouring world's range of the same flat array, is then reported.
#include "hip/hip_runtime.h"
#include "../../bounce_amd/csrc/bx_api.hip"

#include <mutex>
#include <set>

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

namespace bxcpu {
extern void (*g_worldHook)(const void* view, uint32_t w);
extern void (*g_blockEnd)();
extern void (*g_hostAccessEnd)(void* p, size_t n);
extern void (*g_onFree)(void* p);
bool FindAlloc(const void* p, char** base, size_t* n);
}  // namespace bxcpu

namespace {

std::mutex g_checkMutex;
std::map<char*, size_t> g_poisoned;  // whole arrays under the per-world check: start -> bytes
thread_local const WorldInfo* t_world = nullptr;
thread_local std::vector<std::pair<char*, size_t>> t_open;

void PoisonWhole(const void* p) {
    char* base;
    size_t n;
    if (!p || !bxcpu::FindAlloc(p, &base, &n)) return;
    std::lock_guard<std::mutex> lk(g_checkMutex);
    if (g_poisoned.count(base)) return;
    g_poisoned[base] = n;
    __asan_poison_memory_region(base, n);
}

void Open(const void* p, size_t n) {
    if (!n) return;
    char* q = static_cast<char*>(const_cast<void*>(p));
    __asan_unpoison_memory_region(q, n);
    t_open.emplace_back(q, n);
}

void Close() {
    for (auto& r : t_open) __asan_poison_memory_region(r.first, r.second);
    t_open.clear();
    t_world = nullptr;
}

__attribute__((no_sanitize("address"))) void WorldHook(const void* pv, uint32_t w) {
    const BatchView& B = *static_cast<const BatchView*>(pv);
    const WorldInfo* wp = B.worlds + w;
    if (wp == t_world) return;
    Close();
    for (const ArraySpec& a : kArrays) PoisonWhole(ArrayPtr(B, a));
    PoisonWhole(B.worlds);
    PoisonWhole(B.wstate);
    const WorldInfo wi = *wp;
    u32 units[kSpaces], base[kSpaces];
    WorldUnits(wi, units);
    GetBases(wi, base);
    for (const ArraySpec& a : kArrays) {
        char* p = ArrayPtr(B, a);
        const u64 unit = u64(a.perUnit) * a.elemSize;
        Open(p + u64(base[a.space]) * unit, u64(units[a.space]) * unit);
    }
    Open(wp, sizeof(WorldInfo));
    Open(B.wstate + w, sizeof(WorldState));
    t_world = wp;
}

void BlockEnd() { Close(); }

void HostAccessEnd(void* p, size_t n) {
    char* c = static_cast<char*>(p);
    std::lock_guard<std::mutex> lk(g_checkMutex);
    auto it = g_poisoned.upper_bound(c);
    if (it == g_poisoned.begin()) return;
    --it;
    if (c >= it->first + it->second) return;
    __asan_poison_memory_region(c, n);
}

void OnFree(void* p) {
    std::lock_guard<std::mutex> lk(g_checkMutex);
    g_poisoned.erase(static_cast<char*>(p));
}

struct Install {
    Install() {
        bxcpu::g_worldHook = WorldHook;
        bxcpu::g_blockEnd = BlockEnd;
        bxcpu::g_hostAccessEnd = HostAccessEnd;
        bxcpu::g_onFree = OnFree;
    }
} g_install;

}  // namespace

// Self-test of the check: reads moved[] of world w at index fixtureCount + past, as a kernel working on
// world w would. past 0 is the world's own last unit of the fixture space, past 1 the next world's
// first (reported).
extern "C" int bxcpu_probe_world_range(bx_batch* b, uint32_t w, uint32_t past) {
    WorldHook(&b->view, w);
    const WorldInfo& wi = b->view.worlds[w];
    volatile u32 x = b->view.moved[wi.fixtureBase + wi.fixtureCount + past];
    (void)x;
    BlockEnd();
    return 0;
}
#endif  // BXCPU_ASAN
