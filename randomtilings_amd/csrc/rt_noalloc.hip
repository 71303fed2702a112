// rt_noalloc.hip — the device memory behind rt_lozenge_walk_noalloc: one file-scope array per
// device, so that entry point allocates nothing (for callers that cannot call hipMalloc).
//
// The array lives in this library's code object: every device the library runs on holds
// kNoAllocScratchBytes of it from the code object's load until the process ends.  Host code
// never names the array: its device address is taken once per device, with that device
// current, by noalloc_setup (called from rt_create), and only that address is used.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "rt_internal.hpp"

static __device__ __align__(256) unsigned char scratch[rt::kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
bool g_done[kMaxDevices] = {};
void* g_ptr[kMaxDevices] = {};
hipError_t g_err[kMaxDevices] = {};

}  // namespace

namespace rt {

void noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_done[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_err[device] = e;
    g_ptr[device] = e == hipSuccess ? p : nullptr;
    g_done[device] = true;
}

void* noalloc_scratch(int device, std::string& why) {
    if (device < 0 || device >= kMaxDevices) {
        why = "device index out of range";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_done[device]) {
        why = "scratch address not set up on this device (rt_create does it)";
        return nullptr;
    }
    if (!g_ptr[device]) why = std::string("hipGetSymbolAddress: ") + hipGetErrorString(g_err[device]);
    return g_ptr[device];
}

}  // namespace rt
