// Device scratch of the allocation-free analog entry point
// (yaqs_run_analog_noalloc): one file-scope device array per GPU, whose
// device address each context looks up once at creation.
#include <array>
#include <mutex>

#include "engine.h"
#include "noalloc.h"

namespace yq {

namespace {

// Every GPU the library runs on holds its own copy of this array (256 MiB)
// from the first use of the library's code on that device until the process
// ends.  Host code never passes `scratch` itself to a kernel or a copy: its
// name is a host-side stand-in, not a device address; the address comes from
// hipGetSymbolAddress below.
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
std::array<void *, kMaxDevices> g_ptr{};         // device address of `scratch`, per device
std::array<std::mutex, kMaxDevices> g_run_mu;    // serialises the entry point per device

}  // namespace

int noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device]) return 0;
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) return -2;  // the device must be current
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return -3;
    g_ptr[device] = p;
    return 0;
}

void *noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}

std::mutex &noalloc_lock(int device) { return g_run_mu[(unsigned)device % kMaxDevices]; }

}  // namespace yq
