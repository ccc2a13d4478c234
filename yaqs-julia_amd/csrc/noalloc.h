// Device scratch of the allocation-free analog entry point.
#pragma once

#include <cstddef>
#include <mutex>

namespace yq {

// Size of the per-device scratch array (256 MiB).
constexpr size_t kNoAllocScratchBytes = (size_t)256 << 20;

// With `device` current: looks up the device address of the scratch array
// (hipGetSymbolAddress) and keeps it for the process.  Returns 0 on
// success, < 0 when the device is not current or the lookup fails.
int noalloc_init(int device);
// The address kept by noalloc_init, or nullptr.
void *noalloc_scratch(int device);
// Held for the whole of one allocation-free call on `device`: calls from
// several threads of the process run one after the other.
std::mutex &noalloc_lock(int device);

}  // namespace yq
