// Device scratch of the allocation-free group entry point
// (tanx_group_process_device_noalloc): the launch descriptors and block
// offsets of one group call live in a file-scope device array instead of a
// buffer the library allocates, so the entry point calls no allocator.
//
// Each GPU holds its own copy of the array, at its own address, until the
// process ends.  Host code never uses the array's name as an address: the
// per-device address comes from hipGetSymbolAddress, taken once per device
// with that device current, at context creation.  Every call on one GPU
// shares that GPU's copy, which keeps the last call's contents; calls that
// overlap on one GPU (two streams, or two host threads on different
// contexts) race on it.
#include "tan_internal.h"

#include <mutex>

namespace tanx {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex g_mtx;
void *g_ptr[kNoAllocMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mtx);
    if (g_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p || (reinterpret_cast<uintptr_t>(p) & 255)) return hipErrorInvalidValue;
    g_ptr[device] = p;
    return hipSuccess;
}

void *noalloc_scratch(int device)
{
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mtx);
    return g_ptr[device];
}

}  // namespace tanx
