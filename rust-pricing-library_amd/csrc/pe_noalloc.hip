// Fixed device workspace of the allocation-free entry point
// (pe_plan_evaluate_noalloc): one file-scope array per GPU, sized once, and
// its device address per device.
//
// In host code the name `scratch` is only a host-side stand-in for the
// symbol; kernels must be given the address that hipGetSymbolAddress returns
// for the current device, which scratch_setup() records once per device.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pe_plan.hpp"

namespace pe {

namespace {

constexpr size_t kScratchBytes = static_cast<size_t>(8) << 20;   // 8 MiB per GPU, held until the process ends
constexpr int kScratchDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kScratchDevices] = {};

}  // namespace

size_t scratch_bytes() { return kScratchBytes; }

// Records the device address of `scratch` for `device`, which must be the
// current device.  Idempotent per device.
hipError_t scratch_setup(int device) {
    if (device < 0 || device >= kScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return hipSuccess;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

// Device address of this device's workspace (nullptr before scratch_setup).
void* scratch_pointer(int device) {
    if (device < 0 || device >= kScratchDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

}  // namespace pe
