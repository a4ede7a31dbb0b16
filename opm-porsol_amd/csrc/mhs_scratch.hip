// The per-GPU scratch array of the entry points that allocate nothing
// (mhs_sum_outflux_noalloc).  Its name is a host-side stand-in: host code
// passes the address that hipGetSymbolAddress returns for the current device,
// never the name itself.
#include "mhs_scratch.h"

#include <mutex>

static __device__ __align__(256) unsigned char scratch[mhs::kScratchBytes];

namespace mhs {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t scratch_address(int device, void **ptr) {
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int current = -1;
        hipError_t e = hipGetDevice(&current);
        if (e != hipSuccess) return e;
        if (current != device) return hipErrorInvalidDevice;   // the caller makes the device current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace mhs
