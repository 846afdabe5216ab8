// ekg_scratch.hip -- the fixed device workspace of the allocation-free batch
// entry point (ekg_simulate_batch_device_noalloc).  It is a file-scope device
// array, so it exists on every GPU the library's code object is loaded on,
// from module load until the process ends, and no call allocates it.  Host
// code never uses the array's name as an address: the name is a host-side
// stand-in.  scratch_ptr() resolves each device's own copy once, with that
// device current, and every later use (kernel arguments, copies) goes
// through the resolved pointer.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ekg_kernels.hpp"

namespace ekgk {

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_ptr[MAX_DEVICES] = {};
}  // namespace

hipError_t scratch_ptr(int device, void** out)
{
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) {
            e = hipSetDevice(device);
            if (e != hipSuccess) return e;
        }
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_ptr[device] = p;
    }
    *out = g_ptr[device];
    return hipSuccess;
}

}  // namespace ekgk
