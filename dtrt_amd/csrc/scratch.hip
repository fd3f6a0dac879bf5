// Device scratch of the allocation-free render entry point (dtrt_render_device_noalloc):
// one file-scope array per GPU, loaded with the library's code object and held until
// the process ends. Host code never uses the array's name as an address (on the host
// it is only a stand-in for the symbol): scratchArea looks the device address up with
// hipGetSymbolAddress, once per device, and every user takes that address.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

#include "launch.h"

namespace dt {

static __device__ __align__(256) unsigned char scratch[DTRT_SCRATCH_BYTES];

hipError_t scratchArea(void **ptr, uint64_t *bytes) {
    static std::mutex mu;
    static std::map<int, void *> addr;   // device -> its copy of the array
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it == addr.end()) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        it = addr.emplace(dev, p).first;
    }
    *ptr = it->second;
    *bytes = DTRT_SCRATCH_BYTES;
    return hipSuccess;
}

}  // namespace dt
