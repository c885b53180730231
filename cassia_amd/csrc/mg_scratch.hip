// Static device scratch of the no-allocation render entry point: a fixed
// block in the library's code object, which every device the library runs
// on holds until the process ends. Host code never uses the array's name as
// an address (on the host it is only a stand-in): the device address of each
// device's copy comes from hipGetSymbolAddress, looked up once per device
// with that device current.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mg_scratch.h"

namespace mg {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

bool scratch_for_device(int device, void **out, std::string &err) {
    static std::mutex m;
    static void *ptrs[kMaxScratchDevices] = {};
    std::lock_guard<std::mutex> g(m);
    if (device < 0 || device >= kMaxScratchDevices) {
        err = "device index out of range for the no-allocation scratch";
        return false;
    }
    if (!ptrs[device]) {
        int prev = -1;
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        hipError_t e = hipSetDevice(device);
        void *p = nullptr;
        if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (prev >= 0) (void)hipSetDevice(prev);
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress of the no-allocation scratch: ") + hipGetErrorString(e);
            return false;
        }
        ptrs[device] = p;
    }
    *out = ptrs[device];
    return true;
}

}  // namespace mg
