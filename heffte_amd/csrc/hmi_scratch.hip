// Static device workspace of the no-allocation entry point (heffte_mi355_execute_c2c_noalloc).
//
// `scratch` is a file-scope __device__ array: every GPU the library runs on holds its own copy
// (allocated when the runtime loads this code object on that device, kept until the process
// ends).  Host code never passes the array's name to a kernel -- on the host that name is only a
// stand-in -- but the device address hipGetSymbolAddress returns for the current device, looked
// up once per device and kept in scratch_address()'s table.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "hmi_kernels.h"

namespace hmi {

namespace {

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;

}  // namespace

void *scratch_address() {
    static std::mutex mtx;
    static void *addr[kMaxDevices] = {};
    int dev = -1;
    const hipError_t ge = hipGetDevice(&dev);
    if (ge != hipSuccess) throw std::runtime_error(std::string("hipGetDevice: ") + hipGetErrorString(ge));
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range for the scratch table");
    std::lock_guard<std::mutex> lock(mtx);
    if (!addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        addr[dev] = p;
    }
    return addr[dev];
}

}  // namespace hmi
