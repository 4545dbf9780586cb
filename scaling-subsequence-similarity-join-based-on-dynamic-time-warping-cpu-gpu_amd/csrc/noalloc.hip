// Fixed device workspace of the no-allocation entry point (dtwj_top1_device_noalloc): one
// file-scope device array, of which every GPU the library runs on holds its own copy until the
// process ends.  In host code the array's name is only a stand-in for the symbol; the device
// address of a GPU's copy is looked up once per device (with that device current) and kept,
// and only that address is ever handed to kernels and copies.
#include <mutex>

#include "internal.hpp"

namespace dtwj {

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per device
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

size_t noalloc_scratch_bytes() { return kScratchBytes; }

int noalloc_scratch(int device, void **ptr) {
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) {
        set_error("device ordinal out of range for the fixed workspace");
        return DTWJ_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int prev = 0;
        DTWJ_HIP(hipGetDevice(&prev));
        DTWJ_HIP(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(prev);
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        if (!p) {
            set_error("hipGetSymbolAddress(scratch) returned no address");
            return DTWJ_EDEVICE;
        }
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return 0;
}

}  // namespace dtwj
