// taa_resolve_frame_noalloc: taa_resolve_frame for callers that cannot allocate device memory.
// The only device memory taa_resolve_frame allocates is the intermediate screen of a split screen
// with CAS; here it lives in a file-scope device array, `scratch`, that every call on a GPU shares.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "taa_internal.hpp"

namespace taastar {
namespace {

// The intermediate screen: TAA_NOALLOC_SCRATCH_BYTES per device, held from the first call on the
// device until the process ends. Host code never uses the array's name as an address: the
// device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[TAA_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

// Device address of this device's `scratch` (resolved on the first call with the device current).
int scratch_address(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_last_error(TAA_ERR_HIP, hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return set_last_error(TAA_ERR_NO_DEVICE, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char msg[160];
            std::snprintf(msg, sizeof(msg), "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return set_last_error(TAA_ERR_HIP, msg);
        }
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return TAA_OK;
}

}  // namespace
}  // namespace taastar

extern "C" int32_t taa_resolve_frame_noalloc(const taa_uniforms_t* uniforms, const taa_frame_images_t* images,
                                             const taa_cas_consts_t* cas, int32_t arith, void* stream) {
    using namespace taastar;
    if (!(uniforms && images && cas && uniforms->splitScreen))  // no intermediate: taa_resolve_frame itself
        return resolve_frame_with_scratch(uniforms, images, cas, arith, stream, nullptr, 0);
    const size_t need = split_cas_scratch_bytes(images);
    if (need > (size_t)TAA_NOALLOC_SCRATCH_BYTES) {
        char msg[200];
        std::snprintf(msg, sizeof(msg),
                      "split screen with CAS needs a %zu-byte intermediate screen; taa_resolve_frame_noalloc holds %zu "
                      "(TAA_NOALLOC_SCRATCH_BYTES)",
                      need, (size_t)TAA_NOALLOC_SCRATCH_BYTES);
        return set_last_error(TAA_ERR_CAPACITY, msg);
    }
    void* p = nullptr;
    if (int rc = scratch_address(&p)) return rc;
    return resolve_frame_with_scratch(uniforms, images, cas, arith, stream, p, (size_t)TAA_NOALLOC_SCRATCH_BYTES);
}
