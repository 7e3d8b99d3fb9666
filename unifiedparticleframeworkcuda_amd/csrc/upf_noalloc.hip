// upf_noalloc.hip -- the stable key sort for callers that cannot allocate
// device memory (code under stream capture, serving loops, harnesses that
// forbid hipMalloc): keys, sorted keys and permutation in caller-owned device
// memory, the sort's working memory in a fixed device array of this library.
#include <hip/hip_runtime.h>

#include <mutex>

#include "upf.h"
#include "upf_sort.h"

namespace upf {

// Working memory of one sort of up to UPF_NOALLOC_SORT_MAX keys: per-tile
// digit counts (256 x tiles of 1024 keys), digit totals, and the ping-pong
// (key, value) buffers of the passes.  Every device the library runs on
// holds its own copy until the process ends.
constexpr size_t kTiles = (UPF_NOALLOC_SORT_MAX + 1023) / 1024;
constexpr size_t kScratchBytes = (256 * kTiles + 256) * 4 + 2 * (size_t)UPF_NOALLOC_SORT_MAX * 4;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// The array's address on each device, looked up once per device with that
// device current (in host code `scratch` names a host-side stand-in, never
// the device memory).
constexpr int kMaxDevices = 64;
static void* g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_ptr(void** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return UPF_ERR_NO_DEVICE;
    if (dev >= kMaxDevices) return UPF_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return UPF_ERR_HIP;
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return UPF_OK;
}

}  // namespace upf

extern "C" int upf_radix_sort_noalloc(int n, const uint32_t* keys_in, uint32_t* keys_out, int32_t* perm_out,
                                      void* stream) {
    if (n < 0 || n > UPF_NOALLOC_SORT_MAX) return UPF_ERR_ARG;
    if (n == 0) return UPF_OK;
    if (!keys_in || !keys_out || !perm_out) return UPF_ERR_ARG;
    if (upf::radix_sort_temp_bytes(n) > upf::kScratchBytes) return UPF_ERR_ARG;
    void* tmp = nullptr;
    const int rc = upf::scratch_ptr(&tmp);
    if (rc) return rc;
    // every buffer the passes read is written first (counts by rs_hist,
    // offsets by rs_scan, ping-pong by rs_scatter): nothing to clear
    if (upf::radix_sort(keys_in, nullptr, keys_out, reinterpret_cast<uint32_t*>(perm_out), n, 32, tmp,
                        static_cast<hipStream_t>(stream)) != hipSuccess)
        return UPF_ERR_HIP;
    return UPF_OK;
}
