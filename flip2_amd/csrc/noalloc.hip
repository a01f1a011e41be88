// flip2_sort_pairs_u32_noalloc: the key/value radix sort of
// flip2_sort_pairs_u32 with its workspace (digit histograms, digit offsets,
// tile counters and the passes' look-back status words) in a fixed device
// array of this library instead of caller- or library-allocated memory, for
// callers that cannot allocate (code under HIP graph capture, serving loops
// that forbid hipMalloc).  Nothing here calls hipMalloc, hipFree or an
// allocating library routine; the sort's kernels clear the words that must
// start at zero (a fill kernel, no hipMemsetAsync), so every call starts from
// a clean workspace.
//
// `scratch` is one array per GPU, alive until the process ends.  In host
// code its name is a host-side stand-in, not a device address: the device
// address is taken once per device with hipGetSymbolAddress (the device
// current) and only that pointer is passed to kernels.  Calls that overlap on
// one GPU (two host threads, or two streams without ordering) share the
// array and race.
#include "../../include/flip2.h"

#include <functional>
#include <mutex>

#include "common.hpp"
#include "sort.hpp"

namespace flip2 {
int capi_guarded(const std::function<void()>& f);  // capi.cpp: status code + flip2_last_error
}

namespace {

constexpr size_t kScratchBytes = FLIP2_NOALLOC_SORT_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mutex;
void* g_scratch[kMaxDevices] = {};

// Device address of this GPU's `scratch` (looked up on the first call made
// with that device current).
void* scratch_address() {
    int dev = 0;
    FLIP2_HIP(hipGetDevice(&dev));
    FLIP2_REQUIRE(dev >= 0 && dev < kMaxDevices, flip2::kErrArgument, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        FLIP2_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        FLIP2_REQUIRE(p != nullptr, flip2::kErrInternal, "no device address for the sort scratch array");
        g_scratch[dev] = p;
    }
    return g_scratch[dev];
}

}  // namespace

extern "C" int flip2_sort_pairs_u32_noalloc(const uint32_t* d_keys_in, const uint32_t* d_vals_in,
                                            uint32_t* d_keys_out, uint32_t* d_vals_out, uint32_t* d_keys_alt,
                                            uint32_t* d_vals_alt, uint32_t n, uint32_t begin_bit, uint32_t end_bit,
                                            void* stream) {
    return flip2::capi_guarded([&] {
        FLIP2_REQUIRE(end_bit <= 32 && begin_bit <= end_bit, flip2::kErrArgument, "sort: bad bit range");
        const size_t need = flip2::sort_workspace_bytes(n, begin_bit, end_bit);
        FLIP2_REQUIRE(need <= kScratchBytes, flip2::kErrRange,
                      "sort without allocation: " + std::to_string(n) + " keys over bits [" +
                          std::to_string(begin_bit) + ", " + std::to_string(end_bit) + ") need " +
                          std::to_string(need) + " bytes of workspace, the fixed array holds " +
                          std::to_string(kScratchBytes));
        if (n == 0) return;
        void* ws = scratch_address();
        flip2::sort_pairs_u32(d_keys_in, d_vals_in, d_keys_out, d_vals_out, d_keys_alt, d_vals_alt, n, begin_bit,
                              end_bit, ws, kScratchBytes, reinterpret_cast<hipStream_t>(stream));
    });
}
