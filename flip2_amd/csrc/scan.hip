// Plain-array instantiation of the decoupled look-back scan (scan_impl.hpp).
#include "scan_impl.hpp"

namespace flip2 {

void scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, bool exclusive, void* ws, size_t wsBytes,
              uint32_t* totalDev, hipStream_t stream) {
    FLIP2_REQUIRE(n <= 0x7fffffffu, kErrRange, "scan: n must be below 2^31");
    FLIP2_REQUIRE(n == 0 || (in && out && ws), kErrArgument, "scan: null buffer");
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15u) == 0;
    scan_launch(ScanArrayInput{in}, out, n, exclusive, vec, ws, wsBytes, totalDev, stream);
}

}  // namespace flip2
