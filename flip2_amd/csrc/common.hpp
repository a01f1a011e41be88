// Shared definitions for the MI355X (gfx950) particle pipeline library.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "timing.hpp"

namespace flip2 {

// Status codes returned through the C ABI (negative = error).
enum Status : int {
    kOk = 0,
    kErrArgument = -1,
    kErrHip = -2,
    kErrState = -3,
    kErrNoMemory = -4,
    kErrRange = -5,
    kErrInternal = -6,
};

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }
private:
    int code_;
};

inline void hip_check(hipError_t e, const char* what, const char* file, int line) {
    if (e != hipSuccess) {
        int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? kErrNoMemory : kErrHip;
        throw Error(code, std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" +
                              std::to_string(line) + ")");
    }
}

#define FLIP2_HIP(call) ::flip2::hip_check((call), #call, __FILE__, __LINE__)
#define FLIP2_LAUNCH_CHECK() ::flip2::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)
#define FLIP2_REQUIRE(cond, code, msg)                                  \
    do {                                                                \
        if (!(cond)) throw ::flip2::Error((code), std::string(msg));    \
    } while (0)

// Workgroups of `kernel` (block threads, no dynamic LDS) that the current
// device's CUs hold at once, rounded down to a multiple of 8 (one per XCD),
// at least 8: the grid of a kernel whose workgroups loop over a fixed share
// of the work, where a workgroup that waits for a free CU would run its
// whole share after the first ones finish.
template <typename Kernel>
inline uint32_t resident_groups(Kernel kernel, int block) {
    int dev = 0, cus = 0, perCu = 0;
    FLIP2_HIP(hipGetDevice(&dev));
    FLIP2_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    FLIP2_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, block, 0));
    return std::max(8u, (uint32_t)(cus * perCu) & ~7u);
}

// Domain description with the reference's field order and types
// (engine/grid.hpp:8-38): FP32 corner and cell size, u32 cell counts.
struct Grid {
    float negX, negY, negZ;
    uint32_t sizeX, sizeY, sizeZ;
    float cellSize;
};

// Launch constants of the reference (engine/typedefs.h:7-8) kept for the
// API surface; the kernels here choose their own shapes.
constexpr int kBlockSize = 256;
constexpr int kWorkSize = kBlockSize << 1;

// Reference-compatible node geometry: radius 2, refinement level 1
// (engine/particles.cu:15-38) -> 8^3 voxels per node, sub-voxel = cell/4.
constexpr uint32_t kRcVoxels1D = 8;
constexpr uint32_t kRcVoxelsPerNode = 512;
constexpr uint32_t kRcApron = 2;
constexpr double kRcRadius = 2.0;
constexpr uint32_t kRcRefinement = 1;

inline unsigned ceil_div(uint64_t a, uint64_t b) { return (unsigned)((a + b - 1) / b); }

// Stream-ordered device copy by a kernel (util.hip).
void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s);

}  // namespace flip2
