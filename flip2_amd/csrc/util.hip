// Device copy kernel (16-byte vectors), used where a stream-ordered copy by
// a kernel rather than a DMA engine is wanted (and as the reference kernel of
// known traffic when hardware counters are calibrated).
#include "common.hpp"

namespace flip2 {

__global__ void copy16_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t count) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) dst[i] = src[i];
}

__global__ void copy1_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, size_t count) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = src[i];
}

void launch_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    const bool aligned = ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15u) == 0;
    const size_t vec = aligned ? bytes / 16 : 0;
    if (vec) {
        const size_t blocks = std::min<size_t>((vec + 255) / 256, 256u * 1024u);
        FLIP2_KERNEL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                           reinterpret_cast<const uint4*>(src), reinterpret_cast<uint4*>(dst), vec);
        FLIP2_LAUNCH_CHECK();
    }
    const size_t rest = bytes - vec * 16;
    if (rest) {
        FLIP2_KERNEL(copy1_kernel, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0, s,
                           reinterpret_cast<const uint8_t*>(src) + vec * 16, reinterpret_cast<uint8_t*>(dst) + vec * 16,
                           rest);
        FLIP2_LAUNCH_CHECK();
    }
}

}  // namespace flip2
