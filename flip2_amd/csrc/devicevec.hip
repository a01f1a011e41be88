// DeviceVec::getMax: the reference's sampled maximum (cudaVec.cu:174-233).
//
// getMaxFromArray compares shared[t] with shared[t + blockDim.x] at every step
// of its loop (cudaVec.cu:190-196), so a 512-element tile reduces to
// pick(tile[0], tile[256]) and the next level works on those results, tiles
// padded with 0, until one value is left.  The result therefore depends on at
// most 2^L elements of the input (L = number of levels, <= 4 for 32-bit
// sizes): one thread evaluates that binary tree directly.
#include "device.hpp"

namespace flip2 {
namespace {

template <typename T>
__device__ inline T magnitude(T v) {
    return v;  // unsigned and char: the reference's abs() is the identity (cudaVec.cu:180-186)
}
template <>
__device__ inline double magnitude<double>(double v) {
    return fabs(v);
}
template <>
__device__ inline float magnitude<float>(float v) {
    return fabsf(v);
}
template <>
__device__ inline int32_t magnitude<int32_t>(int32_t v) {
    return v < 0 ? -v : v;
}

template <typename T>
__device__ inline T pick(T a, T b, bool absolute) {
    if (absolute) return magnitude(a) > magnitude(b) ? a : b;
    return a > b ? a : b;
}

constexpr int kMaxLevels = 6;

template <typename T>
__global__ void sampled_max_kernel(const T* __restrict__ a, uint32_t n, bool absolute, T* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    // sizes[l]: elements at level l (level 0 is the input)
    uint64_t sizes[kMaxLevels + 1];
    sizes[0] = n;
    sizes[1] = n / 512u + 1u;
    int L = 1;
    while (sizes[L] > 1 && L < kMaxLevels) {
        sizes[L + 1] = sizes[L] / 512u + 1u;
        ++L;
    }
    // indices of the nodes the result depends on, level L down to 0:
    // node k of level l has children 2k, 2k+1 at level l-1, at the element
    // indices 512 t and 512 t + 256 of its own index t
    uint64_t idx[1 << kMaxLevels];
    T val[1 << kMaxLevels];
    idx[0] = 0;
    int count = 1;
    for (int l = L; l >= 1; --l) {
        for (int k = count - 1; k >= 0; --k) {
            const uint64_t t = idx[k];
            idx[2 * k] = 512u * t;
            idx[2 * k + 1] = 512u * t + 256u;
        }
        count *= 2;
    }
    for (int k = 0; k < count; ++k) val[k] = idx[k] < sizes[0] ? a[idx[k]] : T(0);
    // combine upward; a node outside its level's size reads as 0
    for (int l = 1; l <= L; ++l) {
        count /= 2;
        for (int k = 0; k < count; ++k) val[k] = pick(val[2 * k], val[2 * k + 1], absolute);
        // the nodes of level l that lie outside [0, sizes[l]) read as 0
        if (l < L) {
            // element index of node k at level l: walk down from the root
            for (int k = 0; k < count; ++k) {
                uint64_t t = 0;
                for (int m = L; m > l; --m) {
                    const int bit = (k >> (m - l - 1)) & 1;
                    t = 512u * t + (bit ? 256u : 0u);
                }
                if (t >= sizes[l]) val[k] = T(0);
            }
        }
    }
    out[0] = val[0];
}

}  // namespace

template <typename T>
T sampled_max(const T* d, uint32_t n, bool absolute, T* scratch, hipStream_t s) {
    FLIP2_REQUIRE(d && n > 0, kErrArgument, "sampled_max: empty input");
    hipLaunchKernelGGL(sampled_max_kernel<T>, dim3(1), dim3(1), 0, s, d, n, absolute, scratch);
    FLIP2_LAUNCH_CHECK();
    T out{};
    FLIP2_HIP(hipMemcpyAsync(&out, scratch, sizeof(T), hipMemcpyDeviceToHost, s));
    FLIP2_HIP(hipStreamSynchronize(s));
    return out;
}

template double sampled_max<double>(const double*, uint32_t, bool, double*, hipStream_t);
template float sampled_max<float>(const float*, uint32_t, bool, float*, hipStream_t);
template uint32_t sampled_max<uint32_t>(const uint32_t*, uint32_t, bool, uint32_t*, hipStream_t);
template int32_t sampled_max<int32_t>(const int32_t*, uint32_t, bool, int32_t*, hipStream_t);
template char sampled_max<char>(const char*, uint32_t, bool, char*, hipStream_t);
template int8_t sampled_max<int8_t>(const int8_t*, uint32_t, bool, int8_t*, hipStream_t);

}  // namespace flip2
