// Learned preconditioner host side (see unet.hpp).
#include "unet.hpp"

#include "unet_kernels.hpp"

#include <cmath>
#include <cstring>

namespace flip2 {

namespace nn {

namespace {
__device__ inline uint16_t to_bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline void dense_coords(uint32_t brick, uint32_t l, int n, int& x, int& y, int& z) {
    const uint32_t nb = (uint32_t)n >> 3;
    x = (int)((brick % nb) * 8u + (l & 7u));
    y = (int)(((brick / nb) % nb) * 8u + ((l >> 3) & 7u));
    z = (int)((brick / (nb * nb)) * 8u + (l >> 6));
}

inline dim3 g1(long long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }
}  // namespace

__global__ void zero_red_kernel(double* red, int count) {
    if ((int)threadIdx.x < count) red[threadIdx.x] = 0.0;
}

__global__ void clear_input_kernel(uint32_t* __restrict__ in0c, uint8_t* __restrict__ mask, long long nvox) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    in0c[v] = 0u;
    mask[v] = 0;
}

// The reductions below run grid-stride over a bounded number of workgroups
// and add one partial pair per workgroup: atomics on one address serialize,
// so their count, not the data, would otherwise set the time.
constexpr int kRedBlocks = 512;

__device__ inline void block_add2(double a, double c, double* red) {
    __shared__ double s_a[4], s_c[4];
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        c += __shfl_xor(c, d, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_a[wave] = a;
        s_c[wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tc = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            ta += s_a[w];
            tc += s_c[w];
        }
        if (ta != 0.0 || tc != 0.0) {
            atomicAdd(&red[0], ta);
            atomicAdd(&red[1], tc);
        }
    }
}

// Fixed-order variant: each workgroup stores its pair in partials[2 b, 2 b + 1]
// and sum_pairs_kernel adds them up in workgroup order (deterministic).
__device__ inline void block_pair_partial(double a, double c, double* partials) {
    __shared__ double s_a[4], s_c[4];
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        c += __shfl_xor(c, d, 64);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_a[wave] = a;
        s_c[wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tc = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            ta += s_a[w];
            tc += s_c[w];
        }
        partials[2 * blockIdx.x] = ta;
        partials[2 * blockIdx.x + 1] = tc;
    }
}

__device__ inline float rms_scale(const double* red);

// red[0..1] = sums of the pairs (one workgroup of up to 256 threads, a
// fixed order: thread t takes pairs t, t + blockDim, ... four at a time);
// scale != null: also *scale = RMS scale of the network input (rms_scale).
// stop != null and set: returns at once (work queued after the solver's
// convergence).
__global__ void __launch_bounds__(256) sum_pairs_kernel(const double* __restrict__ partials, int blocks,
                                                        double* __restrict__ red, float* __restrict__ scale = nullptr,
                                                        const uint32_t* __restrict__ stop = nullptr) {
    __shared__ double s_a[4], s_c[4];
    if (stop && *stop) return;
    const int nt = (int)blockDim.x;
    double a4[4] = {0.0, 0.0, 0.0, 0.0}, c4[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b0 = (int)threadIdx.x; b0 < blocks; b0 += 4 * nt) {
        double va[4], vc[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int b = b0 + k * nt;
            va[k] = b < blocks ? partials[2 * b] : 0.0;
            vc[k] = b < blocks ? partials[2 * b + 1] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            a4[k] += va[k];
            c4[k] += vc[k];
        }
    }
    double a = (a4[0] + a4[1]) + (a4[2] + a4[3]), c = (c4[0] + c4[1]) + (c4[2] + c4[3]);
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, 64);
        c += __shfl_xor(c, d, 64);
    }
    const int wave = (int)threadIdx.x >> 6, waves = (nt + 63) >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_a[wave] = a;
        s_c[wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tc = 0.0;
        for (int w = 0; w < waves; ++w) {
            ta += s_a[w];
            tc += s_c[w];
        }
        red[0] = ta;
        red[1] = tc;
        if (scale) scale[0] = rms_scale(red);
    }
}

inline dim3 red_grid(long long n) { return dim3((unsigned)std::min<long long>(kRedBlocks, (n + 255) / 256 > 0 ? (n + 255) / 256 : 1)); }

// per workgroup: (sum r^2, fluid count) over its cells (active layout)
__global__ void __launch_bounds__(256) active_norm_kernel(const uint8_t* __restrict__ cellType,
                                                          const float* __restrict__ r, uint32_t cells,
                                                          double* __restrict__ partials) {
    double a = 0.0, c = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += gridDim.x * blockDim.x) {
        if (cellType[i] == 1) {
            a += (double)r[i] * (double)r[i];
            c += 1.0;
        }
    }
    block_pair_partial(a, c, partials);
}

__device__ inline float rms_scale(const double* red) {
    const double cnt = red[1] > 0.0 ? red[1] : 1.0;
    const double s = sqrt(red[0] / cnt);
    return s > 0.0 ? (float)s : 1.0f;
}

__global__ void scatter_input_kernel(const uint32_t* __restrict__ activeBricks, const uint8_t* __restrict__ cellType,
                                     const float* __restrict__ r, uint32_t cells, int n, uint16_t* __restrict__ in0c,
                                     uint8_t* __restrict__ mask, float* __restrict__ scale, const double* red) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const float s = rms_scale(red);
    if (i == 0) scale[0] = s;
    if (i >= cells || cellType[i] != 1) return;
    int x, y, z;
    dense_coords(activeBricks[i >> 9], i & 511u, n, x, y, z);
    const long long v = ((long long)z * n + y) * n + x;
    in0c[v * 2] = to_bf(r[i] / s);
    in0c[v * 2 + 1] = to_bf(1.0f);
    mask[v] = 1;
}

__global__ void dense_system_kernel(const uint32_t* __restrict__ activeBricks, const uint16_t* __restrict__ coef,
                                    const double* __restrict__ b, const double* __restrict__ x, uint32_t cells, int n,
                                    float* __restrict__ bD, float* __restrict__ xD, uint16_t* __restrict__ coefD) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const uint16_t cf = coef[i];
    if (!cf) return;
    int xx, yy, zz;
    dense_coords(activeBricks[i >> 9], i & 511u, n, xx, yy, zz);
    const long long v = ((long long)zz * n + yy) * n + xx;
    bD[v] = (float)b[i];
    if (xD) xD[v] = (float)x[i];
    coefD[v] = cf;
}

__device__ inline float dense_A(const uint16_t* coefD, const float* x, int n, long long v) {
    const uint16_t cf = coefD[v];
    if (!cf) return 0.f;
    const long long st[3] = {1, n, (long long)n * n};
    float s = (float)(cf & 7u) * x[v];
    const uint32_t m = cf >> 3;
#pragma unroll
    for (int d = 0; d < 6; ++d)
        if ((m >> d) & 1u) s -= x[v + ((d & 1) ? st[d >> 1] : -st[d >> 1])];
    return s;
}

__global__ void dense_apply_kernel(const uint16_t* __restrict__ coefD, const float* __restrict__ x, float* __restrict__ y,
                                   int n) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long long)n * n * n) return;
    y[v] = dense_A(coefD, x, n, v);
}

__global__ void __launch_bounds__(256) residual_kernel(const uint16_t* __restrict__ coefD, const float* __restrict__ z,
                                                       const float* __restrict__ bD, float* __restrict__ res, int n,
                                                       double* __restrict__ red) {
    double a = 0.0, c = 0.0;
    const long long nvox = (long long)n * n * n;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
         v += (long long)gridDim.x * blockDim.x) {
        float rv = 0.f;
        if (coefD[v]) {
            rv = dense_A(coefD, z, n, v) - bD[v];
            c += (double)bD[v] * (double)bD[v];
        }
        res[v] = rv;
        a += (double)rv * (double)rv;
    }
    block_add2(a, c, red);
}

__global__ void loss_grad_kernel(const uint16_t* __restrict__ coefD, const float* __restrict__ res,
                                 const double* __restrict__ red, float* __restrict__ dOut, int n) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long long)n * n * n) return;
    const float k = red[1] > 0.0 ? (float)(2.0 / red[1]) : 0.f;
    dOut[v] = k * dense_A(coefD, res, n, v);
}

__global__ void __launch_bounds__(256) dense_norm_kernel(const float* __restrict__ bD,
                                                         const uint16_t* __restrict__ coefD, long long nvox,
                                                         double* __restrict__ red) {
    double a = 0.0, c = 0.0;
    for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
         v += (long long)gridDim.x * blockDim.x) {
        if (coefD[v]) {
            a += (double)bD[v] * (double)bD[v];
            c += 1.0;
        }
    }
    block_add2(a, c, red);
}

__global__ void dense_input_kernel(const float* __restrict__ bD, const uint16_t* __restrict__ coefD, long long nvox,
                                   uint16_t* __restrict__ in0, uint16_t* __restrict__ in0c, uint8_t* __restrict__ mask,
                                   float* __restrict__ scale, const double* __restrict__ red) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const float s = rms_scale(red);
    if (v == 0) scale[0] = s;
    if (v >= nvox) return;
    const bool f = coefD[v] != 0;
    const uint16_t c0 = f ? to_bf(bD[v] / s) : (uint16_t)0, c1 = f ? to_bf(1.0f) : (uint16_t)0;
    if (in0) {
        in0[v * 16] = c0;
        in0[v * 16 + 1] = c1;
    }
    in0c[v * 2] = c0;
    in0c[v * 2 + 1] = c1;
    mask[v] = f ? 1 : 0;
}

__device__ inline float dense_jac(const uint16_t* coefD, const float* r, float omega, long long v) {
    const uint16_t c = coefD[v];
    return c ? omega * r[v] / (float)(c & 7u) : 0.f;
}

// r1 = r - A (w D^-1 r) on fluid cells, 0 elsewhere
__global__ void dense_presmooth_kernel(const uint16_t* __restrict__ coefD, const float* __restrict__ r, float omega,
                                       float* __restrict__ r1, int n) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long long)n * n * n) return;
    const uint16_t cf = coefD[v];
    if (!cf) {
        r1[v] = 0.f;
        return;
    }
    const long long st[3] = {1, n, (long long)n * n};
    float s = (float)(cf & 7u) * dense_jac(coefD, r, omega, v);
    const uint32_t m = cf >> 3;
#pragma unroll
    for (int d = 0; d < 6; ++d)
        if ((m >> d) & 1u) s -= dense_jac(coefD, r, omega, v + ((d & 1) ? st[d >> 1] : -st[d >> 1]));
    r1[v] = r[v] - s;
}

// z = x + w D^-1 (r - A x), x = w D^-1 r + y
__global__ void dense_postsmooth_kernel(const uint16_t* __restrict__ coefD, const float* __restrict__ r,
                                        const float* __restrict__ y, float omega, float* __restrict__ z, int n) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (long long)n * n * n) return;
    const uint16_t cf = coefD[v];
    if (!cf) {
        z[v] = 0.f;
        return;
    }
    const long long st[3] = {1, n, (long long)n * n};
    auto xf = [&](long long u) { return coefD[u] ? dense_jac(coefD, r, omega, u) + y[u] : 0.f; };
    const float xv = xf(v);
    float s = (float)(cf & 7u) * xv;
    const uint32_t m = cf >> 3;
#pragma unroll
    for (int d = 0; d < 6; ++d)
        if ((m >> d) & 1u) s -= xf(v + ((d & 1) ? st[d >> 1] : -st[d >> 1]));
    z[v] = xv + omega * (r[v] - s) / (float)(cf & 7u);
}

void launch_dense_presmooth(const uint16_t* coefD, const float* r, float omega, float* r1, int n, hipStream_t s) {
    FLIP2_KERNEL(dense_presmooth_kernel, g1((long long)n * n * n, 256), dim3(256), 0, s, coefD, r, omega, r1, n);
    FLIP2_LAUNCH_CHECK();
}

void launch_dense_postsmooth(const uint16_t* coefD, const float* r, const float* y, float omega, float* z, int n,
                             hipStream_t s) {
    FLIP2_KERNEL(dense_postsmooth_kernel, g1((long long)n * n * n, 256), dim3(256), 0, s, coefD, r, y, omega, z,
                       n);
    FLIP2_LAUNCH_CHECK();
}

void launch_scatter_input(uint32_t numActive, const uint32_t* activeBricks, const uint8_t* cellType, const float* r,
                          int n, uint16_t* in0c, uint8_t* mask, float* scale, double* red, bool clear, hipStream_t s) {
    const long long nvox = (long long)n * n * n;
    const uint32_t cells = numActive * 512u;
    if (clear)
        FLIP2_KERNEL(clear_input_kernel, g1(nvox, 256), dim3(256), 0, s, reinterpret_cast<uint32_t*>(in0c), mask,
                           nvox);
    const dim3 rg = red_grid(cells);
    FLIP2_KERNEL(active_norm_kernel, rg, dim3(256), 0, s, cellType, r, cells, red + 8);
    FLIP2_KERNEL(sum_pairs_kernel, dim3(1), dim3(256), 0, s, red + 8, (int)rg.x, red, nullptr, nullptr);
    FLIP2_KERNEL(scatter_input_kernel, g1(cells ? cells : 1, 256), dim3(256), 0, s, activeBricks, cellType, r,
                       cells, n, in0c, mask, scale, red);
    FLIP2_LAUNCH_CHECK();
}

void launch_input_scale(uint32_t numActive, const uint8_t* cellType, const float* r, double* red, float* scale,
                        hipStream_t s) {
    const uint32_t cells = numActive * 512u;
    const dim3 rg = red_grid(cells);
    FLIP2_KERNEL(active_norm_kernel, rg, dim3(256), 0, s, cellType, r, cells, red + 8);
    FLIP2_KERNEL(sum_pairs_kernel, dim3(1), dim3(256), 0, s, red + 8, (int)rg.x, red, scale, nullptr);
    FLIP2_LAUNCH_CHECK();
}

void launch_pairs_scale(const double* pairs, uint32_t count, double* red, float* scale, hipStream_t s,
                        const uint32_t* stop) {
    FLIP2_KERNEL(sum_pairs_kernel, dim3(1), dim3(256), 0, s, pairs, (int)count, red, scale, stop);
    FLIP2_LAUNCH_CHECK();
}

__global__ void dense_vector_kernel(const uint32_t* __restrict__ activeBricks, const uint16_t* __restrict__ coef,
                                    const float* __restrict__ v, uint32_t cells, int n, float* __restrict__ vD,
                                    uint16_t* __restrict__ coefD) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cells) return;
    const uint16_t cf = coef[i];
    if (!cf) return;
    int xx, yy, zz;
    dense_coords(activeBricks[i >> 9], i & 511u, n, xx, yy, zz);
    const long long d = ((long long)zz * n + yy) * n + xx;
    vD[d] = v[i];
    coefD[d] = cf;
}

void launch_dense_vector(uint32_t numActive, const uint32_t* activeBricks, const uint16_t* coef, const float* v, int n,
                         float* vD, uint16_t* coefD, hipStream_t s) {
    const uint32_t cells = numActive * 512u;
    if (!cells) return;
    FLIP2_KERNEL(dense_vector_kernel, g1(cells, 256), dim3(256), 0, s, activeBricks, coef, v, cells, n, vD, coefD);
    FLIP2_LAUNCH_CHECK();
}

void launch_dense_system(uint32_t numActive, const uint32_t* activeBricks, const uint16_t* coef, const double* b,
                         const double* x, int n, float* bD, float* xD, uint16_t* coefD, hipStream_t s) {
    const uint32_t cells = numActive * 512u;
    if (!cells) return;
    FLIP2_KERNEL(dense_system_kernel, g1(cells, 256), dim3(256), 0, s, activeBricks, coef, b, x, cells, n, bD, xD,
                       coefD);
    FLIP2_LAUNCH_CHECK();
}

void launch_dense_apply_A(const uint16_t* coefD, const float* x, float* y, int n, hipStream_t s) {
    FLIP2_KERNEL(dense_apply_kernel, g1((long long)n * n * n, 256), dim3(256), 0, s, coefD, x, y, n);
    FLIP2_LAUNCH_CHECK();
}

void launch_residual(const uint16_t* coefD, const float* z, const float* bD, float* res, int n, double* red,
                     hipStream_t s) {
    FLIP2_KERNEL(zero_red_kernel, dim3(1), dim3(64), 0, s, red, 2);
    FLIP2_KERNEL(residual_kernel, red_grid((long long)n * n * n), dim3(256), 0, s, coefD, z, bD, res, n, red);
    FLIP2_LAUNCH_CHECK();
}

void launch_loss_grad(const uint16_t* coefD, const float* res, const double* red, float* dOut, int n, hipStream_t s) {
    FLIP2_KERNEL(loss_grad_kernel, g1((long long)n * n * n, 256), dim3(256), 0, s, coefD, res, red, dOut, n);
    FLIP2_LAUNCH_CHECK();
}

void launch_dense_input(const float* bD, const uint16_t* coefD, int n, uint16_t* in0, uint16_t* in0c, uint8_t* mask,
                        float* scale, double* red, hipStream_t s) {
    const long long nvox = (long long)n * n * n;
    FLIP2_KERNEL(zero_red_kernel, dim3(1), dim3(64), 0, s, red, 2);
    FLIP2_KERNEL(dense_norm_kernel, red_grid(nvox), dim3(256), 0, s, bD, coefD, nvox, red);
    FLIP2_KERNEL(dense_input_kernel, g1(nvox, 256), dim3(256), 0, s, bD, coefD, nvox, in0, in0c, mask, scale,
                       red);
    FLIP2_LAUNCH_CHECK();
}

}  // namespace nn

// ---------------------------------------------------------------------------
// UNet
// ---------------------------------------------------------------------------
namespace {
// SplitMix64 + Box-Muller for deterministic initialisation.
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() { return std::sqrt(-2.0 * std::log(uniform())) * std::cos(6.283185307179586 * uniform()); }
};
}  // namespace

UNet::UNet(int n, hipStream_t stream) : n_(n), n1_(n / 2), n2_(n / 4), stream_(stream) {
    FLIP2_REQUIRE(n >= 8 && n % 8 == 0, kErrArgument, "U-Net grid size must be a positive multiple of 8");
    // the largest activation (the level-0 decoder input, 48 channels per
    // voxel) must count its elements in 32 bits: n <= 440
    FLIP2_REQUIRE((uint64_t)n * n * n * 48u <= 0xFFFFFFFFull, kErrRange,
                  "U-Net grid size too large (at most 440: 48 n^3 activations must fit 32-bit counts)");
    const int spec[10][3] = {{16, 16, 0}, {16, 16, 0}, {16, 32, 1}, {32, 32, 1}, {32, 64, 2},
                             {64, 64, 2}, {96, 32, 1}, {32, 32, 1}, {48, 16, 0}, {16, 16, 0}};
    size_t off = 0;
    for (auto& l : spec) {
        UNetLayer L{l[0], l[1], l[2], off, 0};
        off += (size_t)27 * l[0] * l[1];
        L.b = off;
        off += (size_t)l[1];
        layers_.push_back(L);
    }
    headW_ = off;
    off += 16;
    headB_ = off;
    off += 1;
    numParams_ = off;
    for (auto* a : {&params_, &grad_, &adamM_, &adamV_}) a->resizeAsync((uint32_t)numParams_, stream_);
    paramsBf_.resizeAsync((uint32_t)numParams_, stream_);
    const long long N0 = (long long)n_ * n_ * n_, N1 = (long long)n1_ * n1_ * n1_, N2 = (long long)n2_ * n2_ * n2_;
    in0_.resizeAsync((uint32_t)(N0 * 16), stream_);
    FLIP2_HIP(hipMemsetAsync(in0_.devPtr(), 0, in0_.bytes(), stream_));
    in0c_.resizeAsync((uint32_t)(N0 * 2), stream_);
    FLIP2_HIP(hipMemsetAsync(in0c_.devPtr(), 0, in0c_.bytes(), stream_));
    a0_.resizeAsync((uint32_t)(N0 * 16), stream_);
    b0_.resizeAsync((uint32_t)(N0 * 16), stream_);
    p0_.resizeAsync((uint32_t)(N1 * 16), stream_);
    a1_.resizeAsync((uint32_t)(N1 * 32), stream_);
    b1_.resizeAsync((uint32_t)(N1 * 32), stream_);
    p1_.resizeAsync((uint32_t)(N2 * 32), stream_);
    a2_.resizeAsync((uint32_t)(N2 * 64), stream_);
    b2_.resizeAsync((uint32_t)(N2 * 64), stream_);
    u1_.resizeAsync((uint32_t)(N1 * 96), stream_);
    c1_.resizeAsync((uint32_t)(N1 * 32), stream_);
    d1_.resizeAsync((uint32_t)(N1 * 32), stream_);
    u0_.resizeAsync((uint32_t)(N0 * 48), stream_);
    c0_.resizeAsync((uint32_t)(N0 * 16), stream_);
    d0_.resizeAsync((uint32_t)(N0 * 16), stream_);
    mask_.resizeAsync((uint32_t)N0, stream_);
    mask1_.resizeAsync((uint32_t)N1, stream_);
    mask2_.resizeAsync((uint32_t)N2, stream_);
    const int levelN[3] = {n_, n1_, n2_};
    for (int l = 0; l < 3; ++l) {
        const uint32_t tiles = nn::conv_tiles(levelN[l]);
        for (auto* f : {&tileFlags_[0][l], &tileFlags_[1][l], &tileOnes_[l]}) f->resizeAsync(tiles, stream_);
        FLIP2_HIP(hipMemsetAsync(tileOnes_[l].devPtr(), 1, tiles, stream_));
        listCompute_[l].resizeAsync(tiles, stream_);
        listClear_[l].resizeAsync(tiles, stream_);
        listCounts_[l].resizeAsync(2, stream_);
        nn::conv_tile_dims(levelN[l], tl_[l].tilesX, tl_[l].tilesY);
        tl_[l].compute = listCompute_[l].devPtr();
        tl_[l].clear = listClear_[l].devPtr();
        tl_[l].numCompute = listCounts_[l].devPtr();
        tl_[l].numClear = listCounts_[l].devPtr() + 1;
    }
    scale_.resizeAsync(4, stream_);
    out_.resizeAsync((uint32_t)N0, stream_);
    red_.resizeAsync(8 + 2 * nn::kRedBlocks, stream_);
    initParams(1234);
}

UNet::~UNet() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    else (void)hipDeviceSynchronize();
}

void UNet::initParams(uint64_t seed) {
    std::vector<float> p(numParams_, 0.f);
    Rng rng{seed};
    for (const auto& L : layers_) {
        const double std = std::sqrt(2.0 / (27.0 * L.cin));
        for (size_t i = 0; i < (size_t)27 * L.cin * L.cout; ++i) p[L.w + i] = (float)(std * rng.normal());
    }
    for (int c = 0; c < 16; ++c) p[headW_ + c] = (float)(0.1 * rng.normal());
    setParams(p.data(), p.size());
    step_ = 0;
    FLIP2_HIP(hipMemsetAsync(adamM_.devPtr(), 0, adamM_.bytes(), stream_));
    FLIP2_HIP(hipMemsetAsync(adamV_.devPtr(), 0, adamV_.bytes(), stream_));
}

void UNet::setParams(const float* host, size_t count) {
    FLIP2_REQUIRE(count == numParams_, kErrRange, "parameter count mismatch");
    params_.uploadFrom(host, (uint32_t)count, stream_);
    nn::launch_to_bf16(params_.devPtr(), paramsBf_.devPtr(), (long long)numParams_, stream_);
    FLIP2_HIP(hipStreamSynchronize(stream_));
}

void UNet::getParams(float* host, size_t count) {
    FLIP2_REQUIRE(count >= numParams_, kErrRange, "buffer too small");
    params_.downloadTo(host, (uint32_t)numParams_, stream_);
    FLIP2_HIP(hipStreamSynchronize(stream_));
}

void UNet::conv(int layer, const uint16_t* X, uint16_t* Y, const nn::ConvMask& cm, hipStream_t s) {
    const UNetLayer& L = layers_[layer];
    const int n = L.level == 0 ? n_ : (L.level == 1 ? n1_ : n2_);
    nn::launch_conv3_fwd(X, paramsBf_.devPtr() + L.w, params_.devPtr() + L.b, Y, n, L.cin, L.cout, true, cm, s);
}

void UNet::setInputMasks(hipStream_t s, bool sparse) {
    // Level masks (every activation is zero outside its level's mask).  A
    // sparse forward computes only the tiles that hold masked-in voxels; a
    // tile computed by an earlier forward and not by this one is cleared.
    nn::launch_mask_pool(mask_.devPtr(), n_, mask1_.devPtr(), s);
    nn::launch_mask_pool(mask1_.devPtr(), n1_, mask2_.devPtr(), s);
    const uint8_t* lm[3] = {mask_.devPtr(), mask1_.devPtr(), mask2_.devPtr()};
    const int levelN[3] = {n_, n1_, n2_};
    if (!sparse) {
        for (int l = 0; l < 3; ++l) cm_[l] = nn::ConvMask{lm[l], nullptr, nullptr};
        return;
    }
    flagSet_ ^= 1;
    for (int l = 0; l < 3; ++l) {
        nn::launch_tile_flags(lm[l], levelN[l], tileFlags_[flagSet_][l].devPtr(), s);
        const uint8_t* prev = tilesDirty_ ? tileOnes_[l].devPtr() : tileFlags_[flagSet_ ^ 1][l].devPtr();
        cm_[l] = nn::ConvMask{lm[l], tileFlags_[flagSet_][l].devPtr(), prev};
        nn::launch_tile_lists(tileFlags_[flagSet_][l].devPtr(), prev, nn::conv_tiles(levelN[l]),
                              listCompute_[l].devPtr(), listClear_[l].devPtr(), listCounts_[l].devPtr(), s);
    }
    tilesDirty_ = false;
}

void UNet::forward(hipStream_t s, bool keepActivations) {
    FLIP2_REQUIRE(!keepActivations || !cm_[0].tiles, kErrState, "U-Net: stored activations need a dense forward");
    const nn::ConvMask* cm = cm_;
    if (!cm[0].tiles) tilesDirty_ = true;  // a dense forward writes every tile
    // the first layer reads the two non-zero input channels only
    if (cm[0].tiles) {
        // The apply.  Every convolution runs on a fixed grid of workgroups
        // over the list of its level's tiles with fluid (1, 2 or 4
        // workgroups per tile on levels 0, 1, 2: output channels split), the
        // next step's input halo and weights prefetched during the current
        // step's MFMAs.  The decoder's concatenated inputs are read in place,
        // and the output head runs in the last convolution.
        nn::launch_conv3_in2_list(in0c_.devPtr(), slotIn_, paramsBf_.devPtr() + layers_[0].w,
                                  params_.devPtr() + layers_[0].b, a0_.devPtr(), n_, true, cm[0].voxel, tl_[0], s);
        const int levelN[3] = {n_, n1_, n2_};
        const int split[3] = {1, 2, 4};
        auto lconv = [&](int layer, const uint16_t* X, const uint16_t* Xc, int ca, uint16_t* Y, uint16_t* pooled,
                         const nn::HeadArgs& h) {
            const UNetLayer& L = layers_[layer];
            nn::launch_conv3_list(X, Xc, paramsBf_.devPtr() + L.w, params_.devPtr() + L.b, Y, pooled, levelN[L.level],
                                  L.cin, ca, L.cout, split[L.level], true, h, cm[L.level].voxel, tl_[L.level], s);
        };
        const nn::HeadArgs none{};
        // the encoder's second convolution on levels 0 and 1 also writes the
        // pooled input of the next level
        lconv(1, a0_.devPtr(), nullptr, 0, b0_.devPtr(), p0_.devPtr(), none);
        lconv(2, p0_.devPtr(), nullptr, 0, a1_.devPtr(), nullptr, none);
        lconv(3, a1_.devPtr(), nullptr, 0, b1_.devPtr(), p1_.devPtr(), none);
        lconv(4, p1_.devPtr(), nullptr, 0, a2_.devPtr(), nullptr, none);
        lconv(5, a2_.devPtr(), nullptr, 0, b2_.devPtr(), nullptr, none);
        lconv(6, b1_.devPtr(), b2_.devPtr(), 64, c1_.devPtr(), nullptr, none);
        lconv(7, c1_.devPtr(), nullptr, 0, d1_.devPtr(), nullptr, none);
        lconv(8, b0_.devPtr(), d1_.devPtr(), 32, c0_.devPtr(), nullptr, none);
        const nn::HeadArgs head{params_.devPtr() + headW_, params_.devPtr() + headB_, mask_.devPtr(), scale_.devPtr(),
                                slotOut_ ? nullptr : out_.devPtr(), slotOut_, slotIn_.brickSlot};
        lconv(9, c0_.devPtr(), nullptr, 0, nullptr, nullptr, head);
        return;
    }
    nn::launch_conv3_in2(in0c_.devPtr(), paramsBf_.devPtr() + layers_[0].w, params_.devPtr() + layers_[0].b,
                         a0_.devPtr(), n_, true, cm[0], s);
    conv(1, a0_.devPtr(), b0_.devPtr(), cm[0], s);
    nn::launch_avgpool(b0_.devPtr(), p0_.devPtr(), n_, 16, s);
    conv(2, p0_.devPtr(), a1_.devPtr(), cm[1], s);
    conv(3, a1_.devPtr(), b1_.devPtr(), cm[1], s);
    nn::launch_avgpool(b1_.devPtr(), p1_.devPtr(), n1_, 32, s);
    conv(4, p1_.devPtr(), a2_.devPtr(), cm[2], s);
    conv(5, a2_.devPtr(), b2_.devPtr(), cm[2], s);
    if (!keepActivations) {
        const UNetLayer& L6 = layers_[6];
        const UNetLayer& L8 = layers_[8];
        const UNetLayer& L9 = layers_[9];
        nn::launch_conv3_fwd_upcat(b1_.devPtr(), b2_.devPtr(), paramsBf_.devPtr() + L6.w, params_.devPtr() + L6.b,
                                   c1_.devPtr(), n1_, 64, 32, 32, true, cm[1], s);
        conv(7, c1_.devPtr(), d1_.devPtr(), cm[1], s);
        nn::launch_conv3_fwd_upcat(b0_.devPtr(), d1_.devPtr(), paramsBf_.devPtr() + L8.w, params_.devPtr() + L8.b,
                                   c0_.devPtr(), n_, 32, 16, 16, true, cm[0], s);
        const nn::HeadArgs head{params_.devPtr() + headW_, params_.devPtr() + headB_, mask_.devPtr(), scale_.devPtr(),
                                out_.devPtr()};
        nn::launch_conv3_fwd_head(c0_.devPtr(), paramsBf_.devPtr() + L9.w, params_.devPtr() + L9.b, nullptr, n_, 16,
                                  true, head, cm[0], s);
        return;
    }
    nn::launch_upcat(b2_.devPtr(), b1_.devPtr(), u1_.devPtr(), n1_, 64, 32, s);
    conv(6, u1_.devPtr(), c1_.devPtr(), cm[1], s);
    conv(7, c1_.devPtr(), d1_.devPtr(), cm[1], s);
    nn::launch_upcat(d1_.devPtr(), b0_.devPtr(), u0_.devPtr(), n_, 32, 16, s);
    conv(8, u0_.devPtr(), c0_.devPtr(), cm[0], s);
    conv(9, c0_.devPtr(), d0_.devPtr(), cm[0], s);
    nn::launch_head_fwd(d0_.devPtr(), params_.devPtr() + headW_, params_.devPtr() + headB_, mask_.devPtr(),
                        scale_.devPtr(), out_.devPtr(), (long long)n_ * n_ * n_, 16, s);
}

// dY: gradient w.r.t. this layer's (post-ReLU) output Y; turned into the
// pre-activation gradient in place; weight/bias gradients accumulated; dX
// (if not null) receives the gradient w.r.t. the layer input X.
void UNet::convBackward(int layer, float* dY, const uint16_t* Y, const uint16_t* X, float* dX, hipStream_t s) {
    const UNetLayer& L = layers_[layer];
    const int n = L.level == 0 ? n_ : (L.level == 1 ? n1_ : n2_);
    const long long nvox = (long long)n * n * n;
    if (dPb_.size() < (uint32_t)(nvox * L.cout)) dPb_.resizeAsync((uint32_t)(nvox * L.cout), s);
    if (WT_.size() < (uint32_t)(27 * L.cin * L.cout)) WT_.resizeAsync((uint32_t)(27 * L.cin * L.cout), s);
    nn::launch_relu_mask_bf16(dY, Y, dPb_.devPtr(), nvox * L.cout, s);
    nn::launch_conv3_wgrad(dPb_.devPtr(), dY, X, grad_.devPtr() + L.w, grad_.devPtr() + L.b, n, L.cin, L.cout, s);
    if (dX) nn::launch_conv3_dgrad(dPb_.devPtr(), params_.devPtr() + L.w, WT_.devPtr(), dX, n, L.cin, L.cout, s);
}

void UNet::backward(const float* dOut, hipStream_t s) {
    const long long N0 = (long long)n_ * n_ * n_, N1 = (long long)n1_ * n1_ * n1_, N2 = (long long)n2_ * n2_ * n2_;
    auto ensure = [&](DeviceVec<float>& v, long long count) {
        if (v.size() < count) v.resizeAsync((uint32_t)count, s);
    };
    ensure(gd0_, N0 * 16); ensure(gc0_, N0 * 16); ensure(gu0_, N0 * 48); ensure(gb0_, N0 * 16); ensure(ga0_, N0 * 16);
    ensure(gd1_, N1 * 32); ensure(gc1_, N1 * 32); ensure(gu1_, N1 * 96); ensure(gb1_, N1 * 32); ensure(ga1_, N1 * 32);
    ensure(gp0_, N1 * 16); ensure(gb2_, N2 * 64); ensure(ga2_, N2 * 64); ensure(gp1_, N2 * 32);

    nn::launch_head_bwd(dOut, d0_.devPtr(), params_.devPtr() + headW_, mask_.devPtr(), scale_.devPtr(), gd0_.devPtr(),
                        grad_.devPtr() + headW_, grad_.devPtr() + headB_, N0, 16, s);
    convBackward(9, gd0_.devPtr(), d0_.devPtr(), c0_.devPtr(), gc0_.devPtr(), s);
    convBackward(8, gc0_.devPtr(), c0_.devPtr(), u0_.devPtr(), gu0_.devPtr(), s);
    nn::launch_zero_f32(gd1_.devPtr(), N1 * 32, s);
    nn::launch_zero_f32(gb0_.devPtr(), N0 * 16, s);
    nn::launch_upcat_bwd(gu0_.devPtr(), gd1_.devPtr(), gb0_.devPtr(), n_, 32, 16, s);
    convBackward(7, gd1_.devPtr(), d1_.devPtr(), c1_.devPtr(), gc1_.devPtr(), s);
    convBackward(6, gc1_.devPtr(), c1_.devPtr(), u1_.devPtr(), gu1_.devPtr(), s);
    nn::launch_zero_f32(gb2_.devPtr(), N2 * 64, s);
    nn::launch_zero_f32(gb1_.devPtr(), N1 * 32, s);
    nn::launch_upcat_bwd(gu1_.devPtr(), gb2_.devPtr(), gb1_.devPtr(), n1_, 64, 32, s);
    convBackward(5, gb2_.devPtr(), b2_.devPtr(), a2_.devPtr(), ga2_.devPtr(), s);
    convBackward(4, ga2_.devPtr(), a2_.devPtr(), p1_.devPtr(), gp1_.devPtr(), s);
    nn::launch_avgpool_bwd(gp1_.devPtr(), gb1_.devPtr(), n1_, 32, s);
    convBackward(3, gb1_.devPtr(), b1_.devPtr(), a1_.devPtr(), ga1_.devPtr(), s);
    convBackward(2, ga1_.devPtr(), a1_.devPtr(), p0_.devPtr(), gp0_.devPtr(), s);
    nn::launch_avgpool_bwd(gp0_.devPtr(), gb0_.devPtr(), n_, 16, s);
    convBackward(1, gb0_.devPtr(), b0_.devPtr(), a0_.devPtr(), ga0_.devPtr(), s);
    convBackward(0, ga0_.devPtr(), a0_.devPtr(), in0_.devPtr(), nullptr, s);
}

void UNet::zeroGrad(hipStream_t s) { nn::launch_zero_f32(grad_.devPtr(), (long long)numParams_, s); }

void UNet::adamStep(float lr, hipStream_t s) {
    ++step_;
    nn::launch_adam(params_.devPtr(), grad_.devPtr(), adamM_.devPtr(), adamV_.devPtr(), paramsBf_.devPtr(),
                    (int)numParams_, lr, 0.9f, 0.999f, 1e-8f, step_, s);
}

void UNet::applyActive(uint32_t numActive, const uint32_t* activeBricks, const int32_t* brickSlot,
                       const uint8_t* cellType, const float* r, float* z, int gridN, bool newSystem,
                       const double* normPairs, hipStream_t s, const uint32_t* stop) {
    FLIP2_REQUIRE(gridN == n_ && n_ % 8 == 0, kErrArgument, "U-Net grid size differs from the simulation grid");
    if (newSystem) {
        // the dense fluid mask (level masks, tile lists, the head's mask)
        nn::launch_scatter_input(numActive, activeBricks, cellType, r, n_, in0c_.devPtr(), mask_.devPtr(),
                                 scale_.devPtr(), red_.devPtr(), true, s);
        setInputMasks(s, true);
        nn::launch_zero_f32(z, (long long)numActive * 512, s);
    } else if (normPairs) {
        nn::launch_pairs_scale(normPairs, numActive, red_.devPtr(), scale_.devPtr(), s, stop);
    } else {
        nn::launch_input_scale(numActive, cellType, r, red_.devPtr(), scale_.devPtr(), s);
    }
    slotIn_ = nn::SlotInput{r, cellType, brickSlot, scale_.devPtr()};
    slotOut_ = z;
    for (auto& t : tl_) t.stop = stop;
    auto reset = [&] {
        slotIn_ = nn::SlotInput{};
        slotOut_ = nullptr;
        for (auto& t : tl_) t.stop = nullptr;
    };
    try {
        forward(s, false);
    } catch (...) {
        reset();
        throw;
    }
    reset();
}

}  // namespace flip2
