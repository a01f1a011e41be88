// Kernels of the learned preconditioner: a 3-level 3D U-Net on a dense
// n^3 grid, channels-last activations ([z][y][x][C]).
//
// Forward convolutions are implicit GEMMs on MFMA (v_mfma_f32_16x16x16_bf16):
// BF16 weights and activations, FP32 accumulation, bias + ReLU fused in the
// epilogue.  D[cout][voxel] = sum_{tap,cin} W[tap][cout][cin] X[voxel+tap][cin]
// with M = 16 output channels, N = 16 voxels, K = 16 input channels per
// instruction.  A wave owns 64 consecutive voxels and all output channels.
//
// Backward kernels (data and weight gradients) run in FP32 on the vector
// ALUs; the weight gradient reduces over voxels with one FP32 atomic add per
// weight per workgroup.
#include "unet_kernels.hpp"

namespace flip2 {
namespace nn {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---------------------------------------------------------------------------
// conv3x3x3 forward, zero padding
// ---------------------------------------------------------------------------
template <int CIN, int COUT>
__global__ void __launch_bounds__(256)
conv3_fwd_kernel(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Wt, const float* __restrict__ bias,
                 uint16_t* __restrict__ Y, int n, int relu) {
    constexpr int MB = COUT / 16;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long nvox = (long long)n * n * n;
    const long long v0 = ((long long)blockIdx.x * 4 + wave) * 64;
    if (v0 >= nvox) return;
    const int col = lane & 15, kq = lane >> 4;
    int vx[4], vy[4], vz[4];
    bool valid[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long long v = v0 + 16 * j + col;
        valid[j] = v < nvox;
        const long long vv = valid[j] ? v : 0;
        vx[j] = (int)(vv % n);
        vy[j] = (int)((vv / n) % n);
        vz[j] = (int)(vv / ((long long)n * n));
    }
    v4f acc[MB][4];
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[m][j] = v4f{0.f, 0.f, 0.f, 0.f};

    for (int tap = 0; tap < 27; ++tap) {
        const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
        long long src[4];
        bool inb[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int xs = vx[j] + dx, ys = vy[j] + dy, zs = vz[j] + dz;
            inb[j] = valid[j] && xs >= 0 && ys >= 0 && zs >= 0 && xs < n && ys < n && zs < n;
            src[j] = inb[j] ? (((long long)zs * n + ys) * n + xs) * CIN : 0;
        }
#pragma unroll
        for (int c0 = 0; c0 < CIN; c0 += 16) {
            v4s b[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (inb[j]) {
                    const uint2 raw = *reinterpret_cast<const uint2*>(X + src[j] + c0 + 4 * kq);
                    b[j] = __builtin_bit_cast(v4s, raw);
                } else {
                    b[j] = v4s{0, 0, 0, 0};
                }
            }
#pragma unroll
            for (int m = 0; m < MB; ++m) {
                const uint2 raw =
                    *reinterpret_cast<const uint2*>(Wt + ((size_t)(tap * COUT + 16 * m + col) * CIN) + c0 + 4 * kq);
                const v4s a = __builtin_bit_cast(v4s, raw);
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[j], acc[m][j], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int m = 0; m < MB; ++m) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (!valid[j]) continue;
            const long long v = v0 + 16 * j + col;
            const int co = 16 * m + 4 * kq;
            uint16_t o[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float y = acc[m][j][r] + bias[co + r];
                if (relu) y = fmaxf(y, 0.f);
                o[r] = f2bf(y);
            }
            uint2 packed;
            packed.x = (uint32_t)o[0] | ((uint32_t)o[1] << 16);
            packed.y = (uint32_t)o[2] | ((uint32_t)o[3] << 16);
            *reinterpret_cast<uint2*>(Y + v * COUT + co) = packed;
        }
    }
}

// ---------------------------------------------------------------------------
// FP32 backward of conv3x3x3
// ---------------------------------------------------------------------------

// dX[v][ci] = sum_tap sum_co W[tap][co][ci] * dP[v - off(tap)][co]
// (dP = gradient w.r.t. the pre-activation of this layer's output).
// One thread per (voxel, ci); weights of the workgroup's channel staged in LDS.
__global__ void __launch_bounds__(256)
conv3_dgrad_kernel(const float* __restrict__ dP, const float* __restrict__ W32, float* __restrict__ dX, int n, int cin,
                   int cout) {
    extern __shared__ float s_w[];  // [27][cout] for this block's ci
    const int ci = blockIdx.y;
    for (int t = threadIdx.x; t < 27 * cout; t += blockDim.x) {
        const int tap = t / cout, co = t % cout;
        s_w[t] = W32[((size_t)tap * cout + co) * cin + ci];
    }
    __syncthreads();
    const long long nvox = (long long)n * n * n;
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((long long)n * n));
    float acc = 0.f;
    for (int tap = 0; tap < 27; ++tap) {
        const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
        // output voxel u with u + off = v  ->  u = v - off
        const int xo = x - dx, yo = y - dy, zo = z - dz;
        if (xo < 0 || yo < 0 || zo < 0 || xo >= n || yo >= n || zo >= n) continue;
        const float* g = dP + (((long long)zo * n + yo) * n + xo) * cout;
        const float* w = s_w + tap * cout;
        for (int co = 0; co < cout; ++co) acc = fmaf(w[co], g[co], acc);
    }
    dX[v * cin + ci] = acc;
}

// dW[tap][co][ci] += sum_v dP[v][co] * X[v + off(tap)][ci]; dB[co] += sum_v dP.
// Workgroup: one tap, a chunk of voxels; thread t handles (co, ci) pairs
// t, t+256, ...; voxels staged through LDS in batches of 16.
__global__ void __launch_bounds__(256)
conv3_wgrad_kernel(const float* __restrict__ dP, const uint16_t* __restrict__ X, float* __restrict__ dW,
                   float* __restrict__ dB, int n, int cin, int cout, int chunk) {
    extern __shared__ float s_mem[];
    constexpr int B = 16;
    float* s_g = s_mem;              // [B][cout]
    float* s_x = s_mem + B * cout;   // [B][cin]
    const int tap = blockIdx.y;
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    const long long nvox = (long long)n * n * n;
    const long long vBeg = (long long)blockIdx.x * chunk;
    const long long vEnd = vBeg + chunk < nvox ? vBeg + chunk : nvox;
    const int pairs = cin * cout;
    constexpr int MAXP = 24;  // pairs per thread (cin*cout <= 6144)
    float acc[MAXP];
#pragma unroll
    for (int k = 0; k < MAXP; ++k) acc[k] = 0.f;
    float bacc = 0.f;
    for (long long vb = vBeg; vb < vEnd; vb += B) {
        __syncthreads();
        for (int t = threadIdx.x; t < B * cout; t += blockDim.x) {
            const long long v = vb + t / cout;
            s_g[t] = v < vEnd ? dP[v * cout + t % cout] : 0.f;
        }
        for (int t = threadIdx.x; t < B * cin; t += blockDim.x) {
            const long long v = vb + t / cin;
            float xv = 0.f;
            if (v < vEnd) {
                const int x = (int)(v % n) + dx, y = (int)((v / n) % n) + dy, z = (int)(v / ((long long)n * n)) + dz;
                if (x >= 0 && y >= 0 && z >= 0 && x < n && y < n && z < n)
                    xv = bf2f(X[(((long long)z * n + y) * n + x) * cin + t % cin]);
            }
            s_x[t] = xv;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < MAXP; ++k) {
            const int pidx = threadIdx.x + k * 256;
            if (pidx >= pairs) break;
            const int co = pidx / cin, ci = pidx % cin;
            float a = acc[k];
            for (int b = 0; b < B; ++b) a = fmaf(s_g[b * cout + co], s_x[b * cin + ci], a);
            acc[k] = a;
        }
        if (tap == 13 && (int)threadIdx.x < cout)
            for (int b = 0; b < B; ++b) bacc += s_g[b * cout + threadIdx.x];
    }
#pragma unroll
    for (int k = 0; k < MAXP; ++k) {
        const int pidx = threadIdx.x + k * 256;
        if (pidx >= pairs) break;
        const int co = pidx / cin, ci = pidx % cin;
        atomicAdd(&dW[((size_t)tap * cout + co) * cin + ci], acc[k]);
    }
    if (tap == 13 && (int)threadIdx.x < cout) atomicAdd(&dB[threadIdx.x], bacc);
}

// g <- g * (act > 0)  (ReLU derivative from the stored BF16 activation)
__global__ void relu_mask_kernel(float* __restrict__ g, const uint16_t* __restrict__ act, long long count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    if (!(bf2f(act[i]) > 0.f)) g[i] = 0.f;
}

// ---------------------------------------------------------------------------
// Pooling, upsampling + concatenation, input assembly, 1x1 head
// ---------------------------------------------------------------------------

// P[vp][c] = mean of the 8 children (n = fine size).
__global__ void avgpool_kernel(const uint16_t* __restrict__ X, uint16_t* __restrict__ P, int n, int C) {
    const int nc = n / 2;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)nc * nc * nc * C;
    if (i >= total) return;
    const int c = (int)(i % C);
    const long long vp = i / C;
    const int x = (int)(vp % nc), y = (int)((vp / nc) % nc), z = (int)(vp / ((long long)nc * nc));
    float s = 0.f;
    for (int k = 0; k < 8; ++k) {
        const int xf = 2 * x + (k & 1), yf = 2 * y + ((k >> 1) & 1), zf = 2 * z + (k >> 2);
        s += bf2f(X[(((long long)zf * n + yf) * n + xf) * C + c]);
    }
    P[i] = f2bf(s * 0.125f);
}

__global__ void avgpool_bwd_kernel(const float* __restrict__ dP, float* __restrict__ dX, int n, int C) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)n * n * n * C;
    if (i >= total) return;
    const int c = (int)(i % C);
    const long long v = i / C;
    const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((long long)n * n));
    const int nc = n / 2;
    dX[i] += 0.125f * dP[(((long long)(z / 2) * nc + y / 2) * nc + x / 2) * C + c];
}

// U[v][0:Ca] = A[parent(v)] (nearest upsampling), U[v][Ca:Ca+Cb] = B[v].
__global__ void upcat_kernel(const uint16_t* __restrict__ A, const uint16_t* __restrict__ Bsk, uint16_t* __restrict__ U,
                             int n, int Ca, int Cb) {
    const int C = Ca + Cb;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)n * n * n * C;
    if (i >= total) return;
    const int c = (int)(i % C);
    const long long v = i / C;
    if (c < Ca) {
        const int x = (int)(v % n), y = (int)((v / n) % n), z = (int)(v / ((long long)n * n));
        const int nc = n / 2;
        U[i] = A[(((long long)(z / 2) * nc + y / 2) * nc + x / 2) * Ca + c];
    } else {
        U[i] = Bsk[v * Cb + (c - Ca)];
    }
}

// dA[parent][c] += sum of the 8 children's dU[child][c]; dB[v][c] += dU[v][Ca+c]
__global__ void upcat_bwd_kernel(const float* __restrict__ dU, float* __restrict__ dA, float* __restrict__ dBsk, int n,
                                 int Ca, int Cb) {
    const int C = Ca + Cb;
    const int nc = n / 2;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long coarse = (long long)nc * nc * nc * Ca;
    const long long fine = (long long)n * n * n * Cb;
    if (i < coarse) {
        const int c = (int)(i % Ca);
        const long long vp = i / Ca;
        const int x = (int)(vp % nc), y = (int)((vp / nc) % nc), z = (int)(vp / ((long long)nc * nc));
        float s = 0.f;
        for (int k = 0; k < 8; ++k) {
            const int xf = 2 * x + (k & 1), yf = 2 * y + ((k >> 1) & 1), zf = 2 * z + (k >> 2);
            s += dU[(((long long)zf * n + yf) * n + xf) * C + c];
        }
        dA[i] += s;
    } else if (i < coarse + fine) {
        const long long j = i - coarse;
        const int c = (int)(j % Cb);
        const long long v = j / Cb;
        dBsk[j] += dU[v * C + Ca + c];
    }
}

// out[v] = (sum_c X[v][c] w[c] + b) * scale * mask[v]
__global__ void head_fwd_kernel(const uint16_t* __restrict__ X, const float* __restrict__ w, const float* __restrict__ b,
                                const uint8_t* __restrict__ mask, const float* __restrict__ scale, float* __restrict__ out,
                                long long nvox, int C) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    float s = b[0];
    for (int c = 0; c < C; ++c) s = fmaf(bf2f(X[v * C + c]), w[c], s);
    out[v] = mask[v] ? s * scale[0] : 0.f;
}

// dX[v][c] = dOut_eff[v] w[c]; dw[c] += sum_v dOut_eff X[v][c]; db += sum dOut_eff
// with dOut_eff = dOut * scale * mask.  One workgroup reduces a chunk.
__global__ void __launch_bounds__(256)
head_bwd_kernel(const float* __restrict__ dOut, const uint16_t* __restrict__ X, const float* __restrict__ w,
                const uint8_t* __restrict__ mask, const float* __restrict__ scale, float* __restrict__ dX,
                float* __restrict__ dw, float* __restrict__ db, long long nvox, int C) {
    __shared__ float s_red[17][256];
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float g = 0.f;
    float part[16];
    for (int c = 0; c < 16; ++c) part[c] = 0.f;
    if (v < nvox) {
        g = mask[v] ? dOut[v] * scale[0] : 0.f;
        for (int c = 0; c < C; ++c) {
            dX[v * C + c] = g * w[c];
            part[c] = g * bf2f(X[v * C + c]);
        }
    }
    for (int c = 0; c < 16; ++c) s_red[c][threadIdx.x] = part[c];
    s_red[16][threadIdx.x] = g;
    __syncthreads();
    for (int s = 128; s >= 1; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 17; ++c) s_red[c][threadIdx.x] += s_red[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < (unsigned)C) atomicAdd(&dw[threadIdx.x], s_red[threadIdx.x][0]);
    if (threadIdx.x == 16) atomicAdd(db, s_red[16][0]);
}

// Adam on FP32 master weights, BF16 copy refreshed.
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, uint16_t* __restrict__ wbf, int count, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wn = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wn;
    if (wbf) wbf[i] = f2bf(wn);
}

__global__ void to_bf16_kernel(const float* __restrict__ a, uint16_t* __restrict__ b, long long count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) b[i] = f2bf(a[i]);
}

__global__ void bf16_to_f32_kernel(const uint16_t* __restrict__ a, float* __restrict__ b, long long count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) b[i] = bf2f(a[i]);
}

__global__ void zero_f32_kernel(float* __restrict__ a, long long count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = 0.f;
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
namespace {
inline dim3 g1(long long n, int bs) { return dim3((unsigned)((n + bs - 1) / bs)); }
}

void launch_conv3_fwd(const uint16_t* X, const uint16_t* W, const float* bias, uint16_t* Y, int n, int cin, int cout,
                      bool relu, hipStream_t s) {
    const long long nvox = (long long)n * n * n;
    const dim3 grid((unsigned)((nvox + 255) / 256));
#define CONV_CASE(CI, CO)                                                                                  \
    if (cin == CI && cout == CO) {                                                                         \
        hipLaunchKernelGGL((conv3_fwd_kernel<CI, CO>), grid, dim3(256), 0, s, X, W, bias, Y, n, relu ? 1 : 0); \
        FLIP2_LAUNCH_CHECK();                                                                              \
        return;                                                                                            \
    }
    CONV_CASE(16, 16)
    CONV_CASE(16, 32)
    CONV_CASE(32, 32)
    CONV_CASE(32, 64)
    CONV_CASE(64, 64)
    CONV_CASE(96, 32)
    CONV_CASE(48, 16)
    CONV_CASE(32, 16)
    CONV_CASE(64, 32)
#undef CONV_CASE
    throw Error(kErrArgument, "conv3: unsupported channel pair");
}

void launch_conv3_dgrad(const float* dP, const float* W32, float* dX, int n, int cin, int cout, hipStream_t s) {
    const long long nvox = (long long)n * n * n;
    hipLaunchKernelGGL(conv3_dgrad_kernel, dim3((unsigned)((nvox + 255) / 256), cin), dim3(256),
                       27 * cout * sizeof(float), s, dP, W32, dX, n, cin, cout);
    FLIP2_LAUNCH_CHECK();
}

void launch_conv3_wgrad(const float* dP, const uint16_t* X, float* dW, float* dB, int n, int cin, int cout,
                        hipStream_t s) {
    FLIP2_REQUIRE(cin * cout <= 24 * 256, kErrArgument, "wgrad: too many channel pairs");
    const long long nvox = (long long)n * n * n;
    const int chunk = nvox >= (1 << 20) ? 8192 : (nvox >= 4096 ? 1024 : 64);
    const unsigned blocks = (unsigned)((nvox + chunk - 1) / chunk);
    hipLaunchKernelGGL(conv3_wgrad_kernel, dim3(blocks, 27), dim3(256), 16 * (cin + cout) * sizeof(float), s, dP, X, dW,
                       dB, n, cin, cout, chunk);
    FLIP2_LAUNCH_CHECK();
}

void launch_relu_mask(float* g, const uint16_t* act, long long count, hipStream_t s) {
    hipLaunchKernelGGL(relu_mask_kernel, g1(count, 256), dim3(256), 0, s, g, act, count);
    FLIP2_LAUNCH_CHECK();
}
void launch_avgpool(const uint16_t* X, uint16_t* P, int n, int C, hipStream_t s) {
    const long long total = (long long)(n / 2) * (n / 2) * (n / 2) * C;
    hipLaunchKernelGGL(avgpool_kernel, g1(total, 256), dim3(256), 0, s, X, P, n, C);
    FLIP2_LAUNCH_CHECK();
}
void launch_avgpool_bwd(const float* dP, float* dX, int n, int C, hipStream_t s) {
    const long long total = (long long)n * n * n * C;
    hipLaunchKernelGGL(avgpool_bwd_kernel, g1(total, 256), dim3(256), 0, s, dP, dX, n, C);
    FLIP2_LAUNCH_CHECK();
}
void launch_upcat(const uint16_t* A, const uint16_t* B, uint16_t* U, int n, int Ca, int Cb, hipStream_t s) {
    const long long total = (long long)n * n * n * (Ca + Cb);
    hipLaunchKernelGGL(upcat_kernel, g1(total, 256), dim3(256), 0, s, A, B, U, n, Ca, Cb);
    FLIP2_LAUNCH_CHECK();
}
void launch_upcat_bwd(const float* dU, float* dA, float* dB, int n, int Ca, int Cb, hipStream_t s) {
    const long long total = (long long)(n / 2) * (n / 2) * (n / 2) * Ca + (long long)n * n * n * Cb;
    hipLaunchKernelGGL(upcat_bwd_kernel, g1(total, 256), dim3(256), 0, s, dU, dA, dB, n, Ca, Cb);
    FLIP2_LAUNCH_CHECK();
}
void launch_head_fwd(const uint16_t* X, const float* w, const float* b, const uint8_t* mask, const float* scale,
                     float* out, long long nvox, int C, hipStream_t s) {
    hipLaunchKernelGGL(head_fwd_kernel, g1(nvox, 256), dim3(256), 0, s, X, w, b, mask, scale, out, nvox, C);
    FLIP2_LAUNCH_CHECK();
}
void launch_head_bwd(const float* dOut, const uint16_t* X, const float* w, const uint8_t* mask, const float* scale,
                     float* dX, float* dw, float* db, long long nvox, int C, hipStream_t s) {
    FLIP2_REQUIRE(C <= 16, kErrArgument, "head: at most 16 channels");
    hipLaunchKernelGGL(head_bwd_kernel, g1(nvox, 256), dim3(256), 0, s, dOut, X, w, mask, scale, dX, dw, db, nvox, C);
    FLIP2_LAUNCH_CHECK();
}
void launch_adam(float* w, const float* g, float* m, float* v, uint16_t* wbf, int count, float lr, float b1, float b2,
                 float eps, int t, hipStream_t s) {
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    hipLaunchKernelGGL(adam_kernel, g1(count, 256), dim3(256), 0, s, w, g, m, v, wbf, count, lr, b1, b2, eps, c1, c2);
    FLIP2_LAUNCH_CHECK();
}
void launch_to_bf16(const float* a, uint16_t* b, long long count, hipStream_t s) {
    hipLaunchKernelGGL(to_bf16_kernel, g1(count, 256), dim3(256), 0, s, a, b, count);
    FLIP2_LAUNCH_CHECK();
}
void launch_bf16_to_f32(const uint16_t* a, float* b, long long count, hipStream_t s) {
    hipLaunchKernelGGL(bf16_to_f32_kernel, g1(count, 256), dim3(256), 0, s, a, b, count);
    FLIP2_LAUNCH_CHECK();
}
void launch_zero_f32(float* a, long long count, hipStream_t s) {
    if (count <= 0) return;
    hipLaunchKernelGGL(zero_f32_kernel, g1(count, 256), dim3(256), 0, s, a, count);
    FLIP2_LAUNCH_CHECK();
}

}  // namespace nn
}  // namespace flip2
