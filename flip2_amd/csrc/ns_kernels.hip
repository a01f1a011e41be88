// North-star particle->grid transfer on 8^3 bricks (FP32 state):
//   binning to brick-tiled cell keys (packing the particle into a 32-byte
//   record), gather of the records through the sort permutation, per-cell
//   counts of the sorted keys, and trilinear MAC P2G with one workgroup per
//   brick accumulating in an LDS tile.
//
// Brick-tiled layout: cell (i,j,k) of an n^3 grid (n a multiple of 8) has
//   index ((bk*nbY + bj)*nbX + bi)*512 + (k&7)*64 + (j&7)*8 + (i&7)
// so a brick's 512 cells, and after sorting by that index its particles,
// are contiguous.  MAC faces use the same index: u(i,j,k) is the velocity on
// the low-x face of cell (i,j,k) at ((i)h, (j+1/2)h, (k+1/2)h); the faces on
// the high domain boundary are walls and are not stored.
//
// Particle record (2 x float4): (x, y, z, vx), (vy, vz, 0, 0).  One record is
// one 32-byte access, so gathers through the permutation and the reads of
// neighbouring cells' particles in P2G touch one memory segment per particle.
#include "ns_kernels.hpp"

namespace flip2 {
namespace ns {

__device__ inline uint32_t tiled_index(uint32_t i, uint32_t j, uint32_t k, uint32_t nbX, uint32_t nbY) {
    return (((k >> 3) * nbY + (j >> 3)) * nbX + (i >> 3)) * 512u + (k & 7u) * 64u + (j & 7u) * 8u + (i & 7u);
}

__device__ inline uint32_t clamp_cell(double q, uint32_t n) {
    if (!(q > 0.0)) return 0u;
    const double f = floor(q);
    return f >= (double)(n - 1) ? n - 1 : (uint32_t)f;
}

__device__ inline uint32_t cell_key(float x, float y, float z, const Grid& g) {
    const double inv = 1.0 / (double)g.cellSize;
    const uint32_t i = clamp_cell(((double)x - (double)g.negX) * inv, g.sizeX);
    const uint32_t j = clamp_cell(((double)y - (double)g.negY) * inv, g.sizeY);
    const uint32_t k = clamp_cell(((double)z - (double)g.negZ) * inv, g.sizeZ);
    return tiled_index(i, j, k, g.sizeX >> 3, g.sizeY >> 3);
}

// Cell keys of FP32 positions (clamped into the grid).
__global__ void bin_kernel(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                           uint32_t n, Grid g, uint32_t* __restrict__ keys) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    keys[p] = cell_key(x[p], y[p], z[p], g);
}

// Cell keys plus the packed particle record from SoA arrays.
__global__ void bin_pack_kernel(Float6 s, uint32_t n, Grid g, uint32_t* __restrict__ keys, float4* __restrict__ rec) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float x = s.p[0][p], y = s.p[1][p], z = s.p[2][p];
    keys[p] = cell_key(x, y, z, g);
    rec[2 * p] = make_float4(x, y, z, s.p[3][p]);
    rec[2 * p + 1] = make_float4(s.p[4][p], s.p[5][p], 0.f, 0.f);
}

// Cell keys from records.
__global__ void bin_rec_kernel(const float4* __restrict__ rec, uint32_t n, Grid g, uint32_t* __restrict__ keys) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float4 a = rec[2 * p];
    keys[p] = cell_key(a.x, a.y, a.z, g);
}

// dst[i] = src[perm[i]] (records).
__global__ void gather_rec_kernel(const uint32_t* __restrict__ perm, uint32_t n, const float4* __restrict__ src,
                                  float4* __restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = perm[i];
    if (j >= n) return;
    const float4 a = src[2 * j];
    const float4 b = src[2 * j + 1];
    dst[2 * i] = a;
    dst[2 * i + 1] = b;
}

// SoA -> records and back.
__global__ void pack_kernel(Float6 s, uint32_t n, float4* __restrict__ rec) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    rec[2 * p] = make_float4(s.p[0][p], s.p[1][p], s.p[2][p], s.p[3][p]);
    rec[2 * p + 1] = make_float4(s.p[4][p], s.p[5][p], 0.f, 0.f);
}

__global__ void unpack_kernel(const float4* __restrict__ rec, uint32_t n, Float6Out d) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float4 a = rec[2 * p], b = rec[2 * p + 1];
    d.p[0][p] = a.x; d.p[1][p] = a.y; d.p[2][p] = a.z;
    d.p[3][p] = a.w; d.p[4][p] = b.x; d.p[5][p] = b.y;
}

__global__ void zero_u32_kernel(uint32_t* __restrict__ a, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = 0u;
}

// counts[key] = length of key's run in the sorted key array (written by the
// run's first element; cells without particles keep the zero fill).
__global__ void run_length_kernel(const uint32_t* __restrict__ keys, uint32_t n, uint32_t numCells,
                                  uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = keys[i];
    if (i > 0 && keys[i - 1] == k) return;
    uint32_t e = i + 1;
    while (e < n && keys[e] == k) ++e;
    if (k < numCells) counts[k] = e - i;
}

namespace {

// Trilinear weights and base node of a MAC component: the node lattice is
// offset by (0,1/2,1/2), (1/2,0,1/2) or (1/2,1/2,0) cells.
struct Stencil {
    int base[3];
    float w0[3], w1[3];  // weights of node base and base+1 per axis
};

__device__ inline Stencil mac_stencil(const double q[3], int c) {
    Stencil s;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double v = (d == c) ? q[d] : q[d] - 0.5;
        const double f = floor(v);
        s.base[d] = (int)f;
        const float t = (float)(v - f);
        s.w1[d] = t;
        s.w0[d] = 1.0f - t;
    }
    return s;
}

// Morton (Z-order) decode of a 3x10-bit index.
__device__ inline uint32_t compact3(uint32_t v) {
    v &= 0x09249249u;
    v = (v ^ (v >> 2)) & 0x030c30c3u;
    v = (v ^ (v >> 4)) & 0x0300f00fu;
    v = (v ^ (v >> 8)) & 0xff0000ffu;
    v = (v ^ (v >> 16)) & 0x000003ffu;
    return v;
}

}  // namespace

// P2G of one brick: particles of the brick's cells and of the one-cell shell
// around it scatter trilinear weights into the brick's owned faces (the low
// faces of its 512 cells, per component) held in LDS; each owned face is
// written once: u = sum(w v) / sum(w), 0 where no particle reaches it.
// Workgroups are dealt to the 8 XCDs round-robin; each XCD gets a contiguous
// range of bricks, in Z-order when the brick grid is a power-of-two cube, so
// the shell cells a workgroup reads were loaded by its neighbours into the
// same L2.
__global__ void __launch_bounds__(256)
p2g_brick_kernel(const float4* __restrict__ rec, const uint32_t* __restrict__ cellStart, Grid g,
                 uint32_t numBricks, int morton, float* __restrict__ U, float* __restrict__ V, float* __restrict__ W,
                 float* __restrict__ WU, float* __restrict__ WV, float* __restrict__ WW) {
    __shared__ float s_acc[3][512];
    __shared__ float s_wgt[3][512];
    const uint32_t tid = threadIdx.x;
    const uint32_t perXcd = (numBricks + 7u) >> 3;
    const uint32_t order = (blockIdx.x & 7u) * perXcd + (blockIdx.x >> 3);
    if (order >= numBricks) return;
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    uint32_t bi, bj, bk;
    if (morton) {
        bi = compact3(order);
        bj = compact3(order >> 1);
        bk = compact3(order >> 2);
    } else {
        bi = order % nbX;
        bj = (order / nbX) % nbY;
        bk = order / (nbX * nbY);
    }
    const uint32_t brick = (bk * nbY + bj) * nbX + bi;
    const int lo[3] = {(int)(bi * 8u), (int)(bj * 8u), (int)(bk * 8u)};
    for (uint32_t t = tid; t < 512u; t += 256u) {
        s_acc[0][t] = 0.f; s_acc[1][t] = 0.f; s_acc[2][t] = 0.f;
        s_wgt[0][t] = 0.f; s_wgt[1][t] = 0.f; s_wgt[2][t] = 0.f;
    }
    __syncthreads();
    const double inv = 1.0 / (double)g.cellSize;
    const double neg[3] = {(double)g.negX, (double)g.negY, (double)g.negZ};

    auto scatter = [&](uint32_t p) {
        const float4 a = rec[2 * p];
        const float4 b = rec[2 * p + 1];
        const double q[3] = {((double)a.x - neg[0]) * inv, ((double)a.y - neg[1]) * inv, ((double)a.z - neg[2]) * inv};
        const float vel[3] = {a.w, b.x, b.y};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const Stencil st = mac_stencil(q, c);
#pragma unroll
            for (int ax = 0; ax < 2; ++ax) {
                const int li = st.base[0] + ax - lo[0];
                if ((unsigned)li >= 8u) continue;
                const float wa = ax ? st.w1[0] : st.w0[0];
#pragma unroll
                for (int by = 0; by < 2; ++by) {
                    const int lj = st.base[1] + by - lo[1];
                    if ((unsigned)lj >= 8u) continue;
                    const float wab = wa * (by ? st.w1[1] : st.w0[1]);
#pragma unroll
                    for (int bz = 0; bz < 2; ++bz) {
                        const int lk = st.base[2] + bz - lo[2];
                        if ((unsigned)lk >= 8u) continue;
                        const float wt = wab * (bz ? st.w1[2] : st.w0[2]);
                        const int node = lk * 64 + lj * 8 + li;
                        atomicAdd(&s_wgt[c][node], wt);
                        atomicAdd(&s_acc[c][node], wt * vel[c]);
                    }
                }
            }
        }
    };

    // the brick's own particles: one contiguous range
    const uint32_t pBegin = cellStart[brick * 512u];
    const uint32_t pEnd = cellStart[brick * 512u + 512u];
    for (uint32_t p = pBegin + tid; p < pEnd; p += 256u) scatter(p);

    // shell cells: the 10^3 block around the brick minus the brick
    for (uint32_t t = tid; t < 1000u; t += 256u) {
        const int sx = (int)(t % 10u) - 1, sy = (int)((t / 10u) % 10u) - 1, sz = (int)(t / 100u) - 1;
        if (sx >= 0 && sx < 8 && sy >= 0 && sy < 8 && sz >= 0 && sz < 8) continue;
        const int ci = lo[0] + sx, cj = lo[1] + sy, ck = lo[2] + sz;
        if (ci < 0 || cj < 0 || ck < 0 || ci >= (int)g.sizeX || cj >= (int)g.sizeY || ck >= (int)g.sizeZ) continue;
        const uint32_t cell = tiled_index((uint32_t)ci, (uint32_t)cj, (uint32_t)ck, nbX, nbY);
        const uint32_t b0 = cellStart[cell], b1 = cellStart[cell + 1];
        for (uint32_t p = b0; p < b1; ++p) scatter(p);
    }
    __syncthreads();
    float* out[3] = {U, V, W};
    float* outW[3] = {WU, WV, WW};
    for (uint32_t t = tid; t < 512u; t += 256u) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float wsum = s_wgt[c][t];
            out[c][brick * 512u + t] = wsum > 0.f ? s_acc[c][t] / wsum : 0.f;
            if (outW[c]) outW[c][brick * 512u + t] = wsum;
        }
    }
}

// ---------------------------------------------------------------------------
namespace {
inline dim3 grid1d(uint32_t n, uint32_t bs) { return dim3(ceil_div(n ? n : 1, bs)); }
}

void launch_bin(const float* x, const float* y, const float* z, uint32_t n, Grid g, uint32_t* keys, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(bin_kernel, grid1d(n, 256), dim3(256), 0, s, x, y, z, n, g, keys);
    FLIP2_LAUNCH_CHECK();
}

void launch_bin_pack(Float6 src, uint32_t n, Grid g, uint32_t* keys, float4* rec, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(bin_pack_kernel, grid1d(n, 256), dim3(256), 0, s, src, n, g, keys, rec);
    FLIP2_LAUNCH_CHECK();
}

void launch_bin_rec(const float4* rec, uint32_t n, Grid g, uint32_t* keys, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(bin_rec_kernel, grid1d(n, 256), dim3(256), 0, s, rec, n, g, keys);
    FLIP2_LAUNCH_CHECK();
}

void launch_gather_rec(const uint32_t* perm, uint32_t n, const float4* src, float4* dst, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(gather_rec_kernel, grid1d(n, 256), dim3(256), 0, s, perm, n, src, dst);
    FLIP2_LAUNCH_CHECK();
}

void launch_pack(Float6 src, uint32_t n, float4* rec, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(pack_kernel, grid1d(n, 256), dim3(256), 0, s, src, n, rec);
    FLIP2_LAUNCH_CHECK();
}

void launch_unpack(const float4* rec, uint32_t n, Float6Out dst, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(unpack_kernel, grid1d(n, 256), dim3(256), 0, s, rec, n, dst);
    FLIP2_LAUNCH_CHECK();
}

void launch_cell_counts(const uint32_t* sortedKeys, uint32_t n, uint32_t numCells, uint32_t* counts, hipStream_t s) {
    hipLaunchKernelGGL(zero_u32_kernel, grid1d(numCells + 1, 256), dim3(256), 0, s, counts, numCells + 1);
    FLIP2_LAUNCH_CHECK();
    if (!n) return;
    hipLaunchKernelGGL(run_length_kernel, grid1d(n, 256), dim3(256), 0, s, sortedKeys, n, numCells, counts);
    FLIP2_LAUNCH_CHECK();
}

void launch_p2g(const float4* sortedRec, const uint32_t* cellStart, Grid g, float* U, float* V, float* W, float* WU,
                float* WV, float* WW, hipStream_t s) {
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3, nbZ = g.sizeZ >> 3;
    const uint32_t numBricks = nbX * nbY * nbZ;
    if (!numBricks) return;
    const bool pow2 = (nbX & (nbX - 1)) == 0;
    const int morton = (nbX == nbY && nbY == nbZ && pow2 && nbX <= 1024) ? 1 : 0;
    const uint32_t blocks = ((numBricks + 7u) >> 3) * 8u;
    hipLaunchKernelGGL(p2g_brick_kernel, dim3(blocks), dim3(256), 0, s, sortedRec, cellStart, g, numBricks, morton, U,
                       V, W, WU, WV, WW);
    FLIP2_LAUNCH_CHECK();
}

}  // namespace ns
}  // namespace flip2
