// North-star particle->grid transfer on 8^3 bricks (FP32 state):
//   binning to brick-tiled cell keys (packing the particle into a 32-byte
//   record), gather of the records through the sort permutation, per-cell
//   counts of the sorted keys, and trilinear MAC P2G with one workgroup per
//   brick accumulating in an LDS tile.
//
// P2G accumulates in 64-bit fixed point with integer LDS atomics: on gfx950
// an LDS float atomic add costs tens of times an integer one, and integer
// sums are exact, so the result does not depend on the order in which
// particles arrive (it is deterministic).  The scale is chosen per brick
// from the number of particles the brick's workgroup scatters (a face's
// weight sum cannot exceed it) and the largest particle speed of the step
// (reduced by the gather kernel), so that no sum can exceed 2^62; at ~1000
// particles per brick neighbourhood and |v| <= 1 the resolution is ~2^-52.
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

#include "face_prep.hpp"

namespace flip2 {
namespace ns {

// cell_key, clamp_cell: face_prep.hpp (shared with the G2P + RK2 kernel,
// which writes the next step's keys).

// Cell keys of FP32 positions (clamped into the grid).
__global__ void bin_kernel(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                           uint32_t n, Grid g, uint32_t* __restrict__ keys) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    keys[p] = cell_key(x[p], y[p], z[p], g);
}

// Cell keys plus the packed particle record from SoA arrays.
__global__ void bin_pack_kernel(Float6 s, uint32_t n, Grid g, uint32_t* __restrict__ keys, float4* __restrict__ rec,
                                uint32_t* __restrict__ stats) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && stats) stats[0] = 0u;
    if (p >= n) return;
    const float x = s.p[0][p], y = s.p[1][p], z = s.p[2][p];
    keys[p] = cell_key(x, y, z, g);
    rec[2 * p] = make_float4(x, y, z, s.p[3][p]);
    rec[2 * p + 1] = make_float4(s.p[4][p], s.p[5][p], 0.f, 0.f);
}

// Cell keys from records.
__global__ void bin_rec_kernel(const float4* __restrict__ rec, uint32_t n, Grid g, uint32_t* __restrict__ keys,
                               uint32_t* __restrict__ stats) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0 && stats) stats[0] = 0u;
    if (p >= n) return;
    const float4 a = rec[2 * p];
    keys[p] = cell_key(a.x, a.y, a.z, g);
}

// Block maximum of a non-negative u32 (bit pattern of a non-negative float
// orders like the float), folded into *target with at most one atomic per
// block, and none once *target is already larger (the common case after the
// first blocks, which keeps the single address uncontended).
// Bits of max(|vx|, |vy|, |vz|) for the speed statistic.  fmaxf returns its
// non-NaN operand, so a NaN component is mapped to the quiet-NaN pattern
// 0x7fc00000 explicitly: as an unsigned integer it is larger than every
// finite or infinite magnitude, so it wins the u32 maximum and the statistic
// reads back as NaN.
__device__ inline uint32_t speed_bits(float vx, float vy, float vz) {
    if (isnan(vx) || isnan(vy) || isnan(vz)) return 0x7fc00000u;
    return __float_as_uint(fmaxf(fabsf(vx), fmaxf(fabsf(vy), fabsf(vz))));
}

__device__ inline void block_max_to(uint32_t v, uint32_t* target) {
    __shared__ uint32_t s_part[16];
    for (int d = 32; d >= 1; d >>= 1) v = max(v, (uint32_t)__shfl_xor(v, d, 64));
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0) s_part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t m = 0;
        for (uint32_t w = 0; w < (blockDim.x + 63u) >> 6; ++w) m = max(m, s_part[w]);
        if (m > __hip_atomic_load(target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(target, m);
    }
}

// dst[i] = src[perm[i]] (records); stats[0] = max |velocity component|
// (float bits) when stats is given.
// The reads through the permutation fetch a whole 128-byte line per 32-byte
// record; with one record per lane the kernel runs at the HBM rate of that
// traffic (several records per lane measured slower).
__global__ void gather_rec_kernel(const uint32_t* __restrict__ perm, uint32_t n, const float4* __restrict__ src,
                                  float4* __restrict__ dst, uint32_t* __restrict__ stats) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t vmax = 0u;
    if (i < n) {
        const uint32_t j = perm[i];
        if (j < n) {
            const float4 a = src[2 * j];
            const float4 b = src[2 * j + 1];
            dst[2 * i] = a;
            dst[2 * i + 1] = b;
            vmax = speed_bits(a.w, b.x, b.y);
        }
    }
    if (stats) block_max_to(vmax, &stats[0]);
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

// counts[0 .. numCells] from the sorted keys, after a zero fill of the whole
// array: the first element of each run of equal keys writes its cell's run
// length.  (Filling the empty cells between runs from the run starts instead
// serialises a sparse scene's long empty stretches -- tens of millions of
// cells between two fluid bodies -- into single threads.)  Keys >= numCells
// count as numCells, which keeps 0.
__global__ void run_length_kernel(const uint32_t* __restrict__ keys, uint32_t n, uint32_t numCells,
                                  uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = min(keys[i], numCells);
    if (k >= numCells) return;
    if (i > 0 && min(keys[i - 1], numCells) == k) return;
    // end of the run: a short walk covers the usual few particles per cell;
    // a longer run is bracketed by doubling steps and bisected, so even one
    // run of all n particles costs O(log n)
    uint32_t e = i + 1;
    const uint32_t walkEnd = min(n, i + 16u);
    while (e < walkEnd && min(keys[e], numCells) == k) ++e;
    if (e == walkEnd && e < n && min(keys[e], numCells) == k) {
        uint32_t lo = e, step = 16u;  // keys[lo] == k
        uint32_t hi = lo;
        while (true) {
            hi = (n - lo > step) ? lo + step : n;
            if (hi == n || min(keys[hi], numCells) != k) break;
            lo = hi;
            step = step <= 0x40000000u ? step * 2u : step;
        }
        // keys[lo] == k; hi == n or keys[hi] != k: the run ends in (lo, hi]
        while (hi - lo > 1u) {
            const uint32_t mid = lo + (hi - lo) / 2u;
            if (min(keys[mid], numCells) == k) lo = mid;
            else hi = mid;
        }
        e = hi;
    }
    counts[k] = e - i;
}

// Statistics for a P2G call whose inputs did not come through the pipeline:
// stats[0] = max |velocity component| (float bits).
__global__ void p2g_stats_kernel(const float4* __restrict__ rec, const uint32_t* __restrict__ cellStart,
                                 uint32_t numCells, uint32_t* __restrict__ stats) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t n = cellStart[numCells];
    uint32_t vmax = 0u;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float4 a = rec[2 * i];
        const float4 b = rec[2 * i + 1];
        vmax = max(vmax, speed_bits(a.w, b.x, b.y));
    }
    block_max_to(vmax, &stats[0]);
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

// The one-cell shell around a brick, as 222 runs of cells that are
// contiguous in the tiled order (so each run is one contiguous range of
// sorted particles): the two z-face layers (64 cells each), the y faces (8
// runs of 8 cells each), the four yz edges (8 cells), and the x-side cells,
// which are single cells.  Run r's cells are returned as a tiled range.
__device__ inline bool shell_run(int r, const int bb[3], uint32_t nbX, uint32_t nbY, uint32_t nbZ, uint32_t& c0,
                                 uint32_t& c1) {
    int acc = 0;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!dx && !dy && !dz) continue;
                const int nr = dx ? (dy ? 1 : 8) * (dz ? 1 : 8) : (dy && !dz ? 8 : 1);
                if (r >= acc + nr) {
                    acc += nr;
                    continue;
                }
                const int k = r - acc;
                const int bx = bb[0] + dx, by = bb[1] + dy, bz = bb[2] + dz;
                if (bx < 0 || by < 0 || bz < 0 || bx >= (int)nbX || by >= (int)nbY || bz >= (int)nbZ) return false;
                const uint32_t base = (((uint32_t)bz * nbY + (uint32_t)by) * nbX + (uint32_t)bx) * 512u;
                const int ci = dx < 0 ? 7 : 0, cj = dy < 0 ? 7 : 0, ck = dz < 0 ? 7 : 0;
                uint32_t lo, len;
                if (!dx && !dy) { lo = ck * 64; len = 64; }
                else if (!dx && !dz) { lo = k * 64 + cj * 8; len = 8; }
                else if (!dx) { lo = ck * 64 + cj * 8; len = 8; }
                else {
                    const int j2 = dy ? cj : (k & 7);
                    const int k2 = dz ? ck : (dy ? k : (k >> 3));
                    lo = k2 * 64 + j2 * 8 + ci;
                    len = 1;
                }
                c0 = base + lo;
                c1 = c0 + len;
                return true;
            }
    return false;
}

constexpr int kShellRuns = 222;

// P2G of one brick: the particles of the brick's cells and of the one-cell
// shell around it (run 0 = the brick, runs 1..222 = the shell) are handed out
// one per thread; each scatters trilinear weights into the brick's owned
// faces (the low faces of its 512 cells, per component) held in LDS as
// 64-bit fixed-point sums.  Each owned face is written once:
// u = sum(w v) / sum(w), 0 where no particle reaches it.
struct P2GShared {
    unsigned long long w[3][512];
    unsigned long long m[3][512];
    uint32_t first[kShellRuns + 1];
    uint32_t off[kShellRuns + 2];
    uint32_t wave[4];
};

// Scatter of brick bb (index `brick`) into sh.w, sh.m (cleared here) by the
// 256 threads of the workgroup; scaleW, scaleM: the fixed-point scales.
// Ends with the sums complete and visible to the workgroup.
// occ (may be null): brick occupancy flags; a run in a brick without
// particles is empty without reading cellStart (which then need only hold
// the cells of occupied bricks).
__device__ inline void p2g_scatter_brick(P2GShared& sh, const int bb[3], uint32_t brick, const float4* __restrict__ rec,
                                         const uint32_t* __restrict__ cellStart, Grid g, uint32_t nbX, uint32_t nbY,
                                         uint32_t nbZ, const uint32_t* __restrict__ stats, float& scaleW,
                                         float& scaleM, const uint8_t* __restrict__ occ = nullptr) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t t = tid; t < 1536u; t += 256u) {
        (&sh.w[0][0])[t] = 0ull;
        (&sh.m[0][0])[t] = 0ull;
    }
    // run table and its exclusive scan
    uint32_t first = 0, cnt = 0;
    if (tid <= (uint32_t)kShellRuns) {
        uint32_t c0 = brick * 512u, c1 = c0 + 512u;
        bool ok = tid == 0 || shell_run((int)tid - 1, bb, nbX, nbY, nbZ, c0, c1);
        if (ok && occ) ok = occ[c0 >> 9] != 0;
        if (ok) {
            first = cellStart[c0];
            cnt = cellStart[c1] - first;
        }
        sh.first[tid] = first;
    }
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    uint32_t incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= (uint32_t)d) incl += o;
    }
    if (lane == 63u) sh.wave[wave] = incl;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < wave; ++w) base += sh.wave[w];
    const uint32_t total = sh.wave[0] + sh.wave[1] + sh.wave[2] + sh.wave[3];
    if (tid <= (uint32_t)kShellRuns) sh.off[tid] = base + incl - cnt;
    if (tid == 0) sh.off[kShellRuns + 1] = total;
    __syncthreads();

    // fixed-point scales: a face's weight sum is at most the number of
    // particles this workgroup scatters (each w <= 1)
    const float bound = (float)(total > 0u ? total : 1u);
    const float vmax = __uint_as_float(stats[0]);
    const int ew = ilogbf(bound) + 1;
    scaleW = ldexpf(1.0f, 62 - ew);
    const int em = vmax > 0.f ? ilogbf(bound * vmax) + 1 : ew;
    scaleM = ldexpf(1.0f, min(62 - em, 126));

    const double inv = 1.0 / (double)g.cellSize;
    const double org[3] = {(double)g.negX + (double)(bb[0] * 8) * (double)g.cellSize,
                           (double)g.negY + (double)(bb[1] * 8) * (double)g.cellSize,
                           (double)g.negZ + (double)(bb[2] * 8) * (double)g.cellSize};
    // Wave w takes the contiguous quarter [w Q, w Q + Q) of the particles
    // and its lanes stride through it R apart, so the lanes of one atomic
    // instruction sit in different cells: consecutive particles (same or
    // adjacent cells) would hit the same LDS words and serialise.
    const uint32_t Q = (total + 3u) >> 2;
    const uint32_t R = (Q + 63u) >> 6;
    const uint32_t qBeg = wave * Q, qEnd = min(total, qBeg + Q);
    // A lane's particles are consecutive (sIdx = lBeg + i): one binary search
    // for the first, then the run index only moves forward.  Records are
    // loaded kPre at a time before any is scattered, so their memory
    // latencies overlap.
    constexpr uint32_t kPre = 4;
    const uint32_t lBeg = qBeg + lane * R;
    const uint32_t lCnt = lBeg < qEnd ? min(R, qEnd - lBeg) : 0u;
    int run = 0;
    if (lCnt) {
        int b = kShellRuns + 1;  // last run with s_off <= lBeg
        while (b - run > 1) {
            const int m = (run + b) >> 1;
            if (sh.off[m] <= lBeg) run = m;
            else b = m;
        }
    }
    for (uint32_t i0 = 0; i0 < lCnt; i0 += kPre) {
        float4 pr0[kPre], pr1[kPre];
#pragma unroll
        for (uint32_t k = 0; k < kPre; ++k) {
            if (i0 + k >= lCnt) break;
            const uint32_t sIdx = lBeg + i0 + k;
            while (sh.off[run + 1] <= sIdx) ++run;  // sh.off[kShellRuns + 1] = total > sIdx
            const uint32_t p = sh.first[run] + (sIdx - sh.off[run]);
            pr0[k] = rec[2 * p];
            pr1[k] = rec[2 * p + 1];
        }
#pragma unroll
        for (uint32_t k = 0; k < kPre; ++k) {
            if (i0 + k >= lCnt) break;
            const float4 r0 = pr0[k];
            const float4 r1 = pr1[k];
            // position in cells relative to the brick's low corner
            const float q[3] = {(float)(((double)r0.x - org[0]) * inv), (float)(((double)r0.y - org[1]) * inv),
                                (float)(((double)r0.z - org[2]) * inv)};
            const float vel[3] = {r0.w, r1.x, r1.y};
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                int bs[3];
                float t1[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const float v = (d == c) ? q[d] : q[d] - 0.5f;
                    const float f = floorf(v);
                    bs[d] = (int)f;
                    t1[d] = v - f;
                }
                const float mv = vel[c] * scaleM;
#pragma unroll
                for (int ax = 0; ax < 2; ++ax) {
                    const int li = bs[0] + ax;
                    if ((unsigned)li >= 8u) continue;
                    const float wa = ax ? t1[0] : 1.f - t1[0];
#pragma unroll
                    for (int by = 0; by < 2; ++by) {
                        const int lj = bs[1] + by;
                        if ((unsigned)lj >= 8u) continue;
                        const float wab = wa * (by ? t1[1] : 1.f - t1[1]);
#pragma unroll
                        for (int bz = 0; bz < 2; ++bz) {
                            const int lk = bs[2] + bz;
                            if ((unsigned)lk >= 8u) continue;
                            const float wt = wab * (bz ? t1[2] : 1.f - t1[2]);
                            const int node = lk * 64 + lj * 8 + li;
                            atomicAdd(&sh.w[c][node], (unsigned long long)__float2ull_rn(wt * scaleW));
                            atomicAdd(&sh.m[c][node], (unsigned long long)__float2ll_rn(wt * mv));
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
}

// The face value of component c at cell t of the scattered brick.
__device__ inline float p2g_face(const P2GShared& sh, int c, uint32_t t, double iw, double im, double& ws) {
    ws = (double)sh.w[c][t] * iw;
    const double ms = (double)(long long)sh.m[c][t] * im;
    return ws > 0.0 ? (float)(ms / ws) : 0.f;
}

// Bricks are dealt to the 8 XCDs round-robin by workgroup; each XCD gets a
// contiguous range of bricks, in Z-order when the brick grid is a
// power-of-two cube, so the shell cells a workgroup reads were loaded by its
// neighbours into the same L2.
__global__ void __launch_bounds__(256)
p2g_brick_kernel(const float4* __restrict__ rec, const uint32_t* __restrict__ cellStart, Grid g,
                 uint32_t numBricks, int morton, const uint32_t* __restrict__ stats, float* __restrict__ U,
                 float* __restrict__ V, float* __restrict__ W, float* __restrict__ WU, float* __restrict__ WV,
                 float* __restrict__ WW) {
    __shared__ P2GShared sh;
    const uint32_t tid = threadIdx.x;
    const uint32_t perXcd = (numBricks + 7u) >> 3;
    const uint32_t order = (blockIdx.x & 7u) * perXcd + (blockIdx.x >> 3);
    if (order >= numBricks) return;
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3, nbZ = g.sizeZ >> 3;
    int bb[3];
    if (morton) {
        bb[0] = (int)compact3(order);
        bb[1] = (int)compact3(order >> 1);
        bb[2] = (int)compact3(order >> 2);
    } else {
        bb[0] = (int)(order % nbX);
        bb[1] = (int)((order / nbX) % nbY);
        bb[2] = (int)(order / (nbX * nbY));
    }
    const uint32_t brick = ((uint32_t)bb[2] * nbY + (uint32_t)bb[1]) * nbX + (uint32_t)bb[0];
    float scaleW, scaleM;
    p2g_scatter_brick(sh, bb, brick, rec, cellStart, g, nbX, nbY, nbZ, stats, scaleW, scaleM);
    float* out[3] = {U, V, W};
    float* outW[3] = {WU, WV, WW};
    const double iw = 1.0 / (double)scaleW, im = 1.0 / (double)scaleM;
    for (uint32_t t = tid; t < 512u; t += 256u) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double ws;
            out[c][brick * 512u + t] = p2g_face(sh, c, t, iw, im, ws);
            if (outW[c]) outW[c][brick * 512u + t] = (float)ws;
        }
    }
}

// ---------------------------------------------------------------------------
// Transfer fused with the prepare stage (walls, solid box, gravity): only
// bricks that particles reach are scattered; every other brick's faces are
// the prepared values of a zero transfer.
// ---------------------------------------------------------------------------

// First index of a key >= v in the sorted keys[0, n).
__device__ inline uint32_t lower_bound_u32(const uint32_t* __restrict__ keys, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Brick starts of the sorted cell keys (brick-tiled: brick = key >> 9):
// brickStart[b] = first particle of brick b or later, brickStart[numBricks]
// = n; occ[b] = 1 if brick b holds a particle.  One thread per brick, two
// binary searches.
__global__ void brick_starts_kernel(const uint32_t* __restrict__ keys, uint32_t n, uint32_t numBricks,
                                    uint32_t* __restrict__ brickStart, uint8_t* __restrict__ occ) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= numBricks) return;
    const uint32_t s0 = lower_bound_u32(keys, n, b * 512u);
    const uint32_t s1 = b + 1u < numBricks ? lower_bound_u32(keys, n, (b + 1u) * 512u) : n;
    brickStart[b] = s0;
    if (b + 1u == numBricks) brickStart[numBricks] = n;
    occ[b] = s1 > s0 ? 1 : 0;
}

// cellStart of the occupied bricks of the touched list: the cell counts of a
// brick's particles (a histogram in LDS over its range of the sorted keys),
// exclusive-scanned and offset by the brick's start; entry (b + 1) * 512 is
// the next brick's start, which that brick (if occupied) writes too, with
// the same value.  The other bricks' entries are left as they are: readers
// test occupancy first.
__global__ void __launch_bounds__(256)
sparse_cell_starts_kernel(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                          const uint8_t* __restrict__ occ, const uint32_t* __restrict__ brickStart,
                          const uint32_t* __restrict__ keys, uint32_t* __restrict__ cellStart) {
    __shared__ uint32_t hist[512];
    __shared__ uint32_t waveSum[4];
    const uint32_t total = *count;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    for (uint32_t e = blockIdx.x; e < total; e += gridDim.x) {
        const uint32_t b = list[e];
        if (!occ[b]) continue;  // the same for the whole workgroup
        const uint32_t p0 = brickStart[b], p1 = brickStart[b + 1u];
        hist[tid] = 0u;
        hist[tid + 256u] = 0u;
        __syncthreads();
        for (uint32_t i = p0 + tid; i < p1; i += 256u) atomicAdd(&hist[keys[i] & 511u], 1u);
        __syncthreads();
        // exclusive scan of the 512 counts: thread t holds cells 2t, 2t + 1
        const uint32_t c0 = hist[2u * tid], c1 = hist[2u * tid + 1u];
        uint32_t incl = c0 + c1;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d, 64);
            if (lane >= (uint32_t)d) incl += o;
        }
        if (lane == 63u) waveSum[wave] = incl;
        __syncthreads();
        uint32_t base = p0;
        for (uint32_t w = 0; w < wave; ++w) base += waveSum[w];
        const uint32_t ex = base + incl - c0 - c1;
        cellStart[b * 512u + 2u * tid] = ex;
        cellStart[b * 512u + 2u * tid + 1u] = ex + c0;
        if (tid == 0) cellStart[b * 512u + 512u] = p1;
        __syncthreads();  // hist and waveSum are reused by the next brick
    }
}

// touched[b] = 1 if brick b or one of its 26 neighbours holds a particle (a
// superset of the bricks whose faces a particle reaches: the P2G shell is one
// cell deep); the touched bricks are appended to list.  Thread t looks at
// the t-th brick in Z-order when the brick grid is a power-of-two cube (else
// in index order), and each wave appends its bricks in thread order, so the
// list runs through space in compact stretches: the transfer kernel's
// workgroups on one XCD then read their bricks' shells through one L2.  (The
// transfer's result does not depend on the order.)
__global__ void touched_bricks_kernel(const uint8_t* __restrict__ occ, uint32_t nbX, uint32_t nbY, uint32_t nbZ,
                                      int morton, uint8_t* __restrict__ touched, uint32_t* __restrict__ list,
                                      uint32_t* __restrict__ count) {
    const uint32_t numBricks = nbX * nbY * nbZ;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    bool hit = false;
    uint32_t b = t;
    if (t < numBricks) {
        int bx, by, bz;
        if (morton) {
            bx = (int)compact3(t);
            by = (int)compact3(t >> 1);
            bz = (int)compact3(t >> 2);
            b = ((uint32_t)bz * nbY + (uint32_t)by) * nbX + (uint32_t)bx;
        } else {
            bx = (int)(t % nbX);
            by = (int)((t / nbX) % nbY);
            bz = (int)(t / (nbX * nbY));
        }
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    const int x = bx + dx, y = by + dy, z = bz + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= (int)nbX || y >= (int)nbY || z >= (int)nbZ) continue;
                    hit |= occ[((uint32_t)z * nbY + (uint32_t)y) * nbX + (uint32_t)x] != 0;
                }
        touched[b] = hit ? 1 : 0;
    }
    const unsigned long long mask = __ballot(hit);
    if (!mask) return;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(count, (uint32_t)__popcll(mask));
    base = __shfl(base, 0, 64);
    if (hit) list[base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = b;
}

// Faces of the untouched bricks: prepared values of a zero transfer (with
// prevTouched, only for the bricks the previous run touched).
__global__ void __launch_bounds__(256)
empty_faces_kernel(const uint8_t* __restrict__ touched, const uint8_t* __restrict__ prevTouched, uint32_t numBricks,
                   Grid g, const uint8_t* __restrict__ solid, SolidBox box, float gdt, float* __restrict__ U,
                   float* __restrict__ V, float* __restrict__ W, float* __restrict__ U0, float* __restrict__ V0,
                   float* __restrict__ W0) {
    // 64 bricks per workgroup: the first wave tests their flags once (a
    // lane per brick), then the workgroup writes the bricks that have faces
    // to write (few in steady state; a workgroup per brick spent the kernel
    // dispatching workgroups that exit at once)
    __shared__ unsigned long long s_mask;
    const uint32_t b0 = blockIdx.x * 64u;
    if (threadIdx.x < 64u) {
        const uint32_t bt = b0 + threadIdx.x;
        const bool hit = bt < numBricks && !touched[bt] && (!prevTouched || prevTouched[bt]);
        const unsigned long long m = __ballot(hit);
        if (threadIdx.x == 0) s_mask = m;
    }
    __syncthreads();
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    for (unsigned long long m = s_mask; m; m &= m - 1ull) {
        const uint32_t b = b0 + (uint32_t)__builtin_ctzll(m);
        for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) {
            const uint32_t c = b * 512u + l;
            const int gc[3] = {(int)((b % nbX) * 8u + (l & 7u)), (int)(((b / nbX) % nbY) * 8u + ((l >> 3) & 7u)),
                               (int)((b / (nbX * nbY)) * 8u + (l >> 6))};
            const float in[3] = {0.f, 0.f, 0.f};
            float saved[3], out[3];
            prepare_cell_faces(in, gc, g, solid, box, gdt, nbX, nbY, saved, out);
            U0[c] = saved[0];
            V0[c] = saved[1];
            W0[c] = saved[2];
            U[c] = out[0];
            V[c] = out[1];
            W[c] = out[2];
        }
    }
}

// Scatter + prepare of the touched bricks; a fixed grid of workgroups loops
// over the list, which is cut into chunks of 64 entries (a 4x4x4 block of
// bricks in Z-order) dealt round robin to the 8 XCDs: a chunk's shells are
// read through one L2, and every XCD gets a share of the dense and the
// sparse parts of the scene (contiguous eighths of a Z-ordered list gave one
// XCD each octant, i.e. most of a blob's particles to one or two XCDs).  A
// workgroup takes the next entry of its XCD's share from that XCD's counter
// (counters[xcd], zero at launch), one entry ahead, so workgroups with
// light bricks take more of them.
__global__ void __launch_bounds__(256)
p2g_prepare_kernel(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                   const float4* __restrict__ rec, const uint32_t* __restrict__ cellStart, Grid g,
                   const uint32_t* __restrict__ stats, const uint8_t* __restrict__ solid, SolidBox box, float gdt,
                   float* __restrict__ U, float* __restrict__ V, float* __restrict__ W, float* __restrict__ U0,
                   float* __restrict__ V0, float* __restrict__ W0, const uint8_t* __restrict__ occ,
                   uint32_t* __restrict__ counters) {
    __shared__ P2GShared sh;
    __shared__ uint32_t s_next;
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3, nbZ = g.sizeZ >> 3;
    const uint32_t total = *count;
    const uint32_t xcd = blockIdx.x & 7u;  // workgroups are dealt to the XCDs round robin
    float* F[3] = {U, V, W};
    float* F0[3] = {U0, V0, W0};
    if (threadIdx.x == 0) s_next = atomicAdd(&counters[xcd], 1u);
    __syncthreads();
    for (uint32_t m = s_next;; m = s_next) {
        // the XCD's m-th entry: offset m % 64 of its (m / 64)-th chunk
        const uint32_t e = ((m >> 6) * 8u + xcd) * 64u + (m & 63u);
        if (e >= total) break;  // chunks of an XCD are in increasing order (the same for the workgroup)
        __syncthreads();        // every thread has read s_next
        if (threadIdx.x == 0) s_next = atomicAdd(&counters[xcd], 1u);  // the next entry, during this brick
        const uint32_t brick = list[e];
        const int bb[3] = {(int)(brick % nbX), (int)((brick / nbX) % nbY), (int)(brick / (nbX * nbY))};
        float scaleW, scaleM;
        p2g_scatter_brick(sh, bb, brick, rec, cellStart, g, nbX, nbY, nbZ, stats, scaleW, scaleM, occ);
        const double iw = 1.0 / (double)scaleW, im = 1.0 / (double)scaleM;
        for (uint32_t t = threadIdx.x; t < 512u; t += 256u) {
            float in[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double ws;
                in[c] = p2g_face(sh, c, t, iw, im, ws);
            }
            const int gc[3] = {bb[0] * 8 + (int)(t & 7u), bb[1] * 8 + (int)((t >> 3) & 7u), bb[2] * 8 + (int)(t >> 6)};
            float saved[3], out[3];
            prepare_cell_faces(in, gc, g, solid, box, gdt, nbX, nbY, saved, out);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                F0[c][brick * 512u + t] = saved[c];
                F[c][brick * 512u + t] = out[c];
            }
        }
        __syncthreads();  // the next brick clears the sums; s_next is written
    }
}

// ---------------------------------------------------------------------------
namespace {
inline dim3 grid1d(uint32_t n, uint32_t bs) { return dim3(ceil_div(n ? n : 1, bs)); }
}

void launch_bin(const float* x, const float* y, const float* z, uint32_t n, Grid g, uint32_t* keys, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(bin_kernel, grid1d(n, 256), dim3(256), 0, s, x, y, z, n, g, keys);
    FLIP2_LAUNCH_CHECK();
}

void launch_bin_pack(Float6 src, uint32_t n, Grid g, uint32_t* keys, float4* rec, uint32_t* stats, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(bin_pack_kernel, grid1d(n, 256), dim3(256), 0, s, src, n, g, keys, rec, stats);
    FLIP2_LAUNCH_CHECK();
}

void launch_bin_rec(const float4* rec, uint32_t n, Grid g, uint32_t* keys, uint32_t* stats, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(bin_rec_kernel, grid1d(n, 256), dim3(256), 0, s, rec, n, g, keys, stats);
    FLIP2_LAUNCH_CHECK();
}

void launch_clear_stats(uint32_t* stats, hipStream_t s) {
    FLIP2_KERNEL(zero_u32_kernel, dim3(1), dim3(64), 0, s, stats, 1u);
    FLIP2_LAUNCH_CHECK();
}

void launch_gather_rec(const uint32_t* perm, uint32_t n, const float4* src, float4* dst, uint32_t* stats,
                       hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(gather_rec_kernel, grid1d(n, 256), dim3(256), 0, s, perm, n, src, dst, stats);
    FLIP2_LAUNCH_CHECK();
}

void launch_pack(Float6 src, uint32_t n, float4* rec, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(pack_kernel, grid1d(n, 256), dim3(256), 0, s, src, n, rec);
    FLIP2_LAUNCH_CHECK();
}

void launch_unpack(const float4* rec, uint32_t n, Float6Out dst, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(unpack_kernel, grid1d(n, 256), dim3(256), 0, s, rec, n, dst);
    FLIP2_LAUNCH_CHECK();
}

void launch_cell_counts(const uint32_t* sortedKeys, uint32_t n, uint32_t numCells, uint32_t* counts, hipStream_t s) {
    if (!n) {
        FLIP2_KERNEL(zero_u32_kernel, grid1d(numCells + 1, 256), dim3(256), 0, s, counts, numCells + 1);
        FLIP2_LAUNCH_CHECK();
        return;
    }
    FLIP2_KERNEL(zero_u32_kernel, grid1d(numCells + 1, 256), dim3(256), 0, s, counts, numCells + 1);
    FLIP2_LAUNCH_CHECK();
    FLIP2_KERNEL(run_length_kernel, grid1d(n, 256), dim3(256), 0, s, sortedKeys, n, numCells, counts);
    FLIP2_LAUNCH_CHECK();
}

void launch_p2g_stats(const float4* sortedRec, const uint32_t* cellStart, uint32_t numCells, uint32_t* stats,
                      hipStream_t s) {
    FLIP2_KERNEL(zero_u32_kernel, dim3(1), dim3(64), 0, s, stats, 1u);
    FLIP2_LAUNCH_CHECK();
    FLIP2_KERNEL(p2g_stats_kernel, dim3(1024), dim3(256), 0, s, sortedRec, cellStart, numCells, stats);
    FLIP2_LAUNCH_CHECK();
}

void launch_p2g(const float4* sortedRec, const uint32_t* cellStart, Grid g, const uint32_t* stats, float* U, float* V,
                float* W, float* WU, float* WV, float* WW, hipStream_t s) {
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3, nbZ = g.sizeZ >> 3;
    const uint32_t numBricks = nbX * nbY * nbZ;
    if (!numBricks) return;
    const bool pow2 = (nbX & (nbX - 1)) == 0;
    const int morton = (nbX == nbY && nbY == nbZ && pow2 && nbX <= 1024) ? 1 : 0;
    const uint32_t blocks = ((numBricks + 7u) >> 3) * 8u;
    FLIP2_KERNEL(p2g_brick_kernel, dim3(blocks), dim3(256), 0, s, sortedRec, cellStart, g, numBricks, morton,
                       stats, U, V, W, WU, WV, WW);
    FLIP2_LAUNCH_CHECK();
}

void launch_p2g_prepare(const float4* sortedRec, const uint32_t* sortedKeys, uint32_t n, uint32_t* cellStart,
                        uint32_t* brickStart, Grid g, const uint32_t* stats, const FacePrep& prep, float* U, float* V,
                        float* W, TouchedBricks t, hipStream_t s) {
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3, nbZ = g.sizeZ >> 3;
    const uint32_t numBricks = nbX * nbY * nbZ;
    if (!numBricks) return;
    FLIP2_REQUIRE(prep.U0 && prep.V0 && prep.W0 && t.occ && t.touched && t.list && t.count && brickStart &&
                      cellStart && (sortedKeys || n == 0),
                  kErrArgument, "transfer + prepare: null array");
    FLIP2_KERNEL(zero_u32_kernel, dim3(1), dim3(64), 0, s, t.count, 1u);
    FLIP2_LAUNCH_CHECK();
    FLIP2_KERNEL(brick_starts_kernel, grid1d(numBricks, 256), dim3(256), 0, s, sortedKeys, n, numBricks, brickStart,
                 t.occ);
    FLIP2_LAUNCH_CHECK();
    const bool pow2 = (nbX & (nbX - 1)) == 0;
    const int morton = (nbX == nbY && nbY == nbZ && pow2 && nbX <= 1024) ? 1 : 0;
    FLIP2_KERNEL(touched_bricks_kernel, grid1d(numBricks, 256), dim3(256), 0, s, t.occ, nbX, nbY, nbZ, morton,
                 t.touched, t.list, t.count);
    FLIP2_LAUNCH_CHECK();
    FLIP2_KERNEL(sparse_cell_starts_kernel, dim3(std::min<uint32_t>(numBricks, 2048u)), dim3(256), 0, s, t.list,
                 t.count, t.occ, brickStart, sortedKeys, cellStart);
    FLIP2_LAUNCH_CHECK();
    FLIP2_REQUIRE(prep.full || t.prevTouched, kErrArgument, "transfer + prepare: no previous touched flags");
    FLIP2_KERNEL(empty_faces_kernel, dim3((numBricks + 63u) / 64u), dim3(256), 0, s, t.touched, prep.full ? nullptr : t.prevTouched,
                 numBricks, g, prep.solid, prep.box, prep.gdt, U, V, W, prep.U0, prep.V0, prep.W0);
    FLIP2_LAUNCH_CHECK();
    // as many workgroups as the CUs hold at once (registers and LDS), 8 per XCD row
    const uint32_t groups = std::min<uint32_t>(resident_groups(p2g_prepare_kernel, 256), 8u * ((numBricks + 7u) >> 3));
    FLIP2_KERNEL(zero_u32_kernel, dim3(1), dim3(64), 0, s, t.counters, 8u);
    FLIP2_LAUNCH_CHECK();
    FLIP2_KERNEL(p2g_prepare_kernel, dim3(groups), dim3(256), 0, s, t.list, t.count, sortedRec, cellStart, g, stats,
                 prep.solid, prep.box, prep.gdt, U, V, W, prep.U0, prep.V0, prep.W0, t.occ, t.counters);
    FLIP2_LAUNCH_CHECK();
}

}  // namespace ns
}  // namespace flip2
