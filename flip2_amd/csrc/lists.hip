// Sub-cell lists and per-voxel particle lists: the reference's
// cudaFindSubCell (particleToGridFunctions.cu:69-262) and
// cudaSumParticlesPerNodeAndWriteNumUsedVoxels / cudaParticleListCreate
// (perVoxelParticleListFunctions.cu:6-189) on gfx950.  The reference builds
// these but never calls them from its step; results are bit-exact integer
// lists.
//
// Sub-cells: one thread per particle counts, then (after inclusive scans of
// the counts) writes, the sub-cells of the 4^3 window around its own whose
// staggered point lies within radius sub-cells, per velocity component, in
// the reference's loop order and its mixed float/double arithmetic.
//
// Per-voxel lists: one wave per node.  Voxel histogram and the exclusive
// scans (used-voxel rank, list offset) in LDS; then the node's particles in
// chunks of 64 (one per lane): each lane sets its bit in a per-voxel 64-bit
// mask for every voxel its particle touches, and the particle's place in a
// voxel's list is that list's running length plus the number of lower lanes
// in the voxel's mask -- the lists come out in ascending particle order with
// integer atomics only (order independent).
#pragma clang fp contract(off)

#include "lists.hpp"

#include "scan_impl.hpp"

namespace flip2 {
namespace rc {
namespace {

__device__ inline uint32_t f2u_sat(float v) {
    if (!(v > 0.f)) return 0u;
    if (v >= 4294967296.f) return 0xffffffffu;
    return (uint32_t)v;
}

// float square root correctly rounded: the double root rounded to float
// (53 >= 2 * 24 + 2 bits, so the double rounding is exact)
__device__ inline float sqrt_f32(float v) { return (float)sqrt((double)v); }

// Calls emit(dim, subCell) for the touched sub-cells of one particle in the
// reference's order (x outer, then y, then z).
template <class Emit>
__device__ inline void visit_sub_cells(double px, double py, double pz, Grid g, uint32_t key, uint32_t level,
                                       double radius, Emit emit) {
    const uint32_t xySize = g.sizeX * g.sizeY;
    const uint32_t mod = key % xySize;
    const uint32_t gz = key / xySize, gy = mod / g.sizeX, gx = mod % g.sizeX;
    const uint32_t apron = f2u_sat(floorf((float)radius));
    const double scw = (double)(g.cellSize / (2.0f * (float)(1u << level)));
    const double ox = ((px - (double)g.negX) - (double)((float)gx * g.cellSize)) + (double)apron * scw;
    const double oy = ((py - (double)g.negY) - (double)((float)gy * g.cellSize)) + (double)apron * scw;
    const double oz = ((pz - (double)g.negZ) - (double)((float)gz * g.cellSize)) + (double)apron * scw;
    const uint32_t sx = f2u_sat(floorf((float)(ox / scw)));
    const uint32_t sy = f2u_sat(floorf((float)(oy / scw)));
    const uint32_t sz = f2u_sat(floorf((float)(oz / scw)));
    const double half = scw / 2.0;
    const double rad = radius * scw;
    const uint32_t nv = 2u * apron + (2u << level);
    for (uint32_t x = sx - apron; x < sx + apron; ++x)
        for (uint32_t y = sy - apron; y < sy + apron; ++y)
            for (uint32_t z = sz - apron; z < sz + apron; ++z) {
                const double dx = ox - (double)x * scw, dy = oy - (double)y * scw, dz = oz - (double)z * scw;
                const double dxh = dx + half, dyh = dy + half, dzh = dz + half;
                const double xc = (double)sqrt_f32((float)(dx * dx + dyh * dyh + dzh * dzh));
                const double yc = (double)sqrt_f32((float)(dxh * dxh + dy * dy + dzh * dzh));
                const double zc = (double)sqrt_f32((float)(dxh * dxh + dyh * dyh + dz * dz));
                const uint32_t id = x + y * nv + z * nv * nv;
                if (xc < rad) emit(0, id);
                if (yc < rad) emit(1, id);
                if (zc < rad) emit(2, id);
            }
}

__global__ void __launch_bounds__(256)
sub_cell_count_kernel(const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
                      uint32_t n, Grid g, const uint32_t* __restrict__ keys, uint32_t level, double radius,
                      uint32_t* __restrict__ cx, uint32_t* __restrict__ cy, uint32_t* __restrict__ cz) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c[3] = {0u, 0u, 0u};
    visit_sub_cells(px[i], py[i], pz[i], g, keys[i], level, radius, [&](int d, uint32_t) { ++c[d]; });
    cx[i] = c[0];
    cy[i] = c[1];
    cz[i] = c[2];
}

__global__ void __launch_bounds__(256)
sub_cell_list_kernel(const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
                     uint32_t n, Grid g, const uint32_t* __restrict__ keys, uint32_t level, double radius,
                     const uint32_t* __restrict__ ix, const uint32_t* __restrict__ iy, const uint32_t* __restrict__ iz,
                     uint32_t* __restrict__ lx, uint32_t* __restrict__ ly, uint32_t* __restrict__ lz) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t at[3] = {i ? ix[i - 1] : 0u, i ? iy[i - 1] : 0u, i ? iz[i - 1] : 0u};
    const uint32_t end[3] = {ix[i], iy[i], iz[i]};
    uint32_t* out[3] = {lx, ly, lz};
    visit_sub_cells(px[i], py[i], pz[i], g, keys[i], level, radius, [&](int d, uint32_t id) {
        if (at[d] < end[d]) out[d][at[d]++] = id;  // the counts were made by the same visit
    });
}

constexpr uint32_t kNV = 512;  // voxels per node

// Node b's particle range and entry range of particle p.
__device__ inline uint32_t node_end(uint32_t b, uint32_t B, uint32_t n, const uint32_t* first) {
    return b + 1 < B ? first[b + 1] : n;
}

// Per node: list entries and distinct voxels of its particles (unscanned).
__global__ void __launch_bounds__(64)
particles_per_node_kernel(uint32_t B, uint32_t n, const uint32_t* __restrict__ first,
                          const uint32_t* __restrict__ touched, const uint32_t* __restrict__ cells,
                          uint32_t* __restrict__ nonZero, uint32_t* __restrict__ entries) {
    __shared__ uint32_t s_cnt[kNV];
    const uint32_t b = blockIdx.x, lane = threadIdx.x;
    for (uint32_t v = lane; v < kNV; v += 64u) s_cnt[v] = 0u;
    __syncthreads();
    const uint32_t p0 = first[b], pe = node_end(b, B, n, first);
    for (uint32_t p = p0 + lane; p < pe; p += 64u) {
        const uint32_t e1 = touched[p];
        for (uint32_t e = p ? touched[p - 1] : 0u; e < e1; ++e) {
            const uint32_t v = cells[e];
            if (v < kNV) atomicAdd(&s_cnt[v], 1u);
        }
    }
    __syncthreads();
    uint32_t sum = 0u, nz = 0u;
    for (uint32_t v = lane; v < kNV; v += 64u) {
        sum += s_cnt[v];
        nz += s_cnt[v] != 0u;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        sum += __shfl_xor(sum, d, 64);
        nz += __shfl_xor(nz, d, 64);
    }
    if (lane == 0) {
        nonZero[b] = nz;
        entries[b] = sum;
    }
}

__global__ void __launch_bounds__(64)
particle_list_kernel(uint32_t B, uint32_t n, const uint32_t* __restrict__ usedIncl,
                     const uint32_t* __restrict__ listIncl, const uint32_t* __restrict__ touched,
                     const uint32_t* __restrict__ cells, const uint32_t* __restrict__ first,
                     uint32_t* __restrict__ voxelIDs, uint32_t* __restrict__ listStart, uint32_t* __restrict__ lists) {
    __shared__ uint32_t s_off[kNV];
    __shared__ unsigned long long s_mask[kNV];
    const uint32_t b = blockIdx.x, lane = threadIdx.x;
    for (uint32_t v = lane; v < kNV; v += 64u) {
        s_off[v] = 0u;
        s_mask[v] = 0ull;
    }
    __syncthreads();
    const uint32_t p0 = first[b], pe = node_end(b, B, n, first);
    for (uint32_t p = p0 + lane; p < pe; p += 64u) {
        const uint32_t e1 = touched[p];
        for (uint32_t e = p ? touched[p - 1] : 0u; e < e1; ++e) {
            const uint32_t v = cells[e];
            if (v < kNV) atomicAdd(&s_off[v], 1u);
        }
    }
    __syncthreads();
    // exclusive scans over the 512 voxels (lane l owns voxels 8l .. 8l+7):
    // list offsets and used-voxel ranks
    uint32_t cnt[8], lsum = 0u, lused = 0u;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        cnt[k] = s_off[8u * lane + (uint32_t)k];
        lsum += cnt[k];
        lused += cnt[k] != 0u;
    }
    uint32_t incS = lsum, incU = lused;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t a = __shfl_up(incS, d, 64), u = __shfl_up(incU, d, 64);
        if ((int)lane >= d) {
            incS += a;
            incU += u;
        }
    }
    uint32_t offS = incS - lsum, rank = incU - lused;
    const uint32_t vBase = b ? usedIncl[b - 1] : 0u;
    const uint32_t lBase = b ? listIncl[b - 1] : 0u;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t v = 8u * lane + (uint32_t)k;
        s_off[v] = offS;
        if (cnt[k]) {
            voxelIDs[vBase + rank] = v;
            listStart[vBase + rank] = lBase + offS;
            ++rank;
        }
        offS += cnt[k];
    }
    __syncthreads();
    // lists: chunks of 64 particles, one per lane, in ascending order
    for (uint32_t c0 = p0; c0 < pe; c0 += 64u) {
        const uint32_t p = c0 + lane;
        const bool has = p < pe;
        const uint32_t eb = has ? (p ? touched[p - 1] : 0u) : 0u, ee = has ? touched[p] : 0u;
        for (uint32_t e = eb; e < ee; ++e) {
            const uint32_t v = cells[e];
            if (v < kNV) atomicOr(&s_mask[v], 1ull << lane);
        }
        __syncthreads();
        const unsigned long long below = (1ull << lane) - 1ull;
        for (uint32_t e = eb; e < ee; ++e) {
            const uint32_t v = cells[e];
            if (v < kNV) lists[lBase + s_off[v] + (uint32_t)__popcll(s_mask[v] & below)] = p;
        }
        __syncthreads();
        for (uint32_t v = lane; v < kNV; v += 64u) {
            s_off[v] += (uint32_t)__popcll(s_mask[v]);
            s_mask[v] = 0ull;
        }
        __syncthreads();
    }
}

}  // namespace

void launch_sub_cell_counts(const double* px, const double* py, const double* pz, uint32_t n, Grid g,
                            const uint32_t* keys, uint32_t level, double radius, uint32_t* cx, uint32_t* cy,
                            uint32_t* cz, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(sub_cell_count_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, px, py, pz, n, g, keys, level,
                 radius, cx, cy, cz);
    FLIP2_LAUNCH_CHECK();
}

void launch_sub_cell_lists(const double* px, const double* py, const double* pz, uint32_t n, Grid g,
                           const uint32_t* keys, uint32_t level, double radius, const uint32_t* ix, const uint32_t* iy,
                           const uint32_t* iz, uint32_t* lx, uint32_t* ly, uint32_t* lz, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(sub_cell_list_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, px, py, pz, n, g, keys, level, radius,
                 ix, iy, iz, lx, ly, lz);
    FLIP2_LAUNCH_CHECK();
}

void launch_particles_per_node(uint32_t B, uint32_t n, const uint32_t* first, const uint32_t* touchedIncl,
                               const uint32_t* cells, uint32_t* nonZero, uint32_t* entries, hipStream_t s) {
    if (!B) return;
    FLIP2_KERNEL(particles_per_node_kernel, dim3(B), dim3(64), 0, s, B, n, first, touchedIncl, cells, nonZero,
                 entries);
    FLIP2_LAUNCH_CHECK();
}

void launch_particle_lists(uint32_t B, uint32_t n, const uint32_t* usedIncl, const uint32_t* listIncl,
                           const uint32_t* touchedIncl, const uint32_t* cells, const uint32_t* first,
                           uint32_t* voxelIDs, uint32_t* listStart, uint32_t* lists, hipStream_t s) {
    if (!B) return;
    FLIP2_KERNEL(particle_list_kernel, dim3(B), dim3(64), 0, s, B, n, usedIncl, listIncl, touchedIncl, cells, first,
                 voxelIDs, listStart, lists);
    FLIP2_LAUNCH_CHECK();
}

}  // namespace rc
}  // namespace flip2
