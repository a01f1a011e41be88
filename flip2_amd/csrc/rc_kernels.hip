// Reference-compatible stage kernels (FP64) of the per-step particle
// pipeline: binning, node table, sparse voxel build, P2G, gravity, divergence,
// red-black SOR, pressure-gradient update, G2P, advection and solid
// relocation.  They restate the semantics of engine/particles.cu and
// engine/algorithms/{particleToGridFunctions,voxelSolveFunctions}.cu,
// including the behaviours the reference's outputs depend on, with a
// deterministic accumulation order.
//
// Layout: one workgroup per occupied node (coarse cell) for every per-node
// stage; the node's 8^3 voxel tile lives in LDS.  FMA contraction is off in
// this file so each product and sum is rounded exactly as written.
#pragma clang fp contract(off)

#include "rc_kernels.hpp"

#include <cmath>

namespace flip2 {
namespace rc {

namespace {

constexpr uint32_t N1 = kRcVoxels1D;
constexpr uint32_t NV = kRcVoxelsPerNode;

// double -> u32 with saturation (NaN/negative -> 0, >= 2^32 -> 0xffffffff),
// the semantics of a GPU round-toward-zero convert.
__device__ inline uint32_t sat_u32(double v) {
    if (!(v > 0.0)) return 0u;
    if (v >= 4294967296.0) return 0xffffffffu;
    return (uint32_t)v;
}

// distance() (particles.cu:84-86).
__device__ inline double ref_distance(double dx, double dy, double dz) {
    return sqrt(fma(dx, dx, fma(dy, dy, dz * dz)));
}

// weightFromDistance (particles.cu:370-377).
__device__ inline double ref_weight(double d, double r) { return (d > r) ? 0.0 : 1.0 - d / r; }

// calcSubCellWidth (voxelSolveFunctions.hu:10-12): float division.
__device__ inline double sub_cell_width(const Grid& g) {
    return (double)(g.cellSize / (float)(2u << kRcRefinement));
}

struct Frame {
    double pin[3];
    uint32_t sub[3];
    uint32_t g[3];
};

// Particle position relative to its node's apron corner and its sub-voxel
// (particles.cu:133-146 and copies).  The cell offset gridID*cellSize is an
// FP32 product in the reference.
__device__ inline Frame make_frame(const Grid& g, uint32_t key, double px, double py, double pz) {
    Frame f;
    const uint32_t xy = g.sizeX * g.sizeY;
    const uint32_t m = key % xy;
    f.g[2] = key / xy;
    f.g[1] = m / g.sizeX;
    f.g[0] = m % g.sizeX;
    const double scw = sub_cell_width(g);
    const double pos[3] = {px, py, pz};
    const float neg[3] = {g.negX, g.negY, g.negZ};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float cellOff = (float)f.g[d] * g.cellSize;
        f.pin[d] = ((pos[d] - (double)neg[d]) - (double)cellOff) + (double)kRcApron * scw;
        f.sub[d] = sat_u32(floor(f.pin[d] / scw));
    }
    return f;
}

// True when the reference's stencil loops `for (int x = s - 2; x < s + 2; ++x)`
// (compared against unsigned bounds) execute for all three dimensions.
__device__ inline bool stencil_runs(const Frame& f) {
    return f.sub[0] >= kRcApron && f.sub[1] >= kRcApron && f.sub[2] >= kRcApron &&
           f.sub[0] <= 0xfffffffdu && f.sub[1] <= 0xfffffffdu && f.sub[2] <= 0xfffffffdu;
}

__device__ inline uint32_t node_last(const uint32_t* first, uint32_t b, uint32_t B, uint32_t n) {
    return (b == B - 1) ? n : first[b + 1];
}

// Staggered sample point offsets of component c (particles.cu:449-460).
__device__ inline void stagger(int c, double hs, double& ox, double& oy, double& oz) {
    ox = (c == 0) ? 0.0 : hs;
    oy = (c == 1) ? 0.0 : hs;
    oz = (c == 2) ? 0.0 : hs;
}

}  // namespace

// rootCell (particleToGridFunctions.cu:19-45).
__global__ void root_cell_kernel(double* __restrict__ px, double* __restrict__ py, double* __restrict__ pz,
                                 uint32_t n, Grid g, uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float up[3] = {fmaf((float)g.sizeX, g.cellSize, g.negX), fmaf((float)g.sizeY, g.cellSize, g.negY),
                         fmaf((float)g.sizeZ, g.cellSize, g.negZ)};
    const float neg[3] = {g.negX, g.negY, g.negZ};
    double* P[3] = {px, py, pz};
    uint32_t c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double v = P[d][i];
        if (v < (double)neg[d]) { v = (double)neg[d]; P[d][i] = v; }
        else if (v > (double)up[d]) { v = (double)up[d]; P[d][i] = v; }
        c[d] = (uint32_t)floorf((float)((v - (double)neg[d]) / (double)g.cellSize));
    }
    keys[i] = c[0] + c[1] * g.sizeX + c[2] * g.sizeX * g.sizeY;
}

// Gather of the three position arrays through the sort permutation
// (particles.cu:69-71).
__global__ void gather3_kernel(const uint32_t* __restrict__ perm, uint32_t n, const double* __restrict__ ax,
                               const double* __restrict__ ay, const double* __restrict__ az,
                               double* __restrict__ bx, double* __restrict__ by, double* __restrict__ bz) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = perm[i];
    if (j >= n) return;
    bx[i] = ax[j];
    by[i] = ay[j];
    bz[i] = az[j];
}

// mapNodeIndicesToParticles (particleToGridFunctions.cu:329-341).
__global__ void map_nodes_kernel(const uint32_t* __restrict__ uniq, uint32_t n, uint32_t* __restrict__ first) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == 0) first[0] = 0;
    else if (uniq[i] != uniq[i - 1] && uniq[i] - 1 < n) first[uniq[i] - 1] = i;
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ a, uint32_t count, uint32_t value) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = value;
}

// getYDimPositionFirstInRow (particleToGridFunctions.cu:355-372): the row
// table stores the cell key of the first node of each row.
__global__ void row_table_kernel(uint32_t B, const uint32_t* __restrict__ first, const uint32_t* __restrict__ keys,
                                 uint32_t* __restrict__ rows, uint32_t sizeX) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const uint32_t prev = (b != 0) ? keys[first[b - 1]] : B;
    const uint32_t cur = keys[first[b]];
    if (prev / sizeX != cur / sizeX) rows[cur / sizeX] = cur;
}

// getNumberVoxelsUsedPerNode / generateVoxelIDs marking (particles.cu:88-178,
// 192-295): one wave per node.  Lane l owns stencil offset (l>>4, (l>>2)&3,
// l&3); particles of the node are visited in order.  Output: 512-bit used and
// solid masks and the used count per node.
__global__ void __launch_bounds__(64)
mark_voxels_kernel(uint32_t B, uint32_t n, Grid g, const uint32_t* __restrict__ keys,
                   const uint32_t* __restrict__ first, const double* __restrict__ px,
                   const double* __restrict__ py, const double* __restrict__ pz,
                   uint64_t* __restrict__ usedMask, uint64_t* __restrict__ solidMask,
                   uint32_t* __restrict__ counts, uint32_t* __restrict__ oobEvents) {
    __shared__ uint8_t s_flag[NV];
    __shared__ uint8_t s_solid[NV];
    const uint32_t b = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    // Wall voxels: node coordinates from keys[b] (a particle index used as a
    // node index) and the y index equal to the x index.
    {
        const uint32_t xy = g.sizeX * g.sizeY;
        const uint32_t gid = keys[b];
        const uint32_t gx = gid % g.sizeX, gy = (gid % xy) / g.sizeX, gz = gid / xy;
        const double fr = floor(kRcRadius);
        for (uint32_t i = lane; i < NV; i += 64) {
            const uint32_t vx = i % N1;
            const uint32_t vy = (i % N1 * N1) / N1;
            const uint32_t vz = i / (N1 * N1);
            const uint32_t gvx = gx * N1 + vx, gvy = gy * N1 + vy, gvz = gz * N1 + vz;
            bool wall = false;
            if ((double)gvx < fr || (double)gvy < fr || (double)gvz < fr) wall = true;
            else if ((double)gvx > (double)(g.sizeX * N1) + fr || (double)gvy > (double)(g.sizeY * N1) + fr ||
                     (double)gvz > (double)(g.sizeZ * N1) + fr)
                wall = true;
            s_flag[i] = wall ? 1 : 0;
            s_solid[i] = wall ? 1 : 0;
        }
    }
    __syncthreads();
    const uint32_t ox = lane >> 4, oy = (lane >> 2) & 3u, oz = lane & 3u;
    const double scw = sub_cell_width(g);
    const double hs = scw / 2.0;
    const double rs = kRcRadius * scw;
    const uint32_t last = node_last(first, b, B, n);
    uint32_t oob = 0;
    for (uint32_t idx = first[b]; idx < last; ++idx) {
        const Frame f = make_frame(g, keys[idx], px[idx], py[idx], pz[idx]);
        if (!stencil_runs(f)) continue;
        const uint32_t x = f.sub[0] - kRcApron + ox, y = f.sub[1] - kRcApron + oy, z = f.sub[2] - kRcApron + oz;
        const double dpx = f.pin[0] - (double)(int32_t)x * scw;
        const double dpy = f.pin[1] - (double)(int32_t)y * scw;
        const double dpz = f.pin[2] - (double)(int32_t)z * scw;
        const double xc = ref_distance(dpx, dpy + hs, dpz + hs);
        const double yc = ref_distance(dpx + hs, dpy, dpz + hs);
        const double zc = ref_distance(dpx + hs, dpy + hs, dpz);
        if (xc < rs || yc < rs || zc < rs) {
            if (x < N1 && y < N1 && z < N1) s_flag[x + y * N1 + z * N1 * N1] = 1;
            else ++oob;
        }
    }
    __syncthreads();
    uint32_t count = 0;
    for (uint32_t w = 0; w < NV / 64; ++w) {
        const uint64_t m = __ballot(s_flag[w * 64 + lane] != 0);
        const uint64_t sm = __ballot(s_solid[w * 64 + lane] != 0);
        if (lane == 0) {
            usedMask[(size_t)b * 8 + w] = m;
            solidMask[(size_t)b * 8 + w] = sm;
        }
        count += (uint32_t)__popcll(m);
    }
    if (lane == 0) counts[b] = count;
    if (oobEvents) {
        uint32_t t = oob;
        for (int dd = 32; dd >= 1; dd >>= 1) t += __shfl_xor(t, dd, 64);
        if (lane == 0 && t) atomicAdd(oobEvents, t);
    }
}

// generateVoxelIDs compaction (particles.cu:300-320): ascending local voxel
// ids and their solid flags into the node's segment.
__global__ void __launch_bounds__(64)
compact_voxels_kernel(const uint64_t* __restrict__ usedMask, const uint64_t* __restrict__ solidMask,
                      const uint32_t* __restrict__ nodeOff, uint32_t* __restrict__ voxelIDs,
                      char* __restrict__ solids) {
    const uint32_t b = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint32_t start = (b == 0) ? 0u : nodeOff[b - 1];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t prefix = 0;
    for (uint32_t w = 0; w < NV / 64; ++w) {
        const uint64_t m = usedMask[(size_t)b * 8 + w];
        const uint64_t sm = solidMask[(size_t)b * 8 + w];
        if ((m >> lane) & 1ull) {
            const uint32_t r = prefix + (uint32_t)__popcll(m & lt);
            voxelIDs[start + r] = w * 64 + lane;
            solids[start + r] = (char)((sm >> lane) & 1ull);
        }
        prefix += (uint32_t)__popcll(m);
    }
}

// gatherParticleVelsToVoxels (particles.cu:379-488), all three components.
// One wave per node; lane = stencil offset, particles in order, so every
// voxel accumulates its contributions in particle order.
__global__ void __launch_bounds__(64)
p2g_kernel(uint32_t B, uint32_t n, Grid g, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ first,
           const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
           const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
           const uint32_t* __restrict__ nodeOff, const uint32_t* __restrict__ voxelIDs,
           const char* __restrict__ solids, double* __restrict__ Ux, double* __restrict__ Uy,
           double* __restrict__ Uz) {
    __shared__ double s_u[3][NV];
    __shared__ double s_w[3][NV];
    __shared__ uint8_t s_sol[NV];
    const uint32_t b = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < NV; i += 64) {
        s_u[0][i] = 0.0; s_u[1][i] = 0.0; s_u[2][i] = 0.0;
        s_w[0][i] = 0.0; s_w[1][i] = 0.0; s_w[2][i] = 0.0;
        s_sol[i] = 0;
    }
    __syncthreads();
    const uint32_t vs = (b == 0) ? 0u : nodeOff[b - 1];
    const uint32_t ve = nodeOff[b];
    for (uint32_t k = vs + lane; k < ve; k += 64) s_sol[voxelIDs[k]] = (uint8_t)solids[k];
    __syncthreads();
    const uint32_t ox = lane >> 4, oy = (lane >> 2) & 3u, oz = lane & 3u;
    const double scw = sub_cell_width(g);
    const double hs = scw / 2.0;
    const double rs = kRcRadius * scw;
    const uint32_t last = node_last(first, b, B, n);
    const double* vel[3] = {vx, vy, vz};
    for (uint32_t idx = first[b]; idx < last; ++idx) {
        const Frame f = make_frame(g, keys[idx], px[idx], py[idx], pz[idx]);
        if (!stencil_runs(f)) continue;
        const uint32_t x = f.sub[0] - kRcApron + ox, y = f.sub[1] - kRcApron + oy, z = f.sub[2] - kRcApron + oz;
        if (!(x < N1 && y < N1 && z < N1)) continue;
        const uint32_t v = x + y * N1 + z * N1 * N1;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double sx, sy, sz;
            stagger(c, hs, sx, sy, sz);
            const double bx = (double)(int32_t)x * scw + sx;
            const double by = (double)(int32_t)y * scw + sy;
            const double bz = (double)(int32_t)z * scw + sz;
            const double w = ref_weight(ref_distance(f.pin[0] - bx, f.pin[1] - by, f.pin[2] - bz), rs);
            s_w[c][v] = s_w[c][v] + w;
            s_u[c][v] = s_u[c][v] + w * vel[c][idx];
        }
    }
    __syncthreads();
    double* out[3] = {Ux, Uy, Uz};
    for (uint32_t k = vs + lane; k < ve; k += 64) {
        const uint32_t id = voxelIDs[k];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double u = s_sol[id] ? 0.0 : s_u[c][id];
            out[c][k] = u / (s_w[c][id] + 0.0000001);
        }
    }
}

// applyGravityKernel / removeGravityKernel (voxelSolveFunctions.cu:160-175).
__global__ void gravity_kernel(uint32_t V, const char* __restrict__ solids, double* __restrict__ Uy, double gdt,
                               int remove) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V || solids[i]) return;
    Uy[i] = remove ? Uy[i] - gdt : Uy[i] + gdt;
}

namespace {

// getNeighboringApronCellData (voxelSolveFunctions.cu:63-121) for the three
// component tiles at once.
__device__ inline void add_neighbor(uint32_t nb, const uint32_t* nodeOff, const uint32_t* voxelIDs,
                                    const double* const* data, double (*tile)[NV], int xo, int yo, int zo) {
    const uint32_t n2 = N1 * N1;
    const uint32_t vs = (nb == 0) ? 0u : nodeOff[nb - 1];
    const uint32_t ve = nodeOff[nb];
    // In the z-negative branch the target index does not depend on the
    // source x (Q17), so several sources share a target: that phase runs on
    // one thread in source order.  Every other phase maps sources to
    // distinct targets.
    const bool serial = zo < 0;
    if (serial && threadIdx.x != 0) return;
    const uint32_t step = serial ? 1u : blockDim.x;
    for (uint32_t k = vs + (serial ? 0u : threadIdx.x); k < ve; k += step) {
        const uint32_t id = voxelIDs[k];
        const uint32_t sx = id % N1, sy = (id % n2) / N1, sz = id / n2;
        uint32_t tx = sx, ty = sy, tz = sz;
        int target = -1;
        if (xo < 0) {
            if (sx >= N1 + (uint32_t)xo) { tx = sx - N1 - (uint32_t)xo; target = (int)(tx + ty * N1 + tz * n2); }
        } else if (xo > 0) {
            if (sx < (uint32_t)xo) { tx = N1 - (uint32_t)xo + sx; target = (int)(tx + ty * N1 + tz * n2); }
        } else if (yo < 0) {
            if (sy >= N1 + (uint32_t)yo) { ty = sy - N1 - (uint32_t)yo; target = (int)(tx + ty * N1 + tz * n2); }
        } else if (yo > 0) {
            if (sy < (uint32_t)yo) { ty = N1 - (uint32_t)yo + sy; target = (int)(tx + ty * N1 + tz * n2); }
        } else if (zo < 0) {
            if (sz >= N1 + (uint32_t)zo) { tz = sz - N1 - (uint32_t)zo; target = (int)(tz + ty * N1 + tz * n2); }
        } else if (zo > 0) {
            if (sz < (uint32_t)zo) { tz = N1 - (uint32_t)zo + sz; target = (int)(tx + ty * N1 + tz * n2); }
        }
        if (target >= 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) tile[c][target] = tile[c][target] + data[c][k];
        }
    }
}

}  // namespace

// cudaCalcDivU (voxelSolveFunctions.cu:245-271) for all three components in
// one pass: loadVoxelDataForThisBlock (:97-158) per component tile, then the
// forward differences summed in x, y, z order.
__global__ void __launch_bounds__(256)
divu_kernel(uint32_t B, Grid g, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ first,
            const uint32_t* __restrict__ rows, const uint32_t* __restrict__ nodeOff,
            const uint32_t* __restrict__ voxelIDs, const double* __restrict__ Ux, const double* __restrict__ Uy,
            const double* __restrict__ Uz, double* __restrict__ divU) {
    __shared__ double s_t[3][NV];
    const uint32_t b = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < NV; i += blockDim.x) { s_t[0][i] = 0.0; s_t[1][i] = 0.0; s_t[2][i] = 0.0; }
    __syncthreads();
    const double* data[3] = {Ux, Uy, Uz};
    const uint32_t vs = (b == 0) ? 0u : nodeOff[b - 1];
    const uint32_t ve = nodeOff[b];
    for (uint32_t k = vs + tid; k < ve; k += blockDim.x) {
        const uint32_t id = voxelIDs[k];
        for (int c = 0; c < 3; ++c) s_t[c][id] = s_t[c][id] + data[c][k];
    }
    const uint32_t cur = keys[first[b]];
    const uint32_t xy = g.sizeX * g.sizeY;
    const uint32_t curX = cur % g.sizeX, curY = (cur % xy) / g.sizeX, curZ = cur / xy;
    const int32_t xl = (int32_t)curX - 1, xr = (int32_t)curX + 1;
    const int32_t yl = (int32_t)curY - 1, yr = (int32_t)curY + 1;
    const int32_t zl = (int32_t)curZ - 1, zr = (int32_t)curZ + 1;
    const int apron = 2 * (int)(uint32_t)kRcRadius;
    uint32_t nb;
    __syncthreads();
    if (xl >= 0 && b > 0 && keys[first[b - 1]] == cur - 1) add_neighbor(b - 1, nodeOff, voxelIDs, data, s_t, -apron, 0, 0);
    __syncthreads();
    if ((uint32_t)xr < g.sizeX && b < B - 1 && keys[first[b + 1]] == cur + 1)
        add_neighbor(b + 1, nodeOff, voxelIDs, data, s_t, apron, 0, 0);
    __syncthreads();
    if (yl >= 0 && (nb = rows[(uint32_t)yl + curZ * g.sizeY]) < B) add_neighbor(nb, nodeOff, voxelIDs, data, s_t, 0, -apron, 0);
    __syncthreads();
    if ((uint32_t)yr < g.sizeY && (nb = rows[(uint32_t)yr + curZ * g.sizeY]) < B)
        add_neighbor(nb, nodeOff, voxelIDs, data, s_t, 0, apron, 0);
    __syncthreads();
    if (zl >= 0 && (nb = rows[(uint32_t)zl * g.sizeY + curY]) < B) add_neighbor(nb, nodeOff, voxelIDs, data, s_t, 0, 0, -apron);
    __syncthreads();
    if ((uint32_t)zr < g.sizeZ && (nb = rows[(uint32_t)zr * g.sizeY + curY]) < B)
        add_neighbor(nb, nodeOff, voxelIDs, data, s_t, 0, 0, apron);
    __syncthreads();
    for (uint32_t k = vs + tid; k < ve; k += blockDim.x) {
        const uint32_t i = voxelIDs[k];
        const double dx = (i % N1 < N1 - 1) ? s_t[0][i + 1] - s_t[0][i] : 0.0;
        const double dy = (i % (N1 * N1) / N1 < N1 - 1) ? s_t[1][i + N1] - s_t[1][i] : 0.0;
        const double dz = (i / (N1 * N1) < N1 - 1) ? s_t[2][i + N1 * N1] - s_t[2][i] : 0.0;
        double acc = 0.0;
        acc = acc + dx;
        acc = acc + dy;
        acc = acc + dz;
        divU[k] = acc;
    }
}

namespace {

// isApronCell (voxelSolveFunctions.cu:275-284): one layer, "y" == x index.
__device__ inline bool is_apron(uint32_t id) {
    const uint32_t x = id % N1;
    const uint32_t y = (id % N1 * N1) / N1;
    const uint32_t z = id / (N1 * N1);
    const uint32_t rb = N1 - 1u;
    return x < 1u || x >= rb || y < 1u || y >= rb || z < 1u || z >= rb;
}

// isRed (voxelSolveFunctions.cu:286-291).
__device__ inline bool is_red(uint32_t id) {
    const uint32_t x = id % N1, y = (id % (N1 * N1)) / N1, z = id / (N1 * N1);
    return !((x + y + z) & 1u);
}

}  // namespace

// `iters` red/black iterations of GSiteration (voxelSolveFunctions.cu:293-389)
// with the node tile resident in LDS.  Each thread owns voxel slots tid and
// tid+256.  A half-sweep computes every new value from the tile as it was at
// the start of the half-sweep, then stores them.  The residual written is the
// one of the last (black) half-sweep: non-updated voxels hold 0.
__global__ void __launch_bounds__(256)
sor_kernel(const uint32_t* __restrict__ nodeOff, const uint32_t* __restrict__ voxelIDs,
           const char* __restrict__ solids, const double* __restrict__ divU, double* __restrict__ p,
           double* __restrict__ residuals, double scale, double w, uint32_t iters) {
    __shared__ double s_p[NV];
    __shared__ double s_d[NV];
    __shared__ uint8_t s_sol[NV];
    const uint32_t b = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < NV; i += 256) { s_p[i] = 0.0; s_d[i] = 0.0; s_sol[i] = 0; }
    __syncthreads();
    const uint32_t vs = (b == 0) ? 0u : nodeOff[b - 1];
    const uint32_t cnt = nodeOff[b] - vs;
    for (uint32_t k = tid; k < cnt; k += 256) {
        const uint32_t id = voxelIDs[vs + k];
        s_p[id] = p[vs + k];
        s_d[id] = divU[vs + k];
        s_sol[id] = (uint8_t)solids[vs + k];
    }
    __syncthreads();

    // Per-slot coefficients (divU and solids are constant during the solve).
    uint32_t id[2];
    int colour[2];  // 1 red, 0 black, -1 not updated
    double A[2][6], Adiag[2], dv[2], res[2];
    const int32_t off[6] = {-1, 1, -(int32_t)N1, (int32_t)N1, -(int32_t)(N1 * N1), (int32_t)(N1 * N1)};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const uint32_t k = tid + 256u * s;
        colour[s] = -1;
        id[s] = 0;
        Adiag[s] = 0.0;
        dv[s] = 0.0;
        res[s] = 0.0;
        for (int j = 0; j < 6; ++j) A[s][j] = 0.0;
        if (k < cnt) {
            const uint32_t v = voxelIDs[vs + k];
            id[s] = v;
            dv[s] = s_d[v];
            if (!s_sol[v] && !is_apron(v)) {
                colour[s] = is_red(v) ? 1 : 0;
                double ad = 0.0;
                for (int j = 0; j < 6; ++j) {
                    const uint32_t nbv = (uint32_t)((int32_t)v + off[j]);
                    if (!s_sol[nbv]) {
                        if (s_d[nbv] != 0.0) A[s][j] = -scale;
                        ad = ad + scale;
                    }
                }
                Adiag[s] = ad;
            }
        }
    }

    for (uint32_t it = 0; it < iters; ++it) {
        for (int red = 1; red >= 0; --red) {
            double pn[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                res[s] = 0.0;
                pn[s] = 0.0;
                if (colour[s] == red) {
                    const uint32_t v = id[s];
                    const double sum = A[s][0] * s_p[v - 1] + A[s][1] * s_p[v + 1] + A[s][2] * s_p[v - N1] +
                                       A[s][3] * s_p[v + N1] + A[s][4] * s_p[v - N1 * N1] +
                                       A[s][5] * s_p[v + N1 * N1];
                    const double pv = s_p[v];
                    const double dg = Adiag[s] + 0.000000001;
                    const double q = pv + w * ((-dv[s] - sum) / dg - pv);
                    pn[s] = q;
                    res[s] = (-dv[s] - dg * q + sum);
                }
            }
            __syncthreads();
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (colour[s] == red) s_p[id[s]] = pn[s];
            __syncthreads();
        }
    }
    for (int s = 0; s < 2; ++s) {
        const uint32_t k = tid + 256u * s;
        if (k < cnt) {
            p[vs + k] = s_p[id[s]];
            if (iters) residuals[vs + k] = res[s];
        }
    }
}

// getMaxFromArray/getMax with abs=true (cudaVec.cu:173-233) on `count`
// arrays at once (one workgroup per array): level 0 over the data, then the
// reference's recursion over the tile results.  tmp holds 2 * (n/512+1)
// doubles per array.
__global__ void __launch_bounds__(256)
sampled_absmax_kernel(Ptr3 arrays, uint32_t n, double* __restrict__ tmp, double* __restrict__ result) {
    const uint32_t a = blockIdx.x;
    const double* in = arrays.p[a];
    const uint32_t k0 = n / 512u + 1u;
    double* bufA = tmp + (size_t)a * 2 * k0;
    double* bufB = bufA + k0;
    for (uint32_t t = threadIdx.x; t < k0; t += blockDim.x) {
        const uint64_t i0 = (uint64_t)t * 512u, i1 = i0 + 256u;
        const double x = (i0 < n) ? in[i0] : 0.0;
        const double y = (i1 < n) ? in[i1] : 0.0;
        bufA[t] = (fabs(x) > fabs(y)) ? x : y;
    }
    __syncthreads();
    for (uint32_t i = k0; i > 1; i = i / 512u + 1u) {
        const uint32_t outN = i / 512u + 1u;
        for (uint32_t t = threadIdx.x; t < outN; t += blockDim.x) {
            const uint64_t i0 = (uint64_t)t * 512u, i1 = i0 + 256u;
            const double x = (i0 < i) ? bufA[i0] : 0.0;
            const double y = (i1 < i) ? bufA[i1] : 0.0;
            bufB[t] = (fabs(x) > fabs(y)) ? x : y;
        }
        __syncthreads();
        double* t = bufA; bufA = bufB; bufB = t;
    }
    if (threadIdx.x == 0) result[a] = bufA[0];
}

// pressureToAcceleration (voxelSolveFunctions.cu:415-479), three components.
__global__ void __launch_bounds__(256)
velocity_update_kernel(const uint32_t* __restrict__ nodeOff, const uint32_t* __restrict__ voxelIDs,
                       const char* __restrict__ solids, const double* __restrict__ p, double* __restrict__ Ux,
                       double* __restrict__ Uy, double* __restrict__ Uz, double scale) {
    __shared__ double s_p[NV];
    __shared__ uint8_t s_sol[NV];
    const uint32_t b = blockIdx.x;
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < NV; i += blockDim.x) { s_p[i] = 0.0; s_sol[i] = 0; }
    __syncthreads();
    const uint32_t vs = (b == 0) ? 0u : nodeOff[b - 1];
    const uint32_t ve = nodeOff[b];
    for (uint32_t k = vs + tid; k < ve; k += blockDim.x) {
        const uint32_t id = voxelIDs[k];
        s_p[id] = p[k];
        s_sol[id] = (uint8_t)solids[k];
    }
    __syncthreads();
    double* U[3] = {Ux, Uy, Uz};
    const uint32_t stride[3] = {1u, N1, N1 * N1};
    for (uint32_t k = vs + tid; k < ve; k += blockDim.x) {
        const uint32_t id = voxelIDs[k];
        if (s_sol[id] || is_apron(id)) continue;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double dp = 0.0;
            if (!s_sol[id - stride[c]]) dp = s_p[id] - s_p[id - stride[c]];
            U[c][k] = U[c][k] - dp * scale;
        }
    }
}

// gatherVoxelVelsToParticles (particles.cu:562-662), three components.
// Lane = particle; the stencil is summed in the reference's loop order.
__global__ void __launch_bounds__(64)
g2p_kernel(uint32_t B, uint32_t n, Grid g, const uint32_t* __restrict__ keys, const uint32_t* __restrict__ first,
           const double* __restrict__ px, const double* __restrict__ py, const double* __restrict__ pz,
           double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
           const uint32_t* __restrict__ nodeOff, const uint32_t* __restrict__ voxelIDs,
           const double* __restrict__ Ux, const double* __restrict__ Uy, const double* __restrict__ Uz) {
    __shared__ double s_t[3][NV];
    const uint32_t b = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < NV; i += 64) { s_t[0][i] = 0.0; s_t[1][i] = 0.0; s_t[2][i] = 0.0; }
    __syncthreads();
    const uint32_t vs = (b == 0) ? 0u : nodeOff[b - 1];
    const uint32_t ve = nodeOff[b];
    for (uint32_t k = vs + lane; k < ve; k += 64) {
        const uint32_t id = voxelIDs[k];
        s_t[0][id] = Ux[k]; s_t[1][id] = Uy[k]; s_t[2][id] = Uz[k];
    }
    __syncthreads();
    const double scw = sub_cell_width(g);
    const double hs = scw / 2.0;
    const double rs = kRcRadius * scw;
    const uint32_t last = node_last(first, b, B, n);
    double* vel[3] = {vx, vy, vz};
    for (uint32_t idx = first[b] + lane; idx < last; idx += 64) {
        const Frame f = make_frame(g, keys[idx], px[idx], py[idx], pz[idx]);
        const bool runs = stencil_runs(f);
        for (int c = 0; c < 3; ++c) {
            double sx, sy, sz;
            stagger(c, hs, sx, sy, sz);
            double pv = 0.0, ws = 0.0;
            if (runs) {
                for (uint32_t x = f.sub[0] - kRcApron; x < N1 && x < f.sub[0] + kRcApron; ++x)
                    for (uint32_t y = f.sub[1] - kRcApron; y < N1 && y < f.sub[1] + kRcApron; ++y)
                        for (uint32_t z = f.sub[2] - kRcApron; z < N1 && z < f.sub[2] + kRcApron; ++z) {
                            const double bx = (double)(int32_t)x * scw + sx;
                            const double by = (double)(int32_t)y * scw + sy;
                            const double bz = (double)(int32_t)z * scw + sz;
                            const double wt =
                                ref_weight(ref_distance(f.pin[0] - bx, f.pin[1] - by, f.pin[2] - bz), rs);
                            const double u = s_t[c][x + y * N1 + z * N1 * N1];
                            if (u != 0.0) {
                                ws = ws + wt;
                                pv = pv + wt * u;
                            }
                        }
            }
            vel[c][idx] = pv / (ws + 0.00000001);
        }
    }
}

// advectParticlePositions (particles.cu:664-669), three components fused.
__global__ void advect_kernel(uint32_t n, double dt, double* __restrict__ px, double* __restrict__ py,
                              double* __restrict__ pz, const double* __restrict__ vx, const double* __restrict__ vy,
                              const double* __restrict__ vz) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    px[i] = px[i] + dt * vx[i];
    py[i] = py[i] + dt * vy[i];
    pz[i] = pz[i] + dt * vz[i];
}

// Philox4x32-10 (Salmon et al., SC'11).
__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        const uint32_t n3 = (uint32_t)p0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        if (r != 9) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    }
}

__global__ void philox_kat_kernel(const uint32_t* __restrict__ in, uint32_t count, uint32_t* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t c[4] = {in[i * 6 + 0], in[i * 6 + 1], in[i * 6 + 2], in[i * 6 + 3]};
    philox4x32_10(c, in[i * 6 + 4], in[i * 6 + 5]);
    for (int j = 0; j < 4; ++j) out[i * 4 + j] = c[j];
}

// moveSolids (particles.cu:674-740): particles whose sub-voxel is outside
// the tile or solid move into the node's first solid voxel (ascending id;
// voxel 512 when there is none) plus a Philox jitter.
__global__ void __launch_bounds__(64)
move_solids_kernel(uint32_t B, uint32_t n, Grid g, const uint32_t* __restrict__ keys,
                   const uint32_t* __restrict__ first, double* __restrict__ px, double* __restrict__ py,
                   double* __restrict__ pz, const uint32_t* __restrict__ nodeOff,
                   const uint32_t* __restrict__ voxelIDs, const char* __restrict__ solids, uint64_t counter,
                   uint32_t* __restrict__ relocated) {
    __shared__ uint8_t s_sol[NV];
    const uint32_t b = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < NV; i += 64) s_sol[i] = 0;
    __syncthreads();
    const uint32_t vs = (b == 0) ? 0u : nodeOff[b - 1];
    const uint32_t ve = nodeOff[b];
    uint32_t firstSolid = NV;
    for (uint32_t base = vs; base < ve; base += 64) {
        const uint32_t k = base + lane;
        const bool sol = (k < ve) && solids[k] != 0;
        if (k < ve) s_sol[voxelIDs[k]] = sol ? 1 : 0;
        const uint64_t m = __ballot(sol);
        if (m && firstSolid == NV) firstSolid = voxelIDs[base + (uint32_t)__builtin_ctzll(m)];
    }
    __syncthreads();
    const double scw = sub_cell_width(g);
    const int32_t apron = (int32_t)kRcApron;
    const int32_t nx = (int32_t)(firstSolid % N1);
    const int32_t ny = (int32_t)((firstSolid % (N1 * N1)) / N1);
    const int32_t nz = (int32_t)(firstSolid / (N1 * N1));
    const uint32_t last = node_last(first, b, B, n);
    uint32_t moved = 0;
    for (uint32_t idx = first[b] + lane; idx < last; idx += 64) {
        const Frame f = make_frame(g, keys[idx], px[idx], py[idx], pz[idx]);
        const bool out = f.sub[0] >= N1 || f.sub[1] >= N1 || f.sub[2] >= N1;
        if (!out && !s_sol[f.sub[0] + f.sub[1] * N1 + f.sub[2] * N1 * N1]) continue;
        uint32_t c[4] = {idx, (uint32_t)counter, (uint32_t)(counter >> 32), 0u};
        philox4x32_10(c, 1291293829u, 0u);
        const float bxf = fmaf((float)f.g[0], g.cellSize, g.negX);
        const float byf = fmaf((float)f.g[1], g.cellSize, g.negY);
        const float bzf = fmaf((float)f.g[2], g.cellSize, g.negZ);
        const double inv = 1.0 / 4294967296.0;
        px[idx] = ((double)bxf + (double)(nx - apron) * scw) + (((double)c[0] + 1.0) * inv) * scw;
        py[idx] = ((double)byf + (double)(ny - apron) * scw) + (((double)c[1] + 1.0) * inv) * scw;
        pz[idx] = ((double)bzf + (double)(nz - apron) * scw) + (((double)c[2] + 1.0) * inv) * scw;
        ++moved;
    }
    if (relocated && moved) atomicAdd(relocated, moved);
}

}  // namespace rc
}  // namespace flip2

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace flip2 {
namespace rc {

namespace {
inline dim3 grid1d(uint32_t n, uint32_t bs) { return dim3(ceil_div(n ? n : 1, bs)); }
}  // namespace

void launch_root_cell(double* px, double* py, double* pz, uint32_t n, Grid g, uint32_t* keys, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(root_cell_kernel, grid1d(n, 256), dim3(256), 0, s, px, py, pz, n, g, keys);
    FLIP2_LAUNCH_CHECK();
}

void launch_gather3(const uint32_t* perm, uint32_t n, const double* ax, const double* ay, const double* az,
                    double* bx, double* by, double* bz, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(gather3_kernel, grid1d(n, 256), dim3(256), 0, s, perm, n, ax, ay, az, bx, by, bz);
    FLIP2_LAUNCH_CHECK();
}

void launch_map_nodes(const uint32_t* uniq, uint32_t n, uint32_t* first, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(map_nodes_kernel, grid1d(n, 256), dim3(256), 0, s, uniq, n, first);
    FLIP2_LAUNCH_CHECK();
}

void launch_fill_u32(uint32_t* a, uint32_t count, uint32_t value, hipStream_t s) {
    if (!count) return;
    FLIP2_KERNEL(fill_u32_kernel, grid1d(count, 256), dim3(256), 0, s, a, count, value);
    FLIP2_LAUNCH_CHECK();
}

void launch_row_table(uint32_t B, const uint32_t* first, const uint32_t* keys, uint32_t* rows, uint32_t sizeX,
                      hipStream_t s) {
    if (!B) return;
    FLIP2_KERNEL(row_table_kernel, grid1d(B, 256), dim3(256), 0, s, B, first, keys, rows, sizeX);
    FLIP2_LAUNCH_CHECK();
}

void launch_mark_voxels(uint32_t B, uint32_t n, Grid g, const uint32_t* keys, const uint32_t* first,
                        const double* px, const double* py, const double* pz, uint64_t* usedMask,
                        uint64_t* solidMask, uint32_t* counts, uint32_t* oobEvents, hipStream_t s) {
    if (!B) return;
    FLIP2_KERNEL(mark_voxels_kernel, dim3(B), dim3(64), 0, s, B, n, g, keys, first, px, py, pz, usedMask,
                       solidMask, counts, oobEvents);
    FLIP2_LAUNCH_CHECK();
}

void launch_compact_voxels(uint32_t B, const uint64_t* usedMask, const uint64_t* solidMask, const uint32_t* nodeOff,
                           uint32_t* voxelIDs, char* solids, hipStream_t s) {
    if (!B) return;
    FLIP2_KERNEL(compact_voxels_kernel, dim3(B), dim3(64), 0, s, usedMask, solidMask, nodeOff, voxelIDs,
                       solids);
    FLIP2_LAUNCH_CHECK();
}

void launch_p2g(uint32_t B, uint32_t n, Grid g, const uint32_t* keys, const uint32_t* first, const double* px,
                const double* py, const double* pz, const double* vx, const double* vy, const double* vz,
                const uint32_t* nodeOff, const uint32_t* voxelIDs, const char* solids, double* Ux, double* Uy,
                double* Uz, hipStream_t s) {
    if (!B) return;
    FLIP2_KERNEL(p2g_kernel, dim3(B), dim3(64), 0, s, B, n, g, keys, first, px, py, pz, vx, vy, vz, nodeOff,
                       voxelIDs, solids, Ux, Uy, Uz);
    FLIP2_LAUNCH_CHECK();
}

void launch_gravity(uint32_t V, const char* solids, double* Uy, double gdt, bool remove, hipStream_t s) {
    if (!V) return;
    FLIP2_KERNEL(gravity_kernel, grid1d(V, 256), dim3(256), 0, s, V, solids, Uy, gdt, remove ? 1 : 0);
    FLIP2_LAUNCH_CHECK();
}

void launch_divu(uint32_t B, Grid g, const uint32_t* keys, const uint32_t* first, const uint32_t* rows,
                 const uint32_t* nodeOff, const uint32_t* voxelIDs, const double* Ux, const double* Uy,
                 const double* Uz, double* divU, hipStream_t s) {
    if (!B) return;
    FLIP2_KERNEL(divu_kernel, dim3(B), dim3(256), 0, s, B, g, keys, first, rows, nodeOff, voxelIDs, Ux, Uy,
                       Uz, divU);
    FLIP2_LAUNCH_CHECK();
}

void launch_sor(uint32_t B, const uint32_t* nodeOff, const uint32_t* voxelIDs, const char* solids,
                const double* divU, double* p, double* residuals, double scale, double w, uint32_t iters,
                hipStream_t s) {
    if (!B) return;
    FLIP2_KERNEL(sor_kernel, dim3(B), dim3(256), 0, s, nodeOff, voxelIDs, solids, divU, p, residuals, scale,
                       w, iters);
    FLIP2_LAUNCH_CHECK();
}

size_t sampled_absmax_tmp_doubles(uint32_t n, int count) {
    return (size_t)count * 2 * ((size_t)n / 512u + 1u);
}

void launch_sampled_absmax(Ptr3 arrays, int count, uint32_t n, double* tmp, double* result, hipStream_t s) {
    FLIP2_REQUIRE(count >= 1 && count <= 3, kErrArgument, "sampled max: 1..3 arrays");
    FLIP2_KERNEL(sampled_absmax_kernel, dim3(count), dim3(256), 0, s, arrays, n, tmp, result);
    FLIP2_LAUNCH_CHECK();
}

void launch_velocity_update(uint32_t B, const uint32_t* nodeOff, const uint32_t* voxelIDs, const char* solids,
                            const double* p, double* Ux, double* Uy, double* Uz, double scale, hipStream_t s) {
    if (!B) return;
    FLIP2_KERNEL(velocity_update_kernel, dim3(B), dim3(256), 0, s, nodeOff, voxelIDs, solids, p, Ux, Uy, Uz,
                       scale);
    FLIP2_LAUNCH_CHECK();
}

void launch_g2p(uint32_t B, uint32_t n, Grid g, const uint32_t* keys, const uint32_t* first, const double* px,
                const double* py, const double* pz, double* vx, double* vy, double* vz, const uint32_t* nodeOff,
                const uint32_t* voxelIDs, const double* Ux, const double* Uy, const double* Uz, hipStream_t s) {
    if (!B) return;
    FLIP2_KERNEL(g2p_kernel, dim3(B), dim3(64), 0, s, B, n, g, keys, first, px, py, pz, vx, vy, vz, nodeOff,
                       voxelIDs, Ux, Uy, Uz);
    FLIP2_LAUNCH_CHECK();
}

void launch_advect(uint32_t n, double dt, double* px, double* py, double* pz, const double* vx, const double* vy,
                   const double* vz, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(advect_kernel, grid1d(n, 256), dim3(256), 0, s, n, dt, px, py, pz, vx, vy, vz);
    FLIP2_LAUNCH_CHECK();
}

void launch_move_solids(uint32_t B, uint32_t n, Grid g, const uint32_t* keys, const uint32_t* first, double* px,
                        double* py, double* pz, const uint32_t* nodeOff, const uint32_t* voxelIDs,
                        const char* solids, uint64_t counter, uint32_t* relocated, hipStream_t s) {
    if (!B) return;
    FLIP2_KERNEL(move_solids_kernel, dim3(B), dim3(64), 0, s, B, n, g, keys, first, px, py, pz, nodeOff,
                       voxelIDs, solids, counter, relocated);
    FLIP2_LAUNCH_CHECK();
}

void launch_philox_kat(const uint32_t* in, uint32_t count, uint32_t* out, hipStream_t s) {
    if (!count) return;
    FLIP2_KERNEL(philox_kat_kernel, grid1d(count, 64), dim3(64), 0, s, in, count, out);
    FLIP2_LAUNCH_CHECK();
}

}  // namespace rc
}  // namespace flip2
