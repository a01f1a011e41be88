// Kernels of the learned preconditioner: a 3-level 3D U-Net on a dense
// n^3 grid, channels-last activations ([z][y][x][C]).
//
// Forward convolutions are implicit GEMMs on MFMA (v_mfma_f32_16x16x32_bf16):
// BF16 weights and activations, FP32 accumulation, bias + ReLU fused in the
// epilogue.  D[cout][voxel] = sum_{tap,cin} W[tap][cout][cin] X[voxel+tap][cin]
// with M = 16 output channels, N = 16 voxels, K = 2 taps x 16 input channels
// per instruction.  A workgroup owns a 16 x 4 x 4 block of output voxels
// (whose input, with halo, is staged in LDS), each wave a 16 x 4 slice.
//
// Backward: the data gradient is the same kernel on the BF16 gradient with
// transposed, mirrored weights (FP32 output); the weight gradient is a GEMM
// over the voxels on MFMA (BF16 operands, FP32 accumulation, FP32 atomics per
// chunk); bias gradients are FP32 column sums.
#include "unet_kernels.hpp"

namespace flip2 {
namespace nn {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// Round to BF16, nearest even: gfx950's v_cvt_pk_bf16_f32 (two values per
// instruction when the compiler pairs them, as in f2bf2).
__device__ inline uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
__device__ inline uint32_t f2bf2(float lo, float hi) {
    const v2bf v = {(__bf16)lo, (__bf16)hi};
    return __builtin_bit_cast(uint32_t, v);
}
// Four values as BF16, packed low to high; all zero unless keep (a mask,
// not a branch around the conversions).
__device__ inline uint2 f2bf4(float a, float b, float c, float d, bool keep) {
    const uint32_t m = keep ? 0xffffffffu : 0u;
    return make_uint2(f2bf2(a, b) & m, f2bf2(c, d) & m);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// ---------------------------------------------------------------------------
// conv3x3x3 forward, zero padding
// ---------------------------------------------------------------------------
// MODE 0: Y (BF16) = act(conv + bias); MODE 1: Yf (FP32) = conv (no bias, no
// activation), used for the data gradient with transposed, mirrored weights.
//
// CA > 0: the input is the decoder's concatenation [up(Xc), X] read in
// place: channels [0, CA) from the half-resolution tensor Xc (nearest
// upsampling, voxel (x, y, z) reads (x/2, y/2, z/2)), channels [CA, CIN) from
// X; the concatenated tensor is never written.
//
// head.out != nullptr (COUT == 16): the 1x1 output head is applied in the
// epilogue, out[v] = mask[v] ? (b + sum_c w[c] Y[v][c]) * scale : 0 with the
// BF16-rounded Y and the same summation order as head_fwd_kernel; Y may then
// be null.
//
// Tile order: workgroups are dealt to the 8 XCDs round robin by linear id, so
// the id is remapped to give each XCD one contiguous range of tiles (z-slabs)
// and the halo rows that neighbouring tiles share are re-read from that XCD's
// L2 instead of HBM.
//
// A workgroup computes a 16 (x) x 4 (y) x 4 (z) block of output voxels;
// wave w owns the z-slice w as four MFMA columns, one 16-voxel x-row each.
// The input block with its one-voxel halo (18 x 6 x 6 voxels) is staged in
// LDS 16 channels at a time, as two planes of 8 channels ([chunk][voxel], 16
// bytes per voxel), so each input voxel is fetched from memory once per
// workgroup instead of once per tap.  MFMA is v_mfma_f32_16x16x32_bf16 with
// K = two taps x 16 channels (taps that share dy, so one B fragment serves
// three output rows).  A B fragment is one
// 16-byte LDS read, and the 16 lanes of a quarter wave read 16 consecutive
// voxels of one plane: 256 contiguous bytes, which cover the LDS banks once
// (with [voxel][16 channels] rows and 8-byte reads, bank conflicts made LDS
// the bound: ~16 stall cycles per read).  Weights (A fragments, 16 bytes per
// lane) are read from global memory, where every wave of the GPU shares them
// through the caches.
constexpr int kTX = 16, kTY = 4, kTZ = 4;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2, kHalo = kHX * kHY * kHZ;  // 648 voxels
constexpr int kStageLoads = (kHalo * 2 + 255) / 256;  // 16-byte loads per thread per input stage (6)

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

inline dim3 conv_grid(int n) {
    return dim3((unsigned)((n + kTX - 1) / kTX), (unsigned)((n + kTY - 1) / kTY), (unsigned)((n + kTZ - 1) / kTZ));
}

// Workgroup id -> tile origin (and output-channel group mg when SPLIT
// workgroups share a tile: grid.x = tiles in x * SPLIT), each XCD taking a
// contiguous range of tiles.
template <int SPLIT = 1>
__device__ inline void conv_tile(int& bx, int& by, int& bz, int& mg) {
    const uint32_t gx = gridDim.x, gxy = gridDim.x * gridDim.y, total = gxy * gridDim.z;
    uint32_t lin = blockIdx.x + gx * blockIdx.y + gxy * blockIdx.z;
    if ((total & 7u) == 0) lin = (lin & 7u) * (total >> 3) + (lin >> 3);
    mg = (int)(lin % SPLIT);
    const uint32_t tile = lin / SPLIT, tilesX = gx / SPLIT;
    bx = (int)(tile % tilesX) * kTX;
    by = (int)((tile / tilesX) % gridDim.y) * kTY;
    bz = (int)(tile / (tilesX * gridDim.y)) * kTZ;
}

// Tiles without masked-in voxels: skipped, or cleared when they may hold
// values from an earlier launch.  Returns true when the tile is not computed.
// With SPLIT workgroups per tile each clears its own channel group.
template <int COUT, int SPLIT = 1>
__device__ inline bool skip_tile(const ConvMask& cm, int bx, int by, int bz, int mg, int n, uint16_t* Y,
                                 float* headOut) {
    if (!cm.tiles) return false;
    const uint32_t tilesX = gridDim.x / SPLIT;
    const uint32_t t = ((uint32_t)(bz / kTZ) * gridDim.y + (uint32_t)(by / kTY)) * tilesX + (uint32_t)(bx / kTX);
    if (cm.tiles[t]) return false;
    if (cm.prevTiles && cm.prevTiles[t]) {
        const int x = bx + (threadIdx.x & 15), y = by + ((threadIdx.x >> 4) & 3), z = bz + (threadIdx.x >> 6);
        if (x < n && y < n && z < n) {
            const long long v = ((long long)z * n + y) * n + x;
            if (Y) {
                constexpr int CW = COUT / SPLIT;
                uint4* dst = reinterpret_cast<uint4*>(Y + v * COUT + mg * CW);
#pragma unroll
                for (int q = 0; q < CW / 8; ++q) dst[q] = make_uint4(0u, 0u, 0u, 0u);
            }
            if (headOut) headOut[v] = 0.f;
        }
    }
    return true;
}

__device__ inline int halo_index(int wave, int j, int col, int tap) {
    const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;  // offsets + 1
    return ((wave + dz) * kHY + (j + dy)) * kHX + (col + dx);
}

// Global address of halo element t (voxel t >> 1, 8-channel half t & 1) of
// input stage c0 for the tile at (bx, by, bz); null outside the grid.
template <int CIN, int CA>
__device__ inline const uint4* stage_src(const uint16_t* X, const uint16_t* Xc, int n, int bx, int by, int bz,
                                         int c0, int t) {
    constexpr int CB = CIN - CA;
    const int hv = t >> 1, half = t & 1;
    const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
    const int x = bx + hx - 1, y = by + hy - 1, z = bz + hz - 1;
    if (x < 0 || y < 0 || z < 0 || x >= n || y >= n || z >= n) return nullptr;
    if (CA > 0 && c0 < CA) {
        const int nc = n >> 1;
        return reinterpret_cast<const uint4*>(Xc + (((long long)(z >> 1) * nc + (y >> 1)) * nc + (x >> 1)) * CA + c0 +
                                              8 * half);
    }
    return reinterpret_cast<const uint4*>(X + (((long long)z * n + y) * n + x) * CB + (c0 - CA) + 8 * half);
}

// The 27 taps of one 16-channel input stage held in LDS (two 8-channel
// planes) into the accumulators of output blocks m0 .. m0 + MB - 1.
// The K = 32 of an MFMA pairs two taps that share dy: the nine (dx, dz)
// combinations s = dx + 3 dz form the pairs (2q, 2q + 1), q = 0..4 (s = 9 is
// zero).  The B fragment of pair q at input row r (the wave's 4 output rows
// read rows 0..5 of the halo) is the same for the output rows j = r - dy,
// dy = 0..2, so six 16-byte LDS reads feed twelve MFMAs (per 16 output
// channels): 30 reads per stage instead of one per MFMA.  Lane (col, kq)
// holds combination 2q + (kq >> 1), channels 8 (kq & 1) .. +7.
template <int CIN, int COUT, int MB>
__device__ inline void conv_stage_mfma(const uint4 (*sIn)[kHalo], const uint16_t* __restrict__ Wt, int c0, int m0,
                                       int wave, int col, int kq, v4f (&acc)[MB][4]) {
    const int chunk = kq & 1, half = kq >> 1;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const int sc = 2 * q + half;
        const bool live = sc < 9;
        const int dx = sc % 3, dz = sc / 3;
        v8bf b[kTY + 2];
#pragma unroll
        for (int r = 0; r < kTY + 2; ++r) {
            const uint4 raw = live ? sIn[chunk][((wave + dz) * kHY + r) * kHX + col + dx] : make_uint4(0u, 0u, 0u, 0u);
            b[r] = __builtin_bit_cast(v8bf, raw);
        }
#pragma unroll
        for (int m = 0; m < MB; ++m) {
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const int tap = dx + 3 * dy + 9 * dz;
                const uint4 raw = live ? *reinterpret_cast<const uint4*>(
                                             Wt + ((size_t)(tap * COUT + 16 * (m0 + m) + col) * CIN) + c0 + 8 * chunk)
                                       : make_uint4(0u, 0u, 0u, 0u);
                const v8bf a = __builtin_bit_cast(v8bf, raw);
#pragma unroll
                for (int j = 0; j < kTY; ++j)
                    acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j + dy], acc[m][j], 0, 0, 0);
            }
        }
    }
}

// The same stage for one 16-channel output block, with the A fragment of
// (pair q, dy) supplied by getA(q, dy): from LDS when the weights change
// between steps, from registers when they do not.
template <class GetA>
__device__ inline void conv_stage_rows(const uint4 (*sIn)[kHalo], GetA getA, int wave, int col, int kq,
                                       v4f (&acc)[4]) {
    const int chunk = kq & 1, half = kq >> 1;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const int sc = 2 * q + half;
        const bool live = sc < 9;
        const int dx = sc % 3, dz = sc / 3;
        v8bf b[kTY + 2];
#pragma unroll
        for (int r = 0; r < kTY + 2; ++r) {
            const uint4 raw = live ? sIn[chunk][((wave + dz) * kHY + r) * kHX + col + dx] : make_uint4(0u, 0u, 0u, 0u);
            b[r] = __builtin_bit_cast(v8bf, raw);
        }
#pragma unroll
        for (int dy = 0; dy < 3; ++dy) {
            const v8bf a = getA(q, dy);
#pragma unroll
            for (int j = 0; j < kTY; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j + dy], acc[j], 0, 0, 0);
        }
    }
}

// Weights of one input stage and 16-channel output block in LDS:
// sW[(tap * 2 + chunk) * 16 + cout] = 8 input channels (16 bytes); A
// fragments read from there make the MFMAs wait on LDS only, never on the
// global loads of a prefetch in flight.
constexpr int kWStage = 27 * 2 * 16;            // 16-byte weight fragments per stage (864)
constexpr int kWLoads = (kWStage + 255) / 256;  // per thread (4)

// Tap of lane half `half` in pair q at row offset dy (27 = none).
__device__ inline int pair_tap(int q, int dy, int half) {
    const int sc = 2 * q + half;
    return sc < 9 ? (sc % 3) + 3 * dy + 9 * (sc / 3) : 27;
}

// Zero that the compiler cannot prove wave-uniform.  A load from a uniform
// address yields a uniform value, which the compiler moves into a scalar
// register right after the load -- with a wait for every vector memory
// operation issued before it, the prefetch of the next step included.
// Adding this zero to the index keeps the value per lane, so the wait falls
// where the value is first used.
__device__ inline uint32_t lane_zero() {
    uint32_t z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

// Inputs of the epilogue (MODE 0) for lane (col, kq): the bias of its
// channels and the masks of its four voxels (bx + col, by + j, bz + wave).
// Loaded before the MFMAs, so their latency overlaps them and a prefetch
// issued later is not waited for.
template <int MB>
struct EpiIn {
    float bias[MB][4];
    bool inside[4];        // the voxel lies inside the grid
    // Loaded mask bytes, kept as data and tested only in the epilogue (a
    // test here would wait for the loads, and with them for every store
    // still in flight): bits 0-7 the voxel mask (1 when there is none),
    // bits 8-15 head.mask (0 without a fused head).
    uint32_t m[4];
    uint32_t slotCell[4];  // head output index in the brick-slot layout (head.outSlots)
    __device__ bool live(int j) const { return inside[j] && (m[j] & 0xffu) != 0u; }          // in the voxel mask
    __device__ bool headLive(int j) const { return inside[j] && (m[j] & 0xff00u) != 0u; }  // fused head's mask
};

template <int COUT, int MB>
__device__ inline EpiIn<MB> epi_load(const float* __restrict__ bias, const uint8_t* __restrict__ vmask,
                                     const HeadArgs& head, int n, int bx, int by, int bz, int m0, int wave, int col,
                                     int kq) {
    EpiIn<MB> e;
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int r = 0; r < 4; ++r) e.bias[m][r] = bias[16 * (m0 + m) + 4 * kq + r];
    // Every load is issued unconditionally and only selects depend on the
    // values: a load under a condition becomes a branch that waits for it on
    // the spot, and that wait also drains the previous tile's output stores.
    // Voxels outside the grid read voxel 0; an absent array is replaced by
    // the bias array (always present), whose value is then ignored.
    const int z = bz + wave, x = bx + col;
    const bool headOn = COUT == 16 && (head.out || head.outSlots);
    const bool slotsOn = headOn && head.outSlots;
    const uint8_t* vmBase = vmask ? vmask : reinterpret_cast<const uint8_t*>(bias);
    const uint8_t* hmBase = headOn ? head.mask : reinterpret_cast<const uint8_t*>(bias);
    const int32_t* slBase = slotsOn ? head.brickSlot : reinterpret_cast<const int32_t*>(bias);
    const int nb = n >> 3;
    uint8_t vm[4], hm[4];
    int32_t slot[4];
    // coordinates clamped into the grid (a voxel outside is never live), so
    // no index computation depends on a per-lane condition either
    const uint32_t xs = (uint32_t)min(x, n - 1), zs = (uint32_t)min(z, n - 1);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t ys = (uint32_t)min(by + j, n - 1);
        const uint32_t v = (zs * (uint32_t)n + ys) * (uint32_t)n + xs;
        const uint32_t b = ((zs >> 3) * (uint32_t)nb + (ys >> 3)) * (uint32_t)nb + (xs >> 3);
        vm[j] = vmBase[vmask ? v : 0u];
        hm[j] = hmBase[headOn ? v : 0u];
        slot[j] = slBase[slotsOn ? b : 0u];
    }
    // Defaults for absent arrays are applied with bit operations, not
    // selects: a select between a loaded value and a constant is turned
    // back into a branch around the load.
    const uint32_t vmSet = vmask ? 0u : 1u;
    const uint32_t hmKeep = headOn ? 0xffu : 0u;
    const uint32_t slKeep = slotsOn ? 0xffffffffu : 0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int y = by + j;
        e.inside[j] = x < n && y < n && z < n;
        e.m[j] = ((uint32_t)vm[j] | vmSet) | (((uint32_t)hm[j] & hmKeep) << 8);
        const uint32_t cell = (uint32_t)slot[j] * 512u + (uint32_t)(((z & 7) * 8 + (y & 7)) * 8 + (x & 7));
        e.slotCell[j] = cell & slKeep;
    }
    return e;
}

// The fused 1x1 head's weights, bias and output scale, loaded once per
// workgroup (before the first prefetch: a load issued inside the epilogue
// would wait for the prefetch in flight).
struct HeadW {
    float w[16];
    float b, scale;
};

template <int COUT>
__device__ inline HeadW head_load(const HeadArgs& head) {
    HeadW h{};
    if (COUT == 16 && (head.out || head.outSlots)) {
#pragma unroll
        for (int c = 0; c < 16; ++c) h.w[c] = head.w[c];
        h.b = head.b[0];
        h.scale = head.scale[0];
    }
    return h;
}

// Epilogue: lane (col, kq) holds output channels 16 (m0 + m) + 4 kq .. +3 of
// voxel (bx + col, by + j, bz + wave).
template <int COUT, int MODE, int MB>
__device__ inline void conv_epilogue(const v4f (&acc)[MB][4], const EpiIn<MB>& e, const HeadW& hw,
                                     uint16_t* __restrict__ Y,
                                     float* __restrict__ Yf, int n, int relu, const HeadArgs& head, int bx, int by,
                                     int bz, int m0, int wave, int col, int kq) {
    const int z = bz + wave, x = bx + col;
    if (MODE == 0 && COUT == 16 && (head.out || head.outSlots)) {
        // fused head: gather the 16 channels of each voxel in its kq == 0 lane
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int y = by + j;
            const bool inside = x < n && y < n && z < n;
            const long long v = ((long long)z * n + y) * n + x;
            float t[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                t[r] = acc[0][j][r] + e.bias[0][r];
                if (relu) t[r] = fmaxf(t[r], 0.f);
            }
            const uint2 packed = f2bf4(t[0], t[1], t[2], t[3], e.live(j));
            const float yv[4] = {__uint_as_float(packed.x << 16), __uint_as_float(packed.x & 0xffff0000u),
                                 __uint_as_float(packed.y << 16), __uint_as_float(packed.y & 0xffff0000u)};
            if (Y && inside) *reinterpret_cast<uint2*>(Y + v * COUT + 4 * kq) = packed;
            float ch[16];
#pragma unroll
            for (int r = 0; r < 4; ++r) ch[r] = yv[r];
#pragma unroll
            for (int q = 1; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) ch[4 * q + r] = __shfl(yv[r], col + 16 * q, 64);
            if (kq == 0 && inside) {
                float s = hw.b;
#pragma unroll
                for (int c = 0; c < 16; ++c) s = fmaf(ch[c], hw.w[c], s);
                if (!head.outSlots) head.out[v] = e.headLive(j) ? s * hw.scale : 0.f;
                else if (e.headLive(j)) head.outSlots[e.slotCell[j]] = s * hw.scale;
            }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int y = by + j;
        if (x >= n || y >= n || z >= n) continue;
        const long long v = ((long long)z * n + y) * n + x;
#pragma unroll
        for (int m = 0; m < MB; ++m) {
            const int co = 16 * (m0 + m) + 4 * kq;
            if (MODE == 1) {
                *reinterpret_cast<float4*>(Yf + v * COUT + co) =
                    make_float4(acc[m][j][0], acc[m][j][1], acc[m][j][2], acc[m][j][3]);
                continue;
            }
            float yv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                yv[r] = acc[m][j][r] + e.bias[m][r];
                if (relu) yv[r] = fmaxf(yv[r], 0.f);
            }
            const uint2 packed = f2bf4(yv[0], yv[1], yv[2], yv[3], e.live(j));
            *reinterpret_cast<uint2*>(Y + v * COUT + co) = packed;
        }
    }
}

template <int CIN, int COUT, int MODE = 0, int CA = 0, int SPLIT = 1>
__global__ void __launch_bounds__(256)
conv3_fwd_kernel(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Xc, const uint16_t* __restrict__ Wt,
                 const float* __restrict__ bias, uint16_t* __restrict__ Y, float* __restrict__ Yf, int n, int relu,
                 HeadArgs head, ConvMask cm) {
    constexpr int MB = COUT / 16 / SPLIT;  // 16-channel output blocks of this workgroup
    constexpr int CB = CIN - CA;           // channels read from X
    static_assert(CA % 16 == 0 && CB % 16 == 0, "channel groups of 16");
    static_assert(MB >= 1 && MB * 16 * SPLIT == COUT, "output channels split into whole 16-channel blocks");
    __shared__ __align__(16) uint4 sIn[2][kHalo];  // [8-channel chunk][halo voxel]
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15, kq = lane >> 4;
    int bx, by, bz, mg;
    conv_tile<SPLIT>(bx, by, bz, mg);
    if (skip_tile<COUT, SPLIT>(cm, bx, by, bz, mg, n, MODE == 0 ? Y : nullptr, MODE == 0 ? head.out : nullptr))
        return;
    const int m0 = mg * MB;  // first output block of this workgroup
    EpiIn<MB> epi{};
    HeadW hw{};
    if (MODE == 0) {
        epi = epi_load<COUT, MB>(bias, cm.voxel, head, n, bx, by, bz, m0, wave, col, kq);
        hw = head_load<COUT>(head);
    }

    v4f acc[MB][4];
#pragma unroll
    for (int m = 0; m < MB; ++m)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[m][j] = v4f{0.f, 0.f, 0.f, 0.f};

    // Each stage (648 voxels x 16 channels, as two 8-channel planes) is
    // loaded into registers while the previous stage's MFMAs run.
    uint4 pre[kStageLoads];
    auto load_stage = [&](int c0) {
#pragma unroll
        for (int k = 0; k < kStageLoads; ++k) {
            const int t = (int)threadIdx.x + 256 * k;
            const uint4* src = t < kHalo * 2 ? stage_src<CIN, CA>(X, Xc, n, bx, by, bz, c0, t) : nullptr;
            pre[k] = src ? *src : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    load_stage(0);
#pragma unroll 1
    for (int c0 = 0; c0 < CIN; c0 += 16) {
        if (c0) __syncthreads();  // every wave is done with the previous stage
#pragma unroll
        for (int k = 0; k < kStageLoads; ++k) {
            const int t = (int)threadIdx.x + 256 * k;
            if (t < kHalo * 2) sIn[t & 1][t >> 1] = pre[k];
        }
        __syncthreads();
        if (c0 + 16 < CIN) load_stage(c0 + 16);
        conv_stage_mfma<CIN, COUT, MB>(sIn, Wt, c0, m0, wave, col, kq, acc);
    }
    conv_epilogue<COUT, MODE, MB>(acc, epi, hw, Y, Yf, n, relu, head, bx, by, bz, m0, wave, col, kq);
}

// 2x2x2 average pooling fused into a list convolution's epilogue: the
// tile's 16 x 4 x 4 output voxels are the children of 8 x 2 x 2 voxels of the
// next level, whose pooled values (the BF16 outputs averaged, children summed
// in avgpool_kernel's order) go to P ([voxel][PC] BF16, channels c0 ..
// c0 + 15).  Each pooled voxel belongs to one tile, so the pooled tensor is
// complete once every computed tile has pooled and every cleared tile has
// zeroed its part.  Contains a workgroup barrier.
__device__ inline void pool_tile(const v4f (&acc)[4], const EpiIn<1>& e, int relu, uint16_t (*sPool)[kTY][kTX][16],
                                 uint16_t* __restrict__ P, int n, int PC, int c0, int bx, int by, int bz, int wave,
                                 int col, int kq) {
#pragma unroll
    for (int j = 0; j < kTY; ++j) {
        float t[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            t[r] = acc[j][r] + e.bias[0][r];
            if (relu) t[r] = fmaxf(t[r], 0.f);
        }
        *reinterpret_cast<uint2*>(&sPool[wave][j][col][4 * kq]) = f2bf4(t[0], t[1], t[2], t[3], e.live(j));
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int cv = (int)threadIdx.x >> 1, h = (int)threadIdx.x & 1;  // coarse voxel of the tile, channel half
        const int cx = cv & 7, cy = (cv >> 3) & 1, cz = cv >> 4;
        const int nc = n >> 1;
        const int X = (bx >> 1) + cx, Yc = (by >> 1) + cy, Z = (bz >> 1) + cz;
        if (X < nc && Yc < nc && Z < nc) {
            float sum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const uint4 v =
                    *reinterpret_cast<const uint4*>(&sPool[2 * cz + (k >> 2)][2 * cy + ((k >> 1) & 1)][2 * cx + (k & 1)][8 * h]);
                const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    sum[2 * q] += bf2f((uint16_t)(w[q] & 0xffffu));
                    sum[2 * q + 1] += bf2f((uint16_t)(w[q] >> 16));
                }
            }
            uint32_t o[4];
#pragma unroll
            for (int q = 0; q < 4; ++q)
                o[q] = f2bf2(sum[2 * q] * 0.125f, sum[2 * q + 1] * 0.125f);
            *reinterpret_cast<uint4*>(P + (((long long)Z * nc + Yc) * nc + X) * PC + c0 + 8 * h) =
                make_uint4(o[0], o[1], o[2], o[3]);
        }
    }
}

// The same convolution over a list of tiles (the tiles of a level that hold
// masked-in voxels), for one 16-channel output block per workgroup: a fixed
// grid of workgroups loops over the list, one (tile, channel group, input
// stage) step at a time.  The next step's input halo and weights are loaded
// into registers before the current step's MFMAs and stored into the other of
// two LDS buffers after them, so memory latency overlaps compute.  Each XCD
// takes one contiguous part of the list (neighbouring tiles, shared halo rows
// in its L2).  Tiles in `clear` (computed by an earlier forward, not by this
// one) are zeroed first.  P != null: the output is also average-pooled into
// P (the next level's input, [voxel][COUT] at half resolution; see
// pool_tile).
template <int CIN, int COUT, int CA = 0, int SPLIT = 1>
__global__ void __launch_bounds__(256)
conv3_list_kernel(const uint16_t* __restrict__ X, const uint16_t* __restrict__ Xc, const uint16_t* __restrict__ Wt,
                  const float* __restrict__ bias, uint16_t* __restrict__ Y, uint16_t* __restrict__ P, int n, int relu,
                  HeadArgs head, const uint8_t* __restrict__ vmask, TileList tl) {
    if (tl.stop && *tl.stop) return;
    constexpr int MB = COUT / 16 / SPLIT;
    constexpr int STAGES = CIN / 16;
    static_assert(MB == 1 && 16 * SPLIT == COUT, "one 16-channel output block per workgroup");
    // one input stage and one channel group: the weights are the same at
    // every step and stay in registers
    constexpr bool kFixedW = STAGES == 1 && SPLIT == 1;
    constexpr int kWBuf = kFixedW ? 1 : 2;
    __shared__ __align__(16) uint4 sIn[2][2][kHalo];  // [buffer][8-channel chunk][halo voxel]
    __shared__ __align__(16) uint4 sW[kWBuf][kFixedW ? 1 : kWStage];  // [buffer][(tap, chunk, cout)]
    __shared__ __align__(16) uint16_t sPool[kTZ][kTY][kTX][16];       // outputs to pool
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15, kq = lane >> 4;
    const uint32_t tilesX = tl.tilesX, tilesXY = tl.tilesX * tl.tilesY;
    // XCD x = blockIdx.x % 8 owns units [x U / 8, (x + 1) U / 8) of the U
    // (tile, channel group) units; its workgroups stride through them.
    const uint32_t xcd = blockIdx.x & 7u, perXcd = gridDim.x >> 3, w = blockIdx.x >> 3;

    // clear the listed tiles (unit = tile x channel group)
    {
        const uint32_t U = *tl.numClear * SPLIT;
        const uint32_t beg = (uint32_t)(((uint64_t)U * xcd) >> 3), end = (uint32_t)(((uint64_t)U * (xcd + 1)) >> 3);
        for (uint32_t u = beg + w; u < end; u += perXcd) {
            const uint32_t t = tl.clear[u / SPLIT];
            const int mg = (int)(u % SPLIT);
            const int bx = (int)(t % tilesX) * kTX, by = (int)((t % tilesXY) / tilesX) * kTY,
                      bz = (int)(t / tilesXY) * kTZ;
            const int x = bx + (threadIdx.x & 15), y = by + ((threadIdx.x >> 4) & 3), z = bz + (threadIdx.x >> 6);
            if (x < n && y < n && z < n) {
                const long long v = ((long long)z * n + y) * n + x;
                if (Y) {
                    uint4* dst = reinterpret_cast<uint4*>(Y + v * COUT + mg * 16);
                    dst[0] = make_uint4(0u, 0u, 0u, 0u);
                    dst[1] = make_uint4(0u, 0u, 0u, 0u);
                }
                if (head.out && mg == 0) head.out[v] = 0.f;
            }
            if (P && threadIdx.x < 64) {
                const int cv = (int)threadIdx.x >> 1, h = (int)threadIdx.x & 1;
                const int nc = n >> 1;
                const int X2 = (bx >> 1) + (cv & 7), Y2 = (by >> 1) + ((cv >> 3) & 1), Z2 = (bz >> 1) + (cv >> 4);
                if (X2 < nc && Y2 < nc && Z2 < nc)
                    *reinterpret_cast<uint4*>(P + (((long long)Z2 * nc + Y2) * nc + X2) * COUT + mg * 16 + 8 * h) =
                        make_uint4(0u, 0u, 0u, 0u);
            }
        }
    }

    const uint32_t U = *tl.numCompute * SPLIT;
    const uint32_t beg = (uint32_t)(((uint64_t)U * xcd) >> 3), end = (uint32_t)(((uint64_t)U * (xcd + 1)) >> 3);
    uint32_t u = beg + w;
    if (u >= end) return;
    auto unit_origin = [&](uint32_t unit, int& bx, int& by, int& bz) {
        const uint32_t t = tl.compute[unit / SPLIT];
        bx = (int)(t % tilesX) * kTX;
        by = (int)((t % tilesXY) / tilesX) * kTY;
        bz = (int)(t / tilesXY) * kTZ;
    };
    // the weight fragments as four named registers: as an array they were
    // kept in scratch memory
    uint4 pre[kStageLoads], preW0, preW1, preW2, preW3;
    // Per-thread geometry of the prefetch loads, fixed for the kernel.
    // Element t = threadIdx.x + 256 k of a stage is 8-channel half t & 1 of
    // halo voxel t >> 1.  hOff: that voxel's offset in the grid from the
    // tile's first halo voxel (bx - 1, by - 1, bz - 1); hOffC: the offset of
    // its parent in the half-resolution grid from the parent of (bx, by, bz)
    // (tile origins are even, so both are the same for every tile).  A
    // thread's last element may lie past the halo (t >= 2 kHalo, never
    // stored to LDS): it is clamped to the last element, so its load stays
    // in bounds.  wOff: the offset of the thread's weight fragments within
    // a (stage, channel group) block, clamped the same way.
    constexpr int CB = CIN - CA;
    const int nc = n >> 1;
    const int half = (int)threadIdx.x & 1;
    int hOff[kStageLoads], hOffC[CA > 0 ? kStageLoads : 1];
#pragma unroll
    for (int k = 0; k < kStageLoads; ++k) {
        const int hv = min((int)threadIdx.x + 256 * k, 2 * kHalo - 1) >> 1;
        const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
        hOff[k] = (hz * n + hy) * n + hx;
        if (CA > 0) hOffC[k] = (((hz + 1) / 2 - 1) * nc + ((hy + 1) / 2 - 1)) * nc + ((hx + 1) / 2 - 1);
    }
    static_assert(kWLoads == 4, "four weight fragments per thread and stage");
    int wOff[kWLoads] = {0, 0, 0, 0};
    if (!kFixedW) {
#pragma unroll
        for (int k = 0; k < kWLoads; ++k) {
            const int t = min((int)threadIdx.x + 256 * k, kWStage - 1);
            const int tap = t >> 5, chunk = (t >> 4) & 1, co = t & 15;
            wOff[k] = (tap * COUT + co) * CIN + 8 * chunk;
        }
    }
    // loads of step (unit, stage) into pre / preW0..3
    auto load_step = [&](uint32_t unit, int bx, int by, int bz, int stage) {
        const int c0 = 16 * stage;
        // the halo lies inside the grid: no per-element bounds
        const bool interior =
            bx >= 1 && by >= 1 && bz >= 1 && bx + kTX < n && by + kTY < n && bz + kTZ < n;
        if (interior && CA > 0 && c0 < CA) {
            const uint32_t base = (uint32_t)(((bz >> 1) * nc + (by >> 1)) * nc + (bx >> 1));
#pragma unroll
            for (int k = 0; k < kStageLoads; ++k)
                pre[k] = *reinterpret_cast<const uint4*>(Xc + (size_t)(base + (uint32_t)hOffC[k]) * CA + c0 +
                                                         8 * half);
        } else if (interior) {
            const uint32_t base = (uint32_t)(((bz - 1) * n + (by - 1)) * n + (bx - 1));
#pragma unroll
            for (int k = 0; k < kStageLoads; ++k)
                pre[k] = *reinterpret_cast<const uint4*>(X + (size_t)(base + (uint32_t)hOff[k]) * CB + (c0 - CA) +
                                                         8 * half);
        } else {
#pragma unroll
            for (int k = 0; k < kStageLoads; ++k) {
                const int t = (int)threadIdx.x + 256 * k;
                const uint4* src = t < kHalo * 2 ? stage_src<CIN, CA>(X, Xc, n, bx, by, bz, c0, t) : nullptr;
                pre[k] = src ? *src : make_uint4(0u, 0u, 0u, 0u);
            }
        }
        if (kFixedW) return;
        const uint16_t* w = Wt + (size_t)((int)(unit % SPLIT) * 16 * CIN + c0);
        preW0 = *reinterpret_cast<const uint4*>(w + wOff[0]);
        preW1 = *reinterpret_cast<const uint4*>(w + wOff[1]);
        preW2 = *reinterpret_cast<const uint4*>(w + wOff[2]);
        preW3 = *reinterpret_cast<const uint4*>(w + wOff[3]);
    };
    auto store_step = [&](int b) {
#pragma unroll
        for (int k = 0; k < kStageLoads; ++k) {
            const int t = (int)threadIdx.x + 256 * k;
            if (t < kHalo * 2) sIn[b][t & 1][t >> 1] = pre[k];
        }
        if (kFixedW) return;
        const int t = (int)threadIdx.x;  // t + 512 < kWStage <= t + 768 + 96
        sW[b][t] = preW0;
        sW[b][t + 256] = preW1;
        sW[b][t + 512] = preW2;
        if (t + 768 < kWStage) sW[b][t + 768] = preW3;
    };
    v8bf aw[kFixedW ? 15 : 1];
    if (kFixedW) {
        const int chunk = kq & 1, half = kq >> 1;
#pragma unroll
        for (int q = 0; q < 5; ++q)
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const int tap = pair_tap(q, dy, half);
                const uint4 raw = tap < 27 ? *reinterpret_cast<const uint4*>(Wt + (size_t)(tap * COUT + col) * CIN +
                                                                             8 * chunk)
                                           : make_uint4(0u, 0u, 0u, 0u);
                aw[kFixedW ? q * 3 + dy : 0] = __builtin_bit_cast(v8bf, raw);
            }
    }
    int bx, by, bz;
    unit_origin(u, bx, by, bz);
    load_step(u, bx, by, bz, 0);
    store_step(0);
    // Tile of the next unit, fetched one unit ahead so that the prefetch of
    // a new tile's halo never waits for its tile id (index 0 when there is
    // none: the list is not empty here).
    uint32_t tileNext = tl.compute[u + perXcd < end ? (u + perXcd) / SPLIT : 0u];
    const HeadW hw = head_load<COUT>(head);
    __syncthreads();
    int buf = 0, stage = 0;
    v4f acc[1][4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[0][j] = v4f{0.f, 0.f, 0.f, 0.f};
    for (;;) {
        // next step
        uint32_t nu = u;
        int ns = stage + 1;
        if (ns == STAGES) {
            ns = 0;
            nu = u + perXcd;
        }
        const bool more = nu < end;
        const int m0 = (int)(u % SPLIT);
        const bool newUnit = more && nu != u;
        const uint32_t after = nu + perXcd;
        // the tile id two units ahead, issued before this step's other loads
        // and kept per lane: it is first used after the prefetch has been
        // waited for, so it adds no wait of its own
        const uint32_t tileAfter = tl.compute[(newUnit && after < end ? after / SPLIT : 0u) + lane_zero()];
        // loaded at every step (only the last stage of a unit uses them): a
        // conditional load would be tested, and waited for, on the spot
        const EpiIn<1> epi = epi_load<COUT, 1>(bias, vmask, head, n, bx, by, bz, m0, wave, col, kq);
        // The next step is loaded (and stored into LDS below) also after
        // the last one -- tile 0 of the list, in bounds and never used --
        // so that no path through the loop leaves a prefetch register
        // pending: the compiler's wait analysis does not see that both
        // places test the same condition, and would otherwise wait at the
        // top of every step for the previous step's output stores.
        int nbx = bx, nby = by, nbz = bz;
        if (nu != u) {
            nbx = (int)(tileNext % tilesX) * kTX;
            nby = (int)((tileNext % tilesXY) / tilesX) * kTY;
            nbz = (int)(tileNext / tilesXY) * kTZ;
        }
        load_step(nu, nbx, nby, nbz, ns);
        if constexpr (kFixedW) {
            conv_stage_rows(sIn[buf], [&](int q, int dy) { return aw[q * 3 + dy]; }, wave, col, kq, acc[0]);
        } else {
            const uint4* w = sW[buf];
            const int chunk = kq & 1, half = kq >> 1;
            conv_stage_rows(
                sIn[buf],
                [&](int q, int dy) {
                    const int tap = pair_tap(q, dy, half);
                    return __builtin_bit_cast(v8bf, tap < 27 ? w[(tap * 2 + chunk) * 16 + col]
                                                             : make_uint4(0u, 0u, 0u, 0u));
                },
                wave, col, kq, acc[0]);
        }
        // The prefetched step goes into the other LDS buffer (free since the
        // last barrier) before the epilogue's global stores are issued: the
        // wait for the prefetch then does not also wait for those stores.
        store_step(buf ^ 1);
        if (stage == STAGES - 1) {
            conv_epilogue<COUT, 0, 1>(acc, epi, hw, Y, nullptr, n, relu, head, bx, by, bz, m0, wave, col, kq);
            if (P) pool_tile(acc[0], epi, relu, sPool, P, n, COUT, 16 * m0, bx, by, bz, wave, col, kq);
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[0][j] = v4f{0.f, 0.f, 0.f, 0.f};
        }
        if (!more) break;
        __syncthreads();
        if (newUnit) tileNext = tileAfter;
        buf ^= 1;
        u = nu;
        stage = ns;
        bx = nbx;
        by = nby;
        bz = nbz;
    }
}

// First convolution, whose input has two channels (the scaled residual and
// the fluid mask; the layer's other 14 input channels are zero): X2 is
// [voxel][2] BF16 and the 27 taps x 2 channels form the reduction dimension
// directly, k = 2 tap + channel, 54 of 64 slots in two K = 32 MFMA blocks
// (instead of 27 blocks of 16 channels of which 14 are zero).  Lane (col, kq)
// of K block kb holds taps 16 kb + 4 kq .. +3.  Same tiles, tile order and
// epilogue as conv3_fwd_kernel (MODE 0, 16 output channels).
__global__ void __launch_bounds__(256)
conv3_in2_kernel(const uint16_t* __restrict__ X2, const uint16_t* __restrict__ Wt, const float* __restrict__ bias,
                 uint16_t* __restrict__ Y, int n, int relu, ConvMask cm) {
    __shared__ uint32_t sIn[kHalo];  // [halo voxel] = 2 BF16 channels
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15, kq = lane >> 4;
    int bx, by, bz, mg;
    conv_tile(bx, by, bz, mg);
    if (skip_tile<16>(cm, bx, by, bz, mg, n, Y, nullptr)) return;
    for (int t = threadIdx.x; t < kHalo; t += 256) {
        const int hx = t % kHX, hy = (t / kHX) % kHY, hz = t / (kHX * kHY);
        const int x = bx + hx - 1, y = by + hy - 1, z = bz + hz - 1;
        uint32_t v = 0u;
        if (x >= 0 && y >= 0 && z >= 0 && x < n && y < n && z < n)
            v = reinterpret_cast<const uint32_t*>(X2)[((long long)z * n + y) * n + x];
        sIn[t] = v;
    }
    // A fragments: W[tap][co = col][ci] for the lane's four taps, ci = 0, 1
    v8bf a[2];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        const int t0 = 16 * kb + 4 * kq;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int tap = t0 + (e >> 1);
            const uint16_t w = tap < 27 ? Wt[((size_t)tap * 16 + col) * 16 + (e & 1)] : (uint16_t)0;
            a[kb][e] = __builtin_bit_cast(__bf16, w);
        }
    }
    __syncthreads();
    v4f acc[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        const int t0 = 16 * kb + 4 * kq;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            uint32_t q[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int tap = t0 + e;
                q[e] = tap < 27 ? sIn[halo_index(wave, j, col, tap)] : 0u;
            }
            const v8bf b = __builtin_bit_cast(v8bf, make_uint4(q[0], q[1], q[2], q[3]));
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kb], b, acc[j], 0, 0, 0);
        }
    }
    const int z = bz + wave, x = bx + col;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int y = by + j;
        if (x >= n || y >= n || z >= n) continue;
        const long long v = ((long long)z * n + y) * n + x;
        const bool live = !cm.voxel || cm.voxel[v];
        const int co = 4 * kq;
        float yv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            yv[r] = acc[j][r] + bias[co + r];
            if (relu) yv[r] = fmaxf(yv[r], 0.f);
        }
        *reinterpret_cast<uint2*>(Y + v * 16 + co) = f2bf4(yv[0], yv[1], yv[2], yv[3], live);
    }
}

// The first convolution over the list of level-0 tiles with fluid (the
// preconditioner apply): a fixed grid of workgroups loops over the list, the
// next tile's halo (648 voxels x 2 channels, 4 bytes each) loaded into
// registers during the current tile's MFMAs and stored into the other of two
// LDS buffers; the weights (two A fragments per lane) are loaded once.  The
// arithmetic is conv3_in2_kernel's, in the same order.  Tiles in `clear` are
// zeroed first.
constexpr int kIn2Loads = (kHalo + 255) / 256;  // 3
__global__ void __launch_bounds__(256)
conv3_in2_list_kernel(const uint16_t* __restrict__ X2, SlotInput si, const uint16_t* __restrict__ Wt,
                      const float* __restrict__ bias, uint16_t* __restrict__ Y, int n, int relu,
                      const uint8_t* __restrict__ vmask, TileList tl) {
    __shared__ uint32_t sIn[2][kHalo];
    if (tl.stop && *tl.stop) return;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int col = lane & 15, kq = lane >> 4;
    const uint32_t tilesX = tl.tilesX, tilesXY = tl.tilesX * tl.tilesY;
    const uint32_t xcd = blockIdx.x & 7u, perXcd = gridDim.x >> 3, w = blockIdx.x >> 3;
    auto tile_origin = [&](uint32_t t, int& bx, int& by, int& bz) {
        bx = (int)(t % tilesX) * kTX;
        by = (int)((t % tilesXY) / tilesX) * kTY;
        bz = (int)(t / tilesXY) * kTZ;
    };
    {
        const uint32_t U = *tl.numClear;
        const uint32_t beg = (uint32_t)(((uint64_t)U * xcd) >> 3), end = (uint32_t)(((uint64_t)U * (xcd + 1)) >> 3);
        for (uint32_t u = beg + w; u < end; u += perXcd) {
            int bx, by, bz;
            tile_origin(tl.clear[u], bx, by, bz);
            const int x = bx + (threadIdx.x & 15), y = by + ((threadIdx.x >> 4) & 3), z = bz + (threadIdx.x >> 6);
            if (x < n && y < n && z < n) {
                uint4* dst = reinterpret_cast<uint4*>(Y + (((long long)z * n + y) * n + x) * 16);
                dst[0] = make_uint4(0u, 0u, 0u, 0u);
                dst[1] = make_uint4(0u, 0u, 0u, 0u);
            }
        }
    }
    const uint32_t U = *tl.numCompute;
    const uint32_t beg = (uint32_t)(((uint64_t)U * xcd) >> 3), end = (uint32_t)(((uint64_t)U * (xcd + 1)) >> 3);
    uint32_t u = beg + w;
    if (u >= end) return;
    // A fragments: W[tap][co = col][ci = 0, 1] of the lane's four taps
    v8bf a[2];
#pragma unroll
    for (int kb = 0; kb < 2; ++kb) {
        uint32_t wq[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int tap = 16 * kb + 4 * kq + e;
            wq[e] = tap < 27 ? *reinterpret_cast<const uint32_t*>(Wt + ((size_t)tap * 16 + col) * 16) : 0u;
        }
        a[kb] = __builtin_bit_cast(v8bf, make_uint4(wq[0], wq[1], wq[2], wq[3]));
    }
    // Prefetched halo: the packed input (dense input), or the cell type and
    // residual of each voxel (slot input), converted when stored to LDS.
    // Loads are unconditional (clamped index) with masks applied by bit
    // operations: a load under a per-lane condition is waited for on the
    // spot, and that wait also drains the previous tile's output stores.
    uint32_t pre[kIn2Loads];
    float preR[kIn2Loads];
    const uint32_t one = (uint32_t)f2bf(1.0f) << 16;
    const float scaleIn = si.r ? si.scale[0] : 1.f;
    auto load_tile = [&](int bx, int by, int bz) {
        if (si.r) {
            // brick slots first, then the residual and cell type of each voxel
            int32_t slot[kIn2Loads];
            uint32_t cell[kIn2Loads];
            bool in[kIn2Loads];
#pragma unroll
            for (int k = 0; k < kIn2Loads; ++k) {
                const int t = (int)threadIdx.x + 256 * k;
                const int hx = t % kHX, hy = (t / kHX) % kHY, hz = t / (kHX * kHY);
                const int x = bx + hx - 1, y = by + hy - 1, z = bz + hz - 1;
                in[k] = t < kHalo && x >= 0 && y >= 0 && z >= 0 && x < n && y < n && z < n;
                const int nb = n >> 3;
                slot[k] = si.brickSlot[in[k] ? ((z >> 3) * nb + (y >> 3)) * nb + (x >> 3) : 0];
                cell[k] = in[k] ? (uint32_t)(((z & 7) * 8 + (y & 7)) * 8 + (x & 7)) : 0u;
            }
#pragma unroll
            for (int k = 0; k < kIn2Loads; ++k) {
                const bool valid = in[k] && slot[k] >= 0;
                const uint32_t c = valid ? (uint32_t)slot[k] * 512u + cell[k] : 0u;  // slot 0 exists
                pre[k] = (uint32_t)si.cellType[c] & (valid ? 0xffu : 0u);
                preR[k] = si.r[c];
            }
            return;
        }
#pragma unroll
        for (int k = 0; k < kIn2Loads; ++k) {
            const int t = (int)threadIdx.x + 256 * k;
            const int hx = t % kHX, hy = (t / kHX) % kHY, hz = t / (kHX * kHY);
            const int x = bx + hx - 1, y = by + hy - 1, z = bz + hz - 1;
            const bool in = t < kHalo && x >= 0 && y >= 0 && z >= 0 && x < n && y < n && z < n;
            pre[k] = reinterpret_cast<const uint32_t*>(X2)[in ? ((long long)z * n + y) * n + x : 0] &
                     (in ? 0xffffffffu : 0u);
        }
    };
    auto store_tile = [&](int b) {
#pragma unroll
        for (int k = 0; k < kIn2Loads; ++k) {
            const int t = (int)threadIdx.x + 256 * k;
            const uint32_t v = si.r ? (pre[k] == 1u ? ((uint32_t)f2bf(preR[k] / scaleIn) | one) : 0u) : pre[k];
            if (t < kHalo) sIn[b][t] = v;
        }
    };
    int bx, by, bz;
    tile_origin(tl.compute[u], bx, by, bz);
    load_tile(bx, by, bz);
    store_tile(0);
    // next tile's id fetched one tile ahead (index 0 when there is none)
    uint32_t tileNext = tl.compute[u + perXcd < end ? u + perXcd : 0u];
    __syncthreads();
    int buf = 0;
    for (;;) {
        const uint32_t nu = u + perXcd;
        const bool more = nu < end;
        // the tile id two tiles ahead, first (see lane_zero)
        const uint32_t tileAfter = tl.compute[(more && nu + perXcd < end ? nu + perXcd : 0u) + lane_zero()];
        // this tile's voxel masks before the next tile's loads
        const EpiIn<1> epi = epi_load<16, 1>(bias, vmask, HeadArgs{}, n, bx, by, bz, 0, wave, col, kq);
        // loaded and stored also after the last tile (tile 0 of the list,
        // never used): see conv3_list_kernel
        int nbx = bx, nby = by, nbz = bz;
        tile_origin(tileNext, nbx, nby, nbz);
        load_tile(nbx, nby, nbz);
        v4f acc[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[j] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < 2; ++kb) {
            const int t0 = 16 * kb + 4 * kq;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                uint32_t q[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int tap = t0 + e;
                    q[e] = tap < 27 ? sIn[buf][halo_index(wave, j, col, tap)] : 0u;
                }
                const v8bf b = __builtin_bit_cast(v8bf, make_uint4(q[0], q[1], q[2], q[3]));
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[kb], b, acc[j], 0, 0, 0);
            }
        }
        // next tile into the other buffer before this tile's output stores
        store_tile(buf ^ 1);
        const int z = bz + wave, x = bx + col;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int y = by + j;
            if (x >= n || y >= n || z >= n) continue;
            const long long v = ((long long)z * n + y) * n + x;
            float yv[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                yv[r] = acc[j][r] + epi.bias[0][r];
                if (relu) yv[r] = fmaxf(yv[r], 0.f);
            }
            *reinterpret_cast<uint2*>(Y + v * 16 + 4 * kq) = f2bf4(yv[0], yv[1], yv[2], yv[3], epi.live(j));
        }
        if (!more) break;
        __syncthreads();
        tileNext = tileAfter;
        buf ^= 1;
        u = nu;
        bx = nbx;
        by = nby;
        bz = nbz;
    }
}

__global__ void mask_pool_kernel(const uint8_t* __restrict__ m, int n, uint8_t* __restrict__ out) {
    const int nc = n >> 1;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nc * nc * nc) return;
    const int x = (int)(i % nc), y = (int)((i / nc) % nc), z = (int)(i / ((long long)nc * nc));
    uint8_t any = 0;
#pragma unroll
    for (int c = 0; c < 8; ++c)
        any |= m[((long long)(2 * z + (c >> 2)) * n + 2 * y + ((c >> 1) & 1)) * n + 2 * x + (c & 1)];
    out[i] = any ? 1 : 0;
}

// One thread per convolution tile: any masked-in voxel in its 16x4x4 block.
__global__ void tile_flags_kernel(const uint8_t* __restrict__ m, int n, uint32_t tilesX, uint32_t tilesY,
                                  uint32_t tiles, uint8_t* __restrict__ flags) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tiles) return;
    const int bx = (int)(t % tilesX) * kTX, by = (int)((t / tilesX) % tilesY) * kTY, bz = (int)(t / (tilesX * tilesY)) * kTZ;
    uint8_t any = 0;
    for (int z = bz; z < min(bz + kTZ, n); ++z)
        for (int y = by; y < min(by + kTY, n); ++y)
            for (int x = bx; x < min(bx + kTX, n); ++x) any |= m[((long long)z * n + y) * n + x];
    flags[t] = any ? 1 : 0;
}

// One workgroup of 1024 threads: thread i owns a contiguous chunk of tiles,
// so the lists come out in tile order.
__global__ void __launch_bounds__(1024)
tile_lists_kernel(const uint8_t* __restrict__ cur, const uint8_t* __restrict__ prev, uint32_t tiles,
                  uint32_t* __restrict__ compute, uint32_t* __restrict__ clear, uint32_t* __restrict__ counts) {
    __shared__ uint32_t s_a[1024], s_b[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (tiles + 1023u) / 1024u;
    const uint32_t beg = min(tiles, tid * per), end = min(tiles, beg + per);
    uint32_t na = 0, nb = 0;
    for (uint32_t t = beg; t < end; ++t) {
        const bool c = cur[t] != 0;
        na += c ? 1u : 0u;
        nb += (!c && prev && prev[t]) ? 1u : 0u;
    }
    s_a[tid] = na;
    s_b[tid] = nb;
    __syncthreads();
    for (uint32_t d = 1; d < 1024u; d <<= 1) {  // inclusive Hillis-Steele scan
        const uint32_t a = tid >= d ? s_a[tid - d] : 0u, b = tid >= d ? s_b[tid - d] : 0u;
        __syncthreads();
        s_a[tid] += a;
        s_b[tid] += b;
        __syncthreads();
    }
    uint32_t oa = s_a[tid] - na, ob = s_b[tid] - nb;
    for (uint32_t t = beg; t < end; ++t) {
        const bool c = cur[t] != 0;
        if (c) compute[oa++] = t;
        else if (prev && prev[t]) clear[ob++] = t;
    }
    if (tid == 1023u) {
        counts[0] = s_a[1023];
        counts[1] = s_b[1023];
    }
}

// g <- g * (act > 0) in place, and a BF16 copy of the result for the
// MFMA gradient kernels.
__global__ void relu_mask_bf16_kernel(float* __restrict__ g, const uint16_t* __restrict__ act,
                                      uint16_t* __restrict__ gb, long long count) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float v = g[i];
    if (!(bf2f(act[i]) > 0.f)) v = 0.f;
    g[i] = v;
    gb[i] = f2bf(v);
}

// dB[c] += sum_v g[v][c]: grid-stride partial sums per workgroup, one atomic
// per channel per workgroup.
__global__ void __launch_bounds__(256) bias_grad_kernel(const float* __restrict__ g, long long nvox, int C,
                                                        float* __restrict__ dB) {
    __shared__ float s_acc[256];
    const int c = threadIdx.x % C;
    const int lanesPerC = 256 / C;  // C divides 256 (16, 32 or 64)
    const int sub = threadIdx.x / C;
    float acc = 0.f;
    for (long long v = (long long)blockIdx.x * lanesPerC + sub; v < nvox; v += (long long)gridDim.x * lanesPerC)
        acc += g[v * C + c];
    s_acc[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < C) {
        float t = 0.f;
        for (int k = 0; k < lanesPerC; ++k) t += s_acc[k * C + threadIdx.x];
        atomicAdd(&dB[threadIdx.x], t);
    }
}

// WT[t][ci][co] = bf16(W[26 - t][co][ci]): the data gradient is a forward
// convolution of dP with the transposed weights and mirrored taps.
__global__ void transpose_weights_kernel(const float* __restrict__ W, uint16_t* __restrict__ WT, int cin, int cout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 27 * cin * cout) return;
    const int t = i / (cin * cout), r = i % (cin * cout);
    const int ci = r / cout, co = r % cout;
    WT[i] = f2bf(W[((26 - t) * cout + co) * cin + ci]);
}

// dW[tap][co][ci] += sum_v dP[v][co] X[v + off(tap)][ci] on MFMA: a GEMM with
// M = COUT, N = CIN and the voxels as the reduction dimension.  Workgroup
// (chunk, tap): 128-voxel slices of its chunk are staged in LDS transposed
// ([channel][voxel], rows padded against bank conflicts) so each lane's four
// consecutive voxels are one 8-byte read; each wave owns some of the 16x16
// output tiles; the chunk's partial sums go to dW with FP32 atomics.
template <int CIN, int COUT>
__global__ void __launch_bounds__(256)
conv3_wgrad_mfma_kernel(const uint16_t* __restrict__ dP, const uint16_t* __restrict__ X, float* __restrict__ dW,
                        int n, long long chunk) {
    constexpr int KV = 128, KP = KV + 4;
    constexpr int MT = COUT / 16, NT = CIN / 16, T = MT * NT;
    // with fewer than 4 output tiles the waves also split the reduction
    constexpr int KSPLIT = T >= 4 ? 1 : (T == 1 ? 4 : (T == 2 ? 2 : 1));
    constexpr int TW = KSPLIT > 1 ? 1 : (T + 3) / 4;
    __shared__ __align__(16) uint16_t sP[COUT * KP];
    __shared__ __align__(16) uint16_t sX[CIN * KP];
    const int tap = blockIdx.y;
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    const long long nvox = (long long)n * n * n;
    const long long vBeg = (long long)blockIdx.x * chunk;
    const long long vEnd = vBeg + chunk < nvox ? vBeg + chunk : nvox;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, kq = lane >> 4;
    v4f acc[TW];
#pragma unroll
    for (int t = 0; t < TW; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
    for (long long vb = vBeg; vb < vEnd; vb += KV) {
        __syncthreads();
        for (int i = threadIdx.x; i < KV * COUT; i += 256) {
            const int k = i / COUT, c = i % COUT;
            const long long v = vb + k;
            sP[c * KP + k] = v < vEnd ? dP[v * COUT + c] : (uint16_t)0;
        }
        for (int i = threadIdx.x; i < KV * CIN; i += 256) {
            const int k = i / CIN, c = i % CIN;
            const long long v = vb + k;
            uint16_t xv = 0;
            if (v < vEnd) {
                const int x = (int)(v % n) + dx, y = (int)((v / n) % n) + dy, z = (int)(v / ((long long)n * n)) + dz;
                if (x >= 0 && y >= 0 && z >= 0 && x < n && y < n && z < n)
                    xv = X[(((long long)z * n + y) * n + x) * CIN + c];
            }
            sX[c * KP + k] = xv;
        }
        __syncthreads();
        if (KSPLIT > 1) {
            const int t = wave % T, ks = wave / T;
            const int m = t / NT, nn = t % NT;
            if (ks < KSPLIT) {
#pragma unroll
                for (int k0 = 16 * ks; k0 < KV; k0 += 16 * KSPLIT) {
                    const v4s a = *reinterpret_cast<const v4s*>(&sP[(16 * m + col) * KP + k0 + 4 * kq]);
                    const v4s b = *reinterpret_cast<const v4s*>(&sX[(16 * nn + col) * KP + k0 + 4 * kq]);
                    acc[0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[0], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int k0 = 0; k0 < KV; k0 += 16) {
#pragma unroll
                for (int tt = 0; tt < TW; ++tt) {
                    const int t = wave + 4 * tt;
                    if (t >= T) break;
                    const int m = t / NT, nn = t % NT;
                    const v4s a = *reinterpret_cast<const v4s*>(&sP[(16 * m + col) * KP + k0 + 4 * kq]);
                    const v4s b = *reinterpret_cast<const v4s*>(&sX[(16 * nn + col) * KP + k0 + 4 * kq]);
                    acc[tt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[tt], 0, 0, 0);
                }
            }
        }
    }
    if (KSPLIT > 1) {
        // sum the KSPLIT partial tiles through LDS (reusing sP), then one
        // wave per tile adds the result to dW
        __syncthreads();
        float* red = reinterpret_cast<float*>(sP);  // 4 waves x 256 floats <= sizeof(sP)
#pragma unroll
        for (int r = 0; r < 4; ++r) red[wave * 256 + (4 * kq + r) * 16 + col] = acc[0][r];
        __syncthreads();
        if (wave < T) {
            const int t = wave, m = t / NT, nn = t % NT;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float v = 0.f;
                for (int ks = 0; ks < KSPLIT; ++ks) v += red[(t + ks * T) * 256 + (4 * kq + r) * 16 + col];
                const int co = 16 * m + 4 * kq + r, ci = 16 * nn + col;
                atomicAdd(&dW[((size_t)tap * COUT + co) * CIN + ci], v);
            }
        }
        return;
    }
#pragma unroll
    for (int tt = 0; tt < TW; ++tt) {
        const int t = wave + 4 * tt;
        if (t >= T) break;
        const int m = t / NT, nn = t % NT;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int co = 16 * m + 4 * kq + r, ci = 16 * nn + col;
            atomicAdd(&dW[((size_t)tap * COUT + co) * CIN + ci], acc[tt][r]);
        }
    }
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
                      bool relu, const ConvMask& cm, hipStream_t s) {
    const long long nvox = (long long)n * n * n;
    (void)nvox;
    const dim3 grid = conv_grid(n);
    const HeadArgs noHead{};
    // Small grids (the coarse levels) split the output channels over 2 or 4
    // workgroups per tile, so there are enough workgroups to fill the chip.
    const uint32_t tiles = grid.x * grid.y * grid.z;
    const int split = tiles >= 2048u ? 1 : (tiles >= 1024u ? 2 : 4);
#define CONV_LAUNCH(CI, CO, SP)                                                                                     \
    FLIP2_KERNEL((conv3_fwd_kernel<CI, CO, 0, 0, SP>), dim3(grid.x * SP, grid.y, grid.z), dim3(256), 0, s, X, \
                       (const uint16_t*)nullptr, W, bias, Y, (float*)nullptr, n, relu ? 1 : 0, noHead, cm)
#define CONV_CASE(CI, CO)                                                                                   \
    if (cin == CI && cout == CO) {                                                                          \
        const int sp = std::min(split, CO / 16);                                                            \
        if (sp == 4) CONV_LAUNCH(CI, CO, (CO >= 64 ? 4 : 1));                                               \
        else if (sp == 2) CONV_LAUNCH(CI, CO, (CO >= 32 ? 2 : 1));                                          \
        else CONV_LAUNCH(CI, CO, 1);                                                                        \
        FLIP2_LAUNCH_CHECK();                                                                               \
        return;                                                                                             \
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
#undef CONV_LAUNCH
    throw Error(kErrArgument, "conv3: unsupported channel pair");
}

void launch_conv3_fwd_upcat(const uint16_t* Xskip, const uint16_t* Xcoarse, const uint16_t* W, const float* bias,
                            uint16_t* Y, int n, int ca, int cb, int cout, bool relu, const ConvMask& cm,
                            hipStream_t s) {
    FLIP2_REQUIRE(n % 2 == 0, kErrArgument, "conv3 with upsampled input: even grid size");
    const dim3 grid = conv_grid(n);
    if (ca == 64 && cb == 32 && cout == 32) {
        FLIP2_KERNEL((conv3_fwd_kernel<96, 32, 0, 64>), grid, dim3(256), 0, s, Xskip, Xcoarse, W, bias, Y,
                           (float*)nullptr, n, relu ? 1 : 0, HeadArgs{}, cm);
    } else if (ca == 32 && cb == 16 && cout == 16) {
        FLIP2_KERNEL((conv3_fwd_kernel<48, 16, 0, 32>), grid, dim3(256), 0, s, Xskip, Xcoarse, W, bias, Y,
                           (float*)nullptr, n, relu ? 1 : 0, HeadArgs{}, cm);
    } else {
        throw Error(kErrArgument, "conv3 with upsampled input: unsupported channel counts");
    }
    FLIP2_LAUNCH_CHECK();
}

void launch_conv3_in2(const uint16_t* X2, const uint16_t* W, const float* bias, uint16_t* Y, int n, bool relu,
                      const ConvMask& cm, hipStream_t s) {
    const dim3 grid = conv_grid(n);
    FLIP2_KERNEL(conv3_in2_kernel, grid, dim3(256), 0, s, X2, W, bias, Y, n, relu ? 1 : 0, cm);
    FLIP2_LAUNCH_CHECK();
}

void launch_conv3_in2_list(const uint16_t* X2, const SlotInput& slots, const uint16_t* W, const float* bias,
                           uint16_t* Y, int n, bool relu, const uint8_t* vmask, const TileList& tl, hipStream_t s) {
    FLIP2_REQUIRE(!slots.r || n % 8 == 0, kErrArgument, "slot input: n must be a multiple of the 8^3 bricks");
    // four workgroups per CU (the register budget's limit)
    FLIP2_KERNEL(conv3_in2_list_kernel, dim3(256 * 4), dim3(256), 0, s, X2, slots, W, bias, Y, n, relu ? 1 : 0,
                       vmask, tl);
    FLIP2_LAUNCH_CHECK();
}

void launch_tile_lists(const uint8_t* cur, const uint8_t* prev, uint32_t tiles, uint32_t* compute, uint32_t* clear,
                       uint32_t* counts, hipStream_t s) {
    FLIP2_KERNEL(tile_lists_kernel, dim3(1), dim3(1024), 0, s, cur, prev, tiles, compute, clear, counts);
    FLIP2_LAUNCH_CHECK();
}

void conv_tile_dims(int n, uint32_t& tilesX, uint32_t& tilesY) {
    const dim3 g = conv_grid(n);
    tilesX = g.x;
    tilesY = g.y;
}

void launch_conv3_list(const uint16_t* X, const uint16_t* Xcoarse, const uint16_t* W, const float* bias, uint16_t* Y,
                       uint16_t* pooled, int n, int cin, int ca, int cout, int split, bool relu, const HeadArgs& head,
                       const uint8_t* vmask, const TileList& tl, hipStream_t s) {
    FLIP2_REQUIRE(!pooled || (n % 2 == 0 && !head.out), kErrArgument, "pooled list convolution: even n, no head");
    // two workgroups per CU fit the double-buffered halo and weights (69 KB each)
    const dim3 grid(256 * 2), block(256);
    const int r = relu ? 1 : 0;
#define LIST_CASE(CI, CA_, CO, SP)                                                                               \
    if (cin == CI && ca == CA_ && cout == CO && split == SP) {                                                   \
        FLIP2_KERNEL((conv3_list_kernel<CI, CO, CA_, SP>), grid, block, 0, s, X, Xcoarse, W, bias, Y, pooled, n, r, \
                           head, vmask, tl);                                                                      \
        FLIP2_LAUNCH_CHECK();                                                                                    \
        return;                                                                                                  \
    }
    LIST_CASE(16, 0, 16, 1)
    LIST_CASE(16, 0, 32, 2)
    LIST_CASE(32, 0, 32, 2)
    LIST_CASE(32, 0, 64, 4)
    LIST_CASE(64, 0, 64, 4)
    LIST_CASE(96, 64, 32, 2)
    LIST_CASE(48, 32, 16, 1)
#undef LIST_CASE
    throw Error(kErrArgument, "conv3 over a tile list: unsupported channels / split");
}

uint32_t conv_tiles(int n) {
    const dim3 g = conv_grid(n);
    return g.x * g.y * g.z;
}

void launch_mask_pool(const uint8_t* m, int n, uint8_t* out, hipStream_t s) {
    const long long nc = n / 2;
    FLIP2_KERNEL(mask_pool_kernel, g1(nc * nc * nc, 256), dim3(256), 0, s, m, n, out);
    FLIP2_LAUNCH_CHECK();
}

void launch_tile_flags(const uint8_t* m, int n, uint8_t* flags, hipStream_t s) {
    const dim3 g = conv_grid(n);
    const uint32_t tiles = g.x * g.y * g.z;
    FLIP2_KERNEL(tile_flags_kernel, g1(tiles, 256), dim3(256), 0, s, m, n, g.x, g.y, tiles, flags);
    FLIP2_LAUNCH_CHECK();
}

void launch_conv3_fwd_head(const uint16_t* X, const uint16_t* W, const float* bias, uint16_t* Y, int n, int cin,
                           bool relu, const HeadArgs& head, const ConvMask& cm, hipStream_t s) {
    FLIP2_REQUIRE(cin == 16 && head.out && head.w && head.b && head.mask && head.scale, kErrArgument,
                  "conv3 with fused head: 16 input channels and all head arguments");
    const dim3 grid = conv_grid(n);
    FLIP2_KERNEL((conv3_fwd_kernel<16, 16, 0>), grid, dim3(256), 0, s, X, (const uint16_t*)nullptr, W, bias, Y,
                       (float*)nullptr, n, relu ? 1 : 0, head, cm);
    FLIP2_LAUNCH_CHECK();
}

void launch_conv3_dgrad(const uint16_t* dPb, const float* W32, uint16_t* WT, float* dX, int n, int cin, int cout,
                        hipStream_t s) {
    const long long nvox = (long long)n * n * n;
    const int nw = 27 * cin * cout;
    FLIP2_KERNEL(transpose_weights_kernel, g1(nw, 256), dim3(256), 0, s, W32, WT, cin, cout);
    FLIP2_LAUNCH_CHECK();
    (void)nvox;
    const dim3 grid = conv_grid(n);
    // the forward kernel with the roles of the channel counts swapped
#define DGRAD_CASE(CI, CO)                                                                                      \
    if (cin == CI && cout == CO) {                                                                              \
        FLIP2_KERNEL((conv3_fwd_kernel<CO, CI, 1>), grid, dim3(256), 0, s, dPb, (const uint16_t*)nullptr, WT,  \
                           (const float*)nullptr, (uint16_t*)nullptr, dX, n, 0, HeadArgs{}, ConvMask{});       \
        FLIP2_LAUNCH_CHECK();                                                                                   \
        return;                                                                                                 \
    }
    DGRAD_CASE(16, 16)
    DGRAD_CASE(16, 32)
    DGRAD_CASE(32, 32)
    DGRAD_CASE(32, 64)
    DGRAD_CASE(64, 64)
    DGRAD_CASE(96, 32)
    DGRAD_CASE(48, 16)
#undef DGRAD_CASE
    throw Error(kErrArgument, "conv3 dgrad: unsupported channel pair");
}

void launch_conv3_wgrad(const uint16_t* dPb, const float* dP, const uint16_t* X, float* dW, float* dB, int n, int cin,
                        int cout, hipStream_t s) {
    const long long nvox = (long long)n * n * n;
    // bias: column sums of dP
    const int bblocks = (int)std::min<long long>(1024, (nvox + 255) / 256);
    FLIP2_KERNEL(bias_grad_kernel, dim3(bblocks), dim3(256), 0, s, dP, nvox, cout, dB);
    FLIP2_LAUNCH_CHECK();
    // chunks: enough workgroups to fill the GPU, few enough that the atomics
    // (27 * cout * cin per chunk) stay around 4M
    long long chunks = std::max<long long>(1, 4000000LL / (27LL * cin * cout));
    chunks = std::min<long long>(chunks, (nvox + 127) / 128);
    long long chunk = (nvox + chunks - 1) / chunks;
    chunk = (chunk + 127) / 128 * 128;
    chunks = (nvox + chunk - 1) / chunk;
#define WGRAD_CASE(CI, CO)                                                                                      \
    if (cin == CI && cout == CO) {                                                                              \
        FLIP2_KERNEL((conv3_wgrad_mfma_kernel<CI, CO>), dim3((unsigned)chunks, 27), dim3(256), 0, s, dPb, X, \
                           dW, n, chunk);                                                                       \
        FLIP2_LAUNCH_CHECK();                                                                                   \
        return;                                                                                                 \
    }
    WGRAD_CASE(16, 16)
    WGRAD_CASE(16, 32)
    WGRAD_CASE(32, 32)
    WGRAD_CASE(32, 64)
    WGRAD_CASE(64, 64)
    WGRAD_CASE(96, 32)
    WGRAD_CASE(48, 16)
#undef WGRAD_CASE
    throw Error(kErrArgument, "conv3 wgrad: unsupported channel pair");
}

void launch_relu_mask_bf16(float* g, const uint16_t* act, uint16_t* gb, long long count, hipStream_t s) {
    FLIP2_KERNEL(relu_mask_bf16_kernel, g1(count, 256), dim3(256), 0, s, g, act, gb, count);
    FLIP2_LAUNCH_CHECK();
}
void launch_avgpool(const uint16_t* X, uint16_t* P, int n, int C, hipStream_t s) {
    const long long total = (long long)(n / 2) * (n / 2) * (n / 2) * C;
    FLIP2_KERNEL(avgpool_kernel, g1(total, 256), dim3(256), 0, s, X, P, n, C);
    FLIP2_LAUNCH_CHECK();
}
void launch_avgpool_bwd(const float* dP, float* dX, int n, int C, hipStream_t s) {
    const long long total = (long long)n * n * n * C;
    FLIP2_KERNEL(avgpool_bwd_kernel, g1(total, 256), dim3(256), 0, s, dP, dX, n, C);
    FLIP2_LAUNCH_CHECK();
}
void launch_upcat(const uint16_t* A, const uint16_t* B, uint16_t* U, int n, int Ca, int Cb, hipStream_t s) {
    const long long total = (long long)n * n * n * (Ca + Cb);
    FLIP2_KERNEL(upcat_kernel, g1(total, 256), dim3(256), 0, s, A, B, U, n, Ca, Cb);
    FLIP2_LAUNCH_CHECK();
}
void launch_upcat_bwd(const float* dU, float* dA, float* dB, int n, int Ca, int Cb, hipStream_t s) {
    const long long total = (long long)(n / 2) * (n / 2) * (n / 2) * Ca + (long long)n * n * n * Cb;
    FLIP2_KERNEL(upcat_bwd_kernel, g1(total, 256), dim3(256), 0, s, dU, dA, dB, n, Ca, Cb);
    FLIP2_LAUNCH_CHECK();
}
void launch_head_fwd(const uint16_t* X, const float* w, const float* b, const uint8_t* mask, const float* scale,
                     float* out, long long nvox, int C, hipStream_t s) {
    FLIP2_KERNEL(head_fwd_kernel, g1(nvox, 256), dim3(256), 0, s, X, w, b, mask, scale, out, nvox, C);
    FLIP2_LAUNCH_CHECK();
}
void launch_head_bwd(const float* dOut, const uint16_t* X, const float* w, const uint8_t* mask, const float* scale,
                     float* dX, float* dw, float* db, long long nvox, int C, hipStream_t s) {
    FLIP2_REQUIRE(C <= 16, kErrArgument, "head: at most 16 channels");
    FLIP2_KERNEL(head_bwd_kernel, g1(nvox, 256), dim3(256), 0, s, dOut, X, w, mask, scale, dX, dw, db, nvox, C);
    FLIP2_LAUNCH_CHECK();
}
void launch_adam(float* w, const float* g, float* m, float* v, uint16_t* wbf, int count, float lr, float b1, float b2,
                 float eps, int t, hipStream_t s) {
    const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
    FLIP2_KERNEL(adam_kernel, g1(count, 256), dim3(256), 0, s, w, g, m, v, wbf, count, lr, b1, b2, eps, c1, c2);
    FLIP2_LAUNCH_CHECK();
}
void launch_to_bf16(const float* a, uint16_t* b, long long count, hipStream_t s) {
    FLIP2_KERNEL(to_bf16_kernel, g1(count, 256), dim3(256), 0, s, a, b, count);
    FLIP2_LAUNCH_CHECK();
}
void launch_bf16_to_f32(const uint16_t* a, float* b, long long count, hipStream_t s) {
    FLIP2_KERNEL(bf16_to_f32_kernel, g1(count, 256), dim3(256), 0, s, a, b, count);
    FLIP2_LAUNCH_CHECK();
}
void launch_zero_f32(float* a, long long count, hipStream_t s) {
    if (count <= 0) return;
    FLIP2_KERNEL(zero_f32_kernel, g1(count, 256), dim3(256), 0, s, a, count);
    FLIP2_LAUNCH_CHECK();
}

}  // namespace nn
}  // namespace flip2
