// Single-pass u32 prefix scan with decoupled look-back, written for CDNA
// wave64.  Replaces the reference's recursive 512-element Blelloch scan
// (engine/algorithms/parallelPrefixSumKernels.cu:5-77), which launches two
// kernels per level and allocates its block sums per call.
//
// Tile = 256 threads x 32 u32 = 8192 elements.  Wave w of a tile owns the
// 2048 elements [w*2048, w*2048+2048) as 8 rows of 64 lanes x uint4, so every
// load and store is a coalesced 1 KiB row.  Tile ids come from an atomic
// counter, so a tile only ever waits on tiles that started before it.  The
// look-back uses all 256 lanes: one status word of an earlier tile per lane,
// so an inclusive prefix advances 256 tiles per status round trip.
#pragma once

#include "common.hpp"

namespace flip2 {

constexpr int kScanThreads = 256;
constexpr int kScanRows = 8;                      // uint4 rows per wave
constexpr int kScanTile = kScanThreads * 4 * kScanRows;  // 8192

// Look-back status word: bits 32..33 hold the flag, bits 0..31 the value.
constexpr uint64_t kFlagAggregate = 1ull << 32;
constexpr uint64_t kFlagInclusive = 2ull << 32;

__device__ inline uint64_t status_load(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void status_store(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Inclusive scan across the 64 lanes of a wave.
__device__ inline uint32_t wave_inclusive_scan(uint32_t x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

__device__ inline uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// Plain array input.  `vec` selects 16-byte loads (pointer 16-byte aligned).
struct ScanArrayInput {
    const uint32_t* in;
    __device__ inline uint4 load4(uint32_t i, uint32_t n, bool vec) const {
        if (vec && i + 4 <= n) return *reinterpret_cast<const uint4*>(in + i);
        uint4 r;
        r.x = (i + 0 < n) ? in[i + 0] : 0u;
        r.y = (i + 1 < n) ? in[i + 1] : 0u;
        r.z = (i + 2 < n) ? in[i + 2] : 0u;
        r.w = (i + 3 < n) ? in[i + 3] : 0u;
        return r;
    }
};

// Segment head flags of a sorted key array (1 where key[i] != key[i-1] or
// i == 0), i.e. markUniqueGridCells (particleToGridFunctions.cu:303-316)
// fused into the scan input.
struct ScanHeadFlagInput {
    const uint32_t* keys;
    __device__ inline uint4 load4(uint32_t i, uint32_t n, bool vec) const {
        uint32_t k[5];
        k[0] = (i > 0 && i - 1 < n) ? keys[i - 1] : 0u;
        if (vec && i + 4 <= n) {
            uint4 v = *reinterpret_cast<const uint4*>(keys + i);
            k[1] = v.x; k[2] = v.y; k[3] = v.z; k[4] = v.w;
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) k[e + 1] = (i + e < n) ? keys[i + e] : 0u;
        }
        uint4 r;
        r.x = (i + 0 < n) ? ((i == 0 || k[1] != k[0]) ? 1u : 0u) : 0u;
        r.y = (i + 1 < n) ? ((k[2] != k[1]) ? 1u : 0u) : 0u;
        r.z = (i + 2 < n) ? ((k[3] != k[2]) ? 1u : 0u) : 0u;
        r.w = (i + 3 < n) ? ((k[4] != k[3]) ? 1u : 0u) : 0u;
        return r;
    }
};

// Clears the look-back words and the tile counter before a scan.
template <int Unused = 0>
__global__ void scan_clear_kernel(uint64_t* status, uint32_t count, uint32_t* counter) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) status[i] = 0;
    if (i == 0) *counter = 0;
}

// exclusive != 0 -> out[i] = sum(in[0..i-1]); else sum(in[0..i]).
// If total != nullptr the grand total (mod 2^32) is stored there.
template <class Input>
__global__ void __launch_bounds__(kScanThreads)
scan_kernel(Input input, uint32_t* __restrict__ out, uint32_t n, int exclusive, int vec,
            uint64_t* __restrict__ status, uint32_t* __restrict__ counter,
            uint32_t* __restrict__ total) {
    __shared__ uint32_t s_tile;
    __shared__ uint32_t s_wave_total[kScanThreads / 64];
    __shared__ uint32_t s_prefix;
    __shared__ int s_firstIncl[kScanThreads / 64];
    __shared__ uint32_t s_lookSum[kScanThreads / 64];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    if (tid == 0) s_tile = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t waveBase = tile * (uint32_t)kScanTile + (uint32_t)wave * (64u * 4u * kScanRows);

    uint4 v[kScanRows];
    uint32_t rowExcl[kScanRows];
    uint32_t running = 0;
#pragma unroll
    for (int r = 0; r < kScanRows; ++r) {
        const uint32_t i = waveBase + (uint32_t)r * 256u + (uint32_t)lane * 4u;
        v[r] = input.load4(i, n, vec != 0);
        const uint32_t t = v[r].x + v[r].y + v[r].z + v[r].w;
        const uint32_t incl = wave_inclusive_scan(t);
        rowExcl[r] = running + incl - t;
        running += __shfl(incl, 63, 64);
    }
    if (lane == 0) s_wave_total[wave] = running;
    __syncthreads();

    uint32_t waveExcl = 0, aggregate = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; ++w) {
        const uint32_t t = s_wave_total[w];
        if (w < wave) waveExcl += t;
        aggregate += t;
    }

    // Look-back: lane t of the workgroup reads the status of tile (base - t).
    // Every earlier tile publishes its aggregate without waiting on anything,
    // so spinning until all 256 words are set always ends.
    if (tile == 0) {
        if (tid == 0) {
            status_store(&status[0], kFlagInclusive | aggregate);
            s_prefix = 0;
        }
    } else {
        if (tid == 0) status_store(&status[tile], kFlagAggregate | aggregate);
        uint32_t prefix = 0;  // accumulated in thread 0
        int64_t base = (int64_t)tile - 1;
        for (;;) {
            const int64_t idx = base - tid;
            uint64_t s = (idx >= 0) ? status_load(&status[idx]) : kFlagInclusive;
            while ((s >> 32) == 0) {
                __builtin_amdgcn_s_sleep(1);
                s = status_load(&status[idx]);
            }
            const uint64_t inclMask = __ballot((s >> 32) == 2);
            if (lane == 0) s_firstIncl[wave] = inclMask ? __builtin_ctzll(inclMask) + wave * 64 : kScanThreads;
            __syncthreads();
            int first = kScanThreads;
#pragma unroll
            for (int w = 0; w < kScanThreads / 64; ++w) first = min(first, s_firstIncl[w]);
            const uint32_t ws = wave_sum(tid <= first ? (uint32_t)s : 0u);
            if (lane == 0) s_lookSum[wave] = ws;
            __syncthreads();
            if (tid == 0)
                for (int w = 0; w < kScanThreads / 64; ++w) prefix += s_lookSum[w];
            if (first < kScanThreads) break;
            base -= kScanThreads;
            __syncthreads();  // s_firstIncl / s_lookSum are rewritten next step
        }
        if (tid == 0) {
            status_store(&status[tile], kFlagInclusive | (uint32_t)(prefix + aggregate));
            s_prefix = prefix;
        }
    }
    if (tid == 0 && total && (uint64_t)(tile + 1) * kScanTile >= n) *total = s_prefix + aggregate;
    __syncthreads();
    const uint32_t base = s_prefix + waveExcl;

#pragma unroll
    for (int r = 0; r < kScanRows; ++r) {
        const uint32_t i = waveBase + (uint32_t)r * 256u + (uint32_t)lane * 4u;
        uint32_t acc = base + rowExcl[r];
        uint4 o;
        if (exclusive) {
            o.x = acc; acc += v[r].x;
            o.y = acc; acc += v[r].y;
            o.z = acc; acc += v[r].z;
            o.w = acc;
        } else {
            acc += v[r].x; o.x = acc;
            acc += v[r].y; o.y = acc;
            acc += v[r].z; o.z = acc;
            acc += v[r].w; o.w = acc;
        }
        if (vec && i + 4 <= n) {
            *reinterpret_cast<uint4*>(out + i) = o;
        } else {
            if (i + 0 < n) out[i + 0] = o.x;
            if (i + 1 < n) out[i + 1] = o.y;
            if (i + 2 < n) out[i + 2] = o.z;
            if (i + 3 < n) out[i + 3] = o.w;
        }
    }
}

// Workspace: look-back status words + tile counter.
inline size_t scan_workspace_bytes(uint32_t n) {
    const size_t tiles = (size_t)ceil_div(n ? n : 1, kScanTile);
    return tiles * sizeof(uint64_t) + 256;
}

// Host launcher (defined in scan.hip for the plain-array input; the generic
// form below is used by other translation units for fused inputs).
template <class Input>
void scan_launch(Input input, uint32_t* out, uint32_t n, bool exclusive, bool vec, void* ws,
                 size_t wsBytes, uint32_t* totalDev, hipStream_t stream) {
    if (n == 0) return;
    const uint32_t tiles = ceil_div(n, kScanTile);
    FLIP2_REQUIRE(wsBytes >= scan_workspace_bytes(n), kErrArgument, "scan workspace too small");
    uint64_t* status = reinterpret_cast<uint64_t*>(ws);
    uint32_t* counter = reinterpret_cast<uint32_t*>(status + tiles);
    FLIP2_KERNEL(scan_clear_kernel<0>, dim3(ceil_div(tiles + 1, 256)), dim3(256), 0, stream, status,
                       tiles, counter);
    FLIP2_LAUNCH_CHECK();
    FLIP2_KERNEL(scan_kernel<Input>, dim3(tiles), dim3(kScanThreads), 0, stream, input, out, n,
                       exclusive ? 1 : 0, vec ? 1 : 0, status, counter, totalDev);
    FLIP2_LAUNCH_CHECK();
}

// Plain u32 array scan (scan.hip).
void scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, bool exclusive, void* ws, size_t wsBytes,
              uint32_t* totalDev, hipStream_t stream);

}  // namespace flip2
