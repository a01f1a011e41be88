// North-star transfer kernels on 8^3 bricks (ns_kernels.hip).
#pragma once

#include "common.hpp"
#include "ns_solver.hpp"

namespace flip2 {
namespace ns {

struct Float6 {
    const float* p[6];  // x, y, z, vx, vy, vz
};
using Float6In = Float6;
struct Float6Out {
    float* p[6];
};

// Brick-tiled cell keys of FP32 positions, clamped into the grid.
void launch_bin(const float* x, const float* y, const float* z, uint32_t n, Grid g, uint32_t* keys, hipStream_t s);
// Particle records: 2 x float4 per particle, (x, y, z, vx), (vy, vz, 0, 0).
// stats (u32, may be null) is the P2G statistic of the step: the bin kernels
// zero it, the gather records max |velocity component| (float bits) in it.
void launch_bin_pack(Float6 src, uint32_t n, Grid g, uint32_t* keys, float4* rec, uint32_t* stats, hipStream_t s);
void launch_bin_rec(const float4* rec, uint32_t n, Grid g, uint32_t* keys, uint32_t* stats, hipStream_t s);
// stats[0] = 0 (the binning kernels' reset of the speed statistic, for a run
// whose keys were computed elsewhere).
void launch_clear_stats(uint32_t* stats, hipStream_t s);
// dst[i] = src[perm[i]] (records)
void launch_gather_rec(const uint32_t* perm, uint32_t n, const float4* src, float4* dst, uint32_t* stats,
                       hipStream_t s);
void launch_pack(Float6 src, uint32_t n, float4* rec, hipStream_t s);
void launch_unpack(const float4* rec, uint32_t n, Float6Out dst, hipStream_t s);
// counts[0..numCells] zeroed, then counts[key] = run length of key in the
// sorted keys (counts[numCells] stays 0, so an exclusive scan over
// numCells + 1 entries ends with n).
void launch_cell_counts(const uint32_t* sortedKeys, uint32_t n, uint32_t numCells, uint32_t* counts, hipStream_t s);
// The statistics for a P2G whose inputs did not come through the pipeline.
// (the particle count is cellStart[numCells])
void launch_p2g_stats(const float4* sortedRec, const uint32_t* cellStart, uint32_t numCells, uint32_t* stats,
                      hipStream_t s);
// Trilinear MAC P2G from records sorted by cell key, one workgroup per
// brick, 64-bit fixed-point accumulation scaled from stats.  WU/WV/WW
// (weight sums) may be null.
void launch_p2g(const float4* sortedRec, const uint32_t* cellStart, Grid g, const uint32_t* stats, float* U, float* V,
                float* W, float* WU, float* WV, float* WW, hipStream_t s);

// The prepare stage's inputs and saved field, for the fused transfer.
struct FacePrep {
    float *U0, *V0, *W0;  // pre-projection field (written)
    const uint8_t* solid;  // brick-tiled solid cells, may be null
    SolidBox box;
    float gdt;  // gravity * dt, added to v faces
    // 1: write the faces of every untouched brick; 0: only of those that the
    // previous fused run touched -- the caller guarantees that the others
    // still hold the prepared zero-transfer values of these same parameters
    int full;
};
// Work arrays of the fused transfer: occ, touched (one byte per brick),
// list (one word per brick), count (one word); prevTouched: the previous
// fused run's touched flags (read when FacePrep::full is 0).
struct TouchedBricks {
    uint8_t *occ, *touched;
    uint32_t *list, *count;
    const uint8_t* prevTouched;
    uint32_t* counters;  // 8 words: the transfer's work counter per XCD (cleared by the launch)
};
// launch_p2g followed by launch_prepare_faces (the same U, V, W, U0, V0, W0),
// with the transfer done only for bricks within one brick of a particle and
// the others given the prepared faces of a zero transfer.  Takes the sorted
// keys instead of the dense per-cell offsets and writes brickStart
// (numBricks + 1: first particle of each brick, n at the end), t.occ and
// cellStart of the occupied bricks only (entries b * 512 .. b * 512 + 512);
// the other entries of cellStart keep what they held.
void launch_p2g_prepare(const float4* sortedRec, const uint32_t* sortedKeys, uint32_t n, uint32_t* cellStart,
                        uint32_t* brickStart, Grid g, const uint32_t* stats, const FacePrep& prep, float* U, float* V,
                        float* W, TouchedBricks t, hipStream_t s);

}  // namespace ns
}  // namespace flip2
