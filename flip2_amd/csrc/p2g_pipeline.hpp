// Particle->grid transfer pipeline on 8^3 bricks: bin -> key/value radix
// sort -> gather -> per-cell counts -> exclusive scan -> P2G.
#pragma once

#include "common.hpp"
#include "device.hpp"
#include "ns_kernels.hpp"

#include <string>
#include <vector>

namespace flip2 {

class P2GPipeline {
public:
    // pool: allocator of the pipeline's arrays (null: hipMalloc); it must
    // outlive the pipeline.
    P2GPipeline(uint32_t maxParticles, const Grid& g, DeviceManager* pool = nullptr);
    ~P2GPipeline();
    P2GPipeline(const P2GPipeline&) = delete;
    P2GPipeline& operator=(const P2GPipeline&) = delete;
    // Inputs: device FP32 particle state (unsorted).  Outputs: MAC face
    // velocities U, V, W (n^3 each, brick-tiled) and, inside the pipeline, the
    // sorted state, sort permutation and per-cell start offsets.
    void run(uint32_t n, const float* x, const float* y, const float* z, const float* vx, const float* vy,
             const float* vz, float* U, float* V, float* W, hipStream_t s);
    // Same, from particle records (2 x float4 per particle).  With prep, the
    // transfer is followed by the prepare stage in the same kernels
    // (ns::launch_p2g_prepare): U, V, W then hold the prepared faces.
    // keysReady: keys() already holds the cell key of every record (written
    // by the step's G2P + RK2 kernel), so the binning pass is skipped.
    void runRecords(uint32_t n, const float4* rec, float* U, float* V, float* W, hipStream_t s,
                    const ns::FacePrep* prep = nullptr, bool keysReady = false);
    // The unsorted cell keys of the next run (see runRecords' keysReady).
    uint32_t* keys() { return keys_.devPtr(); }
    size_t get(const std::string& name, void* host, size_t capacity, hipStream_t s);
    // Per-kernel timing of subsequent runs (HIP events on the run's stream):
    // ms[k] accumulates over runs for kernels "bin", "sort_setup" (clear +
    // histogram + offsets), "sort_pass0".. , "gather", "counts", "scan", "p2g".
    // Event sets for `capacity` runs are recorded without host waits; they are
    // read when all sets are used or at finishTiming().
    void enableTiming(int capacity);
    void finishTiming();
    int timedKernels() const { return (int)timingNames_.size(); }
    const char* timedKernelName(int k) const { return timingNames_[k].c_str(); }
    double timedKernelMs(int k) const { return timingMs_[k]; }
    int timedRuns() const { return timedRuns_; }
    void resetTiming();
    const uint32_t* cellStart() const { return cellStart_.devPtr(); }
    // The last fused run's touched bricks: list and (device) count
    // (valid after a runRecords with FacePrep).
    const uint32_t* touchedList() const { return touchedList_.devPtr(); }
    const uint32_t* touchedCount() const { return touchedList_.devPtr() + (numCells_ >> 9); }
    // 8 device counters after the count, for kernels that walk the touched
    // list with a work counter per XCD (they clear them first)
    uint32_t* listCounters() { return touchedList_.devPtr() + (numCells_ >> 9) + 1; }
    // Per-cell counts: written by a plain run only (sparseCells() false).
    const uint32_t* counts() const { return counts_.devPtr(); }
    // The last run was a fused one (runRecords with FacePrep): cellStart
    // holds the cells of the occupied bricks only (occupancy(): one flag per
    // brick), and brickStart() the first particle of every brick.
    bool sparseCells() const { return sparseCells_; }
    const uint8_t* occupancy() const { return occ_.devPtr(); }
    const uint32_t* brickStart() const { return brickStart_.devPtr(); }
    // max |velocity component| of the last run's particles (float bits)
    const uint32_t* maxSpeedBits() const { return stats_.devPtr(); }
    // Sorted particle records of the last run; a caller may swap them into
    // its own state (DeviceVec::swapDevice).
    DeviceVec<float4>& sortedVec() { return sorted_; }
    const float4* sortedRecords() const { return sorted_.devPtr(); }
    uint32_t numCells() const { return numCells_; }
    const Grid& grid() const { return grid_; }
    uint32_t keyBits() const { return keyBits_; }

private:
    Grid grid_;
    uint32_t maxParticles_, numCells_, keyBits_, lastN_ = 0;
    DeviceVec<uint32_t> keys_, keysSorted_, perm_, keysAlt_, permAlt_, counts_, cellStart_;
    DeviceVec<float4> packed_, sorted_;
    DeviceVec<uint8_t> ws_;
    DeviceVec<uint32_t> stats_;  // P2G statistic of the step: max |velocity component| (float bits)
    bool timing_ = false;
    int pending_ = 0;
    std::vector<std::vector<hipEvent_t>> events_;
    std::vector<std::string> timingNames_;
    std::vector<double> timingMs_;
    int timedRuns_ = 0;
    DeviceVec<uint8_t> occ_, touched_, prevTouched_;  // fused transfer + prepare: per-brick flags
    bool prevTouchedValid_ = false;  // prevTouched_ holds the last fused run's flags
    DeviceVec<uint32_t> touchedList_;    // and the touched bricks (list, then count)
    DeviceVec<uint32_t> brickStart_;     // first particle of each brick (fused runs)
    bool sparseCells_ = false;
    void runImpl(uint32_t n, const ns::Float6* soa, const float4* rec, float* U, float* V, float* W, hipStream_t s,
                 const ns::FacePrep* prep, bool keysReady = false);
};

}  // namespace flip2
