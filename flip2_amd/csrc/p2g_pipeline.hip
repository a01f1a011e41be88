#include "p2g_pipeline.hpp"

#include "scan_impl.hpp"
#include "sort.hpp"

namespace flip2 {

P2GPipeline::P2GPipeline(uint32_t maxParticles, const Grid& g, DeviceManager* pool)
    : grid_(g), maxParticles_(maxParticles) {
    FLIP2_REQUIRE(g.sizeX % 8 == 0 && g.sizeY % 8 == 0 && g.sizeZ % 8 == 0 && g.sizeX && g.sizeY && g.sizeZ,
                  kErrArgument, "grid sizes must be positive multiples of 8");
    FLIP2_REQUIRE(g.cellSize > 0.0f, kErrArgument, "cell size must be positive");
    const uint64_t cells = (uint64_t)g.sizeX * g.sizeY * g.sizeZ;
    FLIP2_REQUIRE(cells < (1ull << 31), kErrRange, "grid has more than 2^31 cells");
    FLIP2_REQUIRE(maxParticles > 0 && maxParticles < (1u << 30), kErrRange, "particle count must be in [1, 2^30)");
    numCells_ = (uint32_t)cells;
    for (auto* v : {&keys_, &keysSorted_, &perm_, &keysAlt_, &permAlt_, &counts_, &cellStart_, &stats_}) v->setPool(pool);
    for (auto* v : {&packed_, &sorted_}) v->setPool(pool);
    for (auto* v : {&occ_, &touched_, &prevTouched_}) v->setPool(pool);
    touchedList_.setPool(pool);
    brickStart_.setPool(pool);
    ws_.setPool(pool);
    keyBits_ = std::max<uint32_t>(1, bits_for(numCells_ - 1));
    keys_.resize(maxParticles); keys_.host().clear(); keys_.host().shrink_to_fit();
    keysSorted_.resizeAsync(maxParticles, nullptr);
    perm_.resizeAsync(maxParticles, nullptr);
    keysAlt_.resizeAsync(maxParticles, nullptr);
    permAlt_.resizeAsync(maxParticles, nullptr);
    counts_.resizeAsync(numCells_ + 1, nullptr);
    cellStart_.resizeAsync(numCells_ + 1, nullptr);
    packed_.resizeAsync(2 * maxParticles, nullptr);
    sorted_.resizeAsync(2 * maxParticles, nullptr);
    const size_t ws = std::max(sort_workspace_bytes(maxParticles, 0, keyBits_), scan_workspace_bytes(numCells_ + 1));
    ws_.resizeAsync((uint32_t)ws, nullptr);
    stats_.resizeAsync(1u, nullptr);
    stats_.zeroDevice();
}

void P2GPipeline::run(uint32_t n, const float* x, const float* y, const float* z, const float* vx, const float* vy,
                      const float* vz, float* U, float* V, float* W, hipStream_t s) {
    FLIP2_REQUIRE(n == 0 || (x && y && z && vx && vy && vz), kErrArgument, "null particle array");
    const ns::Float6 src{{x, y, z, vx, vy, vz}};
    runImpl(n, &src, nullptr, U, V, W, s, nullptr);
}

void P2GPipeline::runRecords(uint32_t n, const float4* rec, float* U, float* V, float* W, hipStream_t s,
                             const ns::FacePrep* prep, bool keysReady) {
    FLIP2_REQUIRE(n == 0 || rec, kErrArgument, "null particle records");
    runImpl(n, nullptr, rec, U, V, W, s, prep, keysReady);
}

void P2GPipeline::runImpl(uint32_t n, const ns::Float6* soa, const float4* rec, float* U, float* V, float* W,
                          hipStream_t s, const ns::FacePrep* prep, bool keysReady) {
    FLIP2_REQUIRE(n <= maxParticles_, kErrRange, "more particles than the pipeline was created for");
    FLIP2_REQUIRE(U && V && W, kErrArgument, "null output array");
    lastN_ = n;
    const bool timed = timing_ && n > 0;
    const uint32_t passes = sort_passes(0, keyBits_);
    if (timed && pending_ == (int)events_.size()) finishTiming();
    hipEvent_t* ev = timed ? events_[pending_].data() : nullptr;
    if (timed) FLIP2_HIP(hipEventRecord(ev[0], s));
    const float4* records = rec;
    if (soa) {
        ns::launch_bin_pack(*soa, n, grid_, keys_.devPtr(), packed_.devPtr(), stats_.devPtr(), s);
        records = packed_.devPtr();
    } else if (keysReady) {
        ns::launch_clear_stats(stats_.devPtr(), s);  // what the binning pass also does
    } else {
        ns::launch_bin_rec(rec, n, grid_, keys_.devPtr(), stats_.devPtr(), s);
    }
    if (timed) FLIP2_HIP(hipEventRecord(ev[1], s));
    if (n) {
        sort_pairs_u32(keys_.devPtr(), nullptr, keysSorted_.devPtr(), perm_.devPtr(), keysAlt_.devPtr(),
                       permAlt_.devPtr(), n, 0, keyBits_, ws_.devPtr(), ws_.size(), s, timed ? ev + 2 : nullptr);
        ns::launch_gather_rec(perm_.devPtr(), n, records, sorted_.devPtr(), stats_.devPtr(), s);
    }
    if (timed) FLIP2_HIP(hipEventRecord(ev[3 + passes], s));
    // The fused run (prep) needs the cell offsets of the occupied bricks
    // only, which it derives from the sorted keys itself; the dense per-cell
    // counts and their scan over every cell of the grid are for the plain
    // transfer.
    sparseCells_ = prep != nullptr;
    if (!prep) ns::launch_cell_counts(keysSorted_.devPtr(), n, numCells_, counts_.devPtr(), s);
    if (timed) FLIP2_HIP(hipEventRecord(ev[4 + passes], s));
    if (!prep)
        scan_u32(counts_.devPtr(), cellStart_.devPtr(), numCells_ + 1, true, ws_.devPtr(), ws_.size(), nullptr, s);
    if (timed) FLIP2_HIP(hipEventRecord(ev[5 + passes], s));
    if (prep) {
        const uint32_t numBricks = numCells_ >> 9;
        if (occ_.size() < numBricks) {
            occ_.resizeAsync(numBricks, s);
            touched_.resizeAsync(numBricks, s);
            prevTouched_.resizeAsync(numBricks, s);
            touchedList_.resizeAsync(numBricks + 1 + 8, s);  // list, count, 8 work counters
            brickStart_.resizeAsync(numBricks + 1, s);
            prevTouchedValid_ = false;
        }
        ns::FacePrep fp = *prep;
        if (!prevTouchedValid_) fp.full = 1;
        ns::launch_p2g_prepare(sorted_.devPtr(), keysSorted_.devPtr(), n, cellStart_.devPtr(), brickStart_.devPtr(),
                               grid_, stats_.devPtr(), fp, U, V, W,
                               {occ_.devPtr(), touched_.devPtr(), touchedList_.devPtr(),
                                touchedList_.devPtr() + numBricks, prevTouched_.devPtr(), listCounters()},
                               s);
        touched_.swapDevice(prevTouched_);  // this run's flags are the next run's previous ones
        prevTouchedValid_ = true;
    } else {
        ns::launch_p2g(sorted_.devPtr(), cellStart_.devPtr(), grid_, stats_.devPtr(), U, V, W, nullptr, nullptr,
                       nullptr, s);
    }
    if (timed) {
        FLIP2_HIP(hipEventRecord(ev[6 + passes], s));
        ++pending_;
    }
}

void P2GPipeline::enableTiming(int capacity) {
    finishTiming();
    timing_ = capacity > 0;
    for (auto& set : events_)
        for (auto& e : set) (void)hipEventDestroy(e);
    events_.clear();
    if (timing_) {
        const uint32_t passes = sort_passes(0, keyBits_);
        events_.assign(capacity, std::vector<hipEvent_t>(7 + passes));
        for (auto& set : events_)
            for (auto& e : set) FLIP2_HIP(hipEventCreate(&e));
        timingNames_ = {"bin", "sort_setup"};
        for (uint32_t p = 0; p < passes; ++p) timingNames_.push_back("sort_pass" + std::to_string(p));
        timingNames_.push_back("gather");
        timingNames_.push_back("counts");
        timingNames_.push_back("scan");
        timingNames_.push_back("p2g");
        timingMs_.assign(timingNames_.size(), 0.0);
    }
}

void P2GPipeline::resetTiming() {
    std::fill(timingMs_.begin(), timingMs_.end(), 0.0);
    timedRuns_ = 0;
}

// Waits for the recorded runs and adds each interval to its kernel.
void P2GPipeline::finishTiming() {
    for (int r = 0; r < pending_; ++r) {
        auto& ev = events_[r];
        FLIP2_HIP(hipEventSynchronize(ev.back()));
        for (size_t k = 0; k + 1 < ev.size(); ++k) {
            float ms = 0.f;
            FLIP2_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            timingMs_[k] += ms;
        }
        ++timedRuns_;
    }
    pending_ = 0;
}

P2GPipeline::~P2GPipeline() {
    for (auto& set : events_)
        for (auto& e : set) (void)hipEventDestroy(e);
}

size_t P2GPipeline::get(const std::string& name, void* host, size_t capacity, hipStream_t s) {
    auto out = [&](auto& v, uint32_t count) -> size_t {
        using T = typename std::remove_reference<decltype(*v.devPtr())>::type;
        if (host) {
            FLIP2_REQUIRE(capacity >= count, kErrRange, "buffer too small");
            v.downloadTo(reinterpret_cast<T*>(host), count, s);
            FLIP2_HIP(hipStreamSynchronize(s));
        }
        return count;
    };
    if (name == "keys") return out(keys_, lastN_);
    if (name == "sortedKeys") return out(keysSorted_, lastN_);
    if (name == "perm") return out(perm_, lastN_);
    if (name == "counts" || name == "cellStart")
        FLIP2_REQUIRE(!sparseCells_, kErrState,
                      "counts / cellStart: the last run was a fused transfer, which keeps the offsets of the "
                      "occupied bricks only");
    if (name == "counts") return out(counts_, numCells_ + 1);
    if (name == "cellStart") return out(cellStart_, numCells_ + 1);
    static const char* names[6] = {"x", "y", "z", "vx", "vy", "vz"};
    for (int c = 0; c < 6; ++c)
        if (name == names[c]) {
            if (host) {
                FLIP2_REQUIRE(capacity >= lastN_, kErrRange, "buffer too small");
                std::vector<float4> r(2 * (size_t)lastN_);
                if (lastN_) sorted_.downloadTo(r.data(), 2 * lastN_, s);
                FLIP2_HIP(hipStreamSynchronize(s));
                float* h = reinterpret_cast<float*>(host);
                for (uint32_t i = 0; i < lastN_; ++i) {
                    const float* f = reinterpret_cast<const float*>(&r[2 * (size_t)i]);
                    h[i] = f[c];
                }
            }
            return lastN_;
        }
    throw Error(kErrArgument, "unknown pipeline array " + name);
}

}  // namespace flip2
