// flip2_tester: a C++ harness on the drop-in surface.  It is the friend class
// the reference's Particles grants (engine/particles.hu:94) and reads and
// writes the members by the reference's names (px, gridCell, voxelsUx,
// stream, grid, size, ...; particles.hu:47-92), drives the device vectors
// through the CudaVec method set (host mirror, upload/download on the stream,
// zeroDeviceAsync, getMax, swapDevicePtr; cudaVec.hu:20-43) and calls the
// algorithm free functions with the reference's signatures
// (particleToGridFunctions.hu, voxelSolveFunctions.hu,
// parallelPrefixSumKernels.hu, radixSortKernels.hu).  Every result is checked
// against a host computation or against the Particles stage that computes the
// same thing; the first mismatch ends the run with exit status 1.
//
//   flip2_tester [particles=20000] [cells=24] [seed=7]
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <numeric>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../algorithms.hpp"
#include "../particles.hpp"

namespace flip2 {

namespace {
void check(bool ok, const std::string& what) {
    if (!ok) throw std::runtime_error("check failed: " + what);
}

template <typename T>
std::vector<T> fetch(DeviceVec<T>& v, hipStream_t s) {
    v.download(s);
    FLIP2_HIP(hipStreamSynchronize(s));
    return v.host();
}

// The reference's sampled maximum evaluated over every tile of every level
// (cudaVec.cu:174-233 as written: each 512-tile keeps pick(tile[0],
// tile[256]), levels padded with 0 until one value is left).
template <typename T>
T sampled_max_host(const std::vector<T>& a, bool absolute) {
    auto mag = [](T v) { return v < T(0) ? T(-v) : v; };
    auto pick = [&](T x, T y) { return absolute ? (mag(x) > mag(y) ? x : y) : (x > y ? x : y); };
    std::vector<T> cur = a;
    do {
        std::vector<T> next(cur.size() / 512 + 1);
        for (size_t t = 0; t < next.size(); ++t) {
            const T x = 512 * t < cur.size() ? cur[512 * t] : T(0);
            const T y = 512 * t + 256 < cur.size() ? cur[512 * t + 256] : T(0);
            next[t] = pick(x, y);
        }
        cur.swap(next);
    } while (cur.size() > 1);
    return cur[0];
}
}  // namespace

class ParticleSystemTester {
public:
    ParticleSystemTester(uint32_t n, uint32_t cells, uint64_t seed) : particles(n), cells_(cells), rng_(seed) {}

    void run() {
        Particles& P = particles;
        P.setDomain(0.0, 0.0, 0.0, cells_, cells_, cells_, 1.0 / cells_);
        check(P.size == P.px.size() && P.grid.sizeX == cells_ && P.grid.cellSize == 1.0f / cells_, "domain");

        // positions through the host mirrors: a block in the lower corner,
        // as a dam break; velocities zeroed on the device
        const double L = P.grid.sizeX * (double)P.grid.cellSize;
        std::uniform_real_distribution<double> d(0.05 * L, 0.55 * L);
        for (uint32_t i = 0; i < P.size; ++i) {
            P.px[i] = d(rng_);
            P.py[i] = d(rng_);
            P.pz[i] = d(rng_);
        }
        P.px.upload(P.stream);
        P.py.upload(P.stream);
        P.pz.upload(P.stream);
        P.vx.zeroDeviceAsync(P.stream);
        P.vy.zeroDeviceAsync(P.stream);
        P.vz.zeroDeviceAsync(P.stream);
        const std::vector<double> x0(P.px.host()), y0(P.py.host()), z0(P.pz.host());

        // keys: the same cells as a host binning of the (in-range) positions,
        // the quotient narrowed to float before the floor as the reference
        // does (particleToGridFunctions.cu:40-42)
        P.alignParticlesToGrid();
        std::vector<uint32_t> keys = fetch(P.gridCell, P.stream);
        std::map<uint32_t, uint32_t> histogram;
        const double h = (double)P.grid.cellSize;
        for (uint32_t i = 0; i < P.size; ++i) {
            const uint32_t cx = (uint32_t)std::floor((float)((x0[i] - (double)P.grid.negX) / h));
            const uint32_t cy = (uint32_t)std::floor((float)((y0[i] - (double)P.grid.negY) / h));
            const uint32_t cz = (uint32_t)std::floor((float)((z0[i] - (double)P.grid.negZ) / h));
            check(keys[i] == cx + cy * cells_ + cz * cells_ * cells_, "gridCell of particle " + std::to_string(i));
            ++histogram[keys[i]];
        }

        // sort: ascending, same multiset, positions follow the permutation
        P.sortParticles();
        std::vector<uint32_t> sorted = fetch(P.gridCell, P.stream);
        std::vector<uint32_t> perm = fetch(P.reorderedGridIndices, P.stream);
        std::vector<double> xs = fetch(P.px, P.stream);
        std::map<uint32_t, uint32_t> after;
        for (uint32_t i = 0; i < P.size; ++i) {
            check(i == 0 || sorted[i - 1] < sorted[i] || (sorted[i - 1] == sorted[i] && perm[i - 1] < perm[i]),
                  "stable ascending order at " + std::to_string(i));
            check(sorted[i] == keys[perm[i]] && xs[i] == x0[perm[i]], "permutation at " + std::to_string(i));
            ++after[sorted[i]];
        }
        check(after == histogram, "key counts after the sort");

        // nodes and voxels; the free functions recompute the node tables
        P.generateVoxels();
        DeviceVec<uint32_t> keysCopy(P.size, P.deviceManager()), uniq(P.size, P.deviceManager());
        launch_copy(keysCopy.devPtr(), P.gridCell.devPtr(), P.gridCell.bytes(), P.stream);
        const uint32_t B = cudaMarkUniqueGridCellsAndCount(P.size, keysCopy.devPtr(), uniq.devPtr(), P.stream);
        check(B == P.numUsedGridNodes && B == histogram.size(), "node count");
        check(fetch(uniq, P.stream) == fetch(P.uniqueGridNodeIndices, P.stream), "uniqueGridNodeIndices");
        DeviceVec<uint32_t> first(B, P.deviceManager());
        cudaMapNodeIndicesToParticles(P.size, uniq.devPtr(), first.devPtr(), P.stream);
        const std::vector<uint32_t> firstHost = fetch(first, P.stream);
        check(firstHost == fetch(P.gridNodeIndicesToFirstParticleIndex, P.stream), "first particle of each node");
        {
            uint32_t b = 0, i = 0;
            for (const auto& kv : histogram) {
                check(firstHost[b] == i, "node segment start " + std::to_string(b));
                i += kv.second;
                ++b;
            }
        }
        DeviceVec<uint32_t> rows(cells_ * cells_, P.deviceManager());
        cudaGetFirstNodeInYRows(B, first.devPtr(), P.gridCell.devPtr(), rows.devPtr(), P.grid, P.stream);
        const std::vector<uint32_t> rowsHost = fetch(rows, P.stream);
        const std::vector<uint32_t> rowsP = fetch(P.yDimNumUsedGridNodes, P.stream);
        for (uint32_t r = 0; r < cells_ * cells_; ++r) {
            uint32_t expect = B;  // first key of the row (key / sizeX == r) among the nodes
            for (const auto& kv : histogram)
                if (kv.first / cells_ == r) {
                    expect = kv.first;
                    break;
                }
            check(rowsHost[r] == expect && rowsP[r] == expect, "row table entry " + std::to_string(r));
        }
        const std::vector<uint32_t> nodeOff = fetch(P.nodeIndexUsedVoxels, P.stream);
        check(nodeOff.size() == B && nodeOff.back() == P.voxelIDsUsed.size() && P.numUsedVoxels == nodeOff.back(),
              "inclusive voxel offsets");

        // velocities: a swirl, through the host mirrors in sorted order
        for (uint32_t i = 0; i < P.size; ++i) {
            P.vx[i] = 0.3 * std::sin(6.0 * xs[i]);
            P.vy[i] = -0.2 + 0.1 * std::cos(5.0 * xs[i]);
            P.vz[i] = 0.05 * (double)(i % 7);
        }
        P.vx.upload(P.stream);
        P.vy.upload(P.stream);
        P.vz.upload(P.stream);

        // getMax: the sampled maximum of the reference, both modes and types
        for (bool absolute : {false, true}) {
            check(P.vx.getMax(P.stream, absolute) == sampled_max_host(P.vx.host(), absolute), "getMax(vx)");
            check(P.vz.getMax(P.stream, absolute) == sampled_max_host(P.vz.host(), absolute), "getMax(vz)");
            check(P.gridCell.getMax(P.stream, absolute) == sampled_max_host(sorted, absolute), "getMax(gridCell)");
        }

        // P2G, then gravity, divergence, SOR and the velocity update through
        // the free functions on copies, against the Particles stages
        P.particleVelToVoxels();
        const uint32_t V = P.numUsedVoxels;
        auto copyD = [&](DeviceVec<double>& src) {
            auto* v = new DeviceVec<double>(V, P.deviceManager());
            launch_copy(v->devPtr(), src.devPtr(), (size_t)V * sizeof(double), P.stream);
            return std::unique_ptr<DeviceVec<double>>(v);
        };
        const double dt = 1.0 / 240.0, density = 0.014;
        auto Uy = copyD(P.voxelsUy);
        applyGravity(P.solids, *Uy, dt, P.stream);
        P.applyGravity(dt);
        check(fetch(*Uy, P.stream) == fetch(P.voxelsUy, P.stream), "applyGravity");
        removeGravity(P.solids, *Uy, dt, P.stream);
        P.removeGravity(dt);
        check(fetch(*Uy, P.stream) == fetch(P.voxelsUy, P.stream), "removeGravity");
        P.applyGravity(dt);
        DeviceVec<double> divU(V, P.deviceManager());
        cudaCalcDivU(P.nodeIndexUsedVoxels, P.voxelIDsUsed, P.voxelsUx, P.voxelsUy, P.voxelsUz, P.radius,
                     P.refinementLevel, P.grid, P.yDimNumUsedGridNodes, P.gridNodeIndicesToFirstParticleIndex,
                     P.gridCell, P.numVoxelsPerNode, P.numVoxels1D, divU, P.stream);
        P.calcDivU();
        check(fetch(divU, P.stream) == fetch(P.divU, P.stream), "calcDivU");
        DeviceVec<double> p(V, P.deviceManager()), res(V, P.deviceManager());
        p.zeroDeviceAsync(P.stream);
        P.p.zeroDeviceAsync(P.stream);
        const double rFree = cudaGSiteration(P.numVoxelsPerNode, P.numVoxels1D, P.refinementLevel,
                                             P.nodeIndexUsedVoxels, P.voxelIDsUsed, P.solids, divU, p, res, P.radius,
                                             density, dt, P.grid, 0.01, 3000u, P.stream);
        uint32_t its = 0;
        const double rP = P.gsSolve(dt, 0.01, 3000u, &its);
        check(rFree == rP && its > 0, "GSiteration residual");
        check(fetch(p, P.stream) == fetch(P.p, P.stream), "GSiteration pressure");
        P.dt = dt;
        auto Ux = copyD(P.voxelsUx), Uy2 = copyD(P.voxelsUy), Uz = copyD(P.voxelsUz);
        cudaVelocityUpdate(P.numVoxelsPerNode, P.numVoxels1D, dt, P.radius, density, P.nodeIndexUsedVoxels,
                           P.voxelIDsUsed, P.solids, P.p, *Ux, *Uy2, *Uz, P.refinementLevel, P.grid, P.stream);
        P.updateVoxelVelocities();
        check(fetch(*Ux, P.stream) == fetch(P.voxelsUx, P.stream) &&
                  fetch(*Uy2, P.stream) == fetch(P.voxelsUy, P.stream) &&
                  fetch(*Uz, P.stream) == fetch(P.voxelsUz, P.stream),
              "velocity update");

        // swapDevicePtr adopts a buffer from the pool and frees the old one
        {
            DeviceVec<double> a(V, P.deviceManager());
            double* fresh = nullptr;
            check(!P.deviceManager().cuMalloc(fresh, V) || P.deviceManager().poolBytesReserved() > 0, "cuMalloc");
            if (!fresh) fresh = reinterpret_cast<double*>(P.deviceManager().allocate(sizeof(double) * V));
            launch_copy(fresh, P.divU.devPtr(), (size_t)V * sizeof(double), P.stream);
            a.swapDevicePtr(fresh);
            check(a.devPtr() == fresh && a.size() == V, "swapDevicePtr");
            check(fetch(a, P.stream) == fetch(P.divU, P.stream), "adopted contents");
        }

        // the primitives alone: radix sort and inclusive scan vs std::
        std::uniform_int_distribution<uint32_t> ku(0, 0xffffffffu);
        const uint32_t m = 70001;
        std::vector<uint32_t> in(m);
        for (auto& k : in) k = ku(rng_) >> (ku(rng_) % 20);
        DeviceVec<uint32_t> a(m, P.deviceManager()), o(m, P.deviceManager()), si(m, P.deviceManager()),
            fr(m, P.deviceManager()), bk(m, P.deviceManager()), pm(m, P.deviceManager());
        a.host() = in;
        a.upload(P.stream);
        uint32_t* permPtr = pm.devPtr();
        cudaRadixSortUint(m, a.devPtr(), o.devPtr(), si.devPtr(), fr.devPtr(), bk.devPtr(), P.stream, P.stream,
                          permPtr);
        std::vector<uint32_t> idx(m);
        std::iota(idx.begin(), idx.end(), 0u);
        std::stable_sort(idx.begin(), idx.end(), [&](uint32_t i, uint32_t j) { return in[i] < in[j]; });
        const std::vector<uint32_t> sk = fetch(a, P.stream), sp = fetch(pm, P.stream);
        for (uint32_t i = 0; i < m; ++i) check(sp[i] == idx[i] && sk[i] == in[idx[i]], "radix sort at " + std::to_string(i));
        a.host() = in;
        a.upload(P.stream);
        cudaParallelPrefixSum(m, a.devPtr(), P.stream);
        std::vector<uint32_t> incl(m);
        std::partial_sum(in.begin(), in.end(), incl.begin());
        check(fetch(a, P.stream) == incl, "inclusive prefix sum");

        // whole frames: the pool stops growing once the scene's sizes settle
        P.setSolverLimits(0.01, 2048);
        P.initialize();
        P.solveFrame(240.0);
        const auto chunks = P.deviceManager().chunkAllocations();
        for (int f = 0; f < 3; ++f) P.solveFrame(240.0);
        std::vector<double> xe = fetch(P.px, P.stream);
        for (double v : xe) check(std::isfinite(v) && v >= 0.0 && v <= L, "positions after frames");
        std::printf("frames: %u nodes, %u voxels, pool %llu of %llu bytes in use, %llu chunk allocations (%llu after "
                    "the first frame)\n",
                    P.numUsedGridNodes, P.numUsedVoxels, P.deviceManager().poolBytesInUse(),
                    P.deviceManager().poolBytesReserved(), P.deviceManager().chunkAllocations(), chunks);
        check(P.deviceManager().chunkAllocations() <= chunks + 1, "pool growth in steady state");
    }

    Particles particles;

private:
    uint32_t cells_;
    std::mt19937_64 rng_;
};

}  // namespace flip2

int main(int argc, char** argv) {
    const uint32_t n = argc > 1 ? (uint32_t)std::strtoul(argv[1], nullptr, 10) : 20000u;
    const uint32_t cells = argc > 2 ? (uint32_t)std::strtoul(argv[2], nullptr, 10) : 24u;
    const uint64_t seed = argc > 3 ? std::strtoull(argv[3], nullptr, 10) : 7u;
    try {
        flip2::ParticleSystemTester t(n, cells, seed);
        t.run();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "flip2_tester: %s\n", e.what());
        return 1;
    }
    std::printf("tester ok\n");
    return 0;
}
