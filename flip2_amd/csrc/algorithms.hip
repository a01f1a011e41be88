// The reference's algorithm free functions (algorithms.hpp) on this
// library's kernels.
#include "algorithms.hpp"

#include "lists.hpp"
#include "rc_kernels.hpp"
#include "scan_impl.hpp"
#include "sort.hpp"

#include <cmath>

namespace flip2 {
namespace {

constexpr double kSorOmega = 1.9;          // voxelSolveFunctions.cu:396
constexpr uint32_t kSorCheckEvery = 1024;  // voxelSolveFunctions.cu:395

// Stream-ordered scratch (the reference's cudaMallocAsync / cudaFreeAsync).
class StreamScratch {
public:
    StreamScratch(size_t bytes, hipStream_t s) : s_(s) {
        if (bytes) FLIP2_HIP(hipMallocAsync(&p_, bytes, s));
    }
    ~StreamScratch() {
        if (p_) (void)hipFreeAsync(p_, s_);
    }
    StreamScratch(const StreamScratch&) = delete;
    StreamScratch& operator=(const StreamScratch&) = delete;
    template <typename T>
    T* as() const {
        return reinterpret_cast<T*>(p_);
    }

private:
    void* p_ = nullptr;
    hipStream_t s_;
};

void require_rc_geometry(double radius, uint32_t refinementLevel, uint32_t numVoxelsPerNode, uint32_t numVoxels1D) {
    FLIP2_REQUIRE(radius == kRcRadius && refinementLevel == kRcRefinement && numVoxelsPerNode == kRcVoxelsPerNode &&
                      numVoxels1D == kRcVoxels1D,
                  kErrArgument, "only radius 2, refinement level 1 and 8^3 voxels per node are supported");
}

// dt / (density h_sub^2) with h_sub = cellSize / (2 << refinementLevel) in
// float (calcSubCellWidth, voxelSolveFunctions.hu:10-12).
double solve_scale(const Grid& g, double dt, double density) {
    const double vs = (double)(g.cellSize / (float)(2u << kRcRefinement));
    return dt / (density * vs * vs);
}

uint32_t read_u32(const uint32_t* dev, hipStream_t s) {
    uint32_t v = 0;
    FLIP2_HIP(hipMemcpyAsync(&v, dev, sizeof(v), hipMemcpyDeviceToHost, s));
    FLIP2_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace

void cudaFindGridCell(double* px, double* py, double* pz, uint32_t numParticles, Grid grid, uint32_t* gridPosition,
                      hipStream_t stream) {
    if (numParticles == 0) return;
    rc::launch_root_cell(px, py, pz, numParticles, grid, gridPosition, stream);
    FLIP2_HIP(hipStreamSynchronize(stream));
}

void cudaRadixSortUint(uint32_t numElements, uint32_t* inArray, uint32_t* outArray, uint32_t* sortedIndices,
                       uint32_t* front, uint32_t* back, hipStream_t frontStream, hipStream_t backStream,
                       uint32_t*& reorderedIndicesRelativeToOriginal) {
    (void)backStream;
    if (numElements == 0) return;
    FLIP2_REQUIRE(inArray && outArray && sortedIndices && front && back, kErrArgument, "radix sort: null buffer");
    if (!reorderedIndicesRelativeToOriginal)
        FLIP2_HIP(hipMalloc((void**)&reorderedIndicesRelativeToOriginal, sizeof(uint32_t) * numElements));
    const size_t ws = sort_workspace_bytes(numElements, 0, 32);
    StreamScratch scratch(ws, frontStream);
    // keys: inArray -> outArray (front as the ping-pong buffer); permutation
    // into the caller's buffer (sortedIndices as its ping-pong buffer)
    sort_pairs_u32(inArray, nullptr, outArray, reorderedIndicesRelativeToOriginal, front, sortedIndices, numElements,
                   0, 32, scratch.as<void>(), ws, frontStream);
    launch_copy(inArray, outArray, sizeof(uint32_t) * numElements, frontStream);
    FLIP2_HIP(hipStreamSynchronize(frontStream));
}

void cudaSortParticlesByGridNode(uint32_t numParticles, uint32_t*& gridPosition,
                                 uint32_t*& reorderedIndicesRelativeToOriginal, hipStream_t stream) {
    if (numParticles == 0) return;
    StreamScratch tmp(sizeof(uint32_t) * 4 * (size_t)numParticles, stream);
    uint32_t* t = tmp.as<uint32_t>();
    cudaRadixSortUint(numParticles, gridPosition, t, t + numParticles, t + 2 * (size_t)numParticles,
                      t + 3 * (size_t)numParticles, stream, stream, reorderedIndicesRelativeToOriginal);
}

uint32_t cudaMarkUniqueGridCellsAndCount(uint32_t numParticles, uint32_t* gridPosition, uint32_t* uniqueGridNodes,
                                         hipStream_t stream) {
    if (numParticles == 0) return 0;
    const size_t ws = scan_workspace_bytes(numParticles);
    StreamScratch scratch(ws + 256, stream);
    uint32_t* total = reinterpret_cast<uint32_t*>(scratch.as<char>() + ws);
    const bool vec = ((reinterpret_cast<uintptr_t>(gridPosition) | reinterpret_cast<uintptr_t>(uniqueGridNodes)) &
                      15u) == 0;
    scan_launch(ScanHeadFlagInput{gridPosition}, uniqueGridNodes, numParticles, false, vec, scratch.as<void>(), ws,
                total, stream);
    return read_u32(total, stream);
}

void cudaMapNodeIndicesToParticles(uint32_t numParticles, uint32_t* uniqueGridNodes,
                                   uint32_t* gridNodeIndicesToFirstParticleIndex, hipStream_t stream) {
    if (numParticles == 0) return;
    rc::launch_map_nodes(uniqueGridNodes, numParticles, gridNodeIndicesToFirstParticleIndex, stream);
}

void cudaGetFirstNodeInYRows(uint32_t numUniqueGridNodes, uint32_t* gridNodeIndicesToFirstParticleIndex,
                             uint32_t* gridNodeIDs, uint32_t* yDimFirstNodeIndex, const Grid& grid,
                             hipStream_t stream) {
    rc::launch_fill_u32(yDimFirstNodeIndex, grid.sizeX * grid.sizeY, numUniqueGridNodes, stream);
    FLIP2_HIP(hipStreamSynchronize(stream));
    if (numUniqueGridNodes)
        rc::launch_row_table(numUniqueGridNodes, gridNodeIndicesToFirstParticleIndex, gridNodeIDs, yDimFirstNodeIndex,
                             grid.sizeX, stream);
}

void cudaCalcDivU(const DeviceVec<uint32_t>& nodeIndexToFirstVoxelIndex, const DeviceVec<uint32_t>& voxelIDs,
                  const DeviceVec<double>& voxelsUx, const DeviceVec<double>& voxelsUy,
                  const DeviceVec<double>& voxelsUz, double radius, uint32_t refinementLevel, const Grid& grid,
                  const DeviceVec<uint32_t>& yDimFirstNodeIndex,
                  const DeviceVec<uint32_t>& gridNodeIndicesToFirstParticleIndex, const DeviceVec<uint32_t>& gridNodes,
                  uint32_t numVoxelsPerNode, uint32_t numVoxels1D, DeviceVec<double>& divU, hipStream_t stream) {
    require_rc_geometry(radius, refinementLevel, numVoxelsPerNode, numVoxels1D);
    const uint32_t B = nodeIndexToFirstVoxelIndex.size();
    FLIP2_REQUIRE(gridNodeIndicesToFirstParticleIndex.size() >= B, kErrRange, "calcDivU: node table too short");
    FLIP2_REQUIRE(divU.size() >= voxelIDs.size() && voxelsUx.size() >= voxelIDs.size() &&
                      voxelsUy.size() >= voxelIDs.size() && voxelsUz.size() >= voxelIDs.size(),
                  kErrRange, "calcDivU: voxel arrays shorter than voxelIDs");
    if (B == 0) return;
    rc::launch_divu(B, grid, gridNodes.devPtr(), gridNodeIndicesToFirstParticleIndex.devPtr(),
                    yDimFirstNodeIndex.devPtr(), nodeIndexToFirstVoxelIndex.devPtr(), voxelIDs.devPtr(),
                    voxelsUx.devPtr(), voxelsUy.devPtr(), voxelsUz.devPtr(), divU.devPtr(), stream);
}

namespace rc {
// cudaGSiteration's loop (voxelSolveFunctions.cu:391-412): a check (sampled
// residual maximum) after every iteration whose index is a multiple of 1024;
// the iterations between two checks run inside one LDS-resident launch.
double gs_solve(uint32_t B, uint32_t V, const uint32_t* nodeOff, const uint32_t* voxelIDs, const char* solids,
                const double* divU, double* p, double* residuals, double scale, double threshold,
                uint32_t maxIterations, double* absmaxTmp, double* scalar, hipStream_t stream, uint32_t* itersOut) {
    double maxResidual = 10.0, prevMaxResidual = 100.0;
    uint32_t iterations = 0;
    while (maxResidual > threshold && iterations < maxIterations) {
        uint32_t run;
        if (iterations % kSorCheckEvery == 0) run = 1;
        else run = (iterations / kSorCheckEvery + 1) * kSorCheckEvery - iterations + 1;
        if (iterations + run > maxIterations) run = maxIterations - iterations;
        launch_sor(B, nodeOff, voxelIDs, solids, divU, p, residuals, scale, kSorOmega, run, stream);
        const uint32_t lastIt = iterations + run - 1;
        iterations += run;
        if (lastIt % kSorCheckEvery == 0) {
            prevMaxResidual = maxResidual;
            Ptr3 arrs{{residuals, nullptr, nullptr}};
            launch_sampled_absmax(arrs, 1, V, absmaxTmp, scalar, stream);
            FLIP2_HIP(hipMemcpyAsync(&maxResidual, scalar, sizeof(double), hipMemcpyDeviceToHost, stream));
            FLIP2_HIP(hipStreamSynchronize(stream));
        }
        if (iterations == maxIterations ||
            (std::fabs(prevMaxResidual - maxResidual) < threshold / 1000 && maxResidual > threshold))
            break;
    }
    if (itersOut) *itersOut = iterations;
    return maxResidual;
}
}  // namespace rc

double cudaGSiteration(const uint32_t& numVoxelsPerNode, const uint32_t& numVoxels1D, const uint32_t& refinementLevel,
                       const DeviceVec<uint32_t>& nodeIndexUsedVoxels, const DeviceVec<uint32_t>& voxelIDs,
                       const DeviceVec<char>& solids, const DeviceVec<double>& divU, DeviceVec<double>& p,
                       DeviceVec<double>& residuals, const double& radius, const double& density, const double& dt,
                       const Grid& grid, const double& threshold, const uint32_t& maxIterations, hipStream_t stream) {
    require_rc_geometry(radius, refinementLevel, numVoxelsPerNode, numVoxels1D);
    const uint32_t B = nodeIndexUsedVoxels.size(), V = voxelIDs.size();
    FLIP2_REQUIRE(solids.size() >= V && divU.size() >= V && p.size() >= V && residuals.size() >= V, kErrRange,
                  "GSiteration: voxel arrays shorter than voxelIDs");
    if (B == 0 || V == 0) return 0.0;
    const size_t tmpDoubles = rc::sampled_absmax_tmp_doubles(V, 1) + 1;
    StreamScratch tmp(sizeof(double) * tmpDoubles, stream);
    return rc::gs_solve(B, V, nodeIndexUsedVoxels.devPtr(), voxelIDs.devPtr(), solids.devPtr(), divU.devPtr(),
                        p.devPtr(), residuals.devPtr(), solve_scale(grid, dt, density), threshold, maxIterations,
                        tmp.as<double>() + 1, tmp.as<double>(), stream, nullptr);
}

void cudaVelocityUpdate(uint32_t numVoxelsPerNode, uint32_t numVoxels1D, double dt, double radius, double density,
                        const DeviceVec<uint32_t>& nodeIndexUsedVoxels, const DeviceVec<uint32_t>& voxelIDs,
                        const DeviceVec<char>& solids, const DeviceVec<double>& p, DeviceVec<double>& voxelsUx,
                        DeviceVec<double>& voxelsUy, DeviceVec<double>& voxelsUz, uint32_t refinementLevel, Grid grid,
                        hipStream_t stream) {
    require_rc_geometry(radius, refinementLevel, numVoxelsPerNode, numVoxels1D);
    const uint32_t B = nodeIndexUsedVoxels.size(), V = voxelIDs.size();
    FLIP2_REQUIRE(solids.size() >= V && p.size() >= V && voxelsUx.size() >= V && voxelsUy.size() >= V &&
                      voxelsUz.size() >= V,
                  kErrRange, "velocityUpdate: voxel arrays shorter than voxelIDs");
    if (B == 0) return;
    rc::launch_velocity_update(B, nodeIndexUsedVoxels.devPtr(), voxelIDs.devPtr(), solids.devPtr(), p.devPtr(),
                               voxelsUx.devPtr(), voxelsUy.devPtr(), voxelsUz.devPtr(),
                               solve_scale(grid, dt, density), stream);
    FLIP2_HIP(hipStreamSynchronize(stream));
}

void applyGravity(const DeviceVec<char>& solids, DeviceVec<double>& voxelsUz, double dt, hipStream_t stream) {
    FLIP2_REQUIRE(voxelsUz.size() >= solids.size(), kErrRange, "applyGravity: velocity array shorter than solids");
    if (solids.size()) rc::launch_gravity(solids.size(), solids.devPtr(), voxelsUz.devPtr(), (double)(-9.8f) * dt,
                                          false, stream);
}

void removeGravity(const DeviceVec<char>& solids, DeviceVec<double>& voxelsUz, double dt, hipStream_t stream) {
    FLIP2_REQUIRE(voxelsUz.size() >= solids.size(), kErrRange, "removeGravity: velocity array shorter than solids");
    if (solids.size()) rc::launch_gravity(solids.size(), solids.devPtr(), voxelsUz.devPtr(), (double)(-9.8f) * dt,
                                          true, stream);
}

void cudaFindSubCell(double* px, double* py, double* pz, uint32_t numParticles, Grid grid, uint32_t* gridPosition,
                     uint32_t* subCellsTouchedX, uint32_t* subCellsTouchedY, uint32_t* subCellsTouchedZ,
                     DeviceVec<uint32_t>& subCellPositionX, DeviceVec<uint32_t>& subCellPositionY,
                     DeviceVec<uint32_t>& subCellPositionZ, uint32_t numRefinementLevels, double radius,
                     hipStream_t stream) {
    FLIP2_REQUIRE(numRefinementLevels < 16 && radius >= 0.0 && radius < 64.0, kErrArgument,
                  "findSubCell: refinement level or radius out of range");
    if (numParticles == 0) return;
    rc::launch_sub_cell_counts(px, py, pz, numParticles, grid, gridPosition, numRefinementLevels, radius,
                               subCellsTouchedX, subCellsTouchedY, subCellsTouchedZ, stream);
    uint32_t* touched[3] = {subCellsTouchedX, subCellsTouchedY, subCellsTouchedZ};
    uint32_t total[3];
    for (int d = 0; d < 3; ++d) {
        cudaParallelPrefixSum(numParticles, touched[d], stream);
        FLIP2_HIP(hipMemcpyAsync(&total[d], touched[d] + numParticles - 1, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 stream));
    }
    FLIP2_HIP(hipStreamSynchronize(stream));
    subCellPositionX.resizeAsync(total[0], stream);
    subCellPositionY.resizeAsync(total[1], stream);
    subCellPositionZ.resizeAsync(total[2], stream);
    rc::launch_sub_cell_lists(px, py, pz, numParticles, grid, gridPosition, numRefinementLevels, radius,
                              subCellsTouchedX, subCellsTouchedY, subCellsTouchedZ, subCellPositionX.devPtr(),
                              subCellPositionY.devPtr(), subCellPositionZ.devPtr(), stream);
}

void cudaSumParticlesPerNodeAndWriteNumUsedVoxels(uint32_t numGridNodes, uint32_t numParticles,
                                                  uint32_t* gridNodeIndicesToFirstParticleIndex,
                                                  uint32_t* subCellsTouchedPerParticle, uint32_t* subCellsDim,
                                                  uint32_t* numNonZeroVoxels, uint32_t* numParticlesInVoxelLists,
                                                  uint32_t numVoxelsPerNode, hipStream_t stream) {
    FLIP2_REQUIRE(numVoxelsPerNode == kRcVoxelsPerNode, kErrArgument, "per-voxel lists: 512 voxels per node only");
    if (numGridNodes == 0) return;
    rc::launch_particles_per_node(numGridNodes, numParticles, gridNodeIndicesToFirstParticleIndex,
                                  subCellsTouchedPerParticle, subCellsDim, numNonZeroVoxels, numParticlesInVoxelLists,
                                  stream);
    cudaParallelPrefixSum(numGridNodes, numNonZeroVoxels, stream);
    cudaParallelPrefixSum(numGridNodes, numParticlesInVoxelLists, stream);
    FLIP2_HIP(hipStreamSynchronize(stream));
}

void cudaParticleListCreate(uint32_t totalNumberVoxelsDimension, uint32_t numGridNodes, uint32_t numParticles,
                            uint32_t numVoxelsPerNode, uint32_t* voxelIDs, uint32_t* perVoxelParticleListStartIndices,
                            uint32_t* numUsedVoxelsPerNode, uint32_t* firstParticleListIndexPerNode,
                            uint32_t* subCellsTouchedPerParticle, uint32_t* subCellsDim,
                            uint32_t* firstParticleInNodeIndex, uint32_t* particleLists, hipStream_t stream) {
    (void)totalNumberVoxelsDimension;
    FLIP2_REQUIRE(numVoxelsPerNode == kRcVoxelsPerNode, kErrArgument, "per-voxel lists: 512 voxels per node only");
    if (numGridNodes == 0) return;
    rc::launch_particle_lists(numGridNodes, numParticles, numUsedVoxelsPerNode, firstParticleListIndexPerNode,
                              subCellsTouchedPerParticle, subCellsDim, firstParticleInNodeIndex, voxelIDs,
                              perVoxelParticleListStartIndices, particleLists, stream);
    FLIP2_HIP(hipStreamSynchronize(stream));
}

template <>
void cudaParallelPrefixSum<uint32_t>(uint32_t numElements, uint32_t* array, hipStream_t stream) {
    if (numElements == 0) return;
    const size_t ws = scan_workspace_bytes(numElements);
    StreamScratch scratch(ws, stream);
    scan_u32(array, array, numElements, false, scratch.as<void>(), ws, nullptr, stream);
}

}  // namespace flip2
