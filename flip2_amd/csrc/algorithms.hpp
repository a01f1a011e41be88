// The reference's algorithm free functions with their signatures, on the HIP
// kernels of this library.  Two type names change: cudaStream_t becomes
// hipStream_t and CudaVec<T> becomes flip2::DeviceVec<T>; `uint` is uint32_t
// and `char` solids stay char.  Scratch memory the reference takes with
// cudaMallocAsync is taken stream-ordered here too (hipMallocAsync) and
// released before return.  Every function enqueues on `stream` and, where the
// reference synchronises or returns a device value, synchronises it.
//
// Reference-compatible geometry only: radius 2, refinement level 1, 8^3
// voxels per node (engine/particles.cu:17-38); other values are refused.
#pragma once

#include "common.hpp"
#include "device.hpp"

namespace flip2 {

// particleToGridFunctions.hu:10 -- clamp positions into the domain (written
// back) and compute the root-cell key of each particle.
void cudaFindGridCell(double* px, double* py, double* pz, uint32_t numParticles, Grid grid, uint32_t* gridPosition,
                      hipStream_t stream);

// particleToGridFunctions.hu:16 -- stable ascending sort of the keys in
// gridPosition (sorted in place) with the permutation written to
// reorderedIndicesRelativeToOriginal (reorderedIndices[i] = original index of
// the i-th sorted key).  The reference returns the permutation in a buffer it
// allocates; here it fills the caller's buffer of numParticles elements, or,
// if the pointer is null, allocates one with hipMalloc that the caller frees.
// Both pointers keep their buffers otherwise.
void cudaSortParticlesByGridNode(uint32_t numParticles, uint32_t*& gridPosition,
                                 uint32_t*& reorderedIndicesRelativeToOriginal, hipStream_t stream);

// particleToGridFunctions.hu:18 -- 1-based node id of every particle of a
// sorted key array (inclusive scan of the segment heads); returns the node
// count.
uint32_t cudaMarkUniqueGridCellsAndCount(uint32_t numParticles, uint32_t* gridPosition, uint32_t* uniqueGridNodes,
                                         hipStream_t stream);

// particleToGridFunctions.hu:20 -- first particle index of every node.
void cudaMapNodeIndicesToParticles(uint32_t numParticles, uint32_t* uniqueGridNodes,
                                   uint32_t* gridNodeIndicesToFirstParticleIndex, hipStream_t stream);

// particleToGridFunctions.hu:22 -- row table: entries [0, sizeX*sizeY) set to
// numUniqueGridNodes (as the reference clears them), then for each (y, z) row
// with nodes the cell key of its first node, at index key / sizeX.
void cudaGetFirstNodeInYRows(uint32_t numUniqueGridNodes, uint32_t* gridNodeIndicesToFirstParticleIndex,
                             uint32_t* gridNodeIDs, uint32_t* yDimFirstNodeIndex, const Grid& grid,
                             hipStream_t stream);

// voxelSolveFunctions.hu:14 -- divergence of the voxel velocities (own voxels
// plus the neighbours' apron contributions).
void cudaCalcDivU(const DeviceVec<uint32_t>& nodeIndexToFirstVoxelIndex, const DeviceVec<uint32_t>& voxelIDs,
                  const DeviceVec<double>& voxelsUx, const DeviceVec<double>& voxelsUy,
                  const DeviceVec<double>& voxelsUz, double radius, uint32_t refinementLevel, const Grid& grid,
                  const DeviceVec<uint32_t>& yDimFirstNodeIndex,
                  const DeviceVec<uint32_t>& gridNodeIndicesToFirstParticleIndex, const DeviceVec<uint32_t>& gridNodes,
                  uint32_t numVoxelsPerNode, uint32_t numVoxels1D, DeviceVec<double>& divU, hipStream_t stream);

// voxelSolveFunctions.hu:19 -- brick-local red-black SOR (w = 1.9) until the
// sampled residual maximum, checked every 1024 iterations, drops below
// threshold or stalls, or maxIterations; returns that residual.
double cudaGSiteration(const uint32_t& numVoxelsPerNode, const uint32_t& numVoxels1D, const uint32_t& refinementLevel,
                       const DeviceVec<uint32_t>& nodeIndexUsedVoxels, const DeviceVec<uint32_t>& voxelIDs,
                       const DeviceVec<char>& solids, const DeviceVec<double>& divU, DeviceVec<double>& p,
                       DeviceVec<double>& residuals, const double& radius, const double& density, const double& dt,
                       const Grid& grid, const double& threshold, const uint32_t& maxIterations, hipStream_t stream);

// voxelSolveFunctions.hu:24 -- pressure-gradient update of the three
// velocity components.
void cudaVelocityUpdate(uint32_t numVoxelsPerNode, uint32_t numVoxels1D, double dt, double radius, double density,
                        const DeviceVec<uint32_t>& nodeIndexUsedVoxels, const DeviceVec<uint32_t>& voxelIDs,
                        const DeviceVec<char>& solids, const DeviceVec<double>& p, DeviceVec<double>& voxelsUx,
                        DeviceVec<double>& voxelsUy, DeviceVec<double>& voxelsUz, uint32_t refinementLevel, Grid grid,
                        hipStream_t stream);

// voxelSolveFunctions.hu:25-26 -- add / remove -9.8f * dt on the non-solid
// voxels of the array passed (the reference passes voxelsUy under the
// parameter name voxelsUz).
void applyGravity(const DeviceVec<char>& solids, DeviceVec<double>& voxelsUz, double dt, hipStream_t stream);
void removeGravity(const DeviceVec<char>& solids, DeviceVec<double>& voxelsUz, double dt, hipStream_t stream);

// parallelPrefixSumKernels.hu:53 -- inclusive in-place scan (mod 2^32).
template <typename T>
void cudaParallelPrefixSum(uint32_t numElements, T* array, hipStream_t stream);
template <>
void cudaParallelPrefixSum<uint32_t>(uint32_t numElements, uint32_t* array, hipStream_t stream);

// radixSortKernels.hu:16 -- stable 32-bit LSD sort of inArray.  On return
// inArray holds the sorted keys (as after the reference's 32 passes) and
// reorderedIndicesRelativeToOriginal the permutation (a null pointer gets a
// hipMalloc'd buffer the caller frees).  outArray, sortedIndices, front and
// back (numElements each) are scratch; backStream is not used (one stream
// orders the whole sort).
void cudaRadixSortUint(uint32_t numElements, uint32_t* inArray, uint32_t* outArray, uint32_t* sortedIndices,
                       uint32_t* front, uint32_t* back, hipStream_t frontStream, hipStream_t backStream,
                       uint32_t*& reorderedIndicesRelativeToOriginal);

// particleToGridFunctions.hu:12 -- sub-cells of each particle's root cell
// (with the apron) whose staggered x / y / z velocity point lies within
// radius sub-cells: per-particle counts, inclusive-scanned into
// subCellsTouched* (numParticles each), and the concatenated lists of
// node-local voxel ids resized into subCellPosition*.  The reference never
// calls it from its step.
void cudaFindSubCell(double* px, double* py, double* pz, uint32_t numParticles, Grid grid, uint32_t* gridPosition,
                     uint32_t* subCellsTouchedX, uint32_t* subCellsTouchedY, uint32_t* subCellsTouchedZ,
                     DeviceVec<uint32_t>& subCellPositionX, DeviceVec<uint32_t>& subCellPositionY,
                     DeviceVec<uint32_t>& subCellPositionZ, uint32_t numRefinementLevels, double radius,
                     hipStream_t stream);

// perVoxelParticleListFunctions.hu:8 -- per node, the list entries
// (numParticlesInVoxelLists) and distinct voxels (numNonZeroVoxels) of its
// particles' sub-cell lists, both inclusive-scanned over the nodes.
// Entries outside [0, numVoxelsPerNode) are skipped (the reference indexes
// past its shared array with them).
void cudaSumParticlesPerNodeAndWriteNumUsedVoxels(uint32_t numGridNodes, uint32_t numParticles,
                                                  uint32_t* gridNodeIndicesToFirstParticleIndex,
                                                  uint32_t* subCellsTouchedPerParticle, uint32_t* subCellsDim,
                                                  uint32_t* numNonZeroVoxels, uint32_t* numParticlesInVoxelLists,
                                                  uint32_t numVoxelsPerNode, hipStream_t stream);

// perVoxelParticleListFunctions.hu:13 -- per node, its used voxels in
// ascending order (voxelIDs at numUsedVoxelsPerNode[b-1]), the start of
// each one's particle list (perVoxelParticleListStartIndices) and the lists
// (particleLists: the node's particles touching the voxel, ascending).
// totalNumberVoxelsDimension is unused, as in the reference.
void cudaParticleListCreate(uint32_t totalNumberVoxelsDimension, uint32_t numGridNodes, uint32_t numParticles,
                            uint32_t numVoxelsPerNode, uint32_t* voxelIDs, uint32_t* perVoxelParticleListStartIndices,
                            uint32_t* numUsedVoxelsPerNode, uint32_t* firstParticleListIndexPerNode,
                            uint32_t* subCellsTouchedPerParticle, uint32_t* subCellsDim,
                            uint32_t* firstParticleInNodeIndex, uint32_t* particleLists, hipStream_t stream);

namespace rc {
// The SOR loop of cudaGSiteration on raw arrays (shared with
// Particles::gsSolve): B nodes, V voxels, scale = dt / (density h_sub^2).
// absmaxTmp holds sampled_absmax_tmp_doubles(V, 1) doubles, scalar one.
double gs_solve(uint32_t B, uint32_t V, const uint32_t* nodeOff, const uint32_t* voxelIDs, const char* solids,
                const double* divU, double* p, double* residuals, double scale, double threshold,
                uint32_t maxIterations, double* absmaxTmp, double* scalar, hipStream_t stream, uint32_t* iters);
}  // namespace rc

}  // namespace flip2
