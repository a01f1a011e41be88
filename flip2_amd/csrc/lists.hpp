// Sub-cell lists and per-voxel particle lists (lists.hip): launchers of the
// kernels behind cudaFindSubCell and the per-voxel list functions.
#pragma once

#include "common.hpp"

namespace flip2 {
namespace rc {

// Per-particle counts of touched sub-cells per velocity component.
void launch_sub_cell_counts(const double* px, const double* py, const double* pz, uint32_t n, Grid g,
                            const uint32_t* keys, uint32_t level, double radius, uint32_t* cx, uint32_t* cy,
                            uint32_t* cz, hipStream_t s);
// The lists, at the offsets given by the inclusive scans of the counts.
void launch_sub_cell_lists(const double* px, const double* py, const double* pz, uint32_t n, Grid g,
                           const uint32_t* keys, uint32_t level, double radius, const uint32_t* ix, const uint32_t* iy,
                           const uint32_t* iz, uint32_t* lx, uint32_t* ly, uint32_t* lz, hipStream_t s);
// Per node (512 voxels): distinct voxels and list entries of its particles.
void launch_particles_per_node(uint32_t B, uint32_t n, const uint32_t* first, const uint32_t* touchedIncl,
                               const uint32_t* cells, uint32_t* nonZero, uint32_t* entries, hipStream_t s);
// Used voxels, list starts and per-voxel particle lists of every node.
void launch_particle_lists(uint32_t B, uint32_t n, const uint32_t* usedIncl, const uint32_t* listIncl,
                           const uint32_t* touchedIncl, const uint32_t* cells, const uint32_t* first,
                           uint32_t* voxelIDs, uint32_t* listStart, uint32_t* lists, hipStream_t s);

}  // namespace rc
}  // namespace flip2
