// Host launchers for the reference-compatible FP64 stage kernels
// (rc_kernels.hip).  Every launcher enqueues on `stream` and checks the launch.
#pragma once

#include "common.hpp"

namespace flip2 {
namespace rc {

struct Ptr3 {
    const double* p[3];
};

void launch_root_cell(double* px, double* py, double* pz, uint32_t n, Grid g, uint32_t* keys, hipStream_t s);
void launch_gather3(const uint32_t* perm, uint32_t n, const double* ax, const double* ay, const double* az,
                    double* bx, double* by, double* bz, hipStream_t s);
void launch_map_nodes(const uint32_t* uniq, uint32_t n, uint32_t* first, hipStream_t s);
void launch_fill_u32(uint32_t* a, uint32_t count, uint32_t value, hipStream_t s);
void launch_row_table(uint32_t B, const uint32_t* first, const uint32_t* keys, uint32_t* rows, uint32_t sizeX,
                      hipStream_t s);
void launch_mark_voxels(uint32_t B, uint32_t n, Grid g, const uint32_t* keys, const uint32_t* first,
                        const double* px, const double* py, const double* pz, uint64_t* usedMask,
                        uint64_t* solidMask, uint32_t* counts, uint32_t* oobEvents, hipStream_t s);
void launch_compact_voxels(uint32_t B, const uint64_t* usedMask, const uint64_t* solidMask, const uint32_t* nodeOff,
                           uint32_t* voxelIDs, char* solids, hipStream_t s);
void launch_p2g(uint32_t B, uint32_t n, Grid g, const uint32_t* keys, const uint32_t* first, const double* px,
                const double* py, const double* pz, const double* vx, const double* vy, const double* vz,
                const uint32_t* nodeOff, const uint32_t* voxelIDs, const char* solids, double* Ux, double* Uy,
                double* Uz, hipStream_t s);
void launch_gravity(uint32_t V, const char* solids, double* Uy, double gdt, bool remove, hipStream_t s);
void launch_divu(uint32_t B, Grid g, const uint32_t* keys, const uint32_t* first, const uint32_t* rows,
                 const uint32_t* nodeOff, const uint32_t* voxelIDs, const double* Ux, const double* Uy,
                 const double* Uz, double* divU, hipStream_t s);
void launch_sor(uint32_t B, const uint32_t* nodeOff, const uint32_t* voxelIDs, const char* solids,
                const double* divU, double* p, double* residuals, double scale, double w, uint32_t iters,
                hipStream_t s);
// tmp must hold sampled_absmax_tmp_doubles(n, count) doubles.
size_t sampled_absmax_tmp_doubles(uint32_t n, int count);
void launch_sampled_absmax(Ptr3 arrays, int count, uint32_t n, double* tmp, double* result, hipStream_t s);
void launch_velocity_update(uint32_t B, const uint32_t* nodeOff, const uint32_t* voxelIDs, const char* solids,
                            const double* p, double* Ux, double* Uy, double* Uz, double scale, hipStream_t s);
void launch_g2p(uint32_t B, uint32_t n, Grid g, const uint32_t* keys, const uint32_t* first, const double* px,
                const double* py, const double* pz, double* vx, double* vy, double* vz, const uint32_t* nodeOff,
                const uint32_t* voxelIDs, const double* Ux, const double* Uy, const double* Uz, hipStream_t s);
void launch_advect(uint32_t n, double dt, double* px, double* py, double* pz, const double* vx, const double* vy,
                   const double* vz, hipStream_t s);
void launch_move_solids(uint32_t B, uint32_t n, Grid g, const uint32_t* keys, const uint32_t* first, double* px,
                        double* py, double* pz, const uint32_t* nodeOff, const uint32_t* voxelIDs,
                        const char* solids, uint64_t counter, uint32_t* relocated, hipStream_t s);
// in: count x (ctr0..3, key0, key1); out: count x 4 words.
void launch_philox_kat(const uint32_t* in, uint32_t count, uint32_t* out, hipStream_t s);

}  // namespace rc
}  // namespace flip2
