// Boundary faces, moving solid box and gravity on the MAC grid: the
// "prepare" stage of the FLIP step, shared by the stage kernel
// (prepare_faces_kernel) and the fused transfer + prepare kernels, so both
// compute every face by the same expressions.
#pragma once

#include "common.hpp"
#include "ns_solver.hpp"

namespace flip2 {
namespace ns {

__device__ inline uint32_t face_tiled(uint32_t i, uint32_t j, uint32_t k, uint32_t nbX, uint32_t nbY) {
    return (((k >> 3) * nbY + (j >> 3)) * nbX + (i >> 3)) * 512u + (k & 7u) * 64u + (j & 7u) * 8u + (i & 7u);
}

__device__ inline uint32_t clamp_cell(double q, uint32_t n) {
    if (!(q > 0.0)) return 0u;
    const double f = floor(q);
    return f >= (double)(n - 1) ? n - 1 : (uint32_t)f;
}

// Brick-tiled index of the cell holding an FP32 position (clamped into the
// grid): the particle's sort key.
__device__ inline uint32_t cell_key(float x, float y, float z, const Grid& g) {
    const double inv = 1.0 / (double)g.cellSize;
    const uint32_t i = clamp_cell(((double)x - (double)g.negX) * inv, g.sizeX);
    const uint32_t j = clamp_cell(((double)y - (double)g.negY) * inv, g.sizeY);
    const uint32_t k = clamp_cell(((double)z - (double)g.negZ) * inv, g.sizeZ);
    return face_tiled(i, j, k, g.sizeX >> 3, g.sizeY >> 3);
}

// Cells outside the domain count as solid.
__device__ inline bool cell_solid(int i, int j, int k, Grid g, const uint8_t* solid, uint32_t nbX, uint32_t nbY) {
    if (i < 0 || j < 0 || k < 0 || i >= (int)g.sizeX || j >= (int)g.sizeY || k >= (int)g.sizeZ) return true;
    return solid && solid[face_tiled((uint32_t)i, (uint32_t)j, (uint32_t)k, nbX, nbY)];
}

// The cell's centre lies in the (active) box.
__device__ inline bool in_box(int i, int j, int k, Grid g, const SolidBox& box) {
    if (!box.active) return false;
    // explicit fmas: the same rounding in every kernel that inlines this
    const float c[3] = {fmaf((float)i + 0.5f, g.cellSize, g.negX), fmaf((float)j + 0.5f, g.cellSize, g.negY),
                        fmaf((float)k + 0.5f, g.cellSize, g.negZ)};
    return c[0] >= box.lo[0] && c[0] < box.hi[0] && c[1] >= box.lo[1] && c[1] < box.hi[1] && c[2] >= box.lo[2] &&
           c[2] < box.hi[2];
}

__device__ inline bool in_domain(int i, int j, int k, Grid g) {
    return i >= 0 && j >= 0 && k >= 0 && i < (int)g.sizeX && j < (int)g.sizeY && k < (int)g.sizeZ;
}

// The three low faces of cell gc from their transferred values in[]: faces
// touching the domain boundary take velocity 0, faces touching a solid cell
// the solid's velocity (the box velocity for box cells, else 0); the others
// keep the transferred value, v faces get + g dt.  saved[]: the value before
// gravity (the pre-projection field of the FLIP update), out[]: after.
__device__ inline void prepare_cell_faces(const float in[3], const int gc[3], Grid g, const uint8_t* solid,
                                          const SolidBox& box, float gdt, uint32_t nbX, uint32_t nbY, float saved[3],
                                          float out[3]) {
    const bool here = cell_solid(gc[0], gc[1], gc[2], g, solid, nbX, nbY);
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        int m[3] = {gc[0], gc[1], gc[2]};
        m[axis] -= 1;
        const bool nbSolid = cell_solid(m[0], m[1], m[2], g, solid, nbX, nbY);
        const bool wall = here || nbSolid;
        float vs = 0.f;  // solid velocity of this face
        if (wall && in_domain(m[0], m[1], m[2], g) &&
            ((here && in_box(gc[0], gc[1], gc[2], g, box)) || (nbSolid && in_box(m[0], m[1], m[2], g, box))))
            vs = box.vel[axis];
        float v = wall ? vs : in[axis];
        saved[axis] = v;
        if (axis == 1 && !wall) v += gdt;
        out[axis] = v;
    }
}

}  // namespace ns
}  // namespace flip2
