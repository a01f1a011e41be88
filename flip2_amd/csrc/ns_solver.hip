// North-star pressure projection and grid->particle transfer on 8^3 bricks.
//
// Layout: dense brick-tiled face velocities U, V, W (n^3 each, u(i,j,k) on
// the low-x face of cell (i,j,k)); the Poisson unknowns live only in
// "active" bricks (bricks holding at least one fluid cell), 512 cells per
// slot, with a slot table, a 6-neighbour slot table and one cell-type byte
// per cell.  Cells of inactive bricks are air, or solid where the static
// solid mask says so; cells outside the domain are solid walls.
//
// System (rho = 1): for fluid cell c,
//   sum_{non-solid nb} (p_c - p_nb) = -(h/dt) * D_c,  p_air = 0,
//   D_c = u(i+1)-u(i) + v(j+1)-v(j) + w(k+1)-w(k)
// and the faces are corrected by u -= (dt/h)(p_c - p_{c-1}).
//
// Mixed precision: the operator, the preconditioners and the Krylov vectors
// of the inner solve are FP32; every dot product accumulates in FP64; the
// right-hand side, the solution and the true residual b - A x are FP64
// (iterative refinement).  Reductions are deterministic: every workgroup
// stores its partial; the two PCG kernels of an iteration sum the other's
// partials at their start, other kernels are followed by a one-workgroup
// summing kernel, always in workgroup order.
#include "ns_solver.hpp"

#include "face_prep.hpp"

namespace flip2 {
namespace ns {

namespace {

constexpr int kWG = 256;  // threads per workgroup of the per-slot kernels

__device__ inline uint32_t tiled(uint32_t i, uint32_t j, uint32_t k, uint32_t nbX, uint32_t nbY) {
    return (((k >> 3) * nbY + (j >> 3)) * nbX + (i >> 3)) * 512u + (k & 7u) * 64u + (j & 7u) * 8u + (i & 7u);
}

__device__ inline void brick_coords(uint32_t b, uint32_t nbX, uint32_t nbY, uint32_t& bi, uint32_t& bj, uint32_t& bk) {
    bi = b % nbX;
    bj = (b / nbX) % nbY;
    bk = b / (nbX * nbY);
}

// Wave-level then workgroup-level FP64 sum; result valid in thread 0.
__device__ inline double block_sum(double v, double* s_tmp) {
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) s_tmp[wave] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r += s_tmp[w];
    return r;
}

// Each workgroup stores its partial (of `count` quantities); a one-workgroup
// follow-up kernel sums them in workgroup order (deterministic).  No
// workgroup waits for the others: on a chip whose L2 caches are per XCD, the
// release fence a "last workgroup finishes" scheme needs in every workgroup
// writes back L2 and costs more than the kernel's own work.
__device__ inline void store_partials(const double* mine, int count, double* partials) {
    if (threadIdx.x == 0)
        for (int q = 0; q < count; ++q) partials[(size_t)q * gridDim.x + blockIdx.x] = mine[q];
}

// Orders the LDS accesses of the lanes of one wave (a one-wave workgroup's
// barrier): the wave's LDS operations execute in issue order, so the fences
// only keep the compiler from moving accesses across this point, and no
// instruction waits for a store to complete.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Slot of this workgroup in an XCD-aware order, for kernels that read
// neighbouring slots: the grid has 8 ceil(na / 8) workgroups, dealt to the
// 8 XCDs round robin, and XCD x takes the x-th contiguous eighth of the
// slots, so a slot's neighbours (nearby slot numbers: slots follow the
// brick order) are loaded into the same L2.  Padding workgroups get na.
__device__ inline uint32_t xcd_slot(uint32_t na) {
    const uint32_t per = gridDim.x >> 3;
    const uint32_t s = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    return s < na ? s : na;
}

}  // namespace

// ---------------------------------------------------------------------------
// Active bricks
// ---------------------------------------------------------------------------

// flag[b] = 1 if brick b has a cell with particles that is not solid.
__global__ void __launch_bounds__(64)
classify_bricks_kernel(const uint32_t* __restrict__ counts, const uint8_t* __restrict__ solid,
                       uint32_t* __restrict__ flag) {
    const uint32_t b = blockIdx.x;
    bool any = false;
    for (uint32_t t = threadIdx.x; t < 512u; t += 64u) {
        const uint32_t c = b * 512u + t;
        any |= (counts[c] > 0u) && !(solid && solid[c]);
    }
    const uint64_t m = __ballot(any);
    if (threadIdx.x == 0) flag[b] = m ? 1u : 0u;
}

// The same from the cell offsets of a fused transfer, which exist for the
// occupied bricks only: flag (cleared beforehand) is set for the occupied
// bricks of the touched list that have such a cell.
__global__ void __launch_bounds__(64)
classify_bricks_sparse_kernel(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                              const uint8_t* __restrict__ occ, const uint32_t* __restrict__ cellStart,
                              const uint8_t* __restrict__ solid, uint32_t* __restrict__ flag) {
    const uint32_t total = *count;
    for (uint32_t e = blockIdx.x; e < total; e += gridDim.x) {
        const uint32_t b = list[e];
        if (!occ[b]) continue;
        bool any = false;
        for (uint32_t t = threadIdx.x; t < 512u; t += 64u) {
            const uint32_t c = b * 512u + t;
            any |= (cellStart[c + 1u] > cellStart[c]) && !(solid && solid[c]);
        }
        const uint64_t m = __ballot(any);
        if (threadIdx.x == 0) flag[b] = m ? 1u : 0u;
    }
}

__global__ void build_slots_kernel(const uint32_t* __restrict__ flag, const uint32_t* __restrict__ excl,
                                   uint32_t numBricks, int32_t* __restrict__ brickSlot,
                                   uint32_t* __restrict__ activeBricks) {
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= numBricks) return;
    if (flag[b]) {
        brickSlot[b] = (int32_t)excl[b];
        activeBricks[excl[b]] = b;
    } else {
        brickSlot[b] = -1;
    }
}

// Neighbour slots and cell types of each active brick.
__global__ void __launch_bounds__(kWG)
brick_setup_kernel(const uint32_t* __restrict__ activeBricks, const int32_t* __restrict__ brickSlot,
                   const uint32_t* __restrict__ counts, const uint32_t* __restrict__ cellStart,
                   const uint8_t* __restrict__ solid, Grid g, int32_t* __restrict__ nbrSlot,
                   uint8_t* __restrict__ cellType) {
    const uint32_t slot = blockIdx.x;
    const uint32_t b = activeBricks[slot];
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3, nbZ = g.sizeZ >> 3;
    uint32_t bi, bj, bk;
    brick_coords(b, nbX, nbY, bi, bj, bk);
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        int ni = (int)bi + (d == 0 ? -1 : d == 1 ? 1 : 0);
        int nj = (int)bj + (d == 2 ? -1 : d == 3 ? 1 : 0);
        int nk = (int)bk + (d == 4 ? -1 : d == 5 ? 1 : 0);
        int32_t s = -1;
        if (ni >= 0 && nj >= 0 && nk >= 0 && ni < (int)nbX && nj < (int)nbY && nk < (int)nbZ)
            s = brickSlot[((uint32_t)nk * nbY + (uint32_t)nj) * nbX + (uint32_t)ni];
        nbrSlot[slot * 6u + d] = s;
    }
    for (uint32_t t = threadIdx.x; t < 512u; t += blockDim.x) {
        const uint32_t c = b * 512u + t;
        uint8_t ty = kAir;
        if (solid && solid[c]) ty = kSolid;
        else if (counts ? counts[c] > 0u : cellStart[c + 1u] > cellStart[c]) ty = kFluid;
        cellType[slot * 512u + t] = ty;
    }
}

namespace {

// Type of the neighbour of local cell l (brick coords bc, global cell gc) in
// direction d, looking through the neighbour slot table.
struct NbrRef {
    int32_t slot;   // slot holding the neighbour (-1: inactive brick or outside)
    uint32_t local; // its local index
    uint8_t type;
};

__device__ inline NbrRef neighbour(uint32_t slot, uint32_t l, int d, const int gcoord[3], Grid g,
                                   const int32_t* nbrSlot, const uint8_t* cellType, const uint8_t* solid,
                                   uint32_t nbX, uint32_t nbY) {
    const int axis = d >> 1;
    const int dir = (d & 1) ? 1 : -1;
    int c[3] = {gcoord[0], gcoord[1], gcoord[2]};
    c[axis] += dir;
    NbrRef r;
    const uint32_t n[3] = {g.sizeX, g.sizeY, g.sizeZ};
    if (c[axis] < 0 || c[axis] >= (int)n[axis]) {
        r.slot = -1; r.local = 0; r.type = kSolid;
        return r;
    }
    const uint32_t li = (uint32_t)c[0] & 7u, lj = (uint32_t)c[1] & 7u, lk = (uint32_t)c[2] & 7u;
    const uint32_t nl = lk * 64u + lj * 8u + li;
    const int lc = axis == 0 ? (int)(l & 7u) : axis == 1 ? (int)((l >> 3) & 7u) : (int)(l >> 6);
    const bool inside = (dir < 0) ? lc > 0 : lc < 7;
    int32_t s = inside ? (int32_t)slot : nbrSlot[slot * 6u + d];
    r.local = nl;
    r.slot = s;
    if (s >= 0) {
        r.type = cellType[(uint32_t)s * 512u + nl];
    } else {
        const uint32_t gi = tiled((uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2], nbX, nbY);
        r.type = (solid && solid[gi]) ? kSolid : kAir;
    }
    return r;
}

__device__ inline void cell_coords(uint32_t b, uint32_t l, uint32_t nbX, uint32_t nbY, int gc[3]) {
    uint32_t bi, bj, bk;
    brick_coords(b, nbX, nbY, bi, bj, bk);
    gc[0] = (int)(bi * 8u + (l & 7u));
    gc[1] = (int)(bj * 8u + ((l >> 3) & 7u));
    gc[2] = (int)(bk * 8u + (l >> 6));
}

}  // namespace

// Operator coefficients (per fluid cell: number of non-solid neighbours and a
// mask of fluid neighbours) and the FP64 right-hand side -(h/dt) D_c;
// ||b||^2 reduced into scal[0], the number of fluid cells into scal[1].
__global__ void __launch_bounds__(kWG)
assemble_kernel(const uint32_t* __restrict__ activeBricks, const int32_t* __restrict__ nbrSlot,
                const uint8_t* __restrict__ cellType, const uint8_t* __restrict__ solid, Grid g, double hOverDt,
                const float* __restrict__ U, const float* __restrict__ V, const float* __restrict__ W,
                uint16_t* __restrict__ coef, double* __restrict__ b64, double* __restrict__ partials) {
    __shared__ double s_tmp[16];
    const uint32_t slot = blockIdx.x;
    const uint32_t b = activeBricks[slot];
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    double bb = 0.0, fluid = 0.0;
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) {
        const uint32_t v = slot * 512u + l;
        uint16_t cf = 0;
        double rhs = 0.0;
        if (cellType[v] == kFluid) {
            fluid += 1.0;
            int gc[3];
            cell_coords(b, l, nbX, nbY, gc);
            uint32_t diag = 0, mask = 0;
            for (int d = 0; d < 6; ++d) {
                const NbrRef r = neighbour(slot, l, d, gc, g, nbrSlot, cellType, solid, nbX, nbY);
                const uint8_t ty = r.type;
                if (ty != kSolid) ++diag;
                if (ty == kFluid) mask |= 1u << d;
            }
            cf = (uint16_t)(diag | (mask << 3));
            const uint32_t c0 = b * 512u + l;
            const uint32_t gx = tiled(gc[0] + 1 < (int)g.sizeX ? gc[0] + 1 : 0, gc[1], gc[2], nbX, nbY);
            const uint32_t gy = tiled(gc[0], gc[1] + 1 < (int)g.sizeY ? gc[1] + 1 : 0, gc[2], nbX, nbY);
            const uint32_t gz = tiled(gc[0], gc[1], gc[2] + 1 < (int)g.sizeZ ? gc[2] + 1 : 0, nbX, nbY);
            const double uh = (gc[0] + 1 < (int)g.sizeX) ? (double)U[gx] : 0.0;
            const double vh = (gc[1] + 1 < (int)g.sizeY) ? (double)V[gy] : 0.0;
            const double wh = (gc[2] + 1 < (int)g.sizeZ) ? (double)W[gz] : 0.0;
            const double D = (uh - (double)U[c0]) + (vh - (double)V[c0]) + (wh - (double)W[c0]);
            rhs = -hOverDt * D;
        }
        coef[v] = cf;
        b64[v] = rhs;
        bb += rhs * rhs;
    }
    const double t[2] = {block_sum(bb, s_tmp), block_sum(fluid, s_tmp)};
    store_partials(t, 2, partials);
}

namespace {

// y = A x on FP32 data (x of non-fluid cells is 0).
__device__ inline float apply_A(uint32_t slot, uint32_t l, uint16_t cf, const float* x, const int32_t* nbrSlot) {
    if (!cf) return 0.f;
    const uint32_t diag = cf & 7u, mask = cf >> 3;
    float s = (float)diag * x[slot * 512u + l];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (!((mask >> d) & 1u)) continue;
        const int axis = d >> 1;
        const bool up = d & 1;
        const uint32_t lc = axis == 0 ? (l & 7u) : axis == 1 ? ((l >> 3) & 7u) : (l >> 6);
        const uint32_t stride = axis == 0 ? 1u : axis == 1 ? 8u : 64u;
        uint32_t ns = slot, nl;
        if (up) {
            if (lc < 7u) nl = l + stride;
            else { ns = (uint32_t)nbrSlot[slot * 6u + d]; nl = l - 7u * stride; }
        } else {
            if (lc > 0u) nl = l - stride;
            else { ns = (uint32_t)nbrSlot[slot * 6u + d]; nl = l + 7u * stride; }
        }
        s -= x[ns * 512u + nl];
    }
    return s;
}

// y = A x with x given by a functor of the linear (slot * 512 + local) index.
template <class X>
__device__ inline float apply_A_with(uint32_t slot, uint32_t l, uint16_t cf, const int32_t* nbrSlot, X x) {
    if (!cf) return 0.f;
    const uint32_t diag = cf & 7u, mask = cf >> 3;
    float s = (float)diag * x(slot * 512u + l);
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (!((mask >> d) & 1u)) continue;
        const int axis = d >> 1;
        const bool up = d & 1;
        const uint32_t lc = axis == 0 ? (l & 7u) : axis == 1 ? ((l >> 3) & 7u) : (l >> 6);
        const uint32_t stride = axis == 0 ? 1u : axis == 1 ? 8u : 64u;
        uint32_t ns = slot, nl;
        if (up) {
            if (lc < 7u) nl = l + stride;
            else { ns = (uint32_t)nbrSlot[slot * 6u + d]; nl = l - 7u * stride; }
        } else {
            if (lc > 0u) nl = l - stride;
            else { ns = (uint32_t)nbrSlot[slot * 6u + d]; nl = l + 7u * stride; }
        }
        s -= x(ns * 512u + nl);
    }
    return s;
}

__device__ inline double apply_A64(uint32_t slot, uint32_t l, uint16_t cf, const double* x, const int32_t* nbrSlot) {
    if (!cf) return 0.0;
    const uint32_t diag = cf & 7u, mask = cf >> 3;
    double s = (double)diag * x[slot * 512u + l];
#pragma unroll
    for (int d = 0; d < 6; ++d) {
        if (!((mask >> d) & 1u)) continue;
        const int axis = d >> 1;
        const bool up = d & 1;
        const uint32_t lc = axis == 0 ? (l & 7u) : axis == 1 ? ((l >> 3) & 7u) : (l >> 6);
        const uint32_t stride = axis == 0 ? 1u : axis == 1 ? 8u : 64u;
        uint32_t ns = slot, nl;
        if (up) {
            if (lc < 7u) nl = l + stride;
            else { ns = (uint32_t)nbrSlot[slot * 6u + d]; nl = l - 7u * stride; }
        } else {
            if (lc > 0u) nl = l - stride;
            else { ns = (uint32_t)nbrSlot[slot * 6u + d]; nl = l + 7u * stride; }
        }
        s -= x[ns * 512u + nl];
    }
    return s;
}

}  // namespace

// out[q] = sum over workgroups of partials[q * blocks + b], in a fixed order:
// workgroup q sums quantity q; thread t adds the partials t + T j (T threads)
// into eight accumulators (j mod 8), whose loads are all issued before the
// adds.
constexpr int kReduceThreads = 1024;
__global__ void __launch_bounds__(kReduceThreads)
reduce_partials_kernel(const double* __restrict__ partials, uint32_t blocks, double* __restrict__ out) {
    __shared__ double s_tmp[kReduceThreads / 64];
    const double* p = partials + (size_t)blockIdx.x * blocks;
    double acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0.0;
    for (uint32_t b0 = threadIdx.x; b0 < blocks; b0 += 8u * blockDim.x) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t b = b0 + (uint32_t)k * blockDim.x;
            v[k] = b < blocks ? p[b] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] += v[k];
    }
    const double a = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    const double t = block_sum(a, s_tmp);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kWG)
refine_start_kernel(const double* __restrict__ r64, float* __restrict__ r, float* __restrict__ e,
                    const uint16_t* __restrict__ coef, double* __restrict__ partials, uint32_t* resetStop) {
    __shared__ double s_tmp[16];
    if (resetStop && blockIdx.x == 0 && threadIdx.x == 0) *resetStop = 0u;
    const uint32_t base = blockIdx.x * 512u;
    double rr = 0.0;
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) {
        const float v = coef[base + l] ? (float)r64[base + l] : 0.f;
        r[base + l] = v;
        e[base + l] = 0.f;
        rr += (double)v * (double)v;
    }
    const double t = block_sum(rr, s_tmp);
    store_partials(&t, 1, partials);
}

// q = A p; pq into out[0].
__global__ void __launch_bounds__(kWG)
matvec_dot_kernel(const float* __restrict__ p, const uint16_t* __restrict__ coef, const int32_t* __restrict__ nbrSlot,
                  float* __restrict__ q, double* __restrict__ partials, uint32_t na, const uint32_t* stop) {
    __shared__ double s_tmp[16];
    if (stop && *stop) return;
    const uint32_t slot = xcd_slot(na);
    if (slot >= na) return;
    double pq = 0.0;
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) {
        const uint32_t v = slot * 512u + l;
        const float qv = apply_A(slot, l, coef[v], p, nbrSlot);
        q[v] = qv;
        pq += (double)p[v] * (double)qv;
    }
    const double t = block_sum(pq, s_tmp);
    if (threadIdx.x == 0) partials[slot] = t;
}

// alpha = rz / pq; e += alpha p; r -= alpha q; with Jacobi also z = r/diag
// and rz_new; rr_new always.  scal layout per iteration: see ns_solver.hpp.
__global__ void __launch_bounds__(kWG)
update_kernel(const double* __restrict__ rzOld, const double* __restrict__ pq, const float* __restrict__ p,
              const float* __restrict__ q, const uint16_t* __restrict__ coef, float* __restrict__ e,
              float* __restrict__ r, float* __restrict__ z, int jacobi, double* __restrict__ partials /* [rz_new, rr_new] per workgroup */,
              const uint32_t* stop) {
    __shared__ double s_tmp[16];
    if (stop && *stop) return;
    const double den = *pq;
    const float alpha = (den != 0.0) ? (float)(*rzOld / den) : 0.f;
    const uint32_t base = blockIdx.x * 512u;
    double rz = 0.0, rr = 0.0;
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) {
        const uint32_t v = base + l;
        const uint16_t cf = coef[v];
        if (!cf) continue;
        e[v] = fmaf(alpha, p[v], e[v]);
        const float rv = fmaf(-alpha, q[v], r[v]);
        r[v] = rv;
        rr += (double)rv * (double)rv;
        if (jacobi) {
            const float zv = rv / (float)(cf & 7u);
            z[v] = zv;
            rz += (double)rv * (double)zv;
        }
    }
    double vals[2];
    vals[0] = block_sum(rz, s_tmp);
    vals[1] = block_sum(rr, s_tmp);
    store_partials(vals, 2, partials);
}

// Damped-Jacobi smoothing around an external (learned) correction:
//   presmooth:  r1 = r - A (w D^-1 r)
//   postsmooth: z  = x + w D^-1 (r - A x),  x = w D^-1 r + y
// The operand x of the slot's own 512 cells is evaluated once into LDS;
// neighbours in other slots (the brick's faces) are evaluated from memory.
// pairs != null: workgroup b also stores (sum of r1^2, count) over the cells
// with cellType 1 in pairs[2 b], pairs[2 b + 1].
__global__ void __launch_bounds__(kWG)
presmooth_kernel(const float* __restrict__ r, const uint16_t* __restrict__ coef, const int32_t* __restrict__ nbrSlot,
                 const uint8_t* __restrict__ cellType, float omega, float* __restrict__ r1,
                 double* __restrict__ pairs, const uint32_t* stop) {
    __shared__ float sx[512];
    __shared__ double s_tmp[16];
    if (stop && *stop) return;
    const uint32_t slot = blockIdx.x;
    auto jac = [&](uint32_t i) {
        const uint16_t c = coef[i];
        return c ? omega * r[i] / (float)(c & 7u) : 0.f;
    };
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) sx[l] = jac(slot * 512u + l);
    __syncthreads();
    auto xs = [&](uint32_t i) { return (i >> 9) == slot ? sx[i & 511u] : jac(i); };
    double ss = 0.0, cnt = 0.0;
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) {
        const uint32_t v = slot * 512u + l;
        const uint16_t cf = coef[v];
        const float rv = cf ? r[v] - apply_A_with(slot, l, cf, nbrSlot, xs) : 0.f;
        r1[v] = rv;
        if (pairs && cellType[v] == 1) {
            ss += (double)rv * (double)rv;
            cnt += 1.0;
        }
    }
    if (pairs) {
        const double a = block_sum(ss, s_tmp);
        const double c = block_sum(cnt, s_tmp);
        if (threadIdx.x == 0) {
            pairs[2 * blockIdx.x] = a;
            pairs[2 * blockIdx.x + 1] = c;
        }
    }
}

__global__ void __launch_bounds__(kWG)
postsmooth_kernel(const float* __restrict__ r, const float* __restrict__ y, const uint16_t* __restrict__ coef,
                  const int32_t* __restrict__ nbrSlot, float omega, float* __restrict__ z, const uint32_t* stop) {
    __shared__ float sx[512];
    if (stop && *stop) return;
    const uint32_t slot = blockIdx.x;
    auto xf = [&](uint32_t i) {
        const uint16_t c = coef[i];
        return c ? omega * r[i] / (float)(c & 7u) + y[i] : 0.f;
    };
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) sx[l] = xf(slot * 512u + l);
    __syncthreads();
    auto xs = [&](uint32_t i) { return (i >> 9) == slot ? sx[i & 511u] : xf(i); };
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) {
        const uint32_t v = slot * 512u + l;
        const uint16_t cf = coef[v];
        if (!cf) {
            z[v] = 0.f;
            continue;
        }
        const float xv = sx[l];
        z[v] = xv + omega * (r[v] - apply_A_with(slot, l, cf, nbrSlot, xs)) / (float)(cf & 7u);
    }
}

template <int N>  // kept directions (d.count == N)
__global__ void __launch_bounds__(kWG)
flex_dots_kernel(const float* __restrict__ z, FlexDirections d, double* __restrict__ partials,
                 const uint32_t* stop) {
    if (stop && *stop) return;
    const uint32_t base = blockIdx.x * 512u;
    double acc[N];
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = 0.0;
#pragma unroll
    for (int k = 0; k < 512 / kWG; ++k) {
        const uint32_t v = base + threadIdx.x + k * kWG;
        float q[N];
#pragma unroll
        for (int i = 0; i < N; ++i) q[i] = d.q[i][v];  // all loads issued before any is used
        const double zv = (double)z[v];
#pragma unroll
        for (int i = 0; i < N; ++i) acc[i] += zv * (double)q[i];
    }
    // all N sums with one LDS exchange: wave sums by shuffles, then thread i
    // adds the four waves' sums of quantity i
    __shared__ double s_many[kWG / 64][N];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double v = acc[i];
        for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) s_many[wave][i] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < N) {
        double t = 0.0;
        for (int w = 0; w < kWG / 64; ++w) t += s_many[w][threadIdx.x];
        partials[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

template <int N>  // kept directions (d.count == N, 0 for p = z)
__global__ void __launch_bounds__(kWG)
flex_direction_kernel(const float* __restrict__ z, FlexDirections d, const double* __restrict__ zq,
                      const float* __restrict__ r, float* __restrict__ p, double* __restrict__ partials,
                      const uint32_t* stop) {
    __shared__ double s_tmp[16];
    if (stop && *stop) return;
    float beta[N > 0 ? N : 1];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double den = *d.pq[i];
        beta[i] = den != 0.0 ? (float)(-zq[i] / den) : 0.f;
    }
    const uint32_t base = blockIdx.x * 512u;
    double pr = 0.0;
#pragma unroll
    for (int k = 0; k < 512 / kWG; ++k) {
        const uint32_t v = base + threadIdx.x + k * kWG;
        float pd[N > 0 ? N : 1];
#pragma unroll
        for (int i = 0; i < N; ++i) pd[i] = d.p[i][v];
        float pv = z[v];
#pragma unroll
        for (int i = 0; i < N; ++i) pv = fmaf(beta[i], pd[i], pv);
        p[v] = pv;
        pr += (double)pv * (double)r[v];
    }
    const double t = block_sum(pr, s_tmp);
    store_partials(&t, 1, partials);
}

// x64 += e; r64 = b - A x64 (FP64); ||r64||^2 into out[0].
__global__ void __launch_bounds__(kWG)
accumulate_kernel(const float* __restrict__ e, double* __restrict__ x64, const uint16_t* __restrict__ coef) {
    const uint32_t base = blockIdx.x * 512u;
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x)
        if (coef[base + l]) x64[base + l] += (double)e[base + l];
}

__global__ void __launch_bounds__(kWG)
true_residual_kernel(const double* __restrict__ b64, const double* __restrict__ x64,
                     const uint16_t* __restrict__ coef, const int32_t* __restrict__ nbrSlot,
                     double* __restrict__ r64, double* __restrict__ partials, uint32_t na) {
    __shared__ double s_tmp[16];
    const uint32_t slot = xcd_slot(na);
    if (slot >= na) return;
    double rr = 0.0;
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) {
        const uint32_t v = slot * 512u + l;
        const uint16_t cf = coef[v];
        double rv = 0.0;
        if (cf) rv = b64[v] - apply_A64(slot, l, cf, x64, nbrSlot);
        r64[v] = rv;
        rr += rv * rv;
    }
    const double t = block_sum(rr, s_tmp);
    if (threadIdx.x == 0) partials[slot] = t;
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ a, uint32_t count, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = v;
}
__global__ void fill_f32_kernel(float* __restrict__ a, uint32_t count, float v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = v;
}
__global__ void fill_f64_kernel(double* __restrict__ a, uint32_t count, double v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = v;
}

// ---------------------------------------------------------------------------
// Brick-local MIC(0): one wave per brick, the 512 cells swept in 22
// anti-diagonal planes (i+j+k = s), whose cells are independent.
// ---------------------------------------------------------------------------
struct PlaneTable {
    uint16_t cells[512];
    uint16_t start[23];
    // per plane and lane: the lane's cell c = cells[start[s] + lane] as c + 1
    // (its entry in the MIC sweeps' padded LDS array, less 64), or 0 for a
    // lane past the plane's end (the pad entry 64)
    uint16_t lane_cell[22][64];
};

constexpr PlaneTable make_planes() {
    PlaneTable t{};
    int n = 0;
    for (int s = 0; s <= 21; ++s) {
        t.start[s] = (uint16_t)n;
        for (int k = 0; k < 8; ++k)
            for (int j = 0; j < 8; ++j) {
                const int i = s - j - k;
                if (i >= 0 && i < 8) t.cells[n++] = (uint16_t)(k * 64 + j * 8 + i);
            }
    }
    t.start[22] = (uint16_t)n;
    for (int s = 0; s <= 21; ++s)
        for (int l = 0; l < 64; ++l)
            t.lane_cell[s][l] = t.start[s] + l < t.start[s + 1] ? (uint16_t)(t.cells[t.start[s] + l] + 1) : (uint16_t)0;
    return t;
}

__constant__ PlaneTable c_planes = make_planes();

namespace {
constexpr float kMicTau = 0.97f;
constexpr float kMicSigma = 0.25f;

// A(c, c+e_axis) inside the brick: -1 if both fluid (mask bit), else 0.
__device__ inline float a_plus(uint16_t cf, int axis, uint32_t l) {
    const uint32_t lc = axis == 0 ? (l & 7u) : axis == 1 ? ((l >> 3) & 7u) : (l >> 6);
    if (lc == 7u) return 0.f;
    return ((cf >> 3) >> (2 * axis + 1)) & 1u ? -1.f : 0.f;
}
}  // namespace

// MIC(0) factor of each brick's diagonal block (Bridson's modified
// incomplete Cholesky, tau = 0.97 for the PCG, safety 0.25): precon =
// 1/sqrt(e).  tau = 0 gives the unmodified IC(0) factor.
__global__ void __launch_bounds__(64)
mic_factor_kernel(const uint16_t* __restrict__ coef, float* __restrict__ precon, float tau) {
    __shared__ float s_pc[512];
    __shared__ uint16_t s_cf[512];
    const uint32_t base = blockIdx.x * 512u;
    for (uint32_t l = threadIdx.x; l < 512u; l += 64u) { s_cf[l] = coef[base + l]; s_pc[l] = 0.f; }
    __syncthreads();
    for (int s = 0; s <= 21; ++s) {
        const int beg = c_planes.start[s], end = c_planes.start[s + 1];
        for (int t = beg + (int)threadIdx.x; t < end; t += 64) {
            const uint32_t l = c_planes.cells[t];
            const uint16_t cf = s_cf[l];
            float pc = 0.f;
            if (cf) {
                const float diag = (float)(cf & 7u);
                float e = diag;
                const uint32_t lc[3] = {l & 7u, (l >> 3) & 7u, l >> 6};
                const uint32_t stride[3] = {1u, 8u, 64u};
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    if (lc[ax] == 0u) continue;
                    const uint32_t m = l - stride[ax];
                    const uint16_t cm = s_cf[m];
                    const float a = a_plus(cm, ax, m);
                    const float pm = s_pc[m];
                    e -= (a * pm) * (a * pm);
                    const float o1 = a_plus(cm, (ax + 1) % 3, m), o2 = a_plus(cm, (ax + 2) % 3, m);
                    e -= tau * (a * (o1 + o2) * pm * pm);
                }
                if (e < kMicSigma * diag) e = diag;
                pc = 1.0f / sqrtf(e);
            }
            s_pc[l] = pc;
        }
        __syncthreads();
    }
    __syncthreads();
    for (uint32_t l = threadIdx.x; l < 512u; l += 64u) precon[base + l] = s_pc[l];
}

// ---------------------------------------------------------------------------
// PCG with two kernels per iteration (Jacobi and brick-local MIC(0)).
//
//   direction+matvec (iteration k):  rz_k, rr_k = sums of the update's
//                                    partials;  beta = rz_k / rz_{k-1};
//                                    p = z + beta p';  q = A p;  partial p.q
//   update+precondition:             (p.q)_k = sum of the matvec's partials;
//                                    alpha = rz_k / (p.q)_k;  e += alpha p;
//                                    r -= alpha q;  z = M^-1 r;  partials r.z, r.r
//
// A workgroup is four waves, one slot each (the MIC(0) sweeps run in the
// wave's own LDS), so every slot's memory latency is in flight at once.
// Each kernel starts by summing the previous kernel's per-workgroup partials
// itself (group_sum: the same fixed order in every workgroup, so every
// workgroup -- and the end-of-batch history kernel -- gets the same double);
// workgroup 0 records the sums in the iteration history (rz, rr, pq per
// iteration) for the next iterations and the host's convergence checks.  No
// reduction kernel runs between the two.  p alternates between two buffers:
// the matvec reads the previous direction at neighbour cells owned by other
// workgroups while writing the new one.
// ---------------------------------------------------------------------------
namespace {

// Slots (waves) per workgroup: 4 below kCgGroupedBelow slots, where the
// kernels sum the partials themselves and launches dominate; 1 above, with a
// one-workgroup reduction kernel after each main kernel (more workgroups in
// flight, no partial reads at each workgroup's start).
constexpr uint32_t kCgGroupedBelow = 4096;

// Workgroup index in the XCD-aware order over ng groups (see xcd_slot).
__device__ inline uint32_t cg_group(uint32_t ng) { return xcd_slot(ng); }

// Fixed-order sums of a[0..n) and b[0..n) (b may be null) by the 64 W
// threads of a workgroup: thread t adds a[t + 64 W j] in j order (loads
// issued four at a time before the adds), then the waves' butterfly sums,
// then the W wave sums in order (s_red: 2 W doubles).  Every workgroup
// computes the same doubles.
template <uint32_t W>
__device__ inline void group_sum(const double* __restrict__ a, const double* __restrict__ b, uint32_t n,
                                 double* s_red, double& sa, double& sb) {
    constexpr uint32_t T = 64u * W;
    const uint32_t tid = threadIdx.x;
    double x = 0.0, y = 0.0;
    for (uint32_t i0 = tid; i0 < n; i0 += 4u * T) {
        double va[4], vb[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t idx = i0 + T * (uint32_t)k;
            va[k] = idx < n ? a[idx] : 0.0;
            vb[k] = (b && idx < n) ? b[idx] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            x += va[k];
            y += vb[k];
        }
    }
    for (int d = 32; d >= 1; d >>= 1) {
        x += __shfl_xor(x, d, 64);
        y += __shfl_xor(y, d, 64);
    }
    if ((tid & 63u) == 0) {
        s_red[tid >> 6] = x;
        s_red[W + (tid >> 6)] = y;
    }
    __syncthreads();
    sa = s_red[0];
    sb = s_red[W];
#pragma unroll
    for (uint32_t w = 1; w < W; ++w) {
        sa += s_red[w];
        sb += s_red[W + w];
    }
    __syncthreads();
}

// The workgroup's partial: its W waves' values in wave order (s_red: W
// doubles).
template <uint32_t W>
__device__ inline void group_partial(double mine, double* s_red, double* out) {
    if constexpr (W == 1) {
        if (threadIdx.x == 0) *out = mine;
    } else {
        if ((threadIdx.x & 63u) == 0) s_red[threadIdx.x >> 6] = mine;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = s_red[0];
#pragma unroll
            for (uint32_t w = 1; w < W; ++w) t += s_red[w];
            *out = t;
        }
    }
}

}  // namespace

// Iteration `it`: rz_it, rr_it from the update's partials (rz at
// partials[na_g..], rr at partials[2 na_g..], na_g = number of groups);
// p = z + beta p_old (p = z when first), q = A p; partial p.q per group at
// partials[0..).  Wave = slot, lane = column (i, j) of the brick holding the
// eight cells k = 0..7 (cell l = 64 k + lane, so every load is one coalesced
// row per k).  All inputs -- the column's coefficients, z and p_old, and the
// same for the cells of the neighbouring slots it touches -- are loaded
// before the partial sums are read; the new p goes to LDS for the x and y
// neighbours, z neighbours are the lane's own registers.
template <uint32_t W>
__global__ void __launch_bounds__(64 * W)
cg_direction_matvec_kernel(CgFused c, uint32_t it, int first, const float* __restrict__ pOld,
                           float* __restrict__ pNew) {
    __shared__ float s_pw[W][512];
    __shared__ double s_red[2 * W];
    const uint32_t ng = (c.na + W - 1u) / W;
    const uint32_t grp = cg_group(ng);
    if (grp >= ng) return;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t slotRaw = grp * W + wave;
    const bool live = slotRaw < c.na;
    const uint32_t slot = live ? slotRaw : c.na - 1u;  // a padding wave mirrors the last slot and stores nothing
    const uint32_t base = slot * 512u;
    if (*c.stop) return;  // converged earlier in this refinement: no loads either
    float* s_p = s_pw[wave];
    const uint32_t i = lane & 7u, j = lane >> 3;
    int32_t nb[6];
#pragma unroll
    for (int d = 0; d < 6; ++d) nb[d] = c.nbrSlot[slot * 6u + d];
    // neighbour slot and its cell index for the column's x and y sides
    const int32_t xs = i == 0u ? nb[0] : (i == 7u ? nb[1] : -1);
    const int32_t ys = j == 0u ? nb[2] : (j == 7u ? nb[3] : -1);
    const uint32_t xl = (i == 0u ? 7u : 0u) + 8u * j, yl = i + (j == 0u ? 56u : 0u);
    uint16_t cf[8];
    float zv[8], po[8], xz[8], xp[8], yz[8], yp[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t v = base + 64u * (uint32_t)k + lane;
        cf[k] = c.coef[v];
        zv[k] = c.z[v];
        po[k] = first ? 0.f : pOld[v];
        xz[k] = xp[k] = yz[k] = yp[k] = 0.f;
        if (xs >= 0) {
            const uint32_t w = (uint32_t)xs * 512u + 64u * (uint32_t)k + xl;
            xz[k] = c.z[w];
            xp[k] = first ? 0.f : pOld[w];
        }
        if (ys >= 0) {
            const uint32_t w = (uint32_t)ys * 512u + 64u * (uint32_t)k + yl;
            yz[k] = c.z[w];
            yp[k] = first ? 0.f : pOld[w];
        }
    }
    float zlo = 0.f, plo = 0.f, zhi = 0.f, phi = 0.f;  // the z-side neighbours of cells k = 0 and 7
    if (nb[4] >= 0) {
        const uint32_t w = (uint32_t)nb[4] * 512u + 448u + lane;
        zlo = c.z[w];
        plo = first ? 0.f : pOld[w];
    }
    if (nb[5] >= 0) {
        const uint32_t w = (uint32_t)nb[5] * 512u + lane;
        zhi = c.z[w];
        phi = first ? 0.f : pOld[w];
    }
    double rz;
    if constexpr (W > 1) {
        double rr;
        group_sum<W>(c.partials + ng, c.partials + 2u * ng, ng, s_red, rz, rr);
        if (grp == 0 && threadIdx.x == 0) {
            c.hist[4u * it] = rz;
            c.hist[4u * it + 1u] = rr;
        }
        if (rr <= c.thr2 || !(rr == rr)) {  // the same decision in every workgroup
            if (grp == 0 && threadIdx.x == 0) *c.stop = 1u;
            return;
        }
    } else {
        rz = c.hist[4u * it];  // recorded by the reduction after the update
    }
    float beta = 0.f;
    if (!first) {
        const double rzOld = c.hist[4u * (it - 1u)];
        beta = rzOld != 0.0 ? (float)(rz / rzOld) : 0.f;
    }
    auto dir = [&](float zz, float pp) { return first ? zz : fmaf(beta, pp, zz); };
    float pv[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        pv[k] = cf[k] ? dir(zv[k], po[k]) : 0.f;
        s_p[64u * (uint32_t)k + lane] = pv[k];
    }
    wave_lds_sync();
    double pq = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t l = 64u * (uint32_t)k + lane, v = base + l;
        float qv = 0.f;
        if (cf[k]) {
            const uint32_t mask = cf[k] >> 3;
            float sum = (float)(cf[k] & 7u) * pv[k];
            // neighbours in the order of the mask bits: -x, +x, -y, +y, -z, +z
            if (mask & 1u) sum -= i > 0u ? s_p[l - 1u] : dir(xz[k], xp[k]);
            if (mask & 2u) sum -= i < 7u ? s_p[l + 1u] : dir(xz[k], xp[k]);
            if (mask & 4u) sum -= j > 0u ? s_p[l - 8u] : dir(yz[k], yp[k]);
            if (mask & 8u) sum -= j < 7u ? s_p[l + 8u] : dir(yz[k], yp[k]);
            if (mask & 16u) sum -= k > 0 ? pv[k > 0 ? k - 1 : 0] : dir(zlo, plo);
            if (mask & 32u) sum -= k < 7 ? pv[k < 7 ? k + 1 : 7] : dir(zhi, phi);
            qv = sum;
        }
        if (live) {
            pNew[v] = pv[k];
            c.q[v] = qv;
        }
        pq += (double)pv[k] * (double)qv;
    }
    for (int d = 32; d >= 1; d >>= 1) pq += __shfl_xor(pq, d, 64);
    group_partial<W>(live ? pq : 0.0, s_red, c.partials + grp);
}

namespace {
// e, r update of the eight cells base + lane + 64 k of a slot: e += alpha p,
// r -= alpha q (init: r = (float) r64, e = 0); cells outside the fluid keep
// e = r = 0.  Every load is issued before the first use (no load waits on a
// branch), and e, r are stored for every cell of a live slot.  cf, rk: the
// cells' coefficients and new r.  The loads come first (cg_load_cells), then
// the step length is formed from the partial sums, then cg_store_cells.
struct CellLoads {
    uint16_t cf[8];
    float pv[8], ev[8], qv[8], rv[8];
    double r0[8];
};
__device__ inline void cg_load_cells(const CgFused& c, uint32_t base, uint32_t lane, int init,
                                     const double* __restrict__ r64, const float* __restrict__ p, CellLoads& L) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t v = base + 64u * (uint32_t)k + lane;
        L.cf[k] = c.coef[v];
        if (init) {
            L.r0[k] = r64[v];
        } else {
            L.pv[k] = p[v];
            L.ev[k] = c.e[v];
            L.qv[k] = c.q[v];
            L.rv[k] = c.r[v];
        }
    }
}
__device__ inline void cg_store_cells(const CgFused& c, uint32_t base, uint32_t lane, int init, float alpha, bool live,
                                      const CellLoads& L, float rk[8]) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t v = base + 64u * (uint32_t)k + lane;
        float rn, en;
        if (init) {
            rn = (float)L.r0[k];
            en = 0.f;
        } else {
            en = fmaf(alpha, L.pv[k], L.ev[k]);
            rn = fmaf(-alpha, L.qv[k], L.rv[k]);
        }
        if (!L.cf[k]) rn = en = 0.f;
        if (live) {
            c.e[v] = en;
            c.r[v] = rn;
        }
        rk[k] = rn;
    }
}

// An update runs unless the refinement has converged; the refinement's
// start (init) clears the flag instead (workgroup 0).
__device__ inline bool cg_proceed(const CgFused& c, int init, uint32_t grp) {
    if (init) {
        if (grp == 0 && threadIdx.x == 0) *c.stop = 0u;
        return true;
    }
    return *c.stop == 0u;
}

// alpha of iteration it from (p.q)_it and rz_it (recorded by the matvec
// kernel or the reduction after the update); 0 for the refinement's start.
// With W > 1 the workgroup sums (p.q)_it itself and workgroup 0 records it.
template <uint32_t W>
__device__ inline float cg_alpha_from(const CgFused& c, uint32_t it, int init, uint32_t ng, uint32_t grp,
                                      double* s_red) {
    if (init) return 0.f;
    double pq;
    if constexpr (W > 1) {
        double unused;
        group_sum<W>(c.partials, nullptr, ng, s_red, pq, unused);
        if (grp == 0 && threadIdx.x == 0) c.hist[4u * it + 2u] = pq;
    } else {
        pq = c.hist[4u * it + 2u];  // recorded by the reduction after the matvec
    }
    const double rz = c.hist[4u * it];
    return pq != 0.0 ? (float)(rz / pq) : 0.f;
}

// v where bit `bit` of w is set, else +0 (bit for bit: an AND with 0 or ~0)
__device__ inline float keep_if_bit(float v, uint32_t w, int bit) {
    const int m = (int)(w << (31 - bit)) >> 31;
    return __int_as_float(__float_as_int(v) & m);
}
}  // namespace

// Jacobi: e += alpha p, r -= alpha q (init: r = (float) r64, e = 0),
// z = r / diag; partials r.z and r.r per group.  Wave = slot, each lane the
// eight cells lane + 64 k.
template <uint32_t W, int INIT>
__global__ void __launch_bounds__(64 * W)
cg_update_jacobi_kernel(CgFused c, uint32_t it, const double* __restrict__ r64,
                        const float* __restrict__ p) {
    constexpr int init = INIT;  // the refinement's start
    __shared__ double s_red[2 * W];
    const uint32_t ng = (c.na + W - 1u) / W;
    const uint32_t grp = cg_group(ng);
    if (grp >= ng) return;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t slotRaw = grp * W + wave;
    const bool live = slotRaw < c.na;
    const uint32_t slot = live ? slotRaw : c.na - 1u;
    if (!cg_proceed(c, init, grp)) return;
    CellLoads L;
    cg_load_cells(c, slot * 512u, lane, init, r64, p, L);
    const float alpha = cg_alpha_from<W>(c, it, init, ng, grp, s_red);
    float rk[8];
    cg_store_cells(c, slot * 512u, lane, init, alpha, live, L, rk);
    double rz = 0.0, rr = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t v = slot * 512u + 64u * (uint32_t)k + lane;
        const float rv = rk[k];
        const float zv = L.cf[k] ? rv / (float)(L.cf[k] & 7u) : 0.f;
        if (live) c.z[v] = zv;
        rr += (double)rv * (double)rv;
        rz += (double)rv * (double)zv;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        rz += __shfl_xor(rz, d, 64);
        rr += __shfl_xor(rr, d, 64);
    }
    group_partial<W>(live ? rz : 0.0, s_red, c.partials + ng + grp);
    group_partial<W>(live ? rr : 0.0, s_red + W, c.partials + 2u * ng + grp);
}

// Pieces of the MIC(0) update's sweeps, per wave.  LDS layout: (value,
// precondition factor) and the coefficient word of cell c at entry
// kMicPad + c of kMicEntries, zero entries below and above.
constexpr uint32_t kMicPad = 65, kMicEntries = 65 + 512 + 64;

// This lane's cell of each plane (two per word), from the plane table.
__device__ inline void mic_plane_cells(uint32_t lane, uint32_t cellOf2[11]) {
#pragma unroll
    for (int s = 0; s < 11; ++s)
        cellOf2[s] = (uint32_t)c_planes.lane_cell[2 * s][lane] | ((uint32_t)c_planes.lane_cell[2 * s + 1][lane] << 16);
}

__device__ inline void mic_clear_pads(float2* s_a, uint16_t* s_c, uint32_t lane) {
    for (uint32_t e = lane; e < kMicPad; e += 64u) {
        s_a[e] = make_float2(0.f, 0.f);
        s_c[e] = 0;
    }
    s_a[kMicPad + 512u + lane] = make_float2(0.f, 0.f);
    s_c[kMicPad + 512u + lane] = 0;
}

// The slot's r (rk), factors and in-brick coupling words into LDS; rr += r.r
// of this lane's cells.
__device__ inline void mic_stage(float2* s_a, uint16_t* s_c, uint32_t lane, const uint16_t cf[8], const float rk[8],
                                 const float pc[8], double& rr) {
    const uint32_t i = lane & 7u, j = lane >> 3;
    // in-brick coupling bits of the mask (-x, +x, -y, +y, -z, +z at bits 3..8)
    const uint32_t keepXY = ~((i == 0u ? 1u << 3 : 0u) | (i == 7u ? 1u << 4 : 0u) |
                              (j == 0u ? 1u << 5 : 0u) | (j == 7u ? 1u << 6 : 0u));
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t l = lane + 64u * (uint32_t)k;
        const float rv = rk[k];
        rr += (double)rv * (double)rv;
        const uint32_t keep = keepXY & ~((k == 0 ? 1u << 7 : 0u) | (k == 7 ? 1u << 8 : 0u));
        s_c[kMicPad + l] = (uint16_t)(cf[k] & keep);
        s_a[kMicPad + l] = make_float2(rv, pc[k]);  // r, replaced by the forward sweep's q plane by plane
    }
}

// Forward then backward sweep: z = M^-1 r in place of r.
__device__ inline void mic_sweeps(float2* s_a, const uint16_t* s_c, const uint32_t cellOf2[11]) {
    auto cellOf = [&](int s) { return (s & 1) ? cellOf2[s >> 1] >> 16 : cellOf2[s >> 1] & 0xffffu; };
#pragma unroll
    for (int s = 0; s < 22; ++s) {
        const uint32_t b = cellOf(s);  // own entry b + 64
        const float2 own = s_a[b + 64u], mx = s_a[b + 63u], my = s_a[b + 56u], mz = s_a[b];
        const uint32_t w = s_c[b + 64u];
        float acc = own.x;
        acc += keep_if_bit(mx.y * mx.x, w, 3);
        acc += keep_if_bit(my.y * my.x, w, 5);
        acc += keep_if_bit(mz.y * mz.x, w, 7);
        s_a[b + 64u].x = acc * own.y;  // 0 off the fluid (pc = 0), and on the zero entry
        wave_lds_sync();
    }
#pragma unroll
    for (int s = 21; s >= 0; --s) {
        const uint32_t b = cellOf(s);
        const float2 own = s_a[b + 64u];
        const float ux = s_a[b + 65u].x, uy = s_a[b + 72u].x, uz = s_a[b + 128u].x;
        const uint32_t w = s_c[b + 64u];
        float acc = own.x;
        acc += keep_if_bit(own.y * ux, w, 4);
        acc += keep_if_bit(own.y * uy, w, 6);
        acc += keep_if_bit(own.y * uz, w, 8);
        s_a[b + 64u].x = acc * own.y;
        wave_lds_sync();
    }
}

// MIC(0): the same update, then z = M^-1 r by the brick's forward and
// backward sweeps over its 22 anti-diagonal planes, all in the wave's LDS.
// LDS holds (value, precondition factor) and the coefficient word of cell c
// at entry 65 + c, with 65 zero entries below and 64 above, so a cell's six
// neighbours are at fixed offsets (-1, -8, -64, +1, +8, +64) and every read
// of a plane is issued at once, without a branch: a lane with no cell on the
// plane works on the zero entry 64.  The coefficient words keep only
// couplings inside the brick, so a neighbour offset that wraps to another
// row or plane meets a cleared bit.  The coupling A(c, m) = -1 of fluid
// neighbours is the matching bit of c's own word; the terms (A pc_m) q_m and
// (A pc_c) z_m are formed as -(pc q) (negation is exact), i.e. added where
// the bit is set.  The wave's LDS operations execute in order, so a plane's
// reads see the previous plane's writes without waiting for them
// (wave_lds_sync).
template <uint32_t W, int INIT>
__global__ void __launch_bounds__(64 * W)
cg_update_mic_kernel(CgFused c, uint32_t it, const double* __restrict__ r64, const float* __restrict__ p) {
    constexpr int init = INIT;  // the refinement's start
    __shared__ float2 s_aw[W][kMicEntries];
    __shared__ uint16_t s_cw[W][kMicEntries];
    __shared__ double s_red[2 * W];
    const uint32_t ng = (c.na + W - 1u) / W;
    const uint32_t grp = cg_group(ng);
    if (grp >= ng) return;
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t slotRaw = grp * W + wave;
    const bool live = slotRaw < c.na;
    const uint32_t slot = live ? slotRaw : c.na - 1u;
    const uint32_t base = slot * 512u;
    float2* s_a = s_aw[wave];
    uint16_t* s_c = s_cw[wave];
    // W > 1: the partial sum first (fewer registers live across it); W == 1:
    // the slot's loads first (the step length is two history words)
    if (!cg_proceed(c, init, grp)) return;
    // this lane's cell of each plane (two per word).  W == 1 (large systems,
    // every slot's wave in flight) fetches it first, with the slot's loads;
    // W > 1 after the update, when the loaded values are dead (fewer
    // registers: more waves per SIMD, which the small systems need more).
    uint32_t cellOf2[11];
    if constexpr (W == 1) mic_plane_cells(lane, cellOf2);
    float alpha = 0.f;
    if constexpr (W > 1) alpha = cg_alpha_from<W>(c, it, init, ng, grp, s_red);
    float pc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) pc[k] = c.precon[base + lane + 64u * (uint32_t)k];
    CellLoads L;
    cg_load_cells(c, base, lane, init, r64, p, L);
    if constexpr (W == 1) alpha = cg_alpha_from<W>(c, it, init, ng, grp, s_red);
    float rk[8];  // this lane's r values (cells lane + 64 k)
    cg_store_cells(c, base, lane, init, alpha, live, L, rk);
    if constexpr (W > 1) mic_plane_cells(lane, cellOf2);
    double rz = 0.0, rr = 0.0;
    mic_clear_pads(s_a, s_c, lane);
    mic_stage(s_a, s_c, lane, L.cf, rk, pc, rr);
    wave_lds_sync();
    mic_sweeps(s_a, s_c, cellOf2);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t l = lane + 64u * (uint32_t)k;
        const float zv = s_a[kMicPad + l].x;
        if (live) c.z[base + l] = zv;
        rz += (double)rk[k] * (double)zv;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        rz += __shfl_xor(rz, d, 64);
        rr += __shfl_xor(rr, d, 64);
    }
    group_partial<W>(live ? rz : 0.0, s_red, c.partials + ng + grp);
    group_partial<W>(live ? rr : 0.0, s_red + W, c.partials + 2u * ng + grp);
}

// Smoothing by a brick-local incomplete Cholesky factor around the learned
// correction (the unmodified IC(0) factor: with the MIC(0) one, whose
// M^-1 A has eigenvalues from 1 up to O(1/h), the undamped step grows
// errors), one wave per slot with the MIC(0) update's sweeps: out = a + b +
// M^-1 in on the fluid cells (a, b may be null), 0 elsewhere.  Returns at
// once while *stop is set.
__global__ void __launch_bounds__(64)
mic_smooth_kernel(const float* __restrict__ in, const float* __restrict__ a, const float* __restrict__ b,
                  const uint16_t* __restrict__ coef, const float* __restrict__ precon, float* __restrict__ out,
                  uint32_t na, const uint32_t* __restrict__ stop) {
    __shared__ float2 s_a[kMicEntries];
    __shared__ uint16_t s_c[kMicEntries];
    if (stop && *stop) return;
    const uint32_t slot = xcd_slot(na);
    if (slot >= na) return;
    const uint32_t lane = threadIdx.x, base = slot * 512u;
    uint16_t cf[8];
    float rk[8], pc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t v = base + lane + 64u * (uint32_t)k;
        cf[k] = coef[v];
        rk[k] = in[v];
        pc[k] = precon[v];
    }
    uint32_t cellOf2[11];
    mic_plane_cells(lane, cellOf2);
    double rr = 0.0;  // (not needed here)
    mic_clear_pads(s_a, s_c, lane);
    mic_stage(s_a, s_c, lane, cf, rk, pc, rr);
    wave_lds_sync();
    mic_sweeps(s_a, s_c, cellOf2);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t l = lane + 64u * (uint32_t)k, v = base + l;
        float z = s_a[kMicPad + l].x;
        if (a) z += a[v];
        if (b) z += b[v];
        out[v] = cf[k] ? z : 0.f;
    }
}

// r1 = r - A (a + b) on the fluid cells (b may be null), 0 elsewhere; the
// slot's own operand staged in LDS as in presmooth_kernel, whose pairs
// (sum of r1^2, count over cellType 1) it also writes when pairs != null.
__global__ void __launch_bounds__(kWG)
smooth_residual_kernel(const float* __restrict__ r, const float* __restrict__ a, const float* __restrict__ b,
                       const uint16_t* __restrict__ coef, const int32_t* __restrict__ nbrSlot,
                       const uint8_t* __restrict__ cellType, float* __restrict__ r1, double* __restrict__ pairs,
                       const uint32_t* __restrict__ stop) {
    __shared__ float sx[512];
    __shared__ double s_tmp[16];
    if (stop && *stop) return;
    const uint32_t slot = blockIdx.x;
    auto xf = [&](uint32_t i) { return coef[i] ? a[i] + (b ? b[i] : 0.f) : 0.f; };
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) sx[l] = xf(slot * 512u + l);
    __syncthreads();
    auto xs = [&](uint32_t i) { return (i >> 9) == slot ? sx[i & 511u] : xf(i); };
    double ss = 0.0, cnt = 0.0;
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) {
        const uint32_t v = slot * 512u + l;
        const uint16_t cf = coef[v];
        const float rv = cf ? r[v] - apply_A_with(slot, l, cf, nbrSlot, xs) : 0.f;
        r1[v] = rv;
        if (pairs && cellType[v] == 1) {
            ss += (double)rv * (double)rv;
            cnt += 1.0;
        }
    }
    if (pairs) {
        const double sa = block_sum(ss, s_tmp);
        const double sc = block_sum(cnt, s_tmp);
        if (threadIdx.x == 0) {
            pairs[2 * blockIdx.x] = sa;
            pairs[2 * blockIdx.x + 1] = sc;
        }
    }
}

// hist[4 it], hist[4 it + 1] = rz, rr from the update's partials (the sums
// the next matvec kernel forms), for the host's check at the end of a batch.
template <uint32_t W>
__global__ void __launch_bounds__(64 * W) cg_history_kernel(CgFused c, uint32_t it) {
    __shared__ double s_red[2 * W];
    if (*c.stop) return;
    const uint32_t ng = (c.na + W - 1u) / W;  // the grouped kernels' workgroups
    double rz, rr;
    group_sum<W>(c.partials + ng, c.partials + 2u * ng, ng, s_red, rz, rr);
    if (threadIdx.x == 0) {
        c.hist[4u * it] = rz;
        c.hist[4u * it + 1u] = rr;
        if (rr <= c.thr2 || !(rr == rr)) *c.stop = 1u;
    }
}

// The one-slot-per-workgroup kernels' sums (as reduce_partials_kernel),
// skipped once the refinement has converged; with checkRR, quantity 1 is an
// entry's rr and sets the flag when it meets the tolerance.
__global__ void __launch_bounds__(kReduceThreads)
cg_reduce_kernel(const double* __restrict__ partials, uint32_t blocks, double* __restrict__ out, uint32_t* stop,
                 double thr2, int checkRR) {
    __shared__ double s_tmp[kReduceThreads / 64];
    if (*stop) return;
    const double* p = partials + (size_t)blockIdx.x * blocks;
    double acc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) acc[k] = 0.0;
    for (uint32_t b0 = threadIdx.x; b0 < blocks; b0 += 8u * blockDim.x) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t b = b0 + (uint32_t)k * blockDim.x;
            v[k] = b < blocks ? p[b] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) acc[k] += v[k];
    }
    const double a = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    const double t = block_sum(a, s_tmp);
    if (threadIdx.x == 0) {
        out[blockIdx.x] = t;
        if (checkRR && blockIdx.x == 1 && (t <= thr2 || !(t == t))) *stop = 1u;
    }
}

// ---------------------------------------------------------------------------
// Velocity update: u -= (dt/h)(p_c - p_{c-1}) on faces between two
// non-solid cells with at least one fluid side.  Each active brick updates
// the low faces of its cells and, where the next brick is inactive, the low
// faces of that brick's first layer.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kWG)
velocity_update_kernel(const uint32_t* __restrict__ activeBricks, const int32_t* __restrict__ nbrSlot,
                       const uint8_t* __restrict__ cellType, const uint8_t* __restrict__ solid, Grid g,
                       const double* __restrict__ x64, float dtOverH, float* __restrict__ U,
                       float* __restrict__ V, float* __restrict__ W) {
    const uint32_t slot = blockIdx.x;
    const uint32_t b = activeBricks[slot];
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    float* F[3] = {U, V, W};
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) {
        int gc[3];
        cell_coords(b, l, nbX, nbY, gc);
        const uint8_t ty = cellType[slot * 512u + l];
        const double pc = (ty == kFluid) ? x64[slot * 512u + l] : 0.0;
        const uint32_t c0 = b * 512u + l;
        for (int axis = 0; axis < 3; ++axis) {
            // low face of this cell
            if (gc[axis] > 0) {
                const NbrRef r = neighbour(slot, l, 2 * axis, gc, g, nbrSlot, cellType, solid, nbX, nbY);
                if (ty != kSolid && r.type != kSolid && (ty == kFluid || r.type == kFluid)) {
                    const double pm = (r.type == kFluid) ? x64[(uint32_t)r.slot * 512u + r.local] : 0.0;
                    F[axis][c0] -= (float)((double)dtOverH * (pc - pm));
                }
            }
            // low face of the next cell when it lies in an inactive brick
            const uint32_t lc = axis == 0 ? (l & 7u) : axis == 1 ? ((l >> 3) & 7u) : (l >> 6);
            if (lc == 7u && ty == kFluid) {
                const NbrRef r = neighbour(slot, l, 2 * axis + 1, gc, g, nbrSlot, cellType, solid, nbX, nbY);
                if (r.slot < 0 && r.type == kAir) {
                    int c[3] = {gc[0], gc[1], gc[2]};
                    c[axis] += 1;
                    const uint32_t gi = tiled((uint32_t)c[0], (uint32_t)c[1], (uint32_t)c[2], nbX, nbY);
                    F[axis][gi] -= (float)((double)dtOverH * (0.0 - pc));
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Body force, FLIP state and boundary faces
// ---------------------------------------------------------------------------
__global__ void rasterize_solid_kernel(const uint8_t* __restrict__ staticMask, SolidBox box, Grid g,
                                       uint8_t* __restrict__ solid, uint32_t numCells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= numCells) return;
    int gc[3];
    cell_coords(c >> 9, c & 511u, g.sizeX >> 3, g.sizeY >> 3, gc);
    solid[c] = ((staticMask && staticMask[c]) || in_box(gc[0], gc[1], gc[2], g, box)) ? 1 : 0;
}

// Faces touching the domain boundary take velocity 0, faces touching a solid
// cell the solid's velocity (the box velocity for box cells, else 0); the
// others keep the P2G value, v faces get + g dt.  The result is also saved
// as the pre-projection field for the FLIP update.
__global__ void prepare_faces_kernel(float* __restrict__ U, float* __restrict__ V, float* __restrict__ W,
                                     float* __restrict__ U0, float* __restrict__ V0, float* __restrict__ W0,
                                     const uint8_t* __restrict__ solid, Grid g, float gdt, SolidBox box,
                                     uint32_t numCells) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= numCells) return;
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    int gc[3];
    cell_coords(c >> 9, c & 511u, nbX, nbY, gc);
    const float in[3] = {U[c], V[c], W[c]};
    float saved[3], out[3];
    prepare_cell_faces(in, gc, g, solid, box, gdt, nbX, nbY, saved, out);
    U0[c] = saved[0];
    V0[c] = saved[1];
    W0[c] = saved[2];
    U[c] = out[0];
    V[c] = out[1];
    W[c] = out[2];
}

// ---------------------------------------------------------------------------
// G2P (FLIP blend) and RK2 advection
// ---------------------------------------------------------------------------
namespace {

// Trilinear stencils of the MAC components at q (cell units), built per
// axis.  The brick-tiled index is separable,
//   ((kb nbY + jb) nbX + ib) 512 + (k & 7) 64 + (j & 7) 8 + (i & 7),
// so a corner's index is the sum of one offset per axis, and its weight the
// product of one weight per axis.  A corner outside [0, size) reads 0
// (walls): its axis weight is 0 and its offset that of a stored face, so the
// eight loads of a stencil need no branch and are in flight together;
// adding 0 * (finite value) leaves a running sum unchanged, so the sums equal
// those that skip such corners.  Component c is normal to axis c: it uses
// the coordinate q[d] on axis c and q[d] - 0.5 on the other two, so one
// position needs six axis stencils for all three components.
struct AxisStencil {
    uint32_t off[2];
    float w[2];
};

__device__ inline AxisStencil axis_stencil(double v, int size, uint32_t brickStride, uint32_t cellStride) {
    const double f = floor(v);
    const int base = (int)f;
    const float t = (float)(v - f);
    AxisStencil a;
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        const int c = base + o;
        const uint32_t cc = (uint32_t)min(max(c, 0), size - 1);
        a.off[o] = (cc >> 3) * brickStride + (cc & 7u) * cellStride;
        a.w[o] = (c >= 0 && c < size) ? (o ? t : 1.f - t) : 0.f;
    }
    return a;
}

struct PosStencil {
    AxisStencil ax[3][2];  // [axis][0]: at q[axis]; [axis][1]: at q[axis] - 0.5
};

__device__ inline void pos_stencil(const double q[3], Grid g, uint32_t nbX, uint32_t nbY, PosStencil& ps) {
    const int size[3] = {(int)g.sizeX, (int)g.sizeY, (int)g.sizeZ};
    const uint32_t bs[3] = {512u, nbX * 512u, nbX * nbY * 512u}, cs[3] = {1u, 8u, 64u};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        ps.ax[d][0] = axis_stencil(q[d], size[d], bs[d], cs[d]);
        ps.ax[d][1] = axis_stencil(q[d] - 0.5, size[d], bs[d], cs[d]);
    }
}

__device__ inline AxisStencil pick(const AxisStencil& a, const AxisStencil& b, bool first) {
    AxisStencil r;
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        r.off[o] = first ? a.off[o] : b.off[o];
        r.w[o] = first ? a.w[o] : b.w[o];
    }
    return r;
}

// The eight corners of component c, (a, b, e) order: indices and weights
// ((w_a w_b) w_e, the product order of the per-corner form).
__device__ inline void component_corners(const PosStencil& ps, int c, uint32_t idx[8], float w[8]) {
    const AxisStencil X = pick(ps.ax[0][0], ps.ax[0][1], c == 0);
    const AxisStencil Y = pick(ps.ax[1][0], ps.ax[1][1], c == 1);
    const AxisStencil Z = pick(ps.ax[2][0], ps.ax[2][1], c == 2);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int k = a * 4 + b * 2 + e;
                idx[k] = X.off[a] + Y.off[b] + Z.off[e];
                w[k] = X.w[a] * Y.w[b] * Z.w[e];
            }
}

__device__ inline float interp(const float* F, int comp, const PosStencil& ps) {
    uint32_t idx[8];
    float w[8], v[8];
    component_corners(ps, comp, idx, w);
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = F[idx[k]];
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) acc = fmaf(w[k], v[k], acc);
    return acc;
}

// interp(F) and interp(F - F0) on one stencil
__device__ inline void interp_pic_delta(const float* F, const float* F0, int comp, const PosStencil& ps, float& pic,
                                        float& delta) {
    uint32_t idx[8];
    float w[8], v[8], v0[8];
    component_corners(ps, comp, idx, w);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[k] = F[idx[k]];
        v0[k] = F0[idx[k]];
    }
    float a = 0.f, d = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a = fmaf(w[k], v[k], a);
        d = fmaf(w[k], v[k] - v0[k], d);
    }
    pic = a;
    delta = d;
}

// flip (v + delta) + (1 - flip) pic with one explicit fma: left to the
// compiler, the contraction of the plain expression depends on the code
// around it, and the kernels that share this update must round it alike.
__device__ inline float flip_blend(float flip, float v, float delta, float pic) {
    return fmaf(flip, v + delta, (1.f - flip) * pic);
}

__device__ inline void to_cell_units(float x, float y, float z, Grid g, double q[3]) {
    const double inv = 1.0 / (double)g.cellSize;
    q[0] = ((double)x - (double)g.negX) * inv;
    q[1] = ((double)y - (double)g.negY) * inv;
    q[2] = ((double)z - (double)g.negZ) * inv;
}

}  // namespace

// v_p = flip * (v_p + interp(u - u0)) + (1 - flip) * interp(u)
__global__ void __launch_bounds__(256) g2p_flip_kernel(uint32_t n, float4* __restrict__ rec, const float* __restrict__ U,
                                const float* __restrict__ V, const float* __restrict__ W,
                                const float* __restrict__ U0, const float* __restrict__ V0,
                                const float* __restrict__ W0, Grid g, float flip) {
    const uint32_t p = xcd_slot((n + 255u) >> 8) * 256u + threadIdx.x;  // particles sorted by brick
    if (p >= n) return;
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    float4 a = rec[2 * p], b = rec[2 * p + 1];
    double q[3];
    to_cell_units(a.x, a.y, a.z, g, q);
    PosStencil ps;
    pos_stencil(q, g, nbX, nbY, ps);
    const float* F[3] = {U, V, W};
    const float* F0[3] = {U0, V0, W0};
    float vel[3] = {a.w, b.x, b.y};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float pic, dlt;
        interp_pic_delta(F[c], F0[c], c, ps, pic, dlt);
        vel[c] = flip_blend(flip, vel[c], dlt, pic);
    }
    a.w = vel[0];
    b.x = vel[1];
    b.y = vel[2];
    rec[2 * p] = a;
    rec[2 * p + 1] = b;
}

namespace {

// Midpoint (RK2) advection of x0 through the grid velocity, u1 being the
// velocity at x0; then clamping into [lo, hi] per axis and out of solid
// cells: a particle that ends inside the solid box is pushed out through the
// box face nearest to it, one in another solid cell returns to its start
// point.
// The midpoint of the RK2 step and its position in cell units.
__device__ inline void advect_midpoint(const float x0[3], const float u1[3], Grid g, float dt, double q[3]) {
    const float xm[3] = {fmaf(0.5f * dt, u1[0], x0[0]), fmaf(0.5f * dt, u1[1], x0[1]), fmaf(0.5f * dt, u1[2], x0[2])};
    to_cell_units(xm[0], xm[1], xm[2], g, q);
}

// x0 + dt u2 (u2: the velocity at the midpoint), clamped into [lo, hi] per
// axis and out of solid cells: a particle that ends inside the solid box is
// pushed out through the box face nearest to it, one in another solid cell
// returns to its start point.
__device__ inline void advect_finish(const float x0[3], const float u2[3], const uint8_t* solid, Grid g, float dt,
                                     float margin, const SolidBox& box, float xn[3]) {
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    const float lo[3] = {g.negX + margin, g.negY + margin, g.negZ + margin};
    // explicit fmas: the same rounding in every kernel that inlines this
    const float hi[3] = {fmaf((float)g.sizeX, g.cellSize, g.negX) - margin,
                         fmaf((float)g.sizeY, g.cellSize, g.negY) - margin,
                         fmaf((float)g.sizeZ, g.cellSize, g.negZ) - margin};
#pragma unroll
    for (int c = 0; c < 3; ++c) xn[c] = fminf(fmaxf(fmaf(dt, u2[c], x0[c]), lo[c]), hi[c]);
    if (box.active && xn[0] >= box.lo[0] && xn[0] < box.hi[0] && xn[1] >= box.lo[1] && xn[1] < box.hi[1] &&
        xn[2] >= box.lo[2] && xn[2] < box.hi[2]) {
        int best = -1;
        float bestD = 3.4e38f, target = 0.f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float dl = xn[c] - box.lo[c], dh = box.hi[c] - xn[c];
            if (box.lo[c] - margin >= lo[c] && dl < bestD) { bestD = dl; best = c; target = box.lo[c] - margin; }
            if (box.hi[c] + margin <= hi[c] && dh < bestD) { bestD = dh; best = c; target = box.hi[c] + margin; }
        }
        if (best >= 0) xn[best] = target;
    }
    if (solid) {
        double q[3];
        to_cell_units(xn[0], xn[1], xn[2], g, q);
        const int ci = (int)floor(q[0]), cj = (int)floor(q[1]), ck = (int)floor(q[2]);
        if (cell_solid(ci, cj, ck, g, solid, nbX, nbY) && !in_box(ci, cj, ck, g, box)) {
            xn[0] = x0[0];
            xn[1] = x0[1];
            xn[2] = x0[2];
        }
    }
}

// Midpoint (RK2) advection of x0 through the grid velocity, u1 being the
// velocity at x0.
__device__ inline void advect_from(const float x0[3], const float u1[3], const float* U, const float* V,
                                   const float* W, const uint8_t* solid, Grid g, float dt, float margin,
                                   const SolidBox& box, float xn[3]) {
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    const float* F[3] = {U, V, W};
    double q[3];
    advect_midpoint(x0, u1, g, dt, q);
    PosStencil ps;
    pos_stencil(q, g, nbX, nbY, ps);
    float u2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) u2[c] = interp(F[c], c, ps);
    advect_finish(x0, u2, solid, g, dt, margin, box, xn);
}

// The stencils of a position with offsets into a 10 x 10 x 10 region of
// cells (x fastest) whose first cell is grid cell org: the corners and
// weights of pos_stencil (the same clamping, so the same face values);
// false if a corner lies outside the region (its offset is then clamped
// into the region, and the caller uses the grid instead).
constexpr int kRegion = 10;
__device__ inline bool region_stencil(const double q[3], Grid g, const int org[3], PosStencil& ps) {
    const int size[3] = {(int)g.sizeX, (int)g.sizeY, (int)g.sizeZ};
    const uint32_t st[3] = {1u, (uint32_t)kRegion, (uint32_t)(kRegion * kRegion)};
    bool inside = true;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int sh = 0; sh < 2; ++sh) {
            const double v = sh ? q[d] - 0.5 : q[d];
            const double f = floor(v);
            const int base = (int)f;
            const float t = (float)(v - f);
#pragma unroll
            for (int o = 0; o < 2; ++o) {
                const int c = base + o;
                const int r = min(max(c, 0), size[d] - 1) - org[d];
                inside = inside && r >= 0 && r < kRegion;
                ps.ax[d][sh].off[o] = (uint32_t)min(max(r, 0), kRegion - 1) * st[d];
                ps.ax[d][sh].w[o] = (c >= 0 && c < size[d]) ? (o ? t : 1.f - t) : 0.f;
            }
        }
    return inside;
}

}  // namespace

__global__ void __launch_bounds__(256) advect_rk2_kernel(uint32_t n, float4* __restrict__ rec, const float* __restrict__ U,
                                  const float* __restrict__ V, const float* __restrict__ W,
                                  const uint8_t* __restrict__ solid, Grid g, float dt, float margin, SolidBox box) {
    const uint32_t p = xcd_slot((n + 255u) >> 8) * 256u + threadIdx.x;  // particles sorted by brick
    if (p >= n) return;
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    float4 a = rec[2 * p];
    const float x0[3] = {a.x, a.y, a.z};
    double q[3];
    to_cell_units(x0[0], x0[1], x0[2], g, q);
    PosStencil ps;
    pos_stencil(q, g, nbX, nbY, ps);
    const float* F[3] = {U, V, W};
    float u1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) u1[c] = interp(F[c], c, ps);
    float xn[3];
    advect_from(x0, u1, U, V, W, solid, g, dt, margin, box, xn);
    a.x = xn[0];
    a.y = xn[1];
    a.z = xn[2];
    rec[2 * p] = a;
}

// G2P followed by RK2 advection in one pass over the records: the same
// results as g2p_flip_kernel then advect_rk2_kernel (advection reads the
// grid velocity and the position, which G2P leaves unchanged; its first
// interpolation is G2P's PIC term).
__global__ void __launch_bounds__(256) g2p_advect_kernel(uint32_t n, float4* __restrict__ rec, const float* __restrict__ U,
                                  const float* __restrict__ V, const float* __restrict__ W,
                                  const float* __restrict__ U0, const float* __restrict__ V0,
                                  const float* __restrict__ W0, const uint8_t* __restrict__ solid, Grid g, float flip,
                                  float dt, float margin, SolidBox box) {
    const uint32_t p = xcd_slot((n + 255u) >> 8) * 256u + threadIdx.x;  // particles sorted by brick
    if (p >= n) return;
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    float4 a = rec[2 * p], b = rec[2 * p + 1];
    const float x0[3] = {a.x, a.y, a.z};
    double q[3];
    to_cell_units(x0[0], x0[1], x0[2], g, q);
    PosStencil ps;
    pos_stencil(q, g, nbX, nbY, ps);
    const float* F[3] = {U, V, W};
    const float* F0[3] = {U0, V0, W0};
    float vel[3] = {a.w, b.x, b.y}, u1[3];
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
        float dlt;
        interp_pic_delta(F[c], F0[c], c, ps, u1[c], dlt);
        vel[c] = flip_blend(flip, vel[c], dlt, u1[c]);
    }
    float xn[3];
    advect_from(x0, u1, U, V, W, solid, g, dt, margin, box, xn);
    rec[2 * p] = make_float4(xn[0], xn[1], xn[2], vel[0]);
    b.x = vel[1];
    b.y = vel[2];
    rec[2 * p + 1] = b;
}

// G2P + RK2 by brick, for the step (the particles sorted by brick, the
// touched-brick list and brick starts of the fused transfer): a workgroup copies the faces
// of a brick's cells and their one-cell shell (10^3 cells, the six fields,
// 24 KB) into LDS once, then interpolates for the brick's particles from
// LDS -- a particle's G2P stencil lies in its brick's region, and the RK2
// midpoint's usually does (else it reads the grid) -- instead of gathering
// every corner from memory.  The same results as g2p_advect_kernel (the
// same weights, corners and summation order over the same face values).
// The list is cut into 64-entry chunks dealt to the XCDs round robin (as
// p2g_prepare_kernel walks it); a workgroup takes the next entry of its
// XCD's share from that XCD's counter (counters[xcd], zero at launch), one
// entry ahead, so workgroups with light bricks take more of them.
__global__ void __launch_bounds__(256)
g2p_advect_brick_kernel(const uint32_t* __restrict__ list, const uint32_t* __restrict__ count,
                        uint32_t* __restrict__ counters, const uint32_t* __restrict__ brickStart, float4* __restrict__ rec, const float* __restrict__ U,
                        const float* __restrict__ V, const float* __restrict__ W, const float* __restrict__ U0,
                        const float* __restrict__ V0, const float* __restrict__ W0, const uint8_t* __restrict__ solid,
                        Grid g, float flip, float dt, float margin, SolidBox box, uint32_t* __restrict__ keys) {
    __shared__ float sF[6][kRegion * kRegion * kRegion];
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    const uint32_t total = *count;
    __shared__ uint32_t s_next;
    const uint32_t xcd = blockIdx.x & 7u;  // workgroups are dealt to the XCDs round robin
    const float* Fg[6] = {U, V, W, U0, V0, W0};
    const float* F[3] = {U, V, W};
    const float* F0[3] = {U0, V0, W0};
    if (threadIdx.x == 0) s_next = atomicAdd(&counters[xcd], 1u);
    __syncthreads();
    for (uint32_t m = s_next;; m = s_next) {
        // the XCD's m-th entry: offset m % 64 of its (m / 64)-th chunk
        const uint32_t e = ((m >> 6) * 8u + xcd) * 64u + (m & 63u);
        if (e >= total) break;  // (the same for the whole workgroup) later entries are past it too
        __syncthreads();        // every thread has read s_next
        if (threadIdx.x == 0) s_next = atomicAdd(&counters[xcd], 1u);  // the next entry, during this brick
        const uint32_t brick = list[e];
        const uint32_t p0 = brickStart[brick], p1 = brickStart[brick + 1u];
        if (p0 == p1) {  // a touched brick without particles
            __syncthreads();  // s_next is written
            continue;
        }
        const int org[3] = {(int)(brick % nbX) * 8 - 1, (int)((brick / nbX) % nbY) * 8 - 1,
                            (int)(brick / (nbX * nbY)) * 8 - 1};
        for (uint32_t t = threadIdx.x; t < (uint32_t)(kRegion * kRegion * kRegion); t += 256u) {
            const int x = org[0] + (int)(t % kRegion), y = org[1] + (int)((t / kRegion) % kRegion),
                      z = org[2] + (int)(t / (kRegion * kRegion));
            const bool in = in_domain(x, y, z, g);
            const uint32_t idx = in ? tiled((uint32_t)x, (uint32_t)y, (uint32_t)z, nbX, nbY) : 0u;
#pragma unroll
            for (int a = 0; a < 6; ++a) sF[a][t] = in ? Fg[a][idx] : 0.f;
        }
        __syncthreads();
        // each iteration loads the next one's records before its own work
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
        if (p0 + threadIdx.x < p1) {
            a = rec[2 * (p0 + threadIdx.x)];
            b = rec[2 * (p0 + threadIdx.x) + 1];
        }
        for (uint32_t p = p0 + threadIdx.x; p < p1; p += 256u) {
            float4 an = a, bn = b;
            if (p + 256u < p1) {
                an = rec[2 * (p + 256u)];
                bn = rec[2 * (p + 256u) + 1];
            }
            const float x0[3] = {a.x, a.y, a.z};
            double q[3];
            to_cell_units(x0[0], x0[1], x0[2], g, q);
            PosStencil ps;
            const bool local = region_stencil(q, g, org, ps);
            if (!local) pos_stencil(q, g, nbX, nbY, ps);
            float vel[3] = {a.w, b.x, b.y}, u1[3];
#pragma unroll 1
            for (int c = 0; c < 3; ++c) {
                float dlt;
                if (local)
                    interp_pic_delta(sF[c], sF[3 + c], c, ps, u1[c], dlt);
                else
                    interp_pic_delta(F[c], F0[c], c, ps, u1[c], dlt);
                vel[c] = flip_blend(flip, vel[c], dlt, u1[c]);
            }
            double qm[3];
            advect_midpoint(x0, u1, g, dt, qm);
            PosStencil pm;
            const bool localM = region_stencil(qm, g, org, pm);
            if (!localM) pos_stencil(qm, g, nbX, nbY, pm);
            float u2[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) u2[c] = localM ? interp(sF[c], c, pm) : interp(F[c], c, pm);
            float xn[3];
            advect_finish(x0, u2, solid, g, dt, margin, box, xn);
            rec[2 * p] = make_float4(xn[0], xn[1], xn[2], vel[0]);
            b.x = vel[1];
            b.y = vel[2];
            rec[2 * p + 1] = b;
            if (keys) keys[p] = cell_key(xn[0], xn[1], xn[2], g);  // the next transfer's sort key
            a = an;
            b = bn;
        }
        __syncthreads();  // the next brick overwrites the region
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace {
inline dim3 g1(uint32_t n, uint32_t bs) { return dim3(ceil_div(n ? n : 1, bs)); }
// grid of a kernel that maps workgroups to slots with xcd_slot
inline dim3 xcd_grid(uint32_t na) { return dim3(8u * ceil_div(na ? na : 1, 8)); }
// out[0..count) = sums of the workgroup partials of the kernel just launched
// With red.stop: skipped once converged (cg_reduce_kernel); checkRR: sum 1
// is a residual norm that sets the flag.
void reduce_to(const Reduce& red, int count, uint32_t blocks, double* out, hipStream_t s, bool checkRR = false) {
    // one load per thread per 8 partials: a small sum needs fewer threads
    const int threads = blocks <= 2048u ? 256 : kReduceThreads;
    if (red.stop)
        FLIP2_KERNEL(cg_reduce_kernel, dim3(count), dim3(threads), 0, s, red.partials, blocks, out, red.stop, red.thr2,
                     checkRR ? 1 : 0);
    else
        FLIP2_KERNEL(reduce_partials_kernel, dim3(count), dim3(threads), 0, s, red.partials, blocks, out);
    FLIP2_LAUNCH_CHECK();
}
}  // namespace

void launch_classify_bricks(const uint32_t* counts, const uint8_t* solid, uint32_t numBricks, uint32_t* flag,
                            hipStream_t s) {
    FLIP2_KERNEL(classify_bricks_kernel, dim3(numBricks), dim3(64), 0, s, counts, solid, flag);
    FLIP2_LAUNCH_CHECK();
}
void launch_classify_bricks_sparse(const uint32_t* list, const uint32_t* count, const uint8_t* occ,
                                   const uint32_t* cellStart, const uint8_t* solid, uint32_t numBricks, uint32_t* flag,
                                   hipStream_t s) {
    FLIP2_KERNEL(fill_u32_kernel, g1(numBricks, 256), dim3(256), 0, s, flag, numBricks, 0u);
    FLIP2_LAUNCH_CHECK();
    FLIP2_KERNEL(classify_bricks_sparse_kernel, dim3(std::min<uint32_t>(numBricks, 4096u)), dim3(64), 0, s, list,
                 count, occ, cellStart, solid, flag);
    FLIP2_LAUNCH_CHECK();
}
void launch_build_slots(const uint32_t* flag, const uint32_t* excl, uint32_t numBricks, int32_t* brickSlot,
                        uint32_t* activeBricks, hipStream_t s) {
    FLIP2_KERNEL(build_slots_kernel, g1(numBricks, 256), dim3(256), 0, s, flag, excl, numBricks, brickSlot,
                       activeBricks);
    FLIP2_LAUNCH_CHECK();
}
void launch_brick_setup(uint32_t numActive, const uint32_t* activeBricks, const int32_t* brickSlot,
                        const uint32_t* counts, const uint8_t* solid, Grid g, int32_t* nbrSlot, uint8_t* cellType,
                        hipStream_t s, const uint32_t* cellStart) {
    if (!numActive) return;
    FLIP2_REQUIRE(counts || cellStart, kErrArgument, "brick setup: no cell counts or offsets");
    FLIP2_KERNEL(brick_setup_kernel, dim3(numActive), dim3(kWG), 0, s, activeBricks, brickSlot,
                 counts ? counts : nullptr, counts ? nullptr : cellStart, solid, g, nbrSlot, cellType);
    FLIP2_LAUNCH_CHECK();
}
void launch_assemble(uint32_t numActive, const uint32_t* activeBricks, const int32_t* nbrSlot, const uint8_t* cellType,
                     const uint8_t* solid, Grid g, double hOverDt, const float* U, const float* V, const float* W,
                     uint16_t* coef, double* b64, Reduce red, double* out, hipStream_t s) {
    if (!numActive) return;
    FLIP2_KERNEL(assemble_kernel, dim3(numActive), dim3(kWG), 0, s, activeBricks, nbrSlot, cellType, solid, g,
                       hOverDt, U, V, W, coef, b64, red.partials);
    FLIP2_LAUNCH_CHECK();
    reduce_to(red, 2, numActive, out, s);
}
void launch_refine_start(uint32_t numActive, const double* r64, float* r, float* e, const uint16_t* coef, Reduce red,
                         double* out, hipStream_t s, uint32_t* resetStop) {
    FLIP2_KERNEL(refine_start_kernel, dim3(numActive), dim3(kWG), 0, s, r64, r, e, coef, red.partials, resetStop);
    FLIP2_LAUNCH_CHECK();
    red.stop = nullptr;  // the refinement's first sum is always taken
    reduce_to(red, 1, numActive, out, s);
}
void launch_matvec_dot(uint32_t numActive, const float* p, const uint16_t* coef, const int32_t* nbrSlot, float* q,
                       Reduce red, double* out, hipStream_t s) {
    FLIP2_KERNEL(matvec_dot_kernel, xcd_grid(numActive), dim3(kWG), 0, s, p, coef, nbrSlot, q, red.partials,
                 numActive, red.stop);
    FLIP2_LAUNCH_CHECK();
    reduce_to(red, 1, numActive, out, s);
}
void launch_update(uint32_t numActive, const double* rzOld, const double* pq, const float* p, const float* q,
                   const uint16_t* coef, float* e, float* r, float* z, bool jacobi, Reduce red, double* out,
                   hipStream_t s) {
    FLIP2_KERNEL(update_kernel, dim3(numActive), dim3(kWG), 0, s, rzOld, pq, p, q, coef, e, r, z,
                       jacobi ? 1 : 0, red.partials, red.stop);
    FLIP2_LAUNCH_CHECK();
    reduce_to(red, 2, numActive, out, s, true);
}
void launch_presmooth(uint32_t numActive, const float* r, const uint16_t* coef, const int32_t* nbrSlot,
                      const uint8_t* cellType, float omega, float* r1, double* pairs, hipStream_t s,
                      const uint32_t* stop) {
    FLIP2_KERNEL(presmooth_kernel, dim3(numActive), dim3(kWG), 0, s, r, coef, nbrSlot, cellType, omega, r1,
                       pairs, stop);
    FLIP2_LAUNCH_CHECK();
}
void launch_postsmooth(uint32_t numActive, const float* r, const float* y, const uint16_t* coef, const int32_t* nbrSlot,
                       float omega, float* z, hipStream_t s, const uint32_t* stop) {
    FLIP2_KERNEL(postsmooth_kernel, dim3(numActive), dim3(kWG), 0, s, r, y, coef, nbrSlot, omega, z, stop);
    FLIP2_LAUNCH_CHECK();
}
void launch_mic_smooth(uint32_t numActive, const float* in, const float* a, const float* b, const uint16_t* coef,
                       const float* precon, float* out, hipStream_t s, const uint32_t* stop) {
    if (!numActive) return;
    FLIP2_KERNEL(mic_smooth_kernel, xcd_grid(numActive), dim3(64), 0, s, in, a, b, coef, precon, out, numActive,
                 stop);
    FLIP2_LAUNCH_CHECK();
}
void launch_smooth_residual(uint32_t numActive, const float* r, const float* a, const float* b, const uint16_t* coef,
                            const int32_t* nbrSlot, const uint8_t* cellType, float* r1, double* pairs, hipStream_t s,
                            const uint32_t* stop) {
    if (!numActive) return;
    FLIP2_KERNEL(smooth_residual_kernel, dim3(numActive), dim3(kWG), 0, s, r, a, b, coef, nbrSlot, cellType, r1,
                 pairs, stop);
    FLIP2_LAUNCH_CHECK();
}
void launch_flex_dots(uint32_t numActive, const float* z, const FlexDirections& d, Reduce red, double* out,
                      hipStream_t s) {
    FLIP2_REQUIRE(d.count >= 0 && d.count <= kMaxFlexDirections, kErrArgument, "flexible CG: too many directions");
    if (d.count == 0) return;
    switch (d.count) {
#define FLEX_CASE(N) \
    case N: FLIP2_KERNEL(flex_dots_kernel<N>, dim3(numActive), dim3(kWG), 0, s, z, d, red.partials, red.stop); break;
        FLEX_CASE(1) FLEX_CASE(2) FLEX_CASE(3) FLEX_CASE(4) FLEX_CASE(5) FLEX_CASE(6) FLEX_CASE(7) FLEX_CASE(8)
        FLEX_CASE(9) FLEX_CASE(10) FLEX_CASE(11) FLEX_CASE(12) FLEX_CASE(13) FLEX_CASE(14) FLEX_CASE(15) FLEX_CASE(16)
#undef FLEX_CASE
    }
    FLIP2_LAUNCH_CHECK();
    reduce_to(red, d.count, numActive, out, s);
}

void launch_flex_direction(uint32_t numActive, const float* z, const FlexDirections& d, const double* zq,
                           const float* r, float* p, Reduce red, double* pr, hipStream_t s) {
    FLIP2_REQUIRE(d.count >= 0 && d.count <= kMaxFlexDirections, kErrArgument, "flexible CG: too many directions");
    for (int i = 0; i < d.count; ++i)
        FLIP2_REQUIRE(d.p[i] != p, kErrArgument, "flexible CG: output aliases a kept direction");
    switch (d.count) {
#define FLEX_CASE(N)                                                                                              \
    case N:                                                                                                       \
        FLIP2_KERNEL(flex_direction_kernel<N>, dim3(numActive), dim3(kWG), 0, s, z, d, zq, r, p, red.partials, \
                     red.stop);                                                                                   \
        break;
        FLEX_CASE(0) FLEX_CASE(1) FLEX_CASE(2) FLEX_CASE(3) FLEX_CASE(4) FLEX_CASE(5) FLEX_CASE(6) FLEX_CASE(7)
        FLEX_CASE(8) FLEX_CASE(9) FLEX_CASE(10) FLEX_CASE(11) FLEX_CASE(12) FLEX_CASE(13) FLEX_CASE(14) FLEX_CASE(15)
        FLEX_CASE(16)
#undef FLEX_CASE
    }
    FLIP2_LAUNCH_CHECK();
    reduce_to(red, 1, numActive, pr, s);
}

void launch_accumulate(uint32_t numActive, const float* e, double* x64, const uint16_t* coef, hipStream_t s) {
    FLIP2_KERNEL(accumulate_kernel, dim3(numActive), dim3(kWG), 0, s, e, x64, coef);
    FLIP2_LAUNCH_CHECK();
}
void launch_true_residual(uint32_t numActive, const double* b64, const double* x64, const uint16_t* coef,
                          const int32_t* nbrSlot, double* r64, Reduce red, double* out, hipStream_t s) {
    FLIP2_KERNEL(true_residual_kernel, xcd_grid(numActive), dim3(kWG), 0, s, b64, x64, coef, nbrSlot, r64,
                 red.partials, numActive);
    FLIP2_LAUNCH_CHECK();
    reduce_to(red, 1, numActive, out, s);
}
void launch_fill_f32(float* a, uint32_t count, float v, hipStream_t s) {
    if (!count) return;
    FLIP2_KERNEL(fill_f32_kernel, g1(count, 256), dim3(256), 0, s, a, count, v);
    FLIP2_LAUNCH_CHECK();
}
void launch_fill_f64(double* a, uint32_t count, double v, hipStream_t s) {
    if (!count) return;
    FLIP2_KERNEL(fill_f64_kernel, g1(count, 256), dim3(256), 0, s, a, count, v);
    FLIP2_LAUNCH_CHECK();
}
void launch_mic_factor(uint32_t numActive, const uint16_t* coef, float* precon, hipStream_t s, bool modified) {
    FLIP2_KERNEL(mic_factor_kernel, dim3(numActive), dim3(64), 0, s, coef, precon, modified ? kMicTau : 0.0f);
    FLIP2_LAUNCH_CHECK();
}
namespace {
int cg_reduce_threads(uint32_t blocks) { return blocks <= 2048u ? 256 : kReduceThreads; }
}  // namespace
namespace {
// Waves (slots) per workgroup of the grouped shape for na slots; 0 for the
// one-slot shape with reduction kernels (at 8042 slots, C3: 76 us per
// iteration against 84 with four slots and 90 with eight per workgroup).
uint32_t cg_group_waves(uint32_t na) { return na < kCgGroupedBelow ? 4u : 0u; }

template <uint32_t W>
void launch_matvec_grouped(const CgFused& c, uint32_t it, bool first, const float* pOld, float* pNew, hipStream_t s) {
    FLIP2_KERNEL(cg_direction_matvec_kernel<W>, xcd_grid((c.na + W - 1u) / W), dim3(64 * W), 0, s, c, it,
                 first ? 1 : 0, pOld, pNew);
    FLIP2_LAUNCH_CHECK();
}
template <uint32_t W>
void launch_update_grouped(const CgFused& c, uint32_t it, bool init, bool mic, const double* r64, const float* p,
                           hipStream_t s) {
    const dim3 grid = xcd_grid((c.na + W - 1u) / W), block(64 * W);
    if (mic && init)
        FLIP2_KERNEL((cg_update_mic_kernel<W, 1>), grid, block, 0, s, c, it, r64, p);
    else if (mic)
        FLIP2_KERNEL((cg_update_mic_kernel<W, 0>), grid, block, 0, s, c, it, r64, p);
    else if (init)
        FLIP2_KERNEL((cg_update_jacobi_kernel<W, 1>), grid, block, 0, s, c, it, r64, p);
    else
        FLIP2_KERNEL((cg_update_jacobi_kernel<W, 0>), grid, block, 0, s, c, it, r64, p);
    FLIP2_LAUNCH_CHECK();
}
}  // namespace

uint32_t cg_partials_needed(uint32_t na) {
    const uint32_t w = cg_group_waves(na);
    return 3u * (w ? (na + w - 1u) / w : na);
}
void launch_cg_direction_matvec(const CgFused& c, uint32_t it, bool first, const float* pOld, float* pNew,
                                hipStream_t s) {
    FLIP2_REQUIRE(c.na > 0 && c.stop, kErrArgument, "PCG: no active slots or no convergence flag");
    switch (cg_group_waves(c.na)) {
    case 4: launch_matvec_grouped<4>(c, it, first, pOld, pNew, s); return;
    default: break;
    }
    FLIP2_KERNEL(cg_direction_matvec_kernel<1>, xcd_grid(c.na), dim3(64), 0, s, c, it, first ? 1 : 0, pOld, pNew);
    FLIP2_LAUNCH_CHECK();
    FLIP2_KERNEL(cg_reduce_kernel, dim3(1), dim3(cg_reduce_threads(c.na)), 0, s, c.partials, c.na,
                 c.hist + 4u * it + 2u, c.stop, c.thr2, 0);  // (p.q)_it
    FLIP2_LAUNCH_CHECK();
}
void launch_cg_update_precond(const CgFused& c, uint32_t it, bool init, bool mic, const double* r64, const float* p,
                              hipStream_t s) {
    FLIP2_REQUIRE(c.na > 0 && c.stop, kErrArgument, "PCG: no active slots or no convergence flag");
    switch (cg_group_waves(c.na)) {
    case 4: launch_update_grouped<4>(c, it, init, mic, r64, p, s); return;
    default: break;
    }
    const dim3 grid = xcd_grid(c.na);
    if (mic && init)
        FLIP2_KERNEL((cg_update_mic_kernel<1, 1>), grid, dim3(64), 0, s, c, it, r64, p);
    else if (mic)
        FLIP2_KERNEL((cg_update_mic_kernel<1, 0>), grid, dim3(64), 0, s, c, it, r64, p);
    else if (init)
        FLIP2_KERNEL((cg_update_jacobi_kernel<1, 1>), grid, dim3(64), 0, s, c, it, r64, p);
    else
        FLIP2_KERNEL((cg_update_jacobi_kernel<1, 0>), grid, dim3(64), 0, s, c, it, r64, p);
    FLIP2_LAUNCH_CHECK();
    // rz, rr of the next entry (partials at [na, 3 na))
    FLIP2_KERNEL(cg_reduce_kernel, dim3(2), dim3(cg_reduce_threads(c.na)), 0, s, c.partials + c.na, c.na,
                 c.hist + 4u * (init ? it : it + 1u), c.stop, c.thr2, 1);
    FLIP2_LAUNCH_CHECK();
}
void launch_cg_history(const CgFused& c, uint32_t it, hipStream_t s) {
    switch (cg_group_waves(c.na)) {
    case 4: FLIP2_KERNEL(cg_history_kernel<4>, dim3(1), dim3(256), 0, s, c, it); break;
    default: return;  // the reduction after each update recorded it
    }
    FLIP2_LAUNCH_CHECK();
}
void launch_velocity_update(uint32_t numActive, const uint32_t* activeBricks, const int32_t* nbrSlot,
                            const uint8_t* cellType, const uint8_t* solid, Grid g, const double* x64, float dtOverH,
                            float* U, float* V, float* W, hipStream_t s) {
    if (!numActive) return;
    FLIP2_KERNEL(velocity_update_kernel, dim3(numActive), dim3(kWG), 0, s, activeBricks, nbrSlot, cellType,
                       solid, g, x64, dtOverH, U, V, W);
    FLIP2_LAUNCH_CHECK();
}
void launch_prepare_faces(float* U, float* V, float* W, float* U0, float* V0, float* W0, const uint8_t* solid, Grid g,
                          float gdt, SolidBox box, hipStream_t s) {
    const uint32_t cells = g.sizeX * g.sizeY * g.sizeZ;
    FLIP2_KERNEL(prepare_faces_kernel, g1(cells, 256), dim3(256), 0, s, U, V, W, U0, V0, W0, solid, g, gdt, box,
                       cells);
    FLIP2_LAUNCH_CHECK();
}
void launch_rasterize_solid(const uint8_t* staticMask, SolidBox box, Grid g, uint8_t* solid, hipStream_t s) {
    const uint32_t cells = g.sizeX * g.sizeY * g.sizeZ;
    FLIP2_KERNEL(rasterize_solid_kernel, g1(cells, 256), dim3(256), 0, s, staticMask, box, g, solid, cells);
    FLIP2_LAUNCH_CHECK();
}
void launch_g2p_flip(uint32_t n, float4* rec, const float* U, const float* V, const float* W, const float* U0,
                     const float* V0, const float* W0, Grid g, float flip, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(g2p_flip_kernel, xcd_grid(ceil_div(n, 256)), dim3(256), 0, s, n, rec, U, V, W, U0, V0, W0, g, flip);
    FLIP2_LAUNCH_CHECK();
}
void launch_advect_rk2(uint32_t n, float4* rec, const float* U, const float* V, const float* W, const uint8_t* solid,
                       Grid g, float dt, float margin, SolidBox box, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(advect_rk2_kernel, xcd_grid(ceil_div(n, 256)), dim3(256), 0, s, n, rec, U, V, W, solid, g, dt, margin, box);
    FLIP2_LAUNCH_CHECK();
}
void launch_g2p_advect_bricks(const uint32_t* list, const uint32_t* count, uint32_t* counters,
                              const uint32_t* brickStart, float4* rec,
                              const float* U, const float* V, const float* W, const float* U0, const float* V0,
                              const float* W0, const uint8_t* solid, Grid g, float flip, float dt, float margin,
                              SolidBox box, hipStream_t s, uint32_t* keys) {
    const uint32_t numBricks = (g.sizeX >> 3) * (g.sizeY >> 3) * (g.sizeZ >> 3);
    // as many workgroups as the CUs hold at once (registers and LDS), 8 per XCD row
    const uint32_t groups =
        std::min<uint32_t>(resident_groups(g2p_advect_brick_kernel, 256), 8u * ((numBricks + 7u) >> 3));
    FLIP2_KERNEL(fill_u32_kernel, dim3(1), dim3(64), 0, s, counters, 8u, 0u);
    FLIP2_LAUNCH_CHECK();
    FLIP2_KERNEL(g2p_advect_brick_kernel, dim3(groups), dim3(256), 0, s, list, count, counters, brickStart, rec, U, V,
                 W, U0, V0, W0, solid, g, flip, dt, margin, box, keys);
    FLIP2_LAUNCH_CHECK();
}
void launch_g2p_advect(uint32_t n, float4* rec, const float* U, const float* V, const float* W, const float* U0,
                       const float* V0, const float* W0, const uint8_t* solid, Grid g, float flip, float dt,
                       float margin, SolidBox box, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(g2p_advect_kernel, xcd_grid(ceil_div(n, 256)), dim3(256), 0, s, n, rec, U, V, W, U0, V0, W0, solid, g, flip, dt,
                 margin, box);
    FLIP2_LAUNCH_CHECK();
}

}  // namespace ns
}  // namespace flip2
