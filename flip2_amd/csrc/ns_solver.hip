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
                   const uint32_t* __restrict__ counts, const uint8_t* __restrict__ solid, Grid g,
                   int32_t* __restrict__ nbrSlot, uint8_t* __restrict__ cellType) {
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
        else if (counts[c] > 0u) ty = kFluid;
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

// r32 = (float) r64; e = 0; rr into scal[out].
// out[q] = sum over workgroups of partials[q * blocks + b], fixed order:
// workgroup q sums quantity q, thread t the partials t, t + 256, ...
__global__ void __launch_bounds__(kWG)
reduce_partials_kernel(const double* __restrict__ partials, uint32_t blocks, double* __restrict__ out) {
    __shared__ double s_tmp[16];
    const double* p = partials + (size_t)blockIdx.x * blocks;
    double acc = 0.0;
    for (uint32_t b = threadIdx.x; b < blocks; b += blockDim.x) acc += p[b];
    const double t = block_sum(acc, s_tmp);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

__global__ void __launch_bounds__(kWG)
refine_start_kernel(const double* __restrict__ r64, float* __restrict__ r, float* __restrict__ e,
                    const uint16_t* __restrict__ coef, double* __restrict__ partials) {
    __shared__ double s_tmp[16];
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
                  float* __restrict__ q, double* __restrict__ partials) {
    __shared__ double s_tmp[16];
    const uint32_t slot = blockIdx.x;
    double pq = 0.0;
    for (uint32_t l = threadIdx.x; l < 512u; l += blockDim.x) {
        const uint32_t v = slot * 512u + l;
        const float qv = apply_A(slot, l, coef[v], p, nbrSlot);
        q[v] = qv;
        pq += (double)p[v] * (double)qv;
    }
    const double t = block_sum(pq, s_tmp);
    store_partials(&t, 1, partials);
}

// alpha = rz / pq; e += alpha p; r -= alpha q; with Jacobi also z = r/diag
// and rz_new; rr_new always.  scal layout per iteration: see ns_solver.hpp.
__global__ void __launch_bounds__(kWG)
update_kernel(const double* __restrict__ rzOld, const double* __restrict__ pq, const float* __restrict__ p,
              const float* __restrict__ q, const uint16_t* __restrict__ coef, float* __restrict__ e,
              float* __restrict__ r, float* __restrict__ z, int jacobi, double* __restrict__ partials /* [rz_new, rr_new] per workgroup */) {
    __shared__ double s_tmp[16];
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
                 double* __restrict__ pairs) {
    __shared__ float sx[512];
    __shared__ double s_tmp[16];
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
                  const int32_t* __restrict__ nbrSlot, float omega, float* __restrict__ z) {
    __shared__ float sx[512];
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
flex_dots_kernel(const float* __restrict__ z, FlexDirections d, double* __restrict__ partials) {
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
                      const float* __restrict__ r, float* __restrict__ p, double* __restrict__ partials) {
    __shared__ double s_tmp[16];
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
                     double* __restrict__ r64, double* __restrict__ partials) {
    __shared__ double s_tmp[16];
    const uint32_t slot = blockIdx.x;
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
    store_partials(&t, 1, partials);
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
// incomplete Cholesky, tau = 0.97, safety 0.25): precon = 1/sqrt(e).
__global__ void __launch_bounds__(64)
mic_factor_kernel(const uint16_t* __restrict__ coef, float* __restrict__ precon) {
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
                    e -= kMicTau * (a * (o1 + o2) * pm * pm);
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
// PCG with two main kernels per iteration (Jacobi and brick-local MIC(0)).
//
//   direction+matvec (iteration k):  beta = rz_k / rz_{k-1};  p = z + beta p';
//                                    q = A p;  partial p.q
//   update+precondition:             alpha = rz_k / (p.q)_k;  e += alpha p;
//                                    r -= alpha q;  z = M^-1 r;  partials r.z, r.r
// One workgroup (direction+matvec, Jacobi update) or one wave (MIC(0)
// update: the brick's forward and backward plane sweeps run in LDS) per
// slot, so every slot's memory latency is in flight at once; each main
// kernel is followed by a one-workgroup kernel that sums its partials in
// slot order into the iteration history (rz, rr, pq per iteration:
// deterministic), which the next kernel reads.  p alternates between two
// buffers: the matvec reads the previous direction at neighbour cells owned
// by other workgroups while writing the new one.
// ---------------------------------------------------------------------------

// p = z + beta p_old (p = z when first), q = A p; partial p.q per slot.
__global__ void __launch_bounds__(kWG)
cg_direction_matvec_kernel(CgFused c, uint32_t it, int first, const float* __restrict__ pOld,
                           float* __restrict__ pNew) {
    __shared__ double s_tmp[16];
    __shared__ float s_p[512];
    const uint32_t slot = blockIdx.x, base = slot * 512u;
    float beta = 0.f;
    if (!first) {
        const double rz = c.hist[4u * it], rzOld = c.hist[4u * (it - 1u)];
        beta = rzOld != 0.0 ? (float)(rz / rzOld) : 0.f;
    }
    for (uint32_t l = threadIdx.x; l < 512u; l += kWG) {
        const uint32_t v = base + l;
        s_p[l] = c.coef[v] ? (first ? c.z[v] : fmaf(beta, pOld[v], c.z[v])) : 0.f;
    }
    __syncthreads();
    auto px = [&](uint32_t i) {
        if ((i >> 9) == slot) return s_p[i & 511u];
        return first ? c.z[i] : fmaf(beta, pOld[i], c.z[i]);  // a fluid neighbour in another slot
    };
    double pq = 0.0;
    for (uint32_t l = threadIdx.x; l < 512u; l += kWG) {
        const uint32_t v = base + l;
        const float qv = apply_A_with(slot, l, c.coef[v], c.nbrSlot, px);
        const float pv = s_p[l];
        pNew[v] = pv;
        c.q[v] = qv;
        pq += (double)pv * (double)qv;
    }
    const double t = block_sum(pq, s_tmp);
    if (threadIdx.x == 0) c.partials[slot] = t;
}

namespace {
// e, r update of one cell (init: r = (float) r64, e = 0); returns the new r.
__device__ inline float cg_update_cell(const CgFused& c, uint32_t v, uint16_t cf, int init, float alpha,
                                       const double* __restrict__ r64, const float* __restrict__ p) {
    if (!cf) {
        if (init) {
            c.r[v] = 0.f;
            c.e[v] = 0.f;
        }
        return 0.f;
    }
    if (init) {
        const float rv = (float)r64[v];
        c.r[v] = rv;
        c.e[v] = 0.f;
        return rv;
    }
    c.e[v] = fmaf(alpha, p[v], c.e[v]);
    const float rv = fmaf(-alpha, c.q[v], c.r[v]);
    c.r[v] = rv;
    return rv;
}

__device__ inline float cg_alpha(const CgFused& c, uint32_t it, int init) {
    if (init) return 0.f;
    const double pq = c.hist[4u * it + 2u], rz = c.hist[4u * it];
    return pq != 0.0 ? (float)(rz / pq) : 0.f;
}
}  // namespace

// Jacobi: e += alpha p, r -= alpha q (init: r = (float) r64, e = 0),
// z = r / diag; partials r.z and r.r per slot.
__global__ void __launch_bounds__(kWG)
cg_update_jacobi_kernel(CgFused c, uint32_t it, int init, const double* __restrict__ r64,
                        const float* __restrict__ p) {
    __shared__ double s_tmp[16];
    const float alpha = cg_alpha(c, it, init);
    const uint32_t slot = blockIdx.x;
    double rz = 0.0, rr = 0.0;
    for (uint32_t l = threadIdx.x; l < 512u; l += kWG) {
        const uint32_t v = slot * 512u + l;
        const uint16_t cf = c.coef[v];
        const float rv = cg_update_cell(c, v, cf, init, alpha, r64, p);
        const float zv = cf ? rv / (float)(cf & 7u) : 0.f;
        c.z[v] = zv;
        rr += (double)rv * (double)rv;
        rz += (double)rv * (double)zv;
    }
    const double t0 = block_sum(rz, s_tmp);
    const double t1 = block_sum(rr, s_tmp);
    if (threadIdx.x == 0) {
        c.partials[slot] = t0;
        c.partials[c.na + slot] = t1;
    }
}

// MIC(0): the same update, then z = M^-1 r by the brick's forward and
// backward sweeps over its 22 anti-diagonal planes, all in LDS; one wave
// per slot.
__global__ void __launch_bounds__(64)
cg_update_mic_kernel(CgFused c, uint32_t it, int init, const double* __restrict__ r64, const float* __restrict__ p) {
    __shared__ float sq[512];
    __shared__ float spc[512];
    __shared__ uint16_t scf[512];
    const float alpha = cg_alpha(c, it, init);
    const uint32_t lane = threadIdx.x, slot = blockIdx.x, base = slot * 512u;
    const uint32_t stride[3] = {1u, 8u, 64u};
    double rz = 0.0, rr = 0.0;
    float rk[8];  // this lane's r values (cells lane + 64 k)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t l = lane + 64u * (uint32_t)k;
        const uint16_t cf = c.coef[base + l];
        const float pc = c.precon[base + l];
        const float rv = cg_update_cell(c, base + l, cf, init, alpha, r64, p);
        rr += (double)rv * (double)rv;
        rk[k] = rv;
        scf[l] = cf;
        spc[l] = pc;
        sq[l] = rv;  // r, replaced by the forward sweep's q plane by plane
    }
    __syncthreads();
    for (int s = 0; s <= 21; ++s) {
        const int pb = c_planes.start[s], pe = c_planes.start[s + 1];
        for (int t = pb + (int)lane; t < pe; t += 64) {
            const uint32_t l = c_planes.cells[t];
            const uint16_t cf = scf[l];
            float qv = 0.f;
            if (cf) {
                float acc = sq[l];
                const uint32_t lc[3] = {l & 7u, (l >> 3) & 7u, l >> 6};
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    if (lc[ax] == 0u) continue;
                    const uint32_t m = l - stride[ax];
                    acc -= a_plus(scf[m], ax, m) * spc[m] * sq[m];
                }
                qv = acc * spc[l];
            }
            sq[l] = qv;
        }
        __syncthreads();
    }
    for (int s = 21; s >= 0; --s) {
        const int pb = c_planes.start[s], pe = c_planes.start[s + 1];
        for (int t = pb + (int)lane; t < pe; t += 64) {
            const uint32_t l = c_planes.cells[t];
            const uint16_t cf = scf[l];
            float zz = 0.f;
            if (cf) {
                float acc = sq[l];
                const uint32_t lc[3] = {l & 7u, (l >> 3) & 7u, l >> 6};
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    if (lc[ax] == 7u) continue;
                    const uint32_t m = l + stride[ax];
                    acc -= a_plus(cf, ax, l) * spc[l] * sq[m];
                }
                zz = acc * spc[l];
            }
            sq[l] = zz;
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint32_t l = lane + 64u * (uint32_t)k;
        const float zv = sq[l];
        c.z[base + l] = zv;
        rz += (double)rk[k] * (double)zv;
    }
    for (int d = 32; d >= 1; d >>= 1) {
        rz += __shfl_xor(rz, d, 64);
        rr += __shfl_xor(rr, d, 64);
    }
    if (lane == 0) {
        c.partials[slot] = rz;
        c.partials[c.na + slot] = rr;
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
__device__ inline bool cell_solid(int i, int j, int k, Grid g, const uint8_t* solid, uint32_t nbX, uint32_t nbY) {
    if (i < 0 || j < 0 || k < 0 || i >= (int)g.sizeX || j >= (int)g.sizeY || k >= (int)g.sizeZ) return true;
    return solid && solid[tiled((uint32_t)i, (uint32_t)j, (uint32_t)k, nbX, nbY)];
}

__device__ inline bool in_box(int i, int j, int k, Grid g, const SolidBox& box) {
    if (!box.active) return false;
    const float c[3] = {g.negX + ((float)i + 0.5f) * g.cellSize, g.negY + ((float)j + 0.5f) * g.cellSize,
                        g.negZ + ((float)k + 0.5f) * g.cellSize};
    return c[0] >= box.lo[0] && c[0] < box.hi[0] && c[1] >= box.lo[1] && c[1] < box.hi[1] && c[2] >= box.lo[2] &&
           c[2] < box.hi[2];
}

__device__ inline bool in_domain(int i, int j, int k, Grid g) {
    return i >= 0 && j >= 0 && k >= 0 && i < (int)g.sizeX && j < (int)g.sizeY && k < (int)g.sizeZ;
}

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
    const uint32_t b = c >> 9, l = c & 511u;
    int gc[3];
    cell_coords(b, l, nbX, nbY, gc);
    const bool here = cell_solid(gc[0], gc[1], gc[2], g, solid, nbX, nbY);
    float* F[3] = {U, V, W};
    float* F0[3] = {U0, V0, W0};
    for (int axis = 0; axis < 3; ++axis) {
        int m[3] = {gc[0], gc[1], gc[2]};
        m[axis] -= 1;
        const bool nbSolid = cell_solid(m[0], m[1], m[2], g, solid, nbX, nbY);
        const bool wall = here || nbSolid;
        float vs = 0.f;  // solid velocity of this face
        if (wall && in_domain(m[0], m[1], m[2], g) &&
            ((here && in_box(gc[0], gc[1], gc[2], g, box)) || (nbSolid && in_box(m[0], m[1], m[2], g, box))))
            vs = box.vel[axis];
        float v = wall ? vs : F[axis][c];
        F0[axis][c] = v;
        if (axis == 1 && !wall) v += gdt;
        F[axis][c] = v;
    }
}

// ---------------------------------------------------------------------------
// G2P (FLIP blend) and RK2 advection
// ---------------------------------------------------------------------------
namespace {

// Trilinear interpolation of a MAC component at q (cell units); faces
// outside the stored range read 0 (walls).
__device__ inline float interp(const float* F, int comp, const double q[3], Grid g, uint32_t nbX, uint32_t nbY) {
    int base[3];
    float t[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double v = (d == comp) ? q[d] : q[d] - 0.5;
        const double f = floor(v);
        base[d] = (int)f;
        t[d] = (float)(v - f);
    }
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int i = base[0] + a, j = base[1] + b, k = base[2] + e;
                if (i < 0 || j < 0 || k < 0 || i >= (int)g.sizeX || j >= (int)g.sizeY || k >= (int)g.sizeZ) continue;
                const float w = (a ? t[0] : 1.f - t[0]) * (b ? t[1] : 1.f - t[1]) * (e ? t[2] : 1.f - t[2]);
                acc = fmaf(w, F[tiled((uint32_t)i, (uint32_t)j, (uint32_t)k, nbX, nbY)], acc);
            }
    return acc;
}

__device__ inline float interp_delta(const float* F, const float* F0, int comp, const double q[3], Grid g,
                                     uint32_t nbX, uint32_t nbY) {
    int base[3];
    float t[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double v = (d == comp) ? q[d] : q[d] - 0.5;
        const double f = floor(v);
        base[d] = (int)f;
        t[d] = (float)(v - f);
    }
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int i = base[0] + a, j = base[1] + b, k = base[2] + e;
                if (i < 0 || j < 0 || k < 0 || i >= (int)g.sizeX || j >= (int)g.sizeY || k >= (int)g.sizeZ) continue;
                const float w = (a ? t[0] : 1.f - t[0]) * (b ? t[1] : 1.f - t[1]) * (e ? t[2] : 1.f - t[2]);
                const uint32_t idx = tiled((uint32_t)i, (uint32_t)j, (uint32_t)k, nbX, nbY);
                acc = fmaf(w, F[idx] - F0[idx], acc);
            }
    return acc;
}

__device__ inline void to_cell_units(float x, float y, float z, Grid g, double q[3]) {
    const double inv = 1.0 / (double)g.cellSize;
    q[0] = ((double)x - (double)g.negX) * inv;
    q[1] = ((double)y - (double)g.negY) * inv;
    q[2] = ((double)z - (double)g.negZ) * inv;
}

}  // namespace

// v_p = flip * (v_p + interp(u - u0)) + (1 - flip) * interp(u)
__global__ void g2p_flip_kernel(uint32_t n, float4* __restrict__ rec, const float* __restrict__ U,
                                const float* __restrict__ V, const float* __restrict__ W,
                                const float* __restrict__ U0, const float* __restrict__ V0,
                                const float* __restrict__ W0, Grid g, float flip) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    float4 a = rec[2 * p], b = rec[2 * p + 1];
    double q[3];
    to_cell_units(a.x, a.y, a.z, g, q);
    const float* F[3] = {U, V, W};
    const float* F0[3] = {U0, V0, W0};
    float vel[3] = {a.w, b.x, b.y};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float pic = interp(F[c], c, q, g, nbX, nbY);
        const float dlt = interp_delta(F[c], F0[c], c, q, g, nbX, nbY);
        vel[c] = flip * (vel[c] + dlt) + (1.f - flip) * pic;
    }
    a.w = vel[0];
    b.x = vel[1];
    b.y = vel[2];
    rec[2 * p] = a;
    rec[2 * p + 1] = b;
}

// Midpoint (RK2) advection through the grid velocity, then clamping into
// [lo, hi] per axis and out of solid cells: a particle that ends inside the
// solid box is pushed out through the box face nearest to it, one in another
// solid cell returns to its start point.
__global__ void advect_rk2_kernel(uint32_t n, float4* __restrict__ rec, const float* __restrict__ U,
                                  const float* __restrict__ V, const float* __restrict__ W,
                                  const uint8_t* __restrict__ solid, Grid g, float dt, float margin, SolidBox box) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    float4 a = rec[2 * p];
    const float x0[3] = {a.x, a.y, a.z};
    double q[3];
    to_cell_units(x0[0], x0[1], x0[2], g, q);
    const float* F[3] = {U, V, W};
    float u1[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) u1[c] = interp(F[c], c, q, g, nbX, nbY);
    const float xm[3] = {fmaf(0.5f * dt, u1[0], x0[0]), fmaf(0.5f * dt, u1[1], x0[1]), fmaf(0.5f * dt, u1[2], x0[2])};
    to_cell_units(xm[0], xm[1], xm[2], g, q);
    float xn[3];
    const float lo[3] = {g.negX + margin, g.negY + margin, g.negZ + margin};
    const float hi[3] = {g.negX + (float)g.sizeX * g.cellSize - margin, g.negY + (float)g.sizeY * g.cellSize - margin,
                         g.negZ + (float)g.sizeZ * g.cellSize - margin};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float u2 = interp(F[c], c, q, g, nbX, nbY);
        xn[c] = fminf(fmaxf(fmaf(dt, u2, x0[c]), lo[c]), hi[c]);
    }
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
        to_cell_units(xn[0], xn[1], xn[2], g, q);
        const int ci = (int)floor(q[0]), cj = (int)floor(q[1]), ck = (int)floor(q[2]);
        if (cell_solid(ci, cj, ck, g, solid, nbX, nbY) && !in_box(ci, cj, ck, g, box)) {
            xn[0] = x0[0]; xn[1] = x0[1]; xn[2] = x0[2];
        }
    }
    a.x = xn[0];
    a.y = xn[1];
    a.z = xn[2];
    rec[2 * p] = a;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
namespace {
inline dim3 g1(uint32_t n, uint32_t bs) { return dim3(ceil_div(n ? n : 1, bs)); }
// out[0..count) = sums of the workgroup partials of the kernel just launched
void reduce_to(const Reduce& red, int count, uint32_t blocks, double* out, hipStream_t s) {
    FLIP2_KERNEL(reduce_partials_kernel, dim3(count), dim3(kWG), 0, s, red.partials, blocks, out);
    FLIP2_LAUNCH_CHECK();
}
}  // namespace

void launch_classify_bricks(const uint32_t* counts, const uint8_t* solid, uint32_t numBricks, uint32_t* flag,
                            hipStream_t s) {
    FLIP2_KERNEL(classify_bricks_kernel, dim3(numBricks), dim3(64), 0, s, counts, solid, flag);
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
                        hipStream_t s) {
    if (!numActive) return;
    FLIP2_KERNEL(brick_setup_kernel, dim3(numActive), dim3(kWG), 0, s, activeBricks, brickSlot, counts, solid, g,
                       nbrSlot, cellType);
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
                         double* out, hipStream_t s) {
    FLIP2_KERNEL(refine_start_kernel, dim3(numActive), dim3(kWG), 0, s, r64, r, e, coef, red.partials);
    FLIP2_LAUNCH_CHECK();
    reduce_to(red, 1, numActive, out, s);
}
void launch_matvec_dot(uint32_t numActive, const float* p, const uint16_t* coef, const int32_t* nbrSlot, float* q,
                       Reduce red, double* out, hipStream_t s) {
    FLIP2_KERNEL(matvec_dot_kernel, dim3(numActive), dim3(kWG), 0, s, p, coef, nbrSlot, q, red.partials);
    FLIP2_LAUNCH_CHECK();
    reduce_to(red, 1, numActive, out, s);
}
void launch_update(uint32_t numActive, const double* rzOld, const double* pq, const float* p, const float* q,
                   const uint16_t* coef, float* e, float* r, float* z, bool jacobi, Reduce red, double* out,
                   hipStream_t s) {
    FLIP2_KERNEL(update_kernel, dim3(numActive), dim3(kWG), 0, s, rzOld, pq, p, q, coef, e, r, z,
                       jacobi ? 1 : 0, red.partials);
    FLIP2_LAUNCH_CHECK();
    reduce_to(red, 2, numActive, out, s);
}
void launch_presmooth(uint32_t numActive, const float* r, const uint16_t* coef, const int32_t* nbrSlot,
                      const uint8_t* cellType, float omega, float* r1, double* pairs, hipStream_t s) {
    FLIP2_KERNEL(presmooth_kernel, dim3(numActive), dim3(kWG), 0, s, r, coef, nbrSlot, cellType, omega, r1,
                       pairs);
    FLIP2_LAUNCH_CHECK();
}
void launch_postsmooth(uint32_t numActive, const float* r, const float* y, const uint16_t* coef, const int32_t* nbrSlot,
                       float omega, float* z, hipStream_t s) {
    FLIP2_KERNEL(postsmooth_kernel, dim3(numActive), dim3(kWG), 0, s, r, y, coef, nbrSlot, omega, z);
    FLIP2_LAUNCH_CHECK();
}
void launch_flex_dots(uint32_t numActive, const float* z, const FlexDirections& d, Reduce red, double* out,
                      hipStream_t s) {
    FLIP2_REQUIRE(d.count >= 0 && d.count <= kMaxFlexDirections, kErrArgument, "flexible CG: too many directions");
    if (d.count == 0) return;
    switch (d.count) {
#define FLEX_CASE(N) \
    case N: FLIP2_KERNEL(flex_dots_kernel<N>, dim3(numActive), dim3(kWG), 0, s, z, d, red.partials); break;
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
        FLIP2_KERNEL(flex_direction_kernel<N>, dim3(numActive), dim3(kWG), 0, s, z, d, zq, r, p, red.partials); \
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
    FLIP2_KERNEL(true_residual_kernel, dim3(numActive), dim3(kWG), 0, s, b64, x64, coef, nbrSlot, r64,
                       red.partials);
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
void launch_mic_factor(uint32_t numActive, const uint16_t* coef, float* precon, hipStream_t s) {
    FLIP2_KERNEL(mic_factor_kernel, dim3(numActive), dim3(64), 0, s, coef, precon);
    FLIP2_LAUNCH_CHECK();
}
void launch_cg_direction_matvec(const CgFused& c, uint32_t it, bool first, const float* pOld, float* pNew,
                                hipStream_t s) {
    FLIP2_KERNEL(cg_direction_matvec_kernel, dim3(c.na), dim3(kWG), 0, s, c, it, first ? 1 : 0, pOld, pNew);
    FLIP2_LAUNCH_CHECK();
    reduce_to(Reduce{c.partials}, 1, c.na, c.hist + 4u * it + 2u, s);  // (p.q)_it
}
void launch_cg_update_precond(const CgFused& c, uint32_t it, bool init, bool mic, const double* r64, const float* p,
                              hipStream_t s) {
    if (mic)
        FLIP2_KERNEL(cg_update_mic_kernel, dim3(c.na), dim3(64), 0, s, c, it, init ? 1 : 0, r64, p);
    else
        FLIP2_KERNEL(cg_update_jacobi_kernel, dim3(c.na), dim3(kWG), 0, s, c, it, init ? 1 : 0, r64, p);
    FLIP2_LAUNCH_CHECK();
    reduce_to(Reduce{c.partials}, 2, c.na, c.hist + 4u * (init ? it : it + 1u), s);  // rz, rr of the next entry
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
    FLIP2_KERNEL(g2p_flip_kernel, g1(n, 256), dim3(256), 0, s, n, rec, U, V, W, U0, V0, W0, g, flip);
    FLIP2_LAUNCH_CHECK();
}
void launch_advect_rk2(uint32_t n, float4* rec, const float* U, const float* V, const float* W, const uint8_t* solid,
                       Grid g, float dt, float margin, SolidBox box, hipStream_t s) {
    if (!n) return;
    FLIP2_KERNEL(advect_rk2_kernel, g1(n, 256), dim3(256), 0, s, n, rec, U, V, W, solid, g, dt, margin, box);
    FLIP2_LAUNCH_CHECK();
}

}  // namespace ns
}  // namespace flip2
