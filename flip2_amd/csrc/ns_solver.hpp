// North-star pressure projection, G2P and advection kernels (ns_solver.hip).
#pragma once

#include "common.hpp"

namespace flip2 {
namespace ns {

constexpr uint8_t kAir = 0;
constexpr uint8_t kFluid = 1;
constexpr uint8_t kSolid = 2;

// Deterministic grid-wide reduction scratch: one double per workgroup per
// quantity (a follow-up kernel sums them in workgroup order).
struct Reduce {
    double* partials;
    // Flexible CG's device-side convergence (null: off): the kernels return
    // at once while *stop is set, the sums skip as cg_reduce_kernel does, and
    // the update's residual sum sets *stop when it meets thr2 (or is NaN).
    uint32_t* stop = nullptr;
    double thr2 = 0.0;
};

// A moving solid box (e.g. a wave paddle): cells whose centre lies in
// [lo, hi) are solid and the faces they touch carry the box velocity.
struct SolidBox {
    float lo[3], hi[3], vel[3];
    int active;
};

void launch_classify_bricks(const uint32_t* counts, const uint8_t* solid, uint32_t numBricks, uint32_t* flag,
                            hipStream_t s);
// The same after a fused transfer, whose cell offsets cover the occupied
// bricks of its touched list only (flag is cleared first).
void launch_classify_bricks_sparse(const uint32_t* list, const uint32_t* count, const uint8_t* occ,
                                   const uint32_t* cellStart, const uint8_t* solid, uint32_t numBricks, uint32_t* flag,
                                   hipStream_t s);
void launch_build_slots(const uint32_t* flag, const uint32_t* excl, uint32_t numBricks, int32_t* brickSlot,
                        uint32_t* activeBricks, hipStream_t s);
// Cell types from counts, or (counts null) from the cell offsets cellStart.
void launch_brick_setup(uint32_t numActive, const uint32_t* activeBricks, const int32_t* brickSlot,
                        const uint32_t* counts, const uint8_t* solid, Grid g, int32_t* nbrSlot, uint8_t* cellType,
                        hipStream_t s, const uint32_t* cellStart = nullptr);
void launch_assemble(uint32_t numActive, const uint32_t* activeBricks, const int32_t* nbrSlot, const uint8_t* cellType,
                     const uint8_t* solid, Grid g, double hOverDt, const float* U, const float* V, const float* W,
                     uint16_t* coef, double* b64, Reduce red, double* out, hipStream_t s);
// r = (float) r64 on the fluid cells, e = 0, ||r||^2 into out[0]; clears
// *resetStop (if given) for the refinement it starts.
void launch_refine_start(uint32_t numActive, const double* r64, float* r, float* e, const uint16_t* coef, Reduce red,
                         double* out, hipStream_t s, uint32_t* resetStop = nullptr);
void launch_matvec_dot(uint32_t numActive, const float* p, const uint16_t* coef, const int32_t* nbrSlot, float* q,
                       Reduce red, double* out, hipStream_t s);
void launch_update(uint32_t numActive, const double* rzOld, const double* pq, const float* p, const float* q,
                   const uint16_t* coef, float* e, float* r, float* z, bool jacobi, Reduce red, double* out,
                   hipStream_t s);
// Flexible CG with the last `count` directions kept (Notay's truncated FCG):
// p_i, q_i = A p_i and pq_i = p_i . q_i (device scalar) of each kept direction.
constexpr int kMaxFlexDirections = 16;
struct FlexDirections {
    const float* p[kMaxFlexDirections];
    const float* q[kMaxFlexDirections];
    const double* pq[kMaxFlexDirections];
    int count;
};
// out[i] = z . q_i for the kept directions.
void launch_flex_dots(uint32_t numActive, const float* z, const FlexDirections& d, Reduce red, double* out,
                      hipStream_t s);
// p = z - sum_i (zq[i] / pq_i) p_i (p must not be one of the p_i); pr = p . r.
void launch_flex_direction(uint32_t numActive, const float* z, const FlexDirections& d, const double* zq,
                           const float* r, float* p, Reduce red, double* pr, hipStream_t s);
// PCG (Jacobi or brick MIC(0)) with two main kernels per iteration, each
// followed by a one-workgroup fixed-order sum of its per-slot partials into
// the iteration history (ns_solver.hip).  History entry k (4 doubles at
// hist + 4k): rz_k, rr_k, (p.q)_k.
struct CgFused {
    const uint16_t* coef;
    const int32_t* nbrSlot;
    const float* precon;  // MIC(0) factors (null for Jacobi)
    float* q;
    float* e;
    float* r;
    float* z;
    double* partials;  // cg_partials_needed(na) doubles
    double* hist;
    uint32_t na;
    // Convergence on the device: *stop is cleared by the refinement's start
    // and set once an entry's rr <= thr2 (or is not a number); from then on
    // every kernel of the refinement returns at once, so the host may queue
    // iterations ahead of its checks.
    uint32_t* stop;
    double thr2;
};
// Doubles of CgFused::partials for na slots.
uint32_t cg_partials_needed(uint32_t na);
// Iteration `it`: history rz_it, rr_it (entry it) from the previous update's
// partials; p = z (first) or z + (rz_it / rz_{it-1}) pOld into pNew;
// q = A p; partials of p.q.
void launch_cg_direction_matvec(const CgFused& c, uint32_t it, bool first, const float* pOld, float* pNew,
                                hipStream_t s);
// Iteration `it`: history (p.q)_it from the matvec's partials;
// alpha = rz_it / (p.q)_it; e += alpha p; r -= alpha q; z = M^-1 r;
// partials of rz, rr (history entry it + 1, recorded by the next matvec or
// launch_cg_history).  init: instead r = (float) r64 (0 off the fluid),
// e = 0 (p unused), partials of entry `it` -- the start of a refinement.
void launch_cg_update_precond(const CgFused& c, uint32_t it, bool init, bool mic, const double* r64, const float* p,
                              hipStream_t s);
// history rz, rr of entry `it` from the last update's partials (the values
// the next matvec would record), for a check at the end of a batch
void launch_cg_history(const CgFused& c, uint32_t it, hipStream_t s);
void launch_accumulate(uint32_t numActive, const float* e, double* x64, const uint16_t* coef, hipStream_t s);
void launch_true_residual(uint32_t numActive, const double* b64, const double* x64, const uint16_t* coef,
                          const int32_t* nbrSlot, double* r64, Reduce red, double* out, hipStream_t s);
void launch_fill_f32(float* a, uint32_t count, float v, hipStream_t s);
void launch_fill_f64(double* a, uint32_t count, double v, hipStream_t s);
// modified: the MIC(0) factor (the PCG preconditioner); false: IC(0)
// (the learned preconditioner's smoother).
void launch_mic_factor(uint32_t numActive, const uint16_t* coef, float* precon, hipStream_t s, bool modified = true);
void launch_velocity_update(uint32_t numActive, const uint32_t* activeBricks, const int32_t* nbrSlot,
                            const uint8_t* cellType, const uint8_t* solid, Grid g, const double* x64, float dtOverH,
                            float* U, float* V, float* W, hipStream_t s);
// Damped-Jacobi smoothing around an external correction y (see ns_solver.hip):
// r1 = r - A (w D^-1 r);  z = x + w D^-1 (r - A x) with x = w D^-1 r + y.
// pairs != null (2 numActive doubles): per slot (sum of r1^2, count) over the
// cells with cellType 1, the input scale of the learned correction.
// stop (if given): the kernels return at once while *stop is set.
void launch_presmooth(uint32_t numActive, const float* r, const uint16_t* coef, const int32_t* nbrSlot,
                      const uint8_t* cellType, float omega, float* r1, double* pairs, hipStream_t s,
                      const uint32_t* stop = nullptr);
void launch_postsmooth(uint32_t numActive, const float* r, const float* y, const uint16_t* coef, const int32_t* nbrSlot,
                       float omega, float* z, hipStream_t s, const uint32_t* stop = nullptr);
// IC(0) smoothing around the learned correction: out = a + b + M^-1 in on
// the fluid cells (a, b may be null; M the brick-local factors of
// launch_mic_factor with modified = false), and r1 = r - A (a + b) (b may be
// null) with the network input's pairs as launch_presmooth.
void launch_mic_smooth(uint32_t numActive, const float* in, const float* a, const float* b, const uint16_t* coef,
                       const float* precon, float* out, hipStream_t s, const uint32_t* stop = nullptr);
void launch_smooth_residual(uint32_t numActive, const float* r, const float* a, const float* b, const uint16_t* coef,
                            const int32_t* nbrSlot, const uint8_t* cellType, float* r1, double* pairs, hipStream_t s,
                            const uint32_t* stop = nullptr);
void launch_prepare_faces(float* U, float* V, float* W, float* U0, float* V0, float* W0, const uint8_t* solid, Grid g,
                          float gdt, SolidBox box, hipStream_t s);
// solid = staticMask | box (brick-tiled bytes); staticMask may be null.
void launch_rasterize_solid(const uint8_t* staticMask, SolidBox box, Grid g, uint8_t* solid, hipStream_t s);
// rec: particle records, 2 x float4 per particle (x, y, z, vx), (vy, vz, 0, 0).
void launch_g2p_flip(uint32_t n, float4* rec, const float* U, const float* V, const float* W, const float* U0,
                     const float* V0, const float* W0, Grid g, float flip, hipStream_t s);
void launch_advect_rk2(uint32_t n, float4* rec, const float* U, const float* V, const float* W, const uint8_t* solid,
                       Grid g, float dt, float margin, SolidBox box, hipStream_t s);
// launch_g2p_advect by brick over the fused transfer's touched-brick list
// (list, count: entries and their number on the device; counters: 8 device
// words, cleared here, the work counters of the 8 XCDs; brickStart: the
// transfer's first sorted record of each brick, numBricks + 1 entries); the
// same results.  keys (may be null): also the cell key of each particle's
// new position (what the next transfer's binning computes).
void launch_g2p_advect_bricks(const uint32_t* list, const uint32_t* count, uint32_t* counters,
                              const uint32_t* brickStart, float4* rec,
                              const float* U, const float* V, const float* W, const float* U0, const float* V0,
                              const float* W0, const uint8_t* solid, Grid g, float flip, float dt, float margin,
                              SolidBox box, hipStream_t s, uint32_t* keys = nullptr);
// launch_g2p_flip then launch_advect_rk2 (same results) in one kernel
void launch_g2p_advect(uint32_t n, float4* rec, const float* U, const float* V, const float* W, const float* U0,
                       const float* V0, const float* W0, const uint8_t* solid, Grid g, float flip, float dt,
                       float margin, SolidBox box, hipStream_t s);

}  // namespace ns
}  // namespace flip2
