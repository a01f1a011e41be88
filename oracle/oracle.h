/*
 * CPU oracle for the flip2 per-step particle pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; it is the checker that the HIP path is compared against
 * (tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py).
 *
 * Two families live here:
 *   - "rc_*": a serial FP64 restatement of the reference implementation
 *     (Aidan-F-Ryan/flip2, engine/particles.cu and the engine/algorithms sources),
 *     including the behaviours the reference's outputs depend on.  Each
 *     function cites the reference file:line it follows.
 *   - "ns_*" (ns_oracle.c): the north-star method (trilinear MAC P2G,
 *     Jacobi / brick-MIC(0) PCG, FLIP 0.95, RK2, conv3d U-Net) in FP64.
 *
 * Pins: what checks each function against something other than itself
 * (tests/test_oracle_primitives.py = P, test_oracle_rc.py = R,
 * test_oracle_rc_brute.py = B, test_oracle_ns.py = N, test_golden.py = G):
 *   orc_find_grid_cell         P: hand cases at cell faces +-1 ulp, clamping, float narrowing
 *   orc_scan_u32               P: numpy cumsum mod 2^32
 *   orc_radix_sort_u32/_pairs  P: numpy stable argsort; partial bit ranges
 *   orc_sampled_absmax         P: hand-computed sampled maxima (Q12)
 *   orc_philox4x32_10          P: the published known-answer vectors
 *   orc_set_domain, orc_align_particles_to_grid
 *                              P, B: FP32 narrowing decides the offset-domain cases
 *   orc_sort_particles         B: positions gathered by the stable order, velocities kept (Q4)
 *   orc_generate_voxels        B: brute force over every voxel and particle of
 *                                 the reference's marking and compaction (Q6-Q9,
 *                                 Q24), segments, the hand-built row table
 *   orc_particle_vel_to_voxels R: single-particle hat weights
 *   orc_get_courant_dt         B: closed form over a sampled maximum
 *   orc_apply/remove_gravity   B: closed form on non-solid voxels
 *   orc_calc_div_u             R: linear field; B: brute force across nodes (Q17)
 *   orc_gs_iterations          R: converged iteration = dense solve
 *   orc_gs_solve               B: checks every 1024 iterations (Q14)
 *   orc_pressure_retry/_solve  B: scripted residual sequences (Q20)
 *   orc_update_voxel_velocities R: linear pressure closed form
 *   orc_voxel_vels_to_particles R: constant field; B: exact-zero voxels take no weight (Q21)
 *   orc_advect_particles       R: Euler closed form
 *   orc_move_solid_particles_only  parity unpinned: the reference's relocation
 *                                 draws from curand XORWOW (particles.cu:681,
 *                                 735-737), replaced by Philox4x32-10 here
 *   orc_move_solid_particles   G2P part as above; relocation unpinned
 *   orc_initialize, orc_solve_frame
 *                              composition of the pinned stages in the order of
 *                              particles.cu:784-812; G: fixtures stop drift
 *   orc_ns_tiled_index, orc_ns_bin  N: bijection and hand cases, clamping
 *   orc_ns_p2g                 N: uniform field, staggered hat weights, partition
 *                                 of unity, face point, wall-truncated weights
 *   orc_ns_prepare_faces       N: closed form with solids
 *   orc_ns_assemble, orc_ns_apply_A, orc_ns_pcg_jacobi
 *                              N: scipy sparse operator from the cell types,
 *                                 numpy MAC divergence, scipy spsolve
 *   orc_ns_velocity_update     N: linear pressure closed form, free surface
 *   orc_ns_g2p_flip            N: linear field exact; FLIP 0 / 1 / 0.95 blends
 *   orc_ns_advect_rk2          N: uniform field, midpoint closed form, clamp, solids
 *   orc_ns_transfer_step, orc_ns_step
 *                              N: equal to the stages run one by one
 *   orc_find_sub_cell          L: exact lattice geometry (integer distances),
 *                                 closed-form count at a sub-cell centre
 *   orc_sum_particles_per_node, orc_particle_list_create
 *                              L: numpy bincount / unique / lexsort of the
 *                                 (node, voxel, particle) incidences
 *   (L = tests/test_oracle_lists.py.)  Every entry point also runs under
 *   AddressSanitizer and UndefinedBehaviorSanitizer (sanitize_check.c,
 *   tests/test_oracle_sanitized.py).
 *
 * Build: oracle/Makefile -> oracle/liboracle.so (gcc -O2 -ffp-contract=off).
 * FMA contraction is disabled so every product/sum is an individually rounded
 * IEEE operation; where the reference calls fma() explicitly, so does this.
 */
#ifndef FLIP2_ORACLE_H
#define FLIP2_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and types as the reference's Grid (engine/grid.hpp:35-37). */
typedef struct orc_grid {
    float negX, negY, negZ;
    uint32_t sizeX, sizeY, sizeZ;
    float cellSize;
} orc_grid;

/* ------------------------------------------------------------------------ */
/* Primitives                                                               */
/* ------------------------------------------------------------------------ */

/* rootCell (engine/algorithms/particleToGridFunctions.cu:19-45): clamp
 * positions into the box (written back) and compute the u32 cell key. */
void orc_find_grid_cell(double* px, double* py, double* pz, uint32_t n,
                        orc_grid g, uint32_t* keys);

/* Inclusive u32 scan mod 2^32 (parallelPrefixSumKernels.cu:5-77), and the
 * exclusive form used by the per-node helper (parallelPrefixSumKernels.hu:16-48). */
void orc_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, int exclusive);

/* Stable LSD sort of u32 keys by 1-bit splits over bits [0, nbits)
 * (radixSortKernels.cu:6-123).  keys is sorted in place; perm[i] receives the
 * original index of the i-th sorted key. */
void orc_radix_sort_u32(uint32_t* keys, uint32_t* perm, uint32_t n, uint32_t nbits);

/* Key/value form: values travel with the keys (stable). */
void orc_radix_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t nbits);

/* getMaxFromArray + CudaVec::getMax with abs=true (engine/cudaVec.cu:173-233):
 * the signed element of largest magnitude among the elements the reference's
 * tree samples (indices 0 and 256 of each 512-tile, recursively). */
double orc_sampled_absmax(const double* a, uint32_t n);

/* Philox4x32-10 (Salmon et al., SC'11), used for the relocation jitter in
 * moveSolids and for the seeded synthetic scenes. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ------------------------------------------------------------------------ */
/* Reference-compatible particle system (class Particles, particles.hu)    */
/* ------------------------------------------------------------------------ */

typedef struct orc_particles {
    uint32_t n;
    double *px, *py, *pz, *vx, *vy, *vz;
    uint32_t* gridCell;                 /* keys, sorted in place */
    uint32_t* reorderedGridIndices;     /* sort permutation */
    uint32_t* uniqueGridNodeIndices;    /* inclusive scan of head flags */
    uint32_t numUsedGridNodes;          /* B */
    uint32_t* gridNodeIndicesToFirstParticleIndex; /* [B] */
    uint32_t yDimRows;                  /* allocated rows of the row table */
    uint32_t* yDimNumUsedGridNodes;     /* row table */
    uint32_t* nodeIndexUsedVoxels;      /* [B] inclusive voxel offsets */
    uint32_t numUsedVoxels;             /* V */
    uint32_t* voxelIDsUsed;             /* [V] */
    int8_t* solids;                     /* [V] */
    double *voxelsUx, *voxelsUy, *voxelsUz, *divU, *p, *residuals; /* [V] */
    orc_grid grid;
    double radius;          /* 2 */
    uint32_t refinementLevel; /* 1 */
    uint32_t numVoxels1D;   /* 8 */
    uint32_t numVoxelsPerNode; /* 512 */
    double frameDt, dt, elapsedTime, elapsedTimeThisFrame;
    /* statistics of the last pressure solve */
    uint32_t lastSorIterations;   /* total red/black iterations, all attempts */
    uint32_t lastSolveAttempts;   /* number of SOR runs (1 + dt halvings) */
    double lastResidual;
    uint32_t outOfTileEvents;     /* stencil voxels outside the 8^3 tile (reference printf) */
    uint64_t moveSolidsCounter;   /* Philox counter word for moveSolids */
    uint32_t relocatedParticles;
    double sorThreshold;          /* 0.01 in the reference (particles.cu:509) */
    uint32_t sorMaxIterations;    /* 16384 in the reference (particles.cu:510) */
} orc_particles;

orc_particles* orc_particles_create(uint32_t n);
void orc_particles_destroy(orc_particles* s);
void orc_set_domain(orc_particles* s, double nx, double ny, double nz,
                    uint32_t x, uint32_t y, uint32_t z, double cellSize);

void orc_align_particles_to_grid(orc_particles* s);     /* particles.cu:48-50   */
void orc_sort_particles(orc_particles* s);              /* particles.cu:52-78   */
void orc_generate_voxels(orc_particles* s);             /* particles.cu:323-368 */
void orc_particle_vel_to_voxels(orc_particles* s);      /* particles.cu:500-505 */
double orc_get_courant_dt(orc_particles* s);            /* particles.cu:493-498 */
void orc_apply_gravity(orc_particles* s, double dt);    /* voxelSolveFunctions.cu:177-180 */
void orc_remove_gravity(orc_particles* s, double dt);   /* voxelSolveFunctions.cu:182-185 */
void orc_calc_div_u(orc_particles* s);                  /* voxelSolveFunctions.cu:245-271 */
/* cudaGSiteration (voxelSolveFunctions.cu:391-412); returns the terminating
 * sampled residual, *iters receives the number of red/black iterations. */
double orc_gs_solve(orc_particles* s, double dt, double threshold, uint32_t maxIterations,
                    uint32_t* iters);
/* Run exactly `count` red/black iterations (no convergence logic). */
void orc_gs_iterations(orc_particles* s, double dt, uint32_t count);
void orc_pressure_solve(orc_particles* s);              /* particles.cu:507-555 */
void orc_update_voxel_velocities(orc_particles* s);     /* particles.cu:557-560 */
void orc_voxel_vels_to_particles(orc_particles* s);     /* particles.cu:742-749 */
void orc_advect_particles(orc_particles* s);            /* particles.cu:751-760 */
void orc_move_solid_particles_only(orc_particles* s);   /* particles.cu:674-740 (kernel only) */
void orc_move_solid_particles(orc_particles* s);        /* particles.cu:762-769 */
void orc_initialize(orc_particles* s);                  /* particles.cu:805-812 */
/* The dt-halving retry control of pressureSolve (particles.cu:528-549) on
 * caller-supplied solve and halve steps. */
typedef double (*orc_solve_cb)(void* ctx, double dt);
typedef void (*orc_halve_cb)(void* ctx, double oldDt, double newDt);
uint32_t orc_pressure_retry(double* dt, double threshold, orc_solve_cb solve, orc_halve_cb halve, void* ctx,
                            double* last);
void orc_solve_frame(orc_particles* s, double fps);     /* particles.cu:784-803 */

/* Sub-cell lists (cudaFindSubCell, particleToGridFunctions.cu:69-262) and
 * per-voxel particle lists (perVoxelParticleListFunctions.cu:6-189): code
 * the reference builds but never calls from its step.  touched* receive the
 * inclusive scans of the per-particle counts; list* may be NULL (counts
 * only).  Per-voxel lists skip entries outside [0, numVoxelsPerNode). */
void orc_find_sub_cell(const double* px, const double* py, const double* pz, uint32_t n, orc_grid g,
                       const uint32_t* gridPosition, uint32_t refinementLevel, double radius, uint32_t* touchedX,
                       uint32_t* touchedY, uint32_t* touchedZ, uint32_t* listX, uint32_t* listY, uint32_t* listZ);
void orc_sum_particles_per_node(uint32_t B, uint32_t n, const uint32_t* firstParticle, const uint32_t* touchedIncl,
                                const uint32_t* subCellsDim, uint32_t numVoxelsPerNode, uint32_t* numNonZeroVoxels,
                                uint32_t* numParticlesInVoxelLists);
void orc_particle_list_create(uint32_t B, uint32_t n, uint32_t numVoxelsPerNode, const uint32_t* numUsedVoxelsPerNode,
                              const uint32_t* firstParticleListIndexPerNode, const uint32_t* touchedIncl,
                              const uint32_t* subCellsDim, const uint32_t* firstParticle, uint32_t* voxelIDs,
                              uint32_t* listStart, uint32_t* particleLists);

/* ------------------------------------------------------------------------ */
/* North-star method (ns_oracle.c)                                          */
/* ------------------------------------------------------------------------ */
uint32_t orc_ns_tiled_index(uint32_t i, uint32_t j, uint32_t k, orc_grid g);
void orc_ns_bin(const float* x, const float* y, const float* z, uint32_t n, orc_grid g, uint32_t* keys);
void orc_ns_p2g(const float* x, const float* y, const float* z, const float* vx, const float* vy, const float* vz,
                uint32_t n, orc_grid g, double* U, double* V, double* W, double* WU, double* WV, double* WW);

void orc_ns_transfer_step(const float* x, const float* y, const float* z, const float* vx, const float* vy,
                          const float* vz, uint32_t n, orc_grid g, uint32_t keyBits, uint32_t* keys, uint32_t* perm,
                          float* sorted, uint32_t* counts, uint32_t* cellStart, double* U, double* V, double* W);

void orc_ns_prepare_faces(double* U, double* V, double* W, double* U0, double* V0, double* W0, const uint8_t* solid,
                          orc_grid g, double gdt);
void orc_ns_assemble(const uint32_t* counts, const uint8_t* solid, orc_grid g, double hOverDt, const double* U,
                     const double* V, const double* W, uint8_t* type, uint16_t* coef, double* b);
void orc_ns_apply_A(const uint16_t* coef, orc_grid g, const double* x, double* y);
uint32_t orc_ns_pcg_jacobi(const uint16_t* coef, orc_grid g, const double* b, double* x, double tol, uint32_t maxIt,
                           double* relOut);
void orc_ns_velocity_update(const uint8_t* type, const uint8_t* solid, orc_grid g, const double* p, double dtOverH,
                            double* U, double* V, double* W);
void orc_ns_g2p_flip(uint32_t n, const float* x, const float* y, const float* z, double* vx, double* vy, double* vz,
                     const double* U, const double* V, const double* W, const double* U0, const double* V0,
                     const double* W0, orc_grid g, double flip);
void orc_ns_advect_rk2(uint32_t n, double* x, double* y, double* z, const double* U, const double* V, const double* W,
                       const uint8_t* solid, orc_grid g, double dt, double margin);

/* One whole serial step of the north-star method (transfer, walls and
 * gravity, assembly, FP64 Jacobi-PCG to tol, pressure gradient, FLIP G2P,
 * RK2); the FP32 particle arrays are replaced by the sorted, updated state.
 * Returns the PCG iteration count; *relOut the final relative residual. */
typedef struct orc_ns_work orc_ns_work;
orc_ns_work* orc_ns_work_create(uint32_t n, orc_grid g);
void orc_ns_work_destroy(orc_ns_work* w);
uint32_t orc_ns_step(orc_ns_work* w, float* x, float* y, float* z, float* vx, float* vy, float* vz, orc_grid g,
                     const uint8_t* solid, double dt, double gravity, double flip, double tol, uint32_t maxIt,
                     double margin, double* relOut);

/* Resize voxel arrays to V (used by stage tests that inject state). */
void orc_resize_voxels(orc_particles* s, uint32_t V);
void orc_resize_nodes(orc_particles* s, uint32_t B);

#ifdef __cplusplus
}
#endif
#endif
