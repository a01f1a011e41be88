/*
 * flip2 per-step particle pipeline on MI355X (gfx950): C ABI.
 *
 * Shipped as build/engine/libparticles.so, the path the reference's Blender
 * add-on opens (plugin/plugin.py:4-5).  All functions return 0 on success and
 * a negative status on failure; flip2_last_error() gives the message of the
 * last failure on the calling thread.  Nothing here calls exit().
 *
 * Pointers named d_* are device pointers (e.g. from flip2_malloc); all other
 * array pointers are host memory that is copied and not retained.  `stream`
 * is a hipStream_t (NULL = the default stream).
 *
 * Each declaration names the reference interface it replaces
 * (paths relative to the reference repository root).
 */
#ifndef FLIP2_H
#define FLIP2_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLIP2_OK 0
#define FLIP2_ERR_ARGUMENT (-1)
#define FLIP2_ERR_HIP (-2)
#define FLIP2_ERR_STATE (-3)
#define FLIP2_ERR_NOMEM (-4)
#define FLIP2_ERR_RANGE (-5)
#define FLIP2_ERR_INTERNAL (-6)

/* Same layout as class Grid (engine/grid.hpp:8-38). */
typedef struct flip2_grid {
    float negX, negY, negZ;
    uint32_t sizeX, sizeY, sizeZ;
    float cellSize;
} flip2_grid;

/* Message of the last failed call on this thread ("" if none).
 * Replaces gpuAssert's stderr print + exit (engine/typedefs.h:12-20). */
const char* flip2_last_error(void);
int flip2_version(void);

/* ---------------------------------------------------------------------- */
/* Device runtime (replaces engine/cudaDeviceManager.hu:26-107 and the    */
/* allocation/copy surface of engine/cudaVec.cu:20-163)                   */
/* ---------------------------------------------------------------------- */
int flip2_device_count(int* count);
/* Device name and total memory (replaces Particles::setupCudaDevices,
 * engine/particles.cu:771-782, which prints them). */
int flip2_device_info(int device, char* name, size_t nameLen, uint64_t* totalBytes);
int flip2_malloc(void** d_ptr, size_t bytes);                     /* cudaVec.cu:89-96   */
int flip2_free(void* d_ptr);                                      /* cudaVec.cu:125-134 */
int flip2_memcpy_htod(void* d_dst, const void* src, size_t bytes); /* cudaVec.cu:49-52  */
int flip2_memcpy_dtoh(void* dst, const void* d_src, size_t bytes); /* cudaVec.cu:54-60  */
int flip2_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes);
int flip2_memset(void* d_ptr, int value, size_t bytes);           /* zeroArray, cudaVec.hu:9-15 */
/* Stream-ordered device-to-device copy done by a kernel (16-byte vectors
 * when both pointers are 16-byte aligned); stream NULL = default stream. */
int flip2_copy(void* d_dst, const void* d_src, size_t bytes, void* stream);
int flip2_stream_create(void** stream);
int flip2_stream_destroy(void* stream);
int flip2_stream_synchronize(void* stream);
/* Event timer on a stream (the reference has none). */
int flip2_timer_create(void** timer);
int flip2_timer_destroy(void* timer);
int flip2_timer_start(void* timer, void* stream);
int flip2_timer_stop(void* timer, void* stream);
int flip2_timer_elapsed_ms(void* timer, float* ms);   /* waits for the stop event */

/* ---------------------------------------------------------------------- */
/* Parallel primitives on device pointers                                 */
/* ---------------------------------------------------------------------- */

/* u32 prefix scan, out-of-place or in place (d_in == d_out), mod 2^32.
 * exclusive=0 is the reference's inclusive scan.  d_total (may be NULL)
 * receives the grand total on the device.  Workspace from
 * flip2_scan_workspace_bytes; d_ws == NULL uses an internal workspace.
 * Replaces cudaParallelPrefixSum (parallelPrefixSumKernels.cu:65-77,
 * parallelPrefixSumKernels.hu:53-54). */
size_t flip2_scan_workspace_bytes(uint32_t n);
int flip2_scan_u32(const uint32_t* d_in, uint32_t* d_out, uint32_t n, int exclusive, uint32_t* d_total,
                   void* d_ws, size_t ws_bytes, void* stream);

/* Stable LSD radix sort of u32 keys with u32 values over key bits
 * [begin_bit, end_bit).  d_vals_in == NULL sorts the permutation (values
 * 0..n-1), which is the reference's reorderedGridIndices.  d_keys_alt /
 * d_vals_alt are ping-pong buffers of n elements (unused when one 8-bit pass
 * suffices).  Inputs are left unchanged and must not alias the outputs.
 * Replaces cudaRadixSortUint (radixSortKernels.cu:85-123, radixSortKernels.hu:16-17)
 * and cudaSortParticlesByGridNode (particleToGridFunctions.cu:270-301). */
size_t flip2_sort_workspace_bytes(uint32_t n, uint32_t begin_bit, uint32_t end_bit);
int flip2_sort_pairs_u32(const uint32_t* d_keys_in, const uint32_t* d_vals_in, uint32_t* d_keys_out,
                         uint32_t* d_vals_out, uint32_t* d_keys_alt, uint32_t* d_vals_alt, uint32_t n,
                         uint32_t begin_bit, uint32_t end_bit, void* d_ws, size_t ws_bytes, void* stream);

/* flip2_sort_pairs_u32 without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops that forbid hipMalloc).  The
 * workspace is a fixed device array of FLIP2_NOALLOC_SORT_SCRATCH_BYTES per
 * GPU, held until the process ends (enough for 2^25 keys over 32 bits);
 * larger requests return FLIP2_ERR_RANGE.  Same arguments and results as
 * flip2_sort_pairs_u32 otherwise (caller-provided key/value and ping-pong
 * buffers).  Call it with the stream's device current.  Calls that overlap on
 * one GPU share the array and race: order them on one stream. */
#define FLIP2_NOALLOC_SORT_SCRATCH_BYTES (34u << 20)
int flip2_sort_pairs_u32_noalloc(const uint32_t* d_keys_in, const uint32_t* d_vals_in, uint32_t* d_keys_out,
                                 uint32_t* d_vals_out, uint32_t* d_keys_alt, uint32_t* d_vals_alt, uint32_t n,
                                 uint32_t begin_bit, uint32_t end_bit, void* stream);

/* Clamp FP64 positions into the domain (written back) and compute u32 cell
 * keys x + y*sx + z*sx*sy.  Replaces cudaFindGridCell / rootCell
 * (particleToGridFunctions.cu:19-45, 216-219; particleToGridFunctions.hu:10). */
int flip2_find_grid_cell(double* d_px, double* d_py, double* d_pz, uint32_t n, flip2_grid grid, uint32_t* d_keys,
                         void* stream);

/* Sub-cell lists (cudaFindSubCell, particleToGridFunctions.cu:69-262,
 * particleToGridFunctions.hu:12): for each particle (sorted keys and clamped
 * positions of the reference-compatible step), the sub-cells of its root
 * cell plus apron whose staggered x / y / z velocity point lies within
 * `radius` sub-cells.  d_touched_x/y/z (n each) receive the inclusive scans
 * of the per-particle counts and totals[3] (host) their last elements; when
 * d_list_x/y/z are non-NULL (capacities >= totals) the concatenated lists
 * of node-local voxel ids are written too.  The reference builds these but
 * never calls them from its step. */
int flip2_find_sub_cell(const double* d_px, const double* d_py, const double* d_pz, uint32_t n, flip2_grid grid,
                        const uint32_t* d_keys, uint32_t refinement_level, double radius, uint32_t* d_touched_x,
                        uint32_t* d_touched_y, uint32_t* d_touched_z, uint32_t* d_list_x, uint32_t* d_list_y,
                        uint32_t* d_list_z, const uint32_t* capacities, uint32_t* totals, void* stream);
/* Per-voxel particle lists of one component's sub-cell list
 * (perVoxelParticleListFunctions.cu:6-189, .hu:8-16).  Per node (b <
 * num_nodes, particles [d_first[b], d_first[b+1])): d_num_nonzero_voxels and
 * d_num_entries receive the inclusive scans over the nodes of its distinct
 * voxels and list entries (cudaSumParticlesPerNodeAndWriteNumUsedVoxels);
 * flip2_particle_list_create then writes the used voxels (ascending, at the
 * node's offset), each one's list start and the lists (the node's particles
 * touching the voxel, ascending) (cudaParticleListCreate).  512 voxels per
 * node; list entries outside [0, 512) are skipped. */
int flip2_sum_particles_per_node(uint32_t num_nodes, uint32_t n, const uint32_t* d_first, const uint32_t* d_touched,
                                 const uint32_t* d_list, uint32_t* d_num_nonzero_voxels, uint32_t* d_num_entries,
                                 void* stream);
int flip2_particle_list_create(uint32_t num_nodes, uint32_t n, const uint32_t* d_num_nonzero_voxels,
                               const uint32_t* d_num_entries, const uint32_t* d_touched, const uint32_t* d_list,
                               const uint32_t* d_first, uint32_t* d_voxel_ids, uint32_t* d_list_start,
                               uint32_t* d_particle_lists, void* stream);

/* Philox4x32-10 on the device: in = count x (c0,c1,c2,c3,k0,k1), out = count x 4
 * (host arrays).  Known-answer check of the generator used by moveSolids. */
int flip2_philox4x32_10(const uint32_t* in, uint32_t count, uint32_t* out);

/* ---------------------------------------------------------------------- */
/* Reference-compatible particle system: class Particles                  */
/* (engine/particles.hu:11-95), FP64, same stage semantics                */
/* ---------------------------------------------------------------------- */
typedef struct flip2_particles flip2_particles;

typedef struct flip2_rc_stats {
    uint32_t numUsedGridNodes;
    uint32_t numUsedVoxels;
    uint32_t lastSorIterations;
    uint32_t lastSolveAttempts;
    double lastResidual;
    double dt;
    double frameDt;
    double elapsedTime;
    double elapsedTimeThisFrame;
    uint32_t outOfTileEvents;
    uint32_t relocatedParticles;
    uint32_t substeps;
} flip2_rc_stats;

int flip2_particles_create(uint32_t n, flip2_particles** out);      /* Particles::Particles, particles.cu:15-40 */
int flip2_particles_destroy(flip2_particles* p);                     /* Particles::~Particles, particles.cu:830 */
int flip2_set_domain(flip2_particles* p, double nx, double ny, double nz, uint32_t x, uint32_t y, uint32_t z,
                     double cell_size);                                /* particles.cu:42-46 */
int flip2_set_positions(flip2_particles* p, const double* px, const double* py, const double* pz);
int flip2_set_velocities(flip2_particles* p, const double* vx, const double* vy, const double* vz);
int flip2_get_positions(flip2_particles* p, double* px, double* py, double* pz);
int flip2_get_velocities(flip2_particles* p, double* vx, double* vy, double* vz);
int flip2_align_particles_to_grid(flip2_particles* p);   /* particles.cu:48-50   */
int flip2_sort_particles(flip2_particles* p);            /* particles.cu:52-78   */
int flip2_generate_voxels(flip2_particles* p);           /* particles.cu:323-368 */
int flip2_particle_vel_to_voxels(flip2_particles* p);    /* particles.cu:500-505 */
int flip2_get_courant_dt(flip2_particles* p, double* dt); /* particles.cu:493-498 */
int flip2_apply_gravity(flip2_particles* p, double dt);  /* voxelSolveFunctions.cu:177-180 */
int flip2_remove_gravity(flip2_particles* p, double dt); /* voxelSolveFunctions.cu:182-185 */
int flip2_calc_div_u(flip2_particles* p);                /* voxelSolveFunctions.cu:245-271 */
/* cudaGSiteration (voxelSolveFunctions.cu:391-412). */
int flip2_gs_solve(flip2_particles* p, double dt, double threshold, uint32_t max_iterations, double* residual,
                   uint32_t* iterations);
/* `count` red/black iterations without convergence checks. */
int flip2_gs_iterations(flip2_particles* p, double dt, uint32_t count);
int flip2_pressure_solve(flip2_particles* p);            /* particles.cu:507-555 */
int flip2_update_voxel_velocities(flip2_particles* p);   /* particles.cu:557-560 */
int flip2_voxel_vels_to_particles(flip2_particles* p);   /* particles.cu:742-749 */
int flip2_advect_particles(flip2_particles* p);          /* particles.cu:751-760 */
int flip2_move_solids(flip2_particles* p);               /* moveSolids kernel, particles.cu:674-740 */
int flip2_move_solid_particles(flip2_particles* p);      /* particles.cu:762-769 */
int flip2_initialize(flip2_particles* p);                /* particles.cu:805-812 */
int flip2_solve_frame(flip2_particles* p, double fps);   /* particles.cu:784-803 */
int flip2_write_positions_to_file(flip2_particles* p, const char* path); /* particles.cu:817-828 */
int flip2_get_stats(flip2_particles* p, flip2_rc_stats* stats);
/* Internal arrays by their reference member names (engine/particles.hu:47-68),
 * the access the friend harness engine/testing.h:27-75 has.  *count receives
 * the element count; data is copied when host != NULL and capacity suffices. */
int flip2_get_array(flip2_particles* p, const char* name, void* host, size_t capacity, size_t* count);
int flip2_set_array(flip2_particles* p, const char* name, const void* host, size_t count);
/* SOR threshold and iteration cap used by flip2_pressure_solve; the defaults
 * are the reference's constants 0.01 and 16384 (particles.cu:509-510). */
int flip2_set_solver_limits(flip2_particles* p, double threshold, uint32_t max_iterations);
int flip2_set_time(flip2_particles* p, double dt, double frame_dt, double elapsed_time,
                   double elapsed_time_this_frame);

/* ---------------------------------------------------------------------- */
/* Particle -> grid transfer on 8^3 bricks (north-star method; FP32)      */
/* ---------------------------------------------------------------------- */
/* Grids for these entry points have sizes that are multiples of 8.  Cells
 * and MAC faces are stored brick-tiled: cell (i,j,k) at
 * ((k/8*nbY + j/8)*nbX + i/8)*512 + (k%8)*64 + (j%8)*8 + i%8, u(i,j,k) on the
 * low-x face of cell (i,j,k).  The reference keys cells x + y*sx + z*sx*sy
 * (particleToGridFunctions.cu:40-43) and keeps a node list instead of bricks. */

/* Brick-tiled cell keys of FP32 positions, clamped into the grid. */
int flip2_ns_bin(const float* d_x, const float* d_y, const float* d_z, uint32_t n, flip2_grid grid, uint32_t* d_keys,
                 void* stream);
/* Trilinear MAC P2G onto bricks from particle records sorted by cell key
 * (8 floats per particle: x, y, z, vx, vy, vz, 0, 0; 16-byte aligned), with
 * d_cell_start[c] the first particle of cell c (n^3 + 1 entries).  d_wu,
 * d_wv, d_ww (weight sums) may be NULL.  Replaces the reference's per-node
 * P2G (gatherParticleVelsToVoxels, particles.cu:379-488) for the north-star
 * method. */
int flip2_ns_p2g(const float* d_records, const uint32_t* d_cell_start, flip2_grid grid, float* d_u, float* d_v,
                 float* d_w, float* d_wu, float* d_wv, float* d_ww, void* stream);

/* The whole transfer as one object: bin -> key/value sort -> gather of the
 * particle state -> per-cell counts -> exclusive scan -> P2G.  Mirrors the
 * reference's alignParticlesToGrid / sortParticles / generateVoxels /
 * particleVelToVoxels sequence (particles.cu:48-78, 323-368, 500-505). */
typedef struct flip2_p2g_pipeline flip2_p2g_pipeline;
int flip2_p2g_pipeline_create(uint32_t max_particles, flip2_grid grid, flip2_p2g_pipeline** out);
int flip2_p2g_pipeline_destroy(flip2_p2g_pipeline* pl);
int flip2_p2g_pipeline_run(flip2_p2g_pipeline* pl, uint32_t n, const float* d_x, const float* d_y, const float* d_z,
                           const float* d_vx, const float* d_vy, const float* d_vz, float* d_u, float* d_v,
                           float* d_w, void* stream);
/* Same from particle records (8 floats per particle: x, y, z, vx, vy, vz, 0,
 * 0; 16-byte aligned), the layout the FLIP step keeps its particles in. */
int flip2_p2g_pipeline_run_records(flip2_p2g_pipeline* pl, uint32_t n, const float* d_records, float* d_u, float* d_v,
                                   float* d_w, void* stream);
/* Intermediates of the last run: "keys", "sortedKeys", "perm", "counts",
 * "cellStart", "x", "y", "z", "vx", "vy", "vz" (sorted state). */
int flip2_p2g_pipeline_get(flip2_p2g_pipeline* pl, const char* name, void* host, size_t capacity, size_t* count);
/* Per-kernel HIP-event timing of the following runs (capacity = runs recorded
 * between host reads; 0 disables).  flip2_p2g_pipeline_timing returns, for
 * kernel k, its name and the summed milliseconds over *runs timed runs
 * (FLIP2_ERR_RANGE past the last kernel). */
int flip2_p2g_pipeline_enable_timing(flip2_p2g_pipeline* pl, int capacity);
int flip2_p2g_pipeline_timing(flip2_p2g_pipeline* pl, int k, char* name, size_t name_len, double* total_ms,
                              int* runs);

/* ---------------------------------------------------------------------- */
/* North-star FLIP step on 8^3 bricks: transfer -> wall faces + gravity   */
/* -> active-brick Poisson system -> mixed-precision PCG (FP32 operator   */
/* and preconditioner, FP64 dots, FP64 true residual and iterative        */
/* refinement) -> pressure-gradient update -> FLIP G2P -> RK2 advection.  */
/* The reference's step is Particles::solveFrame (particles.cu:784-803)   */
/* with a brick-local SOR; this is the configs' plain method.             */
/* ---------------------------------------------------------------------- */
#define FLIP2_PREC_JACOBI 0
#define FLIP2_PREC_MIC0 1
#define FLIP2_PREC_LEARNED 2

typedef struct flip2_fluid_config {
    flip2_grid grid;          /* sizes multiples of 8 */
    float dt;                 /* fixed step */
    float gravity;            /* added to v faces: gravity * dt */
    float flip;               /* FLIP blend (0.95) */
    int preconditioner;       /* FLIP2_PREC_* */
    double tolerance;         /* ||b - A p||_2 / ||b||_2 in FP64 */
    uint32_t max_iterations;  /* inner PCG iterations per solve */
    uint32_t check_every;     /* iterations between residual checks */
    float margin;             /* particles kept margin*h inside the box */
} flip2_fluid_config;

typedef struct flip2_solve_stats {
    uint32_t iterations;
    uint32_t refinements;
    double rel_residual;
    double b_norm;
    uint32_t active_bricks;
    uint32_t fluid_cells;
    double solve_ms;
    int converged;
    int used_fallback;
} flip2_solve_stats;

typedef struct flip2_fluid flip2_fluid;
void flip2_fluid_default_config(flip2_fluid_config* cfg);
int flip2_fluid_create(uint32_t max_particles, const flip2_fluid_config* cfg, flip2_fluid** out);
int flip2_fluid_destroy(flip2_fluid* f);
/* Replaces the configuration (the grid must not change); the learned
 * preconditioner's settings made by flip2_fluid_set_unet are kept. */
int flip2_fluid_set_config(flip2_fluid* f, const flip2_fluid_config* cfg);
/* The current configuration (after flip2_fluid_load_snapshot: the restored one). */
int flip2_fluid_get_config(flip2_fluid* f, flip2_fluid_config* cfg);
int flip2_fluid_set_particles(flip2_fluid* f, uint32_t n, const float* x, const float* y, const float* z,
                              const float* vx, const float* vy, const float* vz);
int flip2_fluid_get_particles(flip2_fluid* f, float* x, float* y, float* z, float* vx, float* vy, float* vz);
/* Number of particles the simulation holds (set by flip2_fluid_set_particles
 * or flip2_fluid_load_snapshot): the length of each flip2_fluid_get_particles
 * output array. */
int flip2_fluid_num_particles(flip2_fluid* f, uint32_t* n);
/* Static solid cells: n^3 bytes indexed x + y*nx + z*nx*ny (NULL clears). */
int flip2_fluid_set_solid(flip2_fluid* f, const uint8_t* mask);
/* A moving solid box (e.g. a wave paddle): cells whose centre lies in
 * [lo, hi) are solid in addition to the static mask, and the faces they touch
 * carry velocity vel (may be NULL: at rest).  Call before each step that
 * moves it; lo == NULL removes the box.  No counterpart in the reference
 * (its solids are the low domain walls). */
int flip2_fluid_set_solid_box(flip2_fluid* f, const float* lo, const float* hi, const float* vel);
/* Binary snapshot of the simulation state for resuming a run: grid, the
 * physical settings of the configuration, the particle records in their
 * current order, the static solid mask and the solid box.  Loading it into
 * a simulation of the same grid dimensions (and at least as many particles)
 * and stepping gives the same particles, bit for bit, as stepping the saved
 * one; the learned preconditioner's settings are not part of it.  No
 * counterpart in the reference, which writes positions only
 * (writePositionsToFile, particles.cu:817-828) and cannot resume. */
int flip2_fluid_save_snapshot(flip2_fluid* f, const char* path);
int flip2_fluid_load_snapshot(flip2_fluid* f, const char* path);
int flip2_fluid_step(flip2_fluid* f);
/* Adaptive time step: advances by frame_dt in substeps with
 * dt = min(remaining, cfl * h / max|v_c|), max|v_c| the exact maximum over
 * all particles' velocity components at the start of each substep (the
 * max_substeps-th substep takes the remaining time); *substeps (may be
 * NULL) receives their number.  The configured dt is left unchanged.
 * Replaces the reference's Courant step, dt = 0.7 h_sub / max|v| over a
 * sampled maximum (getCourantDt, particles.cu:493-498; cudaVec.cu:173-233). */
int flip2_fluid_step_frame(flip2_fluid* f, double frame_dt, double cfl, uint32_t max_substeps, uint32_t* substeps);
/* One stage: "transfer", "prepare", "system", "solve", "project", "g2p",
 * "advect", or the fused pairs the step runs, with the same results:
 * "transfer_prepare" (transfer then prepare) and "g2p_advect". */
int flip2_fluid_stage(flip2_fluid* f, const char* stage);
/* Solve the current system with the given preconditioner (the system must
 * have been built by the "system" stage). */
int flip2_fluid_solve_with(flip2_fluid* f, int preconditioner, flip2_solve_stats* stats);
int flip2_fluid_last_solve(flip2_fluid* f, flip2_solve_stats* stats);
/* "U","V","W","U0","V0","W0","counts","cellStart","activeBricks","brickSlot",
 * "cellType","coef","b","p". */
int flip2_fluid_get_array(flip2_fluid* f, const char* name, void* host, size_t capacity, size_t* count);

/* Per-kernel timing with HIP events on the simulation's stream, for the
 * following flip2_fluid_step / _step_frame / _stage / _solve_with calls:
 * each launch of a kernel whose name is in the comma-separated list
 * `kernels` (NULL: every kernel) is bracketed by two events.  Launches
 * captured into a HIP graph are not timed.  flip2_fluid_timing returns, for
 * timed kernel k (in name order), its name, the summed milliseconds and the
 * launch count (FLIP2_ERR_RANGE past the last one); enabling resets the
 * totals.  The reference has no timing (it prints host wall times only,
 * particles.cu:550). */
int flip2_fluid_enable_timing(flip2_fluid* f, const char* kernels);
/* After flip2_fluid_enable_timing: record only the first of every `every`
 * launches of each timed kernel (the events cost the stream a few
 * microseconds per recorded launch).  No counterpart in the reference. */
int flip2_fluid_set_timing_sample(flip2_fluid* f, uint32_t every);
int flip2_fluid_disable_timing(flip2_fluid* f);
int flip2_fluid_timing(flip2_fluid* f, int k, char* name, size_t name_len, double* total_ms, uint64_t* launches);

/* ---------------------------------------------------------------------- */
/* Learned preconditioner: 3-level 3D U-Net (16/32/64 channels, 3^3       */
/* convolutions, BF16 weights/activations, FP32 accumulation on MFMA) on  */
/* the dense n^3 grid of a fluid simulation; trained on the GPU with the  */
/* library's own forward/backward kernels and Adam on (b, p) pairs that   */
/* the solver's PCG solves produce; used inside flexible CG.  No          */
/* counterpart in the reference (its solve is a brick-local SOR).         */
/* ---------------------------------------------------------------------- */
typedef struct flip2_unet flip2_unet;
int flip2_unet_create(int n, flip2_unet** out);
int flip2_unet_destroy(flip2_unet* net);
int flip2_unet_num_params(flip2_unet* net, size_t* count);
int flip2_unet_init(flip2_unet* net, uint64_t seed);
/* FP32 master parameters: per conv layer W[27][cout][cin] then b[cout], in
 * network order, then the 1x1 head w[16], b[1]. */
int flip2_unet_set_params(flip2_unet* net, const float* host, size_t count);
int flip2_unet_get_params(flip2_unet* net, float* host, size_t count);
/* z = net(b) for a dense system (host arrays of n^3: b, operator
 * coefficients u16 diag | fluid-neighbour mask << 3 in [z][y][x] order). */
int flip2_unet_apply_dense(flip2_unet* net, const float* b, const uint16_t* coef, float* z);
/* Loss ||A z - b||^2 / ||b||^2 with z = net(b) and its parameter gradient
 * (host array of num_params floats, may be NULL). */
int flip2_unet_loss_grad(flip2_unet* net, const float* b, const uint16_t* coef, double* loss, float* grad);
/* The preconditioner the network defines: with omega > 0 (default 2/3) it is
 * wrapped in damped-Jacobi smoothing, z1 = w D^-1 r, y = net(r - A z1),
 * z = (z1 + y) + w D^-1 (r - A (z1 + y)); omega = 0 uses the network alone.
 * Training, apply_dense and the solver use the same definition. */
int flip2_unet_set_smoothing(flip2_unet* net, float omega);
/* Training data: the current system and solution of a fluid simulation
 * (after its "solve" stage), kept on the device. */
int flip2_unet_add_frame(flip2_unet* net, flip2_fluid* f);
/* Residual capture: every following solve of `f` (any preconditioner) adds
 * its residual r_k = b - A x_k after each listed inner iteration count k of
 * its first refinement (0 = b itself) as a training frame.  count = 0 stops
 * the capture.  The network must outlive the capture. */
int flip2_unet_capture_residuals(flip2_unet* net, flip2_fluid* f, const uint32_t* iterations, int count);
int flip2_unet_num_frames(flip2_unet* net, int* count);
int flip2_unet_clear_frames(flip2_unet* net);
/* `epochs` passes of Adam over the captured frames (frames [first, last)),
 * learning rate lr; *mean_loss = mean loss of the last epoch. */
int flip2_unet_train(flip2_unet* net, int epochs, float lr, int first, int last, uint64_t seed, double* mean_loss);
/* Mean loss over frames [first, last) without training. */
int flip2_unet_eval(flip2_unet* net, int first, int last, double* mean_loss);
/* Use the network as the FLIP2_PREC_LEARNED preconditioner of a fluid
 * simulation (NULL detaches); fallback_iterations > 0 reruns a solve with
 * Jacobi-PCG when flexible CG needs more iterations than that. */
int flip2_fluid_set_unet(flip2_fluid* f, flip2_unet* net, uint32_t fallback_iterations);
/* Number of earlier search directions (1..16, default 8) that flexible CG
 * orthogonalises each new direction against when the learned
 * preconditioner is used (1 is FCG(1)).  New in this library; the reference
 * solver is red-black SOR (engine/algorithms/voxelSolveFunctions.cu:275-412). */
int flip2_fluid_set_flex_directions(flip2_fluid* f, uint32_t directions);
/* Iterations (1..1024, default 1) between the host's residual checks in
 * flexible CG; the other solvers check every config.check_every
 * iterations.  No counterpart in the reference. */
int flip2_fluid_set_flex_check_every(flip2_fluid* f, uint32_t iterations);
/* The smoothing step around the learned correction (when the network has
 * a smoothing weight): FLIP2_SMOOTH_JACOBI (0) damped Jacobi with that
 * weight, FLIP2_SMOOTH_IC0 (1, default) a brick-local incomplete Cholesky
 * factor (t = M^-1 r, network on r - A t, then z = x + M^-1 (r - A x) with
 * x the corrected t).  No counterpart in the reference. */
#define FLIP2_SMOOTH_JACOBI 0
#define FLIP2_SMOOTH_IC0 1
int flip2_fluid_set_learned_smoother(flip2_fluid* f, int smoother);
/* IC(0) smoothing steps (1, default, or 2) before and after the network. */
int flip2_fluid_set_learned_smooth_steps(flip2_fluid* f, uint32_t steps);
/* enable != 0 (default): the learned preconditioner's kernels are captured
 * once per solve into a HIP graph and replayed; 0 launches them one by one.
 * Same results either way.  No counterpart in the reference. */
int flip2_fluid_set_learned_graph(flip2_fluid* f, int enable);

#ifdef __cplusplus
}
#endif

#endif /* FLIP2_H */
