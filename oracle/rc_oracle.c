/*
 * Serial FP64 restatement of the reference's per-step particle pipeline.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -ffp-contract=off.
 *
 * Every function cites the reference file:line whose semantics it restates
 * (paths relative to the reference repository root).
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* small helpers                                                            */
/* ------------------------------------------------------------------------ */

static void* xcalloc(size_t n, size_t sz) {
    void* p = calloc(n ? n : 1, sz);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

/* double -> u32 conversion with the saturating semantics of a GPU
 * round-toward-zero convert: NaN and negatives give 0, values >= 2^32 give
 * 0xffffffff.  The reference converts floor(double) to uint
 * (particles.cu:144-146 and the three other copies of that code). */
static uint32_t sat_u32(double v) {
    if (!(v > 0.0)) return 0u;
    if (v >= 4294967296.0) return 0xffffffffu;
    return (uint32_t)v;
}

/* distance() (particles.cu:84-86): sqrt(fma(dx,dx,fma(dy,dy,dz*dz))). */
static double ref_distance(double dx, double dy, double dz) {
    return sqrt(fma(dx, dx, fma(dy, dy, dz * dz)));
}

/* weightFromDistance (particles.cu:370-377). */
static double ref_weight(double d, double r) {
    if (d > r) return 0.0;
    return 1.0 - d / r;
}

/* calcSubCellWidth (voxelSolveFunctions.hu:10-12): float division by
 * (2 << refinementLevel), widened to double. */
static double sub_cell_width(const orc_grid* g, uint32_t refinementLevel) {
    float w = g->cellSize / (float)(2u << refinementLevel);
    return (double)w;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10                                                            */
/* ------------------------------------------------------------------------ */

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        if (r != 9) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------ */
/* Primitives                                                               */
/* ------------------------------------------------------------------------ */

/* rootCell (particleToGridFunctions.cu:19-45).  The upper bound
 * negX + sizeX*cellSize is an FP32 expression in the reference; it is
 * evaluated here as one fused multiply-add, which is how an FMA-contracting
 * compiler evaluates it (exact for the power-of-two grids used here either
 * way).  The quotient is a double that is narrowed to float by floorf. */
void orc_find_grid_cell(double* px, double* py, double* pz, uint32_t n,
                        orc_grid g, uint32_t* keys) {
    const float upX = fmaf((float)g.sizeX, g.cellSize, g.negX);
    const float upY = fmaf((float)g.sizeY, g.cellSize, g.negY);
    const float upZ = fmaf((float)g.sizeZ, g.cellSize, g.negZ);
    for (uint32_t i = 0; i < n; ++i) {
        if (px[i] < (double)g.negX) px[i] = g.negX;
        else if (px[i] > (double)upX) px[i] = upX;
        if (py[i] < (double)g.negY) py[i] = g.negY;
        else if (py[i] > (double)upY) py[i] = upY;
        if (pz[i] < (double)g.negZ) pz[i] = g.negZ;
        else if (pz[i] > (double)upZ) pz[i] = upZ;
        uint32_t x = (uint32_t)floorf((float)((px[i] - (double)g.negX) / (double)g.cellSize));
        uint32_t y = (uint32_t)floorf((float)((py[i] - (double)g.negY) / (double)g.cellSize));
        uint32_t z = (uint32_t)floorf((float)((pz[i] - (double)g.negZ) / (double)g.cellSize));
        keys[i] = x + y * g.sizeX + z * g.sizeX * g.sizeY;
    }
}

void orc_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, int exclusive) {
    uint32_t acc = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t v = in[i];
        if (exclusive) { out[i] = acc; acc += v; }
        else { acc += v; out[i] = acc; }
    }
}

/* cudaRadixSortUint (radixSortKernels.cu:85-123): for each bit, front/back
 * flags (:6-22), their inclusive scans, the split into a gather index
 * (coalesceFrontBack :25-56), the gather of keys (:59-68) and the
 * composition of the running permutation (:106-115). */
void orc_radix_sort_u32(uint32_t* keys, uint32_t* perm, uint32_t n, uint32_t nbits) {
    uint32_t* idx = (uint32_t*)xcalloc(n, sizeof(uint32_t));
    uint32_t* tmpk = (uint32_t*)xcalloc(n, sizeof(uint32_t));
    uint32_t* tmpp = (uint32_t*)xcalloc(n, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    for (uint32_t b = 0; b < nbits; ++b) {
        uint32_t zeros = 0;
        for (uint32_t i = 0; i < n; ++i) zeros += ((keys[i] >> b) & 1u) ^ 1u;
        uint32_t f = 0, k = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (((keys[i] >> b) & 1u) == 0u) idx[f++] = i;
            else idx[zeros + k++] = i;
        }
        for (uint32_t i = 0; i < n; ++i) { tmpk[i] = keys[idx[i]]; tmpp[i] = perm[idx[i]]; }
        memcpy(keys, tmpk, (size_t)n * sizeof(uint32_t));
        memcpy(perm, tmpp, (size_t)n * sizeof(uint32_t));
    }
    free(idx); free(tmpk); free(tmpp);
}

void orc_radix_sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t nbits) {
    uint32_t* perm = (uint32_t*)xcalloc(n, sizeof(uint32_t));
    uint32_t* tmp = (uint32_t*)xcalloc(n, sizeof(uint32_t));
    orc_radix_sort_u32(keys, perm, n, nbits);
    for (uint32_t i = 0; i < n; ++i) tmp[i] = vals[perm[i]];
    memcpy(vals, tmp, (size_t)n * sizeof(uint32_t));
    free(perm); free(tmp);
}

/* getMaxFromArray with absolute=true (cudaVec.cu:173-203): each 512-tile
 * reduces to (|a[0]| > |a[256]|) ? a[0] : a[256], out-of-range entries read
 * as 0.  CudaVec::getMax (cudaVec.cu:205-233) repeats that over the tile
 * results until one value remains. */
static void sampled_level(const double* in, uint32_t n, double* out) {
    uint32_t tiles = n / 512u + 1u;
    for (uint32_t t = 0; t < tiles; ++t) {
        uint64_t i0 = (uint64_t)t * 512u, i1 = i0 + 256u;
        double a = (i0 < n) ? in[i0] : 0.0;
        double b = (i1 < n) ? in[i1] : 0.0;
        out[t] = (fabs(a) > fabs(b)) ? a : b;
    }
}

double orc_sampled_absmax(const double* a, uint32_t n) {
    uint32_t k = n / 512u + 1u;
    double* buf = (double*)xcalloc(k, sizeof(double));
    double* buf2 = (double*)xcalloc(k, sizeof(double));
    sampled_level(a, n, buf);
    for (uint32_t i = k; i > 1; i = i / 512u + 1u) {
        sampled_level(buf, i, buf2);
        double* t = buf; buf = buf2; buf2 = t;
    }
    double r = buf[0];
    free(buf); free(buf2);
    return r;
}

/* ------------------------------------------------------------------------ */
/* Particle system                                                          */
/* ------------------------------------------------------------------------ */

orc_particles* orc_particles_create(uint32_t n) {
    orc_particles* s = (orc_particles*)xcalloc(1, sizeof(orc_particles));
    s->n = n;
    s->px = (double*)xcalloc(n, sizeof(double));
    s->py = (double*)xcalloc(n, sizeof(double));
    s->pz = (double*)xcalloc(n, sizeof(double));
    s->vx = (double*)xcalloc(n, sizeof(double));
    s->vy = (double*)xcalloc(n, sizeof(double));
    s->vz = (double*)xcalloc(n, sizeof(double));
    s->gridCell = (uint32_t*)xcalloc(n, sizeof(uint32_t));
    s->reorderedGridIndices = (uint32_t*)xcalloc(n, sizeof(uint32_t));
    s->uniqueGridNodeIndices = (uint32_t*)xcalloc(n, sizeof(uint32_t));
    /* Particles::Particles (particles.cu:15-40) */
    s->radius = 2.0;
    s->refinementLevel = 1;
    s->numVoxels1D = 2u * (uint32_t)floor(s->radius) + (2u << s->refinementLevel);
    s->numVoxelsPerNode = s->numVoxels1D * s->numVoxels1D * s->numVoxels1D;
    s->frameDt = (double)(1.0f / 24.0f);
    s->elapsedTime = 0.0; /* left uninitialised by the reference (particles.hu:87) */
    s->elapsedTimeThisFrame = 0.0;
    s->sorThreshold = 0.01;        /* particles.cu:509 */
    s->sorMaxIterations = 16384;   /* particles.cu:510 */
    return s;
}

static void free_voxels(orc_particles* s) {
    free(s->voxelIDsUsed); free(s->solids);
    free(s->voxelsUx); free(s->voxelsUy); free(s->voxelsUz);
    free(s->divU); free(s->p); free(s->residuals);
    s->voxelIDsUsed = NULL; s->solids = NULL; s->voxelsUx = s->voxelsUy = s->voxelsUz = NULL;
    s->divU = s->p = s->residuals = NULL;
}

void orc_particles_destroy(orc_particles* s) {
    if (!s) return;
    free(s->px); free(s->py); free(s->pz); free(s->vx); free(s->vy); free(s->vz);
    free(s->gridCell); free(s->reorderedGridIndices); free(s->uniqueGridNodeIndices);
    free(s->gridNodeIndicesToFirstParticleIndex); free(s->yDimNumUsedGridNodes);
    free(s->nodeIndexUsedVoxels);
    free_voxels(s);
    free(s);
}

/* Number of rows the row table needs so that every row index the reference
 * computes (key / sizeX, including keys of particles clamped onto the upper
 * faces) is in bounds.  The reference sizes it y*z (particles.cu:45). */
static uint32_t row_table_rows(const orc_grid* g) {
    uint64_t maxKey = (uint64_t)g->sizeX + (uint64_t)g->sizeY * g->sizeX +
                      (uint64_t)g->sizeZ * g->sizeX * g->sizeY;
    uint64_t rows = maxKey / (g->sizeX ? g->sizeX : 1) + 1;
    uint64_t yz = (uint64_t)g->sizeY * g->sizeZ;
    uint64_t xy = (uint64_t)g->sizeX * g->sizeY;
    if (yz > rows) rows = yz;
    if (xy > rows) rows = xy;
    return (uint32_t)rows;
}

/* Particles::setDomain (particles.cu:42-46) -> Grid setters (grid.hpp:22-33):
 * corner and cell size are narrowed to float. */
void orc_set_domain(orc_particles* s, double nx, double ny, double nz,
                    uint32_t x, uint32_t y, uint32_t z, double cellSize) {
    s->grid.negX = (float)nx; s->grid.negY = (float)ny; s->grid.negZ = (float)nz;
    s->grid.sizeX = x; s->grid.sizeY = y; s->grid.sizeZ = z;
    s->grid.cellSize = (float)cellSize;
    free(s->yDimNumUsedGridNodes);
    s->yDimRows = row_table_rows(&s->grid);
    s->yDimNumUsedGridNodes = (uint32_t*)xcalloc(s->yDimRows, sizeof(uint32_t));
}

void orc_align_particles_to_grid(orc_particles* s) {
    orc_find_grid_cell(s->px, s->py, s->pz, s->n, s->grid, s->gridCell);
}

/* Particles::sortParticles (particles.cu:52-78): sort keys with the
 * permutation, gather px/py/pz only (velocities keep their order). */
void orc_sort_particles(orc_particles* s) {
    uint32_t n = s->n;
    orc_radix_sort_u32(s->gridCell, s->reorderedGridIndices, n, 32);
    double* t = (double*)xcalloc(n, sizeof(double));
    double* arrs[3] = {s->px, s->py, s->pz};
    for (int a = 0; a < 3; ++a) {
        for (uint32_t i = 0; i < n; ++i) t[i] = arrs[a][s->reorderedGridIndices[i]];
        memcpy(arrs[a], t, (size_t)n * sizeof(double));
    }
    free(t);
}

void orc_resize_nodes(orc_particles* s, uint32_t B) {
    free(s->gridNodeIndicesToFirstParticleIndex);
    free(s->nodeIndexUsedVoxels);
    s->numUsedGridNodes = B;
    s->gridNodeIndicesToFirstParticleIndex = (uint32_t*)xcalloc(B, sizeof(uint32_t));
    s->nodeIndexUsedVoxels = (uint32_t*)xcalloc(B, sizeof(uint32_t));
}

void orc_resize_voxels(orc_particles* s, uint32_t V) {
    free_voxels(s);
    s->numUsedVoxels = V;
    s->voxelIDsUsed = (uint32_t*)xcalloc(V, sizeof(uint32_t));
    s->solids = (int8_t*)xcalloc(V, sizeof(int8_t));
    s->voxelsUx = (double*)xcalloc(V, sizeof(double));
    s->voxelsUy = (double*)xcalloc(V, sizeof(double));
    s->voxelsUz = (double*)xcalloc(V, sizeof(double));
    s->divU = (double*)xcalloc(V, sizeof(double));
    s->p = (double*)xcalloc(V, sizeof(double));
    s->residuals = (double*)xcalloc(V, sizeof(double));
}

/* Per-particle sub-voxel frame shared by every per-node kernel
 * (particles.cu:133-149, 250-267, 423-440, 601-618, 711-724). */
typedef struct {
    double pin[3];    /* position relative to the node's apron corner */
    uint32_t sub[3];  /* sub-voxel index (saturating double->uint) */
    uint32_t g[3];    /* node coordinates decoded from the particle's key */
} particle_frame;

static void make_frame(const orc_particles* s, uint32_t idx, particle_frame* f) {
    const orc_grid* g = &s->grid;
    uint32_t key = s->gridCell[idx];
    uint32_t xy = g->sizeX * g->sizeY;
    uint32_t m = key % xy;
    f->g[2] = key / xy;
    f->g[1] = m / g->sizeX;
    f->g[0] = m % g->sizeX;
    uint32_t apron = (uint32_t)floor(s->radius);
    double scw = sub_cell_width(g, s->refinementLevel);
    const double pos[3] = {s->px[idx], s->py[idx], s->pz[idx]};
    const float neg[3] = {g->negX, g->negY, g->negZ};
    for (int d = 0; d < 3; ++d) {
        float cellOff = (float)f->g[d] * g->cellSize;           /* uint*float: FP32 */
        f->pin[d] = ((pos[d] - (double)neg[d]) - (double)cellOff) + (double)apron * scw;
        f->sub[d] = sat_u32(floor(f->pin[d] / scw));
    }
}

/* Wall marking of getNumberVoxelsUsedPerNode / generateVoxelIDs
 * (particles.cu:97-130, 203-247).  The node coordinates come from
 * gridPosition[blockIdx.x] (a particle index used as a node index) and the
 * "y" voxel index evaluates to the x index ((i % n*n)/n == i % n). */
static void mark_walls(const orc_particles* s, uint32_t b, uint8_t* flag, uint8_t* solid) {
    const orc_grid* g = &s->grid;
    uint32_t n1 = s->numVoxels1D;
    uint32_t xy = g->sizeX * g->sizeY;
    uint32_t gridID = s->gridCell[b];
    uint32_t gx = gridID % g->sizeX;
    uint32_t gy = (gridID % xy) / g->sizeX;
    uint32_t gz = gridID / xy;
    double fr = floor(s->radius);
    for (uint32_t i = 0; i < s->numVoxelsPerNode; ++i) {
        uint32_t vx = i % n1;
        uint32_t vy = (i % n1 * n1) / n1;
        uint32_t vz = i / (n1 * n1);
        uint32_t gvx = gx * n1 + vx, gvy = gy * n1 + vy, gvz = gz * n1 + vz;
        if ((double)gvx < fr || (double)gvy < fr || (double)gvz < fr) {
            flag[i] = 1; solid[i] = 1;
        } else if ((double)gvx > (double)(g->sizeX * n1) + fr ||
                   (double)gvy > (double)(g->sizeY * n1) + fr ||
                   (double)gvz > (double)(g->sizeZ * n1) + fr) {
            flag[i] = 1; solid[i] = 1;
        }
    }
}

/* Particle marking of used voxels (particles.cu:132-178, 249-295): every
 * voxel of the 4^3 stencil around the particle's sub-voxel is used when one
 * of three half-voxel-staggered FP64 distances is below radius*subCellWidth.
 * A sub-voxel index below the apron count makes the loop start negative,
 * which against the unsigned bound skips the particle entirely. */
static void mark_particles(orc_particles* s, uint32_t b, uint8_t* flag) {
    uint32_t n1 = s->numVoxels1D;
    uint32_t first = s->gridNodeIndicesToFirstParticleIndex[b];
    uint32_t last = (b == s->numUsedGridNodes - 1) ? s->n : s->gridNodeIndicesToFirstParticleIndex[b + 1];
    uint32_t apron = (uint32_t)floor(s->radius);
    double scw = sub_cell_width(&s->grid, s->refinementLevel);
    double hs = scw / 2.0;
    double rs = s->radius * scw;
    for (uint32_t idx = first; idx < last; ++idx) {
        particle_frame f;
        make_frame(s, idx, &f);
        if (f.sub[0] < apron || f.sub[1] < apron || f.sub[2] < apron) continue;
        for (uint32_t x = f.sub[0] - apron; x < f.sub[0] + apron; ++x)
            for (uint32_t y = f.sub[1] - apron; y < f.sub[1] + apron; ++y)
                for (uint32_t z = f.sub[2] - apron; z < f.sub[2] + apron; ++z) {
                    double dpx = f.pin[0] - (double)(int32_t)x * scw;
                    double dpy = f.pin[1] - (double)(int32_t)y * scw;
                    double dpz = f.pin[2] - (double)(int32_t)z * scw;
                    double xc = ref_distance(dpx, dpy + hs, dpz + hs);
                    double yc = ref_distance(dpx + hs, dpy, dpz + hs);
                    double zc = ref_distance(dpx + hs, dpy + hs, dpz);
                    if (xc < rs || yc < rs || zc < rs) {
                        if (x < n1 && y < n1 && z < n1) flag[x + y * n1 + z * n1 * n1] = 1;
                        else s->outOfTileEvents++;
                    }
                }
    }
}

/* Particles::generateVoxels (particles.cu:323-368) with
 * cudaMarkUniqueGridCellsAndCount (particleToGridFunctions.cu:303-327),
 * cudaMapNodeIndicesToParticles (:329-345), cudaGetFirstNodeInYRows
 * (:347-378), getNumberVoxelsUsedPerNode (particles.cu:88-190),
 * cudaParallelPrefixSum over the counts and generateVoxelIDs (:192-321). */
void orc_generate_voxels(orc_particles* s) {
    uint32_t n = s->n;
    const orc_grid* g = &s->grid;
    /* head flags + inclusive scan */
    for (uint32_t i = 0; i < n; ++i)
        s->uniqueGridNodeIndices[i] = (i == 0 || s->gridCell[i] != s->gridCell[i - 1]) ? 1u : 0u;
    orc_scan_u32(s->uniqueGridNodeIndices, s->uniqueGridNodeIndices, n, 0);
    uint32_t B = n ? s->uniqueGridNodeIndices[n - 1] : 0;
    orc_resize_nodes(s, B);
    for (uint32_t i = 0; i < n; ++i) {
        if (i == 0) s->gridNodeIndicesToFirstParticleIndex[0] = 0;
        else if (s->uniqueGridNodeIndices[i] != s->uniqueGridNodeIndices[i - 1])
            s->gridNodeIndicesToFirstParticleIndex[s->uniqueGridNodeIndices[i] - 1] = i;
    }
    /* row table: sentinel B over x*y entries (the reference's zeroYDimArray
     * extent), here over every allocated row */
    for (uint32_t r = 0; r < s->yDimRows; ++r) s->yDimNumUsedGridNodes[r] = B;
    for (uint32_t b = 0; b < B; ++b) {
        uint32_t prev = (b != 0) ? s->gridCell[s->gridNodeIndicesToFirstParticleIndex[b - 1]] : B;
        uint32_t cur = s->gridCell[s->gridNodeIndicesToFirstParticleIndex[b]];
        uint32_t prevRow = prev / g->sizeX, row = cur / g->sizeX;
        if (prevRow != row) s->yDimNumUsedGridNodes[row] = cur;
    }
    /* per-node used-voxel flags */
    uint32_t nv = s->numVoxelsPerNode;
    uint8_t* flags = (uint8_t*)xcalloc((size_t)B * nv, 1);
    uint8_t* solid = (uint8_t*)xcalloc((size_t)B * nv, 1);
    for (uint32_t b = 0; b < B; ++b) {
        mark_walls(s, b, flags + (size_t)b * nv, solid + (size_t)b * nv);
        mark_particles(s, b, flags + (size_t)b * nv);
    }
    for (uint32_t b = 0; b < B; ++b) {
        uint32_t c = 0;
        for (uint32_t i = 0; i < nv; ++i) c += flags[(size_t)b * nv + i];
        s->nodeIndexUsedVoxels[b] = c;
    }
    orc_scan_u32(s->nodeIndexUsedVoxels, s->nodeIndexUsedVoxels, B, 0);
    uint32_t V = B ? s->nodeIndexUsedVoxels[B - 1] : 0;
    orc_resize_voxels(s, V); /* p is zeroed (particles.cu:352) */
    for (uint32_t b = 0; b < B; ++b) {
        uint32_t k = (b == 0) ? 0 : s->nodeIndexUsedVoxels[b - 1];
        for (uint32_t i = 0; i < nv; ++i) {
            if (flags[(size_t)b * nv + i]) {
                s->voxelIDsUsed[k] = i;
                s->solids[k] = (int8_t)solid[(size_t)b * nv + i];
                ++k;
            }
        }
    }
    free(flags); free(solid);
}

static uint32_t node_voxel_start(const orc_particles* s, uint32_t b) {
    return b == 0 ? 0 : s->nodeIndexUsedVoxels[b - 1];
}

/* gatherParticleVelsToVoxels (particles.cu:379-488) for all three
 * components: hat weights to the staggered face points, node-local
 * accumulation in particle order, solids zeroed, normalised with +1e-7. */
void orc_particle_vel_to_voxels(orc_particles* s) {
    uint32_t n1 = s->numVoxels1D, nv = s->numVoxelsPerNode, B = s->numUsedGridNodes;
    uint32_t apron = (uint32_t)floor(s->radius);
    double scw = sub_cell_width(&s->grid, s->refinementLevel);
    double hs = scw / 2.0;
    double rs = s->radius * scw;
    double* U = (double*)xcalloc(3 * (size_t)nv, sizeof(double));
    double* W = (double*)xcalloc(3 * (size_t)nv, sizeof(double));
    uint8_t* sol = (uint8_t*)xcalloc(nv, 1);
    const double* vel[3] = {s->vx, s->vy, s->vz};
    double* out[3] = {s->voxelsUx, s->voxelsUy, s->voxelsUz};
    for (uint32_t b = 0; b < B; ++b) {
        memset(U, 0, 3 * (size_t)nv * sizeof(double));
        memset(W, 0, 3 * (size_t)nv * sizeof(double));
        memset(sol, 0, nv);
        uint32_t vs = node_voxel_start(s, b), ve = s->nodeIndexUsedVoxels[b];
        for (uint32_t k = vs; k < ve; ++k) sol[s->voxelIDsUsed[k]] = (uint8_t)s->solids[k];
        uint32_t first = s->gridNodeIndicesToFirstParticleIndex[b];
        uint32_t last = (b == B - 1) ? s->n : s->gridNodeIndicesToFirstParticleIndex[b + 1];
        for (uint32_t idx = first; idx < last; ++idx) {
            particle_frame f;
            make_frame(s, idx, &f);
            if (f.sub[0] < apron || f.sub[1] < apron || f.sub[2] < apron) continue;
            for (uint32_t x = f.sub[0] - apron; x < f.sub[0] + apron; ++x)
                for (uint32_t y = f.sub[1] - apron; y < f.sub[1] + apron; ++y)
                    for (uint32_t z = f.sub[2] - apron; z < f.sub[2] + apron; ++z) {
                        if (!(x < n1 && y < n1 && z < n1)) { s->outOfTileEvents++; continue; }
                        uint32_t v = x + y * n1 + z * n1 * n1;
                        for (int c = 0; c < 3; ++c) {
                            double bx = (double)(int32_t)x * scw;
                            double by = (double)(int32_t)y * scw;
                            double bz = (double)(int32_t)z * scw;
                            if (c == 0) { by += hs; bz += hs; }
                            else if (c == 1) { bx += hs; bz += hs; }
                            else { by += hs; bx += hs; }
                            double w = ref_weight(ref_distance(f.pin[0] - bx, f.pin[1] - by, f.pin[2] - bz), rs);
                            W[c * nv + v] += w;
                            U[c * nv + v] += w * vel[c][idx];
                        }
                    }
        }
        for (int c = 0; c < 3; ++c) {
            for (uint32_t i = 0; i < nv; ++i) if (sol[i]) U[c * nv + i] = 0.0;
            for (uint32_t k = vs; k < ve; ++k) {
                uint32_t id = s->voxelIDsUsed[k];
                out[c][k] = U[c * nv + id] / (W[c * nv + id] + 0.0000001);
            }
        }
    }
    free(U); free(W); free(sol);
}

/* Particles::getCourantDt (particles.cu:493-498). */
double orc_get_courant_dt(orc_particles* s) {
    double a = fabs(orc_sampled_absmax(s->vx, s->n));
    double b = fabs(orc_sampled_absmax(s->vy, s->n));
    double c = fabs(orc_sampled_absmax(s->vz, s->n));
    double maxVel = fmax(a, fmax(b, c));
    double voxelSize = (double)s->grid.cellSize / ((double)s->numVoxels1D - 2.0 * floor(s->radius));
    return 0.7 * (voxelSize / maxVel);
}

/* applyGravityKernel / removeGravityKernel (voxelSolveFunctions.cu:160-185):
 * Uy += -9.8f*dt on non-solid voxels. */
void orc_apply_gravity(orc_particles* s, double dt) {
    double gdt = (double)(-9.8f) * dt;
    for (uint32_t i = 0; i < s->numUsedVoxels; ++i) if (!s->solids[i]) s->voxelsUy[i] += gdt;
}
void orc_remove_gravity(orc_particles* s, double dt) {
    double gdt = (double)(-9.8f) * dt;
    for (uint32_t i = 0; i < s->numUsedVoxels; ++i) if (!s->solids[i]) s->voxelsUy[i] -= gdt;
}

/* getNeighboringApronCellData (voxelSolveFunctions.cu:63-121): adds the
 * neighbour node's voxels that fall into this node's apron, shifted by the
 * offset; the z-negative branch writes its z target into the x slot. */
static void add_neighbor(const orc_particles* s, uint32_t nb, double* tile, const double* data,
                         int xo, int yo, int zo) {
    uint32_t n1 = s->numVoxels1D, n2 = n1 * n1;
    uint32_t vs = node_voxel_start(s, nb), ve = s->nodeIndexUsedVoxels[nb];
    for (uint32_t k = vs; k < ve; ++k) {
        uint32_t id = s->voxelIDsUsed[k];
        uint32_t sx = id % n1, sy = (id % n2) / n1, sz = id / n2;
        uint32_t tx = sx, ty = sy, tz = sz;
        if (xo < 0) {
            if (sx >= n1 + (uint32_t)xo) { tx = sx - n1 - (uint32_t)xo; tile[tx + ty * n1 + tz * n2] += data[k]; }
        } else if (xo > 0) {
            if (sx < (uint32_t)xo) { tx = n1 - (uint32_t)xo + sx; tile[tx + ty * n1 + tz * n2] += data[k]; }
        } else if (yo < 0) {
            if (sy >= n1 + (uint32_t)yo) { ty = sy - n1 - (uint32_t)yo; tile[tx + ty * n1 + tz * n2] += data[k]; }
        } else if (yo > 0) {
            if (sy < (uint32_t)yo) { ty = n1 - (uint32_t)yo + sy; tile[tx + ty * n1 + tz * n2] += data[k]; }
        } else if (zo < 0) {
            if (sz >= n1 + (uint32_t)zo) { tz = sz - n1 - (uint32_t)zo; tile[tz + ty * n1 + tz * n2] += data[k]; }
        } else if (zo > 0) {
            if (sz < (uint32_t)zo) { tz = n1 - (uint32_t)zo + sz; tile[tx + ty * n1 + tz * n2] += data[k]; }
        }
    }
}

/* loadVoxelDataForThisBlock (voxelSolveFunctions.cu:97-158): own voxels,
 * then the x-left/x-right neighbours (adjacent in the node list) and the
 * y/z neighbours looked up through the row table, whose entries are cell
 * keys compared with, and used as, node indices.  All adds are applied in
 * phase order (the reference has no barrier between the x-right and y-left
 * phases). */
static void load_tile(const orc_particles* s, uint32_t b, double* tile, const double* data) {
    const orc_grid* g = &s->grid;
    uint32_t nv = s->numVoxelsPerNode, B = s->numUsedGridNodes;
    for (uint32_t i = 0; i < nv; ++i) tile[i] = 0.0;
    uint32_t vs = node_voxel_start(s, b), ve = s->nodeIndexUsedVoxels[b];
    for (uint32_t k = vs; k < ve; ++k) tile[s->voxelIDsUsed[k]] += data[k];
    uint32_t cur = s->gridCell[s->gridNodeIndicesToFirstParticleIndex[b]];
    uint32_t xy = g->sizeX * g->sizeY;
    uint32_t curX = cur % g->sizeX, curY = (cur % xy) / g->sizeX, curZ = cur / xy;
    int32_t xl = (int32_t)curX - 1, xr = (int32_t)curX + 1;
    int32_t yl = (int32_t)curY - 1, yr = (int32_t)curY + 1;
    int32_t zl = (int32_t)curZ - 1, zr = (int32_t)curZ + 1;
    int apronTotal = 2 * (int)(uint32_t)s->radius;
    uint32_t nb;
    if (xl >= 0 && b > 0) {
        if (s->gridCell[s->gridNodeIndicesToFirstParticleIndex[b - 1]] == cur - 1)
            add_neighbor(s, b - 1, tile, data, -apronTotal, 0, 0);
    }
    if ((uint32_t)xr < g->sizeX && b < B - 1) {
        if (s->gridCell[s->gridNodeIndicesToFirstParticleIndex[b + 1]] == cur + 1)
            add_neighbor(s, b + 1, tile, data, apronTotal, 0, 0);
    }
    if (yl >= 0 && (nb = s->yDimNumUsedGridNodes[(uint32_t)yl + curZ * g->sizeY]) < B)
        add_neighbor(s, nb, tile, data, 0, -apronTotal, 0);
    if ((uint32_t)yr < g->sizeY && (nb = s->yDimNumUsedGridNodes[(uint32_t)yr + curZ * g->sizeY]) < B)
        add_neighbor(s, nb, tile, data, 0, apronTotal, 0);
    if (zl >= 0 && (nb = s->yDimNumUsedGridNodes[(uint32_t)zl * g->sizeY + curY]) < B)
        add_neighbor(s, nb, tile, data, 0, 0, -apronTotal);
    if ((uint32_t)zr < g->sizeZ && (nb = s->yDimNumUsedGridNodes[(uint32_t)zr * g->sizeY + curY]) < B)
        add_neighbor(s, nb, tile, data, 0, 0, apronTotal);
}

/* cudaCalcDivU (voxelSolveFunctions.cu:245-271) -> calculateDivU (:188-243)
 * per component: forward difference inside the tile, accumulated into divU
 * in x, y, z order. */
void orc_calc_div_u(orc_particles* s) {
    uint32_t n1 = s->numVoxels1D, nv = s->numVoxelsPerNode, B = s->numUsedGridNodes;
    double* tile = (double*)xcalloc(nv, sizeof(double));
    double* d = (double*)xcalloc(nv, sizeof(double));
    for (uint32_t i = 0; i < s->numUsedVoxels; ++i) s->divU[i] = 0.0;
    const double* comps[3] = {s->voxelsUx, s->voxelsUy, s->voxelsUz};
    for (int c = 0; c < 3; ++c) {
        for (uint32_t b = 0; b < B; ++b) {
            load_tile(s, b, tile, comps[c]);
            for (uint32_t i = 0; i < nv; ++i) {
                d[i] = 0.0;
                if (c == 0) { if (i % n1 < n1 - 1) d[i] = tile[i + 1] - tile[i]; }
                else if (c == 1) { if (i % (n1 * n1) / n1 < n1 - 1) d[i] = tile[i + n1] - tile[i]; }
                else { if (i / (n1 * n1) < n1 - 1) d[i] = tile[i + n1 * n1] - tile[i]; }
            }
            uint32_t vs = node_voxel_start(s, b), ve = s->nodeIndexUsedVoxels[b];
            for (uint32_t k = vs; k < ve; ++k) s->divU[k] += d[s->voxelIDsUsed[k]];
        }
    }
    free(tile); free(d);
}

/* isApronCell (voxelSolveFunctions.cu:275-284): one layer, with the "y"
 * index evaluating to the x index. */
static int is_apron(uint32_t id, uint32_t n1) {
    uint32_t x = id % n1;
    uint32_t y = (id % n1 * n1) / n1;
    uint32_t z = id / (n1 * n1);
    uint32_t rb = n1 - 1u;
    return x < 1u || x >= rb || y < 1u || y >= rb || z < 1u || z >= rb;
}

/* isRed (voxelSolveFunctions.cu:286-291) */
static int is_red(uint32_t id, uint32_t n1) {
    uint32_t x = id % n1, y = (id % (n1 * n1)) / n1, z = id / (n1 * n1);
    return !((x + y + z) & 1u);
}

static double gs_scale(const orc_particles* s, double dt, double density) {
    double vs = sub_cell_width(&s->grid, s->refinementLevel);
    return dt / (density * vs * vs);
}

/* `count` red+black iterations of GSiteration (voxelSolveFunctions.cu:293-389)
 * on every node (the node tiles are independent).  Updated voxels are the
 * non-solid, non-apron ones of the half-sweep's colour; their coefficients
 * are fixed for the solve (divU and solids do not change).  Every read of a
 * half-sweep sees the values from its start (the reference reads same-colour
 * neighbours across the y-row wrap without a barrier; this fixes that race to
 * the start-of-half-sweep value).  The residual left behind is that of the
 * last (black) half-sweep; non-updated voxels hold 0. */
void orc_gs_iterations(orc_particles* s, double dt, uint32_t count) {
    uint32_t n1 = s->numVoxels1D, nv = s->numVoxelsPerNode, B = s->numUsedGridNodes;
    const double scale = gs_scale(s, dt, 0.014);
    const double w = 1.9;
    const int32_t off[6] = {-1, 1, -(int32_t)n1, (int32_t)n1, -(int32_t)(n1 * n1), (int32_t)(n1 * n1)};
    double* dv = (double*)xcalloc(nv, sizeof(double));
    double* P = (double*)xcalloc(nv, sizeof(double));
    uint8_t* sol = (uint8_t*)xcalloc(nv, 1);
    uint32_t* upd = (uint32_t*)xcalloc(nv, sizeof(uint32_t));   /* voxel ids, red first */
    double* A = (double*)xcalloc((size_t)nv * 6, sizeof(double));
    double* dg = (double*)xcalloc(nv, sizeof(double));
    double* pn = (double*)xcalloc(nv, sizeof(double));
    double* R = (double*)xcalloc(nv, sizeof(double));
    for (uint32_t b = 0; b < B; ++b) {
        uint32_t vs = node_voxel_start(s, b), ve = s->nodeIndexUsedVoxels[b];
        memset(dv, 0, nv * sizeof(double)); memset(P, 0, nv * sizeof(double));
        memset(sol, 0, nv); memset(R, 0, nv * sizeof(double));
        for (uint32_t k = vs; k < ve; ++k) {
            uint32_t id = s->voxelIDsUsed[k];
            dv[id] = s->divU[k]; P[id] = s->p[k]; sol[id] = (uint8_t)s->solids[k];
        }
        uint32_t nred = 0, nall = 0;
        for (int pass = 1; pass >= 0; --pass) {
            for (uint32_t k = vs; k < ve; ++k) {
                uint32_t id = s->voxelIDsUsed[k];
                if (pass != is_red(id, n1) || sol[id] || is_apron(id, n1)) continue;
                double ad = 0.0;
                for (int j = 0; j < 6; ++j) {
                    uint32_t nb = (uint32_t)((int32_t)id + off[j]);
                    A[(size_t)nall * 6 + j] = 0.0;
                    if (!sol[nb]) {
                        if (dv[nb] != 0.0) A[(size_t)nall * 6 + j] = -scale;
                        ad += scale;
                    }
                }
                dg[nall] = ad + 0.000000001;
                upd[nall++] = id;
            }
            if (pass == 1) nred = nall;
        }
        for (uint32_t it = 0; it < count; ++it) {
            for (int red = 1; red >= 0; --red) {
                uint32_t lo = red ? 0 : nred, hi = red ? nred : nall;
                for (uint32_t i = lo; i < hi; ++i) {
                    uint32_t id = upd[i];
                    const double* a = A + (size_t)i * 6;
                    double sum = a[0] * P[id - 1] + a[1] * P[id + 1] + a[2] * P[id - n1] + a[3] * P[id + n1] +
                                 a[4] * P[id - n1 * n1] + a[5] * P[id + n1 * n1];
                    double q = P[id] + w * ((-dv[id] - sum) / dg[i] - P[id]);
                    pn[i] = q;
                    if (!red) R[id] = (-dv[id] - dg[i] * q + sum);
                }
                for (uint32_t i = lo; i < hi; ++i) P[upd[i]] = pn[i];
            }
        }
        for (uint32_t k = vs; k < ve; ++k) {
            uint32_t id = s->voxelIDsUsed[k];
            s->p[k] = P[id];
            if (count) s->residuals[k] = R[id];
        }
    }
    free(dv); free(P); free(sol); free(upd); free(A); free(dg); free(pn); free(R);
}

/* cudaGSiteration (voxelSolveFunctions.cu:391-412): residual sampled after
 * iteration 0 and then after every 1024th; stops below the threshold, when
 * stalled, or at maxIterations. */
double orc_gs_solve(orc_particles* s, double dt, double threshold, uint32_t maxIterations,
                    uint32_t* itersOut) {
    double maxResidual = 10.0, prevMaxResidual = 100.0;
    uint32_t iterations = 0;
    const uint32_t batch = 1024;
    while (maxResidual > threshold && iterations < maxIterations) {
        /* iterations up to and including the next checked one */
        uint32_t run;
        if (iterations % batch == 0) run = 1;
        else {
            uint32_t nextCheck = (iterations / batch + 1) * batch;
            run = nextCheck - iterations + 1;
        }
        if (iterations + run > maxIterations) run = maxIterations - iterations;
        orc_gs_iterations(s, dt, run);
        uint32_t lastIt = iterations + run - 1;
        iterations += run;
        if (lastIt % batch == 0) {
            prevMaxResidual = maxResidual;
            maxResidual = orc_sampled_absmax(s->residuals, s->numUsedVoxels);
        }
        if (iterations == maxIterations ||
            (fabs(prevMaxResidual - maxResidual) < threshold / 1000 && maxResidual > threshold)) {
            break;
        }
    }
    if (itersOut) *itersOut = iterations;
    return maxResidual;
}

/* The retry loop of Particles::pressureSolve (particles.cu:528-549): solve;
 * while the terminating residual keeps decreasing and is not below the
 * threshold, halve dt (the callback undoes gravity at the old dt, clears p,
 * applies gravity at the new dt and recomputes divU) and solve again.
 * Returns the number of solves; *last receives the final residual. */
uint32_t orc_pressure_retry(double* dt, double threshold, orc_solve_cb solve, orc_halve_cb halve, void* ctx,
                            double* last) {
    double previousTerminatingResidual = 100.0, terminatingResidual;
    uint32_t attempts = 0;
    for (;;) {
        terminatingResidual = solve(ctx, *dt);
        attempts += 1;
        if (!(previousTerminatingResidual - terminatingResidual > 0)) break;
        if (terminatingResidual < threshold) break;
        const double old = *dt;
        *dt /= 2.0;
        halve(ctx, old, *dt);
        previousTerminatingResidual = terminatingResidual;
    }
    if (last) *last = terminatingResidual;
    return attempts;
}

static double rc_retry_solve(void* ctx, double dt) {
    orc_particles* s = (orc_particles*)ctx;
    uint32_t its = 0;
    const double r = orc_gs_solve(s, dt, s->sorThreshold, s->sorMaxIterations, &its);
    s->lastSorIterations += its;
    return r;
}

static void rc_retry_halve(void* ctx, double oldDt, double newDt) {
    orc_particles* s = (orc_particles*)ctx;
    orc_remove_gravity(s, oldDt);
    for (uint32_t i = 0; i < s->numUsedVoxels; ++i) s->p[i] = 0.0;
    orc_apply_gravity(s, newDt);
    orc_calc_div_u(s);
}

/* Particles::pressureSolve (particles.cu:507-555). */
void orc_pressure_solve(orc_particles* s) {
    s->lastSorIterations = 0;
    s->dt = orc_get_courant_dt(s);
    if (s->frameDt - s->elapsedTimeThisFrame < s->dt) s->dt = s->frameDt - s->elapsedTimeThisFrame;
    orc_apply_gravity(s, s->dt);
    orc_calc_div_u(s);
    double last = 0.0;
    s->lastSolveAttempts = orc_pressure_retry(&s->dt, s->sorThreshold, rc_retry_solve, rc_retry_halve, s, &last);
    s->lastResidual = last;
    s->elapsedTime += s->dt;
    s->elapsedTimeThisFrame += s->dt;
}

/* cudaVelocityUpdate -> pressureToAcceleration (voxelSolveFunctions.cu:415-486). */
void orc_update_voxel_velocities(orc_particles* s) {
    uint32_t n1 = s->numVoxels1D, nv = s->numVoxelsPerNode, B = s->numUsedGridNodes;
    double scale = gs_scale(s, s->dt, 0.014);
    double* P = (double*)xcalloc(nv, sizeof(double));
    double* Ut = (double*)xcalloc(nv, sizeof(double));
    uint8_t* sol = (uint8_t*)xcalloc(nv, 1);
    double* comps[3] = {s->voxelsUx, s->voxelsUy, s->voxelsUz};
    const uint32_t stride[3] = {1u, n1, n1 * n1};
    for (int c = 0; c < 3; ++c) {
        for (uint32_t b = 0; b < B; ++b) {
            uint32_t vs = node_voxel_start(s, b), ve = s->nodeIndexUsedVoxels[b];
            memset(P, 0, nv * sizeof(double)); memset(Ut, 0, nv * sizeof(double)); memset(sol, 0, nv);
            for (uint32_t k = vs; k < ve; ++k) {
                uint32_t id = s->voxelIDsUsed[k];
                Ut[id] = comps[c][k]; P[id] = s->p[k]; sol[id] = (uint8_t)s->solids[k];
            }
            for (uint32_t k = vs; k < ve; ++k) {
                uint32_t id = s->voxelIDsUsed[k];
                if (sol[id] || is_apron(id, n1)) continue;
                double dp = 0.0;
                if (!sol[id - stride[c]]) dp = P[id] - P[id - stride[c]];
                Ut[id] = Ut[id] - dp * scale;
            }
            for (uint32_t k = vs; k < ve; ++k) comps[c][k] = Ut[s->voxelIDsUsed[k]];
        }
    }
    free(P); free(Ut); free(sol);
}

/* gatherVoxelVelsToParticles (particles.cu:562-662): PIC gather with the P2G
 * weights, ignoring voxels whose value is exactly 0, normalised with +1e-8.
 * The loops stop at the tile edge (x < 8 in the loop condition). */
void orc_voxel_vels_to_particles(orc_particles* s) {
    uint32_t n1 = s->numVoxels1D, nv = s->numVoxelsPerNode, B = s->numUsedGridNodes;
    uint32_t apron = (uint32_t)floor(s->radius);
    double scw = sub_cell_width(&s->grid, s->refinementLevel);
    double hs = scw / 2.0;
    double rs = s->radius * scw;
    double* T = (double*)xcalloc(3 * (size_t)nv, sizeof(double));
    const double* comps[3] = {s->voxelsUx, s->voxelsUy, s->voxelsUz};
    double* vel[3] = {s->vx, s->vy, s->vz};
    for (int c = 0; c < 3; ++c) {
        for (uint32_t b = 0; b < B; ++b) {
            double* tile = T + (size_t)c * nv;
            memset(tile, 0, nv * sizeof(double));
            uint32_t vs = node_voxel_start(s, b), ve = s->nodeIndexUsedVoxels[b];
            for (uint32_t k = vs; k < ve; ++k) tile[s->voxelIDsUsed[k]] = comps[c][k];
            uint32_t first = s->gridNodeIndicesToFirstParticleIndex[b];
            uint32_t last = (b == B - 1) ? s->n : s->gridNodeIndicesToFirstParticleIndex[b + 1];
            for (uint32_t idx = first; idx < last; ++idx) {
                particle_frame f;
                make_frame(s, idx, &f);
                double pv = 0.0, ws = 0.0;
                if (!(f.sub[0] < apron || f.sub[1] < apron || f.sub[2] < apron)) {
                    for (uint32_t x = f.sub[0] - apron; x < n1 && x < f.sub[0] + apron; ++x)
                        for (uint32_t y = f.sub[1] - apron; y < n1 && y < f.sub[1] + apron; ++y)
                            for (uint32_t z = f.sub[2] - apron; z < n1 && z < f.sub[2] + apron; ++z) {
                                double bx = (double)(int32_t)x * scw;
                                double by = (double)(int32_t)y * scw;
                                double bz = (double)(int32_t)z * scw;
                                if (c == 0) { by += hs; bz += hs; }
                                else if (c == 1) { bx += hs; bz += hs; }
                                else { by += hs; bx += hs; }
                                double w = ref_weight(ref_distance(f.pin[0] - bx, f.pin[1] - by, f.pin[2] - bz), rs);
                                double u = tile[x + y * n1 + z * n1 * n1];
                                if (u != 0.0) { ws += w; pv += w * u; }
                            }
                }
                vel[c][idx] = pv / (ws + 0.00000001);
            }
        }
    }
    free(T);
}

/* advectParticlePositions (particles.cu:664-669): x += dt*v. */
void orc_advect_particles(orc_particles* s) {
    for (uint32_t i = 0; i < s->n; ++i) {
        s->px[i] += s->dt * s->vx[i];
        s->py[i] += s->dt * s->vy[i];
        s->pz[i] += s->dt * s->vz[i];
    }
}

/* Uniform in (0,1] from a Philox word (the reference draws curand_uniform). */
static double philox_unit(uint32_t r) { return ((double)r + 1.0) * (1.0 / 4294967296.0); }

/* moveSolids (particles.cu:674-740).  A particle whose sub-voxel lies
 * outside the tile or in a solid voxel moves into the node's first solid
 * voxel in ascending id order (voxel 512, i.e. (0,0,8), when the node has
 * none) plus a uniform jitter.  The jitter stream is Philox4x32-10 keyed by
 * the reference's seed and counted by (particle index, call counter);
 * the reference's curand XORWOW stream is not reproduced: parity unpinned. */
void orc_move_solid_particles_only(orc_particles* s) {
    const orc_grid* g = &s->grid;
    uint32_t n1 = s->numVoxels1D, nv = s->numVoxelsPerNode, B = s->numUsedGridNodes;
    int32_t apron = (int32_t)floor(s->radius);
    double scw = sub_cell_width(g, s->refinementLevel);
    uint8_t* sol = (uint8_t*)xcalloc(nv, 1);
    const uint32_t key[2] = {1291293829u, 0u};
    s->relocatedParticles = 0;
    for (uint32_t b = 0; b < B; ++b) {
        uint32_t vs = node_voxel_start(s, b), ve = s->nodeIndexUsedVoxels[b];
        memset(sol, 0, nv);
        for (uint32_t k = vs; k < ve; ++k) sol[s->voxelIDsUsed[k]] = (uint8_t)s->solids[k];
        uint32_t firstSolid = nv;
        for (uint32_t k = vs; k < ve; ++k) if (s->solids[k]) { firstSolid = s->voxelIDsUsed[k]; break; }
        uint32_t first = s->gridNodeIndicesToFirstParticleIndex[b];
        uint32_t last = (b == B - 1) ? s->n : s->gridNodeIndicesToFirstParticleIndex[b + 1];
        for (uint32_t idx = first; idx < last; ++idx) {
            particle_frame f;
            make_frame(s, idx, &f);
            int out = f.sub[0] >= n1 || f.sub[1] >= n1 || f.sub[2] >= n1;
            if (!out && !sol[f.sub[0] + f.sub[1] * n1 + f.sub[2] * n1 * n1]) continue;
            int32_t nx = (int32_t)(firstSolid % n1);
            int32_t ny = (int32_t)((firstSolid % (n1 * n1)) / n1);
            int32_t nz = (int32_t)(firstSolid / (n1 * n1));
            uint32_t ctr[4] = {idx, (uint32_t)s->moveSolidsCounter, (uint32_t)(s->moveSolidsCounter >> 32), 0u};
            uint32_t r[4];
            orc_philox4x32_10(ctr, key, r);
            float bxf = fmaf((float)f.g[0], g->cellSize, g->negX);
            float byf = fmaf((float)f.g[1], g->cellSize, g->negY);
            float bzf = fmaf((float)f.g[2], g->cellSize, g->negZ);
            s->px[idx] = ((double)bxf + (double)(nx - apron) * scw) + philox_unit(r[0]) * scw;
            s->py[idx] = ((double)byf + (double)(ny - apron) * scw) + philox_unit(r[1]) * scw;
            s->pz[idx] = ((double)bzf + (double)(nz - apron) * scw) + philox_unit(r[2]) * scw;
            s->relocatedParticles++;
        }
    }
    s->moveSolidsCounter++;
    free(sol);
}

/* Particles::moveSolidParticles (particles.cu:762-769): relocation, then the
 * gather again at the advected positions with the keys of the last binning. */
void orc_move_solid_particles(orc_particles* s) {
    orc_move_solid_particles_only(s);
    orc_voxel_vels_to_particles(s);
}

/* Particles::initialize (particles.cu:805-812). */
void orc_initialize(orc_particles* s) {
    orc_align_particles_to_grid(s);
    orc_sort_particles(s);
    orc_generate_voxels(s);
}

/* Particles::solveFrame (particles.cu:784-803). */
void orc_solve_frame(orc_particles* s, double fps) {
    s->frameDt = 1.0 / fps;
    s->dt = s->frameDt / 10.0;
    s->elapsedTimeThisFrame = 0.0;
    while (s->elapsedTimeThisFrame < s->frameDt) {
        orc_particle_vel_to_voxels(s);
        orc_pressure_solve(s);
        orc_update_voxel_velocities(s);
        orc_voxel_vels_to_particles(s);
        orc_advect_particles(s);
        orc_move_solid_particles(s);
        orc_initialize(s);
    }
}

/* ------------------------------------------------------------------------ */
/* Sub-cell lists and per-voxel particle lists                              */
/* ------------------------------------------------------------------------ */

/* The sub-cell geometry shared by subCellCreateNumSubCellsTouchedEachDimension
 * and subCellCreateLists (particleToGridFunctions.cu:69-139, 141-205): the
 * particle's offset inside its root cell plus the apron, in the reference's
 * mixed float/double arithmetic -- apronCells = floorf(radius); sub-cell
 * width cellSize / (2.0f * 2^level) in float; gridID * cellSize a float
 * product; the sub-cell index floorf of the double quotient narrowed to
 * float; distances through sqrtf of the double sum narrowed to float.  The
 * checks run over x in [s - apron, s + apron) with unsigned bounds (an
 * index below the apron wraps and the loop is empty).  Visits the touched
 * sub-cells of one particle in the reference's loop order (x, then y, then
 * z) and calls emit(dim, sub-cell index). */
typedef void (*orc_sub_cell_emit)(void* ctx, int dim, uint32_t subCell);

static void sub_cell_visit(double px, double py, double pz, orc_grid g, uint32_t key, uint32_t refinementLevel,
                           double radius, orc_sub_cell_emit emit, void* ctx) {
    const uint32_t xySize = g.sizeX * g.sizeY;
    const uint32_t mod = key % xySize;
    const uint32_t gz = key / xySize, gy = mod / g.sizeX, gx = mod % g.sizeX;
    const uint32_t apron = (uint32_t)floorf((float)radius);
    const double scw = (double)(g.cellSize / (2.0f * (float)(1u << refinementLevel)));
    const double ox = ((px - (double)g.negX) - (double)((float)gx * g.cellSize)) + (double)apron * scw;
    const double oy = ((py - (double)g.negY) - (double)((float)gy * g.cellSize)) + (double)apron * scw;
    const double oz = ((pz - (double)g.negZ) - (double)((float)gz * g.cellSize)) + (double)apron * scw;
    const uint32_t sx = sat_u32((double)floorf((float)(ox / scw)));
    const uint32_t sy = sat_u32((double)floorf((float)(oy / scw)));
    const uint32_t sz = sat_u32((double)floorf((float)(oz / scw)));
    const double half = scw / 2.0;
    const double rad = radius * scw;
    const uint32_t nv = 2u * apron + (2u << refinementLevel);
    for (uint32_t x = sx - apron; x < sx + apron; ++x)
        for (uint32_t y = sy - apron; y < sy + apron; ++y)
            for (uint32_t z = sz - apron; z < sz + apron; ++z) {
                const double dx = ox - (double)x * scw, dy = oy - (double)y * scw, dz = oz - (double)z * scw;
                const double xc = (double)sqrtf((float)(dx * dx + (dy + half) * (dy + half) + (dz + half) * (dz + half)));
                const double yc = (double)sqrtf((float)((dx + half) * (dx + half) + dy * dy + (dz + half) * (dz + half)));
                const double zc = (double)sqrtf((float)((dx + half) * (dx + half) + (dy + half) * (dy + half) + dz * dz));
                const uint32_t id = x + y * nv + z * nv * nv;
                if (xc < rad) emit(ctx, 0, id);
                if (yc < rad) emit(ctx, 1, id);
                if (zc < rad) emit(ctx, 2, id);
            }
}

typedef struct {
    uint32_t count[3];
    uint32_t* list[3];
    uint32_t at[3];
} sub_cell_ctx;

static void sub_cell_count(void* c, int dim, uint32_t id) {
    (void)id;
    ((sub_cell_ctx*)c)->count[dim]++;
}
static void sub_cell_write(void* c, int dim, uint32_t id) {
    sub_cell_ctx* s = (sub_cell_ctx*)c;
    s->list[dim][s->at[dim]++] = id;
}

/* cudaFindSubCell (particleToGridFunctions.cu:221-262): per-particle counts
 * of touched sub-cells per velocity component, their inclusive scans
 * (touched*), and the concatenated lists (list* may be NULL: counts only). */
void orc_find_sub_cell(const double* px, const double* py, const double* pz, uint32_t n, orc_grid g,
                       const uint32_t* gridPosition, uint32_t refinementLevel, double radius, uint32_t* touchedX,
                       uint32_t* touchedY, uint32_t* touchedZ, uint32_t* listX, uint32_t* listY, uint32_t* listZ) {
    uint32_t acc[3] = {0, 0, 0};
    uint32_t* touched[3] = {touchedX, touchedY, touchedZ};
    for (uint32_t i = 0; i < n; ++i) {
        sub_cell_ctx c = {{0, 0, 0}, {listX, listY, listZ}, {acc[0], acc[1], acc[2]}};
        if (listX && listY && listZ) sub_cell_visit(px[i], py[i], pz[i], g, gridPosition[i], refinementLevel, radius,
                                                    sub_cell_write, &c);
        else sub_cell_visit(px[i], py[i], pz[i], g, gridPosition[i], refinementLevel, radius, sub_cell_count, &c);
        for (int d = 0; d < 3; ++d) {
            acc[d] = (listX && listY && listZ) ? c.at[d] : acc[d] + c.count[d];
            touched[d][i] = acc[d];
        }
    }
}

/* Entry range of particle i in a concatenated list (inclusive scan). */
static uint32_t entry_begin(const uint32_t* incl, uint32_t i) { return i ? incl[i - 1] : 0u; }

/* sumParticlesPerNode + cudaSumParticlesPerNodeAndWriteNumUsedVoxels
 * (perVoxelParticleListFunctions.cu:6-81): per node, the number of list
 * entries of its particles (numParticlesInVoxelLists) and of distinct
 * voxels among them (numNonZeroVoxels), both then inclusive-scanned over the
 * nodes.  Entries outside [0, numVoxelsPerNode) index past the reference's
 * shared array (undefined there) and are skipped here. */
void orc_sum_particles_per_node(uint32_t B, uint32_t n, const uint32_t* firstParticle, const uint32_t* touchedIncl,
                                const uint32_t* subCellsDim, uint32_t numVoxelsPerNode, uint32_t* numNonZeroVoxels,
                                uint32_t* numParticlesInVoxelLists) {
    uint32_t* cnt = (uint32_t*)xcalloc(numVoxelsPerNode, sizeof(uint32_t));
    uint32_t accV = 0, accP = 0;
    for (uint32_t b = 0; b < B; ++b) {
        memset(cnt, 0, numVoxelsPerNode * sizeof(uint32_t));
        const uint32_t pe = b + 1 < B ? firstParticle[b + 1] : n;
        for (uint32_t i = firstParticle[b]; i < pe; ++i)
            for (uint32_t j = entry_begin(touchedIncl, i); j < touchedIncl[i]; ++j)
                if (subCellsDim[j] < numVoxelsPerNode) cnt[subCellsDim[j]]++;
        uint32_t sum = 0, nz = 0;
        for (uint32_t v = 0; v < numVoxelsPerNode; ++v) {
            sum += cnt[v];
            nz += cnt[v] != 0;
        }
        accV += nz;
        accP += sum;
        numNonZeroVoxels[b] = accV;
        numParticlesInVoxelLists[b] = accP;
    }
    free(cnt);
}

/* createParticleListStartIndices / cudaParticleListCreate
 * (perVoxelParticleListFunctions.cu:85-189): per node, the used voxels in
 * ascending voxel order (voxelIDs, node-local ids, at the node's offset
 * numUsedVoxelsPerNode[b-1]), each with the start of its particle list
 * (perVoxelParticleListStartIndices: the node's list offset
 * firstParticleListIndexPerNode[b-1] plus the exclusive count of entries
 * of lower voxels) and the list itself: the indices of the node's
 * particles that touch the voxel, in ascending particle order. */
void orc_particle_list_create(uint32_t B, uint32_t n, uint32_t numVoxelsPerNode, const uint32_t* numUsedVoxelsPerNode,
                              const uint32_t* firstParticleListIndexPerNode, const uint32_t* touchedIncl,
                              const uint32_t* subCellsDim, const uint32_t* firstParticle, uint32_t* voxelIDs,
                              uint32_t* listStart, uint32_t* particleLists) {
    uint32_t* cnt = (uint32_t*)xcalloc(numVoxelsPerNode, sizeof(uint32_t));
    uint32_t* off = (uint32_t*)xcalloc(numVoxelsPerNode, sizeof(uint32_t));
    for (uint32_t b = 0; b < B; ++b) {
        memset(cnt, 0, numVoxelsPerNode * sizeof(uint32_t));
        const uint32_t p0 = firstParticle[b], pe = b + 1 < B ? firstParticle[b + 1] : n;
        for (uint32_t i = p0; i < pe; ++i)
            for (uint32_t j = entry_begin(touchedIncl, i); j < touchedIncl[i]; ++j)
                if (subCellsDim[j] < numVoxelsPerNode) cnt[subCellsDim[j]]++;
        const uint32_t vBase = b ? numUsedVoxelsPerNode[b - 1] : 0u;
        const uint32_t lBase = b ? firstParticleListIndexPerNode[b - 1] : 0u;
        uint32_t used = 0, acc = 0;
        for (uint32_t v = 0; v < numVoxelsPerNode; ++v) {
            off[v] = acc;
            if (cnt[v]) {
                voxelIDs[vBase + used] = v;
                listStart[vBase + used] = lBase + acc;
                ++used;
            }
            acc += cnt[v];
        }
        for (uint32_t i = p0; i < pe; ++i)
            for (uint32_t j = entry_begin(touchedIncl, i); j < touchedIncl[i]; ++j) {
                const uint32_t v = subCellsDim[j];
                if (v < numVoxelsPerNode) particleLists[lBase + off[v]++] = i;
            }
    }
    free(cnt);
    free(off);
}
