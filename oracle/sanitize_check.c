/* Runs every entry point of the oracle on small seeded inputs so that a
 * build with -fsanitize=address,undefined (oracle/Makefile, target
 * sanitize) reports any out-of-bounds access, use after free, leak or
 * undefined arithmetic.  Test infrastructure only, like the oracle itself.
 * Exit status 0: every call returned and the few invariants below held. */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;
static uint32_t rnd(void) {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (uint32_t)(rng_state >> 16);
}
static double urand(void) { return (rnd() & 0xffffff) / 16777216.0; }

#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            exit(1);                                                     \
        }                                                                \
    } while (0)

static void primitives(void) {
    const uint32_t sizes[] = {1, 255, 256, 257, 513, 4099};
    for (size_t t = 0; t < sizeof(sizes) / sizeof(sizes[0]); ++t) {
        const uint32_t n = sizes[t];
        uint32_t* a = malloc(n * sizeof(uint32_t));
        uint32_t* b = malloc(n * sizeof(uint32_t));
        uint32_t* perm = malloc(n * sizeof(uint32_t));
        double* d = malloc(n * sizeof(double));
        for (uint32_t i = 0; i < n; ++i) {
            a[i] = rnd() % 1000u;
            d[i] = urand() - 0.5;
        }
        orc_scan_u32(a, b, n, 0);
        orc_scan_u32(a, b, n, 1);
        CHECK(b[0] == 0);
        orc_radix_sort_u32(a, perm, n, 10);
        for (uint32_t i = 1; i < n; ++i) CHECK(a[i - 1] <= a[i]);
        for (uint32_t i = 0; i < n; ++i) b[i] = i;
        orc_radix_sort_pairs_u32(a, b, n, 32);
        (void)orc_sampled_absmax(d, n);
        free(a);
        free(b);
        free(perm);
        free(d);
    }
    uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6}, out[4];
    orc_philox4x32_10(ctr, key, out);
}

static void reference_compatible(uint32_t cells, uint32_t n, int frames) {
    orc_particles* s = orc_particles_create(n);
    CHECK(s != NULL);
    orc_set_domain(s, 0.0, 0.0, 0.0, cells, cells, cells, 1.0 / cells);
    for (uint32_t i = 0; i < n; ++i) {
        s->px[i] = 0.05 + 0.4 * urand();
        s->py[i] = 0.02 + 0.5 * urand();
        s->pz[i] = 0.05 + 0.9 * urand();
        s->vx[i] = 0.2 * (urand() - 0.5);
        s->vy[i] = 0.0;
        s->vz[i] = 0.1 * (urand() - 0.5);
    }
    s->px[0] = -1.0;  /* clamped at the low faces */
    s->py[1] = 2.0;   /* clamped at the high faces: the key aliases into the next row */
    orc_initialize(s);
    orc_particle_vel_to_voxels(s);
    (void)orc_get_courant_dt(s);
    orc_apply_gravity(s, 1e-3);
    orc_calc_div_u(s);
    uint32_t its = 0;
    (void)orc_gs_solve(s, 1e-3, 0.01, 3000, &its);
    orc_remove_gravity(s, 1e-3);
    orc_update_voxel_velocities(s);
    orc_voxel_vels_to_particles(s);
    orc_advect_particles(s);
    orc_move_solid_particles(s);
    s->sorMaxIterations = 2048;
    for (int f = 0; f < frames; ++f) orc_solve_frame(s, 240.0);
    for (uint32_t i = 0; i < n; ++i) CHECK(isfinite(s->px[i]) && isfinite(s->vy[i]));
    orc_particles_destroy(s);
}

static void north_star(uint32_t cells, uint32_t n) {
    orc_grid g = {0.0f, 0.0f, 0.0f, cells, cells, cells, 1.0f / cells};
    const size_t nc = (size_t)cells * cells * cells;
    float* f = calloc(6 * (size_t)n, sizeof(float));
    float *x = f, *y = f + n, *z = f + 2 * n, *vx = f + 3 * n, *vy = f + 4 * n, *vz = f + 5 * n;
    for (uint32_t i = 0; i < n; ++i) {
        x[i] = (float)(0.05 + 0.4 * urand());
        y[i] = (float)(0.05 + 0.5 * urand());
        z[i] = (float)(0.05 + 0.9 * urand());
        vx[i] = (float)(urand() - 0.5);
        vy[i] = 0.0f;
        vz[i] = (float)(urand() - 0.5);
    }
    uint32_t* keys = malloc(n * sizeof(uint32_t));
    orc_ns_bin(x, y, z, n, g, keys);
    uint32_t* counts = calloc(nc + 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) counts[keys[i]]++;
    double* F = calloc(12 * nc, sizeof(double));
    double *U = F, *V = F + nc, *W = F + 2 * nc, *U0 = F + 3 * nc, *V0 = F + 4 * nc, *W0 = F + 5 * nc;
    orc_ns_p2g(x, y, z, vx, vy, vz, n, g, U, V, W, F + 6 * nc, F + 7 * nc, F + 8 * nc);
    uint8_t* solid = calloc(nc, 1);
    solid[orc_ns_tiled_index(cells / 2, 1, cells / 2, g)] = 1;
    orc_ns_prepare_faces(U, V, W, U0, V0, W0, solid, g, -9.81 / 240.0);
    uint8_t* type = malloc(nc);
    uint16_t* coef = malloc(nc * sizeof(uint16_t));
    double* b = F + 9 * nc;
    double* p = F + 10 * nc;
    orc_ns_assemble(counts, solid, g, (1.0 / cells) * 240.0, U, V, W, type, coef, b);
    orc_ns_apply_A(coef, g, b, F + 11 * nc);
    double rel = 0.0;
    (void)orc_ns_pcg_jacobi(coef, g, b, p, 1e-8, 5000, &rel);
    orc_ns_velocity_update(type, solid, g, p, (1.0 / 240.0) * cells, U, V, W);
    double* vel = malloc(3 * n * sizeof(double));
    for (uint32_t i = 0; i < n; ++i) {
        vel[i] = vx[i];
        vel[n + i] = vy[i];
        vel[2 * n + i] = vz[i];
    }
    orc_ns_g2p_flip(n, x, y, z, vel, vel + n, vel + 2 * n, U, V, W, U0, V0, W0, g, 0.95);
    double* pos = malloc(3 * n * sizeof(double));
    for (uint32_t i = 0; i < n; ++i) {
        pos[i] = x[i];
        pos[n + i] = y[i];
        pos[2 * n + i] = z[i];
    }
    orc_ns_advect_rk2(n, pos, pos + n, pos + 2 * n, U, V, W, solid, g, 1.0 / 240.0, 1e-3 / cells);
    /* transfer step and whole steps */
    uint32_t* perm = malloc(n * sizeof(uint32_t));
    float* sorted = malloc(6 * n * sizeof(float));
    uint32_t* start = malloc((nc + 1) * sizeof(uint32_t));
    uint32_t bits = 0;
    while ((1ull << bits) < nc) ++bits;
    orc_ns_transfer_step(x, y, z, vx, vy, vz, n, g, bits, keys, perm, sorted, counts, start, U, V, W);
    orc_ns_work* w = orc_ns_work_create(n, g);
    CHECK(w != NULL);
    for (int k = 0; k < 3; ++k)
        (void)orc_ns_step(w, x, y, z, vx, vy, vz, g, solid, 1.0 / 240.0, -9.81, 0.95, 1e-6, 5000, 1e-3 / cells,
                          &rel);
    orc_ns_work_destroy(w);
    for (uint32_t i = 0; i < n; ++i) CHECK(isfinite(x[i]) && isfinite(vy[i]));
    free(f);
    free(keys);
    free(counts);
    free(F);
    free(solid);
    free(type);
    free(coef);
    free(vel);
    free(pos);
    free(perm);
    free(sorted);
    free(start);
}

static double scripted[8];
static int scripted_next;
static double scripted_solve(void* ctx, double dt) {
    (void)ctx;
    (void)dt;
    return scripted[scripted_next++];
}
static void scripted_halve(void* ctx, double oldDt, double newDt) {
    (void)ctx;
    CHECK(newDt == oldDt / 2.0);
}

int main(void) {
    primitives();
    reference_compatible(8, 900, 1);
    reference_compatible(16, 6000, 2);
    north_star(16, 4000);
    north_star(24, 20000);
    const double seq[] = {0.5, 0.3, 0.2, 0.2};
    memcpy(scripted, seq, sizeof(seq));
    scripted_next = 0;
    double dt = 1.0;
    const uint32_t attempts = orc_pressure_retry(&dt, 0.01, scripted_solve, scripted_halve, NULL, NULL);
    CHECK(attempts == 4 && dt == 0.125);
    printf("oracle sanitize check ok\n");
    return 0;
}
