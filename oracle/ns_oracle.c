/*
 * FP64 oracle of the north-star method (no reference counterpart):
 * brick-tiled binning and trilinear MAC P2G.  TEST INFRASTRUCTURE ONLY.
 *
 * Layout (same as the library): cell (i,j,k) of an n^3 grid, n a multiple
 * of 8, has index ((bk*nbY + bj)*nbX + bi)*512 + (k&7)*64 + (j&7)*8 + (i&7);
 * u(i,j,k) lives on the low-x face of cell (i,j,k) at (i, j+1/2, k+1/2)*h.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static uint32_t tiled(uint32_t i, uint32_t j, uint32_t k, uint32_t nbX, uint32_t nbY) {
    return (((k >> 3) * nbY + (j >> 3)) * nbX + (i >> 3)) * 512u + (k & 7u) * 64u + (j & 7u) * 8u + (i & 7u);
}

uint32_t orc_ns_tiled_index(uint32_t i, uint32_t j, uint32_t k, orc_grid g) {
    return tiled(i, j, k, g.sizeX >> 3, g.sizeY >> 3);
}

static uint32_t clamp_cell(double q, uint32_t n) {
    if (!(q > 0.0)) return 0u;
    double f = floor(q);
    return f >= (double)(n - 1) ? n - 1 : (uint32_t)f;
}

void orc_ns_bin(const float* x, const float* y, const float* z, uint32_t n, orc_grid g, uint32_t* keys) {
    double inv = 1.0 / (double)g.cellSize;
    for (uint32_t p = 0; p < n; ++p) {
        uint32_t i = clamp_cell(((double)x[p] - (double)g.negX) * inv, g.sizeX);
        uint32_t j = clamp_cell(((double)y[p] - (double)g.negY) * inv, g.sizeY);
        uint32_t k = clamp_cell(((double)z[p] - (double)g.negZ) * inv, g.sizeZ);
        keys[p] = tiled(i, j, k, g.sizeX >> 3, g.sizeY >> 3);
    }
}

/* Trilinear MAC P2G in FP64: each particle adds w = prod(1 - |q - node|) to
 * the 8 nodes of each component lattice around it; faces outside the grid
 * are dropped; u = sum(w v) / sum(w) where sum(w) > 0, else 0.  U, V, W and
 * the weight sums WU, WV, WW (may be NULL) hold n^3 doubles (tiled). */
void orc_ns_p2g(const float* x, const float* y, const float* z, const float* vx, const float* vy, const float* vz,
                uint32_t n, orc_grid g, double* U, double* V, double* W, double* WU, double* WV, double* WW) {
    uint32_t nx = g.sizeX, ny = g.sizeY, nz = g.sizeZ, nbX = nx >> 3, nbY = ny >> 3;
    size_t cells = (size_t)nx * ny * nz;
    double* acc[3] = {U, V, W};
    double* wsum[3];
    double* own[3] = {WU, WV, WW};
    for (int c = 0; c < 3; ++c) {
        memset(acc[c], 0, cells * sizeof(double));
        wsum[c] = own[c] ? own[c] : (double*)calloc(cells, sizeof(double));
        memset(wsum[c], 0, cells * sizeof(double));
    }
    double inv = 1.0 / (double)g.cellSize;
    const float* vel[3] = {vx, vy, vz};
    for (uint32_t p = 0; p < n; ++p) {
        double q[3] = {((double)x[p] - (double)g.negX) * inv, ((double)y[p] - (double)g.negY) * inv,
                       ((double)z[p] - (double)g.negZ) * inv};
        for (int c = 0; c < 3; ++c) {
            int base[3];
            double t[3];
            for (int d = 0; d < 3; ++d) {
                double v = (d == c) ? q[d] : q[d] - 0.5;
                double f = floor(v);
                base[d] = (int)f;
                t[d] = v - f;
            }
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 2; ++b)
                    for (int e = 0; e < 2; ++e) {
                        int i = base[0] + a, j = base[1] + b, k = base[2] + e;
                        if (i < 0 || j < 0 || k < 0 || i >= (int)nx || j >= (int)ny || k >= (int)nz) continue;
                        double w = (a ? t[0] : 1.0 - t[0]) * (b ? t[1] : 1.0 - t[1]) * (e ? t[2] : 1.0 - t[2]);
                        uint32_t idx = tiled((uint32_t)i, (uint32_t)j, (uint32_t)k, nbX, nbY);
                        wsum[c][idx] += w;
                        acc[c][idx] += w * (double)vel[c][p];
                    }
        }
    }
    for (int c = 0; c < 3; ++c) {
        for (size_t i = 0; i < cells; ++i) acc[c][i] = wsum[c][i] > 0.0 ? acc[c][i] / wsum[c][i] : 0.0;
        if (!own[c]) free(wsum[c]);
    }
}

/* One serial transfer step of the primitive microbench, the work the GPU
 * pipeline does: bin -> stable LSD sort of the keys over `keyBits` bits with
 * the permutation (1-bit splits as in the reference's radix sort) -> gather
 * of the six particle arrays -> per-cell counts -> exclusive scan -> P2G.
 * Scratch arrays are caller-owned; U/V/W (n^3 doubles) receive the faces. */
void orc_ns_transfer_step(const float* x, const float* y, const float* z, const float* vx, const float* vy,
                          const float* vz, uint32_t n, orc_grid g, uint32_t keyBits, uint32_t* keys, uint32_t* perm,
                          float* sorted /* 6*n */, uint32_t* counts /* cells+1 */, uint32_t* cellStart /* cells+1 */,
                          double* U, double* V, double* W) {
    uint32_t cells = g.sizeX * g.sizeY * g.sizeZ;
    orc_ns_bin(x, y, z, n, g, keys);
    orc_radix_sort_u32(keys, perm, n, keyBits);
    const float* src[6] = {x, y, z, vx, vy, vz};
    for (int c = 0; c < 6; ++c)
        for (uint32_t i = 0; i < n; ++i) sorted[(size_t)c * n + i] = src[c][perm[i]];
    memset(counts, 0, ((size_t)cells + 1) * sizeof(uint32_t));
    for (uint32_t i = 0; i < n; ++i) counts[keys[i]]++;
    orc_scan_u32(counts, cellStart, cells + 1, 1);
    orc_ns_p2g(sorted, sorted + n, sorted + 2 * (size_t)n, sorted + 3 * (size_t)n, sorted + 4 * (size_t)n,
               sorted + 5 * (size_t)n, n, g, U, V, W, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------ */
/* Projection, G2P and advection of the north-star step (FP64, dense grid). */
/* Cell types: 0 air, 1 fluid, 2 solid.  solid may be NULL.                 */
/* ------------------------------------------------------------------------ */

static int ns_solid_at(int i, int j, int k, orc_grid g, const uint8_t* solid) {
    if (i < 0 || j < 0 || k < 0 || i >= (int)g.sizeX || j >= (int)g.sizeY || k >= (int)g.sizeZ) return 1;
    return solid ? solid[tiled((uint32_t)i, (uint32_t)j, (uint32_t)k, g.sizeX >> 3, g.sizeY >> 3)] != 0 : 0;
}

static void ns_coords(uint32_t c, orc_grid g, int gc[3]) {
    uint32_t nbX = g.sizeX >> 3, nbY = g.sizeY >> 3, b = c >> 9, l = c & 511u;
    gc[0] = (int)((b % nbX) * 8u + (l & 7u));
    gc[1] = (int)(((b / nbX) % nbY) * 8u + ((l >> 3) & 7u));
    gc[2] = (int)((b / (nbX * nbY)) * 8u + (l >> 6));
}

/* Wall faces take 0, the field before gravity is saved (U0...), v += gdt on
 * non-wall faces. */
void orc_ns_prepare_faces(double* U, double* V, double* W, double* U0, double* V0, double* W0, const uint8_t* solid,
                          orc_grid g, double gdt) {
    uint32_t cells = g.sizeX * g.sizeY * g.sizeZ;
    double* F[3] = {U, V, W};
    double* F0[3] = {U0, V0, W0};
    for (uint32_t c = 0; c < cells; ++c) {
        int gc[3];
        ns_coords(c, g, gc);
        int here = ns_solid_at(gc[0], gc[1], gc[2], g, solid);
        for (int a = 0; a < 3; ++a) {
            int m[3] = {gc[0], gc[1], gc[2]};
            m[a] -= 1;
            int wall = here || ns_solid_at(m[0], m[1], m[2], g, solid);
            double v = wall ? 0.0 : F[a][c];
            F0[a][c] = v;
            if (a == 1 && !wall) v += gdt;
            F[a][c] = v;
        }
    }
}

/* Cell types from per-cell particle counts, operator coefficients
 * (diag = non-solid neighbours, mask = fluid neighbours, as in the library's
 * u16 layout diag | mask << 3) and b = -(h/dt) D for fluid cells. */
void orc_ns_assemble(const uint32_t* counts, const uint8_t* solid, orc_grid g, double hOverDt, const double* U,
                     const double* V, const double* W, uint8_t* type, uint16_t* coef, double* b) {
    uint32_t cells = g.sizeX * g.sizeY * g.sizeZ, nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    for (uint32_t c = 0; c < cells; ++c) {
        int s = solid ? solid[c] != 0 : 0;
        type[c] = s ? 2 : (counts[c] > 0 ? 1 : 0);
    }
    for (uint32_t c = 0; c < cells; ++c) {
        coef[c] = 0;
        b[c] = 0.0;
        if (type[c] != 1) continue;
        int gc[3];
        ns_coords(c, g, gc);
        uint32_t diag = 0, mask = 0;
        for (int d = 0; d < 6; ++d) {
            int n[3] = {gc[0], gc[1], gc[2]};
            n[d >> 1] += (d & 1) ? 1 : -1;
            int ty;
            if (n[0] < 0 || n[1] < 0 || n[2] < 0 || n[0] >= (int)g.sizeX || n[1] >= (int)g.sizeY || n[2] >= (int)g.sizeZ)
                ty = 2;
            else
                ty = type[tiled((uint32_t)n[0], (uint32_t)n[1], (uint32_t)n[2], nbX, nbY)];
            if (ty != 2) ++diag;
            if (ty == 1) mask |= 1u << d;
        }
        coef[c] = (uint16_t)(diag | (mask << 3));
        double uh = gc[0] + 1 < (int)g.sizeX ? U[tiled(gc[0] + 1, gc[1], gc[2], nbX, nbY)] : 0.0;
        double vh = gc[1] + 1 < (int)g.sizeY ? V[tiled(gc[0], gc[1] + 1, gc[2], nbX, nbY)] : 0.0;
        double wh = gc[2] + 1 < (int)g.sizeZ ? W[tiled(gc[0], gc[1], gc[2] + 1, nbX, nbY)] : 0.0;
        double D = (uh - U[c]) + (vh - V[c]) + (wh - W[c]);
        b[c] = -hOverDt * D;
    }
}

/* y = A x over the dense grid (x of non-fluid cells ignored). */
void orc_ns_apply_A(const uint16_t* coef, orc_grid g, const double* x, double* y) {
    uint32_t cells = g.sizeX * g.sizeY * g.sizeZ, nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    for (uint32_t c = 0; c < cells; ++c) {
        uint16_t cf = coef[c];
        if (!cf) { y[c] = 0.0; continue; }
        int gc[3];
        ns_coords(c, g, gc);
        double s = (double)(cf & 7u) * x[c];
        for (int d = 0; d < 6; ++d) {
            if (!(((cf >> 3) >> d) & 1u)) continue;
            int n[3] = {gc[0], gc[1], gc[2]};
            n[d >> 1] += (d & 1) ? 1 : -1;
            s -= x[tiled((uint32_t)n[0], (uint32_t)n[1], (uint32_t)n[2], nbX, nbY)];
        }
        y[c] = s;
    }
}

/* FP64 Jacobi-preconditioned CG to ||b - A x|| <= tol ||b||; returns the
 * iteration count (x starts at 0). */
uint32_t orc_ns_pcg_jacobi(const uint16_t* coef, orc_grid g, const double* b, double* x, double tol, uint32_t maxIt,
                           double* relOut) {
    uint32_t cells = g.sizeX * g.sizeY * g.sizeZ;
    double *r = calloc(cells, sizeof(double)), *z = calloc(cells, sizeof(double));
    double *p = calloc(cells, sizeof(double)), *q = calloc(cells, sizeof(double));
    double bb = 0.0;
    for (uint32_t c = 0; c < cells; ++c) { x[c] = 0.0; r[c] = coef[c] ? b[c] : 0.0; bb += r[c] * r[c]; }
    double bn = sqrt(bb), rz = 0.0;
    uint32_t it = 0;
    for (uint32_t c = 0; c < cells; ++c) { z[c] = coef[c] ? r[c] / (double)(coef[c] & 7u) : 0.0; p[c] = z[c]; rz += r[c] * z[c]; }
    double rel = bn > 0 ? 1.0 : 0.0;
    while (rel > tol && it < maxIt) {
        orc_ns_apply_A(coef, g, p, q);
        double pq = 0.0;
        for (uint32_t c = 0; c < cells; ++c) pq += p[c] * q[c];
        double alpha = rz / pq, rr = 0.0, rzn = 0.0;
        for (uint32_t c = 0; c < cells; ++c) {
            x[c] += alpha * p[c];
            r[c] -= alpha * q[c];
            rr += r[c] * r[c];
            z[c] = coef[c] ? r[c] / (double)(coef[c] & 7u) : 0.0;
            rzn += r[c] * z[c];
        }
        double beta = rzn / rz;
        rz = rzn;
        for (uint32_t c = 0; c < cells; ++c) p[c] = z[c] + beta * p[c];
        rel = sqrt(rr) / bn;
        ++it;
    }
    if (relOut) *relOut = rel;
    free(r); free(z); free(p); free(q);
    return it;
}

/* u -= (dt/h)(p_c - p_{c-1}) on faces between non-solid cells with a fluid
 * side (air p = 0). */
void orc_ns_velocity_update(const uint8_t* type, const uint8_t* solid, orc_grid g, const double* p, double dtOverH,
                            double* U, double* V, double* W) {
    uint32_t cells = g.sizeX * g.sizeY * g.sizeZ, nbX = g.sizeX >> 3, nbY = g.sizeY >> 3;
    double* F[3] = {U, V, W};
    (void)solid;
    for (uint32_t c = 0; c < cells; ++c) {
        int gc[3];
        ns_coords(c, g, gc);
        for (int a = 0; a < 3; ++a) {
            if (gc[a] == 0) continue;
            int m[3] = {gc[0], gc[1], gc[2]};
            m[a] -= 1;
            uint32_t cm = tiled((uint32_t)m[0], (uint32_t)m[1], (uint32_t)m[2], nbX, nbY);
            int t0 = type[c], t1 = type[cm];
            if (t0 == 2 || t1 == 2 || (t0 != 1 && t1 != 1)) continue;
            double pc = t0 == 1 ? p[c] : 0.0, pm = t1 == 1 ? p[cm] : 0.0;
            F[a][c] -= dtOverH * (pc - pm);
        }
    }
}

static double ns_interp(const double* F, int comp, const double q[3], orc_grid g, const double* F0) {
    int base[3];
    double t[3];
    for (int d = 0; d < 3; ++d) {
        double v = (d == comp) ? q[d] : q[d] - 0.5;
        double f = floor(v);
        base[d] = (int)f;
        t[d] = v - f;
    }
    double acc = 0.0;
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int e = 0; e < 2; ++e) {
                int i = base[0] + a, j = base[1] + b, k = base[2] + e;
                if (i < 0 || j < 0 || k < 0 || i >= (int)g.sizeX || j >= (int)g.sizeY || k >= (int)g.sizeZ) continue;
                double w = (a ? t[0] : 1.0 - t[0]) * (b ? t[1] : 1.0 - t[1]) * (e ? t[2] : 1.0 - t[2]);
                uint32_t idx = tiled((uint32_t)i, (uint32_t)j, (uint32_t)k, g.sizeX >> 3, g.sizeY >> 3);
                acc += w * (F0 ? F[idx] - F0[idx] : F[idx]);
            }
    return acc;
}

/* v = flip (v + interp(u - u0)) + (1 - flip) interp(u). */
void orc_ns_g2p_flip(uint32_t n, const float* x, const float* y, const float* z, double* vx, double* vy, double* vz,
                     const double* U, const double* V, const double* W, const double* U0, const double* V0,
                     const double* W0, orc_grid g, double flip) {
    double inv = 1.0 / (double)g.cellSize;
    const double* F[3] = {U, V, W};
    const double* F0[3] = {U0, V0, W0};
    double* vel[3] = {vx, vy, vz};
    for (uint32_t p = 0; p < n; ++p) {
        double q[3] = {((double)x[p] - g.negX) * inv, ((double)y[p] - g.negY) * inv, ((double)z[p] - g.negZ) * inv};
        for (int c = 0; c < 3; ++c) {
            double pic = ns_interp(F[c], c, q, g, NULL);
            double d = ns_interp(F[c], c, q, g, F0[c]);
            vel[c][p] = flip * (vel[c][p] + d) + (1.0 - flip) * pic;
        }
    }
}

/* Midpoint advection, clamped to [neg + margin, neg + size*h - margin]; a
 * particle ending in a solid cell stays where it was. */
void orc_ns_advect_rk2(uint32_t n, double* x, double* y, double* z, const double* U, const double* V, const double* W,
                       const uint8_t* solid, orc_grid g, double dt, double margin) {
    double inv = 1.0 / (double)g.cellSize;
    const double* F[3] = {U, V, W};
    const double neg[3] = {g.negX, g.negY, g.negZ};
    const uint32_t sz[3] = {g.sizeX, g.sizeY, g.sizeZ};
    for (uint32_t p = 0; p < n; ++p) {
        double x0[3] = {x[p], y[p], z[p]}, q[3], xm[3], xn[3];
        for (int d = 0; d < 3; ++d) q[d] = (x0[d] - neg[d]) * inv;
        for (int c = 0; c < 3; ++c) xm[c] = x0[c] + 0.5 * dt * ns_interp(F[c], c, q, g, NULL);
        for (int d = 0; d < 3; ++d) q[d] = (xm[d] - neg[d]) * inv;
        for (int c = 0; c < 3; ++c) {
            double v = x0[c] + dt * ns_interp(F[c], c, q, g, NULL);
            double lo = neg[c] + margin, hi = neg[c] + (double)sz[c] * g.cellSize - margin;
            xn[c] = v < lo ? lo : (v > hi ? hi : v);
        }
        if (solid) {
            int ci = (int)floor((xn[0] - neg[0]) * inv), cj = (int)floor((xn[1] - neg[1]) * inv);
            int ck = (int)floor((xn[2] - neg[2]) * inv);
            if (ns_solid_at(ci, cj, ck, g, solid)) { xn[0] = x0[0]; xn[1] = x0[1]; xn[2] = x0[2]; }
        }
        x[p] = xn[0]; y[p] = xn[1]; z[p] = xn[2];
    }
}

/* ------------------------------------------------------------------------ */
/* One whole serial step of the north-star method, in the library's stage   */
/* order: transfer (bin, 1-bit LSD sort, gather, counts, scan, P2G) ->      */
/* walls + gravity -> assembly -> FP64 Jacobi-PCG to tol -> pressure        */
/* gradient -> FLIP G2P -> RK2.  The particle arrays are replaced by their  */
/* sorted, updated state.  Scratch is owned by the orc_ns_work object.      */
/* ------------------------------------------------------------------------ */
struct orc_ns_work {
    uint32_t n, cells, keyBits;
    uint32_t *keys, *perm, *counts, *start;
    float* sorted;
    double *F[6], *b, *p, *vel[3], *pos[3];
    uint8_t* type;
    uint16_t* coef;
};

orc_ns_work* orc_ns_work_create(uint32_t n, orc_grid g) {
    orc_ns_work* w = calloc(1, sizeof(orc_ns_work));
    if (!w) return NULL;
    w->n = n;
    w->cells = g.sizeX * g.sizeY * g.sizeZ;
    w->keyBits = 1;
    while (w->keyBits < 32 && ((w->cells - 1) >> w->keyBits)) ++w->keyBits;
    size_t nn = n ? n : 1, cc = (size_t)w->cells + 1;
    w->keys = malloc(nn * 4);
    w->perm = malloc(nn * 4);
    w->sorted = malloc(6 * nn * 4);
    w->counts = malloc(cc * 4);
    w->start = malloc(cc * 4);
    for (int c = 0; c < 6; ++c) w->F[c] = malloc(cc * 8);
    w->b = malloc(cc * 8);
    w->p = malloc(cc * 8);
    w->type = malloc(cc);
    w->coef = malloc(cc * 2);
    for (int c = 0; c < 3; ++c) {
        w->vel[c] = malloc(nn * 8);
        w->pos[c] = malloc(nn * 8);
    }
    return w;
}

void orc_ns_work_destroy(orc_ns_work* w) {
    if (!w) return;
    free(w->keys); free(w->perm); free(w->sorted); free(w->counts); free(w->start);
    for (int c = 0; c < 6; ++c) free(w->F[c]);
    free(w->b); free(w->p); free(w->type); free(w->coef);
    for (int c = 0; c < 3; ++c) { free(w->vel[c]); free(w->pos[c]); }
    free(w);
}

uint32_t orc_ns_step(orc_ns_work* w, float* x, float* y, float* z, float* vx, float* vy, float* vz, orc_grid g,
                     const uint8_t* solid, double dt, double gravity, double flip, double tol, uint32_t maxIt,
                     double margin, double* relOut) {
    const uint32_t n = w->n;
    orc_ns_transfer_step(x, y, z, vx, vy, vz, n, g, w->keyBits, w->keys, w->perm, w->sorted, w->counts, w->start,
                         w->F[0], w->F[1], w->F[2]);
    orc_ns_prepare_faces(w->F[0], w->F[1], w->F[2], w->F[3], w->F[4], w->F[5], solid, g, gravity * dt);
    const double h = (double)g.cellSize;
    orc_ns_assemble(w->counts, solid, g, h / dt, w->F[0], w->F[1], w->F[2], w->type, w->coef, w->b);
    uint32_t it = orc_ns_pcg_jacobi(w->coef, g, w->b, w->p, tol, maxIt, relOut);
    orc_ns_velocity_update(w->type, solid, g, w->p, dt / h, w->F[0], w->F[1], w->F[2]);
    const float* s = w->sorted;
    for (uint32_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) w->vel[c][i] = (double)s[(size_t)(3 + c) * n + i];
    orc_ns_g2p_flip(n, s, s + n, s + 2 * (size_t)n, w->vel[0], w->vel[1], w->vel[2], w->F[0], w->F[1], w->F[2],
                    w->F[3], w->F[4], w->F[5], g, flip);
    for (uint32_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) w->pos[c][i] = (double)s[(size_t)c * n + i];
    orc_ns_advect_rk2(n, w->pos[0], w->pos[1], w->pos[2], w->F[0], w->F[1], w->F[2], solid, g, dt, margin);
    float* out[6] = {x, y, z, vx, vy, vz};
    for (uint32_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) out[c][i] = (float)w->pos[c][i];
        for (int c = 0; c < 3; ++c) out[3 + c][i] = (float)w->vel[c][i];
    }
    return it;
}
