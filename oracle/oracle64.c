/*
 * oracle64 — TEST INFRASTRUCTURE ONLY (the oracle).  Nothing in the product links,
 * imports or calls this file.
 *
 * A plain FP64 D2Q9 BGK lattice Boltzmann solver, written independently of the
 * reference's pass structure: generic direction tables, the textbook
 * equilibrium f_k^eq = w_k rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u), and a pull
 * stream with half-way bounce-back.  It reproduces the reference step's
 * boundary semantics (lbm-wgpu, rewritten_shaders/stream/{e_w,n_s,ne_sw,se_nw}_stream.wgsl:31-62):
 *   - a cell is streamed only if it is fluid, x != 0 and y < Y-1;
 *   - column x = 0 holds the inflow equilibrium feq(rho=1, u=(u_in, 0));
 *   - a population whose upstream cell is solid (or lies outside the grid)
 *     is the cell's own opposite population (half-way bounce-back);
 *   - boundary mode 0 ("reference"): the upstream of x = X-1 for the
 *     west-moving populations is taken on the flattened index, i.e. column 0
 *     of the next row (stream/e_w_stream.wgsl:21-23, ne_sw_stream.wgsl:20-22,
 *     se_nw_stream.wgsl:21-23);
 *   - boundary mode 1 ("zero-gradient"): the upstream x is clamped to X-1,
 *     a zero-gradient outflow.
 * Solid cells are not evolved (their values are never read by fluid cells).
 *
 * Memory rows grow with the row index; the "north" populations move to row-1
 * (direction table: lbm-wgpu src/lbm.rs:632-640).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static const int CX[9] = {-1, 0, 1, -1, 0, 1, -1, 0, 1};
static const int CY[9] = {-1, -1, -1, 0, 0, 0, 1, 1, 1};
static const double WK[9] = {1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 9, 4.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 9, 1.0 / 36};

typedef struct {
    int X, Y, mode;
    double omega, u_in;
    double *f[9];  /* pre-collision state */
    double *g[9];  /* post-collision scratch */
    uint8_t *solid;
    uint64_t steps;
} o64;

static void feq(double rho, double ux, double uy, double out[9]) {
    double uu = ux * ux + uy * uy;
    for (int k = 0; k < 9; ++k) {
        /* physical velocity in memory coordinates: c = (CX, CY), u = (ux, uy_mem) */
        double cu = CX[k] * ux + CY[k] * uy;
        out[k] = WK[k] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
    }
}

void o64_free(o64 *h) {
    if (!h) return;
    for (int k = 0; k < 9; ++k) { free(h->f[k]); free(h->g[k]); }
    free(h->solid);
    free(h);
}

o64 *o64_new(int X, int Y, double omega, double u_in, int mode) {
    if (X < 3 || Y < 3) return NULL;
    o64 *h = (o64 *)calloc(1, sizeof(o64));
    h->X = X; h->Y = Y; h->omega = omega; h->u_in = u_in; h->mode = mode;
    size_t n = (size_t)X * Y;
    for (int k = 0; k < 9; ++k) {
        h->f[k] = (double *)malloc(n * sizeof(double));
        h->g[k] = (double *)malloc(n * sizeof(double));
    }
    h->solid = (uint8_t *)calloc(n, 1);
    for (int x = 0; x < X; ++x) { h->solid[x] = 1; h->solid[(size_t)(Y - 1) * X + x] = 1; }
    double e[9];
    feq(1.0, u_in, 0.0, e);
    for (int k = 0; k < 9; ++k)
        for (size_t i = 0; i < n; ++i) h->f[k][i] = e[k];
    return h;
}

void o64_set_omega(o64 *h, double omega) { h->omega = omega; }

void o64_set_solid(o64 *h, const uint8_t *solid) { memcpy(h->solid, solid, (size_t)h->X * h->Y); }

/* uniform equilibrium everywhere; uy is north-positive like the reference's uy field */
void o64_set_equilibrium(o64 *h, double rho, double ux, double uy) {
    double e[9];
    feq(rho, ux, -uy, e);
    size_t n = (size_t)h->X * h->Y;
    for (int k = 0; k < 9; ++k)
        for (size_t i = 0; i < n; ++i) h->f[k][i] = e[k];
    h->steps = 0;
}

/* state as 9 dense planes */
void o64_set_state(o64 *h, const double *f9) {
    size_t n = (size_t)h->X * h->Y;
    for (int k = 0; k < 9; ++k) memcpy(h->f[k], f9 + k * n, n * sizeof(double));
}

void o64_get_state(const o64 *h, double *f9) {
    size_t n = (size_t)h->X * h->Y;
    for (int k = 0; k < 9; ++k) memcpy(f9 + k * n, h->f[k], n * sizeof(double));
}

static void collide_all(o64 *h) {
    const int X = h->X, Y = h->Y;
    const double om = h->omega;
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < Y; ++y)
        for (int x = 0; x < X; ++x) {
            size_t i = (size_t)y * X + x;
            double fl[9], rho = 0, jx = 0, jy = 0;
            for (int k = 0; k < 9; ++k) {
                fl[k] = h->f[k][i];
                rho += fl[k];
                jx += CX[k] * fl[k];
                jy += CY[k] * fl[k];
            }
            double e[9];
            feq(rho, jx / rho, jy / rho, e);
            for (int k = 0; k < 9; ++k) h->g[k][i] = fl[k] + om * (e[k] - fl[k]);
        }
}

static void stream_all(o64 *h) {
    const int X = h->X, Y = h->Y;
    double inflow[9];
    feq(1.0, h->u_in, 0.0, inflow);
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < Y; ++y)
        for (int x = 0; x < X; ++x) {
            size_t i = (size_t)y * X + x;
            if (x == 0) {
                for (int k = 0; k < 9; ++k) h->f[k][i] = inflow[k];
                continue;
            }
            if (h->solid[i] || y >= Y - 1) continue;
            for (int k = 0; k < 9; ++k) {
                if (k == 4) { h->f[4][i] = h->g[4][i]; continue; }
                int ux = x - CX[k], uy = y - CY[k];
                if (ux >= X) {
                    if (h->mode == 0) { ux = 0; uy += 1; }
                    else ux = X - 1;
                }
                int bounce = (uy < 0 || uy >= Y);
                size_t u = 0;
                if (!bounce) {
                    u = (size_t)uy * X + ux;
                    bounce = h->solid[u];
                }
                h->f[k][i] = bounce ? h->g[8 - k][i] : h->g[k][u];
            }
        }
}

void o64_step(o64 *h, int n) {
    for (int s = 0; s < n; ++s) {
        collide_all(h);
        stream_all(h);
        h->steps++;
    }
}

/* rho and velocity u = j / rho; uy is north-positive (the reference's uy field,
 * rewritten_shaders/pre_collision/{corner,cardinal}_pre_collision.wgsl: uy = N - S + corner terms) */
void o64_macro(const o64 *h, double *rho, double *ux, double *uy) {
    size_t n = (size_t)h->X * h->Y;
    for (size_t i = 0; i < n; ++i) {
        double r = 0, jx = 0, jy = 0;
        for (int k = 0; k < 9; ++k) {
            double v = h->f[k][i];
            r += v; jx += CX[k] * v; jy += CY[k] * v;
        }
        rho[i] = r; ux[i] = jx / r; uy[i] = -jy / r;
    }
}
