/*
 * refcpu — TEST INFRASTRUCTURE ONLY (the oracle).  Nothing in the product links,
 * imports or calls this file; tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it to check results and as the timed CPU baseline.
 *
 * A faithful single-precision restatement of the reference lbm-wgpu simulation
 * step: the same 2 x 9 f32 structure-of-arrays planes, the same ping-pong parity,
 * the same eight passes in the same order, the same guards and the same
 * expression trees (built with -ffp-contract=off so no multiply-add is fused).
 * Each pass is one OpenMP parallel loop over all X*Y cells, mirroring one
 * dispatch of ceil(X*Y/256) workgroups.
 *
 * Reference (paths relative to the lbm-wgpu crate):
 *   plane order / pair bindings ........ src/lbm.rs:632-640, src/lbm.rs:754-775
 *   host feq (set_equil) ............... src/lbm.rs:611-643
 *   walls (init_barrier) ............... src/lbm.rs:595-605
 *   step order and parity .............. src/lbm.rs:1112-1134, src/lbm.rs:1174-1252
 *   corner / cardinal pre-collision .... rewritten_shaders/pre_collision/{corner,cardinal}_pre_collision.wgsl:16-21
 *   corner / cardinal collision ........ rewritten_shaders/collision/{corner,cardinal}_collision.wgsl:17-43
 *   four stream passes ................. rewritten_shaders/stream/{e_w,n_s,ne_sw,se_nw}_stream.wgsl:17-62
 *   summary fields ..................... rewritten_shaders/summary_stats/{curl,ux,uy,rho,speed}.wgsl
 *   barrier paint ...................... rewritten_shaders/update_barrier/barrier_draw.wgsl:11-18
 *   resets and pulses .................. src/lbm.rs:1076-1102, src/lbm.rs:1482-1515
 *
 * Two behaviours are not fixed by the reference source and are defined here:
 *   - an out-of-range mask read (the stream passes read mask[index +- offset]
 *     with u32 wrap-around, e.g. mask[X*Y] for cell (X-1, Y-2) in se_nw_stream)
 *     reads as solid (1);
 *   - an out-of-range density read in the curl pass reads as 0.
 * Barrier painting applies the (index, value) pairs in list order, so the last
 * pair for an index wins (the reference dispatches one workgroup per pair, which
 * races for duplicates).
 *
 * Parity pins: see tests/test_oracle.py (closed forms, permutation and
 * bounce-back known-answer tables, conservation, a second numpy restatement;
 * the summary pass against the closed-form curl of linear momentum fields, the
 * flattened-index east neighbour of the last column, the skip rule and both
 * clamp edges of the display density and speed).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* a target with FMA instructions is allowed only together with -ffp-contract=off
 * (REFCPU_ALLOW_FMA, set by the Makefile's native build) */
#if defined(__FP_FAST_FMAF) && !defined(REFCPU_ALLOW_FMA)
#error "refcpu must be built without FMA contraction"
#endif

/* plane indices, src/lbm.rs:632-640 */
enum { NW = 0, N_ = 1, NE = 2, W_ = 3, C_ = 4, E_ = 5, SW = 6, S_ = 7, SE = 8 };

typedef struct {
    uint32_t X, Y, total;
    float omega;
    float *set[2][9];      /* set[p][k]; set[1][C_] exists but the live rest plane is set[0][C_] */
    uint32_t *mask;        /* barrier buffer: only == 1u means solid */
    float *ux, *uy, *rho;  /* density_bg: momentum x, momentum y, density */
    float *output;         /* summary output plane (keeps stale values like the reference) */
    uint64_t compute_step;
} ref_lbm;

static int g_threads = 0;

void ref_set_threads(int n) { g_threads = n; }

int ref_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

#ifdef _OPENMP
#define PAR_FOR _Pragma("omp parallel for schedule(static) num_threads(ref_get_threads())")
#else
#define PAR_FOR
#endif

static inline uint32_t mask_at(const ref_lbm *h, uint32_t idx) {
    return idx < h->total ? h->mask[idx] : 1u;
}

/* set_equil, src/lbm.rs:611-643: returns the 9 plane values of a uniform state */
void ref_host_feq(float ux, float uy, float rho, float out[9]) {
    float ux_2 = ux * ux;
    float uy_2 = uy * uy;
    float u_dot_product = ux_2 + uy_2;
    float u_sum_sq_pos = u_dot_product + 2.0f * (ux * uy);
    float u_sum_sq_neg = u_dot_product - 2.0f * (ux * uy);
    ux *= 3.0f;
    uy *= 3.0f;
    ux_2 *= 4.5f;
    uy_2 *= 4.5f;
    u_dot_product *= 1.5f;
    u_sum_sq_neg *= 4.5f;
    u_sum_sq_pos *= 4.5f;
    float rho_ninth = rho / 9.0f;
    float rho_36th = rho / 36.0f;
    out[0] = rho_36th * (1.0f - ux + uy + u_sum_sq_neg - u_dot_product);
    out[1] = rho_ninth * (1.0f + uy + uy_2 - u_dot_product);
    out[2] = rho_36th * (1.0f + ux + uy + u_sum_sq_pos - u_dot_product);
    out[3] = rho_ninth * (1.0f - ux + ux_2 - u_dot_product);
    out[4] = 4.0f * rho_ninth * (1.0f - u_dot_product);
    out[5] = rho_ninth * (1.0f + ux + ux_2 - u_dot_product);
    out[6] = rho_36th * (1.0f - ux - uy + u_sum_sq_pos - u_dot_product);
    out[7] = rho_ninth * (1.0f - uy - uy_2 - u_dot_product);
    out[8] = rho_36th * (1.0f + ux - uy + u_sum_sq_neg - u_dot_product);
}

static void fill_sets(ref_lbm *h, const float v[9]) {
    for (int p = 0; p < 2; ++p)
        for (int k = 0; k < 9; ++k) {
            float *a = h->set[p][k];
            float x = v[k];
            PAR_FOR
            for (uint32_t i = 0; i < h->total; ++i) a[i] = x;
        }
}

/* init_barrier, src/lbm.rs:595-605 */
void ref_reset_barrier(ref_lbm *h) {
    memset(h->mask, 0, sizeof(uint32_t) * h->total);
    for (uint32_t x = 0; x < h->X; ++x) {
        h->mask[x] = 1u;
        h->mask[x + (h->Y - 1) * h->X] = 1u;
    }
}

void ref_free(ref_lbm *h) {
    if (!h) return;
    for (int p = 0; p < 2; ++p)
        for (int k = 0; k < 9; ++k) free(h->set[p][k]);
    free(h->mask);
    free(h->ux);
    free(h->uy);
    free(h->rho);
    free(h->output);
    free(h);
}

/* LBM::new, src/lbm.rs:726-1049 (state part only) */
ref_lbm *ref_new(uint32_t X, uint32_t Y, float omega) {
    if (X < 3 || Y < 3) return NULL;
    ref_lbm *h = (ref_lbm *)calloc(1, sizeof(ref_lbm));
    if (!h) return NULL;
    h->X = X;
    h->Y = Y;
    h->total = X * Y;
    h->omega = omega;
    size_t nb = sizeof(float) * (size_t)h->total;
    for (int p = 0; p < 2; ++p)
        for (int k = 0; k < 9; ++k) h->set[p][k] = (float *)malloc(nb);
    h->mask = (uint32_t *)malloc(sizeof(uint32_t) * h->total);
    h->ux = (float *)calloc(h->total, sizeof(float));
    h->uy = (float *)calloc(h->total, sizeof(float));
    h->rho = (float *)calloc(h->total, sizeof(float));
    h->output = (float *)calloc(h->total, sizeof(float));
    for (int p = 0; p < 2; ++p)
        for (int k = 0; k < 9; ++k)
            if (!h->set[p][k]) { ref_free(h); return NULL; }
    if (!h->mask || !h->ux || !h->uy || !h->rho || !h->output) { ref_free(h); return NULL; }
    float v[9];
    ref_host_feq(0.1f, 0.0f, 1.0f, v);
    fill_sets(h, v);
    ref_reset_barrier(h);
    h->compute_step = 0;
    return h;
}

void ref_set_omega(ref_lbm *h, float omega) { h->omega = omega; } /* src/lbm.rs:1358-1360 */

/* corner_pre_collision.wgsl:16-21 */
static void pass_pre_corner(ref_lbm *h, int p) {
    const float *ne = h->set[p][NE], *sw = h->set[p][SW], *nw = h->set[p][NW], *se = h->set[p][SE];
    float *ux = h->ux, *uy = h->uy, *rho = h->rho;
    PAR_FOR
    for (uint32_t i = 0; i < h->total; ++i) {
        ux[i] = ne[i] + se[i] - nw[i] - sw[i];
        uy[i] = ne[i] + nw[i] - se[i] - sw[i];
        rho[i] = ne[i] + se[i] + nw[i] + sw[i];
    }
}

/* cardinal_pre_collision.wgsl:16-21 */
static void pass_pre_cardinal(ref_lbm *h, int p) {
    const float *n = h->set[p][N_], *s = h->set[p][S_], *e = h->set[p][E_], *w = h->set[p][W_];
    float *ux = h->ux, *uy = h->uy, *rho = h->rho;
    PAR_FOR
    for (uint32_t i = 0; i < h->total; ++i) {
        ux[i] += e[i] - w[i];
        uy[i] += n[i] - s[i];
        rho[i] += e[i] + n[i] + s[i] + w[i];
    }
}

/* corner_collision.wgsl:17-42 */
static void pass_collide_corner(ref_lbm *h, int p) {
    float *ne = h->set[p][NE], *sw = h->set[p][SW], *nw = h->set[p][NW], *se = h->set[p][SE];
    const float *origin = h->set[0][C_];
    float *ux = h->ux, *uy = h->uy, *rho = h->rho;
    const float omega = h->omega;
    const float c36 = (float)(1.0 / 36.0);
    PAR_FOR
    for (uint32_t i = 0; i < h->total; ++i) {
        rho[i] = rho[i] + origin[i];
        float thisrho = rho[i];
        float thisux = ux[i] / thisrho;
        float thisuy = uy[i] / thisrho;
        float one36thrho = c36 * thisrho;
        float ux3 = 3.0f * thisux;
        float uy3 = 3.0f * thisuy;
        float ux2 = thisux * thisux;
        float uy2 = thisuy * thisuy;
        float uxuy2 = 2.0f * thisux * thisuy;
        float u2 = ux2 + uy2;
        float u215 = 1.5f * u2;
        ne[i] += omega * (one36thrho * (1.0f + ux3 + uy3 + 4.5f * (u2 + uxuy2) - u215) - ne[i]);
        se[i] += omega * (one36thrho * (1.0f + ux3 - uy3 + 4.5f * (u2 - uxuy2) - u215) - se[i]);
        nw[i] += omega * (one36thrho * (1.0f - ux3 + uy3 + 4.5f * (u2 - uxuy2) - u215) - nw[i]);
        sw[i] += omega * (one36thrho * (1.0f - ux3 - uy3 + 4.5f * (u2 + uxuy2) - u215) - sw[i]);
    }
}

/* cardinal_collision.wgsl:17-43 */
static void pass_collide_cardinal(ref_lbm *h, int p) {
    float *n = h->set[p][N_], *s = h->set[p][S_], *e = h->set[p][E_], *w = h->set[p][W_];
    float *origin = h->set[0][C_];
    const float *ux = h->ux, *uy = h->uy, *rho = h->rho;
    const float omega = h->omega;
    const float c9 = (float)(1.0 / 9.0);
    const float c49 = (float)(4.0 / 9.0);
    PAR_FOR
    for (uint32_t i = 0; i < h->total; ++i) {
        float thisrho = rho[i];
        float thisux = ux[i] / thisrho;
        float thisuy = uy[i] / thisrho;
        float one9thrho = c9 * thisrho;
        float ux3 = 3.0f * thisux;
        float uy3 = 3.0f * thisuy;
        float ux2 = thisux * thisux;
        float uy2 = thisuy * thisuy;
        float u2 = ux2 + uy2;
        float u215 = 1.5f * u2;
        origin[i] += omega * (c49 * thisrho * (1.0f - u215) - origin[i]);
        e[i] += omega * (one9thrho * (1.0f + ux3 + 4.5f * ux2 - u215) - e[i]);
        w[i] += omega * (one9thrho * (1.0f - ux3 + 4.5f * ux2 - u215) - w[i]);
        n[i] += omega * (one9thrho * (1.0f + uy3 + 4.5f * uy2 - u215) - n[i]);
        s[i] += omega * (one9thrho * (1.0f - uy3 + 4.5f * uy2 - u215) - s[i]);
    }
}

/* The four stream passes share their guard (stream/{e_w,n_s,ne_sw,se_nw}_stream.wgsl:31-45) and differ in the
 * pair of planes and the index offsets (stream/{e_w,n_s,ne_sw,se_nw}_stream.wgsl:17-23).  `a` is the plane
 * whose population arrives from index + off_a (e.g. post_w[i] <- w[i+1]), `b` the
 * one arriving from index + off_b.  Offsets are applied in u32 arithmetic. */
static void pass_stream_pair(ref_lbm *h, int p, int ka, int kb, uint32_t off_a, uint32_t off_b) {
    const uint32_t X = h->X, Y = h->Y, total = h->total;
    const float *a = h->set[p][ka], *b = h->set[p][kb];
    float *post_a = h->set[1 - p][ka], *post_b = h->set[1 - p][kb];
    PAR_FOR
    for (uint32_t index = 0; index < total; ++index) {
        if (mask_at(h, index) == 1u) continue;
        if (index > total - 1u) continue;
        if (index % X == 0u) continue;
        if (index / X >= Y - 1u) continue;
        uint32_t ia = index + off_a; /* upstream of population a */
        uint32_t ib = index + off_b; /* upstream of population b */
        if (mask_at(h, ia) == 1u) post_a[index] = b[index];
        else post_a[index] = a[ia];
        if (mask_at(h, ib) == 1u) post_b[index] = a[index];
        else post_b[index] = b[ib];
    }
}

/* LBM::collide, src/lbm.rs:1118-1125 */
void ref_collide(ref_lbm *h) {
    int p = (int)(h->compute_step % 2);
    pass_pre_corner(h, p);
    pass_pre_cardinal(h, p);
    pass_collide_corner(h, p);
    pass_collide_cardinal(h, p);
}

/* LBM::stream, src/lbm.rs:1127-1134: e_w, n_s, nw_se, ne_sw */
void ref_stream(ref_lbm *h) {
    int p = (int)(h->compute_step % 2);
    const uint32_t X = h->X;
    /* e_w_stream.wgsl:50-62: post_w[i] <- w[i+1] (bounce e[i]); post_e[i] <- e[i-1] (bounce w[i]) */
    pass_stream_pair(h, p, W_, E_, 1u, (uint32_t)-1);
    /* n_s_stream.wgsl:50-62: post_n[i] <- n[i+X]; post_s[i] <- s[i-X] */
    pass_stream_pair(h, p, N_, S_, X, 0u - X);
    /* se_nw_stream.wgsl:50-62: post_nw[i] <- nw[i+1+X]; post_se[i] <- se[i-1-X] */
    pass_stream_pair(h, p, NW, SE, 1u + X, (uint32_t)-1 - X);
    /* ne_sw_stream.wgsl:50-62: post_ne[i] <- ne[i-1+X]; post_sw[i] <- sw[i+1-X] */
    pass_stream_pair(h, p, NE, SW, X - 1u, 1u - X);
}

/* LBM::compute_step, src/lbm.rs:1112-1116, repeated as in iterate (src/lbm.rs:1065-1068) */
void ref_step(ref_lbm *h, uint32_t n) {
    for (uint32_t s = 0; s < n; ++s) {
        ref_collide(h);
        ref_stream(h);
        h->compute_step += 1;
    }
}

/* custom_speed / reset_to_equilibrium, src/lbm.rs:1076-1102; set_equil takes (ux, uy, rho) */
void ref_reset_equilibrium(ref_lbm *h, float rho, float ux, float uy) {
    float v[9];
    ref_host_feq(ux, uy, rho, v);
    fill_sets(h, v);
    h->compute_step = 0;
    pass_pre_corner(h, 0);
    pass_pre_cardinal(h, 0);
}

/* single_cell / set_single_cell, src/lbm.rs:1482-1515 (no pre-collision pass) */
int ref_single_cell(ref_lbm *h, int k) {
    float v[9];
    ref_host_feq(0.0f, 0.0f, 1.0f, v);
    fill_sets(h, v);
    const uint32_t X = h->X, Y = h->Y;
    uint32_t x = 0, y = 0;
    switch (k) {
    case 0: x = X - 2; y = Y - 2; break;
    case 1: x = 3 * X / 4; y = Y - 2; break;
    case 2: x = X / 3; y = Y - 2; break;
    case 3: x = X - 2; y = Y / 2; break;
    case 4: x = 3 * X / 4; y = Y / 2; break;
    case 5: x = X / 2; y = Y / 2; break;
    case 6: x = X - 2; y = 1; break;
    case 7: x = 3 * X / 4; y = 1; break;
    case 8: x = X / 2; y = 1; break;
    default: k = -1; break;
    }
    if (k >= 0) {
        h->set[0][k][x + y * X] = 4.0f;
        h->set[1][k][x + y * X] = 4.0f;
    }
    h->compute_step = 0;
    return 0;
}

/* draw_barrier_updates + barrier_draw.wgsl:14-17; pairs = [idx0, val0, idx1, val1, ...]
 * (get_points_vector, barrier_shapes/merge_shapes.rs:12-22).  Returns -1 on an
 * out-of-range index or an odd length, 0 otherwise. */
int ref_paint(ref_lbm *h, const uint32_t *pairs, uint64_t n_u32) {
    if (n_u32 % 2) return -1;
    for (uint64_t j = 0; j < n_u32; j += 2)
        if (pairs[j] >= h->total) return -1;
    for (uint64_t j = 0; j < n_u32; j += 2) h->mask[pairs[j]] = pairs[j + 1];
    return 0;
}

void ref_set_mask(ref_lbm *h, const uint32_t *mask) { memcpy(h->mask, mask, sizeof(uint32_t) * h->total); }
void ref_get_mask(const ref_lbm *h, uint32_t *out) { memcpy(out, h->mask, sizeof(uint32_t) * h->total); }

/* Summary passes, summary_stats/{curl,ux,uy,rho,speed}.wgsl.  kind: 0 curl, 1 ux, 2 uy, 3 rho, 4 speed
 * (SummaryStat order, src/lbm.rs:10-16).  The output plane keeps values the pass
 * skips, like the reference's output buffer. */
void ref_summary(ref_lbm *h, int kind, float *out) {
    const uint32_t X = h->X, Y = h->Y, total = h->total;
    const float *ux = h->ux, *uy = h->uy, *rho = h->rho;
    float *o = h->output;
    switch (kind) {
    case 0: /* curl.wgsl:35-49 */
        PAR_FOR
        for (uint32_t i = 0; i < total; ++i) {
            if (i % X == 0u) continue;
            if (i / X >= Y - 1u) continue;
            uint32_t ie = i + 1u, iw = i - 1u, in = i - X, is = i + X;
            float uye = ie < total ? uy[ie] : 0.0f;
            float uyw = iw < total ? uy[iw] : 0.0f;
            float uxn = in < total ? ux[in] : 0.0f;
            float uxs = is < total ? ux[is] : 0.0f;
            o[i] = 10.0f * (uye - uyw - uxn + uxs) / rho[i];
        }
        break;
    case 1: /* ux.wgsl:41 */
        PAR_FOR
        for (uint32_t i = 0; i < total; ++i) o[i] = ux[i];
        break;
    case 2: /* uy.wgsl:41 */
        PAR_FOR
        for (uint32_t i = 0; i < total; ++i) o[i] = uy[i];
        break;
    case 3: /* rho.wgsl:24 */
        PAR_FOR
        for (uint32_t i = 0; i < total; ++i) {
            float v = 0.15f * rho[i];
            v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
            o[i] = 4.0f * v - 0.5f;
        }
        break;
    case 4: /* speed.wgsl:25 */
        PAR_FOR
        for (uint32_t i = 0; i < total; ++i) {
            float v = 5.0f * sqrtf(ux[i] * ux[i] + uy[i] * uy[i]);
            v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
            o[i] = v - 0.5f;
        }
        break;
    default:
        break;
    }
    if (out) memcpy(out, o, sizeof(float) * total);
}

/* The reference state after compute_step steps: planes k != 4 of set
 * (compute_step % 2) and the rest plane of set 0 (src/lbm.rs:775).  out holds
 * 9 dense X*Y planes in plane order. */
void ref_get_state(const ref_lbm *h, float *out) {
    int p = (int)(h->compute_step % 2);
    for (int k = 0; k < 9; ++k) {
        const float *src = (k == C_) ? h->set[0][C_] : h->set[p][k];
        memcpy(out + (size_t)k * h->total, src, sizeof(float) * h->total);
    }
}

/* Raw access to one parity set (9 dense planes); set 1's rest plane is not live. */
void ref_get_set(const ref_lbm *h, int p, float *out) {
    for (int k = 0; k < 9; ++k) memcpy(out + (size_t)k * h->total, h->set[p][k], sizeof(float) * h->total);
}

void ref_set_set(ref_lbm *h, int p, const float *in) {
    for (int k = 0; k < 9; ++k) memcpy(h->set[p][k], in + (size_t)k * h->total, sizeof(float) * h->total);
}

void ref_get_density(const ref_lbm *h, float *ux, float *uy, float *rho) {
    memcpy(ux, h->ux, sizeof(float) * h->total);
    memcpy(uy, h->uy, sizeof(float) * h->total);
    memcpy(rho, h->rho, sizeof(float) * h->total);
}

uint64_t ref_step_count(const ref_lbm *h) { return h->compute_step; }
void ref_set_step_count(ref_lbm *h, uint64_t n) { h->compute_step = n; }
