// Single-domain D2Q9 lattice Boltzmann engine for gfx950, behind the C ABI in
// include/lbm_hip.h.
//
// State representation
// --------------------
// The reference (lbm-wgpu src/lbm.rs:1112-1134) keeps two parity sets of nine
// planes.  Step n+1 collides set p = n % 2 in place (C_{n+1}) and then streams
// it into set 1-p; cells the stream does not write (solid cells, column 0 and
// row ny-1, stream/e_w_stream.wgsl:31-45) keep what set 1-p held.  The state
// after n steps is therefore the pair (S_n = set n%2, set (n+1)%2).
//
// This engine keeps the same two sets but runs each step as ONE pass that pulls
// and collides in registers ("P-form"):
//   S-form   buf[cur] = S_n (pre-collision), buf[1-cur] = the other set.
//   P-form   buf[cur] = C_n = collide(S_{n-1}), buf[1-cur] = the other set; the
//            stream of step n is pending.
// Kernels
//   collide (S -> P)   buf[cur] := collide(buf[cur]) in place.
//   fused   (P -> P)   S_n := pull(buf[cur]) on streamed cells, buf[1-cur] own
//                      values elsewhere; buf[1-cur] := collide(S_n); swap.
//   flush   (P -> S)   buf[1-cur] := S_n (streamed cells only); swap.
//   pair    (P -> P -> P, variants 8 and 9) two fused steps in one pass, the
//                      first one's result staged in LDS (k_pair below); its
//                      last blocks step the cells that read only themselves.
// Since the pull of step n uses the mask at the time the kernel runs, any mask
// change first flushes a pending stream, so results are those of the reference
// step order exactly.  The rest population is stored in both sets (the
// reference keeps one shared plane, lbm.rs:775); it is never streamed, so the
// values are the same.
//
// Memory layout in HBM: one allocation per set, structure of arrays by row: the
// nine population rows of grid row y are contiguous (element (k, x, y) at
// (9 y + k) * pitch + x, pitch = nx rounded up to 64 plus a pad).  A step then
// streams one contiguous window of rows instead of eighteen planes far apart,
// which measured 6.2 TB/s for the pull pattern against 5.8 TB/s for nine
// separate planes.  Mask, the per-cell bounce-back byte and the moment planes
// are separate planes with the same pitch.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "lbm_common.h"
#include "tune_cache.h"
#include "pair.h"

namespace {

// The step kernel of one grid: every cell, CPT consecutive cells per thread.
// P-form: vector pulls patched in place where the bounce-back byte says so;
// S-form (no pull): the cell's own values, CPT = 1.
template <class St, int MODE, bool PFORM, int OP, int CPT, bool NT>
__global__ __launch_bounds__(kBlock) void k_cell(const typename St::T *src, typename St::T *dst,
                                                 const uint8_t *__restrict__ flags,
                                                 const uint8_t *__restrict__ mask, Geom g,
                                                 float *__restrict__ mom, double *__restrict__ partials,
                                                 int blocks_per_row, int nblocks, int remap) {
    static_assert(PFORM || CPT == 1, "S-form passes run one cell per thread");
    const int L = remap ? xcd_remap(blockIdx.x, nblocks) : (int)blockIdx.x;
    const int y = L / blocks_per_row;
    const int x0 = ((L - y * blocks_per_row) * kBlock + threadIdx.x) * CPT;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (x0 < g.X) {
        if constexpr (PFORM)
            patched_tile<St, MODE, OP, CPT, NT>(src, dst, flags, mask, g, y, x0, mom, s0, s1, s2);
        else
            interior_tile<St, MODE, false, OP, CPT, NT>(src, dst, mask, g, y, x0, mom, s0, s1, s2);
    }
    if (OP & OP_REDUCE) block_sums(s0, s1, s2, partials, blockIdx.x);
}

// ---- two steps per launch -------------------------------------------------------
// (the per-block work is pair_block / fixed_cell in pair.h)

// the fixed-cell list: kSlowCell cells per row (one block per row), then the
// (column, row) entries at each row's offset
__global__ __launch_bounds__(kFixedBlock) void k_fixed_count(const uint8_t *__restrict__ flags, Geom g,
                                                             uint32_t *__restrict__ counts) {
    fixed_row_count(flags + (size_t)blockIdx.x * g.pitch, g.X, counts + blockIdx.x);
}

__global__ __launch_bounds__(kFixedBlock) void k_fixed_emit(const uint8_t *__restrict__ flags,
                                                            const uint8_t *__restrict__ mask, Geom g,
                                                            const uint32_t *__restrict__ offsets,
                                                            uint2 *__restrict__ list) {
    const size_t r = (size_t)blockIdx.x * g.pitch;
    fixed_row_emit(flags + r, mask + r, g.X, (int)blockIdx.x, g.Y, blockIdx.x, list + offsets[blockIdx.x]);
}

template <class St, int MODE>
__global__ __launch_bounds__(kBlock) void k_pair(const typename St::T *A, const typename St::T *B,
                                                 typename St::T *__restrict__ A2, typename St::T *__restrict__ B2,
                                                 const uint8_t *__restrict__ flags, Geom g, int nstrips, int rows,
                                                 int nblocks, int remap, const uint2 *__restrict__ fixed,
                                                 uint32_t n_fixed) {
    __shared__ typename St::T ring[kRing][8][kBlock];
    if ((int)blockIdx.x >= nblocks) {
        // the fixed-cell pass: one thread per list entry (column | streamed << 31, row)
        const uint32_t i = (blockIdx.x - nblocks) * kBlock + threadIdx.x;
        if (i < n_fixed) {
            const uint2 e = fixed[i];
            fixed_cell<St>(A, B, g, (int)(e.x & 0x7FFFFFFFu), (int)e.y, (e.x >> 31) != 0, A2, B2);
        }
        return;
    }
    const int L = remap ? xcd_remap(blockIdx.x, nblocks) : (int)blockIdx.x;
    const int chunk = L / nstrips;
    const int r0 = chunk * rows;
    pair_block_any<St, MODE>(ring, A, B, A2, B2, flags, g, L - chunk * nstrips, r0, min(r0 + rows, g.Y));
}

// ---- surrogate support ---------------------------------------------------------

// Network input (NHWC bf16, C channels, nx / F x ny / F pixels) from the
// moments of the current state.  Pixel (px, py) covers the F x F cells from
// (F px, F py): channel 0 is their solid fraction, channels 1..3 the mean over
// their fluid cells of (rho - 1) / s_rho, ux / s_u, uy / s_u (north-positive),
// channels 4, 5 the grid's (tau - 1/2) / 0.2 and inflow speed u_in / s_u.
template <int F>
__global__ void k_pack_net_input(const float *__restrict__ mom, const uint8_t *__restrict__ mask,
                                 uint16_t *__restrict__ xin, Geom g, int C, float inv_s_rho, float inv_s_u,
                                 float u_in) {
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    const int py = blockIdx.y;
    const int w = g.X / F;
    if (px >= w) return;
    float nsolid = 0.0f, nfluid = 0.0f, srho = 0.0f, sux = 0.0f, suy = 0.0f;
#pragma unroll
    for (int j = 0; j < F; ++j)
#pragma unroll
        for (int k = 0; k < F; ++k) {
            const size_t i = (size_t)(F * py + j) * g.pitch + F * px + k;
            if (mask[i] == 1) {
                nsolid += 1.0f;
            } else {
                const float rho = mom[2 * g.pstride + i];
                nfluid += 1.0f;
                srho += rho;
                sux += mom[i] / rho;
                suy += mom[g.pstride + i] / rho;
            }
        }
    float v[6] = {nsolid * (1.0f / (F * F)), 0.0f, 0.0f, 0.0f, (1.0f / g.omega - 0.5f) * 5.0f, u_in * inv_s_u};
    if (nfluid > 0.0f) {
        const float r = 1.0f / nfluid;
        v[1] = (srho * r - 1.0f) * inv_s_rho;
        v[2] = sux * r * inv_s_u;
        v[3] = suy * r * inv_s_u;
    }
    typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
    u16x8 *out = (u16x8 *)(xin + ((size_t)py * w + px) * C);
    u16x8 first = u16x8{};
    for (int c = 0; c < 6; ++c) first[c] = __builtin_bit_cast(uint16_t, (__bf16)v[c]);
    out[0] = first;
    for (int c = 1; c < C / 8; ++c) out[c] = u16x8{};
}

// Streamed fluid cells := feq(rho + d_rho o0, u + d_u (o1, o2)) in set `buf`
// (S-form), where (rho, u) is the cell's own current state and d the network's
// predicted change, bilinearly interpolated from the nx / F x ny / F output
// grid at the cell centre (exactly the cell's own pixel for F = 1).
template <class St, int F>
__global__ void k_apply_jump(const float *__restrict__ net_out, int C, const float *__restrict__ mom,
                             const uint8_t *__restrict__ mask, typename St::T *buf, Geom g, float d_rho, float d_u) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= g.X) return;
    const size_t i = (size_t)y * g.pitch + x;
    if (mask[i] == 1 || x == 0 || y >= g.Y - 1) return;
    const int w = g.X / F, h = g.Y / F;
    const float gx = fmaxf((x + 0.5f) / F - 0.5f, 0.0f), gy = fmaxf((y + 0.5f) / F - 0.5f, 0.0f);
    const int i0 = min((int)gx, w - 1), j0 = min((int)gy, h - 1);
    const int i1 = min(i0 + 1, w - 1), j1 = min(j0 + 1, h - 1);
    const float fx = fminf(gx - (float)i0, 1.0f), fy = fminf(gy - (float)j0, 1.0f);
    float d[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float o00 = net_out[((size_t)j0 * w + i0) * C + c], o01 = net_out[((size_t)j0 * w + i1) * C + c];
        const float o10 = net_out[((size_t)j1 * w + i0) * C + c], o11 = net_out[((size_t)j1 * w + i1) * C + c];
        d[c] = (1.0f - fy) * ((1.0f - fx) * o00 + fx * o01) + fy * ((1.0f - fx) * o10 + fx * o11);
    }
    const float rho0 = mom[2 * g.pstride + i];
    const float rho = rho0 + d_rho * d[0];
    const float ux = mom[i] / rho0 + d_u * d[1];
    const float uym = -(mom[g.pstride + i] / rho0 + d_u * d[2]);  // memory rows grow southwards
    const float uu = ux * ux + uym * uym;
    const size_t sown = (size_t)y * g.ys + x;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const float cu = (float)cx(k) * ux + (float)cy(k) * uym;
        const float feq = weight(k) * rho * (1.0f + 3.0f * cu + 4.5f * cu * cu - 1.5f * uu);
        buf[(size_t)k * g.ks + sown] = St::from_f(feq, k, g.sh);
    }
}

// per-block FP64 sums of |u - u_mark|^2 and |u_mark|^2 over fluid cells
__global__ __launch_bounds__(kBlock) void k_velocity_change(const float *__restrict__ mom,
                                                            const float *__restrict__ mark,
                                                            const uint8_t *__restrict__ mask, Geom g,
                                                            double *__restrict__ partials) {
    const int x = blockIdx.x * kBlock + threadIdx.x;
    const int y = blockIdx.y;
    double s0 = 0.0, s1 = 0.0;
    if (x < g.X) {
        const size_t i = (size_t)y * g.pitch + x;
        if (mask[i] != 1) {
            const float rho = mom[2 * g.pstride + i];
            const double ux = mom[i] / rho, uy = mom[g.pstride + i] / rho;
            const double mx = mark[i], my = mark[g.pstride + i];
            s0 = (ux - mx) * (ux - mx) + (uy - my) * (uy - my);
            s1 = mx * mx + my * my;
        }
    }
    block_sums(s0, s1, 0.0, partials, (size_t)blockIdx.y * gridDim.x + blockIdx.x);
}

// The rollout's check: u the plain LBM's velocity after a cycle, m0 the field
// at the cycle's start (the persistence forecast), m1 the surrogate's forecast;
// over fluid cells s0 = |u - m1|^2, s1 = |u - m0|^2, s2 = |m1|^2.
__global__ __launch_bounds__(kBlock) void k_velocity_skill(const float *__restrict__ mom,
                                                           const float *__restrict__ m0,
                                                           const float *__restrict__ m1,
                                                           const uint8_t *__restrict__ mask, Geom g,
                                                           double *__restrict__ partials) {
    const int x = blockIdx.x * kBlock + threadIdx.x;
    const int y = blockIdx.y;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (x < g.X) {
        const size_t i = (size_t)y * g.pitch + x;
        if (mask[i] != 1) {
            const float rho = mom[2 * g.pstride + i];
            const double ux = mom[i] / rho, uy = mom[g.pstride + i] / rho;
            const double fx = m1[i], fy = m1[g.pstride + i];
            const double px = m0[i], py = m0[g.pstride + i];
            s0 = (ux - fx) * (ux - fx) + (uy - fy) * (uy - fy);
            s1 = (ux - px) * (ux - px) + (uy - py) * (uy - py);
            s2 = fx * fx + fy * fy;
        }
    }
    block_sums(s0, s1, s2, partials, (size_t)blockIdx.y * gridDim.x + blockIdx.x);
}

__global__ void k_velocity_mark(const float *__restrict__ mom, float *__restrict__ mark, Geom g) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= g.X) return;
    const size_t i = (size_t)y * g.pitch + x;
    const float rho = mom[2 * g.pstride + i];
    mark[i] = mom[i] / rho;
    mark[g.pstride + i] = mom[g.pstride + i] / rho;
}

}  // namespace

// provided by unet.hip (same library)
struct lbm_unet;
uint16_t *unet_input_tensor(lbm_unet *n);
float *unet_output_tensor(lbm_unet *n);
void unet_shape(const lbm_unet *n, int *b, int *h, int *w);
void unet_forward_on(lbm_unet *n, hipStream_t s);
void unet_begin_use(lbm_unet *n, hipStream_t s);
void unet_end_use(lbm_unet *n, hipStream_t s);
int unet_device(const lbm_unet *n);

struct lbm_grid {
    int dev = 0;
    hipStream_t stream = nullptr;
    uint32_t X = 0, Y = 0;
    int precision = 0, boundary = 0;
    Geom geom{};
    size_t esize = 4;
    float omega = 1.25f;
    float u_in = 0.1f;  // x velocity of the last uniform reset: the inlet column's speed

    // the inflow speed also centres the F16E storage on the free-stream equilibrium
    void set_inflow(float u) {
        u_in = u;
        geom.sh = make_f16_shift(precision == LBM_PRECISION_F16E, u);
    }
    void *f[2] = {nullptr, nullptr};
    void *spare[2] = {nullptr, nullptr};  // output sets of the two-step kernel, allocated on first use
    uint2 *fixed = nullptr;               // its fixed-cell list: (column | streamed << 31, row) per kSlowCell cell
    uint32_t *fixed_rows = nullptr;       // the list's offset per row
    uint32_t n_fixed = 0;
    bool fixed_dirty = true;              // the list does not follow the flags (a rebuild failed)
    int cur = 0;
    bool pform = false;
    uint8_t *mask = nullptr, *flags = nullptr;
    float *lag = nullptr;   // jx, jy, rho planes: the renderer's density bind group
    float *now = nullptr;   // moments of the current state (scratch)
    float *out = nullptr;   // summary output plane (persistent)
    float *dense = nullptr; // dense f32 staging (4 planes)
    float *rgb = nullptr;   // colour-map output (3 floats per cell), allocated on first use
    size_t rgb_elems = 0;
    double *partials = nullptr;
    double *sums = nullptr;
    size_t npartials = 0;
    uint64_t steps = 0;
    int variant = 0;  // step-kernel variant (cells per thread, non-temporal access)
    std::string err;
    // device snapshot of the evolving state (both sets, fields, form, step count)
    char *snap = nullptr;
    bool snap_valid = false;
    int snap_cur = 0;
    bool snap_pform = false;
    uint64_t snap_steps = 0;
    float *umark = nullptr;  // two slots of velocity planes for the surrogate's residuals
    int umarked = 0;         // bit s: slot s holds a mark

    void ensure_vpart(dim3 grid) {
        if (!vpart) {
            nvpart = (size_t)grid.x * grid.y;
            vpart = dalloc<double>(3 * nvpart);
        }
        if (!sums) sums = dalloc<double>(3);
    }
    double *vpart = nullptr;
    size_t nvpart = 0;

    size_t set_bytes() const { return (size_t)geom.ys * Y * esize; }
    // a set's allocation: two rows more, which the two-step kernel's loads of
    // the last rows (wrapped outlet pulls, the relay lanes) read and discard
    size_t alloc_bytes() const { return set_bytes() + 2 * (size_t)geom.ys * esize; }
    struct Shape {
        int blocks_per_row, nblocks;
    };
    Shape shape(int cpt) const {
        const int groups = ((int)X + cpt - 1) / cpt;
        const int bpr = (groups + kBlock - 1) / kBlock;
        return {bpr, bpr * (int)Y};
    }

    void release() {
        (void)hipSetDevice(dev);
        for (void *p : {f[0], f[1], spare[0], spare[1], (void *)fixed, (void *)fixed_rows, (void *)mask, (void *)flags,
                        (void *)lag, (void *)now, (void *)out, (void *)dense, (void *)partials, (void *)sums,
                        (void *)snap, (void *)umark, (void *)vpart})
            if (p) (void)hipFree(p);
        snap = nullptr;
        umark = nullptr;
        vpart = nullptr;
        f[0] = f[1] = nullptr;
        spare[0] = spare[1] = nullptr;
        fixed = nullptr;
        fixed_rows = nullptr;
        mask = flags = nullptr;
        lag = now = out = dense = nullptr;
        partials = sums = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    // ---- kernel launchers --------------------------------------------------
    // One pass = one launch of the step kernel over every cell; FP64 partial
    // sums (OP_REDUCE) go to part[0, nblocks).
    template <class St, int MODE, bool PFORM, int OP, int CPT, bool NT>
    void go(const void *src, void *dst, float *mom, double *part, int remap) {
        using T = typename St::T;
        const Shape sh = shape(CPT);
        hipLaunchKernelGGL((k_cell<St, MODE, PFORM, OP, CPT, NT>), dim3(sh.nblocks), dim3(kBlock), 0, stream,
                           (const T *)src, (T *)dst, flags, mask, geom, mom, part, sh.blocks_per_row, sh.nblocks,
                           remap);
        HIPCHK(hipGetLastError());
        last_partials = (size_t)sh.nblocks;
    }
    size_t last_partials = 0;

    // cells per thread of the non-step passes
    static constexpr int default_cpt(int precision) { return precision == LBM_PRECISION_F32 ? 1 : 2; }

    // The fused step is the hot kernel and has variants; every other pass runs
    // one fixed shape.  Variants differ only in how memory is moved (cells per
    // thread, non-temporal access, XCD-aware block order), never in
    // arithmetic, so all of them give the same bits.
    //   f32:  v  0 1 2 3 4 5 6 7      f16s: v  0 1 2 3 4 5 6 7
    //   cells    1 1 1 2 2 4 4 4      cells    1 1 1 2 2 4 4 8
    //   nt       1 1 0 1 1 1 1 0      nt       1 1 0 1 1 1 1 1
    //   remap    1 0 0 1 0 1 0 0      remap    1 0 0 1 0 1 0 0
    // and 8, 9: two steps per launch (k_pair, see pair()).
    template <class St, int MODE, bool PFORM, int OP>
    void launch_cell(const void *src, void *dst, float *mom, double *part) {
        constexpr bool f32 = sizeof(typename St::T) == 4;
        constexpr bool hot = PFORM && (OP & OP_COLLIDE);
        if constexpr (hot) {
            switch (variant) {
            case 0: return go<St, MODE, PFORM, OP, 1, true>(src, dst, mom, part, 1);
            case 1: return go<St, MODE, PFORM, OP, 1, true>(src, dst, mom, part, 0);
            case 2: return go<St, MODE, PFORM, OP, 1, false>(src, dst, mom, part, 0);
            case 3: return go<St, MODE, PFORM, OP, 2, true>(src, dst, mom, part, 1);
            case 4: return go<St, MODE, PFORM, OP, 2, true>(src, dst, mom, part, 0);
            case 5: return go<St, MODE, PFORM, OP, 4, true>(src, dst, mom, part, 1);
            case 6: return go<St, MODE, PFORM, OP, 4, true>(src, dst, mom, part, 0);
            case 7:
                if constexpr (f32) return go<St, MODE, PFORM, OP, 4, false>(src, dst, mom, part, 0);
                else return go<St, MODE, PFORM, OP, 8, true>(src, dst, mom, part, 0);
            default:  // single steps of the two-step variants: variant 0 (f32) / 4 (f16s)
                if constexpr (f32) return go<St, MODE, PFORM, OP, 1, true>(src, dst, mom, part, 1);
                else return go<St, MODE, PFORM, OP, 2, true>(src, dst, mom, part, 0);
            }
        }
        go<St, MODE, PFORM, OP, (PFORM && !f32 ? 2 : 1), true>(src, dst, mom, part, 0);
    }

    // Variants 8, 9: two P-form steps per launch (k_pair), blocks of 8 or 16 rows
    // (measured at 8192 x 4096: 2 / 4 / 8 / 16 / 32 rows 362 / 312 / 303 / 314 /
    // 333 us per step).  Two cells per thread (pairs through the ring, 510-column
    // strips) measured slower: f32 408 vs 297 us per step (64 KB of ring, 2 waves
    // per SIMD), binary16 303 vs 255 us.
    static constexpr int kPairFirst = 8;
    static bool pair_variant(int v) { return v >= kPairFirst; }
    int pair_rows() const { return 8 << (variant - kPairFirst); }

    // the two-step kernel's output sets and fixed-cell list (room for every
    // cell); allocated when a two-step variant is set, so that stepping itself
    // does not allocate
    void ensure_pair_buffers() {
        if (spare[0] && spare[1] && fixed) return;
        release_pair_buffers();
        try {
            spare[0] = dalloc<char>(alloc_bytes());
            spare[1] = dalloc<char>(alloc_bytes());
            fixed = dalloc<uint2>((size_t)X * Y);
            fixed_rows = dalloc<uint32_t>((size_t)Y);
            // the rows' padding is never written by this kernel: keep it defined
            HIPCHK(hipMemsetAsync(spare[0], 0, set_bytes(), stream));
            HIPCHK(hipMemsetAsync(spare[1], 0, set_bytes(), stream));
            build_fixed_list();
        } catch (...) {
            release_pair_buffers();
            throw;
        }
    }

    // the fixed-cell list from the current flags (after every mask change
    // while the two-step buffers exist); waits for the stream.  Until it has
    // succeeded the list is dirty, and a two-step launch rebuilds it first.
    void build_fixed_list() {
        fixed_dirty = true;
        hipLaunchKernelGGL(k_fixed_count, dim3(Y), dim3(kFixedBlock), 0, stream, flags, geom, fixed_rows);
        HIPCHK(hipGetLastError());
        std::vector<uint32_t> off(Y);
        HIPCHK(hipMemcpyAsync(off.data(), fixed_rows, Y * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        uint32_t total = 0;
        for (uint32_t &o : off) {
            const uint32_t c = o;
            o = total;
            total += c;
        }
        HIPCHK(hipMemcpyAsync(fixed_rows, off.data(), Y * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_fixed_emit, dim3(Y), dim3(kFixedBlock), 0, stream, flags, mask, geom, fixed_rows, fixed);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));  // `off` is read by the copy above
        n_fixed = total;
        fixed_dirty = false;
    }

    template <class St, int MODE>
    void launch_pair() {
        using T = typename St::T;
        ensure_pair_buffers();
        if (fixed_dirty) build_fixed_list();
        const T *A = (const T *)f[cur], *B = (const T *)f[1 - cur];
        T *A2 = (T *)spare[0], *B2 = (T *)spare[1];
        const int rows = pair_rows();
        const int nstrips = ((int)X + kBlock - 3) / (kBlock - 2);
        const int nblocks = nstrips * (((int)Y + rows - 1) / rows);
        const unsigned nfix = (n_fixed + kBlock - 1) / kBlock;
        hipLaunchKernelGGL((k_pair<St, MODE>), dim3(nblocks + nfix), dim3(kBlock), 0, stream, A, B, A2, B2, flags,
                           geom, nstrips, rows, nblocks, 1, fixed, n_fixed);
        HIPCHK(hipGetLastError());
        spare[0] = f[cur];
        spare[1] = f[1 - cur];
        f[cur] = A2;
        f[1 - cur] = B2;
    }

    // the two-step kernel's output sets and fixed-cell list (the state itself
    // is always in f[0], f[1] between calls)
    void release_pair_buffers() {
        for (void *&p : spare) {
            if (p) (void)hipFree(p);
            p = nullptr;
        }
        if (fixed) (void)hipFree(fixed);
        if (fixed_rows) (void)hipFree(fixed_rows);
        fixed = nullptr;
        fixed_rows = nullptr;
        n_fixed = 0;
    }

    // two P-form steps in one launch; the other set is then whole only on the
    // cells the stream does not write
    void pair() {
        by_store(precision, [&]<class St>() {
            if (boundary == LBM_BOUNDARY_REFERENCE) launch_pair<St, LBM_BOUNDARY_REFERENCE>();
            else launch_pair<St, LBM_BOUNDARY_ZERO_GRADIENT>();
        });
    }

    template <bool PFORM, int OP>
    void cell(const void *src, void *dst, float *mom, double *part) {
        by_store(precision, [&]<class St>() {
            if (boundary == LBM_BOUNDARY_REFERENCE)
                launch_cell<St, LBM_BOUNDARY_REFERENCE, PFORM, OP>(src, dst, mom, part);
            else
                launch_cell<St, LBM_BOUNDARY_ZERO_GRADIENT, PFORM, OP>(src, dst, mom, part);
        });
    }

    void build_flags() {
        dim3 grid((X + 255) / 256, Y);
        if (boundary == LBM_BOUNDARY_REFERENCE)
            hipLaunchKernelGGL(k_build_flags<LBM_BOUNDARY_REFERENCE>, grid, dim3(256), 0, stream, mask, flags, geom);
        else
            hipLaunchKernelGGL(k_build_flags<LBM_BOUNDARY_ZERO_GRADIENT>, grid, dim3(256), 0, stream, mask, flags,
                               geom);
        HIPCHK(hipGetLastError());
        if (fixed) build_fixed_list();
    }

    void fill_both(const float v[9]) {
        const float4 a = make_float4(v[0], v[1], v[2], v[3]);
        const float4 b = make_float4(v[4], v[5], v[6], v[7]);
        const dim3 grid((unsigned)(((size_t)geom.ys * Y + 255) / 256));
        for (int s = 0; s < 2; ++s) {
            by_store(precision, [&]<class St>() {
                hipLaunchKernelGGL(k_fill<St>, grid, dim3(256), 0, stream, (typename St::T *)f[s], geom, a, b, v[8]);
            });
            HIPCHK(hipGetLastError());
        }
    }

    void ensure_now() {
        if (!now) now = dalloc<float>(3 * (size_t)geom.pstride);
    }
    void ensure_dense() {
        if (!dense) dense = dalloc<float>(4 * (size_t)X * Y);
    }
    void ensure_partials() {
        if (!partials) {
            npartials = (size_t)shape(1).nblocks;  // the largest pass
            partials = dalloc<double>(3 * npartials);
            if (!sums) sums = dalloc<double>(3);
        }
    }

    // P-form -> S-form: complete the pending stream with the current mask
    void flush() {
        if (!pform) return;
        cell<true, OP_STORE_S>(f[cur], f[1 - cur], nullptr, nullptr);
        cur = 1 - cur;
        pform = false;
    }

    void step(uint32_t n) {
        for (uint32_t i = 0; i < n; ++i) {
            const bool last = i + 1 == n;
            if (pform && pair_variant(variant) && n - i >= 3) {
                // after a two-step launch the other set is current only on the
                // cells the stream does not write, the only ones the next step
                // reads from it; the last step of a call is always a one-step
                // launch, which writes that set whole (and the moments)
                pair();
                ++i;
                continue;
            }
            if (!pform) {
                if (last)
                    cell<false, OP_COLLIDE | OP_MOMENTS>(f[cur], f[cur], lag, nullptr);
                else
                    cell<false, OP_COLLIDE>(f[cur], f[cur], nullptr, nullptr);
                pform = true;
            } else {
                if (last)
                    cell<true, OP_COLLIDE | OP_MOMENTS>(f[cur], f[1 - cur], lag, nullptr);
                else
                    cell<true, OP_COLLIDE>(f[cur], f[1 - cur], nullptr, nullptr);
                cur = 1 - cur;
            }
        }
        steps += n;
    }

    // moments of the current state into `now` (optionally with the FP64 reduction)
    // flush() and moments_now(false) in one pass where the stored S-form state
    // decodes to exactly the pulled values the moments are taken from (f32,
    // f - w).  F16E re-rounds a bounced population into its new population's
    // shift when it stores it, so its moments come from the stored state in a
    // pass of their own.
    void flush_with_moments() {
        ensure_now();
        if (pform && precision != LBM_PRECISION_F16E) {
            cell<true, OP_STORE_S | OP_MOMENTS>(f[cur], f[1 - cur], now, nullptr);
            cur = 1 - cur;
            pform = false;
            return;
        }
        flush();
        moments_now(false);
    }

    void moments_now(bool reduce) {
        ensure_now();
        if (reduce) ensure_partials();
        if (pform) {
            if (reduce) cell<true, OP_MOMENTS | OP_REDUCE>(f[cur], f[1 - cur], now, partials);
            else cell<true, OP_MOMENTS>(f[cur], f[1 - cur], now, nullptr);
        } else {
            if (reduce) cell<false, OP_MOMENTS | OP_REDUCE>(f[cur], nullptr, now, partials);
            else cell<false, OP_MOMENTS>(f[cur], nullptr, now, nullptr);
        }
    }

    void copy_plane_out(const void *dev_plane, float *host, int k) {
        // decode one pitched population plane into a dense f32 host plane
        ensure_dense();
        dim3 grid((X + 255) / 256, Y);
        by_store(precision, [&]<class St>() {
            hipLaunchKernelGGL(k_decode_plane<St>, grid, dim3(256), 0, stream, (const typename St::T *)dev_plane,
                               dense, geom, k);
        });
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(host, dense, sizeof(float) * X * Y, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }

    void copy_plane_in(const float *host, void *dev_plane, int k) {
        ensure_dense();
        HIPCHK(hipMemcpyAsync(dense, host, sizeof(float) * X * Y, hipMemcpyHostToDevice, stream));
        dim3 grid((X + 255) / 256, Y);
        by_store(precision, [&]<class St>() {
            hipLaunchKernelGGL(k_encode_plane<St>, grid, dim3(256), 0, stream, dense, (typename St::T *)dev_plane,
                               geom, k);
        });
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));
    }

    // element (k, 0, 0) of set s
    void *plane(int s, int k) { return (char *)f[s] + (size_t)k * geom.ks * esize; }
    size_t state_offset(int k, uint32_t x, uint32_t y) const {
        return ((size_t)y * geom.ys + (size_t)k * geom.ks + x) * esize;
    }

    void copy_f32_plane_out(const float *dev, float *host) {
        HIPCHK(hipMemcpy2DAsync(host, sizeof(float) * X, dev, sizeof(float) * geom.pitch, sizeof(float) * X, Y,
                                hipMemcpyDeviceToHost, stream));
    }
};

namespace {

int fail(lbm_grid *g, int code, const std::string &msg) {
    if (g) g->err = msg;
    else g_create_error = msg;
    return code;
}

template <class F>
int guarded(lbm_grid *g, F &&fn) {
    if (!g) return fail(nullptr, LBM_EINVAL, "null grid");
    try {
        HIPCHK(hipSetDevice(g->dev));
        fn();
        g->err.clear();
        return LBM_OK;
    } catch (const HipError &e) {
        return fail(g, e.e == hipErrorOutOfMemory ? LBM_ENOMEM : LBM_EHIP,
                    std::string(e.what) + ": " + hipGetErrorString(e.e));
    } catch (const ArgError &e) {
        return fail(g, LBM_EINVAL, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(g, LBM_ENOMEM, "host allocation failed");
    } catch (...) {
        return fail(g, LBM_EHIP, "unexpected exception");
    }
}

void check(bool ok, const char *msg) {
    if (!ok) throw ArgError{msg};
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int lbm_config_default(lbm_config *cfg, uint32_t nx, uint32_t ny) {
    if (!cfg) return LBM_EINVAL;
    cfg->nx = nx;
    cfg->ny = ny;
    cfg->omega = 1.0f / (0.5f + 0.3f);
    cfg->u_init = 0.1f;
    cfg->precision = LBM_PRECISION_F32;
    cfg->boundary = LBM_BOUNDARY_REFERENCE;
    cfg->device = 0;
    cfg->row_pad = 0;
    return LBM_OK;
}

const char *lbm_last_error(const lbm_grid *g) { return g ? g->err.c_str() : g_create_error.c_str(); }

int lbm_create(const lbm_config *cfg, lbm_grid **out) {
    if (!cfg || !out) return fail(nullptr, LBM_EINVAL, "null argument");
    *out = nullptr;
    if (cfg->nx < 3 || cfg->ny < 3) return fail(nullptr, LBM_EINVAL, "nx and ny must be >= 3");
    if (cfg->row_pad % 8 != 0) return fail(nullptr, LBM_EINVAL, "row_pad must be a multiple of 8 elements");
    if (cfg->ny > 65535) return fail(nullptr, LBM_EINVAL, "ny must be <= 65535 (one launch row per grid row)");
    if ((uint64_t)cfg->nx * cfg->ny > 0xFFFFFFFFull)
        return fail(nullptr, LBM_EINVAL, "nx * ny must fit in 32 bits (the paint wire format is u32)");
    if (cfg->precision != LBM_PRECISION_F32 && cfg->precision != LBM_PRECISION_F16S &&
        cfg->precision != LBM_PRECISION_F16E)
        return fail(nullptr, LBM_EINVAL, "unknown precision");
    if (cfg->boundary != LBM_BOUNDARY_REFERENCE && cfg->boundary != LBM_BOUNDARY_ZERO_GRADIENT)
        return fail(nullptr, LBM_EINVAL, "unknown boundary mode");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, LBM_ENODEV, "no HIP device");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(nullptr, LBM_EINVAL, "bad device ordinal");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) != hipSuccess)
        return fail(nullptr, LBM_ENODEV, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, LBM_ENODEV, std::string("this library is built for gfx950, device is ") + prop.gcnArchName);

    lbm_grid *g = new (std::nothrow) lbm_grid();
    if (!g) return fail(nullptr, LBM_ENOMEM, "host allocation failed");
    g->dev = cfg->device;
    g->X = cfg->nx;
    g->Y = cfg->ny;
    g->precision = cfg->precision;
    g->boundary = cfg->boundary;
    g->esize = cfg->precision == LBM_PRECISION_F32 ? 4 : 2;
    g->omega = cfg->omega;
    const int pitch = (int)((cfg->nx + 63u) / 64u * 64u + cfg->row_pad);
    long long pstride = (long long)pitch * cfg->ny;
    pstride = (pstride + 255) / 256 * 256;
    // state layout: the nine population rows of each grid row are contiguous
    g->geom = Geom{(int)cfg->nx, (int)cfg->ny, pitch, pstride, 9LL * pitch, (long long)pitch, cfg->omega};
    int rc = guarded(g, [&] {
        HIPCHK(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
        g->f[0] = dalloc<char>(g->alloc_bytes());
        g->f[1] = dalloc<char>(g->alloc_bytes());
        g->mask = dalloc<uint8_t>(pstride);
        g->flags = dalloc<uint8_t>(pstride);
        g->lag = dalloc<float>(3 * (size_t)pstride);
        g->out = dalloc<float>(pstride);
        HIPCHK(hipMemsetAsync(g->lag, 0, sizeof(float) * 3 * pstride, g->stream));
        HIPCHK(hipMemsetAsync(g->out, 0, sizeof(float) * pstride, g->stream));
        float v[9];
        host_feq(cfg->u_init, 0.0f, 1.0f, v);
        g->set_inflow(cfg->u_init);
        g->fill_both(v);
        HIPCHK(hipMemsetAsync(g->mask, 0, pstride, g->stream));
        hipLaunchKernelGGL(k_walls, dim3((g->X + 255) / 256), dim3(256), 0, g->stream, g->mask, g->geom);
        HIPCHK(hipGetLastError());
        g->build_flags();
        HIPCHK(hipStreamSynchronize(g->stream));
    });
    if (rc != LBM_OK) {
        g_create_error = g->err;
        g->release();
        delete g;
        return rc;
    }
    *out = g;
    return LBM_OK;
}

int lbm_destroy(lbm_grid *g) {
    if (!g) return LBM_EINVAL;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    g->release();
    delete g;
    return LBM_OK;
}

int lbm_set_omega(lbm_grid *g, float omega) {
    return guarded(g, [&] {
        check(std::isfinite(omega), "omega must be finite");
        g->omega = omega;
        g->geom.omega = omega;
    });
}

int lbm_set_viscosity(lbm_grid *g, float nu) {
    return lbm_set_omega(g, 1.0f / (3.0f * nu + 0.5f));
}

int lbm_get_omega(const lbm_grid *g, float *omega) {
    if (!g || !omega) return LBM_EINVAL;
    *omega = g->omega;
    return LBM_OK;
}

int lbm_reset_equilibrium(lbm_grid *g, float rho, float ux, float uy) {
    return guarded(g, [&] {
        float v[9];
        host_feq(ux, uy, rho, v);
        g->set_inflow(ux);  // (F16E: re-centres the storage on the new free stream; both sets are refilled)
        g->fill_both(v);
        g->cur = 0;
        g->pform = false;
        g->steps = 0;
        // the reference runs the two pre-collision passes after a reset
        g->cell<false, OP_MOMENTS | OP_NOREST>(g->f[0], nullptr, g->lag, nullptr);
    });
}

int lbm_single_cell(lbm_grid *g, int k) {
    return guarded(g, [&] {
        float v[9];
        host_feq(0.0f, 0.0f, 1.0f, v);
        g->fill_both(v);
        g->cur = 0;
        g->pform = false;
        g->steps = 0;
        const uint32_t X = g->X, Y = g->Y;
        uint32_t x = 0, y = 0;
        switch (k) {  // set_single_cell, lbm.rs:1486-1496
        case 0: x = X - 2; y = Y - 2; break;
        case 1: x = 3 * X / 4; y = Y - 2; break;
        case 2: x = X / 3; y = Y - 2; break;
        case 3: x = X - 2; y = Y / 2; break;
        case 4: x = 3 * X / 4; y = Y / 2; break;
        case 5: x = X / 2; y = Y / 2; break;
        case 6: x = X - 2; y = 1; break;
        case 7: x = 3 * X / 4; y = 1; break;
        case 8: x = X / 2; y = 1; break;
        default: return;
        }
        for (int s = 0; s < 2; ++s) {
            char *p = (char *)g->f[s] + g->state_offset(k, x, y);
            by_store(g->precision, [&]<class St>() {
                hipLaunchKernelGGL(k_set_one<St>, dim3(1), dim3(64), 0, g->stream, (typename St::T *)p, 4.0f, k,
                                   g->geom.sh);
            });
            HIPCHK(hipGetLastError());
        }
    });
}

int lbm_set_pulse(lbm_grid *g, int k, uint32_t x, uint32_t y, float v) {
    return guarded(g, [&] {
        check(k >= 0 && k < 9, "k must be in 0..8");
        check(x < g->X && y < g->Y, "cell outside the grid");
        g->flush();
        char *p = (char *)g->f[g->cur] + g->state_offset(k, x, y);
        by_store(g->precision, [&]<class St>() {
            hipLaunchKernelGGL(k_set_one<St>, dim3(1), dim3(64), 0, g->stream, (typename St::T *)p, v, k, g->geom.sh);
        });
        HIPCHK(hipGetLastError());
    });
}

int lbm_reset_barrier(lbm_grid *g) {
    return guarded(g, [&] {
        g->flush();
        HIPCHK(hipMemsetAsync(g->mask, 0, g->geom.pstride, g->stream));
        hipLaunchKernelGGL(k_walls, dim3((g->X + 255) / 256), dim3(256), 0, g->stream, g->mask, g->geom);
        HIPCHK(hipGetLastError());
        g->build_flags();
    });
}

int lbm_paint(lbm_grid *g, const uint32_t *pairs, size_t n_u32) {
    return guarded(g, [&] {
        if (n_u32 == 0) return;
        check(pairs != nullptr, "null pairs");
        check(n_u32 % 2 == 0, "pairs must hold an even number of u32 words");
        const uint64_t total = (uint64_t)g->X * g->Y;
        for (size_t j = 0; j < n_u32; j += 2) check(pairs[j] < total, "cell index out of range");
        // last writer wins: keep the final value per index
        std::unordered_map<uint32_t, uint8_t> last;
        last.reserve(n_u32 / 2);
        for (size_t j = 0; j < n_u32; j += 2) last[pairs[j]] = pairs[j + 1] == 1u ? 1 : 0;
        std::vector<uint32_t> idx;
        std::vector<uint8_t> val;
        idx.reserve(last.size());
        val.reserve(last.size());
        for (const auto &kv : last) {
            idx.push_back(kv.first);
            val.push_back(kv.second);
        }
        g->flush();
        uint32_t *d_idx = dalloc<uint32_t>(idx.size());
        uint8_t *d_val = dalloc<uint8_t>(val.size());
        hipError_t e1 = hipMemcpyAsync(d_idx, idx.data(), idx.size() * 4, hipMemcpyHostToDevice, g->stream);
        hipError_t e2 = hipMemcpyAsync(d_val, val.data(), val.size(), hipMemcpyHostToDevice, g->stream);
        if (e1 == hipSuccess && e2 == hipSuccess) {
            hipLaunchKernelGGL(k_paint, dim3((unsigned)((idx.size() + 255) / 256)), dim3(256), 0, g->stream, d_idx,
                               d_val, idx.size(), g->mask, (int)g->X, g->geom.pitch);
            e1 = hipGetLastError();
        }
        hipError_t e3 = hipStreamSynchronize(g->stream);
        (void)hipFree(d_idx);
        (void)hipFree(d_val);
        HIPCHK(e1);
        HIPCHK(e2);
        HIPCHK(e3);
        g->build_flags();
    });
}

int lbm_set_mask(lbm_grid *g, const uint8_t *solid) {
    return guarded(g, [&] {
        check(solid != nullptr, "null mask");
        std::vector<uint8_t> m((size_t)g->X * g->Y);
        for (size_t i = 0; i < m.size(); ++i) m[i] = solid[i] == 1 ? 1 : 0;
        g->flush();
        HIPCHK(hipMemcpy2DAsync(g->mask, g->geom.pitch, m.data(), g->X, g->X, g->Y, hipMemcpyHostToDevice,
                                g->stream));
        g->build_flags();
        HIPCHK(hipStreamSynchronize(g->stream));
    });
}

int lbm_get_mask(lbm_grid *g, uint8_t *solid) {
    return guarded(g, [&] {
        check(solid != nullptr, "null mask");
        HIPCHK(hipMemcpy2DAsync(solid, g->X, g->mask, g->geom.pitch, g->X, g->Y, hipMemcpyDeviceToHost, g->stream));
        HIPCHK(hipStreamSynchronize(g->stream));
    });
}

int lbm_step(lbm_grid *g, uint32_t n) {
    return guarded(g, [&] { g->step(n); });
}

int lbm_step_timed(lbm_grid *g, uint32_t n, float *ms) {
    return guarded(g, [&] {
        check(ms != nullptr, "null output");
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        hipError_t err = hipEventCreate(&e1);
        if (err != hipSuccess) {
            (void)hipEventDestroy(e0);
            throw HipError{err, "hipEventCreate"};
        }
        try {
            HIPCHK(hipEventRecord(e0, g->stream));
            g->step(n);
            HIPCHK(hipEventRecord(e1, g->stream));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(ms, e0, e1));
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int lbm_set_variant(lbm_grid *g, int variant) {
    return guarded(g, [&] {
        check(variant >= 0 && variant < LBM_NUM_VARIANTS, "unknown kernel variant");
        HIPCHK(hipStreamSynchronize(g->stream));
        if (g->pair_variant(variant)) g->ensure_pair_buffers();  // may fail: then nothing changes
        else g->release_pair_buffers();
        g->variant = variant;
    });
}

int lbm_get_variant(const lbm_grid *g, int *variant) {
    if (!g || !variant) return LBM_EINVAL;
    *variant = g->variant;
    return LBM_OK;
}

int lbm_autotune(lbm_grid *g, uint32_t steps, int *best, float *ms) {
    return guarded(g, [&] {
        check(steps > 0, "steps must be > 0");
        // snapshot the whole evolving state so that tuning leaves no trace
        const size_t sb = g->set_bytes(), lb = 3 * (size_t)g->geom.pstride * sizeof(float);
        char *snap = dalloc<char>(2 * sb + lb);
        const int cur = g->cur;
        const bool pform = g->pform;
        const uint64_t nsteps = g->steps;
        const int variant = g->variant;
        float t[LBM_NUM_VARIANTS];
        int v_best = 0;
        bool saved = false;
        try {
            HIPCHK(hipMemcpyAsync(snap, g->f[0], sb, hipMemcpyDeviceToDevice, g->stream));
            HIPCHK(hipMemcpyAsync(snap + sb, g->f[1], sb, hipMemcpyDeviceToDevice, g->stream));
            HIPCHK(hipMemcpyAsync(snap + 2 * sb, g->lag, lb, hipMemcpyDeviceToDevice, g->stream));
            saved = true;
            hipEvent_t e0, e1;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            for (int v = 0; v < LBM_NUM_VARIANTS; ++v) {
                if (g->pair_variant(v) && steps < 3) {
                    // a call of fewer than three steps never launches the
                    // two-step kernel: it would time the one-step fallback
                    t[v] = INFINITY;
                    continue;
                }
                if (g->pair_variant(v)) {
                    try {
                        g->ensure_pair_buffers();
                    } catch (...) {  // no memory for two more sets: not a candidate
                        t[v] = INFINITY;
                        continue;
                    }
                }
                g->variant = v;
                g->step(2);  // warm-up (and leaves the grid in P-form)
                HIPCHK(hipEventRecord(e0, g->stream));
                g->step(steps);
                HIPCHK(hipEventRecord(e1, g->stream));
                HIPCHK(hipEventSynchronize(e1));
                HIPCHK(hipEventElapsedTime(&t[v], e0, e1));
                if (t[v] < t[v_best]) v_best = v;
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            HIPCHK(hipMemcpyAsync(g->f[0], snap, sb, hipMemcpyDeviceToDevice, g->stream));
            HIPCHK(hipMemcpyAsync(g->f[1], snap + sb, sb, hipMemcpyDeviceToDevice, g->stream));
            HIPCHK(hipMemcpyAsync(g->lag, snap + 2 * sb, lb, hipMemcpyDeviceToDevice, g->stream));
            HIPCHK(hipStreamSynchronize(g->stream));
        } catch (...) {
            // put the grid back as it was (the two-step kernel trades the
            // state sets with its spare pair, so the state may live in any
            // two of the four: the copies restore whichever two are current)
            (void)hipStreamSynchronize(g->stream);
            if (saved) {
                (void)hipMemcpyAsync(g->f[0], snap, sb, hipMemcpyDeviceToDevice, g->stream);
                (void)hipMemcpyAsync(g->f[1], snap + sb, sb, hipMemcpyDeviceToDevice, g->stream);
                (void)hipMemcpyAsync(g->lag, snap + 2 * sb, lb, hipMemcpyDeviceToDevice, g->stream);
                (void)hipStreamSynchronize(g->stream);
            }
            (void)hipFree(snap);
            g->cur = cur;
            g->pform = pform;
            g->steps = nsteps;
            g->variant = variant;
            if (!g->pair_variant(variant)) g->release_pair_buffers();
            throw;
        }
        (void)hipFree(snap);
        g->cur = cur;
        g->pform = pform;
        g->steps = nsteps;
        g->variant = v_best;
        if (!g->pair_variant(v_best)) g->release_pair_buffers();  // tried, not kept
        if (best) *best = v_best;
        if (ms)
            for (int v = 0; v < LBM_NUM_VARIANTS; ++v) ms[v] = t[v] / steps;
    });
}

int lbm_autotune_cached(lbm_grid *g, const char *cache_path, uint32_t steps, int *best, int *hit) {
    std::string key;
    tune_cache::Entry e;
    bool found = false;
    int rc = guarded(g, [&] {
        check(cache_path && *cache_path, "null or empty cache path");
        check(steps > 0, "steps must be > 0");
        key = tune_cache::clean_key(device_tag(g->dev) + "/grid/" + std::to_string(tune_cache::pow2_ceil(g->X)) +
                                    "x" + std::to_string(tune_cache::pow2_ceil(g->Y)) + "/" +
                                    precision_tag(g->precision) +
                                    (g->boundary == LBM_BOUNDARY_REFERENCE ? "/ref" : "/zg") + "/v" +
                                    std::to_string(LBM_NUM_VARIANTS));
        found = tune_cache::lookup(cache_path, key, e) && e.variant >= 0 && e.variant < LBM_NUM_VARIANTS;
    });
    if (rc != LBM_OK) return rc;
    if (hit) *hit = 0;
    // a cached two-step variant whose spare sets cannot be allocated here is a miss
    if (found && lbm_set_variant(g, e.variant) == LBM_OK) {
        if (hit) *hit = 1;
        if (best) *best = e.variant;
        g->err.clear();
        return LBM_OK;
    }
    e.ms.assign(LBM_NUM_VARIANTS, 0.0f);
    e.steps = steps;
    rc = lbm_autotune(g, steps, &e.variant, e.ms.data());
    if (rc != LBM_OK) return rc;
    if (best) *best = e.variant;
    return guarded(g, [&] { check(tune_cache::store(cache_path, key, e), "cannot write the autotune cache file"); });
}

int lbm_snapshot(lbm_grid *g) {
    return guarded(g, [&] {
        const size_t sb = g->set_bytes(), lb = 3 * (size_t)g->geom.pstride * sizeof(float);
        if (!g->snap) g->snap = dalloc<char>(2 * sb + lb);
        HIPCHK(hipMemcpyAsync(g->snap, g->f[0], sb, hipMemcpyDeviceToDevice, g->stream));
        HIPCHK(hipMemcpyAsync(g->snap + sb, g->f[1], sb, hipMemcpyDeviceToDevice, g->stream));
        HIPCHK(hipMemcpyAsync(g->snap + 2 * sb, g->lag, lb, hipMemcpyDeviceToDevice, g->stream));
        g->snap_cur = g->cur;
        g->snap_pform = g->pform;
        g->snap_steps = g->steps;
        g->snap_valid = true;
    });
}

int lbm_restore(lbm_grid *g) {
    return guarded(g, [&] {
        check(g->snap_valid, "no snapshot taken");
        const size_t sb = g->set_bytes(), lb = 3 * (size_t)g->geom.pstride * sizeof(float);
        HIPCHK(hipMemcpyAsync(g->f[0], g->snap, sb, hipMemcpyDeviceToDevice, g->stream));
        HIPCHK(hipMemcpyAsync(g->f[1], g->snap + sb, sb, hipMemcpyDeviceToDevice, g->stream));
        HIPCHK(hipMemcpyAsync(g->lag, g->snap + 2 * sb, lb, hipMemcpyDeviceToDevice, g->stream));
        g->cur = g->snap_cur;
        g->pform = g->snap_pform;
        g->steps = g->snap_steps;
    });
}

int lbm_velocity_mark_slot(lbm_grid *g, int slot) {
    return guarded(g, [&] {
        check(slot == 0 || slot == 1, "slot must be 0 or 1");
        if (!g->umark) g->umark = dalloc<float>(4 * (size_t)g->geom.pstride);  // two slots of (ux, uy)
        g->moments_now(false);
        hipLaunchKernelGGL(k_velocity_mark, dim3((g->X + 255) / 256, g->Y), dim3(256), 0, g->stream, g->now,
                           g->umark + 2 * (size_t)slot * g->geom.pstride, g->geom);
        HIPCHK(hipGetLastError());
        g->umarked |= 1 << slot;
    });
}

int lbm_velocity_mark(lbm_grid *g) { return lbm_velocity_mark_slot(g, 0); }

int lbm_velocity_skill(lbm_grid *g, double out[2]) {
    return guarded(g, [&] {
        check(out != nullptr && g->umarked == 3, "mark both velocity slots first");
        const dim3 grid((g->X + kBlock - 1) / kBlock, g->Y);
        g->ensure_vpart(grid);
        g->moments_now(false);
        hipLaunchKernelGGL(k_velocity_skill, grid, dim3(kBlock), 0, g->stream, g->now, g->umark,
                           g->umark + 2 * (size_t)g->geom.pstride, g->mask, g->geom, g->vpart);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, g->stream, g->vpart, g->nvpart, g->sums);
        HIPCHK(hipGetLastError());
        double s[3];
        HIPCHK(hipMemcpyAsync(s, g->sums, sizeof s, hipMemcpyDeviceToHost, g->stream));
        HIPCHK(hipStreamSynchronize(g->stream));
        auto ratio = [](double num, double den) {
            if (!std::isfinite(num) || !std::isfinite(den)) return (double)INFINITY;
            if (den > 0.0) return std::sqrt(num / den);
            return num > 0.0 ? (double)INFINITY : 0.0;
        };
        out[0] = ratio(s[0], s[1]);  // forecast error over persistence error
        out[1] = ratio(s[0], s[2]);  // forecast error relative to the forecast
    });
}

int lbm_velocity_change(lbm_grid *g, double *rel) {
    return guarded(g, [&] {
        check(rel != nullptr && (g->umarked & 1), "call lbm_velocity_mark first");
        const dim3 grid((g->X + kBlock - 1) / kBlock, g->Y);
        g->ensure_vpart(grid);
        g->moments_now(false);
        hipLaunchKernelGGL(k_velocity_change, grid, dim3(kBlock), 0, g->stream, g->now, g->umark, g->mask, g->geom,
                           g->vpart);
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, g->stream, g->vpart, g->nvpart, g->sums);
        HIPCHK(hipGetLastError());
        double s[3];
        HIPCHK(hipMemcpyAsync(s, g->sums, sizeof s, hipMemcpyDeviceToHost, g->stream));
        HIPCHK(hipStreamSynchronize(g->stream));
        // a non-finite state (a diverged prediction) must read as an infinite change
        if (!std::isfinite(s[0]) || !std::isfinite(s[1])) *rel = INFINITY;
        else if (s[1] > 0.0) *rel = std::sqrt(s[0] / s[1]);
        else *rel = s[0] > 0.0 ? INFINITY : 0.0;
    });
}

int lbm_surrogate_jump(lbm_grid *g, lbm_unet *net, uint32_t steps, float s_rho, float s_u, float d_rho,
                       float d_u) {
    return guarded(g, [&] {
        check(net != nullptr && s_rho > 0.0f && s_u > 0.0f && d_rho > 0.0f && d_u > 0.0f,
              "bad network or scales");
        int b, h, w;
        unet_shape(net, &b, &h, &w);
        const bool fine = b == 1 && h == (int)g->Y && w == (int)g->X;
        const bool coarse = b == 1 && 2 * h == (int)g->Y && 2 * w == (int)g->X;
        check(fine || coarse, "network shape must be (1, ny, nx) or (1, ny / 2, nx / 2)");
        check(unet_device(net) == g->dev, "network and grid on different devices");
        g->flush_with_moments();  // S-form: buf[cur] is the state to predict from and to overwrite
        unet_begin_use(net, g->stream);  // after any use of the network on another stream
        const dim3 pgrid((w + 255) / 256, h);
        if (fine)
            hipLaunchKernelGGL(k_pack_net_input<1>, pgrid, dim3(256), 0, g->stream, g->now, g->mask,
                               unet_input_tensor(net), g->geom, 32, 1.0f / s_rho, 1.0f / s_u, g->u_in);
        else
            hipLaunchKernelGGL(k_pack_net_input<2>, pgrid, dim3(256), 0, g->stream, g->now, g->mask,
                               unet_input_tensor(net), g->geom, 32, 1.0f / s_rho, 1.0f / s_u, g->u_in);
        HIPCHK(hipGetLastError());
        unet_forward_on(net, g->stream);
        const dim3 grid((g->X + 255) / 256, g->Y);
        const float *o = unet_output_tensor(net);
        by_store(g->precision, [&]<class St>() {
            using T = typename St::T;
            if (fine)
                hipLaunchKernelGGL((k_apply_jump<St, 1>), grid, dim3(256), 0, g->stream, o, 32, g->now, g->mask,
                                   (T *)g->f[g->cur], g->geom, d_rho, d_u);
            else
                hipLaunchKernelGGL((k_apply_jump<St, 2>), grid, dim3(256), 0, g->stream, o, 32, g->now, g->mask,
                                   (T *)g->f[g->cur], g->geom, d_rho, d_u);
        });
        HIPCHK(hipGetLastError());
        unet_end_use(net, g->stream);
        g->steps += steps;
    });
}

int lbm_sync(lbm_grid *g) {
    return guarded(g, [&] { HIPCHK(hipStreamSynchronize(g->stream)); });
}

int lbm_step_count(const lbm_grid *g, uint64_t *n) {
    if (!g || !n) return LBM_EINVAL;
    *n = g->steps;
    return LBM_OK;
}

int lbm_get_stream(const lbm_grid *g, void **stream) {
    if (!g || !stream) return LBM_EINVAL;
    *stream = (void *)g->stream;
    return LBM_OK;
}

int lbm_read_moments(lbm_grid *g, float *jx, float *jy, float *rho, int lagged) {
    return guarded(g, [&] {
        const float *src = g->lag;
        if (!lagged) {
            g->moments_now(false);
            src = g->now;
        }
        const size_t ps = g->geom.pstride;
        if (jx) g->copy_f32_plane_out(src, jx);
        if (jy) g->copy_f32_plane_out(src + ps, jy);
        if (rho) g->copy_f32_plane_out(src + 2 * ps, rho);
        HIPCHK(hipStreamSynchronize(g->stream));
    });
}

int lbm_read_fields(lbm_grid *g, float *rho, float *ux, float *uy, float *curl, int lagged) {
    return guarded(g, [&] {
        const float *src = g->lag;
        if (!lagged) {
            g->moments_now(false);
            src = g->now;
        }
        g->ensure_dense();
        hipLaunchKernelGGL(k_fields, dim3((g->X + 255) / 256, g->Y), dim3(256), 0, g->stream, src, g->dense, g->geom);
        HIPCHK(hipGetLastError());
        const size_t dn = (size_t)g->X * g->Y;
        float *dst[4] = {rho, ux, uy, curl};
        for (int i = 0; i < 4; ++i)
            if (dst[i])
                HIPCHK(hipMemcpyAsync(dst[i], g->dense + i * dn, sizeof(float) * dn, hipMemcpyDeviceToHost, g->stream));
        HIPCHK(hipStreamSynchronize(g->stream));
    });
}

int lbm_summary(lbm_grid *g, int kind, float *out) {
    return guarded(g, [&] {
        check(kind >= 0 && kind <= 4, "unknown summary kind");
        hipLaunchKernelGGL(k_summary, dim3((g->X + 255) / 256, g->Y), dim3(256), 0, g->stream, g->lag, g->out, g->mask,
                           (float *)nullptr, g->geom, kind, 0);
        HIPCHK(hipGetLastError());
        if (out) g->copy_f32_plane_out(g->out, out);
        HIPCHK(hipStreamSynchronize(g->stream));
    });
}

int lbm_render(lbm_grid *g, int kind, int colormap, float *rgb) {
    return guarded(g, [&] {
        check(kind >= 0 && kind <= 4, "unknown summary kind");
        check(colormap >= LBM_COLORMAP_INFERNO && colormap <= LBM_COLORMAP_JET, "unknown colour map");
        check(rgb != nullptr, "null output");
        const size_t n = (size_t)g->X * g->Y * 3;
        if (g->rgb_elems < n) {
            if (g->rgb) (void)hipFree(g->rgb);
            g->rgb = nullptr;
            g->rgb = dalloc<float>(n);
            g->rgb_elems = n;
        }
        hipLaunchKernelGGL(k_summary, dim3((g->X + 255) / 256, g->Y), dim3(256), 0, g->stream, g->lag, g->out, g->mask,
                           g->rgb, g->geom, kind, colormap);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(rgb, g->rgb, n * sizeof(float), hipMemcpyDeviceToHost, g->stream));
        HIPCHK(hipStreamSynchronize(g->stream));
    });
}

int lbm_conservation(lbm_grid *g, double out[3]) {
    int rc = guarded(g, [&] {
        check(out != nullptr, "null output");
        g->moments_now(true);
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, g->stream, g->partials, g->last_partials,
                           g->sums);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, g->sums, 3 * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        HIPCHK(hipStreamSynchronize(g->stream));
    });
    if (rc == LBM_OK && !(std::isfinite(out[0]) && std::isfinite(out[1]) && std::isfinite(out[2])))
        return fail(g, LBM_ENONFINITE, "non-finite mass or momentum");
    return rc;
}

}  // extern "C"

// ---- FP64 sums without allocation ------------------------------------------------
//
// lbm_conservation_static's partial sums live in a file-scope device array of a
// fixed size instead of buffers the grid allocates on first use.  In host code
// `scratch` names a host-side stand-in, not device memory: its device address
// comes from hipGetSymbolAddress, taken once per device (with that device
// current) on the entry point's first call there, and only that address is
// ever handed to a kernel.  Every call on one device shares the array, which
// keeps the last call's contents (each call writes every slot it then reads).
constexpr size_t kStaticSlots = (size_t)1 << 18;  // reduction blocks of one pass (6 MB of partial sums)
static __device__ __align__(256) double scratch[3 * kStaticSlots + 4];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};

double *scratch_on(int dev) {
    check(dev >= 0 && dev < kMaxDevices, "device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the caller made `dev` current
        g_scratch_ptr[dev] = (double *)p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int lbm_conservation_static(lbm_grid *g, double out[3]) {
    int rc = guarded(g, [&] {
        check(out != nullptr, "null output");
        // the same pass shape as lbm_conservation's moments pass, so the same
        // per-block partial sums in the same order: the same bits
        const bool f16p = g->pform && g->precision != LBM_PRECISION_F32;
        const size_t blocks = (size_t)g->shape(f16p ? 2 : 1).nblocks;
        if (blocks > kStaticSlots)
            throw ArgError{"grid too large for lbm_conservation_static: its reduction needs " +
                           std::to_string(blocks) + " partial sums, the static scratch holds " +
                           std::to_string(kStaticSlots) + " (use lbm_conservation)"};
        double *part = scratch_on(g->dev);
        double *sums = part + 3 * kStaticSlots;
        if (g->pform) g->cell<true, OP_REDUCE>(g->f[g->cur], g->f[1 - g->cur], nullptr, part);
        else g->cell<false, OP_REDUCE>(g->f[g->cur], nullptr, nullptr, part);
        check(g->last_partials == blocks, "reduction pass shape mismatch");
        hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(1024), 0, g->stream, part, blocks, sums);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out, sums, 3 * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        HIPCHK(hipStreamSynchronize(g->stream));
    });
    if (rc == LBM_OK && !(std::isfinite(out[0]) && std::isfinite(out[1]) && std::isfinite(out[2])))
        return fail(g, LBM_ENONFINITE, "non-finite mass or momentum");
    return rc;
}

int lbm_read_f(lbm_grid *g, float *f9) {
    return guarded(g, [&] {
        check(f9 != nullptr, "null output");
        g->flush();
        const size_t dn = (size_t)g->X * g->Y;
        for (int k = 0; k < 9; ++k) g->copy_plane_out(g->plane(g->cur, k), f9 + k * dn, k);
    });
}

int lbm_write_f(lbm_grid *g, const float *f9) {
    return guarded(g, [&] {
        check(f9 != nullptr, "null input");
        g->pform = false;
        const size_t dn = (size_t)g->X * g->Y;
        for (int s = 0; s < 2; ++s)
            for (int k = 0; k < 9; ++k) g->copy_plane_in(f9 + k * dn, g->plane(s, k), k);
    });
}

// checkpoint image: header, both sets, mask, fields, summary output
namespace {
struct StateHeader {
    char magic[8];
    uint32_t X, Y;
    int32_t precision, boundary;
    float omega;
    int32_t cur, pform;
    uint32_t pitch;
    long long pstride;
    uint64_t steps;
    float u_in;  // the inflow speed of the last uniform reset (the surrogate's input)
    uint32_t reserved;
};
constexpr char kStateMagic[9] = "LBMHIP02";
}  // namespace

int lbm_state_size(const lbm_grid *g, size_t *bytes) {
    if (!g || !bytes) return LBM_EINVAL;
    const size_t ps = (size_t)g->geom.pstride;
    *bytes = sizeof(StateHeader) + 2 * g->set_bytes() + ps + 3 * ps * sizeof(float) + ps * sizeof(float);
    return LBM_OK;
}

int lbm_save_state(lbm_grid *g, void *buf, size_t bytes) {
    return guarded(g, [&] {
        size_t need = 0;
        lbm_state_size(g, &need);
        check(buf != nullptr && bytes >= need, "buffer too small");
        StateHeader h{};
        std::memcpy(h.magic, kStateMagic, 8);
        h.X = g->X;
        h.Y = g->Y;
        h.precision = g->precision;
        h.boundary = g->boundary;
        h.omega = g->omega;
        h.cur = g->cur;
        h.pform = g->pform;
        h.pitch = (uint32_t)g->geom.pitch;
        h.pstride = g->geom.pstride;
        h.steps = g->steps;
        h.u_in = g->u_in;
        char *p = (char *)buf;
        std::memcpy(p, &h, sizeof h);
        p += sizeof h;
        const size_t ps = (size_t)g->geom.pstride;
        HIPCHK(hipMemcpyAsync(p, g->f[0], g->set_bytes(), hipMemcpyDeviceToHost, g->stream));
        p += g->set_bytes();
        HIPCHK(hipMemcpyAsync(p, g->f[1], g->set_bytes(), hipMemcpyDeviceToHost, g->stream));
        p += g->set_bytes();
        HIPCHK(hipMemcpyAsync(p, g->mask, ps, hipMemcpyDeviceToHost, g->stream));
        p += ps;
        HIPCHK(hipMemcpyAsync(p, g->lag, 3 * ps * sizeof(float), hipMemcpyDeviceToHost, g->stream));
        p += 3 * ps * sizeof(float);
        HIPCHK(hipMemcpyAsync(p, g->out, ps * sizeof(float), hipMemcpyDeviceToHost, g->stream));
        HIPCHK(hipStreamSynchronize(g->stream));
    });
}

int lbm_load_state(lbm_grid *g, const void *buf, size_t bytes) {
    return guarded(g, [&] {
        size_t need = 0;
        lbm_state_size(g, &need);
        check(buf != nullptr && bytes >= need, "buffer too small");
        StateHeader h;
        std::memcpy(&h, buf, sizeof h);
        check(std::memcmp(h.magic, kStateMagic, 8) == 0, "not a state image");
        check(h.X == g->X && h.Y == g->Y && h.precision == g->precision && h.boundary == g->boundary &&
                  (int)h.pitch == g->geom.pitch && h.pstride == g->geom.pstride,
              "state image does not match this grid");
        check((h.cur == 0 || h.cur == 1) && (h.pform == 0 || h.pform == 1) && std::isfinite(h.omega) &&
                  std::isfinite(h.u_in),
              "corrupt state image header");
        const char *p = (const char *)buf + sizeof h;
        const size_t ps = (size_t)g->geom.pstride;
        HIPCHK(hipMemcpyAsync(g->f[0], p, g->set_bytes(), hipMemcpyHostToDevice, g->stream));
        p += g->set_bytes();
        HIPCHK(hipMemcpyAsync(g->f[1], p, g->set_bytes(), hipMemcpyHostToDevice, g->stream));
        p += g->set_bytes();
        HIPCHK(hipMemcpyAsync(g->mask, p, ps, hipMemcpyHostToDevice, g->stream));
        p += ps;
        HIPCHK(hipMemcpyAsync(g->lag, p, 3 * ps * sizeof(float), hipMemcpyHostToDevice, g->stream));
        p += 3 * ps * sizeof(float);
        HIPCHK(hipMemcpyAsync(g->out, p, ps * sizeof(float), hipMemcpyHostToDevice, g->stream));
        g->omega = h.omega;
        g->geom.omega = h.omega;
        g->set_inflow(h.u_in);
        g->cur = h.cur;
        g->pform = h.pform != 0;
        g->steps = h.steps;
        g->build_flags();
        HIPCHK(hipStreamSynchronize(g->stream));
    });
}

// ---- generator -------------------------------------------------------------------

uint64_t lbm_splitmix64(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static uint32_t draw_below(uint64_t *s, uint32_t m) {
    return (uint32_t)(((lbm_splitmix64(s) >> 32) * (uint64_t)m) >> 32);
}

int lbm_gen_circles(uint8_t *solid, uint32_t nx, uint32_t ny, uint64_t seed, uint32_t n_circles, uint32_t r_min,
                    uint32_t r_max) {
    if (!solid || nx < 3 || ny < 3 || r_min > r_max || r_max + 2 > nx) return LBM_EINVAL;
    uint64_t s = seed;
    for (uint32_t c = 0; c < n_circles; ++c) {
        const int64_t r = r_min + draw_below(&s, r_max - r_min + 1);
        const int64_t cxv = r + 1 + draw_below(&s, nx - (uint32_t)r - 1);
        const int64_t cyv = draw_below(&s, ny);
        const int64_t y0 = std::max<int64_t>(0, cyv - r), y1 = std::min<int64_t>(ny - 1, cyv + r);
        const int64_t x0 = std::max<int64_t>(0, cxv - r), x1 = std::min<int64_t>(nx - 1, cxv + r);
        for (int64_t y = y0; y <= y1; ++y)
            for (int64_t x = x0; x <= x1; ++x)
                if ((x - cxv) * (x - cxv) + (y - cyv) * (y - cyv) <= r * r) solid[(size_t)y * nx + x] = 1;
    }
    return LBM_OK;
}

}  // extern "C"
