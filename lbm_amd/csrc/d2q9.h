// D2Q9 lattice constants and the reference-exact BGK collision, shared by the
// single-grid and the batched kernels.
//
// Plane order k = 0 NW, 1 N, 2 NE, 3 W, 4 C, 5 E, 6 SW, 7 S, 8 SE follows the
// reference (lbm-wgpu src/lbm.rs:632-640).  Memory rows grow downwards, so N moves
// to row y-1: c_k = (CX[k], CY[k]) in (column, row) units.
//
// collide_ref() restates the reference's four collide passes for one cell with
// the same f32 expression trees (rewritten_shaders/pre_collision/
// corner_pre_collision.wgsl:19-21, cardinal_pre_collision.wgsl:19-21,
// collision/corner_collision.wgsl:24-42, collision/cardinal_collision.wgsl:25-43).
// Files including it are compiled with -ffp-contract=off so that no multiply-add
// is fused and the results match the reference's arithmetic bit for bit;
// division is IEEE correctly rounded (hipcc's default for f32).
#pragma once

#include <stdint.h>

#include <cmath>

namespace d2q9 {

enum : int { NW = 0, N = 1, NE = 2, W = 3, C = 4, E = 5, SW = 6, S = 7, SE = 8 };

__host__ __device__ constexpr int cx(int k) { return (k % 3) - 1; }
__host__ __device__ constexpr int cy(int k) { return (k / 3) - 1; }
__host__ __device__ constexpr int opp(int k) { return 8 - k; }
// bit of a moving population in the per-cell bounce-back byte
__host__ __device__ constexpr int bit(int k) { return k < 4 ? k : k - 1; }

// lattice weights as f32 (the storage shift of the f16 format)
__host__ __device__ constexpr float weight(int k) {
    return k == 4 ? (float)(4.0 / 9.0) : ((k & 1) ? (float)(1.0 / 9.0) : (float)(1.0 / 36.0));
}

struct Moments {
    float jx, jy, rho;
};

// Pre-collision moments in the reference's summation order.  with_rest = false
// gives the state after the two pre-collision passes only (reset_to_equilibrium
// runs just those, lbm.rs:1076-1088).
__host__ __device__ __forceinline__ Moments moments_ref(const float f[9], bool with_rest = true) {
    float jx = ((f[NE] + f[SE]) - f[NW]) - f[SW];
    float jy = ((f[NE] + f[NW]) - f[SE]) - f[SW];
    float rho = ((f[NE] + f[SE]) + f[NW]) + f[SW];
    jx = jx + (f[E] - f[W]);
    jy = jy + (f[N] - f[S]);
    rho = rho + (((f[E] + f[N]) + f[S]) + f[W]);
    if (with_rest) rho = rho + f[C];
    return {jx, jy, rho};
}

// BGK relaxation f_k += omega (feq_k - f_k) with the reference's expression trees.
__host__ __device__ __forceinline__ Moments collide_ref(float f[9], float omega) {
    const Moments m = moments_ref(f, true);
    const float thisrho = m.rho;
    const float thisux = m.jx / thisrho;
    const float thisuy = m.jy / thisrho;
    const float one36thrho = (float)(1.0 / 36.0) * thisrho;
    const float one9thrho = (float)(1.0 / 9.0) * thisrho;
    const float ux3 = 3.0f * thisux;
    const float uy3 = 3.0f * thisuy;
    const float ux2 = thisux * thisux;
    const float uy2 = thisuy * thisuy;
    const float uxuy2 = 2.0f * thisux * thisuy;
    const float u2 = ux2 + uy2;
    const float u215 = 1.5f * u2;
    f[NE] += omega * (one36thrho * (1.0f + ux3 + uy3 + 4.5f * (u2 + uxuy2) - u215) - f[NE]);
    f[SE] += omega * (one36thrho * (1.0f + ux3 - uy3 + 4.5f * (u2 - uxuy2) - u215) - f[SE]);
    f[NW] += omega * (one36thrho * (1.0f - ux3 + uy3 + 4.5f * (u2 - uxuy2) - u215) - f[NW]);
    f[SW] += omega * (one36thrho * (1.0f - ux3 - uy3 + 4.5f * (u2 + uxuy2) - u215) - f[SW]);
    f[C] += omega * ((float)(4.0 / 9.0) * thisrho * (1.0f - u215) - f[C]);
    f[E] += omega * (one9thrho * (1.0f + ux3 + 4.5f * ux2 - u215) - f[E]);
    f[W] += omega * (one9thrho * (1.0f - ux3 + 4.5f * ux2 - u215) - f[W]);
    f[N] += omega * (one9thrho * (1.0f + uy3 + 4.5f * uy2 - u215) - f[N]);
    f[S] += omega * (one9thrho * (1.0f - uy3 + 4.5f * uy2 - u215) - f[S]);
    return m;
}

// The reference's host equilibrium set_equil(ux, uy, rho) (lbm.rs:611-643), with
// one sign corrected: lbm.rs:639 computes plane 7 (S) with "- uy_2" where the
// D2Q9 equilibrium has "+ 4.5 uy^2".  The reference only calls set_equil with
// uy = 0, where both forms give the same bits; for uy != 0 this one is right.
inline void host_feq(float ux, float uy, float rho, float out[9]) {
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
    const float rho_ninth = rho / 9.0f;
    const float rho_36th = rho / 36.0f;
    out[0] = rho_36th * (1.0f - ux + uy + u_sum_sq_neg - u_dot_product);
    out[1] = rho_ninth * (1.0f + uy + uy_2 - u_dot_product);
    out[2] = rho_36th * (1.0f + ux + uy + u_sum_sq_pos - u_dot_product);
    out[3] = rho_ninth * (1.0f - ux + ux_2 - u_dot_product);
    out[4] = 4.0f * rho_ninth * (1.0f - u_dot_product);
    out[5] = rho_ninth * (1.0f + ux + ux_2 - u_dot_product);
    out[6] = rho_36th * (1.0f - ux - uy + u_sum_sq_pos - u_dot_product);
    out[7] = rho_ninth * (1.0f - uy + uy_2 - u_dot_product);
    out[8] = rho_36th * (1.0f + ux - uy + u_sum_sq_neg - u_dot_product);
}

// BGK in shifted variables g_k = f_k - w_k ("DDF shifting": Lehmann, Krause,
// Amati, Sega, Harting, Gekle, Phys. Rev. E 106, 015308 (2022)).  Since
// sum_k w_k = 1 and sum_k c_k w_k = 0:
//   rho - 1 = sum_k g_k,  j = sum_k c_k g_k,
//   geq_k = feq_k - w_k = w_k (rho - 1 - 1.5 rho u.u + rho c.u (3 + 4.5 c.u)).
// Keeping the O(1) part w_k out of the arithmetic keeps the small deviations at
// full f32 resolution.  Used with the f16 storage, whose error budget is stated
// against an FP64 solver; it is not the reference's expression tree.
__host__ __device__ __forceinline__ Moments moments_shifted(const float g[9], bool with_rest = true) {
    const float corner = ((g[NW] + g[NE]) + (g[SW] + g[SE]));
    const float axis = ((g[N] + g[S]) + (g[W] + g[E]));
    float drho = corner + axis;
    if (with_rest) drho += g[C];
    const float jx = ((g[NE] + g[E]) + g[SE]) - ((g[NW] + g[W]) + g[SW]);
    const float jy = ((g[NW] + g[N]) + g[NE]) - ((g[SW] + g[S]) + g[SE]);
    // rho without the rest population: 1 - w_C + the shifted moving populations
    const float rho = with_rest ? 1.0f + drho : (float)(5.0 / 9.0) + drho;
    return {jx, jy, rho};
}

__device__ __forceinline__ Moments collide_shifted(float g[9], float omega) {
    const Moments m = moments_shifted(g, true);
    const float drho = m.rho - 1.0f;
    const float inv = __builtin_amdgcn_rcpf(m.rho);
    const float ux = m.jx * inv;
    const float uym = -m.jy * inv;  // memory-row velocity: north is row - 1
    const float base = fmaf(-1.5f * m.rho, fmaf(ux, ux, uym * uym), drho);
    const float r3 = 3.0f * m.rho, r45 = 4.5f * m.rho;
    // g + omega (geq - g), not (1 - omega) g + omega geq: near the free stream
    // geq - g is small and exact, and only its product is rounded, where the
    // other form rounds at the scale of g (the F16E format, whose stored
    // deviations there are tiny, lost accuracy against FP64 with it)
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        // c_k . u with its zero terms left out (the compiler keeps 0 * u, whose
        // sign follows u's): the rest population's geq is w_C * base
        float geq = weight(k) * base;
        if (cx(k) != 0 || cy(k) != 0) {
            const float cu = cx(k) == 0   ? (float)cy(k) * uym
                             : cy(k) == 0 ? (float)cx(k) * ux
                                          : (float)cx(k) * ux + (float)cy(k) * uym;
            geq = weight(k) * fmaf(cu, fmaf(r45, cu, r3), base);
        }
        g[k] = fmaf(omega, geq - g[k], g[k]);
    }
    return m;
}

// ---- storage formats ------------------------------------------------------------
// Each format stores a "working" variable per population: F32 the population
// f_k itself (collided with the reference's expression trees, bit for bit);
// F16S g_k = f_k - w_k as IEEE binary16, converted with round-to-nearest-even,
// collided in f32 in shifted form.  to_f / from_f convert to and from f_k for
// uploads, downloads and fills.
//
// The binary16 storage may also hold f_k - s_k with s_k = feq_k(1, (u_s, 0)),
// the free-stream equilibrium of a channel whose inflow speed is u_s: where the
// flow is close to the free stream the stored values are then small, and
// binary16's relative rounding error costs proportionally less velocity.  The
// arithmetic stays in the variables of f - w: a value is shifted by
// d_k = s_k - w_k when it is loaded and back before it is rounded (StoreF16E).
// on = 0 (and d = 0) describes the plain f - w format (StoreF16S).
//
// d_k is rounded to a multiple of 2^-20.  Every binary16 value is a multiple of
// 2^-24, so for stored values below 16 in magnitude the f32 sum h + d_k (and
// h + d_opp - d_k) is exact: decoding loses nothing, re-encoding an unchanged
// value gives back the same bits, and a value that half-way bounce-back hands
// from population opp(k) to k is re-encoded the same way on every code path.
struct F16Shift {
    float d[9];
    int on;
};

inline F16Shift make_f16_shift(bool on, double us) {
    F16Shift s{};
    s.on = on ? 1 : 0;
    if (!on) return s;
    for (int k = 0; k < 9; ++k) {
        const double w = k == 4 ? 4.0 / 9.0 : ((k & 1) ? 1.0 / 9.0 : 1.0 / 36.0);
        const double cu = (double)((k % 3) - 1) * us;  // the shift flow is along +x
        const double d = w * (3.0 * cu + 4.5 * cu * cu - 1.5 * us * us);
        s.d[k] = (float)(std::nearbyint(d * 1048576.0) / 1048576.0);
    }
    return s;
}

struct StoreF32 {
    using T = float;
    static constexpr int bytes = 4;
    static constexpr bool shifted = false;
    __device__ __forceinline__ static float load(T v, int, const F16Shift &) { return v; }
    __device__ __forceinline__ static float load_b(T v, int, bool, const F16Shift &) { return v; }
    __device__ __forceinline__ static T store(float w, int, const F16Shift &) { return w; }
    __device__ __forceinline__ static float to_f(T v, int, const F16Shift &) { return v; }
    __device__ __forceinline__ static T from_f(float f, int, const F16Shift &) { return f; }
    __device__ __forceinline__ static Moments collide(float w[9], float omega) { return collide_ref(w, omega); }
    __device__ __forceinline__ static Moments moments(const float w[9], bool with_rest) {
        return moments_ref(w, with_rest);
    }
};

struct StoreF16S {
    using T = uint16_t;
    static constexpr int bytes = 2;
    static constexpr bool shifted = false;
    // f16 -> f32 is exact, so however the compiler folds the conversion into
    // its users (a mixed-precision FMA, an SDWA operand) the value is the
    // same.  (No asm pin here: an inline asm statement is a scheduling
    // boundary, and pinned loads were waited for one by one.)
    __device__ __forceinline__ static float load(T v, int, const F16Shift &) {
        return (float)__builtin_bit_cast(_Float16, v);
    }
    // (f - w has the same shift for a population and its opposite)
    __device__ __forceinline__ static float load_b(T v, int k, bool, const F16Shift &sh) { return load(v, k, sh); }
    // The format rounds the f32 value to nearest even.  Without a barrier the
    // compiler folds f16(fma(a, b, c)) into one mixed-precision FMA (a single
    // rounding instead of two), in some kernel shapes and not others, and
    // kernels would disagree in the last bit; canonicalize keeps the
    // conversion separate and costs nothing (an inline asm pin did the same
    // but was a scheduling boundary).
    __device__ __forceinline__ static T store(float g, int, const F16Shift &) {
        g = __builtin_elementwise_canonicalize(g);
        return __builtin_bit_cast(uint16_t, (_Float16)g);
    }
    __device__ __forceinline__ static float to_f(T v, int k, const F16Shift &sh) { return load(v, k, sh) + weight(k); }
    __device__ __forceinline__ static T from_f(float f, int k, const F16Shift &sh) {
        return store(f - weight(k), k, sh);
    }
    __device__ __forceinline__ static Moments collide(float w[9], float omega) { return collide_shifted(w, omega); }
    __device__ __forceinline__ static Moments moments(const float w[9], bool with_rest) {
        return moments_shifted(w, with_rest);
    }
};

// Binary16 of f_k - s_k, s_k the free-stream equilibrium (F16Shift above): the
// working variables are those of StoreF16S, shifted by d_k on the way in and
// out (exact in f32 for stored values below 16 in magnitude).
struct StoreF16E : StoreF16S {
    static constexpr bool shifted = true;
    __device__ __forceinline__ static float load(T v, int k, const F16Shift &sh) {
        return StoreF16S::load(v, k, sh) + sh.d[k];
    }
    // population k's working value from a stored value of population opp(k) when
    // `bounced` (half-way bounce-back hands a cell its own opposite population)
    __device__ __forceinline__ static float load_b(T v, int k, bool bounced, const F16Shift &sh) {
        return StoreF16S::load(v, k, sh) + (bounced ? sh.d[opp(k)] : sh.d[k]);
    }
    __device__ __forceinline__ static T store(float g, int k, const F16Shift &sh) {
        return StoreF16S::store(g - sh.d[k], k, sh);
    }
    __device__ __forceinline__ static float to_f(T v, int k, const F16Shift &sh) { return load(v, k, sh) + weight(k); }
    __device__ __forceinline__ static T from_f(float f, int k, const F16Shift &sh) {
        return store(f - weight(k), k, sh);
    }
};

}  // namespace d2q9
