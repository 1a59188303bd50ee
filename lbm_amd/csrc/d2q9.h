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

// ---- storage formats ------------------------------------------------------------
// F32: the population itself.  F16S: g_k = f_k - w_k as IEEE binary16, converted
// with round-to-nearest-even; arithmetic stays f32.

struct StoreF32 {
    using T = float;
    static constexpr int bytes = 4;
    __device__ __forceinline__ static float decode(T v, int) { return v; }
    __device__ __forceinline__ static T encode(float v, int) { return v; }
};

struct StoreF16S {
    using T = uint16_t;
    static constexpr int bytes = 2;
    __device__ __forceinline__ static float decode(T v, int k) {
        return (float)__builtin_bit_cast(_Float16, v) + weight(k);
    }
    __device__ __forceinline__ static T encode(float v, int k) {
        return __builtin_bit_cast(uint16_t, (_Float16)(v - weight(k)));
    }
};

}  // namespace d2q9
