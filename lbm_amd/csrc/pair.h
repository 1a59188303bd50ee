// The two-step kernel's per-block work, shared by the single grid (k_pair in
// lbm_grid.hip) and the batched engine (k_batch_pair in lbm_batch.hip).
#pragma once

#include "lbm_common.h"

namespace {

// ---- two steps per launch -------------------------------------------------------
//
// P-form -> P-form twice in one pass over HBM.  Inputs A = buf[cur] (C_n) and
// B = buf[1-cur]; outputs A2 := C_{n+2} and B2 := C_{n+1}, in two other buffers
// (a block's halo reads its neighbours' inputs, so nothing is updated in place).
// A block owns a strip of kBlock - 2 columns and `rows` rows; thread t works on
// column x0 - 1 + t, so threads 0 and kBlock - 1 compute the strip's halo
// columns.  Walking down the strip, phase 1 computes C_{n+1} = collide(S_{n+1})
// of the next row, including the halo, into a ring of four rows in LDS, and
// phase 2 pulls S_{n+2} of the current row from the ring and stores
// collide(S_{n+2}) to A2.  C_{n+1} goes to B2 only where the next step reads
// it (cells the stream does not write); a call therefore ends with one-step
// launches, which leave both sets whole.  The outlet column's wrapped pulls read column 0, which
// belongs to another strip; its C_{n+1} (column 0 is never streamed) comes
// from k_pair_col0, launched first.  Same per-cell operations as k_cell, so
// the same bits.
constexpr int kRing = 4;

// C_{n+1} of column 0, row r, into col0[k * Y + r] (the outlet's wrapped pulls
// read it; column 0 belongs to another strip)
template <class St, int MODE>
__device__ __forceinline__ void pair_col0_row(const typename St::T *A, const typename St::T *B,
                                              const uint8_t *__restrict__ flags, const uint8_t *__restrict__ mask,
                                              const Geom &g, int r, typename St::T *__restrict__ col0) {
    float f[9];
    gather_cell<St, MODE, true>(A, B, flags, mask, g, 0, r, (size_t)r * g.pitch, (size_t)r * g.ys, f);
    St::collide(f, g.omega);
#pragma unroll
    for (int k = 0; k < 9; ++k) col0[(size_t)k * g.Y + r] = St::store(f[k], k, g.sh);
}

// Barrier for LDS hand-off only: waits for this wave's LDS operations, not for
// its global loads, so the next row's pulls stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// One block's work of the two-step kernel: the strip of kBlock - 2 owner
// columns from strip * (kBlock - 2), rows [r0, r1), with `ring` (C_{n+1} of
// four rows: the eight moving populations; the rest population is only read
// by its own cell, which keeps it in a register) in the block's LDS.
template <class St, int MODE>
__device__ __forceinline__ void pair_block(typename St::T (&ring)[kRing][8][kBlock], const typename St::T *A,
                                           const typename St::T *B, typename St::T *__restrict__ A2,
                                           typename St::T *__restrict__ B2, const typename St::T *__restrict__ col0,
                                           const uint8_t *__restrict__ flags, const uint8_t *__restrict__ mask,
                                           const Geom &g, int strip, int r0, int r1) {
    using T = typename St::T;
    const long long ks = g.ks;
    const int t = (int)threadIdx.x;
    const int x = strip * (kBlock - 2) - 1 + t;
    const bool in_grid = x >= 0 && x < g.X;
    const bool owner = t >= 1 && t <= kBlock - 2 && x < g.X;
    // cells whose pulls leave the row (column 0, the outlet column) take the
    // general per-cell path, as in the one-step kernel
    const bool edge = x == 0 || x + 1 >= g.X;
    // lane columns of the pulls, clamped into the row so every lane loads in
    // bounds (lanes outside the grid, x = -1 or x >= nx, load and discard)
    const unsigned xc = (unsigned)min(max(x, 0), g.X - 1);
    const unsigned xw = (unsigned)min(max(x - 1, 0), g.X - 1);
    const unsigned xe = (unsigned)min(max(x + 1, 0), g.pitch - 1);

    // phase 1, part 1: the nine pulls of row r, issued at clamped addresses
    // (the rows are uniform: clamped once per block); patched in part 2
    auto load1 = [&](int r, T v[9], uint32_t &fl) {
        if (r < 0 || r >= g.Y) return;
        fl = flags[(size_t)r * g.pitch + xc];
        const T *up = A + (size_t)min(r + 1, g.Y - 1) * g.ys;  // cy = -1 pulls from r + 1
        const T *mid = A + (size_t)r * g.ys;
        const T *dn = A + (size_t)max(r - 1, 0) * g.ys;        // cy = +1 pulls from r - 1
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const T *row = cy(k) < 0 ? up : (cy(k) > 0 ? dn : mid);
            const unsigned col = cx(k) < 0 ? xe : (cx(k) > 0 ? xw : xc);
            v[k] = row[(size_t)k * ks + col];
        }
    };
    // phase 1, part 2: S_{n+1} of row r (bounce-back, unstreamed cells, edges),
    // C_{n+1} = collide into the ring; returns the rest population
    auto finish1 = [&](int r, const T v[9], uint32_t fl) -> T {
        if (!in_grid || r < 0 || r >= g.Y) return T{};
        const size_t own = (size_t)r * g.pitch + x;
        const size_t sown = (size_t)r * g.ys + x;
        float f[9];
        bool streamed = true;
        if (edge) {
            streamed = gather_cell<St, MODE, true>(A, B, flags, mask, g, x, r, own, sown, f);
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k) f[k] = St::load(v[k], k, g.sh);
            if (fl == kSlowCell) {
                streamed = mask[own] != 1 && r < g.Y - 1;
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    if (k == C) continue;
                    f[k] = St::load_b(streamed ? A[(size_t)opp(k) * ks + sown] : B[(size_t)k * ks + sown], k,
                                      streamed, g.sh);
                }
            } else if (fl != 0) {
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    if (k == C) continue;
                    if ((fl >> bit(k)) & 1u) f[k] = St::load(A[(size_t)opp(k) * ks + sown], opp(k), g.sh);
                }
            }
        }
        St::collide(f, g.omega);
        const bool store = owner && r >= r0 && r < r1 && !streamed;
        T o[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            o[k] = St::store(f[k], k, g.sh);
            if (k != C) ring[r & (kRing - 1)][bit(k)][t] = o[k];
        }
        if (store) {
            T *dst = B2 + (size_t)r * g.ys;
#pragma unroll
            for (int k = 0; k < 9; ++k) dst[(size_t)k * ks + x] = o[k];
        }
        return o[C];
    };

    // phase 2: S_{n+2} of row y from the ring, collide, store to A2; fl is
    // the row's flag byte, read by phase 1
    auto phase2 = [&](int y, T rest, uint32_t fl) {
        if (!owner) return;
        const int me = y & (kRing - 1);
        float f[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (k == C) continue;
            f[k] = St::load(ring[(y - cy(k)) & (kRing - 1)][bit(k)][t - cx(k)], k, g.sh);
        }
        f[C] = St::load(rest, C, g.sh);
        if (fl == kSlowCell) {
            const size_t sown = (size_t)y * g.ys + x;
            const bool streamed = mask[(size_t)y * g.pitch + x] != 1 && x != 0 && y < g.Y - 1;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                if (k == C) continue;
                f[k] = St::load_b(streamed ? ring[me][bit(opp(k))][t] : A[(size_t)k * ks + sown], k, streamed, g.sh);
            }
        } else if (fl != 0 || x == g.X - 1) {
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                if (k == C) continue;
                if ((fl >> bit(k)) & 1u) {
                    f[k] = St::load(ring[me][bit(opp(k))][t], opp(k), g.sh);
                } else if (x - cx(k) >= g.X) {
                    // the outlet's pull past the last column: the flattened
                    // index wraps to column 0 of the next row, or clamps
                    if constexpr (MODE == LBM_BOUNDARY_REFERENCE)
                        f[k] = St::load(col0[(size_t)k * g.Y + (y - cy(k)) + 1], k, g.sh);
                    else
                        f[k] = St::load(ring[(y - cy(k)) & (kRing - 1)][bit(k)][t], k, g.sh);
                }
            }
        }
        St::collide(f, g.omega);
        T *dst = A2 + (size_t)y * g.ys;
#pragma unroll
        for (int k = 0; k < 9; ++k) __builtin_nontemporal_store(St::store(f[k], k, g.sh), dst + (size_t)k * ks + x);
    };

    // rows r0 - 1 and r0, then per row y: row y + 2's pulls issued, row y + 1
    // finished into the ring, barrier, row y's second step
    T v[9];
    uint32_t fl = 0;
    load1(r0 - 1, v, fl);
    finish1(r0 - 1, v, fl);
    load1(r0, v, fl);
    T rest = finish1(r0, v, fl);
    uint32_t fl_y = fl;
    load1(r0 + 1, v, fl);
    for (int y = r0; y < r1; ++y) {
        T cv[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) cv[k] = v[k];
        const uint32_t cfl = fl;
        if (y + 2 <= r1) load1(y + 2, v, fl);
        const T next = finish1(y + 1, cv, cfl);
        lds_barrier();
        phase2(y, rest, fl_y);
        rest = next;
        fl_y = cfl;
    }
}


}  // namespace
