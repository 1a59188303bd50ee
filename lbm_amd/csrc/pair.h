// The two-step kernel's per-block work, shared by the single grid (k_pair in
// lbm_grid.hip) and the batched engine (k_batch_pair in lbm_batch.hip).
#pragma once

#include <type_traits>

#include "lbm_common.h"

namespace {

// ---- two steps per launch -------------------------------------------------------
//
// P-form -> P-form twice in one pass over HBM.  Inputs A = buf[cur] (C_n) and
// B = buf[1-cur]; outputs A2 := C_{n+2} and B2 := C_{n+1}, in two other buffers
// (a block's halo reads its neighbours' inputs, so nothing is updated in place).
//
// Cells with the flag byte kSlowCell -- solid cells, column 0, the last row
// (none of them streamed) and fluid cells whose eight upstream cells are all
// solid (every population bounced) -- read only themselves over both steps.
// The fixed-cell pass (fixed_cell over a list of them: the launch's blocks
// after the strips, running alongside them) steps them into A2 and B2 and the
// strip kernel skips them: nothing pulls from them but column 1 from column 0
// and the row above from the last row, whose C_{n+1} the strip kernel
// computes itself.  The two touch disjoint cells of A2 and B2 and only read
// A and B.
//
// A block owns a strip of kBlock - 2 columns and `rows` rows; thread t works on
// column x0 - 1 + t, so threads 0 and kBlock - 1 compute the strip's halo
// columns.  Walking down the strip, phase 1 computes C_{n+1} = collide(S_{n+1})
// of the next row, including the halo, into a ring of four rows in LDS, and
// phase 2 pulls S_{n+2} of the current row from the ring and stores
// collide(S_{n+2}) to A2.  C_{n+1} goes to B2 only where the next step reads
// it (cells the stream does not write: the fixed-cell pass's); a call
// therefore ends with one-step launches, which leave both sets whole.
//
// In the strip kernel the lane of column 0 computes C_{n+1} of column 0 into
// the ring, for column 1's pulls, and owns nothing; with the reference outlet
// the lane of column X does the same with column 0 one row down, which the
// outlet's pulls reach through the flattened index.  So no lane takes a
// per-cell path for the grid's edges or its solid cells, and the only loads not
// issued one row ahead are bounce-back reads and the last row's (a wave holding
// an edge or solid cell used to wait on dependent loads in every row, and its
// block with it at every barrier).  Same per-cell operations as k_cell, so the
// same bits.
constexpr int kRing = 4;

// A kSlowCell cell (x, y) two steps on: C_{n+1} into B2 (where the cell is
// not streamed: the next step reads it there), C_{n+2} into A2.  A cell that
// is not streamed takes the other set's moving populations (B for step n+1, A
// for step n+2) and keeps its rest population; a streamed one (every upstream
// cell solid) bounces every population back from its own cell.  `streamed`
// comes with the list entry, so every load is issued at once.
template <class St>
__device__ __forceinline__ void fixed_cell(const typename St::T *A, const typename St::T *B, const Geom &g, int x,
                                           int y, bool streamed, typename St::T *__restrict__ A2,
                                           typename St::T *__restrict__ B2) {
    using T = typename St::T;
    const long long ks = g.ks;
    const size_t sown = (size_t)y * g.ys + x;
    T a[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) a[k] = A[(size_t)k * ks + sown];
    float f[9];
    if (streamed) {
#pragma unroll
        for (int k = 0; k < 9; ++k)
            if (k != C) f[k] = St::load_b(a[opp(k)], k, true, g.sh);
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k)
            if (k != C) f[k] = St::load(B[(size_t)k * ks + sown], k, g.sh);
    }
    f[C] = St::load(a[C], C, g.sh);
    St::collide(f, g.omega);
    T o[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        o[k] = St::store(f[k], k, g.sh);
        if (!streamed) B2[(size_t)k * ks + sown] = o[k];
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        if (k == C) f[k] = St::load(o[C], C, g.sh);
        else f[k] = streamed ? St::load_b(o[opp(k)], k, true, g.sh) : St::load(a[k], k, g.sh);
    }
    St::collide(f, g.omega);
#pragma unroll
    for (int k = 0; k < 9; ++k) A2[(size_t)k * ks + sown] = St::store(f[k], k, g.sh);
}

// ---- the fixed-cell list ------------------------------------------------------
//
// The fixed-cell pass runs one thread per entry of a list of the kSlowCell
// cells, (column and whether it is streamed, row) in row order, built on the
// device after each mask change: a count per row, an exclusive scan of the
// counts on the host, and an ordered emit per row.  One block of kFixedBlock
// threads per row; each thread looks at 16 flag bytes (one aligned 16-byte
// word: the flag plane's rows are padded to a multiple of 64 bytes) per pass
// over the row.
constexpr int kFixedBlock = 256;

// bit i set: column x0 + i < X is a kSlowCell cell
__device__ __forceinline__ uint32_t slow_bits16(const uint8_t *__restrict__ row, int x0, int X) {
    if (x0 >= X) return 0;
    const uint4 w = *(const uint4 *)(row + x0);
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (((ws[j] >> (8 * b)) & 0xFFu) == kSlowCell) m |= 1u << (4 * j + b);
    return X - x0 < 16 ? m & ((1u << (X - x0)) - 1u) : m;
}

// exclusive prefix of c over the block's threads (in `sc`, kFixedBlock
// entries); *total gets the sum
__device__ __forceinline__ uint32_t block_prefix(uint32_t c, uint32_t *sc, uint32_t *total) {
    const int t = (int)threadIdx.x;
    sc[t] = c;
    __syncthreads();
    for (int off = 1; off < kFixedBlock; off <<= 1) {
        const uint32_t v = t >= off ? sc[t - off] : 0u;
        __syncthreads();
        sc[t] += v;
        __syncthreads();
    }
    const uint32_t incl = sc[t];
    *total = sc[kFixedBlock - 1];
    __syncthreads();
    return incl - c;
}

// the number of kSlowCell cells of one row of X columns into *count
__device__ __forceinline__ void fixed_row_count(const uint8_t *__restrict__ row, int X, uint32_t *count) {
    __shared__ uint32_t sc[kFixedBlock];
    uint32_t c = 0;
    for (int x0 = (int)threadIdx.x * 16; x0 < X; x0 += kFixedBlock * 16) c += __popc(slow_bits16(row, x0, X));
    uint32_t total;
    block_prefix(c, sc, &total);
    if (threadIdx.x == 0) *count = total;
}

// the kSlowCell cells of row y (of Y, X columns; `row` and `mrow` its flag and
// mask bytes) as (column | streamed << 31, tag) into out[0 ..), in column order
__device__ __forceinline__ void fixed_row_emit(const uint8_t *__restrict__ row, const uint8_t *__restrict__ mrow,
                                               int X, int y, int Y, uint32_t tag, uint2 *__restrict__ out) {
    __shared__ uint32_t sc[kFixedBlock];
    uint32_t base = 0;
    for (int x00 = 0; x00 < X; x00 += kFixedBlock * 16) {
        const int x0 = x00 + (int)threadIdx.x * 16;
        uint32_t m = slow_bits16(row, x0, X), total;
        uint32_t at = base + block_prefix(__popc(m), sc, &total);
        while (m) {
            const int x = x0 + __builtin_ctz(m);
            m &= m - 1u;
            const bool streamed = mrow[x] != 1 && x != 0 && y < Y - 1;
            out[at++] = make_uint2((uint32_t)x | (streamed ? 0x80000000u : 0u), tag);
        }
        base += total;
    }
}

// Raw buffer loads (gfx9 buffer descriptor, word 3 0x00020000: 32-bit
// untyped data; offsets = lane offset + uniform offset, bounds-checked).
constexpr int kBufferWord3 = 0x00020000;

template <class T>
__device__ __forceinline__ T buffer_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(T) == 2)
        return (T)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// The second step's raw buffer store: non-temporal (cache policy bit 1) for
// 4-byte populations, an ordinary store for binary16.  A strip's 254 columns
// do not fill whole 128-byte lines, so the lines at its edges are shared with
// the next strip; binary16 ones hold twice the columns, and ordinary stores,
// which the L2 merges, are faster there (8192 x 4096: 177 -> 169 us per step),
// while f32 is faster non-temporal (268 against 280 us).
template <class T>
constexpr bool kPairStoreNT = sizeof(T) == 4;

template <class T>
__device__ __forceinline__ void buffer_store_out(T v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(T) == 2)
        __builtin_amdgcn_raw_buffer_store_b16((unsigned short)v, r, voff, soff, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, kPairStoreNT<T> ? 2 : 0);
}

// Barrier for LDS hand-off only: waits for this wave's LDS operations, not for
// its global loads, so the next row's pulls stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Waits for all of this wave's global loads (s_waitcnt vmcnt(0); expcnt and
// lgkmcnt left alone).  A builtin, not asm: the compiler's own wait insertion
// sees it, so a branch that ends with it leaves no load pending at its join.
__device__ __forceinline__ void wait_global_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// One block's work of the two-step kernel: the strip of kBlock - 2 owner
// columns from strip * (kBlock - 2), rows [r0, r1), with `ring` (C_{n+1} of
// four rows: the eight moving populations; the rest population is only read
// by its own cell, which keeps it in a register) in the block's LDS.  EDGE:
// the strip holds column 0, X - 1 or X (pair_block_any picks the instance,
// so that the other blocks run a loop without the relay and wrap code).
template <class St, int MODE, bool EDGE>
__device__ __forceinline__ void pair_block(typename St::T (&ring)[kRing][8][kBlock], const typename St::T *A,
                                           const typename St::T *B, typename St::T *__restrict__ A2,
                                           typename St::T *__restrict__ B2, const uint8_t *__restrict__ flags,
                                           const Geom &g, int strip, int r0, int r1) {
    using T = typename St::T;
    constexpr bool wrap = MODE == LBM_BOUNDARY_REFERENCE;
    const long long ks = g.ks;
    const int t = (int)threadIdx.x;
    const int x = strip * (kBlock - 2) - 1 + t;
    const bool in_grid = x >= 0 && x < g.X;
    // column 0 is the fixed-cell pass's: its lane only computes C_{n+1} for
    // column 1's pulls (as does the lane of column X, one row down, for the
    // reference outlet's); these relay lanes own nothing
    const bool owner = t >= 1 && t <= kBlock - 2 && x >= 1 && x < g.X;
    const bool lane0 = EDGE && x == 0, laneX = EDGE && wrap && x == g.X, outlet = EDGE && x == g.X - 1;
    const bool relay = lane0 || laneX;
    // Byte offsets of a lane's nine loads from the pulled row of A (row r -
    // c_y(k)).  Pulls by the sign of c_x: the row's columns x + 1, x, x - 1,
    // clamped into the row so every lane loads in bounds (lanes outside the
    // grid, x = -1 or x >= nx, load and discard; the zero-gradient outlet's
    // clamped upstream is its own column).  The reference outlet's
    // west-moving pulls wrap to column 0 of the next row (the flattened
    // index): one row on, column 0; in the last rows that is up to two rows
    // past the set, which every set's allocation holds (those pulls are
    // flagged as bounce-back).  The relay lanes load S_{n+1} of column 0 --
    // not streamed: B's moving populations and A's rest population -- of row r
    // (column 0's lane) or r + 1 (column X's), so that phase 1 collides it
    // like any other cell.  The offsets are loop-invariant, so every lane
    // issues the same loads.
    const long long es = (long long)sizeof(T), row_b = g.ys * es;
    const long long dB = (long long)((uintptr_t)B - (uintptr_t)A);
    long long off[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int ux = x - cx(k);
        off[k] = (long long)min(max(ux, 0), g.X - 1) * es;
        if (wrap && outlet && cx(k) < 0) off[k] = row_b;
        if (lane0) off[k] = k == C ? 0 : dB + cy(k) * row_b;
        if (laneX) off[k] = k == C ? row_b : dB + (1 + cy(k)) * row_b;
    }
    const unsigned xc = (unsigned)min(max(x, 0), g.X - 1);
    // a loaded value waits a row in a whole register: packing two binary16
    // values into one made the compiler wait for the loads right after issuing
    // them
    using V = std::conditional_t<sizeof(T) == 2, uint32_t, T>;

    // the strip's rows of A and A2 from row brow as buffers (blocks away from
    // the edges load and store through them): the first row loaded is r0 - 1,
    // whose southward pulls come from r0 - 2
    const int brow = max(r0 - 2, 0);
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void *)(A + (size_t)brow * g.ys), (short)0, 0x7FFFFFFF, kBufferWord3);
    const __amdgpu_buffer_rsrc_t rsrc2 =
        __builtin_amdgcn_make_buffer_rsrc((void *)(A2 + (size_t)brow * g.ys), (short)0, 0x7FFFFFFF, kBufferWord3);

    // phase 1, part 1: the nine loads of row r (the rows are uniform; the row
    // below is clamped at row 0, the row above runs into the padding row)
    auto load1 = [&](int r, V v[9], uint32_t &fl, bool first) {
        if (r < 0 || r >= g.Y) return;
        fl = flags[(size_t)r * g.pitch + xc];
        const char *up = (const char *)(A + (size_t)(r + 1) * g.ys);       // cy = -1 pulls from r + 1
        const char *mid = (const char *)(A + (size_t)r * g.ys);
        const char *dn = (const char *)(A + (size_t)max(r - 1, 0) * g.ys);  // cy = +1 pulls from r - 1
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const char *row = cy(k) < 0 ? up : (cy(k) > 0 ? dn : mid);
            long long o = off[k];
            // row 0's loads from row -1 use the clamped row 0: the wrapped and
            // relay loads there step one row back (only the first rows are
            // loaded before the loop: `first`)
            if (first && cy(k) > 0 && r < 1 && (relay || (wrap && outlet && cx(k) < 0))) o -= row_b;
            if constexpr (EDGE) {
                v[k] = (V) * (const T *)(row + k * ks * es + o);
            } else {
                // away from the edges every pull is a column of a row near the
                // block: a buffer load with the row and plane in the uniform
                // offset and the column in the lane offset (no 64-bit lane
                // address per load)
                const int pr = cy(k) < 0 ? r + 1 : (cy(k) > 0 ? max(r - 1, 0) : r);
                const uint32_t soff = (uint32_t)((long long)(pr - brow) * row_b + k * ks * es);
                v[k] = (V)buffer_load<T>(rsrc, (uint32_t)o, soff);
            }
        }
    };
    // phase 1, part 2: S_{n+1} of row r (bounce-back; the last row),
    // C_{n+1} = collide into the ring; returns the rest population.  A relay
    // lane's loads are already the unstreamed S_{n+1} of column 0 (row r, or
    // r + 1 for column X's lane, whose ring slot r the outlet's wrapped pulls
    // read).
    auto finish1 = [&](int r, const V v[9], uint32_t fl) -> T {
        if (r < 0 || r >= g.Y) return T{};
        // fixed cells are the fixed-cell pass's, and nothing pulls from them
        // but the row above from the last row (and column 1 from column 0)
        if (!relay && (!in_grid || (fl == kSlowCell && r != g.Y - 1))) return T{};
        const size_t sown = (size_t)r * g.ys + x;
        float f[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) f[k] = St::load((T)v[k], k, g.sh);
        if (!relay && r == g.Y - 1) {
            // the last row is not streamed: the other set's moving populations
#pragma unroll
            for (int k = 0; k < 9; ++k)
                if (k != C) f[k] = St::load(B[(size_t)k * ks + sown], k, g.sh);
        } else if (!relay && fl != 0) {
            // the cell's own eight moving populations, all loaded, then the
            // bounced ones selected: loads under per-direction branches were
            // waited for one by one, and one left pending past the branch made
            // every wave wait at the join for all its loads in flight, the
            // next row's pulls included
            T b[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                if (k == C) continue;
                if constexpr (EDGE)
                    b[k] = A[(size_t)opp(k) * ks + sown];
                else
                    b[k] = buffer_load<T>(rsrc, (uint32_t)x * (uint32_t)es,
                                          (uint32_t)((long long)(r - brow) * row_b + opp(k) * ks * es));
            }
            wait_global_loads();
#pragma unroll
            for (int k = 0; k < 9; ++k)
                if (k != C && ((fl >> bit(k)) & 1u)) f[k] = St::load(b[k], opp(k), g.sh);
        }
        St::collide(f, g.omega);
        T o[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            o[k] = St::store(f[k], k, g.sh);
            if (k != C) ring[r & (kRing - 1)][bit(k)][t] = o[k];
        }
        return o[C];
    };

    // phase 2: S_{n+2} of row y from the ring, collide, store to A2; fl is
    // the row's flag byte, read by phase 1
    auto phase2 = [&](int y, T rest, uint32_t fl) {
        if (!owner || fl == kSlowCell) return;  // fixed cells: the fixed-cell pass's
        const int me = y & (kRing - 1);
        float f[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (k == C) continue;
            f[k] = St::load(ring[(y - cy(k)) & (kRing - 1)][bit(k)][t - cx(k)], k, g.sh);
        }
        f[C] = St::load(rest, C, g.sh);
        if (fl != 0 || (!wrap && outlet)) {
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                if (k == C) continue;
                if ((fl >> bit(k)) & 1u) {
                    f[k] = St::load(ring[me][bit(opp(k))][t], opp(k), g.sh);
                } else if (!wrap && x - cx(k) >= g.X) {
                    // the zero-gradient outlet's pull past the last column
                    // clamps to the cell's own column (the reference outlet's
                    // wraps: column X's relay lane holds it)
                    f[k] = St::load(ring[(y - cy(k)) & (kRing - 1)][bit(k)][t], k, g.sh);
                }
            }
        }
        St::collide(f, g.omega);
        if constexpr (EDGE) {
            T *dst = A2 + (size_t)y * g.ys;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                if constexpr (kPairStoreNT<T>)
                    __builtin_nontemporal_store(St::store(f[k], k, g.sh), dst + (size_t)k * ks + x);
                else
                    dst[(size_t)k * ks + x] = St::store(f[k], k, g.sh);
            }
        } else {
#pragma unroll
            for (int k = 0; k < 9; ++k)
                buffer_store_out<T>(St::store(f[k], k, g.sh), rsrc2, (uint32_t)x * (uint32_t)es,
                                   (uint32_t)((long long)(y - brow) * row_b + k * ks * es));
        }
    };

    // rows r0 - 1 and r0, then per row y: row y + 2's pulls issued, row y + 1
    // finished into the ring, barrier, row y's second step
    V v[9];
    uint32_t fl = 0;
    load1(r0 - 1, v, fl, true);
    finish1(r0 - 1, v, fl);
    if (r0 == 0 && laneX) {
        // the outlet's south-west pull in row 0 wraps to column 0 of row 0:
        // column X's slot of row -1
        float f[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) f[k] = St::load(k == C ? A[(size_t)C * ks] : B[(size_t)k * ks], k, g.sh);
        St::collide(f, g.omega);
        ring[(0 - 1) & (kRing - 1)][bit(SW)][t] = St::store(f[SW], SW, g.sh);
    }
    load1(r0, v, fl, true);
    T rest = finish1(r0, v, fl);
    uint32_t fl_y = fl;
    load1(r0 + 1, v, fl, true);
    for (int y = r0; y < r1; ++y) {
        V cv[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) cv[k] = v[k];
        const uint32_t cfl = fl;
        if (y + 2 <= r1) load1(y + 2, v, fl, false);
        const T next = finish1(y + 1, cv, cfl);
        lds_barrier();
        phase2(y, rest, fl_y);
        rest = next;
        fl_y = cfl;
    }
}

template <class St, int MODE>
__device__ __forceinline__ void pair_block_any(typename St::T (&ring)[kRing][8][kBlock], const typename St::T *A,
                                               const typename St::T *B, typename St::T *__restrict__ A2,
                                               typename St::T *__restrict__ B2,
                                               const uint8_t *__restrict__ flags, const Geom &g, int strip,
                                               int r0, int r1) {
    if (strip == 0 || strip * (kBlock - 2) + kBlock - 2 >= g.X - 1)
        pair_block<St, MODE, true>(ring, A, B, A2, B2, flags, g, strip, r0, r1);
    else
        pair_block<St, MODE, false>(ring, A, B, A2, B2, flags, g, strip, r0, r1);
}

}  // namespace
