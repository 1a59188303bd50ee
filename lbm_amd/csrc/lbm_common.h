// Device code shared by the single-grid engine (lbm_grid.hip) and the batched
// multi-domain engine (lbm_batch.hip): geometry, the bounce-back byte and
// segment classification, the interior tile (vectorized pull + collide) and the
// general per-cell pull, FP64 block sums, and host-side error helpers.
// Compiled into each translation unit with internal linkage.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <string>

#include "../../include/lbm_hip.h"
#include "d2q9.h"

using namespace d2q9;

namespace {

// ---------------------------------------------------------------------------
// device side
// ---------------------------------------------------------------------------

struct Geom {
    int X, Y;
    int pitch;          // elements per row of every plane
    long long pstride;  // elements per auxiliary plane (mask, flags, moments)
    long long ys, ks;   // state strides: element (k, x, y) of a set is at y * ys + k * ks + x
    float omega;
    F16Shift sh;        // shift of the binary16 storage (off: f - w)
};

// per-cell bounce-back byte: bit(k) set if population k's upstream is solid or
// outside the grid; 0xFF marks a cell the fast path does not pull (not streamed,
// or streamed with all eight upstream cells solid)
constexpr uint8_t kSlowCell = 0xFF;

// upstream cell of population k at (x, y); mode 0 wraps the flattened index at
// the outlet column, mode 1 clamps it (zero-gradient outflow)
template <int MODE>
__host__ __device__ __forceinline__ void upstream(int k, int x, int y, int X, int &ux, int &uy) {
    ux = x - cx(k);
    uy = y - cy(k);
    if (ux >= X) {
        if (MODE == LBM_BOUNDARY_REFERENCE) {
            ux = 0;
            uy += 1;
        } else {
            ux = X - 1;
        }
    }
}

template <int MODE>
__global__ void k_build_flags(const uint8_t *__restrict__ mask, uint8_t *__restrict__ flags, Geom g) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= g.X) return;
    const size_t own = (size_t)y * g.pitch + x;
    const bool streamed = mask[own] != 1 && x != 0 && y < g.Y - 1;
    uint8_t b = kSlowCell;
    if (streamed) {
        b = 0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (k == C) continue;
            int ux, uy;
            upstream<MODE>(k, x, y, g.X, ux, uy);
            const bool solid = uy < 0 || uy >= g.Y || ux < 0 || mask[(size_t)uy * g.pitch + ux] == 1;
            if (solid) b |= (uint8_t)(1u << bit(k));
        }
    }
    flags[own] = b;
}

// The grid is cut into 64-cell segments of a row.  A segment is "clear" when
// every cell in it is streamed with no bounce-back and lies before the outlet
// column: the interior kernel then pulls all 64 cells with no per-cell test.
// Every other segment that holds cells goes on a list that the boundary kernel
// works through, one wave per segment.  Rows of the segment plane hold
// pitch / 64 + 8 entries; the extra ones stay 0.
constexpr int kSegSlack = 8;

__global__ void k_build_segments(const uint8_t *__restrict__ flags, uint8_t *__restrict__ segs,
                                 uint32_t *__restrict__ list, uint32_t *__restrict__ count, Geom g) {
    const int srow = g.pitch / 64 + kSegSlack;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (s >= srow) return;
    uint8_t clear = 0;
    if (s < g.pitch / 64 && 64 * s + 63 < g.X - 1) {
        const uint4 *p = (const uint4 *)(flags + (size_t)y * g.pitch + 64 * s);
        uint32_t acc = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint4 q = p[i];
            acc |= q.x | q.y | q.z | q.w;
        }
        clear = acc == 0;
    }
    segs[(size_t)y * srow + s] = clear;
    if (list && !clear && 64 * s < g.X) list[atomicAdd(count, 1u)] = (uint32_t)((size_t)y * srow + s);
}

// fn.template operator()<St>() with the storage format of `precision`
template <class F>
void by_store(int precision, F &&fn) {
    if (precision == LBM_PRECISION_F32) fn.template operator()<StoreF32>();
    else if (precision == LBM_PRECISION_F16E) fn.template operator()<StoreF16E>();
    else fn.template operator()<StoreF16S>();
}

enum : int {
    OP_COLLIDE = 1,   // collide the gathered state and store all nine planes to dst
    OP_STORE_S = 2,   // store the gathered (streamed) state to dst
    OP_MOMENTS = 4,   // write jx, jy, rho of the gathered state
    OP_REDUCE = 8,    // FP64 partial sums of rho, jx, jy over fluid cells
    OP_NOREST = 16,   // moments without the rest population
};

constexpr int kBlock = 256;

// CPT consecutive cells of a row per thread, moved as one vector: aligned for the
// cell's own planes, element-aligned for the planes pulled from x - 1 / x + 1
template <class T, int N>
struct Vec {
    typedef T a __attribute__((ext_vector_type(N)));
    typedef T u __attribute__((ext_vector_type(N), aligned(sizeof(T))));
};
template <class T>
struct Vec<T, 1> {
    typedef T a;
    typedef T u;
};

template <int N, class V>
__device__ __forceinline__ auto get_lane(const V &v, int j) {
    if constexpr (N == 1) return v;
    else return v[j];
}

template <int N, class V, class E>
__device__ __forceinline__ void set_lane(V &v, int j, E x) {
    if constexpr (N == 1) v = x;
    else v[j] = x;
}

template <class V, bool NT>
__device__ __forceinline__ V vload(const V *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <class V, bool NT>
__device__ __forceinline__ void vstore(V v, V *p) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// Workgroup ids are handed to the eight XCDs in turn; remap them so that each
// XCD walks one contiguous band of rows.  Neighbouring row segments, which
// share the cache lines of their x - 1 / x + 1 pulls, then meet in one L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
    const int q = n >> 3;
    if (b >= (q << 3)) return b;
    return (b & 7) * q + (b >> 3);
}

// The state S of one cell before collision.  P-form: pull from src with the
// bounce-back byte, or, for cells the stream does not write, keep dst's own
// values; the rest population is src's.  S-form: src's own values.
template <class St, int MODE, bool PFORM>
__device__ __forceinline__ bool gather_cell(const typename St::T *src, const typename St::T *dst,
                                            const uint8_t *__restrict__ flags, const uint8_t *__restrict__ mask,
                                            const Geom &g, int x, int y, size_t own, size_t sown, float f[9]) {
    const long long ks = g.ks;
    bool streamed = false;
    if (PFORM) {
        const uint8_t fl = flags[own];
        if (fl != kSlowCell) {
            streamed = true;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                if (k == C) continue;
                int ux, uy;
                upstream<MODE>(k, x, y, g.X, ux, uy);
                const bool bounce = (fl >> bit(k)) & 1u;
                const size_t idx = bounce ? (size_t)opp(k) * ks + sown : (size_t)k * ks + (size_t)uy * g.ys + ux;
                f[k] = St::load_b(src[idx], k, bounce, g.sh);
            }
        } else {
            // enclosed streamed cell: every population bounces back; otherwise
            // the cell is not streamed and keeps the other set's values
            streamed = mask[own] != 1 && x != 0 && y < g.Y - 1;
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                if (k == C) continue;
                f[k] = St::load_b(streamed ? src[(size_t)opp(k) * ks + sown] : dst[(size_t)k * ks + sown], k,
                                  streamed, g.sh);
            }
        }
        f[C] = St::load(src[(size_t)C * ks + sown], C, g.sh);
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) f[k] = St::load(src[(size_t)k * ks + sown], k, g.sh);
    }
    return streamed;
}

// Per-cell epilogue shared by both kernels: collide and store, store the
// streamed state, write moments, accumulate the FP64 sums.
template <class St, int OP>
__device__ __forceinline__ void finish_cell(float f[9], bool streamed, const typename St::T *src,
                                            typename St::T *dst, const uint8_t *__restrict__ mask, const Geom &g,
                                            size_t own, size_t sown, float *__restrict__ mom, double &s0, double &s1,
                                            double &s2) {
    const long long ps = g.pstride, ks = g.ks;
    Moments m;
    if (OP & OP_STORE_S) {
        if (streamed) {
#pragma unroll
            for (int k = 0; k < 9; ++k)
                if (k != C) dst[(size_t)k * ks + sown] = St::store(f[k], k, g.sh);
        }
        dst[(size_t)C * ks + sown] = src[(size_t)C * ks + sown];
    }
    if (OP & OP_COLLIDE) {
        m = St::collide(f, g.omega);
#pragma unroll
        for (int k = 0; k < 9; ++k) dst[(size_t)k * ks + sown] = St::store(f[k], k, g.sh);
    } else {
        m = St::moments(f, !(OP & OP_NOREST));
    }
    if (OP & OP_MOMENTS) {
        mom[own] = m.jx;
        mom[ps + own] = m.jy;
        mom[2 * ps + own] = m.rho;
    }
    if ((OP & OP_REDUCE) && mask[own] != 1) {
        s0 += (double)m.rho;
        s1 += (double)m.jx;
        s2 += (double)m.jy;
    }
}

// FP64 block sums -> partials[slot * 3 + c]
__device__ __forceinline__ void block_sums(double s0, double s1, double s2, double *__restrict__ partials,
                                           size_t slot) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off, 64);
        s1 += __shfl_xor(s1, off, 64);
        s2 += __shfl_xor(s2, off, 64);
    }
    __shared__ double red[kBlock / 64][3];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[wv][0] = s0;
        red[wv][1] = s1;
        red[wv][2] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += red[w][threadIdx.x];
        partials[slot * 3 + threadIdx.x] = t;
    }
}

// One interior tile: CPT consecutive cells of row y starting at x0, each
// population moved as one vector (the nine pulls of a tile are contiguous in
// their upstream rows, at most one element off alignment).  P-form pulls, for
// tiles the caller knows to be clear (every cell streamed, no bounce-back, not
// at either end of the row); S-form (no pull): the cells' own values.
template <class St, int MODE, bool PFORM, int OP, int CPT, bool NT>
__device__ __forceinline__ void interior_tile(const typename St::T *src, typename St::T *dst,
                                              const uint8_t *__restrict__ mask, const Geom &g, int y, int x0,
                                              float *__restrict__ mom, double &s0, double &s1, double &s2) {
    using T = typename St::T;
    using V = typename Vec<T, CPT>::a;
    using U = typename Vec<T, CPT>::u;
    using VF = typename Vec<float, CPT>::a;
    const long long ps = g.pstride, ks = g.ks;
    const size_t own0 = (size_t)y * g.pitch + x0;  // auxiliary planes
    const size_t sown0 = (size_t)y * g.ys + x0;    // state
    V v[9];
    if (PFORM) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            // clamped into the set: in bounds whatever the caller's list holds
            const int ry = min(max(y - cy(k), 0), g.Y - 1);
            const int rx = min(max(x0 - cx(k), 0), g.pitch - CPT);
            const size_t idx = (size_t)k * ks + (size_t)ry * g.ys + rx;
            if (cx(k) == 0) v[k] = vload<V, NT>((const V *)(src + idx));
            else v[k] = (V)vload<U, NT>((const U *)(src + idx));
        }
    } else {
#pragma unroll
        for (int k = 0; k < 9; ++k) v[k] = vload<V, NT>((const V *)(src + (size_t)k * ks + sown0));
    }
    V o[9];
    VF mjx, mjy, mrho;
    uint64_t fluid_bits = ~0ull;
    if ((OP & OP_REDUCE) && !PFORM) {
        fluid_bits = 0;
#pragma unroll
        for (int j = 0; j < CPT; ++j) fluid_bits |= (uint64_t)(mask[own0 + j] != 1) << j;
    }
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        float f[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) f[k] = St::load(get_lane<CPT>(v[k], j), k, g.sh);
        Moments m;
        if (OP & OP_COLLIDE) {
            m = St::collide(f, g.omega);
#pragma unroll
            for (int k = 0; k < 9; ++k) set_lane<CPT>(o[k], j, St::store(f[k], k, g.sh));
        } else {
            m = St::moments(f, !(OP & OP_NOREST));
        }
        set_lane<CPT>(mjx, j, m.jx);
        set_lane<CPT>(mjy, j, m.jy);
        set_lane<CPT>(mrho, j, m.rho);
        if ((OP & OP_REDUCE) && ((fluid_bits >> j) & 1)) {
            s0 += (double)m.rho;
            s1 += (double)m.jx;
            s2 += (double)m.jy;
        }
    }
    if (OP & OP_COLLIDE) {
#pragma unroll
        for (int k = 0; k < 9; ++k) vstore<V, NT>(o[k], (V *)(dst + (size_t)k * ks + sown0));
    }
    if (OP & OP_STORE_S) {
#pragma unroll
        for (int k = 0; k < 9; ++k) vstore<V, NT>(v[k], (V *)(dst + (size_t)k * ks + sown0));
    }
    if (OP & OP_MOMENTS) {
        *(VF *)(mom + own0) = mjx;
        *(VF *)(mom + ps + own0) = mjy;
        *(VF *)(mom + 2 * ps + own0) = mrho;
    }
}

// One tile of a P-form pass with in-place patching, for any cells: the CPT
// cells' pulls are loaded as vectors with clamped coordinates and the tile's
// bounce-back bytes alongside; only cells whose byte is not 0 then have the
// affected populations re-read one at a time:
//   bit set  -> the cell's own opposite population (half-way bounce-back);
//   0xFF     -> not streamed: the other set's own values (or, if streamed with
//               all eight upstream cells solid, every population bounces back).
// The first and last group of a row are gathered cell by cell (one cell per
// thread) or fix their clamped vector loads with a register shift and load the
// outlet cell's wrapped / clamped upstream (several).  The rest of the work is
// shared by all cells.
template <class St, int MODE, int OP, int CPT, bool NT>
__device__ __forceinline__ void patched_tile(const typename St::T *src, typename St::T *dst,
                                             const uint8_t *__restrict__ flags, const uint8_t *__restrict__ mask,
                                             const Geom &g, int y, int x0, float *__restrict__ mom, double &s0,
                                             double &s1, double &s2) {
    using T = typename St::T;
    using V = typename Vec<T, CPT>::a;
    using U = typename Vec<T, CPT>::u;
    using VF = typename Vec<float, CPT>::a;
    const long long ps = g.pstride, ks = g.ks;
    const size_t own0 = (size_t)y * g.pitch + x0;
    const size_t sown0 = (size_t)y * g.ys + x0;
    V v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const int ry = min(max(y - cy(k), 0), g.Y - 1);
        const int rx = min(max(x0 - cx(k), 0), g.pitch - CPT);
        const size_t idx = (size_t)k * ks + (size_t)ry * g.ys + rx;
        if (cx(k) == 0) v[k] = vload<V, NT>((const V *)(src + idx));
        else v[k] = (V)vload<U, NT>((const U *)(src + idx));
    }
    // (Reading the per-64-cell "clear" byte first and skipping the flag bytes of
    // clear segments saves their 1 B/cell of traffic but measured 2.5% slower:
    // the extra dependent load costs more than it saves.)
    uint64_t fw;
    if constexpr (CPT == 8) fw = *(const uint64_t *)(flags + own0);
    else if constexpr (CPT == 4) fw = *(const uint32_t *)(flags + own0);
    else if constexpr (CPT == 2) fw = *(const uint16_t *)(flags + own0);
    else fw = flags[own0];
    // Groups at either end of a row.  One cell per thread: the edge cell is
    // gathered on its own (measured faster for CPT = 1).  Several cells per
    // thread (measured faster for CPT > 1):
    // the first group's east-moving loads were clamped from x0 - 1 = -1 to 0,
    // one element too far; shift them back (element 0 is column 0, which is
    // never streamed: the flag patch below rewrites it).
    constexpr bool kGatherEdges = CPT == 1;
    const bool edge = x0 == 0 || x0 + CPT >= g.X;
    if (!kGatherEdges && x0 == 0) {
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (cx(k) != 1) continue;
#pragma unroll
            for (int j = CPT - 1; j > 0; --j) set_lane<CPT>(v[k], j, get_lane<CPT>(v[k], j - 1));
        }
    }
    // The last group's west-moving loads are one element short when clamped at
    // the end of the row, and the outlet cell's upstream is the wrap / clamp of
    // MODE (loaded in bounds; a cell whose upstream lies outside the grid is
    // bounced back by the flag patch below).
    if (!kGatherEdges && x0 + CPT >= g.X) {
        const bool clamped = x0 + CPT == g.pitch;
        const int jo = g.X - 1 - x0;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if (cx(k) != -1) continue;
            if (clamped) {
#pragma unroll
                for (int j = 0; j + 1 < CPT; ++j) set_lane<CPT>(v[k], j, get_lane<CPT>(v[k], j + 1));
            }
            int ux, uy;
            upstream<MODE>(k, g.X - 1, y, g.X, ux, uy);
            ux = min(max(ux, 0), g.X - 1);
            uy = min(max(uy, 0), g.Y - 1);
#pragma unroll
            for (int j = 0; j < CPT; ++j)
                if (j == jo) set_lane<CPT>(v[k], j, src[(size_t)k * ks + (size_t)uy * g.ys + ux]);
        }
    }
    // from_opp[j] bit k: lane j of v[k] holds the cell's own population opp(k)
    // (bounce-back), which the shifted binary16 format decodes with opp(k)'s shift
    uint32_t from_opp[CPT];
#pragma unroll
    for (int j = 0; j < CPT; ++j) from_opp[j] = 0;
    [[maybe_unused]] float fe[9];
    [[maybe_unused]] bool gathered = false;
    if (fw != 0 || (kGatherEdges && edge)) {
#pragma unroll
        for (int j = 0; j < CPT; ++j) {
            const int x = x0 + j;
            if (x >= g.X) break;
            const uint32_t fl = (uint32_t)(fw >> (8 * j)) & 0xFFu;
            const size_t sown = sown0 + j;
            if (kGatherEdges && edge) {
                if constexpr (St::shifted) {
                    gather_cell<St, MODE, true>(src, dst, flags, mask, g, x, y, own0 + j, sown, fe);
                    gathered = true;
                } else {
                    float f[9];
                    gather_cell<St, MODE, true>(src, dst, flags, mask, g, x, y, own0 + j, sown, f);
#pragma unroll
                    for (int k = 0; k < 9; ++k) set_lane<CPT>(v[k], j, St::store(f[k], k, g.sh));
                }
            } else if (fl == kSlowCell) {
                const bool streamed = mask[own0 + j] != 1 && x != 0 && y < g.Y - 1;
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    if (k == C) continue;
                    set_lane<CPT>(v[k], j, streamed ? src[(size_t)opp(k) * ks + sown] : dst[(size_t)k * ks + sown]);
                }
                if (streamed) from_opp[j] = 0x1FFu & ~(1u << C);
            } else {
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    if (k == C) continue;
                    if ((fl >> bit(k)) & 1u) {
                        set_lane<CPT>(v[k], j, src[(size_t)opp(k) * ks + sown]);
                        from_opp[j] |= 1u << k;
                    }
                }
            }
        }
    }
    V o[9];
    VF mjx, mjy, mrho;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
        float f[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            if constexpr (kGatherEdges && St::shifted) {
                if (gathered) {  // no store/load round trip: the shifted format does not invert exactly
                    f[k] = fe[k];
                    continue;
                }
            }
            f[k] = St::load_b(get_lane<CPT>(v[k], j), k, (from_opp[j] >> k) & 1u, g.sh);
        }
        if constexpr ((OP & OP_STORE_S) && St::shifted) {
            // the pulled state as stored values of each population: bounced
            // values are re-encoded with their new population's shift (values
            // that were not moved between populations come back bit for bit)
#pragma unroll
            for (int k = 0; k < 9; ++k) set_lane<CPT>(v[k], j, St::store(f[k], k, g.sh));
        }
        Moments m;
        if (OP & OP_COLLIDE) {
            m = St::collide(f, g.omega);
#pragma unroll
            for (int k = 0; k < 9; ++k) set_lane<CPT>(o[k], j, St::store(f[k], k, g.sh));
        } else {
            m = St::moments(f, !(OP & OP_NOREST));
        }
        set_lane<CPT>(mjx, j, m.jx);
        set_lane<CPT>(mjy, j, m.jy);
        set_lane<CPT>(mrho, j, m.rho);
        if ((OP & OP_REDUCE) && x0 + j < g.X && mask[own0 + j] != 1) {
            s0 += (double)m.rho;
            s1 += (double)m.jx;
            s2 += (double)m.jy;
        }
    }
    // whole vectors are stored; elements past the last column land in the row's
    // padding, which holds no state
    if (OP & OP_COLLIDE) {
#pragma unroll
        for (int k = 0; k < 9; ++k) vstore<V, NT>(o[k], (V *)(dst + (size_t)k * ks + sown0));
    }
    if (OP & OP_STORE_S) {
#pragma unroll
        for (int k = 0; k < 9; ++k) vstore<V, NT>(v[k], (V *)(dst + (size_t)k * ks + sown0));
    }
    if (OP & OP_MOMENTS) {
        *(VF *)(mom + own0) = mjx;
        *(VF *)(mom + ps + own0) = mjy;
        *(VF *)(mom + 2 * ps + own0) = mrho;
    }
}

template <class St>
__global__ void k_fill(typename St::T *buf, Geom g, float4 v0123, float4 v4567, float v8) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)g.ys * g.Y) return;
    const float v[9] = {v0123.x, v0123.y, v0123.z, v0123.w, v4567.x, v4567.y, v4567.z, v4567.w, v8};
    const int k = (int)((i / (size_t)g.pitch) % 9);
    buf[i] = St::from_f(v[k], k, g.sh);
}

// population k of a set (plane points at element (k, 0, 0)) <-> a dense f32 plane
template <class St>
__global__ void k_encode_plane(const float *__restrict__ dense, typename St::T *__restrict__ plane, Geom g, int k) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= g.X) return;
    plane[(size_t)y * g.ys + x] = St::from_f(dense[(size_t)y * g.X + x], k, g.sh);
}

template <class St>
__global__ void k_decode_plane(const typename St::T *__restrict__ plane, float *__restrict__ dense, Geom g, int k) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= g.X) return;
    dense[(size_t)y * g.X + x] = St::to_f(plane[(size_t)y * g.ys + x], k, g.sh);
}

template <class St>
__global__ void k_set_one(typename St::T *p, float v, int k, F16Shift sh) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = St::from_f(v, k, sh);
}

__global__ void k_paint(const uint32_t *__restrict__ idx, const uint8_t *__restrict__ val, size_t n,
                        uint8_t *__restrict__ mask, int X, int pitch) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t c = idx[i];
    mask[(size_t)(c / X) * pitch + (c % X)] = val[i];
}

__global__ void k_walls(uint8_t *__restrict__ mask, Geom g) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= g.X) return;
    mask[x] = 1;
    mask[(size_t)(g.Y - 1) * g.pitch + x] = 1;
}

// Colour maps (color_map/{inferno,viridis,jet}.wgsl): v is scaled and clamped
// to [-n, n], its integer block picks two neighbouring control colours and the
// fraction within the block blends them; solid cells are black.  The control
// colours are the reference's, block -n first.
__device__ __forceinline__ void colormap_rgb(int map, float v, bool solid, float *rgb) {
    constexpr float kInferno[5][3] = {{0.98828125f, 1.0f, 0.64453125f},
                                      {0.97265625f, 0.55859375f, 0.0390625f},
                                      {0.73828125f, 0.21875f, 0.33203125f},
                                      {0.34375f, 0.06640625f, 0.43359375f},
                                      {0.0f, 0.0f, 0.01853125f}};
    constexpr float kViridis[5][3] = {{0.9921875f, 0.90625f, 0.1484375f},
                                      {0.3671875f, 0.7890625f, 0.3828125f},
                                      {0.1328125f, 0.56640625f, 0.55078125f},
                                      {0.23046875f, 0.32421875f, 0.546875f},
                                      {0.265625f, 0.0078125f, 0.33203125f}};
    constexpr float kJet[9][3] = {{0.0f, 0.0f, 0.5f}, {0.0f, 0.0f, 1.0f}, {0.0f, 0.5f, 1.0f},
                                  {0.0f, 1.0f, 1.0f}, {0.5f, 1.0f, 0.5f}, {1.0f, 1.0f, 0.0f},
                                  {1.0f, 0.5f, 0.0f}, {1.0f, 0.0f, 0.0f}, {0.5f, 0.0f, 0.0f}};
    const bool jet = map == LBM_COLORMAP_JET;
    const int n = jet ? 4 : 2;
    const float c = fminf(fmaxf((jet ? 20.0f : 15.0f) * v, -(float)n), (float)n);
    const int b = (int)floorf(c);
    const float (*tab)[3] = jet ? kJet : (map == LBM_COLORMAP_VIRIDIS ? kViridis : kInferno);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (solid) {
            rgb[k] = 0.0f;
        } else if (b >= n) {
            rgb[k] = tab[2 * n][k];
        } else {
            const float w = (float)(-b) + c;  // right_weight = -block + color
            rgb[k] = (1.0f - w) * tab[b + n][k] + w * tab[b + n + 1][k];
        }
    }
}

// summary passes (summary_stats/{curl,ux,uy,rho,speed}.wgsl); out keeps the
// values of cells a pass skips.  With rgb, the colour map `map` of the whole
// plane follows in the same pass (3 floats per cell, dense rows).
__global__ void k_summary(const float *__restrict__ mom, float *__restrict__ out, const uint8_t *__restrict__ mask,
                          float *__restrict__ rgb, Geom g, int kind, int map) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= g.X) return;
    const long long ps = g.pstride;
    const float *jx = mom, *jy = mom + ps, *rho = mom + 2 * ps;
    const size_t i = (size_t)y * g.pitch + x;
    float v;
    switch (kind) {
    case LBM_SUMMARY_CURL: {
        if (x == 0 || y >= g.Y - 1) {
            v = out[i];  // not written by the curl pass: the previous value stays
            break;
        }
        // flattened neighbours: x+1 at the last column is column 0 of the next row
        const int xe = x + 1 < g.X ? x + 1 : 0, ye = x + 1 < g.X ? y : y + 1;
        const float uye = jy[(size_t)ye * g.pitch + xe];
        const float uyw = jy[i - 1];
        const float uxn = y > 0 ? jx[i - g.pitch] : 0.0f;
        const float uxs = jx[i + g.pitch];
        v = 10.0f * (uye - uyw - uxn + uxs) / rho[i];
        break;
    }
    case LBM_SUMMARY_UX: v = jx[i]; break;
    case LBM_SUMMARY_UY: v = jy[i]; break;
    case LBM_SUMMARY_RHO: v = 4.0f * fminf(fmaxf(0.15f * rho[i], 0.0f), 1.0f) - 0.5f; break;
    default: v = fminf(fmaxf(5.0f * sqrtf(jx[i] * jx[i] + jy[i] * jy[i]), 0.0f), 1.0f) - 0.5f; break;
    }
    out[i] = v;
    if (rgb) colormap_rgb(map, v, mask[i] == 1, rgb + 3 * ((size_t)y * g.X + x));
}

// rho, u = j / rho, curl into four dense planes
__global__ void k_fields(const float *__restrict__ mom, float *__restrict__ outp, Geom g) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= g.X) return;
    const long long ps = g.pstride;
    const float *jx = mom, *jy = mom + ps, *rho = mom + 2 * ps;
    const size_t i = (size_t)y * g.pitch + x;
    const size_t d = (size_t)y * g.X + x, dn = (size_t)g.X * g.Y;
    const float r = rho[i];
    outp[d] = r;
    outp[dn + d] = jx[i] / r;
    outp[2 * dn + d] = jy[i] / r;
    float curl = 0.0f;
    if (x != 0 && y < g.Y - 1) {
        const int xe = x + 1 < g.X ? x + 1 : 0, ye = x + 1 < g.X ? y : y + 1;
        const float uye = jy[(size_t)ye * g.pitch + xe];
        const float uyw = jy[i - 1];
        const float uxn = y > 0 ? jx[i - g.pitch] : 0.0f;
        const float uxs = jx[i + g.pitch];
        curl = 10.0f * (uye - uyw - uxn + uxs) / r;
    }
    outp[3 * dn + d] = curl;
}

// one block sums the per-block partials in a fixed order (deterministic)
__global__ __launch_bounds__(1024) void k_reduce_final(const double *__restrict__ partials, size_t n,
                                                       double *__restrict__ out) {
    __shared__ double red[1024 / 64][3];
    double s[3] = {0.0, 0.0, 0.0};
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) {
        s[0] += partials[i * 3 + 0];
        s[1] += partials[i * 3 + 1];
        s[2] += partials[i * 3 + 2];
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
        for (int off = 32; off > 0; off >>= 1) s[c] += __shfl_xor(s[c], off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
        for (int c = 0; c < 3; ++c) red[wv][c] = s[c];
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) t += red[w][threadIdx.x];
        out[threadIdx.x] = t;
    }
}

// every element of one set (9 * pitch * ny): row slot i / pitch = 9 y + k
// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

thread_local std::string g_create_error;

struct HipError {
    hipError_t e;
    const char *what;
};

// A failed runtime call is reported by the exception; its status is also
// cleared from the thread's last error so that a later launch check does not
// report it a second time.
#define HIPCHK(expr)                                  \
    do {                                              \
        hipError_t e_ = (expr);                       \
        if (e_ != hipSuccess) {                       \
            (void)hipGetLastError();                  \
            throw HipError{e_, #expr};                \
        }                                             \
    } while (0)

struct ArgError {
    std::string msg;
};

// "<arch name>/<CU count>cu" of a device: the device part of an autotune
// cache key (a partitioned device with fewer CUs tunes apart)
inline std::string device_tag(int dev) {
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    return std::string(p.gcnArchName) + "/" + std::to_string(p.multiProcessorCount) + "cu";
}

inline const char *precision_tag(int precision) {
    return precision == LBM_PRECISION_F32 ? "f32" : precision == LBM_PRECISION_F16S ? "f16s" : "f16e";
}

// Test hook: with LBM_TEST_ALLOC_LIMIT=<bytes> in the environment, a request
// above that many bytes is passed on as an impossible size, so the runtime
// fails it the way it fails a real out-of-memory request.
inline size_t test_alloc_limit() {
    const char *s = std::getenv("LBM_TEST_ALLOC_LIMIT");
    return s && *s ? (size_t)std::strtoull(s, nullptr, 10) : SIZE_MAX;
}

template <class T>
T *dalloc(size_t n) {
    void *p = nullptr;
    if (n == 0) n = 1;
    const size_t bytes = n * sizeof(T);
    hipError_t e = hipMalloc(&p, bytes > test_alloc_limit() ? (size_t)1 << 62 : bytes);
    if (e != hipSuccess) {
        // the runtime keeps a failed call's status as the thread's last error,
        // which the next launch check would report: clear it, the failure is
        // reported here
        (void)hipGetLastError();
        throw HipError{e, "hipMalloc"};
    }
    return (T *)p;
}

}  // namespace
