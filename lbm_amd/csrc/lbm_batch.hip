// Batched engine: many independent D2Q9 domains of uneven size and obstacle
// layout stepped together, one launch pair (interior + boundary) per step for
// all domains, replayed from a hipGraph.
//
// Each domain follows exactly the single-grid semantics (the reference step,
// lbm-wgpu src/lbm.rs:1112-1134, with the S-form / P-form state machine of
// lbm_grid.hip) and uses the same device code (lbm_common.h), so a domain of a
// batch equals a single grid of the same size bit for bit.
//
// Work decomposition.  Every domain is cut into rows of 64-cell segments.  A
// tile is the span of one wave (64 * CPT cells of one row); a tile whose
// segments are all clear (streamed, no bounce-back, before the outlet column)
// goes to the interior list, every segment of any other tile to the boundary
// list.  Both kernels are persistent: a grid sized to the GPU strides over its
// list, each wave taking one tile (interior) or one segment (boundary) and
// reading its domain's geometry from the domain table.  The lists are rebuilt
// on the host after a mask change, in a fixed order.
//
// Memory: one allocation per parity set holding every domain's region in the
// single-grid row-interleaved layout; masks, bounce-back bytes, segment bytes
// and lagged moments likewise packed per domain.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "lbm_common.h"
#include "tune_cache.h"
#include "pair.h"

namespace {

struct DomainDev {
    long long f_off;     // element offset of the domain in each parity set
    long long aux_off;   // element offset in the mask / flags planes (and moment planes)
    long long seg_off;   // offset in the segment plane
    Geom g;
};

// list entries: x = domain, y = (row << 16) | first 64-cell segment
template <class St, int MODE, int OP, int CPT, bool NT>
__global__ __launch_bounds__(kBlock) void k_batch_interior(const typename St::T *src, typename St::T *dst,
                                                           const uint8_t *__restrict__ mask,
                                                           const DomainDev *__restrict__ table,
                                                           const uint2 *__restrict__ list, uint32_t n,
                                                           float *__restrict__ mom, double *__restrict__ partials,
                                                           int chunked) {
    const uint32_t lane = threadIdx.x & 63;
    // wave-uniform list positions: the list entry and the domain's table entry
    // then come through scalar loads, once per wave instead of once per lane
    const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
    const uint32_t waves = gridDim.x * (kBlock / 64);
    uint32_t t = wave, stride = waves, end = n;
    if (chunked) {  // each wave a contiguous run of the list
        const uint32_t per = (n + waves - 1) / waves;
        t = wave * per;
        end = min(n, t + per);
        stride = 1;
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (; t < end; t += stride) {
        const uint2 e = list[t];
        const DomainDev d = table[e.x];
        const int y = (int)(e.y >> 16);
        const int x0 = (int)(e.y & 0xFFFFu) * 64 + (int)lane * CPT;
        if (x0 < d.g.X)
            interior_tile<St, MODE, true, OP, CPT, NT>(src + d.f_off, dst + d.f_off, mask + d.aux_off, d.g, y, x0,
                                                      mom ? mom + 3 * d.aux_off : nullptr, s0, s1, s2);
    }
    if (OP & OP_REDUCE) block_sums(s0, s1, s2, partials, blockIdx.x);
}

template <class St, int MODE, bool PFORM, int OP>
__global__ __launch_bounds__(kBlock) void k_batch_boundary(const typename St::T *src, typename St::T *dst,
                                                           const uint8_t *__restrict__ flags,
                                                           const uint8_t *__restrict__ mask,
                                                           const DomainDev *__restrict__ table,
                                                           const uint2 *__restrict__ list, uint32_t n,
                                                           float *__restrict__ mom, double *__restrict__ partials) {
    const int lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (kBlock / 64);
    // wave-uniform list position (scalar loads of the list and table entries)
    const uint32_t wave = __builtin_amdgcn_readfirstlane(blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6));
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (uint32_t t = wave; t < n; t += waves) {
        const uint2 e = list[t];
        const DomainDev d = table[e.x];
        const int y = (int)(e.y >> 16);
        const int x = (int)(e.y & 0xFFFFu) * 64 + lane;
        if (x >= d.g.X) continue;
        const size_t own = (size_t)y * d.g.pitch + x, sown = (size_t)y * d.g.ys + x;
        float f[9];
        const bool streamed = gather_cell<St, MODE, PFORM>(src + d.f_off, dst + d.f_off, flags + d.aux_off,
                                                           mask + d.aux_off, d.g, x, y, own, sown, f);
        finish_cell<St, OP>(f, streamed, src + d.f_off, dst + d.f_off, mask + d.aux_off, d.g, own, sown,
                            mom ? mom + 3 * d.aux_off : nullptr, s0, s1, s2);
    }
    if (OP & OP_REDUCE) block_sums(s0, s1, s2, partials, blockIdx.x);
}

// Two steps per launch for every domain (the single grid's two-step scheme,
// pair.h): one block per work item (domain, 254-column strip, row chunk) of
// the pair list, blocks of one XCD on one contiguous run of it, then the
// fixed-cell pass's blocks.  A, B: the current sets; A2, B2: the spare sets
// that receive C_{n+2} and C_{n+1}.
template <class St, int MODE>
__global__ __launch_bounds__(kBlock) void k_batch_pair(const typename St::T *A, const typename St::T *B,
                                                       typename St::T *__restrict__ A2, typename St::T *__restrict__ B2,
                                                       const uint8_t *__restrict__ flags,
                                                       const DomainDev *__restrict__ table,
                                                       const uint2 *__restrict__ list, int n_items, int rows,
                                                       const uint2 *__restrict__ fixed,
                                                       const uint2 *__restrict__ rowtab, uint32_t n_fixed) {
    __shared__ typename St::T ring[kRing][8][kBlock];
    if ((int)blockIdx.x >= n_items) {
        // the fixed-cell pass: one thread per list entry (column | streamed << 31, row-table index)
        const uint32_t i = (blockIdx.x - n_items) * kBlock + threadIdx.x;
        if (i < n_fixed) {
            const uint2 f = fixed[i];
            const uint2 r = rowtab[f.y];
            const DomainDev d = table[r.x];
            fixed_cell<St>(A + d.f_off, B + d.f_off, d.g, (int)(f.x & 0x7FFFFFFFu), (int)r.y, (f.x >> 31) != 0,
                           A2 + d.f_off, B2 + d.f_off);
        }
        return;
    }
    const uint2 e = list[xcd_remap((int)blockIdx.x, n_items)];
    const DomainDev d = table[e.x];
    const int strip = (int)(e.y >> 16), r0 = (int)(e.y & 0xFFFFu) * rows;
    pair_block_any<St, MODE>(ring, A + d.f_off, B + d.f_off, A2 + d.f_off, B2 + d.f_off, flags + d.aux_off, d.g,
                             strip, r0, min(r0 + rows, d.g.Y));
}

// the fixed-cell list of every domain (pair.h): one block per entry
// (domain, row) of the row table; entries (column | streamed << 31,
// row-table index)
__global__ __launch_bounds__(kFixedBlock) void k_batch_fixed_count(const uint8_t *__restrict__ flags,
                                                                   const DomainDev *__restrict__ table,
                                                                   const uint2 *__restrict__ rowtab,
                                                                   uint32_t *__restrict__ counts) {
    const uint2 e = rowtab[blockIdx.x];
    const DomainDev d = table[e.x];
    fixed_row_count(flags + d.aux_off + (size_t)e.y * d.g.pitch, d.g.X, counts + blockIdx.x);
}

__global__ __launch_bounds__(kFixedBlock) void k_batch_fixed_emit(const uint8_t *__restrict__ flags,
                                                                  const uint8_t *__restrict__ mask,
                                                                  const DomainDev *__restrict__ table,
                                                                  const uint2 *__restrict__ rowtab,
                                                                  const uint32_t *__restrict__ offsets,
                                                                  uint2 *__restrict__ list) {
    const uint2 e = rowtab[blockIdx.x];
    const DomainDev d = table[e.x];
    const size_t r = d.aux_off + (size_t)e.y * d.g.pitch;
    fixed_row_emit(flags + r, mask + r, d.g.X, (int)e.y, d.g.Y, blockIdx.x, list + offsets[blockIdx.x]);
}

thread_local std::string g_batch_create_error;

}  // namespace

struct lbm_batch {
    int dev = 0;
    hipStream_t stream = nullptr;
    int precision = 0, boundary = 0;
    size_t esize = 4;
    std::vector<lbm_domain> doms;
    std::vector<DomainDev> table;
    size_t f_elems = 0, aux_elems = 0, seg_bytes = 0, tail_elems = 0;
    void *f[2] = {nullptr, nullptr};
    uint8_t *mask = nullptr, *flags = nullptr, *segs = nullptr;
    float *lag = nullptr;  // per domain 3 moment planes at 3 * aux_off
    float *dense = nullptr;
    size_t dense_elems = 0;
    DomainDev *dtable = nullptr;
    uint2 *ilist = nullptr, *blist = nullptr, *alist = nullptr;
    uint32_t n_interior = 0, n_boundary = 0, n_all = 0;
    // two-step variants: the output sets and the (domain, strip, row chunk)
    // work list for `plist_rows`-row chunks
    void *spare[2] = {nullptr, nullptr};
    uint2 *plist = nullptr;
    uint32_t n_pair = 0;
    uint2 *fixed = nullptr, *rowtab = nullptr;  // fixed-cell list (column | streamed << 31, row-table index); (domain, row)
    uint32_t *fixed_rows = nullptr;             // the list's offset per row-table entry
    uint32_t n_fixed = 0, n_rows = 0;
    bool fixed_dirty = true;
    int plist_rows = 0;
    int list_cpt = 0;  // tile width (cells per thread) the interior list was built for
    bool dirty = true;  // masks changed since the lists were built
    int cur = 0;
    bool pform = false;
    uint64_t steps = 0;
    int variant = 0;
    bool use_graph = true;
    // fused steps starting from cur = 0 / 1 and, for the two-step variants, from
    // either arrangement of the state and spare sets (each two-step launch
    // trades them; a graph holds the pointers it was captured with)
    hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int pair_parity = 0;  // flips with every two-step launch
    int graph_variant = -1;
    int grid_blocks = 1024;
    std::string err;
    uint64_t cells = 0;

    void drop_graphs() {
        for (auto &row : graph)
            for (auto &gx : row)
                if (gx) {
                    (void)hipGraphExecDestroy(gx);
                    gx = nullptr;
                }
        graph_variant = -1;
    }

    void release() {
        (void)hipSetDevice(dev);
        drop_graphs();
        release_pair_buffers();
        for (void *p : {f[0], f[1], (void *)mask, (void *)flags, (void *)segs, (void *)lag, (void *)dense,
                        (void *)dtable, (void *)ilist, (void *)blist, (void *)alist})
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    // variants of the interior kernel: cells per thread and list order
    //   v        0  1  2  3  4  5
    //   cells    1  1  2  2  4  4      (f16s: 2 2 4 4 8 8)
    //   chunked  0  1  0  1  0  1
    // variants 6, 7: two steps per launch (k_batch_pair) in chunks of 8 / 16
    // rows; their one-step launches (the first and last steps of a call) run
    // as the two-cells-per-thread variant (f32 2, binary16 0)
    static constexpr int kPairFirst = 6;
    static bool pair_variant(int v) { return v >= kPairFirst; }
    int one_step_variant() const {
        return pair_variant(variant) ? (precision == LBM_PRECISION_F32 ? 2 : 0) : variant;
    }
    int cpt_of(int v) const {
        const int c = v / 2 == 0 ? 1 : (v / 2 == 1 ? 2 : 4);
        return precision == LBM_PRECISION_F32 ? c : 2 * c;
    }

    char *fptr(int s, const DomainDev &d) const { return (char *)f[s] + (size_t)d.f_off * esize; }

    // ---- lists -----------------------------------------------------------------
    void rebuild_lists(int cpt) {
        // bounce-back and segment bytes on the device, lists on the host
        fixed_dirty = true;
        for (size_t i = 0; i < table.size(); ++i) {
            const DomainDev &d = table[i];
            const dim3 gf((d.g.X + 255) / 256, d.g.Y);
            if (boundary == LBM_BOUNDARY_REFERENCE)
                hipLaunchKernelGGL(k_build_flags<LBM_BOUNDARY_REFERENCE>, gf, dim3(256), 0, stream, mask + d.aux_off,
                                   flags + d.aux_off, d.g);
            else
                hipLaunchKernelGGL(k_build_flags<LBM_BOUNDARY_ZERO_GRADIENT>, gf, dim3(256), 0, stream,
                                   mask + d.aux_off, flags + d.aux_off, d.g);
            HIPCHK(hipGetLastError());
            const int srow = d.g.pitch / 64 + kSegSlack;
            hipLaunchKernelGGL(k_build_segments, dim3((srow + 255) / 256, d.g.Y), dim3(256), 0, stream,
                               flags + d.aux_off, segs + d.seg_off, nullptr, nullptr, d.g);
            HIPCHK(hipGetLastError());
        }
        std::vector<uint8_t> hs(seg_bytes);
        HIPCHK(hipMemcpyAsync(hs.data(), segs, seg_bytes, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        std::vector<uint2> il, bl, al;
        const int span = cpt;  // segments per tile (a wave covers 64 * cpt cells)
        for (size_t i = 0; i < table.size(); ++i) {
            const DomainDev &d = table[i];
            const int srow = d.g.pitch / 64 + kSegSlack;
            const int nseg = (d.g.X + 63) / 64;
            for (int y = 0; y < d.g.Y; ++y) {
                const uint8_t *row = hs.data() + d.seg_off + (size_t)y * srow;
                for (int s = 0; s < nseg; ++s) al.push_back(make_uint2((uint32_t)i, ((uint32_t)y << 16) | s));
                for (int s0 = 0; s0 < nseg; s0 += span) {
                    bool clear = true;
                    for (int s = s0; s < s0 + span; ++s) clear = clear && s < nseg && row[s] == 1;
                    if (clear) {
                        il.push_back(make_uint2((uint32_t)i, ((uint32_t)y << 16) | s0));
                    } else {
                        for (int s = s0; s < std::min(nseg, s0 + span); ++s)
                            bl.push_back(make_uint2((uint32_t)i, ((uint32_t)y << 16) | s));
                    }
                }
            }
        }
        auto upload = [&](uint2 *&dptr, const std::vector<uint2> &v) {
            if (dptr) (void)hipFree(dptr);
            dptr = dalloc<uint2>(std::max<size_t>(1, v.size()));
            if (!v.empty()) HIPCHK(hipMemcpyAsync(dptr, v.data(), v.size() * sizeof(uint2), hipMemcpyHostToDevice, stream));
        };
        upload(ilist, il);
        upload(blist, bl);
        upload(alist, al);
        HIPCHK(hipStreamSynchronize(stream));
        n_interior = (uint32_t)il.size();
        n_boundary = (uint32_t)bl.size();
        n_all = (uint32_t)al.size();
        list_cpt = cpt;
        dirty = false;
        drop_graphs();
    }

    void ensure_lists() {
        const int cpt = cpt_of(one_step_variant());
        if (dirty || list_cpt != cpt) rebuild_lists(cpt);
    }

    // ---- launches -----------------------------------------------------------------
    template <class St, int MODE, int OP, int CPT, bool NT>
    void interior(const void *src, void *dst, float *mom, int chunked) {
        using T = typename St::T;
        hipLaunchKernelGGL((k_batch_interior<St, MODE, OP, CPT, NT>), dim3(grid_blocks), dim3(kBlock), 0, stream,
                           (const T *)src, (T *)dst, mask, dtable, ilist, n_interior, mom, nullptr, chunked);
        HIPCHK(hipGetLastError());
    }

    template <class St, int MODE, bool PFORM, int OP>
    void boundary_pass(const void *src, void *dst, const uint2 *list, uint32_t n, float *mom) {
        using T = typename St::T;
        hipLaunchKernelGGL((k_batch_boundary<St, MODE, PFORM, OP>), dim3(grid_blocks), dim3(kBlock), 0, stream,
                           (const T *)src, (T *)dst, flags, mask, dtable, list, n, mom, nullptr);
        HIPCHK(hipGetLastError());
    }

    template <class St, int MODE, int OP>
    void pstep(const void *src, void *dst, float *mom) {
        constexpr bool f32 = sizeof(typename St::T) == 4;
        const int v = one_step_variant();
        const int chunked = v & 1;
        switch (v / 2) {
        case 0: interior<St, MODE, OP, (f32 ? 1 : 2), true>(src, dst, mom, chunked); break;
        case 1: interior<St, MODE, OP, (f32 ? 2 : 4), true>(src, dst, mom, chunked); break;
        default: interior<St, MODE, OP, (f32 ? 4 : 8), true>(src, dst, mom, chunked); break;
        }
        boundary_pass<St, MODE, true, OP>(src, dst, blist, n_boundary, mom);
    }

    template <class St, int MODE>
    void do_step(bool last) {
        if (!pform) {
            // S-form: collide in place, every cell of every domain
            if (last) boundary_pass<St, MODE, false, OP_COLLIDE | OP_MOMENTS>(f[cur], f[cur], alist, n_all, lag);
            else boundary_pass<St, MODE, false, OP_COLLIDE>(f[cur], f[cur], alist, n_all, nullptr);
            pform = true;
        } else {
            if (last) pstep<St, MODE, OP_COLLIDE | OP_MOMENTS>(f[cur], f[1 - cur], lag);
            else pstep<St, MODE, OP_COLLIDE>(f[cur], f[1 - cur], nullptr);
            cur = 1 - cur;
        }
    }

    template <class St, int MODE>
    void flush_t() {
        if (!pform) return;
        pstep<St, MODE, OP_STORE_S>(f[cur], f[1 - cur], nullptr);
        cur = 1 - cur;
        pform = false;
    }

    // ---- two steps per launch ---------------------------------------------------------
    size_t set_bytes() const { return f_elems * esize; }
    // a set's allocation: two rows of the widest domain more, which the
    // two-step kernel's loads in the last domain's last rows read and discard
    size_t alloc_bytes() const { return (f_elems + tail_elems) * esize; }

    void release_pair_buffers() {
        drop_graphs();
        for (void *&q : spare) {
            if (q) (void)hipFree(q);
            q = nullptr;
        }
        if (plist) (void)hipFree(plist);
        plist = nullptr;
        n_pair = 0;
        plist_rows = 0;
        for (void *q : {(void *)fixed, (void *)rowtab, (void *)fixed_rows})
            if (q) (void)hipFree(q);
        fixed = rowtab = nullptr;
        fixed_rows = nullptr;
        n_fixed = 0;
    }

    // allocated when a two-step variant is selected, so that stepping itself
    // does not allocate: the output sets, and the fixed-cell list (room for
    // every cell) with its row table (domain, row) and per-row offsets
    void ensure_pair_buffers() {
        if (spare[0] && spare[1] && fixed) return;
        release_pair_buffers();
        try {
            spare[0] = dalloc<char>(alloc_bytes());
            spare[1] = dalloc<char>(alloc_bytes());
            std::vector<uint2> rt;
            for (size_t i = 0; i < table.size(); ++i)
                for (int y = 0; y < table[i].g.Y; ++y) rt.push_back(make_uint2((uint32_t)i, (uint32_t)y));
            n_rows = (uint32_t)rt.size();
            fixed = dalloc<uint2>((size_t)cells);
            rowtab = dalloc<uint2>(rt.size());
            fixed_rows = dalloc<uint32_t>(rt.size());
            // the rows' padding is never written by the two-step kernel: keep it defined
            HIPCHK(hipMemsetAsync(spare[0], 0, set_bytes(), stream));
            HIPCHK(hipMemsetAsync(spare[1], 0, set_bytes(), stream));
            HIPCHK(hipMemcpyAsync(rowtab, rt.data(), rt.size() * sizeof(uint2), hipMemcpyHostToDevice, stream));
            HIPCHK(hipStreamSynchronize(stream));
            fixed_dirty = true;
        } catch (...) {
            release_pair_buffers();
            throw;
        }
    }

    // the fixed-cell list from the current flags; waits for the stream.  Drops
    // the graphs, whose fixed-cell launches hold the old count.
    void build_fixed_list() {
        hipLaunchKernelGGL(k_batch_fixed_count, dim3(n_rows), dim3(kFixedBlock), 0, stream, flags, dtable, rowtab,
                           fixed_rows);
        HIPCHK(hipGetLastError());
        std::vector<uint32_t> off(n_rows);
        HIPCHK(hipMemcpyAsync(off.data(), fixed_rows, off.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        uint32_t total = 0;
        for (uint32_t &o : off) {
            const uint32_t c = o;
            o = total;
            total += c;
        }
        HIPCHK(hipMemcpyAsync(fixed_rows, off.data(), off.size() * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_batch_fixed_emit, dim3(n_rows), dim3(kFixedBlock), 0, stream, flags, mask, dtable,
                           rowtab, fixed_rows, fixed);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream));  // `off` is read by the copy above
        n_fixed = total;
        fixed_dirty = false;
        drop_graphs();
    }

    int pair_rows() const { return 8 << (variant - kPairFirst); }

    void ensure_pair_list() {
        const int rows = pair_rows();
        if (plist && plist_rows == rows) return;
        std::vector<uint2> pl;
        for (size_t i = 0; i < table.size(); ++i) {
            const Geom &g = table[i].g;
            const int nstrips = (g.X + kBlock - 3) / (kBlock - 2), nchunks = (g.Y + rows - 1) / rows;
            for (int c = 0; c < nchunks; ++c)
                for (int sx = 0; sx < nstrips; ++sx)
                    pl.push_back(make_uint2((uint32_t)i, ((uint32_t)sx << 16) | (uint32_t)c));
        }
        uint2 *d = dalloc<uint2>(pl.size());
        try {
            HIPCHK(hipMemcpyAsync(d, pl.data(), pl.size() * sizeof(uint2), hipMemcpyHostToDevice, stream));
            HIPCHK(hipStreamSynchronize(stream));
        } catch (...) {
            (void)hipFree(d);
            throw;
        }
        if (plist) (void)hipFree(plist);
        plist = d;
        n_pair = (uint32_t)pl.size();
        plist_rows = rows;
        drop_graphs();
    }

    // C_n, C_{n-1} in f[cur], f[1 - cur] -> C_{n+2}, C_{n+1} (the latter only
    // where the next step reads it); the spare sets take the old ones
    template <class St, int MODE>
    void launch_pair() {
        using T = typename St::T;
        const T *A = (const T *)f[cur], *B = (const T *)f[1 - cur];
        T *A2 = (T *)spare[0], *B2 = (T *)spare[1];
        const unsigned nfix = (n_fixed + kBlock - 1) / kBlock;
        hipLaunchKernelGGL((k_batch_pair<St, MODE>), dim3(n_pair + nfix), dim3(kBlock), 0, stream, A, B, A2, B2,
                           flags, dtable, plist, (int)n_pair, pair_rows(), fixed, rowtab, n_fixed);
        HIPCHK(hipGetLastError());
        spare[0] = f[cur];
        spare[1] = f[1 - cur];
        f[cur] = A2;
        f[1 - cur] = B2;
        pair_parity ^= 1;
    }

    void pair() {
        dispatch([&]<class St, int MODE>() { launch_pair<St, MODE>(); });
    }

    template <class F>
    void dispatch(F &&fn) {
        by_store(precision, [&]<class St>() {
            if (boundary == LBM_BOUNDARY_REFERENCE) fn.template operator()<St, LBM_BOUNDARY_REFERENCE>();
            else fn.template operator()<St, LBM_BOUNDARY_ZERO_GRADIENT>();
        });
    }

    void flush() {
        dispatch([&]<class St, int MODE>() { flush_t<St, MODE>(); });
    }

    // two fused steps from cur = c (four for the two-step variants), captured
    // once per arrangement of the sets and replayed
    hipGraphExec_t pair_graph(int c) {
        if (graph_variant != variant) drop_graphs();
        hipGraphExec_t &gx = graph[c][pair_parity];
        if (!gx) {
            const int saved = cur;
            cur = c;
            hipGraph_t gr;
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            try {
                if (pair_variant(variant)) {
                    pair();  // f[] and spare[] trade places ...
                    pair();  // ... and trade back: the graph leaves them as it found them
                } else {
                    dispatch([&]<class St, int MODE>() {
                        do_step<St, MODE>(false);
                        do_step<St, MODE>(false);
                    });
                }
            } catch (...) {
                hipGraph_t dummy;
                (void)hipStreamEndCapture(stream, &dummy);
                if (dummy) (void)hipGraphDestroy(dummy);
                cur = saved;
                throw;
            }
            HIPCHK(hipStreamEndCapture(stream, &gr));
            hipError_t e = hipGraphInstantiate(&gx, gr, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gr);
            cur = saved;
            HIPCHK(e);
            graph_variant = variant;
        }
        return gx;
    }

    void step(uint32_t n) {
        if (n == 0) return;
        ensure_lists();
        const bool two = pair_variant(variant);
        if (two) {
            ensure_pair_buffers();
            ensure_pair_list();
            if (fixed_dirty) build_fixed_list();
        }
        uint32_t left = n;
        if (!pform) {
            dispatch([&]<class St, int MODE>() { do_step<St, MODE>(left == 1); });
            --left;
        }
        // all but the last step from the graph: two steps at a time, or four
        // (two two-step launches) for the two-step variants
        const uint32_t per = two ? 4 : 2;
        if (use_graph) {
            while (left > per) {
                HIPCHK(hipGraphLaunch(pair_graph(cur), stream));
                left -= per;  // the graph leaves cur (and the sets) where it started
            }
        }
        // two-step launches while at least one step is left for the closing
        // one-step launch, which leaves both sets whole
        while (two && left >= 3) {
            pair();
            left -= 2;
        }
        while (left > 0) {
            dispatch([&]<class St, int MODE>() { do_step<St, MODE>(left == 1); });
            --left;
        }
        steps += n;
    }

    void ensure_dense(size_t n) {
        if (dense_elems < n) {
            if (dense) (void)hipFree(dense);
            dense = dalloc<float>(n);
            dense_elems = n;
        }
    }
};

namespace {

int bfail(lbm_batch *b, int code, const std::string &msg) {
    if (b) b->err = msg;
    else g_batch_create_error = msg;
    return code;
}

template <class F>
int bguard(lbm_batch *b, F &&fn) {
    if (!b) return bfail(nullptr, LBM_EINVAL, "null batch");
    try {
        HIPCHK(hipSetDevice(b->dev));
        fn();
        b->err.clear();
        return LBM_OK;
    } catch (const HipError &e) {
        return bfail(b, e.e == hipErrorOutOfMemory ? LBM_ENOMEM : LBM_EHIP,
                     std::string(e.what) + ": " + hipGetErrorString(e.e));
    } catch (const ArgError &e) {
        return bfail(b, LBM_EINVAL, e.msg);
    } catch (const std::bad_alloc &) {
        return bfail(b, LBM_ENOMEM, "host allocation failed");
    } catch (...) {
        return bfail(b, LBM_EHIP, "unexpected exception");
    }
}

void bcheck(bool ok, const char *msg) {
    if (!ok) throw ArgError{msg};
}

}  // namespace

extern "C" {

const char *lbm_batch_last_error(const lbm_batch *b) { return b ? b->err.c_str() : g_batch_create_error.c_str(); }

int lbm_batch_create(const lbm_domain *doms, uint32_t n, int32_t precision, int32_t boundary, int32_t device,
                     lbm_batch **out) {
    if (!doms || !out || n == 0) return bfail(nullptr, LBM_EINVAL, "need at least one domain");
    *out = nullptr;
    if (precision != LBM_PRECISION_F32 && precision != LBM_PRECISION_F16S && precision != LBM_PRECISION_F16E)
        return bfail(nullptr, LBM_EINVAL, "unknown precision");
    if (boundary != LBM_BOUNDARY_REFERENCE && boundary != LBM_BOUNDARY_ZERO_GRADIENT)
        return bfail(nullptr, LBM_EINVAL, "unknown boundary mode");
    for (uint32_t i = 0; i < n; ++i)
        if (doms[i].nx < 3 || doms[i].ny < 3 || doms[i].ny > 65535 || doms[i].nx > 64 * 65535 ||
            !std::isfinite(doms[i].omega))
            return bfail(nullptr, LBM_EINVAL, "domain sizes must be 3..65535 and omega finite");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return bfail(nullptr, LBM_ENODEV, "no HIP device");
    if (device < 0 || device >= ndev) return bfail(nullptr, LBM_EINVAL, "bad device ordinal");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return bfail(nullptr, LBM_ENODEV, "no properties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return bfail(nullptr, LBM_ENODEV, std::string("built for gfx950, device is ") + prop.gcnArchName);

    lbm_batch *b = new (std::nothrow) lbm_batch();
    if (!b) return bfail(nullptr, LBM_ENOMEM, "host allocation failed");
    b->dev = device;
    b->precision = precision;
    b->boundary = boundary;
    b->esize = precision == LBM_PRECISION_F32 ? 4 : 2;
    b->doms.assign(doms, doms + n);
    b->grid_blocks = prop.multiProcessorCount * 8;
    long long fo = 0, ao = 0, so = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const int X = (int)doms[i].nx, Y = (int)doms[i].ny;
        const int pitch = (X + 63) / 64 * 64;
        const long long pst = ((long long)pitch * Y + 255) / 256 * 256;
        DomainDev d;
        d.f_off = fo;
        d.aux_off = ao;
        d.seg_off = so;
        d.g = Geom{X, Y, pitch, pst, 9LL * pitch, (long long)pitch, doms[i].omega};
        d.g.sh = make_f16_shift(precision == LBM_PRECISION_F16E, doms[i].u_in);  // each domain's own free stream
        b->table.push_back(d);
        fo += (9LL * pitch * Y + 255) / 256 * 256;
        ao += pst;
        so += ((long long)(pitch / 64 + kSegSlack) * Y + 15) / 16 * 16;
        b->cells += (uint64_t)X * Y;
        b->tail_elems = std::max(b->tail_elems, 18 * (size_t)pitch);
    }
    b->f_elems = fo;
    b->aux_elems = ao;
    b->seg_bytes = so;
    int rc = bguard(b, [&] {
        HIPCHK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        b->f[0] = dalloc<char>(b->alloc_bytes());
        b->f[1] = dalloc<char>(b->alloc_bytes());
        b->mask = dalloc<uint8_t>(b->aux_elems);
        b->flags = dalloc<uint8_t>(b->aux_elems);
        b->segs = dalloc<uint8_t>(b->seg_bytes);
        b->lag = dalloc<float>(3 * b->aux_elems);
        b->dtable = dalloc<DomainDev>(n);
        HIPCHK(hipMemcpyAsync(b->dtable, b->table.data(), n * sizeof(DomainDev), hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipMemsetAsync(b->lag, 0, 3 * b->aux_elems * sizeof(float), b->stream));
        HIPCHK(hipMemsetAsync(b->mask, 0, b->aux_elems, b->stream));
        for (const DomainDev &d : b->table) {
            hipLaunchKernelGGL(k_walls, dim3((d.g.X + 255) / 256), dim3(256), 0, b->stream, b->mask + d.aux_off, d.g);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(b->stream));
    });
    if (rc != LBM_OK) {
        g_batch_create_error = b->err;
        b->release();
        delete b;
        return rc;
    }
    rc = lbm_batch_reset(b);
    if (rc != LBM_OK) {
        g_batch_create_error = b->err;
        b->release();
        delete b;
        return rc;
    }
    *out = b;
    return LBM_OK;
}

int lbm_batch_destroy(lbm_batch *b) {
    if (!b) return LBM_EINVAL;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    b->release();
    delete b;
    return LBM_OK;
}

int lbm_batch_reset(lbm_batch *b) {
    return bguard(b, [&] {
        for (size_t i = 0; i < b->table.size(); ++i) {
            const DomainDev &d = b->table[i];
            float v[9];
            host_feq(b->doms[i].u_in, 0.0f, 1.0f, v);
            const float4 a = make_float4(v[0], v[1], v[2], v[3]);
            const float4 c = make_float4(v[4], v[5], v[6], v[7]);
            const dim3 grid((unsigned)(((size_t)d.g.ys * d.g.Y + 255) / 256));
            for (int s = 0; s < 2; ++s) {
                by_store(b->precision, [&]<class St>() {
                    hipLaunchKernelGGL(k_fill<St>, grid, dim3(256), 0, b->stream, (typename St::T *)b->fptr(s, d), d.g,
                                       a, c, v[8]);
                });
                HIPCHK(hipGetLastError());
            }
        }
        b->cur = 0;
        b->pform = false;
        b->steps = 0;
        HIPCHK(hipStreamSynchronize(b->stream));
    });
}

int lbm_batch_set_mask(lbm_batch *b, uint32_t d, const uint8_t *solid) {
    return bguard(b, [&] {
        bcheck(d < b->table.size() && solid, "bad domain or null mask");
        const DomainDev &dd = b->table[d];
        std::vector<uint8_t> m((size_t)dd.g.X * dd.g.Y);
        for (size_t i = 0; i < m.size(); ++i) m[i] = solid[i] == 1 ? 1 : 0;
        b->flush();  // the pending stream uses the old mask
        HIPCHK(hipMemcpy2DAsync(b->mask + dd.aux_off, dd.g.pitch, m.data(), dd.g.X, dd.g.X, dd.g.Y,
                                hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
        b->dirty = true;
    });
}

int lbm_batch_step(lbm_batch *b, uint32_t n) {
    return bguard(b, [&] { b->step(n); });
}

int lbm_batch_step_timed(lbm_batch *b, uint32_t n, float *ms) {
    return bguard(b, [&] {
        bcheck(ms != nullptr, "null output");
        b->ensure_lists();
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, b->stream));
        b->step(n);
        HIPCHK(hipEventRecord(e1, b->stream));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(ms, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int lbm_batch_sync(lbm_batch *b) {
    return bguard(b, [&] { HIPCHK(hipStreamSynchronize(b->stream)); });
}

int lbm_batch_step_count(const lbm_batch *b, uint64_t *n) {
    if (!b || !n) return LBM_EINVAL;
    *n = b->steps;
    return LBM_OK;
}

int lbm_batch_info(lbm_batch *b, uint64_t *cells, uint32_t *interior_tiles, uint32_t *boundary_segments) {
    return bguard(b, [&] {
        b->ensure_lists();
        if (cells) *cells = b->cells;
        if (interior_tiles) *interior_tiles = b->n_interior;
        if (boundary_segments) *boundary_segments = b->n_boundary;
    });
}

int lbm_batch_set_graph(lbm_batch *b, int enable) {
    return bguard(b, [&] {
        b->use_graph = enable != 0;
        if (!b->use_graph) b->drop_graphs();
    });
}

int lbm_batch_set_variant(lbm_batch *b, int variant) {
    return bguard(b, [&] {
        bcheck(variant >= 0 && variant < LBM_BATCH_NUM_VARIANTS, "unknown variant");
        HIPCHK(hipStreamSynchronize(b->stream));
        if (lbm_batch::pair_variant(variant)) b->ensure_pair_buffers();  // may fail: then nothing changes
        else b->release_pair_buffers();
        b->variant = variant;
    });
}

int lbm_batch_get_variant(const lbm_batch *b, int *variant) {
    if (!b || !variant) return LBM_EINVAL;
    *variant = b->variant;
    return LBM_OK;
}

int lbm_batch_autotune(lbm_batch *b, uint32_t steps, int *best, float *ms) {
    return bguard(b, [&] {
        bcheck(steps >= 2, "steps must be >= 2");
        const size_t sb = b->f_elems * b->esize, lb = 3 * b->aux_elems * sizeof(float);
        char *snap = dalloc<char>(2 * sb + lb);
        const int cur = b->cur;
        const bool pform = b->pform;
        const uint64_t nsteps = b->steps;
        const int variant0 = b->variant;
        float t[LBM_BATCH_NUM_VARIANTS];
        int v_best = 0;
        try {
            HIPCHK(hipMemcpyAsync(snap, b->f[0], sb, hipMemcpyDeviceToDevice, b->stream));
            HIPCHK(hipMemcpyAsync(snap + sb, b->f[1], sb, hipMemcpyDeviceToDevice, b->stream));
            HIPCHK(hipMemcpyAsync(snap + 2 * sb, b->lag, lb, hipMemcpyDeviceToDevice, b->stream));
            hipEvent_t e0, e1;
            HIPCHK(hipEventCreate(&e0));
            HIPCHK(hipEventCreate(&e1));
            for (int v = 0; v < LBM_BATCH_NUM_VARIANTS; ++v) {
                if (lbm_batch::pair_variant(v)) {
                    if (steps < 3) {  // a call this short never launches two steps at once
                        t[v] = INFINITY;
                        continue;
                    }
                    try {
                        b->ensure_pair_buffers();
                    } catch (...) {  // no memory for two more sets: not a candidate
                        t[v] = INFINITY;
                        continue;
                    }
                }
                b->variant = v;
                b->step(2);  // builds lists and graphs, leaves P-form
                HIPCHK(hipEventRecord(e0, b->stream));
                b->step(steps);
                HIPCHK(hipEventRecord(e1, b->stream));
                HIPCHK(hipEventSynchronize(e1));
                HIPCHK(hipEventElapsedTime(&t[v], e0, e1));
                if (t[v] < t[v_best]) v_best = v;
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            HIPCHK(hipMemcpyAsync(b->f[0], snap, sb, hipMemcpyDeviceToDevice, b->stream));
            HIPCHK(hipMemcpyAsync(b->f[1], snap + sb, sb, hipMemcpyDeviceToDevice, b->stream));
            HIPCHK(hipMemcpyAsync(b->lag, snap + 2 * sb, lb, hipMemcpyDeviceToDevice, b->stream));
            HIPCHK(hipStreamSynchronize(b->stream));
        } catch (...) {
            (void)hipStreamSynchronize(b->stream);
            (void)hipFree(snap);
            b->cur = cur;
            b->pform = pform;
            b->steps = nsteps;
            b->variant = variant0;
            throw;
        }
        (void)hipFree(snap);
        b->cur = cur;
        b->pform = pform;
        b->steps = nsteps;
        b->variant = v_best;
        if (!lbm_batch::pair_variant(v_best)) b->release_pair_buffers();  // tried, not kept
        if (best) *best = v_best;
        if (ms)
            for (int v = 0; v < LBM_BATCH_NUM_VARIANTS; ++v) ms[v] = t[v] / steps;
    });
}

int lbm_batch_autotune_cached(lbm_batch *b, const char *cache_path, uint32_t steps, int *best, int *hit) {
    std::string key;
    tune_cache::Entry e;
    bool found = false;
    int rc = bguard(b, [&] {
        bcheck(cache_path && *cache_path, "null or empty cache path");
        bcheck(steps >= 2, "steps must be >= 2");
        uint64_t cells = 0;
        for (const DomainDev &d : b->table) cells += (uint64_t)d.g.X * d.g.Y;
        const uint32_t mcells = (uint32_t)std::min<uint64_t>((cells + (1u << 20) - 1) >> 20, 1u << 30);
        key = tune_cache::clean_key(device_tag(b->dev) + "/batch/" +
                                    std::to_string(tune_cache::pow2_ceil((uint32_t)b->table.size())) + "d/" +
                                    std::to_string(tune_cache::pow2_ceil(mcells)) + "Mcells/" +
                                    precision_tag(b->precision) +
                                    (b->boundary == LBM_BOUNDARY_REFERENCE ? "/ref" : "/zg") + "/v" +
                                    std::to_string(LBM_BATCH_NUM_VARIANTS));
        found = tune_cache::lookup(cache_path, key, e) && e.variant >= 0 && e.variant < LBM_BATCH_NUM_VARIANTS;
        if (found) b->variant = e.variant;
    });
    if (rc != LBM_OK) return rc;
    if (hit) *hit = found ? 1 : 0;
    if (found) {
        if (best) *best = e.variant;
        return LBM_OK;
    }
    e.ms.assign(LBM_BATCH_NUM_VARIANTS, 0.0f);
    e.steps = steps;
    rc = lbm_batch_autotune(b, steps, &e.variant, e.ms.data());
    if (rc != LBM_OK) return rc;
    if (best) *best = e.variant;
    return bguard(b, [&] { bcheck(tune_cache::store(cache_path, key, e), "cannot write the autotune cache file"); });
}

int lbm_batch_read_f(lbm_batch *b, uint32_t d, float *f9) {
    return bguard(b, [&] {
        bcheck(d < b->table.size() && f9, "bad domain or null output");
        b->ensure_lists();
        b->flush();
        const DomainDev &dd = b->table[d];
        const size_t dn = (size_t)dd.g.X * dd.g.Y;
        b->ensure_dense(dn);
        const dim3 grid((dd.g.X + 255) / 256, dd.g.Y);
        for (int k = 0; k < 9; ++k) {
            const char *plane = b->fptr(b->cur, dd) + (size_t)k * dd.g.ks * b->esize;
            by_store(b->precision, [&]<class St>() {
                hipLaunchKernelGGL(k_decode_plane<St>, grid, dim3(256), 0, b->stream, (const typename St::T *)plane,
                                   b->dense, dd.g, k);
            });
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(f9 + k * dn, b->dense, dn * sizeof(float), hipMemcpyDeviceToHost, b->stream));
        }
        HIPCHK(hipStreamSynchronize(b->stream));
    });
}

int lbm_batch_read_moments(lbm_batch *b, uint32_t d, float *jx, float *jy, float *rho) {
    return bguard(b, [&] {
        bcheck(d < b->table.size(), "bad domain");
        const DomainDev &dd = b->table[d];
        const float *base = b->lag + 3 * dd.aux_off;
        float *dst[3] = {jx, jy, rho};
        for (int c = 0; c < 3; ++c)
            if (dst[c])
                HIPCHK(hipMemcpy2DAsync(dst[c], sizeof(float) * dd.g.X, base + c * dd.g.pstride,
                                        sizeof(float) * dd.g.pitch, sizeof(float) * dd.g.X, dd.g.Y,
                                        hipMemcpyDeviceToHost, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    });
}

// ---- generators for batches (host only) ------------------------------------------

static uint32_t gen_below(uint64_t *s, uint32_t m) {
    return (uint32_t)(((lbm_splitmix64(s) >> 32) * (uint64_t)m) >> 32);
}

static double gen_unit(uint64_t *s) { return (double)(lbm_splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

int lbm_gen_batch(uint64_t seed, uint32_t n, lbm_domain *out) {
    if (!out) return LBM_EINVAL;
    uint64_t s = seed;
    for (uint32_t i = 0; i < n; ++i) {
        out[i].nx = 128 + 32 * gen_below(&s, 29);  // 128 .. 1024
        out[i].ny = 64 + 32 * gen_below(&s, 15);   // 64 .. 512
        const double tau = 0.52 + 0.28 * gen_unit(&s);
        out[i].omega = (float)(1.0 / tau);
        out[i].u_in = (float)(0.02 + 0.08 * gen_unit(&s));
    }
    return LBM_OK;
}

int lbm_gen_shapes(uint8_t *solid, uint32_t nx, uint32_t ny, uint64_t seed, uint32_t n_min, uint32_t n_max) {
    if (!solid || nx < 16 || ny < 16 || n_min > n_max) return LBM_EINVAL;
    uint64_t s = seed;
    const uint32_t count = n_min + gen_below(&s, n_max - n_min + 1);
    for (uint32_t c = 0; c < count; ++c) {
        if (gen_below(&s, 2) == 0) {  // circle
            const int64_t r = 3 + gen_below(&s, std::max<uint32_t>(1, std::min(nx, ny) / 8));
            const int64_t cxv = r + 1 + gen_below(&s, (uint32_t)std::max<int64_t>(1, (int64_t)nx - r - 1));
            const int64_t cyv = gen_below(&s, ny);
            for (int64_t y = std::max<int64_t>(0, cyv - r); y <= std::min<int64_t>(ny - 1, cyv + r); ++y)
                for (int64_t x = std::max<int64_t>(1, cxv - r); x <= std::min<int64_t>(nx - 1, cxv + r); ++x)
                    if ((x - cxv) * (x - cxv) + (y - cyv) * (y - cyv) <= r * r) solid[(size_t)y * nx + x] = 1;
        } else {  // axis-aligned rectangle
            const uint32_t w = 2 + gen_below(&s, std::max<uint32_t>(1, nx / 8));
            const uint32_t h = 2 + gen_below(&s, std::max<uint32_t>(1, ny / 4));
            const uint32_t x0 = 1 + gen_below(&s, std::max<uint32_t>(1, nx - w - 1));
            const uint32_t y0 = gen_below(&s, std::max<uint32_t>(1, ny - h));
            for (uint32_t y = y0; y < std::min(ny, y0 + h); ++y)
                for (uint32_t x = x0; x < std::min(nx, x0 + w); ++x) solid[(size_t)y * nx + x] = 1;
        }
    }
    return LBM_OK;
}

}  // extern "C"
