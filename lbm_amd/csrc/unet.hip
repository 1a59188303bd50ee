// Learned surrogate: a 4-level conv U-Net (3x3 convolutions, 32-64-128-256
// channels) with BF16 weights and activations and FP32 accumulation on the
// gfx950 matrix cores, its training (forward, backward, Adam) and its use as a
// 64-step predictor inside an LBM rollout.  No reference counterpart: the
// reference has no learned component.
//
// Convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16.  Activations are
// NHWC; (tap, input channel) is the GEMM's K.  Both operands of one product
// have the same lane layout -- lane l holds row l % 32 and k = 8 (l / 32) +
// 0..7, i.e. 16 contiguous bytes of one pixel's channels or of one row of
// W[tap][cout][cin] -- so which of them is "A" only decides the layout of the
// result.  The 1x1 kernel puts pixels in A: D (32 x 32 f32), lane l, value r
// holds pixel 8 (r / 4) + 4 (l / 32) + r % 4, cout l % 32.  The 3x3 kernels put
// the weights in A: lane l holds pixel l % 32 and couts 8 (r / 4) + 4 (l / 32)
// + r % 4, four runs of four consecutive channels that are stored as vectors.
// The input gradient is the same forward convolution applied to the output
// gradient with the weights flipped in space and transposed in channels; the
// weight gradient reduces over pixels (K) with both operands gathered per lane.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "lbm_common.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

// host-side input channels per pixel (solid fraction, (rho - 1) / s_rho, ux / s_u,
// uy / s_u, (tau - 1/2) / 0.2, u_in / s_u, two spare zeros) and how many carry data
constexpr int kNetIn = 8, kNetInLive = 6;

__device__ __forceinline__ uint16_t f2bf(float v) {  // round to nearest even
    return __builtin_bit_cast(uint16_t, (__bf16)v);
}
__device__ __forceinline__ float bf2f(uint16_t v) { return (float)__builtin_bit_cast(__bf16, v); }

// ---- convolution forward (and input gradient) ------------------------------------
// Input of a convolution, NHWC bf16: channels [0, ca) come from a, channels
// [ca, ca + cb) from b (cb may be 0).  With a_up, a has half the resolution and
// is read through nearest 2x upsampling: the U-Net's "upsample, concatenate"
// happens inside the convolution's loads and is never materialised.
struct ConvIn {
    const uint16_t *a = nullptr, *b = nullptr;
    int ca = 0, cb = 0, a_up = 0;
};

typedef uint16_t u16x16 __attribute__((ext_vector_type(16)));  // 32 B: one lane's operands for two k-steps

// 1x1 convolution: out[p][co] = act(bias[co] + sum_ci in[p][ci] w[co][ci]),
// w [Cout][Cin] bf16, one plain input (ConvIn.a).  GEMM, M = pixels, N = Cout,
// K = Cin.  A wave computes MI runs of 32 pixels x NI tiles of 32 couts.  K
// advances in chunks of 32 channels; lane half h loads channels 16h .. 16h + 15
// of the chunk as one 32-byte load, the first k-step consumes their first
// eight and the second k-step their last eight (the same permutation of K for
// A and B).  The next chunk's operands are loaded while the current one is
// multiplied.  Requirements: P % (32 * MI) == 0, Cin % 32 == 0,
// Cout % (32 * NI) == 0.  Only channels co < co_store are stored.
template <int MI, int NI, bool RELU, bool OUT_F32>
__global__ __launch_bounds__(256) void k_conv1x1(const uint16_t *__restrict__ in, const uint16_t *__restrict__ w,
                                                 const float *__restrict__ bias, void *__restrict__ out, long long P,
                                                 int Cin, int Cout, int co_store) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int half = lane >> 5, l32 = lane & 31;
    const long long p0 = ((long long)blockIdx.x * 4 + wave) * (32 * MI);
    if (p0 >= P) return;
    const int co0 = blockIdx.y * 32 * NI;
    const int nk = Cin >> 5;
    auto load = [&](int kk, u16x16(&a)[MI], u16x16(&b)[NI]) {
        const int c0 = (kk << 5) + 16 * half;
#pragma unroll
        for (int i = 0; i < MI; ++i) a[i] = *(const u16x16 *)(in + (p0 + 32 * i + l32) * Cin + c0);
#pragma unroll
        for (int j = 0; j < NI; ++j) b[j] = *(const u16x16 *)(w + (size_t)(co0 + 32 * j + l32) * Cin + c0);
    };
    f32x16 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = f32x16{};
    auto mma = [&](const u16x16(&a)[MI], const u16x16(&b)[NI]) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
#pragma unroll
            for (int i = 0; i < MI; ++i) {
                const u16x8 a8 = s == 0 ? __builtin_shufflevector(a[i], a[i], 0, 1, 2, 3, 4, 5, 6, 7)
                                        : __builtin_shufflevector(a[i], a[i], 8, 9, 10, 11, 12, 13, 14, 15);
#pragma unroll
                for (int j = 0; j < NI; ++j) {
                    const u16x8 b8 = s == 0 ? __builtin_shufflevector(b[j], b[j], 0, 1, 2, 3, 4, 5, 6, 7)
                                            : __builtin_shufflevector(b[j], b[j], 8, 9, 10, 11, 12, 13, 14, 15);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a8),
                                                                        __builtin_bit_cast(bf16x8, b8), acc[i][j],
                                                                        0, 0, 0);
                }
            }
        }
    };
    u16x16 A0[MI], B0[NI], A1[MI], B1[NI];
    load(0, A0, B0);
    for (int kk = 0; kk < nk; kk += 2) {
        if (kk + 1 < nk) load(kk + 1, A1, B1);
        mma(A0, B0);
        if (kk + 1 >= nk) break;
        if (kk + 2 < nk) load(kk + 2, A0, B0);
        mma(A1, B1);
    }
#pragma unroll
    for (int i = 0; i < MI; ++i) {
#pragma unroll
        for (int j = 0; j < NI; ++j) {
            const int co = co0 + 32 * j + l32;
            const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const long long p = p0 + 32 * i + 8 * (r / 4) + 4 * half + (r % 4);
                float v = acc[i][j][r] + bv;
                if (RELU) v = fmaxf(v, 0.0f);
                if (co < co_store) {
                    if (OUT_F32) ((float *)out)[p * Cout + co] = v;
                    else ((uint16_t *)out)[p * Cout + co] = f2bf(v);
                }
            }
        }
    }
}

// Epilogue of the 3x3 kernels.  The products are issued with the weights as
// the A operand and the pixels as B, so a wave's 32 x 32 result is
// [cout][pixel]: lane l holds pixel l % 32 and output channels
// 8 (r / 4) + 4 (l / 32) + r % 4, r = 0..15 -- four runs of four consecutive
// channels, stored as four 8-byte (bf16) or 16-byte (f32) pieces.  A wave
// covers MI rows (from y_first) x 32 columns (from x0) x NI tiles of 32
// output channels (from co_first).  With a K split (gridDim.z > 1) the raw
// sums go to part[z][p][co]; otherwise bias, ReLU and the store of channels
// co < co_store.
template <int MI, int NI>
__device__ __forceinline__ void conv_store_tile(const f32x16 (&acc)[MI][NI], const float *__restrict__ bias,
                                                void *__restrict__ out, float *__restrict__ part, int N, int H, int W,
                                                int Cout, int co_store, int relu, int out_f32, long long n, int x0,
                                                int y_first, int co_first, int half, int l32) {
    const long long P = (long long)N * H * W;
    const bool split = gridDim.z > 1;
#pragma unroll
    for (int i = 0; i < MI; ++i) {
        const int y = y_first + i;
        if (y >= H) continue;
        const long long p = (n * H + y) * W + x0 + l32;
#pragma unroll
        for (int j = 0; j < NI; ++j) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int co = co_first + 32 * j + 8 * g + 4 * half;
                float v[4] = {acc[i][j][4 * g], acc[i][j][4 * g + 1], acc[i][j][4 * g + 2], acc[i][j][4 * g + 3]};
                if (split) {
                    *(float4 *)&part[((long long)blockIdx.z * P + p) * Cout + co] = make_float4(v[0], v[1], v[2], v[3]);
                    continue;
                }
                if (bias) {
                    const float4 bv = *(const float4 *)&bias[co];
                    v[0] += bv.x;
                    v[1] += bv.y;
                    v[2] += bv.z;
                    v[3] += bv.w;
                }
                if (relu)
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[e] = fmaxf(v[e], 0.0f);
                if (co + 4 <= co_store) {
                    if (out_f32) {
                        *(float4 *)&((float *)out)[p * Cout + co] = make_float4(v[0], v[1], v[2], v[3]);
                    } else {
                        const uint32_t lo = f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                                       hi = f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                        *(uint2 *)&((uint16_t *)out)[p * Cout + co] = make_uint2(lo, hi);
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (co + e < co_store) {
                            if (out_f32) ((float *)out)[p * Cout + co + e] = v[e];
                            else ((uint16_t *)out)[p * Cout + co + e] = f2bf(v[e]);
                        }
                }
            }
        }
    }
}

// 2 x 2 average pool of a block's bf16 outputs into pout (H / 2 x W / 2 x
// Cout), with k_pool2's arithmetic on the stored values:
// bf16(0.25 ((v(y, x) + v(y, x + 1)) + (v(y + 1, x) + v(y + 1, x + 1)))).
// Horizontal neighbours are adjacent lanes; a wave holding two rows (MI == 2)
// has both vertical neighbours, otherwise (one row per wave, four waves down
// the block) odd waves pass their row to the wave above through `scratch`
// (4 KB per 32-channel tile; free after the main loop's last barrier).
// Every thread of the block must call it.
template <int MI, int NI, int WN>
__device__ __forceinline__ void conv_pool_tile(const f32x16 (&acc)[MI][NI], const float *__restrict__ bias,
                                               uint16_t *__restrict__ pout, uint16_t *scratch, int H, int W, int Cout,
                                               int relu, long long n, int x0, int y_first, int wm, int co_first,
                                               int lane) {
    static_assert(MI == 2 || (MI == 1 && WN == 1), "pooling needs row pairs inside a wave or a block column");
    const int half = lane >> 5, l32 = lane & 31;
    const int Ho = H >> 1, Wo = W >> 1;
    // the stored (bf16-rounded) output of row i, tile j, channels 8 g + 4 half + e
    auto vals = [&](int i, int j, int g, float (&v)[4]) {
        const int co = co_first + 32 * j + 8 * g + 4 * half;
        const float4 bv = bias ? *(const float4 *)&bias[co] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        v[0] = acc[i][j][4 * g] + bv.x;
        v[1] = acc[i][j][4 * g + 1] + bv.y;
        v[2] = acc[i][j][4 * g + 2] + bv.z;
        v[3] = acc[i][j][4 * g + 3] + bv.w;
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = bf2f(f2bf(relu ? fmaxf(v[e], 0.0f) : v[e]));
    };
    auto emit = [&](int y, int j, int g, const float (&top)[4], const float (&bot)[4]) {
        float o[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float t1 = __shfl_xor(top[e], 1), b1 = __shfl_xor(bot[e], 1);
            o[e] = 0.25f * ((top[e] + t1) + (bot[e] + b1));
        }
        if ((l32 & 1) == 0 && y < H) {
            const int co = co_first + 32 * j + 8 * g + 4 * half;
            const long long q = (n * Ho + (y >> 1)) * Wo + ((x0 + l32) >> 1);
            const uint32_t lo = f2bf(o[0]) | ((uint32_t)f2bf(o[1]) << 16), hi = f2bf(o[2]) | ((uint32_t)f2bf(o[3]) << 16);
            *(uint2 *)&pout[q * Cout + co] = make_uint2(lo, hi);
        }
    };
    if constexpr (MI == 2) {
#pragma unroll
        for (int j = 0; j < NI; ++j)
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                float top[4], bot[4];
                vals(0, j, g, top);
                vals(1, j, g, bot);
                emit(y_first, j, g, top, bot);
            }
    } else {
        u16x8 *slot = (u16x8 *)scratch + (size_t)(wm >> 1) * NI * 128;
        if (wm & 1) {
#pragma unroll
            for (int j = 0; j < NI; ++j)
#pragma unroll
                for (int h2 = 0; h2 < 2; ++h2) {
                    float a[4], b[4];
                    vals(0, j, 2 * h2, a);
                    vals(0, j, 2 * h2 + 1, b);
                    u16x8 pk;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        pk[e] = f2bf(a[e]);
                        pk[4 + e] = f2bf(b[e]);
                    }
                    slot[(j * 2 + h2) * 64 + lane] = pk;
                }
        }
        __syncthreads();
        if (!(wm & 1)) {
#pragma unroll
            for (int j = 0; j < NI; ++j)
#pragma unroll
                for (int h2 = 0; h2 < 2; ++h2) {
                    const u16x8 pk = slot[(j * 2 + h2) * 64 + lane];
#pragma unroll
                    for (int gg = 0; gg < 2; ++gg) {
                        float top[4], bot[4];
                        vals(0, j, 2 * h2 + gg, top);
#pragma unroll
                        for (int e = 0; e < 4; ++e) bot[e] = bf2f(pk[4 * gg + e]);
                        emit(y_first, j, 2 * h2 + gg, top, bot);
                    }
                }
        }
    }
}

// A 1x1 convolution applied to a 3x3 convolution's output inside its
// epilogue (the network's head): out[p * cout + o] = b[o] + sum_c w[o][c] y[p][c]
// in f32 for o < store rounded up to a multiple of 4 (outputs are stored four
// at a time), y the stored (bf16) output of a 32-channel layer and w
// [cout = 32][32] bf16.  Null w: none.
struct ConvHead {
    const uint16_t *w = nullptr;
    const float *b = nullptr;
    float *out = nullptr;
    int cout = 0, store = 0;
};

// One more product per row of the wave, on values the wave already holds: the
// B operand is the lane's own 16 channels of its pixel (channels 8 g + 4 half
// + e, two groups per k-step), the A operand the head's weights in the same
// channel order, so the K sum runs over all 32 channels; the result is
// [head output][pixel] like the 3x3 products'.
template <int MI>
__device__ __forceinline__ void conv_head_tile(const f32x16 (&acc)[MI][1], const float *__restrict__ bias, int relu,
                                               const ConvHead &hd, int H, int W, long long n, int x0, int y_first,
                                               int lane) {
    const int half = lane >> 5, l32 = lane & 31;
#pragma unroll
    for (int i = 0; i < MI; ++i) {
        const int y = y_first + i;
        f32x16 d = f32x16{};
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2) {
            u16x8 bop, aop;
#pragma unroll
            for (int gg = 0; gg < 2; ++gg) {
                const int g = 2 * s2 + gg, c = 8 * g + 4 * half;
                const float4 bv = bias ? *(const float4 *)&bias[c] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                float v[4] = {acc[i][0][4 * g] + bv.x, acc[i][0][4 * g + 1] + bv.y, acc[i][0][4 * g + 2] + bv.z,
                              acc[i][0][4 * g + 3] + bv.w};
                const uint2 wv = *(const uint2 *)&hd.w[l32 * 32 + c];
#pragma unroll
                for (int e = 0; e < 4; ++e) bop[4 * gg + e] = f2bf(relu ? fmaxf(v[e], 0.0f) : v[e]);
                aop[4 * gg] = (uint16_t)wv.x;
                aop[4 * gg + 1] = (uint16_t)(wv.x >> 16);
                aop[4 * gg + 2] = (uint16_t)wv.y;
                aop[4 * gg + 3] = (uint16_t)(wv.y >> 16);
            }
            d = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, aop), __builtin_bit_cast(bf16x8, bop),
                                                        d, 0, 0, 0);
        }
        if (y < H) {
            const long long p = (n * H + y) * W + x0 + l32;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int o = 8 * g + 4 * half;  // four consecutive outputs, stored as one vector
                if (o < hd.store) {
                    const float4 hb = *(const float4 *)&hd.b[o];
                    *(float4 *)&hd.out[p * hd.cout + o] =
                        make_float4(d[4 * g] + hb.x, d[4 * g + 1] + hb.y, d[4 * g + 2] + hb.z, d[4 * g + 3] + hb.w);
                }
            }
        }
    }
}

// LDS-staged 3x3 convolution.  A block computes TH = (4 / WN) * MI rows x 32
// columns of output pixels of one image for 32 * WN * NI output channels; its
// four waves form a (4 / WN) x WN grid, each computing MI rows x NI tiles of 32
// channels.  For every 32-channel input chunk the block stages the chunk's
// (TH + 2) x 34 pixel halo in LDS once (80-byte pixel stride: the 16-byte reads
// of 16 consecutive pixels hit distinct banks) and runs all nine taps from it,
// so each input value is fetched from L2 about (TH + 2) * 34 / (32 TH) times
// instead of 9 * Cout / 32 times.  The next chunk's halo is loaded into registers
// while the current one is multiplied (double-buffered LDS); weights are read
// from global memory one tap ahead.  gridDim.z > 1 splits the input chunks
// over blocks, which then write raw f32 sums to part[z][p][co] for
// k_conv_finish.  Requirements: W % 32 == 0, ca % 32 == cb % 32 == 0,
// Cout % (32 * WN * NI) == 0; H need not be a multiple of TH.
template <int MI, int NI, int WN>
__global__ __launch_bounds__(256) void k_conv3_lds(ConvIn src, const uint16_t *__restrict__ w,
                                                   const float *__restrict__ bias, void *__restrict__ out,
                                                   float *__restrict__ part, uint16_t *__restrict__ pout, ConvHead hd, int N, int H,
                                                   int W, int Cout,
                                                   int co_store, int relu, int out_f32) {
    constexpr int WM = 4 / WN, TH = WM * MI, HR = TH + 2, HC = 34, PS = 40;
    constexpr int NPIECE = HR * HC * 4;  // 16-byte pieces of one chunk's halo
    constexpr int PER_T = (NPIECE + 255) / 256;
    __shared__ __align__(16) uint16_t lds[2][HR * HC * PS];
    const int Cin = src.ca + src.cb;
    const int tiles_x = W >> 5, tiles_y = (H + TH - 1) / TH;
    int bid = blockIdx.x;
    const int tx = bid % tiles_x;
    bid /= tiles_x;
    const int ty = bid % tiles_y;
    const long long n = bid / tiles_y;
    const int x0 = tx * 32, y0 = ty * TH;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int half = lane >> 5, l32 = lane & 31;
    const int co0 = (blockIdx.y * WN + wn) * 32 * NI;
    const int up = src.a_up ? 1 : 0, Ha = H >> up, Wa = W >> up;
    const int nck = Cin >> 5;
    const int ck0 = (int)(blockIdx.z * nck / gridDim.z), ck1 = (int)((blockIdx.z + 1) * nck / gridDim.z);

    u16x8 stage[PER_T];
    auto gload = [&](int ck) {
        const int c0 = ck << 5;
#pragma unroll
        for (int t = 0; t < PER_T; ++t) {
            const int q = threadIdx.x + 256 * t;
            stage[t] = u16x8{};
            if (q < NPIECE) {
                const int pi = q >> 2, c = c0 + ((q & 3) << 3);
                const int y = y0 - 1 + pi / HC, x = x0 - 1 + pi % HC;
                if (y >= 0 && y < H && x >= 0 && x < W) {
                    const uint16_t *ptr = c < src.ca
                        ? src.a + ((n * Ha + (y >> up)) * Wa + (x >> up)) * src.ca + c
                        : src.b + ((n * H + y) * W + x) * src.cb + (c - src.ca);
                    stage[t] = *(const u16x8 *)ptr;
                }
            }
        }
    };
    auto swrite = [&](int buf) {
#pragma unroll
        for (int t = 0; t < PER_T; ++t) {
            const int q = threadIdx.x + 256 * t;
            if (q < NPIECE) *(u16x8 *)&lds[buf][(q >> 2) * PS + ((q & 3) << 3)] = stage[t];
        }
    };
    f32x16 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = f32x16{};
    // weights of one tap: lane's 8 channels 16 s + 8 half of the chunk, couts co0 + 32 j + l32
    auto bload = [&](int tap, int c0, u16x8(&b)[NI][2]) {
#pragma unroll
        for (int j = 0; j < NI; ++j)
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2)
                b[j][s2] = *(const u16x8 *)(w + ((size_t)tap * Cout + co0 + 32 * j + l32) * Cin + c0 + 16 * s2 +
                                            8 * half);
    };
    auto compute = [&](int buf, int ck) {
        const int c0 = ck << 5;
        u16x8 bc[NI][2], bn[NI][2];
        bload(0, c0, bc);
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            if (tap < 8) bload(tap + 1, c0, bn);
            const int dy = tap / 3, dx = tap % 3;
#pragma unroll
            for (int s2 = 0; s2 < 2; ++s2) {
#pragma unroll
                for (int i = 0; i < MI; ++i) {
                    const int r = wm * MI + i + dy;
                    const u16x8 a = *(const u16x8 *)&lds[buf][(r * HC + l32 + dx) * PS + 16 * s2 + 8 * half];
#pragma unroll
                    for (int j = 0; j < NI; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, bc[j][s2]), __builtin_bit_cast(bf16x8, a), acc[i][j], 0, 0, 0);
                }
            }
            if (tap < 8) {
#pragma unroll
                for (int j = 0; j < NI; ++j) {
                    bc[j][0] = bn[j][0];
                    bc[j][1] = bn[j][1];
                }
            }
        }
    };
    gload(ck0);
    swrite(0);
    __syncthreads();
    for (int ck = ck0; ck < ck1; ++ck) {
        const int buf = (ck - ck0) & 1;
        if (ck + 1 < ck1) gload(ck + 1);
        compute(buf, ck);
        if (ck + 1 < ck1) swrite(buf ^ 1);
        __syncthreads();
    }
    conv_store_tile<MI, NI>(acc, bias, out, part, N, H, W, Cout, co_store, relu, out_f32, n, x0, y0 + wm * MI, co0,
                            half, l32);
    if constexpr (MI == 2 || (MI == 1 && WN == 1))
        if (pout) conv_pool_tile<MI, NI, WN>(acc, bias, pout, &lds[0][0], H, W, Cout, relu, n, x0, y0 + wm * MI, wm, co0, lane);
    if constexpr (NI == 1 && WN == 1)
        if (hd.w) conv_head_tile<MI>(acc, bias, relu, hd, H, W, n, x0, y0 + wm * MI, lane);
}

// The same convolution with the weights staged in LDS too.  Per chunk the
// block loads the B operands of all nine taps for its 32 * WN * NI output
// channels once, with 16-byte loads along input channels (64 contiguous bytes
// per (tap, cout)), into an 80-byte-stride table beside the halo; the waves
// read them from there.  This replaces every wave's own per-tap loads of 32
// scattered 32-byte segments, which (not the products, nor the halo) bound the
// kernel above once a layer has several input chunks.  Both tables are double
// buffered and filled from registers loaded during the previous chunk's
// products.  Block shapes with 32 * WN * NI <= 64 output channels only (LDS).
// KC input channels per chunk: 32, or 64 (fewer, longer chunks: twice the
// products per wait for the next chunk's loads; 32-channel blocks of four rows
// only, 141 KB of LDS).
template <int MI, int NI, int WN, int KC = 32>
__global__ __launch_bounds__(256) void k_conv3_lw(ConvIn src, const uint16_t *__restrict__ w,
                                                  const float *__restrict__ bias, void *__restrict__ out,
                                                  float *__restrict__ part, uint16_t *__restrict__ pout, int N, int H,
                                                   int W, int Cout,
                                                  int co_store, int relu, int out_f32) {
    constexpr int WM = 4 / WN, TH = WM * MI, HR = TH + 2, HC = 34;
    constexpr int PS = KC + 8;           // element stride of a pixel / a weight row: 16 bytes of padding
    constexpr int PPX = KC / 8;          // 16-byte pieces per pixel / weight row
    constexpr int CB = 32 * NI * WN;     // output channels of the block
    static_assert(KC == 32 ? CB <= 64 : (KC == 64 && CB == 32 && TH <= 4), "tables too large for LDS");
    constexpr int NPIECE = HR * HC * PPX;  // 16-byte pieces of one chunk's halo
    constexpr int PER_T = (NPIECE + 255) / 256;
    constexpr int WPIECE = 9 * CB * PPX;   // 16-byte pieces of one chunk's weights
    constexpr int WPER_T = (WPIECE + 255) / 256;
    constexpr int HALO = HR * HC * PS, WTAB = 9 * CB * PS;
    __shared__ __align__(16) uint16_t lds[2][HALO + WTAB];
    const int Cin = src.ca + src.cb;
    const int tiles_x = W >> 5, tiles_y = (H + TH - 1) / TH;
    int bid = blockIdx.x;
    const int tx = bid % tiles_x;
    bid /= tiles_x;
    const int ty = bid % tiles_y;
    const long long n = bid / tiles_y;
    const int x0 = tx * 32, y0 = ty * TH;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int half = lane >> 5, l32 = lane & 31;
    const int cb0 = blockIdx.y * CB;
    const int up = src.a_up ? 1 : 0, Ha = H >> up, Wa = W >> up;
    const int nck = Cin / KC;
    const int ck0 = (int)(blockIdx.z * nck / gridDim.z), ck1 = (int)((blockIdx.z + 1) * nck / gridDim.z);

    u16x8 stage[PER_T], wstage[WPER_T];
    auto gload = [&](int ck) {
        const int c0 = ck * KC;
#pragma unroll
        for (int t = 0; t < PER_T; ++t) {
            const int q = threadIdx.x + 256 * t;
            stage[t] = u16x8{};
            if (q < NPIECE) {
                const int pi = q / PPX, c = c0 + (q % PPX) * 8;  // a piece lies wholly in a or in b
                const int y = y0 - 1 + pi / HC, x = x0 - 1 + pi % HC;
                if (y >= 0 && y < H && x >= 0 && x < W) {
                    const uint16_t *ptr = c < src.ca
                        ? src.a + ((n * Ha + (y >> up)) * Wa + (x >> up)) * src.ca + c
                        : src.b + ((n * H + y) * W + x) * src.cb + (c - src.ca);
                    stage[t] = *(const u16x8 *)ptr;
                }
            }
        }
        // weights: piece q = (tap * CB + cout) * PPX + 16-byte piece of the chunk
#pragma unroll
        for (int t = 0; t < WPER_T; ++t) {
            const int q = threadIdx.x + 256 * t;
            if (q < WPIECE) {
                const int row = q / PPX, tap = row / CB, cl = row % CB;
                wstage[t] = *(const u16x8 *)(w + ((size_t)tap * Cout + cb0 + cl) * Cin + c0 + (q % PPX) * 8);
            }
        }
    };
    auto swrite = [&](int buf) {
#pragma unroll
        for (int t = 0; t < PER_T; ++t) {
            const int q = threadIdx.x + 256 * t;
            if (q < NPIECE) *(u16x8 *)&lds[buf][(q / PPX) * PS + (q % PPX) * 8] = stage[t];
        }
#pragma unroll
        for (int t = 0; t < WPER_T; ++t) {
            const int q = threadIdx.x + 256 * t;
            if (q < WPIECE) *(u16x8 *)&lds[buf][HALO + (q / PPX) * PS + (q % PPX) * 8] = wstage[t];
        }
    };
    f32x16 acc[MI][NI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NI; ++j) acc[i][j] = f32x16{};
    auto compute = [&](int buf) {
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;
#pragma unroll
            for (int s2 = 0; s2 < KC / 16; ++s2) {
                u16x8 b[NI];
#pragma unroll
                for (int j = 0; j < NI; ++j)
                    b[j] = *(const u16x8 *)&lds[buf][HALO + (tap * CB + (wn * NI + j) * 32 + l32) * PS + 16 * s2 +
                                                     8 * half];
#pragma unroll
                for (int i = 0; i < MI; ++i) {
                    const int r = wm * MI + i + dy;
                    const u16x8 a = *(const u16x8 *)&lds[buf][(r * HC + l32 + dx) * PS + 16 * s2 + 8 * half];
#pragma unroll
                    for (int j = 0; j < NI; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, b[j]), __builtin_bit_cast(bf16x8, a), acc[i][j], 0, 0, 0);
                }
            }
        }
    };
    gload(ck0);
    swrite(0);
    __syncthreads();
    for (int ck = ck0; ck < ck1; ++ck) {
        const int buf = (ck - ck0) & 1;
        if (ck + 1 < ck1) gload(ck + 1);
        compute(buf);
        if (ck + 1 < ck1) swrite(buf ^ 1);
        __syncthreads();
    }
    conv_store_tile<MI, NI>(acc, bias, out, part, N, H, W, Cout, co_store, relu, out_f32, n, x0, y0 + wm * MI,
                            cb0 + wn * NI * 32, half, l32);
    if constexpr (MI == 2 || (MI == 1 && WN == 1))
        if (pout)
            conv_pool_tile<MI, NI, WN>(acc, bias, pout, &lds[0][0], H, W, Cout, relu, n, x0, y0 + wm * MI, wm,
                                       cb0 + wn * NI * 32, lane);
}

// out = act(bias + sum of the split partials), four channels per thread
template <bool RELU, bool OUT_F32>
__global__ void k_conv_finish(const float *__restrict__ part, const float *__restrict__ bias, void *__restrict__ out,
                              long long n, int Cout, int splits) {
    const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i >= n) return;
    float4 s = *(const float4 *)(part + i);
    for (int z = 1; z < splits; ++z) {
        const float4 t = *(const float4 *)(part + (long long)z * n + i);
        s.x += t.x;
        s.y += t.y;
        s.z += t.z;
        s.w += t.w;
    }
    const int c = (int)(i % Cout);
    float v[4] = {s.x, s.y, s.z, s.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] += bias ? bias[c + k] : 0.0f;
        if (RELU) v[k] = fmaxf(v[k], 0.0f);
    }
    if (OUT_F32) {
        *(float4 *)((float *)out + i) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
        const uint32_t lo = f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16), hi = f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
        *(uint2 *)((uint16_t *)out + i) = make_uint2(lo, hi);
    }
}

// ---- weight gradient -------------------------------------------------------------------
// dw[tap][co][ci] += sum_p in[p + off(tap)][ci] * g[p][co]: a GEMM with M = co,
// N = ci, K = pixels.  A block takes one tap, a 32 x 32 (co, ci) tile and a
// range of pixels split over its four waves; the four partial tiles are summed
// in LDS and added to dw (f32) with vector atomics.
template <int KS>
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(ConvIn src, const uint16_t *__restrict__ g,
                                                       float *__restrict__ dw, int N, int H, int W, int Cout,
                                                       int pix_per_block) {
    constexpr int R = KS / 2;
    const int Cin = src.ca + src.cb;
    const int up = src.a_up ? 1 : 0;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long P = (long long)N * H * W;
    const int tiles_ci = Cin / 32;
    const int co0 = (blockIdx.y / tiles_ci) * 32, ci0 = (blockIdx.y % tiles_ci) * 32;
    const int tap = blockIdx.z;
    const int dy = tap / KS - R, dx = tap % KS - R;
    const long long pb = (long long)blockIdx.x * pix_per_block;
    const long long pe = min(P, pb + pix_per_block);
    const int kofs = 8 * (lane >> 5);
    f32x16 acc = f32x16{};
    for (long long k0 = pb + 16 * wave; k0 < pe; k0 += 64) {
        uint16_t av[8], bv[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const long long p = k0 + kofs + j;
            uint16_t gv = 0, iv = 0;
            if (p < pe) {
                gv = g[p * Cout + co0 + (lane & 31)];
                const int x = (int)(p % W), y = (int)((p / W) % H);
                const long long n = p / ((long long)W * H);
                const int sy = y + dy, sx = x + dx;
                if (sy >= 0 && sy < H && sx >= 0 && sx < W) {
                    const int c = ci0 + (lane & 31);
                    iv = c < src.ca ? src.a[((n * (H >> up) + (sy >> up)) * (W >> up) + (sx >> up)) * src.ca + c]
                                    : src.b[((n * H + sy) * W + sx) * src.cb + (c - src.ca)];
                }
            }
            av[j] = gv;
            bv[j] = iv;
        }
        u16x8 a8, b8;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            a8[j] = av[j];
            b8[j] = bv[j];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a8), __builtin_bit_cast(bf16x8, b8),
                                                      acc, 0, 0, 0);
    }
    __shared__ float red[4][32 * 32];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);  // co
        red[wave][row * 32 + (lane & 31)] = acc[r];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 32 * 32; i += 256) {
        const float s = (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]);
        const int co = co0 + i / 32, ci = ci0 + i % 32;
        atomicAdd(&dw[((size_t)tap * Cout + co) * Cin + ci], s);
    }
}

// ---- elementwise helpers ----------------------------------------------------------------
__global__ void k_f32_to_bf16(const float *__restrict__ a, uint16_t *__restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

__global__ void k_bf16_to_f32(const uint16_t *__restrict__ a, float *__restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = bf2f(a[i]);
}

// wt[tap'][ci][co] = w[taps - 1 - tap'][co][ci]: weights of the input-gradient convolution
__global__ void k_flip_transpose(const uint16_t *__restrict__ w, uint16_t *__restrict__ wt, int Cin, int Cout,
                                 int taps) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)taps * Cin * Cout;
    if (i >= n) return;
    const int co = (int)(i % Cout);
    const int ci = (int)((i / Cout) % Cin);
    const int tp = (int)(i / ((size_t)Cout * Cin));
    wt[i] = w[((size_t)(taps - 1 - tp) * Cout + co) * Cin + ci];
}

// gradient through ReLU: g *= (y > 0)
__global__ void k_relu_bwd(uint16_t *__restrict__ g, const uint16_t *__restrict__ y, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !(bf2f(y[i]) > 0.0f)) g[i] = 0;
}

// db[c] += sum_p g[p][c]
__global__ void k_bias_grad(const uint16_t *__restrict__ g, float *__restrict__ db, long long P, int C) {
    const int c = blockIdx.y * 64 + (threadIdx.x & 63);
    if (c >= C) return;
    float s = 0.0f;
    for (long long p = (long long)blockIdx.x * 4 + (threadIdx.x >> 6); p < P; p += (long long)gridDim.x * 4)
        s += bf2f(g[p * C + c]);
    atomicAdd(&db[c], s);
}

// 2x2 average pool, NHWC, eight channels per thread
__global__ void k_pool2(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, int N, int H, int W, int C) {
    const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    const int Ho = H / 2, Wo = W / 2;
    if (i >= (size_t)N * Ho * Wo * C) return;
    const int c = (int)(i % C);
    const size_t q = i / C;
    const int x = (int)(q % Wo), y = (int)((q / Wo) % Ho);
    const size_t n = q / ((size_t)Wo * Ho);
    const size_t b = ((n * H + 2 * y) * W + 2 * x) * C + c;
    const size_t r = (size_t)W * C;
    const u16x8 v0 = *(const u16x8 *)(in + b), v1 = *(const u16x8 *)(in + b + C);
    const u16x8 v2 = *(const u16x8 *)(in + b + r), v3 = *(const u16x8 *)(in + b + r + C);
    u16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = f2bf(0.25f * ((bf2f(v0[k]) + bf2f(v1[k])) + (bf2f(v2[k]) + bf2f(v3[k]))));
    *(u16x8 *)(out + i) = o;
}

// Gradient of a convolution input built as [up2(a) | b] (see ConvIn):
// ga[q] = sum of gk's first ca channels over the 2x2 fine pixels of coarse pixel
// q; gb[p] = gk's last cb channels.  H, W: the fine size; eight channels per thread.
__global__ void k_split_up(const uint16_t *__restrict__ gk, uint16_t *__restrict__ ga, uint16_t *__restrict__ gb,
                           int N, int H, int W, int ca, int cb) {
    const size_t t = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    const size_t na = (size_t)N * (H / 2) * (W / 2) * ca, nb = (size_t)N * H * W * cb;
    const int C = ca + cb;
    if (t < na) {
        const int c = (int)(t % ca);
        const size_t q = t / ca;
        const int x = (int)(q % (W / 2)), y = (int)((q / (W / 2)) % (H / 2));
        const size_t n = q / ((size_t)(W / 2) * (H / 2));
        const size_t b = ((n * H + 2 * y) * W + 2 * x) * C + c;
        const size_t r = (size_t)W * C;
        const u16x8 v0 = *(const u16x8 *)(gk + b), v1 = *(const u16x8 *)(gk + b + C);
        const u16x8 v2 = *(const u16x8 *)(gk + b + r), v3 = *(const u16x8 *)(gk + b + r + C);
        u16x8 o;
#pragma unroll
        for (int k = 0; k < 8; ++k) o[k] = f2bf((bf2f(v0[k]) + bf2f(v1[k])) + (bf2f(v2[k]) + bf2f(v3[k])));
        *(u16x8 *)(ga + t) = o;
    } else if (t - na < nb) {
        const size_t j = t - na;
        const size_t p = j / cb;
        *(u16x8 *)(gb + j) = *(const u16x8 *)(gk + p * C + ca + (j % cb));
    }
}

// gin (+)= the pool gradient; H, W: the input (fine) size
__global__ void k_pool2_bwd(const uint16_t *__restrict__ gout, uint16_t *__restrict__ gin, int N, int H, int W, int C,
                            int accumulate) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * H * W * C) return;
    const int c = (int)(i % C);
    const size_t q = i / C;
    const int x = (int)(q % W), y = (int)((q / W) % H);
    const size_t n = q / ((size_t)W * H);
    float v = 0.25f * bf2f(gout[((n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c]);
    if (accumulate) v += bf2f(gin[i]);
    gin[i] = f2bf(v);
}

// masked mean-square loss on the first three output channels and its gradient:
// loss = sum_p wm[p] sum_{c<3} (out[p][c] - tgt[p][c])^2 / sum_p wm[p]
__global__ void k_mse(const float *__restrict__ out, const float *__restrict__ tgt, const float *__restrict__ wm,
                      uint16_t *__restrict__ gout, float *__restrict__ loss, long long P, int C, float inv_wsum) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (p < P) {
        const float w = wm[p];
        for (int c = 0; c < C; ++c) {
            float gv = 0.0f;
            if (c < 3) {
                const float d = out[p * C + c] - tgt[p * 3 + c];
                l += w * d * d;
                gv = 2.0f * w * d * inv_wsum;
            }
            gout[p * C + c] = f2bf(gv);
        }
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l * inv_wsum);
}

// network input from host fields: x[p] = (the kNetIn host channels, then zeros)
__global__ void k_pack_input(const float *__restrict__ fields, uint16_t *__restrict__ x, long long P, int C) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const float4 f0 = *(const float4 *)(fields + p * 8), f1 = *(const float4 *)(fields + p * 8 + 4);
    u16x8 v;
    v[0] = f2bf(f0.x);
    v[1] = f2bf(f0.y);
    v[2] = f2bf(f0.z);
    v[3] = f2bf(f0.w);
    v[4] = f2bf(f1.x);
    v[5] = f2bf(f1.y);
    v[6] = f2bf(f1.z);
    v[7] = f2bf(f1.w);
    *(u16x8 *)(x + p * C) = v;
    for (int c = 8; c < C; c += 8) *(u16x8 *)(x + p * C + c) = u16x8{};
}

// Adam (Kingma and Ba 2015) on f32 master weights; refreshes the bf16 copy
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, uint16_t *__restrict__ wb, size_t n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

thread_local std::string g_unet_error;

template <class F>
int uguard(F &&fn) {
    try {
        fn();
        return LBM_OK;
    } catch (const HipError &e) {
        g_unet_error = std::string(e.what) + ": " + hipGetErrorString(e.e);
        return e.e == hipErrorOutOfMemory ? LBM_ENOMEM : LBM_EHIP;
    } catch (const ArgError &e) {
        g_unet_error = e.msg;
        return LBM_EINVAL;
    } catch (const std::bad_alloc &) {
        g_unet_error = "host allocation failed";
        return LBM_ENOMEM;
    } catch (...) {
        g_unet_error = "unexpected exception";
        return LBM_EHIP;
    }
}

void ucheck(bool ok, const char *msg) {
    if (!ok) throw ArgError{msg};
}

inline unsigned nblk(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

ConvIn one_input(const uint16_t *a, int c) {
    ConvIn s;
    s.a = a;
    s.ca = c;
    return s;
}

// Block shapes of the LDS-staged 3x3 convolution: {MI, NI, WN}
constexpr int kTiles3[13][3] = {{1, 1, 1}, {2, 1, 1}, {4, 1, 1}, {1, 2, 1}, {2, 2, 1}, {1, 1, 2}, {2, 1, 2},
                                {1, 2, 2}, {2, 2, 2}, {1, 1, 4}, {2, 1, 4}, {1, 2, 4}, {2, 2, 4}};

// Register tile / block shape, K split and kernel of a convolution launch.
// tile: index into kTiles3 plus one (0: choose); split: input-chunk splits (0:
// choose); kind: 0 k_conv3_lds (weights per wave from global memory), 1
// k_conv3_lw (weights in LDS, 32-channel chunks), 2 k_conv3_lw with 64-channel
// chunks, -1 choose.
struct ConvTile {
    int tile = 0, split = 0, kind = -1;
};

struct ConvPlan {
    int tile, split, kind;
};

long long conv3_blocks(int t, int N, int H, int W, int Cout) {
    const int th = (4 / kTiles3[t][2]) * kTiles3[t][0];
    return (long long)N * ((H + th - 1) / th) * (W / 32) * (Cout / (32 * kTiles3[t][1] * kTiles3[t][2]));
}

// Block shape and split by layer shape, from timings of every shape on MI355X
// for the network's layers (batch 1 at 512 x 1024 and batch 8 at 256 x 256):
// few output channels favour short, single-wave-column tiles; wide layers
// favour spreading the channels over the four waves.  A shape that leaves
// fewer than 256 blocks (one per CU) falls back to a narrower one, then to a
// split of the input chunks.
ConvPlan conv3_plan(int N, int H, int W, int Cin, int Cout) {
    // Weights staged in LDS (k_conv3_lw) on one-row-per-wave, 32-channel
    // blocks: fastest for every layer shape of the network except 32 -> 32
    // channels (one input chunk, where the halo-only kernel keeps twice the
    // blocks resident); K split while the blocks cover fewer than half the
    // CUs (at 128..255 blocks the split's extra pass costs more than it gains).
    // Where those blocks leave at most one per CU anyway, 64-channel chunks.
    if (!(Cin == 32 && Cout == 32)) {
        const int kind = Cin % 64 == 0 && conv3_blocks(0, N, H, W, Cout) <= 256 ? 2 : 1;
        int split = 1;
        const int nck = Cin / (kind == 2 ? 64 : 32);
        while (split < 4 && 2 * split <= nck && conv3_blocks(0, N, H, W, Cout) * split < 128) split *= 2;
        return {0, split, kind};
    }
    // preference lists, indices into kTiles3
    static const int p32a[] = {1, 0, 2};           // Cout 32, Cin 32
    static const int p32b[] = {1, 2, 0};           // Cout 32, Cin > 32
    static const int p64a[] = {5, 4, 0};           // Cout 64, Cin 32
    static const int p64b[] = {4, 6, 1, 0};        // Cout 64, Cin > 32
    static const int p128a[] = {10, 6, 1, 0};      // Cout 128, Cin <= 128
    static const int p128b[] = {2, 10, 6, 1, 0};   // Cout 128, Cin > 128
    static const int pwide[] = {10, 6, 1, 9, 0};   // Cout >= 256 (multiples of 128)
    const int *list;
    int len;
    if (Cout == 32) {
        list = Cin == 32 ? p32a : p32b;
        len = 3;
    } else if (Cout == 64) {
        list = Cin == 32 ? p64a : p64b;
        len = Cin == 32 ? 3 : 4;
    } else if (Cout == 128) {
        list = Cin <= 128 ? p128a : p128b;
        len = Cin <= 128 ? 4 : 5;
    } else if (Cout % 128 == 0) {
        list = pwide;
        len = 5;
    } else {
        list = p64b;  // other multiples of 32 (64 k + 32): shapes with 32 channels per block column
        len = 4;
    }
    int best = -1;
    for (int i = 0; i < len && best < 0; ++i) {
        const int t = list[i];
        if (Cout % (32 * kTiles3[t][1] * kTiles3[t][2]) == 0 && conv3_blocks(t, N, H, W, Cout) >= 256) best = t;
    }
    if (best < 0) best = 0;
    int split = 1;
    const int nck = Cin / 32;
    while (split < 4 && 2 * split <= nck && conv3_blocks(best, N, H, W, Cout) * split < 256) split *= 2;
    return {best, split, 0};
}

// f32 elements of split-K partial sums a 3x3 convolution needs (0: no split)
size_t conv_split_elems(int ks, long long P, int N, int H, int W, int Cin, int Cout) {
    if (ks != 3) return 0;
    const ConvPlan pl = conv3_plan(N, H, W, Cin, Cout);
    return pl.split > 1 ? (size_t)pl.split * P * Cout : 0;
}

// Launch a convolution (also used for input gradients): 3x3 on k_conv3_lds or
// k_conv3_lw (conv3_plan), 1x1 on k_conv1x1.  Only channels co < co_store are
// stored (all of them when the K split is taken).  pool_out: also the output's
// 2 x 2 average pool; head: also a 1x1 head on the output (ConvHead); both in
// the epilogue where the block shape allows, else as passes of their own.
void conv_fwd(hipStream_t s, int ks, const ConvIn &in, const uint16_t *w, const float *bias, void *out, int N, int H,
              int W, int Cout, bool relu, bool out_f32, int co_store, float *ws, size_t ws_elems,
              ConvTile force = ConvTile{}, uint16_t *pool_out = nullptr, ConvHead head = ConvHead{}) {
    const long long P = (long long)N * H * W;
    const int Cin = in.ca + in.cb;
    ucheck(W % 32 == 0 && in.ca % 32 == 0 && in.cb % 32 == 0 && Cin > 0 && Cout % 32 == 0,
           "convolution shape: W, channel counts must be multiples of 32");
    ucheck(!in.a_up || H % 2 == 0, "upsampled input needs an even height");
    if (ks == 1) {
        ucheck(!in.a_up && in.cb == 0 && !pool_out && !head.w, "1x1 convolutions take one plain input only");
        int mi = 1, ni = 1;
        if (P % 64 == 0 && P / 64 * (Cout / 32) >= 4096) mi = 2;
        else if (Cout % 64 == 0 && P / 32 * (Cout / 64) >= 4096) ni = 2;
        const dim3 grid((unsigned)((P + 128 * mi - 1) / (128 * mi)), (unsigned)(Cout / (32 * ni)), 1);
#define K1(MI, NI, R, F)                                                                                     \
    hipLaunchKernelGGL((k_conv1x1<MI, NI, R, F>), grid, dim3(256), 0, s, in.a, w, bias, out, P, Cin, Cout, \
                       co_store)
#define PICK(R, F)                       \
    do {                                 \
        if (mi == 2) K1(2, 1, R, F);     \
        else if (ni == 2) K1(1, 2, R, F); \
        else K1(1, 1, R, F);             \
    } while (0)
        if (out_f32) PICK(false, true);
        else if (relu) PICK(true, false);
        else PICK(false, false);
#undef PICK
#undef K1
        HIPCHK(hipGetLastError());
        return;
    }
    // the 3x3 epilogues move four channels at a time (float4 bias and partial
    // sums, 8- or 16-byte output and pooled stores)
    auto aligned = [](const void *q, uintptr_t a) { return ((uintptr_t)q & (a - 1)) == 0; };
    ucheck((!bias || aligned(bias, 16)) && aligned(out, out_f32 ? 16 : 8) && (!ws || aligned(ws, 16)) &&
               (!pool_out || aligned(pool_out, 8)),
           "convolution buffers must be 16-byte aligned (8 for bf16 outputs)");
    ConvPlan pl = conv3_plan(N, H, W, Cin, Cout);
    if (force.tile) {  // a forced block shape runs on the halo-only kernel unless the kind is forced too
        pl.tile = force.tile - 1;
        pl.kind = 0;
    }
    if (force.kind >= 0) pl.kind = force.kind;
    if (force.split) pl.split = force.split;
    const int t = pl.tile;
    ucheck(t >= 0 && t < 13 && Cout % (32 * kTiles3[t][1] * kTiles3[t][2]) == 0, "unsupported block shape");
    ucheck(pl.split == 1 || pl.split == 2 || pl.split == 4, "unsupported split");
    while (pl.split > Cin / (pl.kind == 2 ? 64 : 32)) pl.split /= 2;  // every split gets at least one input chunk
    const bool split = pl.split > 1;
    ucheck(!split || (ws && ws_elems >= (size_t)pl.split * P * Cout), "convolution split workspace too small");
    const dim3 grid((unsigned)conv3_blocks(t, N, H, W, Cout) / (Cout / (32 * kTiles3[t][1] * kTiles3[t][2])),
                    (unsigned)(Cout / (32 * kTiles3[t][1] * kTiles3[t][2])), (unsigned)pl.split);
    const int r = relu ? 1 : 0, f = out_f32 ? 1 : 0;
    const int kind = pl.kind;
    ucheck(kind == 0 || 32 * kTiles3[t][1] * kTiles3[t][2] <= 64, "block shape too wide for LDS weights");
    ucheck(kind != 2 || (t == 0 && Cin % 64 == 0), "64-channel chunks need the (1,1,1) block and Cin % 64 == 0");
    // the 2 x 2 average pool of the output rides in the epilogue when the block
    // holds row pairs and the output is final (no K split); otherwise k_pool2
    ucheck(!pool_out || (!out_f32 && co_store == Cout && H % 2 == 0), "pooled output needs a full bf16 output");
    const bool fuse_pool = pool_out && !split && (kTiles3[t][0] == 2 || (kTiles3[t][0] == 1 && kTiles3[t][2] == 1));
    uint16_t *pout = fuse_pool ? pool_out : nullptr;
    // a 1x1 head rides in the epilogue of the per-wave-weights kernel when one
    // wave holds all 32 output channels of its pixels; otherwise its own pass
    ucheck(!head.w || (Cout == 32 && !out_f32 && co_store == Cout && head.b && head.out && head.store <= 32 &&
                       head.cout == 32 && aligned(head.w, 8) && aligned(head.b, 16) && aligned(head.out, 16)),
           "head: a bf16 32-channel layer, 32 x 32 weights, 32 f32 outputs per pixel");
    const bool fuse_head = head.w && kind == 0 && !split && kTiles3[t][1] == 1 && kTiles3[t][2] == 1;
    const ConvHead hd = fuse_head ? head : ConvHead{};
#define L3(MI, NI, WN)                                                                                             \
    do {                                                                                                           \
        if (kind == 2)                                                                                             \
            hipLaunchKernelGGL((k_conv3_lw<1, 1, 1, 64>), grid, dim3(256), 0, s, in, w, bias, out, ws, pout, N, H, \
                               W, Cout, co_store, r, f);                                                           \
        else if (kind == 1 && 32 * (NI) * (WN) <= 64)                                                              \
            hipLaunchKernelGGL((k_conv3_lw<MI, NI, (32 * NI * WN <= 64 ? WN : 1)>), grid, dim3(256), 0, s, in, w,   \
                               bias, out, ws, pout, N, H, W, Cout, co_store, r, f);                                \
        else                                                                                                       \
            hipLaunchKernelGGL((k_conv3_lds<MI, NI, WN>), grid, dim3(256), 0, s, in, w, bias, out, ws, pout, hd, N, \
                               H, W, Cout, co_store, r, f);                                                        \
    } while (0)
    switch (t) {
        case 0: L3(1, 1, 1); break;
        case 1: L3(2, 1, 1); break;
        case 2: L3(4, 1, 1); break;
        case 3: L3(1, 2, 1); break;
        case 4: L3(2, 2, 1); break;
        case 5: L3(1, 1, 2); break;
        case 6: L3(2, 1, 2); break;
        case 7: L3(1, 2, 2); break;
        case 8: L3(2, 2, 2); break;
        case 9: L3(1, 1, 4); break;
        case 10: L3(2, 1, 4); break;
        case 11: L3(1, 2, 4); break;
        default: L3(2, 2, 4); break;
    }
#undef L3
    HIPCHK(hipGetLastError());
    if (split) {
        const long long n = P * Cout;
        const dim3 fg(nblk((size_t)(n / 4)));
        if (out_f32)
            hipLaunchKernelGGL((k_conv_finish<false, true>), fg, dim3(256), 0, s, ws, bias, out, n, Cout, pl.split);
        else if (relu)
            hipLaunchKernelGGL((k_conv_finish<true, false>), fg, dim3(256), 0, s, ws, bias, out, n, Cout, pl.split);
        else
            hipLaunchKernelGGL((k_conv_finish<false, false>), fg, dim3(256), 0, s, ws, bias, out, n, Cout, pl.split);
        HIPCHK(hipGetLastError());
    }
    if (pool_out && !fuse_pool) {
        hipLaunchKernelGGL(k_pool2, dim3(nblk((size_t)P * Cout / 32)), dim3(256), 0, s, (const uint16_t *)out,
                           pool_out, N, H, W, Cout);
        HIPCHK(hipGetLastError());
    }
    if (head.w && !fuse_head)
        conv_fwd(s, 1, one_input((const uint16_t *)out, 32), head.w, head.b, head.out, N, H, W, head.cout, false, true,
                 head.store, ws, ws_elems);
}

void conv_wgrad(hipStream_t s, int ks, const ConvIn &in, const uint16_t *g, float *dw, int N, int H, int W,
                int Cout) {
    const long long P = (long long)N * H * W;
    const int Cin = in.ca + in.cb;
    const int ppb = 2048;
    const dim3 grid((unsigned)((P + ppb - 1) / ppb), (Cout / 32) * (Cin / 32), ks * ks);
    if (ks == 3)
        hipLaunchKernelGGL(k_conv3x3_wgrad<3>, grid, dim3(256), 0, s, in, g, dw, N, H, W, Cout, ppb);
    else
        hipLaunchKernelGGL(k_conv3x3_wgrad<1>, grid, dim3(256), 0, s, in, g, dw, N, H, W, Cout, ppb);
    HIPCHK(hipGetLastError());
}

// one activation tensor (bf16, NHWC) and its gradient
struct Act {
    uint16_t *a = nullptr, *g = nullptr;
    int N = 0, H = 0, W = 0, C = 0;
    size_t n() const { return (size_t)N * H * W * C; }
    long long P() const { return (long long)N * H * W; }
};

struct Conv {
    int ks, cin, cout;
    bool relu;
    size_t woff, boff;  // offsets of weights [ks*ks][cout][cin] and bias [cout] in the parameter array
    size_t wboff;       // offset of the weights in the bf16 arrays
};

}  // namespace

// The network.  Layers (c: 3x3 conv + ReLU, u: 1x1 conv + ReLU then nearest 2x
// upsampling, p: 2x2 average pool, [a|b]: channel concatenation):
//   e1 = c(c(x: 32))            32 ch, H      p1 = p(e1)
//   e2 = c(c(p1))               64 ch, H/2    p2 = p(e2)
//   e3 = c(c(p2))              128 ch, H/4    p3 = p(e3)
//   e4 = c(c(p3))              256 ch, H/8
//   d3 = c(c([u(e4) | e3]))    128 ch, H/4
//   d2 = c(c([u(d3) | e2]))     64 ch, H/2
//   d1 = c(c([u(d2) | e1]))     32 ch, H
//   out = 1x1 conv(d1)          32 ch (the first 3 used), f32
// A 1x1 convolution and ReLU act per pixel, so they commute with nearest
// upsampling: u runs its 1x1 convolution at the coarse resolution (a quarter of
// the work) and the upsampling and concatenation are folded into the next
// convolution's loads (ConvIn).  Only the gradient of each concatenated input
// (k3, k2, k1) is materialised, in the backward pass.
// Input channels: solid fraction, (rho - 1) / s_rho, ux / s_u, uy / s_u,
// (tau - 1/2) / 0.2, u_in / s_u, then zeros; output: the 64-step change of
// (rho, ux, uy) in the same units.
struct lbm_unet {
    int dev = 0;
    hipStream_t stream = nullptr;
    int B = 0, H = 0, W = 0;
    std::vector<Conv> convs;
    size_t nparam = 0, nweight = 0;
    float *w32 = nullptr, *m = nullptr, *v = nullptr, *g32 = nullptr;
    uint16_t *wb = nullptr, *wt = nullptr;
    float *out32 = nullptr, *tgt = nullptr, *wm = nullptr, *fields = nullptr, *loss = nullptr;
    uint16_t *gout = nullptr;
    Act x, e1a, e1, p1, e2a, e2, p2, e3a, e3, p3, e4a, e4, d3u, k3, d3a, d3, d2u, k2, d2a, d2, d1u, k1, d1a, d1;
    float *ws = nullptr;  // split-K partial sums
    size_t ws_elems = 0;
    std::vector<void *> mem;
    int adam_t = 0;
    // the forward pass as a hipGraph (captured once; a graph runs on any stream)
    bool use_graph = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // the network's tensors are shared by every stream that uses it (its own and
    // the streams of the grids that jump with it): `used` marks the end of the
    // last use enqueued on `used_on`, and a use on another stream waits for it
    hipEvent_t used = nullptr;
    hipStream_t used_on = nullptr;

    void order_after_last_use(hipStream_t s) {
        if (used_on && used_on != s) HIPCHK(hipStreamWaitEvent(s, used, 0));
    }
    void mark_use(hipStream_t s) {
        if (!used) HIPCHK(hipEventCreateWithFlags(&used, hipEventDisableTiming));
        HIPCHK(hipEventRecord(used, s));
        used_on = s;
    }

    void *alloc(size_t bytes) {
        void *p = dalloc<char>(bytes);
        mem.push_back(p);
        return p;
    }
    void make(Act &t, int h, int w, int c, bool values = true) {
        t.N = B;
        t.H = h;
        t.W = w;
        t.C = c;
        t.a = values ? (uint16_t *)alloc(t.n() * 2) : nullptr;
        t.g = (uint16_t *)alloc(t.n() * 2);
    }
    void drop_graph() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
    }
    void release() {
        (void)hipSetDevice(dev);
        if (used) {
            (void)hipEventSynchronize(used);
            (void)hipEventDestroy(used);
            used = nullptr;
        }
        drop_graph();
        for (void *p : mem) (void)hipFree(p);
        mem.clear();
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }

    size_t add_conv(int ks, int cin, int cout, bool relu) {
        Conv c{ks, cin, cout, relu, nparam, nparam + (size_t)ks * ks * cin * cout, nweight};
        nparam += (size_t)ks * ks * cin * cout + cout;
        nweight += (size_t)ks * ks * cin * cout;
        convs.push_back(c);
        return convs.size() - 1;
    }

    void build(uint64_t seed) {
        HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        const int h2 = H / 2, w2 = W / 2, h4 = H / 4, w4 = W / 4, h8 = H / 8, w8 = W / 8;
        make(x, H, W, 32);
        make(e1a, H, W, 32);
        make(e1, H, W, 32);
        make(p1, h2, w2, 32);
        make(e2a, h2, w2, 64);
        make(e2, h2, w2, 64);
        make(p2, h4, w4, 64);
        make(e3a, h4, w4, 128);
        make(e3, h4, w4, 128);
        make(p3, h8, w8, 128);
        make(e4a, h8, w8, 256);
        make(e4, h8, w8, 256);
        make(d3u, h8, w8, 128);
        make(k3, h4, w4, 256, false);
        make(d3a, h4, w4, 128);
        make(d3, h4, w4, 128);
        make(d2u, h4, w4, 64);
        make(k2, h2, w2, 128, false);
        make(d2a, h2, w2, 64);
        make(d2, h2, w2, 64);
        make(d1u, h2, w2, 32);
        make(k1, H, W, 64, false);
        make(d1a, H, W, 32);
        make(d1, H, W, 32);
        const int specs[18][4] = {{3, 32, 32, 1},  {3, 32, 32, 1},   {3, 32, 64, 1},   {3, 64, 64, 1},
                                  {3, 64, 128, 1}, {3, 128, 128, 1}, {3, 128, 256, 1}, {3, 256, 256, 1},
                                  {1, 256, 128, 1}, {3, 256, 128, 1}, {3, 128, 128, 1}, {1, 128, 64, 1},
                                  {3, 128, 64, 1},  {3, 64, 64, 1},   {1, 64, 32, 1},   {3, 64, 32, 1},
                                  {3, 32, 32, 1},   {1, 32, 32, 0}};
        for (auto &sp : specs) add_conv(sp[0], sp[1], sp[2], sp[3] != 0);
        w32 = (float *)alloc(nparam * 4);
        m = (float *)alloc(nparam * 4);
        v = (float *)alloc(nparam * 4);
        g32 = (float *)alloc(nparam * 4);
        wb = (uint16_t *)alloc(nweight * 2);
        wt = (uint16_t *)alloc(nweight * 2);
        const size_t P = (size_t)B * H * W;
        out32 = (float *)alloc(P * 32 * 4);
        gout = (uint16_t *)alloc(P * 32 * 2);
        tgt = (float *)alloc(P * 3 * 4);
        wm = (float *)alloc(P * 4);
        fields = (float *)alloc(P * 8 * 4);
        loss = (float *)alloc(4);
        // split-K workspace: the largest need over the forward and input-gradient convolutions
        for (const Conv &c : convs) {
            // level of the layer from its output (forward) / input (gradient) channel count
            for (int lv = 0; lv < 4; ++lv) {
                const int h = H >> lv, w = W >> lv;
                const long long Pl = (long long)B * h * w;
                ws_elems = std::max({ws_elems, conv_split_elems(c.ks, Pl, B, h, w, c.cin, c.cout),
                                     conv_split_elems(c.ks, Pl, B, h, w, c.cout, c.cin)});
            }
        }
        if (ws_elems) ws = (float *)alloc(ws_elems * 4);
        // He-normal weights (fan-in = taps * real input channels), zero biases; the
        // head starts near zero so the untrained network predicts "no change"
        std::vector<float> hw(nparam, 0.0f);
        uint64_t s = seed;
        auto normal = [&]() {
            const double u1 = ((lbm_splitmix64(&s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
            const double u2 = (lbm_splitmix64(&s) >> 11) * (1.0 / 9007199254740992.0);
            return (float)(std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
        };
        for (size_t i = 0; i < convs.size(); ++i) {
            const Conv &c = convs[i];
            const int fan_in = c.ks * c.ks * (i == 0 ? kNetInLive : c.cin);
            const float std_ = (i + 1 == convs.size() ? 0.01f : 1.0f) * (float)std::sqrt(2.0 / fan_in);
            for (int t = 0; t < c.ks * c.ks; ++t)
                for (int co = 0; co < c.cout; ++co)
                    for (int ci = 0; ci < c.cin; ++ci) {
                        const bool live = (i != 0 || ci < kNetIn) && (i + 1 != convs.size() || co < 3);
                        hw[c.woff + ((size_t)t * c.cout + co) * c.cin + ci] = live ? std_ * normal() : 0.0f;
                    }
        }
        HIPCHK(hipMemcpyAsync(w32, hw.data(), nparam * 4, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemsetAsync(m, 0, nparam * 4, stream));
        HIPCHK(hipMemsetAsync(v, 0, nparam * 4, stream));
        refresh_bf16();
        HIPCHK(hipStreamSynchronize(stream));
    }

    // bf16 forward weights and flipped/transposed input-gradient weights from w32
    void refresh_bf16() {
        for (const Conv &c : convs) {
            const size_t n = (size_t)c.ks * c.ks * c.cin * c.cout;
            hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(n)), dim3(256), 0, stream, w32 + c.woff, wb + c.wboff, n);
            hipLaunchKernelGGL(k_flip_transpose, dim3(nblk(n)), dim3(256), 0, stream, wb + c.wboff, wt + c.wboff,
                               c.cin, c.cout, c.ks * c.ks);
        }
        HIPCHK(hipGetLastError());
    }

    static ConvIn cat_up(const Act &coarse, const Act &fine) {
        ConvIn s;
        s.a = coarse.a;
        s.ca = coarse.C;
        s.a_up = 1;
        s.b = fine.a;
        s.cb = fine.C;
        return s;
    }
    void conv(int i, const ConvIn &in, int h, int w, uint16_t *out) {
        const Conv &c = convs[i];
        ucheck(in.ca + in.cb == c.cin, "layer input channels");
        conv_fwd(stream, c.ks, in, wb + c.wboff, w32 + c.boff, out, B, h, w, c.cout, c.relu, false, c.cout, ws,
                 ws_elems);
    }
    void fwd_conv(int i, const Act &in, const Act &out) { conv(i, one_input(in.a, in.C), in.H, in.W, out.a); }
    // conv i into out, and out's 2 x 2 average pool into pooled
    void fwd_conv_pool(int i, const Act &in, const Act &out, const Act &pooled) {
        const Conv &c = convs[i];
        ucheck(in.C == c.cin && pooled.C == c.cout && pooled.H * 2 == in.H && pooled.W * 2 == in.W, "pooled layer");
        conv_fwd(stream, c.ks, one_input(in.a, in.C), wb + c.wboff, w32 + c.boff, out.a, B, in.H, in.W, c.cout, c.relu,
                 false, c.cout, ws, ws_elems, ConvTile{}, pooled.a);
    }

    void forward() {
        fwd_conv(0, x, e1a);
        fwd_conv_pool(1, e1a, e1, p1);
        fwd_conv(2, p1, e2a);
        fwd_conv_pool(3, e2a, e2, p2);
        fwd_conv(4, p2, e3a);
        fwd_conv_pool(5, e3a, e3, p3);
        fwd_conv(6, p3, e4a);
        fwd_conv(7, e4a, e4);
        fwd_conv(8, e4, d3u);
        conv(9, cat_up(d3u, e3), e3.H, e3.W, d3a.a);
        fwd_conv(10, d3a, d3);
        fwd_conv(11, d3, d2u);
        conv(12, cat_up(d2u, e2), e2.H, e2.W, d2a.a);
        fwd_conv(13, d2a, d2);
        fwd_conv(14, d2, d1u);
        conv(15, cat_up(d1u, e1), e1.H, e1.W, d1a.a);
        // the last 3x3 layer carries the 1x1 head (three f32 outputs per pixel) in its epilogue
        const Conv &c16 = convs[16], &h = convs[17];
        ucheck(h.ks == 1 && h.cin == 32 && c16.cout == 32, "head layer");
        ConvHead hd;
        hd.w = wb + h.wboff;
        hd.b = w32 + h.boff;
        hd.out = out32;
        hd.cout = h.cout;
        hd.store = 3;
        conv_fwd(stream, c16.ks, one_input(d1a.a, d1a.C), wb + c16.wboff, w32 + c16.boff, d1.a, B, H, W, c16.cout,
                 c16.relu, false, c16.cout, ws, ws_elems, ConvTile{}, nullptr, hd);
        HIPCHK(hipGetLastError());
    }

    // gradients of conv i: weights and bias into g32; input gradient into gin (if any)
    void bwd_conv(int i, const ConvIn &in, int h, int w, const uint16_t *gout_, uint16_t *gin) {
        const Conv &c = convs[i];
        conv_wgrad(stream, c.ks, in, gout_, g32 + c.woff, B, h, w, c.cout);
        hipLaunchKernelGGL(k_bias_grad, dim3(256, (c.cout + 63) / 64), dim3(256), 0, stream, gout_, g32 + c.boff,
                           (long long)B * h * w, c.cout);
        if (gin)
            conv_fwd(stream, c.ks, one_input(gout_, c.cout), wt + c.wboff, nullptr, gin, B, h, w, c.cin, false, false,
                     c.cin, ws, ws_elems);
    }
    void bwd_conv(int i, const Act &in, const uint16_t *gout_, uint16_t *gin) {
        bwd_conv(i, one_input(in.a, in.C), in.H, in.W, gout_, gin);
    }
    void relu_bwd(const Act &t) {
        hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(t.n())), dim3(256), 0, stream, t.g, t.a, t.n());
    }
    // gradient of [up2(coarse) | fine] -> coarse.g (assigned), fine.g (assigned)
    void split_up(const Act &k, const Act &coarse, const Act &fine) {
        const size_t n = coarse.n() + fine.n();
        hipLaunchKernelGGL(k_split_up, dim3(nblk(n / 8)), dim3(256), 0, stream, k.g, coarse.g, fine.g, B, k.H, k.W,
                           coarse.C, fine.C);
    }
    void pool_bwd(const Act &coarse, const Act &fine) {  // accumulates into fine.g
        hipLaunchKernelGGL(k_pool2_bwd, dim3(nblk(fine.n())), dim3(256), 0, stream, coarse.g, fine.g, B, fine.H,
                           fine.W, fine.C, 1);
    }

    void backward() {
        HIPCHK(hipMemsetAsync(g32, 0, nparam * 4, stream));
        bwd_conv(17, d1, gout, d1.g);
        relu_bwd(d1);
        bwd_conv(16, d1a, d1.g, d1a.g);
        relu_bwd(d1a);
        bwd_conv(15, cat_up(d1u, e1), H, W, d1a.g, k1.g);
        split_up(k1, d1u, e1);
        relu_bwd(d1u);
        bwd_conv(14, d2, d1u.g, d2.g);
        relu_bwd(d2);
        bwd_conv(13, d2a, d2.g, d2a.g);
        relu_bwd(d2a);
        bwd_conv(12, cat_up(d2u, e2), e2.H, e2.W, d2a.g, k2.g);
        split_up(k2, d2u, e2);
        relu_bwd(d2u);
        bwd_conv(11, d3, d2u.g, d3.g);
        relu_bwd(d3);
        bwd_conv(10, d3a, d3.g, d3a.g);
        relu_bwd(d3a);
        bwd_conv(9, cat_up(d3u, e3), e3.H, e3.W, d3a.g, k3.g);
        split_up(k3, d3u, e3);
        relu_bwd(d3u);
        bwd_conv(8, e4, d3u.g, e4.g);
        relu_bwd(e4);
        bwd_conv(7, e4a, e4.g, e4a.g);
        relu_bwd(e4a);
        bwd_conv(6, p3, e4a.g, p3.g);
        pool_bwd(p3, e3);
        relu_bwd(e3);
        bwd_conv(5, e3a, e3.g, e3a.g);
        relu_bwd(e3a);
        bwd_conv(4, p2, e3a.g, p2.g);
        pool_bwd(p2, e2);
        relu_bwd(e2);
        bwd_conv(3, e2a, e2.g, e2a.g);
        relu_bwd(e2a);
        bwd_conv(2, p1, e2a.g, p1.g);
        pool_bwd(p1, e1);
        relu_bwd(e1);
        bwd_conv(1, e1a, e1.g, e1a.g);
        relu_bwd(e1a);
        bwd_conv(0, x, e1a.g, nullptr);
        HIPCHK(hipGetLastError());
    }

    // host inputs -> device; returns the loss
    float loss_and_grad(const float *hx, const float *hy, const float *hwm) {
        const size_t P = (size_t)B * H * W;
        HIPCHK(hipMemcpyAsync(fields, hx, P * 32, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(tgt, hy, P * 12, hipMemcpyHostToDevice, stream));
        HIPCHK(hipMemcpyAsync(wm, hwm, P * 4, hipMemcpyHostToDevice, stream));
        double wsum = 0.0;
        for (size_t i = 0; i < P; ++i) wsum += hwm[i];
        ucheck(wsum > 0.0, "loss weights sum to zero");
        hipLaunchKernelGGL(k_pack_input, dim3(nblk(P)), dim3(256), 0, stream, fields, x.a, (long long)P, 32);
        forward();
        HIPCHK(hipMemsetAsync(loss, 0, 4, stream));
        hipLaunchKernelGGL(k_mse, dim3(nblk(P)), dim3(256), 0, stream, out32, tgt, wm, gout, loss, (long long)P, 32,
                           (float)(1.0 / wsum));
        HIPCHK(hipGetLastError());
        backward();
        float l = 0.0f;
        HIPCHK(hipMemcpyAsync(&l, loss, 4, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return l;
    }

    void adam(float lr) {
        ++adam_t;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - (float)std::pow(b1, adam_t), c2 = 1.0f - (float)std::pow(b2, adam_t);
        hipLaunchKernelGGL(k_adam, dim3(nblk(nparam)), dim3(256), 0, stream, w32, m, v, g32, (uint16_t *)nullptr,
                           nparam, lr, b1, b2, 1e-8f, c1, c2);
        HIPCHK(hipGetLastError());
        refresh_bf16();
    }
};

// hooks for the grid engine's surrogate jump (same library)
uint16_t *unet_input_tensor(lbm_unet *n) { return n->x.a; }
float *unet_output_tensor(lbm_unet *n) { return n->out32; }
void unet_shape(const lbm_unet *n, int *b, int *h, int *w) {
    *b = n->B;
    *h = n->H;
    *w = n->W;
}
// The forward pass on stream s.  Every buffer it touches is fixed for the
// network's lifetime (weights are updated in place), so it is captured once as
// a hipGraph (no memset nodes) and replayed: at surrogate sizes the ~30
// launches cost as much as their work.
void unet_begin_use(lbm_unet *n, hipStream_t s) { n->order_after_last_use(s); }
void unet_end_use(lbm_unet *n, hipStream_t s) { n->mark_use(s); }

void unet_forward_on(lbm_unet *n, hipStream_t s) {
    hipStream_t saved = n->stream;
    n->stream = s;
    try {
        if (!n->use_graph) {
            n->forward();
        } else {
            if (!n->graph_exec) {
                HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                try {
                    n->forward();
                } catch (...) {
                    hipGraph_t g = nullptr;
                    (void)hipStreamEndCapture(s, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                HIPCHK(hipStreamEndCapture(s, &n->graph));
                HIPCHK(hipGraphInstantiate(&n->graph_exec, n->graph, nullptr, nullptr, 0));
            }
            HIPCHK(hipGraphLaunch(n->graph_exec, s));
        }
    } catch (...) {
        n->stream = saved;
        throw;
    }
    n->stream = saved;
}
int unet_device(const lbm_unet *n) { return n->dev; }

extern "C" {

const char *lbm_unet_last_error(void) { return g_unet_error.c_str(); }

int lbm_unet_create(int batch, int height, int width, uint64_t seed, int device, lbm_unet **out) {
    return uguard([&] {
        ucheck(out != nullptr, "null output");
        *out = nullptr;
        ucheck(batch >= 1 && height >= 8 && height % 8 == 0 && width % 256 == 0 && width > 0,
               "need batch >= 1, height a multiple of 8, width a multiple of 256");
        HIPCHK(hipSetDevice(device));
        lbm_unet *n = new lbm_unet();
        n->dev = device;
        n->B = batch;
        n->H = height;
        n->W = width;
        try {
            n->build(seed);
        } catch (...) {
            n->release();
            delete n;
            throw;
        }
        *out = n;
    });
}

int lbm_unet_set_graph(lbm_unet *n, int enable) {
    return uguard([&] {
        ucheck(n != nullptr, "null network");
        HIPCHK(hipSetDevice(n->dev));
        if (n->used) HIPCHK(hipEventSynchronize(n->used));  // a jump may still run the graph
        n->use_graph = enable != 0;
        if (!n->use_graph) n->drop_graph();
    });
}

int lbm_unet_destroy(lbm_unet *n) {
    if (!n) return LBM_EINVAL;
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    n->release();
    delete n;
    return LBM_OK;
}

int lbm_unet_num_params(const lbm_unet *n, size_t *count) {
    if (!n || !count) return LBM_EINVAL;
    *count = n->nparam;
    return LBM_OK;
}

int lbm_unet_get_params(lbm_unet *n, float *params) {
    return uguard([&] {
        ucheck(n && params, "null argument");
        HIPCHK(hipSetDevice(n->dev));
        n->order_after_last_use(n->stream);
        HIPCHK(hipMemcpyAsync(params, n->w32, n->nparam * 4, hipMemcpyDeviceToHost, n->stream));
        HIPCHK(hipStreamSynchronize(n->stream));
    });
}

int lbm_unet_set_params(lbm_unet *n, const float *params) {
    return uguard([&] {
        ucheck(n && params, "null argument");
        HIPCHK(hipSetDevice(n->dev));
        n->order_after_last_use(n->stream);
        HIPCHK(hipMemcpyAsync(n->w32, params, n->nparam * 4, hipMemcpyHostToDevice, n->stream));
        n->refresh_bf16();
        HIPCHK(hipStreamSynchronize(n->stream));
    });
}

int lbm_unet_forward(lbm_unet *n, const float *x, float *y) {
    return uguard([&] {
        ucheck(n && x && y, "null argument");
        HIPCHK(hipSetDevice(n->dev));
        n->order_after_last_use(n->stream);
        const size_t P = (size_t)n->B * n->H * n->W;
        HIPCHK(hipMemcpyAsync(n->fields, x, P * 32, hipMemcpyHostToDevice, n->stream));
        hipLaunchKernelGGL(k_pack_input, dim3(nblk(P)), dim3(256), 0, n->stream, n->fields, n->x.a, (long long)P, 32);
        n->forward();
        HIPCHK(hipMemcpy2DAsync(y, 12, n->out32, 128, 12, P, hipMemcpyDeviceToHost, n->stream));
        HIPCHK(hipStreamSynchronize(n->stream));
    });
}

int lbm_unet_loss_grad(lbm_unet *n, const float *x, const float *y, const float *weights, float *loss, float *grad) {
    return uguard([&] {
        ucheck(n && x && y && weights, "null argument");
        HIPCHK(hipSetDevice(n->dev));
        n->order_after_last_use(n->stream);
        const float l = n->loss_and_grad(x, y, weights);
        if (loss) *loss = l;
        if (grad) {
            HIPCHK(hipMemcpyAsync(grad, n->g32, n->nparam * 4, hipMemcpyDeviceToHost, n->stream));
            HIPCHK(hipStreamSynchronize(n->stream));
        }
    });
}

int lbm_unet_train_step(lbm_unet *n, const float *x, const float *y, const float *weights, float lr, float *loss) {
    return uguard([&] {
        ucheck(n && x && y && weights, "null argument");
        HIPCHK(hipSetDevice(n->dev));
        n->order_after_last_use(n->stream);
        const float l = n->loss_and_grad(x, y, weights);
        n->adam(lr);
        HIPCHK(hipStreamSynchronize(n->stream));
        if (loss) *loss = l;
    });
}

int lbm_unet_time_forward(lbm_unet *n, int reps, float *ms) {
    return uguard([&] {
        ucheck(n && ms && reps > 0, "bad argument");
        HIPCHK(hipSetDevice(n->dev));
        n->order_after_last_use(n->stream);
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        n->forward();
        HIPCHK(hipEventRecord(e0, n->stream));
        for (int r = 0; r < reps; ++r) n->forward();
        HIPCHK(hipEventRecord(e1, n->stream));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipEventElapsedTime(ms, e0, e1));
        *ms /= reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

// Test entry point for the convolution primitives on host arrays (f32 in,
// rounded to bf16 on the device).  mode 0: out = conv(in, w) + bias (no ReLU;
// with bit 12 also its pooled bf16 output); mode 1: out = input gradient for
// output gradient `in`; mode 2: out = weight gradient dw[tap][co][ci] for input
// `in` and output gradient `g`; mode 3: timing.  Bits 4..11 force the block
// shape, split and kernel.
int lbm_conv3x3_test(int mode, int N, int H, int W, int Cin, int Cout, const float *in, const float *g,
                     const float *w, const float *bias, float *out, int device) {
    return uguard([&] {
        const int op = mode & 3;
        ConvTile force;
        force.tile = (mode >> 4) & 15;
        force.split = (mode >> 8) & 7;
        force.kind = (mode >> 13) & 1 ? 2 : (mode >> 11) & 1 ? 1 : -1;
        const bool pooled = (mode >> 12) & 1;
        ucheck(force.tile <= 13 && (mode >> 14) == 0, "mode");
        ucheck(!pooled || ((mode & 3) == 0 && H % 2 == 0), "mode: pooled output needs the forward and an even height");
        ucheck(force.split == 0 || force.split == 1 || force.split == 2 || force.split == 4, "mode: split");
        mode = op;
        ucheck(W % 32 == 0 && Cin % 32 == 0 && Cout % 32 == 0, "need W, Cin, Cout multiples of 32");
        HIPCHK(hipSetDevice(device));
        const size_t P = (size_t)N * H * W;
        const size_t nw = (size_t)9 * Cin * Cout;
        const int cin_in = mode == 1 ? Cout : Cin;  // channels of `in`
        const int cout_out = mode == 1 ? Cin : Cout;
        std::vector<void *> mem;
        auto alloc = [&](size_t bytes) {
            void *p = dalloc<char>(bytes);
            mem.push_back(p);
            return p;
        };
        try {
            float *d32 = (float *)alloc(std::max({P * cin_in, P * Cout, nw}) * sizeof(float));
            uint16_t *din = (uint16_t *)alloc(P * cin_in * 2);
            uint16_t *dw = (uint16_t *)alloc(nw * 2);
            HIPCHK(hipMemcpy(d32, in, P * cin_in * sizeof(float), hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(P * cin_in)), dim3(256), 0, 0, d32, din, P * cin_in);
            if (mode != 2) {
                HIPCHK(hipMemcpy(d32, w, nw * sizeof(float), hipMemcpyHostToDevice));
                hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(nw)), dim3(256), 0, 0, d32, dw, nw);
            }
            float *dout = (float *)alloc(std::max(P * cout_out, nw) * sizeof(float));
            const size_t wsn = (size_t)4 * P * std::max(Cin, Cout);
            float *ws = (float *)alloc(wsn * sizeof(float));
            if (mode == 3) {  // time the forward convolution with a bf16 output, as the network runs it
                hipEvent_t e0, e1;
                HIPCHK(hipEventCreate(&e0));
                HIPCHK(hipEventCreate(&e1));
                const ConvIn ci = one_input(din, Cin);
                for (int r = 0; r < 3; ++r)
                    conv_fwd(0, 3, ci, dw, nullptr, dout, N, H, W, Cout, true, false, Cout, ws, wsn, force);
                HIPCHK(hipEventRecord(e0, 0));
                for (int r = 0; r < 20; ++r)
                    conv_fwd(0, 3, ci, dw, nullptr, dout, N, H, W, Cout, true, false, Cout, ws, wsn, force);
                HIPCHK(hipEventRecord(e1, 0));
                HIPCHK(hipEventSynchronize(e1));
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, e0, e1));
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                out[0] = ms / 20.0f;
            } else if (mode == 0) {
                float *db = nullptr;
                if (bias) {
                    db = (float *)alloc(Cout * sizeof(float));
                    HIPCHK(hipMemcpy(db, bias, Cout * sizeof(float), hipMemcpyHostToDevice));
                }
                if (pooled) {  // bf16 output and its 2 x 2 average pool, both returned as f32
                    uint16_t *o16 = (uint16_t *)alloc(P * Cout * 2), *p16 = (uint16_t *)alloc(P / 4 * Cout * 2);
                    conv_fwd(0, 3, one_input(din, Cin), dw, db, o16, N, H, W, Cout, false, false, Cout, ws, wsn, force,
                             p16);
                    hipLaunchKernelGGL(k_bf16_to_f32, dim3(nblk(P * Cout)), dim3(256), 0, 0, o16, dout, P * Cout);
                    HIPCHK(hipMemcpy(out, dout, P * Cout * sizeof(float), hipMemcpyDeviceToHost));
                    hipLaunchKernelGGL(k_bf16_to_f32, dim3(nblk(P / 4 * Cout)), dim3(256), 0, 0, p16, dout,
                                       P / 4 * Cout);
                    HIPCHK(hipMemcpy(out + P * Cout, dout, P / 4 * Cout * sizeof(float), hipMemcpyDeviceToHost));
                } else {
                    conv_fwd(0, 3, one_input(din, Cin), dw, db, dout, N, H, W, Cout, false, true, Cout, ws, wsn,
                             force);
                    HIPCHK(hipMemcpy(out, dout, P * Cout * sizeof(float), hipMemcpyDeviceToHost));
                }
            } else if (mode == 1) {
                uint16_t *dwt = (uint16_t *)alloc(nw * 2);
                hipLaunchKernelGGL(k_flip_transpose, dim3(nblk(nw)), dim3(256), 0, 0, dw, dwt, Cin, Cout, 9);
                conv_fwd(0, 3, one_input(din, Cout), dwt, nullptr, dout, N, H, W, Cin, false, true, Cin, ws, wsn,
                         force);
                HIPCHK(hipMemcpy(out, dout, P * Cin * sizeof(float), hipMemcpyDeviceToHost));
            } else {
                uint16_t *dg = (uint16_t *)alloc(P * Cout * 2);
                HIPCHK(hipMemcpy(d32, g, P * Cout * sizeof(float), hipMemcpyHostToDevice));
                hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(P * Cout)), dim3(256), 0, 0, d32, dg, P * Cout);
                HIPCHK(hipMemset(dout, 0, nw * sizeof(float)));
                conv_wgrad(0, 3, one_input(din, Cin), dg, dout, N, H, W, Cout);
                HIPCHK(hipMemcpy(out, dout, nw * sizeof(float), hipMemcpyDeviceToHost));
            }
            HIPCHK(hipDeviceSynchronize());
        } catch (...) {
            for (void *p : mem) (void)hipFree(p);
            throw;
        }
        for (void *p : mem) (void)hipFree(p);
    });
}

}  // extern "C"
