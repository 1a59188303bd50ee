// Learned surrogate: a 4-level conv U-Net (3x3 convolutions, 32-64-128-256
// channels) with BF16 weights and activations and FP32 accumulation on the
// gfx950 matrix cores, its training (forward, backward, Adam) and its use as a
// 64-step predictor inside an LBM rollout.  No reference counterpart: the
// reference has no learned component.
//
// Convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16.  Activations are
// NHWC; the pixel dimension is the GEMM's M, output channels its N and
// (tap, input channel) its K.  For one wave:
//   A (32 pixels x 16 k):  lane l holds pixel l % 32, k = 8 (l / 32) + 0..7
//                          -> 16 contiguous bytes of one pixel's channels;
//   B (16 k x 32 couts):   lane l holds cout l % 32, k = 8 (l / 32) + 0..7
//                          -> 16 contiguous bytes of W[tap][cout][cin];
//   D (32 x 32 f32):       lane l, value r: pixel 8 (r / 4) + 4 (l / 32) + r % 4,
//                          cout l % 32.
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

__device__ __forceinline__ uint16_t f2bf(float v) {  // round to nearest even
    return __builtin_bit_cast(uint16_t, (__bf16)v);
}
__device__ __forceinline__ float bf2f(uint16_t v) { return (float)__builtin_bit_cast(__bf16, v); }

// ---- convolution forward (and input gradient) ------------------------------------
// in  [P = N*H*W][Cin] bf16, w [9][Cout][Cin] bf16, bias [Cout] f32 (may be null)
// out [P][Cout] bf16 or f32.  Requirements: W % 32 == 0, Cin % 16 == 0,
// Cout % (32 * NT) == 0.  A wave computes 32 pixels x (32 * NT) couts; a block
// is four waves on consecutive 32-pixel runs.
template <int KS, int NT, bool RELU, bool OUT_F32>
__global__ __launch_bounds__(256) void k_conv3x3(const uint16_t *__restrict__ in, const uint16_t *__restrict__ w,
                                                 const float *__restrict__ bias, void *__restrict__ out, int N, int H,
                                                 int W, int Cin, int Cout) {
    constexpr int TAPS = KS * KS, R = KS / 2;  // KS = 3: 3x3 with zero padding; KS = 1: 1x1
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long P = (long long)N * H * W;
    const long long p0 = ((long long)blockIdx.x * 4 + wave) * 32;  // first pixel of this wave
    if (p0 >= P) return;
    const int co0 = blockIdx.y * 32 * NT;
    // the lane's A pixel
    const long long pa = p0 + (lane & 31);
    const int xa = (int)(pa % W);
    const int ya = (int)((pa / W) % H);
    const long long na = pa / ((long long)W * H);
    const int kofs = 8 * (lane >> 5);
    f32x16 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x16{};
    for (int tap = 0; tap < TAPS; ++tap) {
        const int sy = ya + tap / KS - R, sx = xa + tap % KS - R;
        const bool inside = sy >= 0 && sy < H && sx >= 0 && sx < W;
        const uint16_t *arow = in + ((na * H + (inside ? sy : 0)) * W + (inside ? sx : 0)) * Cin + kofs;
        const uint16_t *wrow = w + ((size_t)tap * Cout + co0 + (lane & 31)) * Cin + kofs;
        for (int c0 = 0; c0 < Cin; c0 += 16) {
            u16x8 av = inside ? *(const u16x8 *)(arow + c0) : u16x8{};
            const bf16x8 a = __builtin_bit_cast(bf16x8, av);
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const bf16x8 b = __builtin_bit_cast(bf16x8, *(const u16x8 *)(wrow + (size_t)t * 32 * Cin + c0));
                acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = co0 + 32 * t + (lane & 31);
        const float bv = bias ? bias[co] : 0.0f;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const long long p = p0 + 8 * (r / 4) + 4 * (lane >> 5) + (r % 4);
            float v = acc[t][r] + bv;
            if (RELU) v = fmaxf(v, 0.0f);
            if (OUT_F32) ((float *)out)[p * Cout + co] = v;
            else ((uint16_t *)out)[p * Cout + co] = f2bf(v);
        }
    }
}

// ---- weight gradient -------------------------------------------------------------------
// dw[tap][co][ci] += sum_p in[p + off(tap)][ci] * g[p][co]: a GEMM with M = co,
// N = ci, K = pixels.  A block takes one tap, a 32 x 32 (co, ci) tile and a
// range of pixels split over its four waves; the four partial tiles are summed
// in LDS and added to dw (f32) with vector atomics.
template <int KS>
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const uint16_t *__restrict__ in,
                                                       const uint16_t *__restrict__ g, float *__restrict__ dw,
                                                       int N, int H, int W, int Cin, int Cout, int pix_per_block) {
    constexpr int R = KS / 2;
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
                if (sy >= 0 && sy < H && sx >= 0 && sx < W) iv = in[((n * H + sy) * W + sx) * Cin + ci0 + (lane & 31)];
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

// 2x2 average pool / its gradient, NHWC
__global__ void k_pool2(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, int N, int H, int W, int C) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    if (i >= (size_t)N * Ho * Wo * C) return;
    const int c = (int)(i % C);
    const size_t q = i / C;
    const int x = (int)(q % Wo), y = (int)((q / Wo) % Ho);
    const size_t n = q / ((size_t)Wo * Ho);
    const size_t b = ((n * H + 2 * y) * W + 2 * x) * C + c;
    const float s = (bf2f(in[b]) + bf2f(in[b + C])) + (bf2f(in[b + (size_t)W * C]) + bf2f(in[b + (size_t)W * C + C]));
    out[i] = f2bf(0.25f * s);
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

// network input from host fields: x[p] = (mask, (rho - 1) / s_rho, ux / s_u, uy / s_u, 0...)
__global__ void k_pack_input(const float *__restrict__ fields4, uint16_t *__restrict__ x, long long P, int C) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    for (int c = 0; c < C; ++c) x[p * C + c] = c < 4 ? f2bf(fields4[p * 4 + c]) : (uint16_t)0;
}

// nearest 2x upsample / its gradient, NHWC; H, W: the coarse size
__global__ void k_up2(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, int N, int H, int W, int C) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * 4 * H * W * C) return;
    const int c = (int)(i % C);
    const size_t q = i / C;
    const int x = (int)(q % (2 * W)), y = (int)((q / (2 * W)) % (2 * H));
    const size_t n = q / ((size_t)4 * W * H);
    out[i] = in[((n * H + y / 2) * W + x / 2) * C + c];
}

__global__ void k_up2_bwd(const uint16_t *__restrict__ gout, uint16_t *__restrict__ gin, int N, int H, int W, int C) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * H * W * C) return;
    const int c = (int)(i % C);
    const size_t q = i / C;
    const int x = (int)(q % W), y = (int)((q / W) % H);
    const size_t n = q / ((size_t)W * H);
    const size_t b = ((n * 2 * H + 2 * y) * 2 * W + 2 * x) * C + c;
    const size_t r = (size_t)2 * W * C;
    const float s = (bf2f(gout[b]) + bf2f(gout[b + C])) + (bf2f(gout[b + r]) + bf2f(gout[b + r + C]));
    gin[i] = f2bf(s);
}

// out[p] = [a[p] (Ca channels), b[p] (Cb channels)]; and back
__global__ void k_concat(const uint16_t *__restrict__ a, const uint16_t *__restrict__ b, uint16_t *__restrict__ out,
                         long long P, int Ca, int Cb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int C = Ca + Cb;
    if (i >= (size_t)P * C) return;
    const int c = (int)(i % C);
    const size_t p = i / C;
    out[i] = c < Ca ? a[p * Ca + c] : b[p * Cb + (c - Ca)];
}

__global__ void k_split(const uint16_t *__restrict__ in, uint16_t *__restrict__ a, uint16_t *__restrict__ b,
                        long long P, int Ca, int Cb, int accumulate_b) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int C = Ca + Cb;
    if (i >= (size_t)P * C) return;
    const int c = (int)(i % C);
    const size_t p = i / C;
    if (c < Ca) {
        a[p * Ca + c] = in[i];
    } else {
        const size_t j = p * Cb + (c - Ca);
        b[j] = accumulate_b ? f2bf(bf2f(b[j]) + bf2f(in[i])) : in[i];
    }
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

// launch the forward convolution with the widest cout tile that divides Cout
template <int KS>
void conv_fwd_ks(hipStream_t s, const uint16_t *in, const uint16_t *w, const float *bias, void *out, int N, int H,
                 int W, int Cin, int Cout, bool relu, bool out_f32) {
    const long long P = (long long)N * H * W;
    const unsigned gx = (unsigned)((P + 127) / 128);
#define LAUNCH(NT, R, F)                                                                                          \
    hipLaunchKernelGGL((k_conv3x3<KS, NT, R, F>), dim3(gx, Cout / (32 * NT)), dim3(256), 0, s, in, w, bias, out, N, \
                       H, W, Cin, Cout)
    if (out_f32) {
        LAUNCH(1, false, true);
    } else if (Cout % 64 == 0) {
        if (relu) LAUNCH(2, true, false);
        else LAUNCH(2, false, false);
    } else {
        if (relu) LAUNCH(1, true, false);
        else LAUNCH(1, false, false);
    }
#undef LAUNCH
    HIPCHK(hipGetLastError());
}

void conv_fwd(hipStream_t s, int ks, const uint16_t *in, const uint16_t *w, const float *bias, void *out, int N, int H,
              int W, int Cin, int Cout, bool relu, bool out_f32) {
    if (ks == 3) conv_fwd_ks<3>(s, in, w, bias, out, N, H, W, Cin, Cout, relu, out_f32);
    else conv_fwd_ks<1>(s, in, w, bias, out, N, H, W, Cin, Cout, relu, out_f32);
}

void conv_wgrad(hipStream_t s, int ks, const uint16_t *in, const uint16_t *g, float *dw, int N, int H, int W, int Cin,
                int Cout) {
    const long long P = (long long)N * H * W;
    const int ppb = 2048;
    const dim3 grid((unsigned)((P + ppb - 1) / ppb), (Cout / 32) * (Cin / 32), ks * ks);
    if (ks == 3)
        hipLaunchKernelGGL(k_conv3x3_wgrad<3>, grid, dim3(256), 0, s, in, g, dw, N, H, W, Cin, Cout, ppb);
    else
        hipLaunchKernelGGL(k_conv3x3_wgrad<1>, grid, dim3(256), 0, s, in, g, dw, N, H, W, Cin, Cout, ppb);
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

// The network.  Layers (c: 3x3 conv + ReLU, u: nearest 2x upsample then 1x1 conv +
// ReLU, p: 2x2 average pool, [a|b]: channel concatenation):
//   e1 = c(c(x: 32))            32 ch, H      p1 = p(e1)
//   e2 = c(c(p1))               64 ch, H/2    p2 = p(e2)
//   e3 = c(c(p2))              128 ch, H/4    p3 = p(e3)
//   e4 = c(c(p3))              256 ch, H/8
//   d3 = c(c([u(e4) | e3]))    128 ch, H/4
//   d2 = c(c([u(d3) | e2]))     64 ch, H/2
//   d1 = c(c([u(d2) | e1]))     32 ch, H
//   out = 1x1 conv(d1)          32 ch (the first 3 used), f32
// Input channels: solid mask, (rho - 1) / s_rho, ux / s_u, uy / s_u, then zeros;
// output: the 64-step change of (rho, ux, uy) in the same units.
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
    Act x, e1a, e1, p1, e2a, e2, p2, e3a, e3, p3, e4a, e4, u3, d3u, k3, d3a, d3, u2, d2u, k2, d2a, d2, u1, d1u, k1,
        d1a, d1;
    std::vector<void *> mem;
    int adam_t = 0;

    void *alloc(size_t bytes) {
        void *p = dalloc<char>(bytes);
        mem.push_back(p);
        return p;
    }
    void make(Act &t, int h, int w, int c) {
        t.N = B;
        t.H = h;
        t.W = w;
        t.C = c;
        t.a = (uint16_t *)alloc(t.n() * 2);
        t.g = (uint16_t *)alloc(t.n() * 2);
    }
    void release() {
        (void)hipSetDevice(dev);
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
        make(u3, h4, w4, 256);
        make(d3u, h4, w4, 128);
        make(k3, h4, w4, 256);
        make(d3a, h4, w4, 128);
        make(d3, h4, w4, 128);
        make(u2, h2, w2, 128);
        make(d2u, h2, w2, 64);
        make(k2, h2, w2, 128);
        make(d2a, h2, w2, 64);
        make(d2, h2, w2, 64);
        make(u1, H, W, 64);
        make(d1u, H, W, 32);
        make(k1, H, W, 64);
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
        fields = (float *)alloc(P * 4 * 4);
        loss = (float *)alloc(4);
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
            const int fan_in = c.ks * c.ks * (i == 0 ? 4 : c.cin);
            const float std_ = (i + 1 == convs.size() ? 0.01f : 1.0f) * (float)std::sqrt(2.0 / fan_in);
            for (int t = 0; t < c.ks * c.ks; ++t)
                for (int co = 0; co < c.cout; ++co)
                    for (int ci = 0; ci < c.cin; ++ci) {
                        const bool live = (i != 0 || ci < 4) && (i + 1 != convs.size() || co < 3);
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

    void fwd_conv(int i, const Act &in, const Act &out) {
        const Conv &c = convs[i];
        conv_fwd(stream, c.ks, in.a, wb + c.wboff, w32 + c.boff, out.a, B, in.H, in.W, c.cin, c.cout, c.relu, false);
    }
    void pool(const Act &in, const Act &out) {
        hipLaunchKernelGGL(k_pool2, dim3(nblk(out.n())), dim3(256), 0, stream, in.a, out.a, B, in.H, in.W, in.C);
    }
    void up(const Act &in, const Act &out) {
        hipLaunchKernelGGL(k_up2, dim3(nblk(out.n())), dim3(256), 0, stream, in.a, out.a, B, in.H, in.W, in.C);
    }
    void cat(const Act &a, const Act &b, const Act &out) {
        hipLaunchKernelGGL(k_concat, dim3(nblk(out.n())), dim3(256), 0, stream, a.a, b.a, out.a, out.P(), a.C, b.C);
    }

    void forward() {
        fwd_conv(0, x, e1a);
        fwd_conv(1, e1a, e1);
        pool(e1, p1);
        fwd_conv(2, p1, e2a);
        fwd_conv(3, e2a, e2);
        pool(e2, p2);
        fwd_conv(4, p2, e3a);
        fwd_conv(5, e3a, e3);
        pool(e3, p3);
        fwd_conv(6, p3, e4a);
        fwd_conv(7, e4a, e4);
        up(e4, u3);
        fwd_conv(8, u3, d3u);
        cat(d3u, e3, k3);
        fwd_conv(9, k3, d3a);
        fwd_conv(10, d3a, d3);
        up(d3, u2);
        fwd_conv(11, u2, d2u);
        cat(d2u, e2, k2);
        fwd_conv(12, k2, d2a);
        fwd_conv(13, d2a, d2);
        up(d2, u1);
        fwd_conv(14, u1, d1u);
        cat(d1u, e1, k1);
        fwd_conv(15, k1, d1a);
        fwd_conv(16, d1a, d1);
        const Conv &h = convs[17];
        conv_fwd(stream, h.ks, d1.a, wb + h.wboff, w32 + h.boff, out32, B, H, W, h.cin, h.cout, false, true);
        HIPCHK(hipGetLastError());
    }

    // gradients of conv i: weights and bias into g32; input gradient into gin (if any)
    void bwd_conv(int i, const Act &in, const uint16_t *gout_, const Act *gin) {
        const Conv &c = convs[i];
        conv_wgrad(stream, c.ks, in.a, gout_, g32 + c.woff, B, in.H, in.W, c.cin, c.cout);
        hipLaunchKernelGGL(k_bias_grad, dim3(256, (c.cout + 63) / 64), dim3(256), 0, stream, gout_, g32 + c.boff,
                           in.P(), c.cout);
        if (gin)
            conv_fwd(stream, c.ks, gout_, wt + c.wboff, nullptr, gin->g, B, in.H, in.W, c.cout, c.cin, false, false);
    }
    void relu_bwd(const Act &t) {
        hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(t.n())), dim3(256), 0, stream, t.g, t.a, t.n());
    }
    void split(const Act &k, const Act &a, const Act &b) {
        hipLaunchKernelGGL(k_split, dim3(nblk(k.n())), dim3(256), 0, stream, k.g, a.g, b.g, k.P(), a.C, b.C, 0);
    }
    void up_bwd(const Act &fine, const Act &coarse) {
        hipLaunchKernelGGL(k_up2_bwd, dim3(nblk(coarse.n())), dim3(256), 0, stream, fine.g, coarse.g, B, coarse.H,
                           coarse.W, coarse.C);
    }
    void pool_bwd(const Act &coarse, const Act &fine) {  // accumulates into fine.g
        hipLaunchKernelGGL(k_pool2_bwd, dim3(nblk(fine.n())), dim3(256), 0, stream, coarse.g, fine.g, B, fine.H,
                           fine.W, fine.C, 1);
    }

    void backward() {
        HIPCHK(hipMemsetAsync(g32, 0, nparam * 4, stream));
        bwd_conv(17, d1, gout, &d1);
        relu_bwd(d1);
        bwd_conv(16, d1a, d1.g, &d1a);
        relu_bwd(d1a);
        bwd_conv(15, k1, d1a.g, &k1);
        split(k1, d1u, e1);
        relu_bwd(d1u);
        bwd_conv(14, u1, d1u.g, &u1);
        up_bwd(u1, d2);
        relu_bwd(d2);
        bwd_conv(13, d2a, d2.g, &d2a);
        relu_bwd(d2a);
        bwd_conv(12, k2, d2a.g, &k2);
        split(k2, d2u, e2);
        relu_bwd(d2u);
        bwd_conv(11, u2, d2u.g, &u2);
        up_bwd(u2, d3);
        relu_bwd(d3);
        bwd_conv(10, d3a, d3.g, &d3a);
        relu_bwd(d3a);
        bwd_conv(9, k3, d3a.g, &k3);
        split(k3, d3u, e3);
        relu_bwd(d3u);
        bwd_conv(8, u3, d3u.g, &u3);
        up_bwd(u3, e4);
        relu_bwd(e4);
        bwd_conv(7, e4a, e4.g, &e4a);
        relu_bwd(e4a);
        bwd_conv(6, p3, e4a.g, &p3);
        pool_bwd(p3, e3);
        relu_bwd(e3);
        bwd_conv(5, e3a, e3.g, &e3a);
        relu_bwd(e3a);
        bwd_conv(4, p2, e3a.g, &p2);
        pool_bwd(p2, e2);
        relu_bwd(e2);
        bwd_conv(3, e2a, e2.g, &e2a);
        relu_bwd(e2a);
        bwd_conv(2, p1, e2a.g, &p1);
        pool_bwd(p1, e1);
        relu_bwd(e1);
        bwd_conv(1, e1a, e1.g, &e1a);
        relu_bwd(e1a);
        bwd_conv(0, x, e1a.g, nullptr);
        HIPCHK(hipGetLastError());
    }

    // host inputs -> device; returns the loss
    float loss_and_grad(const float *hx, const float *hy, const float *hwm) {
        const size_t P = (size_t)B * H * W;
        HIPCHK(hipMemcpyAsync(fields, hx, P * 16, hipMemcpyHostToDevice, stream));
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
void unet_forward_on(lbm_unet *n, hipStream_t s) {
    hipStream_t saved = n->stream;
    n->stream = s;
    try {
        n->forward();
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
        HIPCHK(hipMemcpyAsync(params, n->w32, n->nparam * 4, hipMemcpyDeviceToHost, n->stream));
        HIPCHK(hipStreamSynchronize(n->stream));
    });
}

int lbm_unet_set_params(lbm_unet *n, const float *params) {
    return uguard([&] {
        ucheck(n && params, "null argument");
        HIPCHK(hipSetDevice(n->dev));
        HIPCHK(hipMemcpyAsync(n->w32, params, n->nparam * 4, hipMemcpyHostToDevice, n->stream));
        n->refresh_bf16();
        HIPCHK(hipStreamSynchronize(n->stream));
    });
}

int lbm_unet_forward(lbm_unet *n, const float *x, float *y) {
    return uguard([&] {
        ucheck(n && x && y, "null argument");
        HIPCHK(hipSetDevice(n->dev));
        const size_t P = (size_t)n->B * n->H * n->W;
        HIPCHK(hipMemcpyAsync(n->fields, x, P * 16, hipMemcpyHostToDevice, n->stream));
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
// rounded to bf16 on the device).  mode 0: out = conv(in, w) + bias (no ReLU);
// mode 1: out = input gradient for output gradient `in`; mode 2: out = weight
// gradient dw[tap][co][ci] for input `in` and output gradient `g`.
int lbm_conv3x3_test(int mode, int N, int H, int W, int Cin, int Cout, const float *in, const float *g,
                     const float *w, const float *bias, float *out, int device) {
    return uguard([&] {
        ucheck(mode >= 0 && mode <= 2, "mode");
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
            if (mode == 0) {
                float *db = nullptr;
                if (bias) {
                    db = (float *)alloc(Cout * sizeof(float));
                    HIPCHK(hipMemcpy(db, bias, Cout * sizeof(float), hipMemcpyHostToDevice));
                }
                conv_fwd(0, 3, din, dw, db, dout, N, H, W, Cin, Cout, false, true);
                HIPCHK(hipMemcpy(out, dout, P * Cout * sizeof(float), hipMemcpyDeviceToHost));
            } else if (mode == 1) {
                uint16_t *dwt = (uint16_t *)alloc(nw * 2);
                hipLaunchKernelGGL(k_flip_transpose, dim3(nblk(nw)), dim3(256), 0, 0, dw, dwt, Cin, Cout, 9);
                conv_fwd(0, 3, din, dwt, nullptr, dout, N, H, W, Cout, Cin, false, true);
                HIPCHK(hipMemcpy(out, dout, P * Cin * sizeof(float), hipMemcpyDeviceToHost));
            } else {
                uint16_t *dg = (uint16_t *)alloc(P * Cout * 2);
                HIPCHK(hipMemcpy(d32, g, P * Cout * sizeof(float), hipMemcpyHostToDevice));
                hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(P * Cout)), dim3(256), 0, 0, d32, dg, P * Cout);
                HIPCHK(hipMemset(dout, 0, nw * sizeof(float)));
                conv_wgrad(0, 3, din, dg, dout, N, H, W, Cin, Cout);
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
