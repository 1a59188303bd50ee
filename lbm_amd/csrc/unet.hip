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
template <int NT, bool RELU, bool OUT_F32>
__global__ __launch_bounds__(256) void k_conv3x3(const uint16_t *__restrict__ in, const uint16_t *__restrict__ w,
                                                 const float *__restrict__ bias, void *__restrict__ out, int N, int H,
                                                 int W, int Cin, int Cout) {
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
    for (int tap = 0; tap < 9; ++tap) {
        const int sy = ya + tap / 3 - 1, sx = xa + tap % 3 - 1;
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
__global__ __launch_bounds__(256) void k_conv3x3_wgrad(const uint16_t *__restrict__ in,
                                                       const uint16_t *__restrict__ g, float *__restrict__ dw,
                                                       int N, int H, int W, int Cin, int Cout, int pix_per_block) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long P = (long long)N * H * W;
    const int tiles_ci = Cin / 32;
    const int co0 = (blockIdx.y / tiles_ci) * 32, ci0 = (blockIdx.y % tiles_ci) * 32;
    const int tap = blockIdx.z;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
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

// wt[tap'][ci][co] = w[8 - tap'][co][ci]: weights of the input-gradient convolution
__global__ void k_flip_transpose(const uint16_t *__restrict__ w, uint16_t *__restrict__ wt, int Cin, int Cout) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)9 * Cin * Cout;
    if (i >= n) return;
    const int co = (int)(i % Cout);
    const int ci = (int)((i / Cout) % Cin);
    const int tp = (int)(i / ((size_t)Cout * Cin));
    wt[i] = w[((size_t)(8 - tp) * Cout + co) * Cin + ci];
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

__global__ void k_pool2_bwd(const uint16_t *__restrict__ gout, uint16_t *__restrict__ gin, int N, int H, int W,
                            int C) {  // H, W: the input (fine) size
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)N * H * W * C) return;
    const int c = (int)(i % C);
    const size_t q = i / C;
    const int x = (int)(q % W), y = (int)((q / W) % H);
    const size_t n = q / ((size_t)W * H);
    gin[i] = f2bf(0.25f * bf2f(gout[((n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c]));
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
void conv_fwd(hipStream_t s, const uint16_t *in, const uint16_t *w, const float *bias, void *out, int N, int H, int W,
              int Cin, int Cout, bool relu, bool out_f32) {
    const long long P = (long long)N * H * W;
    const unsigned gx = (unsigned)((P + 127) / 128);
#define LAUNCH(NT, R, F)                                                                                      \
    hipLaunchKernelGGL((k_conv3x3<NT, R, F>), dim3(gx, Cout / (32 * NT)), dim3(256), 0, s, in, w, bias, out, N, H, \
                       W, Cin, Cout)
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

void conv_wgrad(hipStream_t s, const uint16_t *in, const uint16_t *g, float *dw, int N, int H, int W, int Cin,
                int Cout) {
    const long long P = (long long)N * H * W;
    const int ppb = 2048;
    hipLaunchKernelGGL(k_conv3x3_wgrad, dim3((unsigned)((P + ppb - 1) / ppb), (Cout / 32) * (Cin / 32), 9),
                       dim3(256), 0, s, in, g, dw, N, H, W, Cin, Cout, ppb);
    HIPCHK(hipGetLastError());
}

}  // namespace

extern "C" {

const char *lbm_unet_last_error(void) { return g_unet_error.c_str(); }

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
                conv_fwd(0, din, dw, db, dout, N, H, W, Cin, Cout, false, true);
                HIPCHK(hipMemcpy(out, dout, P * Cout * sizeof(float), hipMemcpyDeviceToHost));
            } else if (mode == 1) {
                uint16_t *dwt = (uint16_t *)alloc(nw * 2);
                hipLaunchKernelGGL(k_flip_transpose, dim3(nblk(nw)), dim3(256), 0, 0, dw, dwt, Cin, Cout);
                conv_fwd(0, din, dwt, nullptr, dout, N, H, W, Cout, Cin, false, true);
                HIPCHK(hipMemcpy(out, dout, P * Cin * sizeof(float), hipMemcpyDeviceToHost));
            } else {
                uint16_t *dg = (uint16_t *)alloc(P * Cout * 2);
                HIPCHK(hipMemcpy(d32, g, P * Cout * sizeof(float), hipMemcpyHostToDevice));
                hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(P * Cout)), dim3(256), 0, 0, d32, dg, P * Cout);
                HIPCHK(hipMemset(dout, 0, nw * sizeof(float)));
                conv_wgrad(0, din, dg, dout, N, H, W, Cin, Cout);
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
