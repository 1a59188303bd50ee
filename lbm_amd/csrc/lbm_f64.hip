// FP64 D2Q9 solver on the GPU, for measuring the error of the f32 and
// FP16-shifted grids at sizes a CPU FP64 solver cannot reach in time.
//
// It is the textbook scheme (BGK collide of every cell with
// feq_k = w_k rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u), then a pull stream
// with half-way bounce-back) with the reference step's boundary rules
// (lbm-wgpu rewritten_shaders/stream/{e_w,n_s,ne_sw,se_nw}_stream.wgsl:31-62):
// a cell is streamed only if it is fluid, x != 0 and y < ny - 1; column 0
// holds the inflow equilibrium feq(1, (u_in, 0)); an upstream cell that is
// solid or outside the grid bounces back; the outlet's west-moving upstream is
// the next row's column 0 (reference mode) or x clamped to nx - 1
// (zero-gradient).  Every operation is written in the order of the textbook
// solver the tests compare it with, in IEEE double without contraction, so the
// two agree bit for bit.  No reference counterpart: the reference is f32 only.
//
// Layout: nine dense planes per state (element (k, x, y) at k * nx * ny +
// y * nx + x), a pre-collision state f and a post-collision scratch g.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"

namespace {

__constant__ int kCx[9] = {-1, 0, 1, -1, 0, 1, -1, 0, 1};
__constant__ int kCy[9] = {-1, -1, -1, 0, 0, 0, 1, 1, 1};
__constant__ double kW[9] = {1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 9, 4.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 9, 1.0 / 36};

__device__ __forceinline__ void feq64(double rho, double ux, double uy, double out[9]) {
    const double uu = ux * ux + uy * uy;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const double cu = kCx[k] * ux + kCy[k] * uy;
        out[k] = kW[k] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
    }
}

__global__ void k64_collide(const double *__restrict__ f, double *__restrict__ g, size_t n, double om) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double fl[9], rho = 0.0, jx = 0.0, jy = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        fl[k] = f[k * n + i];
        rho += fl[k];
        jx += kCx[k] * fl[k];
        jy += kCy[k] * fl[k];
    }
    double e[9];
    feq64(rho, jx / rho, jy / rho, e);
#pragma unroll
    for (int k = 0; k < 9; ++k) g[k * n + i] = fl[k] + om * (e[k] - fl[k]);
}

__global__ void k64_stream(const double *__restrict__ g, double *__restrict__ f, const uint8_t *__restrict__ solid,
                           int X, int Y, int mode, double u_in) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= X) return;
    const size_t n = (size_t)X * Y;
    const size_t i = (size_t)y * X + x;
    if (x == 0) {
        double in[9];
        feq64(1.0, u_in, 0.0, in);
#pragma unroll
        for (int k = 0; k < 9; ++k) f[k * n + i] = in[k];
        return;
    }
    if (solid[i] || y >= Y - 1) return;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        if (k == 4) {
            f[4 * n + i] = g[4 * n + i];
            continue;
        }
        int ux = x - kCx[k], uy = y - kCy[k];
        if (ux >= X) {
            if (mode == LBM_BOUNDARY_REFERENCE) {
                ux = 0;
                uy += 1;
            } else {
                ux = X - 1;
            }
        }
        bool bounce = uy < 0 || uy >= Y;
        size_t u = 0;
        if (!bounce) {
            u = (size_t)uy * X + ux;
            bounce = solid[u] != 0;
        }
        f[k * n + i] = bounce ? g[(8 - k) * n + i] : g[k * n + u];
    }
}

// rho, u = j / rho (uy north-positive, as the reference's uy field)
__global__ void k64_macro(const double *__restrict__ f, double *__restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r = 0.0, jx = 0.0, jy = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const double v = f[k * n + i];
        r += v;
        jx += kCx[k] * v;
        jy += kCy[k] * v;
    }
    out[i] = r;
    out[n + i] = jx / r;
    out[2 * n + i] = -jy / r;
}

struct HErr {
    hipError_t e;
    const char *what;
};
struct AErr {
    const char *msg;
};
#define H64(x)                                  \
    do {                                        \
        hipError_t e_ = (x);                    \
        if (e_ != hipSuccess) {                 \
            (void)hipGetLastError();            \
            throw HErr{e_, #x};                 \
        }                                       \
    } while (0)

thread_local std::string g_f64_error;

}  // namespace

struct lbm_f64 {
    int dev = 0;
    hipStream_t stream = nullptr;
    int X = 0, Y = 0, mode = 0;
    double omega = 1.0, u_in = 0.0;
    double *f = nullptr, *g = nullptr, *scratch = nullptr;
    uint8_t *solid = nullptr;
    uint64_t steps = 0;

    size_t n() const { return (size_t)X * Y; }
    void release() {
        (void)hipSetDevice(dev);
        for (void *p : {(void *)f, (void *)g, (void *)scratch, (void *)solid})
            if (p) (void)hipFree(p);
        f = g = scratch = nullptr;
        solid = nullptr;
        if (stream) (void)hipStreamDestroy(stream);
        stream = nullptr;
    }
};

namespace {

template <class F>
int fguard(F &&fn) {
    try {
        fn();
        return LBM_OK;
    } catch (const HErr &e) {
        g_f64_error = std::string(e.what) + ": " + hipGetErrorString(e.e);
        return e.e == hipErrorOutOfMemory ? LBM_ENOMEM : LBM_EHIP;
    } catch (const AErr &e) {
        g_f64_error = e.msg;
        return LBM_EINVAL;
    } catch (const std::bad_alloc &) {
        g_f64_error = "host allocation failed";
        return LBM_ENOMEM;
    } catch (...) {
        g_f64_error = "unexpected exception";
        return LBM_EHIP;
    }
}

void fcheck(bool ok, const char *msg) {
    if (!ok) throw AErr{msg};
}

}  // namespace

extern "C" {

const char *lbm_f64_last_error(void) { return g_f64_error.c_str(); }

int lbm_f64_create(uint32_t nx, uint32_t ny, double omega, double u_in, int32_t boundary, int32_t device,
                   lbm_f64 **out) {
    return fguard([&] {
        fcheck(out != nullptr, "null output");
        *out = nullptr;
        fcheck(nx >= 3 && ny >= 3 && nx <= 65536 && ny <= 65535, "grid size out of range");
        fcheck(boundary == LBM_BOUNDARY_REFERENCE || boundary == LBM_BOUNDARY_ZERO_GRADIENT, "unknown boundary");
        H64(hipSetDevice(device));
        lbm_f64 *h = new lbm_f64();
        h->dev = device;
        h->X = (int)nx;
        h->Y = (int)ny;
        h->mode = boundary;
        h->omega = omega;
        h->u_in = u_in;
        try {
            H64(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
            const size_t n = h->n();
            H64(hipMalloc(&h->f, 9 * n * sizeof(double)));
            H64(hipMalloc(&h->g, 9 * n * sizeof(double)));
            H64(hipMalloc(&h->solid, n));
            // walls on the first and last rows; everything at feq(1, (u_in, 0))
            std::vector<uint8_t> s(n, 0);
            for (uint32_t x = 0; x < nx; ++x) {
                s[x] = 1;
                s[(size_t)(ny - 1) * nx + x] = 1;
            }
            H64(hipMemcpy(h->solid, s.data(), n, hipMemcpyHostToDevice));
            const int cx[9] = {-1, 0, 1, -1, 0, 1, -1, 0, 1};
            const double w[9] = {1.0 / 36, 1.0 / 9, 1.0 / 36, 1.0 / 9, 4.0 / 9, 1.0 / 9, 1.0 / 36, 1.0 / 9, 1.0 / 36};
            std::vector<double> plane(n);
            for (int k = 0; k < 9; ++k) {
                const double uu = u_in * u_in + 0.0 * 0.0;
                const double cu = cx[k] * u_in + 0.0;
                const double e = w[k] * 1.0 * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
                std::fill(plane.begin(), plane.end(), e);
                H64(hipMemcpy(h->f + k * n, plane.data(), n * sizeof(double), hipMemcpyHostToDevice));
            }
        } catch (...) {
            h->release();
            delete h;
            throw;
        }
        *out = h;
    });
}

int lbm_f64_destroy(lbm_f64 *h) {
    if (!h) return LBM_EINVAL;
    h->release();
    delete h;
    return LBM_OK;
}

int lbm_f64_set_mask(lbm_f64 *h, const uint8_t *solid) {
    return fguard([&] {
        fcheck(h && solid, "null argument");
        H64(hipSetDevice(h->dev));
        std::vector<uint8_t> s(solid, solid + h->n());
        for (auto &v : s) v = v == 1 ? 1 : 0;
        H64(hipMemcpyAsync(h->solid, s.data(), h->n(), hipMemcpyHostToDevice, h->stream));
        H64(hipStreamSynchronize(h->stream));
    });
}

int lbm_f64_step(lbm_f64 *h, uint32_t n) {
    return fguard([&] {
        fcheck(h != nullptr, "null handle");
        H64(hipSetDevice(h->dev));
        const size_t cells = h->n();
        const dim3 sg((h->X + 255) / 256, h->Y);
        for (uint32_t s = 0; s < n; ++s) {
            hipLaunchKernelGGL(k64_collide, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, h->stream, h->f,
                               h->g, cells, h->omega);
            hipLaunchKernelGGL(k64_stream, sg, dim3(256), 0, h->stream, h->g, h->f, h->solid, h->X, h->Y, h->mode,
                               h->u_in);
        }
        H64(hipGetLastError());
        h->steps += n;
    });
}

int lbm_f64_sync(lbm_f64 *h) {
    return fguard([&] {
        fcheck(h != nullptr, "null handle");
        H64(hipStreamSynchronize(h->stream));
    });
}

int lbm_f64_macro(lbm_f64 *h, double *rho, double *ux, double *uy) {
    return fguard([&] {
        fcheck(h != nullptr, "null handle");
        H64(hipSetDevice(h->dev));
        const size_t n = h->n();
        if (!h->scratch) H64(hipMalloc(&h->scratch, 3 * n * sizeof(double)));
        hipLaunchKernelGGL(k64_macro, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream, h->f, h->scratch,
                           n);
        H64(hipGetLastError());
        double *dst[3] = {rho, ux, uy};
        for (int c = 0; c < 3; ++c)
            if (dst[c]) H64(hipMemcpyAsync(dst[c], h->scratch + c * n, n * sizeof(double), hipMemcpyDeviceToHost,
                                           h->stream));
        H64(hipStreamSynchronize(h->stream));
    });
}

int lbm_f64_read_f(lbm_f64 *h, double *f9) {
    return fguard([&] {
        fcheck(h && f9, "null argument");
        H64(hipSetDevice(h->dev));
        H64(hipMemcpyAsync(f9, h->f, 9 * h->n() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        H64(hipStreamSynchronize(h->stream));
    });
}

int lbm_f64_step_count(const lbm_f64 *h, uint64_t *n) {
    if (!h || !n) return LBM_EINVAL;
    *n = h->steps;
    return LBM_OK;
}

}  // extern "C"
