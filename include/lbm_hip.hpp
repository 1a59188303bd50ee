/* C++ host mirror of the lbm-wgpu crate's LBM type (src/lbm.rs:32-98) over the
 * C ABI of lbm_hip.h: the same method names and argument meaning, with RAII
 * ownership of the grid handle and exceptions in place of the reference's
 * panics (lbm::Error carries the LBM_E* code and lbm_last_error's text).
 *
 *   lbm::LBM g(1.0f / 0.56f, 400, 160);       // LBM::new(&driver, omega, x, y), lbm.rs:726
 *   g.set_mask(mask);                          // or g.draw_shape(pairs), lbm.rs:1337
 *   g.custom_speed(0.1f);                      // lbm.rs:1090
 *   g.iterate(15);                             // lbm.rs:1065
 *   std::vector<float> curl = g.calculate_summary();
 *
 * Header-only; link with liblbm_hip.so.  One grid per object; calls on one
 * object must not overlap (one HIP stream per grid). */
#ifndef LBM_HIP_HPP
#define LBM_HIP_HPP

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "lbm_hip.h"

namespace lbm {

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

/* order of the reference enums (src/lbm.rs:10-16, 20-24) */
enum class SummaryStat : int { Curl = LBM_SUMMARY_CURL, Ux, Uy, Rho, Speed };
enum class ColorMap : int { Inferno = LBM_COLORMAP_INFERNO, Viridis, Jet };

struct Options {
    float u_init = 0.1f;  // LBM::new's initial speed (lbm.rs:740-745)
    int precision = LBM_PRECISION_F32;
    int boundary = LBM_BOUNDARY_REFERENCE;
    int device = 0;
};

class LBM {
  public:
    /* LBM::new (lbm.rs:726): walls on rows 0 and y-1, uniform equilibrium */
    LBM(float omega, uint32_t x, uint32_t y, const Options &opt = Options()) : x_(x), y_(y) {
        lbm_config cfg;
        check(lbm_config_default(&cfg, x, y), nullptr, "lbm_config_default");
        cfg.omega = omega;
        cfg.u_init = opt.u_init;
        cfg.precision = opt.precision;
        cfg.boundary = opt.boundary;
        cfg.device = opt.device;
        check(lbm_create(&cfg, &g_), nullptr, "lbm_create");
    }
    ~LBM() {
        if (g_) lbm_destroy(g_);
    }
    LBM(const LBM &) = delete;
    LBM &operator=(const LBM &) = delete;
    LBM(LBM &&o) noexcept : g_(std::exchange(o.g_, nullptr)), x_(o.x_), y_(o.y_), stat_(o.stat_) {}
    LBM &operator=(LBM &&o) noexcept {
        if (this != &o) {
            if (g_) lbm_destroy(g_);
            g_ = std::exchange(o.g_, nullptr);
            x_ = o.x_;
            y_ = o.y_;
            stat_ = o.stat_;
        }
        return *this;
    }

    uint32_t x() const { return x_; }
    uint32_t y() const { return y_; }
    lbm_grid *handle() const { return g_; }

    /* ---- the reference's surface ------------------------------------------------ */
    /* iterate (lbm.rs:1065-1074): n steps; the summary field is read on demand */
    void iterate(uint32_t n) { step(n); }
    void update_omega_buffer(float omega) { call(lbm_set_omega(g_, omega)); }  // lbm.rs:1358
    void update_viscosity(float nu) { call(lbm_set_viscosity(g_, nu)); }       // lib.rs:181-186
    void reset_barrier() { call(lbm_reset_barrier(g_)); }                      // lbm.rs:1362
    void reset_to_equilibrium() { custom_speed(0.1f); }                         // lbm.rs:1076
    void custom_speed(float ux) { call(lbm_reset_equilibrium(g_, 1.0f, ux, 0.0f)); }  // lbm.rs:1090
    void single_cell(int k) { call(lbm_single_cell(g_, k)); }                  // lbm.rs:1502
    void set_summary(SummaryStat s) { stat_ = s; }                              // lbm.rs:1061
    /* draw_shape (lbm.rs:1337): get_points_vector's [x + y * nx, 1|0, ...] pairs */
    void draw_shape(const std::vector<uint32_t> &pairs) { call(lbm_paint(g_, pairs.data(), pairs.size())); }
    uint64_t get_compute_num() const {  // lbm.rs:1166
        uint64_t n = 0;
        call(lbm_step_count(g_, &n));
        return n;
    }
    std::vector<float> calculate_summary() { return calculate_summary(stat_); }  // lbm.rs:1051
    std::vector<float> calculate_summary(SummaryStat s) {
        std::vector<float> out((size_t)x_ * y_);
        call(lbm_summary(g_, (int)s, out.data()));
        return out;
    }

    /* ---- beyond the reference ---------------------------------------------------- */
    void step(uint32_t n) { call(lbm_step(g_, n)); }
    void sync() { call(lbm_sync(g_)); }
    float step_timed(uint32_t n) {  // milliseconds on the grid's stream
        float ms = 0.0f;
        call(lbm_step_timed(g_, n, &ms));
        return ms;
    }
    int autotune(uint32_t steps, std::vector<float> *ms_per_variant = nullptr) {
        int best = 0;
        std::vector<float> ms(LBM_NUM_VARIANTS);
        call(lbm_autotune(g_, steps, &best, ms.data()));
        if (ms_per_variant) *ms_per_variant = ms;
        return best;
    }
    // autotune through a JSON cache file; *hit tells whether the variant came from the file
    int autotune_cached(const std::string &cache_path, uint32_t steps, bool *hit = nullptr) {
        int best = 0, h = 0;
        call(lbm_autotune_cached(g_, cache_path.c_str(), steps, &best, &h));
        if (hit) *hit = h != 0;
        return best;
    }
    void set_variant(int v) { call(lbm_set_variant(g_, v)); }
    int variant() const {
        int v = 0;
        call(lbm_get_variant(g_, &v));
        return v;
    }
    void set_mask(const std::vector<uint8_t> &solid) {
        if (solid.size() != (size_t)x_ * y_) throw Error(LBM_EINVAL, "set_mask: need nx * ny bytes");
        call(lbm_set_mask(g_, solid.data()));
    }
    std::vector<uint8_t> mask() {
        std::vector<uint8_t> m((size_t)x_ * y_);
        call(lbm_get_mask(g_, m.data()));
        return m;
    }
    /* colour plane of a summary (3 floats per cell, solid cells black) */
    std::vector<float> render(SummaryStat s, ColorMap c) {
        std::vector<float> rgb((size_t)3 * x_ * y_);
        call(lbm_render(g_, (int)s, (int)c, rgb.data()));
        return rgb;
    }
    std::array<double, 3> conservation() {  // FP64 sums of rho, jx, jy over fluid cells
        std::array<double, 3> s{};
        call(lbm_conservation(g_, s.data()));
        return s;
    }
    std::vector<float> read_f() {  // S_n as 9 dense nx * ny planes
        std::vector<float> f((size_t)9 * x_ * y_);
        call(lbm_read_f(g_, f.data()));
        return f;
    }

  private:
    static void check(int rc, const lbm_grid *g, const char *what) {
        if (rc != LBM_OK) {
            const char *msg = lbm_last_error(g);
            throw Error(rc, std::string(what) + ": " + (msg ? msg : ""));
        }
    }
    void call(int rc) const {
        if (rc != LBM_OK) {
            const char *msg = lbm_last_error(g_);
            throw Error(rc, msg ? msg : "lbm error");
        }
    }

    lbm_grid *g_ = nullptr;
    uint32_t x_ = 0, y_ = 0;
    SummaryStat stat_ = SummaryStat::Curl;
};

/* seeded circle mask (lbm_gen_circles): walls on rows 0 and ny-1, then n circles */
inline std::vector<uint8_t> gen_circles(uint32_t nx, uint32_t ny, uint64_t seed, uint32_t n, uint32_t r_min,
                                        uint32_t r_max) {
    std::vector<uint8_t> m((size_t)nx * ny, 0);
    for (uint32_t x = 0; x < nx; ++x) m[x] = m[(size_t)(ny - 1) * nx + x] = 1;
    if (lbm_gen_circles(m.data(), nx, ny, seed, n, r_min, r_max) != LBM_OK)
        throw Error(LBM_EINVAL, "gen_circles: invalid arguments");
    return m;
}

}  // namespace lbm

#endif
