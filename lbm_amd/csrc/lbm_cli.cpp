// Command-line driver over the C++ mirror of the reference's LBM type
// (include/lbm_hip.hpp): the event loop of lbm-wgpu src/lib.rs:108-204 without
// a window.  Builds a grid, paints obstacles, runs frames of steps and prints
// one JSON line per frame (step count, MLUPS of the frame, FP64 mass and
// momentum), optionally dumping the final state S_n as raw f32 planes.
//
//   lbm_cli --nx 400 --ny 160 --viscosity 0.02 --speed 0.1 --circle 80,80,16 \
//           --frames 4 --steps-per-frame 500
//
// Obstacles: --circle x,y,r (repeatable) or --circles N (seeded, lbm_gen_circles),
// and/or a barrier preset (--barrier, include/lbm_shapes.hpp); walls on rows 0
// and ny-1 always.  omega = 1 / (3 nu + 1/2) (lib.rs:181-186).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lbm_hip.hpp"
#include "lbm_shapes.hpp"

namespace {

struct Args {
    uint32_t nx = 400, ny = 160;
    float viscosity = 0.02f, speed = 0.1f;
    std::vector<std::array<int, 3>> circles;
    uint32_t n_circles = 0, r_min = 8, r_max = 64;
    uint64_t seed = 2024;
    int precision = LBM_PRECISION_F32, boundary = LBM_BOUNDARY_REFERENCE;
    uint32_t frames = 1, steps_per_frame = 15;  // COMPUTE_PER_RENDER (lib.rs:17)
    int variant = -1;                            // -1: autotune
    std::string barrier;                         // curl, chaos or welcome (lbm.rs:1367-1480)
    std::string dump;
    std::string tune_cache;                      // JSON autotune cache file
};

void usage(FILE *to) {
    std::fprintf(to,
                 "usage: lbm_cli [--nx N] [--ny N] [--viscosity NU] [--speed U] [--circle X,Y,R]...\n"
                 "               [--circles N [--seed S] [--r-min R] [--r-max R]] [--precision f32|f16s|f16e]\n"
                 "               [--boundary reference|zero-gradient] [--frames F] [--steps-per-frame K]\n"
                 "               [--barrier curl|chaos|welcome] [--variant V] [--tune-cache PATH] [--dump-f PATH]\n");
}

bool parse(int argc, char **argv, Args &a) {
    for (int i = 1; i < argc; ++i) {
        const std::string k = argv[i];
        auto val = [&]() -> const char * {
            if (i + 1 >= argc) throw lbm::Error(LBM_EINVAL, "missing value for " + k);
            return argv[++i];
        };
        if (k == "--help" || k == "-h") {
            usage(stdout);
            std::exit(0);
        } else if (k == "--nx") a.nx = (uint32_t)std::strtoul(val(), nullptr, 10);
        else if (k == "--ny") a.ny = (uint32_t)std::strtoul(val(), nullptr, 10);
        else if (k == "--viscosity") a.viscosity = std::strtof(val(), nullptr);
        else if (k == "--speed") a.speed = std::strtof(val(), nullptr);
        else if (k == "--circle") {
            std::array<int, 3> c{};
            if (std::sscanf(val(), "%d,%d,%d", &c[0], &c[1], &c[2]) != 3)
                throw lbm::Error(LBM_EINVAL, "--circle takes x,y,r");
            a.circles.push_back(c);
        } else if (k == "--circles") a.n_circles = (uint32_t)std::strtoul(val(), nullptr, 10);
        else if (k == "--seed") a.seed = std::strtoull(val(), nullptr, 10);
        else if (k == "--r-min") a.r_min = (uint32_t)std::strtoul(val(), nullptr, 10);
        else if (k == "--r-max") a.r_max = (uint32_t)std::strtoul(val(), nullptr, 10);
        else if (k == "--precision") {
            const std::string v = val();
            if (v == "f32") a.precision = LBM_PRECISION_F32;
            else if (v == "f16s") a.precision = LBM_PRECISION_F16S;
            else if (v == "f16e") a.precision = LBM_PRECISION_F16E;
            else throw lbm::Error(LBM_EINVAL, "--precision is f32, f16s or f16e");
        } else if (k == "--boundary") {
            const std::string v = val();
            if (v != "reference" && v != "zero-gradient")
                throw lbm::Error(LBM_EINVAL, "--boundary is reference or zero-gradient");
            a.boundary = v == "reference" ? LBM_BOUNDARY_REFERENCE : LBM_BOUNDARY_ZERO_GRADIENT;
        } else if (k == "--frames") a.frames = (uint32_t)std::strtoul(val(), nullptr, 10);
        else if (k == "--steps-per-frame") a.steps_per_frame = (uint32_t)std::strtoul(val(), nullptr, 10);
        else if (k == "--variant") a.variant = std::atoi(val());
        else if (k == "--dump-f") a.dump = val();
        else if (k == "--tune-cache") a.tune_cache = val();
        else if (k == "--barrier") {
            a.barrier = val();
            if (a.barrier != "curl" && a.barrier != "chaos" && a.barrier != "welcome")
                throw lbm::Error(LBM_EINVAL, "--barrier is curl, chaos or welcome");
        } else {
            std::fprintf(stderr, "unknown option %s\n", k.c_str());
            usage(stderr);
            return false;
        }
    }
    return true;
}

std::vector<uint8_t> build_mask(const Args &a) {
    if (a.n_circles) return lbm::gen_circles(a.nx, a.ny, a.seed, a.n_circles, a.r_min, a.r_max);
    std::vector<uint8_t> m((size_t)a.nx * a.ny, 0);
    for (uint32_t x = 0; x < a.nx; ++x) m[x] = m[(size_t)(a.ny - 1) * a.nx + x] = 1;
    for (const auto &c : a.circles)
        for (uint32_t y = 0; y < a.ny; ++y)
            for (uint32_t x = 0; x < a.nx; ++x) {
                const long long dx = (long long)x - c[0], dy = (long long)y - c[1];
                if (dx * dx + dy * dy <= (long long)c[2] * c[2]) m[(size_t)y * a.nx + x] = 1;
            }
    return m;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        Args a;
        if (!parse(argc, argv, a)) return 2;
        lbm::Options opt;
        opt.u_init = a.speed;
        opt.precision = a.precision;
        opt.boundary = a.boundary;
        lbm::LBM g(1.0f / (3.0f * a.viscosity + 0.5f), a.nx, a.ny, opt);
        g.set_mask(build_mask(a));
        if (a.barrier == "curl") lbm::shapes::draw_shape(g, lbm::shapes::curl_barrier(a.nx, a.ny));
        else if (a.barrier == "chaos") lbm::shapes::draw_shape(g, lbm::shapes::chaos_barrier(a.nx, a.ny));
        else if (a.barrier == "welcome") lbm::shapes::draw_shape(g, lbm::shapes::welcome_barrier(a.nx, a.ny));
        g.custom_speed(a.speed);
        bool hit = false;
        const int variant = a.variant >= 0             ? (g.set_variant(a.variant), a.variant)
                            : !a.tune_cache.empty() ? g.autotune_cached(a.tune_cache, 60, &hit)
                                                    : g.autotune(60);
        if (!a.tune_cache.empty() && a.variant < 0)
            std::fprintf(stderr, "lbm_cli: variant %d %s %s\n", variant, hit ? "read from" : "tuned and stored in",
                         a.tune_cache.c_str());
        for (uint32_t f = 0; f < a.frames; ++f) {
            g.sync();
            const auto t0 = std::chrono::steady_clock::now();
            g.iterate(a.steps_per_frame);
            g.sync();
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            const auto c = g.conservation();
            std::printf("{\"frame\": %u, \"step\": %llu, \"variant\": %d, \"mlups\": %.1f, "
                        "\"mass\": %.17g, \"momentum\": [%.17g, %.17g]}\n",
                        f, (unsigned long long)g.get_compute_num(), variant,
                        (double)a.nx * a.ny * a.steps_per_frame / s / 1e6, c[0], c[1], c[2]);
            std::fflush(stdout);
        }
        if (!a.dump.empty()) {
            const std::vector<float> f9 = g.read_f();
            FILE *fh = std::fopen(a.dump.c_str(), "wb");
            if (!fh || std::fwrite(f9.data(), sizeof(float), f9.size(), fh) != f9.size()) {
                if (fh) std::fclose(fh);
                throw lbm::Error(LBM_EINVAL, "cannot write " + a.dump);
            }
            std::fclose(fh);
        }
        return 0;
    } catch (const lbm::Error &e) {
        std::fprintf(stderr, "lbm_cli: %s (code %d)\n", e.what(), e.code());
        return 1;
    }
}
