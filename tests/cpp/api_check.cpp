// Exercises the C++ mirror (include/lbm_hip.hpp) on a GPU: the reference's
// surface, the summary and colour planes, variants, state read-back and the
// error path (lbm::Error with the C ABI's code).  Prints "ok" and exits 0, or
// names the first failed check and exits 1.
#include <cmath>
#include <cstdio>

#include "lbm_shapes.hpp"

#define CHECK(c)                                               \
    do {                                                       \
        if (!(c)) {                                            \
            std::printf("failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                          \
        }                                                      \
    } while (0)

int main() {
    try {
        const uint32_t X = 200, Y = 80;
        lbm::LBM g(1.0f / 0.56f, X, Y);
        CHECK(g.x() == X && g.y() == Y && g.get_compute_num() == 0);
        lbm::shapes::draw_shape(g, lbm::shapes::curl_barrier(X, Y));
        const std::vector<uint8_t> m = g.mask();
        CHECK(m[(size_t)(Y / 4) * X + 4 * X / 10] == 1 && m[(size_t)(Y / 2) * X + 4 * X / 10] == 1);
        g.custom_speed(0.1f);
        g.iterate(15);
        CHECK(g.get_compute_num() == 15);
        g.set_summary(lbm::SummaryStat::Speed);
        const std::vector<float> sp = g.calculate_summary();
        CHECK(sp.size() == (size_t)X * Y);
        const std::vector<float> rgb = g.render(lbm::SummaryStat::Curl, lbm::ColorMap::Viridis);
        CHECK(rgb.size() == (size_t)3 * X * Y);
        for (float v : rgb) CHECK(v >= 0.0f && v <= 1.0f);
        const auto c0 = g.conservation();
        CHECK(std::isfinite(c0[0]) && c0[0] > 0.9 * X * (Y - 2) * 0.9);
        // every variant continues the same state bit for bit
        lbm::LBM h(1.0f / 0.56f, X, Y);
        lbm::shapes::draw_shape(h, lbm::shapes::curl_barrier(X, Y));
        h.custom_speed(0.1f);
        h.iterate(15);
        for (int v = 0; v < LBM_NUM_VARIANTS; ++v) {
            g.set_variant(v);
            CHECK(g.variant() == v);
            g.step(5);
            h.step(5);
            CHECK(g.read_f() == h.read_f());
        }
        // errors carry the ABI's code and leave the grid usable
        bool threw = false;
        try {
            g.draw_shape(std::vector<uint32_t>{(uint32_t)(X * Y), 1});  // index past the grid
        } catch (const lbm::Error &e) {
            threw = e.code() == LBM_EINVAL;
        }
        CHECK(threw);
        threw = false;
        try {
            g.set_variant(LBM_NUM_VARIANTS);
        } catch (const lbm::Error &e) {
            threw = e.code() == LBM_EINVAL;
        }
        CHECK(threw);
        g.step(3);
        CHECK(g.get_compute_num() == 15 + 5 * LBM_NUM_VARIANTS + 3);
        lbm::LBM moved(std::move(g));
        CHECK(moved.handle() != nullptr && g.handle() == nullptr);
        std::printf("ok\n");
        return 0;
    } catch (const lbm::Error &e) {
        std::printf("failed: lbm::Error %d: %s\n", e.code(), e.what());
        return 1;
    }
}
