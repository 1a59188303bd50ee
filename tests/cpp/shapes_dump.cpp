// Prints the point sets of lbm_shapes.hpp for the cases the Python tests name
// on the command line, one "x y v" line per point, each case ended by "--".
//   shapes_dump line X1 Y1 X2 Y2 NX NY ERASED | preset NAME NX NY | bresenham X1 Y1 X2 Y2
#include <cstdio>
#include <cstdlib>
#include <string>

#include "lbm_shapes.hpp"

using namespace lbm::shapes;

static void dump(const Shape &s) {
    for (const auto &p : s.points) std::printf("%lld %lld %d\n", std::get<0>(p), std::get<1>(p), std::get<2>(p) ? 1 : 0);
    std::printf("--\n");
}

int main(int argc, char **argv) {
    int i = 1;
    auto num = [&]() { return std::atoll(argv[i++]); };
    while (i < argc) {
        const std::string kind = argv[i++];
        if (kind == "line") {
            const long long x1 = num(), y1 = num(), x2 = num(), y2 = num(), nx = num(), ny = num(), er = num();
            try {
                dump(Line({x1, y1}, {x2, y2}, nx, ny, er != 0));
            } catch (const lbm::Error &) {
                std::printf("invalid\n--\n");
            }
        } else if (kind == "preset") {
            const std::string name = argv[i++];
            const long long nx = num(), ny = num();
            dump(name == "curl" ? curl_barrier(nx, ny) : name == "chaos" ? chaos_barrier(nx, ny) : welcome_barrier(nx, ny));
        } else if (kind == "bresenham") {
            const long long x1 = num(), y1 = num(), x2 = num(), y2 = num();
            for (const auto &p : bresenham({x1, y1}, {x2, y2})) std::printf("%lld %lld\n", p.first, p.second);
            std::printf("--\n");
        } else {
            return 2;
        }
    }
    return 0;
}
