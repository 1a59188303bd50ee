// Drives lbm_amd/csrc/tune_cache.h (host code only) from the command line for
// tests/test_tune_cache.py:
//   tune_cache_check store PATH KEY VARIANT STEPS [MS|inf ...]   exit 0 / 1
//   tune_cache_check lookup PATH KEY      prints "VARIANT STEPS MS..." or "miss"
//   tune_cache_check pow2 N               prints pow2_ceil(N)
//   tune_cache_check clean KEY            prints clean_key(KEY)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../../lbm_amd/csrc/tune_cache.h"

int main(int argc, char **argv) {
    if (argc < 3) return 2;
    const std::string cmd = argv[1];
    if (cmd == "store" && argc >= 6) {
        tune_cache::Entry e;
        e.variant = std::atoi(argv[4]);
        e.steps = (uint32_t)std::strtoul(argv[5], nullptr, 10);
        for (int i = 6; i < argc; ++i)
            e.ms.push_back(std::string(argv[i]) == "inf" ? INFINITY : std::strtof(argv[i], nullptr));
        return tune_cache::store(argv[2], argv[3], e) ? 0 : 1;
    }
    if (cmd == "lookup" && argc == 4) {
        tune_cache::Entry e;
        if (!tune_cache::lookup(argv[2], argv[3], e)) {
            std::printf("miss\n");
            return 0;
        }
        std::printf("%d %u", e.variant, e.steps);
        for (float m : e.ms) std::printf(std::isfinite(m) ? " %g" : " inf", (double)m);
        std::printf("\n");
        return 0;
    }
    if (cmd == "pow2") {
        std::printf("%u\n", tune_cache::pow2_ceil((uint32_t)std::strtoul(argv[2], nullptr, 10)));
        return 0;
    }
    if (cmd == "clean") {
        std::printf("%s\n", tune_cache::clean_key(argv[2]).c_str());
        return 0;
    }
    return 2;
}
