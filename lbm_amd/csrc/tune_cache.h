// Autotune cache: the fastest kernel variant per (device, shape class,
// precision), kept in a JSON file so that a process that builds a grid or a
// batch of a class already tuned on this kind of device selects its variant
// without timing every candidate again.  Host code only (no HIP calls), so
// that it also builds into plain C++ checks.
//
// The file is one JSON object, one entry per line, as store() writes it:
//   {
//   "gfx950:sramecc+:xnack-/256cu/grid/8192x4096/f32/ref/v10": {"variant": 8, "steps": 20, "ms": [0.93, ...]},
//   ...
//   }
// "ms" holds the per-step time of every candidate (null for one that could not
// run).  A line that does not read as an entry is ignored by load() and is
// dropped the next time store() rewrites the file.
#pragma once

#include <unistd.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

namespace tune_cache {

struct Entry {
    int variant = -1;
    uint32_t steps = 0;
    std::vector<float> ms;
};

// smallest power of two >= n: grids of 300 and 500 columns share a class
inline uint32_t pow2_ceil(uint32_t n) {
    uint32_t p = 1;
    while (p < n && p < (1u << 31)) p <<= 1;
    return p;
}

// a key holds [A-Za-z0-9:+-_./] only, so it needs no escaping in JSON
inline std::string clean_key(const std::string &s) {
    std::string k = s;
    for (char &c : k) {
        const bool ok = (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') ||
                        std::strchr(":+-_./", c) != nullptr;
        if (!ok || c == '\0') c = '_';
    }
    return k;
}

inline const char *skip_ws(const char *p) {
    while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
    return p;
}

// one line of the file -> (key, entry); false when the line is not an entry
inline bool parse_line(const std::string &line, std::string &key, Entry &e) {
    const size_t a = line.find('"');
    if (a == std::string::npos) return false;
    const size_t b = line.find('"', a + 1);
    if (b == std::string::npos || b == a + 1) return false;
    key = line.substr(a + 1, b - a - 1);
    if (clean_key(key) != key) return false;
    const char *s = line.c_str() + b + 1;
    const char *p = std::strstr(s, "\"variant\":");
    if (!p) return false;
    char *end = nullptr;
    const long v = std::strtol(p + 10, &end, 10);
    if (end == p + 10 || v < 0 || v > 4096) return false;
    p = std::strstr(end, "\"steps\":");
    if (!p) return false;
    const unsigned long st = std::strtoul(p + 8, &end, 10);
    if (end == p + 8 || st > 0xffffffffUL) return false;
    p = std::strstr(end, "\"ms\":");
    if (!p) return false;
    p = skip_ws(p + 5);
    if (*p != '[') return false;
    p = skip_ws(p + 1);
    std::vector<float> ms;
    while (*p != ']') {
        if (std::strncmp(p, "null", 4) == 0) {
            ms.push_back(INFINITY);
            p += 4;
        } else {
            const float x = std::strtof(p, &end);
            if (end == p) return false;
            ms.push_back(x);
            p = end;
        }
        p = skip_ws(p);
        if (*p == ',') p = skip_ws(p + 1);
        else if (*p != ']') return false;
    }
    e.variant = (int)v;
    e.steps = (uint32_t)st;
    e.ms = std::move(ms);
    return true;
}

inline std::map<std::string, Entry> load(const std::string &path) {
    std::map<std::string, Entry> m;
    std::ifstream in(path);
    std::string line;
    while (in && std::getline(in, line)) {
        std::string k;
        Entry e;
        if (parse_line(line, k, e)) m[k] = std::move(e);
    }
    return m;
}

inline bool lookup(const std::string &path, const std::string &key, Entry &e) {
    const auto m = load(path);
    const auto it = m.find(key);
    if (it == m.end()) return false;
    e = it->second;
    return true;
}

// adds or replaces `key`: writes <path>.tmp.<pid> and renames it over the
// file, so a reader never sees half a file.  Two processes storing at once:
// the later rename wins and the other entry is lost (re-tuned next time).
inline bool store(const std::string &path, const std::string &key, const Entry &e) {
    if (clean_key(key) != key || key.empty()) return false;
    auto m = load(path);
    m[key] = e;
    const std::string tmp = path + ".tmp." + std::to_string((long)getpid());
    FILE *f = std::fopen(tmp.c_str(), "w");
    if (!f) return false;
    std::fputs("{\n", f);
    size_t i = 0;
    for (const auto &[k, v] : m) {
        std::fprintf(f, "\"%s\": {\"variant\": %d, \"steps\": %u, \"ms\": [", k.c_str(), v.variant, v.steps);
        for (size_t j = 0; j < v.ms.size(); ++j) {
            if (std::isfinite(v.ms[j])) std::fprintf(f, "%s%.6g", j ? ", " : "", (double)v.ms[j]);
            else std::fprintf(f, "%snull", j ? ", " : "");
        }
        std::fprintf(f, "]}%s\n", ++i < m.size() ? "," : "");
    }
    std::fputs("}\n", f);
    bool ok = std::ferror(f) == 0;
    ok = std::fclose(f) == 0 && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        return false;
    }
    return true;
}

}  // namespace tune_cache
