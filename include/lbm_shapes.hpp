/* Obstacle geometry for the C++ host side (lbm-wgpu src/barrier_shapes/ line.rs,
 * blob.rs, curve.rs, curve_collection.rs and merge_shapes.rs, and the barrier
 * presets of src/lbm.rs:1367-1480): Bresenham lines with diagonal
 * filling, curves, blobs, merging, get_points_vector's [x + y * nx, 1|0, ...]
 * wire format for lbm_paint, and the curl / chaos / welcome presets.
 *
 * Lines follow Bresenham's algorithm as the line_drawing crate 1.0.0 publishes
 * it (the reference's dependency): an octant transform to the first octant,
 * then the integer error walk from the start point to the end point
 * inclusive.  Points are (x, y, solid) triples; C++ integer division truncates
 * toward zero like Rust's, which the presets rely on for negative heights. */
#ifndef LBM_SHAPES_HPP
#define LBM_SHAPES_HPP

#include <algorithm>
#include <cstdint>
#include <functional>
#include <set>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "lbm_hip.hpp"

namespace lbm {
namespace shapes {

using XY = std::pair<long long, long long>;
using Point = std::tuple<long long, long long, bool>;  // (x, y, solid)
using PointSet = std::set<Point>;

namespace detail {
inline int octant(XY s, XY e) {
    long long dx = e.first - s.first, dy = e.second - s.second;
    int v = 0;
    if (dy < 0) {
        dx = -dx;
        dy = -dy;
        v += 4;
    }
    if (dx < 0) {
        const long long t = dx;
        dx = dy;
        dy = -t;
        v += 2;
    }
    if (dx < dy) v += 1;
    return v;
}
inline XY to_octant(int o, XY p) {
    const long long x = p.first, y = p.second;
    switch (o) {
    case 0: return {x, y};
    case 1: return {y, x};
    case 2: return {y, -x};
    case 3: return {-x, y};
    case 4: return {-x, -y};
    case 5: return {-y, -x};
    case 6: return {-y, x};
    default: return {x, -y};
    }
}
inline XY from_octant(int o, XY p) {
    const long long x = p.first, y = p.second;
    switch (o) {
    case 0: return {x, y};
    case 1: return {y, x};
    case 2: return {-y, x};
    case 3: return {-x, y};
    case 4: return {-x, -y};
    case 5: return {-y, -x};
    case 6: return {y, -x};
    default: return {x, -y};
    }
}
inline bool valid(XY a, XY b, long long xdim, long long ydim) {
    return a.first >= 0 && a.first < xdim && a.second >= 0 && a.second < ydim && b.first >= 0 &&
           b.first < xdim && b.second >= 0 && b.second < ydim;
}
}  // namespace detail

/* the points of the line from start to end inclusive */
inline std::vector<XY> bresenham(XY start, XY end) {
    const int o = detail::octant(start, end);
    XY p = detail::to_octant(o, start);
    const XY q = detail::to_octant(o, end);
    const long long dx = q.first - p.first, dy = q.second - p.second;
    long long err = dy - dx;
    std::vector<XY> out;
    while (p.first <= q.first) {
        out.push_back(detail::from_octant(o, p));
        if (err >= 0) {
            p.second += 1;
            err -= dx;
        }
        p.first += 1;
        err += dy;
    }
    return out;
}

struct Shape {
    PointSet points;
    bool is_empty() const { return points.empty(); }
};

/* a point keeps its latest value (blob.rs:31-43) */
struct Blob : Shape {
    Blob &join(const Shape &s) {
        for (const auto &p : s.points) set(p);
        return *this;
    }
    Blob &add(const std::vector<Point> &pts, long long xdim, long long ydim) {
        for (const auto &p : pts)
            if (std::get<0>(p) >= 0 && std::get<0>(p) < xdim && std::get<1>(p) >= 0 && std::get<1>(p) < ydim) set(p);
        return *this;
    }
    void empty() { points.clear(); }

  private:
    void set(const Point &p) {
        points.erase(Point(std::get<0>(p), std::get<1>(p), !std::get<2>(p)));
        points.insert(p);
    }
};

/* mod.rs:12-18 */
inline Blob negate(const Shape &s) {
    Blob b;
    for (const auto &p : s.points) b.points.insert(Point(std::get<0>(p), std::get<1>(p), !std::get<2>(p)));
    return b;
}

/* merge (merge_shapes.rs:5-10): s2's values win */
inline Blob merge(const Shape &s1, const Shape &s2) {
    Blob b;
    b.join(s1).join(s2);
    return b;
}

/* Line::new / Line::new_erased (line.rs:22-84): the line and its parallel
 * neighbours (1 pair drawn, 29 pairs erased) with every diagonal step filled */
struct Line : Shape {
    Line(XY p1, XY p2, long long xdim, long long ydim, bool erased = false) {
        if (!detail::valid(p1, p2, xdim, ydim))
            throw Error(LBM_EINVAL, "line endpoints outside the " + std::to_string(xdim) + " x " +
                                        std::to_string(ydim) + " grid");
        const bool value = !erased;
        const int thickness = erased ? 30 : 2;
        const XY a = p1.first > p2.first ? p1 : p2, b = p1.first > p2.first ? p2 : p1;
        std::vector<std::pair<XY, XY>> ends{{a, b}};
        for (long long i = 1; i < thickness; ++i) {
            const XY l0 = a.second > b.second ? XY{a.first, a.second - i} : XY{a.first, a.second + i};
            const XY l1{a.first + i, a.second};
            ends.push_back({l0, XY{b.first - i, b.second}});
            ends.push_back({l1, XY{b.first, b.second - i}});
        }
        for (const auto &e : ends) {
            if (!detail::valid(e.first, e.second, xdim, ydim)) continue;
            XY prev = e.first;
            for (const XY &p : bresenham(e.first, e.second)) {
                points.insert(Point(p.first, p.second, value));
                if (prev.first != p.first && prev.second != p.second) {  // diagonal step
                    const long long ddx = p.first - prev.first, ddy = p.second - prev.second;
                    points.insert(Point(p.first - ddx, p.second, value));
                    points.insert(Point(p.first, p.second - ddy, value));
                }
                prev = p;
            }
        }
    }
};

/* a stroke of joined segments (curve.rs:19-56) */
struct Curve : Shape {
    bool has_last = false;
    XY last{0, 0};
    void add_segment(XY next, long long xdim, long long ydim) { segment(next, xdim, ydim, false); }
    void erase_segment(XY next, long long xdim, long long ydim) { segment(next, xdim, ydim, true); }
    void empty() {
        points.clear();
        has_last = false;
    }

  private:
    void segment(XY next, long long xdim, long long ydim, bool erased) {
        if (has_last) {
            const Line l(last, next, xdim, ydim, erased);
            points.insert(l.points.begin(), l.points.end());
        } else {
            points.insert(Point(next.first, next.second, !erased));
        }
        last = next;
        has_last = true;
    }
};

/* the finished strokes of one drag (curve_collection.rs:4-29) */
struct CurveCollection : Shape {
    bool has_last = false;
    XY last{0, 0};
    void add_curve(const Curve &c) {
        has_last = c.has_last;
        last = c.last;
        points.insert(c.points.begin(), c.points.end());
    }
};

/* get_points_vector (merge_shapes.rs:12-22), points in row-major order */
inline std::vector<uint32_t> points_vector(const Shape &s, uint32_t xdim) {
    std::vector<Point> pts(s.points.begin(), s.points.end());
    std::sort(pts.begin(), pts.end(), [](const Point &a, const Point &b) {
        return std::make_tuple(std::get<1>(a), std::get<0>(a), std::get<2>(a)) <
               std::make_tuple(std::get<1>(b), std::get<0>(b), std::get<2>(b));
    });
    std::vector<uint32_t> out;
    out.reserve(2 * pts.size());
    for (const auto &p : pts) {
        out.push_back((uint32_t)(std::get<0>(p) + std::get<1>(p) * (long long)xdim));
        out.push_back(std::get<2>(p) ? 1u : 0u);
    }
    return out;
}

/* LBM::draw_shape (lbm.rs:1337-1339) for a shape */
inline void draw_shape(LBM &g, const Shape &s) { g.draw_shape(points_vector(s, g.x())); }

/* ---- barrier presets (lbm.rs:1367-1480), for an x by y grid ---------------------- */

inline Blob curl_barrier(long long x, long long y) {
    Blob b;
    b.join(Line({4 * x / 10, y / 4}, {4 * x / 10, y / 2}, x, y));
    return b;
}

inline Blob chaos_barrier(long long x, long long y) {
    Blob b;
    b.join(Line({x / 2, 9 * y / 20}, {x / 2, 0}, x, y));
    b.join(Line({x / 2, 11 * y / 20}, {x / 2, y - 1}, x, y));
    b.join(Line({3 * x / 5, y / 2}, {3 * x / 4, 3 * y / 4}, x, y));
    b.join(Line({3 * x / 5, y / 2}, {3 * x / 4, y / 4}, x, y));
    return b;
}

/* the word "Welcome!" in line segments */
inline Blob welcome_barrier(long long x, long long y) {
    const long long height = -(y / 4), bottom = y / 2, space = x / 50, lw = x / 13;
    long long cx = x / 5;
    std::vector<std::pair<XY, XY>> segs;
    auto seg = [&](XY a, XY b) { segs.push_back({a, b}); };
    auto letter_e = [&](long long c) {
        seg({c, bottom + height / 2}, {c, bottom});
        seg({c, bottom}, {c + lw, bottom});
        seg({c, bottom + height / 4}, {c + lw, bottom + height / 4});
        seg({c + lw, bottom + height / 2}, {c + lw, bottom + height / 4});
        seg({c, bottom + height / 2}, {c + lw, bottom + height / 2});
    };
    // W
    seg({cx, bottom + height}, {cx, bottom});
    seg({cx, bottom}, {cx + lw / 2, bottom + height / 2});
    cx += lw / 2;
    seg({cx, bottom + height / 2}, {cx + lw / 2, bottom});
    cx += lw / 2;
    seg({cx, bottom + height}, {cx, bottom});
    cx += space;
    letter_e(cx);  // e
    cx += lw + space;
    seg({cx, bottom}, {cx, bottom + height});  // l
    cx += space;
    seg({cx, bottom + height / 2}, {cx, bottom});  // c
    seg({cx, bottom}, {cx + lw, bottom});
    seg({cx, bottom + height / 2}, {cx + lw, bottom + height / 2});
    cx += lw + space;
    seg({cx, bottom + height / 2}, {cx, bottom});  // o
    seg({cx, bottom}, {cx + lw, bottom});
    seg({cx, bottom + height / 2}, {cx + lw, bottom + height / 2});
    seg({cx + lw, bottom + height / 2}, {cx + lw, bottom});
    cx += lw + space;
    seg({cx, bottom + height / 2}, {cx, bottom});  // m
    seg({cx, bottom + height / 2}, {cx + lw, bottom + height / 2});
    seg({cx + lw, bottom + height / 2}, {cx + lw, bottom});
    seg({cx + lw / 2, bottom}, {cx + lw / 2, bottom + height / 2});
    cx += lw + space;
    letter_e(cx);  // e
    cx += lw + space;
    seg({cx, height / 10 + bottom}, {cx, bottom});  // !
    seg({cx, height / 5 + bottom}, {cx, bottom + height});
    Blob b;
    for (const auto &s : segs) b.join(Line(s.first, s.second, x, y));
    return b;
}

}  // namespace shapes
}  // namespace lbm

#endif
