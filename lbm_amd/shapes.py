"""Obstacle geometry that produces the paint wire format (lbm-wgpu
src/barrier_shapes/*.rs): lines with diagonal filling, curves, blobs, the
barrier presets, painted through ``LBM.draw_shape`` (get_points_vector's
[index, value] pairs for ``lbm_paint``).

A shape is a set of (x, y, solid) points.  Lines are drawn with Bresenham's
algorithm as published by the line_drawing crate 1.0.0 (the reference's
dependency, not vendored): an octant transform to the first octant, then the
integer error walk from the start point to the end point inclusive.
"""
from __future__ import annotations


def _trunc_div(a: int, b: int) -> int:
    """Integer division rounding toward zero, as Rust's `/` on isize."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def _octant(start, end) -> int:
    dx, dy = end[0] - start[0], end[1] - start[1]
    value = 0
    if dy < 0:
        dx, dy, value = -dx, -dy, value + 4
    if dx < 0:
        dx, dy, value = dy, -dx, value + 2
    if dx < dy:
        value += 1
    return value


def _to(o, p):
    x, y = p
    return [(x, y), (y, x), (y, -x), (-x, y), (-x, -y), (-y, -x), (-y, x), (x, -y)][o]


def _from(o, p):
    x, y = p
    return [(x, y), (y, x), (-y, x), (-x, y), (-x, -y), (-y, -x), (y, -x), (x, -y)][o]


def bresenham(start, end):
    """Points of the line from start to end inclusive (line_drawing 1.0.0's Bresenham)."""
    o = _octant(start, end)
    (x, y), (ex, ey) = _to(o, start), _to(o, end)
    dx, dy = ex - x, ey - y
    err = dy - dx
    while x <= ex:
        yield _from(o, (x, y))
        if err >= 0:
            y += 1
            err -= dx
        x += 1
        err += dy


class Shape:
    """A set of (x, y, solid) points (barrier_shapes/mod.rs:9-18)."""

    def __init__(self, points=None):
        self.points = set(points or ())

    def is_empty(self) -> bool:
        return not self.points

    def negate(self) -> "Blob":
        return Blob((x, y, not v) for x, y, v in self.points)


class Blob(Shape):
    """Union of shapes in which a point keeps its latest value (blob.rs:31-43)."""

    def join(self, shape: Shape) -> "Blob":
        for x, y, v in shape.points:
            self.points.discard((x, y, not v))
            self.points.add((x, y, v))
        return self

    def add(self, points, xdim: int, ydim: int) -> "Blob":
        for x, y, v in points:
            if 0 <= x < xdim and 0 <= y < ydim:
                self.points.discard((x, y, not v))
                self.points.add((x, y, v))
        return self

    def empty(self) -> None:
        self.points.clear()


def _valid(p1, p2, xdim, ydim) -> bool:
    return all(0 <= p[0] < xdim and 0 <= p[1] < ydim for p in (p1, p2))


class Line(Shape):
    """Line::new / Line::new_erased (line.rs:22-84): the line between two grid
    points, thickened by parallel lines offset towards the corner the line
    turns from (1 pair for a drawn line, 29 for an eraser stroke), with every
    diagonal step filled so the result has no diagonal gaps a population could
    leak through.  Endpoints outside the grid raise ValueError."""

    def __init__(self, p1, p2, xdim: int, ydim: int, erased: bool = False):
        super().__init__()
        p1, p2 = tuple(p1), tuple(p2)
        if not _valid(p1, p2, xdim, ydim):
            raise ValueError(f"Endpoints {p1} {p2} are invalid with dimensions {xdim} and {ydim}")
        value = not erased
        for a, b in self._endpoints(p1, p2, 30 if erased else 2):
            if not _valid(a, b, xdim, ydim):
                continue
            prev = a
            for p in bresenham(a, b):
                self.points.add((p[0], p[1], value))
                if prev[0] != p[0] and prev[1] != p[1]:  # diagonal step
                    ddx, ddy = p[0] - prev[0], p[1] - prev[1]
                    self.points.add((p[0] - ddx, p[1], value))
                    self.points.add((p[0], p[1] - ddy, value))
                prev = p

    @staticmethod
    def _endpoints(p1, p2, thickness: int):
        """generate_endpoints / generate_endpoints_variable (line.rs:92-143)."""
        a, b = (p1, p2) if p1[0] > p2[0] else (p2, p1)
        out = [(a, b)]
        for i in range(1, thickness):
            if a[1] > b[1]:
                left = ((a[0], a[1] - i), (a[0] + i, a[1]))
            else:
                left = ((a[0], a[1] + i), (a[0] + i, a[1]))
            right = ((b[0] - i, b[1]), (b[0], b[1] - i))
            out.append((left[0], right[0]))
            out.append((left[1], right[1]))
        return out


class Curve(Shape):
    """A stroke of joined line segments (curve.rs:19-56)."""

    def __init__(self):
        super().__init__()
        self.last_point = None

    def _segment(self, nxt, xdim, ydim, erased):
        nxt = tuple(nxt)
        if self.last_point is not None:
            self.points |= Line(self.last_point, nxt, xdim, ydim, erased).points
        else:
            self.points.add((nxt[0], nxt[1], not erased))
        self.last_point = nxt

    def add_segment(self, nxt, xdim: int, ydim: int) -> None:
        self._segment(nxt, xdim, ydim, False)

    def erase_segment(self, nxt, xdim: int, ydim: int) -> None:
        self._segment(nxt, xdim, ydim, True)

    def empty(self) -> None:
        self.points.clear()
        self.last_point = None


class CurveCollection(Shape):
    """Finished strokes of one drag gesture (curve_collection.rs:4-29)."""

    def __init__(self):
        super().__init__()
        self.last_point = None

    def add_curve(self, curve: Curve) -> None:
        self.last_point = curve.last_point
        self.points |= curve.points


def merge(s1: Shape, s2: Shape) -> Blob:
    """merge (merge_shapes.rs:5-10): s2's values win where both hold a point."""
    return Blob().join(s1).join(s2)


def ordered_points(shape: Shape):
    """The points in row-major order.  The reference paints a hash set's
    iteration order in one parallel dispatch; a set that holds both values of
    one cell (a curve drawn and erased over itself) is therefore painted in an
    unspecified order there, and in this fixed order here."""
    return sorted(shape.points, key=lambda p: (p[1], p[0], p[2]))


# ---- barrier presets (lbm.rs:1367-1480) ------------------------------------------

def curl_barrier(x: int, y: int) -> Shape:
    return Line((4 * x // 10, y // 4), (4 * x // 10, y // 2), x, y)


def chaos_barrier(x: int, y: int) -> Shape:
    b = Blob()
    for p1, p2 in [((x // 2, 9 * y // 20), (x // 2, 0)), ((x // 2, 11 * y // 20), (x // 2, y - 1)),
                   ((3 * x // 5, y // 2), (3 * x // 4, 3 * y // 4)), ((3 * x // 5, y // 2), (3 * x // 4, y // 4))]:
        b.join(Line(p1, p2, x, y))
    return b


def welcome_barrier(x: int, y: int) -> Shape:
    """The word "Welcome!" drawn in line segments (lbm.rs:1388-1480)."""
    segs = []
    height = -(y // 4)
    bottom = y // 2
    space = x // 50
    cx = x // 5
    lw = x // 13
    h2, h4, h5, h10 = (_trunc_div(height, d) for d in (2, 4, 5, 10))
    # W
    segs += [((cx, bottom + height), (cx, bottom)), ((cx, bottom), (cx + lw // 2, bottom + h2))]
    cx += lw // 2
    segs.append(((cx, bottom + h2), (cx + lw // 2, bottom)))
    cx += lw // 2
    segs.append(((cx, bottom + height), (cx, bottom)))
    cx += space

    def e(cx):
        return [((cx, bottom + h2), (cx, bottom)), ((cx, bottom), (cx + lw, bottom)),
                ((cx, bottom + h4), (cx + lw, bottom + h4)), ((cx + lw, bottom + h2), (cx + lw, bottom + h4)),
                ((cx, bottom + h2), (cx + lw, bottom + h2))]
    segs += e(cx)
    cx += lw + space
    # l
    segs.append(((cx, bottom), (cx, bottom + height)))
    cx += space
    # c
    segs += [((cx, bottom + h2), (cx, bottom)), ((cx, bottom), (cx + lw, bottom)),
             ((cx, bottom + h2), (cx + lw, bottom + h2))]
    cx += lw + space
    # o
    segs += [((cx, bottom + h2), (cx, bottom)), ((cx, bottom), (cx + lw, bottom)),
             ((cx, bottom + h2), (cx + lw, bottom + h2)), ((cx + lw, bottom + h2), (cx + lw, bottom))]
    cx += lw + space
    # m
    segs += [((cx, bottom + h2), (cx, bottom)), ((cx, bottom + h2), (cx + lw, bottom + h2)),
             ((cx + lw, bottom + h2), (cx + lw, bottom)), ((cx + lw // 2, bottom), (cx + lw // 2, bottom + h2))]
    cx += lw + space
    # e
    segs += e(cx)
    cx += lw + space
    # !
    segs += [((cx, h10 + bottom), (cx, bottom)), ((cx, h5 + bottom), (cx, bottom + height))]
    b = Blob()
    for p1, p2 in segs:
        b.join(Line(p1, p2, x, y))
    return b


# ---- pointer interaction with undo history (lib.rs:220-425) ------------------------

class ClickType:
    """Pointer modes (lib.rs:220-225)."""
    Line, Erase, Draw, Inactive = range(4)


class ClickHandler:
    """Turns pointer events into shapes to paint, with a per-cell history that
    undo rolls back (lib.rs:227-425).  ``current_blob`` holds what the next
    redraw paints; ``undo`` returns the cells to repaint, or None.

    One difference: an eraser drag validates its strokes against the grid's
    real height (lib.rs:355 passes the width twice, so strokes near the bottom
    edge of a wide grid reach past it)."""

    def __init__(self, x: int, y: int):
        self.x, self.y = int(x), int(y)
        self.current_type = ClickType.Draw
        self.line_points = []
        self.current_blob = Blob()
        self.current_curve = Curve()
        self.contiguous_curve = CurveCollection()
        self.undo_stack = []
        self.history = {}

    def clear_barrier(self) -> None:
        self.current_curve = Curve()
        self.contiguous_curve = CurveCollection()
        self.undo_stack.clear()
        self.history.clear()
        self.current_blob.empty()
        self.line_points.clear()

    def _stroke(self, location) -> None:
        if self.current_type == ClickType.Erase:
            self.current_curve.erase_segment(location, self.x, self.y)
        elif self.current_type == ClickType.Draw:
            self.current_curve.add_segment(location, self.x, self.y)

    def handle_movement(self, location) -> None:
        self._stroke(location)

    def handle_click(self, location) -> None:
        if self.current_type == ClickType.Line:
            self._line_click(location)
        else:
            self._stroke(location)

    def handle_release(self) -> None:
        if self.current_type in (ClickType.Erase, ClickType.Draw):
            self.current_blob.join(self.current_curve)
            done, self.current_curve = self.current_curve, Curve()
            self.contiguous_curve.add_curve(done)
            stroke, self.contiguous_curve = self.contiguous_curve, CurveCollection()
            self._add_to_history(stroke)

    def empty_all(self) -> None:
        self.current_blob.empty()
        self.current_curve.empty()

    def switch_click_type(self, click_type: int) -> None:
        self.current_type = click_type
        self.line_points.clear()

    def update(self, pressed: bool, location, click_type: int) -> None:
        """After a redraw painted the current curve: start a new piece of it at
        the pointer (lib.rs:302-310, 346-358).  ``click_type`` is the mode last
        selected, which the next button event makes current."""
        self.current_blob.empty()
        if pressed and click_type in (ClickType.Draw, ClickType.Erase):
            done, self.current_curve = self.current_curve, Curve()
            self.contiguous_curve.add_curve(done)
            if click_type == ClickType.Draw:
                self.current_curve.add_segment(location, self.x, self.y)
            else:
                self.current_curve.erase_segment(location, self.x, self.y)

    def undo(self):
        if self.current_blob.is_empty() and self.current_curve.is_empty():
            if self.undo_stack:
                return self._remove_shape(self.undo_stack.pop())
            return None
        self.current_blob.empty()
        self.current_curve.empty()
        return None

    def _line_click(self, location) -> None:
        self.line_points.append(tuple(location))
        if len(self.line_points) >= 2:
            shape = Line(self.line_points[0], self.line_points[1], self.x, self.y)
            self.current_blob.join(shape)
            self._add_to_history(shape)
            self.line_points.clear()

    def validate_click(self, location):
        """Pointer position -> cell (lib.rs:360-377): negative coordinates count
        back from the far edge, coordinates past it clamp to the last cell."""
        def one(v, n):
            return n + v if v < 0 else (n - 1 if v >= n else v)
        return (one(int(location[0]), self.x), one(int(location[1]), self.y))

    def _add_to_history(self, shape: Shape) -> None:
        for x, y, v in shape.points:
            self.history.setdefault((x, y), []).append(v)
        self.undo_stack.append(shape)

    def _remove_shape(self, shape: Shape) -> Blob:
        points = []
        for x, y, _ in shape.points:
            vals = self.history.get((x, y))
            if vals is None:
                points.append((x, y, False))
                continue
            vals.pop()
            points.append((x, y, vals[-1] if vals else False))
            if not vals:
                del self.history[(x, y)]
        points.sort()
        return Blob().add(points, self.x, self.y)
