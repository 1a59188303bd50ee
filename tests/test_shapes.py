"""Obstacle geometry (lbm_amd/shapes.py, following lbm-wgpu src/barrier_shapes/*.rs
and the presets at src/lbm.rs:1367-1480), checked on the CPU.

Bresenham comes from the line_drawing crate 1.0.0, which the reference does not
vendor; its tie-breaking (which cell a line takes where the ideal line passes
between two) has no fixture here, so these tests pin it by what fixes a line
drawing: both endpoints, one cell per major-axis step, 8-connected steps,
distance under one cell from the ideal line, and exact axis-aligned and 45
degree lines.  The cell-exact choice on other slopes is parity unpinned.
"""
import random
from collections import deque

import pytest

from lbm_amd import shapes as S


def _octant_cases():
    rng = random.Random(7)
    cases = [((0, 0), (5, 2)), ((5, 2), (0, 0)), ((0, 0), (2, 5)), ((3, 3), (3, 3))]
    for _ in range(400):
        cases.append(((rng.randint(-20, 20), rng.randint(-20, 20)), (rng.randint(-20, 20), rng.randint(-20, 20))))
    return cases


@pytest.mark.parametrize("start,end", _octant_cases()[:4])
def test_bresenham_known_small_cases(start, end):
    pts = list(S.bresenham(start, end))
    assert pts[0] == start and pts[-1] == end


def test_bresenham_line_properties_all_octants():
    for start, end in _octant_cases():
        pts = list(S.bresenham(start, end))
        dx, dy = end[0] - start[0], end[1] - start[1]
        major = max(abs(dx), abs(dy))
        assert pts[0] == start and pts[-1] == end
        assert len(pts) == major + 1
        assert len(set(pts)) == len(pts)
        for (x0, y0), (x1, y1) in zip(pts, pts[1:]):
            assert max(abs(x1 - x0), abs(y1 - y0)) == 1
        if major:
            for k, (x, y) in enumerate(pts):
                # the major coordinate advances by exactly one per point
                if abs(dx) >= abs(dy):
                    assert x - start[0] == k * (1 if dx > 0 else -1)
                else:
                    assert y - start[1] == k * (1 if dy > 0 else -1)
                # distance along the minor axis from the ideal line stays under one cell
                cross = (x - start[0]) * dy - (y - start[1]) * dx
                assert abs(cross) < major


@pytest.mark.parametrize("start,end", [((4, 1), (4, 9)), ((9, 3), (2, 3)), ((1, 1), (7, 7)), ((7, 1), (1, 7))])
def test_bresenham_exact_on_axes_and_diagonals(start, end):
    n = max(abs(end[0] - start[0]), abs(end[1] - start[1]))
    sx = (end[0] > start[0]) - (end[0] < start[0])
    sy = (end[1] > start[1]) - (end[1] < start[1])
    assert list(S.bresenham(start, end)) == [(start[0] + k * sx, start[1] + k * sy) for k in range(n + 1)]


def _four_connected(cells):
    cells = set(cells)
    start = next(iter(cells))
    seen, todo = {start}, deque([start])
    while todo:
        x, y = todo.popleft()
        for n in ((x + 1, y), (x - 1, y), (x, y + 1), (x, y - 1)):
            if n in cells and n not in seen:
                seen.add(n)
                todo.append(n)
    return len(seen) == len(cells)


def test_line_is_four_connected_and_contains_its_bresenham_walk():
    rng = random.Random(11)
    nx, ny = 64, 48
    for _ in range(200):
        p1 = (rng.randrange(nx), rng.randrange(ny))
        p2 = (rng.randrange(nx), rng.randrange(ny))
        line = S.Line(p1, p2, nx, ny)
        cells = {(x, y) for x, y, v in line.points}
        assert all(v for _, _, v in line.points)
        assert all(0 <= x < nx and 0 <= y < ny for x, y in cells)
        a, b = (p1, p2) if p1[0] > p2[0] else (p2, p1)
        assert set(S.bresenham(a, b)) <= cells
        # no diagonal gap a D2Q9 population could stream through
        assert _four_connected(cells)


def test_vertical_line_is_three_columns_closed_form():
    # curl preset on 400 x 160: main column 160 over rows 40..80, the offset
    # lines from (160,79) to (159,40) and from (161,80) to (160,39)
    line = S.curl_barrier(400, 160)
    cells = {(x, y) for x, y, _ in line.points}
    assert {(160, y) for y in range(39, 81)} <= cells
    assert {x for x, _ in cells} == {159, 160, 161}
    assert min(y for _, y in cells) == 39 and max(y for _, y in cells) == 80
    # the two offset lines each take one step sideways; the diagonal fill adds
    # both corners, so each contributes two cells outside column 160
    assert (161, 80) in cells and (159, 40) in cells


def test_line_rejects_endpoints_outside_the_grid():
    with pytest.raises(ValueError):
        S.Line((0, 0), (10, 5), 10, 10)
    with pytest.raises(ValueError):
        S.Line((-1, 0), (3, 3), 10, 10)


def test_offset_lines_outside_the_grid_are_skipped():
    # a line along the top row: the y+1 offset is inside, the y-1 one is not
    line = S.Line((0, 0), (9, 0), 10, 10)
    cells = {(x, y) for x, y, _ in line.points}
    assert all(0 <= x < 10 and 0 <= y < 10 for x, y in cells)
    assert {(x, 0) for x in range(10)} <= cells


def test_erased_line_is_thick_and_clears():
    line = S.Line((40, 40), (60, 45), 100, 100, erased=True)
    assert line.points and not any(v for _, _, v in line.points)
    drawn = S.Line((40, 40), (60, 45), 100, 100)
    er = {(x, y) for x, y, _ in line.points}
    assert {(x, y) for x, y, _ in drawn.points} <= er
    # the start point is the right one, above the end: 29 offsets below it and
    # 29 to its right (line.rs:113-143)
    assert (60, 45 - 29) in er and (60 + 29, 45) in er
    assert (60, 45 - 30) not in er and (60 + 30, 45) not in er


def test_blob_join_keeps_the_latest_value_and_add_bounds():
    line = S.Line((2, 2), (8, 5), 16, 16)
    b = S.Blob().join(line)
    assert b.points == line.points
    b.join(line.negate())
    assert b.points == {(x, y, False) for x, y, _ in line.points}
    b.empty()
    assert b.is_empty()
    b.add([(1, 1, True), (-1, 2, True), (16, 3, True), (3, 16, True), (1, 1, False)], 16, 16)
    assert b.points == {(1, 1, False)}


def test_merge_second_shape_wins():
    s1 = S.Shape({(1, 1, True), (2, 2, True)})
    s2 = S.Shape({(2, 2, False), (3, 3, True)})
    assert S.merge(s1, s2).points == {(1, 1, True), (2, 2, False), (3, 3, True)}


def test_curve_segments_and_collection():
    c = S.Curve()
    c.add_segment((5, 5), 32, 32)
    assert c.points == {(5, 5, True)} and c.last_point == (5, 5)
    c.add_segment((12, 9), 32, 32)
    assert c.points >= S.Line((5, 5), (12, 9), 32, 32).points
    c.erase_segment((20, 9), 32, 32)
    assert (20, 9, False) in c.points and c.last_point == (20, 9)
    cc = S.CurveCollection()
    cc.add_curve(c)
    assert cc.points == c.points and cc.last_point == (20, 9)
    c.empty()
    assert c.is_empty() and c.last_point is None


def test_ordered_points_is_row_major():
    pts = S.ordered_points(S.chaos_barrier(120, 80))
    keys = [(y, x) for x, y, _ in pts]
    assert keys == sorted(keys)


@pytest.mark.parametrize("nx,ny", [(400, 160), (256, 164), (1024, 512), (130, 90)])
def test_presets_in_bounds_and_solid(nx, ny):
    for fn in (S.curl_barrier, S.chaos_barrier, S.welcome_barrier):
        pts = fn(nx, ny).points
        assert pts and all(v for _, _, v in pts)
        assert all(0 <= x < nx and 0 <= y < ny for x, y, _ in pts)


def test_chaos_preset_closed_form():
    nx, ny = 400, 160
    cells = {(x, y) for x, y, _ in S.chaos_barrier(nx, ny).points}
    # two vertical walls at x = nx/2 with a gap between rows 9ny/20 and 11ny/20
    assert {(200, y) for y in range(0, 72 + 1)} <= cells
    assert {(200, y) for y in range(88, ny)} <= cells
    assert not any(x == 200 and 73 < y < 87 for x, y in cells)
    # a chevron from (3nx/5, ny/2) to (3nx/4, ny/4) and (3nx/4, 3ny/4)
    assert {(240, 80), (300, 40), (300, 120)} <= cells


def test_welcome_uses_truncating_division():
    # ny = 164: height = -41; Rust's isize division gives height/2 = -20 (not -21),
    # so the short letters reach row bottom - 20 = 62 and the offset lines one row more
    cells = {(x, y) for x, y, _ in S.welcome_barrier(400, 164).points}
    bottom, height = 82, -41
    assert min(y for _, y in cells) == bottom + height - 1
    # the "l" spans the full height, the "c" only half of it
    space, lw = 8, 30
    x_l = 80 + lw // 2 * 2 + space + lw + space
    assert {(x_l, y) for y in range(bottom + height, bottom + 1)} <= cells
    x_c = x_l + space
    col = [y for x, y in cells if x == x_c]
    assert min(col) == bottom - 20 - 1


# ---- pointer interaction and undo (lib.rs:227-425) -----------------------------------

def _cells(shape, value=None):
    return {(x, y) for x, y, v in shape.points if value is None or v == value}


def test_validate_click_wraps_negative_and_clamps_past_the_edge():
    ch = S.ClickHandler(40, 30)
    assert ch.validate_click((5, 7)) == (5, 7)
    assert ch.validate_click((-1, -2)) == (39, 28)
    assert ch.validate_click((40, 99)) == (39, 29)


def test_line_mode_two_clicks_make_a_line_and_undo_clears_it():
    ch = S.ClickHandler(64, 48)
    ch.switch_click_type(S.ClickType.Line)
    ch.handle_click((10, 10))
    assert ch.current_blob.is_empty()
    ch.handle_click((30, 20))
    line = S.Line((10, 10), (30, 20), 64, 48)
    assert ch.current_blob.points == line.points and len(ch.undo_stack) == 1
    # a pending blob is dropped by the first undo, which returns nothing
    assert ch.undo() is None and ch.current_blob.is_empty() and len(ch.undo_stack) == 1
    u = ch.undo()
    assert _cells(u) == _cells(line) and not _cells(u, True)
    assert ch.history == {} and ch.undo() is None


def test_draw_stroke_then_erase_over_it_undo_restores_the_drawing():
    ch = S.ClickHandler(64, 48)
    # a drawn stroke (10,24) -> (50,24)
    ch.handle_click((10, 24))
    ch.handle_movement((50, 24))
    ch.handle_release()
    drawn = S.Line((10, 24), (50, 24), 64, 48)
    assert _cells(drawn) <= _cells(ch.current_blob, True)
    ch.empty_all()
    # an eraser stroke across it
    ch.switch_click_type(S.ClickType.Erase)
    ch.handle_click((30, 10))
    ch.handle_movement((30, 40))
    ch.handle_release()
    erased = _cells(ch.current_blob, False)
    assert (30, 24) in erased
    ch.empty_all()
    u = ch.undo()
    # cells the eraser took from the drawing come back solid, the rest clear
    for x, y, v in u.points:
        assert v == ((x, y) in _cells(drawn) or (x, y) == (10, 24))
    assert (30, 24, True) in u.points
    u2 = ch.undo()
    assert not _cells(u2, True) and ch.history == {}


def test_update_starts_a_new_piece_of_the_stroke_at_the_pointer():
    ch = S.ClickHandler(64, 48)
    ch.handle_click((5, 5))
    ch.handle_movement((15, 9))
    ch.current_blob.join(ch.current_curve)
    ch.update(True, (15, 9), S.ClickType.Draw)
    assert ch.current_blob.is_empty()
    assert ch.current_curve.points == {(15, 9, True)} and ch.current_curve.last_point == (15, 9)
    assert (5, 5, True) in ch.contiguous_curve.points
    ch.handle_movement((25, 9))
    ch.handle_release()
    # the whole drag is one undo step
    assert len(ch.undo_stack) == 1
    assert {(5, 5), (15, 9), (25, 9)} <= _cells(ch.undo_stack[0])


def test_clear_barrier_forgets_history():
    ch = S.ClickHandler(32, 32)
    ch.switch_click_type(S.ClickType.Line)
    ch.handle_click((1, 1))
    ch.handle_click((9, 9))
    ch.clear_barrier()
    assert ch.undo_stack == [] and ch.history == {} and ch.current_blob.is_empty()


# ---- the C++ host side's geometry (include/lbm_shapes.hpp) equals this module's --------

def _cpp_dump(tmp_path, args):
    import os
    import subprocess
    root = os.path.join(os.path.dirname(__file__), "..")
    exe = tmp_path / "shapes_dump"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(root, "include"),
                    os.path.join(root, "tests", "cpp", "shapes_dump.cpp"), "-o", str(exe)], check=True, timeout=120)
    out = subprocess.run([str(exe)] + [str(a) for a in args], check=True, capture_output=True, text=True,
                         timeout=60).stdout
    cases, cur = [], []
    for line in out.splitlines():
        if line == "--":
            cases.append(cur)
            cur = []
        else:
            cur.append(line)
    return cases


def test_cpp_geometry_equals_python(tmp_path):
    rng = random.Random(5)
    args, expect = [], []
    for _ in range(60):
        nx, ny = rng.randint(10, 90), rng.randint(10, 70)
        p1 = (rng.randrange(nx), rng.randrange(ny))
        p2 = (rng.randrange(nx), rng.randrange(ny))
        erased = rng.random() < 0.3
        args += ["line", *p1, *p2, nx, ny, int(erased)]
        expect.append(S.Line(p1, p2, nx, ny, erased=erased).points)
    args += ["line", 0, 0, 10, 5, 10, 10, 0]
    expect.append(None)  # endpoint outside: rejected by both
    for name, fn in (("curl", S.curl_barrier), ("chaos", S.chaos_barrier), ("welcome", S.welcome_barrier)):
        for nx, ny in ((400, 160), (256, 164), (130, 90)):
            args += ["preset", name, nx, ny]
            expect.append(fn(nx, ny).points)
    for _ in range(40):
        a = (rng.randint(-15, 15), rng.randint(-15, 15))
        b = (rng.randint(-15, 15), rng.randint(-15, 15))
        args += ["bresenham", *a, *b]
        expect.append(list(S.bresenham(a, b)))
    got = _cpp_dump(tmp_path, args)
    assert len(got) == len(expect)
    for g, e in zip(got, expect):
        if e is None:
            assert g == ["invalid"]
        elif isinstance(e, list):
            assert [tuple(int(v) for v in line.split()) for line in g] == e
        else:
            pts = {(int(x), int(y), v == "1") for x, y, v in (line.split() for line in g)}
            assert pts == e
