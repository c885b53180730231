"""Oracle KATs: mold path flattening (mold/src/path/mod.rs:181-403) and
surpass styles (surpass/src/painter/style.rs:434-643, painter/mod.rs)."""
import math

import numpy as np
import pytest

from oracle import mold
from oracle.mold import BlendMode, GradientBuilder, GradientType, Path, Point


def segs(path, t=None):
    return [tuple(map(float, s)) for s in path.segments(t)]


def test_close_open_path():
    p = Path().line(Point(0, 0), Point(1, 1)).line(Point(1, 1), Point(2, 0)).close()
    assert segs(p) == [(0, 0, 1, 1), (1, 1, 2, 0), (2, 0, 0, 0)]


def test_close_closed_path():
    p = Path().line(Point(0, 0), Point(1, 1)).line(Point(1, 1), Point(2, 0))
    p.line(Point(2, 0), Point(0, 0)).close()
    assert segs(p) == [(0, 0, 1, 1), (1, 1, 2, 0), (2, 0, 0, 0)]


def test_scale_translate():
    p = Path().quad(Point(0, 0), Point(1, 1), Point(2, 0))
    s = segs(p, [2.0, 0.0, 2.0, 0.0, 2.0, 2.0, 0.0, 0.0, 1.0])
    assert s[0][:2] == (2.0, 2.0)
    assert s[-1][2:] == (6.0, 2.0)


def test_scale_translate_t22_not_1():
    p = Path().quad(Point(0, 0), Point(1, 1), Point(2, 0))
    s = segs(p, [2.0, 0.0, 2.0, 0.0, 2.0, 2.0, 0.0, 0.0, 2.0])
    assert s[0][:2] == (1.0, 1.0)
    assert s[-1][2:] == (3.0, 1.0)


def test_perspective_quad_becomes_rational():
    # perspective_transform_quad (path/mod.rs:300-315): weights 1, 1.5, 2;
    # the flattened curve starts at (0,0) and ends at (2,0)/2 = (1,0).
    p = Path().quad(Point(0, 0), Point(1, 1), Point(2, 0))
    s = segs(p, [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.5, 0.0, 1.0])
    assert s[0][:2] == (0.0, 0.0)
    assert s[-1][2:] == (1.0, 0.0)


def test_rational_bezier_with_weights_1():
    a = Path().quad(Point(2, 0), Point(2, 2), Point(0, 2))
    b = Path().rat_quad((Point(2, 0), 1.0), (Point(2, 2), 1.0), (Point(0, 2), 1.0))
    assert segs(a) == segs(b)


def test_circle_rational_bezier():
    r = 10.0
    w = float(np.float32(np.sqrt(np.float32(2.0))) / np.float32(2.0))
    p = Path().rat_quad((Point(r, 0), 1.0), (Point(r * w, r * w), w), (Point(0, r), 1.0))
    s = p.segments()
    for seg in s[1:]:
        for x, y in ((seg[0], seg[1]), (seg[2], seg[3])):
            assert np.float32(x) * np.float32(x) + np.float32(y) * np.float32(y) - np.float32(r * r) < 1e-4


def test_quad_split_into_segments():
    p = Path().quad(Point(0, 0), Point(1, 1), Point(2, 0))
    assert len(p.segments()) == 3


def test_wang_count_closed_form():
    # n = ceil(sqrt(sqrt(dev2) * 2 / 0.25)) - 1 for a quad with deviation (0, -2h)
    for h in (1.0, 3.0, 17.0, 120.0):
        p = Path().quad(Point(0, 0), Point(5, h), Point(10, 0))
        dev = 2 * h
        n = math.ceil(math.sqrt(dev * 2 / 0.25)) - 1
        assert len(p.segments()) == n
        s = p.segments()
        assert tuple(s[0][:2]) == (0.0, 0.0) and tuple(s[-1][2:]) == (10.0, 0.0)


def test_small_deviation_is_single_line():
    p = Path().quad(Point(0, 0), Point(1, 0.05), Point(2, 0))
    assert segs(p) == [(0.0, 0.0, 2.0, 0.0)]


def test_cubic_endpoints_and_continuity():
    p = Path().cubic(Point(0, 0), Point(10, 30), Point(30, -10), Point(40, 5))
    s = p.segments()
    assert tuple(s[0][:2]) == (0.0, 0.0) and tuple(s[-1][2:]) == (40.0, 5.0)
    assert np.array_equal(s[1:, :2], s[:-1, 2:])


GRID = [0.0, 0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875, 1.0]


@pytest.mark.parametrize("mode", list(BlendMode))
def test_blend_mode_scalar_equals_simd_on_grid(orc, mode):
    # style.rs:469-490
    for d in GRID:
        for s in GRID:
            simd = orc.blend_simd(mode, [d] * 3, [s] * 3)
            f = np.float32(orc.blend_fn(mode, d, s))
            assert all(np.float32(v) == f for v in simd), (mode, d, s, simd, f)


def test_blend_modes_closed_forms(orc):
    # textbook definitions at a few off-grid points (catches swapped operands)
    d, s = 0.3, 0.6
    assert np.isclose(orc.blend_fn(BlendMode.Multiply, d, s), d * s)
    assert np.isclose(orc.blend_fn(BlendMode.Screen, d, s), d + s - d * s)
    assert np.isclose(orc.blend_fn(BlendMode.Overlay, d, s), 2 * s - 1 + d - (2 * s - 1) * d)
    assert np.isclose(orc.blend_fn(BlendMode.HardLight, d, s), 2 * d * s)
    assert np.isclose(orc.blend_fn(BlendMode.Darken, d, s), 0.3)
    assert np.isclose(orc.blend_fn(BlendMode.Lighten, d, s), 0.6)
    assert np.isclose(orc.blend_fn(BlendMode.Difference, d, s), 0.3)
    assert np.isclose(orc.blend_fn(BlendMode.Exclusion, d, s), d + s - 2 * d * s)
    assert np.isclose(orc.blend_fn(BlendMode.ColorDodge, d, s), min(1, s / (1 - d)))
    assert np.isclose(orc.blend_fn(BlendMode.ColorBurn, d, s), 1 - min(1, (1 - s) / d))
    assert np.isclose(orc.blend_fn(BlendMode.SoftLight, d, s), s - (1 - 2 * d) * s * (1 - s))


def test_workbench_blend_kats(orc):
    # layer_workbench.rs:913 (Solid([0.1875,..,0.75])) and :955 (Solid([0.5625,..,1.0]))
    c = orc.blend(BlendMode.Over, [0, 0, 0, 0], [0.5] * 4)
    c = orc.blend(BlendMode.Multiply, c, [0.5] * 4)
    assert list(c) == [0.1875, 0.1875, 0.1875, 0.75]
    c = orc.blend(BlendMode.Multiply, [1.0] * 4, [0.5] * 4)
    c = orc.blend(BlendMode.Multiply, c, [0.5] * 4)
    assert list(c) == [0.5625, 0.5625, 0.5625, 1.0]


def test_srgb_closed_form(orc):
    # srgb(0.5) * 255 -> 187 (RED_GREEN_50 = 0xBB, composition.rs:345)
    assert np.float32(orc.linear_to_srgb(0.5)) * np.float32(255) < 188
    assert round(float(orc.linear_to_srgb_simd(0.5)) * 255) == 0xBB
    assert list(orc.to_bytes([1.0, 0.0, 0.0, 1.0])) == [255, 0, 0, 255]
    # both forms agree
    for v in np.linspace(0, 1, 1001, dtype=np.float32):
        assert np.float32(orc.linear_to_srgb(v)) == np.float32(orc.linear_to_srgb_simd(v))
    # the approximation is within 1e-3 of the exact sRGB curve on [0, 1]
    for v in np.linspace(0.004, 1, 200):
        exact = 1.055 * v ** (1 / 2.4) - 0.055
        assert abs(orc.linear_to_srgb(v) - exact) < 2e-3


def _gradient_colors(orc, g, x, y):
    keep = []
    p = mold.c_props(mold.Props(func=mold.Func.Draw(mold.Style(fill=mold.Fill.Gradient(g)))), keep=keep)
    out = np.zeros(32, np.float32)
    orc.lib().or_gradient_color_at(p, x, y, orc.ptr(out, orc._f))
    return [[out[c * 8 + r] for c in range(4)] for r in range(8)]


def _eq(col):
    assert col[0] == col[1] == col[2] == col[3]
    return col[0]


def test_linear_gradient(orc):
    # style.rs:527-562
    g = GradientBuilder([0.0, 7.0], [7.0, 0.0])
    for v in (0.25, 0.75, 0.25, 0.75, 0.25):
        g.color([v] * 4)
    g = g.build()
    col = _gradient_colors(orc, g, 0.0, 0.0)
    assert col[0] == [0.25] * 4
    assert _eq(col[1]) < _eq(col[2]) < _eq(col[3])
    assert _eq(col[4]) > _eq(col[5]) > _eq(col[6])
    assert col[7] == [0.25] * 4
    col = _gradient_colors(orc, g, 3.0, 0.0)
    assert _eq(col[0]) < 0.75
    assert _eq(col[1]) > _eq(col[2]) > _eq(col[3])
    assert col[3] == [0.25] * 4
    assert _eq(col[3]) < _eq(col[4]) < _eq(col[5]) < _eq(col[6])
    assert _eq(col[7]) < 0.75
    col = _gradient_colors(orc, g, 7.0, 0.0)
    assert col[0] == [0.25] * 4 and col[7] == [0.25] * 4


def test_radial_gradient(orc):
    # style.rs:564-611
    r = float(np.float32(7.0) * (np.float32(1.0) / np.sqrt(np.float32(2.0))))
    g = GradientBuilder([0.0, 0.0], [r, r]).type(GradientType.Radial)
    g.color([0.25] * 4).color([0.75] * 4)
    g = g.build()
    col = _gradient_colors(orc, g, 0.0, 0.0)
    assert col[0] == [0.25] * 4
    assert all(_eq(col[i]) < _eq(col[i + 1]) for i in range(1, 6))
    assert col[7] == [0.75] * 4
    col = _gradient_colors(orc, g, 4.0, 0.0)
    assert all(_eq(col[i]) < _eq(col[i + 1]) for i in range(0, 5))
    assert col[6] == [0.75] * 4 and col[7] == [0.75] * 4
    col = _gradient_colors(orc, g, 7.0, 0.0)
    assert all(c == [0.75] * 4 for c in col)


def test_gradient_builder_stops():
    assert GradientBuilder([0, 0], [1, 1]).color([1, 0, 0, 1]).build() is None
    g = GradientBuilder([0, 0], [1, 1]).color([1, 0, 0, 1]).color([0, 1, 0, 1]).color([0, 0, 1, 1]).build()
    assert [s for _, s in g.stops] == [0.0, 0.5, 1.0]
    with pytest.raises(ValueError):
        GradientBuilder([0, 0], [1, 1]).color_with_stop([0, 0, 0, 1], 1.5)
