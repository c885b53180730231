"""Oracle KATs for LinesBuilder::build and the rasterizer.

Ported from surpass/src/segment.rs:304-538 (active tests) and the commented
tests of surpass/src/rasterizer/mod.rs:301-698 (TILE_SIZE -> 8, +256 bias).
"""
import numpy as np
import pytest

EPS = np.float32(np.finfo(np.float32).eps)
SX, SY = 2.0, 3.0


def build(orc, segs, layers=None):
    a = np.asarray(segs, dtype=np.float32).reshape(-1, 4)
    return orc.lines_build(a[:, 0], a[:, 1], a[:, 2], a[:, 3], np.zeros(len(a), np.uint16), layers)


def approx(a, b):
    assert abs(np.float32(a) - np.float32(b)) <= EPS, (a, b)


# (dx, dy, start, octant, slope, swap)  -- segment.rs:322-500
OCTANTS = [
    (3, 2, SX, 0, 2 / 3, False), (2, 3, SY, 1, 2 / 3, True), (-2, 3, SY, 2, -2 / 3, True),
    (-3, 2, SX - 3, 3, -2 / 3, False), (-3, -2, SX - 3, 4, 2 / 3, False),
    (-2, -3, SY - 3, 5, 2 / 3, True), (2, -3, SY - 3, 6, -2 / 3, True),
    (3, -2, SX, 7, -2 / 3, False),
]


@pytest.mark.parametrize("dx,dy,start,octant,slope,swap", OCTANTS)
def test_line_octant(orc, dx, dy, start, octant, slope, swap):
    l = build(orc, [SX, SY, SX + dx, SY + dy])
    assert l["starts"][0] == int(start)
    assert l["lengths"][0] == 3
    approx(l["slopes"][0], slope)
    approx(l["slopes_recip"][0], 1 / slope)
    assert l["octants"][0] == octant
    s, sr = np.float32(l["slopes"][0]), np.float32(l["slopes_recip"][0])
    c0, c1 = np.float32(l["py_slope_pxs"][0]), np.float32(l["px_slope_recip_pys"][0])
    # intercept reconstruction (segment.rs:339-342)
    if not swap:
        xm = SX + np.sign(dx) * 1.5
        approx(s * np.float32(xm) + c0, SY + np.sign(dy) * 1.0)
        approx(sr * np.float32(SY + np.sign(dy) * 1.0) + c1, xm)
    else:
        ym = SY + np.sign(dy) * 1.5
        approx(s * np.float32(ym) + c0, SX + np.sign(dx) * 1.0)
        approx(sr * np.float32(SX + np.sign(dx) * 1.0) + c1, ym)


def test_octant_borders(orc):
    segs = [[0, 0, x, y] for x in (-10.0, 0.0, 10.0) for y in (-10.0, 0.0, 10.0) if x and y]
    l = build(orc, segs)
    assert (l["lengths"] == 10).all()


def test_empty_segment(orc):
    l = build(orc, [2.0, 2.0, 2.0, 2.0])
    assert list(l["lengths"]) == [0]


def test_disabled_layer_and_order(orc):
    layers = (orc.Layer * 3)()
    layers[1].exists, layers[1].is_enabled = 1, 0
    layers[2].exists, layers[2].is_enabled, layers[2].has_order, layers[2].order = 1, 1, 1, 77
    a = np.array([[0, 0, 3, 2]] * 3, np.float32)
    l = orc.lines_build(a[:, 0], a[:, 1], a[:, 2], a[:, 3], np.array([0, 1, 2], np.uint16), layers)
    assert list(l["lengths"]) == [3, 0, 3]
    assert list(l["orders"]) == [0, 0, 77]


def raster(orc, p0, p1):
    l = build(orc, [p0[0], p0[1], p1[0], p1[1]])
    return orc.rasterize(l)


def areas_and_covers(orc, segs):
    out = []
    for s in segs:
        f = orc.seg_fields(s)
        out.append(None if f["is_none"] else (f["area"], f["cover"]))
    return out


# rasterizer/mod.rs:341-456 (commented KATs)
AREA_COVER = {
    (3.0, 2.0): [(11 * 16 // 2, 11), None, (5 * 8 // 2 + 5 * 8, 5), (5 * 8 // 2, 5), (11 * 16 // 2, 11), None],
    (2.0, 3.0): [(16 * 11 // 2 + 16 * 5, 16), None, (8 * 5 // 2, 8), (8 * 5 // 2 + 8 * 11, 8), (16 * 11 // 2, 16), None],
    (-2.0, 3.0): [(16 * 11 // 2, 16), None, (8 * 5 // 2 + 8 * 11, 8), (8 * 5 // 2, 8), (16 * 11 // 2 + 16 * 5, 16), None],
    (-3.0, 2.0): [(11 * 16 // 2, 11), None, (5 * 8 // 2 + 5 * 8, 5), (5 * 8 // 2, 5), (11 * 16 // 2, 11), None],
    (-3.0, -2.0): [(-(11 * 16 // 2), -11), None, (-(5 * 8 // 2 + 5 * 8), -5), (-(5 * 8 // 2), -5), (-(11 * 16 // 2), -11), None],
    (-2.0, -3.0): [(-(16 * 11 // 2 + 16 * 5), -16), None, (-(8 * 5 // 2), -8), (-(8 * 5 // 2 + 8 * 11), -8), (-(16 * 11 // 2), -16), None],
    (2.0, -3.0): [(-(16 * 11 // 2), -16), None, (-(8 * 5 // 2 + 8 * 11), -8), (-(8 * 5 // 2), -8), (-(16 * 11 // 2 + 16 * 5), -16), None],
    (3.0, -2.0): [(-(11 * 16 // 2), -11), None, (-(5 * 8 // 2 + 5 * 8), -5), (-(5 * 8 // 2), -5), (-(11 * 16 // 2), -11), None],
    (1.0, 0.0): [],
    (1.0, 1.0): [(16 * 16 // 2, 16), None],
    (0.0, 1.0): [(16 * 16, 16), None],
    (-1.0, 1.0): [(16 * 16 // 2, 16), None],
    (-1.0, 0.0): [],
    (-1.0, -1.0): [(-(16 * 16 // 2), -16), None],
    (0.0, -1.0): [(-(16 * 16), -16), None],
}


@pytest.mark.parametrize("p1", list(AREA_COVER))
def test_area_cover(orc, p1):
    assert areas_and_covers(orc, raster(orc, (0.0, 0.0), p1)) == AREA_COVER[p1]


T = 8  # TILE_WIDTH == TILE_HEIGHT at this snapshot
# rasterizer/mod.rs:480-676 with TILE_SIZE = 8; tile_i is stored biased
TILES = [
    ((T, T), (T + 3, T + 2), [(1, 1, 0, 0), None, (1, 1, 1, 0), (1, 1, 1, 1), (1, 1, 2, 1), None]),
    ((T, T), (T + 2, T + 3), [(1, 1, 0, 0), None, (1, 1, 0, 1), (1, 1, 1, 1), (1, 1, 1, 2), None]),
    ((-T, T), (-T - 2, T + 3), [(-2, 1, T - 1, 0), None, (-2, 1, T - 1, 1), (-2, 1, T - 2, 1), (-2, 1, T - 2, 2), None]),
    ((-T, T), (-T - 3, T + 2), [(-2, 1, T - 3, 1), None, (-2, 1, T - 2, 1), (-2, 1, T - 2, 0), (-2, 1, T - 1, 0), None]),
    ((-T, -T), (-T - 3, -T - 2), [(-2, -2, T - 3, T - 2), None, (-2, -2, T - 2, T - 2), (-2, -2, T - 2, T - 1), (-2, -2, T - 1, T - 1), None]),
    ((-T, -T), (-T - 2, -T - 3), [(-2, -2, T - 2, T - 3), None, (-2, -2, T - 2, T - 2), (-2, -2, T - 1, T - 2), (-2, -2, T - 1, T - 1), None]),
    ((T, -T), (T + 2, -T - 3), [(1, -2, 1, T - 3), None, (1, -2, 1, T - 2), (1, -2, 0, T - 2), (1, -2, 0, T - 1), None]),
    ((T, -T), (T + 3, -T - 2), [(1, -2, 0, T - 1), None, (1, -2, 1, T - 1), (1, -2, 1, T - 2), (1, -2, 2, T - 2), None]),
]


@pytest.mark.parametrize("p0,p1,expected", TILES)
def test_tiles(orc, p0, p1, expected):
    got = []
    for s in raster(orc, p0, p1):
        f = orc.seg_fields(s)
        got.append(None if f["is_none"] else (f["tile_i"], f["tile_j"], f["tile_x"], f["tile_y"]))
    # the wire format stores max(tile_i, -1) + 256 (rasterizer/mod.rs:188)
    exp = [None if e is None else (max(e[0], -1) + 256, e[1], e[2], e[3]) for e in expected]
    assert got == exp


def test_start_and_end_not_on_pixel_border(orc):
    # rasterizer/mod.rs:682-696
    assert areas_and_covers(orc, raster(orc, (0.5, 0.25), (4.0, 2.0)))[0] == (4 * 8 // 2, 4)
    assert areas_and_covers(orc, raster(orc, (0.0, 0.0), (3.5, 1.75)))[6] == (4 * 8 // 2 + 4 * 8, 4)


def test_lengths_to_indices_layout(orc):
    # rasterizer/mod.rs:301-319: emission order is line order x pixel order,
    # two slots per line-pixel (slot 1 NONE unless split).
    segs = [[0, 0, 1, 1], [0, 0, 2, 2], [0, 0, 3, 3]]
    l = build(orc, segs)
    assert list(l["lengths"]) == [1, 2, 3]
    out = orc.rasterize(l)
    assert len(out) == 2 * 6
    assert all(int(s) == orc.NONE for s in out[1::2])


def test_seg_bitfield_roundtrip(orc):
    rng = np.random.default_rng(0)
    for _ in range(200):
        v = [0, int(rng.integers(-2048, 2048)), int(rng.integers(-4096, 4096)),
             int(rng.integers(0, 65536)), int(rng.integers(0, 8)), int(rng.integers(0, 8)),
             int(rng.integers(-512, 512)), int(rng.integers(-32, 32))]
        s = orc.seg_new(*v)
        f = orc.seg_fields(s)
        assert [f[k] for k in ("is_none", "tile_j", "tile_i", "layer", "tile_y", "tile_x", "area", "cover")] == v
    assert orc.seg_new(1, 0, 0, 0, 0, 0, 0, 0) == orc.NONE


def test_closed_path_cover_sums_to_zero(orc):
    """Invariant: for a closed polygon, per (tile row, pixel row) sum of covers is 0."""
    rng = np.random.default_rng(1)
    for _ in range(20):
        k = int(rng.integers(3, 12))
        pts = rng.uniform(-20, 60, size=(k, 2)).astype(np.float32)
        segs = np.concatenate([pts, np.roll(pts, -1, axis=0)], axis=1)
        l = build(orc, segs)
        out = orc.rasterize(l)
        d = orc.decode(out[out != np.uint64(orc.NONE)])
        rows = d["tile_j"] * 8 + d["tile_y"]
        for r in np.unique(rows):
            assert d["cover"][rows == r].sum() == 0


def test_brute_force_grid_rectangle_area(orc):
    """Axis-aligned rectangle with vertices on the 1/16 grid: per-pixel A
    (16 * prefix cover + area) equals 256 x covered fraction exactly."""
    rng = np.random.default_rng(2)
    for _ in range(20):
        x0, x1 = sorted(rng.integers(0, 40 * 16, size=2) / 16.0)
        y0, y1 = sorted(rng.integers(0, 40 * 16, size=2) / 16.0)
        if x0 == x1 or y0 == y1:
            continue
        segs = [[x0, y0, x0, y1], [x0, y1, x1, y1], [x1, y1, x1, y0], [x1, y0, x0, y0]]
        l = build(orc, segs)
        out = orc.rasterize(l)
        d = orc.decode(out[out != np.uint64(orc.NONE)])
        W = 48
        area = np.zeros((W, W), np.int64)
        cover = np.zeros((W, W + 1), np.int64)
        px = (d["tile_i"] - 256) * 8 + d["tile_x"]
        py = d["tile_j"] * 8 + d["tile_y"]
        for X, Y, a, c in zip(px, py, d["area"], d["cover"]):
            area[Y, X] += a
            cover[Y, X + 1] += c
        A = 16 * np.cumsum(cover, axis=1)[:, :W] + area
        xs = np.arange(W)
        fx = np.clip(np.minimum(xs + 1, x1) - np.maximum(xs, x0), 0, None)
        fy = np.clip(np.minimum(xs + 1, y1) - np.maximum(xs, y0), 0, None)
        expected = 256 * np.outer(fy, fx)
        assert np.array_equal(np.abs(A), expected.astype(np.int64))
