"""Oracle KATs for the tile painter and LayerWorkbench.

Ported from the commented tests of surpass/src/painter/mod.rs:599-1078,
layer_workbench.rs:489-1077 and buffer_layout.rs:233-477 (TILE_SIZE -> 8,
lanes = 8 tile rows, tile_i biased by +256 on the wire).
"""
import ctypes as C

import numpy as np
import pytest

from oracle import mold
from oracle.mold import BlendMode, Fill, FillRule, Func, Props, Style

T = 8
BLACK = [0.0, 0.0, 0.0, 1.0]
BLACK_TRANSPARENT = [0.0, 0.0, 0.0, 0.5]
RED = [1.0, 0.0, 0.0, 1.0]
RED_50 = [0.5, 0.0, 0.0, 1.0]
GREEN = [0.0, 1.0, 0.0, 1.0]
GREEN_50 = [0.0, 0.5, 0.0, 1.0]
RED_GREEN_50 = [0.5, 0.5, 0.0, 1.0]
BLUE = [0.0, 0.0, 1.0, 1.0]


def props_table(orc, d):
    """d: {order: mold.Props} -> ctypes array"""
    n = max(d) + 1 if d else 1
    keep = []
    t = (orc.Props * n)()
    for k, v in d.items():
        t[k] = mold.c_props(v, keep=keep)
    t._keep = keep
    return t


def solid(c, rule=FillRule.NonZero, clipped=False, mode=BlendMode.Over):
    return Props(fill_rule=rule, func=Func.Draw(Style(fill=Fill.Solid(c), is_clipped=clipped, blend_mode=mode)))


def line_segments(orc, points, same_layer=False):
    """painter/mod.rs:670-692: lines -> sorted segments cut at the first NONE."""
    a = np.array([[p0[0], p0[1], p1[0], p1[1]] for p0, p1 in points], np.float32)
    ids = np.zeros(len(a), np.uint16) if same_layer else np.arange(len(a), dtype=np.uint16)
    l = orc.lines_build(a[:, 0], a[:, 1], a[:, 2], a[:, 3], ids)
    s = np.sort(orc.rasterize(l))
    return s[s != np.uint64(orc.NONE)]


def carries(orc, lst):
    arr = (orc.Carry * max(1, len(lst)))()
    for i, (layer, cov) in enumerate(lst):
        arr[i].layer = layer
        for y in range(8):
            arr[i].cover[y] = cov[y]
    return arr, len(lst)


def drive(orc, carry_list, segs, ptable, clear=BLACK, passes=7, stop=0, tile_i=0, prev=None):
    arr, n = carries(orc, carry_list)
    segs = np.ascontiguousarray(segs, dtype=np.uint64)
    res = orc.TileResult()
    out = (orc.Carry * 1024)()
    clr = np.array(clear, np.float32)
    prev_p = C.byref(prev) if prev is not None else None
    orc.lib().or_drive_tile(arr, n, orc.ptr(segs, orc._u64), len(segs), ptable, len(ptable),
                            orc.ptr(clr, orc._f), tile_i, 0, prev_p, passes, stop, out, 1024,
                            C.byref(res))
    res._carries = [(out[k].layer, list(out[k].cover)) for k in range(res.n_carries_out)]
    return res


def colors(res):
    c = np.frombuffer(res.colors, dtype=np.float32).reshape(64, 4)
    return [list(map(float, v)) for v in c]


def test_carry_cover(orc):
    # painter/mod.rs:727-757
    segs = line_segments(orc, [((0.0, 0.0), (0.0, float(T)))])
    pt = props_table(orc, {0: solid(GREEN), 1: solid(RED)})
    cov = [0, 16, 0, 0, 0, 0, 0, 0]
    res = drive(orc, [(1, cov)], segs, pt)
    assert colors(res)[0:2] == [GREEN, RED]


def test_overlapping_triangles(orc):
    # painter/mod.rs:759-807
    segs = line_segments(orc, [((0.0, 0.0), (float(T), float(T))), ((float(T), 0.0), (0.0, float(T)))])
    pt = props_table(orc, {0: solid(GREEN), 1: solid(RED)})
    c = colors(drive(orc, [], segs, pt))
    rs, re = T // 2 - 2, T // 2 + 2
    col = (T // 2 - 2) * T
    assert c[col + rs:col + re] == [GREEN_50, BLACK, BLACK, RED_50]
    col += T
    assert c[col + rs:col + re] == [GREEN, GREEN_50, RED_50, RED]
    col += T
    assert c[col + rs:col + re] == [GREEN, RED_GREEN_50, RED, RED]
    col += T
    assert c[col + rs:col + re] == [RED_GREEN_50, RED, RED, RED]


def test_transparent_overlay(orc):
    # painter/mod.rs:809-836
    segs = line_segments(orc, [((0.0, 0.0), (0.0, float(T))), ((0.0, 0.0), (0.0, float(T)))])
    pt = props_table(orc, {0: solid(RED), 1: solid(BLACK_TRANSPARENT)})
    assert colors(drive(orc, [], segs, pt))[0] == RED_50


@pytest.mark.parametrize("v,nz_empty,eo_empty,nz_full,eo_full", [
    (0, True, True, False, False), (1, False, False, False, False),
    (-1, False, False, False, False), (16, False, False, True, True),
    (-16, False, False, True, True), (32, False, True, False, False),
    (-32, False, True, False, False), (48, False, False, False, True),
    (-48, False, False, False, True)])
def test_cover_is_empty_is_full(orc, v, nz_empty, eo_empty, nz_full, eo_full):
    # painter/mod.rs:859-979, observed through the carry queue: an empty carry
    # is dropped, a full opaque carry makes the tile Solid.
    for rule, empty, full in ((FillRule.NonZero, nz_empty, nz_full), (FillRule.EvenOdd, eo_empty, eo_full)):
        pt = props_table(orc, {0: solid(RED, rule=rule)})
        res = drive(orc, [(0, [v] * 8)], np.zeros(0, np.uint64), pt)
        assert (res.op == orc.OP_SOLID and list(res.solid) == RED) == full, (rule, v)
        assert (len(res._carries) == 0) == empty, (rule, v)


def test_clip(orc):
    # painter/mod.rs:981-1078
    segs = line_segments(orc, [((0.0, 0.0), (float(T), float(T))), ((0.0, 0.0), (0.0, float(T))),
                               ((0.0, 0.0), (0.0, float(T)))])
    pt = props_table(orc, {
        0: Props(func=Func.Clip(2)),
        1: solid(GREEN, clipped=True),
        2: solid(RED, clipped=True),
        3: solid(GREEN),
    })
    res = drive(orc, [], segs, pt)
    c = colors(res)
    col = [BLACK] * T
    for i in range(T):
        col[i] = [0.5, 0.25, 0.0, 1.0]
        if i >= 1:
            col[i - 1] = RED
        assert c[i * T:(i + 1) * T] == col
    # second tile: carries from the first plus one line at x = T
    segs2 = line_segments(orc, [((float(T), 0.0), (float(T), float(T)))])
    res2 = drive(orc, res._carries, segs2, pt, tile_i=1)
    assert colors(res2) == [RED] * 64


def wb_ids(res):
    return [res.ids[k] for k in range(res.n_ids) if res.masked[k]]


def seg(layer):
    return orc_seg(layer)


PARTIAL = [1] * 8
FULL = [16] * 8


def test_populate_layers(orc):
    # layer_workbench.rs:534-589
    segs = np.array([orc.seg_new(0, 0, 0, l, 0, 0, 0, 0) for l in (0, 1, 1, 2, 5, 5, 5)], np.uint64)
    pt = props_table(orc, {i: Props() for i in range(6)})
    res = drive(orc, [(0, PARTIAL), (3, PARTIAL), (4, PARTIAL)], segs, pt, passes=0, stop=1)
    assert list(res.ids[:res.n_ids]) == [0, 1, 2, 3, 4, 5]


def test_skip_unchanged(orc):
    # layer_workbench.rs:591-660
    p = {i: Props() for i in range(6)}
    tbl = props_table(orc, p)
    for i in range(5):
        tbl[i].is_unchanged = 1
    segs = np.array([orc.seg_new(0, 0, 0, l, 0, 0, 0, 0) for l in range(5)], np.uint64)
    layers = C.c_int32(4)
    res = drive(orc, [], segs, tbl, passes=1, stop=1, prev=layers)
    assert res.op == -1 and layers.value == 5
    res = drive(orc, [], segs, tbl, passes=1, stop=1, prev=layers)
    assert res.op == orc.OP_NONE and layers.value == 5
    segs = np.array([orc.seg_new(0, 0, 0, l, 0, 0, 0, 0) for l in range(1, 6)], np.uint64)
    res = drive(orc, [], segs, tbl, passes=1, stop=1, prev=layers)
    assert res.op == -1 and layers.value == 5


def test_skip_full_clip(orc):
    # layer_workbench.rs:662-713
    p = {0: Props(func=Func.Draw(Style(is_clipped=False))), 1: Props(func=Func.Clip(1)),
         2: Props(func=Func.Draw(Style(is_clipped=True))), 3: Props(func=Func.Clip(1))}
    res = drive(orc, [(0, PARTIAL), (1, FULL), (2, PARTIAL), (3, FULL)], np.zeros(0, np.uint64),
                props_table(orc, p), passes=2, stop=1)
    assert wb_ids(res) == [0, 2]
    skip = {res.ids[k] for k in range(res.n_ids) if res.skip_clipping[k]}
    assert 0 not in skip and 2 in skip


def test_skip_layer_outside_of_clip(orc):
    # layer_workbench.rs:715-744
    p = {0: Props(func=Func.Draw(Style(is_clipped=True))), 1: Props(func=Func.Draw(Style(is_clipped=True)))}
    res = drive(orc, [(0, PARTIAL), (1, PARTIAL)], np.zeros(0, np.uint64), props_table(orc, p),
                passes=2, stop=1)
    assert wb_ids(res) == []


def test_skip_without_layer_usage(orc):
    # layer_workbench.rs:746-786
    p = {0: Props(), 1: Props(func=Func.Clip(1)), 3: Props(), 4: Props(func=Func.Clip(1))}
    res = drive(orc, [(0, PARTIAL), (1, PARTIAL), (3, PARTIAL), (4, PARTIAL)], np.zeros(0, np.uint64),
                props_table(orc, p), passes=2, stop=1)
    assert wb_ids(res) == [0, 3]


def test_skip_everything_below_opaque(orc):
    # layer_workbench.rs:788-822
    p = {i: Props() for i in range(4)}
    segs = np.array([orc.seg_new(0, 0, 0, 3, 0, 0, 0, 0)], np.uint64)
    res = drive(orc, [(0, PARTIAL), (1, PARTIAL), (2, FULL)], segs, props_table(orc, p), passes=4, stop=1)
    assert res.op == -1
    assert wb_ids(res) == [2, 3]


def test_blend_top_full_layers(orc):
    # layer_workbench.rs:824-857
    p = {0: Props(func=Func.Draw(Style(fill=Fill.Solid([0.5] * 4), blend_mode=BlendMode.Over))),
         1: Props(func=Func.Draw(Style(fill=Fill.Solid([0.5] * 4), blend_mode=BlendMode.Multiply)))}
    res = drive(orc, [(0, FULL), (1, FULL)], np.zeros(0, np.uint64), props_table(orc, p),
                clear=[0.0] * 4, passes=4, stop=1)
    assert res.op == orc.OP_SOLID and list(res.solid) == [0.1875, 0.1875, 0.1875, 0.75]


def test_blend_top_full_layers_with_clear_color(orc):
    # layer_workbench.rs:859-890
    p = {i: Props(func=Func.Draw(Style(fill=Fill.Solid([0.5] * 4), blend_mode=BlendMode.Multiply)))
         for i in range(2)}
    res = drive(orc, [(0, FULL), (1, FULL)], np.zeros(0, np.uint64), props_table(orc, p),
                clear=[1.0] * 4, passes=4, stop=1)
    assert res.op == orc.OP_SOLID and list(res.solid) == [0.5625, 0.5625, 0.5625, 1.0]


def test_skip_fully_covered_layers_clip(orc):
    # layer_workbench.rs:892-925
    p = {0: Props(func=Func.Clip(1)), 1: Props(func=Func.Draw(Style(blend_mode=BlendMode.Multiply)))}
    res = drive(orc, [(0, PARTIAL), (1, FULL)], np.zeros(0, np.uint64), props_table(orc, p),
                clear=[1.0] * 4, passes=4, stop=1)
    assert res.op == -1


def test_skip_clip_then_blend(orc):
    # layer_workbench.rs:927-1076
    p = {0: Props(func=Func.Clip(1)),
         1: Props(func=Func.Draw(Style(fill=Fill.Solid([0.5] * 4), blend_mode=BlendMode.Multiply)))}
    res = drive(orc, [(0, PARTIAL), (1, FULL)], np.zeros(0, np.uint64), props_table(orc, p),
                clear=[1.0] * 4)
    assert res.op == orc.OP_SOLID and list(res.solid) == [0.75, 0.75, 0.75, 1.0]


def c_seg(orc, j, i, layer, ty, tx, area, cover):
    """CompactSegment::new as the pre-bias tests wrote it; the wire stores i + 256."""
    return orc.seg_new(0, j, max(i, -1) + 256, layer, ty, tx, area, cover)


def print_(orc, buf, width, segs, clear, props, crop=None, tile_ops=None):
    segs = np.sort(np.asarray(segs, dtype=np.uint64) >> np.uint64(22), kind="stable")  # key order
    flat = buf.reshape(-1)
    dbg = orc.Debug()
    if tile_ops is not None:
        dbg.tile_ops = tile_ops.ctypes.data_as(C.POINTER(C.c_int8))
    rect = orc.Rect(*crop) if crop is not None else None
    clr = np.array(clear, np.float32)
    return dbg, rect, clr, flat


def run_print(orc, buf, width, segs, clear, ptable, crop=None):
    s = np.asarray(segs, dtype=np.uint64)
    s = s[np.argsort(s >> np.uint64(22), kind="stable")]
    flat = buf.reshape(-1)
    row_len = (width + 7) // 8
    tile_rows = (len(buf) // width + 7) // 8
    ops = np.full(row_len * tile_rows, -1, np.int8)
    dbg = orc.Debug()
    dbg.tile_ops = ops.ctypes.data_as(C.POINTER(C.c_int8))
    rect = orc.Rect(*crop) if crop is not None else None
    clr = np.array(clear, np.float32)
    rc = orc.lib().or_print(orc.ptr(flat, orc._u8), len(buf), width, width, None, 0,
                            orc.ptr(s, orc._u64), len(s), orc.ptr(clr, orc._f),
                            C.byref(rect) if rect is not None else None, ptable, len(ptable),
                            C.byref(dbg))
    assert rc == 0
    return ops


def tiles_of(buf, width):
    """The layout order of buffer_layout.rs:99-127: tile rows, tiles, then rows."""
    h = len(buf) // width
    out = []
    for j in range(0, h, 8):
        for i in range(0, width, 8):
            for y in range(j, min(j + 8, h)):
                out.append([list(p) for p in buf[y * width + i:y * width + min(i + 8, width)]])
    return out


def test_flusher(orc):
    # buffer_layout.rs:253-285: every written slice is flushed
    size = T + T // 2
    buf = np.zeros((size * size, 4), np.uint8)
    segs = [c_seg(orc, 0, 0, 0, 0, 0, 0, 0), c_seg(orc, 0, 1, 0, 0, 0, 0, 0),
            c_seg(orc, 1, 0, 0, 0, 0, 0, 0), c_seg(orc, 1, 1, 0, 0, 0, 0, 0)]
    ops = run_print(orc, buf, size, segs, [0.0] * 4, props_table(orc, {0: Props()}))
    assert (ops != orc.OP_NONE).all() and (ops >= 0).all()


def test_skip_opaque_tiles(orc):
    # buffer_layout.rs:310-381
    buf = np.zeros((T * T * 3, 4), np.uint8)
    segs = [c_seg(orc, 0, -1, 2, y, T - 1, 0, 16) for y in range(T)]
    segs.append(c_seg(orc, 0, -1, 0, 0, T - 1, 0, 16))
    segs.append(c_seg(orc, 0, 0, 1, 1, 0, 0, 16))
    segs += [c_seg(orc, 0, 1, 2, y, T - 1, 0, -16) for y in range(T)]
    pt = props_table(orc, {0: solid(BLUE), 1: solid(GREEN), 2: solid(RED)})
    run_print(orc, buf, T * 3, segs, BLACK, pt)
    R, G, B, K = [255, 0, 0, 255], [0, 255, 0, 255], [0, 0, 255, 255], [0, 0, 0, 255]
    expected = [[R] * T] * T + [[R] * T] * T + [[B] * T, [G] * T] + [[K] * T] * (T - 2)
    assert tiles_of(buf, T * 3) == expected


def test_crop(orc):
    # buffer_layout.rs:383-476
    buf = np.zeros((T * T * 9, 4), np.uint8)
    segs = [c_seg(orc, j, 0, 0, y, T - 1, 0, 16) for j in range(3) for y in range(T)]
    pt = props_table(orc, {0: solid(BLUE)})
    r = mold.Rect(range(T, T * 2 + T // 2), range(T, T * 2))
    run_print(orc, buf, T * 3, segs, RED, pt,
              crop=(r.horizontal.start, r.horizontal.stop, r.vertical.start, r.vertical.stop))
    Z, B = [0, 0, 0, 0], [0, 0, 255, 255]
    expected = [[Z] * T] * (T * 3) + [[Z] * T] * T + [[B] * T] * (T * 2) + [[Z] * T] * (T * 3)
    assert tiles_of(buf, T * 3) == expected
