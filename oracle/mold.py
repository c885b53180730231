"""mold's public scene API (mold/src/lib.rs:15-25) over the CPU oracle.

TEST INFRASTRUCTURE ONLY. A Python restatement of mold's bookkeeping
(composition.rs, layer.rs, buffer.rs, path/mod.rs) that hands the heavy
lifting (flatten, build, rasterize, sort, print) to ``liboracle.so``.
"""
import ctypes as C
import enum
from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np

from . import (Cmd, Debug, Layer as _CLayer, Props as _CProps, Rect as _CRect, f32, lib, ptr,
               _f, _u8, _u16, _u64, _i32, _sz, OP_NONE)

TILE_WIDTH = 8
TILE_HEIGHT = 8
IDENTITY = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
LINES_GARBAGE_THRESHOLD = 2  # composition.rs:16


class FillRule(enum.IntEnum):
    NonZero = 0
    EvenOdd = 1


class BlendMode(enum.IntEnum):
    Over = 0
    Multiply = 1
    Screen = 2
    Overlay = 3
    Darken = 4
    Lighten = 5
    ColorDodge = 6
    ColorBurn = 7
    HardLight = 8
    SoftLight = 9
    Difference = 10
    Exclusion = 11


class GradientType(enum.IntEnum):
    Linear = 0
    Radial = 1


def _f32t(v):
    return tuple(float(x) for x in np.asarray(v, dtype=np.float32).reshape(-1))


@dataclass(frozen=True)
class Gradient:
    type: GradientType
    start: Tuple[float, float]
    end: Tuple[float, float]
    stops: Tuple  # ((r, g, b, a), stop)


class GradientBuilder:
    """style.rs:29-85"""

    def __init__(self, start, end):
        self._type = GradientType.Linear
        self._start = _f32t(start)
        self._end = _f32t(end)
        self._stops = []

    def type(self, t):
        self._type = GradientType(t)
        return self

    def color(self, color):
        self._stops.append((_f32t(color), -0.0))
        return self

    def color_with_stop(self, color, stop):
        stop = float(np.float32(stop))
        if stop < 0.0 or 1.0 < stop:
            raise ValueError("gradient stops must be between 0.0 and 1.0")
        self._stops.append((_f32t(color), stop))
        return self

    def build(self) -> Optional[Gradient]:
        if len(self._stops) < 2:
            return None
        inc = np.float32(1.0) / np.float32(len(self._stops) - 1)
        stops = []
        for i, (c, s) in enumerate(self._stops):
            if s == -0.0:  # NO_STOP; +0.0 compares equal too
                s = float(np.float32(i) * inc)
            stops.append((c, s))
        return Gradient(self._type, self._start, self._end, tuple(stops))


@dataclass(frozen=True)
class Fill:
    kind: str = "solid"
    color: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 1.0)
    gradient: Optional[Gradient] = None

    @staticmethod
    def Solid(color):
        return Fill("solid", _f32t(color), None)

    @staticmethod
    def Gradient(gradient):
        return Fill("gradient", (0.0, 0.0, 0.0, 1.0), gradient)


@dataclass(frozen=True)
class Style:
    fill_rule: FillRule = FillRule.NonZero
    is_clipped: bool = False
    fill: Fill = field(default_factory=lambda: Fill.Solid((0.0, 0.0, 0.0, 1.0)))
    blend_mode: BlendMode = BlendMode.Over


@dataclass(frozen=True)
class Func:
    kind: str = "draw"
    style: Optional[Style] = None
    layers: int = 0

    @staticmethod
    def Draw(style=None):
        return Func("draw", style if style is not None else Style(), 0)

    @staticmethod
    def Clip(layers):
        return Func("clip", None, int(layers))


@dataclass(frozen=True)
class Props:
    fill_rule: FillRule = FillRule.NonZero
    func: Func = field(default_factory=Func.Draw)


@dataclass(frozen=True)
class Point:
    x: float
    y: float


class Path:
    """mold/src/path/mod.rs:24-179"""

    def __init__(self):
        self.commands = []

    @staticmethod
    def _pts(*ps):
        out = []
        for p in ps:
            out += [p.x, p.y]
        return _f32t(out)

    @staticmethod
    def _wpts(*ps):
        out = []
        for p, w in ps:
            w32 = np.float32(w)
            out += [np.float32(p.x) * w32, np.float32(p.y) * w32, w32]
        return _f32t(out)

    def line(self, p0, p1):
        self.commands.append((0, self._pts(p0, p1)))
        return self

    def quad(self, p0, p1, p2):
        self.commands.append((1, self._pts(p0, p1, p2)))
        return self

    def rat_quad(self, p0, p1, p2):
        self.commands.append((2, self._wpts(p0, p1, p2)))
        return self

    def cubic(self, p0, p1, p2, p3):
        self.commands.append((3, self._pts(p0, p1, p2, p3)))
        return self

    def rat_cubic(self, p0, p1, p2, p3):
        self.commands.append((4, self._wpts(p0, p1, p2, p3)))
        return self

    def close(self):
        self.commands.append((5, ()))
        return self

    def segments(self, transform=None):
        from . import path_flatten
        return path_flatten(self.commands, transform)


class LayerId(int):
    pass


class Layer:
    """mold/src/layer.rs:144-244"""

    def __init__(self):
        self.is_enabled_ = True
        self.affine = None
        self.order_ = None
        self.props_ = Props()
        self.is_unchanged_ = 0
        self.len = 0

    def is_enabled(self):
        return self.is_enabled_

    def set_is_enabled(self, v):
        self.is_enabled_ = bool(v)
        return self

    def disable(self):
        self.is_enabled_ = False
        return self

    def enable(self):
        self.is_enabled_ = True
        return self

    def transform(self):
        return self.affine if self.affine is not None else IDENTITY

    def set_transform(self, t):
        t = _f32t(t)
        if t == IDENTITY:
            affine = None
        else:
            a = np.asarray(t, dtype=np.float32)
            if (a[0] * a[0] + a[2] * a[2] > np.float32(1.001)
                    or a[1] * a[1] + a[3] * a[3] > np.float32(1.001)):
                raise ValueError("Layer's scaling on each axis must be between -1.0 and 1.0")
            affine = t
        if self.affine != affine:
            self.is_unchanged_ = 0
            self.affine = affine
        return self

    def order(self):
        if self.order_ is None:
            raise ValueError("Layers should always have orders")
        return self.order_

    def set_order(self, order):
        order = int(order) & 0xFFFF
        if self.order_ != order:
            self.is_unchanged_ = 0
            self.order_ = order
        return self

    def props(self):
        return self.props_

    def set_props(self, props):
        if self.props_ != props:
            self.is_unchanged_ = 0
            self.props_ = props
        return self

    def is_unchanged(self, cache_id):
        return (self.is_unchanged_ >> cache_id) & 1 == 1

    def set_is_unchanged(self, cache_id, v):
        if cache_id >= 32:
            return False
        if v:
            self.is_unchanged_ |= 1 << cache_id
        else:
            self.is_unchanged_ &= ~(1 << cache_id)
        return True


class BufferLayerCache:
    """mold/src/buffer.rs:38-72"""

    def __init__(self, cache_id, owner):
        self.id = cache_id
        self.layers_per_tile = np.zeros(0, dtype=np.int32)
        self._owner = owner

    def clear(self):
        self.layers_per_tile[:] = -1

    def drop(self):
        self._owner._caches &= ~(1 << self.id)


class Rect:
    """buffer_layout.rs:37-50: pixel ranges to tile ranges."""

    def __init__(self, horizontal: range, vertical: range):
        self.horizontal = range(horizontal.start // TILE_WIDTH,
                                (horizontal.stop + TILE_WIDTH - 1) // TILE_WIDTH)
        self.vertical = range(vertical.start // TILE_HEIGHT,
                              (vertical.stop + TILE_HEIGHT - 1) // TILE_HEIGHT)


class Buffer:
    """mold/src/buffer.rs:17-36; `buffer` is a (len, 4) uint8 array."""

    def __init__(self, buffer, width, width_stride=None, layer_cache=None, flusher=None):
        self.buffer = buffer
        self.width = width
        self.width_stride = width_stride
        self.layer_cache = layer_cache
        self.flusher = flusher

    def tiles_len(self):
        bw = self.width_stride if self.width_stride is not None else self.width
        tw = (self.width + TILE_WIDTH - 1) // TILE_WIDTH
        bh = len(self.buffer) // bw
        th = (bh + TILE_HEIGHT - 1) // TILE_HEIGHT
        return tw * th


def clear_buffer(buffer, color):
    buffer[:] = np.asarray(color, dtype=np.uint8)


def c_props(props: Props, is_unchanged=False, keep=None):
    p = _CProps()
    p.defined = 1
    p.fill_rule = int(props.fill_rule)
    p.is_unchanged = 1 if is_unchanged else 0
    f = props.func
    if f.kind == "clip":
        p.func = 1
        p.clip_layers = f.layers & 0xFFFFFFFF
        return p
    s = f.style
    p.func = 0
    p.is_clipped = 1 if s.is_clipped else 0
    p.blend_mode = int(s.blend_mode)
    if s.fill.kind == "solid":
        p.fill_kind = 0
        for i in range(4):
            p.color[i] = s.fill.color[i]
    else:
        g = s.fill.gradient
        p.fill_kind = 1
        p.grad_type = int(g.type)
        p.grad_start[0], p.grad_start[1] = g.start
        p.grad_end[0], p.grad_end[1] = g.end
        stops = np.array([list(c) + [st] for c, st in g.stops], dtype=np.float32).reshape(-1)
        if keep is not None:
            keep.append(stops)
        p.n_stops = len(g.stops)
        p.stops = stops.ctypes.data_as(_f)
    return p


class Composition:
    """mold/src/composition.rs:25-329"""

    def __init__(self):
        self._lines = []  # chunks of (n, 4) float32
        self._ids = []    # chunks of uint16
        self._n_lines = 0
        self.layers_ = {}
        self._free = []   # released ids (min-heap)
        self._next_id = 0
        self.orders_to_layers = {}
        self._layouts = {}
        self._caches = 0

    # ---- lines builder ----
    def _concat(self):
        if len(self._lines) > 1:
            self._lines = [np.concatenate(self._lines)]
            self._ids = [np.concatenate(self._ids)]
        if not self._lines:
            return np.zeros((0, 4), np.float32), np.zeros(0, np.uint16)
        return self._lines[0], self._ids[0]

    def builder_len(self):
        return self._n_lines

    # ---- ids (layer.rs:35-99): the smallest free id is handed out ----
    def create_layer(self) -> Optional[LayerId]:
        import heapq
        if self._free:
            return LayerId(heapq.heappop(self._free))
        if self._next_id > 0xFFFF:
            return None
        i = self._next_id
        self._next_id += 1
        return LayerId(i)

    def _release(self, i):
        import heapq
        heapq.heappush(self._free, int(i))

    def insert_lines(self, layer_id, segs):
        segs = np.ascontiguousarray(segs, dtype=np.float32).reshape(-1, 4)
        self._lines.append(segs)
        self._ids.append(np.full(len(segs), int(layer_id), dtype=np.uint16))
        self._n_lines += len(segs)
        layer = self.layers_.setdefault(int(layer_id), Layer())
        layer.order_ = int(layer_id)      # composition.rs:74
        layer.is_enabled_ = True
        layer.len += len(segs)
        return layer

    def insert_in_layer(self, layer_id, path: Path):
        return self.insert_lines(layer_id, path.segments())

    def insert_in_layer_transformed(self, layer_id, path: Path, transform):
        return self.insert_lines(layer_id, path.segments(transform))

    def get(self, layer_id):
        return self.layers_.get(int(layer_id))

    def get_mut(self, layer_id):
        return self.layers_.get(int(layer_id))

    def layers(self):
        return list(self.layers_.values())

    def actual_len(self):
        return sum(l.len for l in self.layers_.values() if l.is_enabled_)

    def remove_disabled(self):
        """composition.rs:123-142"""
        if self._n_lines >= self.actual_len() * LINES_GARBAGE_THRESHOLD:
            lines, ids = self._concat()
            keep = np.array([self.layers_.get(int(i)) is not None and self.layers_[int(i)].is_enabled_
                             for i in np.unique(ids)], dtype=bool)
            uniq = np.unique(ids)
            enabled = dict(zip(uniq.tolist(), keep.tolist()))
            mask = np.array([enabled[int(i)] for i in ids], dtype=bool) if len(ids) else np.zeros(0, bool)
            self._lines = [lines[mask]]
            self._ids = [ids[mask]]
            self._n_lines = int(mask.sum())
            for lid in [k for k, l in self.layers_.items() if not l.is_enabled_]:
                del self.layers_[lid]
                self._release(lid)

    def create_buffer_layer_cache(self):
        slot = 0
        while (self._caches >> slot) & 1:
            slot += 1
        if slot >= 32:
            return None
        self._caches |= 1 << slot
        return BufferLayerCache(slot, self)

    def _layer_table(self):
        n = (max(self.layers_) + 1) if self.layers_ else 1
        table = (_CLayer * n)()
        for lid, l in self.layers_.items():
            e = table[lid]
            e.exists = 1
            e.is_enabled = 1 if l.is_enabled_ else 0
            if l.affine is not None:
                e.has_affine = 1
                for i in range(6):
                    e.affine[i] = l.affine[i]
            if l.order_ is not None:
                e.has_order = 1
                e.order = l.order_
        return table, n

    def _props_table(self, cache_id, keep):
        n = (max(self.orders_to_layers) + 1) if self.orders_to_layers else 1
        table = (_CProps * n)()
        for order, lid in self.orders_to_layers.items():
            l = self.layers_.get(lid)
            if l is None:
                continue
            unchanged = cache_id is not None and l.is_unchanged(cache_id)
            table[order] = c_props(l.props_, unchanged, keep)
        return table, n

    def _line_ptrs(self):
        lines, ids = self._concat()
        cols = [np.ascontiguousarray(lines[:, k]) for k in range(4)]
        return cols, np.ascontiguousarray(ids)

    def pixel_segments(self):
        """composition.rs:153-188: the whole sorted buffer incl. the NONE tail."""
        self.remove_disabled()
        cols, ids = self._line_ptrs()
        table, nl = self._layer_table()
        L = lib()
        n = L.or_pixel_segments(len(ids), *(ptr(c, _f) for c in cols), ptr(ids, _u16), table, nl,
                                None, 0)
        out = np.zeros(max(1, n), dtype=np.uint64)
        L.or_pixel_segments(len(ids), *(ptr(c, _f) for c in cols), ptr(ids, _u16), table, nl,
                            ptr(out, _u64), n)
        return out[:n]

    def compute_orders(self):
        self.remove_disabled()
        for lid in sorted(self.layers_):
            l = self.layers_[lid]
            if l.is_enabled_:
                self.orders_to_layers[l.order()] = lid

    def render_mt(self, buffer: Buffer, background_color, n_threads):
        """Multi-threaded render (no layer cache, crop or flusher): the CPU
        baseline. Returns the number of non-NONE pixel segments."""
        self.remove_disabled()
        for lid in sorted(self.layers_):
            l = self.layers_[lid]
            if l.is_enabled_:
                self.orders_to_layers[l.order()] = lid
        keep = []
        ptable, np_ = self._props_table(None, keep)
        ltable, nl = self._layer_table()
        cols, ids = self._line_ptrs()
        data = buffer.buffer
        flat = data.reshape(-1)
        stride = buffer.width_stride or buffer.width
        nseg = C.c_size_t(0)
        clear = f32(background_color)
        rc = lib().or_render_mt(int(n_threads), len(ids), *(ptr(c, _f) for c in cols), ptr(ids, _u16),
                                ltable, nl, ptr(flat, _u8), len(data), buffer.width, stride,
                                ptr(clear, _f), ptable, np_, C.byref(nseg))
        if rc != 0:
            raise RuntimeError("oracle render failed")
        return nseg.value

    def render_coverage(self, width, height, background_color, crop: Optional[Rect] = None, cap=1 << 16):
        """The painter's per-pixel A values of every (tile, workbench id) of a
        render into a fresh width x height buffer (painter/mod.rs:229-243,
        347-362): (A int32 (n, 64) column-major x*8+y, keys uint32 (n, 3) =
        (tile row, tile column, layer))."""
        while True:
            a = np.zeros((cap, 64), np.int32)
            k = np.zeros((cap, 3), np.uint32)
            dbg = Debug()
            dbg.cov_a = a.ctypes.data_as(C.POINTER(C.c_int32))
            dbg.cov_key = k.ctypes.data_as(C.POINTER(C.c_uint32))
            dbg.cap_records = cap
            buf = np.zeros((width * height, 4), np.uint8)
            self.render(Buffer(buf, width), background_color, crop, debug=dbg)
            if dbg.n_records <= cap:
                return a[:dbg.n_records], k[:dbg.n_records]
            cap = int(dbg.n_records)

    def render(self, buffer: Buffer, background_color, crop: Optional[Rect] = None, debug=None):
        """composition.rs:212-328"""
        cache = buffer.layer_cache
        if cache is not None:
            tl = buffer.tiles_len()
            if len(cache.layers_per_tile) != tl:
                cache.layers_per_tile = np.full(tl, -1, dtype=np.int32)
                cache.clear()
        self.remove_disabled()
        for lid in sorted(self.layers_):
            l = self.layers_[lid]
            if l.is_enabled_:
                self.orders_to_layers[l.order()] = lid
        data = buffer.buffer
        key = (data.ctypes.data, len(data))
        if key not in self._layouts:  # layout cache keyed by (ptr, len)
            self._layouts[key] = (buffer.width, buffer.width_stride)
        width, stride = self._layouts[key]
        stride = stride if stride is not None else width
        keep = []
        cache_id = cache.id if cache is not None else None
        ptable, np_ = self._props_table(cache_id, keep)
        ltable, nl = self._layer_table()
        cols, ids = self._line_ptrs()
        rect = None
        if crop is not None:
            rect = _CRect(crop.horizontal.start, crop.horizontal.stop, crop.vertical.start,
                          crop.vertical.stop)
        row_len = (width + TILE_WIDTH - 1) // TILE_WIDTH
        rows_px = (len(data) + stride - 1) // stride
        tile_rows = (rows_px + TILE_HEIGHT - 1) // TILE_HEIGHT
        tile_ops = np.full(row_len * tile_rows, -1, dtype=np.int8)
        dbg = debug if debug is not None else Debug()
        dbg.tile_ops = tile_ops.ctypes.data_as(C.POINTER(C.c_int8))
        nseg = C.c_size_t(0)
        flat = data.reshape(-1)
        assert flat.flags["C_CONTIGUOUS"] and data.dtype == np.uint8
        lpt = cache.layers_per_tile if cache is not None else None
        clear = f32(background_color)
        rc = lib().or_render(len(ids), *(ptr(c, _f) for c in cols), ptr(ids, _u16), ltable, nl,
                             ptr(flat, _u8), len(data), width, stride, ptr(lpt, _i32),
                             len(lpt) if lpt is not None else 0, ptr(clear, _f),
                             C.byref(rect) if rect is not None else None, ptable, np_,
                             C.byref(nseg), C.byref(dbg))
        if rc != 0:
            raise RuntimeError("oracle render failed (reference panic condition)")
        if buffer.flusher is not None:
            for t in np.nonzero((tile_ops >= 0) & (tile_ops != OP_NONE))[0]:
                j, i = divmod(int(t), row_len)
                for y in range(8 * j, min(8 * j + 8, rows_px)):
                    x0 = y * stride + 8 * i
                    x1 = y * stride + min(8 * i + 8, width)
                    buffer.flusher(data[x0:x1])
        if cache is not None:
            for l in self.layers_.values():
                l.set_is_unchanged(cache.id, l.is_enabled_)
        self.last_segment_count = nseg.value
        self.last_tile_ops = tile_ops
        return nseg.value
