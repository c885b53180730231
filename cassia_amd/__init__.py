"""Python binding of the MI355X mold renderer (include/mold_gpu.h).

Mirrors mold's public scene API (mold/src/lib.rs:15-25): Composition, Layer,
LayerId, Path, Point, Props, Style, Fill, FillRule, Func, BlendMode,
Gradient*, Rect, Buffer, BufferLayerCache, clear_buffer. Every render runs
on the GPU through ``libmoldgpu.so``; importing fails loudly when the library
has not been built, and rendering raises when no gfx950 device is present.
"""
import ctypes as C
import itertools
import enum
import os
from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmoldgpu.so")

TILE_WIDTH = 8
TILE_HEIGHT = 8
NONE = 1 << 63


class MoldError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"mold_gpu status {status}: {msg}")
        self.status = status


class ScaleError(MoldError, ValueError):
    pass


class GradientStopError(MoldError, ValueError):
    pass


if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
                      " (or python cassia_amd/build.py)")
_lib = C.CDLL(LIB_PATH)


class _Props(C.Structure):
    _fields_ = [("fill_rule", C.c_uint32), ("func", C.c_uint32), ("clip_layers", C.c_uint32),
                ("is_clipped", C.c_uint32), ("style_fill_rule", C.c_uint32), ("blend_mode", C.c_uint32),
                ("fill_kind", C.c_uint32), ("color", C.c_float * 4), ("gradient_type", C.c_uint32),
                ("gradient_start", C.c_float * 2), ("gradient_end", C.c_float * 2),
                ("n_stops", C.c_uint32), ("stops", C.POINTER(C.c_float))]


class _Rect(C.Structure):
    _fields_ = [("x0", C.c_uint32), ("x1", C.c_uint32), ("y0", C.c_uint32), ("y1", C.c_uint32)]


FLUSHER = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_uint8), C.c_size_t)


class _Buffer(C.Structure):
    _fields_ = [("rgba", C.POINTER(C.c_uint8)), ("len_px", C.c_size_t), ("width", C.c_size_t),
                ("width_stride", C.c_size_t), ("layer_cache", C.c_int32), ("flusher", FLUSHER),
                ("flusher_user", C.c_void_p)]


class FrameStats(C.Structure):
    _fields_ = [("lines", C.c_uint64), ("line_pixels", C.c_uint64), ("segments", C.c_uint64),
                ("segments_painted", C.c_uint64), ("tiles", C.c_uint64), ("sort_passes", C.c_uint32),
                ("radix_bits", C.c_uint32), ("device_ms", C.c_double), ("setup_ms", C.c_double),
                ("emit_ms", C.c_double), ("hist_ms", C.c_double), ("sort_ms", C.c_double),
                ("runs_ms", C.c_double), ("walk_ms", C.c_double), ("paint_ms", C.c_double),
                ("runs", C.c_uint64), ("records", C.c_uint64), ("walk_threads", C.c_uint32), ("deep_tiles", C.c_uint32),
                ("walk_lds", C.c_uint32), ("walk_rewalks", C.c_uint32), ("sort_pass_ms", C.c_double * 16)]

WALK_SHARE, WALK_HALF, WALK_WHOLE, WALK_GLOBAL = 0, 1, 2, 3


def _props_dtype():
    """numpy record type with the layout of mg_props (bulk props arrays)."""
    names, formats, offsets = [], [], []
    for name, _t in _Props._fields_:
        names.append(name)
        offsets.append(getattr(_Props, name).offset)
        if name == "stops":
            formats.append(np.uint64)
        elif name == "color":
            formats.append((np.float32, 4))
        elif name in ("gradient_start", "gradient_end"):
            formats.append((np.float32, 2))
        else:
            formats.append(np.uint32)
    return np.dtype(dict(names=names, formats=formats, offsets=offsets, itemsize=C.sizeof(_Props)))


_PROPS_DTYPE = _props_dtype()


class _BatchFrame(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("background", C.c_float * 4),
                ("has_crop", C.c_int32), ("crop", _Rect)]


# numpy view of an array of _BatchFrame
_BATCH_FRAME_DTYPE = np.dtype({"names": ["width", "height", "background", "has_crop", "crop"],
                               "formats": ["<u4", "<u4", ("<f4", 4), "<i4", ("<u4", 4)],
                               "offsets": [_BatchFrame.width.offset, _BatchFrame.height.offset,
                                           _BatchFrame.background.offset, _BatchFrame.has_crop.offset,
                                           _BatchFrame.crop.offset],
                               "itemsize": C.sizeof(_BatchFrame)})


class BatchTuning(C.Structure):
    _fields_ = [("n_candidates", C.c_uint32), ("radix_bits", C.c_uint32 * 4), ("device_ms", C.c_double * 4),
                ("chosen_radix_bits", C.c_uint32), ("n_walk_candidates", C.c_uint32), ("walk_threads", C.c_uint32 * 3),
                ("_pad", C.c_uint32), ("walk_ms", C.c_double * 3), ("chosen_walk_threads", C.c_uint32),
                ("_pad2", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("engines", C.c_uint32), ("max_batch", C.c_uint32)]


class RequestResult(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("status", C.c_int32), ("batch_size", C.c_uint32), ("engine", C.c_uint32),
                ("_pad", C.c_uint32), ("latency_ms", C.c_double), ("queue_ms", C.c_double),
                ("footprint", C.c_uint64)]


class ServerCounters(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("failed", C.c_uint64), ("batches", C.c_uint64),
                ("mean_batch", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("max_concurrent_batches", C.c_uint32), ("_pad", C.c_uint32), ("queued", C.c_uint64),
                ("inflight_bytes", C.c_uint64)]


class CassiaStyling(C.Structure):
    _fields_ = [("fill", C.c_float * 4), ("fillRule", C.c_uint32), ("blendMode", C.c_uint32),
                ("_padding", C.c_uint32 * 2)]


_f = C.POINTER(C.c_float)
_vp = C.c_void_p


def _sig(name, res, *args):
    fn = getattr(_lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


_sig("mg_last_error", C.c_char_p)
_sig("mg_device_count", C.c_int)
_sig("mg_path_create", C.c_int, C.POINTER(_vp))
_sig("mg_path_destroy", None, _vp)
for _n in ("line", "quad", "rat_quad", "cubic", "rat_cubic"):
    _sig("mg_path_" + _n, C.c_int, _vp, _f)
_sig("mg_path_close", C.c_int, _vp)
_sig("mg_path_flatten", C.c_int, _vp, _f, _f, C.c_size_t, C.POINTER(C.c_size_t))
_sig("mg_composition_create", C.c_int, C.POINTER(_vp))
_sig("mg_composition_destroy", None, _vp)
_sig("mg_create_layer", C.c_int, _vp, C.POINTER(C.c_uint16))
_sig("mg_insert_in_layer", C.c_int, _vp, C.c_uint16, _vp)
_sig("mg_insert_in_layer_transformed", C.c_int, _vp, C.c_uint16, _vp, _f)
_sig("mg_insert_lines", C.c_int, _vp, C.c_uint16, _f, C.c_size_t)
_sig("mg_insert_lines_multi", C.c_int, _vp, C.POINTER(C.c_uint16), _f, C.c_size_t)
_sig("mg_layer_exists", C.c_int, _vp, C.c_uint16, C.POINTER(C.c_int))
_sig("mg_layer_is_enabled", C.c_int, _vp, C.c_uint16, C.POINTER(C.c_int))
_sig("mg_layer_set_enabled", C.c_int, _vp, C.c_uint16, C.c_int)
_sig("mg_layer_transform", C.c_int, _vp, C.c_uint16, _f)
_sig("mg_layer_set_transform", C.c_int, _vp, C.c_uint16, _f)
_sig("mg_layer_order", C.c_int, _vp, C.c_uint16, C.POINTER(C.c_uint16))
_sig("mg_layer_set_order", C.c_int, _vp, C.c_uint16, C.c_uint16)
_sig("mg_layer_props", C.c_int, _vp, C.c_uint16, C.POINTER(_Props))
_sig("mg_layer_set_props", C.c_int, _vp, C.c_uint16, C.POINTER(_Props))
_sig("mg_create_layers", C.c_int, _vp, C.c_size_t, C.POINTER(C.c_uint16))
_sig("mg_layers_set_props", C.c_int, _vp, C.POINTER(C.c_uint16), _vp, C.c_size_t)
_sig("mg_layers_set_transform", C.c_int, _vp, C.POINTER(C.c_uint16), _f, C.c_size_t)
_sig("mg_insert_paths_device", C.c_int, _vp, C.POINTER(_vp), C.POINTER(C.c_uint16), _f, C.POINTER(C.c_uint8),
     C.c_size_t)
_sig("mg_insert_paths", C.c_int, _vp, C.POINTER(_vp), C.POINTER(C.c_uint16), _f, C.POINTER(C.c_uint8),
     C.c_size_t)
_sig("mg_lines", C.c_int, _vp, _f, C.POINTER(C.c_uint16), C.c_size_t, C.POINTER(C.c_size_t))
_sig("mg_layers_set_order", C.c_int, _vp, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16), C.c_size_t)
_sig("mg_render_animated", C.c_int, _vp, C.c_uint32, C.c_uint32, _f, C.POINTER(_Rect), C.POINTER(_vp),
     C.POINTER(C.c_int))
_sig("mg_copy_to_host", C.c_int, _vp, _vp, C.c_size_t)
_sig("mg_stream_copy", C.c_int, C.c_size_t, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_uint64))
_sig("mg_remove_disabled", C.c_int, _vp)
_sig("mg_lines_len", C.c_int, _vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t))
_sig("mg_create_buffer_layer_cache", C.c_int, _vp, C.POINTER(C.c_int32))
_sig("mg_buffer_layer_cache_clear", C.c_int, _vp, C.c_int32)
_sig("mg_buffer_layer_cache_release", C.c_int, _vp, C.c_int32)
_sig("mg_render", C.c_int, _vp, C.POINTER(_Buffer), _f, C.POINTER(_Rect))
_sig("mg_render_device", C.c_int, _vp, C.c_uint32, C.c_uint32, _f, C.POINTER(_Rect), C.POINTER(_vp))
_sig("mg_render_coverage", C.c_int, _vp, C.c_uint32, C.c_uint32, _f, C.POINTER(_Rect), C.POINTER(C.c_int32),
     C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_size_t))
_sig("mg_set_tuning", C.c_int, _vp, C.c_uint32)
_sig("mg_set_walk_threads", C.c_int, _vp, C.c_uint32)
_sig("mg_batch_set_walk_threads", C.c_int, _vp, C.c_uint32)
_sig("mg_pixel_segments", C.c_int, _vp, C.POINTER(C.POINTER(C.c_uint64)), C.POINTER(C.c_size_t))
_sig("mg_last_frame_stats", C.c_int, _vp, C.POINTER(FrameStats))
_sig("mg_batch_create", C.c_int, C.POINTER(_vp))
_sig("mg_batch_destroy", None, _vp)
_sig("mg_batch_render", C.c_int, _vp, C.POINTER(_vp), C.POINTER(_BatchFrame), C.c_size_t, C.POINTER(_vp),
     C.POINTER(C.POINTER(C.c_uint64)))
_sig("mg_batch_read", C.c_int, _vp, C.c_size_t, C.POINTER(C.c_uint8), C.c_size_t)
_sig("mg_batch_set_tuning", C.c_int, _vp, C.c_uint32)
_sig("mg_batch_last_tuning", C.c_int, _vp, C.POINTER(BatchTuning))
_sig("mg_batch_stats", C.c_int, _vp, C.POINTER(FrameStats))
_sig("mg_server_create", C.c_int, C.POINTER(ServerConfig), C.POINTER(_vp))
_sig("mg_server_destroy", None, _vp)
_sig("mg_server_config_of", C.c_int, _vp, C.POINTER(ServerConfig))
_sig("mg_server_submit", C.c_int, _vp, _vp, C.POINTER(_BatchFrame), C.POINTER(C.c_uint8), C.c_size_t,
     C.POINTER(C.c_uint64))
_sig("mg_server_wait", C.c_int, _vp, C.c_uint64, C.POINTER(RequestResult))
_sig("mg_server_stats", C.c_int, _vp, C.c_int, C.POINTER(ServerCounters))
_sig("cassia_init", None, C.c_uint32, C.c_uint32)
_sig("cassia_render", None, C.POINTER(C.c_uint64), C.c_size_t, C.POINTER(CassiaStyling), C.c_size_t)
_sig("cassia_shutdown", None)
_sig("cassia_read_rgba8", C.c_int, C.POINTER(C.c_uint8), C.c_size_t)


def _check(status):
    if status != 0:
        msg = _lib.mg_last_error().decode(errors="replace")
        if status == 2:
            raise ScaleError(status, msg)
        if status == 8:
            raise GradientStopError(status, msg)
        raise MoldError(status, msg)


def device_count():
    return _lib.mg_device_count()


def last_error():
    return _lib.mg_last_error().decode(errors="replace")


def _f32(v, n=None):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1))
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} floats")
    return a


def _fp(a):
    return a.ctypes.data_as(_f)


def _f32t(v):
    return tuple(float(x) for x in np.asarray(v, dtype=np.float32).reshape(-1))


# ------------------------------------------------------------------ styles

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


@dataclass(frozen=True)
class Gradient:
    type: GradientType
    start: Tuple[float, float]
    end: Tuple[float, float]
    stops: Tuple


class GradientBuilder:
    """surpass::painter::GradientBuilder (style.rs:29-85)"""

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
            raise GradientStopError(8, "gradient stops must be between 0.0 and 1.0")
        self._stops.append((_f32t(color), stop))
        return self

    def build(self):
        if len(self._stops) < 2:
            return None
        return Gradient(self._type, self._start, self._end, tuple(self._stops))


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


def _to_c_props(p: Props, keep):
    c = _Props()
    c.fill_rule = int(p.fill_rule)
    if p.func.kind == "clip":
        c.func = 1
        c.clip_layers = p.func.layers & 0xFFFFFFFF
        return c
    s = p.func.style
    c.func = 0
    c.is_clipped = 1 if s.is_clipped else 0
    c.style_fill_rule = int(s.fill_rule)
    c.blend_mode = int(s.blend_mode)
    if s.fill.kind == "solid":
        c.fill_kind = 0
        for i in range(4):
            c.color[i] = s.fill.color[i]
    else:
        g = s.fill.gradient
        c.fill_kind = 1
        c.gradient_type = int(g.type)
        c.gradient_start[0], c.gradient_start[1] = g.start
        c.gradient_end[0], c.gradient_end[1] = g.end
        stops = np.array([list(col) + [st] for col, st in g.stops], dtype=np.float32).reshape(-1)
        keep.append(stops)
        c.n_stops = len(g.stops)
        c.stops = _fp(stops)
    return c


@dataclass(frozen=True)
class Point:
    x: float
    y: float


# ------------------------------------------------------------------- paths

class Path:
    """mold::Path (mold/src/path/mod.rs:24-179) over an mg_path handle."""

    def __init__(self):
        h = _vp()
        _check(_lib.mg_path_create(C.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.mg_path_destroy(self._h)
            self._h = None

    def _cmd(self, name, vals):
        a = _f32(vals)
        _check(getattr(_lib, "mg_path_" + name)(self._h, _fp(a)))
        return self

    def line(self, p0, p1):
        return self._cmd("line", [p0.x, p0.y, p1.x, p1.y])

    def quad(self, p0, p1, p2):
        return self._cmd("quad", [p0.x, p0.y, p1.x, p1.y, p2.x, p2.y])

    def rat_quad(self, p0, p1, p2):
        return self._cmd("rat_quad", [v for p, w in (p0, p1, p2) for v in (p.x, p.y, w)])

    def cubic(self, p0, p1, p2, p3):
        return self._cmd("cubic", [v for p in (p0, p1, p2, p3) for v in (p.x, p.y)])

    def rat_cubic(self, p0, p1, p2, p3):
        return self._cmd("rat_cubic", [v for p, w in (p0, p1, p2, p3) for v in (p.x, p.y, w)])

    def close(self):
        _check(_lib.mg_path_close(self._h))
        return self

    def segments(self, transform=None):
        t = _f32(transform, 9) if transform is not None else None
        n = C.c_size_t(0)
        _check(_lib.mg_path_flatten(self._h, _fp(t) if t is not None else None, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 4), np.float32)
        _check(_lib.mg_path_flatten(self._h, _fp(t) if t is not None else None, _fp(out), n.value, C.byref(n)))
        return out[:n.value]


# ------------------------------------------------------------- composition

class LayerId(int):
    pass


class Layer:
    """Handle on mold::Layer (mold/src/layer.rs:152-231); setters chain."""

    def __init__(self, comp, layer_id):
        self._c = comp
        self.id = int(layer_id)

    def is_enabled(self):
        v = C.c_int(0)
        _check(_lib.mg_layer_is_enabled(self._c._h, self.id, C.byref(v)))
        return bool(v.value)

    def set_is_enabled(self, enabled):
        _check(_lib.mg_layer_set_enabled(self._c._h, self.id, 1 if enabled else 0))
        return self

    def disable(self):
        return self.set_is_enabled(False)

    def enable(self):
        return self.set_is_enabled(True)

    def transform(self):
        out = np.zeros(6, np.float32)
        _check(_lib.mg_layer_transform(self._c._h, self.id, _fp(out)))
        return tuple(float(x) for x in out)

    def set_transform(self, t):
        a = _f32(t, 6)
        _check(_lib.mg_layer_set_transform(self._c._h, self.id, _fp(a)))
        return self

    def order(self):
        v = C.c_uint16(0)
        _check(_lib.mg_layer_order(self._c._h, self.id, C.byref(v)))
        return v.value

    def set_order(self, order):
        _check(_lib.mg_layer_set_order(self._c._h, self.id, int(order) & 0xFFFF))
        return self

    def props(self):
        p = self._c._props.get(self.id)
        if p is None:
            p = self._c._bulk_props(self.id)
        return p if p is not None else Props()

    def set_props(self, props):
        keep = []
        c = _to_c_props(props, keep)
        _check(_lib.mg_layer_set_props(self._c._h, self.id, C.byref(c)))
        self._c._props[self.id] = props
        return self


class BufferLayerCache:
    """mold::BufferLayerCache (buffer.rs:46-72)."""

    def __init__(self, comp, cache_id):
        self._c = comp
        self.id = cache_id

    def clear(self):
        _check(_lib.mg_buffer_layer_cache_clear(self._c._h, self.id))

    def drop(self):
        _check(_lib.mg_buffer_layer_cache_release(self._c._h, self.id))


class Rect:
    """surpass::painter::Rect::new(horizontal, vertical) in pixels (buffer_layout.rs:43-50)."""

    def __init__(self, horizontal: range, vertical: range):
        self.px = (horizontal.start, horizontal.stop, vertical.start, vertical.stop)
        self.horizontal = range(horizontal.start // TILE_WIDTH, (horizontal.stop + TILE_WIDTH - 1) // TILE_WIDTH)
        self.vertical = range(vertical.start // TILE_HEIGHT, (vertical.stop + TILE_HEIGHT - 1) // TILE_HEIGHT)


class Buffer:
    """mold::Buffer (buffer.rs:17-25); `buffer` is a (len, 4) uint8 numpy array."""

    def __init__(self, buffer, width, width_stride=None, layer_cache=None, flusher=None):
        self.buffer = buffer
        self.width = width
        self.width_stride = width_stride
        self.layer_cache = layer_cache
        self.flusher = flusher


def clear_buffer(buffer, color):
    """mold::clear_buffer (mold/src/utils.rs:5-9)"""
    buffer[:] = np.asarray(color, dtype=np.uint8)


class Composition:
    """mold::Composition (composition.rs:25-329) over an mg_composition."""

    def __init__(self):
        h = _vp()
        _check(_lib.mg_composition_create(C.byref(h)))
        self._h = h
        self._props = {}
        self._bulk = []  # (id -> row, colors, rules, blends) of set_solid_props calls, latest last

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.mg_composition_destroy(self._h)
            self._h = None

    def create_layer(self):
        v = C.c_uint16(0)
        st = _lib.mg_create_layer(self._h, C.byref(v))
        if st == 3:
            return None
        _check(st)
        return LayerId(v.value)

    def insert_in_layer(self, layer_id, path):
        _check(_lib.mg_insert_in_layer(self._h, int(layer_id), path._h))
        return Layer(self, layer_id)

    def insert_in_layer_transformed(self, layer_id, path, transform):
        t = _f32(transform, 9)
        _check(_lib.mg_insert_in_layer_transformed(self._h, int(layer_id), path._h, _fp(t)))
        return Layer(self, layer_id)

    def insert_lines(self, layer_id, lines):
        a = np.ascontiguousarray(np.asarray(lines, dtype=np.float32).reshape(-1, 4))
        _check(_lib.mg_insert_lines(self._h, int(layer_id), _fp(a), len(a)))
        return Layer(self, layer_id)

    def insert_lines_multi(self, layer_ids, lines):
        a = np.ascontiguousarray(np.asarray(lines, dtype=np.float32).reshape(-1, 4))
        ids = np.ascontiguousarray(np.asarray(layer_ids, dtype=np.uint16).reshape(-1))
        assert len(ids) == len(a)
        _check(_lib.mg_insert_lines_multi(self._h, ids.ctypes.data_as(C.POINTER(C.c_uint16)), _fp(a), len(a)))

    def insert_paths_device(self, layer_ids, paths, transforms=None, device=True):
        """insert_in_layer[_transformed] of many paths, flattened on the GPU
        (mg_insert_paths_device; device=False: mg_insert_paths, on the host);
        transforms[i] is a 9-float transform or None."""
        n = len(paths)
        if n != len(layer_ids):
            raise ValueError("one layer id per path")
        if n == 0:
            return
        hs = (_vp * n)(*[p._h for p in paths])
        ids = np.ascontiguousarray(np.asarray(layer_ids, dtype=np.uint16).reshape(-1))
        t, flags = None, None
        if transforms is not None:
            t = np.zeros((n, 9), np.float32)
            flags = np.zeros(n, np.uint8)
            for i, tr in enumerate(transforms):
                if tr is not None:
                    t[i] = np.asarray(tr, np.float32).reshape(9)
                    flags[i] = 1
        fn = _lib.mg_insert_paths_device if device else _lib.mg_insert_paths
        _check(fn(self._h, hs, ids.ctypes.data_as(C.POINTER(C.c_uint16)), _fp(t) if t is not None else None,
                  flags.ctypes.data_as(C.POINTER(C.c_uint8)) if flags is not None else None, n))

    def lines(self):
        """The composition's lines: ((n, 4) float32 x0, y0, x1, y1; (n,) uint16 layer ids)."""
        n = C.c_size_t(0)
        _check(_lib.mg_lines(self._h, None, None, 0, C.byref(n)))
        pts = np.zeros((max(n.value, 1), 4), np.float32)
        ids = np.zeros(max(n.value, 1), np.uint16)
        _check(_lib.mg_lines(self._h, _fp(pts), ids.ctypes.data_as(C.POINTER(C.c_uint16)), n.value, C.byref(n)))
        return pts[:n.value], ids[:n.value]

    def create_layers(self, n):
        """n create_layer calls in one: returns the ids (uint16 array)."""
        out = np.zeros(n, np.uint16)
        _check(_lib.mg_create_layers(self._h, n, out.ctypes.data_as(C.POINTER(C.c_uint16))))
        return out

    def set_solid_props(self, layer_ids, colors, fill_rules, blend_modes):
        """Layer::set_props on many layers with Draw(Style{fill: Solid(colors[i]),
        blend_mode}) props and fill rule `fill_rules[i]`, in one call."""
        ids = np.ascontiguousarray(np.asarray(layer_ids, dtype=np.uint16).reshape(-1))
        n = len(ids)
        a = np.zeros(n, _PROPS_DTYPE)
        a["fill_rule"] = np.broadcast_to(np.asarray(fill_rules, np.uint32), (n,))
        a["blend_mode"] = np.broadcast_to(np.asarray(blend_modes, np.uint32), (n,))
        a["color"] = np.asarray(colors, np.float32).reshape(n, 4)
        _check(_lib.mg_layers_set_props(self._h, ids.ctypes.data_as(C.POINTER(C.c_uint16)), a.ctypes.data, n))
        for i in ids.tolist():
            self._props.pop(i, None)
        self._bulk.append(({int(v): k for k, v in enumerate(ids.tolist())}, a))

    def _bulk_props(self, layer_id):
        for rows, a in reversed(self._bulk):
            k = rows.get(int(layer_id))
            if k is not None:
                r = a[k]
                style = Style(fill=Fill.Solid([float(v) for v in r["color"]]), blend_mode=BlendMode(int(r["blend_mode"])))
                return Props(fill_rule=FillRule(int(r["fill_rule"])), func=Func.Draw(style))
        return None

    def get(self, layer_id):
        v = C.c_int(0)
        _check(_lib.mg_layer_exists(self._h, int(layer_id), C.byref(v)))
        return Layer(self, layer_id) if v.value else None

    get_mut = get

    def builder_len(self):
        b, a = C.c_size_t(0), C.c_size_t(0)
        _check(_lib.mg_lines_len(self._h, C.byref(b), C.byref(a)))
        return b.value

    def actual_len(self):
        b, a = C.c_size_t(0), C.c_size_t(0)
        _check(_lib.mg_lines_len(self._h, C.byref(b), C.byref(a)))
        return a.value

    def remove_disabled(self):
        _check(_lib.mg_remove_disabled(self._h))

    def create_buffer_layer_cache(self):
        v = C.c_int32(-1)
        st = _lib.mg_create_buffer_layer_cache(self._h, C.byref(v))
        if st == 10:
            return None
        _check(st)
        return BufferLayerCache(self, v.value)

    def pixel_segments(self):
        p = C.POINTER(C.c_uint64)()
        n = C.c_size_t(0)
        _check(_lib.mg_pixel_segments(self._h, C.byref(p), C.byref(n)))
        if n.value == 0:
            return np.zeros(0, np.uint64)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    def render(self, buffer: Buffer, background_color, crop: Optional[Rect] = None):
        data = buffer.buffer
        assert data.dtype == np.uint8 and data.flags["C_CONTIGUOUS"]
        b = _Buffer()
        b.rgba = data.ctypes.data_as(C.POINTER(C.c_uint8))
        b.len_px = len(data)
        b.width = buffer.width
        b.width_stride = buffer.width_stride or 0
        b.layer_cache = buffer.layer_cache.id if buffer.layer_cache is not None else -1
        cb = None
        if buffer.flusher is not None:
            user = buffer.flusher

            def _flush(_u, ptr, n):
                arr = np.ctypeslib.as_array(ptr, shape=(n * 4,)).reshape(n, 4)
                user(arr)

            cb = FLUSHER(_flush)
            b.flusher = cb
        bg = _f32(background_color, 4)
        r = None
        if crop is not None:
            r = _Rect(*crop.px[:2], *crop.px[2:])
        _check(_lib.mg_render(self._h, C.byref(b), _fp(bg), C.byref(r) if r is not None else None))
        return self.stats().segments

    def render_device(self, width, height, background_color, crop: Optional[Rect] = None):
        """Render into a device-resident frame; returns its device address."""
        bg = _f32(background_color, 4)
        r = _Rect(*crop.px[:2], *crop.px[2:]) if crop is not None else None
        out = _vp()
        _check(_lib.mg_render_device(self._h, width, height, _fp(bg), C.byref(r) if r is not None else None,
                                     C.byref(out)))
        return out.value

    def render_coverage(self, width, height, background_color, crop: Optional[Rect] = None, cap=1 << 16):
        """Per-pixel coverage values A of every (tile, workbench id) of a render:
        returns (A int32 (n, 64) column-major x*8+y, keys uint32 (n, 4) =
        (scene, tile row, tile column, layer)), in no particular order."""
        bg = _f32(background_color, 4)
        r = _Rect(*crop.px[:2], *crop.px[2:]) if crop is not None else None
        a = np.zeros((cap, 64), np.int32)
        k = np.zeros((cap, 4), np.uint32)
        n = C.c_size_t(0)
        _check(_lib.mg_render_coverage(self._h, width, height, _fp(bg), C.byref(r) if r is not None else None,
                                       a.ctypes.data_as(C.POINTER(C.c_int32)),
                                       k.ctypes.data_as(C.POINTER(C.c_uint32)), cap, C.byref(n)))
        if n.value > cap:
            return self.render_coverage(width, height, background_color, crop, cap=int(n.value))
        return a[:n.value], k[:n.value]

    def render_animated(self, width, height, background_color, crop: Optional[Rect] = None):
        """render_device through a captured hipGraph (mg_render_animated):
        returns (device address of the frame, whether the graph was replayed)."""
        bg = _f32(background_color, 4)
        r = _Rect(*crop.px[:2], *crop.px[2:]) if crop is not None else None
        out = _vp()
        rep = C.c_int(0)
        _check(_lib.mg_render_animated(self._h, width, height, _fp(bg), C.byref(r) if r is not None else None,
                                       C.byref(out), C.byref(rep)))
        return out.value, bool(rep.value)

    def set_transforms(self, layer_ids, transforms):
        """Layer::set_transform on many layers: transforms is (n, 6) float32."""
        ids = np.ascontiguousarray(np.asarray(layer_ids, dtype=np.uint16).reshape(-1))
        t = np.ascontiguousarray(np.asarray(transforms, dtype=np.float32).reshape(len(ids), 6))
        _check(_lib.mg_layers_set_transform(self._h, ids.ctypes.data_as(C.POINTER(C.c_uint16)), _fp(t), len(ids)))

    def set_orders(self, layer_ids, orders):
        """Layer::set_order on many layers."""
        ids = np.ascontiguousarray(np.asarray(layer_ids, dtype=np.uint16).reshape(-1))
        o = np.ascontiguousarray(np.asarray(orders, dtype=np.uint16).reshape(len(ids)))
        _check(_lib.mg_layers_set_order(self._h, ids.ctypes.data_as(C.POINTER(C.c_uint16)),
                                        o.ctypes.data_as(C.POINTER(C.c_uint16)), len(ids)))

    def set_tuning(self, radix_bits=0):
        """Result-invariant knob: the sort's digit width in bits (4, 6, 8, 9; 0 = automatic)."""
        _check(_lib.mg_set_tuning(self._h, radix_bits))

    def set_walk_threads(self, threads=0):
        """Result-invariant knob: threads per tile row of the row walk (64, 256, 1024; 0 = automatic)."""
        _check(_lib.mg_set_walk_threads(self._h, threads))

    def stats(self):
        s = FrameStats()
        _check(_lib.mg_last_frame_stats(self._h, C.byref(s)))
        return s


def device_frame(address, width, height):
    """Copy of a device-resident RGBA8 frame as a (height * width, 4) uint8 array."""
    out = np.zeros((width * height, 4), np.uint8)
    _check(_lib.mg_copy_to_host(out.ctypes.data, address, out.nbytes))
    return out


def stream_copy(nbytes, repeats=1):
    """`repeats` streaming device-to-device copies of `nbytes` between buffers
    the library allocates (mg_stream_copy): (average milliseconds of one copy,
    32-bit words found wrong afterwards)."""
    ms, bad = C.c_double(), C.c_uint64()
    _check(_lib.mg_stream_copy(nbytes, repeats, C.byref(ms), C.byref(bad)))
    return ms.value, bad.value


class Batch:
    """Many compositions rendered as one device frame (mg_batch_render): frame
    i of a render is the bytes `comps[i].render_device(w, h, bg, crop)` gives,
    inside the crop."""

    def __init__(self):
        h = _vp()
        _check(_lib.mg_batch_create(C.byref(h)))
        self._h = h
        self._frames = []
        self._offsets = []

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.mg_batch_destroy(self._h)
            self._h = None

    def render(self, comps, frames):
        """`frames[i]` = (width, height, background[, crop Rect]). Returns the
        device address of the frame buffer; `offsets` gives each frame's byte
        offset in it."""
        n = len(comps)
        if n != len(frames):
            raise ValueError("one frame per composition")
        # the tables are packed with numpy (a per-field ctypes loop over
        # thousands of frames costs milliseconds of host time per render)
        hs = np.fromiter((c._h.value for c in comps), dtype=np.uint64, count=n)
        fs = np.zeros(max(n, 1), dtype=_BATCH_FRAME_DTYPE)
        if n:
            chain = itertools.chain.from_iterable
            wh = np.fromiter(chain((fr[0], fr[1]) for fr in frames), dtype=np.uint32, count=2 * n)
            fs["width"][:n] = wh[0::2]
            fs["height"][:n] = wh[1::2]
            fs["background"][:n] = np.fromiter(chain(fr[2] for fr in frames), dtype=np.float32,
                                               count=4 * n).reshape(n, 4)
            for i, fr in enumerate(frames):
                if len(fr) > 3 and fr[3] is not None:
                    px = fr[3].px
                    fs["has_crop"][i] = 1
                    fs["crop"][i] = (px[0], px[1], px[2], px[3])
        dev = _vp()
        offs = C.POINTER(C.c_uint64)()
        _check(_lib.mg_batch_render(self._h, hs.ctypes.data_as(C.POINTER(_vp)),
                                    fs.ctypes.data_as(C.POINTER(_BatchFrame)), n, C.byref(dev), C.byref(offs)))
        self._frames = [(fr[0], fr[1]) for fr in frames]
        self._offsets = np.ctypeslib.as_array(offs, shape=(n,)).tolist() if n else []
        return dev.value

    @property
    def offsets(self):
        return list(self._offsets)

    def read(self, i):
        """Frame i of the last render as a (height * width, 4) uint8 array."""
        if not 0 <= i < len(self._frames):
            raise IndexError(f"frame {i} of a batch of {len(self._frames)}")
        w, h = self._frames[i]
        out = np.zeros((w * h, 4), np.uint8)
        _check(_lib.mg_batch_read(self._h, i, out.ctypes.data_as(C.POINTER(C.c_uint8)), out.nbytes))
        return out

    def set_tuning(self, radix_bits=0, walk_threads=None):
        """Sort digit width (4, 6, 8 or 9) and, if given, row-walk threads (64,
        256, 1024); 0 = autotune on the next render."""
        _check(_lib.mg_batch_set_tuning(self._h, radix_bits))
        if walk_threads is not None:
            _check(_lib.mg_batch_set_walk_threads(self._h, walk_threads))

    def tuning(self):
        t = BatchTuning()
        _check(_lib.mg_batch_last_tuning(self._h, C.byref(t)))
        return t

    def stats(self):
        s = FrameStats()
        _check(_lib.mg_batch_stats(self._h, C.byref(s)))
        return s


class Server:
    """Render server (mg_server_*): requests admitted by a device-memory
    budget and rendered as batched frames by `engines` workers, each on its
    own HIP stream (continuous batching)."""

    def __init__(self, hbm_budget=0, engines=2, max_batch=8192):
        cfg = ServerConfig(hbm_budget, engines, max_batch)
        h = _vp()
        _check(_lib.mg_server_create(C.byref(cfg), C.byref(h)))
        self._h = h
        self._live = {}

    def close(self):
        if getattr(self, "_h", None):
            _lib.mg_server_destroy(self._h)
            self._h = None
            self._live.clear()

    __del__ = close

    def config(self):
        c = ServerConfig()
        _check(_lib.mg_server_config_of(self._h, C.byref(c)))
        return c

    def submit(self, comp, width, height, background, crop: Optional[Rect] = None, out=None):
        """Queue a render of `comp` (kept unchanged until wait); returns the ticket."""
        fr = _BatchFrame()
        fr.width, fr.height = width, height
        for k in range(4):
            fr.background[k] = float(background[k])
        if crop is not None:
            fr.has_crop = 1
            fr.crop = _Rect(*crop.px[:2], *crop.px[2:])
        if out is None:
            out = np.zeros((width * height, 4), np.uint8)
        t = C.c_uint64(0)
        _check(_lib.mg_server_submit(self._h, comp._h, C.byref(fr), out.ctypes.data_as(C.POINTER(C.c_uint8)),
                                     out.nbytes, C.byref(t)))
        self._live[t.value] = (comp, out)
        return t.value

    def wait(self, ticket):
        """Blocks until the request is done: (frame (width * height, 4) uint8, RequestResult)."""
        r = RequestResult()
        st = _lib.mg_server_wait(self._h, ticket, C.byref(r))
        comp, out = self._live.pop(ticket, (None, None))
        _check(st)
        return out, r

    def stats(self, reset=False):
        c = ServerCounters()
        _check(_lib.mg_server_stats(self._h, 1 if reset else 0, C.byref(c)))
        return c


# -------------------------------------------------------- cassia-compatible

def cassia_init(width, height):
    _lib.cassia_init(width, height)


def cassia_render(psegments, stylings):
    segs = np.ascontiguousarray(psegments, dtype=np.uint64)
    sty = (CassiaStyling * max(1, len(stylings)))()
    for i, s in enumerate(stylings):
        for k in range(4):
            sty[i].fill[k] = float(s[0][k])
        sty[i].fillRule = int(s[1])
        sty[i].blendMode = int(s[2])
    _lib.cassia_render(segs.ctypes.data_as(C.POINTER(C.c_uint64)), len(segs), sty, len(stylings))


def cassia_read(width, height):
    out = np.zeros((height * width, 4), np.uint8)
    _check(_lib.cassia_read_rgba8(out.ctypes.data_as(C.POINTER(C.c_uint8)), out.nbytes))
    return out


def cassia_shutdown():
    _lib.cassia_shutdown()
