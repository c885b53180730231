"""CPU oracle for the mold/surpass rasterize -> sort -> paint path.

TEST INFRASTRUCTURE ONLY. Only the tests, the smoke check and the bench's
CPU-baseline leg may import this package, and only as the checker (or the
timed CPU baseline). The product never imports it.

Low-level ctypes bindings to ``liboracle.so`` (built from ``oracle.c`` by
``oracle/Makefile``); the mold-shaped model lives in :mod:`oracle.mold`.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NONE = 1 << 63


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


class Layer(C.Structure):
    _fields_ = [("exists", C.c_uint8), ("is_enabled", C.c_uint8), ("has_affine", C.c_uint8),
                ("has_order", C.c_uint8), ("order", C.c_uint16), ("_pad", C.c_uint16),
                ("affine", C.c_float * 6)]


class Props(C.Structure):
    _fields_ = [("fill_rule", C.c_uint32), ("func", C.c_uint32), ("clip_layers", C.c_uint32),
                ("is_clipped", C.c_uint32), ("blend_mode", C.c_uint32), ("fill_kind", C.c_uint32),
                ("color", C.c_float * 4), ("grad_type", C.c_uint32), ("n_stops", C.c_uint32),
                ("grad_start", C.c_float * 2), ("grad_end", C.c_float * 2),
                ("stops", C.POINTER(C.c_float)), ("is_unchanged", C.c_uint32),
                ("defined", C.c_uint32)]


class Carry(C.Structure):
    _fields_ = [("cover", C.c_int8 * 8), ("layer", C.c_uint16), ("_pad", C.c_uint16)]


class TileResult(C.Structure):
    _fields_ = [("op", C.c_int32), ("solid", C.c_float * 4), ("colors", C.c_float * 256),
                ("n_carries_out", C.c_size_t), ("n_ids", C.c_size_t), ("ids", C.c_uint16 * 1024),
                ("masked", C.c_uint8 * 1024), ("skip_clipping", C.c_uint8 * 1024)]


class Rect(C.Structure):
    _fields_ = [("h0", C.c_uint32), ("h1", C.c_uint32), ("v0", C.c_uint32), ("v1", C.c_uint32)]


class Debug(C.Structure):
    _fields_ = [("cov_a", C.POINTER(C.c_int32)), ("cov_key", C.POINTER(C.c_uint32)),
                ("cap_records", C.c_size_t), ("n_records", C.c_size_t),
                ("tile_ops", C.POINTER(C.c_int8))]


class Cmd(C.Structure):
    _fields_ = [("kind", C.c_int32), ("v", C.c_float * 12)]


OP_NONE, OP_SOLID, OP_COLOR_BUFFER = 0, 1, 2

_f = C.POINTER(C.c_float)
_u8 = C.POINTER(C.c_uint8)
_u16 = C.POINTER(C.c_uint16)
_u64 = C.POINTER(C.c_uint64)
_i32 = C.POINTER(C.c_int32)
_sz = C.POINTER(C.c_size_t)


def _declare(L):
    L.or_seg_new.restype = C.c_uint64
    L.or_seg_new.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint32, C.c_uint32, C.c_uint32,
                             C.c_int32, C.c_int32]
    L.or_seg_fields.argtypes = [C.c_uint64, _i32]
    L.or_path_flatten.restype = C.c_size_t
    L.or_path_flatten.argtypes = [C.POINTER(Cmd), C.c_size_t, _f, _f, C.c_size_t]
    L.or_rasterize.restype = C.c_size_t
    L.or_sort_canonical.argtypes = [_u64, C.c_size_t]
    L.or_blend_fn.restype = C.c_float
    L.or_blend_fn.argtypes = [C.c_uint32, C.c_float, C.c_float]
    L.or_blend_simd.argtypes = [C.c_uint32, _f, _f, _f]
    L.or_blend.argtypes = [C.c_uint32, _f, _f, _f]
    L.or_gradient_resolve_stops.argtypes = [_f, C.c_uint32]
    L.or_gradient_color_at.argtypes = [C.POINTER(Props), C.c_float, C.c_float, _f]
    L.or_linear_to_srgb.restype = C.c_float
    L.or_linear_to_srgb.argtypes = [C.c_float]
    L.or_linear_to_srgb_simd.restype = C.c_float
    L.or_linear_to_srgb_simd.argtypes = [C.c_float]
    L.or_to_bytes.argtypes = [_f, _u8]
    L.or_drive_tile.restype = C.c_int
    L.or_drive_tile.argtypes = [C.POINTER(Carry), C.c_size_t, _u64, C.c_size_t, C.POINTER(Props),
                                C.c_size_t, _f, C.c_uint32, C.c_uint32, _i32, C.c_uint32, C.c_int,
                                C.POINTER(Carry), C.c_size_t, C.POINTER(TileResult)]
    L.or_print.restype = C.c_int
    L.or_print.argtypes = [_u8, C.c_size_t, C.c_size_t, C.c_size_t, _i32, C.c_size_t, _u64,
                           C.c_size_t, _f, C.POINTER(Rect), C.POINTER(Props), C.c_size_t,
                           C.POINTER(Debug)]
    L.or_render.restype = C.c_int
    L.or_render.argtypes = [C.c_size_t, _f, _f, _f, _f, _u16, C.POINTER(Layer), C.c_size_t, _u8,
                            C.c_size_t, C.c_size_t, C.c_size_t, _i32, C.c_size_t, _f,
                            C.POINTER(Rect), C.POINTER(Props), C.c_size_t, _sz, C.POINTER(Debug)]
    L.or_render_mt.restype = C.c_int
    L.or_render_mt.argtypes = [C.c_int, C.c_size_t, _f, _f, _f, _f, _u16, C.POINTER(Layer),
                               C.c_size_t, _u8, C.c_size_t, C.c_size_t, C.c_size_t, _f,
                               C.POINTER(Props), C.c_size_t, _sz]
    L.or_pixel_segments.restype = C.c_size_t
    L.or_pixel_segments.argtypes = [C.c_size_t, _f, _f, _f, _f, _u16, C.POINTER(Layer),
                                    C.c_size_t, _u64, C.c_size_t]


def ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def seg_new(is_none, tile_j, tile_i, layer, tile_y, tile_x, area, cover):
    return lib().or_seg_new(is_none, tile_j, tile_i, layer, tile_y, tile_x, area, cover)


def seg_fields(seg):
    out = (C.c_int32 * 8)()
    lib().or_seg_fields(C.c_uint64(int(seg)), out)
    return dict(zip(("is_none", "tile_j", "tile_i", "layer", "tile_y", "tile_x", "area", "cover"),
                    list(out)))


def decode(segs):
    """Vectorised field decode of a uint64 array (same layout as seg_fields)."""
    s = np.asarray(segs, dtype=np.uint64)

    def sx(v, bits):
        v = (v & np.uint64((1 << bits) - 1)).astype(np.int64)
        return np.where(v >> (bits - 1), v - (1 << bits), v)

    return dict(is_none=(s >> np.uint64(63)).astype(np.int64),
                tile_j=sx(s >> np.uint64(51), 12), tile_i=sx(s >> np.uint64(38), 13),
                layer=((s >> np.uint64(22)) & np.uint64(0xFFFF)).astype(np.int64),
                tile_y=((s >> np.uint64(19)) & np.uint64(7)).astype(np.int64),
                tile_x=((s >> np.uint64(16)) & np.uint64(7)).astype(np.int64),
                area=sx(s >> np.uint64(6), 10), cover=sx(s, 6))


def path_flatten(cmds, t9=None):
    """cmds: list of (kind, [floats]). Returns (n, 4) float32 segments."""
    arr = (Cmd * max(1, len(cmds)))()
    for i, (k, v) in enumerate(cmds):
        arr[i].kind = k
        for j, x in enumerate(v):
            arr[i].v[j] = x
    t = f32(t9) if t9 is not None else None
    n = lib().or_path_flatten(arr, len(cmds), ptr(t, _f), None, 0)
    out = np.zeros((max(n, 1), 4), dtype=np.float32)
    lib().or_path_flatten(arr, len(cmds), ptr(t, _f), ptr(out, _f), n)
    return out[:n]


def lines_build(p0x, p0y, p1x, p1y, layer_ids, layers=None, default_transform=None):
    """LinesBuilder::build; returns a dict of SoA arrays."""
    n = len(p0x)
    p = [f32(a) for a in (p0x, p0y, p1x, p1y)]
    ids = np.ascontiguousarray(layer_ids, dtype=np.uint16)
    out = dict(orders=np.zeros(n, np.uint16), starts=np.zeros(n, np.int32),
               starts_f32=np.zeros(n, np.float32), ends_f32=np.zeros(n, np.float32),
               lengths=np.zeros(n, np.int64), slopes=np.zeros(n, np.float32),
               slopes_recip=np.zeros(n, np.float32), py_slope_pxs=np.zeros(n, np.float32),
               px_slope_recip_pys=np.zeros(n, np.float32), octants=np.zeros(n, np.uint8))

    class Lines(C.Structure):
        _fields_ = [("orders", _u16), ("starts", _i32), ("starts_f32", _f), ("ends_f32", _f),
                    ("lengths", C.POINTER(C.c_int64)), ("slopes", _f), ("slopes_recip", _f),
                    ("py_slope_pxs", _f), ("px_slope_recip_pys", _f), ("octants", _u8)]

    ls = Lines(*(out[k].ctypes.data_as(t) for k, t in (
        ("orders", _u16), ("starts", _i32), ("starts_f32", _f), ("ends_f32", _f),
        ("lengths", C.POINTER(C.c_int64)), ("slopes", _f), ("slopes_recip", _f),
        ("py_slope_pxs", _f), ("px_slope_recip_pys", _f), ("octants", _u8))))
    dt = f32(default_transform) if default_transform is not None else None
    lt = layers
    L = lib()
    L.or_lines_build.argtypes = [C.c_size_t, _f, _f, _f, _f, _u16, C.POINTER(Layer), C.c_size_t,
                                 _f, C.POINTER(Lines)]
    L.or_lines_build(n, *(ptr(a, _f) for a in p), ptr(ids, _u16), lt,
                     len(lt) if lt is not None else 0, ptr(dt, _f), C.byref(ls))
    out["_struct"] = ls
    return out


def rasterize(lines):
    """Rasterizer::rasterize on the dict from lines_build (uint64, generation order)."""
    n = len(lines["orders"])
    total = int(np.clip(lines["lengths"], 0, None).sum())
    out = np.zeros(max(1, 2 * total), dtype=np.uint64)
    L = lib()
    L.or_rasterize.argtypes = [C.c_size_t, C.c_void_p, _u64, C.c_size_t]
    k = L.or_rasterize(n, C.addressof(lines["_struct"]), ptr(out, _u64), 2 * total)
    return out[:k]


def sort_canonical(segs):
    s = np.ascontiguousarray(segs, dtype=np.uint64).copy()
    lib().or_sort_canonical(ptr(s, _u64), len(s))
    return s


def blend_fn(mode, dst, src):
    return lib().or_blend_fn(int(mode), dst, src)


def blend_simd(mode, dst3, src3):
    d, s, o = f32(dst3), f32(src3), np.zeros(3, np.float32)
    lib().or_blend_simd(int(mode), ptr(d, _f), ptr(s, _f), ptr(o, _f))
    return o


def blend(mode, dst4, src4):
    d, s, o = f32(dst4), f32(src4), np.zeros(4, np.float32)
    lib().or_blend(int(mode), ptr(d, _f), ptr(s, _f), ptr(o, _f))
    return o


def linear_to_srgb(v):
    return lib().or_linear_to_srgb(v)


def linear_to_srgb_simd(v):
    return lib().or_linear_to_srgb_simd(v)


def to_bytes(color):
    c, o = f32(color), np.zeros(4, np.uint8)
    lib().or_to_bytes(ptr(c, _f), ptr(o, _u8))
    return o
