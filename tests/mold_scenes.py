"""mold's composition tests (mold/src/composition.rs:331-798), written once
against a module `m` exposing mold's API, so the same scene code runs on the
CPU oracle (`oracle.mold`) and on the GPU product (`cassia_amd`).
TILE_SIZE -> 8 (surpass/src/lib.rs:27-33)."""
import numpy as np

T = 8
BLACK = [0x00, 0x00, 0x00, 0xFF]
BLACKF = [0.0, 0.0, 0.0, 1.0]
RED = [0xFF, 0x00, 0x00, 0xFF]
REDF = [1.0, 0.0, 0.0, 1.0]
GREEN = [0x00, 0xFF, 0x00, 0xFF]
GREENF = [0.0, 1.0, 0.0, 1.0]
RED_GREEN_50 = [0xBB, 0xBB, 0x00, 0xFF]


def buf(color, n):
    return np.array([color] * n, dtype=np.uint8)


def pixel_path(m, x, y):
    p = m.Path()
    P = m.Point
    p.line(P(x, y), P(x, y + 1))
    p.line(P(x, y + 1), P(x + 1, y + 1))
    p.line(P(x + 1, y + 1), P(x + 1, y))
    p.line(P(x + 1, y), P(x, y))
    return p


def solid(m, color):
    return m.Props(func=m.Func.Draw(m.Style(fill=m.Fill.Solid(color))))


def background_color_clear(m):
    b = buf(GREEN, 1)
    c = m.Composition()
    c.render(m.Buffer(b, 1), REDF, None)
    assert b.tolist() == [RED]


def one_pixel(m):
    b = buf(GREEN, 3)
    c = m.Composition()
    lid = c.create_layer()
    c.insert_in_layer(lid, pixel_path(m, 1, 0)).set_props(solid(m, REDF))
    c.render(m.Buffer(b, 3), GREENF, None)
    assert b.tolist() == [GREEN, RED, GREEN]


def two_pixels_same_layer(m):
    b = buf(GREEN, 3)
    c = m.Composition()
    lid = c.create_layer()
    c.insert_in_layer(lid, pixel_path(m, 1, 0)).set_props(solid(m, REDF))
    c.insert_in_layer(lid, pixel_path(m, 2, 0))
    c.render(m.Buffer(b, 3), GREENF, None)
    assert b.tolist() == [GREEN, RED, RED]


def one_pixel_translated(m):
    b = buf(GREEN, 3)
    c = m.Composition()
    lid = c.create_layer()
    c.insert_in_layer(lid, pixel_path(m, 1, 0)).set_props(solid(m, REDF)).set_transform(
        [1.0, 0.0, 0.0, 1.0, 0.5, 0.0])
    c.render(m.Buffer(b, 3), GREENF, None)
    assert b.tolist() == [GREEN, RED_GREEN_50, RED_GREEN_50]


def one_pixel_rotated(m):
    b = buf(GREEN, 3)
    c = m.Composition()
    angle = np.float32(-np.pi / 2.0)
    lid = c.create_layer()
    cs, sn = float(np.cos(angle, dtype=np.float32)), float(np.sin(angle, dtype=np.float32))
    c.insert_in_layer(lid, pixel_path(m, -1, 1)).set_props(solid(m, REDF)).set_transform(
        [cs, -sn, sn, cs, 0.0, 0.0])
    c.render(m.Buffer(b, 3), GREENF, None)
    assert b.tolist() == [GREEN, RED, GREEN]


def remove_and_resize(m):
    b = buf(GREEN, 4)
    c = m.Composition()
    l0, l1, l2 = c.create_layer(), c.create_layer(), c.create_layer()
    c.insert_in_layer(l0, pixel_path(m, 0, 0)).set_props(solid(m, REDF))
    c.insert_in_layer(l1, pixel_path(m, 1, 0)).set_props(solid(m, REDF))
    c.insert_in_layer(l2, pixel_path(m, 2, 0)).set_props(solid(m, REDF))
    c.insert_in_layer(l2, pixel_path(m, 3, 0)).set_props(solid(m, REDF))
    c.render(m.Buffer(b, 4), GREENF, None)
    assert b.tolist() == [RED] * 4
    assert c.builder_len() == 16 and c.actual_len() == 16
    b[:] = GREEN
    c.get_mut(l0).disable()
    # the layout cache is keyed by the buffer, so width 3 reuses width 4
    c.render(m.Buffer(b, 3), GREENF, None)
    assert b.tolist() == [GREEN, RED, RED, RED]
    assert c.builder_len() == 16 and c.actual_len() == 12
    b[:] = GREEN
    c.get_mut(l2).disable()
    c.render(m.Buffer(b, 3), GREENF, None)
    assert b.tolist() == [GREEN, RED, GREEN, GREEN]
    assert c.builder_len() == 4 and c.actual_len() == 4


def remove_twice(m):
    c = m.Composition()
    lid = c.create_layer()
    c.insert_in_layer(lid, pixel_path(m, 0, 0)).set_props(solid(m, REDF))
    assert c.actual_len() == 4
    c.get_mut(lid).disable()
    assert c.actual_len() == 0
    c.get_mut(lid).disable()
    assert c.actual_len() == 0


def render_change_layers_only(m):
    b = buf(BLACK, 3 * T * T)
    c = m.Composition()
    cache = c.create_buffer_layer_cache()
    lid = c.create_layer()
    c.insert_in_layer(lid, pixel_path(m, 0, 0)).set_props(solid(m, REDF))
    c.insert_in_layer(lid, pixel_path(m, T, 0))
    lid = c.create_layer()
    c.insert_in_layer(lid, pixel_path(m, T + 1, 0)).set_props(solid(m, GREENF))
    c.insert_in_layer(lid, pixel_path(m, 2 * T, 0))
    c.render(m.Buffer(b, 3 * T, layer_cache=cache), BLACKF, None)
    assert b[0].tolist() == RED and b[T].tolist() == RED
    assert b[T + 1].tolist() == GREEN and b[2 * T].tolist() == GREEN
    b[:] = BLACK
    c.get_mut(lid).set_props(solid(m, REDF))
    c.render(m.Buffer(b, 3 * T, layer_cache=cache), BLACKF, None)
    assert b[0].tolist() == BLACK
    assert b[T].tolist() == RED and b[T + 1].tolist() == RED and b[2 * T].tolist() == RED


def clear_emptied_tiles(m):
    b = buf(BLACK, 2 * T * T)
    c = m.Composition()
    cache = c.create_buffer_layer_cache()
    lid = c.create_layer()
    c.insert_in_layer(lid, pixel_path(m, 0, 0)).set_props(solid(m, REDF))
    c.insert_in_layer(lid, pixel_path(m, T, 0))
    c.render(m.Buffer(b, 2 * T, layer_cache=cache), BLACKF, None)
    assert b[0].tolist() == RED
    c.get_mut(lid).set_transform([1.0, 0.0, 0.0, 1.0, float(T), 0.0])
    c.render(m.Buffer(b, 2 * T, layer_cache=cache), BLACKF, None)
    assert b[0].tolist() == BLACK
    c.get_mut(lid).set_transform([1.0, 0.0, 0.0, 1.0, -float(T), 0.0])
    c.render(m.Buffer(b, T, layer_cache=cache), BLACKF, None)
    assert b[0].tolist() == RED
    c.get_mut(lid).set_transform([1.0, 0.0, 0.0, 1.0, 0.0, float(T)])
    c.render(m.Buffer(b, T, layer_cache=cache), BLACKF, None)
    assert b[0].tolist() == BLACK


def separare_layer_caches(m):
    b = buf(BLACK, T * T)
    c = m.Composition()
    c0 = c.create_buffer_layer_cache()
    c1 = c.create_buffer_layer_cache()
    lid = c.create_layer()
    c.insert_in_layer(lid, pixel_path(m, 0, 0)).set_props(solid(m, REDF))
    c.render(m.Buffer(b, T, layer_cache=c0), BLACKF, None)
    assert b[0].tolist() == RED
    b[:] = BLACK
    c.render(m.Buffer(b, T, layer_cache=c0), BLACKF, None)
    assert b[0].tolist() == BLACK
    c.render(m.Buffer(b, T, layer_cache=c1), BLACKF, None)
    assert b[0].tolist() == RED
    c.get_mut(lid).set_transform([1.0, 0.0, 0.0, 1.0, 1.0, 0.0])
    c.render(m.Buffer(b, T, layer_cache=c0), BLACKF, None)
    assert b[0].tolist() == BLACK and b[1].tolist() == RED
    b[:] = BLACK
    c.render(m.Buffer(b, T, layer_cache=c1), BLACKF, None)
    assert b[0].tolist() == BLACK and b[1].tolist() == RED


def even_odd(m):
    P = m.Point
    p = m.Path()
    p.line(P(0.0, 0.0), P(0.0, T))
    p.line(P(0.0, T), P(3.0 * T, T))
    p.line(P(3.0 * T, T), P(3.0 * T, 0.0))
    p.line(P(3.0 * T, 0.0), P(T, 0.0))
    p.line(P(T, 0.0), P(T, T))
    p.line(P(T, T), P(2.0 * T, T))
    p.line(P(2.0 * T, T), P(2.0 * T, 0.0))
    p.line(P(2.0 * T, 0.0), P(0.0, 0.0))
    b = buf(BLACK, 3 * T * T)
    c = m.Composition()
    lid = c.create_layer()
    c.insert_in_layer(lid, p).set_props(m.Props(fill_rule=m.FillRule.EvenOdd,
                                               func=m.Func.Draw(m.Style(fill=m.Fill.Solid(REDF)))))
    c.render(m.Buffer(b, 3 * T), BLACKF, None)
    assert b[0].tolist() == RED
    assert b[T].tolist() == BLACK
    assert b[2 * T].tolist() == RED


def clear_to_red(m):
    # mold/src/utils.rs:18-23
    b = buf(GREEN, 3)
    m.clear_buffer(b, RED)
    assert b.tolist() == [RED] * 3


ALL = [background_color_clear, one_pixel, two_pixels_same_layer, one_pixel_translated,
       one_pixel_rotated, remove_and_resize, remove_twice, render_change_layers_only,
       clear_emptied_tiles, separare_layer_caches, even_odd, clear_to_red]
