"""Flusher::flush on the GPU path (buffer_layout.rs:52-54, painter/mod.rs:
483-492): once per written tile slice (a row of at most 8 pixels), after the
slice holds its bytes; the flusher may change the slice. Tiles the layer
cache leaves unchanged (TileWriteOp::None) are not written and not flushed.

Bar: the reference's KATs (buffer_layout.rs:253-313, `flusher` and
`flush_background`), and, against the oracle, the same set of flushed slices
(position, length, bytes seen) and the same final buffer, across frames with
a layer cache, a row stride and a crop."""
import numpy as np
import pytest

import scenegen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import cassia_amd
    assert cassia_amd.device_count() > 0, "no gfx950 device: the HIP path cannot run"
    return cassia_amd


@pytest.fixture(scope="module")
def cpu():
    from oracle import mold
    return mold


@pytest.mark.parametrize("size", [12, 8])
def test_flusher_kats(gpu, size):
    # `flusher` (12 x 12: four tiles, two of them partial) and
    # `flush_background` (one 8 x 8 tile, no segments): every written slice
    # goes through the flusher, which whitens it
    buf = np.zeros((size * size, 4), np.uint8)

    def white(s):
        s[:] = 255

    gpu.Composition().render(gpu.Buffer(buf, size, flusher=white), [0.0, 0.0, 0.0, 0.0], None)
    assert (buf == 255).all()


def flushed_frames(m, stride, crop):
    """Three frames of one scene into one buffer with a layer cache: the
    first, an unchanged one, and one after two layers changed. Returns per
    frame the sorted (pixel offset, length, bytes seen) of every flushed
    slice, and the final buffer. The flusher inverts the slice's bytes."""
    w, h = 83, 61
    comp = scenegen.build_scene(m, 920, w, h, 30, transforms=True)
    cache = comp.create_buffer_layer_cache()
    buf = np.full((stride * h, 4), 9, np.uint8)
    base = buf.ctypes.data
    frames = []
    for f in range(3):
        seen = []

        def flush(s):
            seen.append(((s.ctypes.data - base) // 4, len(s), s.tobytes()))
            s[:] = 255 - s

        if f == 2:
            for lid in (3, 11):
                layer = comp.get_mut(lid)
                if layer is not None:
                    layer.set_props(m.Props(func=m.Func.Draw(m.Style(fill=m.Fill.Solid([0.1, 0.7, 0.2, 0.9])))))
        c = m.Rect(range(crop[0], crop[1]), range(crop[2], crop[3])) if crop else None
        comp.render(m.Buffer(buf, w, width_stride=stride, layer_cache=cache, flusher=flush), [1.0, 1.0, 1.0, 1.0], c)
        frames.append(sorted(seen))
    return frames, buf


@pytest.mark.parametrize("stride,crop", [(83, None), (100, (9, 70, 5, 41))])
def test_flushed_slices_equal_oracle(gpu, cpu, stride, crop):
    ref, ref_buf = flushed_frames(cpu, stride, crop)
    got, got_buf = flushed_frames(gpu, stride, crop)
    for f, (a, b) in enumerate(zip(ref, got)):
        assert [x[:2] for x in a] == [x[:2] for x in b], f"frame {f}: flushed slices differ"
        assert a == b, f"frame {f}: bytes of the flushed slices differ"
    assert len(ref[0]) > 0 and len(ref[1]) == 0, "the unchanged frame writes (and flushes) no tile"
    assert 0 < len(ref[2]) < len(ref[0])
    assert np.array_equal(ref_buf, got_buf)
