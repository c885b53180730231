"""Batched rendering (mg_batch_render): many compositions in one device frame.

Bar: every batched frame is byte-identical to the oracle's render of that
composition alone and to the plain per-composition GPU render, for every sort
radix and for the autotuner's choice. Pixels outside a frame's crop are not
written by a batch, so cropped frames are compared on the tiles the crop
covers (Rect::new rounds the crop out to whole 8 x 8 tiles)."""
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


def crop_window(w, h, crop):
    if crop is None:
        return slice(0, h), slice(0, w)
    x0, x1, y0, y1 = crop
    return slice(8 * (y0 // 8), min(h, 8 * ((y1 + 7) // 8))), slice(8 * (x0 // 8), min(w, 8 * ((x1 + 7) // 8)))


def render_plain(m, comp, w, h, bg, crop):
    buf = np.full((w * h, 4), 7, np.uint8)
    c = m.Rect(range(crop[0], crop[1]), range(crop[2], crop[3])) if crop else None
    comp.render(m.Buffer(buf, w), list(bg), c)
    return buf


def assert_frame(ref, got, w, h, crop, what):
    ys, xs = crop_window(w, h, crop)
    a = ref.reshape(h, w, 4)[ys, xs]
    b = got.reshape(h, w, 4)[ys, xs]
    if not np.array_equal(a, b):
        d = np.argwhere((a != b).any(axis=2))
        raise AssertionError(f"{what}: {len(d)} pixels differ, first at (y, x) = {tuple(d[0])}: "
                             f"{a[tuple(d[0])].tolist()} vs {b[tuple(d[0])].tolist()}")


# (seed, w, h, layers, scene features, crop in pixels or None)
MIXED = [
    (1, 83, 61, 20, dict(), None),
    (2, 40, 40, 12, dict(clips=True), None),
    (3, 129, 64, 30, dict(gradients=True, transforms=True), None),
    (4, 17, 9, 5, dict(), None),
    (5, 96, 120, 25, dict(curves=True), (13, 70, 9, 100)),
    (6, 64, 64, 1, dict(), None),
    (7, 200, 33, 40, dict(clips=True, gradients=True), (0, 200, 8, 24)),
    (8, 8, 8, 3, dict(), None),
]


def build_mixed(m):
    return [scenegen.build_scene(m, seed, w, h, n, **kw) for seed, w, h, n, kw, _ in MIXED]


def frames_of(spec, bg=(1.0, 1.0, 1.0, 1.0), rect=None):
    out = []
    for seed, w, h, n, kw, crop in spec:
        fr = (w, h, bg)
        if crop is not None:
            fr = fr + (rect(range(crop[0], crop[1]), range(crop[2], crop[3])),)
        out.append(fr)
    return out


def test_mixed_batch_matches_oracle(gpu, cpu):
    bg = (0.25, 0.5, 0.75, 1.0)
    comps = build_mixed(gpu)
    b = gpu.Batch()
    b.render(comps, frames_of(MIXED, bg, gpu.Rect))
    offs = b.offsets
    assert offs == sorted(offs) and all(o % 256 == 0 for o in offs)
    for i, (seed, w, h, n, kw, crop) in enumerate(MIXED):
        ref = render_plain(cpu, scenegen.build_scene(cpu, seed, w, h, n, **kw), w, h, bg, crop)
        assert_frame(ref, b.read(i), w, h, crop, f"frame {i} (seed {seed}) vs oracle")


def test_mixed_batch_matches_plain_gpu(gpu):
    bg = (1.0, 1.0, 1.0, 1.0)
    comps = build_mixed(gpu)
    b = gpu.Batch()
    b.render(comps, frames_of(MIXED, bg, gpu.Rect))
    batched = [b.read(i) for i in range(len(MIXED))]
    for i, (seed, w, h, n, kw, crop) in enumerate(MIXED):
        plain = render_plain(gpu, comps[i], w, h, bg, crop)
        assert_frame(plain, batched[i], w, h, crop, f"frame {i} (seed {seed}) vs plain")


@pytest.fixture(scope="module")
def glyphs(gpu):
    from cassia_amd import workloads
    scenes = workloads.glyph_batch(seed=11, n_scenes=40)
    return scenes, [s.build(gpu) for s in scenes]


def test_glyph_batch_matches_plain_and_oracle(gpu, cpu, glyphs):
    scenes, comps = glyphs
    b = gpu.Batch()
    b.set_tuning(6)
    b.render(comps, [(s.width, s.height, s.background) for s in scenes])
    st = b.stats()
    assert st.lines == sum(len(s.lines) for s in scenes)
    for i, s in enumerate(scenes):
        plain = render_plain(gpu, comps[i], s.width, s.height, s.background, None)
        assert_frame(plain, b.read(i), s.width, s.height, None, f"glyph scene {i} vs plain")
    for i in (0, 7, 19, 39):
        s = scenes[i]
        ref = render_plain(cpu, s.build(cpu), s.width, s.height, s.background, None)
        assert_frame(ref, b.read(i), s.width, s.height, None, f"glyph scene {i} vs oracle")


@pytest.mark.parametrize("radix", [4, 6, 8, 9])
def test_batch_radix_invariant(gpu, glyphs, radix):
    scenes, comps = glyphs
    frames = [(s.width, s.height, s.background) for s in scenes]
    b6 = gpu.Batch()
    b6.set_tuning(6)
    b6.render(comps, frames)
    b = gpu.Batch()
    b.set_tuning(radix)
    b.render(comps, frames)
    assert b.stats().radix_bits == radix
    for i in range(len(scenes)):
        assert np.array_equal(b.read(i), b6.read(i)), f"scene {i} differs at radix {radix}"


def test_autotuner_choice_is_result_invariant(gpu, glyphs):
    scenes, comps = glyphs
    frames = [(s.width, s.height, s.background) for s in scenes]
    b = gpu.Batch()
    b.render(comps, frames)  # autotune
    t = b.tuning()
    assert t.n_candidates == 4
    assert list(t.radix_bits) == [4, 6, 8, 9]
    assert all(ms > 0 for ms in t.device_ms)
    assert t.chosen_radix_bits == list(t.radix_bits)[int(np.argmin(list(t.device_ms)))]
    assert b.stats().radix_bits == t.chosen_radix_bits
    ref = gpu.Batch()
    ref.set_tuning(8)
    ref.render(comps, frames)
    for i in range(len(scenes)):
        assert np.array_equal(b.read(i), ref.read(i))


def test_batch_follows_composition_changes(gpu, cpu):
    spec = MIXED[:4]
    comps = [scenegen.build_scene(gpu, seed, w, h, n, **kw) for seed, w, h, n, kw, _ in spec]
    refs = [scenegen.build_scene(cpu, seed, w, h, n, **kw) for seed, w, h, n, kw, _ in spec]
    bg = (1.0, 1.0, 1.0, 1.0)
    b = gpu.Batch()
    b.render(comps, frames_of(spec, bg))
    # change props, order, enabled state and lines of some compositions
    for m, cs in ((gpu, comps), (cpu, refs)):
        cs[0].get_mut(0).set_props(m.Props(func=m.Func.Draw(m.Style(fill=m.Fill.Solid([0.1, 0.9, 0.2, 0.6])))))
        cs[1].get_mut(2).set_order(200)
        cs[2].get_mut(1).disable()
        lid = cs[3].create_layer()
        cs[3].insert_lines(lid, np.array([[1, 1, 15, 2], [15, 2, 3, 8], [3, 8, 1, 1]], np.float32))
    b.render(comps, frames_of(spec, bg))
    for i, (seed, w, h, n, kw, crop) in enumerate(spec):
        ref = render_plain(cpu, refs[i], w, h, bg, crop)
        assert_frame(ref, b.read(i), w, h, crop, f"changed frame {i}")
    # a different frame set over the same compositions: sizes and order swapped
    spec2 = [spec[2], spec[0]]
    b.render([comps[2], comps[0]], frames_of(spec2, bg))
    for k, i in enumerate((2, 0)):
        seed, w, h, n, kw, crop = spec[i]
        ref = render_plain(cpu, refs[i], w, h, bg, crop)
        assert_frame(ref, b.read(k), w, h, crop, f"reordered frame {k}")


def test_batch_argument_errors(gpu):
    b = gpu.Batch()
    comp = scenegen.build_scene(gpu, 1, 16, 16, 2)
    with pytest.raises(gpu.MoldError):
        b.render([], [])
    with pytest.raises(gpu.MoldError):
        b.render([comp], [(0, 16, (1, 1, 1, 1))])
    with pytest.raises(gpu.MoldError):
        b.set_tuning(5)
    with pytest.raises(IndexError):
        b.read(0)
