"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bit-exact bar: PixelSegment stream (canonical order) and RGBA8 bytes."""
import os

import numpy as np
import pytest

import mold_scenes
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


@pytest.mark.parametrize("scene", mold_scenes.ALL, ids=lambda f: f.__name__)
def test_mold_kats_on_gpu(gpu, scene):
    scene(gpu)


def render_both(gpu, cpu, seed, w, h, n_layers, bg=(1.0, 1.0, 1.0, 1.0), stride=None, crop=None, **kw):
    outs = []
    for m in (cpu, gpu):
        comp = scenegen.build_scene(m, seed, w, h, n_layers, **kw)
        st = stride or w
        rows = h
        buf = np.full((st * rows, 4), 7, dtype=np.uint8)
        c = m.Rect(range(crop[0], crop[1]), range(crop[2], crop[3])) if crop else None
        comp.render(m.Buffer(buf, w, width_stride=stride), list(bg), c)
        outs.append(buf)
    return outs


def assert_same_image(a, b, w):
    if not np.array_equal(a, b):
        d = np.nonzero((a != b).any(axis=1))[0]
        y, x = d // w, d % w
        raise AssertionError(f"{len(d)} pixels differ; first at (x={x[0]}, y={y[0]}): "
                             f"oracle {a[d[0]].tolist()} gpu {b[d[0]].tolist()}")


@pytest.mark.parametrize("seed", range(6))
def test_random_scene_rgba8(gpu, cpu, seed):
    w, h = 83, 61
    a, b = render_both(gpu, cpu, seed, w, h, 24)
    assert_same_image(a, b, w)


@pytest.mark.parametrize("seed", range(4))
def test_random_scene_transforms_curves(gpu, cpu, seed):
    w, h = 120, 96
    a, b = render_both(gpu, cpu, 100 + seed, w, h, 40, transforms=True, curves=True)
    assert_same_image(a, b, w)


@pytest.mark.parametrize("seed", range(4))
def test_random_scene_clips_gradients(gpu, cpu, seed):
    w, h = 64, 72
    a, b = render_both(gpu, cpu, 200 + seed, w, h, 30, clips=True, gradients=True)
    assert_same_image(a, b, w)


def test_stride_crop_and_transparent_background(gpu, cpu):
    w, h = 77, 50
    a, b = render_both(gpu, cpu, 300, w, h, 20, bg=(0.2, 0.4, 0.6, 0.5), stride=90, crop=(9, 60, 5, 41))
    assert_same_image(a, b, 90)


def test_many_layers_deep_tile(gpu, cpu):
    # 300 overlapping layers over a small canvas: long id lists per tile
    w, h = 48, 40
    a, b = render_both(gpu, cpu, 400, w, h, 300)
    assert_same_image(a, b, w)


def deep_scene(m, n_layers, w, h, seed):
    """n_layers quads that each cover most of a small canvas: every tile holds
    about n_layers layer runs (more than the row walk keeps in LDS)."""
    rng = np.random.default_rng(seed)
    comp = m.Composition()
    for i in range(n_layers):
        lid = comp.create_layer()
        x0, y0 = rng.uniform(-6, 3, 2)
        x1, y1 = rng.uniform(w - 3, w + 6), rng.uniform(h - 3, h + 6)
        q = np.array([[x0, y0], [x1, y0 + rng.uniform(-2, 2)], [x1, y1], [x0 + rng.uniform(-2, 2), y1]], np.float32)
        comp.insert_lines(lid, np.concatenate([q, np.roll(q, -1, 0)], 1))
        style = m.Style(fill=m.Fill.Solid([float(np.float32(v)) for v in rng.uniform(0, 1, 3)] + [0.02]),
                        blend_mode=m.BlendMode(int(rng.integers(0, 12))))
        comp.get_mut(lid).set_props(m.Props(fill_rule=m.FillRule(int(rng.integers(0, 2))), func=m.Func.Draw(style)))
    return comp


@pytest.mark.parametrize("n_layers", [3000, 9000])
def test_tiles_deeper_than_lds(gpu, cpu, n_layers):
    # 3000 runs per tile exceed the LDS walk's run capacity (2048) and 9000
    # its queue capacity: the walk reruns with its arrays in global memory
    w, h = 24, 20
    outs = []
    for m in (cpu, gpu):
        buf = np.full((w * h, 4), 7, np.uint8)
        comp = deep_scene(m, n_layers, w, h, 77)
        comp.render(m.Buffer(buf, w), [1.0, 1.0, 1.0, 1.0], None)
        outs.append(buf)
    assert_same_image(outs[0], outs[1], w)
    # 9000 records per tile: the four-wave deep painter (8,192 or more)
    deep = comp.stats().deep_tiles
    assert deep > 0 if n_layers >= 8192 else deep == 0


def render_deep(m, n_layers, order_step, seed, frames=1, w=24, h=20):
    """deep_scene with layer i at order i * order_step (orders beyond 2,048
    select the large-frame row walk; layers without lines do not exist in
    mold, so only orders of drawn layers count); the buffers of `frames`
    renders and the composition."""
    comp = deep_scene(m, n_layers, w, h, seed)
    for i in range(n_layers):
        comp.get_mut(i).set_order(i * order_step)
    bufs = []
    for _ in range(frames):
        buf = np.full((w * h, 4), 7, np.uint8)
        comp.render(m.Buffer(buf, w), [1.0, 1.0, 1.0, 1.0], None)
        bufs.append(buf)
    return bufs, comp


def test_tiles_deeper_than_half_lds(gpu, cpu):
    # 1,500 layers covering the canvas at orders up to 2,998: the
    # large-frame row walk, which first takes half the LDS (1,024 runs per
    # tile); 1,500 runs per tile make the first frame walk again with the
    # whole LDS, and the composition's next frame starts there
    w, h = 24, 20
    (ref,), _ = render_deep(cpu, 1500, 2, 78)
    (f0, f1), g = render_deep(gpu, 1500, 2, 78, frames=2)
    assert_same_image(ref, f0, w)
    assert_same_image(ref, f1, w)
    st = g.stats()
    assert st.walk_lds == gpu.WALK_WHOLE and st.walk_rewalks == 0, (st.walk_lds, st.walk_rewalks)
    # a new composition of the same shape starts with half the LDS again
    (f2,), g2 = render_deep(gpu, 1500, 2, 78)
    assert_same_image(ref, f2, w)
    st2 = g2.stats()
    assert st2.walk_lds == gpu.WALK_WHOLE and st2.walk_rewalks == 1, (st2.walk_lds, st2.walk_rewalks)


def test_large_walk_in_half_lds(gpu, cpu):
    # 600 covering layers at orders up to 2,396: the large-frame walk, whose
    # tiles fit half the LDS, rendered after a frame that needed the whole LDS
    w, h = 24, 20
    render_deep(gpu, 1500, 2, 78)
    (ref,), _ = render_deep(cpu, 600, 4, 79)
    (out, out2), g = render_deep(gpu, 600, 4, 79, frames=2)
    assert_same_image(ref, out, w)
    assert_same_image(ref, out2, w)
    st = g.stats()
    assert st.walk_lds == gpu.WALK_HALF and st.walk_rewalks == 0, (st.walk_lds, st.walk_rewalks)


def test_pixel_segments_canonical_stream(gpu, cpu):
    for seed in range(3):
        s_cpu = scenegen.build_scene(cpu, 500 + seed, 90, 70, 25, transforms=True, curves=True).pixel_segments()
        s_gpu = scenegen.build_scene(gpu, 500 + seed, 90, 70, 25, transforms=True, curves=True).pixel_segments()
        assert len(s_cpu) == len(s_gpu)
        assert np.array_equal(s_cpu, s_gpu)


@pytest.mark.parametrize("radix", [4, 6, 8, 9])
def test_tuning_is_result_invariant(gpu, cpu, radix):
    w, h = 150, 70
    comp = scenegen.build_scene(cpu, 600, w, h, 60, transforms=True, clips=True)
    ref = np.zeros((w * h, 4), np.uint8)
    comp.render(cpu.Buffer(ref, w), [1.0, 1.0, 1.0, 1.0], None)
    g = scenegen.build_scene(gpu, 600, w, h, 60, transforms=True, clips=True)
    g.set_tuning(radix)
    out = np.zeros((w * h, 4), np.uint8)
    g.render(gpu.Buffer(out, w), [1.0, 1.0, 1.0, 1.0], None)
    assert_same_image(ref, out, w)
    s_cpu = scenegen.build_scene(cpu, 601, 90, 50, 20).pixel_segments()
    g2 = scenegen.build_scene(gpu, 601, 90, 50, 20)
    g2.set_tuning(radix)
    assert np.array_equal(s_cpu, g2.pixel_segments())


@pytest.mark.parametrize("walk", [64, 256, 512, 1024])
@pytest.mark.parametrize("n_layers", [60, 700])
def test_walk_threads_are_result_invariant(gpu, cpu, walk, n_layers):
    # 60 layers: the small-LDS walk; 700: more than 512 orders (the automatic
    # choice would be 1024 threads with the whole LDS)
    w, h = 150, 70
    comp = scenegen.build_scene(cpu, 610 + n_layers, w, h, n_layers, clips=True)
    ref = np.zeros((w * h, 4), np.uint8)
    comp.render(cpu.Buffer(ref, w), [1.0, 1.0, 1.0, 1.0], None)
    g = scenegen.build_scene(gpu, 610 + n_layers, w, h, n_layers, clips=True)
    g.set_walk_threads(walk)
    out = np.zeros((w * h, 4), np.uint8)
    g.render(gpu.Buffer(out, w), [1.0, 1.0, 1.0, 1.0], None)
    assert g.stats().walk_threads == walk
    assert_same_image(ref, out, w)


def test_full_size_circles_frame(gpu, cpu):
    """The CPU-runnable configuration at full size: 1024x1024, 10,000 circles."""
    from cassia_amd import workloads
    scene = workloads.circles()
    bufs = []
    for m in (cpu, gpu):
        comp = scene.build(m)
        b = np.zeros((1024 * 1024, 4), np.uint8)
        comp.render(m.Buffer(b, 1024), scene.background, None)
        bufs.append(b)
    assert_same_image(bufs[0], bufs[1], 1024)
    assert len(np.unique(bufs[1].view(np.uint32))) > 5000


def test_empty_composition(gpu, cpu):
    for m in (cpu, gpu):
        buf = np.zeros((16 * 9, 4), np.uint8)
        m.Composition().render(m.Buffer(buf, 16), [0.0, 1.0, 0.0, 1.0], None)
        assert (buf == [0, 255, 0, 255]).all()


def test_circles_fixture_via_cassia(gpu, orc):
    """The reference-held circles.segs stream (with the +256 bias added,
    cassia/tools/cassia_sim.py:37-43), painted by cassia_render vs the oracle."""
    import ctypes as C
    from oracle import mold as omold
    from test_oracle_fixtures import biased_sorted_circles, load_circles
    raw, sty = load_circles()
    d = orc.decode(raw)
    ti = np.maximum(d["tile_i"], -1) + 256
    biased = (raw & ~np.uint64(0x1FFF << 38)) | ((ti.astype(np.int64) & 0x1FFF).astype(np.uint64) << np.uint64(38))
    biased[d["is_none"] == 1] = np.uint64(orc.NONE)
    W = H = 1000
    gpu.cassia_init(W, H)
    gpu.cassia_render(biased, [(sty[i, :4], 0, 0) for i in range(len(sty))])
    img = gpu.cassia_read(W, H)
    gpu.cassia_shutdown()
    segs, _ = biased_sorted_circles(orc)
    keep = []
    pt = (orc.Props * 10)()
    for i in range(10):
        pt[i] = omold.c_props(omold.Props(func=omold.Func.Draw(omold.Style(fill=omold.Fill.Solid(sty[i, :4])))), keep=keep)
    ref = np.zeros((W * H, 4), np.uint8)
    clr = np.ones(4, np.float32)
    assert orc.lib().or_print(orc.ptr(ref.reshape(-1), orc._u8), W * H, W, W, None, 0, orc.ptr(segs, orc._u64),
                              len(segs), orc.ptr(clr, orc._f), None, pt, 10, None) == 0
    assert_same_image(ref, img, W)


def sorted_coverage(a, k, gpu_keys):
    """(key, A row) pairs in key order; GPU keys drop the scene index."""
    keys = [tuple(r[1:]) if gpu_keys else tuple(r) for r in k.tolist()]
    return sorted(zip(keys, a.tolist()))


@pytest.mark.parametrize("case", ["plain-0", "plain-1", "curves", "deep", "crop"])
def test_per_pixel_coverage_bit_exact(gpu, cpu, case):
    """A = 16 * P + area of every pixel of every (tile, workbench id), against
    the oracle's coverage dump."""
    w, h, n, kw, crop = 83, 61, 30, dict(transforms=True), None
    seed = {"plain-0": 700, "plain-1": 701, "curves": 702, "deep": 703, "crop": 704}[case]
    if case == "curves":
        kw["curves"] = True
    if case == "deep":
        w, h, n = 48, 40, 300
    outs = []
    for m, is_gpu in ((cpu, False), (gpu, True)):
        comp = scenegen.build_scene(m, seed, w, h, n, **kw)
        c = m.Rect(range(9, 70), range(5, 41)) if case == "crop" else None
        a, k = comp.render_coverage(w, h, [1.0, 1.0, 1.0, 1.0], c)
        outs.append(sorted_coverage(a, k, is_gpu))
    ref, got = outs
    assert len(ref) == len(got) > 0
    for (kr, ar), (kg, ag) in zip(ref, got):
        assert kr == kg, (kr, kg)
        assert ar == ag, f"tile/layer {kr}: A differs"


def fan(rng, cx, cy, k, r):
    """k thin spikes of radius r around (cx, cy): 2k edges through its pixel."""
    ang = np.sort(rng.uniform(0, 2 * np.pi, k))
    L = []
    for a in ang:
        d = rng.uniform(0.05, 0.3)
        p1 = (cx + r * np.cos(a), cy + r * np.sin(a))
        p2 = (cx + r * np.cos(a + d), cy + r * np.sin(a + d))
        L += [[cx, cy, *p1], [*p1, *p2], [*p2, cx, cy]]
    return np.array(L, np.float32)


@pytest.mark.parametrize("seed", [810, 811])
def test_dense_pixels_and_long_runs(gpu, cpu, seed):
    """Per-pixel A and RGBA8 against the oracle where one layer puts from a
    few to a few hundred segments into one tile and up to ~2k edges through
    one pixel (runs on both sides of the painter's 256-word limit for runs
    accumulated in groups, cells with many area carries), under and over
    many small polygons (tiles with more than eight runs, carried covers)."""
    w, h = 64, 48
    rng = np.random.default_rng(seed)
    paths = []
    for i, k in enumerate((3, 12, 40, 70, 100, 160, 400, 1000)):
        cx = 4.5 + 8 * (i % 7) + rng.uniform(-2, 2)
        cy = 4.5 + 8 * (i // 7) * 3 + rng.uniform(-2, 2)
        paths.append(fan(rng, cx, cy, k, float(rng.uniform(1.5, 6.0))))
    for _ in range(90):
        paths.append(scenegen.random_polygon(rng, w, h, int(rng.integers(0, 2))))
    order = rng.permutation(len(paths))
    cols = rng.uniform(0.05, 1.0, (len(paths), 4)).astype(np.float32)
    rules = rng.integers(0, 2, len(paths))
    out = {}
    for m in (cpu, gpu):
        comp = m.Composition()
        for i in order:
            lid = comp.create_layer()
            comp.insert_lines(lid, paths[i])
            fr = m.FillRule.EvenOdd if rules[i] else m.FillRule.NonZero
            comp.get_mut(lid).set_props(m.Props(fill_rule=fr,
                                                func=m.Func.Draw(m.Style(fill=m.Fill.Solid(cols[i].tolist())))))
        a, k = comp.render_coverage(w, h, [1.0, 1.0, 1.0, 1.0], None)
        buf = np.zeros((w * h, 4), np.uint8)
        comp.render(m.Buffer(buf, w), [1.0, 1.0, 1.0, 1.0], None)
        out[m] = (sorted_coverage(a, k, m is gpu), buf)
    ref, got = out[cpu][0], out[gpu][0]
    assert len(ref) == len(got) > 0
    for (kr, ar), (kg, ag) in zip(ref, got):
        assert kr == kg, (kr, kg)
        assert ar == ag, f"tile/layer {kr}: A differs"
    assert_same_image(out[cpu][1], out[gpu][1], w)


def test_solid_blend_special_colours(gpu, cpu):
    """RGBA8 against the oracle for solid layers in the blend modes whose
    source-only terms the painter specializes (Darken, Lighten, Soft Light
    with terms carried by the props table, Overlay), with colour channels
    that are NaN, infinite, negative or above one, beside ordinary ones."""
    w, h = 64, 40
    rng = np.random.default_rng(77)
    special = [float("nan"), float("inf"), -float("inf"), -0.5, 1.75, 0.0, 1.0, 0.25, 0.5, -0.0]
    modes = ["Darken", "Lighten", "SoftLight", "Overlay", "ColorDodge", "Multiply"]
    layers = []
    for i in range(36):
        col = rng.uniform(0.0, 1.0, 4)
        for c in range(3):
            if rng.uniform() < 0.3:
                col[c] = special[int(rng.integers(len(special)))]
        col[3] = rng.uniform(0.3, 1.0)
        layers.append((scenegen.random_polygon(rng, w, h, int(rng.integers(0, 2))), col.astype(np.float32).tolist(),
                       modes[i % len(modes)]))
    out = {}
    for m in (cpu, gpu):
        comp = m.Composition()
        for lines, col, mode in layers:
            lid = comp.create_layer()
            comp.insert_lines(lid, lines)
            comp.get_mut(lid).set_props(m.Props(func=m.Func.Draw(m.Style(fill=m.Fill.Solid(col),
                                                                          blend_mode=getattr(m.BlendMode, mode)))))
        buf = np.zeros((w * h, 4), np.uint8)
        comp.render(m.Buffer(buf, w), [0.6, 0.4, 0.2, 1.0], None)
        out[m] = buf
    assert_same_image(out[cpu], out[gpu], w)


def test_line_changes_between_frames(gpu, cpu):
    """A composition changed between frames, against the oracle after each
    change: lines appended to an existing layer and to a new layer with a
    higher id, layer orders permuted, layers disabled and compacted away
    (remove_disabled), and curves inserted (flattened on the device on the
    GPU side)."""
    w, h = 96, 64
    comps = {m: scenegen.build_scene(m, 930, w, h, 20, transforms=True) for m in (cpu, gpu)}
    rng = np.random.default_rng(931)
    quad = np.array([[5, 7], [60, 11], [70, 50], [9, 44]], np.float32)
    quad_lines = np.concatenate([quad, np.roll(quad, -1, 0)], 1)

    def frame(what):
        bufs = []
        for m, c in comps.items():
            b = np.full((w * h, 4), 7, np.uint8)
            c.render(m.Buffer(b, w), [1.0, 1.0, 1.0, 1.0], None)
            bufs.append(b)
        try:
            assert_same_image(bufs[0], bufs[1], w)
        except AssertionError as e:
            raise AssertionError(f"after {what}: {e}")

    frame("the first render")
    style = dict(color=[0.2, 0.4, 0.9, 0.7], rule=1, blend=3)
    for m, c in comps.items():
        c.insert_lines(3, quad_lines + np.float32(4.0))
        ids = [c.create_layer() for _ in range(40)]
        c.insert_lines(int(ids[-1]), quad_lines * np.float32(0.8))
        from cassia_amd.workloads import props_of
        c.get_mut(int(ids[-1])).set_props(props_of(m, style))
    frame("lines appended to an old and a new layer")
    perm = rng.permutation(21)
    layer_ids = list(range(20)) + [59]
    for m, c in comps.items():
        for lid, o in zip(layer_ids, perm):
            c.get_mut(lid).set_order(int(o) * 3 + 1)
    frame("orders permuted")
    for m, c in comps.items():
        for lid in range(0, 20, 2):
            c.get_mut(lid).disable()
        c.remove_disabled()
    frame("layers disabled and removed")
    P = {m: m.Point for m in comps}
    for m, c in comps.items():
        paths = []
        for k in range(3):
            p = m.Path()
            a, b2, cc = (P[m](10.0 + 20 * k, 5.0), P[m](80.0, 30.0 + 5 * k), P[m](15.0, 60.0))
            p.quad(a, b2, cc)
            p.line(cc, a)
            paths.append(p)
        targets = [1, 5, 59]
        if m is gpu:
            c.insert_paths_device(np.array(targets, np.uint16), paths)
        else:
            for t, p in zip(targets, paths):
                c.insert_in_layer(t, p)
        for t, o in zip(targets, (70, 71, 72)):
            c.get_mut(t).set_order(o)
    frame("paths flattened on the device")


def far_coordinate_lines():
    """Lines whose sub-pixel coordinates (16 v + 0.5) leave i32 and saturate
    to i32::MAX / MIN (to_sub_pixel's `as i32`, surpass/src/segment.rs:13-16),
    lines whose tile indices wrap the 12- and 13-bit PixelSegment fields, and
    lines with NaN (converted to 0) or infinite endpoints. Every line spans at
    most ~60 pixels after conversion."""
    nan, inf = float("nan"), float("inf")
    L = []
    for x in (1.4e8, 2.0e9, 3.0e38, -1.4e8, -2.0e9, -3.0e38):  # x saturates: y-major lines
        L += [[x, 3.0, x + 4.0, 40.0], [x, 41.5, x - 7.0, 9.25]]
    for y in (1.4e8, -2.0e9):  # y saturates at both ends: no pixels
        L += [[2.0, y, 50.0, y + 6.0]]
    for x, y in ((5.0e5, 10.0), (3.0e6, -4.0), (-3.0e6, 20.0), (1.0e7, 7.5), (12.0, 2.5e5), (30.0, -3.0e6)):
        L += [[x, y, x + 37.0, y + 11.0], [x + 3.0, y + 20.0, x - 9.0, y - 25.0]]  # field wraps
    L += [[nan, 5.0, 10.0, 30.0], [5.0, nan, 20.0, 40.0], [nan, 2.0, nan, 33.0],
          [inf, 3.0, inf, 30.0], [-inf, 8.0, -inf, 12.0]]
    return np.array(L, np.float32)


def test_far_and_non_finite_coordinates(gpu, cpu):
    """Canonical PixelSegment stream and RGBA8 against the oracle for lines
    far outside any canvas (saturated sub-pixel coordinates, wrapped tile
    fields, NaN and infinite endpoints) beside an ordinary polygon, also
    through a rotating layer transform."""
    lines = far_coordinate_lines()
    w, h = 48, 40
    out = {}
    for m in (cpu, gpu):
        comp = m.Composition()
        a, b, c = comp.create_layer(), comp.create_layer(), comp.create_layer()
        comp.insert_lines(a, lines)
        comp.insert_lines(b, scenegen.random_polygon(np.random.default_rng(5), w, h, 0))
        comp.insert_lines(c, lines[:14])
        for lid, col in ((a, [0.9, 0.2, 0.1, 0.7]), (b, [0.1, 0.5, 0.9, 0.8]), (c, [0.3, 0.9, 0.3, 0.6])):
            comp.get_mut(lid).set_props(m.Props(fill_rule=m.FillRule.EvenOdd,
                                                func=m.Func.Draw(m.Style(fill=m.Fill.Solid(col)))))
        comp.get_mut(c).set_transform([0.6, -0.8, 0.8, 0.6, 5.0, 7.0])
        segs = comp.pixel_segments()
        buf = np.zeros((w * h, 4), np.uint8)
        comp.render(m.Buffer(buf, w), [1.0, 1.0, 1.0, 1.0], None)
        out[m] = (np.asarray(segs), buf)
    assert len(out[cpu][0]) > 500
    assert np.array_equal(out[cpu][0], out[gpu][0])
    assert_same_image(out[cpu][1], out[gpu][1], w)


def _lines_with_line_pixels(rng, target, w, h):
    """Vertical lines (x0 = x1) whose line-pixel counts, per the oracle's
    LinesBuilder::build, add up to exactly `target`."""
    import oracle
    # line-pixels of a vertical line of k pixels' height starting inside a pixel
    ks = np.arange(1, h - 2)
    probe = oracle.lines_build(np.full(len(ks), 3.25), np.full(len(ks), 0.5), np.full(len(ks), 3.25),
                               0.5 + ks.astype(np.float64), np.zeros(len(ks), np.uint16))
    by_len = {}
    for k, n in zip(ks, probe["lengths"]):
        by_len.setdefault(int(n), int(k))
    lines, total = [], 0
    while total < target:
        need = target - total
        n = min(need, int(rng.integers(40, h - 8)))
        while n not in by_len:
            n -= 1
        k = by_len[n]
        x = float(rng.integers(1, w - 1)) + 0.25
        y0 = float(rng.integers(0, h - k - 1)) + 0.5
        lines.append((x, y0, x, y0 + k))
        total += n
    return np.array(lines, np.float32)


@pytest.mark.parametrize("target", [2047, 2048, 2049, 4095, 4096, 4097, 6143, 6145, 8191, 8193, 10247])
def test_sort_partition_boundaries(gpu, cpu, target):
    """Frames whose line-pixel counts sit on either side of the emission
    partition (2,048 line-pixels, 4,096 word slots) and sort partition (two
    emission partitions, 8,192 slots) edges: a sort partition with one
    emission partition, partial emission waves, partial sort partitions."""
    import oracle
    w, h = 96, 128
    rng = np.random.default_rng(target)
    lines = _lines_with_line_pixels(rng, target, w, h)
    got = oracle.lines_build(lines[:, 0], lines[:, 1], lines[:, 2], lines[:, 3], np.zeros(len(lines), np.uint16))
    assert int(got["lengths"].sum()) == target
    props = scenegen.random_props(cpu, np.random.default_rng(target), 6)
    gprops = scenegen.random_props(gpu, np.random.default_rng(target), 6)
    layer_of = rng.integers(0, 6, len(lines))
    outs = []
    for m, pr in ((cpu, props), (gpu, gprops)):
        comp = m.Composition()
        ids = [comp.create_layer() for _ in range(6)]
        for i, lid in enumerate(ids):
            sel = lines[layer_of == i]
            if len(sel):
                comp.insert_lines(lid, sel)
            comp.get_mut(lid).set_props(pr[i])
        buf = np.zeros((w * h, 4), np.uint8)
        comp.render(m.Buffer(buf, w), [1.0, 1.0, 1.0, 1.0], None)
        outs.append(buf)
        if m is gpu:
            assert comp.stats().line_pixels == target
    assert_same_image(outs[0], outs[1], w)
