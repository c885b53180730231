"""GPU parity at the headline configuration: the single large frame of
200,000 star/polygon paths on a 3840x2160 canvas (65,536 layers, 16-bit
layer keys, 9 + 9 tile bits, about 3.2e8 pixel segments), against the oracle.

- The whole RGBA8 frame, byte for byte, against the oracle's threaded render.
- Per-pixel coverage A = 16 * P + area of every (tile, workbench id) of one
  tile row, against the oracle's coverage dump of the lines that reach it
  (a tile row's records depend only on the segments of that row).
- The canonical PixelSegment stream of the same generator at 20,000 paths,
  with layer orders spread over the whole 16-bit range."""
import os

import numpy as np
import pytest

from test_gpu_parity import assert_same_image

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def gpu():
    import cassia_amd
    assert cassia_amd.device_count() > 0, "no gfx950 device: the HIP path cannot run"
    return cassia_amd


@pytest.fixture(scope="module")
def cpu():
    from oracle import mold
    return mold


@pytest.fixture(scope="module")
def polygons():
    from cassia_amd import workloads
    return workloads.polygons()


def test_polygons_full_frame_rgba8(gpu, cpu, polygons):
    sc = polygons
    W, H = sc.width, sc.height
    print(f"polygons {W}x{H}: {len(sc.lines)} lines, {sc.n_layers} layers", flush=True)
    g = sc.build(gpu)
    out = np.zeros((W * H, 4), np.uint8)
    g.render(gpu.Buffer(out, W), sc.background, None)
    st = g.stats()
    # the configuration this test exists for: 16 layer bits, 3 six-bit digit
    # passes over the 18 row and column bits, hundreds of millions of words
    assert st.sort_passes == 3 and st.radix_bits == 6, (st.sort_passes, st.radix_bits)
    assert st.segments > 3e8 and st.records > 5e7, (st.segments, st.records)
    assert st.walk_threads == 512 and st.walk_lds in (gpu.WALK_HALF, gpu.WALK_WHOLE)
    c = sc.build(cpu)
    ref = np.zeros((W * H, 4), np.uint8)
    n = c.render_mt(cpu.Buffer(ref, W), sc.background, THREADS)
    assert n == st.segments, (n, st.segments)
    assert_same_image(ref, out, W)
    # a second frame of the same composition: the same bytes
    out2 = np.zeros((W * H, 4), np.uint8)
    g.render(gpu.Buffer(out2, W), sc.background, None)
    assert_same_image(ref, out2, W)


def band_scene(sc, y0, y1, margin=2.0):
    """The scene's lines whose y range reaches pixel rows [y0, y1) (with a
    margin): every segment of those rows comes from them."""
    from cassia_amd.workloads import Scene
    lo = np.minimum(sc.lines[:, 1], sc.lines[:, 3])
    hi = np.maximum(sc.lines[:, 1], sc.lines[:, 3])
    keep = (hi >= y0 - margin) & (lo <= y1 + margin)
    props = dict(color=sc.colors, rule=sc.rules, blend=sc.blends)
    return Scene(sc.width, sc.height, sc.lines[keep], sc.layer_ids[keep], props, sc.background,
                 orders=sc.orders, name=sc.name + "-band")


def sorted_records(a, keys):
    """(keys, A) in key order (tile row, tile column, layer)."""
    order = np.lexsort((keys[:, 2], keys[:, 1], keys[:, 0]))
    return keys[order], a[order]


@pytest.mark.parametrize("row", [135, 20])
def test_polygons_coverage_tile_row(gpu, cpu, polygons, row):
    sc = polygons
    W, H = sc.width, sc.height
    y0, y1 = 8 * row, 8 * row + 8
    g = sc.build(gpu)
    a_g, k_g = g.render_coverage(W, H, sc.background, gpu.Rect(range(0, W), range(y0, y1)), cap=1 << 19)
    k_g = k_g[:, 1:]  # drop the scene index
    band = band_scene(sc, y0, y1).build(cpu)
    a_c, k_c = band.render_coverage(W, H, sc.background, cpu.Rect(range(0, W), range(y0, y1)), cap=1 << 19)
    print(f"tile row {row}: {len(k_g)} records", flush=True)
    assert len(k_g) == len(k_c) > 100_000
    kg, ag = sorted_records(a_g, k_g)
    kc, ac = sorted_records(a_c, k_c)
    assert np.array_equal(kg, kc.astype(kg.dtype))
    bad = np.nonzero((ag != ac).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} records differ; first (row, column, layer) {kg[bad[0]].tolist()}"


def test_polygons_canonical_stream_spread_orders(gpu, cpu):
    from cassia_amd import workloads
    sc = workloads.polygons(count=20_000)
    rng = np.random.default_rng(9)
    sc.orders = rng.choice(65536, sc.n_layers, replace=False).astype(np.uint16)
    assert sc.orders.max() > 65000
    s_gpu = sc.build(gpu).pixel_segments()
    s_cpu = sc.build(cpu).pixel_segments()
    print(f"canonical stream: {len(s_gpu)} words", flush=True)
    assert len(s_cpu) == len(s_gpu) > 4e7
    assert np.array_equal(s_cpu, s_gpu)
