"""C ABI checks that need no GPU: every declared symbol is exported, host-side
path flattening equals the oracle's, mold's bookkeeping, and a loud failure
(no CPU fallback) when no gfx950 device is present."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mold_gpu.h")).read()
    names = re.findall(r"MG_EXPORT\s+[\w\s\*]+?\b(\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    import cassia_amd
    names = declared_symbols()
    assert len(names) >= 39
    lib = C.CDLL(cassia_amd.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def random_path(m, rng):
    P = m.Point
    p = m.Path()
    pts = [P(*rng.uniform(-50, 150, 2)) for _ in range(8)]
    p.line(pts[0], pts[1])
    p.quad(pts[1], pts[2], pts[3])
    p.cubic(pts[3], pts[4], pts[5], pts[6])
    p.rat_quad((pts[6], 1.0), (pts[7], float(rng.uniform(0.3, 3))), (pts[0], 1.0))
    p.rat_cubic((pts[0], 1.0), (pts[2], 0.7), (pts[4], 1.4), (pts[6], 1.0))
    p.close()
    return p


@pytest.mark.parametrize("seed", range(10))
def test_path_flatten_matches_oracle(seed):
    import cassia_amd
    from oracle import mold
    rng = np.random.default_rng(seed)
    a = random_path(cassia_amd, np.random.default_rng(seed))
    b = random_path(mold, np.random.default_rng(seed))
    t = None
    if seed % 3 == 1:
        t = [float(v) for v in rng.uniform(-2, 2, 9)]
    elif seed % 3 == 2:
        t = [1.5, 0.2, 3.0, -0.1, 0.8, -4.0, 0.001, 0.002, 1.0]
    sa, sb = a.segments(t), b.segments(t)
    assert sa.shape == sb.shape
    assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))


def test_bookkeeping():
    import cassia_amd as m
    c = m.Composition()
    ids = [c.create_layer() for _ in range(5)]
    assert ids == [0, 1, 2, 3, 4]
    p = m.Path().line(m.Point(0, 0), m.Point(0, 1)).line(m.Point(0, 1), m.Point(1, 1))
    c.insert_in_layer(ids[0], p)
    c.insert_in_layer(ids[1], p)
    assert c.builder_len() == 4 and c.actual_len() == 4
    c.get_mut(ids[0]).disable()
    assert c.actual_len() == 2
    c.remove_disabled()  # 4 >= 2 * 2: garbage collected
    assert c.builder_len() == 2 and c.get(ids[0]) is None
    assert c.create_layer() == 0  # released id is reused
    assert c.get(7) is None
    with pytest.raises(m.ScaleError):
        c.get_mut(ids[1]).set_transform([2.0, 0.0, 0.0, 1.0, 0.0, 0.0])
    c.get_mut(ids[1]).set_transform([0.6, 0.8, -0.8, 0.6, 3.0, 4.0])
    assert c.get_mut(ids[1]).transform() == tuple(np.float32([0.6, 0.8, -0.8, 0.6, 3.0, 4.0]).tolist())
    c.get_mut(ids[1]).set_order(99)
    assert c.get_mut(ids[1]).order() == 99
    c.insert_in_layer(ids[1], p)  # insert resets the order to the id (composition.rs:74)
    assert c.get_mut(ids[1]).order() == ids[1]
    g = m.GradientBuilder([0, 0], [1, 1]).color([1, 0, 0, 1]).color([0, 0, 1, 1]).build()
    c.get_mut(ids[1]).set_props(m.Props(func=m.Func.Draw(m.Style(fill=m.Fill.Gradient(g)))))
    caches = [c.create_buffer_layer_cache() for _ in range(33)]
    assert caches[-1] is None and [x.id for x in caches[:3]] == [0, 1, 2]


def test_ids_exhausted():
    import cassia_amd as m
    c = m.Composition()
    for _ in range(65536):
        assert c.create_layer() is not None
    assert c.create_layer() is None


def test_render_without_device_fails_loudly():
    import cassia_amd as m
    if m.device_count() > 0:
        pytest.skip("a GPU is present")
    c = m.Composition()
    lid = c.create_layer()
    c.insert_lines(lid, [[0, 0, 0, 1]])
    buf = np.zeros((4, 4), np.uint8)
    with pytest.raises(m.MoldError) as e:
        c.render(m.Buffer(buf, 2), [1, 1, 1, 1])
    assert e.value.status == 6
