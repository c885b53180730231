"""Path flattening on the GPU (mg_insert_paths_device).

Bar: the lines the device writes equal, bit for bit and in order, the host
path form's (Composition::insert_in_layer[_transformed]) and the oracle's
PathSegments; renders of both compositions are identical; host inserts,
device inserts and compaction interleave correctly."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import cassia_amd
    assert cassia_amd.device_count() > 0, "no gfx950 device: the HIP path cannot run"
    return cassia_amd


def random_path(m, rng, scale=100.0):
    """Every command kind, several subpaths, closes with and without a gap."""
    P = m.Point
    p = m.Path()
    for sub in range(int(rng.integers(1, 4))):
        pts = [P(*rng.uniform(-0.5 * scale, 1.5 * scale, 2)) for _ in range(8)]
        kinds = rng.integers(0, 5, int(rng.integers(1, 6)))
        cur = pts[0]
        for k, kind in enumerate(kinds):
            a, b, c = pts[(k + 1) % 8], pts[(k + 2) % 8], pts[(k + 3) % 8]
            if kind == 0:
                p.line(cur, a)
                cur = a
            elif kind == 1:
                p.quad(cur, a, b)
                cur = b
            elif kind == 2:
                p.cubic(cur, a, b, c)
                cur = c
            elif kind == 3:
                p.rat_quad((cur, 1.0), (a, float(rng.uniform(0.3, 3))), (b, 1.0))
                cur = b
            else:
                p.rat_cubic((cur, 1.0), (a, float(rng.uniform(0.4, 2))), (b, float(rng.uniform(0.4, 2))), (c, 1.0))
                cur = c
        if rng.random() < 0.8:
            p.close()
    return p


def random_transform(rng, kind):
    if kind == 0:
        return None
    if kind == 1:  # affine
        return [float(v) for v in rng.uniform(-2, 2, 6)] + [0.0, 0.0, 1.0]
    return [1.5, 0.2, 3.0, -0.1, 0.8, -4.0, 0.001, 0.002, 1.0]  # perspective


def host_comp(m, paths, ids, transforms):
    comp = m.Composition()
    for _ in range(int(max(ids)) + 1):
        comp.create_layer()
    for p, i, t in zip(paths, ids, transforms):
        if t is None:
            comp.insert_in_layer(int(i), p)
        else:
            comp.insert_in_layer_transformed(int(i), p, t)
    return comp


@pytest.mark.parametrize("seed", range(4))
def test_device_lines_equal_host_and_oracle(gpu, seed):
    from oracle import mold
    rng = np.random.default_rng(seed)
    n = 300
    paths = [random_path(gpu, np.random.default_rng(1000 * seed + i)) for i in range(n)]
    ids = rng.integers(0, 40, n).astype(np.uint16)
    transforms = [random_transform(rng, int(k)) for k in rng.integers(0, 3, n)]
    dev = gpu.Composition()
    for _ in range(40):
        dev.create_layer()
    dev.insert_paths_device(ids, paths, transforms)
    host = host_comp(gpu, paths, ids, transforms)
    dp, di = dev.lines()
    hp, hi = host.lines()
    assert dp.shape == hp.shape and len(dp) > n
    assert np.array_equal(dp.view(np.uint32), hp.view(np.uint32)), "device lines differ from the host path form"
    assert np.array_equal(di, hi)
    # the oracle's PathSegments, path after path
    ref = []
    for i in range(n):
        op = random_path(mold, np.random.default_rng(1000 * seed + i))
        ref.append(op.segments(transforms[i]))
    ref = np.concatenate(ref).astype(np.float32)
    assert np.array_equal(dp.view(np.uint32), ref.view(np.uint32)), "device lines differ from the oracle"


def test_render_and_interleaving(gpu):
    from oracle import mold
    w, h = 160, 120
    groups = []
    for g in range(3):
        paths = [random_path(gpu, np.random.default_rng(50 + 10 * g + i), scale=w) for i in range(10)]
        groups.append(([10 * g + i for i in range(10)], paths))

    def build(m, device):
        rng = np.random.default_rng(7)
        comp = m.Composition()
        for _ in range(30):
            comp.create_layer()
        for gi, (ids, _) in enumerate(groups):
            ps = [random_path(m, np.random.default_rng(50 + 10 * gi + i), scale=w) for i in range(10)]
            if device and gi != 1:  # device, host, device
                comp.insert_paths_device(ids, ps)
            else:
                for i, p in zip(ids, ps):
                    comp.insert_in_layer(i, p)
        for lid in range(30):
            style = m.Style(fill=m.Fill.Solid([float(np.float32(v)) for v in rng.uniform(0, 1, 3)] + [0.7]),
                            blend_mode=m.BlendMode(int(lid % 12)))
            comp.get_mut(lid).set_props(m.Props(fill_rule=m.FillRule(lid % 2), func=m.Func.Draw(style)))
        return comp

    ref = build(mold, False)
    dev = build(gpu, True)
    host = build(gpu, False)
    bufs = []
    for m, comp in ((mold, ref), (gpu, dev), (gpu, host)):
        buf = np.zeros((w * h, 4), np.uint8)
        comp.render(m.Buffer(buf, w), [1.0, 1.0, 1.0, 1.0], None)
        bufs.append(buf)
    assert np.array_equal(bufs[0], bufs[1]), "device-flattened composition renders differently from the oracle"
    assert np.array_equal(bufs[1], bufs[2])
    # disable most layers: the compaction works on the device lines too
    for lid in range(25):
        for comp in (ref, dev):
            comp.get_mut(lid).disable()
    for comp in (ref, dev):
        comp.remove_disabled()
    out = []
    for m, comp in ((mold, ref), (gpu, dev)):
        buf = np.zeros((w * h, 4), np.uint8)
        comp.render(m.Buffer(buf, w), [1.0, 1.0, 1.0, 1.0], None)
        out.append(buf)
    assert np.array_equal(out[0], out[1])
    assert dev.builder_len() == len(dev.lines()[0])


def test_many_glyph_paths(gpu):
    # 20,000 glyph-like loops of quadratics: device lines equal the host's
    rng = np.random.default_rng(11)
    P = gpu.Point
    paths = []
    for i in range(20000):
        p = gpu.Path()
        pts = rng.uniform(0, 256, (int(rng.integers(2, 7)), 2, 2))
        start = P(*pts[0, 0])
        cur = start
        for k in range(len(pts)):
            nxt = P(*pts[(k + 1) % len(pts), 0]) if k + 1 < len(pts) else start
            p.quad(cur, P(*pts[k, 1]), nxt)
            cur = nxt
        p.close()
        paths.append(p)
    ids = (np.arange(20000) % 200).astype(np.uint16)
    dev = gpu.Composition()
    dev.create_layers(200)
    dev.insert_paths_device(ids, paths)
    host = gpu.Composition()
    host.create_layers(200)
    for i, p in zip(ids, paths):
        host.insert_in_layer(int(i), p)
    dp, di = dev.lines()
    hp, hi = host.lines()
    assert np.array_equal(dp.view(np.uint32), hp.view(np.uint32)) and np.array_equal(di, hi)
