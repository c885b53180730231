"""Reference-held fixtures: demo/circles.segs + demo/circles.stylings."""
import os

import numpy as np

HERE = os.path.join(os.path.dirname(__file__), "golden")


def load_circles():
    segs = np.fromfile(os.path.join(HERE, "circles.segs"), dtype="<u8")
    sty = np.fromfile(os.path.join(HERE, "circles.stylings"), dtype="<f4").reshape(-1, 8)
    return segs, sty


def test_circles_fixture_layout(orc):
    segs, sty = load_circles()
    assert len(segs) == 5304 and sty.shape == (10, 8)
    d = orc.decode(segs)
    live = d["is_none"] == 0
    # [seg, seg|NONE] pairs in generation order (rasterizer/mod.rs:195,272-279)
    assert (d["is_none"][0::2] == 0).all()
    assert int(live.sum()) == 5304 - 2001
    assert sorted(set(d["layer"][live].tolist())) == list(range(10))
    assert np.abs(d["cover"][live]).max() <= 16
    assert np.abs(d["area"][live]).max() <= 256
    # closed circles: per (layer, pixel row) the covers sum to zero
    rows = d["tile_j"] * 8 + d["tile_y"]
    key = d["layer"][live] * 100000 + rows[live]
    sums = np.bincount(np.unique(key, return_inverse=True)[1], weights=d["cover"][live])
    assert np.all(sums == 0)
    # stylings: alpha 1.0, NonZero, Over, colours are multiples of 2^-24
    assert (sty[:, 3] == 1.0).all()
    assert (sty[:, 4:].view(np.uint32) == 0).all()
    assert np.all((sty[:, :3] * 2 ** 24) == np.round(sty[:, :3] * 2 ** 24))


def biased_sorted_circles(orc):
    """The fixture predates the +256 tile_i bias: add it, then sort
    canonically (as cassia/tools/cassia_sim.py:37-43 prepares it)."""
    segs, sty = load_circles()
    d = orc.decode(segs)
    live = segs[d["is_none"] == 0]
    dl = orc.decode(live)
    ti = np.maximum(dl["tile_i"], -1) + 256
    field = (ti.astype(np.int64) & 0x1FFF).astype(np.uint64) << np.uint64(38)
    live = (live & ~np.uint64(0x1FFF << 38)) | field
    return orc.sort_canonical(live), sty


def test_circles_paint_is_sane(orc):
    import ctypes as C
    from oracle import mold
    segs, sty = biased_sorted_circles(orc)
    keep = []
    pt = (orc.Props * 10)()
    for i in range(10):
        pt[i] = mold.c_props(mold.Props(func=mold.Func.Draw(mold.Style(fill=mold.Fill.Solid(sty[i, :4])))), keep=keep)
    W = H = 1000
    b = np.zeros((W * H, 4), np.uint8)
    clr = np.ones(4, np.float32)
    rc = orc.lib().or_print(orc.ptr(b.reshape(-1), orc._u8), W * H, W, W, None, 0,
                            orc.ptr(segs, orc._u64), len(segs), orc.ptr(clr, orc._f), None, pt, 10, None)
    assert rc == 0
    img = b.reshape(H, W, 4)
    assert (img[..., 3] == 255).all()
    painted = (img[..., :3] != 255).any(axis=2)
    # ten opaque circles of radius 10..50 px: a few thousand to ~80k pixels
    assert 3000 < painted.sum() < 80000
