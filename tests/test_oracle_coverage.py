"""The oracle's per-pixel coverage dump (A = 16 * P + area per pixel,
painter/mod.rs:229-243, 347-362) pinned by closed forms: A / 256 is the
signed covered area of the pixel for grid-aligned shapes."""
import numpy as np

from oracle import mold


def rect(x0, y0, x1, y1):
    pts = np.array([[x0, y0], [x1, y0], [x1, y1], [x0, y1]], np.float32)
    return np.concatenate([pts, np.roll(pts, -1, 0)], 1)


def coverage_image(comp, w, h):
    """Sum of A over layers per pixel (a single layer per pixel here)."""
    a, k = comp.render_coverage(w, h, [1, 1, 1, 1])
    img = np.zeros((h, w), np.int64)
    for rec, (tj, ti, _) in zip(a, k):
        for x in range(8):
            for y in range(8):
                px, py = 8 * int(ti) + x, 8 * int(tj) + y
                if px < w and py < h:
                    img[py, px] += rec[x * 8 + y]
    return img


def test_grid_rectangle_full_pixels():
    # a rectangle crossing tile borders: full pixels inside, carries across tiles
    w, h = 40, 24
    c = mold.Composition()
    lid = c.create_layer()
    c.insert_lines(lid, rect(3, 2, 29, 19))
    img = coverage_image(c, w, h)
    inside = np.zeros((h, w), bool)
    inside[2:19, 3:29] = True
    assert (np.abs(img[inside]) == 256).all()
    assert len(np.unique(img[inside])) == 1
    assert (img[~inside] == 0).all()


def test_fractional_columns_and_rows():
    # x from 2.5 to 13.25, y from 1 to 10.75: half / quarter / three-quarter pixels
    w, h = 24, 16
    c = mold.Composition()
    lid = c.create_layer()
    c.insert_lines(lid, rect(2.5, 1.0, 13.25, 10.75))
    img = np.abs(coverage_image(c, w, h))
    assert (img[1:10, 3:13] == 256).all()
    assert (img[1:10, 2] == 128).all()      # half column
    assert (img[1:10, 13] == 64).all()      # quarter column
    assert (img[10, 3:13] == 192).all()     # three-quarter row
    assert img[10, 2] == 96 and img[10, 13] == 48
    assert img[0].sum() == 0 and img[11:].sum() == 0


def test_nested_even_odd_counts_windings():
    # two nested same-orientation squares in one layer: winding 2 inside the
    # inner one, so A = 2 * 256 there (EvenOdd resolves it to empty, NonZero full)
    w, h = 32, 32
    c = mold.Composition()
    lid = c.create_layer()
    c.insert_lines(lid, np.concatenate([rect(2, 2, 30, 30), rect(10, 10, 20, 20)]))
    img = np.abs(coverage_image(c, w, h))
    assert (img[10:20, 10:20] == 512).all()
    assert (img[2:10, 2:30] == 256).all()
