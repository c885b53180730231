"""Seeded synthetic workloads of the benchmark configurations.

All geometry is generated in float32 with numpy (PCG64, fixed seeds) so the
same scene can be handed to any implementation of the mold API. Paths made
only of lines are inserted through the bulk line entry point; an affine
transform is applied to the vertices in float32 exactly as mold's
`insert_in_layer_transformed` does for line commands
(mold/src/path/transform.rs:17-22: x' = (t0*x + t1*y) + t2, no FMA).
"""
import numpy as np

SRGB_TO_LINEAR = None


def srgb_u8_to_linear(u8):
    """sRGB byte -> linear float32 (IEC 61966-2-1 decode, float32 ops)."""
    c = np.asarray(u8, dtype=np.float32) / np.float32(255.0)
    lo = c / np.float32(12.92)
    hi = ((c + np.float32(0.055)) / np.float32(1.055)) ** np.float32(2.4)
    return np.where(c <= np.float32(0.04045), lo, hi).astype(np.float32)


def closed_polyline(pts):
    """(k, 2) vertices -> (k, 4) closing line segments."""
    pts = np.asarray(pts, dtype=np.float32)
    return np.concatenate([pts, np.roll(pts, -1, axis=0)], axis=1)


def affine_path_transform(pts, t9):
    """Transform::transform for Line commands (transform.rs:57-93), t22 = 1."""
    t = np.asarray(t9, dtype=np.float32)
    r = np.float32(1.0) / t[8]
    t = t * r
    x, y = pts[..., 0], pts[..., 1]
    return np.stack([t[0] * x + t[1] * y + t[2], t[3] * x + t[4] * y + t[5]], axis=-1).astype(np.float32)


class Scene:
    """Lines with per-line layer ids plus per-layer props as plain data."""

    def __init__(self, width, height, lines, layer_ids, props, background, orders=None, name=""):
        self.width, self.height = width, height
        self.lines = np.ascontiguousarray(lines, dtype=np.float32)
        self.layer_ids = np.ascontiguousarray(layer_ids, dtype=np.uint16)
        self.props = props  # list of dicts per layer id
        self.background = list(background)
        self.orders = orders
        self.name = name

    @property
    def n_layers(self):
        return len(self.props)

    def build(self, m):
        """Composition of module `m` (mold-API implementation) holding the scene."""
        comp = m.Composition()
        for _ in range(self.n_layers):
            comp.create_layer()
        if hasattr(comp, "insert_lines_multi"):
            comp.insert_lines_multi(self.layer_ids, self.lines)
        else:
            order = np.argsort(self.layer_ids, kind="stable")
            ids = self.layer_ids[order]
            bounds = np.flatnonzero(np.diff(ids.astype(np.int64))) + 1
            for chunk in np.split(order, bounds):
                comp.insert_lines(int(self.layer_ids[chunk[0]]), self.lines[chunk])
        for lid, p in enumerate(self.props):
            layer = comp.get_mut(lid)
            if layer is None:
                continue
            layer.set_props(props_of(m, p))
            if self.orders is not None:
                layer.set_order(int(self.orders[lid]))
        return comp


def props_of(m, p):
    style = m.Style(fill=m.Fill.Solid(p["color"]), blend_mode=m.BlendMode(p["blend"]))
    return m.Props(fill_rule=m.FillRule(p["rule"]), func=m.Func.Draw(style))


def circles(seed=1, width=1024, height=1024, count=10_000, segments=64, rmin=2.0, rmax=64.0):
    """Plain method, CPU-runnable: `count` circles, one layer each (order =
    index), opaque sRGB colour, NonZero, Over, white background."""
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0, width, count).astype(np.float32)
    cy = rng.uniform(0, height, count).astype(np.float32)
    r = np.exp(rng.uniform(np.log(rmin), np.log(rmax), count)).astype(np.float32)
    ang = (np.arange(segments, dtype=np.float32) * np.float32(2 * np.pi / segments)).astype(np.float32)
    ca, sa = np.cos(ang).astype(np.float32), np.sin(ang).astype(np.float32)
    px = cx[:, None] + r[:, None] * ca[None, :]
    py = cy[:, None] + r[:, None] * sa[None, :]
    pts = np.stack([px, py], axis=-1).astype(np.float32)
    lines = np.concatenate([pts, np.roll(pts, -1, axis=1)], axis=-1).reshape(-1, 4)
    ids = np.repeat(np.arange(count, dtype=np.uint16), segments)
    cols = srgb_u8_to_linear(rng.integers(0, 256, (count, 3)))
    props = [dict(color=[float(c) for c in cols[i]] + [1.0], rule=0, blend=0) for i in range(count)]
    return Scene(width, height, lines, ids, props, [1.0, 1.0, 1.0, 1.0], name="circles-1024")


def polygons(seed=2, width=3840, height=2160, count=200_000, kmin=5, kmax=50, smin=4.0, smax=512.0,
             max_layers=65536, scale_range=(0.5, 2.0)):
    """Single large frame: star/polygon paths with random affine transforms,
    50/50 fill rules, uniform blend modes, alpha in [0.2, 1]. Path i goes to
    layer i mod max_layers (layer ids are u16, mold/src/layer.rs:13)."""
    rng = np.random.default_rng(seed)
    k = rng.integers(kmin, kmax + 1, count)
    side = np.exp(rng.uniform(np.log(smin), np.log(smax), count)).astype(np.float32)
    star = rng.random(count) < 0.5
    theta = rng.uniform(0, 2 * np.pi, count)
    scale = rng.uniform(*scale_range, count)
    tx = rng.uniform(0, width, count)
    ty = rng.uniform(0, height, count)
    total = int(k.sum())
    first = np.concatenate([[0], np.cumsum(k)[:-1]])
    path_of = np.repeat(np.arange(count), k)
    local = (rng.random((total, 2)).astype(np.float32) - np.float32(0.5)) * side[path_of, None]
    # simple polygons: vertices sorted by angle around the box centre
    ang = np.arctan2(local[:, 1], local[:, 0])
    key = path_of * 8.0 + np.where(star[path_of], 0.0, ang + 4.0)
    idx = np.lexsort((np.arange(total), key)) if True else None
    local = local[idx]
    c = (np.cos(theta) * scale).astype(np.float32)
    s = (np.sin(theta) * scale).astype(np.float32)
    t0, t1, t2 = c[path_of], -s[path_of], tx.astype(np.float32)[path_of]
    t3, t4, t5 = s[path_of], c[path_of], ty.astype(np.float32)[path_of]
    x, y = local[:, 0], local[:, 1]
    gx = (t0 * x + t1 * y + t2).astype(np.float32)
    gy = (t3 * x + t4 * y + t5).astype(np.float32)
    nxt = np.arange(total) + 1
    last = first + k - 1
    nxt[last] = first
    lines = np.stack([gx, gy, gx[nxt], gy[nxt]], axis=1).astype(np.float32)
    n_layers = min(count, max_layers)
    ids = (path_of % max_layers).astype(np.uint16)
    lrng = np.random.default_rng(seed + 1000)
    cols = lrng.random((n_layers, 3)).astype(np.float32)
    alpha = lrng.uniform(0.2, 1.0, n_layers).astype(np.float32)
    rules = lrng.integers(0, 2, n_layers)
    blends = lrng.integers(0, 12, n_layers)
    props = [dict(color=[float(v) for v in cols[i]] + [float(alpha[i])], rule=int(rules[i]),
                  blend=int(blends[i])) for i in range(n_layers)]
    return Scene(width, height, lines, ids, props, [1.0, 1.0, 1.0, 1.0],
                 name=f"polygons-{width}x{height}-{count}")
