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
    """Lines with per-line layer ids plus per-layer solid props as plain data.

    `props` is a list of dicts (color, rule, blend) per layer id, or one dict
    of arrays with those keys ((n, 4) colors, (n,) fill rules, (n,) blend modes).
    """

    def __init__(self, width, height, lines, layer_ids, props, background, orders=None, name=""):
        self.width, self.height = width, height
        self.lines = np.ascontiguousarray(lines, dtype=np.float32)
        self.layer_ids = np.ascontiguousarray(layer_ids, dtype=np.uint16)
        if isinstance(props, dict):
            self.colors = np.asarray(props["color"], np.float32).reshape(-1, 4)
            self.rules = np.asarray(props["rule"], np.uint32).reshape(-1)
            self.blends = np.asarray(props["blend"], np.uint32).reshape(-1)
        else:
            self.colors = np.array([p["color"] for p in props], np.float32).reshape(-1, 4)
            self.rules = np.array([p["rule"] for p in props], np.uint32)
            self.blends = np.array([p["blend"] for p in props], np.uint32)
        self.background = list(background)
        self.orders = orders
        self.name = name

    @property
    def n_layers(self):
        return len(self.colors)

    @property
    def props(self):
        return [dict(color=[float(v) for v in self.colors[i]], rule=int(self.rules[i]), blend=int(self.blends[i]))
                for i in range(self.n_layers)]

    def build(self, m):
        """Composition of module `m` (mold-API implementation) holding the scene."""
        comp = m.Composition()
        n = self.n_layers
        if hasattr(comp, "create_layers"):
            comp.create_layers(n)
        else:
            for _ in range(n):
                comp.create_layer()
        if hasattr(comp, "insert_lines_multi"):
            comp.insert_lines_multi(self.layer_ids, self.lines)
        else:
            order = np.argsort(self.layer_ids, kind="stable")
            ids = self.layer_ids[order]
            bounds = np.flatnonzero(np.diff(ids.astype(np.int64))) + 1
            for chunk in np.split(order, bounds):
                comp.insert_lines(int(self.layer_ids[chunk[0]]), self.lines[chunk])
        present = np.zeros(n, bool)
        present[np.unique(self.layer_ids)] = True
        if hasattr(comp, "set_solid_props"):
            # layers without lines do not exist in mold (get_mut gives None)
            ids = np.flatnonzero(present).astype(np.uint16)
            comp.set_solid_props(ids, self.colors[ids], self.rules[ids], self.blends[ids])
        else:
            for lid, p in enumerate(self.props):
                if present[lid]:
                    comp.get_mut(lid).set_props(props_of(m, p))
        if self.orders is not None and hasattr(comp, "set_orders"):
            ids = np.flatnonzero(present).astype(np.uint16)
            comp.set_orders(ids, np.asarray(self.orders, np.uint16)[ids])
        elif self.orders is not None:
            for lid in range(n):
                layer = comp.get_mut(lid)
                if layer is not None:
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


def quad_flatten_points(p0, c, p1, n):
    """Points B(t_j), j = 0..n, of quadratic Béziers (p0, c, p1) at t_j = j / n.

    float32, left to right, no FMA: mt = 1 - t; x = ((mt*mt)*x0 + (2*mt)*t*cx) + (t*t)*x1.
    `p0`, `c`, `p1` are (k, 2) arrays, `n` a (k,) array of line counts; returns the
    (sum(n + 1), 2) points curve after curve. B(0) = p0 and B(1) = p1 exactly.
    """
    n = np.asarray(n, dtype=np.int64)
    curve = np.repeat(np.arange(len(n)), n + 1)
    start = np.concatenate([[0], np.cumsum(n + 1)[:-1]])
    j = np.arange(curve.size) - start[curve]
    t = (j.astype(np.float32) / n[curve].astype(np.float32)).astype(np.float32)
    mt = (np.float32(1.0) - t).astype(np.float32)
    a = (mt * mt).astype(np.float32)
    b = ((np.float32(2.0) * mt).astype(np.float32) * t).astype(np.float32)
    cc = (t * t).astype(np.float32)
    out = np.empty((curve.size, 2), np.float32)
    for k in range(2):
        x0, cx, x1 = p0[curve, k], c[curve, k], p1[curve, k]
        out[:, k] = ((a * x0).astype(np.float32) + (b * cx).astype(np.float32)).astype(np.float32) + \
            (cc * x1).astype(np.float32)
    return out


def glyph_batch(seed=3, n_scenes=4096, gamma=3.0, min_side=32, max_side=512, paths=(1, 200), curves=(2, 6),
                lines_per_curve=(8, 32)):
    """Many independent small scenes: canvas side round(32 * 16^(u^gamma)),
    u ~ U[0, 1) (32..512 px, square), each holding U{1..200} glyph-like paths.
    A path is a closed loop of U{2..6} quadratic Béziers whose anchor and
    control points are uniform over the canvas, each flattened to U{8..32}
    lines (quad_flatten_points). One layer per path (order = index), EvenOdd,
    opaque solid sRGB colour, Over, white background. Returns a list of Scene."""
    rng = np.random.default_rng(seed)
    u = rng.random(n_scenes)
    side = np.clip(np.round(min_side * (max_side / min_side) ** (u ** gamma)), min_side, max_side).astype(np.int64)
    n_paths = rng.integers(paths[0], paths[1] + 1, n_scenes)
    P = int(n_paths.sum())
    scene_of_path = np.repeat(np.arange(n_scenes), n_paths)
    m = rng.integers(curves[0], curves[1] + 1, P)
    Cn = int(m.sum())
    path_of_curve = np.repeat(np.arange(P), m)
    sc_side = side[scene_of_path[path_of_curve]].astype(np.float32)[:, None]
    anchors = (rng.random((Cn, 2)).astype(np.float32) * sc_side).astype(np.float32)
    ctrl = (rng.random((Cn, 2)).astype(np.float32) * sc_side).astype(np.float32)
    first = np.concatenate([[0], np.cumsum(m)[:-1]])
    nxt = np.arange(Cn) + 1
    nxt[first + m - 1] = first
    nl = rng.integers(lines_per_curve[0], lines_per_curve[1] + 1, Cn)
    pts = quad_flatten_points(anchors, ctrl, anchors[nxt], nl)
    # lines j -> j+1 inside each curve
    cstart = np.concatenate([[0], np.cumsum(nl + 1)[:-1]])
    li = np.repeat(cstart, nl) + (np.arange(int(nl.sum())) - np.repeat(np.concatenate([[0], np.cumsum(nl)[:-1]]), nl))
    lines = np.concatenate([pts[li], pts[li + 1]], axis=1).astype(np.float32)
    curve_of_line = np.repeat(np.arange(Cn), nl)
    path_of_line = path_of_curve[curve_of_line]
    path_first = np.concatenate([[0], np.cumsum(n_paths)[:-1]])
    scene_of_line = scene_of_path[path_of_line]
    local_id = (path_of_line - path_first[scene_of_line]).astype(np.uint16)
    line_bounds = np.concatenate([[0], np.cumsum(np.bincount(scene_of_line, minlength=n_scenes))])
    crng = np.random.default_rng(seed + 1000)
    cols = srgb_u8_to_linear(crng.integers(0, 256, (P, 3)))
    scenes = []
    for s in range(n_scenes):
        a, b = line_bounds[s], line_bounds[s + 1]
        p0 = path_first[s]
        k = int(n_paths[s])
        props = dict(color=np.concatenate([cols[p0:p0 + k], np.ones((k, 1), np.float32)], axis=1),
                     rule=np.ones(k, np.uint32), blend=np.zeros(k, np.uint32))
        w = int(side[s])
        scenes.append(Scene(w, w, lines[a:b], local_id[a:b], props, [1.0, 1.0, 1.0, 1.0], name=f"glyphs-{s}"))
    return scenes


SERVE_CANVASES = ((64, 0.5), (256, 0.3), (1024, 0.15), (2048, 0.05))


def class_counts(n, probs):
    """`n` items split over classes in proportion to `probs` (largest
    remainder), every class at least once when n allows it."""
    probs = np.asarray(probs, dtype=np.float64) / np.sum(probs)
    exact = n * probs
    counts = np.floor(exact).astype(np.int64)
    for k in np.argsort(-(exact - counts), kind="stable")[: n - int(counts.sum())]:
        counts[k] += 1
    if n >= len(probs):
        while (counts == 0).any():
            counts[int(np.argmin(counts))] += 1
            counts[int(np.argmax(counts))] -= 1
    return counts


def serve_pool(seed=5, n=32, lines=(1e2, 1e6), canvases=SERVE_CANVASES):
    """Request scenes of the serving workload: square canvases of the sides
    in `canvases` (side, probability) in proportion to the probabilities
    (every side present, class_counts), in seeded random order; line count
    log-uniform in `lines`, generated as the single large frame's polygons
    (5-50 edges, box 4-512 px, affine, mixed fill rules, blend modes and
    alpha) on that canvas."""
    rng = np.random.default_rng(seed)
    counts = class_counts(n, [p for _, p in canvases])
    sides = np.repeat(np.array([c for c, _ in canvases]), counts)
    rng.shuffle(sides)
    out = []
    for i in range(n):
        side = int(sides[i])
        n_lines = float(np.exp(rng.uniform(np.log(lines[0]), np.log(lines[1]))))
        count = max(1, int(round(n_lines / 27.5)))  # 27.5 edges per path on average
        out.append(polygons(seed=seed * 1000 + i, width=side, height=side, count=count))
    return out


class Animation:
    """Keyframed animation over one composition: `paths` random polygons
    (U{16..256} edges, 50% star / 50% simple, box side log-uniform in
    [4, 512] px, centred on the origin), one layer per path with random
    colour, alpha in [0.2, 1], fill rule and blend mode. Every path has an
    affine transform (angle, scale in [0.5, 1], translation over the canvas)
    interpolated linearly between keyframes every `key_every` frames and set
    with Layer::set_transform; every frame also swaps the orders of `swap`
    random layer pairs (orders stay a permutation) and gives `restyle`
    random layers new props. All values are float32 and seeded."""

    def __init__(self, seed=4, width=1920, height=1080, paths=2000, edges=(16, 256), frames=240, key_every=30,
                 swap_frac=0.05, restyle_frac=0.02, box=(4.0, 512.0)):
        self.width, self.height, self.paths, self.frames = width, height, paths, frames
        rng = np.random.default_rng(seed)
        k = rng.integers(edges[0], edges[1] + 1, paths)
        side = np.exp(rng.uniform(np.log(box[0]), np.log(box[1]), paths)).astype(np.float32)
        star = rng.random(paths) < 0.5
        total = int(k.sum())
        path_of = np.repeat(np.arange(paths), k)
        first = np.concatenate([[0], np.cumsum(k)[:-1]])
        local = (rng.random((total, 2)).astype(np.float32) - np.float32(0.5)) * side[path_of, None]
        ang = np.arctan2(local[:, 1], local[:, 0])
        order = np.lexsort((np.arange(total), path_of * 8.0 + np.where(star[path_of], 0.0, ang + 4.0)))
        local = local[order]
        nxt = np.arange(total) + 1
        nxt[first + k - 1] = first
        self.lines = np.concatenate([local, local[nxt]], axis=1).astype(np.float32)
        self.layer_ids = path_of.astype(np.uint16)
        self._rng = np.random.default_rng(seed + 1)
        self.colors, self.rules, self.blends = self._styles(paths)
        n_keys = frames // key_every + 2
        self.key_every = key_every
        self.theta = self._rng.uniform(0, 2 * np.pi, (n_keys, paths)).astype(np.float32)
        self.scale = self._rng.uniform(0.5, 1.0, (n_keys, paths)).astype(np.float32)
        self.tx = self._rng.uniform(0, width, (n_keys, paths)).astype(np.float32)
        self.ty = self._rng.uniform(0, height, (n_keys, paths)).astype(np.float32)
        self.n_swap = int(round(paths * swap_frac / 2))
        self.n_restyle = int(round(paths * restyle_frac))
        self.orders = np.arange(paths, dtype=np.uint16)
        self.seed = seed

    def _styles(self, n):
        cols = self._rng.random((n, 4)).astype(np.float32)
        cols[:, 3] = self._rng.uniform(0.2, 1.0, n).astype(np.float32)
        return cols, self._rng.integers(0, 2, n).astype(np.uint32), self._rng.integers(0, 12, n).astype(np.uint32)

    def scene(self):
        props = dict(color=self.colors, rule=self.rules, blend=self.blends)
        return Scene(self.width, self.height, self.lines, self.layer_ids, props, [1.0, 1.0, 1.0, 1.0],
                     name=f"animation-{self.width}x{self.height}-{self.paths}")

    def transforms(self, f):
        """(paths, 6) float32 [a, b, c, d, tx, ty] of frame f: keyframes
        interpolated as k0 + (k1 - k0) * w in float32."""
        i = f // self.key_every
        w = np.float32((f % self.key_every) / self.key_every)
        lerp = lambda a: (a[i] + (a[i + 1] - a[i]) * w).astype(np.float32)
        th, sc, tx, ty = lerp(self.theta), lerp(self.scale), lerp(self.tx), lerp(self.ty)
        c = (np.cos(th).astype(np.float32) * sc).astype(np.float32)
        s = (np.sin(th).astype(np.float32) * sc).astype(np.float32)
        return np.stack([c, -s, s, c, tx, ty], axis=1).astype(np.float32)

    def updates(self, f):
        """Changes of frame f >= 1 (frame 0 only sets transforms): layer order
        swaps (ids, new orders) and restyled layers (ids, colors, rules, blends).
        Advances the animation state; call once per frame, in order."""
        if f == 0:
            return (np.zeros(0, np.uint16), np.zeros(0, np.uint16)), (np.zeros(0, np.uint16),) + self._styles(0)
        pick = self._rng.choice(self.paths, 2 * self.n_swap, replace=False)
        a, b = pick[: self.n_swap], pick[self.n_swap:]
        oa, ob = self.orders[a].copy(), self.orders[b].copy()
        self.orders[a], self.orders[b] = ob, oa
        ids = np.concatenate([a, b]).astype(np.uint16)
        restyle = self._rng.choice(self.paths, self.n_restyle, replace=False).astype(np.uint16)
        cols, rules, blends = self._styles(self.n_restyle)
        self.colors[restyle], self.rules[restyle], self.blends[restyle] = cols, rules, blends
        return (ids, self.orders[ids].copy()), (restyle, cols, rules, blends)

    def apply(self, m, comp, f):
        """Frame f's changes on composition `comp` of module `m`."""
        t = self.transforms(f)
        (oid, orders), (sid, cols, rules, blends) = self.updates(f)
        ids = np.arange(self.paths, dtype=np.uint16)
        if hasattr(comp, "set_transforms"):
            comp.set_transforms(ids, t)
            if len(oid):
                comp.set_orders(oid, orders)
            if len(sid):
                comp.set_solid_props(sid, cols, rules, blends)
        else:
            for i in range(self.paths):
                comp.get_mut(i).set_transform([float(v) for v in t[i]])
            for i, o in zip(oid.tolist(), orders.tolist()):
                comp.get_mut(i).set_order(int(o))
            for k, i in enumerate(sid.tolist()):
                comp.get_mut(i).set_props(props_of(m, dict(color=[float(v) for v in cols[k]], rule=int(rules[k]),
                                                           blend=int(blends[k]))))
