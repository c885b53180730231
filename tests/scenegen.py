"""Seeded random scenes applied identically to the oracle (`oracle.mold`) and
the GPU product (`cassia_amd`), for parity tests."""
import numpy as np


def random_props(m, rng, n_layers, clips=False, gradients=False, opaque_frac=0.3):
    props = []
    for i in range(n_layers):
        rule = m.FillRule(int(rng.integers(0, 2)))
        if clips and rng.random() < 0.15:
            props.append(m.Props(fill_rule=rule, func=m.Func.Clip(int(rng.integers(1, 4)))))
            continue
        mode = m.BlendMode(int(rng.integers(0, 12)))
        alpha = 1.0 if rng.random() < opaque_frac else float(np.float32(rng.uniform(0.2, 1.0)))
        if rng.random() < 0.3:
            mode = m.BlendMode.Over
        color = [float(np.float32(v)) for v in rng.uniform(0, 1, 3)] + [alpha]
        if gradients and rng.random() < 0.25:
            g = m.GradientBuilder(list(rng.uniform(0, 64, 2)), list(rng.uniform(0, 64, 2)))
            g.type(m.GradientType(int(rng.integers(0, 2))))
            for _ in range(int(rng.integers(2, 5))):
                g.color(list(rng.uniform(0, 1, 3)) + [float(np.float32(rng.uniform(0.3, 1)))])
            fill = m.Fill.Gradient(g.build())
        else:
            fill = m.Fill.Solid(color)
        style = m.Style(fill=fill, blend_mode=mode, is_clipped=bool(clips and rng.random() < 0.3))
        props.append(m.Props(fill_rule=rule, func=m.Func.Draw(style)))
    return props


def random_polygon(rng, w, h, kind):
    k = int(rng.integers(3, 14))
    size = float(np.exp(rng.uniform(np.log(2), np.log(max(w, h) * 0.6))))
    cx, cy = rng.uniform(-0.2 * w, 1.2 * w), rng.uniform(-0.2 * h, 1.2 * h)
    if kind == 0:  # star: random vertex order
        pts = np.stack([cx + rng.uniform(-size, size, k), cy + rng.uniform(-size, size, k)], 1)
    else:  # simple polygon: angle-sorted
        ang = np.sort(rng.uniform(0, 2 * np.pi, k))
        r = size * rng.uniform(0.3, 1.0, k)
        pts = np.stack([cx + r * np.cos(ang), cy + r * np.sin(ang)], 1)
    pts = pts.astype(np.float32)
    return np.concatenate([pts, np.roll(pts, -1, 0)], 1)


def build_scene(m, seed, w, h, n_layers, clips=False, gradients=False, transforms=False, orders=True,
                curves=False):
    """Returns the composition; the same seed gives the same scene in any module."""
    rng = np.random.default_rng(seed)
    comp = m.Composition()
    props = random_props(m, rng, n_layers, clips, gradients)
    ids = []
    for i in range(n_layers):
        lid = comp.create_layer()
        ids.append(lid)
        n_paths = int(rng.integers(1, 3))
        for _ in range(n_paths):
            if curves and rng.random() < 0.5:
                P = m.Point
                p = m.Path()
                c = rng.uniform(-0.1, 1.1, 2) * [w, h]
                s = rng.uniform(2, max(w, h) / 2)
                pts = [P(*(c + rng.uniform(-s, s, 2))) for _ in range(7)]
                p.quad(pts[0], pts[1], pts[2])
                p.cubic(pts[2], pts[3], pts[4], pts[5])
                p.rat_quad((pts[5], 1.0), (pts[6], float(rng.uniform(0.5, 2))), (pts[0], 1.0))
                p.close()
                if rng.random() < 0.5:
                    t = [float(v) for v in rng.uniform(-0.5, 1.5, 9)]
                    t[0], t[4], t[8] = abs(t[0]) + 0.5, abs(t[4]) + 0.5, 1.0
                    t[6] = t[7] = 0.0
                    comp.insert_in_layer_transformed(lid, p, t)
                else:
                    comp.insert_in_layer(lid, p)
            else:
                comp.insert_lines(lid, random_polygon(rng, w, h, int(rng.integers(0, 2))))
        layer = comp.get_mut(lid)
        layer.set_props(props[i])
        if transforms and rng.random() < 0.4:
            th = rng.uniform(0, 2 * np.pi)
            sc = rng.uniform(0.5, 1.0)
            layer.set_transform([sc * np.cos(th), -sc * np.sin(th), sc * np.sin(th), sc * np.cos(th),
                                 rng.uniform(-w / 4, w / 4), rng.uniform(-h / 4, h / 4)])
    if orders:
        perm = rng.permutation(n_layers)
        for i, lid in enumerate(ids):
            comp.get_mut(lid).set_order(int(perm[i]))
    return comp
