"""mold's composition KATs against the CPU oracle (mold/src/composition.rs:331-798)."""
import pytest

import mold_scenes
from oracle import mold


@pytest.mark.parametrize("scene", mold_scenes.ALL, ids=lambda f: f.__name__)
def test_composition_kat(scene):
    scene(mold)


def test_id_set_acquire_release():
    # layer.rs:258-290 (ids exhausted -> None)
    c = mold.Composition()
    ids = [c.create_layer() for _ in range(65536)]
    assert len(set(ids)) == 65536
    assert c.create_layer() is None
    c._release(17)
    assert c.create_layer() == 17


def test_set_transform_scale_panics():
    l = mold.Layer()
    with pytest.raises(ValueError):
        l.set_transform([2.0, 0.0, 0.0, 1.0, 0.0, 0.0])


def test_render_mt_matches_render():
    """The threaded restatement (the bench's CPU baseline) paints the same
    bytes and counts the same segments as the single-threaded one."""
    import numpy as np
    from cassia_amd.workloads import polygons
    scene = polygons(seed=5, width=160, height=96, count=300, smax=64.0)
    outs = []
    for threads in (None, 4):
        comp = scene.build(mold)
        buf = np.zeros((scene.width * scene.height, 4), np.uint8)
        if threads is None:
            n = comp.render(mold.Buffer(buf, scene.width), scene.background)
        else:
            n = comp.render_mt(mold.Buffer(buf, scene.width), scene.background, threads)
        outs.append((n, buf))
    assert outs[0][0] == outs[1][0] > 0
    assert np.array_equal(outs[0][1], outs[1][1])
