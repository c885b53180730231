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
