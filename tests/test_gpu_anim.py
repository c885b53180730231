"""Animation frames through a captured hipGraph (mg_render_animated).

Bar: every replayed frame equals the eager render of the same state
(mg_render_device) byte for byte, sampled frames equal the oracle's render,
and a frame beyond the captured capacities falls back to an eager render with
the same bytes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = dict(seed=7, width=320, height=200, paths=150, edges=(16, 64), frames=20, key_every=5, box=(4.0, 128.0))


@pytest.fixture(scope="module")
def gpu():
    import cassia_amd
    assert cassia_amd.device_count() > 0, "no gfx950 device: the HIP path cannot run"
    return cassia_amd


def play(gpu, animated, frames=None):
    from cassia_amd import workloads
    A = workloads.Animation(**SMALL)
    comp = A.scene().build(gpu)
    out, replayed = [], []
    for f in range(A.frames if frames is None else frames):
        A.apply(gpu, comp, f)
        if animated:
            addr, rep = comp.render_animated(A.width, A.height, [1.0, 1.0, 1.0, 1.0])
        else:
            addr, rep = comp.render_device(A.width, A.height, [1.0, 1.0, 1.0, 1.0]), False
        out.append(gpu.device_frame(addr, A.width, A.height))
        replayed.append(rep)
    return out, replayed


def test_replayed_frames_equal_eager(gpu):
    g, rep = play(gpu, True)
    e, _ = play(gpu, False)
    assert not rep[0], "the first frame renders eagerly and captures"
    assert sum(rep) >= len(rep) - 3, f"graph replayed only {sum(rep)} of {len(rep)} frames"
    for f, (a, b) in enumerate(zip(g, e)):
        assert np.array_equal(a, b), f"frame {f} (replayed={rep[f]}) differs from the eager render"


def test_sampled_frames_equal_oracle(gpu):
    from cassia_amd import workloads
    from oracle import mold
    g, _ = play(gpu, True)
    A = workloads.Animation(**SMALL)
    comp = A.scene().build(mold)
    for f in range(A.frames):
        A.apply(mold, comp, f)
        if f in (0, 6, 19):
            buf = np.zeros((A.width * A.height, 4), np.uint8)
            comp.render(mold.Buffer(buf, A.width), [1.0, 1.0, 1.0, 1.0], None)
            assert np.array_equal(buf, g[f]), f"frame {f} differs from the oracle"


def test_capacity_overflow_falls_back(gpu):
    # frames at scale 0.5, then every path at scale 1: about twice the
    # line-pixels of the captured frame, beyond its 25% headroom
    from cassia_amd import workloads
    A = workloads.Animation(**SMALL)
    comp = A.scene().build(gpu)
    ids = np.arange(A.paths, dtype=np.uint16)
    rng = np.random.default_rng(3)
    tx = rng.uniform(0, A.width, A.paths).astype(np.float32)
    ty = rng.uniform(0, A.height, A.paths).astype(np.float32)

    def frame(scale):
        t = np.zeros((A.paths, 6), np.float32)
        t[:, 0] = t[:, 3] = np.float32(scale)
        t[:, 4], t[:, 5] = tx, ty
        comp.set_transforms(ids, t)
        addr, rep = comp.render_animated(A.width, A.height, [1.0, 1.0, 1.0, 1.0])
        got = gpu.device_frame(addr, A.width, A.height)
        eager = gpu.device_frame(comp.render_device(A.width, A.height, [1.0, 1.0, 1.0, 1.0]), A.width, A.height)
        assert np.array_equal(got, eager), f"scale {scale}: animated frame differs from eager"
        return rep

    # render_device between animated frames uses the same engine: the graph
    # stays valid because its key (lines, image, canvas) is unchanged
    assert not frame(0.5)
    assert frame(0.5)
    assert not frame(1.0), "a frame beyond the captured capacity must not be replayed"
    assert frame(1.0), "the fallback captures anew"
