"""GPU parity at the batching, animation and serving configurations, against
the oracle (and, for every batched frame, against the plain render):

- batching: the 4,096-scene glyph batch (power-law canvases 32..512 px,
  1-200 glyph paths per scene) as one batched frame; every frame equals its
  plain per-composition render, and 64 seeded frames equal the oracle's;
- animation: the 1920x1080, 2,000-path keyframed animation through the
  captured hipGraph; its first frames (the eager capture frame and replays)
  equal the oracle's render of the same state;
- serving: a mix of requests over all four canvas classes (64^2 .. 2048^2),
  one of them above 5e5 lines, submitted as one burst so that they batch
  and overlap on the device; every result equals the oracle's render."""
import os

import numpy as np
import pytest

from test_gpu_parity import assert_same_image

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


@pytest.fixture(scope="module")
def gpu():
    import cassia_amd
    assert cassia_amd.device_count() > 0, "no gfx950 device: the HIP path cannot run"
    return cassia_amd


@pytest.fixture(scope="module")
def cpu():
    from oracle import mold
    return mold


def oracle_frame(cpu, scene):
    buf = np.zeros((scene.width * scene.height, 4), np.uint8)
    scene.build(cpu).render_mt(cpu.Buffer(buf, scene.width), scene.background, THREADS)
    return buf


def test_glyph_batch_full_config(gpu, cpu):
    from cassia_amd import workloads
    scenes = workloads.glyph_batch()
    assert len(scenes) == 4096
    comps = [s.build(gpu) for s in scenes]
    b = gpu.Batch()
    b.render(comps, [(s.width, s.height, s.background) for s in scenes])
    st = b.stats()
    print(f"glyph batch: {st.lines} lines, {st.segments} segments, radix {st.radix_bits}", flush=True)
    assert st.lines == sum(len(s.lines) for s in scenes)
    for i, (s, c) in enumerate(zip(scenes, comps)):
        plain = gpu.device_frame(c.render_device(s.width, s.height, s.background), s.width, s.height)
        got = b.read(i)
        if not np.array_equal(plain, got):
            raise AssertionError(f"glyph scene {i} ({s.width}x{s.height}): batched frame differs from plain")
    sample = np.random.default_rng(17).choice(len(scenes), 64, replace=False)
    # the largest canvases too
    sample = sorted(set(sample.tolist()) | set(np.argsort([-s.width for s in scenes])[:4].tolist()))
    for i in sample:
        assert_same_image(oracle_frame(cpu, scenes[i]), b.read(i), scenes[i].width)


def test_animation_full_config(gpu, cpu):
    from cassia_amd import workloads
    Ag, Ac = workloads.Animation(), workloads.Animation()
    assert (Ag.width, Ag.height, Ag.paths) == (1920, 1080, 2000)
    g = Ag.scene().build(gpu)
    c = Ac.scene().build(cpu)
    replayed = []
    for f in range(4):
        Ag.apply(gpu, g, f)
        Ac.apply(cpu, c, f)
        addr, rep = g.render_animated(Ag.width, Ag.height, [1.0, 1.0, 1.0, 1.0])
        replayed.append(rep)
        got = gpu.device_frame(addr, Ag.width, Ag.height)
        ref = np.zeros((Ac.width * Ac.height, 4), np.uint8)
        c.render_mt(cpu.Buffer(ref, Ac.width), [1.0, 1.0, 1.0, 1.0], THREADS)
        print(f"animation frame {f}: replayed {rep}", flush=True)
        assert_same_image(ref, got, Ag.width)
    assert not replayed[0] and sum(replayed) >= 2, replayed


def serving_mix():
    """Eight pool requests over all canvas classes, a 2048x2048 request and a
    64x64 request of about 6e5 lines (the polygons generator)."""
    from cassia_amd import workloads
    mix = workloads.serve_pool(seed=11, n=8, lines=(1e2, 2e5))
    mix.append(workloads.polygons(seed=1101, width=2048, height=2048, count=4000))
    mix.append(workloads.polygons(seed=1102, width=64, height=64, count=22_000))
    return mix


def test_serving_mix_against_oracle(gpu, cpu):
    mix = serving_mix()
    sides = sorted({s.width for s in mix})
    assert sides == [64, 256, 1024, 2048], sides
    assert max(len(s.lines) for s in mix) >= 5e5
    comps = [s.build(gpu) for s in mix]
    srv = gpu.Server(engines=2)
    tickets = [srv.submit(c, s.width, s.height, s.background) for s, c in zip(mix, comps)]
    results = [srv.wait(t) for t in tickets]
    srv.close()
    assert max(r.batch_size for _, r in results) > 1, "the burst should be batched"
    for k, (s, (frame, res)) in enumerate(zip(mix, results)):
        assert res.status == 0
        print(f"request {k}: {s.width}x{s.width}, {len(s.lines)} lines, batch of {res.batch_size}", flush=True)
        assert_same_image(oracle_frame(cpu, s), frame, s.width)
