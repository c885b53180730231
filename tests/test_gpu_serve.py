"""The render server (mg_server_*): continuous batching under a device-memory
budget, two engines (HIP streams) in flight.

Bar: every request's RGBA8 equals the render of its composition alone (the
plain per-composition path) byte for byte, whatever batch it lands in and
whichever requests share the device with it; every request completes; the
budget is respected (a request larger than the budget is refused)."""
import threading
import time

import numpy as np
import pytest

import scenegen

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import cassia_amd
    assert cassia_amd.device_count() > 0, "no gfx950 device: the HIP path cannot run"
    return cassia_amd


@pytest.fixture(scope="module")
def pool(gpu):
    """(composition, width, height, background, RGBA8 rendered alone)."""
    from cassia_amd import workloads
    items = []
    for s in workloads.serve_pool(seed=9, n=10, lines=(1e2, 3e4)):
        items.append((s.build(gpu), s.width, s.height, s.background))
    for seed, w, h, n, kw in [(21, 83, 61, 20, dict(clips=True)), (22, 40, 72, 12, dict(gradients=True)),
                              (23, 129, 64, 30, dict(transforms=True, curves=True))]:
        items.append((scenegen.build_scene(gpu, seed, w, h, n, **kw), w, h, (1.0, 1.0, 1.0, 1.0)))
    out = []
    for comp, w, h, bg in items:
        buf = np.zeros((w * h, 4), np.uint8)
        comp.render(gpu.Buffer(buf, w), list(bg), None)
        out.append((comp, w, h, bg, buf))
    return out


def test_requests_equal_alone(gpu, pool):
    srv = gpu.Server(engines=2)
    order = np.random.default_rng(1).integers(0, len(pool), 60)
    tickets = [(int(i), srv.submit(pool[i][0], pool[i][1], pool[i][2], pool[i][3])) for i in order]
    batch_sizes = []
    for i, t in tickets:
        frame, res = srv.wait(t)
        assert res.status == 0
        batch_sizes.append(res.batch_size)
        ref = pool[i][4]
        if not np.array_equal(frame, ref):
            d = np.nonzero((frame != ref).any(axis=1))[0]
            raise AssertionError(f"request of scene {i} (batch of {res.batch_size}): {len(d)} pixels differ")
    st = srv.stats()
    assert st.completed == len(tickets) and st.failed == 0
    assert max(batch_sizes) > 1, "a burst of submissions should be batched"
    assert st.p99_ms >= st.p50_ms > 0
    srv.close()


def test_concurrent_clients_and_budget(gpu, pool):
    """Two client threads submit with random gaps under a budget that holds
    only a few requests at once: admission waits for memory, results stay
    identical to the renders alone."""
    fps = []
    probe = gpu.Server(engines=1)
    for comp, w, h, bg, _ in pool:
        t = probe.submit(comp, w, h, bg)
        _, r = probe.wait(t)
        fps.append(r.footprint)
    probe.close()
    budget = 3 * max(fps)
    srv = gpu.Server(hbm_budget=budget, engines=2)
    errors = []

    def client(seed):
        rng = np.random.default_rng(seed)
        mine = []
        for _ in range(25):
            i = int(rng.integers(0, len(pool)))
            mine.append((i, srv.submit(pool[i][0], pool[i][1], pool[i][2], pool[i][3])))
            time.sleep(float(rng.exponential(0.002)))
        for i, t in mine:
            frame, res = srv.wait(t)
            if not np.array_equal(frame, pool[i][4]):
                errors.append((seed, i, res.batch_size))

    threads = [threading.Thread(target=client, args=(s,)) for s in (1, 2)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, f"requests differing from their render alone: {errors[:5]}"
    st = srv.stats()
    assert st.completed == 50 and st.failed == 0
    assert st.inflight_bytes == 0 and st.queued == 0
    srv.close()


def test_request_over_budget_is_refused(gpu, pool):
    srv = gpu.Server(hbm_budget=1 << 16, engines=1)
    comp, w, h, bg, _ = pool[0]
    with pytest.raises(gpu.MoldError):
        srv.submit(comp, w, h, bg)
    srv.close()


def test_heavy_request_is_a_batch_of_its_own(gpu, pool):
    # 100 lines of 200,000 pixels each (mostly off a 64 x 64 canvas): about 2e7
    # estimated line-pixels, above the server's heavy mark (2^24). Submitted
    # between light requests, it renders alone and the light ones are batched
    # without it; every frame still equals its render alone.
    heavy = gpu.Composition()
    heavy.create_layers(4)
    rng = np.random.default_rng(5)
    for k in range(100):
        y0, y1 = rng.uniform(-10, 74, 2)
        p = gpu.Path()
        p.line(gpu.Point(-100000.0, float(y0)), gpu.Point(100000.0, float(y1)))
        p.line(gpu.Point(100000.0, float(y1)), gpu.Point(100000.0, float(y1) + 3.0))
        p.line(gpu.Point(100000.0, float(y1) + 3.0), gpu.Point(-100000.0, float(y0) + 3.0))
        p.close()
        heavy.insert_in_layer(k % 4, p)
    for lid in range(4):
        heavy.get_mut(lid).set_props(gpu.Props(fill_rule=gpu.FillRule(lid % 2), func=gpu.Func.Draw(
            gpu.Style(fill=gpu.Fill.Solid([0.2 * lid, 0.5, 0.3, 0.6]), blend_mode=gpu.BlendMode(lid)))))
    ref = np.zeros((64 * 64, 4), np.uint8)
    heavy.render(gpu.Buffer(ref, 64), [1.0, 1.0, 1.0, 1.0], None)
    srv = gpu.Server(engines=2)
    light = [i for i in range(len(pool)) if pool[i][1] * pool[i][2] <= 256 * 256][:6]
    tickets = []
    for n, i in enumerate(light):
        tickets.append((i, srv.submit(pool[i][0], pool[i][1], pool[i][2], pool[i][3])))
        if n == 1:
            tickets.append((-1, srv.submit(heavy, 64, 64, (1.0, 1.0, 1.0, 1.0))))
    for i, t in tickets:
        frame, res = srv.wait(t)
        assert res.status == 0
        if i < 0:
            assert res.batch_size == 1, "a heavy request renders alone"
            assert np.array_equal(frame, ref)
        else:
            assert np.array_equal(frame, pool[i][4])
    srv.close()
