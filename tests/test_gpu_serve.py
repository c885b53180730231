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
