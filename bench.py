#!/usr/bin/env python3
"""Benchmark of the mold rasterize -> sort -> paint path on one MI355X.

Workload (default): the single large frame -- 3840x2160, 200,000 star /
polygon paths of 5-50 edges, random affine transforms (rotation, scale
0.5-2), 50/50 fill rules, uniform blend modes, alpha in [0.2, 1]
(cassia_amd.workloads.polygons). One step = one full frame: line setup,
scan, pixel-segment emission, the LSD radix sort and the tile-row paint into
an RGBA8 frame. Lines, layer tables and the frame stay resident in HBM; no
host<->device copy is inside the timed region.

Batched workload (the "batched" object of the JSON line): 4,096 independent
glyph scenes of power-law canvas sizes 32..512 px, 1-200 paths of 2-6
quadratic Beziers flattened to 8-32 lines each (cassia_amd.workloads.
glyph_batch). Plain = one mg_render_device frame per scene; batched = the
same 4,096 compositions as one mg_batch_render frame, with the autotuner's
sort radix. Every batched frame is compared byte for byte with its plain
render.

Prints one JSON line (see the keys in `main`).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frames/s & pixel-segments/s batched vs plain on 1 MI355X, bit-exact; p99 ms"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak
# vector-ALU issue: 256 CUs x 4 SIMDs at 2.4 GHz. A wave64 f32 or 3-operand
# instruction takes 4 cycles of its 16-lane SIMD; 2-operand 32-bit integer and
# logic instructions issue about every 2 cycles on this device.
SIMD_CYCLES_PER_S = 256 * 4 * 2.4e9
VALU_PEAK_PER_S = SIMD_CYCLES_PER_S / 4


# kernel -> stats field with its per-frame time (hipEvents on the library's
# stream, bracketing only that kernel's launches; the sort passes' own events
# in sort_pass_ms)
KERNELS = {"k_emit": "emit_ms", "k_sort_pass": "sort_pass_ms", "k_tile_runs": "runs_ms",
           "k_row_walk": "walk_ms", "k_paint": "paint_ms"}


def kernel_ms(st, field):
    """Per-frame ms of one KERNELS entry: the stage time, or the sum of the
    digit pass kernels."""
    if field == "sort_pass_ms":
        return sum(st.sort_pass_ms[p] for p in range(st.sort_passes))
    return getattr(st, field)


def kernel_bytes(st, width, height, n_layers):
    """(launches per frame, algorithmic HBM bytes per launch) per kernel.
    L lines, S kept pixel segments, P digit passes, T tiles, R tile runs:
      k_emit       18 L + 8 S   read each line (16 B points + 2 B id), write each segment
      k_sort_pass  16 S         read + write every key once per digit pass
      k_tile_runs  8 S + 8 T    read the sorted stream and the tile ranges
      k_row_walk   8 R          the cover carry: 8 row covers per run
      k_paint      8 S + 8 T + 4 W H + 32 L   accumulate (stream, tile ranges) and
                                composite (RGBA8 frame, layer styles)"""
    L, S, P, T, R = st.lines, st.segments_painted, st.sort_passes, st.tiles, st.runs
    return {
        "k_emit": (1, 18 * L + 8 * S),
        "k_sort_pass": (P, 16 * S),
        "k_tile_runs": (1, 8 * S + 8 * T),
        "k_row_walk": (1, 8 * R),
        "k_paint": (1, 8 * S + 8 * T + 4 * width * height + 32 * n_layers),
    }


def host_cpu():
    """Host CPU facts for the baseline: logical CPUs, the CPUs this process may
    run on, the cgroup CPU quota, and the model name."""
    nproc = os.cpu_count() or 1
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = nproc
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": nproc, "affinity": affinity, "cgroup_quota": quota, "model": model}


def baseline_threads(info, share=16):
    """Worker threads of the multi-threaded baseline: the CPUs this process
    may use, bounded by the cgroup quota and by one GPU's share of the host
    (`share` CPUs on the GPU box, whatever nproc shows)."""
    n = info["affinity"]
    if info["cgroup_quota"]:
        n = min(n, max(1, int(info["cgroup_quota"])))
    return max(1, min(n, share))


def cpu_baseline(circles_scene, frame_scene, gpu_frame, runs=5):
    """The CPU baseline and the headline frame's parity check, both on the
    oracle (the C restatement of the reference CPU path, kind "port"):
    - the plain 1024x1024 circles frame rendered single-threaded and on
      `baseline_threads` threads, median of `runs` timed renders after one
      warm-up each (render() from the inserted composition to RGBA8);
    - the large frame rendered once on the same threads, compared byte for
      byte with the GPU's timed frame (`gpu_frame`)."""
    from oracle import mold
    info = host_cpu()
    threads = baseline_threads(info)
    sc = circles_scene
    comp = sc.build(mold)
    res = {}
    for label, th in (("single", 1), ("multi", threads)):
        times, segs = [], 0
        for r in range(runs + 1):
            buf = np.zeros((sc.width * sc.height, 4), np.uint8)
            t0 = time.perf_counter()
            segs = comp.render_mt(mold.Buffer(buf, sc.width), sc.background, th)
            dt = time.perf_counter() - t0
            if r > 0:
                times.append(dt)
        med = float(np.median(times))
        res[label] = {"threads": th, "median_s": med, "runs_s": [round(t, 4) for t in times],
                      "frames_per_s": 1.0 / med, "pixel_segments_per_s": segs / med, "pixel_segments": int(segs)}
        progress(f"cpu baseline {label} ({th} threads): {med * 1e3:.1f} ms per circles frame")
    fs = frame_scene
    fc = fs.build(mold)
    ref = np.zeros((fs.width * fs.height, 4), np.uint8)
    t0 = time.perf_counter()
    fsegs = fc.render_mt(mold.Buffer(ref, fs.width), fs.background, threads)
    fdt = time.perf_counter() - t0
    same = bool(np.array_equal(ref, gpu_frame))
    diff = 0 if same else int((ref != gpu_frame).any(axis=1).sum())
    progress(f"large frame on the CPU: {fdt:.2f} s, bit-identical to the GPU frame: {same}")
    mt = res["multi"]
    return {
        "value": mt["pixel_segments_per_s"], "unit": "pixel-segments/s", "cores": threads, "kind": "port",
        "sample": (f"plain 1024x1024 frame of 10,000 circles x 64 lines ({mt['pixel_segments']} pixel segments), "
                   f"median of {runs} renders after a warm-up, {threads} threads; C restatement of the "
                   f"reference CPU path (oracle/oracle.c)"),
        "frames_per_s": mt["frames_per_s"],
        "single_thread": res["single"], "multi_thread": mt,
        "host": info,
        "large_frame": {"seconds": fdt, "threads": threads, "pixel_segments": int(fsegs),
                        "pixel_segments_per_s": fsegs / fdt, "bit_identical_to_gpu": same,
                        "pixels_differing": diff},
    }


def progress(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def bench_batched(cassia_amd, workloads, steps, warmup, n_scenes, plain=True):
    """Plain per-scene frames vs one batched frame over the same compositions
    (`plain` False: the batched frame alone, for profiling it)."""
    t = time.perf_counter()
    scenes = workloads.glyph_batch(n_scenes=n_scenes)
    comps = [s.build(cassia_amd) for s in scenes]
    t_build = time.perf_counter() - t
    progress(f"{n_scenes} glyph scenes built in {t_build:.1f} s")
    frames = [(s.width, s.height, s.background) for s in scenes]
    # plain: one frame per scene (a first pass sizes each composition's buffers)
    plain_dev_ms = 0.0
    plain_segs = 0
    plain_s = float("nan")
    if plain:
        for c, (w, h, bg) in zip(comps, frames):
            c.render_device(w, h, bg)
        t = time.perf_counter()
        for c, (w, h, bg) in zip(comps, frames):
            c.render_device(w, h, bg)
        plain_s = time.perf_counter() - t
        progress(f"plain: {n_scenes} frames in {plain_s:.2f} s")
        for c in comps:
            st = c.stats()
            plain_dev_ms += st.device_ms
            plain_segs += st.segments
    # batched: the first render runs the autotuner
    b = cassia_amd.Batch()
    t = time.perf_counter()
    b.render(comps, frames)
    tune_s = time.perf_counter() - t
    tn = b.tuning()
    progress(f"autotuned in {tune_s:.2f} s: radix {tn.chosen_radix_bits}")
    for _ in range(warmup):
        b.render(comps, frames)
    lat = []
    dev_ms = 0.0
    t = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        b.render(comps, frames)
        lat.append((time.perf_counter() - a) * 1e3)
        dev_ms += b.stats().device_ms
    batch_s = (time.perf_counter() - t) / steps
    st = b.stats()
    progress(f"batched: {batch_s * 1e3:.2f} ms per batch")
    # bit-identity: every batched frame against its plain render
    same = 0
    for i, (c, (w, h, bg)) in enumerate(zip(comps, frames)):
        if not plain:
            break
        hb = np.zeros((w * h, 4), np.uint8)
        c.render(cassia_amd.Buffer(hb, w), bg, None)
        same += int(np.array_equal(hb, b.read(i)))
    n = len(scenes)
    return {
        "workload": (f"{n} glyph scenes, canvas round(32*16^(u^3)) px square, 1-200 paths of 2-6 quadratic "
                     "Beziers x 8-32 lines, EvenOdd, opaque solid, Over"),
        "scenes": n,
        "lines": int(st.lines),
        "pixels": int(sum(w * h for w, h, _ in frames)),
        "pixel_segments": int(st.segments),
        "plain_frames_per_s": n / plain_s,
        "plain_ms_per_batch": plain_s * 1e3,
        "plain_device_ms_per_batch": plain_dev_ms,
        "plain_pixel_segments_per_s": plain_segs / plain_s,
        "batched_frames_per_s": n / batch_s,
        "batched_ms_per_batch": batch_s * 1e3,
        "batched_device_ms": dev_ms / steps,
        "batched_pixel_segments_per_s": st.segments / batch_s,
        "batched_p50_ms": float(np.percentile(lat, 50)),
        "batched_p99_ms": float(np.percentile(lat, 99)),
        "speedup_frames_per_s": plain_s / batch_s,
        "bit_identical_frames": same,
        "bit_identical": same == n if plain else None,
        "radix_bits": int(st.radix_bits),
        "sort_passes": int(st.sort_passes),
        "autotune": {"candidates": {int(tn.radix_bits[k]): round(tn.device_ms[k], 3) for k in range(tn.n_candidates)},
                     "chosen": int(tn.chosen_radix_bits),
                     "walk_candidates": {int(tn.walk_threads[k]): round(tn.walk_ms[k], 3)
                                         for k in range(tn.n_walk_candidates)},
                     "chosen_walk_threads": int(tn.chosen_walk_threads), "seconds": round(tune_s, 3)},
        "walk_threads": int(st.walk_threads), "walk_lds": int(st.walk_lds),
        # dispatches of one batched frame (kernels, clears and status reads),
        # from the committed kernel trace of this workload
        "dispatches_per_frame": profile_value("kernel-trace_glyphs-4096-scenes", "dispatches_per_frame"),
        "stage_ms": {k: round(getattr(st, k), 3) for k in ("setup_ms", "emit_ms", "hist_ms", "sort_ms", "runs_ms",
                                                             "walk_ms", "paint_ms")},
        "scene_build_s": round(t_build, 2),
    }


def bench_serving(cassia_amd, workloads, n_requests=160, fractions=(0.25, 0.5, 0.8), seed=5):
    """Poisson request streams against the render server: a pool of 32
    request scenes (serve_pool), each request one of them drawn at random.
    The capacity is the rate a closed loop with four requests in flight
    sustains over every scene four times (a burst of the same requests, which
    batches more of them at once, is reported beside it); each stream then
    arrives at a fraction of the capacity. Latency = submission -> RGBA8 in the
    caller's buffer; every result is compared with the scene's render alone."""
    import threading
    t = time.perf_counter()
    scenes = workloads.serve_pool(seed=seed, n=32)
    comps = [s.build(cassia_amd) for s in scenes]
    progress(f"serving pool: {len(scenes)} scenes, {sum(len(s.lines) for s in scenes)} lines "
             f"built in {time.perf_counter() - t:.1f} s")
    refs, alone = [], []
    for s, c in zip(scenes, comps):
        buf = np.zeros((s.width * s.height, 4), np.uint8)
        c.render(cassia_amd.Buffer(buf, s.width), s.background, None)
        refs.append(buf)
        c.render_device(s.width, s.height, s.background)
        st = c.stats()
        alone.append({"side": s.width, "lines": int(st.lines), "line_pixels": int(st.line_pixels),
                      "segments": int(st.segments), "painted": int(st.segments_painted),
                      "device_ms": round(st.device_ms, 3)})
    progress("pool alone: " + "; ".join(f"{a['side']}px {a['lines']}L {a['device_ms']}ms" for a in alone))
    srv = cassia_amd.Server(engines=2)

    def run_stream(arrivals, picks):
        done = [None] * len(picks)
        tickets = [None] * len(picks)
        ready = threading.Semaphore(0)

        def waiter():
            for k in range(len(picks)):
                ready.acquire()
                try:
                    frame, res = srv.wait(tickets[k])
                    done[k] = (np.array_equal(frame, refs[picks[k]]), res.latency_ms, res.batch_size)
                except cassia_amd.MoldError as e:
                    progress(f"request {k} (scene {picks[k]}) failed: {e}")
                    done[k] = (False, float("nan"), 0)

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for k, (a, i) in enumerate(zip(arrivals, picks)):
            dt = t0 + a - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            s = scenes[i]
            tickets[k] = srv.submit(comps[i], s.width, s.height, s.background)
            ready.release()
        th.join()
        return done, time.perf_counter() - t0

    def run_closed(picks, inflight):
        """Each completion submits the next request, `inflight` outstanding."""
        tickets = {}
        ok = True
        t0 = time.perf_counter()
        for k in range(min(inflight, len(picks))):
            s = scenes[picks[k]]
            tickets[k] = srv.submit(comps[picks[k]], s.width, s.height, s.background)
        for k in range(len(picks)):
            frame, _ = srv.wait(tickets.pop(k))
            ok = ok and np.array_equal(frame, refs[picks[k]])
            n = k + inflight
            if n < len(picks):
                s = scenes[picks[n]]
                tickets[n] = srv.submit(comps[picks[n]], s.width, s.height, s.background)
        return ok, time.perf_counter() - t0

    rng = np.random.default_rng(seed)
    picks = [i for _ in range(4) for i in range(len(scenes))]
    # a burst (everything submitted at once) batches many requests together
    srv.stats(reset=True)
    done, el = run_stream([0.0] * len(picks), picks)
    st = srv.stats(reset=True)
    progress(f"serving burst: {len(picks) / el:.0f} requests/s ({len(picks)} requests)")
    burst = {"requests": len(picks), "requests_per_s": len(picks) / el, "p50_ms": st.p50_ms, "p99_ms": st.p99_ms,
             "mean_batch": st.mean_batch, "identical": all(d[0] for d in done)}
    # capacity: the rate a closed loop of two requests in flight per engine
    # sustains, in batches as small as an open stream near that rate forms
    inflight = 4
    ok, el = run_closed(picks, inflight)
    st = srv.stats(reset=True)
    capacity = len(picks) / el
    progress(f"serving capacity: {capacity:.0f} requests/s (closed loop, {inflight} in flight)")
    closed = {"requests": len(picks), "in_flight": inflight, "requests_per_s": capacity, "p50_ms": st.p50_ms,
              "p99_ms": st.p99_ms, "mean_batch": st.mean_batch, "identical": ok}
    rates = []
    for frac in fractions:
        lam = frac * capacity
        gaps = rng.exponential(1.0 / lam, n_requests)
        arrivals = np.cumsum(gaps) - gaps[0]
        picks = [int(i) for i in rng.integers(0, len(scenes), n_requests)]
        done, el = run_stream(list(arrivals), picks)
        st = srv.stats(reset=True)
        lat = np.array([d[1] for d in done])
        side = np.array([scenes[i].width for i in picks])
        per_class = {}
        for c in sorted(set(side.tolist())):
            lc = lat[side == c]
            per_class[f"{c}x{c}"] = {"requests": int(len(lc)), "p50_ms": float(np.percentile(lc, 50)),
                                     "p99_ms": float(np.percentile(lc, 99)), "max_ms": float(lc.max())}
        rates.append({"arrival_rate": lam, "requests": n_requests, "requests_per_s": n_requests / el,
                      "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                      "max_ms": float(lat.max()), "mean_batch": st.mean_batch,
                      "max_concurrent_batches": int(st.max_concurrent_batches),
                      "per_canvas": per_class,
                      "identical": all(d[0] for d in done)})
        progress(f"serving at {lam:.0f}/s: p50 {rates[-1]['p50_ms']:.2f} ms, p99 {rates[-1]['p99_ms']:.2f} ms")
    srv.close()
    sides = [s.width for s in scenes]
    return {
        "workload": ("Poisson request streams; request scene = canvas {64,256,1024,2048}^2 in proportion "
                     ".5/.3/.15/.05, log-uniform 1e2..1e6 lines generated as the large frame's polygons; "
                     "pool of 32 scenes, each request one of them drawn uniformly"),
        "pool_lines": int(sum(len(s.lines) for s in scenes)),
        "pool_canvases": {str(c): sides.count(c) for c in sorted(set(sides))},
        "pool_alone_device_ms": round(sum(a["device_ms"] for a in alone), 3),
        "pool_alone": alone,
        "engines": 2,
        "burst": burst,
        "capacity": closed,
        "rates": rates,
        "identical": burst["identical"] and closed["identical"] and all(r["identical"] for r in rates),
    }


def bench_animation(cassia_amd, workloads, frames=240, sample_every=10):
    """The keyframed animation (workloads.Animation: 1920x1080, 2,000 paths of
    16-256 lines, per-frame transforms of every layer, 5% of orders swapped,
    2% restyled) played twice from the same seed: through the captured
    hipGraph (mg_render_animated) and eagerly (mg_render_device). A frame's
    time covers its host-side layer updates and the render; every
    `sample_every`-th frame of both runs is compared byte for byte."""
    out = {}
    samples = {}
    for mode in ("graph", "eager"):
        A = workloads.Animation(frames=frames)
        comp = A.scene().build(cassia_amd)
        lat, dev, replayed = [], [], 0
        t0 = time.perf_counter()
        for f in range(frames):
            a = time.perf_counter()
            A.apply(cassia_amd, comp, f)
            if mode == "graph":
                addr, rep = comp.render_animated(A.width, A.height, [1.0, 1.0, 1.0, 1.0])
                replayed += int(rep)
            else:
                addr = comp.render_device(A.width, A.height, [1.0, 1.0, 1.0, 1.0])
            lat.append((time.perf_counter() - a) * 1e3)
            dev.append(comp.stats().device_ms)
            if f % sample_every == 0:
                c0 = time.perf_counter()
                samples.setdefault(f, []).append(cassia_amd.device_frame(addr, A.width, A.height))
                t0 += time.perf_counter() - c0  # the sample copy is not frame time
        el = time.perf_counter() - t0
        st = comp.stats()
        out[mode] = {"frames_per_s": frames / el, "ms_per_frame": el * 1e3 / frames,
                     "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                     "device_ms_per_frame": float(np.mean(dev)), "replayed_frames": replayed,
                     "pixel_segments_last_frame": int(st.segments), "lines": int(st.lines)}
        progress(f"animation {mode}: {out[mode]['ms_per_frame']:.2f} ms/frame")
    same = all(np.array_equal(v[0], v[1]) for v in samples.values())
    return {"workload": ("keyframed animation, 1920x1080, 2000 random polygons of 16-256 lines, per-frame affine of "
                         "every layer, 5% order swaps and 2% restyles per frame"),
            "frames": frames, **out,
            "speedup_graph_vs_eager": out["eager"]["ms_per_frame"] / out["graph"]["ms_per_frame"],
            "sampled_frames_identical": same, "sampled_frames": len(samples)}


def bench_flatten(cassia_amd, n_paths=100_000, seed=12):
    """Curve ingestion: glyph-like closed loops of 2-6 quadratic Beziers over
    a 512-px box (as the batched scenes' paths, before flattening), inserted
    into one composition with the flattening on the GPU
    (mg_insert_paths_device) and on the host (mg_insert_paths, C++); the
    two line sets are compared bit for bit."""
    rng = np.random.default_rng(seed)
    P = cassia_amd.Point
    paths = []
    for _ in range(n_paths):
        p = cassia_amd.Path()
        pts = rng.uniform(0, 512, (int(rng.integers(2, 7)), 2, 2)).astype(np.float32)
        start = P(*pts[0, 0])
        cur = start
        for k in range(len(pts)):
            nxt = P(*pts[k + 1, 0]) if k + 1 < len(pts) else start
            p.quad(cur, P(*pts[k, 1]), nxt)
            cur = nxt
        p.close()
        paths.append(p)
    ids = (np.arange(n_paths) % 4096).astype(np.uint16)
    out = {}
    lines = {}
    for mode in ("device", "host"):
        best = None
        for rep in range(3):
            comp = cassia_amd.Composition()
            comp.create_layers(4096)
            t = time.perf_counter()
            comp.insert_paths_device(ids, paths, device=(mode == "device"))
            if mode == "host":
                comp.render_device(8, 8, [1.0, 1.0, 1.0, 1.0])  # the lines reach the device too
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        n = comp.builder_len()
        out[mode] = {"ms": best * 1e3, "lines": int(n), "lines_per_s": n / best}
        lines[mode] = comp.lines()
    same = bool(np.array_equal(lines["device"][0].view(np.uint32), lines["host"][0].view(np.uint32)) and
                np.array_equal(lines["device"][1], lines["host"][1]))
    progress(f"flatten: device {out['device']['ms']:.2f} ms, host {out['host']['ms']:.2f} ms")
    return {"workload": f"{n_paths} closed loops of 2-6 quadratic Beziers in a 512-px box, 4096 layers",
            **out, "speedup": out["host"]["ms"] / out["device"]["ms"], "identical": same}


def profile_value(summary, key, default=None):
    """The JSON value of the `key:` line of a committed rocprofv3 summary
    (profiles/<summary>.md), or `default` when there is none."""
    path = os.path.join(ROOT, "profiles", f"{summary}.md")
    if not os.path.exists(path):
        return default
    for line in open(path):
        if line.startswith(key + ":"):
            return json.loads(line.split(":", 1)[1])
    return default


def measured_traffic(workload):
    """HBM bytes per launch per kernel from the committed rocprofv3 PMC
    summary of this workload (FETCH_SIZE and WRITE_SIZE runs, each scaled by
    the counter's reading for a 1 GiB device copy), or {} if there is none."""
    return profile_value(f"hbm-traffic_{workload}", "traffic_bytes_per_launch", {})


def measured_valu(workload):
    """Wave-level VALU instructions per launch per kernel from the committed
    rocprofv3 SQ counter summary of this workload, or {}."""
    return profile_value(f"sq-instructions_{workload}", "valu_per_launch", {})


def measured_valu_f32(workload):
    """Of those, f32 arithmetic instructions (add, fma, mul, cvt, trans), or {}."""
    return profile_value(f"sq-instructions_{workload}", "valu_f32_per_launch", {})


def bench_circles(cassia_amd, workloads, steps, warmup):
    """The plain configuration on the GPU: 1024x1024, 10,000 circles x 64
    lines, one frame per step into a device-resident frame."""
    sc = workloads.circles()
    comp = sc.build(cassia_amd)
    for _ in range(warmup):
        comp.render_device(sc.width, sc.height, sc.background)
    dev = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        comp.render_device(sc.width, sc.height, sc.background)
        dev += comp.stats().device_ms
    el = time.perf_counter() - t0
    st = comp.stats()
    progress(f"circles: {el * 1e3 / steps:.3f} ms per frame")
    return {"workload": "plain method: 1024x1024, 10,000 circles x 64 lines, opaque, NonZero, Over",
            "frames_per_s": steps / el, "ms_per_frame": el * 1e3 / steps, "device_ms_per_frame": dev / steps,
            "pixel_segments": int(st.segments), "pixel_segments_per_s": st.segments * steps / el,
            "lines": int(st.lines)}, sc


def copy_calibration(nbytes):
    """One device-to-device streaming copy of `nbytes` (mg_stream_copy, the
    library's non-temporal copy kernel k_copy_stream): the reference dispatch
    whose FETCH_SIZE / WRITE_SIZE counters scale the renderer's. Then 20
    timed copies: the streaming-copy rate (read + write bytes per second), the
    measured HBM ceiling the bandwidth-bound kernels compare with."""
    import cassia_amd
    cassia_amd.stream_copy(nbytes, 1)
    ms, bad = cassia_amd.stream_copy(nbytes, 20)
    print(json.dumps({"copy_bytes": nbytes, "copy_kernel": "mg::k_copy_stream", "copy_ms": round(ms, 4),
                      "copy_gbs": round(2 * nbytes / ms / 1e6, 1), "copy_exact": bad == 0}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["polygons", "circles"], default="polygons")
    ap.add_argument("--paths", type=int, default=200_000)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline and the oracle check of the large frame")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched many-scenes measurement")
    ap.add_argument("--batch-scenes", type=int, default=4096)
    ap.add_argument("--no-frame", action="store_true",
                    help="measure only the batched many-scenes workload (prints its object alone)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-anim", action="store_true", help="skip the animation measurement")
    ap.add_argument("--no-flatten", action="store_true", help="skip the curve flattening measurement")
    ap.add_argument("--only-anim", action="store_true", help="measure only the animation workload")
    ap.add_argument("--only-flatten", action="store_true", help="measure only the curve flattening")
    ap.add_argument("--only-serve", action="store_true", help="measure only the serving workload")
    ap.add_argument("--no-plain", action="store_true",
                    help="with --no-frame: skip the plain per-scene frames and the identity check")
    ap.add_argument("--radix", type=int, default=0, help="sort digit width (0 = automatic)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only copy a device buffer of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_bytes:
        copy_calibration(args.copy_bytes)
        return
    if args.gpus != 1:
        print("bench.py: this build renders on one GPU only; multi-GPU runs (tile-row bands "
              "exchanged with one all-to-all) are not implemented", file=sys.stderr)
        sys.exit(2)

    import cassia_amd
    from cassia_amd import workloads
    if cassia_amd.device_count() < 1:
        raise SystemExit("no gfx950 device")
    if args.only_flatten:
        print(json.dumps({"flatten": bench_flatten(cassia_amd)}))
        return
    if args.only_anim:
        print(json.dumps({"animation": bench_animation(cassia_amd, workloads)}))
        return
    if args.only_serve:
        print(json.dumps({"serving": bench_serving(cassia_amd, workloads)}))
        return
    if args.no_frame:
        print(json.dumps({"batched": bench_batched(cassia_amd, workloads, args.steps, args.warmup,
                                                   args.batch_scenes, plain=not args.no_plain)}))
        return
    if args.workload == "polygons":
        scene = workloads.polygons(count=args.paths)
        workload = (f"single large frame, plain GPU path: {scene.width}x{scene.height}, "
                    f"{args.paths} star/polygon paths, 5-50 edges, affine, mixed fill rules and blend modes")
        profile_name = "polygons-3840x2160" if args.paths == 200_000 else None
    else:
        scene = workloads.circles()
        workload = "plain method: 1024x1024, 10,000 circles x 64 lines, opaque, NonZero, Over"
        profile_name = None
    t_build = time.perf_counter()
    comp = scene.build(cassia_amd)
    t_build = time.perf_counter() - t_build
    progress(f"{scene.name} built in {t_build:.1f} s")
    W, H = scene.width, scene.height
    comp.set_tuning(args.radix)

    for _ in range(args.warmup):
        comp.render_device(W, H, scene.background)
    lat = []
    kms = {k: 0.0 for k in KERNELS}
    pass_ms = [0.0] * 16
    sort_stage_ms = 0.0
    dev_ms = 0.0
    addr = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        addr = comp.render_device(W, H, scene.background)
        lat.append((time.perf_counter() - a) * 1e3)
        st = comp.stats()
        for k, f in KERNELS.items():
            kms[k] += kernel_ms(st, f)
        for p in range(st.sort_passes):
            pass_ms[p] += st.sort_pass_ms[p]
        sort_stage_ms += st.hist_ms + st.sort_ms
        dev_ms += st.device_ms
    elapsed = time.perf_counter() - t0
    progress(f"{args.steps} frames in {elapsed:.2f} s")
    # the last timed frame, for the oracle check in the CPU-baseline leg
    gpu_frame = cassia_amd.device_frame(addr, W, H) if args.workload == "polygons" else None
    st = comp.stats()
    K = args.steps
    kms = {k: v / K for k, v in kms.items()}
    ms_per_step = elapsed * 1e3 / K
    fps = K / elapsed
    segs = st.segments
    kb = kernel_bytes(st, W, H, scene.n_layers)
    kernels = {}
    for k, ms in kms.items():
        n, b = kb[k]
        per_launch_ms = ms / n if n else 0.0
        kernels[k] = {"ms_per_frame": round(ms, 4), "launches": n, "ms_per_launch": round(per_launch_ms, 4),
                      "bytes_per_launch": b,
                      "gbs": round(b / (per_launch_ms * 1e-3) / 1e9, 2) if per_launch_ms > 0 else None}
    kernels["k_sort_pass"]["pass_ms"] = [round(pass_ms[p] / K, 4) for p in range(st.sort_passes)]
    # the whole sort stage: the passes, the later passes' digit counts and
    # the digit-count scans, against the passes' algorithmic bytes
    n_sort, b_sort = kb["k_sort_pass"]
    stage = sort_stage_ms / K
    sort_stage = {"ms_per_frame": round(stage, 4), "bytes_per_frame": n_sort * b_sort,
                  "gbs": round(n_sort * b_sort / (stage * 1e-3) / 1e9, 2) if stage > 0 else None,
                  "includes": "digit passes, the later passes' digit counts (k_digit_count) and the "
                              "digit-count scans (k_scan_u32)"}
    sort_stage["hbm_frac"] = round(sort_stage["gbs"] / HBM_PEAK_GBS, 4) if sort_stage["gbs"] else None
    valu = measured_valu(profile_name) if profile_name else {}
    valu_f32 = measured_valu_f32(profile_name) if profile_name else {}
    for k, v in kernels.items():
        v["hbm_frac"] = round(v["gbs"] / HBM_PEAK_GBS, 4) if v["gbs"] else None
        if k in valu and k in valu_f32 and v["ms_per_launch"] > 0:
            # VALU issue floor (f32 at 4 cycles, the rest at 2) over the measured time
            floor_ms = (4 * valu_f32[k] + 2 * (valu[k] - valu_f32[k])) / SIMD_CYCLES_PER_S * 1e3
            v["valu_frac"] = round(floor_ms / v["ms_per_launch"], 4)
    dominant = max(kms, key=kms.get)
    d = kernels[dominant]
    traffic = measured_traffic(profile_name).get(dominant) if profile_name else None
    hbm_frac = d["gbs"] / HBM_PEAK_GBS
    roofline = {"bound": "hbm", "achieved": d["gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(hbm_frac, 4), "traffic": traffic, "kernel": dominant,
                "ms_per_launch": d["ms_per_launch"], "bytes_per_launch": d["bytes_per_launch"]}
    if profile_name:
        # the streaming-copy rate measured on this device (mg_stream_copy)
        roofline["copy_ceiling_gbs"] = profile_value(f"hbm-traffic_{profile_name}", "copy_ceiling_gbs")
    if dominant in valu:
        # vector-ALU issue: the kernel's wave-level VALU instructions (SQ
        # counters of the committed profile) per second of its measured time,
        # and the shortest time that instruction stream can issue in (f32
        # instructions at 4 cycles, all others at 2); frac = that floor over
        # the measured time
        rate = valu[dominant] / (d["ms_per_launch"] * 1e-3)
        f32 = measured_valu_f32(profile_name).get(dominant)
        roofline["valu"] = {"achieved": rate, "unit": "wave64 VALU instructions/s",
                            "cadence_4_cycles": VALU_PEAK_PER_S, "instructions_per_launch": valu[dominant],
                            "source": f"profiles/sq-instructions_{profile_name}.md"}
        if f32 is not None:
            floor_ms = (4 * f32 + 2 * (valu[dominant] - f32)) / SIMD_CYCLES_PER_S * 1e3
            roofline["valu"].update({"f32_per_launch": f32, "issue_floor_ms": round(floor_ms, 4),
                                     "frac": round(floor_ms / d["ms_per_launch"], 4)})
            if floor_ms / d["ms_per_launch"] > hbm_frac:
                roofline["bound"] = "valu"
    out = {
        "metric": METRIC,
        "value": segs * fps,
        "unit": "pixel-segments/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded numpy PCG64 scene generator)",
        "config": {"workload": workload, "width": W, "height": H, "paths": args.paths if args.workload == "polygons" else 10000,
                   "lines": int(st.lines), "layers": scene.n_layers},
        "frames_per_s": fps,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "device_ms_per_frame": dev_ms / K,
        "pixel_segments": int(segs),
        "pixel_segments_painted": int(st.segments_painted),
        "line_pixels": int(st.line_pixels),
        "sort_passes": int(st.sort_passes),
        "radix_bits": int(st.radix_bits),
        "tiles": int(st.tiles),
        "tile_runs": int(st.runs),
        "tile_records": int(st.records),
        "walk_threads": int(st.walk_threads),
        "walk_lds": ["share", "half", "whole", "global"][int(st.walk_lds)],
        "setup_hist_ms": round(st.setup_ms + st.hist_ms, 4),
        # dispatches of one frame (kernels, clears and status reads), from
        # the committed kernel trace of this workload
        "dispatches_per_frame": (profile_value(f"kernel-trace_{profile_name}", "dispatches_per_frame")
                                 if profile_name else None),
        "kernels": kernels,
        "sort_stage": sort_stage,
        "roofline": roofline,
        "scene_build_s": round(t_build, 2),
    }
    # PCIe-inclusive rate: the frame rendered into a caller-owned host buffer
    # (upload of the caller's pixels, render, download), as mg_render does
    hb = np.zeros((W * H, 4), np.uint8)
    host_ms = []
    for _ in range(3):
        a = time.perf_counter()
        comp.render(cassia_amd.Buffer(hb, W), scene.background, None)
        host_ms.append((time.perf_counter() - a) * 1e3)
    out["host_buffer_ms_per_frame"] = round(float(np.median(host_ms)), 3)
    del comp
    out["circles"], circles_scene = bench_circles(cassia_amd, workloads, args.steps, args.warmup)
    if not args.no_batch:
        out["batched"] = bench_batched(cassia_amd, workloads, args.steps, args.warmup, args.batch_scenes)
    if not args.no_anim:
        out["animation"] = bench_animation(cassia_amd, workloads)
    if not args.no_flatten:
        out["flatten"] = bench_flatten(cassia_amd)
    if not args.no_serve:
        out["serving"] = bench_serving(cassia_amd, workloads)
    if not args.no_cpu and gpu_frame is not None:
        cb = cpu_baseline(circles_scene, scene, gpu_frame)
        out["bit_identical"] = cb["large_frame"]["bit_identical_to_gpu"]
        out["cpu_baseline"] = cb
    else:
        out["bit_identical"] = None
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
