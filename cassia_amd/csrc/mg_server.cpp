// The serving loop: render requests (a composition and a frame) admitted by
// a device-memory budget and rendered as batched frames by worker threads,
// one per engine (its own HIP stream and scratch), so that a batch admitted
// while another is on the GPU runs beside it. Continuous batching: a worker
// whose engine is free takes every queued request that fits the budget left,
// up to max_batch, as its next batch. A heavy request (estimated above
// kHeavyLinePixels) is a batch of its own and ends the batch before it, so
// light requests do not wait inside a heavy batch; the next free engine
// takes the light requests queued behind it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mold_gpu.h"
#include "mg_internal.h"
#include "mg_runtime.h"

namespace {

using Clock = std::chrono::steady_clock;

// Estimated line-pixels one batch may take: half of a frame's limit (2^29),
// so that estimates below the actual count still fit.
constexpr uint64_t kBatchLinePixels = 1ull << 28;
// A request above this many estimated line-pixels renders alone (a 1,024 x
// 1,024 scene of 170,000 polygon lines is about 7 M; a 64 x 64 scene of a
// million lines of large polygons, 43 M and 9 ms).
constexpr uint64_t kHeavyLinePixels = 1ull << 24;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    uint64_t ticket = 0;
    mg_composition *comp = nullptr;
    mg_batch_frame frame{};
    uint8_t *out = nullptr;
    size_t len = 0;
    uint64_t footprint = 0;
    uint64_t line_pixels = 0;  // estimate
    Clock::time_point t_submit, t_start, t_done;
    mg_status status = MG_OK;
    std::string error;
    uint32_t batch_size = 0;
    uint32_t engine = 0;
    bool done = false;
};

}  // namespace

struct mg_server {
    mg_server_config cfg{};
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
    uint64_t next_ticket = 1;
    uint64_t inflight_bytes = 0;
    uint32_t busy_engines = 0;
    bool stop = false;
    std::vector<std::shared_ptr<mg::Renderer>> engines;
    std::vector<mg_batch *> batches;
    std::vector<std::thread> workers;
    // statistics since the last reset
    std::vector<double> latencies;
    uint64_t n_batches = 0, n_done = 0, n_failed = 0, batched_requests = 0;
    uint32_t max_concurrent = 0;

    bool admissible() const {
        if (pending.empty()) return false;
        const uint64_t fp = pending.front()->footprint;
        return inflight_bytes + fp <= cfg.hbm_budget;
    }

    void worker(uint32_t k) {
        std::vector<std::shared_ptr<Request>> take;
        std::vector<mg_composition *> comps;
        std::vector<mg_batch_frame> frames;
        for (;;) {
            uint64_t bytes = 0;
            {
                std::unique_lock<std::mutex> lk(m);
                cv_work.wait(lk, [&] { return (stop && pending.empty()) || admissible(); });
                if (stop && pending.empty()) return;
                if (!admissible()) continue;
                take.clear();
                uint64_t lp = 0;
                while (!pending.empty() && take.size() < cfg.max_batch) {
                    const auto &r = pending.front();
                    if (inflight_bytes + bytes + r->footprint > cfg.hbm_budget) break;
                    if (!take.empty() && lp + r->line_pixels > kBatchLinePixels) break;
                    const bool heavy = r->line_pixels > kHeavyLinePixels;
                    if (!take.empty() && heavy) break;
                    bytes += r->footprint;
                    lp += r->line_pixels;
                    take.push_back(r);
                    pending.pop_front();
                    if (heavy) break;
                }
                inflight_bytes += bytes;
                ++busy_engines;
                max_concurrent = std::max(max_concurrent, busy_engines);
            }
            // the rest of the queue may fit beside this batch on another engine
            cv_work.notify_one();
            const Clock::time_point t_start = Clock::now();
            comps.clear();
            frames.clear();
            for (const auto &r : take) {
                comps.push_back(r->comp);
                frames.push_back(r->frame);
            }
            std::vector<mg_status> st(take.size(), MG_OK);
            std::vector<std::string> errs(take.size());
            mg_status s = mgi::batch_render(batches[k], comps.data(), frames.data(), take.size(), false, nullptr,
                                            nullptr);
            if (!s) {
                for (size_t i = 0; i < take.size(); ++i) {
                    if (!take[i]->out) continue;
                    st[i] = mgi::batch_download(batches[k], i, take[i]->out, take[i]->len);
                    if (st[i]) errs[i] = mgi::last_error();
                }
            } else {
                // a request the batch could not hold fails alone: render each by itself
                for (size_t i = 0; i < take.size(); ++i) {
                    st[i] = take.size() == 1 ? s
                                             : mgi::batch_render(batches[k], &comps[i], &frames[i], 1, false, nullptr,
                                                                 nullptr);
                    if (!st[i] && take[i]->out) st[i] = mgi::batch_download(batches[k], 0, take[i]->out, take[i]->len);
                    if (st[i]) errs[i] = mgi::last_error();
                }
            }
            const Clock::time_point t_done = Clock::now();
            {
                std::lock_guard<std::mutex> lk(m);
                inflight_bytes -= bytes;
                --busy_engines;
                ++n_batches;
                batched_requests += take.size();
                for (size_t i = 0; i < take.size(); ++i) {
                    Request &r = *take[i];
                    r.t_start = t_start;
                    r.t_done = t_done;
                    r.status = st[i];
                    r.error = errs[i];
                    r.batch_size = (uint32_t)take.size();
                    r.engine = k;
                    r.done = true;
                    if (st[i]) {
                        ++n_failed;
                    } else {
                        ++n_done;
                        latencies.push_back(ms_between(r.t_submit, t_done));
                    }
                }
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }

    ~mg_server() {
        {
            std::lock_guard<std::mutex> lk(m);
            stop = true;
        }
        cv_work.notify_all();
        for (std::thread &t : workers) t.join();
        for (mg_batch *b : batches) mg_batch_destroy(b);
    }
};

static double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

extern "C" {

MG_EXPORT mg_status mg_server_create(const mg_server_config *cfg, mg_server **out) {
    if (!out) return mgi::set_error(MG_ERR_INVALID_ARG, "null argument");
    mg_server_config c = cfg ? *cfg : mg_server_config{};
    if (c.engines == 0) c.engines = 2;
    if (c.engines > 4) return mgi::set_error(MG_ERR_INVALID_ARG, "at most 4 engines");
    if (c.max_batch == 0 || c.max_batch > 8192) c.max_batch = 8192;
    auto s = std::make_unique<mg_server>();
    for (uint32_t k = 0; k < c.engines; ++k) {
        auto r = std::make_shared<mg::Renderer>();
        std::string err;
        if (!r->init(err)) return mgi::set_error(MG_ERR_DEVICE, err);
        if (k == 0 && c.hbm_budget == 0) {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return mgi::set_error(MG_ERR_DEVICE, "hipMemGetInfo");
            c.hbm_budget = free_b / 2;
        }
        s->engines.push_back(r);
        s->batches.push_back(mgi::batch_new(r));
    }
    s->cfg = c;
    mg_server *p = s.release();
    for (uint32_t k = 0; k < c.engines; ++k) p->workers.emplace_back([p, k] { p->worker(k); });
    *out = p;
    return MG_OK;
}

MG_EXPORT void mg_server_destroy(mg_server *s) { delete s; }

MG_EXPORT mg_status mg_server_config_of(const mg_server *s, mg_server_config *out) {
    if (!s || !out) return mgi::set_error(MG_ERR_INVALID_ARG, "null argument");
    *out = s->cfg;
    return MG_OK;
}

MG_EXPORT mg_status mg_server_submit(mg_server *s, mg_composition *comp, const mg_batch_frame *frame,
                                     uint8_t *rgba_out, size_t len, uint64_t *ticket) {
    if (!s || !comp || !frame || !ticket) return mgi::set_error(MG_ERR_INVALID_ARG, "null argument");
    if (frame->width == 0 || frame->height == 0) return mgi::set_error(MG_ERR_INVALID_ARG, "empty frame");
    if (rgba_out && len < (size_t)frame->width * frame->height * 4)
        return mgi::set_error(MG_ERR_BUFFER, "buffer shorter than the frame");
    auto r = std::make_shared<Request>();
    r->t_submit = Clock::now();
    if (mg_status ps = mgi::prepare(comp)) return ps;
    r->comp = comp;
    r->frame = *frame;
    r->out = rgba_out;
    r->len = len;
    r->footprint = mgi::footprint(comp, *frame, &r->line_pixels);
    if (r->footprint > s->cfg.hbm_budget)
        return mgi::set_error(MG_ERR_CAPACITY, "request exceeds the server's device-memory budget");
    {
        std::lock_guard<std::mutex> lk(s->m);
        if (s->stop) return mgi::set_error(MG_ERR_INVALID_ARG, "server is stopping");
        r->ticket = s->next_ticket++;
        s->live[r->ticket] = r;
        s->pending.push_back(r);
    }
    s->cv_work.notify_one();
    *ticket = r->ticket;
    return MG_OK;
}

MG_EXPORT mg_status mg_server_wait(mg_server *s, uint64_t ticket, mg_request_result *out) {
    if (!s) return mgi::set_error(MG_ERR_INVALID_ARG, "null argument");
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->m);
        auto it = s->live.find(ticket);
        if (it == s->live.end()) return mgi::set_error(MG_ERR_INVALID_ARG, "unknown ticket");
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->done; });
        s->live.erase(ticket);
    }
    if (out) {
        out->ticket = r->ticket;
        out->status = r->status;
        out->latency_ms = ms_between(r->t_submit, r->t_done);
        out->queue_ms = ms_between(r->t_submit, r->t_start);
        out->batch_size = r->batch_size;
        out->engine = r->engine;
        out->footprint = r->footprint;
    }
    if (r->status) return mgi::set_error(r->status, r->error);
    return MG_OK;
}

MG_EXPORT mg_status mg_server_stats(mg_server *s, int reset, mg_server_counters *out) {
    if (!s || !out) return mgi::set_error(MG_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->m);
    *out = mg_server_counters{};
    out->completed = s->n_done;
    out->failed = s->n_failed;
    out->batches = s->n_batches;
    out->mean_batch = s->n_batches ? (double)s->batched_requests / (double)s->n_batches : 0.0;
    out->p50_ms = percentile(s->latencies, 0.50);
    out->p99_ms = percentile(s->latencies, 0.99);
    out->max_ms = s->latencies.empty() ? 0.0 : *std::max_element(s->latencies.begin(), s->latencies.end());
    out->max_concurrent_batches = s->max_concurrent;
    out->queued = s->pending.size();
    out->inflight_bytes = s->inflight_bytes;
    if (reset) {
        s->latencies.clear();
        s->n_done = s->n_failed = s->n_batches = s->batched_requests = 0;
        s->max_concurrent = s->busy_engines;
    }
    return MG_OK;
}

}  // extern "C"
