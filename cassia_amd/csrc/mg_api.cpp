// mold's scene API over the MI355X renderer, exported as the C ABI of
// include/mold_gpu.h. Bookkeeping follows mold/src/{composition,layer,buffer}.rs;
// every pixel operation runs on the GPU (mg_runtime.hip, mg_kernels.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <queue>
#include <string>
#include <vector>

#include "../../include/mold_gpu.h"
#include "mg_internal.h"
#include "mg_path.h"
#include "mg_runtime.h"
#include "mg_scratch.h"
#include "mg_shared.h"
#include "mg_exact.h"

namespace {

thread_local std::string g_error;

// A solid Soft Light entry with finite colour channels carries the source
// terms of its channels (soft_light_terms: u0 u1 u2 v0 v1 v2 in the six
// words from grad_start on, which a solid fill does not use) and
// PF_SOFT_TERMS; the painter then reads them instead of computing them per
// record. Finite channels give finite terms, so no NaN is made here.
void set_soft_light_terms(mg::PropsEntry &e) {
    static_assert(offsetof(mg::PropsEntry, n_stops) + 4 == offsetof(mg::PropsEntry, grad_start) + 24,
                  "six words from grad_start");
    if ((e.flags & (mg::PF_GRADIENT | mg::PF_CLIP)) || ((e.flags >> mg::PF_BLEND_SHIFT) & 15u) != mg::BM_SOFT_LIGHT)
        return;
    float t[6];
    for (int i = 0; i < 3; ++i) {
        if (!std::isfinite(e.color[i])) return;
        mg::soft_light_terms(e.color[i], t[i], t[3 + i]);
    }
    std::memcpy(reinterpret_cast<char *>(&e) + offsetof(mg::PropsEntry, grad_start), t, sizeof t);
    e.flags |= mg::PF_SOFT_TERMS;
}

mg_status fail(mg_status s, const std::string &msg) {
    g_error = msg;
    return s;
}

const float kIdentity[6] = {1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f};

// Owner ids of frame tables: the device engine keeps the tables of the last
// frame it ran and reuses them when owner and version both match.
uint64_t next_owner_id() {
    static std::atomic<uint64_t> n{1};
    return n.fetch_add(1);
}

// surpass::painter::Props with an owned stop list.
struct Props {
    mg_props p{};
    std::vector<float> stops;

    Props() {
        // Props::default(): NonZero, Draw(Style::default()) with
        // Fill::Solid([0, 0, 0, 1]) and BlendMode::Over (style.rs:183-187)
        p.color[3] = 1.0f;
    }
    bool operator==(const Props &o) const {
        // derived PartialEq over f32 fields (-0.0 == 0.0, NaN != NaN)
        const mg_props &a = p, &b = o.p;
        if (a.fill_rule != b.fill_rule || a.func != b.func) return false;
        if (a.func == MG_FUNC_CLIP) return a.clip_layers == b.clip_layers;
        if (a.is_clipped != b.is_clipped || a.style_fill_rule != b.style_fill_rule ||
            a.blend_mode != b.blend_mode || a.fill_kind != b.fill_kind)
            return false;
        if (a.fill_kind == MG_FILL_SOLID) {
            for (int i = 0; i < 4; ++i)
                if (!(a.color[i] == b.color[i])) return false;
            return true;
        }
        if (a.gradient_type != b.gradient_type) return false;
        for (int i = 0; i < 2; ++i)
            if (!(a.gradient_start[i] == b.gradient_start[i]) || !(a.gradient_end[i] == b.gradient_end[i]))
                return false;
        if (stops.size() != o.stops.size()) return false;
        for (size_t i = 0; i < stops.size(); ++i)
            if (!(stops[i] == o.stops[i])) return false;
        return true;
    }
};

mg_status make_props(const mg_props *in, Props &out) {
    if (!in) return fail(MG_ERR_INVALID_ARG, "props is null");
    if (in->fill_rule > 1 || in->func > 1 || in->blend_mode > 11 || in->fill_kind > 1 || in->gradient_type > 1)
        return fail(MG_ERR_INVALID_ARG, "props enum out of range");
    Props p;
    p.p = *in;
    p.p.stops = nullptr;
    if (in->func == MG_FUNC_DRAW && in->fill_kind == MG_FILL_GRADIENT) {
        if (in->n_stops < 2 || !in->stops)
            return fail(MG_ERR_INVALID_ARG, "a gradient needs at least two stops (style.rs:66-69)");
        p.stops.assign(in->stops, in->stops + 5 * (size_t)in->n_stops);
        for (uint32_t i = 0; i < in->n_stops; ++i) {
            float s = p.stops[5 * i + 4];
            if (s < 0.0f || 1.0f < s) return fail(MG_ERR_GRADIENT, "gradient stops must be between 0.0 and 1.0");
        }
        // GradientBuilder::build: NO_STOP (-0.0, and +0.0 which compares
        // equal) becomes i / (n - 1) (style.rs:71-76)
        float inc = 1.0f / (float)(in->n_stops - 1);
        for (uint32_t i = 0; i < in->n_stops; ++i)
            if (p.stops[5 * i + 4] == -0.0f) p.stops[5 * i + 4] = (float)i * inc;
    }
    out = std::move(p);
    return MG_OK;
}

// mold::Layer (mold/src/layer.rs:144-244)
struct Layer {
    bool enabled = true;
    bool has_affine = false;
    float affine[6] = {0, 0, 0, 0, 0, 0};
    bool has_order = false;
    uint16_t order = 0;
    Props props;
    uint32_t is_unchanged = 0;  // SmallBitSet of cache ids
    size_t len = 0;
};

}  // namespace

struct mg_path {
    mg::Path path;
};

struct mg_composition {
    std::vector<float> pts;      // 4 floats per line (LinesBuilder SoA, interleaved)
    std::vector<uint16_t> ids;
    size_t uploaded = 0;         // lines already resident on the device
    // lines appended on the device (flattened there) after the `uploaded`
    // host lines, not yet copied into pts / ids (see ensure_mirror)
    size_t dev_extra = 0;
    std::map<uint16_t, Layer> layers;
    std::priority_queue<uint32_t, std::vector<uint32_t>, std::greater<uint32_t>> free_ids;
    uint32_t next_id = 0;
    std::map<uint16_t, uint16_t> orders_to_layers;
    std::map<std::pair<uintptr_t, size_t>, std::pair<size_t, size_t>> layouts;
    uint32_t caches = 0;
    std::shared_ptr<mg::Renderer> rdr;  // the device's shared frame engine
    mg::LineStore dlines;               // resident lines
    mg::LayerCaches dcaches;            // BufferLayerCache arrays
    mg::DevBuf frame;                   // device frame of the last render
    const uint64_t owner = next_owner_id();
    std::vector<uint64_t> pixsegs;
    mg_frame_stats stats{};
    uint64_t version = 1;      // bumped by every change of layers, props or lines
    uint32_t radix_bits = 0;
    uint32_t walk_threads = 0;
    // the row walk's need on the last render (largest runs of a tile, queue
    // + runs), which picks its LDS configuration on the next one
    uint32_t walk_need_runs = 0, walk_need_ids = 0;
    // frame tables (layers by id, props by order) of the last render, valid
    // while `version` and the layer cache id are unchanged
    mg::FrameSpec spec;
    uint64_t spec_version = 0;
    int32_t spec_cache_id = -2;
    bool spec_clips = false;
    bool spec_opaque = true;  // some layer of the tables has PF_OPAQUE

    size_t n_lines() const { return ids.size() + dev_extra; }

    uint32_t scene_flags() const { return (spec_clips ? mg::SF_CLIPS : 0) | (spec_opaque ? 0 : mg::SF_NO_OPAQUE); }

    // Copies device-appended lines into the host arrays, for host work on
    // them (inserting host lines after them, compaction, estimates).
    mg_status ensure_mirror() {
        if (!dev_extra) return MG_OK;
        const size_t n0 = ids.size();
        pts.resize(4 * (n0 + dev_extra));
        ids.resize(n0 + dev_extra);
        if (hipMemcpy(&pts[4 * n0], dlines.points() + n0, dev_extra * 16, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(&ids[n0], dlines.layer_ids() + n0, dev_extra * 2, hipMemcpyDeviceToHost) != hipSuccess) {
            pts.resize(4 * n0);
            ids.resize(n0);
            return fail(MG_ERR_DEVICE, "download of device lines failed");
        }
        uploaded = ids.size();
        dev_extra = 0;
        return MG_OK;
    }

    size_t actual_len() const {
        size_t n = 0;
        for (auto &kv : layers)
            if (kv.second.enabled) n += kv.second.len;
        return n;
    }

    // Composition::remove_disabled (composition.rs:123-142)
    void remove_disabled() {
        if (n_lines() >= actual_len() * 2 && n_lines() > 0) {
            if (ensure_mirror()) return;  // compaction stays for a later call
            size_t w = 0;
            for (size_t i = 0; i < ids.size(); ++i) {
                auto it = layers.find(ids[i]);
                if (it == layers.end() || !it->second.enabled) continue;
                if (w != i) {
                    std::memcpy(&pts[4 * w], &pts[4 * i], 16);
                    ids[w] = ids[i];
                }
                ++w;
            }
            if (w != ids.size()) {
                ++version;
                pts.resize(4 * w);
                ids.resize(w);
                uploaded = 0;
            }
            for (auto it = layers.begin(); it != layers.end();) {
                if (!it->second.enabled) {
                    free_ids.push(it->first);
                    it = layers.erase(it);
                } else {
                    ++it;
                }
            }
        }
    }

    ~mg_composition() {
        if (!rdr) return;
        std::lock_guard<std::mutex> g(rdr->mu);
        (void)hipSetDevice(rdr->device());
        (void)hipStreamSynchronize(rdr->stream());
        dlines.release();
        dcaches.release_all();
        frame.release();
    }

    mg_status ensure_renderer() {
        if (rdr) return MG_OK;
        std::string err;
        rdr = mg::Renderer::shared(err);
        if (!rdr) return fail(MG_ERR_DEVICE, err);
        return MG_OK;
    }

    // Called with the engine locked.
    mg_status sync_lines() {
        std::string err;
        if (uploaded > ids.size()) uploaded = 0;
        if (uploaded < ids.size()) {
            if (uploaded == 0) dlines.ranked = 0;  // the lines are uploaded anew: rank them anew
            if (hipSetDevice(rdr->device()) != hipSuccess ||
                !dlines.upload(&pts[4 * uploaded], &ids[uploaded], uploaded, ids.size() - uploaded, rdr->stream(), err))
                return fail(MG_ERR_OOM, err.empty() ? "hipSetDevice failed" : err);
            uploaded = ids.size();
        }
        return MG_OK;
    }

    // Ranks of the resident lines among the lines of their layers (k_setup
    // places line i of a layer at the layer's base + its rank), computed on
    // the device for the lines not ranked yet: once per change of the lines.
    // Called with the engine locked, after sync_lines.
    mg_status ensure_ranks() {
        std::string err;
        const size_t n = n_lines();
        const uint32_t n_ids = layers.empty() ? 1u : (uint32_t)layers.rbegin()->first + 1u;
        bool reset = dlines.ranked > n;
        if (n_ids > dlines.count_ids) {
            if (!dlines.counts.reserve((size_t)n_ids * 4)) return fail(MG_ERR_OOM, "out of device memory (line ranks)");
            dlines.count_ids = n_ids;
            reset = true;
        }
        if (dlines.rank.bytes < n * 4) {
            if (!dlines.rank.reserve(std::max(n, dlines.cap) * 4)) return fail(MG_ERR_OOM, "out of device memory (line ranks)");
            reset = true;
        }
        if (reset) dlines.ranked = 0;
        if (dlines.ranked == n) return MG_OK;
        if (!rdr->rank_lines(dlines.layer_ids(), dlines.counts.as<uint32_t>(), dlines.count_ids,
                             dlines.rank.as<uint32_t>(), dlines.ranked, n, dlines.ranked == 0, err))
            return fail(MG_ERR_DEVICE, err);
        dlines.ranked = n;
        return MG_OK;
    }

    // Device frame of `bytes`, owned by the composition.
    uint8_t *frame_image(size_t bytes) {
        if (hipSetDevice(rdr->device()) != hipSuccess || !frame.reserve(bytes)) return nullptr;
        return frame.as<uint8_t>();
    }

    void insert(uint16_t id, const float *lines, size_t n) {
        ensure_mirror();  // host lines go after the device-appended ones
        ++version;
        pts.insert(pts.end(), lines, lines + 4 * n);
        ids.insert(ids.end(), n, id);
        Layer &l = layers[id];
        // insert_segments (composition.rs:72-78)
        l.has_order = true;
        l.order = id;
        l.enabled = true;
        l.len += n;
    }

    // Layer table (by id) and props table (by order) of the frame.
    void tables(int32_t cache_id, mg::FrameSpec &f, bool &has_clips) {
        uint32_t n_layers = layers.empty() ? 1 : (uint32_t)layers.rbegin()->first + 1;
        f.layers.assign(n_layers, mg::LayerEntry{});
        for (auto &kv : layers) {
            mg::LayerEntry &e = f.layers[kv.first];
            const Layer &l = kv.second;
            e.flags = mg::LF_EXISTS | (l.enabled ? mg::LF_ENABLED : 0) | (l.has_affine ? mg::LF_AFFINE : 0) |
                      (l.has_order ? mg::LF_ORDER : 0);
            e.order = l.order;
            std::memcpy(e.affine, l.affine, sizeof e.affine);
        }
        uint32_t n_props = orders_to_layers.empty() ? 1 : (uint32_t)orders_to_layers.rbegin()->first + 1;
        f.props.assign(n_props, mg::PropsEntry{});
        f.stops.clear();
        has_clips = false;
        spec_opaque = false;
        uint32_t max_order = 0;
        for (auto &kv : orders_to_layers) {
            auto it = layers.find(kv.second);
            if (it == layers.end()) continue;
            const Layer &l = it->second;
            const mg_props &p = l.props.p;
            mg::PropsEntry &e = f.props[kv.first];
            uint32_t fl = mg::PF_DEFINED;
            if (p.fill_rule == MG_EVENODD) fl |= mg::PF_EVENODD;
            if (p.func == MG_FUNC_CLIP) {
                fl |= mg::PF_CLIP;
                e.clip_layers = p.clip_layers;
                has_clips = true;
            } else {
                if (p.is_clipped) {
                    fl |= mg::PF_CLIPPED;
                    has_clips = true;
                }
                fl |= (p.blend_mode & 15u) << mg::PF_BLEND_SHIFT;
                std::memcpy(e.color, p.color, sizeof e.color);
                if (p.fill_kind == MG_FILL_GRADIENT) {
                    fl |= mg::PF_GRADIENT;
                    if (p.gradient_type == MG_GRADIENT_RADIAL) fl |= mg::PF_RADIAL;
                    std::memcpy(e.grad_start, p.gradient_start, sizeof e.grad_start);
                    std::memcpy(e.grad_end, p.gradient_end, sizeof e.grad_end);
                    e.stops_offset = (uint32_t)f.stops.size();
                    e.n_stops = p.n_stops;
                    f.stops.insert(f.stops.end(), l.props.stops.begin(), l.props.stops.end());
                } else if (p.blend_mode == MG_BLEND_OVER && p.color[3] == 1.0f) {
                    fl |= mg::PF_OPAQUE;
                    spec_opaque = true;
                }
            }
            if (cache_id >= 0 && ((l.is_unchanged >> cache_id) & 1u)) fl |= mg::PF_UNCHANGED;
            e.flags = fl;
            set_soft_light_terms(e);
            if (l.enabled) max_order = std::max<uint32_t>(max_order, kv.first);
        }
        f.max_order = std::max<uint32_t>(max_order, n_props - 1);
        // where render frames place each layer's lines: enabled layers in
        // their render order (ties by id), then disabled layers (whose lines
        // set up with no pixels); a layer's lines start at its base
        std::vector<std::pair<uint32_t, uint16_t>> on, off;
        for (auto &kv : layers) {
            const Layer &l = kv.second;
            (l.enabled ? on : off).push_back({l.has_order ? (uint32_t)l.order : (uint32_t)kv.first, kv.first});
        }
        std::sort(on.begin(), on.end());
        uint64_t base = 0;
        for (const auto *v : {&on, &off})
            for (const auto &p : *v) {
                f.layers[p.second].base = (uint32_t)base;
                base += layers[p.second].len;
            }
        lines_placed = base;
    }
    size_t lines_placed = 0;  // lines the last tables placed (equals n_lines() while consistent)
};

static mg_status check_comp(const mg_composition *c) {
    return c ? MG_OK : fail(MG_ERR_INVALID_ARG, "composition is null");
}

extern "C" {

MG_EXPORT const char *mg_last_error(void) { return g_error.c_str(); }

MG_EXPORT int mg_device_count(void) { return mg::device_count(); }

// ------------------------------------------------------------------ paths

MG_EXPORT mg_status mg_path_create(mg_path **out) {
    if (!out) return fail(MG_ERR_INVALID_ARG, "out is null");
    *out = new (std::nothrow) mg_path();
    return *out ? MG_OK : fail(MG_ERR_OOM, "out of memory");
}

MG_EXPORT void mg_path_destroy(mg_path *p) { delete p; }

#define MG_PATH_CMD(fn, call, n)                                              \
    MG_EXPORT mg_status fn(mg_path *p, const float pts[n]) {                  \
        if (!p || !pts) return fail(MG_ERR_INVALID_ARG, "null argument");      \
        p->path.call(pts);                                                     \
        return MG_OK;                                                          \
    }
MG_PATH_CMD(mg_path_line, line, 4)
MG_PATH_CMD(mg_path_quad, quad, 6)
MG_PATH_CMD(mg_path_rat_quad, rat_quad, 9)
MG_PATH_CMD(mg_path_cubic, cubic, 8)
MG_PATH_CMD(mg_path_rat_cubic, rat_cubic, 12)

MG_EXPORT mg_status mg_path_close(mg_path *p) {
    if (!p) return fail(MG_ERR_INVALID_ARG, "null argument");
    p->path.close();
    return MG_OK;
}

MG_EXPORT mg_status mg_path_flatten(const mg_path *p, const float *t9, float *out, size_t cap, size_t *n) {
    if (!p || !n) return fail(MG_ERR_INVALID_ARG, "null argument");
    std::vector<float> segs;
    p->path.flatten(t9, segs);
    *n = segs.size() / 4;
    if (out) std::memcpy(out, segs.data(), std::min(*n, cap) * 16);
    return MG_OK;
}

// ------------------------------------------------------------ composition

MG_EXPORT mg_status mg_composition_create(mg_composition **out) {
    if (!out) return fail(MG_ERR_INVALID_ARG, "out is null");
    *out = new (std::nothrow) mg_composition();
    return *out ? MG_OK : fail(MG_ERR_OOM, "out of memory");
}

MG_EXPORT void mg_composition_destroy(mg_composition *c) { delete c; }

MG_EXPORT mg_status mg_create_layer(mg_composition *c, mg_layer_id *out) {
    if (check_comp(c) || !out) return fail(MG_ERR_INVALID_ARG, "null argument");
    if (!c->free_ids.empty()) {
        *out = (mg_layer_id)c->free_ids.top();
        c->free_ids.pop();
        return MG_OK;
    }
    if (c->next_id > 0xFFFFu) return fail(MG_ERR_IDS_EXHAUSTED, "all 65536 layer ids are in use");
    *out = (mg_layer_id)c->next_id++;
    return MG_OK;
}

MG_EXPORT mg_status mg_insert_lines(mg_composition *c, mg_layer_id id, const float *lines, size_t n) {
    if (check_comp(c) || (n && !lines)) return fail(MG_ERR_INVALID_ARG, "null argument");
    c->insert(id, lines, n);
    return MG_OK;
}

MG_EXPORT mg_status mg_insert_lines_multi(mg_composition *c, const mg_layer_id *ids, const float *lines, size_t n) {
    if (check_comp(c) || (n && (!lines || !ids))) return fail(MG_ERR_INVALID_ARG, "null argument");
    if (mg_status ms = c->ensure_mirror()) return ms;
    ++c->version;
    c->pts.insert(c->pts.end(), lines, lines + 4 * n);
    c->ids.insert(c->ids.end(), ids, ids + n);
    for (size_t i = 0; i < n; ++i) {
        Layer &l = c->layers[ids[i]];
        l.has_order = true;
        l.order = ids[i];
        l.enabled = true;
        l.len += 1;
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_insert_in_layer(mg_composition *c, mg_layer_id id, const mg_path *p) {
    if (check_comp(c) || !p) return fail(MG_ERR_INVALID_ARG, "null argument");
    std::vector<float> segs;
    p->path.flatten(nullptr, segs);
    c->insert(id, segs.data(), segs.size() / 4);
    return MG_OK;
}

MG_EXPORT mg_status mg_insert_in_layer_transformed(mg_composition *c, mg_layer_id id, const mg_path *p,
                                                   const float t9[9]) {
    if (check_comp(c) || !p || !t9) return fail(MG_ERR_INVALID_ARG, "null argument");
    std::vector<float> segs;
    p->path.flatten(t9, segs);
    c->insert(id, segs.data(), segs.size() / 4);
    return MG_OK;
}

MG_EXPORT mg_status mg_insert_paths_device(mg_composition *c, const mg_path *const *paths, const mg_layer_id *ids,
                                           const float *t9s, const uint8_t *transformed, size_t n) {
    if (check_comp(c) || (n && (!paths || !ids))) return fail(MG_ERR_INVALID_ARG, "null argument");
    for (size_t i = 0; i < n; ++i)
        if (!paths[i]) return fail(MG_ERR_INVALID_ARG, "path is null");
    if (!n) return MG_OK;
    mg_status s = c->ensure_renderer();
    if (s) return s;
    std::vector<mg::PathCommand> cmds;
    std::vector<mg::flat::PathDesc> descs(n);
    for (size_t i = 0; i < n; ++i) {
        const auto &pc = paths[i]->path.commands();
        mg::flat::PathDesc &d = descs[i];
        d = mg::flat::PathDesc{};
        d.first_cmd = (uint32_t)cmds.size();
        d.n_cmds = (uint32_t)pc.size();
        d.layer = ids[i];
        d.has_xf = t9s && (!transformed || transformed[i]) ? 1u : 0u;
        if (d.has_xf) std::memcpy(d.t9, t9s + 9 * i, sizeof d.t9);
        cmds.insert(cmds.end(), pc.begin(), pc.end());
    }
    if (cmds.size() > 0xFFFFFFFFull) return fail(MG_ERR_CAPACITY, "too many path commands");
    std::lock_guard<std::mutex> g(c->rdr->mu);
    // device lines follow every host line: those go up first
    if ((s = c->sync_lines())) return s;
    const size_t base = c->n_lines();
    std::vector<uint32_t> counts;
    std::string err;
    if (!c->rdr->flatten_paths(cmds.data(), cmds.size(), descs.data(), n, c->dlines, base, counts, err))
        return fail(err.find("memory") != std::string::npos ? MG_ERR_OOM : MG_ERR_DEVICE, err);
    size_t added = 0;
    for (size_t i = 0; i < n; ++i) {
        // insert_segments (composition.rs:72-78), as for a host insert
        Layer &l = c->layers[ids[i]];
        l.has_order = true;
        l.order = ids[i];
        l.enabled = true;
        l.len += counts[i];
        added += counts[i];
    }
    c->dev_extra += added;
    ++c->version;
    return MG_OK;
}

MG_EXPORT mg_status mg_insert_paths(mg_composition *c, const mg_path *const *paths, const mg_layer_id *ids,
                                    const float *t9s, const uint8_t *transformed, size_t n) {
    if (check_comp(c) || (n && (!paths || !ids))) return fail(MG_ERR_INVALID_ARG, "null argument");
    for (size_t i = 0; i < n; ++i)
        if (!paths[i]) return fail(MG_ERR_INVALID_ARG, "path is null");
    std::vector<float> segs;
    for (size_t i = 0; i < n; ++i) {
        segs.clear();
        const bool xf = t9s && (!transformed || transformed[i]);
        paths[i]->path.flatten(xf ? t9s + 9 * i : nullptr, segs);
        c->insert(ids[i], segs.data(), segs.size() / 4);
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_lines(mg_composition *c, float *pts, uint16_t *ids, size_t cap, size_t *n) {
    if (check_comp(c) || !n || (cap && (!pts || !ids))) return fail(MG_ERR_INVALID_ARG, "null argument");
    if (mg_status s = c->ensure_mirror()) return s;
    *n = c->ids.size();
    const size_t m = std::min(cap, c->ids.size());
    if (m) {
        std::memcpy(pts, c->pts.data(), m * 16);
        std::memcpy(ids, c->ids.data(), m * 2);
    }
    return MG_OK;
}

static Layer *find_layer(const mg_composition *c, mg_layer_id id) {
    auto it = const_cast<mg_composition *>(c)->layers.find(id);
    return it == c->layers.end() ? nullptr : &it->second;
}

#define MG_LAYER(c, id, l)                                                        \
    if (check_comp(c)) return MG_ERR_INVALID_ARG;                                 \
    Layer *l = find_layer(c, id);                                                 \
    if (!l) return fail(MG_ERR_NO_LAYER, "no layer with this id");                \
    ++const_cast<mg_composition *>(c)->version;

MG_EXPORT mg_status mg_layer_exists(const mg_composition *c, mg_layer_id id, int *out) {
    if (check_comp(c) || !out) return fail(MG_ERR_INVALID_ARG, "null argument");
    *out = find_layer(c, id) != nullptr;
    return MG_OK;
}

MG_EXPORT mg_status mg_layer_is_enabled(const mg_composition *c, mg_layer_id id, int *out) {
    MG_LAYER(c, id, l)
    if (!out) return fail(MG_ERR_INVALID_ARG, "null argument");
    *out = l->enabled;
    return MG_OK;
}

MG_EXPORT mg_status mg_layer_set_enabled(mg_composition *c, mg_layer_id id, int enabled) {
    MG_LAYER(c, id, l)
    l->enabled = enabled != 0;
    return MG_OK;
}

MG_EXPORT mg_status mg_layer_transform(const mg_composition *c, mg_layer_id id, float out[6]) {
    MG_LAYER(c, id, l)
    if (!out) return fail(MG_ERR_INVALID_ARG, "null argument");
    std::memcpy(out, l->has_affine ? l->affine : kIdentity, 6 * sizeof(float));
    return MG_OK;
}

MG_EXPORT mg_status mg_layer_set_transform(mg_composition *c, mg_layer_id id, const float t[6]) {
    MG_LAYER(c, id, l)
    if (!t) return fail(MG_ERR_INVALID_ARG, "null argument");
    bool identity = true;
    for (int i = 0; i < 6; ++i) identity = identity && t[i] == kIdentity[i];
    bool has = !identity;
    if (has) {
        // layer.rs:186-190
        if (t[0] * t[0] + t[2] * t[2] > 1.001f || t[1] * t[1] + t[3] * t[3] > 1.001f)
            return fail(MG_ERR_LAYER_SCALE, "Layer's scaling on each axis must be between -1.0 and 1.0");
    }
    bool same = has == l->has_affine;
    if (same && has)
        for (int i = 0; i < 6; ++i) same = same && t[i] == l->affine[i];
    if (!same) {
        l->is_unchanged = 0;
        l->has_affine = has;
        if (has) std::memcpy(l->affine, t, sizeof l->affine);
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_layer_order(const mg_composition *c, mg_layer_id id, uint16_t *out) {
    MG_LAYER(c, id, l)
    if (!out) return fail(MG_ERR_INVALID_ARG, "null argument");
    if (!l->has_order) return fail(MG_ERR_INVALID_ARG, "Layers should always have orders");
    *out = l->order;
    return MG_OK;
}

MG_EXPORT mg_status mg_layer_set_order(mg_composition *c, mg_layer_id id, uint16_t order) {
    MG_LAYER(c, id, l)
    if (!(l->has_order && l->order == order)) {
        l->is_unchanged = 0;
        l->has_order = true;
        l->order = order;
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_layer_props(const mg_composition *c, mg_layer_id id, mg_props *out) {
    MG_LAYER(c, id, l)
    if (!out) return fail(MG_ERR_INVALID_ARG, "null argument");
    *out = l->props.p;
    out->stops = l->props.stops.empty() ? nullptr : l->props.stops.data();
    return MG_OK;
}

MG_EXPORT mg_status mg_layer_set_props(mg_composition *c, mg_layer_id id, const mg_props *props) {
    MG_LAYER(c, id, l)
    Props p;
    mg_status s = make_props(props, p);
    if (s) return s;
    if (!(l->props == p)) {
        l->is_unchanged = 0;
        l->props = std::move(p);
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_create_layers(mg_composition *c, size_t n, mg_layer_id *out) {
    if (check_comp(c) || (n && !out)) return fail(MG_ERR_INVALID_ARG, "null argument");
    for (size_t i = 0; i < n; ++i) {
        mg_status s = mg_create_layer(c, &out[i]);
        if (s) return s;
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_layers_set_transform(mg_composition *c, const mg_layer_id *ids, const float *t6, size_t n) {
    if (check_comp(c) || (n && (!ids || !t6))) return fail(MG_ERR_INVALID_ARG, "null argument");
    for (size_t i = 0; i < n; ++i) {
        mg_status s = mg_layer_set_transform(c, ids[i], &t6[6 * i]);
        if (s) return s;
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_layers_set_order(mg_composition *c, const mg_layer_id *ids, const uint16_t *orders, size_t n) {
    if (check_comp(c) || (n && (!ids || !orders))) return fail(MG_ERR_INVALID_ARG, "null argument");
    for (size_t i = 0; i < n; ++i) {
        mg_status s = mg_layer_set_order(c, ids[i], orders[i]);
        if (s) return s;
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_layers_set_props(mg_composition *c, const mg_layer_id *ids, const mg_props *props, size_t n) {
    if (check_comp(c) || (n && (!ids || !props))) return fail(MG_ERR_INVALID_ARG, "null argument");
    for (size_t i = 0; i < n; ++i) {
        mg_status s = mg_layer_set_props(c, ids[i], &props[i]);
        if (s) return s;
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_remove_disabled(mg_composition *c) {
    if (check_comp(c)) return MG_ERR_INVALID_ARG;
    c->remove_disabled();
    return MG_OK;
}

MG_EXPORT mg_status mg_lines_len(const mg_composition *c, size_t *builder_len, size_t *actual_len) {
    if (check_comp(c)) return MG_ERR_INVALID_ARG;
    if (builder_len) *builder_len = c->n_lines();
    if (actual_len) *actual_len = c->actual_len();
    return MG_OK;
}

MG_EXPORT mg_status mg_create_buffer_layer_cache(mg_composition *c, int32_t *cache_id) {
    if (check_comp(c) || !cache_id) return fail(MG_ERR_INVALID_ARG, "null argument");
    // SmallBitSet::first_empty_slot (layer.rs:137-141)
    uint32_t slot = 0;
    while (slot < 32 && ((c->caches >> slot) & 1u)) ++slot;
    if (slot >= 32) return fail(MG_ERR_NO_CACHE, "all buffer layer caches are in use");
    c->caches |= 1u << slot;
    *cache_id = (int32_t)slot;
    return MG_OK;
}

MG_EXPORT mg_status mg_buffer_layer_cache_clear(mg_composition *c, int32_t id) {
    if (check_comp(c) || id < 0 || id >= 32 || !((c->caches >> id) & 1u))
        return fail(MG_ERR_INVALID_ARG, "unknown cache");
    if (c->rdr) {
        std::lock_guard<std::mutex> g(c->rdr->mu);
        c->dcaches.clear(id, c->rdr->stream());
        if (hipGetLastError() != hipSuccess) return fail(MG_ERR_DEVICE, "cache clear failed");
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_buffer_layer_cache_release(mg_composition *c, int32_t id) {
    if (check_comp(c) || id < 0 || id >= 32) return fail(MG_ERR_INVALID_ARG, "unknown cache");
    c->caches &= ~(1u << id);
    if (c->rdr) {
        std::lock_guard<std::mutex> g(c->rdr->mu);
        (void)hipStreamSynchronize(c->rdr->stream());
        c->dcaches.release(id);
    }
    return MG_OK;
}

// Frame tables of a composition (remove_disabled, orders_to_layers and the
// layer / props tables, composition.rs:229-290), rebuilt only when the
// composition changed: every change of lines, layers, props or cache state
// bumps `version`.
static void prepare_tables(mg_composition *c, int32_t cache_id) {
    if (c->spec_version == c->version && c->spec_cache_id == cache_id) return;
    c->remove_disabled();
    for (auto &kv : c->layers)
        if (kv.second.enabled && kv.second.has_order) c->orders_to_layers[kv.second.order] = kv.first;
    c->spec = mg::FrameSpec{};
    c->tables(cache_id, c->spec, c->spec_clips);
    c->spec_version = c->version;
    c->spec_cache_id = cache_id;
}

static void fill_stats(mg_frame_stats &st, const mg::FrameResult &r, size_t lines) {
    st.lines = lines;
    st.line_pixels = r.line_pixels;
    st.segments = r.segments;
    st.segments_painted = r.kept;
    st.tiles = r.tiles;
    st.sort_passes = r.passes;
    st.radix_bits = r.radix_bits;
    st.device_ms = r.device_ms;
    st.setup_ms = r.stage_ms[0];
    st.emit_ms = r.stage_ms[1];
    st.hist_ms = r.stage_ms[2];
    st.sort_ms = r.stage_ms[3];
    st.runs_ms = r.stage_ms[4];
    st.walk_ms = r.stage_ms[5];
    st.paint_ms = r.stage_ms[6];
    st.runs = r.runs;
    st.records = r.records;
    st.walk_threads = r.walk_threads;
    st.deep_tiles = r.deep_tiles;
    st.walk_lds = r.walk_lds;
    st.walk_rewalks = r.walk_rewalks;
    for (uint32_t p = 0; p < 16; ++p) st.sort_pass_ms[p] = p < r.passes ? r.pass_ms[p] : 0.0;
}

// Scene descriptor of one frame of `width` x `rows_px` pixels at `stride`
// (Rect::new for the crop, buffer_layout.rs:43-50).
static mg::SceneDesc scene_desc(size_t width, size_t stride, size_t rows_px, const float bg[4], const mg_rect *crop) {
    const size_t row_len = (width + 7) / 8, tile_rows = (rows_px + 7) / 8;
    mg::SceneDesc sd{};
    sd.width = (uint32_t)width;
    sd.rows_px = (uint32_t)rows_px;
    sd.stride = (uint32_t)stride;
    sd.row_len = (uint32_t)row_len;
    sd.tile_rows = (uint32_t)tile_rows;
    sd.crop_h0 = 0;
    sd.crop_h1 = (uint32_t)row_len;
    sd.crop_v0 = 0;
    sd.crop_v1 = (uint32_t)tile_rows;
    if (crop) {
        sd.crop_h0 = crop->x0 / 8;
        sd.crop_h1 = (uint32_t)std::min<uint64_t>(((uint64_t)crop->x1 + 7) / 8, row_len);
        sd.crop_v0 = crop->y0 / 8;
        sd.crop_v1 = (uint32_t)std::min<uint64_t>(((uint64_t)crop->y1 + 7) / 8, tile_rows);
        if (sd.crop_h0 > sd.crop_h1) sd.crop_h0 = sd.crop_h1;
        if (sd.crop_v0 > sd.crop_v1) sd.crop_v0 = sd.crop_v1;
    }
    std::memcpy(sd.clear, bg, sizeof sd.clear);
    return sd;
}

// Composition::render for one frame of `width` x `rows_px` pixels at
// `stride` into the device image `img` (composition.rs:212-328).
static mg_status render_frame(mg_composition *c, size_t width, size_t stride, size_t rows_px, const float bg[4],
                              const mg_rect *crop, int32_t cache_id, int32_t *prev, uint8_t *img, size_t img_bytes,
                              bool want_tile_ops, mg::FrameResult &r, uint32_t coverage_cap = 0,
                              bool *graph_replayed = nullptr) {
    std::string err;
    prepare_tables(c, cache_id);
    mg::FrameSpec &f = c->spec;
    f.scenes.clear();
    f.scene_line_base.clear();
    f.tables_owner = c->owner;
    f.tables_version = c->version * 64 + (uint64_t)(cache_id + 1);
    mg::SceneDesc sd = scene_desc(width, stride, rows_px, bg, crop);
    sd.line_base = 0;
    sd.n_lines = (uint32_t)c->n_lines();
    sd.layer_base = 0;
    sd.n_layer_entries = (uint32_t)f.layers.size();
    sd.props_base = 0;
    sd.n_props = (uint32_t)f.props.size();
    sd.row_base = 0;
    sd.tile_base = 0;
    sd.flags = c->scene_flags() | (cache_id >= 0 ? mg::SF_CACHE : 0);
    sd.out_offset = 0;
    f.scenes.push_back(sd);
    f.scene_line_base.push_back(0);
    f.mode = 0;
    f.image_bytes = img_bytes;
    f.want_tile_ops = want_tile_ops;
    f.want_coverage = coverage_cap > 0;
    f.coverage_cap = coverage_cap;
    f.radix_bits = c->radix_bits;
    f.walk_threads = c->walk_threads;
    f.walk_need_runs = c->walk_need_runs;
    f.walk_need_ids = c->walk_need_ids;
    mg_status s;
    if ((s = c->sync_lines()) || (s = c->ensure_ranks())) return s;
    if (c->lines_placed != c->n_lines()) return fail(MG_ERR_DEVICE, "layer line counts disagree with the lines");
    const uint32_t *ranks = c->dlines.rank.as<uint32_t>();
    const bool ok = graph_replayed
                        ? c->rdr->run_graph(f, c->dlines.points(), c->dlines.layer_ids(), ranks, c->n_lines(), img, r,
                                            *graph_replayed, err)
                        : c->rdr->run(f, c->dlines.points(), c->dlines.layer_ids(), ranks, c->n_lines(), img, prev, r,
                                      err);
    if (!ok) return fail(err.find("capacity") != std::string::npos ? MG_ERR_CAPACITY : MG_ERR_DEVICE, err);
    c->walk_need_runs = r.walk_max_runs;
    c->walk_need_ids = r.walk_max_ids;
    if (cache_id >= 0) {
        bool changed = false;
        for (auto &kv : c->layers) {
            uint32_t before = kv.second.is_unchanged;
            if (kv.second.enabled) kv.second.is_unchanged |= 1u << cache_id;
            else kv.second.is_unchanged &= ~(1u << cache_id);
            changed = changed || before != kv.second.is_unchanged;
        }
        if (changed) ++c->version;
    }
    fill_stats(c->stats, r, c->n_lines());
    return MG_OK;
}

MG_EXPORT mg_status mg_render(mg_composition *c, const mg_buffer *buf, const float bg[4], const mg_rect *crop) {
    if (check_comp(c) || !buf || !bg || !buf->rgba) return fail(MG_ERR_INVALID_ARG, "null argument");
    mg_status s = c->ensure_renderer();
    if (s) return s;
    std::lock_guard<std::mutex> g(c->rdr->mu);
    std::string err;
    const int32_t cache_id = buf->layer_cache;
    if (cache_id >= 0 && (cache_id >= 32 || !((c->caches >> cache_id) & 1u)))
        return fail(MG_ERR_INVALID_ARG, "unknown layer cache");
    // the layout cache is keyed by (pointer, length) (composition.rs:238-241)
    auto key = std::make_pair((uintptr_t)buf->rgba, buf->len_px);
    auto lit = c->layouts.find(key);
    if (lit == c->layouts.end())
        lit = c->layouts.emplace(key, std::make_pair(buf->width, buf->width_stride ? buf->width_stride : buf->width)).first;
    const size_t width = lit->second.first, stride = lit->second.second;
    const size_t len = buf->len_px;
    // BufferLayoutBuilder::build checks (buffer_layout.rs:81-92, 103)
    if (width > len || stride > len || width > stride || width == 0)
        return fail(MG_ERR_BUFFER, "width / width_stride do not fit the buffer");
    const size_t rows_px = (len + stride - 1) / stride;
    if (len - (rows_px - 1) * stride < width) return fail(MG_ERR_BUFFER, "last buffer row is shorter than width");
    const size_t row_len = (width + 7) / 8, tile_rows = (rows_px + 7) / 8;
    int32_t *prev = nullptr;
    if (cache_id >= 0) {
        // Buffer::tiles_len (buffer.rs:37-44) and the resize/clear of render
        const size_t bw = buf->width_stride ? buf->width_stride : buf->width;
        const size_t tiles_len = ((buf->width + 7) / 8) * (((len / bw) + 7) / 8);
        if (tiles_len != row_len * tile_rows) return fail(MG_ERR_BUFFER, "layer cache size mismatch");
        bool resized = false;
        prev = c->dcaches.array(cache_id, tiles_len, c->rdr->stream(), resized, err);
        if (!prev) return fail(MG_ERR_OOM, err);
    }
    const size_t bytes = len * 4;
    uint8_t *img = c->frame_image(bytes);
    if (!img) return fail(MG_ERR_OOM, "out of device memory (frame)");
    // The caller's pixels go first where the frame leaves some of them as
    // they are: tiles the layer cache skips, tiles outside the crop, and
    // pixels beyond `width` in each row (stride padding, a longer last row).
    // Without those, the painter writes every pixel and nothing is uploaded.
    const mg::SceneDesc cover = scene_desc(width, stride, rows_px, bg, crop);
    const bool writes_all = cache_id < 0 && stride == width && len == rows_px * width && cover.crop_h0 == 0 &&
                            cover.crop_v0 == 0 && cover.crop_h1 == row_len && cover.crop_v1 == tile_rows;
    if (!writes_all &&
        hipMemcpyAsync(img, buf->rgba, bytes, hipMemcpyHostToDevice, c->rdr->stream()) != hipSuccess)
        return fail(MG_ERR_DEVICE, "upload of the frame buffer failed");
    mg::FrameResult r;
    if ((s = render_frame(c, width, stride, rows_px, bg, crop, cache_id, prev, img, bytes, buf->flusher != nullptr, r)))
        return s;
    if (hipMemcpyAsync(buf->rgba, img, bytes, hipMemcpyDeviceToHost, c->rdr->stream()) != hipSuccess ||
        hipStreamSynchronize(c->rdr->stream()) != hipSuccess)
        return fail(MG_ERR_DEVICE, "download of the frame failed");
    if (buf->flusher) {
        std::vector<int8_t> ops(row_len * tile_rows);
        if (!c->rdr->download_tile_ops(ops.data(), ops.size(), err)) return fail(MG_ERR_DEVICE, err);
        for (size_t t = 0; t < ops.size(); ++t) {
            if (ops[t] != 1 && ops[t] != 2) continue;
            size_t j = t / row_len, i = t % row_len;
            for (size_t y = 8 * j; y < std::min(8 * j + 8, rows_px); ++y) {
                size_t x0 = 8 * i, x1 = std::min(8 * i + 8, width);
                buf->flusher(buf->flusher_user, buf->rgba + 4 * (y * stride + x0), x1 - x0);
            }
        }
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_render_device(mg_composition *c, uint32_t width, uint32_t height, const float bg[4],
                                     const mg_rect *crop, void **device_rgba) {
    if (check_comp(c) || !bg || !device_rgba || width == 0 || height == 0)
        return fail(MG_ERR_INVALID_ARG, "bad argument");
    mg_status s = c->ensure_renderer();
    if (s) return s;
    std::lock_guard<std::mutex> g(c->rdr->mu);
    std::string err;
    const size_t bytes = (size_t)width * height * 4;
    uint8_t *img = c->frame_image(bytes);
    if (!img) return fail(MG_ERR_OOM, "out of device memory (frame)");
    mg::FrameResult r;
    if ((s = render_frame(c, width, width, height, bg, crop, -1, nullptr, img, bytes, false, r))) return s;
    *device_rgba = img;
    return MG_OK;
}

// mg_render_device with no device allocation: the frame engine of the
// callers that cannot allocate takes every buffer from the static scratch
// array of mg_scratch.hip (kNoAllocScratchBytes per device), the output is
// the caller's device buffer, and the composition's lines are read in place
// when they are resident and current, else copied into the scratch.
MG_EXPORT mg_status mg_render_device_noalloc(mg_composition *c, uint32_t width, uint32_t height, const float bg[4],
                                             const mg_rect *crop, void *device_rgba, size_t device_bytes) {
    if (check_comp(c) || !bg || !device_rgba || width == 0 || height == 0)
        return fail(MG_ERR_INVALID_ARG, "bad argument");
    const size_t bytes = (size_t)width * height * 4;
    if (device_bytes < bytes) return fail(MG_ERR_BUFFER, "device frame shorter than width * height * 4 bytes");
    std::string err;
    const int dev = mg::render_device();
    if (dev < 0) return fail(MG_ERR_DEVICE, "no gfx950 device");
    void *base = nullptr;
    if (!mg::scratch_for_device(dev, &base, err)) return fail(MG_ERR_DEVICE, err);
    std::shared_ptr<mg::Renderer> rdr = mg::Renderer::shared_arena(base, mg::kNoAllocScratchBytes, err);
    if (!rdr) return fail(MG_ERR_DEVICE, err);
    std::lock_guard<std::mutex> g(rdr->mu);
    if (hipSetDevice(rdr->device()) != hipSuccess) return fail(MG_ERR_DEVICE, "hipSetDevice failed");
    prepare_tables(c, -1);  // remove_disabled may compact the host lines first
    const size_t n = c->n_lines();
    rdr->arena_begin();
    if (c->lines_placed != n) return fail(MG_ERR_DEVICE, "layer line counts disagree with the lines");
    const float4 *pts = nullptr;
    const uint16_t *ids = nullptr;
    uint32_t *ranks = nullptr;
    const bool resident = n > 0 && c->rdr && c->rdr->device() == rdr->device() && c->uploaded == c->ids.size() &&
                          c->dlines.cap >= n;
    if (resident) {
        pts = c->dlines.points();
        ids = c->dlines.layer_ids();
    } else if (n) {
        mg_status s = c->ensure_mirror();
        if (s) return s;
        void *p = rdr->arena_take(n * 16), *q = rdr->arena_take(n * 2);
        if (!p || !q)
            return fail(MG_ERR_CAPACITY, "the composition's lines do not fit the no-allocation scratch (" +
                                             std::to_string(mg::kNoAllocScratchBytes >> 20) + " MiB)");
        if (hipMemcpyAsync(p, c->pts.data(), n * 16, hipMemcpyHostToDevice, rdr->stream()) != hipSuccess ||
            hipMemcpyAsync(q, c->ids.data(), n * 2, hipMemcpyHostToDevice, rdr->stream()) != hipSuccess)
            return fail(MG_ERR_DEVICE, "upload of the lines failed");
        pts = static_cast<const float4 *>(p);
        ids = static_cast<const uint16_t *>(q);
    }
    if (n) {
        // the lines' ranks among their layers' lines, computed in the scratch
        const uint32_t n_ids = c->layers.empty() ? 1u : (uint32_t)c->layers.rbegin()->first + 1u;
        uint32_t *counts = static_cast<uint32_t *>(rdr->arena_take((size_t)n_ids * 4));
        ranks = static_cast<uint32_t *>(rdr->arena_take(n * 4));
        if (!counts || !ranks)
            return fail(MG_ERR_CAPACITY, "the composition's lines do not fit the no-allocation scratch (" +
                                             std::to_string(mg::kNoAllocScratchBytes >> 20) + " MiB)");
        if (!rdr->rank_lines(ids, counts, n_ids, ranks, 0, n, true, err)) return fail(MG_ERR_DEVICE, err);
    }
    mg::FrameSpec &f = c->spec;
    f.scenes.clear();
    f.scene_line_base.clear();
    f.tables_owner = c->owner;
    f.tables_version = c->version * 64;
    mg::SceneDesc sd = scene_desc(width, width, height, bg, crop);
    sd.n_lines = (uint32_t)n;
    sd.n_layer_entries = (uint32_t)f.layers.size();
    sd.n_props = (uint32_t)f.props.size();
    sd.flags = c->scene_flags();
    f.scenes.push_back(sd);
    f.scene_line_base.push_back(0);
    f.mode = 0;
    f.image_bytes = bytes;
    f.want_tile_ops = false;
    f.want_coverage = false;
    f.coverage_cap = 0;
    f.radix_bits = c->radix_bits;
    f.walk_threads = c->walk_threads;
    f.walk_need_runs = c->walk_need_runs;
    f.walk_need_ids = c->walk_need_ids;
    mg::FrameResult r;
    if (!rdr->run(f, pts, ids, ranks, n, static_cast<uint8_t *>(device_rgba), nullptr, r, err)) {
        if (rdr->arena().short_by)
            return fail(MG_ERR_CAPACITY, "the frame needs more device scratch than the no-allocation entry point's " +
                                             std::to_string(mg::kNoAllocScratchBytes >> 20) + " MiB (at least " +
                                             std::to_string(rdr->arena().short_by) + " more bytes)");
        return fail(err.find("capacity") != std::string::npos ? MG_ERR_CAPACITY : MG_ERR_DEVICE, err);
    }
    c->walk_need_runs = r.walk_max_runs;
    c->walk_need_ids = r.walk_max_ids;
    fill_stats(c->stats, r, n);
    return MG_OK;
}

MG_EXPORT mg_status mg_render_animated(mg_composition *c, uint32_t width, uint32_t height, const float bg[4],
                                       const mg_rect *crop, void **device_rgba, int *replayed) {
    if (check_comp(c) || !bg || !device_rgba || width == 0 || height == 0)
        return fail(MG_ERR_INVALID_ARG, "bad argument");
    mg_status s = c->ensure_renderer();
    if (s) return s;
    std::lock_guard<std::mutex> g(c->rdr->mu);
    const size_t bytes = (size_t)width * height * 4;
    uint8_t *img = c->frame_image(bytes);
    if (!img) return fail(MG_ERR_OOM, "out of device memory (frame)");
    mg::FrameResult r;
    bool rep = false;
    if ((s = render_frame(c, width, width, height, bg, crop, -1, nullptr, img, bytes, false, r, 0, &rep))) return s;
    *device_rgba = img;
    if (replayed) *replayed = rep ? 1 : 0;
    return MG_OK;
}

MG_EXPORT mg_status mg_copy_to_host(void *dst, const void *device_src, size_t bytes) {
    if ((!dst || !device_src) && bytes) return fail(MG_ERR_INVALID_ARG, "null argument");
    if (bytes && hipMemcpy(dst, device_src, bytes, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(MG_ERR_DEVICE, "device-to-host copy failed");
    return MG_OK;
}

MG_EXPORT mg_status mg_stream_copy(size_t bytes, uint32_t repeats, double *ms_per_copy, uint64_t *mismatches) {
    if (!ms_per_copy || !mismatches || bytes == 0 || bytes % 16 || repeats == 0)
        return fail(MG_ERR_INVALID_ARG, "bad argument (bytes: a nonzero multiple of 16)");
    std::string err;
    auto r = mg::Renderer::shared(err);
    if (!r) return fail(MG_ERR_DEVICE, err);
    std::lock_guard<std::mutex> g(r->mu);
    if (!r->stream_copy(bytes, repeats, *ms_per_copy, *mismatches, err)) return fail(MG_ERR_DEVICE, err);
    return MG_OK;
}

MG_EXPORT mg_status mg_render_coverage(mg_composition *c, uint32_t width, uint32_t height, const float bg[4],
                                       const mg_rect *crop, int32_t *a_out, uint32_t *key_out, size_t cap,
                                       size_t *n) {
    if (check_comp(c) || !bg || !n || width == 0 || height == 0 || (cap && (!a_out || !key_out)))
        return fail(MG_ERR_INVALID_ARG, "bad argument");
    if (cap > 0xFFFFFFFFull) return fail(MG_ERR_INVALID_ARG, "cap too large");
    mg_status s = c->ensure_renderer();
    if (s) return s;
    std::lock_guard<std::mutex> g(c->rdr->mu);
    std::string err;
    const size_t bytes = (size_t)width * height * 4;
    uint8_t *img = c->frame_image(bytes);
    if (!img) return fail(MG_ERR_OOM, "out of device memory (frame)");
    mg::FrameResult r;
    if ((s = render_frame(c, width, width, height, bg, crop, -1, nullptr, img, bytes, false, r,
                          (uint32_t)std::max<size_t>(cap, 1))))
        return s;
    std::vector<int32_t> a;
    std::vector<uint32_t> key;
    uint32_t count = 0;
    if (!c->rdr->download_coverage(a, key, count, err)) return fail(MG_ERR_DEVICE, err);
    const size_t m = std::min<size_t>({(size_t)count, cap, key.size() / 4});
    if (m) {
        std::memcpy(a_out, a.data(), m * 64 * sizeof(int32_t));
        std::memcpy(key_out, key.data(), m * 4 * sizeof(uint32_t));
    }
    *n = count;
    return MG_OK;
}

MG_EXPORT mg_status mg_pixel_segments(mg_composition *c, const uint64_t **out, size_t *n) {
    if (check_comp(c) || !out || !n) return fail(MG_ERR_INVALID_ARG, "null argument");
    mg_status s = c->ensure_renderer();
    if (s) return s;
    std::lock_guard<std::mutex> g(c->rdr->mu);
    c->remove_disabled();
    mg::FrameSpec f;
    bool has_clips = false;
    c->tables(-1, f, has_clips);
    mg::SceneDesc sd{};
    sd.n_lines = (uint32_t)c->n_lines();
    sd.n_layer_entries = (uint32_t)f.layers.size();
    sd.n_props = (uint32_t)f.props.size();
    f.scenes.push_back(sd);
    f.scene_line_base.push_back(0);
    f.mode = 1;
    if (c->radix_bits) f.radix_bits = c->radix_bits;
    if ((s = c->sync_lines())) return s;
    std::string err;
    mg::FrameResult r;
    if (!c->rdr->run(f, c->dlines.points(), c->dlines.layer_ids(), nullptr, c->n_lines(), nullptr, nullptr, r, err))
        return fail(MG_ERR_DEVICE, err);
    c->pixsegs.assign(2 * r.line_pixels, mg::kNone);
    if (!c->rdr->download_sorted(c->pixsegs.data(), r.kept, err)) return fail(MG_ERR_DEVICE, err);
    *out = c->pixsegs.data();
    *n = c->pixsegs.size();
    return MG_OK;
}

MG_EXPORT mg_status mg_set_tuning(mg_composition *c, uint32_t radix_bits) {
    if (check_comp(c)) return MG_ERR_INVALID_ARG;
    if (radix_bits != 0 && radix_bits != 4 && radix_bits != 6 && radix_bits != 8 && radix_bits != 9)
        return fail(MG_ERR_INVALID_ARG, "radix_bits must be 0, 4, 6, 8 or 9");
    c->radix_bits = radix_bits;
    return MG_OK;
}

MG_EXPORT mg_status mg_set_walk_threads(mg_composition *c, uint32_t threads) {
    if (check_comp(c)) return MG_ERR_INVALID_ARG;
    if (threads != 0 && threads != 64 && threads != 256 && threads != 512 && threads != 1024)
        return fail(MG_ERR_INVALID_ARG, "walk threads must be 0, 64, 256, 512 or 1024");
    c->walk_threads = threads;
    return MG_OK;
}

MG_EXPORT mg_status mg_last_frame_stats(const mg_composition *c, mg_frame_stats *out) {
    if (check_comp(c) || !out) return fail(MG_ERR_INVALID_ARG, "null argument");
    *out = c->stats;
    return MG_OK;
}

// ------------------------------------------------------------------ batches

}  // extern "C"

struct mg_batch {
    std::shared_ptr<mg::Renderer> rdr;
    mg::LineStore dlines;  // lines of all scenes, scene after scene
    mg::DevBuf frame;      // all frames, 256-byte aligned
    const uint64_t owner = next_owner_id();
    // what the scene descriptors and tables were packed from
    std::vector<uint64_t> owners, versions;
    std::vector<mg_batch_frame> packed_frames;
    uint64_t tables_version = 0;
    mg::FrameSpec spec;
    size_t n_lines = 0;
    std::vector<uint64_t> offsets, bytes;
    uint32_t radix_bits = 0;    // 0 = autotune (or the runtime's rule without the autotuner)
    uint32_t walk_threads = 0;  // 0 = autotune (or the runtime's rule without the autotuner)
    bool autotune = true;       // false for serving batches, whose shapes change every batch
    // autotuner: the choice holds for batches of the same shape
    uint32_t tuned_bits = 0, tuned_walk = 0, tuned_for = ~0u;
    uint32_t walk_need_runs = 0, walk_need_ids = 0;  // as mg_composition's
    mg_batch_tuning tuning{};
    mg_frame_stats stats{};

    ~mg_batch() {
        if (!rdr) return;
        std::lock_guard<std::mutex> g(rdr->mu);
        (void)hipSetDevice(rdr->device());
        (void)hipStreamSynchronize(rdr->stream());
        frame.release();
    }
};

namespace {

bool same_frame(const mg_batch_frame &a, const mg_batch_frame &b) {
    return std::memcmp(&a, &b, sizeof a) == 0;
}

// Packs the compositions into one frame: scene descriptors with their line,
// layer, props, tile-row and output bases, concatenated tables and (when a
// composition changed) the concatenated resident lines.
// Tables and resident lines of a composition for batched frames. The
// composition's lines live in its own device buffer (uploaded as they are
// inserted); a batch reads them in place through each scene's line address.
// `held`: the engine whose mutex the caller holds.
mg_status prepare_for_batch(mg_composition *c, const mg::Renderer *held) {
    mg_status s = c->ensure_renderer();
    if (s) return s;
    prepare_tables(c, -1);
    if (c->lines_placed != c->n_lines()) return fail(MG_ERR_DEVICE, "layer line counts disagree with the lines");
    if (c->rdr.get() == held) {
        if ((s = c->sync_lines())) return s;
        return c->ensure_ranks();  // on the batch's own stream, before its frame
    }
    std::lock_guard<std::mutex> g(c->rdr->mu);
    if ((s = c->sync_lines()) || (s = c->ensure_ranks())) return s;
    // the ranks are read by another engine's stream
    if (hipStreamSynchronize(c->rdr->stream()) != hipSuccess) return fail(MG_ERR_DEVICE, "line ranks failed");
    return MG_OK;
}

// Packs the compositions into one frame: scene descriptors with their line,
// layer, props, tile-row and output bases and line addresses, and the
// concatenated tables.
mg_status pack_batch(mg_batch *b, mg_composition *const *comps, const mg_batch_frame *frames, size_t n,
                     bool prepare) {
    bool same = b->owners.size() == n;
    for (size_t i = 0; i < n; ++i) {
        if (prepare) {
            mg_status s = prepare_for_batch(comps[i], b->rdr.get());
            if (s) return s;
        }
        same = same && b->owners[i] == comps[i]->owner && b->versions[i] == comps[i]->version;
    }
    bool same_frames = same && b->packed_frames.size() == n;
    for (size_t i = 0; same_frames && i < n; ++i) same_frames = same_frame(b->packed_frames[i], frames[i]);
    if (same_frames) return MG_OK;
    size_t total = 0;
    for (size_t i = 0; i < n; ++i) total += comps[i]->n_lines();
    if (total > 0xFFFFFFFFull) return fail(MG_ERR_CAPACITY, "a batch holds at most 2^32 lines");
    b->n_lines = total;
    b->owners.resize(n);
    b->versions.resize(n);
    for (size_t i = 0; i < n; ++i) {
        b->owners[i] = comps[i]->owner;
        b->versions[i] = comps[i]->version;
    }
    mg::FrameSpec &f = b->spec;
    f = mg::FrameSpec{};
    f.scenes.resize(n);
    f.scene_line_base.resize(n);
    b->offsets.resize(n);
    b->bytes.resize(n);
    uint64_t line_base = 0, row_base = 0, tile_base = 0, out = 0;
    for (size_t i = 0; i < n; ++i) {
        const mg_composition *c = comps[i];
        const mg::FrameSpec &cs = c->spec;
        const mg_batch_frame &fr = frames[i];
        mg::SceneDesc sd = scene_desc(fr.width, fr.width, fr.height, fr.background, fr.has_crop ? &fr.crop : nullptr);
        sd.line_base = (uint32_t)line_base;
        sd.n_lines = (uint32_t)c->n_lines();
        sd.layer_base = (uint32_t)f.layers.size();
        sd.n_layer_entries = (uint32_t)cs.layers.size();
        sd.props_base = (uint32_t)f.props.size();
        sd.n_props = (uint32_t)cs.props.size();
        sd.row_base = (uint32_t)row_base;
        sd.tile_base = (uint32_t)tile_base;
        sd.lines_pts = (uint64_t)(uintptr_t)c->dlines.points();
        sd.lines_ids = (uint64_t)(uintptr_t)c->dlines.layer_ids();
        sd.lines_rank = (uint64_t)(uintptr_t)c->dlines.rank.as<uint32_t>();
        sd.flags = c->scene_flags();
        sd.out_offset = out;
        const uint32_t stops_base = (uint32_t)f.stops.size();
        for (mg::LayerEntry e : cs.layers) {
            e.base += (uint32_t)line_base;  // the scene's lines follow the earlier scenes'
            f.layers.push_back(e);
        }
        for (mg::PropsEntry e : cs.props) {
            if (e.flags & mg::PF_GRADIENT) e.stops_offset += stops_base;
            f.props.push_back(e);
        }
        f.stops.insert(f.stops.end(), cs.stops.begin(), cs.stops.end());
        f.max_order = std::max(f.max_order, cs.max_order);
        f.scenes[i] = sd;
        f.scene_line_base[i] = (uint32_t)line_base;
        b->offsets[i] = out;
        b->bytes[i] = (uint64_t)fr.width * fr.height * 4;
        line_base += c->n_lines();
        row_base += sd.tile_rows;
        tile_base += (uint64_t)sd.tile_rows * sd.row_len;
        out += (b->bytes[i] + 255) & ~255ull;
    }
    if (row_base > 0xFFFFFFFFull || tile_base > 0xFFFFFFFFull || f.layers.size() > 0xFFFFFFFFull ||
        f.props.size() > 0xFFFFFFFFull)
        return fail(MG_ERR_CAPACITY, "batch too large");
    if (!b->frame.reserve(std::max<uint64_t>(out, 256))) return fail(MG_ERR_OOM, "out of device memory (frames)");
    f.mode = 0;
    f.image_bytes = out;
    f.tables_owner = b->owner;
    f.tables_version = ++b->tables_version;
    b->packed_frames.assign(frames, frames + n);
    return MG_OK;
}

}  // namespace

extern "C" {

MG_EXPORT mg_status mg_batch_create(mg_batch **out) {
    if (!out) return fail(MG_ERR_INVALID_ARG, "null argument");
    std::string err;
    auto r = mg::Renderer::shared(err);
    if (!r) return fail(MG_ERR_DEVICE, err);
    mg_batch *b = new mg_batch();
    b->rdr = std::move(r);
    *out = b;
    return MG_OK;
}

MG_EXPORT void mg_batch_destroy(mg_batch *b) { delete b; }

MG_EXPORT mg_status mg_batch_set_tuning(mg_batch *b, uint32_t radix_bits) {
    if (!b) return fail(MG_ERR_INVALID_ARG, "batch is null");
    if (radix_bits != 0 && radix_bits != 4 && radix_bits != 6 && radix_bits != 8 && radix_bits != 9)
        return fail(MG_ERR_INVALID_ARG, "radix_bits must be 0, 4, 6, 8 or 9");
    b->radix_bits = radix_bits;
    b->tuned_for = ~0u;
    return MG_OK;
}

MG_EXPORT mg_status mg_batch_set_walk_threads(mg_batch *b, uint32_t threads) {
    if (!b) return fail(MG_ERR_INVALID_ARG, "batch is null");
    if (threads != 0 && threads != 64 && threads != 256 && threads != 512 && threads != 1024)
        return fail(MG_ERR_INVALID_ARG, "walk threads must be 0, 64, 256, 512 or 1024");
    b->walk_threads = threads;
    b->tuned_for = ~0u;
    return MG_OK;
}

MG_EXPORT mg_status mg_batch_render(mg_batch *b, mg_composition *const *comps, const mg_batch_frame *frames,
                                    size_t n, void **device_rgba, const uint64_t **offsets) {
    if (!b || (n && (!comps || !frames)) || !device_rgba || !offsets) return fail(MG_ERR_INVALID_ARG, "null argument");
    if (n == 0 || n > 8192) return fail(MG_ERR_INVALID_ARG, "a batch holds 1 to 8192 frames");
    for (size_t i = 0; i < n; ++i) {
        if (!comps[i]) return fail(MG_ERR_INVALID_ARG, "composition is null");
        if (frames[i].width == 0 || frames[i].height == 0) return fail(MG_ERR_INVALID_ARG, "empty frame");
    }
    return mgi::batch_render(b, comps, frames, n, true, device_rgba, offsets);
}

}  // extern "C"

namespace mgi {

mg_status set_error(mg_status s, const std::string &msg) { return fail(s, msg); }
const char *last_error() { return g_error.c_str(); }

mg_batch *batch_new(std::shared_ptr<mg::Renderer> engine) {
    mg_batch *b = new mg_batch();
    b->rdr = std::move(engine);
    b->autotune = false;
    return b;
}

mg_status prepare(mg_composition *c) { return prepare_for_batch(c, nullptr); }

uint64_t footprint(const mg_composition *c, const mg_batch_frame &fr, uint64_t *line_pixels) {
    // 18 B per resident line, 40 + 40 B of line setup (and its ordered
    // copy), two 8-byte sort buffers of 2 slots per line-pixel, the frame and
    // per-tile paint state; line-pixels from each line's major extent
    double lp = 0.0;
    if (const_cast<mg_composition *>(c)->ensure_mirror()) return ~0ull;
    const size_t L = c->ids.size();
    for (size_t i = 0; i < L; ++i) {
        const float *q = &c->pts[4 * i];
        lp += std::max(std::fabs(q[2] - q[0]), std::fabs(q[3] - q[1])) + 2.0;
    }
    if (line_pixels) *line_pixels = (uint64_t)lp;
    const uint64_t tiles = (uint64_t)((fr.width + 7) / 8) * ((fr.height + 7) / 8);
    return 98ull * L + 32ull * (uint64_t)lp + 4ull * fr.width * fr.height + 1024ull * tiles + (1ull << 20);
}

mg_status batch_render(mg_batch *b, mg_composition *const *comps, const mg_batch_frame *frames, size_t n,
                       bool prepare, void **device_rgba, const uint64_t **offsets) {
    std::lock_guard<std::mutex> g(b->rdr->mu);
    mg_status s = pack_batch(b, comps, frames, n, prepare);
    if (s) return s;
    mg::FrameSpec &f = b->spec;
    std::string err;
    mg::FrameResult r;
    auto run = [&](uint32_t bits, uint32_t walk) -> mg_status {
        f.radix_bits = bits;
        f.walk_threads = walk;
        f.walk_need_runs = b->walk_need_runs;
        f.walk_need_ids = b->walk_need_ids;
        if (!b->rdr->run(f, nullptr, nullptr, nullptr, b->n_lines, b->frame.as<uint8_t>(), nullptr, r, err))
            return fail(err.find("capacity") != std::string::npos ? MG_ERR_CAPACITY : MG_ERR_DEVICE, err);
        b->walk_need_runs = r.walk_max_runs;
        b->walk_need_ids = r.walk_max_ids;
        return MG_OK;
    };
    uint32_t bits = b->radix_bits, walk = b->walk_threads;
    if (b->autotune && (!bits || !walk)) {
        // the autotuner: each free knob in turn (radix with the fixed or
        // automatic walk size, then walk sizes with the chosen radix), each
        // candidate timed twice on this batch (the first run also sizes
        // scratch buffers); the choice holds for later batches with the same
        // sort key width and power-of-two class of line counts
        uint32_t lg = 0;
        while (lg < 63 && (b->n_lines >> lg)) ++lg;
        const uint32_t key = mg::key_layout(f).sort_bits * 64 + lg;
        if (b->tuned_for != key) {
            auto time_of = [&](uint32_t rb, uint32_t wt, double *ms) -> mg_status {
                for (int rep = 0; rep < 2; ++rep) {
                    mg_status e = run(rb, wt);
                    if (e) return e;
                    *ms = rep == 0 ? r.device_ms : std::min(*ms, r.device_ms);
                }
                return MG_OK;
            };
            mg_batch_tuning t{};
            uint32_t best_bits = bits;
            if (!bits) {
                static const uint32_t cand[4] = {4, 6, 8, 9};
                t.n_candidates = 4;
                double best = 0.0;
                for (int k = 0; k < 4; ++k) {
                    double ms = 0.0;
                    if ((s = time_of(cand[k], walk, &ms))) return s;
                    t.radix_bits[k] = cand[k];
                    t.device_ms[k] = ms;
                    if (!best_bits || ms < best) {
                        best = ms;
                        best_bits = cand[k];
                    }
                }
            }
            uint32_t best_walk = walk;
            if (!walk) {
                static const uint32_t cand[3] = {64, 256, 1024};
                t.n_walk_candidates = 3;
                double best = 0.0;
                for (int k = 0; k < 3; ++k) {
                    double ms = 0.0;
                    if ((s = time_of(best_bits, cand[k], &ms))) return s;
                    t.walk_threads[k] = cand[k];
                    t.walk_ms[k] = ms;
                    if (!best_walk || ms < best) {
                        best = ms;
                        best_walk = cand[k];
                    }
                }
            }
            t.chosen_radix_bits = best_bits;
            t.chosen_walk_threads = best_walk;
            b->tuning = t;
            b->tuned_bits = best_bits;
            b->tuned_walk = best_walk;
            b->tuned_for = key;
        }
        bits = b->tuned_bits;
        walk = b->tuned_walk;
    }
    if ((s = run(bits, walk))) return s;
    fill_stats(b->stats, r, b->n_lines);
    if (device_rgba) *device_rgba = b->frame.as<uint8_t>();
    if (offsets) *offsets = b->offsets.data();
    return MG_OK;
}

mg_status batch_download(mg_batch *b, size_t i, uint8_t *rgba, size_t len) {
    if (i >= b->offsets.size()) return fail(MG_ERR_INVALID_ARG, "frame index out of range");
    if (len < b->bytes[i]) return fail(MG_ERR_BUFFER, "buffer shorter than the frame");
    std::lock_guard<std::mutex> g(b->rdr->mu);
    if (hipMemcpyAsync(rgba, b->frame.as<uint8_t>() + b->offsets[i], b->bytes[i], hipMemcpyDeviceToHost,
                       b->rdr->stream()) != hipSuccess ||
        hipStreamSynchronize(b->rdr->stream()) != hipSuccess)
        return fail(MG_ERR_DEVICE, "download of the frame failed");
    return MG_OK;
}

}  // namespace mgi

extern "C" {

MG_EXPORT mg_status mg_batch_read(mg_batch *b, size_t i, uint8_t *rgba, size_t len) {
    if (!b || !rgba) return fail(MG_ERR_INVALID_ARG, "null argument");
    return mgi::batch_download(b, i, rgba, len);
}

MG_EXPORT mg_status mg_batch_last_tuning(const mg_batch *b, mg_batch_tuning *out) {
    if (!b || !out) return fail(MG_ERR_INVALID_ARG, "null argument");
    *out = b->tuning;
    return MG_OK;
}

MG_EXPORT mg_status mg_batch_stats(const mg_batch *b, mg_frame_stats *out) {
    if (!b || !out) return fail(MG_ERR_INVALID_ARG, "null argument");
    *out = b->stats;
    return MG_OK;
}

}  // extern "C"

// --------------------------------------------------------------------------
// cassia-compatible entry points (cassia/src/Cassia.h:32-41, Cassia.cpp:240-255)
// --------------------------------------------------------------------------

namespace {

struct CassiaState {
    uint32_t width = 0, height = 0;
    std::shared_ptr<mg::Renderer> rdr;
    mg::DevBuf in, frame;
    uint8_t *image = nullptr;

    ~CassiaState() {
        if (!rdr) return;
        std::lock_guard<std::mutex> g(rdr->mu);
        (void)hipStreamSynchronize(rdr->stream());
        in.release();
        frame.release();
    }
};

CassiaState *g_cassia = nullptr;

}  // namespace

extern "C" {

MG_EXPORT void cassia_init(uint32_t width, uint32_t height) {
    delete g_cassia;
    g_cassia = new CassiaState();
    g_cassia->width = width;
    g_cassia->height = height;
    std::string err;
    auto r = mg::Renderer::shared(err);
    if (!r) {
        g_error = err;
        return;
    }
    std::lock_guard<std::mutex> g(r->mu);
    if (!g_cassia->frame.reserve((size_t)width * height * 4 + 4)) {
        g_error = "out of device memory (frame)";
        return;
    }
    g_cassia->image = g_cassia->frame.as<uint8_t>();
    (void)hipMemsetAsync(g_cassia->image, 0, (size_t)width * height * 4, r->stream());
    (void)hipStreamSynchronize(r->stream());
    g_cassia->rdr = std::move(r);
}

MG_EXPORT void cassia_render(const uint64_t *psegments, size_t n, const CassiaStyling *stylings, size_t m) {
    CassiaState *st = g_cassia;
    if (!st || !st->rdr) {
        g_error = "cassia_render before a successful cassia_init";
        return;
    }
    std::lock_guard<std::mutex> g(st->rdr->mu);
    std::string err;
    if (!st->in.reserve(std::max<size_t>(n, 1) * 8) ||
        (n && hipMemcpyAsync(st->in.ptr, psegments, n * 8, hipMemcpyHostToDevice, st->rdr->stream()) != hipSuccess)) {
        g_error = "upload of the segment stream failed";
        return;
    }
    mg::FrameSpec f;
    f.mode = 2;
    f.in_segs = st->in.as<uint64_t>();
    f.n_in = n;
    f.props.assign(std::max<size_t>(m, 1), mg::PropsEntry{});
    for (size_t i = 0; i < m; ++i) {
        const CassiaStyling &s = stylings[i];
        mg::PropsEntry &e = f.props[i];
        uint32_t blend = s.blendMode <= 11 ? s.blendMode : 0;
        e.flags = mg::PF_DEFINED | (s.fillRule == 1 ? mg::PF_EVENODD : 0) | (blend << mg::PF_BLEND_SHIFT);
        if (blend == MG_BLEND_OVER && s.fill[3] == 1.0f) e.flags |= mg::PF_OPAQUE;
        std::memcpy(e.color, s.fill, sizeof e.color);
        set_soft_light_terms(e);
    }
    f.layers.assign(1, mg::LayerEntry{});
    f.max_order = m ? (uint32_t)m - 1 : 0;
    mg::SceneDesc sd{};
    sd.n_props = (uint32_t)m;
    sd.width = st->width;
    sd.rows_px = st->height;
    sd.stride = st->width;
    sd.row_len = (st->width + 7) / 8;
    sd.tile_rows = (st->height + 7) / 8;
    sd.crop_h1 = sd.row_len;
    sd.crop_v1 = sd.tile_rows;
    // opaque white, the background of the reference demo's captures
    // (demo/src/circles.rs:130)
    for (int q = 0; q < 4; ++q) sd.clear[q] = 1.0f;
    f.scenes.push_back(sd);
    f.scene_line_base.push_back(0);
    f.image_bytes = (size_t)st->width * st->height * 4;
    mg::FrameResult r;
    if (!st->rdr->run(f, nullptr, nullptr, nullptr, 0, st->image, nullptr, r, err)) g_error = err;
}

MG_EXPORT void cassia_shutdown(void) {
    delete g_cassia;
    g_cassia = nullptr;
}

MG_EXPORT mg_status cassia_read_rgba8(uint8_t *out, size_t cap) {
    CassiaState *st = g_cassia;
    if (!st || !st->rdr || !out) return fail(MG_ERR_INVALID_ARG, "no cassia frame");
    size_t bytes = std::min<size_t>(cap, (size_t)st->width * st->height * 4);
    if (hipMemcpyAsync(out, st->image, bytes, hipMemcpyDeviceToHost, st->rdr->stream()) != hipSuccess ||
        hipStreamSynchronize(st->rdr->stream()) != hipSuccess)
        return fail(MG_ERR_DEVICE, "download failed");
    return MG_OK;
}

}  // extern "C"
