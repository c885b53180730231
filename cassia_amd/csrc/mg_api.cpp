// mold's scene API over the MI355X renderer, exported as the C ABI of
// include/mold_gpu.h. Bookkeeping follows mold/src/{composition,layer,buffer}.rs;
// every pixel operation runs on the GPU (mg_runtime.hip, mg_kernels.hip).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <queue>
#include <string>
#include <vector>

#include "../../include/mold_gpu.h"
#include "mg_path.h"
#include "mg_runtime.h"
#include "mg_shared.h"

namespace {

thread_local std::string g_error;

mg_status fail(mg_status s, const std::string &msg) {
    g_error = msg;
    return s;
}

const float kIdentity[6] = {1.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f};

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
    std::map<uint16_t, Layer> layers;
    std::priority_queue<uint32_t, std::vector<uint32_t>, std::greater<uint32_t>> free_ids;
    uint32_t next_id = 0;
    std::map<uint16_t, uint16_t> orders_to_layers;
    std::map<std::pair<uintptr_t, size_t>, std::pair<size_t, size_t>> layouts;
    uint32_t caches = 0;
    std::unique_ptr<mg::Renderer> rdr;
    std::vector<uint64_t> pixsegs;
    mg_frame_stats stats{};
    uint64_t version = 1;      // bumped by every change of layers, props or lines
    uint32_t radix_bits = 0;
    uint8_t *device_frame = nullptr;
    size_t device_frame_bytes = 0;
    // frame tables (layers by id, props by order) of the last render, valid
    // while `version` and the layer cache id are unchanged
    mg::FrameSpec spec;
    uint64_t spec_version = 0;
    int32_t spec_cache_id = -2;
    bool spec_clips = false;

    size_t actual_len() const {
        size_t n = 0;
        for (auto &kv : layers)
            if (kv.second.enabled) n += kv.second.len;
        return n;
    }

    // Composition::remove_disabled (composition.rs:123-142)
    void remove_disabled() {
        if (ids.size() >= actual_len() * 2) {
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

    mg_status ensure_renderer() {
        if (rdr) return MG_OK;
        auto r = std::make_unique<mg::Renderer>();
        std::string err;
        if (!r->init(err)) return fail(MG_ERR_DEVICE, err);
        rdr = std::move(r);
        return MG_OK;
    }

    mg_status sync_lines() {
        std::string err;
        if (uploaded > ids.size()) uploaded = 0;
        if (uploaded < ids.size()) {
            if (!rdr->upload_lines(&pts[4 * uploaded], &ids[uploaded], uploaded, ids.size() - uploaded, err))
                return fail(MG_ERR_OOM, err);
            uploaded = ids.size();
        }
        return MG_OK;
    }

    void insert(uint16_t id, const float *lines, size_t n) {
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
                }
            }
            if (cache_id >= 0 && ((l.is_unchanged >> cache_id) & 1u)) fl |= mg::PF_UNCHANGED;
            e.flags = fl;
            if (l.enabled) max_order = std::max<uint32_t>(max_order, kv.first);
        }
        f.max_order = std::max<uint32_t>(max_order, n_props - 1);
    }
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

MG_EXPORT mg_status mg_remove_disabled(mg_composition *c) {
    if (check_comp(c)) return MG_ERR_INVALID_ARG;
    c->remove_disabled();
    return MG_OK;
}

MG_EXPORT mg_status mg_lines_len(const mg_composition *c, size_t *builder_len, size_t *actual_len) {
    if (check_comp(c)) return MG_ERR_INVALID_ARG;
    if (builder_len) *builder_len = c->ids.size();
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
        std::string err;
        if (!c->rdr->cache_clear(id, err)) return fail(MG_ERR_DEVICE, err);
    }
    return MG_OK;
}

MG_EXPORT mg_status mg_buffer_layer_cache_release(mg_composition *c, int32_t id) {
    if (check_comp(c) || id < 0 || id >= 32) return fail(MG_ERR_INVALID_ARG, "unknown cache");
    c->caches &= ~(1u << id);
    if (c->rdr) c->rdr->cache_release(id);
    return MG_OK;
}

// Composition::render for one frame of `width` x `rows_px` pixels at
// `stride` into the device image `img` (composition.rs:212-328).
static mg_status render_frame(mg_composition *c, size_t width, size_t stride, size_t rows_px, const float bg[4],
                              const mg_rect *crop, int32_t cache_id, int32_t *prev, uint8_t *img,
                              bool want_tile_ops, mg::FrameResult &r, uint32_t coverage_cap = 0) {
    std::string err;
    const size_t row_len = (width + 7) / 8, tile_rows = (rows_px + 7) / 8;
    if (c->spec_version != c->version || c->spec_cache_id != cache_id) {
        // every change of lines, layers, props or cache state bumps `version`
        c->remove_disabled();
        for (auto &kv : c->layers)
            if (kv.second.enabled && kv.second.has_order) c->orders_to_layers[kv.second.order] = kv.first;
        c->spec = mg::FrameSpec{};
        c->tables(cache_id, c->spec, c->spec_clips);
        c->spec_version = c->version;
        c->spec_cache_id = cache_id;
    }
    mg::FrameSpec &f = c->spec;
    const bool has_clips = c->spec_clips;
    f.scenes.clear();
    f.scene_line_base.clear();
    f.tables_version = c->version * 64 + (uint64_t)(cache_id + 1);
    mg::SceneDesc sd{};
    sd.line_base = 0;
    sd.n_lines = (uint32_t)c->ids.size();
    sd.layer_base = 0;
    sd.n_layer_entries = (uint32_t)f.layers.size();
    sd.props_base = 0;
    sd.n_props = (uint32_t)f.props.size();
    sd.width = (uint32_t)width;
    sd.rows_px = (uint32_t)rows_px;
    sd.stride = (uint32_t)stride;
    sd.row_len = (uint32_t)row_len;
    sd.tile_rows = (uint32_t)tile_rows;
    sd.row_base = 0;
    sd.crop_h0 = 0;
    sd.crop_h1 = (uint32_t)row_len;
    sd.crop_v0 = 0;
    sd.crop_v1 = (uint32_t)tile_rows;
    if (crop) {
        // Rect::new (buffer_layout.rs:43-50): pixel ranges to tile ranges
        sd.crop_h0 = crop->x0 / 8;
        sd.crop_h1 = (uint32_t)std::min<uint64_t>(((uint64_t)crop->x1 + 7) / 8, row_len);
        sd.crop_v0 = crop->y0 / 8;
        sd.crop_v1 = (uint32_t)std::min<uint64_t>(((uint64_t)crop->y1 + 7) / 8, tile_rows);
        if (sd.crop_h0 > sd.crop_h1) sd.crop_h0 = sd.crop_h1;
        if (sd.crop_v0 > sd.crop_v1) sd.crop_v0 = sd.crop_v1;
    }
    sd.cache_base = 0;
    sd.flags = (has_clips ? mg::SF_CLIPS : 0) | (cache_id >= 0 ? mg::SF_CACHE : 0);
    sd.out_offset = 0;
    std::memcpy(sd.clear, bg, sizeof sd.clear);
    f.scenes.push_back(sd);
    f.scene_line_base.push_back(0);
    f.mode = 0;
    f.want_tile_ops = want_tile_ops;
    f.want_coverage = coverage_cap > 0;
    f.coverage_cap = coverage_cap;
    f.radix_bits = c->radix_bits;
    mg_status s;
    if ((s = c->sync_lines())) return s;
    if (!c->rdr->run(f, c->ids.size(), img, prev, r, err))
        return fail(err.find("capacity") != std::string::npos ? MG_ERR_CAPACITY : MG_ERR_DEVICE, err);
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
    c->stats.lines = c->ids.size();
    c->stats.line_pixels = r.line_pixels;
    c->stats.segments = r.segments;
    c->stats.segments_painted = r.kept;
    c->stats.tiles = r.tiles;
    c->stats.sort_passes = r.passes;
    c->stats.radix_bits = r.radix_bits;
    c->stats.device_ms = r.device_ms;
    c->stats.setup_ms = r.stage_ms[0];
    c->stats.emit_ms = r.stage_ms[1];
    c->stats.hist_ms = r.stage_ms[2];
    c->stats.sort_ms = r.stage_ms[3];
    c->stats.runs_ms = r.stage_ms[4];
    c->stats.walk_ms = r.stage_ms[5];
    c->stats.paint_ms = r.stage_ms[6];
    c->stats.runs = r.runs;
    c->stats.records = r.records;
    return MG_OK;
}

MG_EXPORT mg_status mg_render(mg_composition *c, const mg_buffer *buf, const float bg[4], const mg_rect *crop) {
    if (check_comp(c) || !buf || !bg || !buf->rgba) return fail(MG_ERR_INVALID_ARG, "null argument");
    mg_status s = c->ensure_renderer();
    if (s) return s;
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
        prev = c->rdr->cache_array(cache_id, tiles_len, resized, err);
        if (!prev) return fail(MG_ERR_OOM, err);
    }
    const size_t bytes = len * 4;
    uint8_t *img = c->rdr->frame_image(bytes, err);
    if (!img) return fail(MG_ERR_OOM, err);
    // the caller's pixels go first: tiles the painter skips keep them
    if (hipMemcpyAsync(img, buf->rgba, bytes, hipMemcpyHostToDevice, c->rdr->stream()) != hipSuccess)
        return fail(MG_ERR_DEVICE, "upload of the frame buffer failed");
    mg::FrameResult r;
    if ((s = render_frame(c, width, stride, rows_px, bg, crop, cache_id, prev, img, buf->flusher != nullptr, r)))
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
    std::string err;
    const size_t bytes = (size_t)width * height * 4;
    uint8_t *img = c->rdr->frame_image(bytes, err);
    if (!img) return fail(MG_ERR_OOM, err);
    mg::FrameResult r;
    if ((s = render_frame(c, width, width, height, bg, crop, -1, nullptr, img, false, r))) return s;
    *device_rgba = img;
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
    std::string err;
    const size_t bytes = (size_t)width * height * 4;
    uint8_t *img = c->rdr->frame_image(bytes, err);
    if (!img) return fail(MG_ERR_OOM, err);
    mg::FrameResult r;
    if ((s = render_frame(c, width, width, height, bg, crop, -1, nullptr, img, false, r,
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
    c->remove_disabled();
    mg::FrameSpec f;
    bool has_clips = false;
    c->tables(-1, f, has_clips);
    mg::SceneDesc sd{};
    sd.n_lines = (uint32_t)c->ids.size();
    sd.n_layer_entries = (uint32_t)f.layers.size();
    sd.n_props = (uint32_t)f.props.size();
    f.scenes.push_back(sd);
    f.scene_line_base.push_back(0);
    f.mode = 1;
    if (c->radix_bits) f.radix_bits = c->radix_bits;
    if ((s = c->sync_lines())) return s;
    std::string err;
    mg::FrameResult r;
    if (!c->rdr->run(f, c->ids.size(), nullptr, nullptr, r, err)) return fail(MG_ERR_DEVICE, err);
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

MG_EXPORT mg_status mg_last_frame_stats(const mg_composition *c, mg_frame_stats *out) {
    if (check_comp(c) || !out) return fail(MG_ERR_INVALID_ARG, "null argument");
    *out = c->stats;
    return MG_OK;
}

}  // extern "C"

// --------------------------------------------------------------------------
// cassia-compatible entry points (cassia/src/Cassia.h:32-41, Cassia.cpp:240-255)
// --------------------------------------------------------------------------

namespace {

struct CassiaState {
    uint32_t width = 0, height = 0;
    std::unique_ptr<mg::Renderer> rdr;
    mg::DevBuf in;
    uint8_t *image = nullptr;
};

CassiaState *g_cassia = nullptr;

}  // namespace

extern "C" {

MG_EXPORT void cassia_init(uint32_t width, uint32_t height) {
    delete g_cassia;
    g_cassia = new CassiaState();
    g_cassia->width = width;
    g_cassia->height = height;
    auto r = std::make_unique<mg::Renderer>();
    std::string err;
    if (!r->init(err)) {
        g_error = err;
        return;
    }
    g_cassia->image = r->frame_image((size_t)width * height * 4 + 4, err);
    if (!g_cassia->image) {
        g_error = err;
        return;
    }
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
    mg::FrameResult r;
    if (!st->rdr->run(f, 0, st->image, nullptr, r, err)) g_error = err;
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
