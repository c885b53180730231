// Device runtime of the MI355X mold renderer.
#include "mg_runtime.h"

#include <algorithm>
#include <cstring>

#include "mg_kernels.h"

namespace mg {

// words of the paint flags: overflow flags, record store use, run count,
// deep tiles and the walk's maxima at 0..6, record counter slices from
// kStatSlices on
constexpr uint32_t kOverflowWords = 3 * kStatSlices;  // flags, record slices, run slices

// The frame's status block on the device, cleared by the frame's first
// kernel and read back by one copy at each host synchronisation: counters;
// error flags (0: a line longer than i32 pixels, or line ranks inconsistent
// with the layer bases; 1: a look-back wait gave up; 2: the frame exceeded a
// replayed graph's line-pixel capacity); paint flags.
struct FrameStatus {
    FrameCounters ctr;
    uint32_t err[4];
    uint32_t ov[kOverflowWords];
};
static_assert(sizeof(FrameStatus) % 4 == 0 && sizeof(FrameStatus) <= 512, "status block");
constexpr size_t kStatusHead = offsetof(FrameStatus, ov);  // counters and error flags

// One launch clearing the given ranges of 32-bit words (empty ones skipped).
static void zero_words(hipStream_t s, uint32_t cus, std::initializer_list<std::pair<void *, uint64_t>> ranges) {
    ZeroRanges z{};
    uint64_t most = 0;
    for (const auto &r : ranges) {
        if (!r.first || !r.second) continue;
        z.p[z.count] = static_cast<uint32_t *>(r.first);
        z.n[z.count] = r.second;
        most = std::max(most, r.second);
        ++z.count;
    }
    if (!z.count) return;
    const uint64_t blocks = std::min<uint64_t>((most + 255) / 256, (uint64_t)std::max(cus, 1u) * 8);
    k_zero<<<(uint32_t)std::max<uint64_t>(blocks, 1), 256, 0, s>>>(z);
}

#define MG_HIP(call, what)                                                  \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) {                                             \
            err = std::string(what) + ": " + hipGetErrorString(e_);         \
            return false;                                                   \
        }                                                                   \
    } while (0)

void *Arena::take(size_t n) {
    const size_t want = (std::max<size_t>(n, 256) + 255) & ~size_t(255);
    if (want > size - used) {
        short_by = std::max(short_by, want - (size - used));
        return nullptr;
    }
    void *p = base + used;
    used += want;
    return p;
}

bool DevBuf::reserve(size_t n) {
    if (n <= bytes) return true;
    if (arena) {
        void *p = arena->take(n);
        if (!p) return false;
        ptr = p;
        bytes = (std::max<size_t>(n, 256) + 255) & ~size_t(255);
        return true;
    }
    release();
    size_t want = std::max<size_t>(n, 256);
    if (hipMalloc(&ptr, want) != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        return false;
    }
    bytes = want;
    return true;
}

void DevBuf::release() {
    if (ptr && !arena) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

static bool is_gfx950(int dev) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return false;
    return std::strncmp(p.gcnArchName, "gfx950", 6) == 0;
}

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    int ok = 0;
    for (int d = 0; d < n; ++d)
        if (is_gfx950(d)) ++ok;
    return ok;
}

int render_device() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return -1;
    for (int d = 0; d < n; ++d)
        if (is_gfx950(d)) return d;
    return -1;
}

Renderer::Renderer() {}

std::vector<DevBuf *> Renderer::all_buffers() {
    return {&scenes_, &scene_line_base_, &layers_, &props_, &stops_,
            &setup_, &lengths_, &offsets_, &scan_status_, &status_, &segs_a_,
            &segs_b_, &dig_,
            &row_start_, &row_scene_,
            &tile_ops_, &cov_a_, &cov_key_, &cov_count_, &work_, &part_line_, &cnt0_, &cnt_acc_, &cnt_status_, &tile_start_, &tile_run0_,
            &run_key_, &run_beg_, &run_cov_, &rec_, &rec_len_, &rec_hdr_, &runs_status_, &runs_work_, &tile_nr_,
            &walk_scratch_, &part_kept_, &flat_cmds_, &flat_paths_, &flat_counts_, &flat_offsets_, &flat_status_,
            &deep_list_};
}

Renderer::~Renderer() {
    if (device_ >= 0) {
        (void)hipSetDevice(device_);
        if (stream_) (void)hipStreamSynchronize(stream_);
    }
    drop_graph();
    for (DevBuf *b : all_buffers()) b->release();
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    for (hipEvent_t e : evs_)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : evp_)
        if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
    if (status_pinned_) (void)hipHostFree(status_host_);
}

// Copies the first `bytes` of the status block to its host copy and waits
// for the stream.
bool Renderer::read_status(size_t bytes, std::string &err) {
    MG_HIP(hipMemcpyAsync(status_host_, status_.ptr, bytes, hipMemcpyDeviceToHost, stream_), "read status");
    MG_HIP(hipStreamSynchronize(stream_), "frame");
    return true;
}

bool Renderer::init(std::string &err) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        err = "no HIP device";
        return false;
    }
    device_ = render_device();
    if (device_ < 0) {
        err = "no gfx950 device";
        return false;
    }
    MG_HIP(hipSetDevice(device_), "hipSetDevice");
    hipDeviceProp_t p;
    MG_HIP(hipGetDeviceProperties(&p, device_), "hipGetDeviceProperties");
    cus_ = p.multiProcessorCount;
    lds_max_ = p.sharedMemPerBlock;
    // dynamic LDS limits of the kernels that size their LDS per frame, set
    // once to the device maximum (the attribute is per kernel, shared by
    // every engine of the process)
    for (const void *fn : {(const void *)k_row_walk<64, false>, (const void *)k_row_walk<256, false>,
                           (const void *)k_row_walk<512, false>, (const void *)k_row_walk<1024, false>,
                           (const void *)k_paint,
                           (const void *)k_paint_deep}) {
        hipFuncAttributes a;
        MG_HIP(hipFuncGetAttributes(&a, fn), "kernel attributes");
        MG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)(lds_max_ - a.sharedSizeBytes)),
               "lds limit");
    }
    MG_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    MG_HIP(hipEventCreate(&ev0_), "hipEventCreate");
    MG_HIP(hipEventCreate(&ev1_), "hipEventCreate");
    for (hipEvent_t &e : evs_) MG_HIP(hipEventCreate(&e), "hipEventCreate");
    for (hipEvent_t &e : evp_) MG_HIP(hipEventCreate(&e), "hipEventCreate");
    if (!status_.reserve(sizeof(FrameStatus)) || !cov_count_.reserve(16)) {
        err = "out of device memory";
        return false;
    }
    // the status block's host copy: pinned, except for an engine that must
    // not allocate (arena mode)
    status_host_ = status_local_;
    if (!arena_mode_ && hipHostMalloc(&status_host_, sizeof(FrameStatus), hipHostMallocDefault) == hipSuccess)
        status_pinned_ = true;
    else
        status_host_ = status_local_;
    return true;
}

void Renderer::arena_begin() {
    arena_.used = 0;
    arena_.short_by = 0;
    for (DevBuf *b : all_buffers()) b->release();
    // tables, work lists and the record store are taken anew by the frame
    uploaded_owner_ = uploaded_version_ = 0;
    uploaded_scenes_.clear();
    work_scenes_.clear();
    n_work_ = n_runs_work_ = 0;
    rec_cap_ = 0;
    last_sorted_ = nullptr;
    // the frame's fixed counters (their contents are cleared by every frame)
    (void)(status_.reserve(sizeof(FrameStatus)) && cov_count_.reserve(16));
}

std::shared_ptr<Renderer> Renderer::shared_arena(void *base, size_t bytes, std::string &err) {
    static std::mutex m;
    static std::shared_ptr<Renderer> r;  // lives until the process ends, as the arena does
    std::lock_guard<std::mutex> g(m);
    if (r) return r;
    auto e = std::make_shared<Renderer>();
    e->arena_mode_ = true;
    e->arena_.base = static_cast<unsigned char *>(base);
    e->arena_.size = bytes;
    for (DevBuf *b : e->all_buffers()) b->arena = &e->arena_;
    if (!e->init(err)) return nullptr;
    r = e;
    return r;
}

std::shared_ptr<Renderer> Renderer::shared(std::string &err) {
    static std::mutex m;
    static std::weak_ptr<Renderer> w;
    std::lock_guard<std::mutex> g(m);
    if (auto r = w.lock()) return r;
    auto r = std::make_shared<Renderer>();
    if (!r->init(err)) return nullptr;
    w = r;
    return r;
}

bool LineStore::upload(const float *pts4, const uint16_t *ids_h, size_t first, size_t n, hipStream_t s,
                       std::string &err) {
    size_t need = first + n;
    if (need > cap) {
        size_t ncap = std::max<size_t>(need, cap * 2);
        DevBuf np, ni;
        if (!np.reserve(ncap * 16) || !ni.reserve(ncap * 2)) {
            np.release();
            ni.release();
            err = "out of device memory (lines)";
            return false;
        }
        if (first > 0 && pts.ptr) {
            MG_HIP(hipMemcpyAsync(np.ptr, pts.ptr, first * 16, hipMemcpyDeviceToDevice, s), "copy lines");
            MG_HIP(hipMemcpyAsync(ni.ptr, ids.ptr, first * 2, hipMemcpyDeviceToDevice, s), "copy ids");
        }
        MG_HIP(hipStreamSynchronize(s), "sync");
        pts.release();
        ids.release();
        pts = np;
        ids = ni;
        cap = ncap;
    }
    if (n) {
        MG_HIP(hipMemcpyAsync(pts.as<char>() + first * 16, pts4, n * 16, hipMemcpyHostToDevice, s), "upload lines");
        MG_HIP(hipMemcpyAsync(ids.as<char>() + first * 2, ids_h, n * 2, hipMemcpyHostToDevice, s), "upload ids");
        MG_HIP(hipStreamSynchronize(s), "sync");
    }
    return true;
}

bool LineStore::grow(size_t need, size_t keep, hipStream_t s, std::string &err) {
    if (need <= cap) return true;
    size_t ncap = std::max<size_t>(need, cap * 2);
    DevBuf np, ni;
    if (!np.reserve(ncap * 16) || !ni.reserve(ncap * 2)) {
        np.release();
        ni.release();
        err = "out of device memory (lines)";
        return false;
    }
    if (keep && pts.ptr) {
        MG_HIP(hipMemcpyAsync(np.ptr, pts.ptr, keep * 16, hipMemcpyDeviceToDevice, s), "copy lines");
        MG_HIP(hipMemcpyAsync(ni.ptr, ids.ptr, keep * 2, hipMemcpyDeviceToDevice, s), "copy ids");
    }
    MG_HIP(hipStreamSynchronize(s), "sync");
    pts.release();
    ids.release();
    pts = np;
    ids = ni;
    cap = ncap;
    return true;
}

void LineStore::release() {
    pts.release();
    ids.release();
    rank.release();
    counts.release();
    cap = 0;
    ranked = 0;
    count_ids = 0;
}



static void fill32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) {
    if (!n) return;
    uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    k_fill_u32<<<blocks, 256, 0, s>>>(p, v, n);
}

static void fill64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s) {
    if (!n) return;
    uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    k_fill_u64<<<blocks, 256, 0, s>>>(p, v, n);
}

bool Renderer::rank_lines(const uint16_t *ids, uint32_t *counts, uint32_t n_ids, uint32_t *rank, size_t first,
                          size_t n, bool reset, std::string &err) {
    MG_HIP(hipSetDevice(device_), "hipSetDevice");
    if (reset) fill32(counts, 0, n_ids, stream_);
    if (n > first)
        k_line_ranks<<<(uint32_t)((n - first + 255) / 256), 256, 0, stream_>>>(ids, first, n, counts, rank);
    MG_HIP(hipGetLastError(), "line ranks launch");
    return true;
}

int32_t *LayerCaches::array(int id, size_t n_tiles, hipStream_t s, bool &resized, std::string &err) {
    resized = false;
    if (id < 0) return nullptr;
    if ((size_t)id >= bufs.size()) {
        bufs.resize(id + 1);
        tiles.resize(id + 1, 0);
    }
    if (tiles[id] != n_tiles) {
        if (!bufs[id].reserve(std::max<size_t>(n_tiles, 1) * 4)) {
            err = "out of device memory (cache)";
            return nullptr;
        }
        fill32(bufs[id].as<uint32_t>(), 0xFFFFFFFFu, n_tiles, s);
        tiles[id] = n_tiles;
        resized = true;
    }
    return bufs[id].as<int32_t>();
}

void LayerCaches::clear(int id, hipStream_t s) {
    if (id >= 0 && (size_t)id < bufs.size() && bufs[id].ptr) fill32(bufs[id].as<uint32_t>(), 0xFFFFFFFFu, tiles[id], s);
}

void LayerCaches::release(int id) {
    if (id >= 0 && (size_t)id < bufs.size()) {
        bufs[id].release();
        tiles[id] = 0;
    }
}

void LayerCaches::release_all() {
    for (DevBuf &b : bufs) b.release();
    bufs.clear();
    tiles.clear();
}

static uint32_t bits_for(uint64_t v) {  // bits to represent 0..v
    uint32_t b = 0;
    while (b < 64 && (v >> b)) ++b;
    return b;
}

template <class T>
static bool upload_vec(DevBuf &b, const std::vector<T> &v, hipStream_t s, std::string &err) {
    if (!b.reserve(std::max<size_t>(v.size(), 1) * sizeof(T))) {
        err = "out of device memory";
        return false;
    }
    if (!v.empty() && hipMemcpyAsync(b.ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s) != hipSuccess) {
        err = "upload failed";
        return false;
    }
    return true;
}

KeyLayout key_layout(const FrameSpec &spec) {
    uint32_t global_rows = 0, max_cols = 1;
    for (const SceneDesc &s : spec.scenes) {
        global_rows = std::max(global_rows, s.row_base + s.tile_rows);
        max_cols = std::max(max_cols, std::min(s.crop_h1, s.row_len));
    }
    KeyLayout k;
    k.col_bits = std::max(1u, bits_for(max_cols));
    k.layer_bits = std::max(1u, bits_for(spec.max_order));
    const uint32_t row_bits = global_rows > 1 ? bits_for(global_rows - 1) : 0;
    k.key_bits = spec.mode != 1 ? row_bits + k.col_bits + k.layer_bits : 41;
    // Render frames emit their lines grouped by layer order, so slot order is
    // layer order inside every tile and a stable sort on (row, column) alone
    // gives (row, column, layer). The canonical stream (generation order) and
    // ingested streams sort on the whole key.
    const bool order_lines = spec.mode == 0;
    k.sort_shift = kPayloadBits + (order_lines ? k.layer_bits : 0);
    k.sort_bits = order_lines ? std::max(1u, row_bits + k.col_bits) : k.key_bits;
    return k;
}

// Host check of the scene descriptors the kernels index with: scenes own
// disjoint tile rows and per-tile ranges (in scene order) inside the frame's
// totals, their lines and tables lie inside what is uploaded, their crops
// inside their canvas and their pixels inside the output image.
static bool scenes_valid(const FrameSpec &spec, size_t n_lines, const uint8_t *image, uint64_t tiles,
                         std::string &err) {
    uint64_t row_end = 0, tile_end = 0;
    for (size_t i = 0; i < spec.scenes.size(); ++i) {
        const SceneDesc &s = spec.scenes[i];
        bool ok = s.row_base >= row_end && s.tile_base >= tile_end;
        row_end = (uint64_t)s.row_base + s.tile_rows;
        tile_end = (uint64_t)s.tile_base + (uint64_t)s.tile_rows * s.row_len;
        ok = ok && (uint64_t)s.layer_base + s.n_layer_entries <= spec.layers.size() &&
             (uint64_t)s.props_base + s.n_props <= spec.props.size();
        if (spec.mode != 2)
            ok = ok && (uint64_t)s.line_base + s.n_lines <= n_lines && i < spec.scene_line_base.size() &&
                 spec.scene_line_base[i] == s.line_base;
        ok = ok && s.row_len == (s.width + 7) / 8 && s.tile_rows == (s.rows_px + 7) / 8 && s.stride >= s.width &&
             s.crop_h0 <= s.crop_h1 && s.crop_h1 <= s.row_len && s.crop_v0 <= s.crop_v1 && s.crop_v1 <= s.tile_rows;
        if (image && s.rows_px && s.width)
            ok = ok && s.out_offset + ((uint64_t)(s.rows_px - 1) * s.stride + s.width) * 4 <= spec.image_bytes;
        if (!ok) {
            err = "invalid scene descriptor " + std::to_string(i);
            return false;
        }
    }
    if (tile_end > tiles) {
        err = "scene tile ranges exceed the frame's tiles";
        return false;
    }
    return true;
}

static bool props_valid(const FrameSpec &spec, std::string &err) {
    for (const PropsEntry &p : spec.props)
        if ((p.flags & PF_GRADIENT) && (uint64_t)p.stops_offset + 5ull * p.n_stops > spec.stops.size()) {
            err = "gradient stops outside the stops table";
            return false;
        }
    return true;
}

// Scene descriptors, layer / props / stops tables and the row -> scene map of
// a frame, uploaded unless the engine holds them already.
bool Renderer::upload_tables(const FrameSpec &spec, uint32_t global_rows, std::string &err) {
    const bool same_scenes = spec.tables_version != 0 && spec.tables_owner == uploaded_owner_ &&
                             spec.tables_version == uploaded_version_ &&
                             uploaded_scenes_.size() == spec.scenes.size() &&
                             std::memcmp(uploaded_scenes_.data(), spec.scenes.data(),
                                         spec.scenes.size() * sizeof(SceneDesc)) == 0;
    if (same_scenes) return true;
    if (!props_valid(spec, err)) return false;
    if (!upload_vec(scenes_, spec.scenes, stream_, err) ||
        !upload_vec(scene_line_base_, spec.scene_line_base, stream_, err) ||
        !upload_vec(layers_, spec.layers, stream_, err) || !upload_vec(props_, spec.props, stream_, err) ||
        !upload_vec(stops_, spec.stops, stream_, err))
        return false;
    std::vector<uint32_t> rs(std::max<uint32_t>(global_rows, 1), 0);
    for (uint32_t s = 0; s < spec.scenes.size(); ++s)
        for (uint32_t r = 0; r < spec.scenes[s].tile_rows; ++r) rs[spec.scenes[s].row_base + r] = s;
    if (!upload_vec(row_scene_, rs, stream_, err)) return false;
    uploaded_owner_ = spec.tables_owner;
    uploaded_version_ = spec.tables_version;
    uploaded_scenes_ = spec.scenes;
    return true;
}

bool Renderer::run(const FrameSpec &spec, const float4 *pts, const uint16_t *ids, const uint32_t *ranks,
                   size_t n_lines, uint8_t *image, int32_t *prev_layers, FrameResult &res, std::string &err) {
    MG_HIP(hipSetDevice(device_), "hipSetDevice");
    const uint32_t n_scenes = (uint32_t)spec.scenes.size();
    if (n_scenes == 0 || n_scenes > 8192) {  // LineSetup::packed holds a 13-bit scene index
        err = "a frame holds 1 to 8192 scenes";
        return false;
    }
    uint32_t global_rows = 0;
    uint32_t max_cols = 1;
    uint64_t tiles = 0;
    for (const SceneDesc &s : spec.scenes) {
        global_rows = std::max(global_rows, s.row_base + s.tile_rows);
        max_cols = std::max(max_cols, std::min(s.crop_h1, s.row_len));
        tiles += (uint64_t)s.tile_rows * s.row_len;
    }
    if (!scenes_valid(spec, n_lines, image, tiles, err)) return false;
    FrameDims dims{};
    dims.n_scenes = n_scenes;
    dims.n_lines = (uint32_t)n_lines;
    dims.global_rows = global_rows;
    const KeyLayout kl = key_layout(spec);
    dims.col_bits = kl.col_bits;
    dims.layer_bits = kl.layer_bits;
    dims.key_bits = kl.key_bits;
    const bool order_lines = spec.mode == 0;
    dims.sort_shift = kl.sort_shift;
    const uint32_t sort_bits = kl.sort_bits;
    uint32_t R = spec.radix_bits;
    if (R == 0)  // measured per-pass cost on MI355X: 6-bit 2.3 ms, 8-bit 2.6 ms, 9-bit 4.0 ms (3e8 words)
        R = (sort_bits + 7) / 8 < (sort_bits + 5) / 6 ? 8 : 6;
    const uint32_t radix = 1u << R;
    dims.radix_bits = R;
    dims.passes = (sort_bits + R - 1) / R;
    dims.mode = spec.mode;
    if (kPayloadBits + dims.key_bits > 64 || dims.passes > kMaxSortPasses) {
        err = "frame too large for the 64-bit render key";
        return false;
    }
    res = FrameResult{};
    res.passes = dims.passes;
    res.radix_bits = R;
    res.sort_bits = sort_bits;
    res.tiles = tiles;

    if (!upload_tables(spec, global_rows, err)) return false;
    const size_t L = n_lines;
    if (!setup_.reserve(std::max<size_t>(L, 1) * sizeof(LineSetup)) || !lengths_.reserve(std::max<size_t>(L, 1) * 4) ||
        !offsets_.reserve(std::max<size_t>(L, 1) * 4)) {
        err = "out of device memory (lines)";
        return false;
    }
    FrameStatus *fs = status_.as<FrameStatus>();
    FrameCounters *ctr = &fs->ctr;
    uint32_t *stall = fs->err + 1;
    status_final_ = false;
    const bool ingest = spec.mode == 2;
    const bool lines = L && !ingest;
    // render frames place each line at its layer's base plus its rank (lines
    // ordered by scene and layer order); a length left unwritten would only
    // come from inconsistent tables, and stays 0
    if (lines && order_lines && !ranks && !spec.scenes[0].lines_rank) {
        err = "render frame without line ranks";
        return false;
    }
    const uint64_t nparts_scan = (L + kScanTileHost - 1) / kScanTileHost;
    const uint32_t count_rows = radix + 2;  // first-pass digits, kept, culled
    if (!scan_status_.reserve(std::max<uint64_t>(nparts_scan, 1) * 8) ||
        !cnt_acc_.reserve((uint64_t)kCountSlicesHost * 2 * 4)) {
        err = "out of device memory";
        return false;
    }
    MG_HIP(mark(ev0_), "event");
    MG_HIP(mark(evs_[0]), "event");
    // the status block, the lines' lengths, the scan's look-back words and
    // the kept/culled slices the emission adds into
    zero_words(stream_, (uint32_t)cus_,
               {{fs, sizeof(FrameStatus) / 4}, {lengths_.ptr, lines && order_lines ? L : 0},
                {scan_status_.ptr, lines ? 2 * nparts_scan : 0}, {cnt_acc_.ptr, (uint64_t)kCountSlicesHost * 2}});
    if (lines) {
        k_setup<<<(uint32_t)((L + 255) / 256), 256, 0, stream_>>>(
            pts, ids, ranks, layers_.as<LayerEntry>(), scenes_.as<SceneDesc>(), scene_line_base_.as<uint32_t>(), dims,
            setup_.as<LineSetup>(), lengths_.as<uint32_t>(), fs->err, order_lines ? 1u : 0u);
        k_scan_u32<<<(uint32_t)nparts_scan, 256, 0, stream_>>>(lengths_.as<uint32_t>(), offsets_.as<uint32_t>(), L,
                                                             scan_status_.as<uint64_t>(), stall, &ctr->line_pixels);
    }
    MG_HIP(hipGetLastError(), "setup/scan launch");
    const FrameStatus &hs = *static_cast<const FrameStatus *>(status_host_);
    uint64_t P = 0;
    if (!ingest && capture_P_) {
        // frame of a graph: buffers and grids for the capacity, no host read
        k_clamp_line_pixels<<<1, 64, 0, stream_>>>(ctr, capture_P_, fs->err + 2);
        P = capture_P_;
    } else if (!ingest) {
        if (!read_status(kStatusHead, err)) return false;
        if (hs.err[0]) {
            err = (hs.err[0] & 2u) ? "line ranks and layer bases of the frame are inconsistent"
                                   : "a line spans more pixels than an i32 can index";
            return false;
        }
        P = hs.ctr.line_pixels;
    } else {
        P = (spec.n_in + 1) / 2;  // capacity: n_in segments
    }
    if (2 * P >= (1ull << 30)) {
        err = "frame exceeds 2^29 line-pixels";
        return false;
    }
    res.line_pixels = ingest ? 0 : P;
    MG_HIP(mark(evs_[1]), "event");
    const uint64_t cap_segs = std::max<uint64_t>(2 * P, 1);
    const uint64_t nparts_emit = (P + kEmitTileHost - 1) / kEmitTileHost;
    const uint64_t nparts_sort = (cap_segs + kSortTileHost - 1) / kSortTileHost;
    // the next pass's digits, padded to whole partitions (k_digit_count reads whole ones)
    const size_t dig_size = R == 9 ? 2 : 1;
    if (!segs_a_.reserve(cap_segs * 8) || !segs_b_.reserve(cap_segs * 8) ||
        !dig_.reserve(std::max<uint64_t>(nparts_sort, 1) * kSortTileHost * dig_size) ||
        !row_start_.reserve(((size_t)global_rows + 1) * 4)) {
        err = "out of device memory (segments)";
        return false;
    }
    // Per-partition digit counts, digit-major, scanned in place into a
    // pass's output offsets: the emission's for the first pass, then
    // k_digit_count's for each later one (each scan has its own look-back
    // words).
    // (the emission's counts are per emission partition, twice as many as
    // the sort partitions whose counts the later passes scan)
    const uint64_t n_cnt0 = (uint64_t)radix * std::max<uint64_t>(nparts_emit, 1);
    const uint64_t cnt_parts = (n_cnt0 + kScanTileHost - 1) / kScanTileHost;
    if (!cnt0_.reserve(n_cnt0 * 4) || !cnt_status_.reserve(std::max<uint64_t>(cnt_parts, 1) * 8 * dims.passes) ||
        !part_kept_.reserve(std::max<uint64_t>(std::max(4 * nparts_emit, 8 * nparts_sort), 1) * 4)) {
        err = "out of device memory (slot counts)";
        return false;
    }
    // a capacity frame has partitions beyond the frame's: their counts are zero
    if (capture_P_) fill32(cnt0_.as<uint32_t>(), 0, n_cnt0, stream_);
    SlotCounts counts{cnt0_.as<uint32_t>(), cnt_acc_.as<uint32_t>(), part_kept_.as<uint32_t>()};
    const size_t count_lds = (size_t)count_rows * 4;
    // every line-pixel owns two word slots; the ingest path uses the same layout
    if (ingest && spec.n_in) {
        fill64((uint64_t *)&ctr->line_pixels, P, 1, stream_);
        k_ingest<<<(uint32_t)nparts_emit, 256, count_lds, stream_>>>(spec.in_segs, spec.n_in, scenes_.as<SceneDesc>(),
                                                                   dims, ctr, segs_a_.as<uint64_t>(), counts);
    } else if (P) {
        if (!part_line_.reserve(nparts_emit * 4)) { err = "out of device memory"; return false; }
        k_part_lines<<<(uint32_t)((L + 255) / 256), 256, 0, stream_>>>(offsets_.as<uint32_t>(), lengths_.as<uint32_t>(), L,
                                                                      kEmitTileHost, part_line_.as<uint32_t>(),
                                                                      nparts_emit);
        k_emit<<<xcd_grid(nparts_emit), 256, count_lds, stream_>>>(setup_.as<LineSetup>(), offsets_.as<uint32_t>(),
                                                                 scenes_.as<SceneDesc>(), dims, ctr,
                                                                 segs_a_.as<uint64_t>(), part_line_.as<uint32_t>(),
                                                                 counts);
    }
    MG_HIP(hipGetLastError(), "emit launch");
    MG_HIP(mark(evs_[2]), "event");
    // look-back words of the digit-count scans, cleared outside the sort's timing
    zero_words(stream_, (uint32_t)cus_, {{cnt_status_.ptr, P ? 2 * cnt_parts * dims.passes : 0}});
    if (P) {
        k_count_reduce<<<2, 256, 0, stream_>>>(cnt_acc_.as<uint32_t>(), ctr);
        k_scan_u32<<<(uint32_t)cnt_parts, 256, 0, stream_>>>(
            cnt0_.as<uint32_t>(), cnt0_.as<uint32_t>(), n_cnt0, cnt_status_.as<uint64_t>(), stall, &ctr->order_total,
            ScanLimit{&ctr->line_pixels, radix, kEmitTileHost});
    }
    MG_HIP(hipGetLastError(), "count launch");
    MG_HIP(mark(evs_[3]), "event");
    uint64_t *in = segs_a_.as<uint64_t>(), *out = segs_b_.as<uint64_t>();
    const uint32_t sort_grid = xcd_grid(nparts_sort);
    for (uint32_t p = 0; p < dims.passes; ++p) {
        const uint32_t shift = dims.sort_shift + p * R;
        if (p && P) {
            // offsets of this pass: the digits the previous pass stored, counted and scanned
            if (R == 9)
                k_digit_count<9><<<sort_grid, 256, 0, stream_>>>(dig_.as<uint16_t>(), ctr, cnt0_.as<uint32_t>());
            else if (R == 8)
                k_digit_count<8><<<sort_grid, 256, 0, stream_>>>(dig_.as<uint8_t>(), ctr, cnt0_.as<uint32_t>());
            else if (R == 6)
                k_digit_count<6><<<sort_grid, 256, 0, stream_>>>(dig_.as<uint8_t>(), ctr, cnt0_.as<uint32_t>());
            else
                k_digit_count<4><<<sort_grid, 256, 0, stream_>>>(dig_.as<uint8_t>(), ctr, cnt0_.as<uint32_t>());
            k_scan_u32<<<(uint32_t)cnt_parts, 256, 0, stream_>>>(
                cnt0_.as<uint32_t>(), cnt0_.as<uint32_t>(), n_cnt0, cnt_status_.as<uint64_t>() + (size_t)p * cnt_parts,
                stall, &ctr->order_total, ScanLimit{&ctr->kept, radix, kSortTileHost});
        }
        const uint32_t *kept = p == 0 ? part_kept_.as<uint32_t>() : nullptr;
        const bool next = p + 1 < dims.passes;
        const uint32_t *offs = cnt0_.as<uint32_t>();
        MG_HIP(mark(evp_[2 * p]), "event");
        if (R == 9)
            k_sort_pass<9><<<sort_grid, kSortThreadsHost, 0, stream_>>>(in, out, ctr, shift, offs, kept,
                                                                         next ? dig_.as<uint16_t>() : nullptr);
        else if (R == 8)
            k_sort_pass<8><<<sort_grid, kSortThreadsHost, 0, stream_>>>(in, out, ctr, shift, offs, kept,
                                                                         next ? dig_.as<uint8_t>() : nullptr);
        else if (R == 6)
            k_sort_pass<6><<<sort_grid, kSortThreadsHost, 0, stream_>>>(in, out, ctr, shift, offs, kept,
                                                                         next ? dig_.as<uint8_t>() : nullptr);
        else
            k_sort_pass<4><<<sort_grid, kSortThreadsHost, 0, stream_>>>(in, out, ctr, shift, offs, kept,
                                                                         next ? dig_.as<uint8_t>() : nullptr);
        MG_HIP(mark(evp_[2 * p + 1]), "event");
        std::swap(in, out);
    }
    MG_HIP(hipGetLastError(), "sort launch");
    MG_HIP(mark(evs_[4]), "event");
    n_runs_ = n_records_ = 0;
    // paint re-records these at its kernel boundaries
    for (int k = 5; k < 8; ++k) MG_HIP(mark(evs_[k]), "event");
    last_sorted_ = in;
    res.sorted = in;
    if (spec.mode != 1) {
        const uint32_t row_shift = kPayloadBits + dims.col_bits + dims.layer_bits;
        k_row_starts<<<(global_rows + 1 + 255) / 256, 256, 0, stream_>>>(in, ctr, global_rows, row_shift,
                                                                         row_start_.as<uint32_t>());
    }
    if (spec.mode != 1) {
        if (spec.want_tile_ops) {
            if (!tile_ops_.reserve(std::max<uint64_t>(tiles, 1))) { err = "out of device memory"; return false; }
            MG_HIP(hipMemsetAsync(tile_ops_.ptr, 0xFF, std::max<uint64_t>(tiles, 1), stream_), "tile ops");
        }
        if (spec.want_coverage) {
            if (!cov_a_.reserve((size_t)spec.coverage_cap * 64 * 4) || !cov_key_.reserve((size_t)spec.coverage_cap * 16)) {
                err = "out of device memory";
                return false;
            }
            fill32(cov_count_.as<uint32_t>(), 0, 1, stream_);
        }
        if (!paint(spec, global_rows, dims.col_bits, dims.layer_bits, in, cap_segs, image, prev_layers, tiles, res,
                   err))
            return false;
    }
    if (capture_P_) {
        MG_HIP(mark(ev1_), "event");
        return true;  // the replay reads the results
    }
    // (a painted frame's last read was the paint's)
    if (!status_final_) {
        MG_HIP(mark(ev1_), "event");
        if (!read_status(sizeof(FrameStatus), err)) return false;
    }
    if (hs.err[1]) {
        err = "a decoupled look-back wait timed out; the frame is invalid";
        return false;
    }
    float ms = 0.0f;
    MG_HIP(hipEventElapsedTime(&ms, ev0_, ev1_), "elapsed");
    for (int k = 0; k < 7; ++k) {
        float sm = 0.0f;
        MG_HIP(hipEventElapsedTime(&sm, evs_[k], evs_[k + 1]), "elapsed");
        res.stage_ms[k] = sm;
    }
    for (uint32_t p = 0; p < dims.passes; ++p) {
        float sm = 0.0f;
        MG_HIP(hipEventElapsedTime(&sm, evp_[2 * p], evp_[2 * p + 1]), "elapsed");
        res.pass_ms[p] = sm;
    }
    res.segments = hs.ctr.segments;
    res.kept = hs.ctr.kept;
    res.device_ms = ms;
    res.runs = n_runs_;
    res.records = n_records_;
    res.deep_tiles = n_deep_;
    return true;
}

bool Renderer::paint(const FrameSpec &spec, uint32_t global_rows, uint32_t col_bits, uint32_t layer_bits,
                     const uint64_t *sorted, uint64_t words_cap, uint8_t *image, int32_t *prev_layers, size_t tiles,
                     FrameResult &res, std::string &err) {
    if (global_rows == 0) return true;
    // one paint work group per tile inside the crop; the list is rebuilt
    // only when the scene descriptors change
    uint32_t tstride = 2;
    bool clips = false;
    for (const SceneDesc &s : spec.scenes) {
        tstride = std::max(tstride, std::min(s.crop_h1, s.row_len) + 2);
        clips |= (s.flags & SF_CLIPS) != 0;
    }
    if (!upload_work(spec, err)) return false;
    const uint32_t n_work = n_work_;
    if (n_work == 0) return true;
    const size_t n_cols = (size_t)global_rows * tstride;
    if (!rec_hdr_.reserve(std::max<size_t>(tiles, 1) * 8) || !tile_start_.reserve(n_cols * 4) ||
        !tile_run0_.reserve(n_cols * 4) || !tile_nr_.reserve(n_cols * 4) || !run_key_.reserve(words_cap * 4) ||
        !run_beg_.reserve((words_cap + 1) * 4) || !run_cov_.reserve(words_cap * 8) ||
        !runs_status_.reserve(((words_cap + kRunsTileHost - 1) / kRunsTileHost + 1) * 12)) {
        err = "out of device memory (tile runs)";
        return false;
    }

    PaintParams pp{};
    pp.segs = sorted;
    pp.row_start = row_start_.as<uint32_t>();
    pp.row_scene = row_scene_.as<uint32_t>();
    pp.scenes = scenes_.as<SceneDesc>();
    pp.props = props_.as<PropsEntry>();
    pp.stops = stops_.as<float>();
    pp.prev_layers = prev_layers;
    pp.out = image;
    pp.col_bits = col_bits;
    pp.layer_bits = layer_bits;
    FrameStatus *fs = status_.as<FrameStatus>();
    uint32_t *ov_dev = fs->ov;
    pp.overflow = ov_dev;
    pp.tile_ops = spec.want_tile_ops ? tile_ops_.as<int8_t>() : nullptr;
    pp.tile_start = tile_start_.as<uint32_t>();
    pp.tile_run0 = tile_run0_.as<uint32_t>();
    pp.tile_nr = tile_nr_.as<uint32_t>();
    pp.tstride = tstride;
    pp.run_key = run_key_.as<uint32_t>();
    pp.run_beg = run_beg_.as<uint32_t>();
    pp.run_cov = run_cov_.as<uint64_t>();
    pp.rec_hdr = rec_hdr_.as<uint2>();
    pp.rec_used = ov_dev + 1;
    pp.work = work_.as<uint2>();
    if (spec.want_coverage) {
        pp.cov_a = cov_a_.as<int32_t>();
        pp.cov_key = cov_key_.as<uint32_t>();
        pp.cov_count = cov_count_.as<uint32_t>();
        pp.cov_cap = spec.coverage_cap;
    }
    pp.n_runs = ov_dev + 2 * kStatSlices;
    pp.n_records = ov_dev + kStatSlices;
    // Tiles of 8,192 records or more go to the four-wave deep painter (at most
    // two per CU; more stay with the one-wave painter). A tile holds one wave
    // of pixels, so a frame of few deep tiles would otherwise leave the GPU
    // waiting on the serial record loops of a few waves.
    pp.deep_min = spec.want_coverage ? 0u : 8192u;
    pp.deep_cap = 2 * cus_;
    pp.deep_count = ov_dev + 4;
    pp.walk_max = ov_dev + 5;
    if (!deep_list_.reserve((size_t)pp.deep_cap * 8)) {
        err = "out of device memory (deep tile list)";
        return false;
    }
    pp.deep_list = deep_list_.as<uint2>();
    // Tile runs: one wave per tile column while tiles are many and of
    // moderate depth; over the whole sorted stream (work follows the words)
    // when a frame's tiles are deep: more than 8,192 word slots per column,
    // as in small canvases under a million lines, where a few tiles hold
    // most of the words. Both give the same runs.
    const bool stream_runs = words_cap > (uint64_t)8192 * std::max<uint32_t>(n_runs_work_, 1);
    if (!stream_runs) {
        k_tile_starts<<<(uint32_t)((n_cols + 255) / 256), 256, 0, stream_>>>(pp, global_rows, 0u);
        // four waves per work group, at least four columns per wave (a wave
        // clears its LDS cells once) and at most eight work groups per CU,
        // which all fit at once (1024x1024 circles 0.140 -> 0.073 ms,
        // 3840x2160 polygons 1.10 -> 1.09 ms against sixteen per CU)
        const uint32_t blocks = (uint32_t)std::min<uint64_t>((n_runs_work_ + 15) / 16, (uint64_t)cus_ * 8);
        if (blocks) k_tile_runs<<<blocks, 256, 0, stream_>>>(pp, runs_work_.as<uint2>(), n_runs_work_);
    } else {
        // runs over the whole sorted stream, then each tile's first run
        const uint64_t nparts = (words_cap + kRunsTileHost - 1) / kRunsTileHost;
        fill64(runs_status_.as<uint64_t>(), 0, std::max<uint64_t>(nparts, 1), stream_);
        // status words (8 B) then the open-run index (4 B) of each partition
        uint32_t *open = (uint32_t *)(runs_status_.as<uint64_t>() + std::max<uint64_t>(nparts, 1));
        if (nparts)
            k_runs_scan<<<(uint32_t)nparts, 256, 0, stream_>>>(pp, &fs->ctr, runs_status_.as<uint64_t>(),
                                                               fs->err + 1, open);
        k_run_cov<<<(uint32_t)std::max<uint64_t>(std::min<uint64_t>((nparts + 255) / 256, (uint64_t)cus_ * 4), 1), 256,
                    0, stream_>>>(pp, &fs->ctr, open);
        k_tile_starts<<<(uint32_t)((n_cols + 255) / 256), 256, 0, stream_>>>(pp, global_rows, 1u);
        k_tile_nr<<<(uint32_t)((n_cols + 255) / 256), 256, 0, stream_>>>(pp, global_rows);
    }
    MG_HIP(hipGetLastError(), "tile runs launch");
    MG_HIP(mark(evs_[5]), "event");

    // Row-walk LDS: a share per work group (one tile row); runs of one tile
    // up to rcap.
    const size_t lds_max = lds_max_;
    // A tile holds at most one run per layer order and the cover queue one
    // entry per order, so the walk never needs more than runs = orders and
    // queue + runs = 2 orders. Frames whose arrays sized that way fit in a
    // LDS share walk each row with 256 threads (several rows per CU when the
    // share is small); others use 512 threads and half the LDS, or the whole
    // LDS for deeper tiles (measured on MI355X: 256 threads are 12% faster for a
    // 2,000-order frame; a 65,536-order one below). The work-group size is a
    // result-invariant tuning knob (FrameSpec::walk_threads).
    const uint64_t orders = (uint64_t)spec.max_order + 1;
    const uint32_t rcap_fit = (uint32_t)((orders + 63) & ~63ull);
    const bool fits = orders <= 2048 && walk_lds_bytes_host(2 * rcap_fit, rcap_fit, tstride, 1) + 256 <= lds_max;
    uint32_t walk_threads = spec.walk_threads;
    if (walk_threads != 64 && walk_threads != 256 && walk_threads != 512 && walk_threads != 1024) {
        // Each row is a chain of dependent steps, so the fewer rows a frame
        // has per CU, the more threads a row takes. Measured on MI355X: a
        // 64x64 canvas of 330,000 lines (8 rows) walks in 0.34 ms with 1,024
        // threads, 0.49 with 512 (a 1024x1024 one, 128 rows: 0.32 against
        // 0.27 with 512); 3840x2160 with 200,000 polygons (270 rows)
        // 1.81 / 1.71 / 4.1 ms with 1,024 / 512 / 256 threads (256 need the
        // whole LDS there); a 1920x1080 animation (135 rows, small share)
        // 0.32 ms with 512, 0.36 with 256; frames of many more rows than CUs
        // gain nothing from more than 256 threads when their arrays fit a
        // small share.
        const uint32_t cus = (uint32_t)std::max(cus_, 1);
        if (global_rows <= cus / 4) walk_threads = 1024;
        else if (global_rows <= 4 * cus) walk_threads = 512;
        else walk_threads = fits ? 256 : 512;
    }
    const bool small_walk = fits;
    // The large configuration keeps two sets of run arrays so that the
    // next column's runs are stored while the current column is merged (one
    // barrier per column; 3840x2160 walk 2.73 -> 2.61 ms). The small configuration
    // keeps one: its LDS share sets how many rows a CU walks at once.
    // The large configuration first takes half the LDS (1,024 runs per tile,
    // a queue of about 1,700 entries), so that two rows share a CU: with one
    // row per CU, the rows beyond the CU count (270 rows on 256 CUs in a
    // 3840x2160 frame) wait for a whole row to finish (its walk 2.69 -> 2.03
    // ms). A frame that outgrows it is walked again with the whole LDS. The
    // walk reports the largest tile it met, and a frame whose previous render
    // (FrameSpec::walk_need_*) needed more than half the LDS starts with the
    // whole LDS; one that fits again returns to half.
    const uint32_t run_bufs = small_walk ? 1u : 2u;
    pp.walk_run_bufs = run_bufs;
    uint32_t cap, rcap;
    bool half_lds = false;
    auto large_caps = [&](bool half) {
        rcap = half ? 1024u : 2048u;
        const size_t budget = half ? lds_max / 2 - 512 : lds_max;
        const size_t fixed = walk_lds_bytes_host(0, rcap, tstride, run_bufs) + 256;
        if (fixed + 64 * 24 > budget) return false;
        cap = (uint32_t)std::min<size_t>(((budget - fixed) / 24) & ~size_t(63), 0xFFFFu - rcap);
        return true;
    };
    if (small_walk) {
        rcap = rcap_fit;
        cap = 2 * rcap;
    } else {
        half_lds = walk_threads >= 512 && large_caps(true) && spec.walk_need_runs <= rcap &&
                   spec.walk_need_ids <= cap;
        if (!half_lds && !large_caps(false)) {
            err = "canvas too wide for the row walk";
            return false;
        }
    }
    res.walk_threads = walk_threads;
    res.walk_lds = small_walk ? WALK_SHARE : (half_lds ? WALK_HALF : WALK_WHOLE);
    res.walk_rewalks = 0;
    size_t lds_w = walk_lds_bytes_host(cap, rcap, tstride, run_bufs), lds_p = paint_lds_bytes_host(cap, clips);
    if (lds_w > lds_max) {
        err = "canvas too wide for the row walk";
        return false;
    }
    bool global_walk = false;
    pp.cap_ids = cap;
    pp.cap_runs = rcap;
    // id records a row reserves at a time: about a row's worth of tiles
    {
        uint32_t chunk = 1024;
        while (chunk < 16384 && chunk < tstride * 16) chunk <<= 1;
        pp.rec_chunk = chunk;
    }
    if (rec_cap_ < tiles * 64) rec_cap_ = tiles * 64;
    for (int attempt = 0; attempt < 8; ++attempt) {
        if (rec_cap_ > 0xFFFF0000ull) rec_cap_ = 0xFFFF0000ull;
        if (!rec_.reserve(rec_cap_ * 16) || !rec_len_.reserve(rec_cap_ * 4)) {
            err = "out of device memory (tile id records)";
            return false;
        }
        pp.rec = rec_.as<uint4>();
        pp.rec_len = rec_len_.as<uint32_t>();
        pp.rec_cap = rec_cap_;
        // (the frame's first kernel cleared the flags; a walk again restarts
        // them, the record store and its counters)
        if (attempt > 0)
            zero_words(stream_, (uint32_t)cus_,
                       {{ov_dev, 3}, {ov_dev + 4, 3}, {ov_dev + kStatSlices, kStatSlices},
                        {spec.want_coverage ? cov_count_.ptr : nullptr, 1}});
        MG_HIP(mark(evs_[5]), "event");
        const size_t lds = global_walk ? 0 : lds_w;
        if (walk_threads == 64)
            global_walk ? k_row_walk<64, true><<<global_rows, 64, 0, stream_>>>(pp)
                        : k_row_walk<64, false><<<global_rows, 64, lds, stream_>>>(pp);
        else if (walk_threads == 256)
            global_walk ? k_row_walk<256, true><<<global_rows, 256, 0, stream_>>>(pp)
                        : k_row_walk<256, false><<<global_rows, 256, lds, stream_>>>(pp);
        else if (walk_threads == 512)
            global_walk ? k_row_walk<512, true><<<global_rows, 512, 0, stream_>>>(pp)
                        : k_row_walk<512, false><<<global_rows, 512, lds, stream_>>>(pp);
        else
            global_walk ? k_row_walk<1024, true><<<global_rows, 1024, 0, stream_>>>(pp)
                        : k_row_walk<1024, false><<<global_rows, 1024, lds, stream_>>>(pp);
        MG_HIP(mark(evs_[6]), "event");
        k_paint<<<n_work, 64, lds_p, stream_>>>(pp);
        if (pp.deep_min) k_paint_deep<<<pp.deep_cap, 256, paint_deep_lds_bytes_host(), stream_>>>(pp);
        MG_HIP(mark(evs_[7]), "event");
        MG_HIP(hipGetLastError(), "paint launch");
        if (capture_P_) return true;  // one attempt; the replay reads the flags
        MG_HIP(mark(ev1_), "event");
        if (!read_status(sizeof(FrameStatus), err)) return false;
        const uint32_t *ov = static_cast<const FrameStatus *>(status_host_)->ov;
        res.walk_max_runs = ov[5];
        res.walk_max_ids = ov[6];
        if (!ov[0]) {
            n_records_ = 0;
            for (uint32_t k = 0; k < kStatSlices; ++k) {
                n_records_ += ov[kStatSlices + k];
                n_runs_ += ov[2 * kStatSlices + k];
            }
            n_deep_ = std::min(ov[4], pp.deep_cap);
            status_final_ = true;
            return true;
        }
        if (ov[0] & 1u) {
            ++res.walk_rewalks;
            if (half_lds) {
                // outgrew half the LDS: walk again with the whole LDS
                half_lds = false;
                res.walk_lds = WALK_WHOLE;
                large_caps(false);
                pp.cap_ids = cap;
                pp.cap_runs = rcap;
                lds_w = walk_lds_bytes_host(cap, rcap, tstride, run_bufs);
                lds_p = paint_lds_bytes_host(cap, clips);
                continue;
            }
            // a tile holds more layers than the LDS arrays: walk again with
            // them in global memory, sized so that no tile can exceed them
            if (global_walk) break;
            const uint64_t orders = (uint64_t)spec.max_order + 1;
            rcap = (uint32_t)((orders + 63) & ~63ull);
            cap = 2 * rcap;
            lds_p = paint_lds_bytes_host(cap, clips);
            if (lds_p > lds_max) {
                err = "a tile holds more clipped layers than the painter's LDS capacity";
                return false;
            }
            const uint64_t stride = (walk_lds_bytes_host(cap, rcap, tstride, run_bufs) + 255) & ~255ull;
            if (!walk_scratch_.reserve(stride * global_rows)) {
                err = "out of device memory (row walk)";
                return false;
            }
            pp.walk_scratch = walk_scratch_.as<unsigned char>();
            pp.walk_stride = stride;
            pp.cap_ids = cap;
            pp.cap_runs = rcap;
            global_walk = true;
            res.walk_lds = WALK_GLOBAL;
            continue;
        }
        if (rec_cap_ >= 0xFFFF0000ull) break;
        // the walk kept counting after the store filled up: ov[1] covers the
        // frame's records plus at most one partly used chunk per row
        rec_cap_ = std::max<uint64_t>(rec_cap_ * 2, (uint64_t)ov[1] + (uint64_t)global_rows * pp.rec_chunk);
    }
    err = (rec_cap_ >= 0xFFFF0000ull) ? "too many tile id records" : "a tile holds more layers than the row walk's capacity";
    return false;
}

// Paint work list (the tiles inside each crop) and tile-runs work list
// (columns 0..h1 of the same rows), rebuilt when the scenes change.
bool Renderer::upload_work(const FrameSpec &spec, std::string &err) {
    if (work_scenes_.size() != spec.scenes.size() ||
        std::memcmp(work_scenes_.data(), spec.scenes.data(), spec.scenes.size() * sizeof(SceneDesc)) != 0) {
        // paint: the tiles inside each crop; tile runs (per-tile form):
        // columns 0..h1 (colp) of the same rows
        std::vector<uint32_t> work, rwork;
        for (uint32_t sc = 0; sc < spec.scenes.size(); ++sc) {
            const SceneDesc &s = spec.scenes[sc];
            const uint32_t h1 = std::min(s.crop_h1, s.row_len);
            if (s.crop_h0 >= h1) continue;
            for (uint32_t j = s.crop_v0; j < std::min(s.crop_v1, s.tile_rows); ++j) {
                for (uint32_t c = s.crop_h0; c < h1; ++c) {
                    work.push_back(s.row_base + j);
                    work.push_back(c);
                }
                for (uint32_t c = 0; c <= h1; ++c) {
                    rwork.push_back(s.row_base + j);
                    rwork.push_back(c);
                }
            }
        }
        n_work_ = (uint32_t)(work.size() / 2);
        n_runs_work_ = (uint32_t)(rwork.size() / 2);
        if (n_work_ && (!upload_vec(work_, work, stream_, err) || !upload_vec(runs_work_, rwork, stream_, err)))
            return false;
        work_scenes_ = spec.scenes;
    }
    return true;
}

// Every device buffer a frame's kernels address: a captured graph stays valid
// while none of them moved.
std::vector<const void *> Renderer::frame_buffers() const {
    const DevBuf *all[] = {&scenes_, &scene_line_base_, &layers_, &props_, &stops_, &row_scene_, &setup_,
                           &lengths_, &offsets_, &scan_status_, &status_, &segs_a_, &segs_b_, &dig_,
                           &row_start_, &part_line_,
                           &cnt0_, &cnt_acc_,
                           &cnt_status_, &work_, &runs_work_, &tile_nr_, &runs_status_, &tile_start_, &tile_run0_, &run_key_, &run_beg_,
                           &run_cov_, &rec_, &rec_len_, &rec_hdr_, &deep_list_, &part_kept_};
    std::vector<const void *> v;
    for (const DevBuf *b : all) v.push_back(b->ptr);
    return v;
}

void Renderer::drop_graph() {
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    graph_exec_ = nullptr;
    graph_ = nullptr;
}

bool Renderer::run_graph(const FrameSpec &spec, const float4 *pts, const uint16_t *ids, const uint32_t *ranks,
                         size_t n_lines,
                         uint8_t *image, FrameResult &res, bool &replayed, std::string &err) {
    replayed = false;
    MG_HIP(hipSetDevice(device_), "hipSetDevice");
    // everything the captured kernels take as arguments or grid sizes
    GraphKey key;
    key.pts = pts;
    key.ids = ids;
    key.ranks = ranks;
    key.image = image;
    key.n_lines = n_lines;
    key.image_bytes = spec.image_bytes;
    key.radix_bits = spec.radix_bits;
    key.max_order = spec.max_order;
    key.n_props = spec.props.size();
    key.mode = spec.mode;
    key.scenes = spec.scenes;
    if (graph_exec_ && key == graph_key_ && spec.mode == 0 && !spec.want_tile_ops && !spec.want_coverage) {
        uint32_t global_rows = 0;
        for (const SceneDesc &s : spec.scenes) global_rows = std::max(global_rows, s.row_base + s.tile_rows);
        // this frame's tables and work lists (outside the graph) into the
        // buffers it reads
        if (!upload_tables(spec, global_rows, err) || !upload_work(spec, err)) return false;
        if (frame_buffers() == graph_bufs_) {
            MG_HIP(hipEventRecord(ev0_, stream_), "event");
            MG_HIP(hipGraphLaunch(graph_exec_, stream_), "graph launch");
            MG_HIP(hipEventRecord(ev1_, stream_), "event");
            if (!read_status(sizeof(FrameStatus), err)) return false;
            const FrameStatus &hs = *static_cast<const FrameStatus *>(status_host_);
            const FrameCounters &hc = hs.ctr;
            const uint32_t *herr = hs.err, *ov = hs.ov;
            if (!herr[0] && !herr[1] && !herr[2] && !ov[0]) {
                float ms = 0.0f;
                MG_HIP(hipEventElapsedTime(&ms, ev0_, ev1_), "elapsed");
                res = graph_res_;
                res.line_pixels = hc.line_pixels;
                res.segments = hc.segments;
                res.kept = hc.kept;
                res.device_ms = ms;
                res.records = 0;
                res.runs = 0;
                for (uint32_t k = 0; k < kStatSlices; ++k) {
                    res.records += ov[kStatSlices + k];
                    res.runs += ov[2 * kStatSlices + k];
                }
                for (double &v : res.stage_ms) v = 0.0;
                for (double &v : res.pass_ms) v = 0.0;
                replayed = true;
                return true;
            }
            // beyond the captured capacities: render it again below, and capture anew
        }
        drop_graph();
    }
    drop_graph();
    if (!run(spec, pts, ids, ranks, n_lines, image, nullptr, res, err)) return false;
    if (spec.mode != 0 || spec.want_tile_ops || spec.want_coverage) return true;
    // capture the frame with 25% headroom over this one: a run with the
    // capacity sizes first (it sizes every buffer, so the capture allocates
    // nothing), then the captured run
    capture_P_ = std::max<uint64_t>(res.line_pixels + res.line_pixels / 4, 1 << 16);
    if (2 * capture_P_ >= (1ull << 30)) {
        capture_P_ = 0;
        return true;  // too large to capture with headroom: stays eager
    }
    rec_cap_ += rec_cap_ / 4;
    FrameResult tmp;
    std::string cerr;
    bool ok = run(spec, pts, ids, ranks, n_lines, image, nullptr, tmp, cerr);
    {
        // the capacity run must complete like the eager one (same image)
        std::string rerr;
        const FrameStatus &hs = *static_cast<const FrameStatus *>(status_host_);
        ok = ok && read_status(sizeof(FrameStatus), rerr) && !hs.err[0] && !hs.err[1] && !hs.err[2] && !hs.ov[0];
    }
    if (!ok) {
        capture_P_ = 0;
        (void)hipGetLastError();
        return run(spec, pts, ids, ranks, n_lines, image, nullptr, res, err);  // stays eager
    }
    hipGraph_t g = nullptr;
    if (ok) {
        ok = hipStreamBeginCapture(stream_, hipStreamCaptureModeRelaxed) == hipSuccess;
        if (ok) {
            const bool enq = run(spec, pts, ids, ranks, n_lines, image, nullptr, tmp, cerr);
            ok = hipStreamEndCapture(stream_, &g) == hipSuccess && enq;
        }
    }
    capture_P_ = 0;
    if (ok) ok = hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0) == hipSuccess;
    if (!ok) {
        if (g) (void)hipGraphDestroy(g);
        graph_exec_ = nullptr;
        (void)hipGetLastError();
        return true;  // rendered eagerly; later frames try to capture again
    }
    graph_ = g;
    graph_key_ = key;
    graph_bufs_ = frame_buffers();
    graph_res_ = res;
    return true;
}

bool Renderer::flatten_paths(const PathCommand *cmds, size_t n_cmds, const flat::PathDesc *paths, size_t n_paths,
                             LineStore &ls, size_t base, std::vector<uint32_t> &counts, std::string &err) {
    MG_HIP(hipSetDevice(device_), "hipSetDevice");
    counts.assign(n_paths, 0);
    if (!n_paths) return true;
    if (n_paths > 0xFFFFFFFFull) {
        err = "too many paths";
        return false;
    }
    const uint64_t nparts = (n_paths + kScanTileHost - 1) / kScanTileHost;
    if (!flat_cmds_.reserve(std::max<size_t>(n_cmds, 1) * sizeof(PathCommand)) ||
        !flat_paths_.reserve(n_paths * sizeof(flat::PathDesc)) || !flat_counts_.reserve(n_paths * 4) ||
        !flat_offsets_.reserve(n_paths * 4) || !flat_status_.reserve(nparts * 8)) {
        err = "out of device memory (paths)";
        return false;
    }
    if (n_cmds)
        MG_HIP(hipMemcpyAsync(flat_cmds_.ptr, cmds, n_cmds * sizeof(PathCommand), hipMemcpyHostToDevice, stream_),
               "upload commands");
    MG_HIP(hipMemcpyAsync(flat_paths_.ptr, paths, n_paths * sizeof(flat::PathDesc), hipMemcpyHostToDevice, stream_),
           "upload paths");
    FrameStatus *fs = status_.as<FrameStatus>();
    FrameCounters *ctr = &fs->ctr;
    zero_words(stream_, (uint32_t)cus_, {{fs, kStatusHead / 4}, {flat_status_.ptr, 2 * (uint64_t)nparts}});
    const uint32_t grid = (uint32_t)((n_paths + 255) / 256);
    k_path_counts<<<grid, 256, 0, stream_>>>(flat_cmds_.as<PathCommand>(), flat_paths_.as<flat::PathDesc>(),
                                             (uint32_t)n_paths, flat_counts_.as<uint32_t>(), fs->err);
    k_scan_u32<<<(uint32_t)nparts, 256, 0, stream_>>>(flat_counts_.as<uint32_t>(), flat_offsets_.as<uint32_t>(),
                                                    n_paths, flat_status_.as<uint64_t>(), fs->err + 1,
                                                    &ctr->line_pixels);
    MG_HIP(hipGetLastError(), "flatten count launch");
    if (!read_status(kStatusHead, err)) return false;
    const FrameStatus &hs = *static_cast<const FrameStatus *>(status_host_);
    const uint64_t total = hs.ctr.line_pixels;
    const uint32_t herr[2] = {hs.err[0], hs.err[1]};
    if (herr[0] || herr[1] || base + total > 0xFFFFFFFFull) {
        err = herr[1] ? "a look-back wait timed out" : "too many lines";
        return false;
    }
    if (!ls.grow(base + total, base, stream_, err)) return false;
    k_path_emit<<<grid, 256, 0, stream_>>>(flat_cmds_.as<PathCommand>(), flat_paths_.as<flat::PathDesc>(),
                                           (uint32_t)n_paths, flat_offsets_.as<uint32_t>(),
                                           flat_counts_.as<uint32_t>(), ls.pts.as<float4>() + base,
                                           ls.ids.as<uint16_t>() + base);
    MG_HIP(hipGetLastError(), "flatten launch");
    MG_HIP(hipMemcpyAsync(counts.data(), flat_counts_.ptr, n_paths * 4, hipMemcpyDeviceToHost, stream_),
           "read counts");
    MG_HIP(hipStreamSynchronize(stream_), "flatten");
    return true;
}

bool Renderer::stream_copy(size_t bytes, uint32_t repeats, double &ms, uint64_t &mismatches, std::string &err) {
    MG_HIP(hipSetDevice(device_), "hipSetDevice");
    DevBuf src, dst, bad;
    struct Release {
        DevBuf *b[3];
        ~Release() {
            for (DevBuf *x : b) x->release();
        }
    } release{{&src, &dst, &bad}};
    constexpr uint64_t kGuardWords = 64;  // past the copy in dst: must keep their own pattern
    const uint64_t n = bytes / 16, words = bytes / 4;
    if (!src.reserve(bytes) || !dst.reserve(bytes + kGuardWords * 4) || !bad.reserve(8)) {
        err = "out of device memory";
        return false;
    }
    const uint32_t seed = 0x5EEDu;
    k_copy_pattern<<<2048, 256, 0, stream_>>>(src.as<uint32_t>(), words, seed);
    k_copy_pattern<<<2048, 256, 0, stream_>>>(dst.as<uint32_t>(), words + kGuardWords, seed ^ 1u);
    fill64(bad.as<uint64_t>(), 0, 1, stream_);
    // 16,384 work groups (64 per CU) of 4 x 16 B per thread and pass
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(16384, (n + 1023) / 1024));
    MG_HIP(hipEventRecord(ev0_, stream_), "event");
    for (uint32_t r = 0; r < repeats; ++r)
        k_copy_stream<<<grid, 256, 0, stream_>>>(src.as<uint4>(), dst.as<uint4>(), n);
    MG_HIP(hipGetLastError(), "copy launch");
    MG_HIP(hipEventRecord(ev1_, stream_), "event");
    k_copy_check<<<2048, 256, 0, stream_>>>(dst.as<uint32_t>(), 0, words, seed, bad.as<unsigned long long>());
    k_copy_check<<<1, 256, 0, stream_>>>(dst.as<uint32_t>(), words, words + kGuardWords, seed ^ 1u,
                                         bad.as<unsigned long long>());
    MG_HIP(hipGetLastError(), "check launch");
    unsigned long long nbad = 0;
    MG_HIP(hipMemcpyAsync(&nbad, bad.ptr, 8, hipMemcpyDeviceToHost, stream_), "read check");
    MG_HIP(hipStreamSynchronize(stream_), "copy");
    float t = 0.0f;
    MG_HIP(hipEventElapsedTime(&t, ev0_, ev1_), "elapsed");
    ms = (double)t / repeats;
    mismatches = nbad;
    return true;
}

bool Renderer::download_sorted(uint64_t *dst, size_t n, std::string &err) {
    if (!n) return true;
    MG_HIP(hipMemcpyAsync(dst, last_sorted_, n * 8, hipMemcpyDeviceToHost, stream_), "download sorted");
    MG_HIP(hipStreamSynchronize(stream_), "sync");
    return true;
}

bool Renderer::download_tile_ops(int8_t *dst, size_t n, std::string &err) {
    if (!n) return true;
    MG_HIP(hipMemcpyAsync(dst, tile_ops_.ptr, n, hipMemcpyDeviceToHost, stream_), "download tile ops");
    MG_HIP(hipStreamSynchronize(stream_), "sync");
    return true;
}

bool Renderer::download_coverage(std::vector<int32_t> &a, std::vector<uint32_t> &key, uint32_t &count,
                                 std::string &err) {
    MG_HIP(hipMemcpyAsync(&count, cov_count_.ptr, 4, hipMemcpyDeviceToHost, stream_), "cov count");
    MG_HIP(hipStreamSynchronize(stream_), "sync");
    size_t n = std::min<size_t>(count, cov_a_.bytes / 256);
    a.resize(n * 64);
    key.resize(n * 4);
    if (n) {
        MG_HIP(hipMemcpyAsync(a.data(), cov_a_.ptr, n * 256, hipMemcpyDeviceToHost, stream_), "cov a");
        MG_HIP(hipMemcpyAsync(key.data(), cov_key_.ptr, n * 16, hipMemcpyDeviceToHost, stream_), "cov key");
        MG_HIP(hipStreamSynchronize(stream_), "sync");
    }
    return true;
}

}  // namespace mg
