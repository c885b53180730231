// Device runtime of the MI355X mold renderer: resident buffers, frame
// execution (setup -> scan -> emit -> sort -> paint) and timing.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "mg_flatten.h"
#include "mg_shared.h"

namespace mg {

constexpr uint32_t kMaxSortPasses = 16;  // digit passes of the widest sort key

// A fixed block of device memory handed out front to back; nothing is freed
// until reset. An engine in arena mode takes every buffer from one.
struct Arena {
    unsigned char *base = nullptr;
    size_t size = 0, used = 0;
    size_t short_by = 0;  // bytes a refused request lacked (0: none refused)
    void *take(size_t n);
};

// Growable device allocation (hipMalloc), or a block of an arena when
// `arena` is set (no device allocation; the block is dropped, not freed).
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    Arena *arena = nullptr;
    bool reserve(size_t n);  // keeps contents only when not growing
    void release();
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
};

// Device-resident lines of one composition or batch: float4 points and u16
// layer ids, appended as the caller inserts them; and each line's rank among
// the lines of its layer id (the first `ranked` lines ranked; `counts`: lines
// ranked per layer id, `count_ids` entries).
struct LineStore {
    DevBuf pts, ids;
    size_t cap = 0;
    DevBuf rank, counts;
    size_t ranked = 0;
    uint32_t count_ids = 0;
    bool upload(const float *pts4, const uint16_t *ids, size_t first, size_t n, hipStream_t s, std::string &err);
    // Capacity for `need` lines, keeping the first `keep`.
    bool grow(size_t need, size_t keep, hipStream_t s, std::string &err);
    void release();
    const float4 *points() const { return pts.as<float4>(); }
    const uint16_t *layer_ids() const { return ids.as<uint16_t>(); }
};

// Per-tile previous layer counts of the BufferLayerCaches of one
// composition (-1 = None).
struct LayerCaches {
    std::vector<DevBuf> bufs;
    std::vector<size_t> tiles;
    int32_t *array(int id, size_t n_tiles, hipStream_t s, bool &resized, std::string &err);
    void clear(int id, hipStream_t s);
    void release(int id);
    void release_all();
};

// Host-side description of one frame (a batch of scenes).
struct FrameSpec {
    std::vector<SceneDesc> scenes;
    std::vector<uint32_t> scene_line_base;
    std::vector<LayerEntry> layers;     // concatenated per scene
    std::vector<PropsEntry> props;      // concatenated per scene
    std::vector<float> stops;
    uint32_t max_order = 0;             // over all scenes
    uint32_t mode = 0;                  // 0 render, 1 canonical stream, 2 render an ingested stream
    const uint64_t *in_segs = nullptr;  // mode 2: device PixelSegment stream
    uint64_t n_in = 0;
    uint32_t radix_bits = 0;            // sort digit width; 0 = automatic
    uint32_t walk_threads = 0;          // row-walk work-group size (64, 256, 512, 1024); 0 = automatic
    // what the row walk of this frame's previous render needed (FrameResult::
    // walk_max_*), 0 = unknown: a large-frame walk that needed more than half
    // the LDS starts with the whole LDS
    uint32_t walk_need_runs = 0, walk_need_ids = 0;
    bool want_tile_ops = false;
    bool want_coverage = false;
    uint32_t coverage_cap = 0;
    // layers/props/stops/scenes are those of the last upload when both match
    uint64_t tables_owner = 0;
    uint64_t tables_version = 0;
    uint64_t image_bytes = 0;           // size of the output image (every scene's frame must fit)
};

struct FrameResult {
    uint64_t line_pixels = 0, segments = 0, kept = 0;
    uint32_t passes = 0, radix_bits = 0, sort_bits = 0, walk_threads = 0;
    uint32_t deep_tiles = 0;  // tiles painted by the four-wave deep painter
    // row walk: LDS configuration of the final walk (WalkLds), walks repeated
    // for LDS capacity, and the largest runs per tile and queue + runs it met
    uint32_t walk_lds = 0, walk_rewalks = 0;
    uint32_t walk_max_runs = 0, walk_max_ids = 0;
    uint64_t tiles = 0;
    double device_ms = 0.0;
    // setup+scan, emit, histogram+bases, sort passes, tile runs, row walk, paint
    double stage_ms[7] = {0, 0, 0, 0, 0, 0, 0};
    double pass_ms[kMaxSortPasses] = {};  // each sort pass kernel alone
    uint64_t runs = 0, records = 0;     // tile runs and tile id records of the painted frame
    const uint64_t *sorted = nullptr;   // device pointer to the sorted words (mode 1)
};

// LDS configuration of the row walk: a share of the LDS (several rows per
// CU), half the LDS (two rows per CU), the whole LDS, or arrays in global
// memory for tiles deeper than LDS holds.
enum WalkLds : uint32_t { WALK_SHARE = 0, WALK_HALF = 1, WALK_WHOLE = 2, WALK_GLOBAL = 3 };

// Bit layout of a frame's render key: (row, column + 1, layer order) above
// the 22-bit payload, and the bits the LSD sort covers.
struct KeyLayout {
    uint32_t col_bits, layer_bits, key_bits, sort_shift, sort_bits;
};
KeyLayout key_layout(const FrameSpec &spec);

// The frame engine of one device: its stream, the scratch buffers of a frame
// and the uploaded frame tables. Compositions and batches on a device share
// one engine (Renderer::shared); `mu` serialises their frames, and scratch
// contents (sorted words, tile ops, coverage) belong to the last frame run.
class Renderer {
public:
    Renderer();
    ~Renderer();
    bool init(std::string &err);
    static std::shared_ptr<Renderer> shared(std::string &err);
    // An engine whose buffers all come from `bytes` of device memory at
    // `base` (no hipMalloc or hipFree, ever); its frames start with
    // arena_begin(). One per device, shared by the callers that cannot
    // allocate: `device` is the device of the shared engine.
    static std::shared_ptr<Renderer> shared_arena(void *base, size_t bytes, std::string &err);
    // Drops every buffer of the engine (arena mode): the next frame takes
    // them from the arena's start. Cached tables are uploaded again.
    void arena_begin();
    // `bytes` of the arena for the caller's frame inputs (nullptr when the
    // arena cannot hold them); valid until the next arena_begin().
    void *arena_take(size_t bytes) { return arena_.take(bytes); }
    const Arena &arena() const { return arena_; }

    std::mutex mu;

    // Frame over `n_lines` device-resident lines (float4 points, u16 layer
    // ids; unused by ingest frames) into the device image `image`.
    bool run(const FrameSpec &spec, const float4 *pts, const uint16_t *ids, const uint32_t *ranks, size_t n_lines,
             uint8_t *image, int32_t *prev_layers, FrameResult &res, std::string &err);
    // A render frame through a captured hipGraph: replayed while everything
    // its kernels take as arguments (lines, image, canvas, key layout) is
    // unchanged, with the frame's tables uploaded beside it. The first frame
    // of a shape, and any frame beyond the captured capacities, runs eagerly
    // and captures anew (`replayed` false).
    bool run_graph(const FrameSpec &spec, const float4 *pts, const uint16_t *ids, const uint32_t *ranks,
                   size_t n_lines, uint8_t *image, FrameResult &res, bool &replayed, std::string &err);
    // Ranks of lines [first, n) among the lines of their layer ids (ids
    // below n_ids): counts[id] holds the lines of id ranked so far, cleared
    // first when `reset`.
    bool rank_lines(const uint16_t *ids, uint32_t *counts, uint32_t n_ids, uint32_t *rank, size_t first, size_t n,
                    bool reset, std::string &err);
    // Path flattening on the device: the paths' lines appended to `ls` after
    // its first `base` lines, path after path; `counts` receives each
    // path's number of lines.
    bool flatten_paths(const PathCommand *cmds, size_t n_cmds, const flat::PathDesc *paths, size_t n_paths,
                       LineStore &ls, size_t base, std::vector<uint32_t> &counts, std::string &err);
    // `repeats` streaming copies of `bytes` (a multiple of 16) between two
    // buffers of its own on the engine's stream: `ms` receives the average
    // time of one (events), `mismatches` the 32-bit words of the result that
    // differ from the source, or past the copy from what they held before.
    bool stream_copy(size_t bytes, uint32_t repeats, double &ms, uint64_t &mismatches, std::string &err);
    bool download_sorted(uint64_t *dst, size_t n, std::string &err);
    bool download_tile_ops(int8_t *dst, size_t n, std::string &err);
    bool download_coverage(std::vector<int32_t> &a, std::vector<uint32_t> &key, uint32_t &count,
                           std::string &err);
    hipStream_t stream() const { return stream_; }
    int device() const { return device_; }
    int cu_count() const { return cus_; }

private:
    struct GraphKey {
        const void *pts = nullptr, *ids = nullptr, *ranks = nullptr, *image = nullptr;
        size_t n_lines = 0;
        uint64_t image_bytes = 0;
        uint32_t radix_bits = 0, max_order = 0, mode = 0;
        size_t n_props = 0;
        std::vector<SceneDesc> scenes;
        bool operator==(const GraphKey &o) const {
            return pts == o.pts && ids == o.ids && ranks == o.ranks && image == o.image && n_lines == o.n_lines &&
                   image_bytes == o.image_bytes && radix_bits == o.radix_bits && max_order == o.max_order &&
                   mode == o.mode && n_props == o.n_props && scenes.size() == o.scenes.size() &&
                   std::memcmp(scenes.data(), o.scenes.data(), scenes.size() * sizeof(SceneDesc)) == 0;
        }
    };
    hipError_t mark(hipEvent_t e) { return capture_P_ ? hipSuccess : hipEventRecord(e, stream_); }
    bool upload_tables(const FrameSpec &spec, uint32_t global_rows, std::string &err);
    bool upload_work(const FrameSpec &spec, std::string &err);
    std::vector<const void *> frame_buffers() const;
    void drop_graph();
    bool paint(const FrameSpec &spec, uint32_t global_rows, uint32_t col_bits, uint32_t layer_bits,
               const uint64_t *sorted, uint64_t words_cap, uint8_t *image, int32_t *prev_layers, size_t tiles,
               FrameResult &res, std::string &err);

    std::vector<DevBuf *> all_buffers();
    Arena arena_;
    bool arena_mode_ = false;
    int device_ = -1;
    int cus_ = 0;
    size_t lds_max_ = 65536;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    hipEvent_t evs_[8] = {};
    hipEvent_t evp_[2 * kMaxSortPasses] = {};  // before and after each sort pass kernel
    uint64_t uploaded_owner_ = 0, uploaded_version_ = 0;
    std::vector<SceneDesc> uploaded_scenes_;
    DevBuf scenes_, scene_line_base_, layers_, props_, stops_;
    DevBuf setup_, lengths_, offsets_, scan_status_;
    // the frame's status block (counters, error flags, paint flags), its host
    // copy (pinned memory; in arena mode the engine's own `status_local_`) and
    // whether the host copy is of the frame's end
    DevBuf status_;
    void *status_host_ = nullptr;
    bool status_pinned_ = false, status_final_ = false;
    alignas(16) unsigned char status_local_[512] = {};
    bool read_status(size_t bytes, std::string &err);
    DevBuf segs_a_, segs_b_, dig_;  // ping-pong words; the next sort pass's digits
    DevBuf row_start_, row_scene_;
    DevBuf tile_ops_, cov_a_, cov_key_, cov_count_;
    DevBuf part_line_;
    DevBuf cnt0_, cnt_acc_, cnt_status_;
    DevBuf flat_cmds_, flat_paths_, flat_counts_, flat_offsets_, flat_status_;
    DevBuf work_, runs_work_, tile_nr_, runs_status_, walk_scratch_, tile_start_, tile_run0_, run_key_, run_beg_, run_cov_, rec_, rec_len_, rec_hdr_;
    DevBuf deep_list_, part_kept_;
    uint64_t rec_cap_ = 0;  // id records of the last frame's size
    std::vector<SceneDesc> work_scenes_;  // scenes of the resident paint work list
    uint32_t n_work_ = 0, n_runs_work_ = 0;
    uint64_t n_records_ = 0, n_runs_ = 0;
    uint32_t n_deep_ = 0;
    const uint64_t *last_sorted_ = nullptr;
    uint64_t capture_P_ = 0;  // nonzero while a frame is enqueued for capture: its line-pixel capacity
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    GraphKey graph_key_;
    std::vector<const void *> graph_bufs_;
    FrameResult graph_res_;
};

// Number of usable gfx950 devices.
int device_count();
// The device an engine renders on: the first gfx950 device (-1: none).
int render_device();

}  // namespace mg
