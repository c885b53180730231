/*
 * mold_gpu.h -- C ABI of the MI355X-native mold/surpass renderer.
 *
 * The entry points mirror mold's public scene API (mold/src/lib.rs:15-25) as
 * opaque handles plus plain pointers and sizes, so a Rust, Python or C caller
 * can bind them directly (the reference's only FFI precedent is the dlopen'd
 * cassia ABI, cassia/src/Cassia.h:32-41, bound by demo/src/cassia.rs:15-56;
 * its three entry points are reproduced at the end of this file).
 *
 * Conventions
 *   - Every function returns an mg_status (MG_OK == 0). No exception or panic
 *     crosses the ABI; conditions where the reference panics return an error
 *     code and mg_last_error() describes them.
 *   - Colours are linear f32 RGBA; buffers are RGBA8, row-major, 4 bytes per
 *     pixel, `width_stride` pixels per row.
 *   - A composition is single-owner (the reference's Composition is !Send,
 *     composition.rs:33): use one per thread or serialise calls.
 *   - Rendering runs on the GPU (gfx950). There is no CPU fallback: without a
 *     usable device, mg_render returns MG_ERR_DEVICE.
 */
#ifndef MOLD_GPU_H
#define MOLD_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define MG_EXPORT __attribute__((visibility("default")))
#else
#define MG_EXPORT
#endif

typedef enum mg_status {
    MG_OK = 0,
    MG_ERR_INVALID_ARG = 1,      /* bad pointer / size / enum value                     */
    MG_ERR_LAYER_SCALE = 2,      /* Layer::set_transform scale check (layer.rs:186-190)  */
    MG_ERR_IDS_EXHAUSTED = 3,    /* create_layer -> None (layer.rs:61-72)                */
    MG_ERR_NO_LAYER = 4,         /* unknown layer id                                     */
    MG_ERR_OOM = 5,              /* host or device allocation failed                     */
    MG_ERR_DEVICE = 6,           /* no gfx950 device, or a HIP call failed               */
    MG_ERR_BUFFER = 7,           /* buffer/width/stride mismatch (buffer_layout.rs:81-92,
                                    180-185)                                             */
    MG_ERR_GRADIENT = 8,         /* gradient stop outside [0, 1] (style.rs:58-60)        */
    MG_ERR_CAPACITY = 9,         /* a tile exceeds the painter's layer capacity          */
    MG_ERR_NO_CACHE = 10,        /* all 32 buffer layer caches in use (layer.rs:137-141) */
} mg_status;

/* Last error message of the calling thread ("" if none). */
MG_EXPORT const char *mg_last_error(void);

/* Number of usable gfx950 devices (0 when none). */
MG_EXPORT int mg_device_count(void);

/* ---------------------------------------------------------------- styles */

typedef enum mg_fill_rule { MG_NONZERO = 0, MG_EVENODD = 1 } mg_fill_rule;   /* style.rs:9-19 */

typedef enum mg_blend_mode {                                                  /* style.rs:189-203 */
    MG_BLEND_OVER = 0, MG_BLEND_MULTIPLY, MG_BLEND_SCREEN, MG_BLEND_OVERLAY, MG_BLEND_DARKEN,
    MG_BLEND_LIGHTEN, MG_BLEND_COLOR_DODGE, MG_BLEND_COLOR_BURN, MG_BLEND_HARD_LIGHT,
    MG_BLEND_SOFT_LIGHT, MG_BLEND_DIFFERENCE, MG_BLEND_EXCLUSION
} mg_blend_mode;

typedef enum mg_func { MG_FUNC_DRAW = 0, MG_FUNC_CLIP = 1 } mg_func;          /* painter/mod.rs:125-135 */
typedef enum mg_fill_kind { MG_FILL_SOLID = 0, MG_FILL_GRADIENT = 1 } mg_fill_kind;
typedef enum mg_gradient_type { MG_GRADIENT_LINEAR = 0, MG_GRADIENT_RADIAL = 1 } mg_gradient_type;

/* surpass::painter::Props (painter/mod.rs:137-141) with its Style
   (style.rs:426-432) flattened. Gradient stops are (r, g, b, a, stop)
   quintuples; a stop of -0.0 (or +0.0) is resolved to i / (n - 1) as
   GradientBuilder::build does (style.rs:66-84). */
typedef struct mg_props {
    uint32_t fill_rule;        /* mg_fill_rule, the painter's rule            */
    uint32_t func;             /* mg_func                                     */
    uint32_t clip_layers;      /* Func::Clip(n)                               */
    uint32_t is_clipped;       /* Style::is_clipped                           */
    uint32_t style_fill_rule;  /* Style::fill_rule (kept, unused by painting) */
    uint32_t blend_mode;       /* mg_blend_mode                               */
    uint32_t fill_kind;        /* mg_fill_kind                                */
    float color[4];            /* Fill::Solid                                 */
    uint32_t gradient_type;    /* mg_gradient_type                            */
    float gradient_start[2];
    float gradient_end[2];
    uint32_t n_stops;          /* >= 2 for a gradient fill                    */
    const float *stops;        /* n_stops x 5 floats                          */
} mg_props;

/* ------------------------------------------------------------------ paths */

typedef struct mg_path mg_path;

/* Path::new (mold/src/path/mod.rs:38-40) */
MG_EXPORT mg_status mg_path_create(mg_path **out);
MG_EXPORT void mg_path_destroy(mg_path *path);
/* Path::line(p0, p1): pts = {x0, y0, x1, y1} (path/mod.rs:53-55) */
MG_EXPORT mg_status mg_path_line(mg_path *path, const float pts[4]);
/* Path::quad: {x0,y0,x1,y1,x2,y2} (path/mod.rs:69-71) */
MG_EXPORT mg_status mg_path_quad(mg_path *path, const float pts[6]);
/* Path::rat_quad: {x0,y0,w0, x1,y1,w1, x2,y2,w2} (path/mod.rs:86-97) */
MG_EXPORT mg_status mg_path_rat_quad(mg_path *path, const float pts[9]);
/* Path::cubic: 4 points (path/mod.rs:112-114) */
MG_EXPORT mg_status mg_path_cubic(mg_path *path, const float pts[8]);
/* Path::rat_cubic: 4 weighted points (path/mod.rs:130-143) */
MG_EXPORT mg_status mg_path_rat_cubic(mg_path *path, const float pts[12]);
/* Path::close (path/mod.rs:168-170) */
MG_EXPORT mg_status mg_path_close(mg_path *path);
/* Flatten to line segments (PathSegments, path/segments.rs:233-263); `t9`
   may be NULL. Writes at most `cap` segments (4 floats each) and returns the
   total count in *n. */
MG_EXPORT mg_status mg_path_flatten(const mg_path *path, const float *t9, float *out, size_t cap,
                                    size_t *n);

/* ------------------------------------------------------------ composition */

typedef struct mg_composition mg_composition;
typedef uint16_t mg_layer_id;

/* Composition::new (composition.rs:36-48) */
MG_EXPORT mg_status mg_composition_create(mg_composition **out);
MG_EXPORT void mg_composition_destroy(mg_composition *comp);
/* Composition::create_layer -> Option<LayerId> (composition.rs:54-56);
   MG_ERR_IDS_EXHAUSTED for None. Ids are handed out smallest-free-first. */
MG_EXPORT mg_status mg_create_layer(mg_composition *comp, mg_layer_id *out);
/* Composition::insert_in_layer (composition.rs:82-84) */
MG_EXPORT mg_status mg_insert_in_layer(mg_composition *comp, mg_layer_id id, const mg_path *path);
/* Composition::insert_in_layer_transformed (composition.rs:86-94) */
MG_EXPORT mg_status mg_insert_in_layer_transformed(mg_composition *comp, mg_layer_id id,
                                                   const mg_path *path, const float t9[9]);
/* Bulk form of insert_in_layer for a path made only of `n` Path::line
   commands: lines = n x {x0, y0, x1, y1}. Same effect as the path form
   (composition.rs:58-79). */
MG_EXPORT mg_status mg_insert_lines(mg_composition *comp, mg_layer_id id, const float *lines,
                                    size_t n);
/* Bulk form over many layers: lines[i] goes to ids[i]; each distinct layer
   is treated as an insert_in_layer call (order reset to its id). */
MG_EXPORT mg_status mg_insert_lines_multi(mg_composition *comp, const mg_layer_id *ids,
                                          const float *lines, size_t n);

/* Bulk Composition::insert_in_layer / insert_in_layer_transformed
   (composition.rs:58-95) with the flattening (PathSegments,
   mold/src/path/segments.rs:51-263, transform.rs:57-93) on the GPU: path i
   goes to layer ids[i], transformed by t9s[9 i .. 9 i + 9] when t9s is not
   NULL and transformed (NULL = all) has transformed[i] != 0. The lines are
   written straight into the composition's resident line store, path after
   path, in the order the host path form gives them; the host keeps no copy
   until one is needed (mg_lines, host inserts, compaction). */
MG_EXPORT mg_status mg_insert_paths_device(mg_composition *comp, const mg_path *const *paths,
                                           const mg_layer_id *ids, const float *t9s, const uint8_t *transformed,
                                           size_t n);
/* The same with the flattening on the host (one insert_in_layer per path). */
MG_EXPORT mg_status mg_insert_paths(mg_composition *comp, const mg_path *const *paths, const mg_layer_id *ids,
                                    const float *t9s, const uint8_t *transformed, size_t n);
/* The composition's lines (LinesBuilder contents, segment.rs:31-53):
   n x {x0, y0, x1, y1} and their layer ids, at most `cap`; *n receives the
   total. */
MG_EXPORT mg_status mg_lines(mg_composition *comp, float *pts, uint16_t *ids, size_t cap, size_t *n);

/* Layer accessors (mold/src/layer.rs:152-231). MG_ERR_NO_LAYER when the id
   has no layer entry (Composition::get_mut -> None). */
MG_EXPORT mg_status mg_layer_exists(const mg_composition *comp, mg_layer_id id, int *out);
MG_EXPORT mg_status mg_layer_is_enabled(const mg_composition *comp, mg_layer_id id, int *out);
MG_EXPORT mg_status mg_layer_set_enabled(mg_composition *comp, mg_layer_id id, int enabled);
MG_EXPORT mg_status mg_layer_transform(const mg_composition *comp, mg_layer_id id, float out[6]);
/* Layer::set_transform: MG_ERR_LAYER_SCALE where the reference panics */
MG_EXPORT mg_status mg_layer_set_transform(mg_composition *comp, mg_layer_id id, const float t[6]);
MG_EXPORT mg_status mg_layer_order(const mg_composition *comp, mg_layer_id id, uint16_t *out);
MG_EXPORT mg_status mg_layer_set_order(mg_composition *comp, mg_layer_id id, uint16_t order);
MG_EXPORT mg_status mg_layer_props(const mg_composition *comp, mg_layer_id id, mg_props *out);
MG_EXPORT mg_status mg_layer_set_props(mg_composition *comp, mg_layer_id id, const mg_props *props);
/* Bulk forms for scenes with many layers: n calls of create_layer (ids out,
   stops at the first failure) and of Layer::set_props (layer.rs:224-231) on
   ids[i] with props[i]. */
MG_EXPORT mg_status mg_create_layers(mg_composition *comp, size_t n, mg_layer_id *out);
MG_EXPORT mg_status mg_layers_set_props(mg_composition *comp, const mg_layer_id *ids, const mg_props *props,
                                        size_t n);
/* Bulk Layer::set_transform (layer.rs:182-201; t6 = n x 6 floats) and
   Layer::set_order (layer.rs:209-216). */
MG_EXPORT mg_status mg_layers_set_transform(mg_composition *comp, const mg_layer_id *ids, const float *t6, size_t n);
MG_EXPORT mg_status mg_layers_set_order(mg_composition *comp, const mg_layer_id *ids, const uint16_t *orders,
                                        size_t n);
/* Composition::remove_disabled (composition.rs:123-142) */
MG_EXPORT mg_status mg_remove_disabled(mg_composition *comp);
/* Lines held by the builder / lines of enabled layers (composition.rs:116-121) */
MG_EXPORT mg_status mg_lines_len(const mg_composition *comp, size_t *builder_len,
                                 size_t *actual_len);

/* Composition::create_buffer_layer_cache -> Option<BufferLayerCache>
   (composition.rs:144-151). */
MG_EXPORT mg_status mg_create_buffer_layer_cache(mg_composition *comp, int32_t *cache_id);
/* BufferLayerCache::clear (buffer.rs:46-51) */
MG_EXPORT mg_status mg_buffer_layer_cache_clear(mg_composition *comp, int32_t cache_id);
/* Drop of the last BufferLayerCache clone (buffer.rs:53-63) */
MG_EXPORT mg_status mg_buffer_layer_cache_release(mg_composition *comp, int32_t cache_id);

/* Rect::new(horizontal, vertical) in pixels (buffer_layout.rs:43-50). */
typedef struct mg_rect {
    uint32_t x0, x1, y0, y1;
} mg_rect;

/* Called once per written tile slice (row of <= 8 pixels) after the frame is
   in the caller's buffer (Flusher::flush, buffer_layout.rs:52-54). */
typedef void (*mg_flusher_fn)(void *user, uint8_t *slice_rgba, size_t n_pixels);

/* mold::Buffer (buffer.rs:17-25). `len_px` = buffer.len() in pixels. */
typedef struct mg_buffer {
    uint8_t *rgba;            /* host pointer, len_px * 4 bytes          */
    size_t len_px;
    size_t width;
    size_t width_stride;      /* 0 = width                               */
    int32_t layer_cache;      /* -1 = none                               */
    mg_flusher_fn flusher;    /* may be NULL                             */
    void *flusher_user;
} mg_buffer;

/* Composition::render(buffer, background_color, crop) (composition.rs:212-328).
   Blocking: returns when the caller's buffer holds the frame. `crop` may be
   NULL. */
MG_EXPORT mg_status mg_render(mg_composition *comp, const mg_buffer *buffer,
                              const float background[4], const mg_rect *crop);

/* Composition::render into a device-resident RGBA8 frame of width x height
   (stride = width): no host copies, no layer cache, no flusher. `*device_rgba`
   receives the device pointer of the frame, valid until the next render on
   `comp`. For callers that consume frames on the GPU, and for timing the
   path with its inputs and output resident in HBM. */
MG_EXPORT mg_status mg_render_device(mg_composition *comp, uint32_t width, uint32_t height,
                                     const float background[4], const mg_rect *crop,
                                     void **device_rgba);

/* mg_render_device for callers that cannot allocate device memory (tight
   serving loops, harnesses that forbid hipMalloc). No counterpart in the
   reference. Renders `comp` into the
   caller's device frame `device_rgba` (`device_bytes` >= width * height * 4;
   stride = width; tiles outside `crop` keep their contents) and calls no
   hipMalloc, hipFree or allocating library routine: every buffer of the
   frame comes from a static device array of 64 MiB per device. The
   composition's lines are read in place when they are resident on the device
   and current (an earlier mg_render* call uploaded them), else copied into
   that array. Frames that need more than it holds (roughly: 18 bytes per
   line, 32 bytes per line-pixel, the layer and props tables and about 1.3
   KB per tile) return MG_ERR_CAPACITY. Blocking: the frame is complete on
   return. The array keeps the last call's contents; calls serialise on one
   engine per device (calls that overlapped on one GPU would race on the
   array). The first call on a device looks up the array's address and
   creates the engine's stream and events. */
MG_EXPORT mg_status mg_render_device_noalloc(mg_composition *comp, uint32_t width, uint32_t height,
                                             const float background[4], const mg_rect *crop, void *device_rgba,
                                             size_t device_bytes);

/* mg_render_device through a captured hipGraph, for animations: while the
   composition's lines, the canvas and the layer count stay the same, each
   frame uploads its layer tables (transforms, orders, props) and replays the
   graph captured from an earlier frame. The first frame of a shape, and a
   frame whose pixel work exceeds the captured capacities, render eagerly and
   capture anew; `*replayed` (may be NULL) says which. Same bytes as
   mg_render_device. No counterpart in the reference (composition.rs:212-328
   renders every frame from scratch). */
MG_EXPORT mg_status mg_render_animated(mg_composition *comp, uint32_t width, uint32_t height,
                                       const float background[4], const mg_rect *crop, void **device_rgba,
                                       int *replayed);

/* Copies `bytes` of device memory (a frame from mg_render_device or
   mg_render_animated) to host memory. */
MG_EXPORT mg_status mg_copy_to_host(void *dst, const void *device_src, size_t bytes);

/* No counterpart in the reference: the measured HBM ceiling that the
   bandwidth-bound kernels are compared with. Allocates two device buffers of
   `bytes` (a nonzero multiple of 16), copies one to the other `repeats` times
   with the library's streaming (non-temporal) copy kernel on the device's
   engine stream, stores the average milliseconds of one copy (events on that
   stream) and the number of 32-bit words the check found wrong (0 for an
   exact copy that wrote nothing past its end), and frees the buffers. */
MG_EXPORT mg_status mg_stream_copy(size_t bytes, uint32_t repeats, double *ms_per_copy, uint64_t *mismatches);

/* Composition::pixel_segments (composition.rs:153-188): the canonically
   sorted PixelSegment stream (stable by `seg >> 22` over generation order,
   NONE tail included). The pointer stays valid until the next call on `comp`. */
MG_EXPORT mg_status mg_pixel_segments(mg_composition *comp, const uint64_t **out, size_t *n);

/* Per-pixel coverage of a render (the painter's integer accumulation,
   painter/mod.rs:229-243, 347-362): renders like mg_render_device and, for
   every tile inside the crop and every id of its workbench
   (populate_layers, layer_workbench.rs:401-428, before any workbench pass),
   records the 64 values A = 16 * P + area, where P is the wrapping i8 cover
   prefix of the pixel's row and area the wrapping i16 area sum of the pixel.
   Record r: a_out[64 r + (x * 8 + y)] for pixel (x, y) of the tile and
   key_out[4 r ..] = (scene 0, tile row, tile column, layer order). Records
   come in no particular order; at most `cap` are written and *n receives the
   total. Coverage for NonZero / EvenOdd follows from A (painter/mod.rs:
   53-71). No counterpart in the reference's API: the reference keeps these
   values inside the painter; this entry point exists to check them. */
MG_EXPORT mg_status mg_render_coverage(mg_composition *comp, uint32_t width, uint32_t height,
                                       const float background[4], const mg_rect *crop, int32_t *a_out,
                                       uint32_t *key_out, size_t cap, size_t *n);

/* Result-invariant tuning knob of the renderer (0 = automatic): the sort's
   digit width in bits (4, 6, 8 or 9). Every setting produces the same bytes;
   only the speed changes. */
MG_EXPORT mg_status mg_set_tuning(mg_composition *comp, uint32_t radix_bits);
/* Second result-invariant knob: threads per tile row in the row walk (64,
   256, 512 or 1024; 0 = automatic). */
MG_EXPORT mg_status mg_set_walk_threads(mg_composition *comp, uint32_t threads);

/* Statistics of the last render: lines, line-pixels, non-NONE segments,
   segments kept for painting, tiles, digit passes, radix bits. */
typedef struct mg_frame_stats {
    uint64_t lines, line_pixels, segments, segments_painted, tiles;
    uint32_t sort_passes, radix_bits;
    double device_ms;          /* GPU time of the frame (hipEvents)      */
    /* per stage (hipEvents on the library's stream) */
    double setup_ms;           /* line setup + line-pixel scan           */
    double emit_ms;            /* pixel-segment emission                 */
    double hist_ms;            /* first sort pass's digit-count scan     */
    double sort_ms;            /* all LSD digit passes, with the later
                                  passes' digit counts and scans         */
    double runs_ms;            /* row/tile ranges + per-tile layer runs  */
    double walk_ms;            /* row-serial cover carry (id records)    */
    double paint_ms;           /* tile paint: workbench + painter        */
    uint64_t runs, records;    /* tile layer runs, tile id records       */
    uint32_t walk_threads;     /* row-walk work-group size used          */
    uint32_t deep_tiles;       /* tiles painted four waves per tile      */
    /* row-walk arrays of the final walk: 0 a share of the LDS, 1 half the
       LDS, 2 the whole LDS, 3 global memory; walks repeated because a tile
       outgrew the arrays */
    uint32_t walk_lds, walk_rewalks;
    double sort_pass_ms[16];   /* each digit pass kernel alone (hipEvents
                                  around it; 0 past sort_passes)          */
} mg_frame_stats;
MG_EXPORT mg_status mg_last_frame_stats(const mg_composition *comp, mg_frame_stats *out);

/* ------------------------------------------------------------- batching */

/* A batch renders many compositions, each into its own frame, as ONE frame
   of the device pipeline: one launch per stage covers every scene, through a
   per-scene table of line, layer, props, tile-row and output offsets. No
   counterpart in the reference's API, where each Composition::render
   (composition.rs:212-328) is one frame; a caller with many small scenes
   calls mg_batch_render once instead of mg_render per scene. Every frame's
   bytes equal what mg_render_device gives for that composition alone. The
   batch keeps its lines resident while the compositions' contents are
   unchanged; compositions stay owned by the caller. */
typedef struct mg_batch mg_batch;

/* One frame of a batch: width x height pixels (stride = width), cleared to
   `background`, optionally cropped (Rect::new, buffer_layout.rs:43-50). */
typedef struct mg_batch_frame {
    uint32_t width, height;
    float background[4];
    int32_t has_crop;
    mg_rect crop;
} mg_batch_frame;

MG_EXPORT mg_status mg_batch_create(mg_batch **out);
MG_EXPORT void mg_batch_destroy(mg_batch *batch);

/* Renders comps[i] into frames[i] for i < n (n <= 8192; no layer caches, no
   flusher). Blocking. `*device_rgba` receives the device buffer holding all
   frames and `*offsets` (n entries, owned by the batch) the byte offset of
   each frame in it; both stay valid until the next call on `batch`. */
MG_EXPORT mg_status mg_batch_render(mg_batch *batch, mg_composition *const *comps, const mg_batch_frame *frames,
                                    size_t n, void **device_rgba, const uint64_t **offsets);

/* Copies frame i of the last mg_batch_render (width * height * 4 bytes) to
   host memory `rgba` of `len` bytes. */
MG_EXPORT mg_status mg_batch_read(mg_batch *batch, size_t i, uint8_t *rgba, size_t len);

/* Radix bits of the batch's sort (4, 6, 8 or 9) and threads per row of its
   row walk (64, 256 or 1024), or 0 to let the autotuner choose: the next
   render times each radix on the batch itself, then each walk size with the
   fastest radix, and keeps the fastest pair for batches of the same shape.
   Result-invariant. */
MG_EXPORT mg_status mg_batch_set_tuning(mg_batch *batch, uint32_t radix_bits);
MG_EXPORT mg_status mg_batch_set_walk_threads(mg_batch *batch, uint32_t threads);

/* The autotuner's last decision: candidates tried with their frame times
   (ms; 0 = not tried) and the choices. */
typedef struct mg_batch_tuning {
    uint32_t n_candidates;
    uint32_t radix_bits[4];
    double device_ms[4];
    uint32_t chosen_radix_bits;
    uint32_t n_walk_candidates;
    uint32_t walk_threads[3];
    uint32_t _pad;
    double walk_ms[3];
    uint32_t chosen_walk_threads;
    uint32_t _pad2;
} mg_batch_tuning;
MG_EXPORT mg_status mg_batch_last_tuning(const mg_batch *batch, mg_batch_tuning *out);

/* Statistics of the batch's last render (as mg_last_frame_stats, summed over
   its scenes). */
MG_EXPORT mg_status mg_batch_stats(const mg_batch *batch, mg_frame_stats *out);

/* -------------------------------------------------------------- serving */

/* A render server: requests (a composition and a frame) are admitted by a
   device-memory budget and rendered as batched frames (continuous batching)
   by one worker thread per engine, each engine with its own HIP stream, so a
   batch admitted while another is on the GPU runs beside it. No counterpart
   in the reference, whose Composition::render (composition.rs:212-328) is a
   blocking call on the caller's thread; for callers that render many
   independent scenes as requests arrive. Each request's bytes equal what
   mg_render_device gives for that composition alone. */
typedef struct mg_server mg_server;

typedef struct mg_server_config {
    uint64_t hbm_budget;  /* bytes of device memory in-flight requests may take (0 = half of free) */
    uint32_t engines;     /* batches in flight at once, 1..4 (0 = 2) */
    uint32_t max_batch;   /* requests per batch, 1..8192 (0 = 8192) */
} mg_server_config;

typedef struct mg_request_result {
    uint64_t ticket;
    int32_t status;       /* MG_OK or the render's error */
    uint32_t batch_size;  /* requests in the batch that rendered it */
    uint32_t engine;
    uint32_t _pad;
    double latency_ms;    /* submission -> RGBA8 in the caller's buffer */
    double queue_ms;      /* submission -> start of its batch */
    uint64_t footprint;   /* device-memory estimate used for admission */
} mg_request_result;

typedef struct mg_server_counters {
    uint64_t completed, failed, batches;
    double mean_batch;              /* requests per batch */
    double p50_ms, p99_ms, max_ms;  /* latency of completed requests */
    uint32_t max_concurrent_batches;
    uint32_t _pad;
    uint64_t queued, inflight_bytes;
} mg_server_counters;

MG_EXPORT mg_status mg_server_create(const mg_server_config *cfg, mg_server **out);
/* Finishes every submitted request, then stops the workers. */
MG_EXPORT void mg_server_destroy(mg_server *server);
MG_EXPORT mg_status mg_server_config_of(const mg_server *server, mg_server_config *out);
/* Queues a render of `comp` into `frame`; `rgba_out` (len >= width * height
   * 4 bytes, or NULL) receives the frame. `comp` and `rgba_out` must stay
   alive, and `comp` unchanged, until mg_server_wait returns for the ticket.
   MG_ERR_CAPACITY when the request alone exceeds the budget. */
MG_EXPORT mg_status mg_server_submit(mg_server *server, mg_composition *comp, const mg_batch_frame *frame,
                                     uint8_t *rgba_out, size_t len, uint64_t *ticket);
/* Blocks until the request completed; returns its status. */
MG_EXPORT mg_status mg_server_wait(mg_server *server, uint64_t ticket, mg_request_result *out);
/* Counts and latency percentiles since the last reset (reset != 0 clears them). */
MG_EXPORT mg_status mg_server_stats(mg_server *server, int reset, mg_server_counters *out);

/* ------------------------------------------------------ cassia-compatible */

/* cassia_init / cassia_render / cassia_shutdown (cassia/src/Cassia.h:32-41).
   The reference paints into a window; this build paints into a device
   RGBA8 frame of width x height read back with cassia_read_rgba8.
   `psegments` is a PixelSegment stream (any order, NONE entries ignored);
   stylings are indexed by the segment's layer field. */
typedef struct CassiaStyling {
    float fill[4];
    uint32_t fillRule;
    uint32_t blendMode;
    uint32_t _padding[2];
} CassiaStyling;

MG_EXPORT void cassia_init(uint32_t width, uint32_t height);
MG_EXPORT void cassia_render(const uint64_t *psegments, size_t psegmentCount,
                             const CassiaStyling *stylings, size_t stylingCount);
MG_EXPORT void cassia_shutdown(void);
/* Not in the reference: copy the last cassia_render frame (width*height*4 bytes). */
MG_EXPORT mg_status cassia_read_rgba8(uint8_t *out, size_t cap_bytes);

#ifdef __cplusplus
}
#endif

#endif /* MOLD_GPU_H */
