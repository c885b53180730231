/*
 * CPU oracle: a scalar, single-threaded restatement of the mold/surpass
 * rasterize -> sort -> paint path, written from the Rust sources.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls this
 * code; only the tests, the smoke check and the bench's CPU-baseline leg do,
 * and there only as the checker (or as the timed CPU baseline).
 *
 * Arithmetic contract (mirrors Rust on x86_64 with the portable `auto.rs` SIMD
 * backend, surpass/src/simd/mod.rs:7-35):
 *   - binary32 round-to-nearest-even, no contraction (built -ffp-contract=off),
 *     every `mul_add` is an explicit fmaf();
 *   - `as i32` / `as u8` from f32 truncate, saturate, NaN -> 0;
 *   - `round()` is half away from zero; `f32::min/max` ignore NaN;
 *   - integer +,-,* wrap (release build; built -fwrapv);
 *   - i8/i16 accumulators wrap.
 *
 * Every function cites the reference file:line it restates.
 */
#ifndef MOLD_ORACLE_H
#define MOLD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_PIXEL_WIDTH 16
#define OR_TILE_WIDTH 8
#define OR_TILE_HEIGHT 8
#define OR_NONE ((uint64_t)1 << 63)

/* ---- PixelSegment (surpass/src/rasterizer/raster_segment.rs:142-160) ---- */
uint64_t or_seg_new(int32_t is_none, int32_t tile_j, int32_t tile_i, uint32_t layer,
                    uint32_t tile_y, uint32_t tile_x, int32_t area, int32_t cover);
/* out: is_none, tile_j, tile_i (raw sign-extended field), layer, tile_y, tile_x, area, cover */
void or_seg_fields(uint64_t seg, int32_t out[8]);

/* ---- mold path flattening (mold/src/path/{mod,segments,transform}.rs) ---- */
enum { OR_CMD_LINE = 0, OR_CMD_QUAD, OR_CMD_RAT_QUAD, OR_CMD_CUBIC, OR_CMD_RAT_CUBIC, OR_CMD_CLOSE };
typedef struct {
    int32_t kind;
    /* points as (x, y) pairs; rational commands store (x*w, y*w, w) triples
       (premultiplied, mold/src/path/mod.rs:92-96,137-142) */
    float v[12];
} or_cmd;
/* Returns the number of segments produced (4 floats each: p0x p0y p1x p1y),
   writing at most `cap`. `t9` may be NULL (no transform). */
size_t or_path_flatten(const or_cmd *cmds, size_t n, const float *t9, float *out, size_t cap);

/* ---- surpass::Layer as seen by LinesBuilder::build (surpass/src/layer.rs:5-20) ---- */
typedef struct {
    uint8_t exists;     /* `layers(id)` returns Some */
    uint8_t is_enabled;
    uint8_t has_affine;
    uint8_t has_order;
    uint16_t order;
    uint16_t _pad;
    float affine[6];
} or_layer;

/* ---- Lines SoA (surpass/src/segment.rs:240-258) ---- */
typedef struct {
    uint16_t *orders;
    int32_t *starts;
    float *starts_f32;
    float *ends_f32;
    int64_t *lengths;
    float *slopes;
    float *slopes_recip;
    float *py_slope_pxs;
    float *px_slope_recip_pys;
    uint8_t *octants;
} or_lines;

/* LinesBuilder::build (surpass/src/segment.rs:91-237). `layers` is indexed by
   layer id and may be NULL (the `|_| None` closure); `default_transform` may
   be NULL. */
void or_lines_build(size_t n, const float *p0x, const float *p0y, const float *p1x,
                    const float *p1y, const uint16_t *layer_ids, const or_layer *layers,
                    size_t n_layers, const float *default_transform, or_lines *out);

/* Rasterizer::rasterize (surpass/src/rasterizer/mod.rs:214-285): writes
   2 * sum(lengths) segments in generation order ([seg, seg|NONE] per
   line-pixel). Returns the count written or (size_t)-1 if cap is too small. */
size_t or_rasterize(size_t n_lines, const or_lines *lines, uint64_t *out, size_t cap);

/* Canonical sort: stable by the 42-bit key `seg >> 22` over generation order
   (the reference's unstable sort, mod.rs:287-290, made canonical). */
void or_sort_canonical(uint64_t *segs, size_t n);

/* ---- Styles / props (surpass/src/painter/{mod,style}.rs) ---- */
enum { OR_NONZERO = 0, OR_EVENODD = 1 };
enum { OR_FUNC_DRAW = 0, OR_FUNC_CLIP = 1 };
enum { OR_FILL_SOLID = 0, OR_FILL_GRADIENT = 1 };
enum { OR_GRAD_LINEAR = 0, OR_GRAD_RADIAL = 1 };
typedef struct {
    uint32_t fill_rule;
    uint32_t func;
    uint32_t clip_layers; /* Func::Clip(usize) */
    uint32_t is_clipped;
    uint32_t blend_mode;  /* BlendMode as u32, style.rs:189-203 */
    uint32_t fill_kind;
    float color[4];
    uint32_t grad_type;
    uint32_t n_stops;
    float grad_start[2];
    float grad_end[2];
    const float *stops;   /* n_stops x (r, g, b, a, stop), stops already resolved */
    uint32_t is_unchanged;
    uint32_t defined;     /* props exist for this order */
} or_props;

/* Blend a single channel, scalar form (style.rs:206-267) and SIMD form
   (style.rs:288-418). */
float or_blend_fn(uint32_t mode, float dst, float src);
void or_blend_simd(uint32_t mode, const float dst[3], const float src[3], float out[3]);
void or_blend(uint32_t mode, const float dst[4], const float src[4], float out[4]);
/* GradientBuilder::build stop resolution (style.rs:66-84): rewrites stops
   whose value == NO_STOP (-0.0, which also matches +0.0). */
void or_gradient_resolve_stops(float *stops, uint32_t n_stops);
/* Gradient::color_at (style.rs:121-175): 8 rows starting at (x, y). out[ch*8+row] */
void or_gradient_color_at(const or_props *p, float x, float y, float out[32]);
float or_linear_to_srgb(float l);       /* painter/mod.rs:91-106 */
float or_linear_to_srgb_simd(float l);  /* painter/mod.rs:74-88 */
void or_to_bytes(const float color[4], uint8_t out[4]); /* painter/mod.rs:108-123 */

/* ---- Tile-level painter / workbench ---- */
typedef struct {
    int8_t cover[8];
    uint16_t layer;
    uint16_t _pad;
} or_carry;

enum { OR_OP_NONE = 0, OR_OP_SOLID = 1, OR_OP_COLOR_BUFFER = 2 };

/* One call of LayerWorkbench::drive_tile_painting (layer_workbench.rs:430-486)
   on a fresh painter/workbench: `carries` are the queue after init(), `segs`
   the tile's segments (sorted by layer). Linear colours of the painter are
   returned column-major (x*8+y) when the op is ColorBuffer. */
typedef struct {
    int32_t op;
    float solid[4];
    float colors[64 * 4]; /* [pixel][rgba], pixel = x*8+y */
    size_t n_carries_out;
    size_t n_ids;
    uint16_t ids[1024];   /* workbench ids after the passes, in order */
    uint8_t masked[1024]; /* effective mask (i >= skipped && mask) */
    uint8_t skip_clipping[1024];
} or_tile_result;

/* passes: bit0 tile_unchanged, bit1 skip_trivial_clips, bit2 skip_fully_covered;
   stop_after_passes: when set, only populate + the selected passes run and the
   result's op is OR_OP_NONE-with-Continue (-1) or the Break op. */
int or_drive_tile(const or_carry *carries, size_t n_carries, const uint64_t *segs, size_t n_segs,
                  const or_props *props_by_order, size_t n_props, const float clear[4],
                  uint32_t tile_i, uint32_t tile_j, int32_t *previous_layers, uint32_t passes,
                  int stop_after_passes, or_carry *carries_out, size_t carries_cap,
                  or_tile_result *res);

/* ---- BufferLayout::print (buffer_layout.rs:170-230) + paint_tile_row ---- */
typedef struct {
    uint32_t h0, h1, v0, v1; /* crop in tiles: horizontal [h0,h1), vertical [v0,v1) */
} or_rect;

typedef struct {
    /* optional coverage dump: for every processed tile and every workbench id
       the 64 values A = 16*P + area (column-major), see painter/mod.rs:347-362 */
    int32_t *cov_a;       /* cap_records x 64 */
    uint32_t *cov_key;    /* cap_records x 3: tile_j, tile_i, layer */
    size_t cap_records;
    size_t n_records;
    /* optional: tile op per tile (T entries, -1 untouched) */
    int8_t *tile_ops;
} or_debug;

/* `segs` must be the canonical sorted stream cut after the last non-NONE
   segment (composition.rs:299-304). `layers_per_tile` may be NULL (no
   BufferLayerCache), else tiles_len entries with -1 meaning None.
   Returns 0 on success, nonzero on a reference panic condition. */
int or_print(uint8_t *buffer, size_t len_px, size_t width, size_t width_stride,
             int32_t *layers_per_tile, size_t tiles_len, const uint64_t *segs, size_t n_segs,
             const float clear[4], const or_rect *crop, const or_props *props_by_order,
             size_t n_props, or_debug *dbg);

/* Whole render from lines: build -> rasterize -> canonical sort -> cut ->
   print. Returns 0 on success. `n_segments_out` (may be NULL) receives the
   number of non-NONE pixel segments. */
int or_render(size_t n_lines, const float *p0x, const float *p0y, const float *p1x,
              const float *p1y, const uint16_t *layer_ids, const or_layer *layers,
              size_t n_layers, uint8_t *buffer, size_t len_px, size_t width,
              size_t width_stride, int32_t *layers_per_tile, size_t tiles_len,
              const float clear[4], const or_rect *crop, const or_props *props_by_order,
              size_t n_props, size_t *n_segments_out, or_debug *dbg);

/* Sorted stream for Composition::pixel_segments (composition.rs:153-188):
   all 2*sum(lengths) entries, canonical order, NONE tail included. */
size_t or_pixel_segments(size_t n_lines, const float *p0x, const float *p0y, const float *p1x,
                         const float *p1y, const uint16_t *layer_ids, const or_layer *layers,
                         size_t n_layers, uint64_t *out, size_t cap);

/* Multi-threaded variant of or_render (the CPU baseline). Same output. */
int or_render_mt(int n_threads, size_t n_lines, const float *p0x, const float *p0y,
                 const float *p1x, const float *p1y, const uint16_t *layer_ids,
                 const or_layer *layers, size_t n_layers, uint8_t *buffer, size_t len_px,
                 size_t width, size_t width_stride, const float clear[4],
                 const or_props *props_by_order, size_t n_props, size_t *n_segments_out);

#ifdef __cplusplus
}
#endif

#endif
