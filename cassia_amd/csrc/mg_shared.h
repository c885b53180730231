// Host/device shared definitions of the MI355X mold renderer.
//
// PixelSegment wire format (surpass/src/rasterizer/raster_segment.rs:142-154),
// MSB -> LSB:
//   is_none:1 | tile_j:12 (signed) | tile_i:13 (max(tx,-1)+256) | layer:16 |
//   tile_y:3 | tile_x:3 | area:10 (signed) | cover:6 (signed)
//
// Render-order word used between emission, sort and paint (not exported):
//   [63 .. 22] key = global_tile_row << (col_bits + layer_bits)
//                    | (col + 1) << layer_bits | layer      (col = -1: left of canvas)
//   [21 ..  0] the PixelSegment's tile_y | tile_x | area | cover bits, unchanged.
#pragma once

#include <stdint.h>

namespace mg {

constexpr int kPayloadBits = 22;
constexpr uint64_t kNone = 1ull << 63;
constexpr uint32_t kTileW = 8;
constexpr uint32_t kTileH = 8;

// Layer table entry, indexed by scene layer id (surpass/src/layer.rs:5-20).
enum : uint32_t { LF_EXISTS = 1, LF_ENABLED = 2, LF_AFFINE = 4, LF_ORDER = 8 };
struct LayerEntry {
    uint32_t flags;
    uint32_t order;
    // render frames: position of the layer's first line when the frame's
    // lines are ordered by (scene, layer order); line i of the layer goes to
    // base + its rank among the layer's lines
    uint32_t base;
    uint32_t _pad;
    float affine[6];
};

// Props entry, indexed by scene order (painter/mod.rs:137-141, style.rs:426-432).
enum : uint32_t {
    PF_DEFINED = 1u << 0,
    PF_EVENODD = 1u << 1,
    PF_CLIP = 1u << 2,      // Func::Clip
    PF_CLIPPED = 1u << 3,   // Style::is_clipped
    PF_GRADIENT = 1u << 4,  // Fill::Gradient
    PF_UNCHANGED = 1u << 5, // Layer::is_unchanged(cache id) for this frame
    PF_OPAQUE = 1u << 6,    // Draw + Solid + Over + alpha == 1.0
    PF_RADIAL = 1u << 7,
    PF_BLEND_SHIFT = 8,     // 4 bits of BlendMode
    // a solid Soft Light layer with finite colour channels: the six words
    // from grad_start on hold soft_light_terms of channels 0..2 (u0 u1 u2
    // v0 v1 v2), computed when the table is built
    PF_SOFT_TERMS = 1u << 12,
};
struct PropsEntry {
    uint32_t flags;
    uint32_t clip_layers;
    float color[4];
    float grad_start[2];
    float grad_end[2];
    uint32_t stops_offset;  // into the stops array, 5 floats per stop
    uint32_t n_stops;
};

// One scene of a frame (a plain render is a batch of one scene).
// SF_NO_OPAQUE: no layer of the scene has PF_OPAQUE (the painter's search for
// the topmost opaque layer of a tile can stop early)
enum : uint32_t { SF_CLIPS = 1, SF_CACHE = 2, SF_GRADIENTS = 4, SF_NO_OPAQUE = 8 };
struct SceneDesc {
    uint32_t line_base, n_lines;
    uint32_t layer_base, n_layer_entries;
    uint32_t props_base, n_props;
    uint32_t width, rows_px, stride;  // pixels
    uint32_t row_len, tile_rows;
    uint32_t row_base;                // first global tile row of the scene
    uint32_t crop_h0, crop_h1, crop_v0, crop_v1;  // tiles
    uint32_t tile_base;               // first tile of the scene in per-tile arrays (record headers, previous_layers)
    uint32_t flags;
    uint64_t out_offset;              // byte offset of pixel (0, 0) in the output
    float clear[4];
    // device addresses of the scene's own resident lines (float4 points, u16
    // layer ids), line i of the frame being their line i - line_base; 0: the
    // frame's line arrays hold them at line_base
    uint64_t lines_pts, lines_ids;
    // and each line's rank among the lines of its layer id (u32)
    uint64_t lines_rank;
};

// Per-line setup record written by the setup kernel (segment.rs:144-219).
struct LineSetup {
    int32_t start;
    float start_f32, end_f32;
    float slope, slope_recip, py_slope_px, px_slope_recip_py;
    uint32_t packed;  // order | octant << 16 | scene << 19
};

struct FrameDims {
    uint32_t n_scenes;
    uint32_t n_lines;
    uint32_t global_rows;
    uint32_t col_bits, layer_bits, key_bits;
    uint32_t radix_bits, passes;
    uint32_t mode;        // 0: render (cull + render key), 1: canonical full stream, 2: ingest
    uint32_t sort_shift;  // bit of the first sort digit
};

// Device-side counters of a frame.
struct FrameCounters {
    uint64_t line_pixels;     // P
    uint64_t segments;        // non-NONE pixel segments
    uint64_t kept;            // segments emitted for sorting
    uint64_t order_total;     // lines counted by the line-order histogram
};

}  // namespace mg
