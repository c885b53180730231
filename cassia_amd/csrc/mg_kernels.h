// Kernel entry points and parameter blocks of the MI355X mold renderer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mg_flatten.h"
#include "mg_shared.h"

namespace mg {

struct PaintParams {
    const uint64_t *segs;        // sorted render words
    const uint32_t *row_start;   // global_rows + 1
    const uint32_t *row_scene;   // global row -> scene
    const SceneDesc *scenes;
    const PropsEntry *props;
    const float *stops;
    int32_t *prev_layers;        // BufferLayerCache::layers_per_tile (-1 = None)
    uint8_t *out;                // RGBA8 frame(s)
    uint32_t col_bits, layer_bits;
    uint32_t cap_ids, cap_runs;  // row-walk LDS capacities: queue + runs, runs of one tile
    uint32_t rec_chunk;          // id records a row reserves at a time
    unsigned char *walk_scratch; // row-walk arrays in global memory (k_row_walk<T, true>)
    uint64_t walk_stride;        // bytes per row of walk_scratch
    uint32_t walk_run_bufs;      // sets of row-walk run arrays (1, or 2 to overlap columns)
    // tiles of deep_min records or more (0: none) painted by k_paint_deep,
    // listed by the walk: (global row, tile column), at most deep_cap
    uint32_t deep_min, deep_cap;
    uint32_t *deep_count;
    uint2 *deep_list;
    uint32_t *overflow;          // bit 0: LDS capacity, bit 1: record capacity
    uint32_t *walk_max;          // the walk's largest runs of a tile [0] and queue + runs [1] (atomicMax)
    // debug: coverage records (A values) and per-tile write ops
    int32_t *cov_a;
    uint32_t *cov_key;
    uint32_t *cov_count;
    uint32_t cov_cap;
    int8_t *tile_ops;
    // tile columns: first word per (row, colp) (stride tstride), run counts
    uint32_t *tile_start;
    uint32_t *tile_run0;         // index of each (row, colp)'s first run in the run arrays
    uint32_t *tile_nr;           // runs of each (row, colp)
    uint32_t tstride;
    // runs, indexed by their column's first word + run index:
    // layer | even_odd << 16, first word, 8 row cover sums (i8 each)
    uint32_t *run_key;
    uint32_t *run_beg;
    uint64_t *run_cov;
    // per-tile id records (layer | has_run << 16 | has_carry << 17, run first
    // word, carried cover lo, hi) and run lengths; per-tile (offset, count)
    uint4 *rec;
    uint32_t *rec_len;
    uint2 *rec_hdr;
    uint32_t *rec_used;
    uint64_t rec_cap;
    uint32_t *n_runs;            // statistics: runs of the frame, in kStatSlices counters
    uint32_t *n_records;         // statistics: records of the frame, in kStatSlices counters
    const uint2 *work;           // (global row, tile column) per paint work group
};

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n);
__global__ void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n);
constexpr uint32_t kZeroRanges = 6;
struct ZeroRanges {
    uint32_t *p[kZeroRanges];
    uint64_t n[kZeroRanges];  // 32-bit words of each range
    uint32_t count;
};
__global__ void k_zero(ZeroRanges z);
__global__ void k_copy_stream(const uint4 *src, uint4 *dst, uint64_t n);
__global__ void k_copy_pattern(uint32_t *p, uint64_t n, uint32_t seed);
__global__ void k_copy_check(const uint32_t *b, uint64_t first, uint64_t n, uint32_t seed, unsigned long long *bad);
__global__ void k_setup(const float4 *pts, const uint16_t *ids, const uint32_t *ranks, const LayerEntry *layers,
                        const SceneDesc *scenes, const uint32_t *scene_line_base, FrameDims dims,
                        LineSetup *out, uint32_t *lengths, uint32_t *err, uint32_t order_lines);
__global__ void k_line_ranks(const uint16_t *ids, uint64_t first, uint64_t n, uint32_t *counts, uint32_t *rank);
// device-side bound of a scan (see k_scan_u32)
struct ScanLimit {
    const uint64_t *src = nullptr;
    uint32_t mul = 0, div = 1;
};
__global__ void k_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, uint64_t *status,
                           uint32_t *stall, uint64_t *total,
                           ScanLimit lim = ScanLimit{});
// Per-partition slot counts written by the emission: pass 0's digit counts
// (digit-major, one column per partition) and slices of the other totals.
struct SlotCounts {
    uint32_t *cnt0;
    uint32_t *acc;
    uint32_t *part_kept;  // per sort partition and wave: leading slots the first pass reads
};
constexpr uint32_t kCountSlicesHost = 256;
__global__ void k_emit(const LineSetup *lines, const uint32_t *offsets, const SceneDesc *scenes,
                       FrameDims dims, const FrameCounters *counters, uint64_t *out, const uint32_t *part_line,
                       SlotCounts sc);
__global__ void k_path_counts(const PathCommand *cmds, const flat::PathDesc *paths, uint32_t n_paths,
                              uint32_t *counts, uint32_t *err);
__global__ void k_path_emit(const PathCommand *cmds, const flat::PathDesc *paths, uint32_t n_paths,
                            const uint32_t *offsets, const uint32_t *counts, float4 *pts, uint16_t *ids);
__global__ void k_part_lines(const uint32_t *offsets, const uint32_t *lengths, uint64_t n_lines,
                             uint32_t part_size, uint32_t *part_line, uint64_t n_parts);
__global__ void k_clamp_line_pixels(FrameCounters *counters, uint64_t cap, uint32_t *flag);
__global__ void k_row_starts(const uint64_t *sorted, const FrameCounters *counters, uint32_t rows,
                             uint32_t row_shift, uint32_t *row_start);
__global__ void k_ingest(const uint64_t *in, uint64_t n_in, const SceneDesc *scenes, FrameDims dims,
                         const FrameCounters *counters, uint64_t *out, SlotCounts sc);
__global__ void k_count_reduce(const uint32_t *acc, FrameCounters *counters);
// the type a sort pass stores the next pass's R-bit digit in
template <int R> struct DigitOf { using T = uint8_t; };
template <> struct DigitOf<9> { using T = uint16_t; };
template <int R>
__global__ void k_digit_count(const typename DigitOf<R>::T *dig, const FrameCounters *counters, uint32_t *cnt);
template <int R>
__global__ void k_sort_pass(const uint64_t *in, uint64_t *out, const FrameCounters *counters, uint32_t shift,
                            const uint32_t *offsets, const uint32_t *part_kept, typename DigitOf<R>::T *dig_out);
__global__ void k_tile_starts(PaintParams pp, uint32_t global_rows, uint32_t stream_runs);
__global__ void k_tile_runs(PaintParams pp, const uint2 *work, uint32_t n_work);
__global__ void k_tile_nr(PaintParams pp, uint32_t global_rows);
__global__ void k_runs_scan(PaintParams pp, const FrameCounters *counters, uint64_t *status, uint32_t *stall,
                            uint32_t *open);
__global__ void k_run_cov(PaintParams pp, const FrameCounters *counters, const uint32_t *open);
// T = 64, 256, 512 or 1024 threads per tile row; G: queue in global scratch
template <uint32_t T, bool G>
__global__ void k_row_walk(PaintParams pp);
__global__ void k_paint(PaintParams pp);
__global__ void k_paint_deep(PaintParams pp);
size_t paint_deep_lds_bytes_host();
size_t walk_lds_bytes_host(uint32_t cap, uint32_t rcap, uint32_t tstride, uint32_t run_bufs);
size_t paint_lds_bytes_host(uint32_t cap, bool clips);

constexpr uint32_t kStatSlices = 32;
constexpr uint32_t kEmitTileHost = 2048;
constexpr uint32_t kSortTileHost = 8192;
constexpr uint32_t kSortThreadsHost = 512;  // threads per sort pass work group
constexpr uint32_t kXcdsHost = 8;           // XCD-mapped launches round their grids to multiples of this
inline uint32_t xcd_grid(uint64_t parts) {
    return (uint32_t)(parts ? (parts + kXcdsHost - 1) / kXcdsHost * kXcdsHost : kXcdsHost);
}
constexpr uint32_t kScanTileHost = 4096;  // values per k_scan_u32 partition (16 per thread)
constexpr uint32_t kRunsTileHost = 4096;  // words per k_runs_scan partition

}  // namespace mg
