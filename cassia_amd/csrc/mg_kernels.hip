// HIP kernels of the MI355X mold renderer (gfx950, wave64).
//
// Pipeline of one frame (all scenes of a batch in the same launches):
//   k_setup        per line: LinesBuilder::build           (segment.rs:91-237)
//   k_scan_u32     exclusive scan of line lengths          (rasterizer/mod.rs:56-125)
//   k_emit         per line-pixel: 1-2 PixelSegments,      (rasterizer/mod.rs:127-285)
//                  compacted in generation order, with the first sort
//                  pass's per-partition digit counts fused in
//   k_sort_pass    one stable LSD digit pass (block rank, offsets from scanned
//                  per-partition digit counts; k_digit_count counts the next pass's)
//   k_row_starts   first word of every tile row (binary search)
//   k_tile_starts, k_tile_runs, k_row_walk, k_paint       (painter/mod.rs:229-596,
//                  tile ranges, runs with cover sums, the   layer_workbench.rs:134-486)
//                  row-serial cover carry producing per-tile id records, and
//                  the tile-parallel workbench + painter with RGBA8 stores
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mg_exact.h"
#include "mg_kernels.h"
#include "mg_shared.h"

namespace mg {

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// Partitions dealt to the XCDs in contiguous ranges. Work groups go to the
// eight XCDs round-robin in index order, each XCD starting its own in index
// order, so work group b runs on XCD b % 8: giving XCD x the partitions
// [x * per, (x + 1) * per) lets neighbouring partitions share an L2, where
// the partial cache lines they write at their common edges merge instead of
// reaching HBM as partial writes. Only for kernels whose partitions do not
// wait on one another. The launch has 8 * ceil(max partitions / 8) work
// groups; false for a work group without a partition.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ bool xcd_part(uint32_t nparts, uint32_t &part) {
    const uint32_t per = (nparts + kXcds - 1) / kXcds;
    const uint32_t j = blockIdx.x / kXcds;
    part = (blockIdx.x % kXcds) * per + j;
    return j < per && part < nparts;
}

// Look-back status words carry their whole payload (a count and a flag) in
// one atomic word and no other memory is published through them, so relaxed
// device-coherent accesses suffice: no cache invalidation per spin.
__device__ __forceinline__ uint64_t ld_acquire(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_acquire32(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release32(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back status words: 2 flag bits above a 62-bit value.
constexpr uint64_t kStAgg = 1ull << 62;
constexpr uint64_t kStInc = 2ull << 62;
constexpr uint64_t kStVal = (1ull << 62) - 1;

// Partitions are work-group indices. Work groups are dispatched in index
// order (per XCD, and XCDs receive them in a fixed rotation), so the
// predecessors of a running partition are running or finished, and every
// look-back wait ends. The waits are still bounded: after kSpinLimit units of
// 64-cycle sleep (about 7 ms) a look-back gives up, raises *stall and finishes
// with a partial prefix (every write stays inside its buffer; the host
// reports the frame as failed).
constexpr uint32_t kSpinLimit = 1u << 18;

// Sleeps 64 * n cycles for n in 1..32 (s_sleep takes an immediate).
__device__ __forceinline__ void backoff(uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) __builtin_amdgcn_s_sleep(1);
}

// Exclusive prefix of the partitions before `part`; executed by one full
// wave. Each round reads 64 predecessors, consumes the ready ones nearest to
// `part` (up to the first inclusive prefix) and waits, with exponential
// backoff, only on the first one not yet published.
__device__ uint64_t lookback_wave(const uint64_t *status, uint32_t part, uint32_t *stall) {
    const uint32_t lane = lane_id();
    uint64_t excl = 0;
    int64_t top = (int64_t)part - 1;
    uint32_t spins = 0, nap = 1;
    while (top >= 0) {
        const int64_t idx = top - (int64_t)lane;
        const uint64_t s = idx >= 0 ? ld_acquire(status + idx) : kStInc;
        const uint64_t flag = s >> 62;
        const uint64_t not_ready = __ballot(flag == 0);
        const uint64_t inc = __ballot(flag == 2);
        const uint32_t first_inc = inc ? (uint32_t)__ffsll((long long)inc) - 1 : 64u;
        const uint32_t first_nr = not_ready ? (uint32_t)__ffsll((long long)not_ready) - 1 : 64u;
        const uint32_t take = first_inc < first_nr ? first_inc + 1 : first_nr;
        uint64_t v = (lane < take && idx >= 0) ? (s & kStVal) : 0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        excl += v;
        if (first_inc < first_nr) break;
        top -= take;
        if (take < 64) {
            if ((spins += nap) > kSpinLimit) {
                if (lane == 0) atomicOr(stall, 1u);
                break;
            }
            backoff(nap);
            nap = nap < 32 ? nap * 2 : 32;
        } else {
            nap = 1;
        }
    }
    return excl;
}

// Inclusive sum over the 64 lanes of a wave (all lanes active): shifts by
// 1, 2, 4, 8 inside each 16-lane row, then the row totals carried forward
// by the row broadcasts of lanes 15 and 31 (one DPP add each).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);  // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return x;
}

// Block-wide exclusive scan of one u32 per thread over W waves (W <= 64):
// wave scans, then one wave-scan of the wave totals read from LDS. Without
// TrailingSync the caller must pass a barrier before wsum is written again.
template <uint32_t W, bool TrailingSync = true>
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t *wsum /* W */, uint32_t *total) {
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t x = wave_incl_scan(v);
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    const uint32_t t = wave_incl_scan(lane < W ? wsum[lane] : 0u);
    const uint32_t base = w ? (uint32_t)__builtin_amdgcn_readlane((int)t, (int)w - 1) : 0u;
    *total = (uint32_t)__builtin_amdgcn_readlane((int)t, (int)W - 1);
    if (TrailingSync) __syncthreads();
    return base + x - v;
}

// Inclusive maximum over the 64 lanes of a wave (all active), the DPP steps
// of wave_incl_scan.
__device__ __forceinline__ uint32_t wave_incl_max(uint32_t x) {
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, false));  // row_shr:1
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, false));  // row_shr:2
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false));  // row_shr:4
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false));  // row_shr:8
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false));  // row_bcast:15
    x = max(x, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return x;
}

// Block-wide exclusive scan of one u32 per thread (blockDim = 256).
__device__ uint32_t block_excl_scan_256(uint32_t v, uint32_t *wsum /* 4 */, uint32_t *total) {
    return block_excl_scan_u32<4>(v, wsum, total);
}

// Block-wide exclusive scan of one value (u32 or u64) per thread over W waves.
template <uint32_t W, class V>
__device__ V block_excl_scan(V v, V *wsum /* W */, V *total) {
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    V x = v;
    for (int off = 1; off < 64; off <<= 1) {
        V y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    V base = 0, all = 0;
    for (uint32_t k = 0; k < W; ++k) {
        V s = wsum[k];
        if (k < w) base += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return base + x - v;
}

// LDS hand-off between the lanes of one wave: the wave's LDS operations
// complete in issue order, so ordering the compiler (and waiting for the
// wave's own outstanding LDS operations) is all it takes. Waves that work
// independently (a deep tile's stagers, tile-run waves) use it instead of a
// work-group barrier.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ------------------------------------------------------------ fill kernels

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Clears up to kZeroRanges ranges of 32-bit words in one launch (the frame's
// counters, flags, look-back words and histograms).
__global__ void k_zero(ZeroRanges z) {
    const uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x, step = (uint64_t)gridDim.x * blockDim.x;
    for (uint32_t r = 0; r < z.count; ++r)
        for (uint64_t i = t; i < z.n[r]; i += step) z.p[r][i] = 0u;
}

// Streaming copy of n 16-byte words: four non-temporal 16-byte loads in
// flight per thread, then their stores. The device-copy rate it reaches is
// the HBM ceiling the bandwidth-bound kernels are measured against.
__global__ __launch_bounds__(256) void k_copy_stream(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                     uint64_t n) {
    typedef uint32_t V4 __attribute__((ext_vector_type(4)));
    constexpr uint32_t U = 4;
    const V4 *a = (const V4 *)src;
    V4 *b = (V4 *)dst;
    const uint64_t stride = (uint64_t)gridDim.x * 256 * U;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 * U + threadIdx.x; i < n; i += stride) {
        V4 v[U];
#pragma unroll
        for (uint32_t k = 0; k < U; ++k)
            if (i + k * 256 < n) v[k] = __builtin_nontemporal_load(&a[i + k * 256]);
#pragma unroll
        for (uint32_t k = 0; k < U; ++k)
            if (i + k * 256 < n) __builtin_nontemporal_store(v[k], &b[i + k * 256]);
    }
}

// Word i of a copy check's source (a 32-bit mix of i and the seed), and the
// check: words of `b` differing from that pattern are counted into *bad.
__device__ __forceinline__ uint32_t copy_pattern(uint64_t i, uint32_t seed) {
    uint32_t h = (uint32_t)i * 0x9E3779B1u ^ (uint32_t)(i >> 32) ^ seed;
    h ^= h >> 15;
    h *= 0x85EBCA77u;
    return h ^ (h >> 13);
}
__global__ void k_copy_pattern(uint32_t *p, uint64_t n, uint32_t seed) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = copy_pattern(i, seed);
}
__global__ void k_copy_check(const uint32_t *b, uint64_t first, uint64_t n, uint32_t seed,
                             unsigned long long *bad) {
    uint32_t local = 0;
    for (uint64_t i = first + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        local += b[i] != copy_pattern(i, seed) ? 1u : 0u;
    if (local) atomicAdd(bad, (unsigned long long)local);
}

// ----------------------------------------------------------------- setup

__device__ __forceinline__ uint32_t scene_of_line(const uint32_t *scene_line_base, uint32_t n_scenes,
                                                  uint32_t i) {
    uint32_t lo = 0, hi = n_scenes;  // last scene with base <= i
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (scene_line_base[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// Head lane of each run of equal keys among a wave's valid lanes (a prefix
// of the wave); returns the head of this lane's run.
__device__ __forceinline__ uint32_t run_head_lane(bool valid, uint32_t key, uint64_t *heads_out) {
    const uint32_t lane = lane_id();
    uint32_t prev = __shfl_up(key, 1);
    const bool head = valid && (lane == 0 || prev != key);
    const uint64_t heads = __ballot(head);
    *heads_out = heads;
    const uint64_t le = heads & ((2ull << lane) - 1ull);
    return le ? 63u - (uint32_t)__clzll((long long)le) : 0u;
}

// Rank of each line among the lines of its layer id: lines [first, n) are
// ranked after those counted already (counts[id]: lines of id ranked so
// far), one returning atomic per run of equal ids in a wave. Ranks inside a
// layer come in no fixed order: the segments of one layer in one tile are
// summed (wrapping integer adds), so their order never shows in a frame.
__global__ __launch_bounds__(256) void k_line_ranks(const uint16_t *__restrict__ ids, uint64_t first, uint64_t n,
                                                    uint32_t *__restrict__ counts, uint32_t *__restrict__ rank) {
    const uint64_t i = first + blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const uint32_t key = valid ? ids[i] : 0xFFFFFFFFu;
    uint64_t heads;
    const uint32_t head = run_head_lane(valid, key, &heads);
    const uint32_t lane = lane_id();
    const uint32_t vn = (uint32_t)__popcll(__ballot(valid));  // valid lanes are a prefix
    uint32_t at = 0;
    if ((heads >> lane) & 1ull) {
        const uint64_t after = heads & ~((2ull << lane) - 1ull);
        const uint32_t next = after ? (uint32_t)__ffsll((long long)after) - 1u : 64u;
        const uint32_t end = next < vn ? next : vn;
        at = atomicAdd(&counts[key], end - lane);
    }
    const uint32_t r = __shfl(at, (int)head) + (lane - head);
    if (valid) rank[i] = r;
}

// LinesBuilder::build for one line (segment.rs:107-219). Render frames
// (order_lines) write line i to its layer's base + its rank, which orders
// the frame's lines by (scene, layer order): slot order inside every tile is
// then layer order, and the segment sort covers only the row and column
// bits. Other frames keep the line order.
__global__ __launch_bounds__(256) void k_setup(const float4 *__restrict__ pts, const uint16_t *__restrict__ ids,
                                               const uint32_t *__restrict__ ranks,
                                               const LayerEntry *__restrict__ layers,
                                               const SceneDesc *__restrict__ scenes,
                                               const uint32_t *__restrict__ scene_line_base,
                                               FrameDims dims, LineSetup *__restrict__ out,
                                               uint32_t *__restrict__ lengths, uint32_t *__restrict__ err,
                                               uint32_t order_lines) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dims.n_lines) return;
    uint32_t sc = 0;
    if (dims.n_scenes > 1) {
        // the wave's lines are consecutive and mostly of one scene: one
        // search for its first line, then lanes past that scene's lines
        // search on their own
        const uint32_t i0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)i);
        sc = scene_of_line(scene_line_base, dims.n_scenes, i0);
        const uint32_t next = sc + 1 < dims.n_scenes ? scene_line_base[sc + 1] : 0xFFFFFFFFu;
        if (i >= next) sc = scene_of_line(scene_line_base, dims.n_scenes, i);
    }
    const SceneDesc &S = scenes[sc];
    // the scene's own resident lines, or the frame's line arrays
    const float4 *P = S.lines_pts ? (const float4 *)S.lines_pts : pts;
    const uint16_t *I = S.lines_ids ? (const uint16_t *)S.lines_ids : ids;
    const uint32_t *Rk = S.lines_pts ? (const uint32_t *)S.lines_rank : ranks;
    const uint32_t li = S.lines_pts ? i - S.line_base : i;
    uint32_t id = I[li];
    LayerEntry le;
    le.flags = 0;
    if (id < S.n_layer_entries) le = layers[S.layer_base + id];
    LineSetup r;
    r.start = 0;
    r.start_f32 = r.end_f32 = r.slope = r.slope_recip = r.py_slope_px = r.px_slope_recip_py = 0.0f;
    r.packed = sc << 19;
    uint32_t length = 0;
    bool exists = le.flags & LF_EXISTS;
    uint32_t pos = i;
    if (order_lines) {
        // every line of a render frame has a layer entry (its position)
        pos = exists ? le.base + Rk[li] : 0xFFFFFFFFu;
        if (pos >= dims.n_lines) {
            atomicOr(err, 2u);
            return;
        }
    }
    if (!(exists && !(le.flags & LF_ENABLED))) {
        uint32_t order = (exists && (le.flags & LF_ORDER)) ? le.order : id;
        float4 p = P[li];
        float p0x = p.x, p0y = p.y, p1x = p.z, p1y = p.w;
        if (exists && (le.flags & LF_AFFINE)) {
            const float *t = le.affine;
            float a0 = __builtin_fmaf(t[0], p0x, __builtin_fmaf(t[1], p0y, t[4]));
            float b0 = __builtin_fmaf(t[2], p0x, __builtin_fmaf(t[3], p0y, t[5]));
            float a1 = __builtin_fmaf(t[0], p1x, __builtin_fmaf(t[1], p1y, t[4]));
            float b1 = __builtin_fmaf(t[2], p1x, __builtin_fmaf(t[3], p1y, t[5]));
            p0x = a0; p0y = b0; p1x = a1; p1y = b1;
        }
        int32_t x0 = sub_pixel_hw(p0x), y0 = sub_pixel_hw(p0y), x1 = sub_pixel_hw(p1x), y1 = sub_pixel_hw(p1y);
        if (y0 != y1) {
            int32_t dx = wrap_sub(x1, x0), dy = wrap_sub(y1, y0);
            int32_t adx = dx < 0 ? wrap_sub(0, dx) : dx, ady = dy < 0 ? wrap_sub(0, dy) : dy;
            bool x_major = adx > ady;
            bool forward = wrap_add(dx, dy) > 0 || (dx == wrap_sub(0, dy) && dx < dy);
            int32_t s_sub, e_sub;
            if (forward) { s_sub = x_major ? x0 : y0; e_sub = x_major ? x1 : y1; }
            else { s_sub = x_major ? x1 : y1; e_sub = x_major ? x0 : y0; }
            int32_t start = s_sub >> 4;
            int32_t end = (e_sub & 15) == 0 ? (e_sub >> 4) : wrap_add(e_sub >> 4, 1);
            int32_t len = wrap_sub(end, start);
            if (len < 0) atomicOr(err, 1u);  // reference: `as usize` of a negative length
            length = len < 0 ? 0u : (uint32_t)len;
            float fx0 = (float)x0 * 0.0625f, fy0 = (float)y0 * 0.0625f;
            float px = x_major ? fx0 : fy0, py = x_major ? fy0 : fx0;
            float slope = x_major ? (float)dy / (float)dx : (float)dx / (float)dy;
            float slope_recip = 1.0f / slope;
            uint32_t octant;
            bool gx = dx > 0, gy = dy > 0;
            if (x_major) octant = gy ? (gx ? 0u : 3u) : (gx ? 7u : 4u);
            else octant = gy ? (gx ? 1u : 2u) : (gx ? 6u : 5u);
            r.start = start;
            r.start_f32 = (float)s_sub * 0.0625f;
            r.end_f32 = (float)e_sub * 0.0625f;
            r.slope = slope;
            r.slope_recip = slope_recip;
            r.py_slope_px = __builtin_fmaf(slope, -px, py);
            r.px_slope_recip_py = __builtin_fmaf(slope_recip, -py, px);
            r.packed = (order & 0xFFFFu) | (octant << 16) | (sc << 19);
        }
    }
    out[pos] = r;
    lengths[pos] = length;
}

// ------------------------------------------------------- path flattening

// PathSegments on the device (mold/src/path/segments.rs:51-263): one thread
// per path counts its lines, and after a scan of the counts writes them in
// order, with the serial t += 1/n recurrence of each curve.
__global__ __launch_bounds__(256) void k_path_counts(const PathCommand *__restrict__ cmds,
                                                     const flat::PathDesc *__restrict__ paths, uint32_t n_paths,
                                                     uint32_t *__restrict__ counts, uint32_t *__restrict__ err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_paths) return;
    const flat::PathDesc d = paths[i];
    flat::Xform xf;
    if (d.has_xf) xf.init(d.t9);
    const uint64_t n = flat::count_commands(cmds + d.first_cmd, d.n_cmds, d.has_xf ? &xf : nullptr);
    if (n > 0xFFFFFFFFull) atomicOr(err, 1u);
    counts[i] = n > 0xFFFFFFFFull ? 0u : (uint32_t)n;
}

__global__ __launch_bounds__(256) void k_path_emit(const PathCommand *__restrict__ cmds,
                                                   const flat::PathDesc *__restrict__ paths, uint32_t n_paths,
                                                   const uint32_t *__restrict__ offsets,
                                                   const uint32_t *__restrict__ counts, float4 *__restrict__ pts,
                                                   uint16_t *__restrict__ ids) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_paths) return;
    const flat::PathDesc d = paths[i];
    flat::Xform xf;
    if (d.has_xf) xf.init(d.t9);
    const uint64_t base = offsets[i], cnt = counts[i];
    uint64_t k = 0;
    flat::flatten_commands(cmds + d.first_cmd, d.n_cmds, d.has_xf ? &xf : nullptr,
                           [&](float x0, float y0, float x1, float y1) {
                               if (k < cnt) {  // the count pass bounds the writes
                                   pts[base + k] = make_float4(x0, y0, x1, y1);
                                   ids[base + k] = (uint16_t)d.layer;
                               }
                               ++k;
                           });
}

// ------------------------------------------------------------------- scan

// Single-pass exclusive scan of u32 values (decoupled look-back), 2048
// values per partition. out may alias in. total receives the sum (u64).
// With lim.src the scan covers only the first min(n, lim.mul *
// ceil(*lim.src / lim.div)) values: the partitions a frame's device-side
// count leaves (digit counts of a frame smaller than its buffers' capacity).
__global__ __launch_bounds__(256) void k_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n,
                                                  uint64_t *status, uint32_t *stall,
                                                  uint64_t *total, ScanLimit lim) {
    __shared__ uint32_t wsum[4];
    __shared__ uint64_t s_excl;
    constexpr uint32_t IPT = kScanTileHost / 256;
    static_assert(IPT % 4 == 0 && 256 * IPT == kScanTileHost, "whole uint4 per thread");
    if (lim.src) {
        const uint64_t m = (*lim.src + lim.div - 1) / lim.div * lim.mul;
        n = m < n ? m : n;
    }
    const uint32_t part = blockIdx.x;
    if ((uint64_t)part * 256 * IPT >= n && part) return;  // past the values (partition 0 writes total)
    const uint64_t base = (uint64_t)part * 256 * IPT + (uint64_t)threadIdx.x * IPT;
    // a thread's IPT values are contiguous: 16-byte loads and stores where
    // all IPT are in range (the buffers are 256-byte aligned)
    const bool whole = base + IPT <= n;
    uint32_t v[IPT], sum = 0;
    if (whole) {
#pragma unroll
        for (uint32_t q = 0; q < IPT / 4; ++q) {
            const uint4 a = reinterpret_cast<const uint4 *>(in + base)[q];
            v[4 * q] = a.x, v[4 * q + 1] = a.y, v[4 * q + 2] = a.z, v[4 * q + 3] = a.w;
        }
    } else {
        for (uint32_t k = 0; k < IPT; ++k) v[k] = base + k < n ? in[base + k] : 0u;
    }
    for (uint32_t k = 0; k < IPT; ++k) sum += v[k];
    uint32_t agg;
    uint32_t ex = block_excl_scan_256(sum, wsum, &agg);
    if (threadIdx.x == 0) st_release(status + part, (part == 0 ? kStInc : kStAgg) | agg);
    if (threadIdx.x < 64) {
        uint64_t e = part ? lookback_wave(status, part, stall) : 0;
        if (threadIdx.x == 0) {
            s_excl = e;
            if (part) st_release(status + part, kStInc | (e + agg));
            if ((uint64_t)(part + 1) * 256 * IPT >= n) *total = e + agg;
        }
    }
    __syncthreads();
    uint64_t run = s_excl + ex;
    uint32_t o[IPT];
    for (uint32_t k = 0; k < IPT; ++k) {
        o[k] = (uint32_t)run;
        run += v[k];
    }
    if (whole) {
#pragma unroll
        for (uint32_t q = 0; q < IPT / 4; ++q)
            reinterpret_cast<uint4 *>(out + base)[q] = make_uint4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
    } else {
        for (uint32_t k = 0; k < IPT; ++k)
            if (base + k < n) out[base + k] = o[k];
    }
}

// -------------------------------------------------------------- emission

// PixelSegment of one rasterized piece (raster_segment.rs:142-160), without
// branches on the octant (lanes of a wave hold different lines):
//   tiles()     (rasterizer/mod.rs:134-147): tile coordinates swapped for
//               octants 1, 2, 5, 6 (mask 0x66);
//   area_cover() (rasterizer/mod.rs:149-172): x/y swapped for the same
//               octants, the area axis reversed for octants 2, 6 (0x44) and
//               the cover axis for octants 3, 4, 5, 6 (0x78).
// The fields of one piece's PixelSegment, each truncated as its bit field
// of the u64 layout truncates it.
struct SegFields {
    int32_t tj;          // tile row: i16, stored in 12 bits
    int32_t tib;         // tile column + 256 (at least -1 + 256): i16, stored in 13 bits
    uint32_t layer;      // 16 bits
    uint32_t tx, ty;     // pixel in the tile, 3 bits each
    uint32_t area, cover;  // i16 and i8, stored in 10 and 6 bits
};

__device__ __forceinline__ SegFields seg_fields(uint32_t layer, int32_t x0, int32_t x1, int32_t y0, int32_t y1,
                                                int32_t bx, int32_t by, uint32_t octant) {
    const bool sw = (0x66u >> octant) & 1u;
    const bool fa = (0x44u >> octant) & 1u;
    const bool fb = (0x78u >> octant) & 1u;
    const int32_t tbx = sw ? by : bx, tby = sw ? bx : by;
    const int32_t ti = as_i16(tbx >> 3);
    const int32_t a0 = sw ? y0 : x0, a1 = sw ? y1 : x1;
    const int32_t c0 = sw ? x0 : y0, c1 = sw ? x1 : y1;
    const int32_t ax0 = fa ? a1 : a0, ax1 = fa ? a0 : a1;
    const int32_t ay0 = fb ? c1 : c0, ay1 = fb ? c0 : c1;
    const int32_t border = wrap_add(ax0 & ~15, 16);
    const int32_t h = wrap_sub(ay1, ay0);
    const int32_t tri = as_i16(wrap_mul(h, wrap_sub(ax1, ax0)) / 2);
    const int32_t rect = as_i16(wrap_mul(h, wrap_sub(border, ax1)));
    SegFields f;
    f.tj = as_i16(tby >> 3);
    f.tib = as_i16((ti < -1 ? -1 : ti) + 256);
    f.layer = layer & 0xFFFFu;
    f.tx = (uint32_t)(tbx & 7);
    f.ty = (uint32_t)(tby & 7);
    f.area = (uint32_t)as_i16(tri + rect);
    f.cover = (uint32_t)as_i8(h);
    return f;
}

// The PixelSegment (raster_segment.rs:142-160) of one piece.
__device__ __forceinline__ uint64_t pack_segment(const SegFields &f) {
    uint64_t w = ((uint64_t)(uint32_t)f.tj & 0xFFFull) << 51;
    w |= ((uint64_t)(uint32_t)f.tib & 0x1FFFull) << 38;
    w |= (uint64_t)f.layer << 22;
    w |= (uint64_t)f.ty << 19;
    w |= (uint64_t)f.tx << 16;
    w |= ((uint64_t)f.area & 0x3FFull) << 6;
    w |= (uint64_t)f.cover & 0x3Full;
    return w;
}

// Rasterizer::rasterize for one line-pixel (rasterizer/mod.rs:233-281): one
// piece, or two when the span crosses a pixel-row border. Both pieces are
// always computed (no divergence); the return value says how many count.
// `emit` turns a piece's fields into its slot word.
template <class Emit>
__device__ __forceinline__ int rasterize_pixel(const LineSetup &l, int32_t pi, uint64_t out[2], Emit &&emit) {
    const uint32_t layer = l.packed & 0xFFFFu, octant = (l.packed >> 16) & 7u;
    const int32_t x0i = wrap_add(l.start, pi);
    const int32_t x1i = wrap_add(x0i, 1);
    const float x0 = fmax_rs((float)x0i, l.start_f32);
    const float x1 = fmin_rs((float)x1i, l.end_f32);
    const float y0 = __builtin_fmaf(l.slope, x0, l.py_slope_px);
    const float y1 = __builtin_fmaf(l.slope, x1, l.py_slope_px);
    const int32_t x0s = sub_pixel_hw(x0), x1s = sub_pixel_hw(x1), y0s = sub_pixel_hw(y0), y1s = sub_pixel_hw(y1);
    const bool split = (y0s & 15) != 0 && (y1s & 15) != 0 && (y0s >> 4) != (y1s >> 4);
    const int32_t y = (y0s > y1s ? y0s : y1s) >> 4;
    const int32_t ys = (int32_t)((uint32_t)y << 4);
    const float x = __builtin_fmaf(l.slope_recip, (float)y, l.px_slope_recip_py);
    const int32_t xs = sub_pixel_hw(x);
    const int32_t ex = split ? xs : x1s, ey = split ? ys : y1s;
    out[0] = emit(seg_fields(layer, x0s, ex, y0s, ey, x0i, (ey < y0s ? ey : y0s) >> 4, octant));
    out[1] = emit(seg_fields(layer, xs, x1s, ys, y1s, x0i, (ys < y1s ? ys : y1s) >> 4, octant));
    return split ? 2 : 1;
}

constexpr uint32_t kSortThreads = 512;
constexpr uint32_t kSortWaves = kSortThreads / 64;
constexpr uint32_t kSortKPT = 16;
constexpr uint32_t kSortTile = kSortThreads * kSortKPT;  // keys per sort partition
// waves per SIMD the register allocation of a sort pass must allow (its LDS
// share already limits a CU to two work groups, four waves per SIMD)
constexpr uint32_t kSortWavesPerSimd = 4;
constexpr uint32_t kEmitThreads = 256;
constexpr uint32_t kEmitIPT = 8;
constexpr uint32_t kEmitTile = kEmitThreads * kEmitIPT;  // line-pixels of an emission partition
constexpr uint32_t kEmitWaveLp = kEmitTile / 4;          // line-pixels of one emission wave
constexpr uint32_t kSlotPart = 2 * kEmitTile;            // word slots of an emission partition
constexpr uint32_t kSlotParts = kSortTile / kSlotPart;   // emission partitions per sort partition
static_assert(kSortTile % kSlotPart == 0, "a sort partition holds whole emission partitions");
static_assert(2 * kEmitWaveLp == 64 * kSortKPT, "an emission wave's slots are one sort wave's");
static_assert(kEmitTile == kEmitTileHost && kSortTile == kSortTileHost && kSortThreads == kSortThreadsHost &&
                  kXcds == kXcdsHost,
              "host partition sizes");
// Slot sentinels (bit 63 set; no word has it: render words use 57 bits and
// emitted PixelSegments are never NONE).
constexpr uint64_t kSlotNone = ~0ull, kSlotCull = ~1ull;
constexpr uint32_t kMaxRadix = 2048;

// Render-mode filter + render key of one PixelSegment. Returns false when
// the segment can never influence the frame (rows outside the scene, tiles
// right of the crop / canvas).
__device__ __forceinline__ bool render_word(uint64_t w, const SceneDesc &S, const FrameDims &d, uint64_t *out,
                                            uint32_t *grow) {
    int32_t tj = (int32_t)((int64_t)(w << 1) >> 52);  // sign-extended 12-bit field
    if (tj < (int32_t)S.crop_v0 || tj >= (int32_t)S.crop_v1) return false;
    int32_t col = (int32_t)((w >> 38) & 0x1FFFu) - 256;
    if (col >= (int32_t)S.crop_h1) return false;
    uint32_t colp = col < 0 ? 0u : (uint32_t)col + 1u;
    uint32_t layer = (uint32_t)(w >> 22) & 0xFFFFu;
    uint32_t row = S.row_base + (uint32_t)tj;
    uint64_t key = ((uint64_t)row << (d.col_bits + d.layer_bits)) | ((uint64_t)colp << d.layer_bits) | layer;
    *out = (key << kPayloadBits) | (w & ((1ull << kPayloadBits) - 1));
    *grow = row;
    return true;
}

// The scene fields of the render-mode filter, and the filter itself: the
// render word, or kSlotCull when the segment can never influence the frame.
struct SceneCull {
    int32_t v0, v1, h1;
    uint32_t row_base;
};
__device__ __forceinline__ SceneCull scene_cull(const SceneDesc &S) {
    return SceneCull{(int32_t)S.crop_v0, (int32_t)S.crop_v1, (int32_t)S.crop_h1, S.row_base};
}
// The render word of a piece (render_word's key over the payload; rows
// outside the scene and tiles right of its crop culled) straight from its
// fields, truncated as the PixelSegment's bit fields truncate them (tile row
// to 12 bits, sign-extended; biased column to 13 bits).
__device__ __forceinline__ uint64_t render_word_fields(const SegFields &f, const SceneCull &c, const FrameDims &d) {
    const int32_t tj = (int32_t)((uint32_t)f.tj << 20) >> 20;  // sign-extended 12-bit field
    const int32_t col = (int32_t)((uint32_t)f.tib & 0x1FFFu) - 256;
    const uint32_t colp = col < 0 ? 0u : (uint32_t)col + 1u;
    const uint32_t row = c.row_base + (uint32_t)tj;
    const uint64_t key = ((uint64_t)row << (d.col_bits + d.layer_bits)) | ((uint64_t)colp << d.layer_bits) | f.layer;
    const uint32_t payload = f.ty << 19 | f.tx << 16 | (f.area & 0x3FFu) << 6 | (f.cover & 0x3Fu);
    const bool keep = tj >= c.v0 && tj < c.v1 && col < c.h1;
    return keep ? (key << kPayloadBits) | payload : kSlotCull;
}

// Emission: every line-pixel p owns the two slots 2p, 2p+1 of the word
// buffer, the layout of the reference's rasterizer output ([seg, seg|NONE],
// rasterizer/mod.rs:195, 272-279). A slot holds a word, kSlotNone (no second
// segment) or kSlotCull (a render-mode segment that can never reach the
// frame). Slot order is generation order (line, pixel, slot); the first sort
// pass drops the sentinels while it scatters, so no compaction pass and no
// cross-work-group prefix is needed here. One work group per kEmitTile
// consecutive line-pixels; thread t takes line-pixels b0 + t + 256 k, so each
// store instruction of a wave writes 64 consecutive slot pairs (1 KB).
// Lines of a partition staged in LDS (partitions with more lines read them
// from global memory). 256 keeps the emission's LDS near 22 KB, seven work
// groups per CU instead of four at 512 (3840x2160 frame: 2.60 -> 2.10 ms).
constexpr uint32_t kEmitLines = 256;
constexpr uint32_t kCountSlices = 256;  // accumulation slices of the kept and culled counts

// Per-partition counts of the word slots (a partition is kSortTile slots).
// LDS layout `s_cnt`: the first sort pass's digits, then kept words, then
// culled words. At the end of the partition the digit counts go to
// cnt0[digit * nparts + part] (scanned into the first pass's output offsets)
// and the kept and culled counts are added into slice part % kCountSlices of
// `acc` (reduced to the frame counters). The later passes count their digits
// from the byte digits the pass before them writes (k_digit_count).

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ uint32_t count_rows(const FrameDims &d) { return (1u << d.radix_bits) + 2; }

__device__ __forceinline__ void count_zero(uint32_t *s_cnt, const FrameDims &d) {
    for (uint32_t k = threadIdx.x; k < count_rows(d); k += blockDim.x) s_cnt[k] = 0;
}

// Counts the slot of every active lane: its first-pass digit (one LDS atomic
// per kept word), kept words in the thread's `kept`, culled words once per
// wave.
__device__ __forceinline__ void count_slot(uint32_t *s_cnt, uint64_t v, const FrameDims &d, uint32_t &kept) {
    const uint32_t radix = 1u << d.radix_bits;
    const uint32_t lane = lane_id();
    const uint64_t exec = __ballot(true);
    const uint32_t leader0 = (uint32_t)__ffsll((long long)exec) - 1u;
    const uint64_t culled = __ballot(v == kSlotCull);
    if (culled && lane == leader0) atomicAdd(&s_cnt[radix + 1], (uint32_t)__popcll(culled));
    const bool ok = !(v >> 63);
    kept += ok ? 1u : 0u;
    if (ok) atomicAdd(&s_cnt[(uint32_t)(v >> d.sort_shift) & (radix - 1)], 1u);
}

// count_slot for the two slots of every active lane's line-pixel at once.
__device__ __forceinline__ void count_slot_pair(uint32_t *s_cnt, uint64_t v0, uint64_t v1, const FrameDims &d,
                                                uint32_t &kept) {
    const uint32_t radix = 1u << d.radix_bits;
    const uint32_t lane = lane_id();
    const uint64_t exec = __ballot(true);
    const uint32_t leader0 = (uint32_t)__ffsll((long long)exec) - 1u;
    const uint32_t nc = (uint32_t)(__popcll(__ballot(v0 == kSlotCull)) + __popcll(__ballot(v1 == kSlotCull)));
    if (nc && lane == leader0) atomicAdd(&s_cnt[radix + 1], nc);
    const bool ok0 = !(v0 >> 63), ok1 = !(v1 >> 63);
    kept += (ok0 ? 1u : 0u) + (ok1 ? 1u : 0u);
    if (ok0) atomicAdd(&s_cnt[(uint32_t)(v0 >> d.sort_shift) & (radix - 1)], 1u);
    if (ok1) atomicAdd(&s_cnt[(uint32_t)(v1 >> d.sort_shift) & (radix - 1)], 1u);
}

// Called by the whole work group after its last count_slot.
__device__ __forceinline__ void count_flush(uint32_t *s_cnt, const FrameDims &d, const SlotCounts &sc,
                                            uint32_t part, uint32_t nparts, uint32_t kept) {
    const uint32_t radix = 1u << d.radix_bits;
    const uint32_t kw = wave_sum(kept);
    if (lane_id() == 0 && kw) atomicAdd(&s_cnt[radix], kw);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < radix; k += blockDim.x) sc.cnt0[(uint64_t)k * nparts + part] = s_cnt[k];
    uint32_t *a = sc.acc + (uint64_t)(part % kCountSlices) * 2;
    if (threadIdx.x < 2 && s_cnt[radix + threadIdx.x]) atomicAdd(&a[threadIdx.x], s_cnt[radix + threadIdx.x]);
}

// Line-pixels b0 + threadIdx.x + 256 k of [.., b1) (see k_emit); owner[i]
// is the line (relative to l0) holding line-pixel b0 + i. With Staged,
// `offs`, `lines` and `culls` are LDS tables of lines l0 .. (indexed
// l - l0); otherwise `offs` and `lines` are the global arrays and the cull
// bounds come from the scene table.
template <bool Staged>
__device__ __forceinline__ uint32_t emit_span(const LineSetup *lines, const SceneCull *culls, const uint32_t *offs,
                                          uint32_t l0, uint32_t l1,
                                          uint64_t b0, uint64_t b1, uint64_t P,
                                          const SceneDesc *__restrict__ scenes, const FrameDims &dims,
                                          uint64_t *__restrict__ dst, uint32_t *s_cnt, const uint32_t *owner,
                                          uint32_t &kept) {
    const uint32_t base = Staged ? l0 : 0u;
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t lt = lanemask_lt();
    // wave w: line-pixels b0 + kEmitWaveLp w .. in steps of 64; its kept
    // words, in slot order, go to the front of its 2 kEmitWaveLp slots
    const uint64_t wb0 = b0 + (uint64_t)kEmitWaveLp * w;
    const uint32_t nsteps = wb0 < b1 ? (uint32_t)((b1 - wb0 < kEmitWaveLp ? b1 - wb0 : kEmitWaveLp) + 63) / 64 : 0u;
    uint64_t *wdst = dst + 2 * (uint64_t)kEmitWaveLp * w;
    uint32_t written = 0;  // kept words of the wave so far
    for (uint32_t k = 0; k < nsteps; ++k) {
        const uint64_t p = wb0 + 64 * k + lane;
        uint64_t v[2] = {kSlotNone, kSlotNone};
        if (p < b1) {
            const uint32_t li = l0 + owner[p - b0];
            const uint64_t first = offs[li - base];
            const LineSetup L = lines[li - base];
            const SceneCull sc = Staged ? culls[li - base] : scene_cull(scenes[L.packed >> 19]);
            uint64_t wd[2];
            // render frames: the render word straight from the fields; the
            // canonical stream: the PixelSegment itself
            const int n = dims.mode == 0
                              ? rasterize_pixel(L, (int32_t)(p - first), wd,
                                                [&](const SegFields &f) { return render_word_fields(f, sc, dims); })
                              : rasterize_pixel(L, (int32_t)(p - first), wd,
                                                [](const SegFields &f) { return pack_segment(f); });
            v[0] = wd[0];
            v[1] = wd[1];
            if (n < 2) v[1] = kSlotNone;
            count_slot_pair(s_cnt, v[0], v[1], dims, kept);
        }
        // the step's kept words in slot order (line-pixel, then slot)
        const bool k0 = !(v[0] >> 63), k1 = !(v[1] >> 63);
        const uint64_t m0 = __ballot(k0), m1 = __ballot(k1);
        const uint32_t pos = written + (uint32_t)(__popcll(m0 & lt) + __popcll(m1 & lt));
        if (k0) wdst[pos] = v[0];
        if (k1) wdst[pos + (k0 ? 1u : 0u)] = v[1];
        written += (uint32_t)(__popcll(m0) + __popcll(m1));
    }
    return written;
}

__global__ __launch_bounds__(256) void k_emit(const LineSetup *__restrict__ lines,
                                              const uint32_t *__restrict__ offsets,
                                              const SceneDesc *__restrict__ scenes, FrameDims dims,
                                              const FrameCounters *__restrict__ counters,
                                              uint64_t *__restrict__ out, const uint32_t *__restrict__ part_line,
                                              SlotCounts sc) {
    __shared__ uint32_t s_off[kEmitLines + 1];
    __shared__ LineSetup s_lines[kEmitLines];
    __shared__ SceneCull s_culls[kEmitLines];
    __shared__ uint32_t s_owner[kEmitTile];
    __shared__ uint32_t s_wmax[4];
    extern __shared__ uint32_t s_cnt[];
    const uint64_t P = counters->line_pixels;
    const uint32_t nparts = (uint32_t)((P + kEmitTile - 1) / kEmitTile);
    uint32_t part;
    if (!xcd_part(nparts, part)) return;
    count_zero(s_cnt, dims);
    const uint64_t b0 = (uint64_t)part * kEmitTile;
    const uint64_t b1 = b0 + kEmitTile < P ? b0 + kEmitTile : P;
    // lines covering [b0, b1): from the line holding the partition's first
    // line-pixel to the one holding the next partition's (a superset)
    const uint32_t l0 = part_line[part];
    const uint32_t l1 = part + 1 < nparts ? part_line[part + 1] : dims.n_lines - 1;
    const uint32_t nl = l1 - l0 + 1;
    // the partition's kept words go to the front of its 2 x kEmitTile slots
    uint64_t *dst = out + 2 * b0;
    uint32_t kept = 0, written = 0;
    // owner table: the first line-pixel of every line of nonzero length that
    // starts inside the partition gets the line's index (relative to l0,
    // which holds the partition's first line-pixel); a running maximum fills
    // the positions in between. Lines after l0 start after b0.
    for (uint32_t i = threadIdx.x; i < kEmitTile; i += 256) s_owner[i] = 0;
    __syncthreads();
    for (uint32_t k = threadIdx.x + 1; k < nl; k += 256) {
        const uint64_t o = offsets[l0 + k];
        const uint64_t e = l0 + k + 1 < dims.n_lines ? (uint64_t)offsets[l0 + k + 1] : P;
        if (e > o && o < b1) s_owner[o - b0] = k;
    }
    __syncthreads();
    {
        constexpr uint32_t E = kEmitTile / 256;
        uint32_t v[E], m = 0;
#pragma unroll
        for (uint32_t j = 0; j < E; ++j) {
            m = max(m, s_owner[threadIdx.x * E + j]);
            v[j] = m;
        }
        const uint32_t wm = wave_incl_max(m);
        uint32_t before = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)wm, 0x138, 0xF, 0xF, false);  // wave_shr:1
        if (lane_id() == 63) s_wmax[threadIdx.x >> 6] = wm;
        __syncthreads();
        for (uint32_t k = 0; k < (threadIdx.x >> 6); ++k) before = max(before, s_wmax[k]);
#pragma unroll
        for (uint32_t j = 0; j < E; ++j) s_owner[threadIdx.x * E + j] = max(before, v[j]);
    }
    if (nl <= kEmitLines) {
        for (uint32_t k = threadIdx.x; k <= nl; k += blockDim.x)
            s_off[k] = l0 + k < dims.n_lines ? offsets[l0 + k] : (uint32_t)P;
        for (uint32_t k = threadIdx.x; k < nl; k += blockDim.x) {
            const LineSetup L = lines[l0 + k];
            s_lines[k] = L;
            s_culls[k] = scene_cull(scenes[L.packed >> 19]);
        }
        __syncthreads();
        written = emit_span<true>(s_lines, s_culls, s_off, l0, l1, b0, b1, P, scenes, dims, dst, s_cnt, s_owner,
                                  kept);
    } else {  // a partition of very short or zero-length lines
        __syncthreads();
        written = emit_span<false>(lines, nullptr, offsets, l0, l1, b0, b1, P, scenes, dims, dst, s_cnt, s_owner,
                                   kept);
    }
    if (lane_id() == 0) sc.part_kept[4 * part + (threadIdx.x >> 6)] = written;
    count_flush(s_cnt, dims, sc, part, nparts, kept);
}

// Totals of the kept and culled words (into the frame counters): one work
// group per count.
__global__ __launch_bounds__(256) void k_count_reduce(const uint32_t *__restrict__ acc,
                                                      FrameCounters *__restrict__ counters) {
    __shared__ uint32_t wsum[4];
    const uint32_t r = blockIdx.x;
    if (r >= 2) return;
    uint32_t v = 0;
    for (uint32_t k = threadIdx.x; k < kCountSlices; k += 256) v += acc[(uint64_t)k * 2 + r];
    uint32_t all;
    block_excl_scan_256(v, wsum, &all);
    if (threadIdx.x != 0) return;
    if (r == 0) counters->kept = all;
    atomicAdd((unsigned long long *)&counters->segments, (unsigned long long)all);
}

// part_line[p] = the line whose line-pixels contain p * part_size.
// (Partitions from n_parts on are not emitted: a frame replayed with a
// line-pixel capacity stops there.)
__global__ __launch_bounds__(256) void k_part_lines(const uint32_t *__restrict__ offsets,
                                                    const uint32_t *__restrict__ lengths, uint64_t n_lines,
                                                    uint32_t part_size, uint32_t *__restrict__ part_line,
                                                    uint64_t n_parts) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= n_lines) return;
    const uint32_t len = lengths[i];
    if (!len) return;
    const uint64_t off = offsets[i];
    const uint64_t p0 = (off + part_size - 1) / part_size, p1 = (off + len - 1) / part_size;
    for (uint64_t p = p0; p <= p1 && p < n_parts; ++p) part_line[p] = (uint32_t)i;
}

// Frames replayed from a captured graph have buffers and grids sized for a
// line-pixel capacity: a frame beyond it is cut to the capacity (every later
// kernel reads the count from the counters) and flagged, and the host
// renders it again without the graph.
__global__ void k_clamp_line_pixels(FrameCounters *__restrict__ counters, uint64_t cap,
                                    uint32_t *__restrict__ flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && counters->line_pixels > cap) {
        counters->line_pixels = cap;
        *flag = 1u;
    }
}

// First sorted word of every global tile row (row_start[rows] = kept): one
// binary search per row over the sorted render words.
__global__ void k_row_starts(const uint64_t *__restrict__ sorted, const FrameCounters *__restrict__ counters,
                             uint32_t rows, uint32_t row_shift, uint32_t *__restrict__ row_start) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > rows) return;
    const uint64_t n = counters->kept;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if ((sorted[mid] >> row_shift) < r) lo = mid + 1; else hi = mid;
    }
    row_start[r] = (uint32_t)lo;
}

// Ingest of an external PixelSegment stream (the cassia_render entry,
// cassia/src/Cassia.h:34-39): NONE entries and segments of layers without
// styling are dropped, the rest become render words in input order. Scene 0.
__global__ __launch_bounds__(256) void k_ingest(const uint64_t *__restrict__ in, uint64_t n_in,
                                                const SceneDesc *__restrict__ scenes, FrameDims dims,
                                                const FrameCounters *__restrict__ counters,
                                                uint64_t *__restrict__ out, SlotCounts sc) {
    extern __shared__ uint32_t s_cnt[];
    const uint64_t n_slots = 2 * counters->line_pixels;
    const uint32_t nparts = (uint32_t)((n_slots + kSlotPart - 1) / kSlotPart);  // emission-sized
    const uint32_t part = blockIdx.x;
    if (part >= nparts) return;
    count_zero(s_cnt, dims);
    __syncthreads();
    const SceneDesc &S = scenes[0];
    uint32_t kept = 0;
    for (uint32_t k = 0; k < kSlotPart / 256; ++k) {
        const uint64_t i = (uint64_t)part * kSlotPart + k * 256 + threadIdx.x;
        if (i >= n_slots) break;
        const uint64_t w = i < n_in ? in[i] : kNone;
        uint64_t v = kSlotNone;
        if (!(w >> 63)) {
            uint64_t rw;
            uint32_t grow;
            v = kSlotCull;
            if (((uint32_t)(w >> 22) & 0xFFFFu) < S.n_props && render_word(w, S, dims, &rw, &grow)) v = rw;
        }
        out[i] = v;
        count_slot(s_cnt, v, dims, kept);
    }
    // the first pass reads every slot of the partition (sentinels included)
    if (lane_id() == 0) {
        const uint64_t b = (uint64_t)part * kSlotPart + 2 * (uint64_t)kEmitWaveLp * (threadIdx.x >> 6);
        sc.part_kept[4 * part + (threadIdx.x >> 6)] =
            b < n_slots ? (uint32_t)(n_slots - b < 2 * kEmitWaveLp ? n_slots - b : 2 * kEmitWaveLp) : 0u;
    }
    count_flush(s_cnt, dims, sc, part, nparts, kept);
}

// ------------------------------------------------------------------- sort
//
// LSD radix sort of the words, R bits per pass, stable. Pass p scatters the
// words of each partition (kSortTile consecutive input positions) to
// offsets[d * nparts + part]: an exclusive scan, digit-major, of the
// partitions' digit counts. Pass 0's counts come from the emission; each
// later pass's come from k_digit_count over the digits the pass before it
// stored beside its output (one byte per word, two for 9-bit digits). So no
// partition waits on another: the offsets are complete before a pass starts,
// and the partitions are dealt to the XCDs in contiguous ranges (xcd_part).
// A partition writes one run per digit; a run's first and last cache lines
// are shared with the neighbouring partitions' runs of that digit, and with
// the neighbours on the same XCD those lines are completed in its L2 (on
// 3840x2160 render words a pass took 1.30 ms with partitions in index order
// and 0.95 ms dealt this way; 4,096-key partitions 1.01 ms).

// Digit counts of the words the previous pass wrote: partition `part` is
// positions [part * kSortTile, ..) of its output, n = counters->kept words;
// cnt[d * nparts + part] for every partition up to nparts = ceil(n / kSortTile).
template <int R>
__global__ __launch_bounds__(256) void k_digit_count(const typename DigitOf<R>::T *__restrict__ dig,
                                                     const FrameCounters *__restrict__ counters,
                                                     uint32_t *__restrict__ cnt) {
    using T = typename DigitOf<R>::T;
    constexpr uint32_t RADIX = 1u << R;
    constexpr uint32_t DPT = kSortTile / 256;        // digits per thread
    constexpr uint32_t V = DPT * sizeof(T) / 16;     // 16-byte loads per thread
    constexpr uint32_t PER = 16 / sizeof(T);         // digits per load
    static_assert(DPT * sizeof(T) % 16 == 0, "whole 16-byte loads");
    __shared__ uint32_t s_h[4][RADIX];
    const uint64_t n = counters->kept;
    const uint32_t nparts = (uint32_t)((n + kSortTile - 1) / kSortTile);
    uint32_t part;
    if (!xcd_part(nparts, part)) return;
    for (uint32_t k = threadIdx.x; k < 4 * RADIX; k += 256) (&s_h[0][0])[k] = 0;
    const uint64_t base = (uint64_t)part * kSortTile;
    const uint32_t m = n - base < kSortTile ? (uint32_t)(n - base) : kSortTile;
    // the buffer is padded to whole partitions: the loads stay inside it.
    // A thread counts DPT consecutive digits as runs of equal digits (words
    // at consecutive output positions mostly come from one line in one tile
    // and share their next digit), one LDS add per run.
    const uint4 *src = (const uint4 *)(dig + base) + threadIdx.x * V;
    uint4 v[V];
#pragma unroll
    for (uint32_t j = 0; j < V; ++j) v[j] = src[j];
    __syncthreads();
    uint32_t *h = s_h[threadIdx.x >> 6];
    const uint32_t i0 = threadIdx.x * DPT;
    const uint32_t lim = m > i0 ? (m - i0 < DPT ? m - i0 : DPT) : 0u;  // digits of this thread inside n
    uint32_t prev = 0xFFFFFFFFu, run = 0;
#pragma unroll
    for (uint32_t j = 0; j < V; ++j) {
        const uint32_t wd[4] = {v[j].x, v[j].y, v[j].z, v[j].w};
#pragma unroll
        for (uint32_t e = 0; e < PER; ++e) {
            const uint32_t d = (wd[e / (4 / sizeof(T))] >> (8 * sizeof(T) * (e % (4 / sizeof(T))))) & (RADIX - 1);
            if (j * PER + e < lim) {
                if (d != prev) {
                    if (run) atomicAdd(&h[prev], run);
                    prev = d;
                    run = 0;
                }
                ++run;
            }
        }
    }
    if (run) atomicAdd(&h[prev], run);
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < RADIX; d += 256)
        cnt[(uint64_t)d * nparts + part] = s_h[0][d] + s_h[1][d] + s_h[2][d] + s_h[3][d];
}

template __global__ void k_digit_count<4>(const uint8_t *, const FrameCounters *, uint32_t *);
template __global__ void k_digit_count<6>(const uint8_t *, const FrameCounters *, uint32_t *);
template __global__ void k_digit_count<8>(const uint8_t *, const FrameCounters *, uint32_t *);
template __global__ void k_digit_count<9>(const uint16_t *, const FrameCounters *, uint32_t *);

// One stable digit pass. Each wave ranks its 64 * kSortKPT keys in kSortKPT
// steps of 64 (striped, so step-major order equals input order) with
// ballot-based peer matching; the block's keys are reordered by digit in LDS
// and stored as digit runs at their offsets. With part_kept the input is the
// emission's slots: every emission wave's words lie at the front of its
// 2 kEmitWaveLp slots (part_kept[4 p + e] of them for wave e of emission
// partition p); without it the input is
// the previous pass's counters->kept words. With dig_out each word's digit
// of the next pass (bits shift + R ..) is stored at its output position.
template <int R>
__global__ __launch_bounds__(kSortThreads) __attribute__((amdgpu_waves_per_eu(kSortWavesPerSimd))) void k_sort_pass(
    const uint64_t *__restrict__ in, uint64_t *__restrict__ out, const FrameCounters *__restrict__ counters,
    uint32_t shift, const uint32_t *__restrict__ offsets, const uint32_t *__restrict__ part_kept,
    typename DigitOf<R>::T *__restrict__ dig_out) {
    using T = typename DigitOf<R>::T;
    constexpr uint32_t RADIX = 1u << R;
    static_assert(RADIX <= kSortThreads, "one digit per thread");
    __shared__ uint64_t s_keys[kSortTile];
    __shared__ uint32_t s_wcnt[kSortWaves][RADIX];
    __shared__ uint32_t s_dexcl[RADIX];
    __shared__ uint32_t s_glob[RADIX];
    __shared__ uint32_t wsum[kSortWaves];
    const uint64_t n = part_kept ? 2 * counters->line_pixels : counters->kept;
    const uint32_t nparts = (uint32_t)((n + kSortTile - 1) / kSortTile);
    uint32_t part;
    if (!xcd_part(nparts, part)) return;  // uniform over the work group
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    for (uint32_t k = threadIdx.x; k < kSortWaves * RADIX; k += kSortThreads) (&s_wcnt[0][0])[k] = 0;
    // the partition's output offset of each digit (thread = digit): the
    // counts of the slots are per emission partition (kSlotParts to a sort
    // partition), those of a dense input per sort partition
    const uint32_t ostride = part_kept ? (uint32_t)((n + kSlotPart - 1) / kSlotPart) : nparts;
    const uint32_t opart = part_kept ? part * kSlotParts : part;
    const uint32_t glob = threadIdx.x < RADIX ? offsets[(uint64_t)threadIdx.x * ostride + opart] : 0u;
    __syncthreads();
    const uint64_t base = (uint64_t)part * kSortTile;
    const uint32_t valid_n = (uint32_t)(n - base < kSortTile ? n - base : kSortTile);
    // in the slots, this wave's 64 kSortKPT positions are one emission
    // wave's 2 kEmitWaveLp slots, whose leading part_kept[..] hold words
    static_assert(kSortWaves == 4 * kSlotParts, "a sort wave per emission wave");
    const uint32_t wave_lim = !part_kept ? valid_n
                              : w * (64 * kSortKPT) < valid_n ? w * (64 * kSortKPT) + part_kept[part * kSortWaves + w]
                                                              : 0u;
    uint64_t keys[kSortKPT];
    uint32_t rank[kSortKPT];
    // all loads in flight before the first ranking step (bit 63 marks a
    // dropped slot, and the tail past the input)
#pragma unroll
    for (uint32_t i = 0; i < kSortKPT; ++i) {
        const uint32_t li = w * (64 * kSortKPT) + i * 64 + lane;
        keys[i] = li < valid_n && li < wave_lim ? in[base + li] : ~0ull;
    }
    // steps past the wave's last word (the slots of the first pass) rank nothing
    const uint32_t wave_end = (wave_lim < valid_n ? wave_lim : valid_n);
#pragma unroll
    for (uint32_t i = 0; i < kSortKPT; ++i) {
        if (w * (64 * kSortKPT) + i * 64 >= wave_end) {  // uniform over the wave
            rank[i] = 0xFFFFFFFFu;
            continue;
        }
        bool valid = true;
        uint64_t key = keys[i];
        if (key >> 63) {
            valid = false;
            key = 0;
        }
        const uint32_t d = (uint32_t)(key >> shift) & (RADIX - 1);
        // lanes with the same digit: for each bit, the lanes whose bit
        // equals this lane's (m: the bit spread over 32 bits)
        const uint64_t vb = __ballot(valid);
        uint32_t pl = (uint32_t)vb, ph = (uint32_t)(vb >> 32);
#pragma unroll
        for (int b = 0; b < R; ++b) {
            const uint32_t m = (uint32_t)((int32_t)(d << (31 - b)) >> 31);
            const uint64_t bits = __ballot(m != 0u);
            pl &= ~((uint32_t)bits ^ m);
            ph &= ~((uint32_t)(bits >> 32) ^ m);
        }
        const uint64_t peers = (uint64_t)ph << 32 | pl;
        const uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
        const uint32_t c = valid ? s_wcnt[w][d] : 0u;
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) s_wcnt[w][d] = c + (uint32_t)__popcll(peers);
        __builtin_amdgcn_wave_barrier();
        keys[i] = key;
        rank[i] = valid ? c + below : 0xFFFFFFFFu;
    }
    __syncthreads();
    // per digit: the waves' exclusive prefixes, the block's digit totals and
    // their exclusive scan
    uint32_t tot = 0;
    if (threadIdx.x < RADIX) {
        for (uint32_t ww = 0; ww < kSortWaves; ++ww) {
            const uint32_t v = s_wcnt[ww][threadIdx.x];
            s_wcnt[ww][threadIdx.x] = tot;
            tot += v;
        }
    }
    uint32_t all;
    const uint32_t ex = block_excl_scan_u32<kSortWaves>(tot, wsum, &all);
    if (threadIdx.x < RADIX) {
        s_dexcl[threadIdx.x] = ex;
        s_glob[threadIdx.x] = glob - ex;  // output position of block-local index k: s_glob[d] + k
    }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < kSortKPT; ++i) {
        if (rank[i] == 0xFFFFFFFFu) continue;
        const uint32_t d = (uint32_t)(keys[i] >> shift) & (RADIX - 1);
        s_keys[s_dexcl[d] + s_wcnt[w][d] + rank[i]] = keys[i];
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < all; k += kSortThreads) {
        const uint64_t key = s_keys[k];
        const uint32_t g = s_glob[(uint32_t)(key >> shift) & (RADIX - 1)] + k;
        out[g] = key;
        if (dig_out) dig_out[g] = (T)((uint32_t)(key >> (shift + R)) & (RADIX - 1));
    }
}

template __global__ void k_sort_pass<4>(const uint64_t *, uint64_t *, const FrameCounters *, uint32_t,
                                        const uint32_t *, const uint32_t *, uint8_t *);
template __global__ void k_sort_pass<6>(const uint64_t *, uint64_t *, const FrameCounters *, uint32_t,
                                        const uint32_t *, const uint32_t *, uint8_t *);
template __global__ void k_sort_pass<8>(const uint64_t *, uint64_t *, const FrameCounters *, uint32_t,
                                        const uint32_t *, const uint32_t *, uint8_t *);
template __global__ void k_sort_pass<9>(const uint64_t *, uint64_t *, const FrameCounters *, uint32_t,
                                        const uint32_t *, const uint32_t *, uint16_t *);

// ------------------------------------------------------------------ paint
//
// Four kernels over the sorted render words. Tile rows are independent, and a
// tile's words are contiguous and sorted by layer.
//
// k_tile_starts  per (tile row, tile column): index of the column's first
//                word (binary search in the row).
// k_tile_runs    per tile column, one wave: the runs (maximal same-layer
//                groups) with their 8 per-row cover sums, from a segmented
//                byte-wise scan in registers. Run r of the column is stored at
//                index (column's first word + r), so no allocation is needed.
// k_row_walk     per tile row, one 256-thread work group: the cover queue of
//                LayerWorkbench (layer_workbench.rs:134-198, 449-485) carried
//                left to right. Per tile it merges the queue with the tile's
//                runs (merge path) and stores the tile's id list
//                (populate_layers, layer_workbench.rs:401-428) as records
//                (layer, run, carried cover) in layer order.
// k_paint        per tile, one wave (lane = pixel x*8+y): the workbench passes
//                and the painter (painter/mod.rs:229-493) over the tile's
//                records in batches of 64, with the batch's segments staged
//                in LDS; no barrier per layer.

constexpr uint32_t kMaskBit = 1u << 30, kSkipClipBit = 1u << 31;
constexpr uint32_t kRunEvenOdd = 1u << 16;                    // run key: layer | even_odd << 16
constexpr uint32_t kRecRun = 1u << 16, kRecCarry = 1u << 17;  // record key flags
constexpr uint32_t kRecEvenOdd = 1u << 18;                     // record key: the layer's fill rule is even-odd
constexpr uint32_t kStageWin = 256;                           // staged words per paint wave
// Run groups of the paint (group_build): slots of kGroupStride words, runs of
// at most kGroupLenMax words. A stride of 8 mod 32 words puts the 8 slots'
// rows in different banks for the prefix pass.
constexpr uint32_t kGroupRuns = 8, kGroupStride = 72, kGroupLenMax = 256;
// stage window + run-group slots of one wave (the per-run cells share the
// slots): 4.25 KB, so that LDS allows more than the 8 waves per SIMD that the
// paint's 64 VGPRs allow
constexpr uint32_t kPaintWaveLds = kStageWin * 8 + kGroupRuns * kGroupStride * 4;
static_assert(kGroupRuns * kGroupStride >= 2 * 64 && kGroupRuns <= 8, "per-run cells fit in the group slots");
constexpr uint32_t kDeepFlag = 1u << 31;                      // record header: tile of the deep list
constexpr uint32_t kDeepChunk = 192;                          // deep paint: records per round

size_t walk_lds_bytes_host(uint32_t cap, uint32_t rcap, uint32_t tstride, uint32_t run_bufs) {
    return (size_t)cap * 24 + (size_t)rcap * 16 * run_bufs + (size_t)tstride * 12 + 64;
}
size_t paint_lds_bytes_host(uint32_t cap, bool clips) {
    return (size_t)kPaintWaveLds + (clips ? (size_t)cap * 4 : 0);
}

__device__ __forceinline__ int32_t cov_byte(uint64_t c, uint32_t y) { return as_i8((int32_t)(c >> (8 * y))); }

// Byte-wise wrapping add of 8 packed i8 row covers.
__device__ __forceinline__ uint64_t add_bytes(uint64_t a, uint64_t b) {
    constexpr uint64_t H = 0x8080808080808080ull;
    return ((a & ~H) + (b & ~H)) ^ ((a ^ b) & H);
}

__device__ __forceinline__ uint32_t rlane(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ float rlane_f(float v, uint32_t l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)l));
}
__device__ __forceinline__ uint64_t rlane64(uint64_t v, uint32_t l) {
    return (uint64_t)rlane((uint32_t)v, l) | ((uint64_t)rlane((uint32_t)(v >> 32), l) << 32);
}

// Cover::is_empty / is_full over 8 rows (painter/mod.rs:164-189), byte-wise
// on the packed i8 covers. |v| & 31 == 0 holds exactly when v is a multiple
// of 32, i.e. when the byte's low 5 bits are zero (-128 included), and
// |v| & 31 == 16 when they equal 16. |v| == 16 means the byte is 0x10 or
// 0xF0: low 5 bits 16 and bits 7..5 all equal.
__device__ __forceinline__ bool cover_empty(uint64_t c, bool eo) {
    return (c & (eo ? 0x1F1F1F1F1F1F1F1Full : ~0ull)) == 0;
}
__device__ __forceinline__ bool cover_full(uint64_t c, bool eo) {
    if ((c & 0x1F1F1F1F1F1F1F1Full) != 0x1010101010101010ull) return false;
    return eo || ((((c & 0xE0E0E0E0E0E0E0E0ull) >> 1) ^ c) & 0x6060606060606060ull) == 0;
}

// Gradient::color_at for pixel row `row` of the column at (gx, gy)
// (style.rs:96-175).
__device__ void gradient_color(const PropsEntry &p, const float *stops, float gx, float gy, uint32_t row,
                               float out[4]) {
    const float dx = p.grad_end[0] - p.grad_start[0];
    const float dy = p.grad_end[1] - p.grad_start[1];
    const float dot = dx * dx + dy * dy;
    const float dot_recip = 1.0f / dot;
    float t;
    if (!(p.flags & PF_RADIAL)) {
        const float tx = (gx - p.grad_start[0]) * dx * dot_recip;
        const float ty = gy - p.grad_start[1];
        t = __builtin_fmaf(((float)row + ty) * dy, dot_recip, tx);
    } else {
        const float px = gx - p.grad_start[0];
        const float px2 = px * px;
        const float py = (float)row + (gy - p.grad_start[1]);
        t = __builtin_sqrtf(__builtin_fmaf(py, py, px2) * dot_recip);
    }
    const float *st = stops + p.stops_offset;
    uint32_t ch0 = 0, ch1 = 0, ch2 = 0, ch3 = 0;
    bool acc = t <= st[4];
    if (acc) {
        ch0 |= __float_as_uint(st[0]); ch1 |= __float_as_uint(st[1]);
        ch2 |= __float_as_uint(st[2]); ch3 |= __float_as_uint(st[3]);
    }
    float start_stop = 0.0f;
    const float *sc = st;
    for (uint32_t k = 1; k < p.n_stops; ++k) {
        const float *color = st + 5 * k;
        const float end_stop = color[4];
        const bool m = acc ^ (t < end_stop);
        if (m) {
            const float d = end_stop - start_stop;
            const float lt = (t - start_stop) * (1.0f / d);
            ch0 |= __float_as_uint(__builtin_fmaf(lt, color[0], __builtin_fmaf(-lt, sc[0], sc[0])));
            ch1 |= __float_as_uint(__builtin_fmaf(lt, color[1], __builtin_fmaf(-lt, sc[1], sc[1])));
            ch2 |= __float_as_uint(__builtin_fmaf(lt, color[2], __builtin_fmaf(-lt, sc[2], sc[2])));
            ch3 |= __float_as_uint(__builtin_fmaf(lt, color[3], __builtin_fmaf(-lt, sc[3], sc[3])));
            acc = true;
        }
        start_stop = end_stop;
        sc = color;
    }
    if (!acc) {
        const float *last = st + 5 * (p.n_stops - 1);
        ch0 |= __float_as_uint(last[0]); ch1 |= __float_as_uint(last[1]);
        ch2 |= __float_as_uint(last[2]); ch3 |= __float_as_uint(last[3]);
    }
    out[0] = __uint_as_float(ch0);
    out[1] = __uint_as_float(ch1);
    out[2] = __uint_as_float(ch2);
    out[3] = __uint_as_float(ch3);
}

__device__ __forceinline__ uint32_t seg_colp(uint64_t w, uint32_t lb, uint32_t cmask) {
    return (uint32_t)(w >> (kPayloadBits + lb)) & cmask;
}

// Painted column range [h0, h1) of a scene's rows (crop in tiles).
__device__ __forceinline__ uint32_t scene_h1(const SceneDesc &S) {
    return S.crop_h1 < S.row_len ? S.crop_h1 : S.row_len;
}

__global__ __launch_bounds__(256) void k_tile_starts(PaintParams pp, uint32_t global_rows, uint32_t stream_runs) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= global_rows * pp.tstride) return;
    const uint32_t grow = i / pp.tstride, k = i - grow * pp.tstride;
    const uint32_t lb = pp.layer_bits, cmask = (1u << pp.col_bits) - 1u;
    uint32_t lo = pp.row_start[grow], hi = pp.row_start[grow + 1];
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (seg_colp(pp.segs[mid], lb, cmask) < k) lo = mid + 1; else hi = mid;
    }
    pp.tile_start[i] = lo;
    if (!stream_runs) {
        pp.tile_run0[i] = lo;  // k_tile_runs stores a tile's runs from its first word on
        return;
    }
    // the tile's first run: the first run starting at or after its first word
    uint32_t rlo = 0, rhi = *pp.n_runs;
    while (rlo < rhi) {
        const uint32_t mid = (rlo + rhi) >> 1;
        if (pp.run_beg[mid] < lo) rlo = mid + 1; else rhi = mid;
    }
    pp.tile_run0[i] = rlo;
}

// Runs per tile of the stream form: the next column's first run minus this one's.
__global__ __launch_bounds__(256) void k_tile_nr(PaintParams pp, uint32_t global_rows) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= global_rows * pp.tstride) return;
    const uint32_t k = i % pp.tstride;
    pp.tile_nr[i] = k + 1 < pp.tstride ? pp.tile_run0[i + 1] - pp.tile_run0[i] : 0u;
}

// Runs of the sorted render words: maximal ranges of equal (row, column,
// layer) key, i.e. one layer's segments in one tile. Word i starts a run when
// its key differs from word i - 1's. Run heads are ranked over the whole
// stream (block scan of ballots, decoupled look-back over partitions of
// 4,096 words), so the work follows the words, not the tiles: a tile with a
// million words spreads over many work groups. Run g: run_key[g] = layer |
// even-odd flag, run_beg[g] = its first word; run_beg[R] = the word count;
// run_cov[g] = its 8 row cover sums (wrapping i8, packed in 64 bits), from
// the block's covers staged in LDS; a run that continues past the block's end
// is left to k_run_cov (open[part] = g + 1, else 0).
constexpr uint32_t kRunsIPT = 16;
static_assert(256 * kRunsIPT == kRunsTileHost, "k_runs_scan partition size");
__global__ __launch_bounds__(256) void k_runs_scan(PaintParams pp, const FrameCounters *__restrict__ counters,
                                                   uint64_t *__restrict__ status, uint32_t *__restrict__ stall,
                                                   uint32_t *__restrict__ open) {
    __shared__ uint32_t s_cnt[kRunsIPT * 4];
    __shared__ uint64_t s_excl;
    __shared__ uint64_t s_heads[kRunsIPT * 4];     // head bits in word order
    __shared__ uint16_t s_cov[256 * kRunsIPT];     // cover byte | tile_y << 8 per word
    const uint64_t n = counters->kept;
    const uint32_t part = blockIdx.x;
    const uint64_t base = (uint64_t)part * 256 * kRunsIPT;
    if (base >= n) return;  // partitions past the stream (grids sized for a capacity)
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint64_t key[kRunsIPT], hm[kRunsIPT];
#pragma unroll
    for (uint32_t k = 0; k < kRunsIPT; ++k) {
        const uint64_t i = base + k * 256 + threadIdx.x;
        const uint64_t wd = i < n ? pp.segs[i] : ~0ull;
        key[k] = wd >> kPayloadBits;
        s_cov[k * 256 + threadIdx.x] = (uint16_t)(((uint32_t)wd & 0x3Fu) | (((uint32_t)(wd >> 19) & 7u) << 8));
    }
#pragma unroll
    for (uint32_t k = 0; k < kRunsIPT; ++k) {
        const uint64_t i = base + k * 256 + threadIdx.x;
        uint64_t prev = __shfl_up(key[k], 1);
        if (lane == 0) prev = i > 0 && i - 1 < n ? pp.segs[i - 1] >> kPayloadBits : ~0ull;
        const bool head = i < n && (i == 0 || key[k] != prev);
        hm[k] = __ballot(head);
        if (lane == 0) {
            s_cnt[k * 4 + w] = (uint32_t)__popcll(hm[k]);
            s_heads[k * 4 + w] = hm[k];
        }
    }
    __syncthreads();
    // heads before each (step, wave) group, in word order (step-major)
    if (threadIdx.x < 64) {
        const uint32_t c = s_cnt[lane];
        uint32_t x = c;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = __shfl_up(x, off);
            if (lane >= (uint32_t)off) x += y;
        }
        const uint32_t agg = __shfl(x, 63);
        s_cnt[lane] = x - c;
        if (lane == 0) st_release(status + part, (part == 0 ? kStInc : kStAgg) | agg);
        const uint64_t e = part ? lookback_wave(status, part, stall) : 0;
        if (lane == 0) {
            s_excl = e;
            if (part) st_release(status + part, kStInc | (e + agg));
            if (base + 256 * kRunsIPT >= n) {
                *pp.n_runs = (uint32_t)(e + agg);
                pp.run_beg[e + agg] = (uint32_t)n;
            }
        }
    }
    if (threadIdx.x == 0) open[part] = 0;
    __syncthreads();
    const uint64_t excl = s_excl;
    const uint32_t lmask = (1u << pp.layer_bits) - 1u, row_shift = pp.col_bits + pp.layer_bits;
    const uint32_t in_block = n - base < 256 * kRunsIPT ? (uint32_t)(n - base) : 256 * kRunsIPT;
#pragma unroll
    for (uint32_t k = 0; k < kRunsIPT; ++k) {
        if (!((hm[k] >> lane) & 1ull)) continue;
        const uint64_t g = excl + s_cnt[k * 4 + w] + (uint64_t)__popcll(hm[k] & lt);
        const uint32_t layer = (uint32_t)key[k] & lmask;
        const SceneDesc &S = pp.scenes[pp.row_scene[(uint32_t)(key[k] >> row_shift)]];
        const bool eo = pp.props[S.props_base + layer].flags & PF_EVENODD;
        pp.run_key[g] = layer | (eo ? kRunEvenOdd : 0u);
        pp.run_beg[g] = (uint32_t)(base + k * 256 + threadIdx.x);
    }
    // Cover sums: thread t takes the block's words [16 t, 16 t + 16) from
    // LDS. A segmented scan (reset at run heads, byte-wise wrapping adds)
    // over the threads gives every word the sum of its run so far; the last
    // word of each run in the block writes that run's total.
    const uint32_t c0 = threadIdx.x * kRunsIPT;
    const uint32_t hb = (uint32_t)((s_heads[c0 >> 6] >> (c0 & 63)) & 0xFFFFull);
    uint64_t v[kRunsIPT];
    uint64_t tail = 0;
#pragma unroll
    for (uint32_t j = 0; j < kRunsIPT; ++j) {
        const uint32_t c = s_cov[c0 + j];
        const uint64_t byte = (uint64_t)((uint32_t)as_i8((int32_t)((int64_t)((uint64_t)c << 58) >> 58)) & 0xFFu);
        v[j] = c0 + j < in_block ? byte << (8 * (c >> 8)) : 0ull;
        tail = ((hb >> j) & 1u) ? v[j] : add_bytes(tail, v[j]);
    }
    // exclusive segmented scan of (has head, tail) over the threads
    uint32_t f = hb != 0u;
    uint64_t x = tail;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint64_t xp = __shfl_up(x, off);
        const uint32_t fp = __shfl_up(f, off);
        if (lane >= off) {
            if (!f) x = add_bytes(xp, x);
            f |= fp;
        }
    }
    __shared__ uint64_t s_wx[4];
    __shared__ uint32_t s_wf[4];
    if (lane == 63) {
        s_wx[w] = x;
        s_wf[w] = f;
    }
    uint64_t ex = __shfl_up(x, 1);
    uint32_t ef = __shfl_up(f, 1);
    if (lane == 0) {
        ex = 0;
        ef = 0;
    }
    __syncthreads();
    uint64_t carry = 0;  // from the waves before this one
    for (uint32_t q = 0; q < w; ++q) carry = s_wf[q] ? s_wx[q] : add_bytes(carry, s_wx[q]);
    uint64_t run = ef ? ex : add_bytes(carry, ex);
    const uint64_t *heads = s_heads;
#pragma unroll
    for (uint32_t j = 0; j < kRunsIPT; ++j) {
        const uint32_t li = c0 + j;
        run = ((hb >> j) & 1u) ? v[j] : add_bytes(run, v[j]);
        if (li >= in_block) break;
        const uint32_t nx = li + 1;
        const bool last = nx == in_block || ((heads[nx >> 6] >> (nx & 63)) & 1ull);
        if (!last) continue;
        // the run of word li: heads at or before li in this block
        const uint32_t q = li >> 6;
        const uint32_t le = s_cnt[q] + (uint32_t)__popcll(heads[q] & (~0ull >> (63 - (li & 63))));
        if (le == 0) continue;  // a run from an earlier block: its block's open run
        const uint64_t g = excl + le - 1;
        pp.run_cov[g] = run;
        if (nx == in_block && base + 256 * kRunsIPT < n) open[part] = (uint32_t)g + 1u;
    }
}

// Cover sums of the runs that continue past their block's end (one per
// block at most): one thread per block, over the run's words.
__global__ __launch_bounds__(256) void k_run_cov(PaintParams pp, const FrameCounters *__restrict__ counters,
                                                 const uint32_t *__restrict__ open) {
    const uint64_t n_parts = (counters->kept + kRunsTileHost - 1) / kRunsTileHost;
    for (uint64_t part = blockIdx.x * 256 + threadIdx.x; part < n_parts; part += (uint64_t)gridDim.x * 256) {
        const uint32_t o = open[part];
        if (!o) continue;
        const uint32_t g = o - 1u;
        const uint32_t b = pp.run_beg[g], e = pp.run_beg[g + 1];
        uint64_t s = 0;
        for (uint32_t i = b; i < e; ++i) {
            const uint64_t w = pp.segs[i];
            const int32_t cover = (int32_t)((int64_t)(w << 58) >> 58);
            const uint32_t ty = (uint32_t)(w >> 19) & 7u;
            s = add_bytes(s, (uint64_t)((uint32_t)cover & 0xFFu) << (8 * ty));
        }
        pp.run_cov[g] = s;
    }
}


// Per-tile form of the runs, for frames with many tiles of moderate depth:
// one wave per tile column, the tile's runs stored from its first word's
// index on (tile_run0 = tile_start). Each chunk of 64 words adds its covers
// into the wave's LDS cells, one u32 per (run slot, pixel row): slot 0 takes
// the run left open by the earlier chunks, slot s the s-th run starting in
// the chunk; the low byte of a cell is the wrapping i8 sum. One lane per slot
// then packs the slot's 8 row sums into bytes (and clears its cells).
constexpr uint32_t kRunSlots = 65;  // the open run and up to 64 run heads of a chunk
__device__ __forceinline__ uint64_t pack_run_cells(uint32_t *c) {
    const uint4 a = *(const uint4 *)c, b = *(const uint4 *)(c + 4);
    *(uint4 *)c = make_uint4(0, 0, 0, 0);
    *(uint4 *)(c + 4) = make_uint4(0, 0, 0, 0);
    // v_perm_b32 byte selectors 0, 4 of each source pair: low bytes
    const uint32_t lo = __builtin_amdgcn_perm(__builtin_amdgcn_perm(a.w, a.z, 0x0c0c0400u),
                                              __builtin_amdgcn_perm(a.y, a.x, 0x0c0c0400u), 0x05040100u);
    const uint32_t hi = __builtin_amdgcn_perm(__builtin_amdgcn_perm(b.w, b.z, 0x0c0c0400u),
                                              __builtin_amdgcn_perm(b.y, b.x, 0x0c0c0400u), 0x05040100u);
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// Runs of tile column colp (col + 1; 0 gathers every col < 0) of row grow,
// by one wave; `cells` is the wave's kRunSlots x 8 LDS cells, all zero.
// Returns the number of runs.
__device__ __forceinline__ uint32_t tile_runs_one(const PaintParams &pp, uint32_t grow, uint32_t colp,
                                                  uint32_t *cells) {
    const uint32_t lane = lane_id();
    const SceneDesc &S = pp.scenes[pp.row_scene[grow]];
    const uint32_t *ts_row = pp.tile_start + (uint64_t)grow * pp.tstride;
    const uint32_t ts = ts_row[colp], te = ts_row[colp + 1];
    const PropsEntry *props = pp.props + S.props_base;
    const uint32_t lmask = (1u << pp.layer_bits) - 1u;
    uint32_t nr = 0, prev_layer = 0xFFFFFFFFu;
    uint64_t open = 0;  // row cover sums of the run still open after the earlier chunks
    for (uint32_t p = ts; p < te; p += 256) {
        uint64_t w[4];
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
            const uint32_t idx = p + b * 64 + lane;
            w[b] = idx < te ? pp.segs[idx] : 0ull;
        }
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
            const uint32_t base = p + b * 64;
            if (base >= te) break;
            const uint32_t nvalid = te - base < 64 ? te - base : 64;
            const bool valid = lane < nvalid;
            const uint32_t layer = valid ? (uint32_t)(w[b] >> kPayloadBits) & lmask : 0xFFFFFFFFu;
            uint32_t prevl = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)layer, 0x138, 0xF, 0xF, false);  // wave_shr:1
            if (lane == 0) prevl = prev_layer;
            const bool head = valid && layer != prevl;
            const uint64_t heads = __ballot(head);
            const uint32_t nh = (uint32_t)__popcll(heads);
            const uint32_t slot = (uint32_t)__popcll(heads & ((2ull << lane) - 1ull));
            if (valid) {
                const int32_t cover = (int32_t)((int64_t)(w[b] << 58) >> 58);
                const uint32_t ty = (uint32_t)(w[b] >> 19) & 7u;
                atomicAdd(&cells[slot * 8 + ty], (uint32_t)cover);
            }
            if (head) {
                const uint32_t r = nr + slot - 1;
                const bool eo = props[layer].flags & PF_EVENODD;
                pp.run_key[ts + r] = layer | (eo ? kRunEvenOdd : 0u);
                pp.run_beg[ts + r] = base + lane;
            }
            wave_lds_sync();
            const uint64_t pk = lane <= nh ? pack_run_cells(cells + lane * 8) : 0ull;
            const uint64_t pk64 = (lane == 0 && nh == 64) ? pack_run_cells(cells + 64 * 8) : 0ull;
            const uint64_t cont = add_bytes(open, rlane64(pk, 0));  // the open run with this chunk's words
            if (nh == 0) {
                open = cont;
            } else {
                // the open run closed at the chunk's first head; runs 1 .. nh - 1
                // of the chunk are complete, the last stays open
                if (lane == 0 && nr) pp.run_cov[ts + nr - 1] = cont;
                if (lane >= 1 && lane < nh) pp.run_cov[ts + nr + lane - 1] = pk;
                open = nh < 64 ? rlane64(pk, nh) : rlane64(pk64, 0);
            }
            nr += nh;
            prev_layer = rlane(layer, nvalid - 1);
            wave_lds_sync();
        }
    }
    if (lane == 0) {
        pp.tile_nr[(uint64_t)grow * pp.tstride + colp] = nr;
        if (nr) pp.run_cov[ts + nr - 1] = open;
    }
    return nr;
}

// Tile runs of every (row, column) of the work list: each wave takes entries
// wave, wave + waves, ...; waves never synchronise.
__global__ __launch_bounds__(256) void k_tile_runs(PaintParams pp, const uint2 *__restrict__ work, uint32_t n_work) {
    __shared__ __align__(16) uint32_t s_cells[4][kRunSlots * 8];
    uint32_t *cells = s_cells[threadIdx.x >> 6];
    for (uint32_t k = lane_id(); k < kRunSlots * 8; k += 64) cells[k] = 0;
    wave_lds_sync();
    const uint32_t waves = gridDim.x * 4;
    uint32_t runs = 0;
    for (uint32_t t = blockIdx.x * 4 + (threadIdx.x >> 6); t < n_work; t += waves) {
        const uint2 wk = work[t];
        runs += tile_runs_one(pp, wk.x, wk.y, cells);
    }
    // (one of kStatSlices counters per work group: the waves of a frame of
    // many small tiles would otherwise queue on one address)
    if (lane_id() == 0 && runs) atomicAdd(pp.n_runs + (blockIdx.x % kStatSlices), runs);
}

// Merge-path split of two layer-sorted lists at diagonal d: q elements first
// on equal layers. Returns the number of q elements before the split.
__device__ __forceinline__ uint32_t merge_split(const uint32_t *qk, uint32_t nq, const uint32_t *rk, uint32_t nr,
                                                uint32_t d) {
    uint32_t lo = d > nr ? d - nr : 0, hi = d < nq ? d : nq;
    // (an eight-way search, seven probes in flight per step, was slower: the
    // walk is partly bound by instruction issue)
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((qk[mid] & 0xFFFFu) <= (rk[d - 1 - mid] & 0xFFFFu)) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// The cover queue of one tile row, left to right (painter/mod.rs:495-596:
// tiles left of the crop and tile -1 are folded without dropping empty
// covers; inside the crop, covers that become empty are dropped,
// layer_workbench.rs:194,475).
// G: the queue and run arrays live in the row's slice of a global scratch
// buffer instead of LDS (tiles with more layers than LDS holds).
// 1024 threads in LDS: at most 64 VGPRs and few enough SGPRs for eight
// waves per SIMD, two work groups (two tile rows) per CU.
template <uint32_t T, bool G>
__global__ __launch_bounds__(T) void k_row_walk(PaintParams pp) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ uint64_t s_wsum[T / 64];
    __shared__ uint32_t s_rec[3];
    const uint32_t tid = threadIdx.x;
    const uint32_t cap = pp.cap_ids, rcap = pp.cap_runs, tstride = pp.tstride;
    unsigned char *p = G ? pp.walk_scratch + (uint64_t)blockIdx.x * pp.walk_stride : lds_raw;
    // two queues (the next column's is written while the current one is
    // read) and one or two sets of run arrays (pp.walk_run_bufs; with two,
    // the next column's runs are stored while the current ones are read).
    // Array k of a pair follows array 0 (queues: cap entries, runs: rcap), so
    // each is addressed from one base: offsets from lds_raw keep the accesses
    // LDS instructions (pointers picked from an array by a run-time index
    // became generic pointers, whose flat accesses also wait for every
    // outstanding global load and store).
    const uint32_t nb = pp.walk_run_bufs;
    uint64_t *const qc0 = (uint64_t *)p; p += (size_t)cap * 8 * 2;
    uint64_t *const rc0 = (uint64_t *)p; p += (size_t)rcap * 8 * nb;
    uint32_t *const qk0 = (uint32_t *)p; p += (size_t)cap * 4 * 2;
    uint32_t *const rk0 = (uint32_t *)p; p += (size_t)rcap * 4 * nb;
    uint32_t *const rb0 = (uint32_t *)p; p += (size_t)rcap * 4 * nb;
    // (buf is 0 when there is one set of run arrays)
    auto qk = [&](uint32_t k) { return qk0 + (size_t)k * cap; };
    auto qc = [&](uint32_t k) { return qc0 + (size_t)k * cap; };
    auto rks = [&](uint32_t b) { return rk0 + (size_t)b * rcap; };
    auto rbs = [&](uint32_t b) { return rb0 + (size_t)b * rcap; };
    auto rcs = [&](uint32_t b) { return rc0 + (size_t)b * rcap; };
    uint32_t *ts_l = (uint32_t *)p; p += (size_t)tstride * 4;
    uint32_t *nr_l = (uint32_t *)p; p += (size_t)tstride * 4;
    uint32_t *r0_l = (uint32_t *)p;

    const uint32_t grow = blockIdx.x;
    const SceneDesc S = pp.scenes[pp.row_scene[grow]];
    const uint32_t j = grow - S.row_base;
    if (j < S.crop_v0 || j >= S.crop_v1) return;
    const uint32_t h0 = S.crop_h0, h1 = scene_h1(S);
    if (h0 >= h1) return;
    for (uint32_t k = tid; k < tstride; k += T) {
        const uint32_t *r0 = pp.tile_run0 + (uint64_t)grow * tstride;
        ts_l[k] = pp.tile_start[(uint64_t)grow * tstride + k];
        r0_l[k] = r0[k];
        nr_l[k] = k + 1 < tstride ? pp.tile_nr[(uint64_t)grow * tstride + k] : 0u;
    }
    __syncthreads();
    uint32_t nq = 0, cur = 0;
    uint32_t rec_next = 0, rec_end = 0;
    bool counting = false;  // record store full: only count what the frame needs
    uint32_t row_records = 0;  // id records of the row (thread 0)
    // the next column to walk: fold columns without runs change nothing
    auto next_colp = [&](uint32_t c) {
        uint32_t x = c + 1;
        while (x <= h0 && nr_l[x] == 0) ++x;
        return x;
    };
    // Runs are prefetched into registers (one run per thread); runs beyond T
    // are loaded when stored. With two sets of run arrays a column's runs
    // are stored into the other set during the column before it (one barrier
    // per column); with one set, after a barrier at the end of that column.
    struct Prefetch {
        uint32_t key, beg, n;
        uint64_t cov;
    };
    auto issue = [&](Prefetch &f, uint32_t col) {
        f.n = 0;
        if (col > h1) return;
        const uint32_t nr2 = nr_l[col], r02 = r0_l[col];
        f.n = nr2 < T ? nr2 : T;
        if (f.n > rcap) f.n = rcap;
        if (tid < f.n) {
            f.key = pp.run_key[r02 + tid];
            f.beg = pp.run_beg[r02 + tid];
            f.cov = pp.run_cov[r02 + tid];
        }
    };
    // The runs of column col into run arrays `buf`; false when they exceed
    // the arrays (the work group must stop).
    // largest runs of a tile and queue + runs met by this row, published at
    // the row's end or when it stops; only the large configuration (two sets
    // of run arrays) reads them, and a row publishes only what raises them
    // (frames of many rows would otherwise queue on the two addresses)
    uint32_t max_runs = 0, max_ids = 0;
    auto publish_max = [&]() {
        if (tid == 0 && pp.walk_run_bufs == 2) {
            if (max_runs > __hip_atomic_load(pp.walk_max, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMax(pp.walk_max, max_runs);
            if (max_ids > __hip_atomic_load(pp.walk_max + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                atomicMax(pp.walk_max + 1, max_ids);
        }
    };
    auto store = [&](const Prefetch &f, uint32_t col, uint32_t buf) -> bool {
        if (col > h1) return true;
        const uint32_t nr = nr_l[col];
        if (nr > rcap) {
            max_runs = (uint32_t)__builtin_amdgcn_readfirstlane((int)nr);
            if (tid == 0) atomicOr(pp.overflow, 1u);
            publish_max();
            return false;
        }
        if (tid < f.n) {
            rks(buf)[tid] = f.key;
            rbs(buf)[tid] = f.beg;
            rcs(buf)[tid] = f.cov;
        }
        const uint32_t r0 = r0_l[col];
        for (uint32_t k = f.n + tid; k < nr; k += T) {
            rks(buf)[k] = pp.run_key[r0 + k];
            rbs(buf)[k] = pp.run_beg[r0 + k];
            rcs(buf)[k] = pp.run_cov[r0 + k];
        }
        return true;
    };
    // One column: the cover queue entering tile colp - 1 meets its runs (in
    // run arrays `buf`). `mid` runs between the scan and the stores of the
    // records (the two-buffer walk stores the next column's runs there).
    // Returns false when the work group must stop (capacity exceeded). Ends
    // without a barrier.
    auto step = [&](uint32_t colp, uint32_t buf, auto &&mid) -> bool {
        const bool fold = colp <= h0;
        const uint32_t nr = nr_l[colp];
        {
            // uniform over the work group: kept in scalar registers
            const uint32_t r = (uint32_t)__builtin_amdgcn_readfirstlane((int)nr);
            const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)(nq + nr));
            max_runs = r > max_runs ? r : max_runs;
            max_ids = t > max_ids ? t : max_ids;
        }
        if (nq + nr > cap) {
            if (tid == 0) atomicOr(pp.overflow, 1u);
            publish_max();
            return false;
        }
        const uint32_t ts = ts_l[colp], te = ts_l[colp + 1];
        const uint32_t *rk = rks(buf), *rb = rbs(buf);
        const uint64_t *rc = rcs(buf);
        const uint32_t *qkc = qk(cur);
        const uint64_t *qcc = qc(cur);
        const uint32_t total = nq + nr;
        const uint32_t per = (total + T - 1) / T;
        const uint32_t d0 = tid * per < total ? tid * per : total;
        const uint32_t d1 = d0 + per < total ? d0 + per : total;
        const uint32_t i0 = merge_split(qkc, nq, rk, nr, d0);
        // One element of the queue ∪ runs union: layer | kRunEvenOdd, its
        // queue and run indices (-1: none), the carried cover `a` and the run's
        // cover `b`, and the run's words [beg, beg + len). Second halves of
        // equal-layer pairs are skipped.
        struct Elem {
            uint32_t lay_eo;
            int32_t qi, ri;
            uint64_t a, b;
            uint32_t beg, len;
        };
        // The element at merge position (i, r), from one round of LDS reads
        // (positions past an array's end read its entry 0, unused), built with
        // selects: `take_q` when it comes from the queue (else from the runs),
        // `skip` for a run whose layer the previous queue element had (merged
        // there).
        auto element = [&](uint32_t i, uint32_t r, bool &take_q, bool &skip) {
            const uint32_t iq = i < nq ? i : 0u, ir = r < nr ? r : 0u, ip = i > 0 ? i - 1u : 0u;
            const uint32_t ir1 = ir + 1u < nr ? ir + 1u : ir;
            const uint32_t kq = qkc[iq], kr = rk[ir], kp = qkc[ip];
            const uint64_t ca = qcc[iq], cb = rc[ir];
            const uint32_t b0 = rb[ir], b1 = ir + 1u < nr ? rb[ir1] : te;
            const uint32_t lq = kq & 0xFFFFu, lr = kr & 0xFFFFu;
            take_q = i < nq && (r >= nr || lq <= lr);
            const bool with_run = r < nr && (!take_q || lr == lq);
            skip = !take_q && i > 0 && (kp & 0xFFFFu) == lr;
            Elem e;
            e.lay_eo = (take_q ? kq : kr) & (0xFFFFu | kRunEvenOdd);
            e.qi = take_q ? (int32_t)i : -1;
            e.ri = with_run ? (int32_t)r : -1;
            e.a = take_q ? ca : 0ull;
            e.b = with_run ? cb : 0ull;
            e.beg = b0;
            e.len = b1 - b0;
            return e;
        };
        auto walk = [&](auto &&emit) {
            uint32_t i = i0, r = d0 - i0;
            for (uint32_t m = d0; m < d1; ++m) {
                bool take_q, skip;
                const Elem e = element(i, r, take_q, skip);
                if (!skip) emit(e);
                i += take_q ? 1u : 0u;
                r += take_q ? 0u : 1u;
            }
        };
        // next cover of a union element
        auto next_cover = [&](const Elem &e, bool *keep) {
            const uint64_t v = add_bytes(e.a, e.b);
            *keep = fold || !cover_empty(v, (e.lay_eo & kRunEvenOdd) != 0);
            return v;
        };
        // The first two elements of the thread's slice stay in registers for
        // the writing pass (one element per thread is the common case);
        // longer slices are walked again.
        Elem e0{}, e1{};
        uint64_t ev0 = 0, ev1 = 0;
        uint32_t cu = 0, cq = 0, e_keep = 0;
        if (per == 1) {
            // (uniform) at most one element per thread: no loop
            if (d0 < d1) {
                bool take_q, skip, keep;
                e0 = element(i0, d0 - i0, take_q, skip);
                ev0 = next_cover(e0, &keep);
                cu = skip ? 0u : 1u;
                cq = !skip && keep ? 1u : 0u;
                e_keep = cq;
            }
        } else if (per == 2) {
            // (uniform) at most two elements per thread: the second after the
            // first, moved into the first slot when the first is skipped
            bool skip_a = true, skip_b = true, keep_a = false, keep_b = false;
            Elem ea{}, eb{};
            uint64_t va = 0, vb = 0;
            if (d0 < d1) {
                bool take_q;
                ea = element(i0, d0 - i0, take_q, skip_a);
                va = next_cover(ea, &keep_a);
                if (d0 + 1 < d1) {
                    bool take_q_b;
                    const uint32_t i1 = i0 + (take_q ? 1u : 0u);
                    eb = element(i1, d0 + 1 - i1, take_q_b, skip_b);
                    vb = next_cover(eb, &keep_b);
                }
            }
            cu = (skip_a ? 0u : 1u) + (skip_b ? 0u : 1u);
            cq = (!skip_a && keep_a ? 1u : 0u) + (!skip_b && keep_b ? 1u : 0u);
            e0 = skip_a ? eb : ea;
            ev0 = skip_a ? vb : va;
            e1 = eb;
            ev1 = vb;
            const bool k0 = skip_a ? keep_b : keep_a;
            e_keep = (k0 ? 1u : 0u) | (!skip_a && keep_b ? 2u : 0u);
        } else {
            walk([&](const Elem &e) {
                bool keep;
                const uint64_t v = next_cover(e, &keep);
                if (cu == 0) {
                    e0 = e;
                    ev0 = v;
                    e_keep |= keep ? 1u : 0u;
                } else if (cu == 1) {
                    e1 = e;
                    ev1 = v;
                    e_keep |= keep ? 2u : 0u;
                }
                ++cu;
                cq += keep ? 1u : 0u;
            });
        }
        // (union entries, kept queue entries) in the low and high halves:
        // 16 bits each when the LDS arrays bound them (cap < 2^16)
        uint64_t ex, all;
        if (G) {
            ex = block_excl_scan<T / 64>((uint64_t)cu | ((uint64_t)cq << 32), s_wsum, &all);
        } else {
            uint32_t all32;
            // (no trailing barrier: the column ends with one)
            const uint32_t ex32 = block_excl_scan_u32<T / 64, false>(cu | (cq << 16), (uint32_t *)s_wsum, &all32);
            ex = (uint64_t)(ex32 & 0xFFFFu) | ((uint64_t)(ex32 >> 16) << 32);
            all = (uint64_t)(all32 & 0xFFFFu) | ((uint64_t)(all32 >> 16) << 32);
        }
        const uint32_t n_union = (uint32_t)all, nq_new = (uint32_t)(all >> 32);
        uint32_t rec_off = 0;
        if (!fold && counting) {
            if (tid == 0) atomicAdd(pp.rec_used, n_union);
        } else if (!fold) {
            if (rec_next + n_union > rec_end) {
                if (tid == 0) {
                    const uint32_t sz = n_union > pp.rec_chunk ? n_union : pp.rec_chunk;
                    const uint32_t base = atomicAdd(pp.rec_used, sz);
                    s_rec[0] = base;
                    s_rec[1] = base + sz;
                    s_rec[2] = (uint64_t)base + sz > pp.rec_cap ? 1u : 0u;
                    if (s_rec[2]) atomicOr(pp.overflow, 2u);
                }
                __syncthreads();
                counting = s_rec[2] != 0;
                rec_next = s_rec[0];
                rec_end = s_rec[1];
            }
            rec_off = rec_next;
            rec_next += n_union;
        }
        if (!mid()) return false;
        const bool write_rec = !fold && !counting;
        const uint32_t nxt = cur ^ 1u;
        uint32_t ou = rec_off + (uint32_t)ex, oq = (uint32_t)(ex >> 32);
        // the element's id record and, when its cover stays, its queue entry
        // (key: layer | kRunEvenOdd)
        auto put = [&](const Elem &e, uint64_t v, bool keep) {
            if (write_rec) {
                uint32_t key = (e.lay_eo & 0xFFFFu) | ((e.lay_eo & kRunEvenOdd) ? kRecEvenOdd : 0u), beg = 0, len = 0;
                uint64_t cv = 0;
                if (e.ri >= 0) {
                    key |= kRecRun;
                    beg = e.beg;
                    len = e.len;
                }
                if (e.qi >= 0) {
                    key |= kRecCarry;
                    cv = e.a;
                }
                pp.rec[ou] = make_uint4(key, beg, (uint32_t)cv, (uint32_t)(cv >> 32));
                pp.rec_len[ou] = len;
                ++ou;
            }
            if (keep) {
                qk(nxt)[oq] = e.lay_eo;
                qc(nxt)[oq] = v;
                ++oq;
            }
        };
        if (cu <= 2) {
            if (cu > 0) put(e0, ev0, (e_keep & 1u) != 0);
            if (cu > 1) put(e1, ev1, (e_keep & 2u) != 0);
        } else {
            walk([&](const Elem &e) {
                bool keep;
                const uint64_t v = next_cover(e, &keep);
                put(e, v, keep);
            });
        }
        if (write_rec && tid == 0) {
            const uint64_t tile = (uint64_t)S.tile_base + (uint64_t)j * S.row_len + (colp - 1);
            uint32_t cnt = n_union;
            if (pp.deep_min && n_union >= pp.deep_min && !(S.flags & SF_CLIPS)) {
                const uint32_t di = atomicAdd(pp.deep_count, 1u);
                if (di < pp.deep_cap) {
                    pp.deep_list[di] = make_uint2(grow, colp - 1);
                    cnt |= kDeepFlag;
                }
            }
            pp.rec_hdr[tile] = make_uint2(rec_off, cnt);
            row_records += n_union;
        }
        nq = nq_new;
        cur = nxt;
        return true;
    };
    uint32_t c0 = nr_l[0] ? 0u : next_colp(0), c1 = next_colp(c0);
    const auto none = [] { return true; };
    if (nb == 2) {
        // Two sets of run arrays: a column's runs are loaded into registers
        // after the records of the column before it are stored, and stored
        // into the other set in the middle of that column's step, before its
        // own records. Loads waited for while record stores are outstanding
        // wait for those stores too (one counter for both), so no record
        // store lies between a load and its use.
        Prefetch B;
        issue(B, c0);
        if (!store(B, c0, 0)) return;
        issue(B, c1);
        __syncthreads();
        uint32_t buf = 0;
        while (c0 <= h1) {
            // column c0 in arrays buf; B holds c1
            if (!step(c0, buf, [&] { return store(B, c1, buf ^ 1u); })) return;
            c0 = c1;
            c1 = next_colp(c1);
            issue(B, c1);
            __syncthreads();
            buf ^= 1u;
        }
    } else {
        // One set: the next column's runs are stored after a barrier; they
        // are loaded two columns ahead.
        uint32_t c2 = next_colp(c1);
        Prefetch A, B;
        issue(A, c0);
        issue(B, c1);
        if (!store(A, c0, 0)) return;
        issue(A, c2);
        __syncthreads();
        while (c0 <= h1) {
            // column c0; B holds c1, A holds c2
            if (!step(c0, 0, none)) return;
            const uint32_t c3 = next_colp(c2);
            __syncthreads();
            if (!store(B, c1, 0)) return;
            issue(B, c3);
            __syncthreads();
            if (c1 > h1) break;
            // column c1; A holds c2, B holds c3
            if (!step(c1, 0, none)) return;
            const uint32_t c4 = next_colp(c3);
            __syncthreads();
            if (!store(A, c2, 0)) return;
            issue(A, c4);
            __syncthreads();
            c0 = c2;
            c1 = c3;
            c2 = c4;
        }
    }
    // (one of kStatSlices counters per row: rows of a large batch would
    // queue on one address)
    if (tid == 0 && row_records) atomicAdd(pp.n_records + (blockIdx.x % kStatSlices), row_records);
    publish_max();
}
template __global__ void k_row_walk<64, false>(PaintParams);
template __global__ void k_row_walk<256, false>(PaintParams);
template __global__ void k_row_walk<512, false>(PaintParams);
template __global__ void k_row_walk<1024, false>(PaintParams);
template __global__ void k_row_walk<64, true>(PaintParams);
template __global__ void k_row_walk<256, true>(PaintParams);
template __global__ void k_row_walk<512, true>(PaintParams);
template __global__ void k_row_walk<1024, true>(PaintParams);

// Staged window of sorted words for the paint wave.
struct Stage {
    uint64_t *w;
    uint32_t base, end;
};


__device__ __forceinline__ void stage_load(Stage &st, const uint64_t *segs, uint32_t from, uint32_t to) {
    const uint32_t lane = lane_id();
    wave_lds_sync();
    for (uint32_t k = lane; k < to - from; k += 64) st.w[k] = segs[from + k];
    wave_lds_sync();
    // wave-uniform: kept in scalar registers
    st.base = (uint32_t)__builtin_amdgcn_readfirstlane((int)from);
    st.end = (uint32_t)__builtin_amdgcn_readfirstlane((int)to);
}

// A = 16 * P + area for this lane's pixel (painter/mod.rs:229-243, 347-362):
// P is the i8 running cover prefix of row py over columns < px starting from
// the carried cover cy, area the i16 area sum of cell (px, py); the record's
// run is words [beg, beg + len) of the sorted stream, all inside [.., lim).
// The lanes take the run's words 64 at a time and add area and cover into
// the wave's 64 cells (LDS, no return values); each lane then reads its own
// cell and the column prefix is a stride-8 scan across lanes (lane = x*8+y).
// The words come from the staged window when the run lies inside it, else
// from global memory.
__device__ __forceinline__ int32_t run_area(const Stage &st, int32_t *cells, const uint64_t *segs, uint32_t beg,
                                            uint32_t len, int32_t cy) {
    const uint32_t lane = lane_id();
    int32_t *c_area = cells, *c_cov = cells + 64;
    c_area[lane] = 0;
    c_cov[lane] = 0;
    const uint32_t e = beg + len;
    auto add = [&](uint64_t w) {
        const uint32_t cell = ((uint32_t)(w >> 16) & 7u) * 8u + ((uint32_t)(w >> 19) & 7u);
        atomicAdd(&c_area[cell], (int32_t)((int64_t)(w << 48) >> 54));
        atomicAdd(&c_cov[cell], (int32_t)((int64_t)(w << 58) >> 58));
    };
    // separate loops keep the LDS and global reads in their own address spaces
    if (beg >= st.base && e <= st.end) {
        for (uint32_t s = beg + lane; s < e; s += 64) add(st.w[s - st.base]);
    } else {
        for (uint32_t s = beg + lane; s < e; s += 64) add(segs[s]);
    }
    const int32_t ar = c_area[lane], cv = c_cov[lane];
    int32_t inc = cv, t;
    t = __shfl_up(inc, 8);
    if (lane >= 8) inc += t;
    t = __shfl_up(inc, 16);
    if (lane >= 16) inc += t;
    t = __shfl_up(inc, 32);
    if (lane >= 32) inc += t;
    return 16 * as_i8(cy + inc - cv) + as_i16(ar);
}

// One batch of up to 64 records, lane k holding record k0 + k. Only the
// first `cnt` are used: as many as keep the words of their runs within one
// staging window (at least one record).
struct Batch {
    uint32_t key, beg, len, flags;
    uint64_t cov;
    float4 color;
    uint32_t cnt;    // records of the batch (wave-uniform)
    uint32_t first;  // first word of the batch's runs
    uint32_t lim;    // end of the batch's last run (lim == first: no run)
    uint64_t rm;     // lanes of the batch whose record has a run
    uint64_t longm;  // of those, runs longer than kGroupLenMax words
};

// Without Window the batch is simply the next 64 records (for a reader that
// does not stage their runs).
template <bool Window = true>
__device__ __forceinline__ void load_batch(const PaintParams &pp, const PropsEntry *props, const uint32_t *fl,
                                           const uint4 *rec, const uint32_t *rlen, uint32_t n, uint32_t k0,
                                           Batch &b) {
    const uint32_t k = k0 + lane_id();
    const bool in = k < n;
    const uint4 rv = in ? rec[k] : make_uint4(0, 0, 0, 0);
    b.key = rv.x;
    b.beg = rv.y;
    b.cov = (uint64_t)rv.z | ((uint64_t)rv.w << 32);
    b.len = in ? rlen[k] : 0u;
    const uint32_t lay = rv.x & 0xFFFFu;
    if (in) {
        const PropsEntry &pe = props[lay];
        b.flags = fl ? fl[k] : ((pe.flags & ~(kMaskBit | kSkipClipBit)) | kMaskBit);
        b.color = make_float4(pe.color[0], pe.color[1], pe.color[2], pe.color[3]);
    } else {
        b.flags = 0;
        b.color = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const uint32_t avail = n - k0 < 64 ? n - k0 : 64;
    const bool run = in && (rv.x & kRecRun);
    const uint64_t rm = __ballot(run);
    b.cnt = (uint32_t)__builtin_amdgcn_readfirstlane((int)avail);
    b.first = b.lim = 0;
    b.rm = b.longm = 0;
    if (Window && rm) {
        b.first = rlane(b.beg, (uint32_t)__ffsll((long long)rm) - 1u);
        // the first run ending beyond the window closes the batch (a run
        // longer than the window makes a batch of its own)
        const uint64_t over = __ballot(run && b.beg + b.len - b.first > kStageWin);
        uint64_t used = rm;
        if (over) {
            const uint32_t fo = (uint32_t)__ffsll((long long)over) - 1u;
            b.cnt = fo > 0 ? fo : 1u;
            used &= b.cnt >= 64 ? ~0ull : (1ull << b.cnt) - 1ull;
        }
        const uint32_t last = 63u - (uint32_t)__clzll((long long)used);
        b.lim = used ? rlane(b.beg + b.len, last) : b.first;
        b.rm = used;
        b.longm = used & __ballot(b.len > kGroupLenMax);
    }
}

// Stages the batch's runs (when they fit the window and are not staged yet).
__device__ __forceinline__ void stage_batch(Stage &st, const uint64_t *segs, const Batch &b) {
    if (b.lim > b.first && b.lim - b.first <= kStageWin && (b.first < st.base || b.lim > st.end))
        stage_load(st, segs, b.first, b.lim);
}

__device__ __forceinline__ int32_t record_area(const Stage &st, int32_t *cells, const uint64_t *segs, const Batch &b,
                                               uint32_t q, uint32_t py) {
    const uint32_t key = rlane(b.key, q);
    // carried cover of row py (records without a carry hold zero covers)
    const int32_t cy = as_i8((int32_t)(rlane64(b.cov, q) >> (8 * py)));
    if (!(key & kRecRun)) return 16 * cy;
    return run_area(st, cells, segs, rlane(b.beg, q), rlane(b.len, q), cy);
}

// Run groups: the area sums of up to kGroupRuns consecutive runs of a batch
// for every pixel, accumulated in one pass over the runs' staged words (64
// words per step, lane = word) instead of one pass per run. The word of
// (slot, pixel row y, column x), at y*8 + x in the slot, holds
//   bits  0..15  the pixel's area sum mod 2^16 (the painter's i16 area),
//   bits 16..23  the carries out of bit 15: at most one per word of the
//                pixel, so a run of at most kGroupLenMax words never carries
//                into
//   bits 24..31  the cover sum mod 2^8 (the painter's i8 cover),
// and a pass per (slot, pixel row) then replaces each cover byte with the sum
// of the row's covers in the columns to its left, so that a record reads
// 16 * P + area (run_area's value) from one word. The runs of a batch are
// consecutive stretches of the sorted words, one key each, so a word's slot
// is the number of key changes between the group's first word and it.
//
// group_build serves the records of a staged batch from record q on and
// returns the record after those it serves:
//   - no run at or after q: the batch's end (carried covers only);
//   - the first run from q on is longer than kGroupLenMax words: the record
//     after it, which takes run_area (whose cells share the slots);
//   - else it fills the slots with up to kGroupRuns runs from that one on,
//     ending before the next longer run, sets *r0 to the rank of the first
//     (slot 0) and returns the record of the first run after the group.
__device__ __forceinline__ uint32_t group_build(uint32_t *slots, const Stage &st, const Batch &b, uint32_t q,
                                                uint32_t *r0) {
    const uint32_t lane = lane_id();
    const uint64_t from_q = b.rm & (~0ull << q);  // q < 64
    if (!from_q) return b.cnt;
    const uint32_t q1 = (uint32_t)__ffsll((long long)from_q) - 1u;
    if ((b.longm >> q1) & 1ull) return q1 + 1u;
    const uint32_t r = (uint32_t)__popcll(b.rm & ~(~0ull << q1));
    const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(b.rm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)b.rm, 0u));
    uint64_t sel = from_q & __ballot(rank - r < kGroupRuns);
    const uint64_t lg = sel & b.longm;
    if (lg) sel &= (1ull << (__ffsll((long long)lg) - 1)) - 1ull;
    const uint64_t after = from_q & ~sel;
    const uint32_t last = 63u - (uint32_t)__clzll((long long)sel);
    const uint32_t n = (uint32_t)__popcll(sel);
    const uint32_t from = rlane(b.beg, q1), to = rlane(b.beg, last) + rlane(b.len, last);
    // lane = slot*8 + y clears, and later prefixes, the row y of a slot
    const bool slot_lane = (lane >> 3) < kGroupRuns;
    uint4 *row = (uint4 *)(slots + (lane >> 3) * kGroupStride + (lane & 7u) * 8u);
    if (slot_lane) {
        row[0] = make_uint4(0, 0, 0, 0);
        row[1] = make_uint4(0, 0, 0, 0);
    }
    wave_lds_sync();
    const uint64_t *w = st.w - st.base;
    uint32_t prev_key = (uint32_t)(w[from] >> kPayloadBits), sbase = 0;
    for (uint32_t s0 = from; s0 < to; s0 += 64) {
        const uint32_t s = s0 + lane;
        // lanes past the group's last word read that word again: no key
        // change, and they add nothing
        const uint64_t word = w[min(s, to - 1u)];
        const uint32_t hi = (uint32_t)(word >> kPayloadBits), lo = (uint32_t)word;
        // the key of the word before (lane 0: the previous step's last word)
        const uint32_t before = (uint32_t)__builtin_amdgcn_update_dpp((int)prev_key, (int)hi, 0x138, 0xF, 0xF,
                                                                      false);  // wave_shr:1
        const bool start = hi != before;
        const uint64_t sm = __ballot(start);
        uint32_t slot = __builtin_amdgcn_mbcnt_hi((uint32_t)(sm >> 32),
                                                  __builtin_amdgcn_mbcnt_lo((uint32_t)sm, sbase + (start ? 1u : 0u)));
        slot = min(slot, n - 1u);  // (always below n: the runs are consecutive)
        if (s < to) {
            const uint32_t v = ((uint32_t)((int32_t)(lo << 16) >> 22) & 0xFFFFu) | (uint32_t)((int32_t)(lo << 26) >> 2);
            atomicAdd(&slots[__umul24(slot, kGroupStride) + ((lo >> 16) & 63u)], v);
        }
        sbase += (uint32_t)__popcll(sm);
        prev_key = rlane(hi, 63);
    }
    wave_lds_sync();
    // cover bytes -> exclusive prefixes over the row's columns
    if (slot_lane) {
        uint4 c0 = row[0], c1 = row[1];
        uint32_t e = 0;
        auto pre = [&](uint32_t &v) {
            const uint32_t t = v & 0xFF000000u;
            v = (v & 0x00FFFFFFu) | e;
            e += t;
        };
        pre(c0.x), pre(c0.y), pre(c0.z), pre(c0.w), pre(c1.x), pre(c1.y), pre(c1.z), pre(c1.w);
        row[0] = c0;
        row[1] = c1;
    }
    wave_lds_sync();
    *r0 = r;
    return after ? (uint32_t)__ffsll((long long)after) - 1u : b.cnt;
}

// record_area for a record that group_build has served (r0: the rank in slot
// 0): a run of at most kGroupLenMax words reads its pixel's word of the
// slots, a longer one takes run_area.
__device__ __forceinline__ int32_t grouped_area(uint32_t *slots, const Stage &st, const uint64_t *segs,
                                                const Batch &b, uint32_t q, uint32_t r0, uint32_t py) {
    const uint32_t key = rlane(b.key, q);
    const int32_t cy = as_i8((int32_t)(rlane64(b.cov, q) >> (8 * py)));
    if (!(key & kRecRun)) return 16 * cy;
    if ((b.longm >> q) & 1ull) return run_area(st, (int32_t *)slots, segs, rlane(b.beg, q), rlane(b.len, q), cy);
    const uint32_t slot = (uint32_t)__popcll(b.rm & ~(~0ull << q)) - r0, lane = lane_id();
    const uint32_t v = slots[__umul24(slot, kGroupStride) + (lane & 7u) * 8u + (lane >> 3)];  // (y, x)
    return 16 * as_i8(cy + (int32_t)(v >> 24)) + as_i16((int32_t)v);
}

// The source terms of blend_at for one pixel (painter/mod.rs:364-400):
// blend(dst, src) * alpha per channel, one dispatch on the layer's (uniform)
// blend mode for all three channels.
__device__ __forceinline__ void blend_terms(uint32_t mode, float c0, float c1, float c2, float f0, float f1,
                                            float f2, float a, float &b0, float &b1, float &b2) {
    switch (mode) {
#define MG_BLEND_CASE(m)                  \
    case m:                               \
        b0 = blend_lane(m, c0, f0) * a;   \
        b1 = blend_lane(m, c1, f1) * a;   \
        b2 = blend_lane(m, c2, f2) * a;   \
        break;
        MG_BLEND_CASE(BM_MULTIPLY)
        MG_BLEND_CASE(BM_SCREEN)
        MG_BLEND_CASE(BM_OVERLAY)
        MG_BLEND_CASE(BM_DARKEN)
        MG_BLEND_CASE(BM_LIGHTEN)
        MG_BLEND_CASE(BM_COLOR_DODGE)
        MG_BLEND_CASE(BM_COLOR_BURN)
        MG_BLEND_CASE(BM_HARD_LIGHT)
        MG_BLEND_CASE(BM_SOFT_LIGHT)
        MG_BLEND_CASE(BM_DIFFERENCE)
        MG_BLEND_CASE(BM_EXCLUSION)
#undef MG_BLEND_CASE
    default:  // Over, and modes outside the table (blend_lane's default)
        b0 = f0 * a;
        b1 = f1 * a;
        b2 = f2 * a;
        break;
    }
}

// blend_terms for a source colour that is the same over the wave (a solid
// layer): the parts of a formula that depend on the source alone are not
// repeated per lane. Overlay's choice depends on the source, so only the
// chosen side is computed; Soft Light's source terms (polynomial or square
// root, s·(1−s), D(s)−s) are computed once per channel in three active lanes
// and read back: the lanes blend in whole columns of 8, so the first active
// lane and the two after it are active. The operations and their operands
// are those of blend_lane, so the results are the same bits.
__device__ __forceinline__ float overlay_uniform(float d, float s) {
    if (s <= 0.5f) return (d * s) * 2.0f;
    return 2.0f * ((s + d) - __builtin_fmaf(s, d, 0.5f));
}
// The props table seen as constant memory: read at a wave-uniform index, its
// entries come through scalar loads (the table is not written while a frame
// paints), so a record's flags and solid colour need no lane reads.
typedef uint32_t PropsHead __attribute__((ext_vector_type(8)));  // flags, clip_layers, color[4], grad_start
static_assert(offsetof(PropsEntry, color) == 8 && offsetof(PropsEntry, grad_start) == 24 &&
                  sizeof(PropsEntry) == 48, "props entry layout");
typedef const __attribute__((address_space(4))) PropsHead *ScalarPropsHead;

// NaN test on the bits of a (wave-uniform) float: scalar instructions.
__device__ __forceinline__ bool nan_bits(float x) { return (__float_as_uint(x) & 0x7FFFFFFFu) > 0x7F800000u; }

// With PF_SOFT_TERMS (flags `f` of entry `pe`) Soft Light reads the source
// terms of its three channels from the entry (words 6..11, one scalar load of
// words 4..11) instead of computing them. Darken and Lighten with a source
// that is not NaN take one comparison per channel: fmin_rs(d, s) is then
// d <= s ? d : s (a NaN d fails the comparison and gives s), fmax_rs(d, s)
// is d >= s ? d : s.
__device__ __forceinline__ void blend_terms_uniform(uint32_t mode, float c0, float c1, float c2, float f0,
                                                    float f1, float f2, float a, float &b0, float &b1,
                                                    float &b2, uint32_t f, const PropsEntry *pe) {
    const float fs[3] = {f0, f1, f2};
    const float cs[3] = {c0, c1, c2};
    float out[3];
    switch (mode) {
    case BM_OVERLAY:
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) out[ch] = overlay_uniform(cs[ch], fs[ch]) * a;
        break;
    case BM_DARKEN:
    case BM_LIGHTEN:
        if (nan_bits(f0) || nan_bits(f1) || nan_bits(f2)) {
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) out[ch] = blend_lane(mode, cs[ch], fs[ch]) * a;
        } else if (mode == BM_DARKEN) {
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) out[ch] = (cs[ch] <= fs[ch] ? cs[ch] : fs[ch]) * a;
        } else {
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) out[ch] = (cs[ch] >= fs[ch] ? cs[ch] : fs[ch]) * a;
        }
        break;
    case BM_SOFT_LIGHT: {
        float us[3], vs[3];
        if (f & PF_SOFT_TERMS) {
            const PropsHead w = *(ScalarPropsHead)((uintptr_t)pe + 16);
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                us[ch] = __uint_as_float(w[2 + ch]);
                vs[ch] = __uint_as_float(w[5 + ch]);
            }
        } else {
            const uint32_t l0 = (uint32_t)__builtin_ctzll(__ballot(true));
            const uint32_t own = lane_id() - l0;
            const float sl = own == 0 ? f0 : (own == 1 ? f1 : f2);
            float u, v;
            soft_light_terms(sl, u, v);
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                us[ch] = rlane_f(u, l0 + (uint32_t)ch);
                vs[ch] = rlane_f(v, l0 + (uint32_t)ch);
            }
        }
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
            const float k = __builtin_fmaf(2.0f, cs[ch], -1.0f);
            const float lo = __builtin_fmaf(us[ch], k, fs[ch]);
            const float hi = __builtin_fmaf(vs[ch], k, fs[ch]);
            out[ch] = (cs[ch] <= 0.5f ? lo : hi) * a;
        }
        break;
    }
#define MG_BLEND_UNIFORM_CASE(m)                                                     \
    case m:                                                                          \
        _Pragma("unroll") for (int ch = 0; ch < 3; ++ch) out[ch] = blend_lane(m, cs[ch], fs[ch]) * a; \
        break;
        MG_BLEND_UNIFORM_CASE(BM_MULTIPLY)
        MG_BLEND_UNIFORM_CASE(BM_SCREEN)
        MG_BLEND_UNIFORM_CASE(BM_COLOR_DODGE)
        MG_BLEND_UNIFORM_CASE(BM_COLOR_BURN)
        MG_BLEND_UNIFORM_CASE(BM_HARD_LIGHT)
        MG_BLEND_UNIFORM_CASE(BM_DIFFERENCE)
        MG_BLEND_UNIFORM_CASE(BM_EXCLUSION)
#undef MG_BLEND_UNIFORM_CASE
    default:  // Over, and modes outside the table (blend_lane's default)
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) out[ch] = fs[ch] * a;
        break;
    }
    b0 = out[0];
    b1 = out[1];
    b2 = out[2];
}

// dst * (1 - alpha) + term (painter/mod.rs:364-400), as fused multiply-adds
// that update the accumulators in place (the compiler's accumulate form
// writes into the term's register and copies it back on every layer).
__device__ __forceinline__ void fma_in_place(float &c, float m, float t) {
    asm("v_fma_f32 %0, %0, %1, %2" : "+v"(c) : "v"(m), "v"(t));
}
__device__ __forceinline__ void blend_apply(float &c0, float &c1, float &c2, float &ca, float b0, float b1, float b2,
                                            float a) {
    const float inv = 1.0f - a;
    fma_in_place(c0, inv, b0);
    fma_in_place(c1, inv, b1);
    fma_in_place(c2, inv, b2);
    fma_in_place(ca, inv, a);
}

// The first 32 bytes of entry `lay` (flags and solid colour), one scalar load.
__device__ __forceinline__ PropsHead props_head(const PropsEntry *props, uint32_t lay) {
    return *(ScalarPropsHead)((uintptr_t)(props + lay));
}

// The painter over records [skipped, n) of a tile (painter/mod.rs:260-400):
// coverage, clip masks (only in scenes with clips), column skipping and
// blending into this lane's pixel colour.
// `group(b, q, r0)` prepares the area of the batch's records from q on and
// returns the record after those it prepared (setting r0 for them);
// `area(b, k0, q, r0)` then gives the accumulated area A of record k0 + q
// for this lane's pixel. With Staged the batch's runs are staged first.
template <bool Clips, bool Staged, class AreaFn, class GroupFn>
__device__ __forceinline__ void paint_layers(const PaintParams &pp, const PropsEntry *props, const uint32_t *fl,
                                             const uint4 *rec, const uint32_t *rlen, uint32_t n, uint32_t skipped,
                                             Stage &st, const uint64_t *segs, uint32_t px, uint32_t py, uint32_t c,
                                             uint32_t j, float &c0, float &c1, float &c2, float &ca, AreaFn &&area,
                                             GroupFn &&group) {
    bool has_clip = false;
    uint32_t clip_last = 0;
    float clip_mask = 0.0f;
    for (uint32_t k0 = skipped, cnt = 0; k0 < n; k0 += cnt) {
        Batch b;
        load_batch<Staged>(pp, props, fl, rec, rlen, n, k0, b);
        if (Staged) stage_batch(st, segs, b);
        cnt = b.cnt;
        for (uint32_t q = 0; q < cnt;) {
        uint32_t r0 = 0;
        const uint32_t qe = group(b, q, r0);
        for (; q < qe; ++q) {
            // with clips the flags are the clip pass's (per record); without,
            // the layer's, read with its colour through scalar loads, which
            // the coverage computation below leaves time to arrive
            const uint32_t key = rlane(b.key, q);
            const uint32_t lay = key & 0xFFFFu;
            PropsHead ph;
            if (!Clips) ph = props_head(props, lay);
            const uint32_t f = Clips ? rlane(b.flags, q) : ph[0];
            if (Clips && !(f & kMaskBit)) continue;
            if (Clips && has_clip && clip_last < lay) has_clip = false;
            const int32_t A = area(b, k0, q, r0);
            const float cov = coverage_from_area(A, (key & kRecEvenOdd) != 0);
            if (Clips && (f & PF_CLIP)) {
                // clip_at (painter/mod.rs:402-416)
                if (!has_clip) {
                    has_clip = true;
                    clip_last = (lay + props[lay].clip_layers) & 0xFFFFu;
                }
                clip_mask = cov;
                continue;
            }
            const bool apply_clip = Clips && (f & PF_CLIPPED) && !(f & kSkipClipBit);
            const uint64_t nz = __ballot(cov != 0.0f);
            const bool col_zero = ((nz >> (8 * px)) & 0xFFull) == 0;
            if (col_zero || (apply_clip && !has_clip)) continue;
            // blend_at (painter/mod.rs:364-400); a solid layer's colour stays
            // in scalar registers (its own copy of the blend terms), and one
            // accumulation serves both
            const uint32_t mode = (f >> PF_BLEND_SHIFT) & 15u;
            float t0, t1, t2, a;
            if (f & PF_GRADIENT) {
                float g[4];
                gradient_color(props[lay], pp.stops, (float)(px + 1 + c * 8), (float)(j * 8), py, g);
                a = g[3] * cov;
                if (apply_clip) a *= clip_mask;
                blend_terms(mode, c0, c1, c2, g[0], g[1], g[2], a, t0, t1, t2);
            } else {
                const float s0 = Clips ? rlane_f(b.color.x, q) : __uint_as_float(ph[2]);
                const float s1 = Clips ? rlane_f(b.color.y, q) : __uint_as_float(ph[3]);
                const float s2 = Clips ? rlane_f(b.color.z, q) : __uint_as_float(ph[4]);
                const float s3 = Clips ? rlane_f(b.color.w, q) : __uint_as_float(ph[5]);
                a = s3 * cov;
                if (apply_clip) a *= clip_mask;
                blend_terms_uniform(mode, c0, c1, c2, s0, s1, s2, a, t0, t1, t2, f, props + lay);
            }
            blend_apply(c0, c1, c2, ca, t0, t1, t2, a);
        }
        }
    }
}

// Deep tiles (pp.deep_min records or more, no clips, no coverage dump) are
// painted by k_paint_deep, four waves per tile: waves 1..3 compute the
// coverage of the next 192 records while wave 0 blends the current 192. The
// row walk lists them (kDeepFlag in the header's count) up to pp.deep_cap.

// Waves 1..3 store the area A of their 64 records of a round for every pixel,
// reduced so that coverage_from_area gives the same coverage (nonzero: A
// clamped to [-256, 256]; even-odd: A mod 512); wave 0 blends from them.
__device__ __forceinline__ void paint_deep(const PaintParams &pp, const PropsEntry *props, const uint4 *rec,
                                           const uint32_t *rlen, uint32_t n, uint32_t skipped, Stage &st,
                                           int32_t *cells, const uint64_t *segs, uint32_t px, uint32_t py,
                                           uint32_t c, uint32_t j, float &c0, float &c1, float &c2, float &ca,
                                           int16_t *abuf) {
    const uint32_t wv = threadIdx.x >> 6, lane = lane_id();
    const uint32_t nchunks = (n - skipped + kDeepChunk - 1) / kDeepChunk;
    auto produce = [&](uint32_t chunk) {
        int16_t *A = abuf + (size_t)(chunk & 1u) * kDeepChunk * 64;
        const uint32_t base = skipped + chunk * kDeepChunk;
        const uint32_t r0 = base + (wv - 1) * 64;
        const uint32_t r1 = r0 + 64 < n ? r0 + 64 : n;
        for (uint32_t k0 = r0, cnt = 0; k0 < r1; k0 += cnt) {
            Batch b;
            load_batch(pp, props, nullptr, rec, rlen, r1, k0, b);
            stage_batch(st, segs, b);
            cnt = b.cnt;
            for (uint32_t q = 0; q < cnt; ++q) {
                const int32_t a = record_area(st, cells, segs, b, q, py);
                const bool eo = rlane(b.flags, q) & PF_EVENODD;
                const int32_t red = eo ? (a & 511) : (a < -256 ? -256 : (a > 256 ? 256 : a));
                A[(k0 + q - base) * 64 + lane] = (int16_t)red;
            }
        }
    };
    if (wv > 0 && nchunks > 0) produce(0);
    __syncthreads();
    for (uint32_t ch = 0; ch < nchunks; ++ch) {
        if (wv > 0) {
            if (ch + 1 < nchunks) produce(ch + 1);
        } else {
            const int16_t *A = abuf + (size_t)(ch & 1u) * kDeepChunk * 64;
            const uint32_t base = skipped + ch * kDeepChunk;
            const uint32_t end = base + kDeepChunk < n ? base + kDeepChunk : n;
            paint_layers<false, false>(pp, props, nullptr, rec, rlen, end, base, st, segs, px, py, c, j, c0, c1, c2,
                                       ca,
                                       [&](const Batch &, uint32_t k0, uint32_t q, uint32_t) {
                                           return (int32_t)A[(k0 + q - base) * 64 + lane];
                                       },
                                       [](const Batch &b, uint32_t, uint32_t &) { return b.cnt; });
        }
        __syncthreads();
    }
}

// One tile (LayerWorkbench::drive_tile_painting, layer_workbench.rs:430-486,
// and Painter, painter/mod.rs:260-493): one wave, or (Deep) four waves of
// which wave 0 holds the pixels' colours. In Deep every wave runs the
// workbench passes; only thread 0 / wave 0 write.
template <bool Deep>
__device__ __forceinline__ void paint_tile(const PaintParams &pp, uint32_t grow, uint32_t c,
                                           unsigned char *lds_raw) {
    const uint32_t lane = lane_id(), wv = Deep ? threadIdx.x >> 6 : 0u;
    const uint32_t sc = pp.row_scene[grow];
    const SceneDesc S = pp.scenes[sc];
    const uint32_t j = grow - S.row_base;
    const PropsEntry *props = pp.props + S.props_base;
    const uint32_t px = lane >> 3, py = lane & 7u;  // pixel of this lane (x*8+y)
    const bool clips = S.flags & SF_CLIPS;
    const uint64_t tile = (uint64_t)S.tile_base + (uint64_t)j * S.row_len + c;
    const uint2 hdr = pp.rec_hdr[tile];
    // a tile listed for the deep kernel is painted there
    if (!Deep && (hdr.y & kDeepFlag)) return;
    const uint32_t n = hdr.y & ~kDeepFlag;
    const uint4 *rec = pp.rec + hdr.x;
    const uint32_t *rlen = pp.rec_len + hdr.x;
    Stage st;
    st.w = (uint64_t *)(lds_raw + (size_t)wv * kPaintWaveLds);
    st.base = st.end = 0;
    int32_t *cells = (int32_t *)(lds_raw + (size_t)wv * kPaintWaveLds + kStageWin * 8);
    uint32_t *slots = (uint32_t *)cells;  // run-group slots (shared with the per-run cells)
    uint32_t *fl = clips ? (uint32_t *)(lds_raw + kPaintWaveLds) : nullptr;
    const uint64_t *segs = pp.segs;
    if (fl) {
        for (uint32_t k = lane; k < n; k += 64)
            fl[k] = (props[rec[k].x & 0xFFFFu].flags & ~(kMaskBit | kSkipClipBit)) | kMaskBit;
        __syncthreads();
    }

    // ---- coverage dump (debug): A of every id, independent of the passes
    if (pp.cov_a) {
        for (uint32_t k0 = 0, cnt = 0; k0 < n; k0 += cnt) {
            Batch b;
            load_batch(pp, props, fl, rec, rlen, n, k0, b);
            stage_batch(st, segs, b);
            cnt = b.cnt;
            for (uint32_t q = 0, qe = 0, r0 = 0; q < cnt; ++q) {
                if (q == qe) qe = group_build(slots, st, b, q, &r0);
                const int32_t A = grouped_area(slots, st, segs, b, q, r0, py);
                uint32_t r = 0;
                if (lane == 0) r = atomicAdd(pp.cov_count, 1u);
                r = __shfl(r, 0);
                if (r < pp.cov_cap) {
                    pp.cov_a[(uint64_t)r * 64 + lane] = A;
                    if (lane == 0) {
                        pp.cov_key[r * 4 + 0] = sc;
                        pp.cov_key[r * 4 + 1] = j;
                        pp.cov_key[r * 4 + 2] = c;
                        pp.cov_key[r * 4 + 3] = rlane(b.key, q) & 0xFFFFu;
                    }
                }
            }
        }
    }

    auto id_flags = [&](uint32_t k, uint32_t lay) {
        return fl ? fl[k] : ((props[lay].flags & ~(kMaskBit | kSkipClipBit)) | kMaskBit);
    };
    // a carried-only id whose cover is full (Cover::is_full)
    auto is_full = [&](const uint4 &rv, uint32_t f) {
        if ((rv.x & kRecRun) || !(rv.x & kRecCarry)) return false;
        return cover_full((uint64_t)rv.z | ((uint64_t)rv.w << 32), f & PF_EVENODD);
    };

    // ---- workbench passes (layer_workbench.rs:390-399)
    int op = 2;  // 0 None, 1 Solid, 2 ColorBuffer
    float solid0 = 0.0f, solid1 = 0.0f, solid2 = 0.0f, solid3 = 0.0f;
    uint32_t skipped = 0;
    if (S.flags & SF_CACHE) {
        // tile_unchanged_pass (layer_workbench.rs:200-222)
        bool changed = false;
        for (uint32_t k0 = 0; k0 < n; k0 += 64) {
            const uint32_t k = k0 + lane;
            bool ch = false;
            if (k < n) ch = !(props[rec[k].x & 0xFFFFu].flags & PF_UNCHANGED);
            if (__ballot(ch)) changed = true;
        }
        int32_t *prev = pp.prev_layers + tile;
        const int32_t old = *prev;
        const int32_t now = (int32_t)(uint16_t)n;
        __syncthreads();
        if (threadIdx.x == 0) *prev = now;
        if (old >= 0 && old == now && !changed) op = 0;
    }
    if (op != 0 && clips) {
        // skip_trivial_clips_pass (layer_workbench.rs:224-308): serial, uniform
        bool has = false, full = false, used = false;
        uint32_t last = 0, ci = 0;
        for (uint32_t k = 0; k < n; ++k) {
            uint32_t f = fl[k];
            const uint4 rv = rec[k];
            const uint32_t lay = rv.x & 0xFFFFu;
            if (!(f & kMaskBit)) continue;
            if (f & PF_CLIP) {
                const bool isfull = is_full(rv, f);
                has = true;
                full = isfull;
                last = (lay + props[lay].clip_layers) & 0xFFFFu;
                ci = k;
                used = false;
                if (isfull) f &= ~kMaskBit;
            } else if (f & PF_CLIPPED) {
                if (has && lay <= last) {
                    if (full) f |= kSkipClipBit; else used = true;
                } else {
                    f &= ~kMaskBit;
                }
            }
            __syncthreads();
            if (lane == 0) fl[k] = f;
            if (has && lay > last) {
                has = false;
                if (!used && lane == 0) fl[ci] &= ~kMaskBit;
            }
            __syncthreads();
        }
        if (has && !used && lane == 0) fl[ci] &= ~kMaskBit;
        __syncthreads();
    }
    if (op != 0) {
        // skip_fully_covered_layers (layer_workbench.rs:310-388)
        int first = 0;  // 0 unset, 1 opaque, 2 incomplete
        int32_t k_opaque = -1;
        // without opaque layers in the scene, the search only decides
        // whether the tile is solid: it ends at the first incomplete id
        const bool no_opaque = (S.flags & SF_NO_OPAQUE) != 0;
        for (int32_t top = (int32_t)n - 1; top >= 0; top -= 64) {
            const int32_t k = top - (int32_t)lane;
            bool pI = false, pO = false;
            if (k >= 0) {
                const uint4 rv = rec[k];
                const uint32_t f = id_flags((uint32_t)k, rv.x & 0xFFFFu);
                if (f & kMaskBit) {
                    const bool clipped = !(f & PF_CLIP) && (f & PF_CLIPPED) && !(f & kSkipClipBit);
                    pI = clipped || !is_full(rv, f);
                    pO = !pI && (f & PF_OPAQUE);
                }
            }
            const uint64_t bI = __ballot(pI), bO = __ballot(pO);
            const uint64_t before = bO ? ((1ull << (__ffsll((long long)bO) - 1)) - 1ull) : ~0ull;
            if (first == 0 && (bI & before)) first = 2;
            if (no_opaque && first == 2) break;
            if (bO) {
                const uint32_t l = (uint32_t)__ffsll((long long)bO) - 1;
                k_opaque = top - (int32_t)l;
                if (first == 0) first = 1;
                break;
            }
        }
        if (k_opaque >= 0) skipped = (uint32_t)k_opaque;
        if (first != 2) {
            uint32_t from = 0;
            float col[4];
            if (first == 1) {
                const PropsEntry &pe = props[rec[k_opaque].x & 0xFFFFu];
                col[0] = pe.color[0]; col[1] = pe.color[1]; col[2] = pe.color[2]; col[3] = pe.color[3];
                from = (uint32_t)k_opaque + 1;
            } else {
                col[0] = S.clear[0]; col[1] = S.clear[1]; col[2] = S.clear[2]; col[3] = S.clear[3];
            }
            bool ok = true;
            for (uint32_t k0 = from; k0 < n; k0 += 64) {
                const uint32_t k = k0 + lane;
                bool bad = false;
                if (k < n) {
                    const uint32_t f = id_flags(k, rec[k].x & 0xFFFFu);
                    bad = (f & kMaskBit) && ((f & PF_CLIP) || (f & PF_GRADIENT));
                }
                if (__ballot(bad)) ok = false;
            }
            if (ok) {
                for (uint32_t k0 = from; k0 < n; k0 += 64) {
                    Batch b;
                    load_batch(pp, props, fl, rec, rlen, n, k0, b);
                    const uint32_t cnt = (uint32_t)__builtin_amdgcn_readfirstlane((int)(n - k0 < 64 ? n - k0 : 64));
                    for (uint32_t q = 0; q < cnt; ++q) {
                        const uint32_t f = rlane(b.flags, q);
                        if (!(f & kMaskBit)) continue;
                        const float src[4] = {rlane_f(b.color.x, q), rlane_f(b.color.y, q), rlane_f(b.color.z, q),
                                              rlane_f(b.color.w, q)};
                        blend_color((f >> PF_BLEND_SHIFT) & 15u, col, src);
                    }
                }
                op = 1;
                solid0 = col[0]; solid1 = col[1]; solid2 = col[2]; solid3 = col[3];
            }
        }
    }

    // ---- paint (TileWriteOp::ColorBuffer, layer_workbench.rs:449-485) and write
    uint32_t pix = 0;
    if (op == 2) {
        float c0 = S.clear[0], c1 = S.clear[1], c2 = S.clear[2], ca = S.clear[3];
        auto area = [&](const Batch &b, uint32_t, uint32_t q, uint32_t r0) {
            return grouped_area(slots, st, segs, b, q, r0, py);
        };
        auto group = [&](const Batch &b, uint32_t q, uint32_t &r0) { return group_build(slots, st, b, q, &r0); };
        if (Deep)
            paint_deep(pp, props, rec, rlen, n, skipped, st, cells, segs, px, py, c, j, c0, c1, c2, ca,
                       (int16_t *)(lds_raw + 4 * kPaintWaveLds));
        else if (clips)
            paint_layers<true, true>(pp, props, fl, rec, rlen, n, skipped, st, segs, px, py, c, j, c0, c1, c2, ca,
                                     area, group);
        else
            paint_layers<false, true>(pp, props, nullptr, rec, rlen, n, skipped, st, segs, px, py, c, j, c0, c1, c2,
                                      ca, area, group);
        pix = pixel_bytes(c0, c1, c2, ca);
    } else if (op == 1) {
        const float sc4[4] = {solid0, solid1, solid2, solid3};
        pix = solid_bytes(sc4);
    }
    if (op != 0 && wv == 0) {
        const uint32_t gx = c * 8 + px, gy = j * 8 + py;
        if (gx < S.width && gy < S.rows_px) {
            uint32_t *o = (uint32_t *)(pp.out + S.out_offset) + (uint64_t)gy * S.stride + gx;
            *o = pix;
        }
    }
    if (pp.tile_ops && threadIdx.x == 0) pp.tile_ops[tile] = (int8_t)op;
}

// One wave per tile of the work list; at most 64 VGPRs, for 8 waves per SIMD.
__global__ __launch_bounds__(64, 8) void k_paint(PaintParams pp) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    // the row walk stopped early: its records are incomplete
    if (__hip_atomic_load(pp.overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    const uint2 wk = pp.work[blockIdx.x];
    paint_tile<false>(pp, wk.x, wk.y, lds_raw);
}

// Four waves per deep tile: work group t paints entry t of the deep list.
__global__ __launch_bounds__(256) void k_paint_deep(PaintParams pp) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    if (__hip_atomic_load(pp.overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    const uint32_t count = *pp.deep_count;
    if (blockIdx.x >= count || blockIdx.x >= pp.deep_cap) return;
    const uint2 wk = pp.deep_list[blockIdx.x];
    paint_tile<true>(pp, wk.x, wk.y, lds_raw);
}

size_t paint_deep_lds_bytes_host() { return (size_t)4 * kPaintWaveLds + (size_t)2 * kDeepChunk * 64 * 2; }

}  // namespace mg
