// HIP kernels of the MI355X mold renderer (gfx950, wave64).
//
// Pipeline of one frame (all scenes of a batch in the same launches):
//   k_setup        per line: LinesBuilder::build           (segment.rs:91-237)
//   k_scan_u32     exclusive scan of line lengths          (rasterizer/mod.rs:56-125)
//   k_emit         per line-pixel: 1-2 PixelSegments,      (rasterizer/mod.rs:127-285)
//                  compacted in generation order, with all sort-digit
//                  histograms and per-tile-row counts fused in
//   k_digit_bases  histogram -> digit base offsets
//   k_sort_pass    one stable LSD digit pass (onesweep: block rank + decoupled look-back)
//   k_paint        one wave per tile row walks its tiles:  (painter/mod.rs:495-596,
//                  carries, workbench decisions, coverage,  layer_workbench.rs:134-486)
//                  blending and RGBA8 stores
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mg_exact.h"
#include "mg_kernels.h"
#include "mg_shared.h"

namespace mg {

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

__device__ __forceinline__ uint64_t ld_acquire(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_acquire32(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release32(uint32_t *p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Decoupled look-back status words: 2 flag bits above a 62-bit value.
constexpr uint64_t kStAgg = 1ull << 62;
constexpr uint64_t kStInc = 2ull << 62;
constexpr uint64_t kStVal = (1ull << 62) - 1;

// Exclusive prefix of the partitions before `part`; executed by one full
// wave. Predecessors are always running or finished because partitions are
// handed out in order by an atomic counter.
__device__ uint64_t lookback_wave(const uint64_t *status, uint32_t part) {
    const uint32_t lane = lane_id();
    uint64_t excl = 0;
    int64_t top = (int64_t)part - 1;
    while (top >= 0) {
        int64_t idx = top - (int64_t)lane;
        uint64_t s = idx >= 0 ? ld_acquire(status + idx) : kStInc;
        uint64_t flag = s >> 62;
        uint64_t not_ready = __ballot(flag == 0);
        uint64_t inc = __ballot(flag == 2);
        uint64_t first_inc = inc ? (uint64_t)__ffsll((long long)inc) - 1 : 64;
        uint64_t blocked = not_ready & ((first_inc >= 64) ? ~0ull : ((1ull << first_inc) - 1));
        if (blocked) {
            __builtin_amdgcn_s_sleep(1);
            continue;
        }
        uint64_t v = (lane <= first_inc && idx >= 0) ? (s & kStVal) : 0;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        excl += v;
        if (first_inc < 64) break;
        top -= 64;
    }
    return excl;
}

// Block-wide exclusive scan of one u32 per thread (blockDim = 256).
__device__ uint32_t block_excl_scan_256(uint32_t v, uint32_t *wsum /* 4 */, uint32_t *total) {
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(x, off);
        if (lane >= (uint32_t)off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (uint32_t k = 0; k < 4; ++k) {
        uint32_t s = wsum[k];
        if (k < w) base += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return base + x - v;
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

// LinesBuilder::build for one line (segment.rs:107-219).
__global__ __launch_bounds__(256) void k_setup(const float4 *__restrict__ pts, const uint16_t *__restrict__ ids,
                                               const LayerEntry *__restrict__ layers,
                                               const SceneDesc *__restrict__ scenes,
                                               const uint32_t *__restrict__ scene_line_base,
                                               FrameDims dims, LineSetup *__restrict__ out,
                                               uint32_t *__restrict__ lengths, uint32_t *__restrict__ err) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= dims.n_lines) return;
    uint32_t sc = dims.n_scenes == 1 ? 0 : scene_of_line(scene_line_base, dims.n_scenes, i);
    const SceneDesc &S = scenes[sc];
    uint32_t id = ids[i];
    LayerEntry le;
    le.flags = 0;
    if (id < S.n_layer_entries) le = layers[S.layer_base + id];
    LineSetup r;
    r.start = 0;
    r.start_f32 = r.end_f32 = r.slope = r.slope_recip = r.py_slope_px = r.px_slope_recip_py = 0.0f;
    r.packed = sc << 19;
    uint32_t length = 0;
    bool exists = le.flags & LF_EXISTS;
    if (!(exists && !(le.flags & LF_ENABLED))) {
        uint32_t order = (exists && (le.flags & LF_ORDER)) ? le.order : id;
        float4 p = pts[i];
        float p0x = p.x, p0y = p.y, p1x = p.z, p1y = p.w;
        if (exists && (le.flags & LF_AFFINE)) {
            const float *t = le.affine;
            float a0 = __builtin_fmaf(t[0], p0x, __builtin_fmaf(t[1], p0y, t[4]));
            float b0 = __builtin_fmaf(t[2], p0x, __builtin_fmaf(t[3], p0y, t[5]));
            float a1 = __builtin_fmaf(t[0], p1x, __builtin_fmaf(t[1], p1y, t[4]));
            float b1 = __builtin_fmaf(t[2], p1x, __builtin_fmaf(t[3], p1y, t[5]));
            p0x = a0; p0y = b0; p1x = a1; p1y = b1;
        }
        int32_t x0 = sub_pixel(p0x), y0 = sub_pixel(p0y), x1 = sub_pixel(p1x), y1 = sub_pixel(p1y);
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
    out[i] = r;
    lengths[i] = length;
}

// ------------------------------------------------------------------- scan

// Single-pass exclusive scan of u32 values (decoupled look-back), 2048
// values per partition. out may alias in. total receives the sum (u64).
__global__ __launch_bounds__(256) void k_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n,
                                                  uint64_t *status, uint32_t *part_counter,
                                                  uint64_t *total) {
    __shared__ uint32_t s_part, wsum[4];
    __shared__ uint64_t s_excl;
    constexpr uint32_t IPT = 8;
    if (threadIdx.x == 0) s_part = atomicAdd(part_counter, 1u);
    __syncthreads();
    const uint32_t part = s_part;
    const uint64_t base = (uint64_t)part * 256 * IPT + (uint64_t)threadIdx.x * IPT;
    uint32_t v[IPT], sum = 0;
    for (uint32_t k = 0; k < IPT; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        sum += v[k];
    }
    uint32_t agg;
    uint32_t ex = block_excl_scan_256(sum, wsum, &agg);
    if (threadIdx.x == 0) st_release(status + part, (part == 0 ? kStInc : kStAgg) | agg);
    if (threadIdx.x < 64) {
        uint64_t e = part ? lookback_wave(status, part) : 0;
        if (threadIdx.x == 0) {
            s_excl = e;
            if (part) st_release(status + part, kStInc | (e + agg));
            if ((uint64_t)(part + 1) * 256 * IPT >= n) *total = e + agg;
        }
    }
    __syncthreads();
    uint64_t run = s_excl + ex;
    for (uint32_t k = 0; k < IPT; ++k) {
        if (base + k < n) out[base + k] = (uint32_t)run;
        run += v[k];
    }
}

// -------------------------------------------------------------- emission

__device__ __forceinline__ uint64_t pack_segment(uint32_t layer, int32_t x0, int32_t x1, int32_t y0, int32_t y1,
                                                 int32_t bx, int32_t by, uint32_t octant) {
    // tiles() (rasterizer/mod.rs:134-147)
    if (octant == 1 || octant == 2 || octant == 5 || octant == 6) { int32_t t = bx; bx = by; by = t; }
    int32_t ti = as_i16(bx >> 3);
    int32_t tj = as_i16(by >> 3);
    uint32_t tx = (uint32_t)(bx & 7), ty = (uint32_t)(by & 7);
    // area_cover() (rasterizer/mod.rs:149-172)
    int32_t ax0, ax1, ay0, ay1;
    switch (octant) {
    case 1: ax0 = y0; ax1 = y1; ay0 = x0; ay1 = x1; break;
    case 2: ax0 = y1; ax1 = y0; ay0 = x0; ay1 = x1; break;
    case 3: case 4: ax0 = x0; ax1 = x1; ay0 = y1; ay1 = y0; break;
    case 5: ax0 = y0; ax1 = y1; ay0 = x1; ay1 = x0; break;
    case 6: ax0 = y1; ax1 = y0; ay0 = x1; ay1 = x0; break;
    default: ax0 = x0; ax1 = x1; ay0 = y0; ay1 = y1; break;
    }
    int32_t border = wrap_add(ax0 & ~15, 16);
    int32_t h = wrap_sub(ay1, ay0);
    int32_t tri = as_i16(wrap_mul(h, wrap_sub(ax1, ax0)) / 2);
    int32_t rect = as_i16(wrap_mul(h, wrap_sub(border, ax1)));
    int32_t area = as_i16(tri + rect);
    int32_t cover = as_i8(h);
    int32_t tib = as_i16((ti < -1 ? -1 : ti) + 256);
    uint64_t w = ((uint64_t)(uint32_t)tj & 0xFFFull) << 51;
    w |= ((uint64_t)(uint32_t)tib & 0x1FFFull) << 38;
    w |= (uint64_t)(layer & 0xFFFFu) << 22;
    w |= (uint64_t)ty << 19;
    w |= (uint64_t)tx << 16;
    w |= ((uint64_t)(uint32_t)area & 0x3FFull) << 6;
    w |= (uint64_t)(uint32_t)cover & 0x3Full;
    return w;
}

// Rasterizer::rasterize for one line-pixel (rasterizer/mod.rs:233-281).
__device__ __forceinline__ int rasterize_pixel(const LineSetup &l, int32_t pi, uint64_t out[2]) {
    uint32_t layer = l.packed & 0xFFFFu, octant = (l.packed >> 16) & 7u;
    int32_t x0i = wrap_add(l.start, pi);
    int32_t x1i = wrap_add(x0i, 1);
    float x0 = fmax_rs((float)x0i, l.start_f32);
    float x1 = fmin_rs((float)x1i, l.end_f32);
    float y0 = __builtin_fmaf(l.slope, x0, l.py_slope_px);
    float y1 = __builtin_fmaf(l.slope, x1, l.py_slope_px);
    int32_t x0s = sub_pixel(x0), x1s = sub_pixel(x1), y0s = sub_pixel(y0), y1s = sub_pixel(y1);
    if ((y0s & 15) != 0 && (y1s & 15) != 0 && (y0s >> 4) != (y1s >> 4)) {
        int32_t y = (y0s > y1s ? y0s : y1s) >> 4;
        int32_t ys = (int32_t)((uint32_t)y << 4);
        float x = __builtin_fmaf(l.slope_recip, (float)y, l.px_slope_recip_py);
        int32_t xs = sub_pixel(x);
        out[0] = pack_segment(layer, x0s, xs, y0s, ys, x0i, (ys < y0s ? ys : y0s) >> 4, octant);
        out[1] = pack_segment(layer, xs, x1s, ys, y1s, x0i, (ys < y1s ? ys : y1s) >> 4, octant);
        return 2;
    }
    out[0] = pack_segment(layer, x0s, x1s, y0s, y1s, x0i, (y0s < y1s ? y0s : y1s) >> 4, octant);
    return 1;
}

constexpr uint32_t kEmitThreads = 256;
constexpr uint32_t kEmitIPT = 4;
constexpr uint32_t kEmitTile = kEmitThreads * kEmitIPT;
constexpr uint32_t kMaxPasses = 8;
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

// Emission: one thread per kEmitIPT consecutive line-pixels (blocked), a
// persistent grid claiming partitions of kEmitTile line-pixels in order.
// Output positions follow generation order (line, pixel, slot), so the
// stable sort that follows yields the canonical order.
__global__ __launch_bounds__(256) void k_emit(const LineSetup *__restrict__ lines,
                                              const uint32_t *__restrict__ offsets,
                                              const SceneDesc *__restrict__ scenes, FrameDims dims,
                                              FrameCounters *__restrict__ counters, uint64_t *__restrict__ out,
                                              uint64_t capacity, uint64_t *status, uint32_t *part_counter,
                                              uint32_t *__restrict__ hist /* passes x radix */,
                                              uint32_t *__restrict__ row_count) {
    __shared__ uint32_t s_part, wsum[4];
    __shared__ uint64_t s_excl;
    __shared__ uint32_t s_line[2];
    __shared__ uint32_t s_hist[kMaxPasses * 256];
    __shared__ uint32_t s_off[kEmitTile + 1];
    const uint32_t radix = 1u << dims.radix_bits;
    const bool lds_hist = dims.passes * radix <= kMaxPasses * 256;
    for (uint32_t k = threadIdx.x; k < kMaxPasses * 256; k += blockDim.x) s_hist[k] = 0;
    const uint64_t P = counters->line_pixels;
    const uint32_t nparts = (uint32_t)((P + kEmitTile - 1) / kEmitTile);
    const uint32_t lane = lane_id();
    uint64_t seg_total = 0;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_part = atomicAdd(part_counter, 1u);
        __syncthreads();
        const uint32_t part = s_part;
        if (part >= nparts) break;
        const uint64_t b0 = (uint64_t)part * kEmitTile;
        const uint64_t b1 = b0 + kEmitTile < P ? b0 + kEmitTile : P;
        // lines covering [b0, b1): last line with offset <= x
        if (threadIdx.x < 2) {
            uint64_t x = threadIdx.x == 0 ? b0 : b1 - 1;
            uint32_t lo = 0, hi = dims.n_lines;
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if ((uint64_t)offsets[mid] <= x) lo = mid; else hi = mid;
            }
            s_line[threadIdx.x] = lo;
        }
        __syncthreads();
        const uint32_t l0 = s_line[0], l1 = s_line[1];
        const bool in_lds = l1 - l0 + 1 <= kEmitTile;
        if (in_lds)
            for (uint32_t k = threadIdx.x; k <= l1 - l0; k += blockDim.x) s_off[k] = offsets[l0 + k];
        __syncthreads();
        // this thread's line-pixels
        const uint64_t t0 = b0 + (uint64_t)threadIdx.x * kEmitIPT;
        uint64_t segs[2 * kEmitIPT];
        uint32_t rows[2 * kEmitIPT];
        uint32_t cnt = 0;
        if (t0 < b1) {
            uint32_t li;
            if (in_lds) {
                uint32_t lo = 0, hi = l1 - l0 + 1;
                while (hi - lo > 1) {
                    uint32_t mid = (lo + hi) >> 1;
                    if ((uint64_t)s_off[mid] <= t0) lo = mid; else hi = mid;
                }
                li = l0 + lo;
            } else {
                uint32_t lo = l0, hi = l1 + 1;
                while (hi - lo > 1) {
                    uint32_t mid = (lo + hi) >> 1;
                    if ((uint64_t)offsets[mid] <= t0) lo = mid; else hi = mid;
                }
                li = lo;
            }
            LineSetup L = lines[li];
            uint64_t next = li + 1 < dims.n_lines ? offsets[li + 1] : P;
            // lengths of 0 share offsets: skip to the line that owns t0
            while (next <= t0) {
                ++li;
                L = lines[li];
                next = li + 1 < dims.n_lines ? offsets[li + 1] : P;
            }
            uint32_t sc = L.packed >> 19;
            SceneDesc S = scenes[sc];
            for (uint32_t k = 0; k < kEmitIPT; ++k) {
                uint64_t p = t0 + k;
                if (p >= b1) break;
                while (p >= next) {
                    ++li;
                    L = lines[li];
                    next = li + 1 < dims.n_lines ? offsets[li + 1] : P;
                    uint32_t s2 = L.packed >> 19;
                    if (s2 != sc) { sc = s2; S = scenes[sc]; }
                }
                uint64_t w[2];
                int n = rasterize_pixel(L, (int32_t)(p - offsets[li]), w);
                seg_total += (uint64_t)n;
                for (int q = 0; q < n; ++q) {
                    if (dims.mode == 0) {
                        uint64_t rw;
                        uint32_t grow;
                        if (render_word(w[q], S, dims, &rw, &grow)) {
                            segs[cnt] = rw;
                            rows[cnt] = grow;
                            ++cnt;
                        }
                    } else {
                        segs[cnt] = w[q];
                        rows[cnt] = 0;
                        ++cnt;
                    }
                }
            }
        }
        uint32_t agg;
        uint32_t ex = block_excl_scan_256(cnt, wsum, &agg);
        if (threadIdx.x == 0) st_release(status + part, (part == 0 ? kStInc : kStAgg) | agg);
        if (threadIdx.x < 64) {
            uint64_t e = part ? lookback_wave(status, part) : 0;
            if (threadIdx.x == 0) {
                s_excl = e;
                if (part) st_release(status + part, kStInc | (e + agg));
                if (part == nparts - 1) counters->kept = e + agg;
            }
        }
        __syncthreads();
        const uint64_t dst = s_excl + ex;
        for (uint32_t q = 0; q < cnt; ++q) {
            uint64_t w = segs[q];
            if (dst + q < capacity) out[dst + q] = w;
            uint32_t shift = kPayloadBits;
            for (uint32_t ps = 0; ps < dims.passes; ++ps, shift += dims.radix_bits) {
                uint32_t dg = (uint32_t)(w >> shift) & (radix - 1);
                if (lds_hist) atomicAdd(&s_hist[ps * radix + dg], 1u);
                else atomicAdd(&hist[ps * radix + dg], 1u);
            }
        }
        // per-tile-row counts, aggregated across the wave
        if (dims.mode == 0) {
            for (uint32_t q = 0; q < 2 * kEmitIPT; ++q) {
                bool has = q < cnt;
                uint32_t r = has ? rows[q] : 0;
                uint64_t pending = __ballot(has);
                while (pending) {
                    uint32_t leader = (uint32_t)__ffsll((long long)pending) - 1;
                    uint32_t lr = __shfl(r, leader);
                    uint64_t same = __ballot(has && r == lr);
                    if (lane == leader) atomicAdd(&row_count[lr], (uint32_t)__popcll(same));
                    if (has && r == lr) has = false;
                    pending &= ~same;
                }
            }
        }
    }
    __syncthreads();
    if (lds_hist)
        for (uint32_t k = threadIdx.x; k < dims.passes * radix; k += blockDim.x)
            if (s_hist[k]) atomicAdd(&hist[k], s_hist[k]);
    for (int off = 32; off > 0; off >>= 1) seg_total += __shfl_xor(seg_total, off);
    if (lane == 0 && seg_total) atomicAdd((unsigned long long *)&counters->segments, (unsigned long long)seg_total);
}

// Ingest of an external PixelSegment stream (the cassia_render entry,
// cassia/src/Cassia.h:34-39): NONE entries and segments of layers without
// styling are dropped, the rest become render words in input order, with the
// same digit histograms and row counts as k_emit. Scene 0 only.
__global__ __launch_bounds__(256) void k_ingest(const uint64_t *__restrict__ in, uint64_t n_in,
                                                const SceneDesc *__restrict__ scenes, FrameDims dims,
                                                FrameCounters *__restrict__ counters, uint64_t *__restrict__ out,
                                                uint64_t *status, uint32_t *part_counter,
                                                uint32_t *__restrict__ hist, uint32_t *__restrict__ row_count) {
    __shared__ uint32_t s_part, wsum[4];
    __shared__ uint64_t s_excl;
    __shared__ uint32_t s_hist[kMaxPasses * 256];
    const uint32_t radix = 1u << dims.radix_bits;
    const bool lds_hist = dims.passes * radix <= kMaxPasses * 256;
    for (uint32_t k = threadIdx.x; k < kMaxPasses * 256; k += blockDim.x) s_hist[k] = 0;
    const SceneDesc S = scenes[0];
    const uint32_t nparts = (uint32_t)((n_in + kEmitTile - 1) / kEmitTile);
    const uint32_t lane = lane_id();
    uint64_t live = 0;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_part = atomicAdd(part_counter, 1u);
        __syncthreads();
        const uint32_t part = s_part;
        if (part >= nparts) break;
        const uint64_t t0 = (uint64_t)part * kEmitTile + (uint64_t)threadIdx.x * kEmitIPT;
        uint64_t segs[kEmitIPT];
        uint32_t rows[kEmitIPT];
        uint32_t cnt = 0;
        for (uint32_t k = 0; k < kEmitIPT; ++k) {
            uint64_t p = t0 + k;
            if (p >= n_in) break;
            uint64_t w = in[p];
            if (w >> 63) continue;
            ++live;
            if (((uint32_t)(w >> 22) & 0xFFFFu) >= S.n_props) continue;
            uint64_t rw;
            uint32_t grow;
            if (render_word(w, S, dims, &rw, &grow)) {
                segs[cnt] = rw;
                rows[cnt] = grow;
                ++cnt;
            }
        }
        uint32_t agg;
        uint32_t ex = block_excl_scan_256(cnt, wsum, &agg);
        if (threadIdx.x == 0) st_release(status + part, (part == 0 ? kStInc : kStAgg) | agg);
        if (threadIdx.x < 64) {
            uint64_t e = part ? lookback_wave(status, part) : 0;
            if (threadIdx.x == 0) {
                s_excl = e;
                if (part) st_release(status + part, kStInc | (e + agg));
                if (part == nparts - 1) counters->kept = e + agg;
            }
        }
        __syncthreads();
        const uint64_t dst = s_excl + ex;
        for (uint32_t q = 0; q < cnt; ++q) {
            uint64_t w = segs[q];
            out[dst + q] = w;
            uint32_t shift = kPayloadBits;
            for (uint32_t ps = 0; ps < dims.passes; ++ps, shift += dims.radix_bits) {
                uint32_t dg = (uint32_t)(w >> shift) & (radix - 1);
                if (lds_hist) atomicAdd(&s_hist[ps * radix + dg], 1u);
                else atomicAdd(&hist[ps * radix + dg], 1u);
            }
        }
        for (uint32_t q = 0; q < kEmitIPT; ++q) {
            bool has = q < cnt;
            uint32_t r = has ? rows[q] : 0;
            uint64_t pending = __ballot(has);
            while (pending) {
                uint32_t leader = (uint32_t)__ffsll((long long)pending) - 1;
                uint32_t lr = __shfl(r, leader);
                uint64_t same = __ballot(has && r == lr);
                if (lane == leader) atomicAdd(&row_count[lr], (uint32_t)__popcll(same));
                if (has && r == lr) has = false;
                pending &= ~same;
            }
        }
    }
    __syncthreads();
    if (lds_hist)
        for (uint32_t k = threadIdx.x; k < dims.passes * radix; k += blockDim.x)
            if (s_hist[k]) atomicAdd(&hist[k], s_hist[k]);
    for (int off = 32; off > 0; off >>= 1) live += __shfl_xor(live, off);
    if (lane == 0 && live) atomicAdd((unsigned long long *)&counters->segments, (unsigned long long)live);
}

// Exclusive scan of each pass's digit histogram (one block per pass).
__global__ __launch_bounds__(256) void k_digit_bases(const uint32_t *hist, uint32_t *bases, uint32_t radix) {
    __shared__ uint32_t wsum[4];
    const uint32_t *h = hist + blockIdx.x * radix;
    uint32_t *b = bases + blockIdx.x * radix;
    const uint32_t per = (radix + 255) / 256;
    uint32_t local = 0;
    for (uint32_t k = 0; k < per; ++k) {
        uint32_t d = threadIdx.x * per + k;
        local += d < radix ? h[d] : 0u;
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan_256(local, wsum, &tot);
    for (uint32_t k = 0; k < per; ++k) {
        uint32_t d = threadIdx.x * per + k;
        if (d < radix) {
            b[d] = ex;
            ex += h[d];
        }
    }
}

// ------------------------------------------------------------------- sort

constexpr uint32_t kSortThreads = 256;
constexpr uint32_t kSortKPT = 16;
constexpr uint32_t kSortTile = kSortThreads * kSortKPT;  // 4096 keys
constexpr uint32_t kSt32Agg = 1u << 30;
constexpr uint32_t kSt32Inc = 2u << 30;
constexpr uint32_t kSt32Val = (1u << 30) - 1u;

// One stable LSD digit pass. Each wave ranks its 1024 keys in 16 steps of
// 64 (striped, so step-major order equals input order) with ballot-based
// peer matching; the block's digit counts are published for decoupled
// look-back; keys are reordered in LDS and written as digit runs.
template <int R>
__global__ __launch_bounds__(256) void k_sort_pass(const uint64_t *__restrict__ in, uint64_t *__restrict__ out,
                                                   const FrameCounters *__restrict__ counters, uint32_t shift,
                                                   const uint32_t *__restrict__ digit_base, uint32_t *status,
                                                   uint32_t *part_counter) {
    constexpr uint32_t RADIX = 1u << R;
    constexpr uint32_t DPT = (RADIX + 255) / 256;  // digits per thread
    __shared__ uint64_t s_keys[kSortTile];
    __shared__ uint32_t s_wcnt[4][RADIX];
    __shared__ uint32_t s_dexcl[RADIX];
    __shared__ uint32_t s_glob[RADIX];
    __shared__ uint32_t s_part, wsum[4];
    const uint64_t n = counters->kept;
    const uint32_t nparts = (uint32_t)((n + kSortTile - 1) / kSortTile);
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) s_part = atomicAdd(part_counter, 1u);
        for (uint32_t k = threadIdx.x; k < 4 * RADIX; k += 256) (&s_wcnt[0][0])[k] = 0;
        __syncthreads();
        const uint32_t part = s_part;
        if (part >= nparts) break;
        const uint64_t base = (uint64_t)part * kSortTile;
        const uint32_t valid_n = (uint32_t)(n - base < kSortTile ? n - base : kSortTile);
        uint64_t keys[kSortKPT];
        uint32_t rank[kSortKPT];
        for (uint32_t i = 0; i < kSortKPT; ++i) {
            uint32_t li = w * 1024 + i * 64 + lane;
            bool valid = li < valid_n;
            uint64_t key = valid ? in[base + li] : 0ull;
            uint32_t d = (uint32_t)(key >> shift) & (RADIX - 1);
            uint64_t peers = __ballot(valid);
#pragma unroll
            for (int b = 0; b < R; ++b) {
                uint64_t bits = __ballot((d >> b) & 1u);
                peers &= ((d >> b) & 1u) ? bits : ~bits;
            }
            uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
            uint32_t c = valid ? s_wcnt[w][d] : 0u;
            __builtin_amdgcn_wave_barrier();
            if (valid && below == 0) s_wcnt[w][d] = c + (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
            keys[i] = key;
            rank[i] = valid ? c + below : 0xFFFFFFFFu;
        }
        __syncthreads();
        // per digit: wave prefixes and block total
        uint32_t tot_local = 0;
        uint32_t tots[DPT];
        for (uint32_t k = 0; k < DPT; ++k) {
            uint32_t d = threadIdx.x * DPT + k;
            uint32_t t = 0;
            if (d < RADIX) {
                for (uint32_t ww = 0; ww < 4; ++ww) {
                    uint32_t v = s_wcnt[ww][d];
                    s_wcnt[ww][d] = t;
                    t += v;
                }
            }
            tots[k] = t;
            tot_local += t;
        }
        uint32_t all;
        uint32_t ex = block_excl_scan_256(tot_local, wsum, &all);
        for (uint32_t k = 0; k < DPT; ++k) {
            uint32_t d = threadIdx.x * DPT + k;
            if (d < RADIX) {
                s_dexcl[d] = ex;
                ex += tots[k];
                st_release32(status + (uint64_t)part * RADIX + d, (part == 0 ? kSt32Inc : kSt32Agg) | tots[k]);
            }
        }
        // decoupled look-back per digit
        for (uint32_t k = 0; k < DPT; ++k) {
            uint32_t d = threadIdx.x * DPT + k;
            if (d >= RADIX) continue;
            uint32_t excl = 0;
            int64_t p = (int64_t)part - 1;
            while (p >= 0) {
                uint32_t s = ld_acquire32(status + (uint64_t)p * RADIX + d);
                uint32_t f = s >> 30;
                if (f == 0) { __builtin_amdgcn_s_sleep(1); continue; }
                excl += s & kSt32Val;
                if (f == 2) break;
                --p;
            }
            if (part) st_release32(status + (uint64_t)part * RADIX + d, kSt32Inc | (excl + tots[k]));
            s_glob[d] = digit_base[d] + excl - s_dexcl[d];
        }
        __syncthreads();
        for (uint32_t i = 0; i < kSortKPT; ++i) {
            if (rank[i] == 0xFFFFFFFFu) continue;
            uint32_t d = (uint32_t)(keys[i] >> shift) & (RADIX - 1);
            s_keys[s_dexcl[d] + s_wcnt[w][d] + rank[i]] = keys[i];
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < valid_n; k += 256) {
            uint64_t key = s_keys[k];
            uint32_t d = (uint32_t)(key >> shift) & (RADIX - 1);
            out[(uint64_t)s_glob[d] + k] = key;
        }
    }
}

template __global__ void k_sort_pass<4>(const uint64_t *, uint64_t *, const FrameCounters *, uint32_t,
                                        const uint32_t *, uint32_t *, uint32_t *);
template __global__ void k_sort_pass<6>(const uint64_t *, uint64_t *, const FrameCounters *, uint32_t,
                                        const uint32_t *, uint32_t *, uint32_t *);
template __global__ void k_sort_pass<8>(const uint64_t *, uint64_t *, const FrameCounters *, uint32_t,
                                        const uint32_t *, uint32_t *, uint32_t *);

// ------------------------------------------------------------------ paint

// LDS layout of the paint kernel (one wave per work group), `cap` ids /
// carries and `rcap` runs:
//   carries (double-buffered): layer u32, cover u64 (8 x i8, byte y = row y)
//   ids: layer u32, run i32, carry i32, props flags u32, mask u32
//   runs: layer u32, begin u32, end u32, cover sums 8 x i32
//   cells: area i32[64], cover i32[64] (pixel = x*8+y)
struct PaintLds {
    uint32_t *c_layer[2];
    uint64_t *c_cov[2];
    uint32_t *i_layer;
    int32_t *i_run;
    int32_t *i_carry;
    uint32_t *i_flags;
    uint32_t *i_mask;  // bit0 mask, bit1 skip_clipping
    uint32_t *r_layer, *r_beg, *r_end;
    int32_t *r_cov;
    int32_t *cell_area, *cell_cov;
};

size_t paint_lds_bytes_host(uint32_t cap, uint32_t rcap) {
    return (size_t)cap * (2 * 4 + 2 * 8 + 5 * 4) + (size_t)rcap * (3 * 4 + 8 * 4) + 2 * 64 * 4;
}

__device__ __forceinline__ PaintLds paint_lds(unsigned char *lds, uint32_t cap, uint32_t rcap) {
    PaintLds L;
    unsigned char *p = lds;
    L.c_cov[0] = (uint64_t *)p; p += cap * 8;
    L.c_cov[1] = (uint64_t *)p; p += cap * 8;
    L.c_layer[0] = (uint32_t *)p; p += cap * 4;
    L.c_layer[1] = (uint32_t *)p; p += cap * 4;
    L.i_layer = (uint32_t *)p; p += cap * 4;
    L.i_run = (int32_t *)p; p += cap * 4;
    L.i_carry = (int32_t *)p; p += cap * 4;
    L.i_flags = (uint32_t *)p; p += cap * 4;
    L.i_mask = (uint32_t *)p; p += cap * 4;
    L.r_layer = (uint32_t *)p; p += rcap * 4;
    L.r_beg = (uint32_t *)p; p += rcap * 4;
    L.r_end = (uint32_t *)p; p += rcap * 4;
    L.r_cov = (int32_t *)p; p += rcap * 32;
    L.cell_area = (int32_t *)p; p += 256;
    L.cell_cov = (int32_t *)p;
    return L;
}

__device__ __forceinline__ int32_t cov_byte(uint64_t c, uint32_t y) { return as_i8((int32_t)(c >> (8 * y))); }

// Cover::is_empty / is_full over 8 rows (painter/mod.rs:164-189).
__device__ __forceinline__ bool cover_empty(uint64_t c, bool eo) {
    if (!eo) return c == 0;
    for (uint32_t y = 0; y < 8; ++y) {
        int32_t v = cov_byte(c, y);
        int32_t a = as_i8(v < 0 ? -v : v);
        if ((a & 31) != 0) return false;
    }
    return true;
}
__device__ __forceinline__ bool cover_full(uint64_t c, bool eo) {
    for (uint32_t y = 0; y < 8; ++y) {
        int32_t v = cov_byte(c, y);
        int32_t a = as_i8(v < 0 ? -v : v);
        if (eo ? ((a & 31) != 16) : (a != 16)) return false;
    }
    return true;
}
__device__ __forceinline__ uint64_t cover_add(uint64_t a, const int32_t *sums /* 8 */) {
    uint64_t r = 0;
    for (uint32_t y = 0; y < 8; ++y) {
        uint32_t v = (uint32_t)(cov_byte(a, y) + sums[y]) & 0xFFu;
        r |= (uint64_t)v << (8 * y);
    }
    return r;
}

// Gradient::color_at for one pixel row `row` of the column at (gx, gy)
// (style.rs:96-175).
__device__ void gradient_color(const PropsEntry &p, const float *stops, float gx, float gy, uint32_t row,
                               float out[4]) {
    float dx = p.grad_end[0] - p.grad_start[0];
    float dy = p.grad_end[1] - p.grad_start[1];
    float dot = dx * dx + dy * dy;
    float dot_recip = 1.0f / dot;
    float t;
    if (!(p.flags & PF_RADIAL)) {
        float tx = (gx - p.grad_start[0]) * dx * dot_recip;
        float ty = gy - p.grad_start[1];
        t = __builtin_fmaf(((float)row + ty) * dy, dot_recip, tx);
    } else {
        float px = gx - p.grad_start[0];
        float px2 = px * px;
        float py = (float)row + (gy - p.grad_start[1]);
        t = __builtin_sqrtf(__builtin_fmaf(py, py, px2) * dot_recip);
    }
    const float *st = stops + p.stops_offset;
    uint32_t ch[4] = {0, 0, 0, 0};
    bool acc = t <= st[4];
    if (acc)
        for (int c = 0; c < 4; ++c) ch[c] |= __float_as_uint(st[c]);
    float start_stop = 0.0f;
    const float *start_color = st;
    for (uint32_t k = 1; k < p.n_stops; ++k) {
        const float *color = st + 5 * k;
        float end_stop = color[4];
        bool m = acc ^ (t < end_stop);
        if (m) {
            float d = end_stop - start_stop;
            float lt = (t - start_stop) * (1.0f / d);
            for (int c = 0; c < 4; ++c) {
                float v = __builtin_fmaf(lt, color[c], __builtin_fmaf(-lt, start_color[c], start_color[c]));
                ch[c] |= __float_as_uint(v);
            }
            acc = true;
        }
        start_stop = end_stop;
        start_color = color;
    }
    if (!acc) {
        const float *last = st + 5 * (p.n_stops - 1);
        for (int c = 0; c < 4; ++c) ch[c] |= __float_as_uint(last[c]);
    }
    for (int c = 0; c < 4; ++c) out[c] = __uint_as_float(ch[c]);
}

__device__ __forceinline__ uint32_t seg_colp(uint64_t w, uint32_t lb, uint32_t cmask) {
    return (uint32_t)(w >> (kPayloadBits + lb)) & cmask;
}

// Reads the runs (maximal same-layer groups) of tile column `colp` starting
// at *pos; returns the number of runs or -1 on overflow.
__device__ int32_t build_runs(const PaintParams &pp, const PaintLds &L, uint64_t *pos, uint64_t end,
                              uint32_t colp, uint32_t rcap) {
    const uint32_t lane = lane_id();
    const uint32_t lb = pp.layer_bits, lmask = (1u << lb) - 1u;
    const uint32_t cmask = (1u << pp.col_bits) - 1u;
    uint64_t p = *pos;
    int32_t nruns = 0;
    uint32_t prev_layer = 0xFFFFFFFFu;
    for (;;) {
        uint64_t idx = p + lane;
        uint64_t w = idx < end ? pp.segs[idx] : ~0ull;
        bool in = idx < end && seg_colp(w, lb, cmask) == colp;
        uint64_t inmask = __ballot(in);
        uint32_t n_in = (uint32_t)__popcll(inmask);  // contiguous prefix (sorted)
        uint32_t layer = (uint32_t)(w >> kPayloadBits) & lmask;
        uint32_t prevl = __shfl_up(layer, 1);
        if (lane == 0) prevl = prev_layer;
        bool head = in && layer != prevl;
        uint64_t heads = __ballot(head);
        uint32_t nh = (uint32_t)__popcll(heads);
        if ((uint32_t)nruns + nh > rcap) return -1;
        uint32_t my_run = (uint32_t)nruns + (uint32_t)__popcll(heads & ((2ull << lane) - 1ull)) - 1u;
        if (head) {
            L.r_layer[my_run] = layer;
            L.r_beg[my_run] = (uint32_t)(idx - *pos);
            if (my_run > 0) L.r_end[my_run - 1] = (uint32_t)(idx - *pos);
            for (uint32_t y = 0; y < 8; ++y) L.r_cov[my_run * 8 + y] = 0;
        }
        __syncthreads();
        if (in) {
            int32_t cover = (int32_t)((int64_t)(w << 58) >> 58);
            uint32_t ty = (uint32_t)(w >> 19) & 7u;
            atomicAdd(&L.r_cov[my_run * 8 + ty], cover);
        }
        nruns += (int32_t)nh;
        prev_layer = __shfl(layer, n_in ? n_in - 1 : 0);
        if (n_in == 0) prev_layer = 0xFFFFFFFFu;
        p += n_in;
        if (n_in < 64) break;
    }
    if (nruns > 0 && lane == 0) L.r_end[nruns - 1] = (uint32_t)(p - *pos);
    __syncthreads();
    *pos = p;
    return nruns;
}

// ids = sorted union of run layers and carry layers; returns the count.
__device__ uint32_t merge_ids(const PaintLds &L, uint32_t cur, uint32_t nr, uint32_t nc) {
    const uint32_t lane = lane_id();
    const uint32_t n = nr + nc;
    const uint32_t *cl = L.c_layer[cur];
    for (uint32_t k0 = 0; k0 < n; k0 += 64) {
        uint32_t k = k0 + lane;
        if (k < nr) {
            uint32_t v = L.r_layer[k];
            uint32_t lo = 0, hi = nc;  // #carries < v
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (cl[mid] < v) lo = mid + 1; else hi = mid;
            }
            uint32_t s = k + lo;
            L.i_layer[s] = v;
            L.i_run[s] = (int32_t)k;
            L.i_carry[s] = -1;
        } else if (k < n) {
            uint32_t m = k - nr;
            uint32_t v = cl[m];
            uint32_t lo = 0, hi = nr;  // #runs <= v
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (L.r_layer[mid] <= v) lo = mid + 1; else hi = mid;
            }
            uint32_t s = m + lo;
            L.i_layer[s] = v;
            L.i_run[s] = -1;
            L.i_carry[s] = (int32_t)m;
        }
    }
    __syncthreads();
    // drop the carry entry that follows its run, moving its carry index over
    uint32_t out = 0;
    for (uint32_t k0 = 0; k0 < n; k0 += 64) {
        uint32_t k = k0 + lane;
        uint32_t lay = 0;
        int32_t run = -1, car = -1;
        bool keep = false;
        if (k < n) {
            lay = L.i_layer[k];
            run = L.i_run[k];
            car = L.i_carry[k];
            keep = !(k > 0 && L.i_layer[k - 1] == lay);
            if (keep && k + 1 < n && L.i_layer[k + 1] == lay) car = L.i_carry[k + 1];
        }
        uint64_t km = __ballot(keep);
        __syncthreads();
        if (keep) {
            uint32_t d = out + (uint32_t)__popcll(km & lanemask_lt());
            L.i_layer[d] = lay;
            L.i_run[d] = run;
            L.i_carry[d] = car;
        }
        out += (uint32_t)__popcll(km);
        __syncthreads();
    }
    return out;
}

// Cover of id k after the tile: carry in + run sums (i8 wrapping).
__device__ __forceinline__ uint64_t id_cover_out(const PaintLds &L, uint32_t cur, uint32_t k) {
    int32_t sums[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int32_t r = L.i_run[k];
    if (r >= 0)
        for (uint32_t y = 0; y < 8; ++y) sums[y] = L.r_cov[r * 8 + y];
    int32_t c = L.i_carry[k];
    uint64_t base = c >= 0 ? L.c_cov[cur][c] : 0ull;
    return cover_add(base, sums);
}

// New carry list from the ids (filtered: Cover::is_empty drops, as
// layer_workbench.rs:194,475 do; unfiltered while folding the left of the
// row, painter/mod.rs:515-550). Returns the count.
__device__ uint32_t next_carries(const PaintLds &L, uint32_t cur, uint32_t n, bool filter) {
    const uint32_t lane = lane_id();
    const uint32_t nxt = cur ^ 1u;
    uint32_t out = 0;
    for (uint32_t k0 = 0; k0 < n; k0 += 64) {
        uint32_t k = k0 + lane;
        bool keep = false;
        uint64_t c = 0;
        if (k < n) {
            c = id_cover_out(L, cur, k);
            keep = !filter || !cover_empty(c, L.i_flags[k] & PF_EVENODD);
        }
        uint64_t km = __ballot(keep);
        if (keep) {
            uint32_t d = out + (uint32_t)__popcll(km & lanemask_lt());
            L.c_layer[nxt][d] = L.i_layer[k];
            L.c_cov[nxt][d] = c;
        }
        out += (uint32_t)__popcll(km);
    }
    __syncthreads();
    return out;
}

// A = 16 * P + area for this lane's pixel and id k (painter/mod.rs:229-243,
// 347-362): cells from the id's run (segments at tbase + [beg, end)) and its
// carried cover in column 0; P is the i8 running cover prefix.
__device__ __forceinline__ int32_t pixel_area(const PaintParams &pp, const PaintLds &L, uint32_t cur,
                                              uint64_t tbase, uint32_t k, uint32_t px, uint32_t py) {
    const uint32_t lane = lane_id();
    const int32_t run = L.i_run[k];
    const int32_t car = L.i_carry[k];
    const int32_t cy = car >= 0 ? cov_byte(L.c_cov[cur][car], py) : 0;
    if (run < 0) return 16 * cy;
    L.cell_area[lane] = 0;
    L.cell_cov[lane] = 0;
    __syncthreads();
    const uint32_t beg = L.r_beg[run], en = L.r_end[run];
    for (uint32_t s0 = beg; s0 < en; s0 += 64) {
        uint32_t s = s0 + lane;
        if (s < en) {
            uint64_t w = pp.segs[tbase + s];
            int32_t cover = (int32_t)((int64_t)(w << 58) >> 58);
            int32_t area = (int32_t)((int64_t)(w << 48) >> 54);
            uint32_t cell = ((uint32_t)(w >> 16) & 7u) * 8u + ((uint32_t)(w >> 19) & 7u);
            atomicAdd(&L.cell_area[cell], area);
            atomicAdd(&L.cell_cov[cell], cover);
        }
    }
    __syncthreads();
    int32_t P = cy;
    for (uint32_t x = 0; x < 7; ++x)
        if (x < px) P += L.cell_cov[x * 8 + py];
    const int32_t A = 16 * as_i8(P) + as_i16(L.cell_area[lane]);
    __syncthreads();
    return A;
}

// One wave walks the tiles of one tile row.
__global__ __launch_bounds__(64) void k_paint(PaintParams pp) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t lane = lane_id();
    const uint32_t cap = pp.cap_ids, rcap = pp.cap_runs;
    PaintLds L = paint_lds(lds_raw, cap, rcap);
    const uint32_t grow = blockIdx.x;
    const uint32_t sc = pp.row_scene[grow];
    const SceneDesc S = pp.scenes[sc];
    const uint32_t j = grow - S.row_base;
    if (j < S.crop_v0 || j >= S.crop_v1) return;
    const PropsEntry *props = pp.props + S.props_base;
    uint64_t pos = pp.row_start[grow];
    const uint64_t end = pp.row_start[grow + 1];
    const uint32_t lb = pp.layer_bits, cmask = (1u << pp.col_bits) - 1u;
    const uint32_t px = lane >> 3, py = lane & 7u;  // pixel of this lane (x*8+y)
    const uint32_t h0 = S.crop_h0, h1 = S.crop_h1 < S.row_len ? S.crop_h1 : S.row_len;
    const bool clips = S.flags & SF_CLIPS;
    uint32_t cur = 0, nc = 0;

    // ---- fold every tile left of the first painted one, unfiltered
    //      (populate_covers, painter/mod.rs:515-550)
    while (pos < end) {
        const uint32_t c0 = seg_colp(pp.segs[pos], lb, cmask);
        if (c0 > h0) break;
        int32_t nr = build_runs(pp, L, &pos, end, c0, rcap);
        if (nr < 0 || (uint32_t)nr + nc > cap) {
            if (lane == 0) atomicOr(pp.overflow, 1u);
            return;
        }
        const uint32_t n = merge_ids(L, cur, (uint32_t)nr, nc);
        for (uint32_t k0 = 0; k0 < n; k0 += 64) {
            uint32_t k = k0 + lane;
            if (k < n) L.i_flags[k] = props[L.i_layer[k]].flags;
        }
        __syncthreads();
        nc = next_carries(L, cur, n, false);
        cur ^= 1u;
    }

    for (uint32_t c = h0; c < h1; ++c) {
        // ---- segments, runs and ids of tile c (populate_layers, layer_workbench.rs:401-428)
        const uint64_t tbase = pos;
        int32_t nr = build_runs(pp, L, &pos, end, c + 1, rcap);
        if (nr < 0 || (uint32_t)nr + nc > cap) {
            if (lane == 0) atomicOr(pp.overflow, 1u);
            return;
        }
        const uint32_t n = merge_ids(L, cur, (uint32_t)nr, nc);
        for (uint32_t k0 = 0; k0 < n; k0 += 64) {
            uint32_t k = k0 + lane;
            if (k < n) {
                L.i_flags[k] = props[L.i_layer[k]].flags;
                L.i_mask[k] = 1u;
            }
        }
        __syncthreads();

        // ---- coverage dump (debug): A of every id, independent of the passes
        if (pp.cov_a) {
            for (uint32_t k = 0; k < n; ++k) {
                const int32_t A = pixel_area(pp, L, cur, tbase, k, px, py);
                uint32_t rec = 0;
                if (lane == 0) rec = atomicAdd(pp.cov_count, 1u);
                rec = __shfl(rec, 0);
                if (rec < pp.cov_cap) {
                    pp.cov_a[(uint64_t)rec * 64 + lane] = A;
                    if (lane == 0) {
                        pp.cov_key[rec * 4 + 0] = sc;
                        pp.cov_key[rec * 4 + 1] = j;
                        pp.cov_key[rec * 4 + 2] = c;
                        pp.cov_key[rec * 4 + 3] = L.i_layer[k];
                    }
                }
            }
        }

        // ---- workbench passes (layer_workbench.rs:390-399)
        int op = 2;  // 0 None, 1 Solid, 2 ColorBuffer
        float solid[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        uint32_t skipped = 0;
        if (S.flags & SF_CACHE) {
            // tile_unchanged_pass (layer_workbench.rs:200-222)
            bool changed = false;
            for (uint32_t k0 = 0; k0 < n; k0 += 64) {
                uint32_t k = k0 + lane;
                if (__ballot(k < n && !(L.i_flags[k] & PF_UNCHANGED))) changed = true;
            }
            int32_t *prev = pp.prev_layers + S.cache_base + j * S.row_len + c;
            const int32_t old = *prev;
            const int32_t now = (int32_t)(uint16_t)n;
            __syncthreads();
            if (lane == 0) *prev = now;
            if (old >= 0 && old == now && !changed) op = 0;
        }
        if (op != 0 && clips) {
            // skip_trivial_clips_pass (layer_workbench.rs:224-308): serial, uniform
            bool has = false, full = false, used = false;
            uint32_t last = 0, ci = 0;
            for (uint32_t k = 0; k < n; ++k) {
                const uint32_t f = L.i_flags[k];
                const uint32_t lay = L.i_layer[k];
                uint32_t m = L.i_mask[k];
                if (!(m & 1u)) continue;
                if (f & PF_CLIP) {
                    bool isfull = false;
                    if (L.i_run[k] < 0 && L.i_carry[k] >= 0)
                        isfull = cover_full(L.c_cov[cur][L.i_carry[k]], f & PF_EVENODD);
                    has = true;
                    full = isfull;
                    last = (lay + props[lay].clip_layers) & 0xFFFFu;
                    ci = k;
                    used = false;
                    if (isfull) m &= ~1u;
                } else if (f & PF_CLIPPED) {
                    if (has && lay <= last) {
                        if (full) m |= 2u; else used = true;
                    } else {
                        m &= ~1u;
                    }
                }
                __syncthreads();
                if (lane == 0) L.i_mask[k] = m;
                if (has && lay > last) {
                    has = false;
                    if (!used && lane == 0) L.i_mask[ci] &= ~1u;
                }
                __syncthreads();
            }
            if (has && !used && lane == 0) L.i_mask[ci] &= ~1u;
            __syncthreads();
        }
        if (op != 0) {
            // skip_fully_covered_layers (layer_workbench.rs:310-388)
            int first = 0;  // 0 unset, 1 opaque, 2 incomplete
            int32_t k_opaque = -1;
            for (int32_t top = (int32_t)n - 1; top >= 0; top -= 64) {
                const int32_t k = top - (int32_t)lane;
                bool pI = false, pO = false;
                if (k >= 0) {
                    const uint32_t f = L.i_flags[k];
                    const uint32_t m = L.i_mask[k];
                    if (m & 1u) {
                        const bool clipped = !(f & PF_CLIP) && (f & PF_CLIPPED) && !(m & 2u);
                        bool isfull = false;
                        if (L.i_run[k] < 0 && L.i_carry[k] >= 0)
                            isfull = cover_full(L.c_cov[cur][L.i_carry[k]], f & PF_EVENODD);
                        pI = clipped || !isfull;
                        pO = !pI && (f & PF_OPAQUE);
                    }
                }
                const uint64_t bI = __ballot(pI), bO = __ballot(pO);
                const uint64_t before = bO ? ((1ull << (__ffsll((long long)bO) - 1)) - 1ull) : ~0ull;
                if (first == 0 && (bI & before)) first = 2;
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
                    const PropsEntry &po = props[L.i_layer[k_opaque]];
                    for (int q = 0; q < 4; ++q) col[q] = po.color[q];
                    from = (uint32_t)k_opaque + 1;
                } else {
                    for (int q = 0; q < 4; ++q) col[q] = S.clear[q];
                }
                bool ok = true;
                for (uint32_t k0 = from; k0 < n; k0 += 64) {
                    const uint32_t k = k0 + lane;
                    const bool bad = k < n && (L.i_mask[k] & 1u) &&
                                     ((L.i_flags[k] & PF_CLIP) || (L.i_flags[k] & PF_GRADIENT));
                    if (__ballot(bad)) ok = false;
                }
                if (ok) {
                    for (uint32_t k = from; k < n; ++k) {
                        if (!(L.i_mask[k] & 1u)) continue;
                        const PropsEntry &pk = props[L.i_layer[k]];
                        blend_color((pk.flags >> PF_BLEND_SHIFT) & 15u, col, pk.color);
                    }
                    op = 1;
                    for (int q = 0; q < 4; ++q) solid[q] = col[q];
                }
            }
        }

        // ---- paint (TileWriteOp::ColorBuffer, layer_workbench.rs:449-485) and write
        uint32_t pix = 0;
        if (op == 2) {
            float c0 = S.clear[0], c1 = S.clear[1], c2 = S.clear[2], ca = S.clear[3];
            bool has_clip = false;
            uint32_t clip_last = 0;
            float clip_mask = 0.0f;
            for (uint32_t k = skipped; k < n; ++k) {
                const uint32_t m = L.i_mask[k];
                if (!(m & 1u)) continue;
                const uint32_t f = L.i_flags[k];
                const uint32_t lay = L.i_layer[k];
                if (has_clip && clip_last < lay) has_clip = false;
                const int32_t A = pixel_area(pp, L, cur, tbase, k, px, py);
                const float cov = coverage_from_area(A, f & PF_EVENODD);
                if (f & PF_CLIP) {
                    // clip_at (painter/mod.rs:402-416)
                    if (!has_clip) {
                        has_clip = true;
                        clip_last = (lay + props[lay].clip_layers) & 0xFFFFu;
                    }
                    clip_mask = cov;
                    continue;
                }
                const bool apply_clip = (f & PF_CLIPPED) && !(m & 2u);
                const uint64_t nz = __ballot(cov != 0.0f);
                const bool col_zero = ((nz >> (8 * px)) & 0xFFull) == 0;
                if (col_zero || (apply_clip && !has_clip)) continue;
                const PropsEntry &pk = props[lay];
                float fill[4];
                if (f & PF_GRADIENT) {
                    gradient_color(pk, pp.stops, (float)(px + 1 + c * 8), (float)(j * 8), py, fill);
                } else {
                    fill[0] = pk.color[0]; fill[1] = pk.color[1]; fill[2] = pk.color[2]; fill[3] = pk.color[3];
                }
                // blend_at (painter/mod.rs:364-400)
                float a = fill[3] * cov;
                if (apply_clip) a *= clip_mask;
                const float inv = 1.0f - a;
                const uint32_t mode = (f >> PF_BLEND_SHIFT) & 15u;
                const float b0 = blend_lane(mode, c0, fill[0]) * a;
                const float b1 = blend_lane(mode, c1, fill[1]) * a;
                const float b2 = blend_lane(mode, c2, fill[2]) * a;
                c0 = __builtin_fmaf(c0, inv, b0);
                c1 = __builtin_fmaf(c1, inv, b1);
                c2 = __builtin_fmaf(c2, inv, b2);
                ca = __builtin_fmaf(ca, inv, a);
            }
            pix = pixel_bytes(c0, c1, c2, ca);
        } else if (op == 1) {
            pix = solid_bytes(solid);
        }
        if (op != 0) {
            const uint32_t gx = c * 8 + px, gy = j * 8 + py;
            if (gx < S.width && gy < S.rows_px) {
                uint32_t *o = (uint32_t *)(pp.out + S.out_offset) + (uint64_t)gy * S.stride + gx;
                *o = pix;
            }
        }
        if (pp.tile_ops && lane == 0) pp.tile_ops[S.cache_base + j * S.row_len + c] = (int8_t)op;

        // ---- carries for the next tile
        nc = next_carries(L, cur, n, true);
        cur ^= 1u;
    }
}

}  // namespace mg
