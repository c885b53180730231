// Exact binary32 arithmetic with Rust's semantics, for device code.
//
// The renderer must reproduce the reference's bytes, so every operation
// states its rounding: the build disables contraction (-ffp-contract=off) and
// keeps IEEE division and square root; each Rust `mul_add` is an explicit
// __builtin_fmaf; conversions saturate like Rust's `as`.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mg {

// `f32 as i32`: truncate toward zero, saturate, NaN -> 0. Written with
// selects only (the operand is clamped into range before the conversion).
__host__ __device__ __forceinline__ int32_t f2i_sat(float v) {
    const float c = __builtin_fminf(__builtin_fmaxf(v, -2147483648.0f), 2147483520.0f);
    int32_t r = (int32_t)c;
    r = v >= 2147483648.0f ? 2147483647 : r;
    return v == v ? r : 0;
}

// `f32 as u8`: truncate, saturate, NaN -> 0.
__host__ __device__ __forceinline__ uint32_t f2u8_sat(float v) {
    if (!(v > 0.0f)) return 0u;
    if (v >= 255.0f) return 255u;
    return (uint32_t)v;
}

// `f32::round`: ties away from zero (x - trunc(x) is exact).
__host__ __device__ __forceinline__ float round_away(float x) {
    float t = __builtin_truncf(x);
    if (__builtin_fabsf(x - t) >= 0.5f) t += __builtin_copysignf(1.0f, x);
    return t;
}

// `f32::min` / `f32::max`: a NaN operand yields the other; ties keep `a`.
__host__ __device__ __forceinline__ float fmin_rs(float a, float b) {
    return (a != a) ? b : ((b < a) ? b : a);
}
__host__ __device__ __forceinline__ float fmax_rs(float a, float b) {
    return (a != a) ? b : ((b > a) ? b : a);
}

// `f32::clamp`: NaN propagates.
__host__ __device__ __forceinline__ float clamp_rs(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

// to_sub_pixel (surpass/src/segment.rs:13-16)
__host__ __device__ __forceinline__ int32_t sub_pixel(float v) {
    return f2i_sat(__builtin_fmaf(v, 16.0f, __builtin_copysignf(0.5f, v)));
}

// `f32 as i32` in one instruction: v_cvt_i32_f32 truncates toward zero,
// saturates out-of-range values (infinities included) and converts NaN to 0,
// which is exactly Rust's `as` (f2i_sat spells the same out with selects).
__device__ __forceinline__ int32_t f2i_sat_hw(float v) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
__device__ __forceinline__ int32_t sub_pixel_hw(float v) {
    return f2i_sat_hw(__builtin_fmaf(v, 16.0f, __builtin_copysignf(0.5f, v)));
}

__host__ __device__ __forceinline__ int32_t wrap_add(int32_t a, int32_t b) {
    return (int32_t)((uint32_t)a + (uint32_t)b);
}
__host__ __device__ __forceinline__ int32_t wrap_sub(int32_t a, int32_t b) {
    return (int32_t)((uint32_t)a - (uint32_t)b);
}
__host__ __device__ __forceinline__ int32_t wrap_mul(int32_t a, int32_t b) {
    return (int32_t)((uint32_t)a * (uint32_t)b);
}
__host__ __device__ __forceinline__ int32_t as_i8(int32_t v) { return (int32_t)(int8_t)(uint8_t)(uint32_t)v; }
__host__ __device__ __forceinline__ int32_t as_i16(int32_t v) { return (int32_t)(int16_t)(uint16_t)(uint32_t)v; }

// linear_to_srgb_approx{,_simd} (painter/mod.rs:74-106): identical results.
__host__ __device__ __forceinline__ float srgb_approx(float l) {
    const float a = 0.20101772f, b = -0.51280147f, c = 1.344401f, d = -0.030656587f;
    float s = __builtin_sqrtf(l);
    float s3 = l * s;
    float m = l * 12.92f;
    float n = __builtin_fmaf(a, s3, __builtin_fmaf(b, l, __builtin_fmaf(c, s, d)));
    return (l <= 0.0031308f) ? m : n;
}

// from_area (painter/mod.rs:53-71)
__host__ __device__ __forceinline__ float coverage_from_area(int32_t a, bool even_odd) {
    if (!even_odd) return clamp_rs(__builtin_fabsf((float)a * (1.0f / 256.0f)), 0.0f, 1.0f);
    float capped = (float)(a & 255) * (1.0f / 256.0f);
    return ((a >> 8) & 1) == 0 ? capped : 1.0f - capped;
}

enum : uint32_t {
    BM_OVER = 0, BM_MULTIPLY, BM_SCREEN, BM_OVERLAY, BM_DARKEN, BM_LIGHTEN, BM_COLOR_DODGE,
    BM_COLOR_BURN, BM_HARD_LIGHT, BM_SOFT_LIGHT, BM_DIFFERENCE, BM_EXCLUSION
};

// The terms of Soft Light's formula that depend on the source channel s
// alone (style.rs:288-418, as in blend_lane below): u = s·(1−s) and
// v = D(s) − s, D being the polynomial for s ≤ 0.25 and the square root
// above. Host and device compute them with the same IEEE operations.
__host__ __device__ __forceinline__ void soft_light_terms(float s, float &u, float &v) {
    const float poly = __builtin_fmaf(__builtin_fmaf(16.0f, s, -12.0f), s, 4.0f) * s;
    const float dd = s <= 0.25f ? poly : __builtin_sqrtf(s);
    u = s * (1.0f - s);
    v = dd - s;
}

// Per-lane form of blend_function! (style.rs:288-418), used for painted
// pixels. `select(a, b, m)` is `m ? a : b` (simd/auto.rs:459-502).
__host__ __device__ __forceinline__ float blend_lane(uint32_t mode, float d, float s) {
    switch (mode) {
    case BM_OVER: return s;
    case BM_MULTIPLY: return d * s;
    case BM_SCREEN: return __builtin_fmaf(d, -s, d) + s;
    case BM_OVERLAY: {
        float lo = (d * s) * 2.0f;
        float hi = 2.0f * ((s + d) - __builtin_fmaf(s, d, 0.5f));
        return s <= 0.5f ? lo : hi;
    }
    case BM_DARKEN: return fmin_rs(d, s);
    case BM_LIGHTEN: return fmax_rs(d, s);
    case BM_COLOR_DODGE: {
        float v = fmin_rs(1.0f, s / (1.0f - d));
        return s == 0.0f ? 0.0f : v;
    }
    case BM_COLOR_BURN: {
        float v = 1.0f - fmin_rs(1.0f, (1.0f - s) / d);
        return s == 1.0f ? 1.0f : v;
    }
    case BM_HARD_LIGHT: {
        float lo = (d * s) * 2.0f;
        float hi = 2.0f * ((s + d) - __builtin_fmaf(s, d, 0.5f));
        return d <= 0.5f ? lo : hi;
    }
    case BM_SOFT_LIGHT: {
        float poly = __builtin_fmaf(__builtin_fmaf(16.0f, s, -12.0f), s, 4.0f) * s;
        float dd = s <= 0.25f ? poly : __builtin_sqrtf(s);
        float k = __builtin_fmaf(2.0f, d, -1.0f);
        float lo = __builtin_fmaf(s * (1.0f - s), k, s);
        float hi = __builtin_fmaf(dd - s, k, s);
        return d <= 0.5f ? lo : hi;
    }
    case BM_DIFFERENCE: return __builtin_fabsf(d - s);
    case BM_EXCLUSION: return __builtin_fmaf(-2.0f * d, s, d) + s;
    default: return s;
    }
}

// Scalar form BlendMode::blend_fn (style.rs:206-267), used by the
// solid-tile fold.
__host__ __device__ __forceinline__ float blend_scalar(uint32_t mode, float dst, float src) {
    switch (mode) {
    case BM_OVER: return src;
    case BM_MULTIPLY: return dst * src;
    case BM_SCREEN: return (dst + src) - (dst * src);
    case BM_OVERLAY: {
        if (src <= 0.5f) return (2.0f * src) * dst;
        float t = 2.0f * src - 1.0f;
        return (t + dst) - (t * dst);
    }
    case BM_DARKEN: return fmin_rs(dst, src);
    case BM_LIGHTEN: return fmax_rs(dst, src);
    case BM_COLOR_DODGE:
        if (src == 0.0f) return 0.0f;
        if (dst == 1.0f) return 1.0f;
        return fmin_rs(1.0f, src / (1.0f - dst));
    case BM_COLOR_BURN:
        if (src == 1.0f) return 1.0f;
        if (dst == 0.0f) return 0.0f;
        return 1.0f - fmin_rs(1.0f, (1.0f - src) / dst);
    case BM_HARD_LIGHT: {
        if (dst <= 0.5f) return (2.0f * dst) * src;
        float t = 2.0f * dst - 1.0f;
        return (t + src) - (t * src);
    }
    case BM_SOFT_LIGHT: {
        float dv = src <= 0.25f ? ((16.0f * src - 12.0f) * src + 4.0f) * src : __builtin_sqrtf(src);
        if (dst <= 0.5f) return src - ((1.0f - 2.0f * dst) * src) * (1.0f - src);
        return src + (2.0f * dst - 1.0f) * (dv - src);
    }
    case BM_DIFFERENCE: return __builtin_fabsf(dst - src);
    case BM_EXCLUSION: return (dst + src) - (2.0f * dst) * src;
    default: return src;
    }
}

// BlendMode::blend (style.rs:269-285)
__host__ __device__ __forceinline__ void blend_color(uint32_t mode, float dst[4], const float src[4]) {
    float a = src[3];
    float inv = 1.0f - a;
    float c0 = blend_scalar(mode, dst[0], src[0]) * a;
    float c1 = blend_scalar(mode, dst[1], src[1]) * a;
    float c2 = blend_scalar(mode, dst[2], src[2]) * a;
    dst[0] = __builtin_fmaf(dst[0], inv, c0);
    dst[1] = __builtin_fmaf(dst[1], inv, c1);
    dst[2] = __builtin_fmaf(dst[2], inv, c2);
    dst[3] = __builtin_fmaf(dst[3], inv, a);
}

// to_bytes (painter/mod.rs:108-123), packed little-endian RGBA.
__host__ __device__ __forceinline__ uint32_t solid_bytes(const float c[4]) {
    float r = 1.0f / c[3];
    uint32_t b0 = f2u8_sat(__builtin_fmaf(srgb_approx(c[0] * r), 255.0f, 0.5f));
    uint32_t b1 = f2u8_sat(__builtin_fmaf(srgb_approx(c[1] * r), 255.0f, 0.5f));
    uint32_t b2 = f2u8_sat(__builtin_fmaf(srgb_approx(c[2] * r), 255.0f, 0.5f));
    uint32_t b3 = f2u8_sat(__builtin_fmaf(c[3], 255.0f, 0.5f));
    return b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
}

// compute_srgb for one pixel (painter/mod.rs:418-453, simd/auto.rs:73-86)
__host__ __device__ __forceinline__ uint32_t pixel_bytes(float c0, float c1, float c2, float a) {
    uint32_t b0 = f2u8_sat(round_away((srgb_approx(c0) * a) * 255.0f));
    uint32_t b1 = f2u8_sat(round_away((srgb_approx(c1) * a) * 255.0f));
    uint32_t b2 = f2u8_sat(round_away((srgb_approx(c2) * a) * 255.0f));
    uint32_t b3 = f2u8_sat(round_away(__builtin_fmaf(a, 255.0f, 0.5f)));
    return b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
}

}  // namespace mg
