/*
 * CPU oracle for the mold/surpass rasterize -> sort -> paint path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). A scalar, single-threaded
 * restatement of the Rust reference; each function cites the file:line it
 * follows. Built with -ffp-contract=off -fwrapv -fno-fast-math.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Rust numeric semantics                                                    */
/* ------------------------------------------------------------------------ */

/* `f32 as i32`: truncate, saturate, NaN -> 0. */
static inline int32_t rs_f2i(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}

/* `f32 as u8`: truncate, saturate, NaN -> 0. */
static inline uint8_t rs_f2u8(float v) {
    if (v != v) return 0;
    if (v <= 0.0f) return 0;
    if (v >= 255.0f) return 255;
    return (uint8_t)v;
}

/* `f32 as usize` (64-bit): truncate, saturate, NaN -> 0. */
static inline uint64_t rs_f2usize(float v) {
    if (v != v) return 0;
    if (v <= 0.0f) return 0;
    if (v >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)v;
}

/* `f32::round`: half away from zero. x - trunc(x) is exact in binary32. */
static inline float rs_round(float x) {
    float t = truncf(x);
    if (fabsf(x - t) >= 0.5f) t += copysignf(1.0f, x);
    return t;
}

/* `f32::min` / `f32::max`: a NaN operand yields the other one. On equal
   operands (including +0/-0) the first operand is returned. */
static inline float rs_min(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return b < a ? b : a;
}
static inline float rs_max(float a, float b) {
    if (a != a) return b;
    if (b != b) return a;
    return b > a ? b : a;
}

/* `f32::clamp` (NaN propagates). */
static inline float rs_clamp(float x, float lo, float hi) {
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}

/* i32::abs in a release build (wraps on i32::MIN). */
static inline int32_t rs_abs_i32(int32_t x) { return x < 0 ? (int32_t)(0u - (uint32_t)x) : x; }

/* i8::abs wrapping (simd/auto.rs:232-235 in a release build). */
static inline int8_t rs_abs_i8(int8_t x) { return x == -128 ? (int8_t)-128 : (int8_t)(x < 0 ? -x : x); }

static inline int8_t wrap_i8(int32_t x) { return (int8_t)(uint8_t)(uint32_t)x; }
static inline int16_t wrap_i16(int32_t x) { return (int16_t)(uint16_t)(uint32_t)x; }

/* to_sub_pixel (surpass/src/segment.rs:13-16). */
static inline int32_t to_sub_pixel(float v) {
    return rs_f2i(fmaf(v, (float)OR_PIXEL_WIDTH, copysignf(0.5f, v)));
}

/* ------------------------------------------------------------------------ */
/* CompactSegment bitfield (raster_segment.rs:13-160)                         */
/* ------------------------------------------------------------------------ */

uint64_t or_seg_new(int32_t is_none, int32_t tile_j, int32_t tile_i, uint32_t layer,
                    uint32_t tile_y, uint32_t tile_x, int32_t area, int32_t cover) {
    /* `bitfield!(@shift_or)`: each field is `as u64` (sign-extending signed
       fields) and masked to its width, MSB first. */
    uint64_t v = (uint64_t)(int64_t)is_none & 0x1;
    v = (v << 12) | ((uint64_t)(int64_t)tile_j & 0xFFF);
    v = (v << 13) | ((uint64_t)(int64_t)tile_i & 0x1FFF);
    v = (v << 16) | ((uint64_t)layer & 0xFFFF);
    v = (v << 3) | ((uint64_t)tile_y & 0x7);
    v = (v << 3) | ((uint64_t)tile_x & 0x7);
    v = (v << 10) | ((uint64_t)(int64_t)area & 0x3FF);
    v = (v << 6) | ((uint64_t)(int64_t)cover & 0x3F);
    return v;
}

static inline int32_t sext(uint64_t v, int bits) {
    uint64_t m = ((uint64_t)1 << bits) - 1;
    int64_t x = (int64_t)(v & m);
    if (x >> (bits - 1)) x -= (int64_t)1 << bits;
    return (int32_t)x;
}

static inline int32_t seg_tile_j(uint64_t s) { return sext(s >> 51, 12); }
static inline uint32_t seg_tile_i_field(uint64_t s) { return (uint32_t)((s >> 38) & 0x1FFF); }
static inline uint32_t seg_layer(uint64_t s) { return (uint32_t)((s >> 22) & 0xFFFF); }
static inline uint32_t seg_ty(uint64_t s) { return (uint32_t)((s >> 19) & 7); }
static inline uint32_t seg_tx(uint64_t s) { return (uint32_t)((s >> 16) & 7); }
static inline int32_t seg_area(uint64_t s) { return sext(s >> 6, 10); }
static inline int32_t seg_cover(uint64_t s) { return sext(s, 6); }
/* Painter tile column: the +256 bias of rasterizer/mod.rs:188 removed, reading
   the 13-bit field as unsigned so the sorted stream stays monotone in it. */
static inline int32_t seg_col(uint64_t s) { return (int32_t)seg_tile_i_field(s) - 256; }

void or_seg_fields(uint64_t s, int32_t out[8]) {
    out[0] = (int32_t)(s >> 63);
    out[1] = seg_tile_j(s);
    out[2] = sext(s >> 38, 13);
    out[3] = (int32_t)seg_layer(s);
    out[4] = (int32_t)seg_ty(s);
    out[5] = (int32_t)seg_tx(s);
    out[6] = seg_area(s);
    out[7] = seg_cover(s);
}

/* ------------------------------------------------------------------------ */
/* Path flattening (mold/src/path/segments.rs, transform.rs, point.rs)       */
/* ------------------------------------------------------------------------ */

typedef struct { float x, y; } pt;

typedef struct {
    float t[9];
    int is_perspective;
} xform;

/* Transform::new (transform.rs:57-69). */
static void xform_new(xform *x, const float *t9) {
    float r = 1.0f / t9[8];
    for (int i = 0; i < 9; i++) x->t[i] = t9[i] * r;
    x->is_perspective = fabsf(x->t[6]) > 1.1920929e-7f || fabsf(x->t[7]) > 1.1920929e-7f;
}

/* full_transform (transform.rs:9-15). */
static void full_transform(const float p[3], const float *t, float o[3]) {
    o[0] = t[0] * p[0] + t[1] * p[1] + t[2] * p[2];
    o[1] = t[3] * p[0] + t[4] * p[1] + t[5] * p[2];
    o[2] = t[6] * p[0] + t[7] * p[1] + t[8] * p[2];
}

/* Transform::transform (transform.rs:71-93). Returns the transformed command. */
static or_cmd xform_cmd(const xform *x, or_cmd c) {
    or_cmd o = c;
    int npts = 0, rational = 0;
    switch (c.kind) {
    case OR_CMD_LINE: npts = 2; break;
    case OR_CMD_QUAD: npts = 3; break;
    case OR_CMD_RAT_QUAD: npts = 3; rational = 1; break;
    case OR_CMD_CUBIC: npts = 4; break;
    case OR_CMD_RAT_CUBIC: npts = 4; rational = 1; break;
    default: return o;
    }
    const float *t = x->t;
    if (!rational && !x->is_perspective) {
        /* simple_transform (transform.rs:17-22) */
        for (int i = 0; i < npts; i++) {
            float px = c.v[2 * i], py = c.v[2 * i + 1];
            o.v[2 * i] = t[0] * px + t[1] * py + t[2];
            o.v[2 * i + 1] = t[3] * px + t[4] * py + t[5];
        }
        return o;
    }
    if (c.kind == OR_CMD_LINE) {
        /* perspective line: full transform + from_weighted (point.rs:51-53) */
        for (int i = 0; i < 2; i++) {
            float p[3] = {c.v[2 * i], c.v[2 * i + 1], 1.0f}, q[3];
            full_transform(p, t, q);
            float r = 1.0f / q[2];
            o.v[2 * i] = q[0] * r;
            o.v[2 * i + 1] = q[1] * r;
        }
        return o;
    }
    /* full!: quad/cubic with weight 1, or rational with their weights */
    o.kind = (npts == 3) ? OR_CMD_RAT_QUAD : OR_CMD_RAT_CUBIC;
    for (int i = 0; i < npts; i++) {
        float p[3], q[3];
        if (rational) {
            p[0] = c.v[3 * i]; p[1] = c.v[3 * i + 1]; p[2] = c.v[3 * i + 2];
        } else {
            p[0] = c.v[2 * i]; p[1] = c.v[2 * i + 1]; p[2] = 1.0f;
        }
        full_transform(p, t, q);
        o.v[3 * i] = q[0]; o.v[3 * i + 1] = q[1]; o.v[3 * i + 2] = q[2];
    }
    return o;
}

typedef struct {
    const or_cmd *cmds;
    size_t n, pos;
    int has_xform;
    xform xf;
    int has_state;
    size_t subdivisions; /* usize */
    float t, increment;
    pt point;
    int has_last, has_closing;
    pt last, closing;
} path_iter;

static inline float lerpf(float t, float a, float b) { return a * (1.0f - t) + b * t; }
static inline pt lerp_pt(float t, pt a, pt b) {
    pt r = {a.x * (1.0f - t) + b.x * t, a.y * (1.0f - t) + b.y * t};
    return r;
}

/* PathSegments::line (segments.rs:106-114). */
static void it_line(path_iter *it, pt p0, pt p1, float *seg) {
    if (!it->has_closing) { it->has_closing = 1; it->closing = p0; }
    it->has_last = 1;
    it->last = p1;
    seg[0] = p0.x; seg[1] = p0.y; seg[2] = p1.x; seg[3] = p1.y;
}

static pt curve_point(const or_cmd *c, float t) {
    const float *v = c->v;
    switch (c->kind) {
    case OR_CMD_QUAD: {
        pt p0 = {v[0], v[1]}, p1 = {v[2], v[3]}, p2 = {v[4], v[5]};
        return lerp_pt(t, lerp_pt(t, p0, p1), lerp_pt(t, p1, p2));
    }
    case OR_CMD_CUBIC: {
        pt p0 = {v[0], v[1]}, p1 = {v[2], v[3]}, p2 = {v[4], v[5]}, p3 = {v[6], v[7]};
        pt a = lerp_pt(t, lerp_pt(t, p0, p1), lerp_pt(t, p1, p2));
        pt b = lerp_pt(t, lerp_pt(t, p1, p2), lerp_pt(t, p2, p3));
        return lerp_pt(t, a, b);
    }
    case OR_CMD_RAT_QUAD: {
        pt p0 = {v[0], v[1]}, p1 = {v[3], v[4]}, p2 = {v[6], v[7]};
        pt b = lerp_pt(t, lerp_pt(t, p0, p1), lerp_pt(t, p1, p2));
        float w = lerpf(t, lerpf(t, v[2], v[5]), lerpf(t, v[5], v[8]));
        float r = 1.0f / w;
        pt o = {b.x * r, b.y * r};
        return o;
    }
    default: { /* RAT_CUBIC */
        pt p0 = {v[0], v[1]}, p1 = {v[3], v[4]}, p2 = {v[6], v[7]}, p3 = {v[9], v[10]};
        pt a = lerp_pt(t, lerp_pt(t, p0, p1), lerp_pt(t, p1, p2));
        pt bb = lerp_pt(t, lerp_pt(t, p1, p2), lerp_pt(t, p2, p3));
        pt b = lerp_pt(t, a, bb);
        float wa = lerpf(t, lerpf(t, v[2], v[5]), lerpf(t, v[5], v[8]));
        float wb = lerpf(t, lerpf(t, v[5], v[8]), lerpf(t, v[8], v[11]));
        float w = lerpf(t, wa, wb);
        float r = 1.0f / w;
        pt o = {b.x * r, b.y * r};
        return o;
    }
    }
}

/* First/last points and deviations for next_segment
   (segments.rs:86-95,159-206). */
static void curve_params(const or_cmd *c, pt *first, pt *last, float *dx, float *dy) {
    const float *v = c->v;
    pt p[4];
    int n = (c->kind == OR_CMD_QUAD || c->kind == OR_CMD_RAT_QUAD) ? 3 : 4;
    if (c->kind == OR_CMD_RAT_QUAD || c->kind == OR_CMD_RAT_CUBIC) {
        for (int i = 0; i < n; i++) {
            /* Point::from_weighted (point.rs:51-53) */
            float r = 1.0f / v[3 * i + 2];
            p[i].x = v[3 * i] * r;
            p[i].y = v[3 * i + 1] * r;
        }
    } else {
        for (int i = 0; i < n; i++) { p[i].x = v[2 * i]; p[i].y = v[2 * i + 1]; }
    }
    *first = p[0];
    *last = p[n - 1];
    if (n == 3) {
        *dx = p[0].x + p[2].x - 2.0f * p[1].x;
        *dy = p[0].y + p[2].y - 2.0f * p[1].y;
    } else {
        *dx = rs_max(fabsf(p[0].x + p[2].x - 2.0f * p[1].x), fabsf(p[1].x + p[3].x - 2.0f * p[2].x));
        *dy = rs_max(fabsf(p[0].y + p[2].y - 2.0f * p[1].y), fabsf(p[1].y + p[3].y - 2.0f * p[2].y));
    }
}

/* next_segment + subdivisions (segments.rs:51-72,116-157). Returns 1 and
   writes seg. */
static void it_curve(path_iter *it, const or_cmd *c, float *seg) {
    pt first, last;
    float dvx, dvy;
    curve_params(c, &first, &last, &dvx, &dvy);
    if (!it->has_state) {
        float dev2 = dvx * dvx + dvy * dvy;
        const float accuracy = 0.5f / (float)OR_PIXEL_WIDTH;
        if (dev2 < accuracy) { it_line(it, first, last, seg); return; }
        uint64_t count = rs_f2usize(ceilf(sqrtf(sqrtf(dev2) * 2.0f / (8.0f * accuracy))));
        uint64_t sub = (count > 1 ? count : 1) - 1;
        if (sub == 0) { it_line(it, first, last, seg); return; }
        it->has_state = 1;
        it->subdivisions = (size_t)sub;
        it->increment = 1.0f / (float)sub;
        it->t = 0.0f;
        it->point = first;
    }
    if (it->subdivisions == 1) {
        it->has_state = 0;
        it_line(it, it->point, last, seg);
        return;
    }
    it->t += it->increment;
    pt np = curve_point(c, it->t);
    it_line(it, it->point, np, seg);
    it->point = np;
    it->subdivisions -= 1;
}

/* Iterator::next for PathSegments (segments.rs:233-263). Returns 0 at end. */
static int it_next(path_iter *it, float *seg) {
    for (;;) {
        if (it->pos >= it->n) return 0;
        or_cmd c = it->cmds[it->pos];
        if (it->has_xform) c = xform_cmd(&it->xf, c);
        if (c.kind == OR_CMD_CLOSE) {
            it->pos++;
            /* close (segments.rs:208-223) */
            if (it->has_closing && it->has_last) {
                int emitted = 0;
                pt cp = it->closing, lp = it->last;
                int approx = fabsf(cp.x - lp.x) < 1.1920929e-7f && fabsf(cp.y - lp.y) < 1.1920929e-7f;
                if (!approx) { it_line(it, lp, cp, seg); emitted = 1; }
                it->has_closing = 0;
                if (emitted) return 1;
            }
            continue;
        }
        if (c.kind == OR_CMD_LINE) {
            pt p0 = {c.v[0], c.v[1]}, p1 = {c.v[2], c.v[3]};
            it_line(it, p0, p1, seg);
        } else {
            it_curve(it, &c, seg);
        }
        if (!it->has_state) it->pos++;
        return 1;
    }
}

size_t or_path_flatten(const or_cmd *cmds, size_t n, const float *t9, float *out, size_t cap) {
    path_iter it;
    memset(&it, 0, sizeof it);
    it.cmds = cmds;
    it.n = n;
    if (t9) { it.has_xform = 1; xform_new(&it.xf, t9); }
    size_t count = 0;
    float seg[4];
    while (it_next(&it, seg)) {
        if (count < cap) memcpy(out + 4 * count, seg, sizeof seg);
        count++;
    }
    return count;
}

/* ------------------------------------------------------------------------ */
/* LinesBuilder::build (surpass/src/segment.rs:91-237)                       */
/* ------------------------------------------------------------------------ */

static void build_one(size_t i, const float *p0x_, const float *p0y_, const float *p1x_,
                      const float *p1y_, const uint16_t *layer_ids, const or_layer *layers,
                      size_t n_layers, const float *default_transform, or_lines *o) {
    uint16_t id = layer_ids[i];
    const or_layer *layer = (layers && id < n_layers && layers[id].exists) ? &layers[id] : NULL;

    /* disabled layer / horizontal line: Default::default() tuple */
    o->orders[i] = 0;
    o->starts[i] = 0;
    o->starts_f32[i] = 0.0f;
    o->ends_f32[i] = 0.0f;
    o->lengths[i] = 0;
    o->slopes[i] = 0.0f;
    o->slopes_recip[i] = 0.0f;
    o->py_slope_pxs[i] = 0.0f;
    o->px_slope_recip_pys[i] = 0.0f;
    o->octants[i] = 0;

    if (layer && !layer->is_enabled) return;

    uint16_t order = (layer && layer->has_order) ? layer->order : id;

    float p0x = p0x_[i], p0y = p0y_[i], p1x = p1x_[i], p1y = p1y_[i];
    const float *t = (layer && layer->has_affine) ? layer->affine : default_transform;
    if (t) {
        /* transform_point (segment.rs:123-128) */
        float a0 = fmaf(t[0], p0x, fmaf(t[1], p0y, t[4]));
        float b0 = fmaf(t[2], p0x, fmaf(t[3], p0y, t[5]));
        float a1 = fmaf(t[0], p1x, fmaf(t[1], p1y, t[4]));
        float b1 = fmaf(t[2], p1x, fmaf(t[3], p1y, t[5]));
        p0x = a0; p0y = b0; p1x = a1; p1y = b1;
    }

    int32_t p0x_sub = to_sub_pixel(p0x);
    int32_t p0y_sub = to_sub_pixel(p0y);
    int32_t p1x_sub = to_sub_pixel(p1x);
    int32_t p1y_sub = to_sub_pixel(p1y);

    if (p0y_sub == p1y_sub) return;

    int32_t dx = (int32_t)((uint32_t)p1x_sub - (uint32_t)p0x_sub);
    int32_t dy = (int32_t)((uint32_t)p1y_sub - (uint32_t)p0y_sub);

    int dx_dy = rs_abs_i32(dx) > rs_abs_i32(dy);
    int32_t sum = (int32_t)((uint32_t)dx + (uint32_t)dy);
    int32_t ndy = (int32_t)(0u - (uint32_t)dy);
    int sum_0 = sum > 0 || (dx == ndy && dx < dy);

    int32_t start_sub, end_sub;
    if (sum_0 && dx_dy) { start_sub = p0x_sub; end_sub = p1x_sub; }
    else if (sum_0) { start_sub = p0y_sub; end_sub = p1y_sub; }
    else if (dx_dy) { start_sub = p1x_sub; end_sub = p0x_sub; }
    else { start_sub = p1y_sub; end_sub = p0y_sub; }

    int32_t start = start_sub >> 4;
    int32_t end = (end_sub & 15) == 0 ? (end_sub >> 4) : (int32_t)((uint32_t)(end_sub >> 4) + 1u);
    int32_t len32 = (int32_t)((uint32_t)end - (uint32_t)start);
    /* `as usize` of a negative i32 would be a huge length; keep it signed */
    int64_t length = (int64_t)len32;

    float start_f32 = (float)start_sub * (1.0f / (float)OR_PIXEL_WIDTH);
    float end_f32 = (float)end_sub * (1.0f / (float)OR_PIXEL_WIDTH);
    float fp0x = (float)p0x_sub * (1.0f / (float)OR_PIXEL_WIDTH);
    float fp0y = (float)p0y_sub * (1.0f / (float)OR_PIXEL_WIDTH);
    float px = dx_dy ? fp0x : fp0y;
    float py = dx_dy ? fp0y : fp0x;
    float slope = dx_dy ? (float)dy / (float)dx : (float)dx / (float)dy;
    float slope_recip = 1.0f / slope;
    float py_slope_px = fmaf(slope, -px, py);
    float px_slope_recip_py = fmaf(slope_recip, -py, px);

    uint8_t octant;
    int gx = dx > 0, gy = dy > 0;
    if (dx_dy && gx && gy) octant = 0;
    else if (!dx_dy && gx && gy) octant = 1;
    else if (!dx_dy && !gx && gy) octant = 2;
    else if (dx_dy && !gx && gy) octant = 3;
    else if (dx_dy && !gx && !gy) octant = 4;
    else if (!dx_dy && !gx && !gy) octant = 5;
    else if (!dx_dy && gx && !gy) octant = 6;
    else octant = 7;

    o->orders[i] = order;
    o->starts[i] = start;
    o->starts_f32[i] = start_f32;
    o->ends_f32[i] = end_f32;
    o->lengths[i] = length;
    o->slopes[i] = slope;
    o->slopes_recip[i] = slope_recip;
    o->py_slope_pxs[i] = py_slope_px;
    o->px_slope_recip_pys[i] = px_slope_recip_py;
    o->octants[i] = octant;
}

void or_lines_build(size_t n, const float *p0x, const float *p0y, const float *p1x,
                    const float *p1y, const uint16_t *layer_ids, const or_layer *layers,
                    size_t n_layers, const float *default_transform, or_lines *out) {
    for (size_t i = 0; i < n; i++)
        build_one(i, p0x, p0y, p1x, p1y, layer_ids, layers, n_layers, default_transform, out);
}

/* ------------------------------------------------------------------------ */
/* Rasterizer (surpass/src/rasterizer/mod.rs:127-285)                        */
/* ------------------------------------------------------------------------ */

/* intersects_y_border (mod.rs:127-132) */
static inline int intersects_y_border(int32_t y0, int32_t y1) {
    return (y0 & 15) != 0 && (y1 & 15) != 0 && (y0 >> 4) != (y1 >> 4);
}

/* segment() = tiles() + area_cover() + CompactSegment::new (mod.rs:134-189) */
static uint64_t make_segment(uint16_t layer, int32_t x0, int32_t x1, int32_t y0, int32_t y1,
                             int32_t border_x, int32_t border_y, uint8_t octant) {
    /* tiles (mod.rs:134-147) */
    int32_t bx = border_x, by = border_y;
    if (octant == 1 || octant == 2 || octant == 5 || octant == 6) { bx = border_y; by = border_x; }
    int16_t tile_i = (int16_t)(uint16_t)(uint32_t)(bx >> 3);
    int16_t tile_j = (int16_t)(uint16_t)(uint32_t)(by >> 3);
    uint32_t tile_x = (uint32_t)(bx & 7);
    uint32_t tile_y = (uint32_t)(by & 7);

    /* area_cover (mod.rs:149-172) */
    int32_t ax0, ax1, ay0, ay1;
    switch (octant) {
    case 0: ax0 = x0; ax1 = x1; ay0 = y0; ay1 = y1; break;
    case 1: ax0 = y0; ax1 = y1; ay0 = x0; ay1 = x1; break;
    case 2: ax0 = y1; ax1 = y0; ay0 = x0; ay1 = x1; break;
    case 3: ax0 = x0; ax1 = x1; ay0 = y1; ay1 = y0; break;
    case 4: ax0 = x0; ax1 = x1; ay0 = y1; ay1 = y0; break;
    case 5: ax0 = y0; ax1 = y1; ay0 = x1; ay1 = x0; break;
    case 6: ax0 = y1; ax1 = y0; ay0 = x1; ay1 = x0; break;
    default: ax0 = x0; ax1 = x1; ay0 = y0; ay1 = y1; break;
    }
    int32_t border = (int32_t)((uint32_t)(ax0 & ~15) + 16u);
    int32_t height = (int32_t)((uint32_t)ay1 - (uint32_t)ay0);
    int32_t dxw = (int32_t)((uint32_t)ax1 - (uint32_t)ax0);
    int32_t prod = (int32_t)((uint32_t)height * (uint32_t)dxw);
    int16_t triangle = wrap_i16(prod / 2);
    int32_t bw = (int32_t)((uint32_t)border - (uint32_t)ax1);
    int16_t rectangle = wrap_i16((int32_t)((uint32_t)height * (uint32_t)bw));
    int16_t area = wrap_i16((int32_t)triangle + (int32_t)rectangle);
    int8_t cover = wrap_i8(height);

    int16_t ti = tile_i < -1 ? (int16_t)-1 : tile_i;
    ti = wrap_i16((int32_t)ti + 256);
    return or_seg_new(0, tile_j, ti, layer, tile_y, tile_x, area, cover);
}

static void rasterize_pixel(const or_lines *l, size_t li, int32_t pi, uint64_t out[2]) {
    uint16_t layer = l->orders[li];
    float py_slope_px = l->py_slope_pxs[li];
    float px_slope_recip_py = l->px_slope_recip_pys[li];
    uint8_t octant = l->octants[li];

    int32_t x0i = (int32_t)((uint32_t)l->starts[li] + (uint32_t)pi);
    int32_t x1i = (int32_t)((uint32_t)x0i + 1u);
    int32_t border_x = x0i;

    float x0 = rs_max((float)x0i, l->starts_f32[li]);
    float x1 = rs_min((float)x1i, l->ends_f32[li]);
    float y0 = fmaf(l->slopes[li], x0, py_slope_px);
    float y1 = fmaf(l->slopes[li], x1, py_slope_px);

    int32_t x0s = to_sub_pixel(x0), x1s = to_sub_pixel(x1);
    int32_t y0s = to_sub_pixel(y0), y1s = to_sub_pixel(y1);

    if (intersects_y_border(y0s, y1s)) {
        int32_t y = (y0s > y1s ? y0s : y1s) >> 4;
        int32_t ys = (int32_t)((uint32_t)y << 4);
        float x = fmaf(l->slopes_recip[li], (float)y, px_slope_recip_py);
        int32_t xs = to_sub_pixel(x);
        int32_t by0 = (ys < y0s ? ys : y0s) >> 4;
        out[0] = make_segment(layer, x0s, xs, y0s, ys, border_x, by0, octant);
        int32_t by1 = (ys < y1s ? ys : y1s) >> 4;
        out[1] = make_segment(layer, xs, x1s, ys, y1s, border_x, by1, octant);
    } else {
        int32_t by = (y0s < y1s ? y0s : y1s) >> 4;
        out[0] = make_segment(layer, x0s, x1s, y0s, y1s, border_x, by, octant);
        out[1] = OR_NONE;
    }
}

size_t or_rasterize(size_t n_lines, const or_lines *lines, uint64_t *out, size_t cap) {
    size_t total = 0;
    for (size_t li = 0; li < n_lines; li++) {
        int64_t len = lines->lengths[li];
        if (len > 0) total += (size_t)len;
    }
    if (2 * total > cap) return (size_t)-1;
    size_t k = 0;
    for (size_t li = 0; li < n_lines; li++) {
        int64_t len = lines->lengths[li];
        for (int64_t pi = 0; pi < len; pi++) {
            rasterize_pixel(lines, li, (int32_t)pi, out + k);
            k += 2;
        }
    }
    return k;
}

/* ------------------------------------------------------------------------ */
/* Canonical sort: stable LSD radix on the 42-bit key                         */
/* ------------------------------------------------------------------------ */

void or_sort_canonical(uint64_t *segs, size_t n) {
    if (n < 2) return;
    uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *src = segs, *dst = tmp;
    /* key bits [22, 64): 42 bits in 6 passes of 7 bits */
    for (int pass = 0; pass < 6; pass++) {
        int shift = 22 + 7 * pass;
        size_t count[129];
        memset(count, 0, sizeof count);
        for (size_t i = 0; i < n; i++) count[((src[i] >> shift) & 127) + 1]++;
        for (int d = 0; d < 128; d++) count[d + 1] += count[d];
        for (size_t i = 0; i < n; i++) dst[count[(src[i] >> shift) & 127]++] = src[i];
        uint64_t *t = src; src = dst; dst = t;
    }
    /* 6 passes: result is back in segs */
    free(tmp);
}

/* ------------------------------------------------------------------------ */
/* Styles and blending (surpass/src/painter/style.rs)                        */
/* ------------------------------------------------------------------------ */

enum {
    BM_OVER, BM_MULTIPLY, BM_SCREEN, BM_OVERLAY, BM_DARKEN, BM_LIGHTEN, BM_COLOR_DODGE,
    BM_COLOR_BURN, BM_HARD_LIGHT, BM_SOFT_LIGHT, BM_DIFFERENCE, BM_EXCLUSION
};

/* BlendMode::blend_fn (style.rs:206-267) */
float or_blend_fn(uint32_t mode, float dst, float src) {
    switch (mode) {
    case BM_OVER: return src;
    case BM_MULTIPLY: return dst * src;
    case BM_SCREEN: return dst + src - (dst * src);
    case BM_OVERLAY: {
        /* hard_light(src, dst) */
        float a = src, b = dst;
        if (a <= 0.5f) return (2.0f * a) * b;
        float a2 = 2.0f * a - 1.0f;
        return a2 + b - (a2 * b);
    }
    case BM_DARKEN: return rs_min(dst, src);
    case BM_LIGHTEN: return rs_max(dst, src);
    case BM_COLOR_DODGE:
        if (src == 0.0f) return 0.0f;
        if (dst == 1.0f) return 1.0f;
        return rs_min(1.0f, src / (1.0f - dst));
    case BM_COLOR_BURN:
        if (src == 1.0f) return 1.0f;
        if (dst == 0.0f) return 0.0f;
        return 1.0f - rs_min(1.0f, (1.0f - src) / dst);
    case BM_HARD_LIGHT: {
        if (dst <= 0.5f) return (2.0f * dst) * src;
        float a2 = 2.0f * dst - 1.0f;
        return a2 + src - (a2 * src);
    }
    case BM_SOFT_LIGHT: {
        float d = src <= 0.25f ? ((16.0f * src - 12.0f) * src + 4.0f) * src : sqrtf(src);
        if (dst <= 0.5f) return src - (1.0f - 2.0f * dst) * src * (1.0f - src);
        return src + (2.0f * dst - 1.0f) * (d - src);
    }
    case BM_DIFFERENCE: return fabsf(dst - src);
    case BM_EXCLUSION: return dst + src - 2.0f * dst * src;
    default: return src;
    }
}

/* BlendMode::blend (style.rs:269-285) */
void or_blend(uint32_t mode, const float dst[4], const float src[4], float out[4]) {
    float alpha = src[3];
    float inv = 1.0f - alpha;
    float c0 = or_blend_fn(mode, dst[0], src[0]) * alpha;
    float c1 = or_blend_fn(mode, dst[1], src[1]) * alpha;
    float c2 = or_blend_fn(mode, dst[2], src[2]) * alpha;
    out[0] = fmaf(dst[0], inv, c0);
    out[1] = fmaf(dst[1], inv, c1);
    out[2] = fmaf(dst[2], inv, c2);
    out[3] = fmaf(dst[3], inv, alpha);
}

/* blend_function! for one lane (style.rs:288-418); `select(a, b, m)` is
   m ? a : b (simd/auto.rs:459-502). */
static float blend_simd1(uint32_t mode, float d, float s) {
    switch (mode) {
    case BM_OVER: return s;
    case BM_MULTIPLY: return d * s;
    case BM_SCREEN: return fmaf(d, -s, d) + s;
    case BM_OVERLAY:
        return s <= 0.5f ? d * s * 2.0f : 2.0f * (s + d - fmaf(s, d, 0.5f));
    case BM_DARKEN: return rs_min(d, s);
    case BM_LIGHTEN: return rs_max(d, s);
    case BM_COLOR_DODGE: {
        float v = rs_min(1.0f, s / (1.0f - d));
        return s == 0.0f ? 0.0f : v;
    }
    case BM_COLOR_BURN: {
        float v = 1.0f - rs_min(1.0f, (1.0f - s) / d);
        return s == 1.0f ? 1.0f : v;
    }
    case BM_HARD_LIGHT:
        return d <= 0.5f ? d * s * 2.0f : 2.0f * (s + d - fmaf(s, d, 0.5f));
    case BM_SOFT_LIGHT: {
        float poly = fmaf(fmaf(16.0f, s, -12.0f), s, 4.0f) * s;
        float dd = s <= 0.25f ? poly : sqrtf(s);
        float k = fmaf(2.0f, d, -1.0f);
        float lo = fmaf(s * (1.0f - s), k, s);
        float hi = fmaf(dd - s, k, s);
        return d <= 0.5f ? lo : hi;
    }
    case BM_DIFFERENCE: return fabsf(d - s);
    case BM_EXCLUSION: return fmaf(-2.0f * d, s, d) + s;
    default: return s;
    }
}

void or_blend_simd(uint32_t mode, const float dst[3], const float src[3], float out[3]) {
    for (int i = 0; i < 3; i++) out[i] = blend_simd1(mode, dst[i], src[i]);
}

/* GradientBuilder::build (style.rs:66-84) */
void or_gradient_resolve_stops(float *stops, uint32_t n) {
    if (n < 2) return;
    float inc = 1.0f / (float)(n - 1);
    for (uint32_t i = 0; i < n; i++)
        if (stops[5 * i + 4] == -0.0f) stops[5 * i + 4] = (float)i * inc;
}

static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* Gradient::get_t + color_at (style.rs:96-175) */
void or_gradient_color_at(const or_props *p, float x, float y, float out[32]) {
    float dx = p->grad_end[0] - p->grad_start[0];
    float dy = p->grad_end[1] - p->grad_start[1];
    float dot = dx * dx + dy * dy;
    float dot_recip = 1.0f / dot;
    float t[8];
    if (p->grad_type == OR_GRAD_LINEAR) {
        float tx = (x - p->grad_start[0]) * dx * dot_recip;
        float ty = y - p->grad_start[1];
        for (int r = 0; r < 8; r++) t[r] = fmaf(((float)r + ty) * dy, dot_recip, tx);
    } else {
        float px = x - p->grad_start[0];
        float px2 = px * px;
        for (int r = 0; r < 8; r++) {
            float py = (float)r + (y - p->grad_start[1]);
            t[r] = sqrtf(fmaf(py, py, px2) * dot_recip);
        }
    }
    uint32_t ch[4][8];
    memset(ch, 0, sizeof ch);
    const float *st = p->stops;
    uint32_t n = p->n_stops;
    int acc[8];
    int any = 0;
    for (int r = 0; r < 8; r++) { acc[r] = t[r] <= st[4]; any |= acc[r]; }
    if (any)
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 8; r++) ch[c][r] |= acc[r] ? fbits(st[c]) : fbits(0.0f);
    float start_stop = 0.0f;
    const float *start_color = st;
    for (uint32_t k = 1; k < n; k++) {
        const float *color = st + 5 * k;
        float end_stop = color[4];
        int m[8];
        any = 0;
        for (int r = 0; r < 8; r++) { m[r] = acc[r] ^ (t[r] < end_stop); any |= m[r]; }
        if (any) {
            float d = end_stop - start_stop;
            float drec = 1.0f / d;
            for (int c = 0; c < 4; c++) {
                for (int r = 0; r < 8; r++) {
                    float lt = (t[r] - start_stop) * drec;
                    float v = fmaf(lt, color[c], fmaf(-lt, start_color[c], start_color[c]));
                    ch[c][r] |= m[r] ? fbits(v) : fbits(0.0f);
                }
            }
            for (int r = 0; r < 8; r++) acc[r] |= m[r];
        }
        start_stop = end_stop;
        start_color = color;
    }
    any = 0;
    for (int r = 0; r < 8; r++) any |= !acc[r];
    if (any) {
        const float *last = st + 5 * (n - 1);
        for (int c = 0; c < 4; c++)
            for (int r = 0; r < 8; r++) ch[c][r] |= !acc[r] ? fbits(last[c]) : fbits(0.0f);
    }
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 8; r++) out[c * 8 + r] = bitsf(ch[c][r]);
}

/* linear_to_srgb_approx (painter/mod.rs:91-106) */
float or_linear_to_srgb(float l) {
    const float a = 0.20101772f, b = -0.51280147f, c = 1.344401f, d = -0.030656587f;
    float s = sqrtf(l);
    float s2 = l;
    float s3 = s2 * s;
    if (l <= 0.0031308f) return l * 12.92f;
    return fmaf(a, s3, fmaf(b, s2, fmaf(c, s, d)));
}

/* linear_to_srgb_approx_simd (painter/mod.rs:74-88) */
float or_linear_to_srgb_simd(float l) {
    const float a = 0.20101772f, b = -0.51280147f, c = 1.344401f, d = -0.030656587f;
    float s = sqrtf(l);
    float s2 = l;
    float s3 = s2 * s;
    float m = l * 12.92f;
    float n = fmaf(a, s3, fmaf(b, s2, fmaf(c, s, d)));
    return l <= 0.0031308f ? m : n;
}

/* to_byte / to_bytes (painter/mod.rs:108-123) */
static inline uint8_t to_byte(float n) { return rs_f2u8(fmaf(n, 255.0f, 0.5f)); }
void or_to_bytes(const float color[4], uint8_t out[4]) {
    float ar = 1.0f / color[3];
    out[0] = to_byte(or_linear_to_srgb(color[0] * ar));
    out[1] = to_byte(or_linear_to_srgb(color[1] * ar));
    out[2] = to_byte(or_linear_to_srgb(color[2] * ar));
    out[3] = to_byte(color[3]);
}

/* from_area (painter/mod.rs:53-71) */
static inline float from_area(int32_t a, uint32_t rule) {
    if (rule == OR_NONZERO) {
        return rs_clamp(fabsf((float)a * (1.0f / 256.0f)), 0.0f, 1.0f);
    }
    int32_t number = a >> 8;
    float capped = (float)(a & 255) * (1.0f / 256.0f);
    return (number & 1) == 0 ? capped : 1.0f - capped;
}

/* ------------------------------------------------------------------------ */
/* Cover (painter/mod.rs:148-199), 8 lanes = 8 tile rows                      */
/* ------------------------------------------------------------------------ */

static int cover_is_empty(const int8_t c[8], uint32_t rule) {
    for (int y = 0; y < 8; y++) {
        if (rule == OR_NONZERO) { if (c[y] != 0) return 0; }
        else if ((rs_abs_i8(c[y]) & 31) != 0) return 0;
    }
    return 1;
}

static int cover_is_full(const int8_t c[8], uint32_t rule) {
    for (int y = 0; y < 8; y++) {
        if (rule == OR_NONZERO) { if (rs_abs_i8(c[y]) != OR_PIXEL_WIDTH) return 0; }
        else if ((rs_abs_i8(c[y]) & 31) != 16) return 0;
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Painter + LayerWorkbench                                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    int16_t areas[64];  /* x*8+y */
    int8_t covers[72];  /* x*8+y, x in 0..=8 */
    int has_clip;
    float clip_mask[64];
    uint16_t clip_last;
    float c0[64], c1[64], c2[64], alpha[64];
} painter;

typedef struct {
    uint16_t layer;
    int32_t seg_start, seg_end; /* [start, end) in the tile's segment slice, -1 if none */
    int32_t q;                  /* index in queue, -1 if none */
    uint8_t mask;
    uint8_t skip_clipping;
} wb_id;

typedef struct {
    wb_id *ids;
    size_t n_ids, cap_ids;
    size_t skipped;
    or_carry *queue, *next;
    size_t n_queue, n_next, cap_queue;
} workbench;

static const or_props *get_props(const or_props *props, size_t n_props, uint16_t layer) {
    if (layer >= n_props || !props[layer].defined) return NULL;
    return &props[layer];
}

static void wb_reserve_ids(workbench *wb, size_t n) {
    if (n > wb->cap_ids) {
        size_t c = wb->cap_ids ? wb->cap_ids : 64;
        while (c < n) c *= 2;
        wb->ids = (wb_id *)realloc(wb->ids, c * sizeof(wb_id));
        wb->cap_ids = c;
    }
}

static void wb_reserve_queue(workbench *wb, size_t n) {
    if (n > wb->cap_queue) {
        size_t c = wb->cap_queue ? wb->cap_queue : 64;
        while (c < n) c *= 2;
        wb->queue = (or_carry *)realloc(wb->queue, c * sizeof(or_carry));
        wb->next = (or_carry *)realloc(wb->next, c * sizeof(or_carry));
        wb->cap_queue = c;
    }
}

static void wb_free(workbench *wb) {
    free(wb->ids);
    free(wb->queue);
    free(wb->next);
}

/* populate_layers (layer_workbench.rs:401-428): sorted union of segment
   layers and queue layers. */
static void wb_populate(workbench *wb, const uint64_t *segs, size_t n_segs) {
    size_t n_runs = 0;
    for (size_t i = 0; i < n_segs; i++)
        if (i == 0 || seg_layer(segs[i]) != seg_layer(segs[i - 1])) n_runs++;
    wb_reserve_ids(wb, n_runs + wb->n_queue + 1);
    /* both inputs are sorted by layer: merge */
    size_t s = 0, q = 0, k = 0;
    while (s < n_segs || q < wb->n_queue) {
        uint32_t ls = s < n_segs ? seg_layer(segs[s]) : 0x10000u;
        uint32_t lq = q < wb->n_queue ? wb->queue[q].layer : 0x10000u;
        wb_id id;
        id.mask = 1;
        id.skip_clipping = 0;
        id.seg_start = id.seg_end = -1;
        id.q = -1;
        if (ls <= lq) {
            id.layer = (uint16_t)ls;
            size_t e = s;
            while (e < n_segs && seg_layer(segs[e]) == ls) e++;
            id.seg_start = (int32_t)s;
            id.seg_end = (int32_t)e;
            s = e;
            if (lq == ls) { id.q = (int32_t)q; q++; }
        } else {
            id.layer = (uint16_t)lq;
            id.q = (int32_t)q;
            q++;
        }
        wb->ids[k++] = id;
    }
    wb->n_ids = k;
    wb->skipped = 0;
}

static inline int id_masked(const workbench *wb, size_t i) { return i >= wb->skipped && wb->ids[i].mask; }

/* layer_is_full (layer_workbench.rs:164-175) */
static int wb_layer_is_full(const workbench *wb, const wb_id *id, uint32_t rule) {
    if (id->seg_start >= 0) return 0;
    if (id->q < 0) return 0;
    return cover_is_full(wb->queue[id->q].cover, rule);
}

/* cover_carry (layer_workbench.rs:177-198) */
static int wb_cover_carry(const workbench *wb, const wb_id *id, const uint64_t *segs,
                          uint32_t rule, or_carry *out) {
    int8_t acc[8] = {0};
    if (id->seg_start >= 0)
        for (int32_t i = id->seg_start; i < id->seg_end; i++) {
            uint32_t ty = seg_ty(segs[i]);
            acc[ty] = wrap_i8((int32_t)acc[ty] + seg_cover(segs[i]));
        }
    if (id->q >= 0)
        for (int y = 0; y < 8; y++) acc[y] = wrap_i8((int32_t)acc[y] + wb->queue[id->q].cover[y]);
    if (cover_is_empty(acc, rule)) return 0;
    memcpy(out->cover, acc, 8);
    out->layer = id->layer;
    out->_pad = 0;
    return 1;
}

/* tile_unchanged_pass (layer_workbench.rs:200-222); returns 1 on Break(None) */
static int wb_tile_unchanged(workbench *wb, int32_t *previous_layers, const or_props *props,
                             size_t n_props) {
    if (!previous_layers) return 0;
    int32_t layers = (int32_t)(uint16_t)wb->n_ids;
    int unchanged;
    if (*previous_layers >= 0) {
        int32_t old = *previous_layers;
        *previous_layers = layers;
        unchanged = old == layers;
        for (size_t i = 0; unchanged && i < wb->n_ids; i++) {
            const or_props *p = get_props(props, n_props, wb->ids[i].layer);
            unchanged = p && p->is_unchanged;
        }
    } else {
        *previous_layers = layers;
        unchanged = 0;
    }
    return unchanged;
}

/* skip_trivial_clips_pass (layer_workbench.rs:224-308) */
static void wb_skip_trivial_clips(workbench *wb, const or_props *props, size_t n_props) {
    int has_clip = 0, clip_full = 0, clip_used = 0;
    uint16_t clip_last = 0;
    size_t clip_i = 0;
    for (size_t i = 0; i < wb->n_ids; i++) {
        if (!id_masked(wb, i)) continue;
        wb_id *id = &wb->ids[i];
        const or_props *p = get_props(props, n_props, id->layer);
        if (p->func == OR_FUNC_CLIP) {
            int full = wb_layer_is_full(wb, id, p->fill_rule);
            has_clip = 1;
            clip_full = full;
            clip_last = (uint16_t)(id->layer + (uint16_t)p->clip_layers);
            clip_i = i;
            clip_used = 0;
            if (full) id->mask = 0;
        }
        if (p->func == OR_FUNC_DRAW && p->is_clipped) {
            if (has_clip && id->layer <= clip_last) {
                if (clip_full) id->skip_clipping = 1;
                else clip_used = 1;
            } else {
                id->mask = 0;
            }
        }
        if (has_clip && id->layer > clip_last) {
            has_clip = 0;
            if (!clip_used) wb->ids[clip_i].mask = 0;
        }
    }
    if (has_clip && !clip_used) wb->ids[clip_i].mask = 0;
}

/* skip_fully_covered_layers (layer_workbench.rs:310-388).
   Returns 1 with `solid` set on Break(Solid). */
static int wb_skip_fully_covered(workbench *wb, const or_props *props, size_t n_props,
                                 const float clear[4], float solid[4]) {
    enum { UNSET, OPAQUE, INCOMPLETE } first = UNSET;
    float opaque[4] = {0, 0, 0, 0};
    for (size_t k = wb->n_ids; k-- > 0;) {
        if (!id_masked(wb, k)) continue;
        wb_id *id = &wb->ids[k];
        const or_props *p = get_props(props, n_props, id->layer);
        int is_clipped = p->func == OR_FUNC_DRAW && p->is_clipped && !id->skip_clipping;
        if (is_clipped || !wb_layer_is_full(wb, id, p->fill_rule)) {
            if (first == UNSET) first = INCOMPLETE;
        } else if (p->func == OR_FUNC_DRAW && p->fill_kind == OR_FILL_SOLID && p->blend_mode == BM_OVER) {
            if (p->color[3] == 1.0f) {
                if (first == UNSET) { first = OPAQUE; memcpy(opaque, p->color, sizeof opaque); }
                wb->skipped = k;
                break;
            }
        }
    }
    size_t skip;
    float color[4];
    if (first == OPAQUE) { skip = 1; memcpy(color, opaque, sizeof color); }
    else if (first == UNSET) { skip = 0; memcpy(color, clear, sizeof color); }
    else return 0;
    size_t seen = 0;
    for (size_t k = 0; k < wb->n_ids; k++) {
        if (!id_masked(wb, k)) continue;
        if (seen++ < skip) continue;
        const or_props *p = get_props(props, n_props, wb->ids[k].layer);
        if (p->func == OR_FUNC_DRAW && p->fill_kind == OR_FILL_SOLID) {
            float o[4];
            or_blend(p->blend_mode, color, p->color, o);
            memcpy(color, o, sizeof o);
        } else {
            return 0;
        }
    }
    memcpy(solid, color, sizeof color);
    return 1;
}

/* Painter::blend_at (painter/mod.rs:364-400) for column x */
static void painter_blend_column(painter *pt, int x, const float cov[8], int is_clipped,
                                 const float fill[4][8], uint32_t mode) {
    for (int y = 0; y < 8; y++) {
        int i = x * 8 + y;
        float a = fill[3][y] * cov[y];
        if (is_clipped && pt->has_clip) a *= pt->clip_mask[i];
        float inv = 1.0f - a;
        float d[3] = {pt->c0[i], pt->c1[i], pt->c2[i]};
        float s[3] = {fill[0][y], fill[1][y], fill[2][y]};
        float b[3];
        or_blend_simd(mode, d, s, b);
        b[0] *= a; b[1] *= a; b[2] *= a;
        pt->c0[i] = fmaf(pt->c0[i], inv, b[0]);
        pt->c1[i] = fmaf(pt->c1[i], inv, b[1]);
        pt->c2[i] = fmaf(pt->c2[i], inv, b[2]);
        pt->alpha[i] = fmaf(pt->alpha[i], inv, a);
    }
}

/* Painter::paint_layer (painter/mod.rs:260-317). Returns the carried cover.
   When `dump` is non-NULL the 64 coverage values A are written there. */
static void painter_paint_layer(painter *pt, uint32_t tile_i, uint32_t tile_j, uint16_t layer,
                                const or_props *p, int apply_clip, int8_t out_cover[8],
                                int32_t *dump) {
    int8_t covers[8] = {0};
    if (pt->has_clip && pt->clip_last < layer) pt->has_clip = 0;
    for (int x = 0; x <= 8; x++) {
        if (x != 0) {
            int32_t a[8];
            float cov[8];
            for (int y = 0; y < 8; y++) {
                a[y] = 16 * (int32_t)covers[y] + (int32_t)pt->areas[(x - 1) * 8 + y];
                cov[y] = from_area(a[y], p->fill_rule);
                if (dump) dump[(x - 1) * 8 + y] = a[y];
            }
            if (p->func == OR_FUNC_DRAW) {
                int all_zero = 1;
                for (int y = 0; y < 8; y++) all_zero &= cov[y] == 0.0f;
                int skip = all_zero || (apply_clip && !pt->has_clip);
                if (!skip) {
                    float fill[4][8];
                    if (p->fill_kind == OR_FILL_SOLID) {
                        for (int c = 0; c < 4; c++)
                            for (int y = 0; y < 8; y++) fill[c][y] = p->color[c];
                    } else {
                        float g[32];
                        or_gradient_color_at(p, (float)(x + (int)tile_i * 8), (float)(tile_j * 8), g);
                        for (int c = 0; c < 4; c++)
                            for (int y = 0; y < 8; y++) fill[c][y] = g[c * 8 + y];
                    }
                    painter_blend_column(pt, x - 1, cov, apply_clip, (const float(*)[8])fill, p->blend_mode);
                }
            } else {
                /* clip_at (painter/mod.rs:402-416) */
                if (!pt->has_clip) {
                    pt->has_clip = 1;
                    memset(pt->clip_mask, 0, sizeof pt->clip_mask);
                    pt->clip_last = (uint16_t)(layer + (uint16_t)p->clip_layers);
                }
                for (int y = 0; y < 8; y++) pt->clip_mask[(x - 1) * 8 + y] = cov[y];
            }
        }
        for (int y = 0; y < 8; y++) covers[y] = wrap_i8((int32_t)covers[y] + pt->covers[x * 8 + y]);
    }
    memcpy(out_cover, covers, 8);
}

typedef struct {
    int32_t op;
    float solid[4];
} tile_op;

/* LayerWorkbench::drive_tile_painting (layer_workbench.rs:430-486).
   passes: bit0 unchanged, bit1 clips, bit2 fully covered. */
static tile_op wb_drive(workbench *wb, painter *pt, const uint64_t *segs, size_t n_segs,
                        const or_props *props, size_t n_props, const float clear[4],
                        uint32_t tile_i, uint32_t tile_j, int32_t *previous_layers,
                        uint32_t passes, int stop_after_passes, or_debug *dbg,
                        int *continue_flag) {
    tile_op r;
    r.op = OR_OP_COLOR_BUFFER;
    memset(r.solid, 0, sizeof r.solid);
    wb_populate(wb, segs, n_segs);
    wb_reserve_queue(wb, wb->n_ids + 1);

    /* coverage dump of every id (independent of the workbench decisions) */
    if (dbg && dbg->cov_a) {
        for (size_t k = 0; k < wb->n_ids; k++) {
            wb_id *id = &wb->ids[k];
            int16_t areas[64];
            int8_t covers[72];
            memset(areas, 0, sizeof areas);
            memset(covers, 0, sizeof covers);
            if (id->seg_start >= 0)
                for (int32_t s = id->seg_start; s < id->seg_end; s++) {
                    uint32_t x = seg_tx(segs[s]), y = seg_ty(segs[s]);
                    areas[x * 8 + y] = wrap_i16((int32_t)areas[x * 8 + y] + seg_area(segs[s]));
                    covers[(x + 1) * 8 + y] = wrap_i8((int32_t)covers[(x + 1) * 8 + y] + seg_cover(segs[s]));
                }
            if (id->q >= 0)
                for (int y = 0; y < 8; y++) covers[y] = wrap_i8((int32_t)covers[y] + wb->queue[id->q].cover[y]);
            if (dbg->n_records < dbg->cap_records) {
                int32_t *a = dbg->cov_a + 64 * dbg->n_records;
                int8_t run[8] = {0};
                for (int x = 0; x < 8; x++) {
                    for (int y = 0; y < 8; y++) run[y] = wrap_i8((int32_t)run[y] + covers[x * 8 + y]);
                    for (int y = 0; y < 8; y++) a[x * 8 + y] = 16 * (int32_t)run[y] + (int32_t)areas[x * 8 + y];
                }
                dbg->cov_key[3 * dbg->n_records] = tile_j;
                dbg->cov_key[3 * dbg->n_records + 1] = tile_i;
                dbg->cov_key[3 * dbg->n_records + 2] = id->layer;
            }
            dbg->n_records++;
        }
    }

    int brk = 0;
    if ((passes & 1) && wb_tile_unchanged(wb, previous_layers, props, n_props)) {
        brk = 1;
        r.op = OR_OP_NONE;
    }
    if (!brk && (passes & 2)) wb_skip_trivial_clips(wb, props, n_props);
    if (!brk && (passes & 4)) {
        if (wb_skip_fully_covered(wb, props, n_props, clear, r.solid)) {
            brk = 1;
            r.op = OR_OP_SOLID;
        }
    }
    if (continue_flag) *continue_flag = !brk;
    if (stop_after_passes) return r;

    wb->n_next = 0;
    if (brk) {
        for (size_t k = 0; k < wb->n_ids; k++) {
            wb_id *id = &wb->ids[k];
            const or_props *p = get_props(props, n_props, id->layer);
            or_carry cc;
            if (wb_cover_carry(wb, id, segs, p->fill_rule, &cc)) wb->next[wb->n_next++] = cc;
        }
    } else {
        /* Painter::clear (painter/mod.rs:245-258) */
        for (int i = 0; i < 64; i++) {
            pt->c0[i] = clear[0];
            pt->c1[i] = clear[1];
            pt->c2[i] = clear[2];
            pt->alpha[i] = clear[3];
        }
        for (size_t k = 0; k < wb->n_ids; k++) {
            wb_id *id = &wb->ids[k];
            const or_props *p = get_props(props, n_props, id->layer);
            if (id_masked(wb, k)) {
                memset(pt->areas, 0, sizeof pt->areas);
                memset(pt->covers, 0, sizeof pt->covers);
                if (id->seg_start >= 0)
                    for (int32_t s = id->seg_start; s < id->seg_end; s++) {
                        uint32_t x = seg_tx(segs[s]), y = seg_ty(segs[s]);
                        pt->areas[x * 8 + y] = wrap_i16((int32_t)pt->areas[x * 8 + y] + seg_area(segs[s]));
                        pt->covers[(x + 1) * 8 + y] =
                            wrap_i8((int32_t)pt->covers[(x + 1) * 8 + y] + seg_cover(segs[s]));
                    }
                if (id->q >= 0)
                    for (int y = 0; y < 8; y++)
                        pt->covers[y] = wrap_i8((int32_t)pt->covers[y] + wb->queue[id->q].cover[y]);
                int apply_clip = p->func == OR_FUNC_DRAW && p->is_clipped && !id->skip_clipping;
                int8_t cover[8];
                painter_paint_layer(pt, tile_i, tile_j, id->layer, p, apply_clip, cover, NULL);
                if (!cover_is_empty(cover, p->fill_rule)) {
                    or_carry cc;
                    memcpy(cc.cover, cover, 8);
                    cc.layer = id->layer;
                    cc._pad = 0;
                    wb->next[wb->n_next++] = cc;
                }
            } else {
                or_carry cc;
                if (wb_cover_carry(wb, id, segs, p->fill_rule, &cc)) wb->next[wb->n_next++] = cc;
            }
        }
    }
    /* next_tile (layer_workbench.rs:139-148) */
    or_carry *t = wb->queue;
    wb->queue = wb->next;
    wb->next = t;
    wb->n_queue = wb->n_next;
    wb->n_next = 0;
    return r;
}

int or_drive_tile(const or_carry *carries, size_t n_carries, const uint64_t *segs, size_t n_segs,
                  const or_props *props, size_t n_props, const float clear[4], uint32_t tile_i,
                  uint32_t tile_j, int32_t *previous_layers, uint32_t passes,
                  int stop_after_passes, or_carry *carries_out, size_t carries_cap,
                  or_tile_result *res) {
    workbench wb;
    memset(&wb, 0, sizeof wb);
    painter pt;
    memset(&pt, 0, sizeof pt);
    for (int i = 0; i < 64; i++) pt.alpha[i] = 1.0f;
    wb_reserve_queue(&wb, n_carries + 1);
    memcpy(wb.queue, carries, n_carries * sizeof(or_carry));
    wb.n_queue = n_carries;
    int cont = 0;
    tile_op r = wb_drive(&wb, &pt, segs, n_segs, props, n_props, clear, tile_i, tile_j,
                         previous_layers, passes, stop_after_passes, NULL, &cont);
    memset(res, 0, sizeof *res);
    res->op = stop_after_passes && cont ? -1 : r.op;
    memcpy(res->solid, r.solid, sizeof r.solid);
    for (int i = 0; i < 64; i++) {
        res->colors[4 * i] = pt.c0[i];
        res->colors[4 * i + 1] = pt.c1[i];
        res->colors[4 * i + 2] = pt.c2[i];
        res->colors[4 * i + 3] = pt.alpha[i];
    }
    size_t n = wb.n_ids < 1024 ? wb.n_ids : 1024;
    res->n_ids = n;
    for (size_t k = 0; k < n; k++) {
        res->ids[k] = wb.ids[k].layer;
        res->masked[k] = (uint8_t)id_masked(&wb, k);
        res->skip_clipping[k] = wb.ids[k].skip_clipping;
    }
    if (!stop_after_passes) {
        res->n_carries_out = wb.n_queue;
        for (size_t k = 0; k < wb.n_queue && k < carries_cap; k++) carries_out[k] = wb.queue[k];
    }
    wb_free(&wb);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Output (painter/mod.rs:418-493)                                           */
/* ------------------------------------------------------------------------ */

static void write_tile(painter *pt, const tile_op *op, uint8_t *buffer, size_t stride,
                       size_t x0, size_t y0, size_t tw, size_t th) {
    if (op->op == OR_OP_NONE) return;
    if (op->op == OR_OP_SOLID) {
        uint8_t b[4];
        or_to_bytes(op->solid, b);
        for (size_t y = 0; y < th; y++)
            for (size_t x = 0; x < tw; x++) memcpy(buffer + 4 * ((y0 + y) * stride + x0 + x), b, 4);
        return;
    }
    /* compute_srgb */
    uint8_t px[64][4];
    for (int i = 0; i < 64; i++) {
        float a = pt->alpha[i];
        float c0 = (or_linear_to_srgb_simd(pt->c0[i]) * a) * 255.0f;
        float c1 = (or_linear_to_srgb_simd(pt->c1[i]) * a) * 255.0f;
        float c2 = (or_linear_to_srgb_simd(pt->c2[i]) * a) * 255.0f;
        float al = fmaf(a, 255.0f, 0.5f);
        px[i][0] = rs_f2u8(rs_round(c0));
        px[i][1] = rs_f2u8(rs_round(c1));
        px[i][2] = rs_f2u8(rs_round(c2));
        px[i][3] = rs_f2u8(rs_round(al));
    }
    for (size_t y = 0; y < th; y++)
        for (size_t x = 0; x < tw; x++) memcpy(buffer + 4 * ((y0 + y) * stride + x0 + x), px[x * 8 + y], 4);
}

/* ------------------------------------------------------------------------ */
/* BufferLayout::print + Painter::paint_tile_row                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint16_t layer;
    int8_t cover[8];
} left_cover;

/* acc_covers for covers_left_of_row (painter/mod.rs:507-513): a layer-sorted
   map, entries kept even when their sum is empty. */
static size_t acc_left(const uint64_t *segs, size_t n, or_carry **out, size_t *cap) {
    /* segments are sorted by (column, layer): gather per layer */
    size_t count = 0;
    for (size_t i = 0; i < n; i++) {
        uint16_t l = (uint16_t)seg_layer(segs[i]);
        /* find or insert l in the sorted array */
        size_t lo = 0, hi = count;
        while (lo < hi) {
            size_t mid = (lo + hi) / 2;
            if ((*out)[mid].layer < l) lo = mid + 1; else hi = mid;
        }
        if (lo == count || (*out)[lo].layer != l) {
            if (count + 1 > *cap) {
                *cap = *cap ? *cap * 2 : 64;
                *out = (or_carry *)realloc(*out, *cap * sizeof(or_carry));
            }
            memmove(*out + lo + 1, *out + lo, (count - lo) * sizeof(or_carry));
            memset(&(*out)[lo], 0, sizeof(or_carry));
            (*out)[lo].layer = l;
            count++;
        }
        uint32_t ty = seg_ty(segs[i]);
        (*out)[lo].cover[ty] = wrap_i8((int32_t)(*out)[lo].cover[ty] + seg_cover(segs[i]));
    }
    return count;
}

static void paint_tile_row(workbench *wb, painter *pt, uint32_t j, const uint64_t *segs, size_t n,
                           const or_props *props, size_t n_props, const float clear[4],
                           int32_t *layers_row, uint8_t *buffer, size_t width, size_t stride,
                           size_t row_len, size_t tile_h, const or_rect *crop, or_debug *dbg) {
    or_carry *left = NULL;
    size_t left_cap = 0, n_left = 0;
    left = (or_carry *)malloc(64 * sizeof(or_carry));
    left_cap = 64;
    /* populate_covers(None): fold every tile left of the canvas */
    size_t k = 0;
    while (k < n && seg_col(segs[k]) < 0) k++;
    n_left = acc_left(segs, k, &left, &left_cap);
    segs += k;
    n -= k;
    if (crop && crop->h0 > 0) {
        size_t k2 = 0;
        while (k2 < n && seg_col(segs[k2]) < (int32_t)crop->h0) k2++;
        /* merge into the same map */
        or_carry *extra = NULL;
        size_t extra_cap = 0;
        size_t n_extra = acc_left(segs, k2, &extra, &extra_cap);
        for (size_t e = 0; e < n_extra; e++) {
            size_t lo = 0, hi = n_left;
            while (lo < hi) {
                size_t mid = (lo + hi) / 2;
                if (left[mid].layer < extra[e].layer) lo = mid + 1; else hi = mid;
            }
            if (lo == n_left || left[lo].layer != extra[e].layer) {
                if (n_left + 1 > left_cap) {
                    left_cap *= 2;
                    left = (or_carry *)realloc(left, left_cap * sizeof(or_carry));
                }
                memmove(left + lo + 1, left + lo, (n_left - lo) * sizeof(or_carry));
                left[lo] = extra[e];
                n_left++;
            } else {
                for (int y = 0; y < 8; y++)
                    left[lo].cover[y] = wrap_i8((int32_t)left[lo].cover[y] + extra[e].cover[y]);
            }
        }
        free(extra);
        segs += k2;
        n -= k2;
    }
    /* workbench.init */
    wb_reserve_queue(wb, n_left + 1);
    memcpy(wb->queue, left, n_left * sizeof(or_carry));
    wb->n_queue = n_left;
    free(left);

    for (size_t i = 0; i < row_len; i++) {
        if (crop && !(i >= crop->h0 && i < crop->h1)) continue;
        size_t e = 0;
        while (e < n && seg_col(segs[e]) == (int32_t)i) e++;
        const uint64_t *cur = segs;
        size_t n_cur = e;
        segs += e;
        n -= e;
        pt->has_clip = 0;
        tile_op op = wb_drive(wb, pt, cur, n_cur, props, n_props, clear, (uint32_t)i, j,
                              layers_row ? &layers_row[i] : NULL, 7, 0, dbg, NULL);
        if (dbg && dbg->tile_ops) dbg->tile_ops[j * row_len + i] = (int8_t)op.op;
        size_t tw = width - 8 * i < 8 ? width - 8 * i : 8;
        write_tile(pt, &op, buffer, stride, 8 * i, 8 * (size_t)j, tw, tile_h);
    }
}

int or_print(uint8_t *buffer, size_t len_px, size_t width, size_t width_stride,
             int32_t *layers_per_tile, size_t tiles_len, const uint64_t *segs, size_t n_segs,
             const float clear[4], const or_rect *crop, const or_props *props, size_t n_props,
             or_debug *dbg) {
    if (width_stride == 0) width_stride = width;
    if (width > len_px || width_stride > len_px || width > width_stride) return 1;
    size_t rows_px = (len_px + width_stride - 1) / width_stride;
    if (len_px - (rows_px - 1) * width_stride < width) return 1; /* row[..width] panics */
    size_t row_len = (width + 7) / 8;
    size_t tile_rows = (rows_px + 7) / 8;
    if (layers_per_tile && tiles_len != row_len * tile_rows) return 1; /* zip_eq panics */
    if (dbg && dbg->tile_ops) memset(dbg->tile_ops, -1, row_len * tile_rows);

    /* drop negative tile_j (buffer_layout.rs:187-191) */
    size_t n = 0;
    while (n < n_segs && seg_tile_j(segs[n]) >= 0) n++;
    if (n == 0) n = n_segs; /* search finds nothing: unchanged; rows below match none */

    workbench wb;
    memset(&wb, 0, sizeof wb);
    painter pt;
    memset(&pt, 0, sizeof pt);
    for (int i = 0; i < 64; i++) pt.alpha[i] = 1.0f;

    size_t pos = 0;
    for (size_t j = 0; j < tile_rows; j++) {
        /* this row's segments: tile_j == j (sorted, non-negative prefix) */
        while (pos < n && seg_tile_j(segs[pos]) >= 0 && (size_t)seg_tile_j(segs[pos]) < j) pos++;
        size_t end = pos;
        while (end < n && seg_tile_j(segs[end]) == (int32_t)j) end++;
        if (crop && !(j >= crop->v0 && j < crop->v1)) { pos = end; continue; }
        size_t tile_h = rows_px - 8 * j < 8 ? rows_px - 8 * j : 8;
        paint_tile_row(&wb, &pt, (uint32_t)j, segs + pos, end - pos, props, n_props, clear,
                       layers_per_tile ? layers_per_tile + j * row_len : NULL, buffer, width,
                       width_stride, row_len, tile_h, crop, dbg);
        pos = end;
    }
    wb_free(&wb);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Whole-frame entry points                                                  */
/* ------------------------------------------------------------------------ */

typedef struct {
    or_lines l;
    void *mem;
} lines_buf;

static int lines_alloc(lines_buf *b, size_t n) {
    size_t sz = n * (2 + 4 + 4 + 4 + 8 + 4 + 4 + 4 + 4 + 1) + 64;
    b->mem = malloc(sz ? sz : 1);
    if (!b->mem) return 1;
    char *p = (char *)b->mem;
    b->l.lengths = (int64_t *)p; p += n * 8;
    b->l.starts = (int32_t *)p; p += n * 4;
    b->l.starts_f32 = (float *)p; p += n * 4;
    b->l.ends_f32 = (float *)p; p += n * 4;
    b->l.slopes = (float *)p; p += n * 4;
    b->l.slopes_recip = (float *)p; p += n * 4;
    b->l.py_slope_pxs = (float *)p; p += n * 4;
    b->l.px_slope_recip_pys = (float *)p; p += n * 4;
    b->l.orders = (uint16_t *)p; p += n * 2;
    b->l.octants = (uint8_t *)p;
    return 0;
}

static uint64_t *build_and_rasterize(size_t n_lines, const float *p0x, const float *p0y,
                                     const float *p1x, const float *p1y, const uint16_t *ids,
                                     const or_layer *layers, size_t n_layers, size_t *n_out) {
    lines_buf lb;
    if (lines_alloc(&lb, n_lines)) return NULL;
    or_lines_build(n_lines, p0x, p0y, p1x, p1y, ids, layers, n_layers, NULL, &lb.l);
    size_t total = 0;
    for (size_t i = 0; i < n_lines; i++)
        if (lb.l.lengths[i] > 0) total += (size_t)lb.l.lengths[i];
    uint64_t *segs = (uint64_t *)malloc((2 * total + 1) * sizeof(uint64_t));
    if (!segs) { free(lb.mem); return NULL; }
    *n_out = or_rasterize(n_lines, &lb.l, segs, 2 * total);
    free(lb.mem);
    return segs;
}

size_t or_pixel_segments(size_t n_lines, const float *p0x, const float *p0y, const float *p1x,
                         const float *p1y, const uint16_t *ids, const or_layer *layers,
                         size_t n_layers, uint64_t *out, size_t cap) {
    size_t n = 0;
    uint64_t *segs = build_and_rasterize(n_lines, p0x, p0y, p1x, p1y, ids, layers, n_layers, &n);
    if (!segs) return (size_t)-1;
    or_sort_canonical(segs, n);
    if (n <= cap) memcpy(out, segs, n * sizeof(uint64_t));
    free(segs);
    return n;
}

int or_render(size_t n_lines, const float *p0x, const float *p0y, const float *p1x,
              const float *p1y, const uint16_t *ids, const or_layer *layers, size_t n_layers,
              uint8_t *buffer, size_t len_px, size_t width, size_t width_stride,
              int32_t *layers_per_tile, size_t tiles_len, const float clear[4],
              const or_rect *crop, const or_props *props, size_t n_props,
              size_t *n_segments_out, or_debug *dbg) {
    size_t n = 0;
    uint64_t *segs = build_and_rasterize(n_lines, p0x, p0y, p1x, p1y, ids, layers, n_layers, &n);
    if (!segs) return 2;
    or_sort_canonical(segs, n);
    /* cut after the last non-NONE segment (composition.rs:299-304) */
    size_t live = 0;
    while (live < n && !(segs[live] >> 63)) live++;
    if (n_segments_out) *n_segments_out = live;
    int rc = or_print(buffer, len_px, width, width_stride, layers_per_tile, tiles_len, segs, live,
                      clear, crop, props, n_props, dbg);
    free(segs);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Multi-threaded variant (CPU baseline): the rayon decomposition of the      */
/* reference -- lines, line-pixels, a parallel sort, tile rows -- on pthreads */
/* ------------------------------------------------------------------------ */

#include <pthread.h>

typedef struct {
    int tid, nt;
    void (*fn)(void *ctx, int tid, int nt);
    void *ctx;
} mt_job;

static void *mt_trampoline(void *p) {
    mt_job *j = (mt_job *)p;
    j->fn(j->ctx, j->tid, j->nt);
    return NULL;
}

static void mt_run(int nt, void (*fn)(void *, int, int), void *ctx) {
    pthread_t th[256];
    mt_job jobs[256];
    memset(jobs, 0, sizeof jobs);
    if (nt < 1) nt = 1;
    if (nt > 256) nt = 256;
    for (int t = 0; t < nt; t++) {
        jobs[t].tid = t; jobs[t].nt = nt; jobs[t].fn = fn; jobs[t].ctx = ctx;
        if (t) pthread_create(&th[t], NULL, mt_trampoline, &jobs[t]);
    }
    mt_trampoline(&jobs[0]);
    for (int t = 1; t < nt; t++) pthread_join(th[t], NULL);
}

typedef struct {
    size_t n_lines;
    const float *p0x, *p0y, *p1x, *p1y;
    const uint16_t *ids;
    const or_layer *layers;
    size_t n_layers;
    or_lines *l;
    size_t *offsets; /* per line exclusive prefix of lengths */
    uint64_t *segs, *tmp;
    size_t n_segs;
    size_t (*hist)[256];
    int shift;
    /* print */
    uint8_t *buffer;
    size_t width, stride, row_len, tile_rows, rows_px;
    const size_t *row_start, *row_end;
    const or_props *props;
    size_t n_props;
    const float *clear;
    volatile int next_row;
    pthread_mutex_t lock;
} mt_ctx;

static void mt_build(void *p, int tid, int nt) {
    mt_ctx *c = (mt_ctx *)p;
    size_t a = c->n_lines * tid / nt, b = c->n_lines * (tid + 1) / nt;
    for (size_t i = a; i < b; i++)
        build_one(i, c->p0x, c->p0y, c->p1x, c->p1y, c->ids, c->layers, c->n_layers, NULL, c->l);
}

static void mt_raster(void *p, int tid, int nt) {
    mt_ctx *c = (mt_ctx *)p;
    size_t a = c->n_lines * tid / nt, b = c->n_lines * (tid + 1) / nt;
    for (size_t li = a; li < b; li++) {
        int64_t len = c->l->lengths[li];
        uint64_t *o = c->segs + 2 * c->offsets[li];
        for (int64_t pi = 0; pi < len; pi++) rasterize_pixel(c->l, li, (int32_t)pi, o + 2 * pi);
    }
}

static void mt_hist(void *p, int tid, int nt) {
    mt_ctx *c = (mt_ctx *)p;
    size_t a = c->n_segs * tid / nt, b = c->n_segs * (tid + 1) / nt;
    size_t *h = c->hist[tid];
    memset(h, 0, 256 * sizeof(size_t));
    for (size_t i = a; i < b; i++) h[(c->segs[i] >> c->shift) & 255]++;
}

static void mt_scatter(void *p, int tid, int nt) {
    mt_ctx *c = (mt_ctx *)p;
    size_t a = c->n_segs * tid / nt, b = c->n_segs * (tid + 1) / nt;
    size_t *h = c->hist[tid];
    for (size_t i = a; i < b; i++) c->tmp[h[(c->segs[i] >> c->shift) & 255]++] = c->segs[i];
}

static void mt_print(void *p, int tid, int nt) {
    mt_ctx *c = (mt_ctx *)p;
    workbench wb;
    memset(&wb, 0, sizeof wb);
    painter pt;
    memset(&pt, 0, sizeof pt);
    for (int i = 0; i < 64; i++) pt.alpha[i] = 1.0f;
    for (;;) {
        pthread_mutex_lock(&c->lock);
        int j = c->next_row++;
        pthread_mutex_unlock(&c->lock);
        if ((size_t)j >= c->tile_rows) break;
        size_t tile_h = c->rows_px - 8 * (size_t)j < 8 ? c->rows_px - 8 * (size_t)j : 8;
        paint_tile_row(&wb, &pt, (uint32_t)j, c->segs + c->row_start[j], c->row_end[j] - c->row_start[j],
                       c->props, c->n_props, c->clear, NULL, c->buffer, c->width, c->stride,
                       c->row_len, tile_h, NULL, NULL);
    }
    wb_free(&wb);
}

int or_render_mt(int nt, size_t n_lines, const float *p0x, const float *p0y, const float *p1x,
                 const float *p1y, const uint16_t *ids, const or_layer *layers, size_t n_layers,
                 uint8_t *buffer, size_t len_px, size_t width, size_t width_stride,
                 const float clear[4], const or_props *props, size_t n_props,
                 size_t *n_segments_out) {
    if (nt < 1) nt = 1;
    if (width_stride == 0) width_stride = width;
    if (width > len_px || width_stride > len_px || width > width_stride) return 1;
    mt_ctx c;
    memset(&c, 0, sizeof c);
    c.n_lines = n_lines; c.p0x = p0x; c.p0y = p0y; c.p1x = p1x; c.p1y = p1y;
    c.ids = ids; c.layers = layers; c.n_layers = n_layers;
    lines_buf lb;
    if (lines_alloc(&lb, n_lines)) return 2;
    c.l = &lb.l;
    mt_run(nt, mt_build, &c);
    c.offsets = (size_t *)malloc((n_lines + 1) * sizeof(size_t));
    size_t total = 0;
    for (size_t i = 0; i < n_lines; i++) {
        c.offsets[i] = total;
        if (lb.l.lengths[i] > 0) total += (size_t)lb.l.lengths[i];
    }
    c.n_segs = 2 * total;
    c.segs = (uint64_t *)malloc((c.n_segs + 1) * sizeof(uint64_t));
    c.tmp = (uint64_t *)malloc((c.n_segs + 1) * sizeof(uint64_t));
    mt_run(nt, mt_raster, &c);
    free(lb.mem);
    /* parallel stable LSD radix sort on bits [22, 64) in 8-bit digits */
    c.hist = (size_t(*)[256])malloc((size_t)nt * sizeof(size_t[256]));
    for (int shift = 22; shift < 64; shift += 8) {
        c.shift = shift;
        mt_run(nt, mt_hist, &c);
        size_t run = 0;
        for (int d = 0; d < 256; d++)
            for (int t = 0; t < nt; t++) {
                size_t v = c.hist[t][d];
                c.hist[t][d] = run;
                run += v;
            }
        mt_run(nt, mt_scatter, &c);
        uint64_t *tt = c.segs; c.segs = c.tmp; c.tmp = tt;
    }
    free(c.hist);
    size_t live = 0;
    while (live < c.n_segs && !(c.segs[live] >> 63)) live++;
    if (n_segments_out) *n_segments_out = live;
    /* rows */
    c.rows_px = (len_px + width_stride - 1) / width_stride;
    c.row_len = (width + 7) / 8;
    c.tile_rows = (c.rows_px + 7) / 8;
    size_t *rs = (size_t *)calloc(c.tile_rows + 1, sizeof(size_t));
    size_t *re = (size_t *)calloc(c.tile_rows + 1, sizeof(size_t));
    size_t n = 0;
    while (n < live && seg_tile_j(c.segs[n]) >= 0) n++;
    size_t pos = 0;
    for (size_t j = 0; j < c.tile_rows; j++) {
        while (pos < n && (size_t)seg_tile_j(c.segs[pos]) < j) pos++;
        size_t e = pos;
        while (e < n && seg_tile_j(c.segs[e]) == (int32_t)j) e++;
        rs[j] = pos; re[j] = e; pos = e;
    }
    c.row_start = rs; c.row_end = re;
    c.buffer = buffer; c.width = width; c.stride = width_stride;
    c.props = props; c.n_props = n_props; c.clear = clear;
    pthread_mutex_init(&c.lock, NULL);
    mt_run(nt, mt_print, &c);
    pthread_mutex_destroy(&c.lock);
    free(rs); free(re); free(c.segs); free(c.tmp); free(c.offsets);
    return 0;
}
