// mold's path flattening (mold/src/path/segments.rs:51-263, transform.rs:9-93,
// point.rs:46-58) as host/device code: the host Path::flatten and the GPU
// flattening kernel run these same binary32 operations (no contraction,
// correctly rounded division and square root), so both give the same lines.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mg {

struct PathCommand {
    enum Kind : uint8_t { Line, Quad, RatQuad, Cubic, RatCubic, Close } kind;
    // Points as (x, y); rational commands as premultiplied (x*w, y*w, w).
    float v[12];
};

namespace flat {

struct V2 {
    float x, y;
};

__host__ __device__ inline V2 lerp2(float t, V2 a, V2 b) {
    // Point::lerp: p0 * (1 - t) + p1 * t (point.rs:47-49)
    float u = 1.0f - t;
    return V2{a.x * u + b.x * t, a.y * u + b.y * t};
}

__host__ __device__ inline float lerp1(float t, float a, float b) { return a * (1.0f - t) + b * t; }

__host__ __device__ inline float max_nan(float a, float b) { return (a != a) ? b : ((b > a) ? b : a); }

__host__ __device__ inline float absf(float v) { return __builtin_fabsf(v); }

// Transform::new: normalised by t22, perspective when |t6| or |t7| > eps
// (transform.rs:57-69).
struct Xform {
    float t[9];
    bool perspective;
    __host__ __device__ void init(const float *t9) {
        float r = 1.0f / t9[8];
        for (int i = 0; i < 9; ++i) t[i] = t9[i] * r;
        const float eps = 1.1920929e-7f;
        perspective = absf(t[6]) > eps || absf(t[7]) > eps;
    }
    __host__ __device__ void full(const float p[3], float o[3]) const {
        o[0] = t[0] * p[0] + t[1] * p[1] + t[2] * p[2];
        o[1] = t[3] * p[0] + t[4] * p[1] + t[5] * p[2];
        o[2] = t[6] * p[0] + t[7] * p[1] + t[8] * p[2];
    }
    // Transform::transform (transform.rs:71-93)
    __host__ __device__ PathCommand apply(const PathCommand &c) const {
        PathCommand o = c;
        int n = 0;
        bool rational = false;
        switch (c.kind) {
        case PathCommand::Line: n = 2; break;
        case PathCommand::Quad: n = 3; break;
        case PathCommand::RatQuad: n = 3; rational = true; break;
        case PathCommand::Cubic: n = 4; break;
        case PathCommand::RatCubic: n = 4; rational = true; break;
        default: return o;
        }
        if (!rational && !perspective) {
            for (int i = 0; i < n; ++i) {
                float x = c.v[2 * i], y = c.v[2 * i + 1];
                o.v[2 * i] = t[0] * x + t[1] * y + t[2];
                o.v[2 * i + 1] = t[3] * x + t[4] * y + t[5];
            }
            return o;
        }
        if (c.kind == PathCommand::Line) {
            for (int i = 0; i < 2; ++i) {
                float p[3] = {c.v[2 * i], c.v[2 * i + 1], 1.0f}, q[3];
                full(p, q);
                float r = 1.0f / q[2];
                o.v[2 * i] = q[0] * r;
                o.v[2 * i + 1] = q[1] * r;
            }
            return o;
        }
        o.kind = n == 3 ? PathCommand::RatQuad : PathCommand::RatCubic;
        for (int i = 0; i < n; ++i) {
            float p[3], q[3];
            if (rational) {
                p[0] = c.v[3 * i]; p[1] = c.v[3 * i + 1]; p[2] = c.v[3 * i + 2];
            } else {
                p[0] = c.v[2 * i]; p[1] = c.v[2 * i + 1]; p[2] = 1.0f;
            }
            full(p, q);
            o.v[3 * i] = q[0]; o.v[3 * i + 1] = q[1]; o.v[3 * i + 2] = q[2];
        }
        return o;
    }
};

// The curve being subdivided: control points, weights and the endpoint /
// deviation parameters of next_segment (segments.rs:159-206).
struct Curve {
    int n;
    bool rational;
    V2 p[4];     // premultiplied for rational curves
    float w[4];  // weights (1 for plain curves)
    V2 first, last;
    float dev_x, dev_y;

    __host__ __device__ void init(const PathCommand &c) {
        rational = c.kind == PathCommand::RatQuad || c.kind == PathCommand::RatCubic;
        n = (c.kind == PathCommand::Quad || c.kind == PathCommand::RatQuad) ? 3 : 4;
        V2 q[4];
        for (int i = 0; i < n; ++i) {
            if (rational) {
                p[i] = V2{c.v[3 * i], c.v[3 * i + 1]};
                w[i] = c.v[3 * i + 2];
                float r = 1.0f / w[i];  // Point::from_weighted (point.rs:51-53)
                q[i] = V2{p[i].x * r, p[i].y * r};
            } else {
                p[i] = V2{c.v[2 * i], c.v[2 * i + 1]};
                w[i] = 1.0f;
                q[i] = p[i];
            }
        }
        first = q[0];
        last = q[n - 1];
        if (n == 3) {
            dev_x = q[0].x + q[2].x - 2.0f * q[1].x;
            dev_y = q[0].y + q[2].y - 2.0f * q[1].y;
        } else {
            dev_x = max_nan(absf(q[0].x + q[2].x - 2.0f * q[1].x), absf(q[1].x + q[3].x - 2.0f * q[2].x));
            dev_y = max_nan(absf(q[0].y + q[2].y - 2.0f * q[1].y), absf(q[1].y + q[3].y - 2.0f * q[2].y));
        }
    }

    // bezier! / lerp! of the control net at t (segments.rs:19-29,74-84)
    __host__ __device__ V2 at(float t) const {
        V2 b;
        if (n == 3) {
            b = lerp2(t, lerp2(t, p[0], p[1]), lerp2(t, p[1], p[2]));
        } else {
            V2 a = lerp2(t, lerp2(t, p[0], p[1]), lerp2(t, p[1], p[2]));
            V2 c = lerp2(t, lerp2(t, p[1], p[2]), lerp2(t, p[2], p[3]));
            b = lerp2(t, a, c);
        }
        if (!rational) return b;
        float wt;
        if (n == 3) {
            wt = lerp1(t, lerp1(t, w[0], w[1]), lerp1(t, w[1], w[2]));
        } else {
            float wa = lerp1(t, lerp1(t, w[0], w[1]), lerp1(t, w[1], w[2]));
            float wb = lerp1(t, lerp1(t, w[1], w[2]), lerp1(t, w[2], w[3]));
            wt = lerp1(t, wa, wb);
        }
        float r = 1.0f / wt;
        return V2{b.x * r, b.y * r};
    }
};

// `f32 as usize` (saturating, NaN -> 0)
__host__ __device__ inline uint64_t to_usize(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 18446744073709551616.0f) return ~0ull;
    return (uint64_t)v;
}

// Lines a curve subdivides into: subdivisions (segments.rs:51-72) with the
// "deviation below accuracy: one line" case.
__host__ __device__ inline uint64_t curve_lines(const Curve &c) {
    const float accuracy = 0.5f / 16.0f;
    float dev2 = c.dev_x * c.dev_x + c.dev_y * c.dev_y;
    if (dev2 < accuracy) return 1;
    uint64_t count = to_usize(ceilf(sqrtf(sqrtf(dev2) * 2.0f / (8.0f * accuracy))));
    uint64_t sub = (count > 1 ? count : 1) - 1;
    return sub == 0 ? 1 : sub;
}

// PathSegments over the commands of one path (segments.rs:106-223), each
// line given to emit(x0, y0, x1, y1) in order. `xf` may be null.
template <class Emit>
__host__ __device__ inline void flatten_commands(const PathCommand *cmds, uint64_t n, const Xform *xf, Emit &&emit) {
    bool has_closing = false, has_last = false;
    V2 closing{0.0f, 0.0f}, last{0.0f, 0.0f};
    // PathSegments::line (segments.rs:106-114)
    auto line = [&](V2 a, V2 b) {
        if (!has_closing) {
            has_closing = true;
            closing = a;
        }
        has_last = true;
        last = b;
        emit(a.x, a.y, b.x, b.y);
    };
    for (uint64_t i = 0; i < n; ++i) {
        const PathCommand c = xf ? xf->apply(cmds[i]) : cmds[i];
        if (c.kind == PathCommand::Close) {
            // close (segments.rs:208-223)
            if (has_closing && has_last) {
                const float eps = 1.1920929e-7f;
                bool near = absf(closing.x - last.x) < eps && absf(closing.y - last.y) < eps;
                if (!near) line(last, closing);
                has_closing = false;
            }
        } else if (c.kind == PathCommand::Line) {
            line(V2{c.v[0], c.v[1]}, V2{c.v[2], c.v[3]});
        } else {
            Curve cv;
            cv.init(c);
            const uint64_t sub = curve_lines(cv);
            if (sub == 1) {
                line(cv.first, cv.last);
                continue;
            }
            // serial t += 1/n; the last point is the curve's end (segments.rs:116-157)
            float inc = 1.0f / (float)sub;
            float t = 0.0f;
            V2 pt = cv.first;
            for (uint64_t left = sub; left > 1; --left) {
                t += inc;
                V2 np = cv.at(t);
                line(pt, np);
                pt = np;
            }
            line(pt, cv.last);
        }
    }
}

// Number of lines flatten_commands gives for the commands, without
// evaluating the curves: a curve's lines run from its first to its last
// point, which is all the close logic looks at.
__host__ __device__ inline uint64_t count_commands(const PathCommand *cmds, uint64_t n, const Xform *xf) {
    bool has_closing = false, has_last = false;
    V2 closing{0.0f, 0.0f}, last{0.0f, 0.0f};
    uint64_t count = 0;
    auto span = [&](V2 a, V2 b, uint64_t lines) {
        if (!has_closing) {
            has_closing = true;
            closing = a;
        }
        has_last = true;
        last = b;
        count += lines;
    };
    for (uint64_t i = 0; i < n; ++i) {
        const PathCommand c = xf ? xf->apply(cmds[i]) : cmds[i];
        if (c.kind == PathCommand::Close) {
            if (has_closing && has_last) {
                const float eps = 1.1920929e-7f;
                bool near = absf(closing.x - last.x) < eps && absf(closing.y - last.y) < eps;
                if (!near) span(last, closing, 1);
                has_closing = false;
            }
        } else if (c.kind == PathCommand::Line) {
            span(V2{c.v[0], c.v[1]}, V2{c.v[2], c.v[3]}, 1);
        } else {
            Curve cv;
            cv.init(c);
            span(cv.first, cv.last, curve_lines(cv));
        }
    }
    return count;
}

// One path of a device flattening batch: its commands, layer and transform.
struct PathDesc {
    uint32_t first_cmd, n_cmds;
    uint32_t layer, has_xf;
    float t9[9];
    uint32_t _pad[3];
};

}  // namespace flat
}  // namespace mg
