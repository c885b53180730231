// Path flattening, restated from mold/src/path/segments.rs:51-263,
// transform.rs:9-93 and point.rs:46-58.
#include "mg_path.h"

#include <cmath>
#include <cstring>

namespace mg {

namespace {

struct V2 {
    float x, y;
};

inline V2 lerp2(float t, V2 a, V2 b) {
    // Point::lerp: p0 * (1 - t) + p1 * t (point.rs:47-49)
    float u = 1.0f - t;
    return V2{a.x * u + b.x * t, a.y * u + b.y * t};
}

inline float lerp1(float t, float a, float b) { return a * (1.0f - t) + b * t; }

inline float max_nan(float a, float b) { return (a != a) ? b : ((b > a) ? b : a); }

// Transform::new: normalised by t22, perspective when |t6| or |t7| > eps
// (transform.rs:57-69).
struct Xform {
    float t[9];
    bool perspective;
    explicit Xform(const float *t9) {
        float r = 1.0f / t9[8];
        for (int i = 0; i < 9; ++i) t[i] = t9[i] * r;
        const float eps = 1.1920929e-7f;
        perspective = std::fabs(t[6]) > eps || std::fabs(t[7]) > eps;
    }
    void full(const float p[3], float o[3]) const {
        o[0] = t[0] * p[0] + t[1] * p[1] + t[2] * p[2];
        o[1] = t[3] * p[0] + t[4] * p[1] + t[5] * p[2];
        o[2] = t[6] * p[0] + t[7] * p[1] + t[8] * p[2];
    }
    // Transform::transform (transform.rs:71-93)
    PathCommand apply(const PathCommand &c) const {
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
    int n = 0;
    bool rational = false;
    V2 p[4];      // premultiplied for rational curves
    float w[4];   // weights (1 for plain curves)
    V2 first, last;
    float dev_x = 0, dev_y = 0;

    explicit Curve(const PathCommand &c) {
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
            dev_x = max_nan(std::fabs(q[0].x + q[2].x - 2.0f * q[1].x), std::fabs(q[1].x + q[3].x - 2.0f * q[2].x));
            dev_y = max_nan(std::fabs(q[0].y + q[2].y - 2.0f * q[1].y), std::fabs(q[1].y + q[3].y - 2.0f * q[2].y));
        }
    }

    // bezier! / lerp! of the control net at t (segments.rs:19-29,74-84)
    V2 at(float t) const {
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
inline uint64_t to_usize(float v) {
    if (!(v > 0.0f)) return 0;
    if (v >= 18446744073709551616.0f) return UINT64_MAX;
    return (uint64_t)v;
}

class Flattener {
public:
    Flattener(const std::vector<PathCommand> &cmds, const float *t9, std::vector<float> &out)
        : cmds_(cmds), out_(out) {
        if (t9) {
            xf_ = new Xform(t9);
        }
    }
    ~Flattener() { delete xf_; }

    void run() {
        size_t i = 0;
        while (i < cmds_.size()) {
            PathCommand c = xf_ ? xf_->apply(cmds_[i]) : cmds_[i];
            switch (c.kind) {
            case PathCommand::Close:
                close();
                ++i;
                break;
            case PathCommand::Line:
                emit(V2{c.v[0], c.v[1]}, V2{c.v[2], c.v[3]});
                ++i;
                break;
            default:
                curve(Curve(c));
                ++i;
                break;
            }
        }
    }

private:
    // PathSegments::line (segments.rs:106-114)
    void emit(V2 a, V2 b) {
        if (!has_closing_) {
            has_closing_ = true;
            closing_ = a;
        }
        has_last_ = true;
        last_ = b;
        out_.push_back(a.x);
        out_.push_back(a.y);
        out_.push_back(b.x);
        out_.push_back(b.y);
    }

    // close (segments.rs:208-223)
    void close() {
        if (has_closing_ && has_last_) {
            const float eps = 1.1920929e-7f;
            bool near = std::fabs(closing_.x - last_.x) < eps && std::fabs(closing_.y - last_.y) < eps;
            if (!near) emit(last_, closing_);
            has_closing_ = false;
        }
    }

    // subdivisions + next_segment driven to completion (segments.rs:51-72,116-157)
    void curve(const Curve &c) {
        const float accuracy = 0.5f / 16.0f;
        float dev2 = c.dev_x * c.dev_x + c.dev_y * c.dev_y;
        if (dev2 < accuracy) {
            emit(c.first, c.last);
            return;
        }
        uint64_t count = to_usize(std::ceil(std::sqrt(std::sqrt(dev2) * 2.0f / (8.0f * accuracy))));
        uint64_t sub = (count > 1 ? count : 1) - 1;
        if (sub == 0) {
            emit(c.first, c.last);
            return;
        }
        float inc = 1.0f / (float)sub;
        float t = 0.0f;
        V2 pt = c.first;
        for (uint64_t left = sub; left > 1; --left) {
            t += inc;
            V2 np = c.at(t);
            emit(pt, np);
            pt = np;
        }
        emit(pt, c.last);
    }

    const std::vector<PathCommand> &cmds_;
    std::vector<float> &out_;
    Xform *xf_ = nullptr;
    bool has_closing_ = false, has_last_ = false;
    V2 closing_{0, 0}, last_{0, 0};
};

}  // namespace

void Path::line(const float p[4]) {
    PathCommand c{};
    c.kind = PathCommand::Line;
    std::memcpy(c.v, p, 4 * sizeof(float));
    cmds_.push_back(c);
}

void Path::quad(const float p[6]) {
    PathCommand c{};
    c.kind = PathCommand::Quad;
    std::memcpy(c.v, p, 6 * sizeof(float));
    cmds_.push_back(c);
}

void Path::rat_quad(const float p[9]) {
    // (p.0 * p.1, p.1) per point (path/mod.rs:92-96)
    PathCommand c{};
    c.kind = PathCommand::RatQuad;
    for (int i = 0; i < 3; ++i) {
        float w = p[3 * i + 2];
        c.v[3 * i] = p[3 * i] * w;
        c.v[3 * i + 1] = p[3 * i + 1] * w;
        c.v[3 * i + 2] = w;
    }
    cmds_.push_back(c);
}

void Path::cubic(const float p[8]) {
    PathCommand c{};
    c.kind = PathCommand::Cubic;
    std::memcpy(c.v, p, 8 * sizeof(float));
    cmds_.push_back(c);
}

void Path::rat_cubic(const float p[12]) {
    PathCommand c{};
    c.kind = PathCommand::RatCubic;
    for (int i = 0; i < 4; ++i) {
        float w = p[3 * i + 2];
        c.v[3 * i] = p[3 * i] * w;
        c.v[3 * i + 1] = p[3 * i + 1] * w;
        c.v[3 * i + 2] = w;
    }
    cmds_.push_back(c);
}

void Path::close() {
    PathCommand c{};
    c.kind = PathCommand::Close;
    cmds_.push_back(c);
}

void Path::flatten(const float *t9, std::vector<float> &out) const {
    Flattener f(cmds_, t9, out);
    f.run();
}

}  // namespace mg
