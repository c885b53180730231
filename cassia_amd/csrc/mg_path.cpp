// mold::Path: the command list (mold/src/path/mod.rs:38-178); flattening in
// mg_flatten.h, shared with the GPU kernel.
#include "mg_path.h"

#include <cmath>
#include <cstring>

namespace mg {


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
    flat::Xform xf;
    if (t9) xf.init(t9);
    flat::flatten_commands(cmds_.data(), cmds_.size(), t9 ? &xf : nullptr,
                           [&](float x0, float y0, float x1, float y1) {
                               out.push_back(x0);
                               out.push_back(y0);
                               out.push_back(x1);
                               out.push_back(y1);
                           });
}

}  // namespace mg
