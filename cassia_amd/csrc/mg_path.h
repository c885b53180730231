// mold::Path and its flattening to lines (mold/src/path/{mod,segments,transform}.rs).
//
// Host code, binary32 throughout and built with -ffp-contract=off: the
// flattened endpoints are inputs of the bit-exact pipeline.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "mg_flatten.h"

namespace mg {

class Path {
public:
    void line(const float p[4]);
    void quad(const float p[6]);
    void rat_quad(const float p[9]);   // (x, y, w) per point, premultiplied here
    void cubic(const float p[8]);
    void rat_cubic(const float p[12]);
    void close();
    // Appends the flattened segments (x0, y0, x1, y1) to `out`; t9 may be null.
    void flatten(const float *t9, std::vector<float> &out) const;
    const std::vector<PathCommand> &commands() const { return cmds_; }

private:
    std::vector<PathCommand> cmds_;
};

}  // namespace mg
