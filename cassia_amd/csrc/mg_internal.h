// Library-internal entry points of the C ABI's implementation (mg_api.cpp)
// used by the serving loop (mg_server.cpp).
#pragma once

#include <memory>
#include <string>

#include "../../include/mold_gpu.h"
#include "mg_runtime.h"

namespace mgi {

mg_status set_error(mg_status s, const std::string &msg);
const char *last_error();

// A batch bound to `engine` (its own stream and scratch), for one serving
// worker: no autotuner (serving batches change shape every batch), the
// runtime's automatic radix and walk size.
mg_batch *batch_new(std::shared_ptr<mg::Renderer> engine);

// Frame tables and resident lines of a composition for batched frames (done
// at submission, so workers only read compositions).
mg_status prepare(mg_composition *c);

// Device-memory estimate of rendering `c` into `frame` alone: resident lines,
// line setup, sort buffers of the estimated slots, the frame, per-tile state.
// `line_pixels` receives the estimate of the composition's line-pixels.
uint64_t footprint(const mg_composition *c, const mg_batch_frame &frame, uint64_t *line_pixels);

// mg_batch_render on an engine locked by the call; `prepare` false when the
// compositions' tables are current.
mg_status batch_render(mg_batch *b, mg_composition *const *comps, const mg_batch_frame *frames, size_t n,
                       bool prepare, void **device_rgba, const uint64_t **offsets);

// Frame i of the batch's last render into host memory.
mg_status batch_download(mg_batch *b, size_t i, uint8_t *rgba, size_t len);

}  // namespace mgi
