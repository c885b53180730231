// Static device scratch of the no-allocation render entry point.
#pragma once

#include <stddef.h>

#include <string>

namespace mg {

// Size of the static scratch array (per device): the largest frame the
// no-allocation entry point renders needs its line inputs, per-line and
// per-tile tables and two 16-byte sort slots per line-pixel inside it.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int kMaxScratchDevices = 64;

// Device address of `device`'s copy of the scratch array (looked up once per
// device with hipGetSymbolAddress, with the device current).
bool scratch_for_device(int device, void **out, std::string &err);

}  // namespace mg
