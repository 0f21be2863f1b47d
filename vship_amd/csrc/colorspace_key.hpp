// colorspace_key.hpp - a map key for a Vship_Colorspace_t built from its named
// fields. The struct has 4 bytes of tail padding (84 bytes of members in 88),
// which C callers filling a descriptor on the stack may leave uninitialised;
// keying on the raw bytes would then split identical descriptors.
// Host-only, so that the CPU tests can check it.
#pragma once

#include <stdint.h>

#include <string>

#include "../../include/VshipColor.h"

namespace vship {

inline std::string colorspace_key(const Vship_Colorspace_t& cs)
{
    const int64_t f[17] = {cs.width,
                           cs.height,
                           cs.target_width,
                           cs.target_height,
                           (int64_t)cs.sample,
                           (int64_t)cs.range,
                           (int64_t)cs.subsampling.subw,
                           (int64_t)cs.subsampling.subh,
                           (int64_t)cs.chromaLocation,
                           (int64_t)cs.colorFamily,
                           (int64_t)cs.YUVMatrix,
                           (int64_t)cs.transferFunction,
                           (int64_t)cs.primaries,
                           (int64_t)cs.crop.top,
                           (int64_t)cs.crop.bottom,
                           (int64_t)cs.crop.left,
                           (int64_t)cs.crop.right};
    return std::string(reinterpret_cast<const char*>(f), sizeof(f));
}

}  // namespace vship
