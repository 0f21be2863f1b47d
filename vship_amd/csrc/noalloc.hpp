// noalloc.hpp - the pair entry point that allocates no device memory
// (VshipAMD_ComputeSSIMU2DeviceNoAlloc): its tables, XYB planes and partial
// sums live in a fixed device array of kNoAllocScratchBytes per GPU.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "../../include/VshipColor.h"

namespace vship {

// 32 MiB: the two FP32 pyramids of a 1920x1080 pair take 28.1 MB, of a
// 2048x1080 pair 30.0 MB.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
// blur items per call (their partial sums pass through a 32 KB buffer on the
// calling thread's stack); 1920x1080 has 498
constexpr int kNoAllocMaxItems = 1024;

// Score of one pair of device-resident planes on `device`; throws vship::Error.
double compute_ssimu2_noalloc(const Vship_Colorspace_t& src, const Vship_Colorspace_t& dis,
                              const uint8_t* const p1[3], const uint8_t* const p2[3], const int64_t ls1[3],
                              const int64_t ls2[3], int device);

}  // namespace vship
