// kernels.hpp - host-side launchers of the gfx950 kernels.
#pragma once

#include "ssimu2_common.hpp"

namespace vship {

// LDS bytes the front-end kernel needs for images of this colour description.
size_t frontend_lds_bytes(const FrontColor& fc);

template <typename StoreT>
hipError_t launch_frontend(const FrontJob* jobs, int njobs, const int* total_tiles, int grid,
                           const FrontColor* colors, const PairGeom* geoms, size_t lds_bytes, hipStream_t stream);

template <typename StoreT>
hipError_t launch_blur(const MapJob* jobs, int njobs, const int* total_tiles, int grid, const BlurTaps& taps,
                       float* partials, hipStream_t stream);

hipError_t launch_score(const ScoreJob* pairs, const int* npairs, int grid, const MapJob* jobs,
                        const float* partials, double* features, double* scores, hipStream_t stream);

}  // namespace vship
