// Exact-sweep transmittance map (sweep.hip).
#pragma once

#include "kernels.hpp"

namespace beams {

struct SweepParams {
  GridParams g;       // field and coordinates
  MapParams m;        // map geometry and output; m.ms is the slice-contiguous layout
  ShadeParams s;      // transfer function (alpha / max alpha) and scalar range
  int32_t axis;       // sweep axis
  int32_t a1, a2;     // transverse axes (a1 fastest in the slice and map layout)
  int32_t step;       // +1: slices in increasing index, -1: decreasing
  float h;            // reference light step |size| / 128
  float inv_h;
  float2* prev;       // (tau, sigma) of the previous slice, mp[a1] * mp[a2]
  float2* cur;        // (tau, sigma) of the current slice
};

// Enqueues one launch per slice on st.
hipError_t launch_sweep(SweepParams sp, hipStream_t st);

}  // namespace beams
