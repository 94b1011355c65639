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

// Launch shape of a sweep: the halo R (vertices a light ray can move across one slice, plus
// one), slices per blocked launch nt, the tile width, their LDS, and whether the blocked shape
// applies.
struct SweepPlan {
  int32_t R, nt, tile;
  size_t lds;
  bool blocked;
};
SweepPlan sweep_plan(const SweepParams& sp, const double* bounds);

// Enqueues the sweep on st: blocked (sigma pass, then ceil(n0 / nt) launches) when plan.blocked
// and sig (one float per map vertex) is given, else one launch per slice.
hipError_t launch_sweep(SweepParams sp, const SweepPlan& plan, float* sig, hipStream_t st);

}  // namespace beams
