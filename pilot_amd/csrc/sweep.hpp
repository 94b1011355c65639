// Exact-sweep transmittance map (sweep.hip).
#pragma once

#include "kernels.hpp"

namespace beams {

// The map's geometry in the sweep's axis order (0: the sweep axis, 1 and 2: a1 and a2): origin,
// spacing, reciprocal spacing, the light (position or direction) and the vertex counts.
struct SweepFrame {
  float org[3], sp[3], inv[3], light[3];
  int32_t n[3];
};

struct SweepParams {
  GridParams g;       // field and coordinates
  MapParams m;        // map geometry and output; m.ms is the slice-contiguous layout
  ShadeParams s;      // transfer function (alpha / max alpha) and scalar range
  int32_t axis;       // sweep axis
  int32_t a1, a2;     // transverse axes (a1 fastest in the slice and map layout)
  int32_t step;       // +1: slices in increasing index, -1: decreasing
  float h;            // reference light step |size| / 128
  float inv_h;
  SweepFrame f;       // sweep_frame(m, axis, a1, a2)
  float2* prev;       // (tau, sigma) of the previous slice, mp[a1] * mp[a2]
  float2* cur;        // (tau, sigma) of the current slice
};

SweepFrame sweep_frame(const MapParams& m, int axis, int a1, int a2);

// Launch shape of a sweep: how far a light ray's bilinear stencil reaches across one slice on
// the low and high side of each transverse axis (a1, a2), in vertices; slices per blocked
// launch nt, the tile width, their LDS, and whether the blocked shape applies.
struct SweepPlan {
  int32_t lo[2], hi[2];
  int32_t nt, tile;
  size_t lds;
  bool blocked;
};
SweepPlan sweep_plan(const SweepParams& sp);

// Enqueues the sweep on st: blocked (sigma pass, then ceil(n0 / nt) launches) when plan.blocked
// and sig (one float per map vertex) is given, else one launch per slice.
hipError_t launch_sweep(SweepParams sp, const SweepPlan& plan, float* sig, hipStream_t st);

}  // namespace beams
