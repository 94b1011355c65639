// Pinhole camera rays shared by the host ray generator and the image kernel, so
// both produce the same FP32 rays: FP64 set-up rounded once to FP32.
#pragma once

#include "kernels.hpp"

namespace beams {

BEAMS_HD void camera_ray(const CameraParams& c, int px, int py, float* org, float* dir,
                         double* org64 = nullptr, double* dir64 = nullptr)
{
  const double u =
    (((double)px + 0.5) / (double)c.width * 2.0 - 1.0) * c.tan_half_fov * c.aspect;
  const double v = (1.0 - ((double)py + 0.5) / (double)c.height * 2.0) * c.tan_half_fov;
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = c.fwd[k] + u * c.right[k] + v * c.up[k];
  const double len = __builtin_sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int k = 0; k < 3; ++k) {
    const double dk = d[k] / len;
    dir[k] = (float)dk;
    org[k] = (float)c.eye[k];
    if (dir64) dir64[k] = dk;
    if (org64) org64[k] = c.eye[k];
  }
}

}  // namespace beams
