// Seeded synthetic scalar fields: a sum of isotropic Gaussian blobs plus lattice
// value noise, normalised to [0, 1].  The per-point arithmetic is FP64 add/mul
// only (the exponentials are tabulated per axis on the host), so the host and
// gfx950 evaluations of one recipe produce identical bits.
#pragma once

#include "common.hpp"

namespace beams {

struct BlobTables {
  int32_t nblobs;
  int64_t n[3];            // point dims of the global grid
  const double* amp;       // [nblobs]
  const double* axis[3];   // axis[d][k * n[d] + i] = exp(-(x_i - c_kd)^2 / (2 sigma_k^2))
};

struct NoiseParams {
  int32_t octaves;
  double amp;              // amplitude of octave 0
  double freq;             // base lattice frequency
  double phase;            // lattice offset (advances with time)
  uint64_t salt;
};

// Unnormalised field value at global point (i, j, k).
BEAMS_HD double field_raw(const BlobTables& t, const NoiseParams& np, int64_t i, int64_t j,
                          int64_t k)
{
  double v = 0.0;
  for (int32_t b = 0; b < t.nblobs; ++b) {
    double e = t.amp[b] * t.axis[0][b * t.n[0] + i];
    e = e * t.axis[1][b * t.n[1] + j];
    e = e * t.axis[2][b * t.n[2] + k];
    v = v + e;
  }
  if (np.octaves > 0) {
    const double x = (double)i / (double)(t.n[0] - 1);
    const double y = (double)j / (double)(t.n[1] - 1);
    const double z = (double)k / (double)(t.n[2] - 1);
    double f = np.freq, a = np.amp;
    for (int32_t o = 0; o < np.octaves; ++o) {
      const double X = x * f + np.phase, Y = y * f + np.phase, Z = z * f + np.phase;
      const double fx = __builtin_floor(X), fy = __builtin_floor(Y), fz = __builtin_floor(Z);
      const double tx = X - fx, ty = Y - fy, tz = Z - fz;
      const double sx = tx * tx * (3.0 - 2.0 * tx);
      const double sy = ty * ty * (3.0 - 2.0 * ty);
      const double sz = tz * tz * (3.0 - 2.0 * tz);
      const int64_t ix = (int64_t)fx, iy = (int64_t)fy, iz = (int64_t)fz;
      const uint64_t salt = np.salt + (uint64_t)o * 0x2545f4914f6cdd1dull;
      double c[8];
      for (int q = 0; q < 8; ++q)
        c[q] = lattice_value(ix + (q & 1), iy + ((q >> 1) & 1), iz + ((q >> 2) & 1), salt);
      const double x00 = c[0] + sx * (c[1] - c[0]);
      const double x10 = c[2] + sx * (c[3] - c[2]);
      const double x01 = c[4] + sx * (c[5] - c[4]);
      const double x11 = c[6] + sx * (c[7] - c[6]);
      const double y0 = x00 + sy * (x10 - x00);
      const double y1 = x01 + sy * (x11 - x01);
      v = v + a * (y0 + sz * (y1 - y0));
      f = f * 2.0;
      a = a * 0.5;
    }
  }
  return v;
}

// Normalised value -> storage.  FP16/BF16 round via FP32 (double -> float -> 16 bit).
template <typename T> BEAMS_HD T store_from_f64(double v);
template <> BEAMS_HD double store_from_f64<double>(double v) { return v; }
template <> BEAMS_HD float store_from_f64<float>(double v) { return (float)v; }

}  // namespace beams
