// Internal interface between the C-ABI layer (abi.hip) and the HIP kernels.
// Parameter blocks are plain structs passed by value to the kernels.
#pragma once

#include "common.hpp"

namespace beams {

enum Dtype : int32_t { kF32 = 0, kF16 = 1, kBF16 = 2, kF64 = 3 };

// One block's structured grid: rectilinear coordinates + scalars.
struct GridParams {
  const void* field;   // scalars, x fastest
  int32_t dtype;       // kF32 / kF16 / kBF16
  int32_t assoc;       // 0 points, 1 cells
  int64_t n[3];        // point dims
  int64_t field_stride[3];  // element strides of the stored array (x, y, z)
  const float* x[3];   // coordinate arrays (device)
  const float* r[3];   // reciprocal cell widths 1 / (x[c+1] - x[c]) in FP32 (device; r[n-1] = 0)
  float minp[3], maxp[3];
  float gx0[3], ginv[3];    // cell-guess helpers: x[0] and (n-1)/(x[n-1]-x[0])
  double bounds[6];         // FP64 bounds of the coordinates
  float ext_mag;            // |extent| in FP32 (BoundsMagnitude)
  uint32_t bytes32;         // bytes from field to the end of its array when below 2^32, else 0
  // Uneven coordinates: a bin table per axis (n-1 uniform bins over the axis, each holding the
  // cell its lower edge lies in) follows the reciprocals in the coordinate array, at the
  // coordinates' own offsets plus 2 (n0+n1+n2); binv = bins per unit length.
  int32_t use_bins;
  float binv[3];
};

// Transmittance (opacity) map over the block bounds.
struct MapParams {
  void* map;           // (S+1)^3 opacities
  int32_t dtype;       // kF32 / kF16
  int32_t light_kind;  // 0 point, 1 directional
  int64_t mp[3];       // map point dims
  int64_t ms[3];       // element strides of the map layout (x fastest: 1, mp0, mp0*mp1)
  float morg[3], msp[3], minv[3];
  float step;          // |size| / 128
  float light[3];      // position (point) or travel direction (directional)
  float light_color[3];
  float dir_far;       // directional-light origin distance
  uint32_t bytes32;    // bytes of the map array when below 2^32, else 0
  // Optional gate of a fallback build: when set, the build runs only if the float stored (as
  // bits) at *gate, a check's largest error, exceeds gate_budget (a NaN error exceeds it).
  const unsigned int* gate;
  float gate_budget;
};

// True when a gated build is to be skipped (uniform: every lane reads the same word, which an
// earlier kernel wrote).
__device__ __forceinline__ bool gate_closed(const MapParams& m)
{
  if (!m.gate) return false;
  const float e = __uint_as_float((unsigned)__builtin_amdgcn_readfirstlane((int)*m.gate));
  return e <= m.gate_budget;
}

// Bytes of a map array of mp[0] mp[1] mp[2] values of dtype when below 2^32, else 0.
inline uint32_t map_bytes32(const int64_t* mp, int32_t dtype)
{
  const uint64_t b = (uint64_t)mp[0] * (uint64_t)mp[1] * (uint64_t)mp[2] * (dtype == kF32 ? 4u : 2u);
  return b < ((uint64_t)1 << 32) ? (uint32_t)b : 0u;
}

// Host/device: the element count of a 4 x 4 x 4-blocked copy of n points per axis, and its
// block strides (ybs, zbs).
BEAMS_HD int64_t blocked_count(const int64_t* n)
{
  return ((n[0] + 3) / 4) * ((n[1] + 3) / 4) * ((n[2] + 3) / 4) * 64;
}
BEAMS_HD void blocked_strides(const int64_t* n, int64_t& ybs, int64_t& zbs)
{
  ybs = ((n[0] + 3) / 4) * 64;
  zbs = ybs * ((n[1] + 3) / 4);
}
BEAMS_HD int64_t blocked_index(int64_t i, int64_t j, int64_t k, int64_t ybs, int64_t zbs)
{
  return (i >> 2) * 64 + (i & 3) + (j >> 2) * ybs + (j & 3) * 4 + (k >> 2) * zbs + (k & 3) * 16;
}

struct ShadeParams {
  const float* tf;     // ntf x RGBA (device)
  int32_t ntf;
  float max_density;   // max alpha of the TF
  float min_scalar, inv_delta;
  float sample_dist, mesh_eps, alpha_stop;
  int32_t use_map;
  int32_t alpha_finite;  // every alpha / max alpha is finite (a march seeded at 1 stays at 1)
  int64_t max_steps;   // bound on samples per camera ray (2 diagonals / step + 64)
  unsigned long long* counters;  // optional: [0] composited camera-ray samples, [1] light steps
};

// Bound on light-march steps per map vertex: a segment inside the bounds is at most one
// diagonal, i.e. 128 steps of |size| / 128.
constexpr int64_t kMapMaxSteps = 1024;

struct RaySoA {
  const float* org[3];
  const float* dir[3];
  float* min_d;
  float* max_d;
  float* dist;
  float* rgba;
  int64_t n;
};

struct CameraParams {
  double eye[3];
  double fwd[3], right[3], up[3];
  double tan_half_fov, aspect;
  int32_t width, height;
  float near_d, far_d;
};

struct ImageOut {
  float* rgba;   // W*H*4
  float* depth;  // W*H (may be null)
  float bg[4];   // background blended under the image when use_bg
  int32_t use_bg;
};

// Instantiate BODY for the run-time field (FDT) or map (MDT) storage type; unknown types
// return hipErrorInvalidValue from the enclosing launcher.
#define BEAMS_DISPATCH_FIELD(FDT_VAR, BODY)          \
  switch (FDT_VAR) {                                 \
    case kF32: { constexpr int FDT = kF32; BODY; } break;   \
    case kF16: { constexpr int FDT = kF16; BODY; } break;   \
    case kBF16: { constexpr int FDT = kBF16; BODY; } break; \
    default: return hipErrorInvalidValue;            \
  }

#define BEAMS_DISPATCH_MAP(MDT_VAR, BODY)            \
  switch (MDT_VAR) {                                 \
    case kF32: { constexpr int MDT = kF32; BODY; } break;   \
    case kF16: { constexpr int MDT = kF16; BODY; } break;   \
    default: return hipErrorInvalidValue;            \
  }

// ---- launchers (return hipError_t) ------------------------------------------------------
hipError_t launch_tmap_ref(const GridParams& g, const MapParams& m, const ShadeParams& s,
                           const float* seed, hipStream_t st);
hipError_t launch_march_rays(const GridParams& g, const MapParams& m, const ShadeParams& s,
                             const RaySoA& r, hipStream_t st);
hipError_t launch_march_image(const GridParams& g, const MapParams& m, const ShadeParams& s,
                              const CameraParams& c, const ImageOut& o, hipStream_t st);

hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t st);

// field generation on the device
struct FieldGenParams;
hipError_t launch_field_minmax(const void* tables, int64_t n0, int64_t n1, int64_t n2,
                               double* block_minmax, int32_t nblocks, hipStream_t st);

}  // namespace beams
