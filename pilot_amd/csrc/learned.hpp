// The learned shadow closure: a coarse optical depth on a 32^3-cell mip of the extinction,
// corrected per map vertex by the network of mlp.hpp, checked against the exact light march
// at a sparse set of vertices.
//
// Extinction: sigma = -ln(1 - a) / h with a = alpha / max alpha of the TF entry and h the
// map step |size| / 128, so that exp(-sigma h) is the transmittance of one step of the
// reference's map march (TransmittanceMap.h:329).  Optical depth tau = -ln(1 - opacity).
//
// Features of a map vertex v (kFeat = 32):
//   [0]       min(tau_c(v), 8) / 4, tau_c trilinear in the coarse optical-depth grid
//   [1..27]   the density patch: -ln(1 - min(a, 1 - 2^-24)) / 4 (the optical depth of one
//             map step, scaled) at the 3 x 3 x 3 field vertices around the field vertex
//             nearest v (index clamped to the grid; dz slowest, dx fastest)
//   [28..30]  the light's direction of travel at v
//   [31]      0
// Target / output y = min(tau_c, 8) - min(tau_exact, 8); the learned opacity is
// 1 - exp(-max(min(tau_c, 8) - y, 0)).  Depths beyond 8 (transmittance < 3.4e-4) are
// clamped: the map's shading weight (1 - opacity) is below any error budget there.
#pragma once

#include "bricks.hpp"
#include "kernels.hpp"

struct beams_mlp;

namespace beams {

constexpr int kCoarseCells = 32;
constexpr float kTauMax = 8.0f;

struct LearnedParams {
  GridParams g;
  MapParams m;         // the learned map (x-fastest layout) and the light
  ShadeParams s;
  float* coarse_sigma;  // cn^3 mean extinction around each coarse vertex, x fastest
  float* coarse_tau;    // cn^3 optical depth from each coarse vertex to the light
  int32_t* coarse_ranges;  // [3][cn][2]: first field index and count nearest each coarse index
  int32_t cn;           // coarse vertices per axis
  float corg[3], csp[3], cinv[3];
  float inv_h;          // 1 / map step
  const float* params;  // network parameters (FP32 packed); null: coarse optical depth only
  unsigned int* err;    // [0] max |learned - exact|, [1] max |coarse only - exact| (float bits)
  int32_t check_stride; // check vertices: every check_stride-th map index per axis + the last
  int32_t field_small;  // the field's storage spans less than 4 GiB (32-bit byte offsets)
};

hipError_t launch_coarse(const LearnedParams& p, hipStream_t st);
hipError_t launch_learned_map(const LearnedParams& p, hipStream_t st);
hipError_t launch_learned_check(const LearnedParams& p, hipStream_t st);
// n features (n x kFeat) and targets at map vertices drawn uniformly (SplitMix64(seed));
// the exact opacities come from `exact` (layout and storage type of em).
hipError_t launch_learned_samples(const LearnedParams& p, const MapParams& em, int64_t n,
                                  uint64_t seed, float* x, float* t, hipStream_t st);
int64_t learned_check_count(const MapParams& m, int32_t stride);

// Bricks: p describes the whole volume (field, coarse grids, light, map step); each brick's
// map_final receives its learned map and err[b] (float bits, zeroed by the launcher) its
// check.  The fallback's work list is built on the device from the checks (launch_learned_redo,
// no host round trip), and the fallback re-marches the listed bricks through the whole volume.
//
// Work list layout (int32, learned_redo_words(nb) words): [0] the largest check error (float
// bits), [1] the number of bricks over the budget, [2] their map items, [3, 4 + count) the first
// item of each listed brick in a contiguous renumbering (count + 1 entries), [4 + nb, 4 + nb +
// count) the listed bricks in ascending order.
inline size_t learned_redo_words(int nb) { return 4 + 2 * (size_t)nb; }
hipError_t launch_bricks_learned_map(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                     hipStream_t st);
hipError_t launch_bricks_learned_check(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                       unsigned int* err, hipStream_t st);
hipError_t launch_learned_redo(const BrickFrame& f, const unsigned int* err, float budget, int32_t* list,
                               hipStream_t st);
hipError_t launch_bricks_exact_map(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                   const int32_t* list, hipStream_t st);

hipError_t launch_mlp_infer(const float* params, const float* x, int64_t n, float* y, hipStream_t st);
hipError_t launch_mlp_grad(const float* params, const float* x, const float* t, int64_t n,
                           float inv_n, float* grads, float* loss, hipStream_t st);
const float* mlp_device_params(const beams_mlp* h);

}  // namespace beams
