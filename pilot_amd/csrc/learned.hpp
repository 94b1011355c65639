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
  float* verr;          // optional: |learned - exact| per check vertex (the check tiles' input)
  const uint8_t* tfidx; // optional: the TF entry of every field point (dense field, at most 256
                        // entries), written this frame by launch_tf_index_field
  int32_t check_stride; // check vertices: every check_stride-th map index per axis + the last
  int32_t field_small;  // the field's storage spans less than 4 GiB (32-bit byte offsets)
};

hipError_t launch_coarse(const LearnedParams& p, hipStream_t st);
hipError_t launch_learned_map(const LearnedParams& p, hipStream_t st);
// The TF entry of every point of p.g's field into out (one byte each): the feature patches read
// it instead of the field.  Refused (hipErrorInvalidValue) unless the field is dense, below 2^32
// points and 16-byte aligned (as out), and the TF has at most 256 entries.
hipError_t launch_tf_index_field(const LearnedParams& p, uint8_t* out, hipStream_t st);
// Whether launch_tf_index_field applies to p (for a 16-byte aligned out).
inline bool tf_index_field_ok(const LearnedParams& p)
{
  const int64_t n = p.g.n[0] * p.g.n[1] * p.g.n[2];
  return p.s.ntf <= 256 && p.g.field_stride[0] == 1 && p.g.field_stride[1] == p.g.n[0] &&
         p.g.field_stride[2] == p.g.n[0] * p.g.n[1] && n < ((int64_t)1 << 32) && ((uintptr_t)p.g.field & 15) == 0;
}
hipError_t launch_learned_check(const LearnedParams& p, hipStream_t st);
// n features (n x kFeat) and targets at map vertices drawn uniformly (SplitMix64(seed));
// the exact opacities come from `exact` (layout and storage type of em).
hipError_t launch_learned_samples(const LearnedParams& p, const MapParams& em, int64_t n,
                                  uint64_t seed, float* x, float* t, hipStream_t st);
int64_t learned_check_count(const MapParams& m, int32_t stride);

// Bricks: p describes the whole volume (field, coarse grids, light, map step), and each
// brick's map_final receives its learned map.  The check vertices of a brick (every stride-th
// map vertex per axis, and the last) cut its map into tiles of up to stride^3 cells: tile t
// along an axis holds the vertices [t s, (t + 1) s) and the last tile also the last vertex.
// The check writes |learned - exact| per check vertex; a tile is listed on the device when the
// largest error over its own and its 26 neighbours' corners exceeds the budget (a NaN exceeds
// it), and the fallback re-marches the listed tiles' vertices through the whole volume.  No host round trip: the launches
// can be captured in one graph.
struct LearnedBricks {
  const int32_t* chk_off;   // nbricks + 1: first check slot of each brick (counts padded to 256)
  const int32_t* tile_off;  // nbricks + 1: first tile of each brick
  float* verr;              // per check slot: |learned - exact|
  int32_t* list;            // [0] tiles listed, [1] largest check error (float bits), [2, ...) the tiles
  uint8_t* tile_fail;       // optional, per tile: 1 when the tile is listed (re-marched), else 0
  int32_t n_chk_items;      // chk_off[nbricks] / 256
  int32_t n_tiles;          // tile_off[nbricks]
  int32_t stride;
  float budget;
};
// Check vertices along an axis of `cells` cells: every stride-th vertex and the last.
BEAMS_HD int64_t check_count(int64_t cells, int32_t stride)
{
  return cells / stride + 1 + (cells % stride != 0 ? 1 : 0);
}
// at_checks: the learned value at every check vertex of lb only (before the check); else at
// every map vertex, skipping the waves whose vertices all lie in tiles lb.tile_fail marks
// (the fallback re-marches them), when lb and its flags are given.
hipError_t launch_bricks_learned_map(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                     const LearnedBricks* lb, bool at_checks, hipStream_t st);
// Zeroes the list's two counters, then the check and the tile listing.
hipError_t launch_bricks_learned_check(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                       const LearnedBricks& lb, hipStream_t st);
hipError_t launch_bricks_exact_tiles(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                     const LearnedBricks& lb, hipStream_t st);
// The tile listing alone (the list's counters zeroed by the caller).  A single block uses it,
// and the fallback above, as a frame of one brick whose map is the whole volume's.
hipError_t launch_learned_tiles(const BrickFrame& f, const LearnedBricks& lb, hipStream_t st);

hipError_t launch_mlp_infer(const float* params, const float* x, int64_t n, float* y, hipStream_t st);
hipError_t launch_mlp_grad(const float* params, const float* x, const float* t, int64_t n,
                           float inv_n, float* grads, float* loss, hipStream_t st);
const float* mlp_device_params(const beams_mlp* h);

}  // namespace beams
