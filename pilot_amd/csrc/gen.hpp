// Host-side interface of the synthetic-input generators (gen.hip).
#pragma once

#include "field_gen.hpp"
#include "kernels.hpp"

#include <string>
#include <vector>

namespace beams {

struct FieldRecipe {
  uint64_t seed;
  int32_t nblobs;
  int32_t noise_octaves;
  double noise_amp;
  double noise_freq;
  double time;
  double drift_speed;
  double noise_phase_rate;
};

void blob_tables_host(const FieldRecipe& r, const int64_t dims[3], std::vector<double>& amp,
                      std::vector<double> axis[3], NoiseParams& np);

// Returns 0, or a BEAMS_ERR_* code with err set.
int gen_field(const FieldRecipe& r, const int64_t dims[3], int32_t dtype, void* out, bool device,
              hipStream_t st, std::string& err);

void gen_colormap(uint64_t seed, int32_t npoints, int64_t n, float* rgba32, double* rgba64);

// Seeded kd decomposition of a dims grid into nbricks boxes whose cell extents lie in
// [min_cells, max_cells]; ranges receives 6 inclusive point-index bounds per brick.
int gen_kd_bricks(uint64_t seed, const int64_t dims[3], int32_t nbricks, int64_t min_cells,
                  int64_t max_cells, int64_t* ranges, std::string& err);

}  // namespace beams
