// Shared host/device helpers for the Beams MI355X renderer.
//
// Everything here is compiled by hipcc for both the x86-64 host and gfx950 with
// -ffp-contract=off, so a function used on both sides produces the same bits on
// both: IEEE add/mul/div/sqrt round to nearest-even on either side, and a fused
// multiply-add appears only where it is written out (fmaf / fma).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define BEAMS_HD __host__ __device__ __forceinline__

namespace beams {

// ---- storage conversions (round to nearest even) -----------------------------------
BEAMS_HD uint32_t f32_bits(float f) { return __builtin_bit_cast(uint32_t, f); }
BEAMS_HD float bits_f32(uint32_t u) { return __builtin_bit_cast(float, u); }

BEAMS_HD uint16_t f32_to_bf16(float f)
{
  uint32_t u = f32_bits(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
BEAMS_HD float bf16_to_f32(uint16_t b) { return bits_f32((uint32_t)b << 16); }

BEAMS_HD uint16_t f32_to_f16(float f) { return __builtin_bit_cast(uint16_t, (_Float16)f); }
BEAMS_HD float f16_to_f32(uint16_t h) { return (float)__builtin_bit_cast(_Float16, h); }

// ---- SplitMix64 stream -> 53-bit uniforms --------------------------------------------
struct SplitMix64 {
  uint64_t s;
  BEAMS_HD explicit SplitMix64(uint64_t seed) : s(seed) {}
  BEAMS_HD uint64_t next()
  {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  BEAMS_HD double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Integer lattice hash for value noise (a 64-bit finaliser over packed coordinates).
BEAMS_HD double lattice_value(int64_t i, int64_t j, int64_t k, uint64_t salt)
{
  uint64_t h = salt ^ 0x632be59bd9b4e019ull;
  h ^= (uint64_t)i * 0x9e3779b97f4a7c15ull;
  h = (h ^ (h >> 29)) * 0xbf58476d1ce4e5b9ull;
  h ^= (uint64_t)j * 0xc2b2ae3d27d4eb4full;
  h = (h ^ (h >> 32)) * 0x94d049bb133111ebull;
  h ^= (uint64_t)k * 0x165667b19e3779f9ull;
  h = (h ^ (h >> 29)) * 0xbf58476d1ce4e5b9ull;
  h ^= h >> 32;
  return (double)(h >> 11) * (2.0 / 9007199254740992.0) - 1.0;  // [-1, 1)
}

}  // namespace beams
