// Device functions of the exact (reference-order FP32) path.
//
// Each function follows one reference routine operation by operation; the file
// is compiled with -ffp-contract=off, so products and sums round separately as in
// the reference's CPU build, and fused multiply-adds appear only where the
// reference's own code (lcl's lerp) uses them.
#pragma once

#include "kernels.hpp"

namespace beams {

// ---- storage loads --------------------------------------------------------------------
__device__ __forceinline__ float load_scalar(const void* p, int32_t dtype, int64_t i)
{
  if (dtype == kF16) return f16_to_f32(((const uint16_t*)p)[i]);
  if (dtype == kBF16) return bf16_to_f32(((const uint16_t*)p)[i]);
  return ((const float*)p)[i];
}

template <int DT> __device__ __forceinline__ float load_t(const void* p, int64_t i);
template <> __device__ __forceinline__ float load_t<kF32>(const void* p, int64_t i)
{
  return ((const float*)p)[i];
}
template <> __device__ __forceinline__ float load_t<kF16>(const void* p, int64_t i)
{
  return f16_to_f32(((const uint16_t*)p)[i]);
}
template <> __device__ __forceinline__ float load_t<kBF16>(const void* p, int64_t i)
{
  return bf16_to_f32(((const uint16_t*)p)[i]);
}

template <int DT> __device__ __forceinline__ void store_t(void* p, int64_t i, float v);
template <> __device__ __forceinline__ void store_t<kF32>(void* p, int64_t i, float v)
{
  ((float*)p)[i] = v;
}
template <> __device__ __forceinline__ void store_t<kF16>(void* p, int64_t i, float v)
{
  ((uint16_t*)p)[i] = f32_to_f16(v);
}

// The 8 corner scalars of a field cell (hex order 0:(i,j,k) 1:(i+1,j,k) 2:(i+1,j+1,k)
// 3:(i,j+1,k), 4-7 the same at k+1).  A field view of fewer than 2^32 bytes is read through a
// raw buffer resource: the cell's byte offset is the only per-lane address and the corners'
// offsets from it are uniform scalar operands; larger views take 64-bit addresses.  Every
// lane of a wave reads the same view (the kernels march one grid, or one brick per
// workgroup), which the constructor states to the compiler so that the resource stays in
// scalar registers.
// FAST: the caller guarantees a view below 2^32 bytes and strides below 2^24, so the offset
// takes full-rate 24-bit multiplies and the load has no branch (the conversions of FP16
// corners can then fuse into the interpolation's first multiply-adds).
template <int FDT, bool FAST = false>
struct CornerFetch {
  __amdgpu_buffer_rsrc_t rsrc;
  int so[8];  // corner byte offsets
  int esz;
  const void* field;
  int sx, sy, sz;
  bool w32;
  __device__ __forceinline__ explicit CornerFetch(const void* f, uint32_t bytes32, int sx_, int sy_, int sz_)
  {
    const uint64_t fp = (uint64_t)f;
    const uint32_t plo = __builtin_amdgcn_readfirstlane((uint32_t)fp);
    const uint32_t phi = __builtin_amdgcn_readfirstlane((uint32_t)(fp >> 32));
    field = (const void*)(((uint64_t)phi << 32) | plo);
    bytes32 = __builtin_amdgcn_readfirstlane(bytes32);
    sx = __builtin_amdgcn_readfirstlane(sx_);
    sy = __builtin_amdgcn_readfirstlane(sy_);
    sz = __builtin_amdgcn_readfirstlane(sz_);
    w32 = FAST || bytes32 != 0;
    esz = FDT == kF32 ? 4 : 2;
    rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(field), (short)0, (int)bytes32, 0x00020000);
    // element offsets below 2^31, byte offsets below 2^32 when w32 (unsigned products)
    const uint32_t e = (uint32_t)esz;
    so[0] = 0;
    so[1] = (int)((uint32_t)sx * e);
    so[2] = (int)((uint32_t)(sx + sy) * e);
    so[3] = (int)((uint32_t)sy * e);
    so[4] = (int)((uint32_t)sz * e);
    so[5] = (int)((uint32_t)(sz + sx) * e);
    so[6] = (int)((uint32_t)(sz + sx + sy) * e);
    so[7] = (int)((uint32_t)(sz + sy) * e);
  }
  __device__ __forceinline__ float buf(int voff, int soff) const
  {
    if (FDT == kF32) return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, voff, soff, 0));
    const uint16_t h = __builtin_amdgcn_raw_buffer_load_b16(rsrc, voff, soff, 0);
    if (FAST && FDT == kF16) {
      // a plain conversion: left to itself the compiler folds it into the interpolation's
      // first multiply-adds as v_fma_mix_f32, which measured slower than convert + v_fma_f32
      float r;
      asm("v_cvt_f32_f16 %0, %1" : "=v"(r) : "v"((uint32_t)h));
      return r;
    }
    return FDT == kF16 ? f16_to_f32(h) : bf16_to_f32(h);
  }
  // The raw storage words of the 8 corners (FP32 bits, or FP16/BF16 bits in the low half),
  // for callers that keep loads in flight and decode them later with decode().
  __device__ __forceinline__ void load_raw(int ci, int cj, int ck, uint32_t* w) const
  {
    if (w32) {
      const uint32_t e = FAST ? __umul24((uint32_t)ci, (uint32_t)sx) + __umul24((uint32_t)cj, (uint32_t)sy) +
                                  __umul24((uint32_t)ck, (uint32_t)sz)
                              : (uint32_t)(ci * sx + cj * sy + ck * sz);
      const int off = (int)(e * (uint32_t)(FDT == kF32 ? 4 : 2));
#pragma unroll
      for (int q = 0; q < 8; ++q)
        w[q] = FDT == kF32 ? __builtin_amdgcn_raw_buffer_load_b32(rsrc, off, so[q], 0)
                           : (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(rsrc, off, so[q], 0);
    } else {
      const int64_t b = (int64_t)ci * sx + (int64_t)cj * sy + (int64_t)ck * sz;
      const int64_t d[8] = { 0, sx, sx + sy, sy, sz, sz + sx, sz + sx + sy, sz + sy };
#pragma unroll
      for (int q = 0; q < 8; ++q)
        w[q] = FDT == kF32 ? ((const uint32_t*)field)[b + d[q]] : (uint32_t)((const uint16_t*)field)[b + d[q]];
    }
  }
  __device__ __forceinline__ static float decode(uint32_t w)
  {
    if (FDT == kF32) return __builtin_bit_cast(float, w);
    if (FDT == kBF16) return bf16_to_f32((uint16_t)w);
    float r;
    // a plain conversion (not folded into the interpolation as v_fma_mix_f32)
    asm("v_cvt_f32_f16 %0, %1" : "=v"(r) : "v"(w));
    return r;
  }
  __device__ __forceinline__ void load(int ci, int cj, int ck, float* v) const
  {
    if (FAST) {
      // cell indices below 2^20 (point dims at most 2^20), strides below 2^24
      const uint32_t e = __umul24((uint32_t)ci, (uint32_t)sx) + __umul24((uint32_t)cj, (uint32_t)sy) +
                         __umul24((uint32_t)ck, (uint32_t)sz);
      const int off = (int)(e * (uint32_t)(FDT == kF32 ? 4 : 2));
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = buf(off, so[q]);
    } else if (w32) {
      const int off = (int)((uint32_t)(ci * sx + cj * sy + ck * sz) * (uint32_t)esz);
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = buf(off, so[q]);
    } else {
      const int64_t b = (int64_t)ci * sx + (int64_t)cj * sy + (int64_t)ck * sz;
      v[0] = load_t<FDT>(field, b);
      v[1] = load_t<FDT>(field, b + sx);
      v[2] = load_t<FDT>(field, b + sx + sy);
      v[3] = load_t<FDT>(field, b + sy);
      v[4] = load_t<FDT>(field, b + sz);
      v[5] = load_t<FDT>(field, b + sz + sx);
      v[6] = load_t<FDT>(field, b + sz + sx + sy);
      v[7] = load_t<FDT>(field, b + sz + sy);
    }
  }
};

__device__ __forceinline__ float lerp_lcl(float a, float b, float t)
{
  return __builtin_fmaf(t, b, __builtin_fmaf(-t, a, a));
}

__device__ __forceinline__ float interp_hex(const float* v, const float* pc)
{
  float a = lerp_lcl(v[0], v[1], pc[0]);
  float b = lerp_lcl(v[3], v[2], pc[0]);
  float c = lerp_lcl(v[4], v[5], pc[0]);
  float d = lerp_lcl(v[7], v[6], pc[0]);
  float lo = lerp_lcl(a, b, pc[1]);
  float hi = lerp_lcl(c, d, pc[1]);
  return lerp_lcl(lo, hi, pc[2]);
}

__device__ __forceinline__ float dot3(const float* a, const float* b)
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Intersections.cxx:8-49 (the slab terms promote to FP64 through vtkm::Bounds).
__device__ __forceinline__ bool segment_aabb(const float* s, const float* e, const double* bb,
                                             float teps, float& tmin_o, float& tmax_o)
{
  float d[3] = { e[0] - s[0], e[1] - s[1], e[2] - s[2] };
  float len = __builtin_sqrtf(dot3(d, d));
  d[0] = d[0] / len;
  d[1] = d[1] / len;
  d[2] = d[2] / len;
  float tmin = 0.0f, tmax = __builtin_inff();
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float t0 = (float)((bb[2 * a] - (double)s[a]) / (double)d[a]);
    float t1 = (float)((bb[2 * a + 1] - (double)s[a]) / (double)d[a]);
    tmin = fmaxf(tmin, fminf(t0, t1));
    tmax = fminf(tmax, fmaxf(t0, t1));
  }
  if (tmin < 0.0f) tmin = teps;
  if (tmax > len) tmax = len;
  tmin_o = tmin;
  tmax_o = tmax;
  return tmin <= tmax;
}

__device__ __forceinline__ bool inside(const GridParams& g, const float* p)
{
  bool in = true;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if (p[d] < g.minp[d] || p[d] > g.maxp[d]) in = false;
  return in;
}

// RectilinearMeshOracle.h:154-176: smallest c in [0, C-1] with v <= x[c+1].  The binary
// search is replaced by a uniform-spacing guess corrected against the coordinates, which
// yields the same unique cell for monotone coordinates.
__device__ __forceinline__ int64_t oracle_cell(const float* x, int64_t ncell, float v, float x0,
                                               float ginv)
{
  float gf = (v - x0) * ginv;
  int64_t c = (gf > 0.0f) ? (int64_t)gf : 0;
  c = c < ncell - 1 ? c : ncell - 1;
  while (c > 0 && v <= x[c]) --c;
  while (c < ncell - 1 && v > x[c + 1]) ++c;
  return c;
}

// The same cell as oracle_cell in 32-bit indices (clamping the guess before truncating it
// gives the same index): for per-step use where the axis has fewer than 2^31 points.
__device__ __forceinline__ int oracle_cell32(const float* x, int ncell, float v, float x0, float ginv)
{
  const float gf = (v - x0) * ginv;
  int c = (gf > 0.0f) ? (int)fminf(gf, (float)(ncell - 1)) : 0;
  while (c > 0 && v <= x[c]) --c;
  while (c < ncell - 1 && v > x[c + 1]) ++c;
  return c;
}

// a / d in FP64 from y = 1 / d rounded to FP64, corrected as div_by_recip (the same theorem
// at 53 bits); zero or tiny numerators and quotients, and d = 0 (fast false), divide.
__device__ __forceinline__ double div64_by_recip(double a, double d, double y, bool fast)
{
  const double q = a * y;
  if (__builtin_expect(fast && __builtin_fabs(a) >= 0x1p-900 && __builtin_fabs(q) >= 0x1p-900, 1))
    return __builtin_fma(__builtin_fma(-q, d, a), y, q);
  return a / d;
}

// a / b from y = 1 / b rounded to FP32: q = a y corrected by one residual step,
// q + (a - q b) y with fused products (Markstein's theorem: y within half an ulp of 1 / b and
// q within one ulp of a / b make the corrected q the correctly rounded quotient).  The
// theorem needs y normal and the residual a - q b representable: widths b lie in
// [2^-72, 2^61] (set_field keeps field cells in [2^-60, 2^61]; map cells split the field's
// extent at most 4096 ways), the quotients here are parametric coordinates near [0, 1], and
// a first quotient of at least 2^-28 then puts |a| above 2^-100, where the residual's
// granularity is above the subnormal step.  Smaller quotients (zeros among them) divide.
__device__ __forceinline__ float div_by_recip(float a, float b, float y)
{
  const float q = a * y;
  if (__builtin_expect(__builtin_fabsf(q) >= 0x1p-28f, 1))
    return __builtin_fmaf(__builtin_fmaf(-q, b, a), y, q);
  return a / b;
}

// segment_aabb for one segment against many boxes: the direction, length and the FP64
// reciprocals of the direction are computed once (segment_prepare); segment_box then gives
// segment_aabb's result for each box with the slab quotients from those reciprocals.
struct Segment {
  float s[3], d[3], len;
  double rd[3];  // 1 / (double)d
};
__device__ __forceinline__ void segment_prepare(const float* s, const float* e, Segment& sg)
{
  float d[3] = { e[0] - s[0], e[1] - s[1], e[2] - s[2] };
  sg.len = __builtin_sqrtf(dot3(d, d));
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    sg.s[a] = s[a];
    sg.d[a] = d[a] / sg.len;
    sg.rd[a] = 1.0 / (double)sg.d[a];
  }
}
__device__ __forceinline__ bool segment_box(const Segment& sg, const double* bb, float teps,
                                            float& tmin_o, float& tmax_o)
{
  float tmin = 0.0f, tmax = __builtin_inff();
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double da = (double)sg.d[a];
    const bool fast = sg.d[a] != 0.0f;
    float t0 = (float)div64_by_recip(bb[2 * a] - (double)sg.s[a], da, sg.rd[a], fast);
    float t1 = (float)div64_by_recip(bb[2 * a + 1] - (double)sg.s[a], da, sg.rd[a], fast);
    tmin = fmaxf(tmin, fminf(t0, t1));
    tmax = fminf(tmax, fmaxf(t0, t1));
  }
  if (tmin < 0.0f) tmin = teps;
  if (tmax > sg.len) tmax = sg.len;
  tmin_o = tmin;
  tmax_o = tmax;
  return tmin <= tmax;
}

// One axis's cell of a point, with its bounding coordinates and reciprocal width
// (x[rofs + c]): axis_guess reads the uniform-spacing guess's cell, oracle_settle /
// walk_settle keep it when the point lies in it (a walk from there would not move) and walk
// otherwise.  The three axes' guesses are read before any is checked.
struct AxisCell {
  int c;
  float lo, hi, rw;
};
// BINS: the guess comes from the axis's bin table (int entries at x + 2 rofs, staged with the
// coordinates) and ginv is the bins per unit length; the walk that follows corrects any guess.
template <bool BINS = false>
__device__ __forceinline__ AxisCell axis_guess(const float* x, int rofs, int ncell, float v,
                                               float x0, float ginv)
{
  AxisCell a;
  // clamp (med3; v is finite inside the bounds, so the guess is never NaN)
  a.c = (int)__builtin_amdgcn_fmed3f((v - x0) * ginv, 0.0f, (float)(ncell - 1));
  if (BINS) {
    // the bin's first cell, then up to two cells on without a branch: a bin spans at most
    // three cells where widths stay within twice the mean; the callers' checks settle the rest
    a.c = min(max(((const int*)(x + 2 * rofs))[a.c], 0), ncell - 1);
    a.c += (a.c < ncell - 1 && v > x[a.c + 1]) ? 1 : 0;
    a.c += (a.c < ncell - 1 && v > x[a.c + 1]) ? 1 : 0;
  }
  a.lo = x[a.c];
  a.hi = x[a.c + 1];
  a.rw = x[rofs + a.c];
  return a;
}

// oracle_cell32's cell: v in (x[c], x[c+1]] (lower cell on an interior plane).
__device__ __forceinline__ bool oracle_in(const AxisCell& a, float v) { return v > a.lo && v <= a.hi; }
__device__ __forceinline__ void oracle_settle(const float* x, int rofs, int ncell, float v, AxisCell& a)
{
  int c = a.c;
  while (c > 0 && v <= x[c]) --c;
  while (c < ncell - 1 && v > x[c + 1]) ++c;
  a.c = c;
  a.lo = x[c];
  a.hi = x[c + 1];
  a.rw = x[rofs + c];
}

// walk_cell32's cell (the camera march): v in [x[c], x[c+1]).
template <bool BINS = false>
__device__ __forceinline__ AxisCell walk_axis(const float* x, int rofs, int ncell, float v,
                                              float x0, float ginv)
{
  AxisCell a = axis_guess<BINS>(x, rofs, ncell, v, x0, ginv);
  if (__builtin_expect(!(v >= a.lo && v < a.hi), 0)) {
    int c = a.c;
    while (c > 0 && v < x[c]) --c;
    while (c < ncell - 1 && v >= x[c + 1]) ++c;
    a.c = c;
    a.lo = x[c];
    a.hi = x[c + 1];
    a.rw = x[rofs + c];
  }
  return a;
}

// LightedVolumeRenderer.cxx:123-164: unique c with x[c] <= v < x[c+1] (v != max).
__device__ __forceinline__ int64_t walk_cell(const float* x, int64_t ncell, float v, float x0,
                                             float ginv)
{
  float gf = (v - x0) * ginv;
  int64_t c = (gf > 0.0f) ? (int64_t)gf : 0;
  c = c < ncell - 1 ? c : ncell - 1;
  while (c > 0 && v < x[c]) --c;
  while (c < ncell - 1 && v >= x[c + 1]) ++c;
  return c;
}

// The map cell of the previous lookup and its corner opacities: consecutive samples of a ray
// usually share a map cell, whose corners need not be gathered again.  Same result as
// map_opacity (the values are the same loads; the interpolation is recomputed each time).
// Corner fetch from a 4 x 4 x 4-blocked copy of a grid (64 values per block, a 128-byte line
// for FP16): point (i, j, k) sits at element
//     X(i) + Y(j) + Z(k),  X(i) = 64 (i >> 2) + (i & 3),  Y(j) = ybs (j >> 2) + 4 (j & 3),
//     Z(k) = zbs (k >> 2) + 16 (k & 3),
// with ybs = 64 x-blocks and zbs = 64 x-blocks y-blocks elements.  A wave's rays sample a few
// voxels across in y and z, which the x-fastest layout spreads over a line per (y, z) row and
// the blocked one keeps within a few lines.  The copy is below 2^32 bytes (one buffer
// resource, 32-bit offsets); ybs and zbs come in the stride arguments.
template <int FDT>
struct BlockFetch {
  __amdgpu_buffer_rsrc_t rsrc;
  uint32_t ybs, zbs;
  __device__ __forceinline__ explicit BlockFetch(const void* f, uint32_t bytes32, int, int ybs_, int zbs_)
  {
    const uint64_t fp = (uint64_t)f;
    const uint32_t plo = __builtin_amdgcn_readfirstlane((uint32_t)fp);
    const uint32_t phi = __builtin_amdgcn_readfirstlane((uint32_t)(fp >> 32));
    const void* field = (const void*)(((uint64_t)phi << 32) | plo);
    bytes32 = __builtin_amdgcn_readfirstlane(bytes32);
    ybs = __builtin_amdgcn_readfirstlane((uint32_t)ybs_);
    zbs = __builtin_amdgcn_readfirstlane((uint32_t)zbs_);
    rsrc = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(field), (short)0, (int)bytes32, 0x00020000);
  }
  __device__ __forceinline__ float buf(uint32_t e) const
  {
    const int off = (int)(e * (uint32_t)(FDT == kF32 ? 4 : 2));
    if (FDT == kF32) return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, off, 0, 0));
    const uint16_t h = __builtin_amdgcn_raw_buffer_load_b16(rsrc, off, 0, 0);
    if (FDT == kF16) {
      float r;
      asm("v_cvt_f32_f16 %0, %1" : "=v"(r) : "v"((uint32_t)h));
      return r;
    }
    return bf16_to_f32(h);
  }
  __device__ __forceinline__ void load(int ci, int cj, int ck, float* v) const
  {
    const uint32_t i = (uint32_t)ci, j = (uint32_t)cj, k = (uint32_t)ck;
    const uint32_t base = ((i >> 2) << 6) + (i & 3u) + __umul24(j >> 2, ybs) + ((j & 3u) << 2) +
                          __umul24(k >> 2, zbs) + ((k & 3u) << 4);
    const uint32_t dx = (i & 3u) == 3u ? 61u : 1u;            // the next x: the next block's first
    const uint32_t dy = (j & 3u) == 3u ? ybs - 12u : 4u;
    const uint32_t dz = (k & 3u) == 3u ? zbs - 48u : 16u;
    v[0] = buf(base);
    v[1] = buf(base + dx);
    v[2] = buf(base + dx + dy);
    v[3] = buf(base + dy);
    v[4] = buf(base + dz);
    v[5] = buf(base + dz + dx);
    v[6] = buf(base + dz + dx + dy);
    v[7] = buf(base + dz + dy);
  }
};

struct MapCellCache {
  int c[3] = { -1, -1, -1 };
  float lo[3], w[3], rw[3];  // the cell's lower coordinates, widths and reciprocal widths
  float v[8];
};

// map_opacity with a cache, in two parts: map_cell_update finds the map cell of p and, when it
// is not the cached one, issues the corner loads (the camera march calls it before the field's
// own loads, so both sets are in flight together); map_cell_opacity interpolates in the cell.
// Same result as map_opacity: the same corner values and parametric coordinates.
template <int MDT, class Fetch>
__device__ __forceinline__ void map_cell_update(const MapParams& m, const Fetch& mf, const float* p,
                                                MapCellCache& mc)
{
  int cell[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int np = (int)m.mp[d];
    float t = (p[d] - m.morg[d]) * m.minv[d];
    if (t < 0.0f) t = 0.0f;
    if (t >= (float)(np - 1)) t = (float)(np - 2);
    cell[d] = (int)t;
  }
  if (cell[0] != mc.c[0] || cell[1] != mc.c[1] || cell[2] != mc.c[2]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float lo = m.morg[d] + m.msp[d] * (float)cell[d];
      const float hi = m.morg[d] + m.msp[d] * (float)(cell[d] + 1);
      mc.lo[d] = lo;
      mc.w[d] = hi - lo;
      mc.rw[d] = 1.0f / mc.w[d];
    }
    mf.load(cell[0], cell[1], cell[2], mc.v);
    mc.c[0] = cell[0];
    mc.c[1] = cell[1];
    mc.c[2] = cell[2];
  }
}

__device__ __forceinline__ float map_cell_opacity(const float* p, const MapCellCache& mc)
{
  float pc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) pc[d] = div_by_recip(p[d] - mc.lo[d], mc.w[d], mc.rw[d]);
  return interp_hex(mc.v, pc);
}

// walk_cell in 32-bit indices.
__device__ __forceinline__ int walk_cell32(const float* x, int ncell, float v, float x0, float ginv)
{
  const float gf = (v - x0) * ginv;
  int c = (gf > 0.0f) ? (int)fminf(gf, (float)(ncell - 1)) : 0;
  while (c > 0 && v < x[c]) --c;
  while (c < ncell - 1 && v >= x[c + 1]) ++c;
  return c;
}

// The reference's TF index (truncate toward zero, then clamp to the table): clamping in float
// first gives the same index for every finite product and keeps it 32-bit.
__device__ __forceinline__ int tf_clamped_index(float t, float tfmax)
{
  // med3 clamps a finite t to [0, tfmax] in one instruction; a NaN t gives 0, as the
  // fmax/fmin form does
  return (int)__builtin_amdgcn_fmed3f(t, 0.0f, tfmax);
}

// TransmittanceMap.h:93-112, 166-229: trilinear map opacity at p.
template <int MDT>
__device__ __forceinline__ float map_opacity(const MapParams& m, const float* p)
{
  // 32-bit cells and conversions (a map axis has at most 4097 points; the same float values
  // as 64-bit integers would give)
  float pc[3];
  int cell[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int np = (int)m.mp[d];
    float t = (p[d] - m.morg[d]) * m.minv[d];
    if (t < 0.0f) t = 0.0f;
    if (t >= (float)(np - 1)) t = (float)(np - 2);
    cell[d] = (int)t;
    float lo = m.morg[d] + m.msp[d] * (float)cell[d];
    float hi = m.morg[d] + m.msp[d] * (float)(cell[d] + 1);
    pc[d] = (p[d] - lo) / (hi - lo);
  }
  const int64_t s0 = m.ms[0], sx = m.ms[1], sxy = m.ms[2];
  const int64_t i0 = (int64_t)cell[2] * sxy + (int64_t)cell[1] * sx + (int64_t)cell[0] * s0;
  float v[8];
  v[0] = load_t<MDT>(m.map, i0);
  v[1] = load_t<MDT>(m.map, i0 + s0);
  v[2] = load_t<MDT>(m.map, i0 + s0 + sx);
  v[3] = load_t<MDT>(m.map, i0 + sx);
  v[4] = load_t<MDT>(m.map, i0 + sxy);
  v[5] = load_t<MDT>(m.map, i0 + sxy + s0);
  v[6] = load_t<MDT>(m.map, i0 + sxy + s0 + sx);
  v[7] = load_t<MDT>(m.map, i0 + sxy + sx);
  return interp_hex(v, pc);
}

// map_opacity for a map every lane of the wave reads (m in scalar registers): the same cell,
// parametric coordinates and corner values, the corners read through a buffer resource with
// 32-bit offsets when the map is below 2^32 bytes.
template <int MDT>
__device__ __forceinline__ float map_opacity_uniform(const MapParams& m, const float* p)
{
  if (m.bytes32 == 0) return map_opacity<MDT>(m, p);
  float pc[3];
  int cell[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int np = (int)m.mp[d];
    float t = (p[d] - m.morg[d]) * m.minv[d];
    if (t < 0.0f) t = 0.0f;
    if (t >= (float)(np - 1)) t = (float)(np - 2);
    cell[d] = (int)t;
    float lo = m.morg[d] + m.msp[d] * (float)cell[d];
    float hi = m.morg[d] + m.msp[d] * (float)(cell[d] + 1);
    pc[d] = (p[d] - lo) / (hi - lo);
  }
  const CornerFetch<MDT> mf(m.map, m.bytes32, (int)m.ms[0], (int)m.ms[1], (int)m.ms[2]);
  float v[8];
  mf.load(cell[0], cell[1], cell[2], v);
  return interp_hex(v, pc);
}

// ArrayPortalUniformPointCoordinates::Get on the map grid.
__device__ __forceinline__ void map_vertex(const MapParams& m, int64_t id, float* v)
{
  const int64_t i = id % m.mp[0];
  const int64_t j = (id / m.mp[0]) % m.mp[1];
  const int64_t k = id / (m.mp[0] * m.mp[1]);
  v[0] = m.morg[0] + m.msp[0] * (float)i;
  v[1] = m.morg[1] + m.msp[1] * (float)j;
  v[2] = m.morg[2] + m.msp[2] * (float)k;
}

// map_vertex for a vertex id below 2^31: the same indices from 32-bit divisions (the 64-bit
// ones are long emulated sequences), so the same position.
__device__ __forceinline__ void map_vertex32(const MapParams& m, uint32_t id, float* v)
{
  const uint32_t mp0 = (uint32_t)m.mp[0], mp1 = (uint32_t)m.mp[1];
  const uint32_t row = id / mp0, i = id - row * mp0;
  const uint32_t k = row / mp1, j = row - k * mp1;
  v[0] = m.morg[0] + m.msp[0] * (float)i;
  v[1] = m.morg[1] + m.msp[1] * (float)j;
  v[2] = m.morg[2] + m.msp[2] * (float)k;
}

__device__ __forceinline__ void light_origin(const MapParams& m, const float* v, float* o)
{
  if (m.light_kind == 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) o[d] = v[d] - m.light[d] * m.dir_far;
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) o[d] = m.light[d];
  }
}

// LightedVolumeRenderer.cxx:666-717
__device__ __forceinline__ float rcp_safe(float f)
{
  return 1.0f / ((__builtin_fabsf(f) < 1e-8f) ? 1e-8f : f);
}

__device__ __forceinline__ void calc_ray_start(const GridParams& g, const float* org,
                                               const float* dir, float& minD, float& dist,
                                               float& maxD)
{
  const float xminb = (float)g.bounds[0], xmaxb = (float)g.bounds[1];
  const float yminb = (float)g.bounds[2], ymaxb = (float)g.bounds[3];
  const float zminb = (float)g.bounds[4], zmaxb = (float)g.bounds[5];
  const float ix = rcp_safe(dir[0]), iy = rcp_safe(dir[1]), iz = rcp_safe(dir[2]);
  const float ox = org[0] * ix, oy = org[1] * iy, oz = org[2] * iz;
  const float xmin = xminb * ix - ox, ymin = yminb * iy - oy, zmin = zminb * iz - oz;
  const float xmax = xmaxb * ix - ox, ymax = ymaxb * iy - oy, zmax = zmaxb * iz - oz;
  minD = fmaxf(fmaxf(fmaxf(fminf(ymin, ymax), fminf(xmin, xmax)), fminf(zmin, zmax)), minD);
  const float ex = fminf(fminf(fmaxf(ymin, ymax), fmaxf(xmin, xmax)), fmaxf(zmin, zmax));
  maxD = fminf(maxD, ex);
  if (maxD < minD)
    minD = -1.0f;
  else
    dist = minD;
}

}  // namespace beams
