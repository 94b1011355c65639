// Device routines of the exact (reference-order FP32) path shared by the single-block
// kernels (march.hip) and the brick kernels (bricks.hip).
#pragma once

#include <type_traits>

#include "march_core.hpp"

namespace beams {

// Load the block's rectilinear coordinates into LDS, xs[0..nc) with nc = n0 + n1 + n2,
// followed by the reciprocal cell widths, xs[nc..2 nc) (coords_rofs(g) past each coordinate).
__device__ __forceinline__ void stage_coords(const GridParams& g, float* xs)
{
  const int64_t n0 = g.n[0], n1 = g.n[1], n2 = g.n[2], nc = n0 + n1 + n2;
  for (int64_t i = threadIdx.x; i < nc; i += blockDim.x) {
    float v, w;
    if (i < n0) {
      v = g.x[0][i];
      w = g.r[0][i];
    } else if (i < n0 + n1) {
      v = g.x[1][i - n0];
      w = g.r[1][i - n0];
    } else {
      v = g.x[2][i - n0 - n1];
      w = g.r[2][i - n0 - n1];
    }
    xs[i] = v;
    xs[nc + i] = w;
  }
}

__device__ __forceinline__ int coords_rofs(const GridParams& g) { return (int)(g.n[0] + g.n[1] + g.n[2]); }

// The bin tables of uneven coordinates (GridParams::use_bins) into LDS after the reciprocals:
// xs[2 nc .. 3 nc) as ints, at the coordinates' offsets.
__device__ __forceinline__ void stage_bins(const GridParams& g, float* xs)
{
  const int64_t nc = g.n[0] + g.n[1] + g.n[2];
  const int* src = (const int*)(g.x[0] + 2 * nc);
  int* dst = (int*)(xs + 2 * nc);
  for (int64_t i = threadIdx.x; i < nc; i += blockDim.x) dst[i] = src[i];
}

// Adds the counts of the lanes active here to *ctr with one atomic per wave: the active lanes'
// values are summed through readlane (a wave-uniform loop over the ballot), then the first
// active lane adds the sum.  Per-lane atomics on one address serialise in L2 (about 10 ns
// each: seconds for a 10^9-vertex map).
__device__ __forceinline__ void count_add(unsigned long long* ctr, uint32_t v)
{
  const unsigned long long active = __ballot(1);
  unsigned long long mask = active, sum = 0;
  while (mask) {
    const int lane = __builtin_ctzll(mask);
    sum += (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
    mask &= mask - 1;
  }
  if ((int)__lane_id() == __builtin_ctzll(active) && sum) atomicAdd(ctr, sum);
}

// ---- TransmittanceMap.h:279-331: one map vertex's light march ----------------------------
// a_tab: alpha / max alpha per TF entry; x: the block's coordinates (LDS).  Returns the
// opacity; *taken receives the number of steps inside the volume.
// march_vertex_at marches from the map vertex at v (map_vertex's position of the vertex).
// FAST: the field view and strides suit CornerFetch<FDT, true>.
template <int FDT, bool BINS = false, bool FAST = false>
__device__ __forceinline__ float march_vertex_at(const GridParams& g, const MapParams& m,
                                                 const ShadeParams& s, const float* a_tab,
                                                 const float* const* x, const float* v,
                                                 float seed_value, int64_t* taken_out)
{
  float o[3], dir[3];
  light_origin(m, v, o);
  float dd[3] = { v[0] - o[0], v[1] - o[1], v[2] - o[2] };
  const float rs = 1.0f / __builtin_sqrtf(dot3(dd, dd));
  dir[0] = dd[0] * rs;
  dir[1] = dd[1] * rs;
  dir[2] = dd[2] * rs;
  float tmin, tmax;
  (void)segment_aabb(o, v, g.bounds, 1e-4f, tmin, tmax);
  float start[3], p[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    start[d] = o[d] + tmin * dir[d];
    p[d] = start[d];
  }
  const float len2 = (tmax - tmin) * (tmax - tmin);
  const float tfmax = (float)(s.ntf - 1);
  // 32-bit cells and strides: set_field keeps a field's xy plane below 2^31 points
  const int c0 = (int)(g.n[0] - 1), c1 = (int)(g.n[1] - 1), c2 = (int)(g.n[2] - 1);
  const int sx = (int)g.field_stride[0], sy = (int)g.field_stride[1], sz = (int)g.field_stride[2];
  const int rofs = coords_rofs(g);
  const CornerFetch<FDT, FAST> corners(g.field, g.bytes32, sx, sy, sz);
  float op = seed_value;
  int64_t taken = 0;
  for (int64_t it = 0; it < kMapMaxSteps; ++it) {
    float df[3] = { p[0] - start[0], p[1] - start[1], p[2] - start[2] };
    if (!(dot3(df, df) <= len2)) break;
    p[0] = p[0] + m.step * dir[0];
    p[1] = p[1] + m.step * dir[1];
    p[2] = p[2] + m.step * dir[2];
    if (!inside(g, p)) break;
    ++taken;
    // FindCell and the parametric coordinates (p - x[c]) / (x[c+1] - x[c])
    AxisCell ax = axis_guess<BINS>(x[0], rofs, c0, p[0], g.gx0[0], BINS ? g.binv[0] : g.ginv[0]);
    AxisCell ay = axis_guess<BINS>(x[1], rofs, c1, p[1], g.gx0[1], BINS ? g.binv[1] : g.ginv[1]);
    AxisCell az = axis_guess<BINS>(x[2], rofs, c2, p[2], g.gx0[2], BINS ? g.binv[2] : g.ginv[2]);
    // all three tests evaluated (no short circuit: the reads above are issued together)
    const bool settled = ((int)oracle_in(ax, p[0]) & (int)oracle_in(ay, p[1]) & (int)oracle_in(az, p[2])) != 0;
    if (__builtin_expect(!settled, 0)) {
      oracle_settle(x[0], rofs, c0, p[0], ax);
      oracle_settle(x[1], rofs, c1, p[1], ay);
      oracle_settle(x[2], rofs, c2, p[2], az);
    }
    float pc[3];
    pc[0] = div_by_recip(p[0] - ax.lo, ax.hi - ax.lo, ax.rw);
    pc[1] = div_by_recip(p[1] - ay.lo, ay.hi - ay.lo, ay.rw);
    pc[2] = div_by_recip(p[2] - az.lo, az.hi - az.lo, az.rw);
    float val[8];
    corners.load(ax.c, ay.c, az.c, val);
    float sc = interp_hex(val, pc);
    sc = (sc - s.min_scalar) * s.inv_delta;
    op = op + (1.0f - op) * a_tab[tf_clamped_index(sc * tfmax, tfmax)];
    if (op == 1.0f) break;  // every later step leaves an opacity of exactly 1 unchanged
  }
  *taken_out = taken;
  return op;
}

// march_vertex_at with the next step's corner loads in flight while the current step is
// interpolated and composited: the same steps, cells, values and operations in the same order
// (so the same opacity bit for bit), with the gather latency off the dependent chain.  Unrolled
// by two with two load buffers, so no registers are copied between steps.
template <int FDT, bool FAST = false>
__device__ __forceinline__ float march_vertex_pipe(const GridParams& g, const MapParams& m,
                                                   const ShadeParams& s, const float* a_tab,
                                                   const float* const* x, const float* v,
                                                   float seed_value, int64_t* taken_out)
{
  float o[3], dir[3];
  light_origin(m, v, o);
  float dd[3] = { v[0] - o[0], v[1] - o[1], v[2] - o[2] };
  const float rs = 1.0f / __builtin_sqrtf(dot3(dd, dd));
  dir[0] = dd[0] * rs;
  dir[1] = dd[1] * rs;
  dir[2] = dd[2] * rs;
  float tmin, tmax;
  (void)segment_aabb(o, v, g.bounds, 1e-4f, tmin, tmax);
  float start[3], q[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    start[d] = o[d] + tmin * dir[d];
    q[d] = start[d];
  }
  const float len2 = (tmax - tmin) * (tmax - tmin);
  const float tfmax = (float)(s.ntf - 1);
  const int c0 = (int)(g.n[0] - 1), c1 = (int)(g.n[1] - 1), c2 = (int)(g.n[2] - 1);
  const int sx = (int)g.field_stride[0], sy = (int)g.field_stride[1], sz = (int)g.field_stride[2];
  const int rofs = coords_rofs(g);
  const CornerFetch<FDT, FAST> corners(g.field, g.bytes32, sx, sy, sz);
  float op = seed_value;
  int64_t taken = 0;
  struct Sample {
    uint32_t raw[8];  // the corners' storage words, decoded when the sample is composited
    float pc[3];
  };
  // the reference loop's test and step for the next sample after position q (q moves to it)
  auto advance = [&](float* pq) -> bool {
    const float df[3] = { pq[0] - start[0], pq[1] - start[1], pq[2] - start[2] };
    if (!(dot3(df, df) <= len2)) return false;
    pq[0] = pq[0] + m.step * dir[0];
    pq[1] = pq[1] + m.step * dir[1];
    pq[2] = pq[2] + m.step * dir[2];
    return inside(g, pq);
  };
  // FindCell, the parametric coordinates and the corner loads of the sample at pq
  auto prepare = [&](const float* pq, Sample& S) {
    AxisCell ax = axis_guess(x[0], rofs, c0, pq[0], g.gx0[0], g.ginv[0]);
    AxisCell ay = axis_guess(x[1], rofs, c1, pq[1], g.gx0[1], g.ginv[1]);
    AxisCell az = axis_guess(x[2], rofs, c2, pq[2], g.gx0[2], g.ginv[2]);
    const bool settled = ((int)oracle_in(ax, pq[0]) & (int)oracle_in(ay, pq[1]) & (int)oracle_in(az, pq[2])) != 0;
    if (__builtin_expect(!settled, 0)) {
      oracle_settle(x[0], rofs, c0, pq[0], ax);
      oracle_settle(x[1], rofs, c1, pq[1], ay);
      oracle_settle(x[2], rofs, c2, pq[2], az);
    }
    corners.load_raw(ax.c, ay.c, az.c, S.raw);
    S.pc[0] = div_by_recip(pq[0] - ax.lo, ax.hi - ax.lo, ax.rw);
    S.pc[1] = div_by_recip(pq[1] - ay.lo, ay.hi - ay.lo, ay.rw);
    S.pc[2] = div_by_recip(pq[2] - az.lo, az.hi - az.lo, az.rw);
  };
  // interpolation and composite of a sample; true once the opacity is exactly 1
  auto consume = [&](const Sample& S) -> bool {
    ++taken;
    float val[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) val[k] = CornerFetch<FDT, FAST>::decode(S.raw[k]);
    float sc = interp_hex(val, S.pc);
    sc = (sc - s.min_scalar) * s.inv_delta;
    op = op + (1.0f - op) * a_tab[tf_clamped_index(sc * tfmax, tfmax)];
    return op == 1.0f;
  };
  Sample A, B;
  int64_t it = 0;  // index of the sample in flight (at most kMapMaxSteps samples)
  bool live = advance(q);
  if (live) prepare(q, A);
  while (live) {
    bool nxt = it + 1 < kMapMaxSteps && advance(q);
    if (nxt) prepare(q, B);
    if (consume(A) || !nxt) break;
    ++it;
    nxt = it + 1 < kMapMaxSteps && advance(q);
    if (nxt) prepare(q, A);
    if (consume(B) || !nxt) break;
    ++it;
  }
  *taken_out = taken;
  return op;
}

template <int FDT, bool BINS = false, bool FAST = false>
__device__ __forceinline__ float march_vertex(const GridParams& g, const MapParams& m,
                                              const ShadeParams& s, const float* a_tab,
                                              const float* const* x, int64_t id, float seed_value,
                                              int64_t* taken_out)
{
  float v[3];
  map_vertex(m, id, v);
  return march_vertex_at<FDT, BINS, FAST>(g, m, s, a_tab, x, v, seed_value, taken_out);
}

// ---- LightedVolumeRenderer.cxx:345-498 ----------------------------------------------------
// BLK: g.field and m.map are 4 x 4 x 4-blocked copies (BlockFetch; their block strides in
// field_stride[1..2] and ms[1..2]).
template <int FDT, int MDT, bool USE_MAP, bool BINS = false, bool FAST = false, bool BLK = false>
__device__ __forceinline__ int64_t sample_point(const GridParams& g, const MapParams& m,
                                             const ShadeParams& s, const float4* tf,
                                             const float* const* x, const float* org,
                                             const float* dir, float minD, float maxD,
                                             float* color)
{
  if (minD == -1.0f) return 0;
  int composited = 0;  // per ray: below 2^31
  float dist = minD + s.mesh_eps;
  float p[3];
  int steps = 0;  // bounded by s.max_steps (below 2^31), which no finite in-box ray reaches
  const int max_steps = (int)s.max_steps;
#pragma unroll
  for (int d = 0; d < 3; ++d) p[d] = org[d] + dist * dir[d];
  while (!inside(g, p) && dist < maxD && ++steps < max_steps) {
    dist += s.sample_dist;
#pragma unroll
    for (int d = 0; d < 3; ++d) p[d] = org[d] + dist * dir[d];
  }
  float bl[3] = { 0.0f, 0.0f, 0.0f };
  bool newCell = true;
  float tx = 0.0f, ty = 0.0f, tz = 0.0f;
  float s0 = 0.0f, s10 = 0.0f, s23 = 0.0f, s3 = 0.0f, s4 = 0.0f, s54 = 0.0f, s67 = 0.0f,
        s7 = 0.0f;
  int cell[3] = { 0, 0, 0 };
  float invsp[3] = { 0.0f, 0.0f, 0.0f };
  const int cms = s.ntf - 1;
  // 32-bit cells and strides: set_field keeps a field's xy plane below 2^31 points
  const int sx = (int)g.field_stride[0], sy = (int)g.field_stride[1], sz = (int)g.field_stride[2];
  const int rofs = coords_rofs(g);
  typedef typename std::conditional<BLK, BlockFetch<FDT>, CornerFetch<FDT, FAST>>::type FieldFetch;
  typedef typename std::conditional<BLK, BlockFetch<MDT>, CornerFetch<MDT, FAST>>::type MapFetch;
  const FieldFetch corners(g.field, g.bytes32, sx, sy, sz);
  const MapFetch map_corners(m.map, m.bytes32, (int)m.ms[0], (int)m.ms[1], (int)m.ms[2]);
  const float sd = s.sample_dist;
  MapCellCache mcache;
  while (inside(g, p) && dist < maxD && ++steps < max_steps) {
    // the shadow map's cell first: its corner loads, if any, overlap the field's
    if (USE_MAP) map_cell_update<MDT>(m, map_corners, p, mcache);
    const float mint = fminf(tx, fminf(ty, tz));
    const float maxt = fmaxf(tx, fmaxf(ty, tz));
    if (maxt > 1.0f || mint < 0.0f) newCell = true;
    if (newCell) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (p[d] == g.maxp[d]) {
          cell[d] = (int)g.n[d] - 2;  // invSpacing keeps its previous value
          bl[d] = x[d][cell[d]];
        } else {
          // the cell and 1 / (x[c+1] - x[c]) (the staged IEEE quotient)
          const AxisCell a = walk_axis<BINS>(x[d], rofs, (int)g.n[d] - 1, p[d], g.gx0[d],
                                             BINS ? g.binv[d] : g.ginv[d]);
          invsp[d] = a.rw;
          cell[d] = a.c;
          bl[d] = a.lo;
        }
      }
      float cv[8];
      corners.load(cell[0], cell[1], cell[2], cv);
      s0 = cv[0];
      const float s1 = cv[1], s2 = cv[2];
      s3 = cv[3];
      s4 = cv[4];
      const float s5 = cv[5], s6 = cv[6];
      s7 = cv[7];
      s67 = s6 - s7;
      s54 = s5 - s4;
      s10 = s1 - s0;
      s23 = s2 - s3;
      tx = (p[0] - bl[0]) * invsp[0];
      ty = (p[1] - bl[1]) * invsp[1];
      tz = (p[2] - bl[2]) * invsp[2];
      newCell = false;
    }
    const float l76 = s7 + tx * s67;
    const float l45 = s4 + tx * s54;
    const float top = l45 + ty * (l76 - l45);
    const float l01 = s0 + tx * s10;
    const float l32 = s3 + tx * s23;
    const float bot = l01 + ty * (l32 - l01);
    float fs = bot + tz * (top - bot);
    fs = (fs - s.min_scalar) * s.inv_delta;
    // truncation toward zero as the reference's static_cast<Id>, after clamping to
    // [-1, cms + 1]: a product out of the table's range stays out of it, in 32 bits
    const float fci = fs * (float)cms;
    const int ci = (int)fminf(fmaxf(fci, -1.0f), (float)cms + 1.0f);
    // The reference re-evaluates the same sample forever for an index outside the
    // table; here the sample is skipped and the ray advances.
    if (ci >= 0 && ci <= cms) {
      float4 c = tf[ci];
      c.w *= (1.0f - color[3]);
      if (USE_MAP) {
        const float tr = 1.0f - map_cell_opacity(p, mcache);
        c.x = c.x * (tr * m.light_color[0]);
        c.y = c.y * (tr * m.light_color[1]);
        c.z = c.z * (tr * m.light_color[2]);
      }
      c.x = c.x < 0.0f ? 0.0f : (c.x > 1.0f ? 1.0f : c.x);
      c.y = c.y < 0.0f ? 0.0f : (c.y > 1.0f ? 1.0f : c.y);
      c.z = c.z < 0.0f ? 0.0f : (c.z > 1.0f ? 1.0f : c.z);
      color[0] = color[0] + c.x * c.w;
      color[1] = color[1] + c.y * c.w;
      color[2] = color[2] + c.z * c.w;
      color[3] = c.w + color[3];
      ++composited;
    }
    dist += sd;
    p[0] = p[0] + sd * dir[0];
    p[1] = p[1] + sd * dir[1];
    p[2] = p[2] + sd * dir[2];
    tx = (p[0] - bl[0]) * invsp[0];
    ty = (p[1] - bl[1]) * invsp[1];
    tz = (p[2] - bl[2]) * invsp[2];
    if (color[3] >= s.alpha_stop) break;
  }
  return composited;
}

// ---- LightedVolumeRenderer.cxx:540-641 (cell-associated scalars, no shadow term) --------
template <int FDT>
__device__ __forceinline__ int64_t sample_cell(const GridParams& g, const ShadeParams& s,
                                            const float4* tf, const float* const* x,
                                            const float* org, const float* dir, float minD,
                                            float maxD, float* color)
{
  if (minD == -1.0f) return 0;
  int composited = 0;  // per ray: below 2^31
  float dist = minD + s.mesh_eps;
  float p[3];
  int steps = 0;
  const int max_steps = (int)s.max_steps;
#pragma unroll
  for (int d = 0; d < 3; ++d) p[d] = org[d] + dist * dir[d];
  while (!inside(g, p) && dist < maxD && ++steps < max_steps) {
    dist += s.sample_dist;
#pragma unroll
    for (int d = 0; d < 3; ++d) p[d] = org[d] + dist * dir[d];
  }
  bool newCell = true;
  float tx = 2.0f, ty = 2.0f, tz = 2.0f;
  float4 sc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float bl[3] = { 0.0f, 0.0f, 0.0f };
  float invsp[3] = { 0.0f, 0.0f, 0.0f };
  int64_t cell[3] = { 0, 0, 0 };
  const int64_t cms = s.ntf - 1;
  const float sd = s.sample_dist;
  while (inside(g, p) && dist < maxD && ++steps < max_steps) {
    const float mint = fminf(tx, fminf(ty, tz));
    const float maxt = fmaxf(tx, fmaxf(ty, tz));
    if (maxt > 1.0f || mint < 0.0f) newCell = true;
    if (newCell) {
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (p[d] == g.maxp[d]) {
          cell[d] = g.n[d] - 2;
        } else {
          const int64_t c = walk_cell(x[d], g.n[d] - 1, p[d], g.gx0[d], g.ginv[d]);
          invsp[d] = 1.0f / (x[d][c + 1] - x[d][c]);
          cell[d] = c;
        }
      }
      const int64_t cid = cell[0] * g.field_stride[0] + cell[1] * g.field_stride[1] +
                          cell[2] * g.field_stride[2];
      const float v = load_t<FDT>(g.field, cid);
      const float ns = (v - s.min_scalar) * s.inv_delta;
      int64_t ci = (int64_t)(ns * (float)cms);
      if (ci < 0) ci = 0;
      if (ci > cms) ci = cms;
      sc = tf[ci];
      bl[0] = x[0][cell[0]];
      bl[1] = x[1][cell[1]];
      bl[2] = x[2][cell[2]];
      tx = (p[0] - bl[0]) * invsp[0];
      ty = (p[1] - bl[1]) * invsp[1];
      tz = (p[2] - bl[2]) * invsp[2];
      newCell = false;
    }
    const float a = sc.w * (1.0f - color[3]);
    color[0] = color[0] + sc.x * a;
    color[1] = color[1] + sc.y * a;
    color[2] = color[2] + sc.z * a;
    color[3] = a + color[3];
    ++composited;
    dist += sd;
    p[0] = p[0] + sd * dir[0];
    p[1] = p[1] + sd * dir[1];
    p[2] = p[2] + sd * dir[2];
    if (color[3] >= s.alpha_stop) break;
    tx = (p[0] - bl[0]) * invsp[0];
    ty = (p[1] - bl[1]) * invsp[1];
    tz = (p[2] - bl[2]) * invsp[2];
  }
  color[0] = fminf(color[0], 1.0f);
  color[1] = fminf(color[1], 1.0f);
  color[2] = fminf(color[2], 1.0f);
  color[3] = fminf(color[3], 1.0f);
  return composited;
}

__device__ __forceinline__ void stage_tf(const ShadeParams& s, float4* tf)
{
  for (int i = threadIdx.x; i < s.ntf; i += blockDim.x)
    tf[i] = make_float4(s.tf[4 * i], s.tf[4 * i + 1], s.tf[4 * i + 2], s.tf[4 * i + 3]);
}

template <int FDT, int MDT, bool USE_MAP, bool BINS = false, bool FAST = false, bool BLK = false>
__device__ __forceinline__ void shade(const GridParams& g, const MapParams& m,
                                      const ShadeParams& s, const float4* tf,
                                      const float* const* x, const float* org, const float* dir,
                                      float minD, float maxD, float* color)
{
  const int64_t n = (g.assoc == 1)
                      ? sample_cell<FDT>(g, s, tf, x, org, dir, minD, maxD, color)
                      : sample_point<FDT, MDT, USE_MAP, BINS, FAST, BLK>(g, m, s, tf, x, org, dir, minD, maxD, color);
  if (s.counters) count_add(&s.counters[0], (uint32_t)n);
}


}  // namespace beams
