// Exact sweep: the transmittance map in one pass over light-aligned slices.
//
// The reference marches a separate light ray for every map vertex (<= 129 steps of
// |size|/128, TransmittanceMap.h:299-331).  The sweep instead visits the map one slice at a
// time, starting at the face the light enters through, and carries optical depth from the
// previous slice: for a vertex v the light ray crosses the previous slice's plane at p,
//     tau(v) = tau(p) + 0.5 (sigma(p) + sigma(v)) |v - p|,
// with tau(p), sigma(p) interpolated bilinearly from the previous slice (staged in LDS) and
// sigma(x) = -ln(1 - alpha(x) / alpha_max) / h the extinction that makes one reference step
// h = |size|/128 composite exactly alpha/alpha_max.  A ray that enters through a side face
// between the two planes starts at that face.  The map stores the opacity
// 1 - exp(-(tau + sigma(v) h / 2)): the reference's last step lands on average half a step
// past the vertex.  Arithmetic is FP32; the result is a re-discretisation of the reference
// map, compared with the FP64 oracle rather than bit for bit.
#include "march_dev.hpp"
#include "sweep.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

// Two launch shapes, the same arithmetic per vertex (bit-identical maps):
// - blocked (the default): one fully parallel pass writes every vertex's extinction sigma
//   (the only field reads), then each launch advances T slices of a square tile in LDS
//   from a halo of T times the stencil's reach per slice (bounded on each side of each
//   transverse axis separately: a light ray moves one way across the slices); each slice's
//   region shrinks by one reach, so no workgroup ever waits on another, and the (tau, sigma)
//   slice reaches HBM once per launch;
// - one launch per slice, reading the previous slice from HBM: for lights so oblique that
//   the halo would not fit in LDS.
// Vertices are addressed in the sweep's own axis order (0: the sweep axis, 1 and 2: the
// transverse axes a1 and a2, a1 fastest in the map and slice layouts), with that order's
// geometry in SweepParams.f.
// Both shapes evaluate a vertex through the same inline functions (sweep_cell, sweep_interp,
// sweep_tau, sweep_opacity) with contraction off and the fused multiply-adds written out, so
// every kernel rounds alike whatever the surrounding code.
#pragma clang fp contract(off)

namespace beams {

namespace {

constexpr int kTile = 16;
constexpr int kTileMin = 32, kTileMax = 64;
constexpr int kHalo = 3;
constexpr int kStage = kTile + 2 * kHalo + 1;

// Staging of the TF's extinction per entry, the coordinates and their reciprocal cell widths
// (LDS: ntf + 2 ncoord floats; xs[ncoord + i] = 1 / (x[i+1] - x[i]) within each axis).
__device__ __forceinline__ void stage_sigma_table(const SweepParams& sp, float* sig_tab)
{
  for (int i = threadIdx.x; i < sp.s.ntf; i += blockDim.x) {
    const float a = fminf(sp.s.tf[4 * i + 3] / sp.s.max_density, 0.99999994f);
    sig_tab[i] = -__logf(1.0f - a) * sp.inv_h;
  }
}
__device__ __forceinline__ void stage_sweep(const SweepParams& sp, float* sig_tab, float* xs)
{
  const GridParams& g = sp.g;
  const int n0 = (int)g.n[0], n01 = (int)(g.n[0] + g.n[1]), ncoord = (int)(g.n[0] + g.n[1] + g.n[2]);
  stage_sigma_table(sp, sig_tab);
  for (int i = threadIdx.x; i < ncoord; i += blockDim.x) {
    const int d = i < n0 ? 0 : (i < n01 ? 1 : 2), j = i - (d == 0 ? 0 : (d == 1 ? n0 : n01));
    xs[i] = g.x[d][j];
    xs[ncoord + i] = g.r[d][j];
  }
}

// A map vertex's coordinates: w[j] along sweep-order axis j, and its position xyz in the
// field's order (a branch on the sweep axis, uniform across the launch).
__device__ __forceinline__ void sweep_w(const SweepParams& sp, int k, int i1, int i2, float* w)
{
  w[0] = __builtin_fmaf(sp.f.sp[0], (float)k, sp.f.org[0]);
  w[1] = __builtin_fmaf(sp.f.sp[1], (float)i1, sp.f.org[1]);
  w[2] = __builtin_fmaf(sp.f.sp[2], (float)i2, sp.f.org[2]);
}
__device__ __forceinline__ void sweep_xyz(const SweepParams& sp, const float* w, float* v)
{
  if (sp.axis == 0) {
    v[0] = w[0];
    v[1] = w[1];
    v[2] = w[2];
  } else if (sp.axis == 1) {
    v[0] = w[1];
    v[1] = w[0];
    v[2] = w[2];
  } else {
    v[0] = w[1];
    v[1] = w[2];
    v[2] = w[0];
  }
}

// The field cell of position v (x, y, z) and its parametric coordinates t: the guess from
// uniform spacing for every axis first (their reads issued together), then a walk where any
// guess missed (v in [x[c], x[c+1]), the last cell also holding the max face, which the
// guess's clamp already lands on).  x[d]: axis d's coordinates, rw[d]: their reciprocal cell
// widths (LDS or global).
struct SweepCell {
  int c[3];
  float t[3];
};
__device__ __forceinline__ SweepCell sweep_cell(const GridParams& g, const float* const* x, const float* const* rw,
                                                const float* v)
{
  int c[3];
  float lo[3], hi[3];
  bool hit = true;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[d] = (int)__builtin_amdgcn_fmed3f((v[d] - g.gx0[d]) * g.ginv[d], 0.0f, (float)(g.n[d] - 2));
    lo[d] = x[d][c[d]];
    hi[d] = x[d][c[d] + 1];
    hit = hit & (v[d] >= lo[d]) & ((v[d] < hi[d]) | (c[d] == (int)g.n[d] - 2));
  }
  if (__builtin_expect(!hit, 0)) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int nc = (int)g.n[d] - 1;
      int cc = c[d];
      while (cc > 0 && v[d] < x[d][cc]) --cc;
      while (cc < nc - 1 && v[d] >= x[d][cc + 1]) ++cc;
      c[d] = cc;
      lo[d] = x[d][cc];
    }
  }
  SweepCell r;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    r.c[d] = c[d];
    r.t[d] = fminf(fmaxf((v[d] - lo[d]) * rw[d][c[d]], 0.0f), 1.0f);
  }
  return r;
}

// Extinction from the cell's eight corner values (x fastest): trilinear value, TF entry, sigma.
__device__ __forceinline__ float sweep_interp(const SweepParams& sp, const float* sig_tab, const float* v,
                                             const float* t)
{
  const float x0 = __builtin_fmaf(t[0], v[1] - v[0], v[0]), x1 = __builtin_fmaf(t[0], v[2] - v[3], v[3]);
  const float x2 = __builtin_fmaf(t[0], v[5] - v[4], v[4]), x3 = __builtin_fmaf(t[0], v[6] - v[7], v[7]);
  const float y0 = __builtin_fmaf(t[1], x1 - x0, x0), y1 = __builtin_fmaf(t[1], x3 - x2, x2);
  const float f = __builtin_fmaf(t[2], y1 - y0, y0);
  const float tfmax = (float)(sp.s.ntf - 1);
  return sig_tab[tf_clamped_index((f - sp.s.min_scalar) * sp.s.inv_delta * tfmax, tfmax)];
}

template <int FDT>
__device__ __forceinline__ CornerFetch<FDT> sweep_corners(const GridParams& g)
{
  return CornerFetch<FDT>(g.field, g.bytes32, (int)g.field_stride[0], (int)g.field_stride[1],
                          (int)g.field_stride[2]);
}

// The eight corner loads of a cell for a field view below 2^32 bytes with strides below 2^24
// (launch_sweep checks it): full-rate 24-bit offset products, and the storage words kept at their own
// width until decoded (the same values as CornerFetch).
template <int FDT>
struct SweepFetch {
  typedef typename std::conditional<FDT == kF32, uint32_t, uint16_t>::type Raw;
  __amdgpu_buffer_rsrc_t rsrc;
  uint32_t sx, sy, sz;
  int so[8];
  __device__ __forceinline__ explicit SweepFetch(const GridParams& g)
  {
    const uint64_t fp = (uint64_t)g.field;
    const uint32_t plo = __builtin_amdgcn_readfirstlane((uint32_t)fp);
    const uint32_t phi = __builtin_amdgcn_readfirstlane((uint32_t)(fp >> 32));
    void* f = (void*)(((uint64_t)phi << 32) | plo);
    rsrc = __builtin_amdgcn_make_buffer_rsrc(f, (short)0, (int)__builtin_amdgcn_readfirstlane(g.bytes32), 0x00020000);
    sx = __builtin_amdgcn_readfirstlane((uint32_t)g.field_stride[0]);
    sy = __builtin_amdgcn_readfirstlane((uint32_t)g.field_stride[1]);
    sz = __builtin_amdgcn_readfirstlane((uint32_t)g.field_stride[2]);
    const uint32_t e = FDT == kF32 ? 4u : 2u;
    const uint32_t d[8] = { 0u, sx, sx + sy, sy, sz, sz + sx, sz + sx + sy, sz + sy };
#pragma unroll
    for (int q = 0; q < 8; ++q) so[q] = (int)(d[q] * e);
  }
  __device__ __forceinline__ void load_raw(int c0, int c1, int c2, Raw* w) const
  {
    const uint32_t e = __umul24((uint32_t)c0, sx) + __umul24((uint32_t)c1, sy) + __umul24((uint32_t)c2, sz);
    const int off = (int)(e * (FDT == kF32 ? 4u : 2u));
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if constexpr (FDT == kF32)
        w[q] = __builtin_amdgcn_raw_buffer_load_b32(rsrc, off, so[q], 0);
      else
        w[q] = __builtin_amdgcn_raw_buffer_load_b16(rsrc, off, so[q], 0);
    }
  }
  __device__ __forceinline__ static void decode(const Raw* w, float* v)
  {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if constexpr (FDT == kF32)
        v[q] = __builtin_bit_cast(float, (uint32_t)w[q]);
      else
        v[q] = FDT == kF16 ? f16_to_f32((uint16_t)w[q]) : bf16_to_f32((uint16_t)w[q]);
    }
  }
};

// A slice's light geometry (sweep-order slice coordinate w0): the light ray back from vertex
// (i1, i2) crosses the previous slice's plane at transverse vertex coordinates
// u_j = a_j i_j + b_j (a point light's move is linear across a slice), after a path of
// tb (directional) or tb |e| (point light, e = w - L the light-to-vertex vector, whose first
// component is e0 across the slice).
struct SweepSlice {
  float a1, b1, a2, b2, tb, e0sq, o1, o2;
};
__device__ __forceinline__ SweepSlice sweep_slice(const SweepParams& sp, int k)
{
  const SweepFrame& f = sp.f;
  SweepSlice g;
  if (sp.m.light_kind == 1) {
    // u_j = i_j - tb d_j / sp_j, tb = sp0 / |d0|
    g.tb = f.sp[0] * __builtin_amdgcn_rcpf(fabsf(f.light[0]));
    g.a1 = 1.0f;
    g.a2 = 1.0f;
    g.b1 = -(g.tb * f.light[1]) * f.inv[1];
    g.b2 = -(g.tb * f.light[2]) * f.inv[2];
    g.e0sq = g.o1 = g.o2 = 0.0f;
  } else {
    // u_j = i_j - (sp0 / |e0|) e_j / sp_j with e_j = sp_j i_j + (org_j - L_j)
    const float e0 = __builtin_fmaf(f.sp[0], (float)k, f.org[0]) - f.light[0];
    g.tb = f.sp[0] * __builtin_amdgcn_rcpf(fabsf(e0));
    g.e0sq = e0 * e0;
    g.o1 = f.org[1] - f.light[1];
    g.o2 = f.org[2] - f.light[2];
    const float k1 = g.tb * f.inv[1], k2 = g.tb * f.inv[2];
    g.a1 = __builtin_fmaf(-k1, f.sp[1], 1.0f);
    g.a2 = __builtin_fmaf(-k2, f.sp[2], 1.0f);
    g.b1 = -k1 * g.o1;
    g.b2 = -k2 * g.o2;
  }
  return g;
}

// Optical depth at vertex (i1, i2) of slice g (not the first) from the previous slice:
// q(c1, c2) is the previous slice's (tau, sigma) at transverse vertex (c1, c2).
template <class Prev>
__device__ __forceinline__ float sweep_tau(const SweepParams& sp, const SweepSlice& g, int i1, int i2, float sig_v,
                                           const Prev& q)
{
  const SweepFrame& f = sp.f;
  const bool point = sp.m.light_kind != 1;
  const float fi1 = (float)i1, fi2 = (float)i2;
  const float u1 = __builtin_fmaf(g.a1, fi1, g.b1), u2 = __builtin_fmaf(g.a2, fi2, g.b2);
  float tb = g.tb, e1 = 0.0f, e2 = 0.0f, len = 1.0f;
  if (point) {
    e1 = __builtin_fmaf(f.sp[1], fi1, g.o1);
    e2 = __builtin_fmaf(f.sp[2], fi2, g.o2);
    const float l2 = __builtin_fmaf(e2, e2, __builtin_fmaf(e1, e1, g.e0sq));
    len = l2 * __builtin_amdgcn_rsqf(l2);
    tb = g.tb * len;
  }
  const int n1 = f.n[1], n2 = f.n[2];
  if (u1 >= 0.0f && u2 >= 0.0f && u1 <= (float)(n1 - 1) && u2 <= (float)(n2 - 1)) {
    // a map has at least 2 vertices per axis, so c + 1 stays in the slice
    const int c1 = min((int)u1, n1 - 2), c2 = min((int)u2, n2 - 2);
    const float f1 = u1 - (float)c1, f2 = u2 - (float)c2;
    const float2 q0 = q(c1, c2), q1 = q(c1 + 1, c2), q2 = q(c1, c2 + 1), q3 = q(c1 + 1, c2 + 1);
    const float ta = __builtin_fmaf(f1, q1.x - q0.x, q0.x), tc = __builtin_fmaf(f1, q3.x - q2.x, q2.x);
    const float sa = __builtin_fmaf(f1, q1.y - q0.y, q0.y), sc = __builtin_fmaf(f1, q3.y - q2.y, q2.y);
    const float tp = __builtin_fmaf(f2, tc - ta, ta), sp_ = __builtin_fmaf(f2, sc - sa, sa);
    return __builtin_fmaf(0.5f * (sp_ + sig_v), tb, tp);
  }
  // entered through a side face between the two planes: the path back to that face
  const float d1 = point ? e1 / len : f.light[1], d2 = point ? e2 / len : f.light[2];
  const float w1 = __builtin_fmaf(f.sp[1], fi1, f.org[1]), w2 = __builtin_fmaf(f.sp[2], fi2, f.org[2]);
  const float hi1 = __builtin_fmaf(f.sp[1], (float)(n1 - 1), f.org[1]);
  const float hi2 = __builtin_fmaf(f.sp[2], (float)(n2 - 1), f.org[2]);
  float te = tb;
  if (d1 > 0.0f) te = fminf(te, (w1 - f.org[1]) / d1);
  if (d1 < 0.0f) te = fminf(te, (hi1 - w1) / -d1);
  if (d2 > 0.0f) te = fminf(te, (w2 - f.org[2]) / d2);
  if (d2 < 0.0f) te = fminf(te, (hi2 - w2) / -d2);
  return sig_v * fmaxf(te, 0.0f);
}

__device__ __forceinline__ float sweep_opacity(const SweepParams& sp, float tau, float sig_v)
{
  return 1.0f - __expf(-__builtin_fmaf(0.5f * sp.h, sig_v, tau));
}

// ---- one launch per slice -----------------------------------------------------------------
template <int FDT, int MDT>
__global__ void __launch_bounds__(256) k_sweep_slice(SweepParams sp, int32_t k, int32_t first)
{
  if (gate_closed(sp.m)) return;
  extern __shared__ float lds[];
  const GridParams& g = sp.g;
  const int ntf = sp.s.ntf;
  float* sig_tab = lds;
  float* xs = lds + ntf;
  const int ncoord = (int)(g.n[0] + g.n[1] + g.n[2]);
  float2* stage = (float2*)(xs + 2 * ncoord);
  stage_sweep(sp, sig_tab, xs);
  const int n1 = sp.f.n[1], n2 = sp.f.n[2];
  const int base1 = blockIdx.x * kTile - kHalo, base2 = blockIdx.y * kTile - kHalo;
  if (!first)
    for (int i = threadIdx.x; i < kStage * kStage; i += blockDim.x) {
      const int u = min(max(base1 + i % kStage, 0), n1 - 1);
      const int w = min(max(base2 + i / kStage, 0), n2 - 1);
      stage[i] = sp.prev[w * n1 + u];
    }
  __syncthreads();
  const float* x[3] = { xs, xs + g.n[0], xs + g.n[0] + g.n[1] };
  const float* rw[3] = { x[0] + ncoord, x[1] + ncoord, x[2] + ncoord };
  const int i1 = blockIdx.x * kTile + (threadIdx.x % kTile);
  const int i2 = blockIdx.y * kTile + (threadIdx.x / kTile);
  if (i1 >= n1 || i2 >= n2) return;
  float w[3], v[3];
  sweep_w(sp, k, i1, i2, w);
  sweep_xyz(sp, w, v);
  const SweepCell cl = sweep_cell(g, x, rw, v);
  float cv[8];
  sweep_corners<FDT>(g).load(cl.c[0], cl.c[1], cl.c[2], cv);
  const float sig_v = sweep_interp(sp, sig_tab, cv, cl.t);
  float tau = 0.0f;
  if (!first) {
    auto q = [&](int c1, int c2) -> float2 {
      const int l1 = c1 - base1, l2 = c2 - base2;
      if (l1 >= 0 && l2 >= 0 && l1 < kStage && l2 < kStage) return stage[l2 * kStage + l1];
      return sp.prev[c2 * n1 + c1];
    };
    tau = sweep_tau(sp, sweep_slice(sp, k), i1, i2, sig_v, q);
  }
  const int p = i2 * n1 + i1;
  sp.cur[p] = make_float2(tau, sig_v);
  store_t<MDT>(sp.m.map, (int64_t)k * n1 * n2 + p, sweep_opacity(sp, tau, sig_v));
}

// ---- blocked -------------------------------------------------------------------------------
// Every vertex's sigma, slice by slice in sweep order (slice s = the s-th one the light crosses),
// a1 fastest: sig[s * n1 n2 + i2 n1 + i1].  Workgroups take kSigmaRuns runs of 256 vertices at
// a time, one per thread each, with every run's LDS reads and then all corner loads in flight
// together.
constexpr int kSigmaRuns = 2;

// FAST: the field suits SweepFetch (launch_sweep checks it), otherwise CornerFetch.
template <int FDT, bool FAST>
__global__ void __launch_bounds__(256) k_sweep_sigma(SweepParams sp, float* sig)
{
  if (gate_closed(sp.m)) return;
  extern __shared__ float lds[];
  float* sig_tab = lds;
  float* xs = lds + sp.s.ntf;
  stage_sweep(sp, sig_tab, xs);
  __syncthreads();
  const GridParams& g = sp.g;
  const int ncoord = (int)(g.n[0] + g.n[1] + g.n[2]);
  const float* x[3] = { xs, xs + g.n[0], xs + g.n[0] + g.n[1] };
  const float* rw[3] = { x[0] + ncoord, x[1] + ncoord, x[2] + ncoord };
  typedef typename std::conditional<FAST, SweepFetch<FDT>, CornerFetch<FDT>>::type Fetch;
  typedef typename std::conditional<FAST, typename SweepFetch<FDT>::Raw, uint32_t>::type Raw;
  constexpr int kLoads = 8;
  const Fetch corners = [&]() {
    if constexpr (FAST)
      return SweepFetch<FDT>(g);
    else
      return sweep_corners<FDT>(g);
  }();
  // a slice has at most 4097^2 vertices, a sweep at most 4097 slices of 65,569 runs
  const int n0 = sp.f.n[0], n1 = sp.f.n[1], n12 = n1 * sp.f.n[2];
  const int runs = (n12 + 255) / 256, items = n0 * runs;
  const float rn1 = 1.0f / (float)n1;
  // item w0 + u is run r + u of slice s (carried into the next slices); the stride between a
  // workgroup's iterations is ds slices and dr runs
  const int stride = (int)gridDim.x * kSigmaRuns, ds = stride / runs, dr = stride - ds * runs;
  int s0 = (int)blockIdx.x * kSigmaRuns / runs, r0 = (int)blockIdx.x * kSigmaRuns - s0 * runs;
  for (int w0 = blockIdx.x * kSigmaRuns; w0 < items; w0 += stride) {
    SweepCell cl[kSigmaRuns];
    int64_t dst[kSigmaRuns];
    bool live[kSigmaRuns];
    int s = s0, r = r0;
#pragma unroll
    for (int u = 0; u < kSigmaRuns; ++u) {
      const int it = w0 + u;
      if (u > 0 && ++r == runs) {
        r = 0;
        ++s;
      }
      const int p = r * 256 + (int)threadIdx.x;
      live[u] = it < items && p < n12;
      const int pl = live[u] ? p : 0;
      // pl / n1: the FP32 quotient is within one of it, corrected either way
      int i2 = (int)((float)pl * rn1);
      i2 -= i2 * n1 > pl ? 1 : 0;
      i2 += (i2 + 1) * n1 <= pl ? 1 : 0;
      const int i1 = pl - i2 * n1;
      const int sl = min(s, n0 - 1);
      float w[3], v[3];
      sweep_w(sp, sp.step > 0 ? sl : n0 - 1 - sl, i1, i2, w);
      sweep_xyz(sp, w, v);
      cl[u] = sweep_cell(g, x, rw, v);
      dst[u] = (int64_t)sl * n12 + pl;
    }
    s0 += ds;
    r0 += dr;
    if (r0 >= runs) {
      r0 -= runs;
      ++s0;
    }
    Raw raw[kSigmaRuns][kLoads];
#pragma unroll
    for (int u = 0; u < kSigmaRuns; ++u) corners.load_raw(cl[u].c[0], cl[u].c[1], cl[u].c[2], raw[u]);
#pragma unroll
    for (int u = 0; u < kSigmaRuns; ++u) {
      float cv[8];
      if constexpr (FAST) {
        SweepFetch<FDT>::decode(raw[u], cv);
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) cv[q] = CornerFetch<FDT>::decode(raw[u][q]);
      }
      const float sg = sweep_interp(sp, sig_tab, cv, cl[u].t);
      if (live[u]) sig[dst[u]] = sg;
    }
  }
}

// Slices s0 .. s0 + nt - 1 of a tile x tile tile.  A light ray's bilinear stencil for vertex
// (i1, i2) lies within previous-slice vertices i1 - lo1 .. i1 + hi1 and i2 - lo2 .. i2 + hi2
// (bounded on each side separately: the plan's reach over the whole map, narrowed per tile by
// tile_reach), so slice t (0-based in the launch) is computed over the tile widened by
// (nt - 1 - t) lo on the low side and (nt - 1 - t) hi on the high side, from the region of
// slice t - 1 alone; slice -1 is the previous launch's last slice, staged from HBM over the
// tile widened by nt lo and nt hi (clamped to the slice).  No workgroup waits on another.
// The (tau, sigma) slices live in LDS (two W1 x W2 buffers); each thread keeps the sigma of its
// next slice's vertices in registers, loaded while the current slice is computed.  The last
// slice's tile goes back to HBM.
struct SweepShape {
  int32_t nt, tile, lo1, hi1, lo2, hi2;
};

constexpr int kSweepThreads = 1024;
constexpr int kSweepPer = 4;  // region vertices per thread and slice, at most

__host__ __device__ inline int sweep_len(int tile, int nt, int lo, int hi, int t)
{
  return tile + (nt - 1 - t) * (lo + hi);
}

// The transverse move bound of sweep_plan over one tile: a point light's move across a slice,
// -(sp0 / sp_j) (w_j - L_j) / |w_0 - L_0|, is extreme at the corners of the tile's window (the
// plan's reach) and this launch's nearest and farthest slices past the first; the same margin
// as the plan's, capped by the plan's reach.  A directional light keeps the plan's.
__device__ __forceinline__ void tile_reach(const SweepParams& sp, const SweepShape& sh, int t1, int t2, int s0,
                                           int nts, int* lo, int* hi)
{
  lo[0] = sh.lo1;
  hi[0] = sh.hi1;
  lo[1] = sh.lo2;
  hi[1] = sh.hi2;
  const SweepFrame& f = sp.f;
  const int sa = max(s0, 1), sb = s0 + nts - 1;
  if (sp.m.light_kind == 1 || sa > sb) return;
  const int ka = sp.step > 0 ? sa : f.n[0] - 1 - sa, kb = sp.step > 0 ? sb : f.n[0] - 1 - sb;
  const float da = fabsf(__builtin_fmaf(f.sp[0], (float)ka, f.org[0]) - f.light[0]);
  const float db = fabsf(__builtin_fmaf(f.sp[0], (float)kb, f.org[0]) - f.light[0]);
  const float dn = fminf(da, db), df = fmaxf(da, db);
  const int tt[2] = { t1, t2 };
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int ax = j + 1, n = f.n[ax];
    const int a = max(tt[j] - sh.nt * lo[j], 0), b = min(tt[j] + sh.tile - 1 + sh.nt * hi[j], n - 1);
    const float r = f.sp[0] * f.inv[ax];
    const float ea = __builtin_fmaf(f.sp[ax], (float)a, f.org[ax]) - f.light[ax];
    const float eb = __builtin_fmaf(f.sp[ax], (float)b, f.org[ax]) - f.light[ax];
    const float m0 = -r * ea / dn, m1 = -r * ea / df, m2 = -r * eb / dn, m3 = -r * eb / df;
    float dmin = fminf(fminf(m0, m1), fminf(m2, m3)), dmax = fmaxf(fmaxf(m0, m1), fmaxf(m2, m3));
    const float eps = 1e-3f + 1e-6f * (float)n;
    dmin -= eps * (1.0f + fabsf(dmin));
    dmax += eps * (1.0f + fabsf(dmax));
    lo[j] = min(lo[j], max(0, -(int)floorf(dmin)));
    hi[j] = min(hi[j], max(0, (int)floorf(dmax) + 1));
  }
}

// Window position of ring-order index q.  Ring 0 is the tile (slice nt - 1's region), ring m
// the vertices of slice (nt - 1 - m)'s region outside slice (nt - m)'s: the region of slice t is
// then the prefix q < |region t|.  b*, L*: slice t's region (b1 + r1, b2 + r2 within it).
__device__ __forceinline__ void ring_position(int q, int tile, int nt, int lo1, int hi1, int lo2, int hi2, int t1,
                                              int t2, int& i1, int& i2)
{
  if (q < tile * tile) {
    const int r2 = q / tile;
    i1 = t1 + q - r2 * tile;
    i2 = t2 + r2;
    return;
  }
  // the ring: outer region of slice nt - 1 - m around the inner one of slice nt - m
  int m = 1, Lo1 = tile + lo1 + hi1, Lo2 = tile + lo2 + hi2, Li1 = tile, Li2 = tile;
  while (m < nt - 1 && q >= Lo1 * Lo2) {
    ++m;
    Li1 = Lo1;
    Li2 = Lo2;
    Lo1 += lo1 + hi1;
    Lo2 += lo2 + hi2;
  }
  const int bo1 = t1 - m * lo1, bo2 = t2 - m * lo2;  // outer region's first vertex
  int j = q - Li1 * Li2;
  if (j < lo2 * Lo1) {  // rows below the inner region
    const int r = j / Lo1;
    i1 = bo1 + j - r * Lo1;
    i2 = bo2 + r;
    return;
  }
  j -= lo2 * Lo1;
  if (j < hi2 * Lo1) {  // rows above it
    const int r = j / Lo1;
    i1 = bo1 + j - r * Lo1;
    i2 = bo2 + lo2 + Li2 + r;
    return;
  }
  j -= hi2 * Lo1;  // its rows: lo1 vertices on the low side, hi1 on the high side
  const int w = lo1 + hi1, r = j / w, c = j - r * w;
  i2 = bo2 + lo2 + r;
  i1 = c < lo1 ? bo1 + c : bo1 + lo1 + Li1 + (c - lo1);
}

template <int MDT>
__global__ void __launch_bounds__(kSweepThreads) k_sweep_block(SweepParams sp, const float* sig, int32_t s0,
                                                               SweepShape sh)
{
  if (gate_closed(sp.m)) return;
  extern __shared__ float2 buf[];
  const int n0 = sp.f.n[0], n1 = sp.f.n[1], n2 = sp.f.n[2];
  const int nt = sh.nt, tile = sh.tile;
  const int t1 = blockIdx.x * tile, t2 = blockIdx.y * tile;  // the tile's first vertex
  const int nts = min(nt, n0 - s0);                           // slices of this launch
  int lo[2], hi[2];
  tile_reach(sp, sh, t1, t2, s0, nts, lo, hi);
  const int lo1 = __builtin_amdgcn_readfirstlane(lo[0]), hi1 = __builtin_amdgcn_readfirstlane(hi[0]);
  const int lo2 = __builtin_amdgcn_readfirstlane(lo[1]), hi2 = __builtin_amdgcn_readfirstlane(hi[1]);
  const int W1 = tile + nt * (lo1 + hi1), W2 = tile + nt * (lo2 + hi2);  // within the plan's LDS
  const int o1 = t1 - nt * lo1, o2 = t2 - nt * lo2;                       // the window's first vertex
  const int tid = (int)threadIdx.x;
  const int n12 = n1 * n2;
  float2* prev = buf;
  float2* cur = buf + W1 * W2;
  if (s0 > 0) {
    // the previous slice over the window, eight loads in flight per thread
    constexpr int kBatch = 8;
    const int d2 = kSweepThreads / W1, d1 = kSweepThreads - d2 * W1;
    int l2 = tid / W1, l1 = tid - l2 * W1;
    for (int base = 0; base < W1 * W2; base += kBatch * kSweepThreads) {
      float2 v[kBatch];
      int j1 = l1, j2 = l2;
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int i1 = min(max(o1 + j1, 0), n1 - 1), i2 = min(max(o2 + j2, 0), n2 - 1);
        v[u] = j2 < W2 ? sp.prev[i2 * n1 + i1] : make_float2(0.0f, 0.0f);
        j1 += d1;
        j2 += d2;
        if (j1 >= W1) {
          j1 -= W1;
          ++j2;
        }
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        if (l2 < W2) prev[l2 * W1 + l1] = v[u];
        l1 += d1;
        l2 += d2;
        if (l1 >= W1) {
          l1 -= W1;
          ++l2;
        }
      }
    }
  }
  // This thread's vertices for the whole launch: ring-order indices tid + kSweepThreads u, in
  // slice t's region while below |region t|; outside the slice they are skipped.
  const int size0 = sweep_len(tile, nt, lo1, hi1, 0) * sweep_len(tile, nt, lo2, hi2, 0);
  int vi1[kSweepPer], vi2[kSweepPer];
  bool vin[kSweepPer];
#pragma unroll
  for (int u = 0; u < kSweepPer; ++u) {
    const int q = tid + kSweepThreads * u;
    int i1 = t1, i2 = t2;
    if (q < size0) ring_position(q, tile, nt, lo1, hi1, lo2, hi2, t1, t2, i1, i2);
    vi1[u] = i1;
    vi2[u] = i2;
    vin[u] = q < size0 && (unsigned)i1 < (unsigned)n1 && (unsigned)i2 < (unsigned)n2;
  }
  const int ntile = tile * tile;
  auto load_sigma = [&](int t, float* out) {
    const int size = sweep_len(tile, nt, lo1, hi1, t) * sweep_len(tile, nt, lo2, hi2, t);
    const float* ss = sig + (int64_t)(s0 + t) * n12;
#pragma unroll
    for (int u = 0; u < kSweepPer; ++u)
      if (tid + kSweepThreads * u < size && vin[u]) out[u] = ss[vi2[u] * n1 + vi1[u]];
  };
  float sg_next[kSweepPer];
  load_sigma(0, sg_next);
  __syncthreads();
  for (int t = 0; t < nts; ++t) {
    float sg[kSweepPer];
#pragma unroll
    for (int u = 0; u < kSweepPer; ++u) sg[u] = sg_next[u];
    if (t + 1 < nts) load_sigma(t + 1, sg_next);
    const int s = s0 + t, k = sp.step > 0 ? s : n0 - 1 - s;
    const bool last = t == nts - 1;
    const int size = sweep_len(tile, nt, lo1, hi1, t) * sweep_len(tile, nt, lo2, hi2, t);
    const int64_t kbase = (int64_t)k * n12;
    const SweepSlice geo = sweep_slice(sp, k);
#pragma unroll
    for (int u = 0; u < kSweepPer; ++u) {
      const int q = tid + kSweepThreads * u;
      if (q < size && vin[u]) {
        const int i1 = vi1[u], i2 = vi2[u];
        const float sig_v = sg[u];
        float tau = 0.0f;
        if (s > 0) {
          auto qv = [&](int c1, int c2) -> float2 { return prev[(c2 - o2) * W1 + (c1 - o1)]; };
          tau = sweep_tau(sp, geo, i1, i2, sig_v, qv);
        }
        cur[(i2 - o2) * W1 + (i1 - o1)] = make_float2(tau, sig_v);
        if (q < ntile) {  // the tile
          const int p = i2 * n1 + i1;
          store_t<MDT>(sp.m.map, kbase + p, sweep_opacity(sp, tau, sig_v));
          if (last) sp.cur[p] = make_float2(tau, sig_v);
        }
      }
    }
    __syncthreads();
    float2* tmp = prev;
    prev = cur;
    cur = tmp;
  }
}

}  // namespace

SweepFrame sweep_frame(const MapParams& m, int axis, int a1, int a2)
{
  SweepFrame f{};
  const int ax[3] = { axis, a1, a2 };
  for (int j = 0; j < 3; ++j) {
    f.org[j] = m.morg[ax[j]];
    f.sp[j] = m.msp[ax[j]];
    f.inv[j] = m.minv[ax[j]];
    f.light[j] = m.light[ax[j]];
    f.n[j] = (int32_t)m.mp[ax[j]];
  }
  return f;
}

SweepPlan sweep_plan(const SweepParams& sp)
{
  const MapParams& m = sp.m;
  const int a0 = sp.axis;
  SweepPlan plan{};
  // The transverse move of a light ray across one slice, in vertices of axis a: the vertex
  // (i1, i2) reads the previous slice at i_a + delta_a, delta_a = -(msp0 / msp_a) d_a / |d_0|.
  // A directional light's delta is one constant; a point light's is (L_a - v_a) (msp0 / msp_a)
  // / |v_0 - L_0|, extreme at the map's transverse faces and its nearest and farthest slices
  // past the first (the first slice has no previous one).
  const double msp0 = (double)m.msp[a0];
  const double lo0 = (double)m.morg[a0], hi0 = lo0 + msp0 * (double)(m.mp[a0] - 1);
  const double L0 = (double)m.light[a0];
  const double near0 = std::min(std::fabs(lo0 - L0), std::fabs(hi0 - L0)) + msp0;
  const double far0 = std::max(std::fabs(lo0 - L0), std::fabs(hi0 - L0));
  const int ax[2] = { sp.a1, sp.a2 };
  bool fits = true;
  for (int s = 0; s < 2; ++s) {
    const int a = ax[s];
    const double mspa = (double)m.msp[a];
    double dmin, dmax;
    if (m.light_kind == 1) {
      dmin = dmax = -(msp0 / mspa) * (double)m.light[a] / std::max(std::fabs((double)m.light[a0]), 1e-300);
    } else {
      const double loa = (double)m.morg[a], hia = loa + mspa * (double)(m.mp[a] - 1);
      dmin = 1e300;
      dmax = -1e300;
      for (double va : { loa, hia })
        for (double D : { near0, std::max(far0, near0) }) {
          const double dl = -(msp0 / mspa) * (va - (double)m.light[a]) / D;
          dmin = std::min(dmin, dl);
          dmax = std::max(dmax, dl);
        }
    }
    // FP32 evaluation of the move on the device: a margin far above its rounding
    const double eps = 1e-3 + 1e-6 * (double)m.mp[a];
    dmin -= eps * (1.0 + std::fabs(dmin));
    dmax += eps * (1.0 + std::fabs(dmax));
    if (!(std::fabs(dmin) < 64.0 && std::fabs(dmax) < 64.0)) {
      fits = false;
      break;
    }
    // the stencil (c, c + 1), c = floor(i + delta): reads i + floor(dmin) .. i + floor(dmax) + 1
    plan.lo[s] = std::max(0, -(int)std::floor(dmin));
    plan.hi[s] = std::max(0, (int)std::floor(dmax) + 1);
  }
  plan.blocked = fits && std::max({ plan.lo[0], plan.hi[0], plan.lo[1], plan.hi[1] }) <= 8;
  if (!plan.blocked) return plan;
  // the tile: the narrowest from 32 whose grid fits the 256 CUs in one wave
  const int n0 = (int)m.mp[a0], n1 = (int)m.mp[sp.a1], n2 = (int)m.mp[sp.a2];
  plan.tile = kTileMax;
  for (int t = kTileMin; t <= kTileMax; ++t)
    if ((int64_t)((n1 + t - 1) / t) * ((n2 + t - 1) / t) <= 256) {
      plan.tile = t;
      break;
    }
  // slices per launch: the least estimated time (a launch costs about as much as eight
  // vertex rounds of its threads), within 150 KB of LDS for the two slice buffers and
  // kSweepPer vertices per thread in the first slice's region
  const int g1 = plan.lo[0] + plan.hi[0], g2 = plan.lo[1] + plan.hi[1];
  double best = 1e300;
  plan.nt = 0;
  for (int nt = 1; nt <= 16; ++nt) {
    const size_t W1 = (size_t)(plan.tile + nt * g1), W2 = (size_t)(plan.tile + nt * g2);
    const size_t lds = 2 * sizeof(float2) * W1 * W2;
    const int64_t first = (int64_t)sweep_len(plan.tile, nt, plan.lo[0], plan.hi[0], 0) *
                          sweep_len(plan.tile, nt, plan.lo[1], plan.hi[1], 0);
    if (lds > 150 * 1024 || first > (int64_t)kSweepPer * kSweepThreads) break;
    double rounds = 8.0;
    for (int t = 0; t < nt; ++t)
      rounds += (double)(((int64_t)sweep_len(plan.tile, nt, plan.lo[0], plan.hi[0], t) *
                              sweep_len(plan.tile, nt, plan.lo[1], plan.hi[1], t) +
                          kSweepThreads - 1) / kSweepThreads);
    const double est = rounds * (double)((n0 + nt - 1) / nt);
    if (est < best) {
      best = est;
      plan.nt = nt;
      plan.lds = lds;
    }
  }
  plan.blocked = plan.nt > 0;
  return plan;
}

hipError_t launch_sweep(SweepParams sp, const SweepPlan& plan, float* sig, hipStream_t st)
{
  const int n0 = sp.f.n[0], n1 = sp.f.n[1], n2 = sp.f.n[2];
  const int ncoord = (int)(sp.g.n[0] + sp.g.n[1] + sp.g.n[2]);
  if (plan.blocked && sig) {
    const int64_t items = (int64_t)n0 * (((int64_t)n1 * n2 + 255) / 256);
    const size_t lds_a = sizeof(float) * (size_t)(sp.s.ntf + 2 * ncoord);
    // 24-bit offset products: a view below 2^32 bytes, strides and cells below 2^24
    const bool fast = sp.g.bytes32 != 0 && sp.g.field_stride[0] < (1 << 24) && sp.g.field_stride[2] < (1 << 24) &&
                      sp.g.field_stride[1] < (1 << 24) && sp.g.n[0] < (1 << 24) && sp.g.n[1] < (1 << 24) &&
                      sp.g.n[2] < (1 << 24);
    const dim3 ga((unsigned)std::max<int64_t>(1, std::min<int64_t>((items + kSigmaRuns - 1) / kSigmaRuns, 2048)));
    if (fast) {
      BEAMS_DISPATCH_FIELD(sp.g.dtype, {
        hipLaunchKernelGGL((k_sweep_sigma<FDT, true>), ga, dim3(256), lds_a, st, sp, sig);
      });
    } else {
      BEAMS_DISPATCH_FIELD(sp.g.dtype, {
        hipLaunchKernelGGL((k_sweep_sigma<FDT, false>), ga, dim3(256), lds_a, st, sp, sig);
      });
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const SweepShape sh{ plan.nt, plan.tile, plan.lo[0], plan.hi[0], plan.lo[1], plan.hi[1] };
    const dim3 grid((unsigned)((n1 + plan.tile - 1) / plan.tile), (unsigned)((n2 + plan.tile - 1) / plan.tile));
    for (int s0 = 0; s0 < n0; s0 += plan.nt) {
      if (sp.m.dtype == kF32)
        hipLaunchKernelGGL((k_sweep_block<kF32>), grid, dim3(kSweepThreads), plan.lds, st, sp, sig, s0, sh);
      else
        hipLaunchKernelGGL((k_sweep_block<kF16>), grid, dim3(kSweepThreads), plan.lds, st, sp, sig, s0, sh);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      float2* t = sp.prev;
      sp.prev = sp.cur;
      sp.cur = t;
    }
    return hipSuccess;
  }
  const dim3 grid((unsigned)((n1 + kTile - 1) / kTile), (unsigned)((n2 + kTile - 1) / kTile));
  const size_t lds = sizeof(float) * (size_t)(sp.s.ntf + 2 * ncoord) + sizeof(float2) * kStage * kStage;
  for (int s = 0; s < n0; ++s) {
    const int k = sp.step > 0 ? s : n0 - 1 - s;
    switch (sp.g.dtype * 2 + sp.m.dtype) {
      case kF32 * 2 + kF32: hipLaunchKernelGGL((k_sweep_slice<kF32, kF32>), grid, dim3(256), lds, st, sp, k, s == 0); break;
      case kF32 * 2 + kF16: hipLaunchKernelGGL((k_sweep_slice<kF32, kF16>), grid, dim3(256), lds, st, sp, k, s == 0); break;
      case kF16 * 2 + kF32: hipLaunchKernelGGL((k_sweep_slice<kF16, kF32>), grid, dim3(256), lds, st, sp, k, s == 0); break;
      case kF16 * 2 + kF16: hipLaunchKernelGGL((k_sweep_slice<kF16, kF16>), grid, dim3(256), lds, st, sp, k, s == 0); break;
      case kBF16 * 2 + kF32: hipLaunchKernelGGL((k_sweep_slice<kBF16, kF32>), grid, dim3(256), lds, st, sp, k, s == 0); break;
      case kBF16 * 2 + kF16: hipLaunchKernelGGL((k_sweep_slice<kBF16, kF16>), grid, dim3(256), lds, st, sp, k, s == 0); break;
      default: return hipErrorInvalidValue;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    float2* t = sp.prev;
    sp.prev = sp.cur;
    sp.cur = t;
  }
  return hipSuccess;
}

}  // namespace beams
