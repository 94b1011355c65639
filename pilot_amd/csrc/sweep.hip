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

// Two launch shapes, the same arithmetic per vertex (bit-identical maps):
// - blocked (the default): one fully parallel pass writes every vertex's extinction sigma
//   (the only field reads), then each launch advances T slices of a square tile in LDS
//   from a halo of T R vertices, R bounding how far (in vertices) a light ray moves across a
//   slice plus one for the bilinear stencil; each slice's region shrinks by R, so no
//   workgroup ever waits on another, and the (tau, sigma) slice reaches HBM once per launch;
// - one launch per slice, reading the previous slice from HBM: for lights so oblique that
//   the halo would not fit in LDS.
// Not an exact-order path: the compiler may fuse multiply-adds; both shapes evaluate a vertex
// through the same inline functions (sweep_pos, sweep_sigma, sweep_tau, sweep_store).
#pragma clang fp contract(fast)

namespace beams {

namespace {

constexpr int kTile = 16;
constexpr int kHalo = 3;
constexpr int kStage = kTile + 2 * kHalo + 1;

// Trilinear field value at a point inside the block (cell containing p, max face -> last cell);
// rw[d] holds the reciprocal cell widths of axis d.
template <int FDT>
__device__ __forceinline__ float field_at(const GridParams& g, const float* const* x, const float* const* rw,
                                          const float* p)
{
  int64_t c[3];
  float t[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[d] = (p[d] >= g.maxp[d]) ? g.n[d] - 2 : walk_cell(x[d], g.n[d] - 1, p[d], g.gx0[d], g.ginv[d]);
    t[d] = fminf(fmaxf((p[d] - x[d][c[d]]) * rw[d][c[d]], 0.0f), 1.0f);
  }
  const int64_t sx = g.field_stride[0], sy = g.field_stride[1], sz = g.field_stride[2];
  const int64_t b = c[0] * sx + c[1] * sy + c[2] * sz;
  const float v0 = load_t<FDT>(g.field, b), v1 = load_t<FDT>(g.field, b + sx);
  const float v2 = load_t<FDT>(g.field, b + sx + sy), v3 = load_t<FDT>(g.field, b + sy);
  const float v4 = load_t<FDT>(g.field, b + sz), v5 = load_t<FDT>(g.field, b + sz + sx);
  const float v6 = load_t<FDT>(g.field, b + sz + sx + sy), v7 = load_t<FDT>(g.field, b + sz + sy);
  const float x0 = v0 + t[0] * (v1 - v0), x1 = v3 + t[0] * (v2 - v3);
  const float x2 = v4 + t[0] * (v5 - v4), x3 = v7 + t[0] * (v6 - v7);
  const float y0 = x0 + t[1] * (x1 - x0), y1 = x2 + t[1] * (x3 - x2);
  return y0 + t[2] * (y1 - y0);
}

// Staging of the TF's extinction per entry, the coordinates and their reciprocal cell widths
// (LDS: ntf + 2 ncoord floats; xs[ncoord + i] = 1 / (x[i+1] - x[i]) within each axis).
__device__ __forceinline__ void stage_sweep(const SweepParams& sp, float* sig_tab, float* xs)
{
  const GridParams& g = sp.g;
  const int ntf = sp.s.ntf, n0 = (int)g.n[0], n01 = (int)(g.n[0] + g.n[1]), ncoord = (int)(g.n[0] + g.n[1] + g.n[2]);
  for (int i = threadIdx.x; i < ntf; i += blockDim.x) {
    const float a = fminf(sp.s.tf[4 * i + 3] / sp.s.max_density, 0.99999994f);
    sig_tab[i] = -__logf(1.0f - a) * sp.inv_h;
  }
  for (int i = threadIdx.x; i < ncoord; i += blockDim.x) {
    const int d = i < n0 ? 0 : (i < n01 ? 1 : 2), j = i - (d == 0 ? 0 : (d == 1 ? n0 : n01));
    const float* xa = g.x[d];
    xs[i] = xa[j];
    xs[ncoord + i] = j + 1 < (int)g.n[d] ? 1.0f / (xa[j + 1] - xa[j]) : 0.0f;
  }
}

// Position of map vertex (k along the sweep axis, i1, i2 across).
__device__ __forceinline__ void sweep_pos(const SweepParams& sp, int k, int i1, int i2, float* v)
{
  const MapParams& m = sp.m;
  v[sp.axis] = m.morg[sp.axis] + m.msp[sp.axis] * (float)k;
  v[sp.a1] = m.morg[sp.a1] + m.msp[sp.a1] * (float)i1;
  v[sp.a2] = m.morg[sp.a2] + m.msp[sp.a2] * (float)i2;
}

// Extinction at a map vertex: the field there, its TF entry, that entry's sigma.
template <int FDT>
__device__ __forceinline__ float sweep_sigma(const SweepParams& sp, const float* const* x, const float* const* rw,
                                             const float* sig_tab, const float* v)
{
  const int ntf = sp.s.ntf;
  const float f = field_at<FDT>(sp.g, x, rw, v);
  int64_t idx = (int64_t)(((f - sp.s.min_scalar) * sp.s.inv_delta) * (float)(ntf - 1));
  idx = idx < ntf - 1 ? idx : ntf - 1;
  idx = idx > 0 ? idx : 0;
  return sig_tab[idx];
}

// Optical depth at vertex v (not on the first slice) from the previous slice: q(c1, c2) is
// the previous slice's (tau, sigma) at transverse vertex (c1, c2).
template <class Prev>
__device__ __forceinline__ float sweep_tau(const SweepParams& sp, const float* v, int i1, int i2, float sig_v,
                                           const Prev& q)
{
  const MapParams& m = sp.m;
  const int a0 = sp.axis, a1 = sp.a1, a2 = sp.a2;
  const int n1 = (int)m.mp[a1], n2 = (int)m.mp[a2];
  float d[3];
  if (m.light_kind == 1) {
    d[0] = m.light[0];
    d[1] = m.light[1];
    d[2] = m.light[2];
  } else {
    const float dd[3] = { v[0] - m.light[0], v[1] - m.light[1], v[2] - m.light[2] };
    const float rs = rsqrtf(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
    d[0] = dd[0] * rs;
    d[1] = dd[1] * rs;
    d[2] = dd[2] * rs;
  }
  const float tb = m.msp[a0] * __builtin_amdgcn_rcpf(fabsf(d[a0]));  // distance back to the previous plane
  const float u1 = (float)i1 - tb * d[a1] * m.minv[a1];
  const float u2 = (float)i2 - tb * d[a2] * m.minv[a2];
  if (u1 >= 0.0f && u2 >= 0.0f && u1 <= (float)(n1 - 1) && u2 <= (float)(n2 - 1)) {
    const int c1 = min((int)u1, n1 - 2 < 0 ? 0 : n1 - 2), c2 = min((int)u2, n2 - 2 < 0 ? 0 : n2 - 2);
    const float f1 = u1 - (float)c1, f2 = u2 - (float)c2;
    const int c1b = min(c1 + 1, n1 - 1), c2b = min(c2 + 1, n2 - 1);
    const float2 q0 = q(c1, c2), q1 = q(c1b, c2), q2 = q(c1, c2b), q3 = q(c1b, c2b);
    const float tp = (q0.x + f1 * (q1.x - q0.x)) + f2 * ((q2.x + f1 * (q3.x - q2.x)) - (q0.x + f1 * (q1.x - q0.x)));
    const float sp_ = (q0.y + f1 * (q1.y - q0.y)) + f2 * ((q2.y + f1 * (q3.y - q2.y)) - (q0.y + f1 * (q1.y - q0.y)));
    return tp + 0.5f * (sp_ + sig_v) * tb;
  }
  // entered through a side face between the two planes
  float te = tb;
  const int ax[2] = { a1, a2 };
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int a = ax[s];
    const float lo = m.morg[a], hi = m.morg[a] + m.msp[a] * (float)(m.mp[a] - 1);
    if (d[a] > 0.0f) te = fminf(te, (v[a] - lo) / d[a]);
    if (d[a] < 0.0f) te = fminf(te, (hi - v[a]) / -d[a]);
  }
  return sig_v * fmaxf(te, 0.0f);
}

template <int MDT>
__device__ __forceinline__ void sweep_store(const SweepParams& sp, int k, int i1, int i2, float tau, float sig_v)
{
  const MapParams& m = sp.m;
  const int64_t mi = (int64_t)i1 * m.ms[sp.a1] + (int64_t)i2 * m.ms[sp.a2] + (int64_t)k * m.ms[sp.axis];
  store_t<MDT>(m.map, mi, 1.0f - __expf(-(tau + 0.5f * sp.h * sig_v)));
}

// ---- one launch per slice -----------------------------------------------------------------
template <int FDT, int MDT>
__global__ void __launch_bounds__(256) k_sweep_slice(SweepParams sp, int32_t k, int32_t first)
{
  if (gate_closed(sp.m)) return;
  extern __shared__ float lds[];
  const GridParams& g = sp.g;
  const MapParams& m = sp.m;
  const int ntf = sp.s.ntf;
  float* sig_tab = lds;
  float* xs = lds + ntf;
  const int ncoord = (int)(g.n[0] + g.n[1] + g.n[2]);
  float2* stage = (float2*)(xs + 2 * ncoord);
  stage_sweep(sp, sig_tab, xs);
  const int n1 = (int)m.mp[sp.a1], n2 = (int)m.mp[sp.a2];
  const int base1 = blockIdx.x * kTile - kHalo, base2 = blockIdx.y * kTile - kHalo;
  if (!first)
    for (int i = threadIdx.x; i < kStage * kStage; i += blockDim.x) {
      const int u = min(max(base1 + i % kStage, 0), n1 - 1);
      const int w = min(max(base2 + i / kStage, 0), n2 - 1);
      stage[i] = sp.prev[(int64_t)w * n1 + u];
    }
  __syncthreads();
  const float* x[3] = { xs, xs + g.n[0], xs + g.n[0] + g.n[1] };
  const float* rw[3] = { x[0] + ncoord, x[1] + ncoord, x[2] + ncoord };
  const int i1 = blockIdx.x * kTile + (threadIdx.x % kTile);
  const int i2 = blockIdx.y * kTile + (threadIdx.x / kTile);
  if (i1 >= n1 || i2 >= n2) return;
  float v[3];
  sweep_pos(sp, k, i1, i2, v);
  const float sig_v = sweep_sigma<FDT>(sp, x, rw, sig_tab, v);
  float tau = 0.0f;
  if (!first) {
    auto q = [&](int c1, int c2) -> float2 {
      const int l1 = c1 - base1, l2 = c2 - base2;
      if (l1 >= 0 && l2 >= 0 && l1 < kStage && l2 < kStage) return stage[l2 * kStage + l1];
      return sp.prev[(int64_t)c2 * n1 + c1];
    };
    tau = sweep_tau(sp, v, i1, i2, sig_v, q);
  }
  sp.cur[(int64_t)i2 * n1 + i1] = make_float2(tau, sig_v);
  sweep_store<MDT>(sp, k, i1, i2, tau, sig_v);
}

// ---- blocked -------------------------------------------------------------------------------
// Every vertex's sigma, slice by slice in sweep order (slice s = the s-th one the light crosses),
// a1 fastest: sig[(s * n2 + i2) * n1 + i1].
template <int FDT>
__global__ void __launch_bounds__(256) k_sweep_sigma(SweepParams sp, float* sig, int64_t nv)
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
  const int n0 = (int)sp.m.mp[sp.axis], n1 = (int)sp.m.mp[sp.a1], n2 = (int)sp.m.mp[sp.a2];
#pragma unroll 2
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nv; q += (int64_t)gridDim.x * blockDim.x) {
    const int i1 = (int)(q % n1), i2 = (int)((q / n1) % n2), s = (int)(q / ((int64_t)n1 * n2));
    const int k = sp.step > 0 ? s : n0 - 1 - s;
    float v[3];
    sweep_pos(sp, k, i1, i2, v);
    sig[q] = sweep_sigma<FDT>(sp, x, rw, sig_tab, v);
  }
}

constexpr int kTileMin = 32, kTileMax = 64;

// Slices s0 .. s0 + nt - 1 of a tile x tile tile: the previous slice's (tau, sigma) over the tile
// plus a halo of nt R vertices is staged from HBM (clamped to the slice), and the sigma of
// every slice's region at once (all loads in flight together); then slice t is computed from
// LDS alone over the tile plus (nt - 1 - t) R vertices, whose stencils lie within the region
// of slice t - 1 because no light ray moves more than R - 1 vertices across a slice.  The
// last slice's tile goes back to HBM for the next launch.
__host__ __device__ inline int sweep_region(int tile, int nt, int R, int t) { return tile + 2 * (nt - 1 - t) * R; }

template <int MDT>
__global__ void __launch_bounds__(1024) k_sweep_block(SweepParams sp, const float* sig, int32_t s0, int32_t nt,
                                                      int32_t R, int32_t tile)
{
  if (gate_closed(sp.m)) return;
  extern __shared__ float2 buf[];
  const MapParams& m = sp.m;
  const int n0 = (int)m.mp[sp.axis], n1 = (int)m.mp[sp.a1], n2 = (int)m.mp[sp.a2];
  const int halo = nt * R, W = tile + 2 * halo;
  const int o1 = blockIdx.x * tile - halo, o2 = blockIdx.y * tile - halo;
  const int nts = min(nt, n0 - s0);  // slices of this launch
  float2* prev = buf;
  float2* cur = buf + W * W;
  float* sg_lds = (float*)(buf + 2 * W * W);  // slice t's region at offset sum of the earlier regions
  // Staging, eight loads in flight per thread before their LDS stores: the previous slice's
  // (tau, sigma) (clamped to the slice), then every slice's sigma region.
  constexpr int kBatch = 8;
  if (s0 > 0)
    for (int base = 0; base < W * W; base += kBatch * (int)blockDim.x) {
      float2 v[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int i = base + u * (int)blockDim.x + (int)threadIdx.x;
        const int u1 = min(max(o1 + i % W, 0), n1 - 1), w2 = min(max(o2 + i / W, 0), n2 - 1);
        v[u] = i < W * W ? sp.prev[(int64_t)w2 * n1 + u1] : make_float2(0.0f, 0.0f);
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int i = base + u * (int)blockDim.x + (int)threadIdx.x;
        if (i < W * W) prev[i] = v[u];
      }
    }
  int total = 0;
  for (int t = 0; t < nts; ++t) total += sweep_region(tile, nt, R, t) * sweep_region(tile, nt, R, t);
  for (int base = 0; base < total; base += kBatch * (int)blockDim.x) {
    float v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      int p = base + u * (int)blockDim.x + (int)threadIdx.x, t = 0, L = sweep_region(tile, nt, R, 0);
      v[u] = 0.0f;
      if (p < total) {
        while (p >= L * L) {
          p -= L * L;
          ++t;
          L = sweep_region(tile, nt, R, t);
        }
        const int l0 = halo - (nt - 1 - t) * R;
        const int i1 = min(max(o1 + l0 + p % L, 0), n1 - 1), i2 = min(max(o2 + l0 + p / L, 0), n2 - 1);
        v[u] = sig[(int64_t)(s0 + t) * n1 * n2 + (int64_t)i2 * n1 + i1];
      }
    }
#pragma unroll
    for (int u = 0; u < kBatch; ++u) {
      const int p = base + u * (int)blockDim.x + (int)threadIdx.x;
      if (p < total) sg_lds[p] = v[u];
    }
  }
  __syncthreads();
  int off = 0;
  for (int t = 0; t < nts; ++t) {
    const int s = s0 + t, k = sp.step > 0 ? s : n0 - 1 - s;
    const int rt = (nt - 1 - t) * R, L = tile + 2 * rt, l0 = halo - rt;
    const bool last = t == nts - 1;
    for (int p = threadIdx.x; p < L * L; p += blockDim.x) {
      const int l1 = l0 + p % L, l2 = l0 + p / L;
      const int i1 = o1 + l1, i2 = o2 + l2;
      if (i1 < 0 || i2 < 0 || i1 >= n1 || i2 >= n2) continue;
      const float sig_v = sg_lds[off + p];
      float tau = 0.0f;
      float v[3];
      sweep_pos(sp, k, i1, i2, v);
      if (s > 0) {
        auto q = [&](int c1, int c2) -> float2 {
          const int a = min(max(c1 - o1, 0), W - 1), b = min(max(c2 - o2, 0), W - 1);
          return prev[b * W + a];
        };
        tau = sweep_tau(sp, v, i1, i2, sig_v, q);
      }
      cur[l2 * W + l1] = make_float2(tau, sig_v);
      if (l1 >= halo && l2 >= halo && l1 < halo + tile && l2 < halo + tile) {
        sweep_store<MDT>(sp, k, i1, i2, tau, sig_v);
        if (last) sp.cur[(int64_t)i2 * n1 + i1] = make_float2(tau, sig_v);
      }
    }
    off += L * L;
    __syncthreads();
    float2* tmp = prev;
    prev = cur;
    cur = tmp;
  }
}

}  // namespace

SweepPlan sweep_plan(const SweepParams& sp, const double* bounds)
{
  const MapParams& m = sp.m;
  const int a0 = sp.axis;
  SweepPlan plan{};
  // the largest transverse move of a light ray across one slice, in vertices
  double worst = 0.0;
  const int ax[2] = { sp.a1, sp.a2 };
  for (int s = 0; s < 2; ++s) {
    const int a = ax[s];
    double ratio;
    if (m.light_kind == 1) {
      ratio = std::fabs((double)m.light[a]) / std::max(std::fabs((double)m.light[a0]), 1e-300);
    } else {
      // nearest plane after the first one the light crosses, and the farthest transverse offset
      const double near0 = std::min(std::fabs(bounds[2 * a0] - (double)m.light[a0]),
                                    std::fabs(bounds[2 * a0 + 1] - (double)m.light[a0])) + (double)m.msp[a0];
      const double far = std::max(std::fabs(bounds[2 * a] - (double)m.light[a]),
                                  std::fabs(bounds[2 * a + 1] - (double)m.light[a]));
      ratio = far / near0;
    }
    worst = std::max(worst, ratio * (double)m.msp[a0] / (double)m.msp[a]);
  }
  const double margin = worst * (1.0 + 1e-4) + 1e-4;
  const int R = margin < 64.0 ? (int)std::ceil(margin) + 1 : 1 << 20;
  plan.R = R;
  plan.blocked = R <= 8;
  // the tile: the narrowest (a multiple of 4 from 32) whose grid fits the 256 CUs in one wave
  const int n1 = (int)m.mp[sp.a1], n2 = (int)m.mp[sp.a2];
  plan.tile = kTileMax;
  for (int t = kTileMin; t <= kTileMax; t += 4)
    if ((int64_t)((n1 + t - 1) / t) * ((n2 + t - 1) / t) <= 256) {
      plan.tile = t;
      break;
    }
  // the most slices per launch (at most 16) whose (tau, sigma) buffers and sigma regions fit in
  // 150 KB of LDS
  for (int nt = std::max(1, std::min(16, 16 / std::max(R, 1))); nt >= 1; --nt) {
    const size_t W = (size_t)(plan.tile + 2 * nt * R);
    size_t lds = sizeof(float2) * 2 * W * W;
    for (int t = 0; t < nt; ++t)
      lds += sizeof(float) * (size_t)sweep_region(plan.tile, nt, R, t) * (size_t)sweep_region(plan.tile, nt, R, t);
    plan.nt = nt;
    plan.lds = lds;
    if (lds <= 150 * 1024) break;
  }
  plan.blocked = plan.blocked && plan.lds <= 150 * 1024;
  return plan;
}

hipError_t launch_sweep(SweepParams sp, const SweepPlan& plan, float* sig, hipStream_t st)
{
  const int n0 = (int)sp.m.mp[sp.axis], n1 = (int)sp.m.mp[sp.a1], n2 = (int)sp.m.mp[sp.a2];
  const int ncoord = (int)(sp.g.n[0] + sp.g.n[1] + sp.g.n[2]);
  if (plan.blocked && sig) {
    const int64_t nv = (int64_t)n0 * n1 * n2;
    const size_t lds_a = sizeof(float) * (size_t)(sp.s.ntf + 2 * ncoord);
    const dim3 ga((unsigned)std::max<int64_t>(1, std::min<int64_t>((nv + 255) / 256, 4096)));
    BEAMS_DISPATCH_FIELD(sp.g.dtype, {
      hipLaunchKernelGGL((k_sweep_sigma<FDT>), ga, dim3(256), lds_a, st, sp, sig, nv);
    });
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const dim3 grid((unsigned)((n1 + plan.tile - 1) / plan.tile), (unsigned)((n2 + plan.tile - 1) / plan.tile));
    for (int s0 = 0; s0 < n0; s0 += plan.nt) {
      if (sp.m.dtype == kF32)
        hipLaunchKernelGGL((k_sweep_block<kF32>), grid, dim3(1024), plan.lds, st, sp, sig, s0, plan.nt, plan.R,
                           plan.tile);
      else
        hipLaunchKernelGGL((k_sweep_block<kF16>), grid, dim3(1024), plan.lds, st, sp, sig, s0, plan.nt, plan.R,
                           plan.tile);
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
