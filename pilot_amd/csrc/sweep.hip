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

// Not an exact-order path: let the compiler fuse multiply-adds here.
#pragma clang fp contract(fast)

namespace beams {

namespace {

constexpr int kTile = 16;
constexpr int kHalo = 3;
constexpr int kStage = kTile + 2 * kHalo + 1;

// Trilinear field value at a point inside the block (cell containing p, max face -> last cell).
template <int FDT>
__device__ __forceinline__ float field_at(const GridParams& g, const float* const* x, const float* p)
{
  int64_t c[3];
  float t[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    c[d] = (p[d] >= g.maxp[d]) ? g.n[d] - 2 : walk_cell(x[d], g.n[d] - 1, p[d], g.gx0[d], g.ginv[d]);
    const float lo = x[d][c[d]], hi = x[d][c[d] + 1];
    t[d] = fminf(fmaxf((p[d] - lo) / (hi - lo), 0.0f), 1.0f);
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
  float2* stage = (float2*)(xs + ((ncoord + 1) & ~1));
  for (int i = threadIdx.x; i < ntf; i += blockDim.x) {
    const float a = fminf(sp.s.tf[4 * i + 3] / sp.s.max_density, 0.99999994f);
    sig_tab[i] = -__logf(1.0f - a) * sp.inv_h;
  }
  for (int i = threadIdx.x; i < ncoord; i += blockDim.x)
    xs[i] = (i < g.n[0]) ? g.x[0][i] : ((i < g.n[0] + g.n[1]) ? g.x[1][i - g.n[0]] : g.x[2][i - g.n[0] - g.n[1]]);
  const int a0 = sp.axis, a1 = sp.a1, a2 = sp.a2;
  const int n1 = (int)m.mp[a1], n2 = (int)m.mp[a2];
  const int base1 = blockIdx.x * kTile - kHalo, base2 = blockIdx.y * kTile - kHalo;
  if (!first)
    for (int i = threadIdx.x; i < kStage * kStage; i += blockDim.x) {
      const int u = min(max(base1 + i % kStage, 0), n1 - 1);
      const int w = min(max(base2 + i / kStage, 0), n2 - 1);
      stage[i] = sp.prev[(int64_t)w * n1 + u];
    }
  __syncthreads();
  const float* x[3] = { xs, xs + g.n[0], xs + g.n[0] + g.n[1] };
  const int i1 = blockIdx.x * kTile + (threadIdx.x % kTile);
  const int i2 = blockIdx.y * kTile + (threadIdx.x / kTile);
  if (i1 >= n1 || i2 >= n2) return;
  float v[3];
  v[a0] = m.morg[a0] + m.msp[a0] * (float)k;
  v[a1] = m.morg[a1] + m.msp[a1] * (float)i1;
  v[a2] = m.morg[a2] + m.msp[a2] * (float)i2;
  // extinction at the vertex
  const float f = field_at<FDT>(g, x, v);
  int64_t idx = (int64_t)(((f - sp.s.min_scalar) * sp.s.inv_delta) * (float)(ntf - 1));
  idx = idx < ntf - 1 ? idx : ntf - 1;
  idx = idx > 0 ? idx : 0;
  const float sig_v = sig_tab[idx];
  float tau = 0.0f;
  if (!first) {
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
    const float tb = m.msp[a0] / fabsf(d[a0]);  // distance back to the previous slice plane
    const float u1 = (float)i1 - tb * d[a1] * m.minv[a1];
    const float u2 = (float)i2 - tb * d[a2] * m.minv[a2];
    if (u1 >= 0.0f && u2 >= 0.0f && u1 <= (float)(n1 - 1) && u2 <= (float)(n2 - 1)) {
      const int c1 = min((int)u1, n1 - 2 < 0 ? 0 : n1 - 2), c2 = min((int)u2, n2 - 2 < 0 ? 0 : n2 - 2);
      const float f1 = u1 - (float)c1, f2 = u2 - (float)c2;
      float2 q[4];
      const int l1 = c1 - base1, l2 = c2 - base2;
      if (l1 >= 0 && l2 >= 0 && l1 + 1 < kStage && l2 + 1 < kStage) {
        q[0] = stage[l2 * kStage + l1];
        q[1] = stage[l2 * kStage + l1 + 1];
        q[2] = stage[(l2 + 1) * kStage + l1];
        q[3] = stage[(l2 + 1) * kStage + l1 + 1];
      } else {
        const int c1b = min(c1 + 1, n1 - 1), c2b = min(c2 + 1, n2 - 1);
        q[0] = sp.prev[(int64_t)c2 * n1 + c1];
        q[1] = sp.prev[(int64_t)c2 * n1 + c1b];
        q[2] = sp.prev[(int64_t)c2b * n1 + c1];
        q[3] = sp.prev[(int64_t)c2b * n1 + c1b];
      }
      const float tp = (q[0].x + f1 * (q[1].x - q[0].x)) +
                       f2 * ((q[2].x + f1 * (q[3].x - q[2].x)) - (q[0].x + f1 * (q[1].x - q[0].x)));
      const float sp_ = (q[0].y + f1 * (q[1].y - q[0].y)) +
                        f2 * ((q[2].y + f1 * (q[3].y - q[2].y)) - (q[0].y + f1 * (q[1].y - q[0].y)));
      tau = tp + 0.5f * (sp_ + sig_v) * tb;
    } else {
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
      tau = sig_v * fmaxf(te, 0.0f);
    }
  }
  sp.cur[(int64_t)i2 * n1 + i1] = make_float2(tau, sig_v);
  const int64_t mi = (int64_t)i1 * m.ms[a1] + (int64_t)i2 * m.ms[a2] + (int64_t)k * m.ms[a0];
  store_t<MDT>(m.map, mi, 1.0f - __expf(-(tau + 0.5f * sp.h * sig_v)));
}

}  // namespace

hipError_t launch_sweep(SweepParams sp, hipStream_t st)
{
  const int n0 = (int)sp.m.mp[sp.axis], n1 = (int)sp.m.mp[sp.a1], n2 = (int)sp.m.mp[sp.a2];
  const dim3 grid((unsigned)((n1 + kTile - 1) / kTile), (unsigned)((n2 + kTile - 1) / kTile));
  const int ncoord = (int)(sp.g.n[0] + sp.g.n[1] + sp.g.n[2]);
  const size_t lds = sizeof(float) * (size_t)(sp.s.ntf + ((ncoord + 1) & ~1)) +
                     sizeof(float2) * kStage * kStage;
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
