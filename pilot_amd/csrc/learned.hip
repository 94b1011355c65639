// Kernels of the learned shadow closure (learned.hpp): the coarse extinction mip and its
// optical depth, the learned map (features + network on the matrix cores per wave), the
// sparse exact check, and training samples drawn from an exact map.
#include "learned.hpp"
#include "march_dev.hpp"
#include "mlp.hpp"
#include "workqueue.hpp"

#include <algorithm>

namespace beams {

namespace {

constexpr int kBlock = 256;

// First index i in [0, n) with x[i] >= v (n when none).
__device__ __forceinline__ int64_t lower_index(const float* x, int64_t n, float v)
{
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (x[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Index of the grid coordinate nearest v (ties to the lower one), v clamped to the grid: from
// the uniform-spacing guess i, walk down while x[i] > v, then up while x[i+1] < v.  The guess's
// two coordinates are read first, without waiting on each other, and the walks run only when
// one of their tests would move i (otherwise neither loop would step), so the index is the
// walk's.
__device__ __forceinline__ int64_t nearest_index(const float* x, int64_t n, float v, float x0, float ginv)
{
  int64_t i = (int64_t)floorf((v - x0) * ginv);
  i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
  float a = x[i], b = x[i + 1];
  if (((int)(i > 0) & (int)(a > v)) | ((int)(i < n - 2) & (int)(b < v))) {
    while (i > 0 && x[i] > v) --i;
    while (i < n - 2 && x[i + 1] < v) ++i;
    a = x[i];
    b = x[i + 1];
  }
  return (v - a <= b - v) ? i : i + 1;
}

// nearest_index on the three axes of v in 32 bits (an axis has fewer than 2^31 points): the
// same indices, with the six guess coordinates read together before any test.
__device__ __forceinline__ void nearest_index3(const GridParams& g, const float* v, int* c)
{
  int i[3], n[3];
  float a[3], b[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    n[d] = (int)g.n[d];
    i[d] = (int)fminf(fmaxf(floorf((v[d] - g.gx0[d]) * g.ginv[d]), 0.0f), (float)(n[d] - 2));
    a[d] = g.x[d][i[d]];
    b[d] = g.x[d][i[d] + 1];
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float* x = g.x[d];
    if (((int)(i[d] > 0) & (int)(a[d] > v[d])) | ((int)(i[d] < n[d] - 2) & (int)(b[d] < v[d]))) {
      while (i[d] > 0 && x[i[d]] > v[d]) --i[d];
      while (i[d] < n[d] - 2 && x[i[d] + 1] < v[d]) ++i[d];
      a[d] = x[i[d]];
      b[d] = x[i[d] + 1];
    }
    c[d] = (v[d] - a[d] <= b[d] - v[d]) ? i[d] : i[d] + 1;
  }
}

// The TF entry of a scalar as the reference indexes it (truncate, then clamp to the table);
// clamping in float first gives the same index for every finite value and keeps it 32-bit.
__device__ __forceinline__ int tf_index(const ShadeParams& s, float val)
{
  const float tfmax = (float)(s.ntf - 1);
  const float t = ((val - s.min_scalar) * s.inv_delta) * tfmax;
  return (int)fminf(fmaxf(t, 0.0f), tfmax);
}

__device__ __forceinline__ void stage_alpha(const ShadeParams& s, float* a_tab)
{
  for (int i = threadIdx.x; i < s.ntf; i += blockDim.x) a_tab[i] = s.tf[4 * i + 3] / s.max_density;
}

// Density-patch feature per TF entry: -ln(1 - min(a, 1 - 2^-24)) / 4.
__device__ __forceinline__ void stage_patch(const ShadeParams& s, float* d_tab)
{
  for (int i = threadIdx.x; i < s.ntf; i += blockDim.x)
    d_tab[i] = -log1pf(-fminf(s.tf[4 * i + 3] / s.max_density, 1.0f - 5.9604645e-8f)) * 0.25f;
}

// Trilinear interpolation in a cn^3 vertex grid (x fastest), p clamped to the grid.
__device__ __forceinline__ float coarse_lerp(const LearnedParams& p, const float* grid, const float* v)
{
  int c[3];
  float t[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float u = (v[d] - p.corg[d]) * p.cinv[d];
    u = fminf(fmaxf(u, 0.0f), (float)(p.cn - 1));
    int i = (int)u;
    i = i > p.cn - 2 ? p.cn - 2 : i;
    c[d] = i;
    t[d] = u - (float)i;
  }
  const int n = p.cn, b = c[0] + n * (c[1] + n * c[2]);
  const float v00 = grid[b] + t[0] * (grid[b + 1] - grid[b]);
  const float v10 = grid[b + n] + t[0] * (grid[b + n + 1] - grid[b + n]);
  const float v01 = grid[b + n * n] + t[0] * (grid[b + n * n + 1] - grid[b + n * n]);
  const float v11 = grid[b + n * n + n] + t[0] * (grid[b + n * n + n + 1] - grid[b + n * n + n]);
  const float w0 = v00 + t[1] * (v10 - v00);
  const float w1 = v01 + t[1] * (v11 - v01);
  return w0 + t[2] * (w1 - w0);
}

// Light direction of travel at v (unit).
__device__ __forceinline__ void light_dir(const MapParams& m, const float* v, float* d)
{
  if (m.light_kind == 1) {
    d[0] = m.light[0];
    d[1] = m.light[1];
    d[2] = m.light[2];
    return;
  }
  float dd[3] = { v[0] - m.light[0], v[1] - m.light[1], v[2] - m.light[2] };
  const float l2 = dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2];
  const float rs = l2 > 0.0f ? 1.0f / sqrtf(l2) : 0.0f;
  d[0] = dd[0] * rs;
  d[1] = dd[1] * rs;
  d[2] = dd[2] * rs;
}

// Features of map vertex v (d_tab: the staged density-patch table); returns min(tau_c, 8).
template <int FDT>
__device__ __forceinline__ float features(const LearnedParams& p, const float* d_tab, const float* v,
                                          float* f)
{
  const GridParams& g = p.g;
  const float tc = fminf(coarse_lerp(p, p.coarse_tau, v), kTauMax);
  f[0] = tc * 0.25f;
  if (p.tfidx) {
    // the TF entries of the field points, precomputed this frame: one byte per gather, through
    // a buffer resource over the index field (fewer than 2^32 points: 32-bit offsets, no
    // 64-bit address per gather); the clamped neighbour offsets per axis
    uint32_t b[3][3];
    int cc[3];
    nearest_index3(g, v, cc);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int n = (int)g.n[d], c = cc[d];
      const uint32_t st = (uint32_t)g.field_stride[d], oc = (uint32_t)c * st;
      b[d][0] = c > 0 ? oc - st : oc;
      b[d][1] = oc;
      b[d][2] = c < n - 1 ? oc + st : oc;
    }
    // readfirstlane returns int: each half goes through uint32_t so the low one is not
    // sign-extended into the high one when the OR forms the 64-bit base
    const uint64_t tp = (uint64_t)p.tfidx;
    const uint32_t tlo = __builtin_amdgcn_readfirstlane((uint32_t)tp);
    const uint32_t thi = __builtin_amdgcn_readfirstlane((uint32_t)(tp >> 32));
    void* tbase = (void*)(((uint64_t)thi << 32) | tlo);
    const uint32_t npts = (uint32_t)(g.n[0] * g.n[1] * g.n[2]);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(tbase, (short)0, (int)npts, 0x00020000);
#pragma unroll
    for (int zy = 0; zy < 9; ++zy) {
      const uint32_t row = b[2][zy / 3] + b[1][zy % 3];
#pragma unroll
      for (int t = 0; t < 3; ++t)
        f[1 + 3 * zy + t] = d_tab[__builtin_amdgcn_raw_buffer_load_b8(rs, (int)(row + b[0][t]), 0, 0)];
    }
    light_dir(p.m, v, f + 28);
    f[31] = 0.0f;
    return tc;
  }
  // element offsets of the clamped neighbour indices per axis
  int64_t o[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int64_t c = nearest_index(g.x[d], g.n[d], v[d], g.gx0[d], g.ginv[d]);
    const int64_t lo = c > 0 ? c - 1 : 0, hi = c < g.n[d] - 1 ? c + 1 : g.n[d] - 1;
    o[d][0] = lo * g.field_stride[d];
    o[d][1] = c * g.field_stride[d];
    o[d][2] = hi * g.field_stride[d];
  }
  float val[27];
  if (p.field_small) {
    // byte offsets below 4 GiB: one 32-bit add per gather from the uniform field base
    constexpr uint32_t es = FDT == kF32 ? 4u : 2u;
    uint32_t b[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int t = 0; t < 3; ++t) b[d][t] = (uint32_t)o[d][t] * es;
    const char* base = (const char*)g.field;
#pragma unroll
    for (int q = 0; q < 27; ++q) {
      const uint32_t off = b[2][q / 9] + b[1][(q / 3) % 3] + b[0][q % 3];
      if constexpr (FDT == kF32)
        val[q] = *(const float*)(base + off);
      else if constexpr (FDT == kF16)
        val[q] = f16_to_f32(*(const uint16_t*)(base + off));
      else
        val[q] = bf16_to_f32(*(const uint16_t*)(base + off));
    }
  } else {
#pragma unroll
    for (int q = 0; q < 27; ++q) val[q] = load_t<FDT>(g.field, o[2][q / 9] + o[1][(q / 3) % 3] + o[0][q % 3]);
  }
#pragma unroll
  for (int q = 0; q < 27; ++q) f[1 + q] = d_tab[tf_index(p.s, val[q])];
  light_dir(p.m, v, f + 28);
  f[31] = 0.0f;
  return tc;
}

__device__ __forceinline__ float sigma_of(float a, float inv_h)
{
  return -log1pf(-fminf(a, 1.0f - 5.9604645e-8f)) * inv_h;
}

// ---- TF entry of every field point (the feature patches' table index), one byte each -------
// Four points per thread: one 8-byte (16-bit storage) or 16-byte (FP32) load and one 4-byte
// store of their four indices (the field and the output 16-byte aligned), the last n % 4 points
// one at a time.
template <int FDT>
__global__ void __launch_bounds__(kBlock) k_tf_index_field(LearnedParams p, uint8_t* out, int64_t n)
{
  const int64_t n4 = n / 4;
  for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n4; q += (int64_t)gridDim.x * kBlock) {
    float v[4];
    if constexpr (FDT == kF32) {
      const float4 w = ((const float4*)p.g.field)[q];
      v[0] = w.x;
      v[1] = w.y;
      v[2] = w.z;
      v[3] = w.w;
    } else {
      const uint2 w = ((const uint2*)p.g.field)[q];
      const uint16_t h[4] = { (uint16_t)(w.x & 0xffffu), (uint16_t)(w.x >> 16), (uint16_t)(w.y & 0xffffu),
                              (uint16_t)(w.y >> 16) };
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = FDT == kF16 ? f16_to_f32(h[u]) : bf16_to_f32(h[u]);
    }
    uint32_t packed = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) packed |= (uint32_t)tf_index(p.s, v[u]) << (8 * u);
    ((uint32_t*)out)[q] = packed;
  }
  for (int64_t q = 4 * n4 + (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock)
    out[q] = (uint8_t)tf_index(p.s, load_t<FDT>(p.g.field, q));
}

// ---- coarse mip: mean extinction over the field vertices nearest each coarse vertex -------
// Per axis and coarse index, the field indices whose coordinate lies within half a coarse
// spacing (one thread each).
__global__ void k_coarse_ranges(LearnedParams p)
{
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= 3 * p.cn) return;
  const int d = t / p.cn, ci = t % p.cn;
  const GridParams& g = p.g;
  const float c = p.corg[d] + p.csp[d] * (float)ci;
  const float half = 0.5f * p.csp[d];
  int64_t lo = lower_index(g.x[d], g.n[d], c - half);
  const int64_t hi = ci == p.cn - 1 ? g.n[d] : lower_index(g.x[d], g.n[d], c + half);
  int64_t cnt = hi - lo;
  if (cnt <= 0) {  // coarser field than mip: the nearest field vertex alone
    lo = nearest_index(g.x[d], g.n[d], c, g.gx0[d], g.ginv[d]);
    cnt = 1;
  }
  p.coarse_ranges[2 * t] = (int32_t)lo;
  p.coarse_ranges[2 * t + 1] = (int32_t)cnt;
}

// One workgroup per coarse (y, z) index pair: it reads the field rows of that pair's boxes
// once, coalesced along x, and spreads each column's sum over the (one or two) coarse x boxes
// holding that column.  A field coarser than the mip (a column in three or more boxes) sums
// each box directly instead.  Field x extent up to kCoarseMaxX points.
// Rows of 16-byte aligned vectors (x contiguous, the row extent and strides multiples of the
// vector): each thread reads 16 bytes per row (8 FP16/BF16 or 4 FP32 columns), sums its
// columns over its share of the rows in registers, and each box then sums its columns.
constexpr int kCoarseMaxX = 4097;

template <int FDT>
__global__ void __launch_bounds__(kBlock) k_coarse_sigma(LearnedParams p)
{
  constexpr int V = FDT == kF32 ? 4 : 8;  // columns per 16-byte vector
  __shared__ float sig_tab[1024];
  __shared__ int16_t first_box[kCoarseMaxX], second_box[kCoarseMaxX];
  __shared__ float colsum[kCoarseMaxX];
  __shared__ float sums[64];
  __shared__ int crowded;
  const GridParams& g = p.g;
  const int cn = p.cn, n0 = (int)g.n[0];
  for (int i = threadIdx.x; i < p.s.ntf; i += blockDim.x)
    sig_tab[i] = sigma_of(p.s.tf[4 * i + 3] / p.s.max_density, p.inv_h);
  for (int i = threadIdx.x; i < n0; i += blockDim.x) {
    first_box[i] = -1;
    second_box[i] = -1;
    colsum[i] = 0.0f;
  }
  for (int i = threadIdx.x; i < cn; i += blockDim.x) sums[i] = 0.0f;
  __syncthreads();
  if (threadIdx.x == 0) {
    crowded = 0;
    for (int cx = 0; cx < cn; ++cx) {
      const int lo = p.coarse_ranges[2 * cx], cnt = p.coarse_ranges[2 * cx + 1];
      for (int i = lo; i < lo + cnt; ++i) {
        if (first_box[i] < 0)
          first_box[i] = (int16_t)cx;
        else if (second_box[i] < 0)
          second_box[i] = (int16_t)cx;
        else
          crowded = 1;
      }
    }
  }
  __syncthreads();
  const int cj = (int)blockIdx.x % cn, ck = (int)blockIdx.x / cn;
  const int lo1 = p.coarse_ranges[2 * (cn + cj)], cnt1 = p.coarse_ranges[2 * (cn + cj) + 1];
  const int lo2 = p.coarse_ranges[2 * (2 * cn + ck)], cnt2 = p.coarse_ranges[2 * (2 * cn + ck) + 1];
  if (crowded) {
    for (int cx = 0; cx < cn; ++cx) {
      const int lo0 = p.coarse_ranges[2 * cx], cnt0 = p.coarse_ranges[2 * cx + 1];
      float acc = 0.0f;
      for (int q = threadIdx.x; q < cnt0 * cnt1 * cnt2; q += blockDim.x) {
        const int64_t i = lo0 + q % cnt0, j = lo1 + (q / cnt0) % cnt1, k = lo2 + q / (cnt0 * cnt1);
        const float val = load_t<FDT>(g.field, i * g.field_stride[0] + j * g.field_stride[1] + k * g.field_stride[2]);
        acc += sig_tab[tf_index(p.s, val)];
      }
      atomicAdd(&sums[cx], acc);
    }
  }
  const bool vec = !crowded && g.field_stride[0] == 1 && n0 % V == 0 && g.field_stride[1] % V == 0 &&
                   g.field_stride[2] % V == 0 && ((uintptr_t)g.field & 15) == 0;
  if (vec) {
    // vector column vc of the row, row phase rp of P (consecutive threads: consecutive vectors)
    const int ncv = n0 / V, cw = ncv < kBlock ? ncv : kBlock, P = kBlock / cw;
    const int rp = (int)threadIdx.x / cw, rows = cnt1 * cnt2;
    for (int vc = (int)threadIdx.x % cw; rp < P && vc < ncv; vc += cw) {
      float acc[V];
#pragma unroll
      for (int c = 0; c < V; ++c) acc[c] = 0.0f;
      int jj = rp % cnt1, kk = rp / cnt1;
      for (int row = rp; row < rows; row += P) {
        const int64_t off = (int64_t)(lo1 + jj) * g.field_stride[1] + (int64_t)(lo2 + kk) * g.field_stride[2] + vc * V;
        const uint4 w = *(const uint4*)((const char*)g.field + off * (FDT == kF32 ? 4 : 2));
        const uint32_t ws[4] = { w.x, w.y, w.z, w.w };
#pragma unroll
        for (int c = 0; c < V; ++c) {
          float val;
          if constexpr (FDT == kF32)
            val = __uint_as_float(ws[c]);
          else if constexpr (FDT == kF16)
            val = f16_to_f32((uint16_t)(ws[c / 2] >> (16 * (c & 1))));
          else
            val = bf16_to_f32((uint16_t)(ws[c / 2] >> (16 * (c & 1))));
          acc[c] += sig_tab[tf_index(p.s, val)];
        }
        jj += P;
        while (jj >= cnt1) {
          jj -= cnt1;
          ++kk;
        }
      }
#pragma unroll
      for (int c = 0; c < V; ++c) atomicAdd(&colsum[vc * V + c], acc[c]);
    }
    __syncthreads();
    for (int cx = threadIdx.x; cx < cn; cx += blockDim.x) {
      const int lo0 = p.coarse_ranges[2 * cx], cnt0 = p.coarse_ranges[2 * cx + 1];
      float acc = 0.0f;
      for (int i = lo0; i < lo0 + cnt0; ++i) acc += colsum[i];
      sums[cx] = acc;
    }
  }
  for (int i = threadIdx.x; i < (crowded || vec ? 0 : n0); i += blockDim.x) {
    float acc = 0.0f;
    for (int row = 0; row < cnt1 * cnt2; ++row) {
      const int64_t j = lo1 + row % cnt1, k = lo2 + row / cnt1;
      const float val = load_t<FDT>(g.field, (int64_t)i * g.field_stride[0] + j * g.field_stride[1] + k * g.field_stride[2]);
      acc += sig_tab[tf_index(p.s, val)];
    }
    if (first_box[i] >= 0) atomicAdd(&sums[first_box[i]], acc);
    if (second_box[i] >= 0) atomicAdd(&sums[second_box[i]], acc);
  }
  __syncthreads();
  for (int cx = threadIdx.x; cx < cn; cx += blockDim.x) {
    const int cnt0 = p.coarse_ranges[2 * cx + 1];
    p.coarse_sigma[cx + (int64_t)cn * (cj + (int64_t)cn * ck)] = sums[cx] / (float)(cnt0 * cnt1 * cnt2);
  }
}

// ---- coarse optical depth: midpoint rule toward the light through the coarse extinction ---
// One wave per coarse vertex; its lanes take every 64th midpoint and the wave sums them.
__global__ void __launch_bounds__(kBlock) k_coarse_tau(LearnedParams p)
{
  const int lane = threadIdx.x & 63;
  const int64_t cn = p.cn, id = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  if (id >= cn * cn * cn) return;  // uniform across the wave
  float v[3];
  v[0] = p.corg[0] + p.csp[0] * (float)(id % cn);
  v[1] = p.corg[1] + p.csp[1] * (float)((id / cn) % cn);
  v[2] = p.corg[2] + p.csp[2] * (float)(id / (cn * cn));
  float d[3];
  light_dir(p.m, v, d);
  // toward the light: u = -d; distance to the bounds' exit, and to a point light
  float len = INFINITY;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float u = -d[a];
    if (u > 0.0f) len = fminf(len, ((float)p.g.bounds[2 * a + 1] - v[a]) / u);
    if (u < 0.0f) len = fminf(len, ((float)p.g.bounds[2 * a] - v[a]) / u);
  }
  if (p.m.light_kind == 0) {
    const float dl[3] = { p.m.light[0] - v[0], p.m.light[1] - v[1], p.m.light[2] - v[2] };
    len = fminf(len, sqrtf(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]));
  }
  float tau = 0.0f;
  if (len > 0.0f && len < INFINITY) {
    const float hstep = 0.5f * fminf(p.csp[0], fminf(p.csp[1], p.csp[2]));
    int n = (int)ceilf(len / hstep);
    n = n < 1 ? 1 : (n > 512 ? 512 : n);
    const float dt = len / (float)n;
    for (int k = lane; k < n; k += 64) {
      const float t = ((float)k + 0.5f) * dt;
      const float q[3] = { v[0] - t * d[0], v[1] - t * d[1], v[2] - t * d[2] };
      tau += coarse_lerp(p, p.coarse_sigma, q) * dt;
    }
  }
#pragma unroll
  for (int mk = 32; mk >= 1; mk >>= 1) tau += __shfl_xor(tau, mk, 64);
  if (lane == 0) p.coarse_tau[id] = tau;
}

// ---- the learned map: 64 vertices per wave, four waves per workgroup -----------------------
// Static part of the learned-map kernels' LDS (38 KB: the waves' activation tiles and the
// biases; the weights live in registers); the density-patch table (ntf floats) is dynamic.
struct MapLds {
  MlpBias w;
  uint16_t buf[kBlock / 64][64 * kLd];
};

template <int FDT, int MDT>
// Work unit: a chunk of 64 consecutive vertices of one map row (x), one per wave; the
// vertex position is map_vertex's, computed from the row and the column without divisions
// of the 64-bit vertex index.  Workgroups are dealt to the 8 XCDs round-robin, so each XCD
// takes one contiguous eighth of the chunks: its waves walk neighbouring rows together and
// the 3 x 3 field rows of their patches stay in that XCD's L2.  The grid is a multiple of 8.
__global__ void __launch_bounds__(kBlock) k_learned_map(LearnedParams p, int32_t chunks_per_row,
                                                        int32_t nchunks)
{
  __shared__ MapLds s;
  extern __shared__ float d_tab[];
  MlpFrags fr;
  if (p.params) {
    mlp_load_frags(p.params, fr);
    mlp_stage_bias(p.params, s.w);
  }
  stage_patch(p.s, d_tab);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int mp0 = (int)p.m.mp[0], mp1 = (int)p.m.mp[1];
  const int xcd = (int)(blockIdx.x % 8), slot = (int)(blockIdx.x / 8);
  const int waves_per_xcd = (int)(gridDim.x / 8) * (kBlock / 64);
  const int first = (int)((int64_t)nchunks * xcd / 8), last = (int)((int64_t)nchunks * (xcd + 1) / 8);
  for (int chunk = first + slot * (kBlock / 64) + wave; chunk < last; chunk += waves_per_xcd) {
    const int row = chunk / chunks_per_row;
    const int i0 = (chunk - row * chunks_per_row) * 64;
    const int j = row % mp1, k = row / mp1;
    const int i = i0 + lane < mp0 ? i0 + lane : mp0 - 1;
    float v[3], f[kFeat];
    v[0] = p.m.morg[0] + p.m.msp[0] * (float)i;
    v[1] = p.m.morg[1] + p.m.msp[1] * (float)j;
    v[2] = p.m.morg[2] + p.m.msp[2] * (float)k;
    const float tc = features<FDT>(p, d_tab, v, f);
    const float y = p.params ? mlp_forward_fast(fr, s.w, s.buf[wave], f) : 0.0f;
    const float tau = fmaxf(tc - y, 0.0f);
    if (i0 + lane < mp0) store_t<MDT>(p.m.map, (int64_t)row * mp0 + i, 1.0f - expf(-tau));
  }
}

// ---- the sparse exact check ----------------------------------------------------------------

template <int FDT, int MDT>
__global__ void __launch_bounds__(kBlock) k_learned_check(LearnedParams p, int64_t ncheck)
{
  extern __shared__ float lds[];
  float* a_tab = lds;
  float* xs = lds + p.s.ntf;
  stage_alpha(p.s, a_tab);
  stage_coords(p.g, xs);
  __syncthreads();
  const float* x[3] = { xs, xs + p.g.n[0], xs + p.g.n[0] + p.g.n[1] };
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= ncheck) return;
  const MapParams& m = p.m;
  const int64_t n0 = check_count(m.mp[0] - 1, p.check_stride);
  const int64_t n1 = check_count(m.mp[1] - 1, p.check_stride);
  const int64_t qi[3] = { q % n0, (q / n0) % n1, q / (n0 * n1) };
  int64_t vi[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int64_t i = qi[d] * p.check_stride;
    vi[d] = i < m.mp[d] - 1 ? i : m.mp[d] - 1;
  }
  const int64_t id = vi[0] + m.mp[0] * (vi[1] + m.mp[1] * vi[2]);
  int64_t taken = 0;
  const float exact = march_vertex<FDT>(p.g, m, p.s, a_tab, x, id, 0.0f, &taken);
  float v[3];
  map_vertex(m, id, v);
  const float tc = fminf(coarse_lerp(p, p.coarse_tau, v), kTauMax);
  const float coarse = 1.0f - expf(-tc);
  const float learned = load_t<MDT>(m.map, id);
  if (p.verr) p.verr[q] = fabsf(learned - exact);
  atomicMax(&p.err[0], f32_bits(fabsf(learned - exact)));
  atomicMax(&p.err[1], f32_bits(fabsf(coarse - exact)));
}

// ---- training samples ---------------------------------------------------------------------
template <int FDT, int EDT>
__global__ void __launch_bounds__(kBlock) k_learned_samples(LearnedParams p, MapParams em, int64_t n,
                                                            uint64_t seed, float* x, float* t)
{
  __shared__ float a_tab[1024];
  stage_patch(p.s, a_tab);
  __syncthreads();
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t nv = em.mp[0] * em.mp[1] * em.mp[2];
  SplitMix64 rng(seed ^ (0x5851f42d4c957f2dull * (uint64_t)(k + 1)));
  const int64_t id = (int64_t)(rng.next() % (uint64_t)nv);
  float v[3], f[kFeat];
  map_vertex(p.m, id, v);
  const float tc = features<FDT>(p, a_tab, v, f);
  const int64_t i = id % em.mp[0], j = (id / em.mp[0]) % em.mp[1], kk = id / (em.mp[0] * em.mp[1]);
  const float op = load_t<EDT>(em.map, i * em.ms[0] + j * em.ms[1] + kk * em.ms[2]);
  const float te = fminf(-logf(fmaxf(1.0f - op, 1e-30f)), kTauMax);
#pragma unroll
  for (int q = 0; q < kFeat; ++q) x[k * kFeat + q] = f[q];
  t[k] = tc - te;
}

// ---- bricks: every brick's map from the whole volume's coarse optical depth --------------
// The brick of slot `v` in a table of nb + 1 ascending offsets: the last b with off[b] <= v.
__device__ __forceinline__ int brick_of(const int32_t* off, int nb, int v)
{
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= v)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// One map work item (256 vertices of one brick) per workgroup iteration, 64 per wave.
// AT_CHECKS: the items are the check slots (256 per item, lb.chk_off) and each lane's vertex
// is its check vertex.  Else, with lb.tile_fail, a wave whose vertices all lie in listed tiles
// (every vertex belongs to one tile: tile t along an axis holds [t s, (t+1) s), the last one
// up to the end) is skipped, since the fallback writes every vertex of those tiles.
template <int FDT, int MDT, bool AT_CHECKS>
__global__ void __launch_bounds__(kBlock) k_bricks_learned_map(LearnedParams p, BrickFrame f, LearnedBricks lb)
{
  __shared__ MapLds s;
  extern __shared__ float d_tab[];
  MlpFrags fr;
  if (p.params) {
    mlp_load_frags(p.params, fr);
    mlp_stage_bias(p.params, s.w);
  }
  stage_patch(p.s, d_tab);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  // The current brick's item range, vertex count, map geometry and output, in scalar registers:
  // a workgroup's items are gridDim.x apart, so it stays on a brick for many items, and the
  // brick is looked up (and its words loaded) again only when an item leaves the range.
  typedef const __attribute__((address_space(4))) int32_t ConstInt;
  ConstInt* item_off = (ConstInt*)(const void*)(AT_CHECKS ? lb.chk_off : f.map_item_off);
  ConstInt* tile_off = (ConstInt*)(const void*)lb.tile_off;
  const int nitems = AT_CHECKS ? lb.n_chk_items : f.n_map_items;
  const int s_chk = lb.stride;
  int b = 0, b_lo = 0, b_hi = 0, n0 = 0, n1 = 0, n2 = 0;
  int64_t nv = 0;
  MapParams dm;
  void* map_final = nullptr;
  for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
    if (item < b_lo || item >= b_hi) {
      if (AT_CHECKS)
        b = brick_of((const int32_t*)lb.chk_off, f.nbricks, item * kBlock);
      else
        b = item_brick(f, item);
      b = __builtin_amdgcn_readfirstlane(b);
      const BrickDesc* bk = f.bricks + b;
      b_lo = AT_CHECKS ? item_off[b] / kBlock : item_off[b];
      b_hi = AT_CHECKS ? item_off[b + 1] / kBlock : item_off[b + 1];
      nv = load_uniform(&bk->nverts);
      dm = load_uniform(&bk->m);
      map_final = load_uniform(&bk->map_final);
      if (AT_CHECKS || lb.tile_fail) {
        n0 = (int)check_count(dm.mp[0] - 1, s_chk);
        n1 = (int)check_count(dm.mp[1] - 1, s_chk);
        n2 = (int)check_count(dm.mp[2] - 1, s_chk);
      }
    }
    int64_t id;
    bool valid;
    if (AT_CHECKS) {
      // check slot q of the brick: check vertex (every s-th per axis and the last)
      const int q = (item - b_lo) * kBlock + (int)threadIdx.x;
      valid = q < n0 * n1 * n2;
      const int qq = valid ? q : 0, q1 = qq / n0, q2 = q1 / n1;
      const int qi[3] = { qq - q1 * n0, q1 - q2 * n1, q2 };
      int vi[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int i = qi[a] * s_chk, last = (int)dm.mp[a] - 1;
        vi[a] = i < last ? i : last;
      }
      id = vi[0] + dm.mp[0] * (vi[1] + dm.mp[1] * vi[2]);
    } else {
      id = (int64_t)(item - b_lo) * kMapItem + wave * 64 + lane;
      valid = id < nv;
      if (lb.tile_fail && nv <= INT32_MAX) {
        // the tile of this lane's vertex; the wave is skipped when every lane's tile is listed
        const uint32_t u = (uint32_t)(valid ? id : nv - 1), mp0 = (uint32_t)dm.mp[0], mp1 = (uint32_t)dm.mp[1];
        const uint32_t row = u / mp0, i = u - row * mp0, k = row / mp1, j = row - k * mp1;
        const int ti = min((int)(i / (uint32_t)s_chk), n0 - 2), tj = min((int)(j / (uint32_t)s_chk), n1 - 2);
        const int tk = min((int)(k / (uint32_t)s_chk), n2 - 2);
        const bool listed = !valid || lb.tile_fail[tile_off[b] + ti + (n0 - 1) * (tj + (n1 - 1) * tk)] != 0;
        if (__ballot(!listed) == 0) continue;
      }
    }
    float v[3], feat[kFeat];
    // (a brick's map holds fewer than 2^31 vertices; the one-block frame's map may not)
    if (nv <= INT32_MAX)
      map_vertex32(dm, (uint32_t)(valid ? id : (AT_CHECKS ? 0 : nv - 1)), v);
    else
      map_vertex(dm, valid ? id : (AT_CHECKS ? 0 : nv - 1), v);
    const float tc = features<FDT>(p, d_tab, v, feat);
    const float y = p.params ? mlp_forward_fast(fr, s.w, s.buf[wave], feat) : 0.0f;
    if (valid) store_t<MDT>(map_final, id, 1.0f - expf(-fmaxf(tc - y, 0.0f)));
  }
}

// The brick's map geometry with the whole volume's light-march step.
__device__ __forceinline__ MapParams whole_volume_march(const MapParams& brick_map, const MapParams& whole)
{
  MapParams m = brick_map;
  m.step = whole.step;
  m.dir_far = whole.dir_far;
  return m;
}

// Per check vertex: |learned - exact|, the exact value being the whole volume's light march
// (the single-block reference march) to the brick's vertex.  One item = 256 check slots of one
// brick (each brick's slots are padded to whole items); the largest error over all bricks goes
// to list[1] (non-negative floats order as their bits, a NaN's bits above every number).
// FAST and PIPE as in the phase-1 map launches; check slots and brick map indices are 32-bit
// (a brick's map holds fewer than 2^31 vertices).
template <int FDT, int MDT, bool FAST, bool PIPE>
__global__ void __launch_bounds__(kBlock) k_bricks_learned_check(LearnedParams p, BrickFrame f, LearnedBricks lb)
{
  extern __shared__ float lds[];
  float* a_tab = lds;
  float* xs = lds + p.s.ntf;
  __shared__ unsigned red[kBlock / 64];
  stage_alpha(p.s, a_tab);
  stage_coords(p.g, xs);
  __syncthreads();
  const float* x[3] = { xs, xs + p.g.n[0], xs + p.g.n[0] + p.g.n[1] };
  typedef const __attribute__((address_space(4))) int32_t ConstInt;
  ConstInt* coff = (ConstInt*)(const void*)lb.chk_off;
  unsigned worst = 0;
  for (int item = blockIdx.x; item < lb.n_chk_items; item += gridDim.x) {
    const int slot0 = item * kBlock;
    int lo = 0, hi = f.nbricks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (coff[mid] <= slot0)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int b = __builtin_amdgcn_readfirstlane(lo);
    const BrickDesc* bk = f.bricks + b;
    const MapParams m = whole_volume_march(load_uniform(&bk->m), p.m);
    const int n0 = (int)check_count(m.mp[0] - 1, lb.stride), n1 = (int)check_count(m.mp[1] - 1, lb.stride);
    const int n2 = (int)check_count(m.mp[2] - 1, lb.stride);
    const int q = slot0 - coff[b] + (int)threadIdx.x;
    if (q < n0 * n1 * n2) {
      const int q1 = q / n0, q2 = q1 / n1;
      const int qi[3] = { q - q1 * n0, q1 - q2 * n1, q2 };
      int vi[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int i = qi[a] * lb.stride, last = (int)m.mp[a] - 1;
        vi[a] = i < last ? i : last;
      }
      const int64_t id = vi[0] + m.mp[0] * (vi[1] + m.mp[1] * vi[2]);
      // map_vertex's position of vertex id
      const float pos[3] = { m.morg[0] + m.msp[0] * (float)vi[0], m.morg[1] + m.msp[1] * (float)vi[1],
                             m.morg[2] + m.msp[2] * (float)vi[2] };
      int64_t taken = 0;
      const float exact = PIPE ? march_vertex_pipe<FDT, FAST>(p.g, m, p.s, a_tab, x, pos, 0.0f, &taken)
                               : march_vertex_at<FDT, false, FAST>(p.g, m, p.s, a_tab, x, pos, 0.0f, &taken);
      const float e = fabsf(load_t<MDT>(bk->map_final, id) - exact);
      lb.verr[slot0 + threadIdx.x] = e;
      worst = max(worst, f32_bits(e));
    }
  }
#pragma unroll
  for (int mk = 32; mk >= 1; mk >>= 1) worst = max(worst, (unsigned)__shfl_xor((int)worst, mk, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = worst;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned w = red[0];
    for (int i = 1; i < kBlock / 64; ++i) w = max(w, red[i]);
    atomicMax((unsigned*)&lb.list[1], w);
  }
}

// One thread per tile: the largest error over the check vertices of the tile and of its 26
// neighbours (the 4 x 4 x 4 check vertices from one before its lower corner to one past its
// upper corner, clamped to the brick) against the budget: an error inside a tile whose own
// corners pass shows, most often, at a neighbour's corners.  The tiles over the budget are
// appended to the list (one atomic per wave).
__global__ void __launch_bounds__(kBlock) k_learned_tiles(BrickFrame f, LearnedBricks lb)
{
  const int g = (int)(blockIdx.x * kBlock + threadIdx.x);
  bool fail = false;
  if (g < lb.n_tiles) {
    const int b = brick_of(lb.tile_off, f.nbricks, g);
    const int64_t* mp = f.bricks[b].m.mp;
    const int n0 = (int)check_count(mp[0] - 1, lb.stride), n1 = (int)check_count(mp[1] - 1, lb.stride);
    const int t = g - lb.tile_off[b];
    const int ti = t % (n0 - 1), tj = (t / (n0 - 1)) % (n1 - 1), tk = t / ((n0 - 1) * (n1 - 1));
    const float* e = lb.verr + lb.chk_off[b];
    const int n2 = (int)check_count(mp[2] - 1, lb.stride);
    unsigned w = 0;
    for (int c = 0; c < 64; ++c) {
      const int ci = min(max(ti - 1 + (c & 3), 0), n0 - 1), cj = min(max(tj - 1 + ((c >> 2) & 3), 0), n1 - 1);
      const int ck = min(max(tk - 1 + (c >> 4), 0), n2 - 1);
      w = max(w, f32_bits(e[ci + n0 * (cj + n1 * ck)]));
    }
    fail = !(bits_f32(w) <= lb.budget);
    if (lb.tile_fail) lb.tile_fail[g] = fail ? 1 : 0;
  }
  const unsigned long long mask = __ballot(fail);
  if (mask) {
    const int lane = (int)(threadIdx.x & 63), leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = (int)atomicAdd((unsigned*)&lb.list[0], (unsigned)__popcll(mask));
    base = __shfl(base, leader, 64);
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
    if (fail) lb.list[2 + base + rank] = g;
  }
}

// Fallback: the whole volume's light march for every vertex of the listed tiles (one tile per
// work item).  The list was written by earlier kernels, so its words go through the scalar cache.
// FAST and PIPE as in the phase-1 map launches: the branch-free corner fetch (the whole field
// below 4 GiB, strides below 2^24) and the next step's loads in flight; the same opacities.
template <int FDT, int MDT, bool FAST, bool PIPE>
__global__ void __launch_bounds__(kBlock) k_bricks_exact_tiles(LearnedParams p, BrickFrame f, LearnedBricks lb)
{
  typedef const __attribute__((address_space(4))) int32_t ConstInt;
  ConstInt* lw = (ConstInt*)(const void*)lb.list;
  ConstInt* toff = (ConstInt*)(const void*)lb.tile_off;
  const int nitems = lw[0];
  if (nitems == 0) return;
  extern __shared__ float lds[];
  float* a_tab = lds;
  float* xs = lds + p.s.ntf;
  stage_alpha(p.s, a_tab);
  stage_coords(p.g, xs);
  __syncthreads();
  const float* x[3] = { xs, xs + p.g.n[0], xs + p.g.n[0] + p.g.n[1] };
  __shared__ ItemQueue queue;
  queue_init(&queue);
  const int s = lb.stride;
  for (;;) {
    const int item = __builtin_amdgcn_readfirstlane(next_item(f.queue + 4 * kQueueSlots, &queue, nitems, 1,
                                                              f.xcd_ranges));
    if (item >= nitems) break;
    const int g = lw[2 + item];
    int lo = 0, hi = f.nbricks - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (toff[mid] <= g)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int b = __builtin_amdgcn_readfirstlane(lo);
    const BrickDesc* bk = f.bricks + b;
    const MapParams m = whole_volume_march(load_uniform(&bk->m), p.m);
    const int mp0 = (int)m.mp[0], mp1 = (int)m.mp[1], mp2 = (int)m.mp[2];
    const int n0 = (int)check_count(mp0 - 1, s), n1 = (int)check_count(mp1 - 1, s);
    const int n2 = (int)check_count(mp2 - 1, s);
    const int t = g - toff[b];
    const int ti = t % (n0 - 1), tj = (t / (n0 - 1)) % (n1 - 1), tk = t / ((n0 - 1) * (n1 - 1));
    const int i0 = ti * s, j0 = tj * s, k0 = tk * s;
    const int c0 = (ti == n0 - 2 ? mp0 : i0 + s) - i0, c1 = (tj == n1 - 2 ? mp1 : j0 + s) - j0;
    const int c2 = (tk == n2 - 2 ? mp2 : k0 + s) - k0;
    for (int v = threadIdx.x; v < c0 * c1 * c2; v += kBlock) {
      const int i = i0 + v % c0, j = j0 + (v / c0) % c1, k = k0 + v / (c0 * c1);
      const int64_t id = i + (int64_t)mp0 * (j + (int64_t)mp1 * k);
      // map_vertex's position of vertex (i, j, k)
      const float pos[3] = { m.morg[0] + m.msp[0] * (float)i, m.morg[1] + m.msp[1] * (float)j,
                             m.morg[2] + m.msp[2] * (float)k };
      int64_t taken = 0;
      const float op = PIPE ? march_vertex_pipe<FDT, FAST>(p.g, m, p.s, a_tab, x, pos, 0.0f, &taken)
                            : march_vertex_at<FDT, false, FAST>(p.g, m, p.s, a_tab, x, pos, 0.0f, &taken);
      store_t<MDT>(bk->map_final, id, op);
    }
  }
}

}  // namespace

// Whether the whole volume's light march takes the branch-free corner fetch: the schedule's
// switch (and the bricks' own conditions) in f.map_fast, the field below 4 GiB, strides below
// 2^24.
static bool whole_field_fast(const LearnedParams& p, const BrickFrame& f)
{
  return f.map_fast && p.g.bytes32 != 0 && p.g.field_stride[1] < (1 << 24) && p.g.field_stride[2] < (1 << 24);
}

hipError_t launch_bricks_learned_map(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                     const LearnedBricks* lb, bool at_checks, hipStream_t st)
{
  if (p.s.ntf > 1024 || (at_checks && !lb)) return hipErrorInvalidValue;
  LearnedBricks l{};
  if (lb) l = *lb;
  const int nitems = at_checks ? l.n_chk_items : f.n_map_items;
  if (nitems == 0) return hipSuccess;
  const dim3 grid((unsigned)std::max(1, std::min(nitems, f.persistent_groups)));
  BEAMS_DISPATCH_FIELD(p.g.dtype, BEAMS_DISPATCH_MAP(map_dtype, {
    if (at_checks)
      hipLaunchKernelGGL((k_bricks_learned_map<FDT, MDT, true>), grid, dim3(kBlock), sizeof(float) * p.s.ntf, st, p, f, l);
    else
      hipLaunchKernelGGL((k_bricks_learned_map<FDT, MDT, false>), grid, dim3(kBlock), sizeof(float) * p.s.ntf, st, p, f, l);
  }));
  return hipGetLastError();
}

hipError_t launch_bricks_learned_check(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                       const LearnedBricks& lb, hipStream_t st)
{
  hipError_t e = launch_zero_u32((unsigned int*)lb.list, 2, st);
  if (e != hipSuccess) return e;
  const size_t lds = sizeof(float) * (size_t)(p.s.ntf + 2 * (p.g.n[0] + p.g.n[1] + p.g.n[2]));
  const dim3 grid((unsigned)std::max(1, std::min(lb.n_chk_items, 2048)));
  const bool fast = whole_field_fast(p, f);
  BEAMS_DISPATCH_FIELD(p.g.dtype, BEAMS_DISPATCH_MAP(map_dtype, {
    if (fast && f.map_pipeline)
      hipLaunchKernelGGL((k_bricks_learned_check<FDT, MDT, true, true>), grid, dim3(kBlock), lds, st, p, f, lb);
    else if (fast)
      hipLaunchKernelGGL((k_bricks_learned_check<FDT, MDT, true, false>), grid, dim3(kBlock), lds, st, p, f, lb);
    else if (f.map_pipeline)
      hipLaunchKernelGGL((k_bricks_learned_check<FDT, MDT, false, true>), grid, dim3(kBlock), lds, st, p, f, lb);
    else
      hipLaunchKernelGGL((k_bricks_learned_check<FDT, MDT, false, false>), grid, dim3(kBlock), lds, st, p, f, lb);
  }));
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return launch_learned_tiles(f, lb, st);
}

hipError_t launch_learned_tiles(const BrickFrame& f, const LearnedBricks& lb, hipStream_t st)
{
  hipLaunchKernelGGL(k_learned_tiles, dim3((unsigned)((lb.n_tiles + kBlock - 1) / kBlock)), dim3(kBlock), 0, st,
                     f, lb);
  return hipGetLastError();
}

hipError_t launch_bricks_exact_tiles(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                     const LearnedBricks& lb, hipStream_t st)
{
  const size_t lds = sizeof(float) * (size_t)(p.s.ntf + 2 * (p.g.n[0] + p.g.n[1] + p.g.n[2]));
  const dim3 grid((unsigned)std::max(1, std::min(lb.n_tiles, f.persistent_groups)));
  const bool fast = whole_field_fast(p, f);
  BEAMS_DISPATCH_FIELD(p.g.dtype, BEAMS_DISPATCH_MAP(map_dtype, {
    if (fast && f.map_pipeline)
      hipLaunchKernelGGL((k_bricks_exact_tiles<FDT, MDT, true, true>), grid, dim3(kBlock), lds, st, p, f, lb);
    else if (fast)
      hipLaunchKernelGGL((k_bricks_exact_tiles<FDT, MDT, true, false>), grid, dim3(kBlock), lds, st, p, f, lb);
    else if (f.map_pipeline)
      hipLaunchKernelGGL((k_bricks_exact_tiles<FDT, MDT, false, true>), grid, dim3(kBlock), lds, st, p, f, lb);
    else
      hipLaunchKernelGGL((k_bricks_exact_tiles<FDT, MDT, false, false>), grid, dim3(kBlock), lds, st, p, f, lb);
  }));
  return hipGetLastError();
}

int64_t learned_check_count(const MapParams& m, int32_t stride)
{
  int64_t c = 1;
  for (int d = 0; d < 3; ++d) c *= check_count(m.mp[d] - 1, stride);
  return c;
}

hipError_t launch_coarse(const LearnedParams& p, hipStream_t st)
{
  if (p.s.ntf > 1024 || p.g.n[0] > kCoarseMaxX || p.cn > 64) return hipErrorInvalidValue;
  const int64_t nc = (int64_t)p.cn * p.cn * p.cn;
  hipLaunchKernelGGL(k_coarse_ranges, dim3((unsigned)((3 * p.cn + 63) / 64)), dim3(64), 0, st, p);
  BEAMS_DISPATCH_FIELD(p.g.dtype, {
    hipLaunchKernelGGL((k_coarse_sigma<FDT>), dim3((unsigned)(p.cn * p.cn)), dim3(kBlock), 0, st, p);
  });
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t per_block = kBlock / 64;
  hipLaunchKernelGGL(k_coarse_tau, dim3((unsigned)((nc + per_block - 1) / per_block)), dim3(kBlock), 0, st, p);
  return hipGetLastError();
}

hipError_t launch_tf_index_field(const LearnedParams& p, uint8_t* out, hipStream_t st)
{
  const int64_t n = p.g.n[0] * p.g.n[1] * p.g.n[2];
  if (p.s.ntf > 256 || p.g.field_stride[0] != 1 || p.g.field_stride[1] != p.g.n[0] ||
      p.g.field_stride[2] != p.g.n[0] * p.g.n[1] || n >= ((int64_t)1 << 32) || ((uintptr_t)p.g.field & 15) != 0 ||
      ((uintptr_t)out & 15) != 0)
    return hipErrorInvalidValue;
  const dim3 grid((unsigned)std::max<int64_t>(1, std::min<int64_t>((n / 4 + kBlock - 1) / kBlock, 16384)));
  BEAMS_DISPATCH_FIELD(p.g.dtype, {
    hipLaunchKernelGGL((k_tf_index_field<FDT>), grid, dim3(kBlock), 0, st, p, out, n);
  });
  return hipGetLastError();
}

hipError_t launch_learned_map(const LearnedParams& p, hipStream_t st)
{
  if (p.s.ntf > 1024) return hipErrorInvalidValue;
  // the map's vertices come from MapParams (mp <= 4097 per axis), so rows and chunks fit 32 bits
  const int64_t cpr = (p.m.mp[0] + 63) / 64;
  const int64_t nchunks = cpr * p.m.mp[1] * p.m.mp[2];
  if (nchunks > INT32_MAX) return hipErrorInvalidValue;
  const int64_t groups = std::max<int64_t>(8, std::min<int64_t>((nchunks + 3) / 4, 256 * 12) / 8 * 8);
  BEAMS_DISPATCH_FIELD(p.g.dtype, BEAMS_DISPATCH_MAP(p.m.dtype, {
    hipLaunchKernelGGL((k_learned_map<FDT, MDT>), dim3((unsigned)groups), dim3(kBlock), sizeof(float) * p.s.ntf,
                       st, p, (int32_t)cpr, (int32_t)nchunks);
  }));
  return hipGetLastError();
}

hipError_t launch_learned_check(const LearnedParams& p, hipStream_t st)
{
  const int64_t nc = learned_check_count(p.m, p.check_stride);
  const size_t lds = sizeof(float) * (size_t)(p.s.ntf + 2 * (p.g.n[0] + p.g.n[1] + p.g.n[2]));
  BEAMS_DISPATCH_FIELD(p.g.dtype, BEAMS_DISPATCH_MAP(p.m.dtype, {
    hipLaunchKernelGGL((k_learned_check<FDT, MDT>), dim3((unsigned)((nc + kBlock - 1) / kBlock)),
                       dim3(kBlock), lds, st, p, nc);
  }));
  return hipGetLastError();
}

hipError_t launch_learned_samples(const LearnedParams& p, const MapParams& em, int64_t n,
                                  uint64_t seed, float* x, float* t, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  if (p.s.ntf > 1024) return hipErrorInvalidValue;
  const dim3 grid((unsigned)((n + kBlock - 1) / kBlock));
  BEAMS_DISPATCH_FIELD(p.g.dtype, {
    if (em.dtype == kF32)
      hipLaunchKernelGGL((k_learned_samples<FDT, kF32>), grid, dim3(kBlock), 0, st, p, em, n, seed, x, t);
    else
      hipLaunchKernelGGL((k_learned_samples<FDT, kF16>), grid, dim3(kBlock), 0, st, p, em, n, seed, x, t);
  });
  return hipGetLastError();
}

}  // namespace beams
