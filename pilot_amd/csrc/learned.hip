// Kernels of the learned shadow closure (learned.hpp): the coarse extinction mip and its
// optical depth, the learned map (features + network on the matrix cores per wave), the
// sparse exact check, and training samples drawn from an exact map.
#include "learned.hpp"
#include "march_dev.hpp"
#include "mlp.hpp"

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

// Index of the grid coordinate nearest v (ties to the lower one), v clamped to the grid.
__device__ __forceinline__ int64_t nearest_index(const float* x, int64_t n, float v, float x0, float ginv)
{
  int64_t i = (int64_t)floorf((v - x0) * ginv);
  i = i < 0 ? 0 : (i > n - 2 ? n - 2 : i);
  while (i > 0 && x[i] > v) --i;
  while (i < n - 2 && x[i + 1] < v) ++i;
  return (v - x[i] <= x[i + 1] - v) ? i : i + 1;
}

__device__ __forceinline__ int64_t tf_index(const ShadeParams& s, float val)
{
  const float sc = (val - s.min_scalar) * s.inv_delta;
  const int64_t tfmax = s.ntf - 1;
  int64_t idx = (int64_t)(sc * (float)tfmax);
  idx = idx < tfmax ? idx : tfmax;
  return idx > 0 ? idx : 0;
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
  int64_t c[3];
  float t[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    float u = (v[d] - p.corg[d]) * p.cinv[d];
    u = fminf(fmaxf(u, 0.0f), (float)(p.cn - 1));
    int64_t i = (int64_t)u;
    i = i > p.cn - 2 ? p.cn - 2 : i;
    c[d] = i;
    t[d] = u - (float)i;
  }
  const int64_t n = p.cn, b = c[0] + n * (c[1] + n * c[2]);
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
  int64_t c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) c[d] = nearest_index(g.x[d], g.n[d], v[d], g.gx0[d], g.ginv[d]);
  int q = 1;
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz) {
    int64_t k = c[2] + dz;
    k = k < 0 ? 0 : (k > g.n[2] - 1 ? g.n[2] - 1 : k);
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      int64_t j = c[1] + dy;
      j = j < 0 ? 0 : (j > g.n[1] - 1 ? g.n[1] - 1 : j);
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        int64_t i = c[0] + dx;
        i = i < 0 ? 0 : (i > g.n[0] - 1 ? g.n[0] - 1 : i);
        const float val = load_t<FDT>(g.field, i * g.field_stride[0] + j * g.field_stride[1] + k * g.field_stride[2]);
        f[q++] = d_tab[tf_index(p.s, val)];
      }
    }
  }
  light_dir(p.m, v, f + 28);
  f[31] = 0.0f;
  return tc;
}

__device__ __forceinline__ float sigma_of(float a, float inv_h)
{
  return -log1pf(-fminf(a, 1.0f - 5.9604645e-8f)) * inv_h;
}

// ---- coarse mip: mean extinction over the field vertices nearest each coarse vertex -------
template <int FDT>
__global__ void __launch_bounds__(kBlock) k_coarse_sigma(LearnedParams p)
{
  __shared__ float sig_tab[1024];
  __shared__ float red[kBlock / 64];
  for (int i = threadIdx.x; i < p.s.ntf; i += blockDim.x)
    sig_tab[i] = sigma_of(p.s.tf[4 * i + 3] / p.s.max_density, p.inv_h);
  __syncthreads();
  const GridParams& g = p.g;
  const int64_t cn = p.cn, id = blockIdx.x;
  const int64_t ci[3] = { id % cn, (id / cn) % cn, id / (cn * cn) };
  int64_t lo[3], cnt[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const float c = p.corg[d] + p.csp[d] * (float)ci[d];
    const float half = 0.5f * p.csp[d];
    lo[d] = lower_index(g.x[d], g.n[d], c - half);
    const int64_t hi = ci[d] == cn - 1 ? g.n[d] : lower_index(g.x[d], g.n[d], c + half);
    cnt[d] = hi - lo[d];
    if (cnt[d] <= 0) {  // coarser field than mip: the nearest field vertex alone
      lo[d] = nearest_index(g.x[d], g.n[d], c, g.gx0[d], g.ginv[d]);
      cnt[d] = 1;
    }
  }
  const int64_t total = cnt[0] * cnt[1] * cnt[2];
  float acc = 0.0f;
  for (int64_t q = threadIdx.x; q < total; q += blockDim.x) {
    const int64_t i = lo[0] + q % cnt[0];
    const int64_t j = lo[1] + (q / cnt[0]) % cnt[1];
    const int64_t k = lo[2] + q / (cnt[0] * cnt[1]);
    const float val = load_t<FDT>(g.field, i * g.field_stride[0] + j * g.field_stride[1] + k * g.field_stride[2]);
    acc += sig_tab[tf_index(p.s, val)];
  }
#pragma unroll
  for (int mk = 32; mk >= 1; mk >>= 1) acc += __shfl_xor(acc, mk, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float sum = 0.0f;
    for (int w = 0; w < kBlock / 64; ++w) sum += red[w];
    p.coarse_sigma[id] = sum / (float)total;
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
// Static part of the learned-map kernels' LDS (51 KB); the density-patch table (ntf floats)
// is dynamic, so that three workgroups fit a CU for the usual 256-entry TF.
struct MapLds {
  MlpLds w;
  uint16_t buf[kBlock / 64][64 * kLd];
};

template <int FDT, int MDT>
__global__ void __launch_bounds__(kBlock) k_learned_map(LearnedParams p, int64_t nv)
{
  __shared__ MapLds s;
  extern __shared__ float d_tab[];
  if (p.params) mlp_stage(p.params, s.w);
  stage_patch(p.s, d_tab);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t base = (int64_t)blockIdx.x * kBlock + wave * 64; base < nv; base += stride) {
    const int64_t id = base + lane;
    const bool valid = id < nv;
    float v[3], f[kFeat];
    map_vertex(p.m, valid ? id : nv - 1, v);
    const float tc = features<FDT>(p, d_tab, v, f);
    const float y = p.params ? mlp_forward_fast(s.w, s.buf[wave], f) : 0.0f;
    const float tau = fmaxf(tc - y, 0.0f);
    if (valid) store_t<MDT>(p.m.map, id, 1.0f - expf(-tau));
  }
}

// ---- the sparse exact check ----------------------------------------------------------------
__device__ __forceinline__ int64_t check_axis(int64_t cells, int32_t stride)
{
  return cells / stride + 1 + (cells % stride != 0 ? 1 : 0);
}

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
  const int64_t n0 = check_axis(m.mp[0] - 1, p.check_stride);
  const int64_t n1 = check_axis(m.mp[1] - 1, p.check_stride);
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
// One map work item (256 vertices of one brick) per workgroup iteration, 64 per wave.
template <int FDT, int MDT>
__global__ void __launch_bounds__(kBlock) k_bricks_learned_map(LearnedParams p, BrickFrame f)
{
  __shared__ MapLds s;
  extern __shared__ float d_tab[];
  if (p.params) mlp_stage(p.params, s.w);
  stage_patch(p.s, d_tab);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int item = blockIdx.x; item < f.n_map_items; item += gridDim.x) {
    const BrickDesc& d = f.bricks[f.map_item_brick[item]];
    const int64_t nv = d.nverts;
    const int64_t id = f.map_item_first[item] + wave * 64 + lane;
    const bool valid = id < nv;
    float v[3], feat[kFeat];
    map_vertex(d.m, valid ? id : nv - 1, v);
    const float tc = features<FDT>(p, d_tab, v, feat);
    const float y = p.params ? mlp_forward_fast(s.w, s.buf[wave], feat) : 0.0f;
    if (valid) store_t<MDT>(d.map_final, id, 1.0f - expf(-fmaxf(tc - y, 0.0f)));
  }
}

// The brick's map geometry with the whole volume's light-march step.
__device__ __forceinline__ MapParams whole_volume_march(const BrickDesc& d, const MapParams& whole)
{
  MapParams m = d.m;
  m.step = whole.step;
  m.dir_far = whole.dir_far;
  return m;
}

// Per brick (one workgroup each): largest |learned - exact| over its check vertices, the exact
// value being the whole volume's light march (the single-block reference march).
template <int FDT, int MDT>
__global__ void __launch_bounds__(kBlock) k_bricks_learned_check(LearnedParams p, BrickFrame f, float* err)
{
  extern __shared__ float lds[];
  float* a_tab = lds;
  float* xs = lds + p.s.ntf;
  __shared__ float red[kBlock / 64];
  stage_alpha(p.s, a_tab);
  stage_coords(p.g, xs);
  __syncthreads();
  const float* x[3] = { xs, xs + p.g.n[0], xs + p.g.n[0] + p.g.n[1] };
  const BrickDesc& d = f.bricks[blockIdx.x];
  const MapParams m = whole_volume_march(d, p.m);
  const int64_t n0 = check_axis(m.mp[0] - 1, p.check_stride);
  const int64_t n1 = check_axis(m.mp[1] - 1, p.check_stride);
  const int64_t n2 = check_axis(m.mp[2] - 1, p.check_stride);
  float worst = 0.0f;
  for (int64_t q = threadIdx.x; q < n0 * n1 * n2; q += blockDim.x) {
    const int64_t qi[3] = { q % n0, (q / n0) % n1, q / (n0 * n1) };
    int64_t vi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int64_t i = qi[a] * p.check_stride;
      vi[a] = i < m.mp[a] - 1 ? i : m.mp[a] - 1;
    }
    const int64_t id = vi[0] + m.mp[0] * (vi[1] + m.mp[1] * vi[2]);
    int64_t taken = 0;
    const float exact = march_vertex<FDT>(p.g, m, p.s, a_tab, x, id, 0.0f, &taken);
    worst = fmaxf(worst, fabsf(load_t<MDT>(d.map_final, id) - exact));
  }
#pragma unroll
  for (int mk = 32; mk >= 1; mk >>= 1) worst = fmaxf(worst, __shfl_xor(worst, mk, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = worst;
  __syncthreads();
  if (threadIdx.x == 0) {
    float w = red[0];
    for (int i = 1; i < kBlock / 64; ++i) w = fmaxf(w, red[i]);
    err[blockIdx.x] = w;
  }
}

// Fallback: the whole volume's light march for every vertex of the bricks flagged in redo.
template <int FDT, int MDT>
__global__ void __launch_bounds__(kBlock) k_bricks_exact_map(LearnedParams p, BrickFrame f, const int32_t* redo)
{
  extern __shared__ float lds[];
  float* a_tab = lds;
  float* xs = lds + p.s.ntf;
  stage_alpha(p.s, a_tab);
  stage_coords(p.g, xs);
  __syncthreads();
  const float* x[3] = { xs, xs + p.g.n[0], xs + p.g.n[0] + p.g.n[1] };
  for (int item = blockIdx.x; item < f.n_map_items; item += gridDim.x) {
    const int b = f.map_item_brick[item];
    if (!redo[b]) continue;
    const BrickDesc& d = f.bricks[b];
    const int64_t id = f.map_item_first[item] + threadIdx.x;
    if (id >= d.nverts) continue;
    const MapParams m = whole_volume_march(d, p.m);
    int64_t taken = 0;
    store_t<MDT>(d.map_final, id, march_vertex<FDT>(p.g, m, p.s, a_tab, x, id, 0.0f, &taken));
  }
}

}  // namespace

hipError_t launch_bricks_learned_map(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                     hipStream_t st)
{
  if (p.s.ntf > 1024) return hipErrorInvalidValue;
  const dim3 grid((unsigned)std::max(1, std::min(f.n_map_items, f.persistent_groups)));
  BEAMS_DISPATCH_FIELD(p.g.dtype, BEAMS_DISPATCH_MAP(map_dtype, {
    hipLaunchKernelGGL((k_bricks_learned_map<FDT, MDT>), grid, dim3(kBlock), sizeof(float) * p.s.ntf, st, p, f);
  }));
  return hipGetLastError();
}

hipError_t launch_bricks_learned_check(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                       float* err, hipStream_t st)
{
  const size_t lds = sizeof(float) * (size_t)(p.s.ntf + p.g.n[0] + p.g.n[1] + p.g.n[2]);
  BEAMS_DISPATCH_FIELD(p.g.dtype, BEAMS_DISPATCH_MAP(map_dtype, {
    hipLaunchKernelGGL((k_bricks_learned_check<FDT, MDT>), dim3((unsigned)f.nbricks), dim3(kBlock), lds,
                       st, p, f, err);
  }));
  return hipGetLastError();
}

hipError_t launch_bricks_exact_map(const LearnedParams& p, const BrickFrame& f, int32_t map_dtype,
                                   const int32_t* redo, hipStream_t st)
{
  const size_t lds = sizeof(float) * (size_t)(p.s.ntf + p.g.n[0] + p.g.n[1] + p.g.n[2]);
  const dim3 grid((unsigned)std::max(1, std::min(f.n_map_items, f.persistent_groups)));
  BEAMS_DISPATCH_FIELD(p.g.dtype, BEAMS_DISPATCH_MAP(map_dtype, {
    hipLaunchKernelGGL((k_bricks_exact_map<FDT, MDT>), grid, dim3(kBlock), lds, st, p, f, redo);
  }));
  return hipGetLastError();
}

int64_t learned_check_count(const MapParams& m, int32_t stride)
{
  int64_t c = 1;
  for (int d = 0; d < 3; ++d) {
    const int64_t cells = m.mp[d] - 1;
    c *= cells / stride + 1 + (cells % stride != 0 ? 1 : 0);
  }
  return c;
}

hipError_t launch_coarse(const LearnedParams& p, hipStream_t st)
{
  if (p.s.ntf > 1024) return hipErrorInvalidValue;
  const int64_t nc = (int64_t)p.cn * p.cn * p.cn;
  BEAMS_DISPATCH_FIELD(p.g.dtype, {
    hipLaunchKernelGGL((k_coarse_sigma<FDT>), dim3((unsigned)nc), dim3(kBlock), 0, st, p);
  });
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int64_t per_block = kBlock / 64;
  hipLaunchKernelGGL(k_coarse_tau, dim3((unsigned)((nc + per_block - 1) / per_block)), dim3(kBlock), 0, st, p);
  return hipGetLastError();
}

hipError_t launch_learned_map(const LearnedParams& p, hipStream_t st)
{
  if (p.s.ntf > 1024) return hipErrorInvalidValue;
  const int64_t nv = p.m.mp[0] * p.m.mp[1] * p.m.mp[2];
  const int64_t groups = std::min<int64_t>((nv + kBlock - 1) / kBlock, 256 * 8);
  BEAMS_DISPATCH_FIELD(p.g.dtype, BEAMS_DISPATCH_MAP(p.m.dtype, {
    hipLaunchKernelGGL((k_learned_map<FDT, MDT>), dim3((unsigned)groups), dim3(kBlock), sizeof(float) * p.s.ntf,
                       st, p, nv);
  }));
  return hipGetLastError();
}

hipError_t launch_learned_check(const LearnedParams& p, hipStream_t st)
{
  const int64_t nc = learned_check_count(p.m, p.check_stride);
  const size_t lds = sizeof(float) * (size_t)(p.s.ntf + p.g.n[0] + p.g.n[1] + p.g.n[2]);
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
