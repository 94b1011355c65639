// Exact-path kernels: the transmittance map built by per-vertex light marches
// (TransmittanceMapGenerator) and the camera-ray march with shadow lookups
// (CalcRayStart + Sampler), in the reference's FP32 operation order.
#include "camera.hpp"
#include "march_dev.hpp"

namespace beams {

namespace {

constexpr int kBlock = 256;

// ---- TransmittanceMap.h:239-340 (TransmittanceMapGenerator over all map vertices) -------
template <int FDT, int MDT, bool BINS>
__global__ void __launch_bounds__(kBlock) k_tmap_ref(GridParams g, MapParams m, ShadeParams s,
                                                     const float* seed, int64_t nv)
{
  if (gate_closed(m)) return;
  extern __shared__ float lds[];
  float* a_tab = lds;          // alpha / max alpha per TF entry
  float* xs = lds + s.ntf;     // coordinates
  for (int i = threadIdx.x; i < s.ntf; i += blockDim.x) a_tab[i] = s.tf[4 * i + 3] / s.max_density;
  stage_coords(g, xs);
  if (BINS) stage_bins(g, xs);
  __syncthreads();
  const float* x[3] = { xs, xs + g.n[0], xs + g.n[0] + g.n[1] };
  const int64_t id = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (id >= nv) return;
  int64_t taken = 0;
  const float op = march_vertex<FDT, BINS>(g, m, s, a_tab, x, id, seed ? seed[id] : 0.0f, &taken);
  store_t<MDT>(m.map, id, op);
  if (s.counters) count_add(&s.counters[1], (uint32_t)taken);
}

template <int FDT, int MDT, bool USE_MAP, bool BINS>
__global__ void __launch_bounds__(kBlock) k_march_rays(GridParams g, MapParams m, ShadeParams s,
                                                       RaySoA r)
{
  extern __shared__ float4 lds4[];
  float4* tf = lds4;
  float* xs = (float*)(lds4 + s.ntf);
  stage_tf(s, tf);
  stage_coords(g, xs);
  if (BINS) stage_bins(g, xs);
  __syncthreads();
  const float* x[3] = { xs, xs + g.n[0], xs + g.n[0] + g.n[1] };
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= r.n) return;
  float org[3] = { r.org[0][i], r.org[1][i], r.org[2][i] };
  float dir[3] = { r.dir[0][i], r.dir[1][i], r.dir[2][i] };
  float minD = r.min_d[i], maxD = r.max_d[i], dist = r.dist[i];
  calc_ray_start(g, org, dir, minD, dist, maxD);
  r.min_d[i] = minD;
  r.max_d[i] = maxD;
  r.dist[i] = dist;
  float4 c4 = ((float4*)r.rgba)[i];
  float color[4] = { c4.x, c4.y, c4.z, c4.w };
  shade<FDT, MDT, USE_MAP, BINS>(g, m, s, tf, x, org, dir, minD, maxD, color);
  ((float4*)r.rgba)[i] = make_float4(color[0], color[1], color[2], color[3]);
}

// 16x16-pixel workgroups made of four 8x8 wave64 tiles; workgroup ids are remapped
// so that each of the 8 XCDs (which receive workgroups round-robin) marches one
// contiguous band of tiles and keeps that band's bricks in its own L2.
template <int FDT, int MDT, bool USE_MAP, bool BINS>
__global__ void __launch_bounds__(kBlock) k_march_image(GridParams g, MapParams m, ShadeParams s,
                                                        CameraParams cam, ImageOut out,
                                                        int32_t tiles_x, int32_t ntiles,
                                                        int32_t per_xcd)
{
  extern __shared__ float4 lds4[];
  float4* tf = lds4;
  float* xs = (float*)(lds4 + s.ntf);
  stage_tf(s, tf);
  stage_coords(g, xs);
  if (BINS) stage_bins(g, xs);
  __syncthreads();
  const int32_t logical = (int32_t)(blockIdx.x % 8u) * per_xcd + (int32_t)(blockIdx.x / 8u);
  if (logical >= ntiles) return;
  const float* x[3] = { xs, xs + g.n[0], xs + g.n[0] + g.n[1] };
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int px = (logical % tiles_x) * 16 + (wave & 1) * 8 + (lane & 7);
  const int py = (logical / tiles_x) * 16 + (wave >> 1) * 8 + (lane >> 3);
  if (px >= cam.width || py >= cam.height) return;
  float org[3], dir[3];
  camera_ray(cam, px, py, org, dir);
  float minD = cam.near_d, maxD = cam.far_d, dist = 0.0f;
  calc_ray_start(g, org, dir, minD, dist, maxD);
  float color[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
  shade<FDT, MDT, USE_MAP, BINS>(g, m, s, tf, x, org, dir, minD, maxD, color);
  if (out.use_bg) {
    const float ia = 1.0f - color[3];
    color[0] = color[0] + out.bg[0] * ia;
    color[1] = color[1] + out.bg[1] * ia;
    color[2] = color[2] + out.bg[2] * ia;
    color[3] = color[3] + out.bg[3] * ia;
  }
  const int64_t pix = (int64_t)py * cam.width + px;
  ((float4*)out.rgba)[pix] = make_float4(color[0], color[1], color[2], color[3]);
  if (out.depth) out.depth[pix] = (minD == -1.0f) ? __builtin_inff() : minD;
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// coordinates and reciprocal cell widths (stage_coords)
// coordinates and reciprocal cell widths (stage_coords), then the bin tables (stage_bins)
size_t lds_bytes_coords(const GridParams& g)
{
  return (g.use_bins ? 3 : 2) * sizeof(float) * (g.n[0] + g.n[1] + g.n[2]);
}
}  // namespace

// ---- launchers -----------------------------------------------------------------------------
hipError_t launch_zero_u64(unsigned long long* p, int n, hipStream_t st)
{
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_tmap_ref(const GridParams& g, const MapParams& m, const ShadeParams& s,
                           const float* seed, hipStream_t st)
{
  const int64_t nv = m.mp[0] * m.mp[1] * m.mp[2];
  const size_t lds = sizeof(float) * s.ntf + lds_bytes_coords(g);
  const dim3 grid((unsigned)((nv + kBlock - 1) / kBlock));
  BEAMS_DISPATCH_FIELD(g.dtype, BEAMS_DISPATCH_MAP(m.dtype, {
    if (g.use_bins)
      hipLaunchKernelGGL((k_tmap_ref<FDT, MDT, true>), grid, dim3(kBlock), lds, st, g, m, s, seed, nv);
    else
      hipLaunchKernelGGL((k_tmap_ref<FDT, MDT, false>), grid, dim3(kBlock), lds, st, g, m, s, seed, nv);
  }));
  return hipGetLastError();
}

hipError_t launch_march_rays(const GridParams& g, const MapParams& m, const ShadeParams& s,
                             const RaySoA& r, hipStream_t st)
{
  if (r.n == 0) return hipSuccess;
  const size_t lds = sizeof(float4) * s.ntf + lds_bytes_coords(g);
  const dim3 grid((unsigned)((r.n + kBlock - 1) / kBlock));
  BEAMS_DISPATCH_FIELD(g.dtype, BEAMS_DISPATCH_MAP(m.dtype, {
    if (s.use_map && g.use_bins)
      hipLaunchKernelGGL((k_march_rays<FDT, MDT, true, true>), grid, dim3(kBlock), lds, st, g, m, s, r);
    else if (s.use_map)
      hipLaunchKernelGGL((k_march_rays<FDT, MDT, true, false>), grid, dim3(kBlock), lds, st, g, m, s, r);
    else if (g.use_bins)
      hipLaunchKernelGGL((k_march_rays<FDT, MDT, false, true>), grid, dim3(kBlock), lds, st, g, m, s, r);
    else
      hipLaunchKernelGGL((k_march_rays<FDT, MDT, false, false>), grid, dim3(kBlock), lds, st, g, m, s, r);
  }));
  return hipGetLastError();
}

hipError_t launch_march_image(const GridParams& g, const MapParams& m, const ShadeParams& s,
                              const CameraParams& c, const ImageOut& o, hipStream_t st)
{
  const int32_t tiles_x = (c.width + 15) / 16, tiles_y = (c.height + 15) / 16;
  const int32_t ntiles = tiles_x * tiles_y;
  const int32_t per_xcd = (ntiles + 7) / 8;
  const dim3 grid((unsigned)(per_xcd * 8));
  const size_t lds = sizeof(float4) * s.ntf + lds_bytes_coords(g);
  BEAMS_DISPATCH_FIELD(g.dtype, BEAMS_DISPATCH_MAP(m.dtype, {
    if (s.use_map && g.use_bins)
      hipLaunchKernelGGL((k_march_image<FDT, MDT, true, true>), grid, dim3(kBlock), lds, st, g, m, s,
                         c, o, tiles_x, ntiles, per_xcd);
    else if (s.use_map)
      hipLaunchKernelGGL((k_march_image<FDT, MDT, true, false>), grid, dim3(kBlock), lds, st, g, m, s,
                         c, o, tiles_x, ntiles, per_xcd);
    else if (g.use_bins)
      hipLaunchKernelGGL((k_march_image<FDT, MDT, false, true>), grid, dim3(kBlock), lds, st, g, m, s,
                         c, o, tiles_x, ntiles, per_xcd);
    else
      hipLaunchKernelGGL((k_march_image<FDT, MDT, false, false>), grid, dim3(kBlock), lds, st, g, m,
                         s, c, o, tiles_x, ntiles, per_xcd);
  }));
  return hipGetLastError();
}

}  // namespace beams
