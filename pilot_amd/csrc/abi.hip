// C-ABI layer: a handle holding the renderer state of
// beams::rendering::LightedVolumeRenderer (LightedVolumeRenderer.h:20-101) with
// device-resident inputs, and the frame driver of RenderOnDevice
// (LightedVolumeRenderer.cxx:910-1372) over the HIP kernels.
#include "../../include/beams_mi355x.h"
#include "bricks.hpp"
#include "camera.hpp"
#include "sweep.hpp"
#include "gen.hpp"
#include "kernels.hpp"
#include "learned.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

using namespace beams;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t n)
  {
    if (n <= bytes && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMalloc(&p, std::max<size_t>(n, 256));
    if (e == hipSuccess) bytes = std::max<size_t>(n, 256);
    return e;
  }
  void release()
  {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

size_t dtype_size(int32_t dt)
{
  switch (dt) {
    case BEAMS_F32: return 4;
    case BEAMS_F16:
    case BEAMS_BF16: return 2;
    case BEAMS_F64: return 8;
    default: return 0;
  }
}

// Scratch of beams_render_rays_noalloc: the transmittance map of callers that cannot
// allocate.  One array per device, held from module load until the process ends; host code
// uses only the address hipGetSymbolAddress reports for the current device (kept in
// g_scratch), never the array's name.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

struct beams_renderer {
  int device = 0;
  void* scratch = nullptr;  // this device's scratch array (g_scratch[device])
  hipStream_t stream = nullptr;
  std::string err;
  // SetData
  DevBuf field;
  int32_t dtype = -1, assoc = 0;
  int64_t dims[3] = { 0, 0, 0 };
  std::vector<float> hx[3];
  bool use_bins = false;  // uneven coordinates: bin tables follow the reciprocals
  float binv[3] = { 0.0f, 0.0f, 0.0f };
  DevBuf coords;
  double bounds[6] = { 0, 0, 0, 0, 0, 0 };
  bool has_field = false;
  double smin = 0.0, smax = 1.0;
  bool has_range = false;
  // SetColorMap
  std::vector<float> htf;
  DevBuf tf;
  // AddLight (light 0 only)
  int32_t light_kind = BEAMS_LIGHT_POINT;
  float light[3] = { 0, 0, 0 }, light_color[3] = { 1, 1, 1 };
  bool has_light = false;
  // shadow map
  bool use_shadow = true;
  int64_t map_cells[3] = { 16, 16, 16 };
  int32_t shadow_mode = BEAMS_SHADOW_REF_MARCH;
  int32_t map_dtype = BEAMS_F32;
  DevBuf map;
  DevBuf sweep_slices;  // (tau, sigma) ping-pong slices of the exact sweep
  DevBuf sweep_sigma;   // per-vertex extinction of the blocked sweep
  bool map_valid = false;
  int32_t built_mode = BEAMS_SHADOW_REF_MARCH;  // the mode (and so the layout) of the map held
  // learned closure
  const beams_mlp* net = nullptr;
  int32_t fallback_mode = BEAMS_SHADOW_REF_MARCH;
  float error_budget = 2e-3f;
  int32_t check_stride = 8;
  DevBuf coarse;  // coarse extinction and optical depth, (kCoarseCells + 1)^3 each
  DevBuf tfidx;   // the TF entry of every field point (the learned map's feature patches)
  DevBuf lerr;    // check errors as float bits
  // the last learned build, whose statistics are read back on request: 0 none, 1 single block
  // (lerr holds its two check errors), 2 bricks (lb_list holds the fallback's tile list)
  int32_t learned_kind = 0;
  float built_budget = 0.0f;
  // march
  float sample_distance = -1.0f;
  float alpha_stop = 1.0f;
  DevBuf rays, image;
  // Per-frame phase timing: events [2p] / [2p+1] bracket phase p.  A ring of frames
  // holds the frames rendered since the last reset.
  struct Frame {
    hipEvent_t ev[2 * BEAMS_NUM_PHASES] = {};
    bool timed[BEAMS_NUM_PHASES] = {};
  };
  static constexpr int kMaxFrames = 256;
  std::vector<Frame> frames;
  int64_t total_frames = 0;  // since the last reset
  int cur = -1;
  // optional sample counters: [0] camera-ray samples, [1] light-march steps
  bool counting = false;
  DevBuf counters;
  // Bricks: one reference block (MPI rank) per brick.  Empty = the field is one block.
  struct Brick {
    int64_t lo[3], hi[3];      // owned point ranges, inclusive (neighbours share planes)
    int64_t slo[3], sdims[3];  // stored region: owned points + ghost layers
    int64_t cells[3];          // map cells
    DevBuf field;
    DevBuf field_blk;  // 4 x 4 x 4-blocked copy of the owned points (the blocked march's)
    int64_t nverts = 0, seed_off = 0;
    void release()
    {
      field.release();
      field_blk.release();
    }
  };
  std::vector<Brick> bricks;
  std::vector<Brick> brick_pool;  // storage of the decomposition a new field replaced
  int32_t ghost = 0;
  DevBuf bmap_local, bmap_final, bseeds, btables, bpartial, bqueue;
  DevBuf brm_list;  // phase 3's list of the vertices to re-march (one uint32 per map vertex at most)
  DevBuf bmap_blk;  // 4 x 4 x 4-blocked copies of the final maps (the blocked march's)
  // learned bricks: check-slot and tile offsets (host copy and device table), per-check-vertex
  // errors, and the fallback's tile list
  std::vector<int32_t> lb_host;
  DevBuf lb_tab, lb_err, lb_list, lb_fail;
  LearnedBricks lb{};
  // learned single block: its fallback tiles through the brick kernels, as a frame of one brick
  // whose map is the whole volume's (device descriptor, queue, tables, errors, tile list)
  std::vector<unsigned char> sb_host;
  DevBuf sb_desc, sb_queue, sb_tab, sb_err, sb_list;
  LearnedBricks sb_lb{};
  BrickFrame sb_f{};
  bool sb_tiles = false;  // the last single-block learned build used the tile fallback
  std::vector<unsigned char> btab_host;  // tables of the last frame (reused while unchanged)
  int32_t n_map_items = 0, n_px_items = 0, max_coords = 0;
  int32_t persistent_groups = 2048;
  // schedule knobs of the brick launches (the same results whatever their values)
  int32_t xcd_ranges = 1, light_rows = 1, march_xcd_tiles = 0, map_fast_fetch = 1, march_fast_fetch = 1;
  int32_t map_pipeline = 1;
  int32_t sweep_block = 1;
  int32_t march_groups_per_cu = 0;
  int32_t march_blocked = 1;
  int32_t composite_order = BEAMS_ORDER_CENTER_DISTANCE;
  DevBuf brows;                          // every brick's map rows in light-sheet order
  std::vector<unsigned char> brows_key;  // the light and brick maps the rows were ordered for
  int64_t btab_version = 0;
  // hipGraph of the frame's launches, replayed while its key (every kernel argument) holds
  bool use_graph = false;
  hipGraphExec_t graph_exec = nullptr;
  std::vector<unsigned char> graph_key;

  int fail(int code, const std::string& msg)
  {
    err = msg;
    return code;
  }
};

#define HIP_TRY(h, expr)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return (h)->fail(BEAMS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHECK_HANDLE(h) \
  if (!(h)) return BEAMS_ERR_INVALID

namespace {

hipMemcpyKind copy_kind(int32_t mem, bool to_device)
{
  if (mem == BEAMS_MEM_DEVICE) return hipMemcpyDeviceToDevice;
  return to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
}

// GridParams for the handle's block (SetData state).
GridParams grid_params(const beams_renderer* h)
{
  GridParams g{};
  g.field = h->field.p;
  g.dtype = h->dtype;
  g.assoc = h->assoc;
  for (int d = 0; d < 3; ++d) g.n[d] = h->dims[d];
  const int64_t sx = (h->assoc == 1) ? h->dims[0] - 1 : h->dims[0];
  const int64_t sy = (h->assoc == 1) ? h->dims[1] - 1 : h->dims[1];
  g.field_stride[0] = 1;
  g.field_stride[1] = sx;
  g.field_stride[2] = sx * sy;
  const int64_t nvals = (h->assoc == 1) ? (h->dims[0] - 1) * (h->dims[1] - 1) * (h->dims[2] - 1)
                                        : h->dims[0] * h->dims[1] * h->dims[2];
  const uint64_t fbytes = (uint64_t)nvals * dtype_size(h->dtype);
  g.bytes32 = fbytes < ((uint64_t)1 << 32) ? (uint32_t)fbytes : 0u;
  const float* base = (const float*)h->coords.p;
  const int64_t nc = h->dims[0] + h->dims[1] + h->dims[2];
  g.x[0] = base;
  g.x[1] = base + h->dims[0];
  g.x[2] = base + h->dims[0] + h->dims[1];
  for (int d = 0; d < 3; ++d) g.r[d] = g.x[d] + nc;
  for (int d = 0; d < 3; ++d) {
    const std::vector<float>& x = h->hx[d];
    g.minp[d] = x.front();
    g.maxp[d] = x.back();
    g.gx0[d] = x.front();
    g.ginv[d] = (float)((double)(h->dims[d] - 1) / ((double)x.back() - (double)x.front()));
  }
  std::memcpy(g.bounds, h->bounds, sizeof(g.bounds));
  g.use_bins = h->use_bins ? 1 : 0;
  for (int d = 0; d < 3; ++d) g.binv[d] = h->binv[d];
  // Math.h:12-20 BoundsMagnitude<Float32>
  const float e0 = (float)(h->bounds[1] - h->bounds[0]);
  const float e1 = (float)(h->bounds[3] - h->bounds[2]);
  const float e2 = (float)(h->bounds[5] - h->bounds[4]);
  g.ext_mag = std::sqrt(e0 * e0 + e1 * e1 + e2 * e2);
  return g;
}

// CreateDataSetForOpacityMap + GenerateEstimator set-up (TransmittanceMap.h:550-621).
// Sweep geometry: the axis along which the light crosses the map (the light must be
// outside the bounds' slab on that axis), its direction and the transverse axes.
bool sweep_axis(const beams_renderer* h, int& axis, int& a1, int& a2, int& step)
{
  if (h->light_kind == BEAMS_LIGHT_DIRECTIONAL) {
    axis = 0;
    for (int d = 1; d < 3; ++d)
      if (std::fabs(h->light[d]) > std::fabs(h->light[axis])) axis = d;
    step = h->light[axis] > 0.0f ? 1 : -1;
  } else {
    double best = 0.0;
    axis = -1;
    for (int d = 0; d < 3; ++d) {
      const double out = std::max((double)h->bounds[2 * d] - h->light[d], (double)h->light[d] - h->bounds[2 * d + 1]);
      if (out > best) {
        best = out;
        axis = d;
      }
    }
    if (axis < 0) return false;
    step = (double)h->light[axis] < h->bounds[2 * axis] ? 1 : -1;
  }
  a1 = axis == 0 ? 1 : 0;
  a2 = axis == 2 ? 1 : 2;
  return true;
}

// Map parameters of a map built by `mode`: the sweep stores slices contiguously, the other
// modes x fastest.
MapParams map_params(const beams_renderer* h, const GridParams& g, int32_t mode)
{
  MapParams m{};
  m.map = h->map.p;
  m.dtype = h->map_dtype;
  m.light_kind = h->light_kind;
  float size[3];
  for (int d = 0; d < 3; ++d) {
    m.mp[d] = h->map_cells[d] + 1;
    m.morg[d] = (float)h->bounds[2 * d];
    size[d] = (float)(h->bounds[2 * d + 1] - h->bounds[2 * d]);
    m.msp[d] = size[d] / (float)h->map_cells[d];
    m.minv[d] = 1.0f / m.msp[d];
    m.light[d] = h->light[d];
    m.light_color[d] = h->light_color[d];
  }
  m.ms[0] = 1;
  m.ms[1] = m.mp[0];
  m.ms[2] = m.mp[0] * m.mp[1];
  int axis, a1, a2, step;
  if (mode == BEAMS_SHADOW_SWEEP && sweep_axis(h, axis, a1, a2, step)) {
    m.ms[a1] = 1;  // slice-contiguous layout, the sweep axis slowest
    m.ms[a2] = m.mp[a1];
    m.ms[axis] = m.mp[a1] * m.mp[a2];
  }
  m.bytes32 = map_bytes32(m.mp, m.dtype);
  m.step = std::sqrt(size[0] * size[0] + size[1] * size[1] + size[2] * size[2]) / 128.0f;
  m.dir_far = 4.0f * g.ext_mag;
  return m;
}

// Bound on a camera ray's samples: two extents' worth of steps plus 64, at most 2^31 - 1 (the
// kernels count samples in 32 bits).
int64_t max_ray_steps(float ext_mag, float sample_dist)
{
  const float q = 2.0f * ext_mag / sample_dist;
  if (!(q < 2147483520.0f)) return 2147483647;
  return (int64_t)q + 64;
}

ShadeParams shade_params(beams_renderer* h, const GridParams& g)
{
  ShadeParams s{};
  s.tf = (const float*)h->tf.p;
  s.ntf = (int32_t)(h->htf.size() / 4);
  float ma = 0.0f;
  bool finite = true;
  for (int32_t i = 0; i < s.ntf; ++i) {
    ma = std::fmax(ma, h->htf[4 * i + 3]);
    finite = finite && std::isfinite(h->htf[4 * i + 3]);
  }
  s.max_density = ma;
  s.alpha_finite = (finite && std::isfinite(ma) && ma > 0.0f) ? 1 : 0;
  const float mn = (float)h->smin, mx = (float)h->smax;
  s.min_scalar = mn;
  s.inv_delta = ((mx - mn) != 0.0f) ? 1.0f / (mx - mn) : mn;
  // Render() stores the default once (LightedVolumeRenderer.cxx:815-819)
  if (h->sample_distance <= 0.0f) h->sample_distance = g.ext_mag / 200.0f;
  s.sample_dist = h->sample_distance;
  s.mesh_eps = g.ext_mag * 0.0001f;  // LightedVolumeRenderer.cxx:1261
  s.alpha_stop = h->alpha_stop;
  s.use_map = h->use_shadow ? 1 : 0;
  s.max_steps = max_ray_steps(g.ext_mag, s.sample_dist);
  s.counters = h->counting ? (unsigned long long*)h->counters.p : nullptr;
  return s;
}

// Pinhole basis of a beams_camera (FP64); false for a degenerate camera.
bool camera_params(const beams_camera* cam, int32_t width, int32_t height, CameraParams& c)
{
  double f[3], len = 0.0;
  for (int d = 0; d < 3; ++d) {
    f[d] = cam->look_at[d] - cam->eye[d];
    len += f[d] * f[d];
  }
  len = std::sqrt(len);
  if (!(len > 0.0)) return false;
  for (int d = 0; d < 3; ++d) f[d] /= len;
  double r[3] = { f[1] * cam->up[2] - f[2] * cam->up[1], f[2] * cam->up[0] - f[0] * cam->up[2],
                  f[0] * cam->up[1] - f[1] * cam->up[0] };
  const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  if (!(rl > 0.0)) return false;
  for (int d = 0; d < 3; ++d) r[d] /= rl;
  const double u[3] = { r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2],
                        r[0] * f[1] - r[1] * f[0] };
  for (int d = 0; d < 3; ++d) {
    c.eye[d] = cam->eye[d];
    c.fwd[d] = f[d];
    c.right[d] = r[d];
    c.up[d] = u[d];
  }
  c.tan_half_fov = std::tan(cam->fov_y_deg * M_PI / 360.0);
  c.aspect = (double)width / (double)height;
  c.width = width;
  c.height = height;
  c.near_d = (float)cam->near_dist;
  c.far_d = cam->far_dist > 0.0 ? (float)cam->far_dist : INFINITY;
  return true;
}

// LDS a workgroup may allocate on gfx950, and the most the kernels' own fixed arrays (work
// queues, hit lists, reductions) take beside what a frame stages.
constexpr size_t kLdsPerGroup = 160 * 1024;
constexpr size_t kLdsFixed = 1024;

int check_ready(beams_renderer* h)
{
  if (!h->has_field) return h->fail(BEAMS_ERR_STATE, "no field: call beams_set_field first");
  if (h->htf.empty()) return h->fail(BEAMS_ERR_STATE, "no colormap: call beams_set_colormap first");
  if (!h->has_range) return h->fail(BEAMS_ERR_STATE, "no scalar range: call beams_set_scalar_range");
  if (h->dtype == BEAMS_F64) return h->fail(BEAMS_ERR_UNSUPPORTED, "FP64 field storage is not rendered");
  // The camera marches stage the colormap as float4 beside the coordinates, their reciprocal
  // widths and (uneven coordinates) the bin tables; the map and sweep kernels stage less.
  const size_t ntf = h->htf.size() / 4;
  const size_t ncoord = (size_t)(h->dims[0] + h->dims[1] + h->dims[2]);
  const size_t staged = 16 * ntf + (h->use_bins ? 12 : 8) * ncoord;
  if (staged + kLdsFixed > kLdsPerGroup)
    return h->fail(BEAMS_ERR_UNSUPPORTED,
                   "colormap and coordinates exceed the LDS a workgroup stages: 16 B per colormap "
                   "entry + 8 B per coordinate point (12 B for unevenly spaced coordinates) must "
                   "stay within 159 KiB (got " + std::to_string(staged) + " B)");
  if (h->use_shadow && h->assoc == 0 && h->shadow_mode == BEAMS_SHADOW_LEARNED && ntf > 1024)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "the learned closure stages at most 1024 colormap entries");
  if (h->use_shadow && h->assoc == 0) {
    if (!h->has_light) return h->fail(BEAMS_ERR_STATE, "shadow map needs a light: call beams_set_light");
    float ma = 0.0f;
    for (size_t i = 3; i < h->htf.size(); i += 4) ma = std::fmax(ma, h->htf[i]);
    // The reference divides by the maximum alpha (TransmittanceMap.h:329) and produces NaN.
    if (!(ma > 0.0f)) return h->fail(BEAMS_ERR_INVALID, "shadow map needs a colormap with nonzero alpha");
  }
  return BEAMS_OK;
}

int frame_begin(beams_renderer* h)
{
  const int slot = (int)(h->total_frames % beams_renderer::kMaxFrames);
  if ((int)h->frames.size() <= slot) {
    h->frames.emplace_back();
    for (auto& e : h->frames.back().ev) HIP_TRY(h, hipEventCreate(&e));
  }
  h->cur = slot;
  ++h->total_frames;
  for (bool& t : h->frames[slot].timed) t = false;
  if (h->counting) {
    HIP_TRY(h, h->counters.ensure(2 * sizeof(unsigned long long)));
    HIP_TRY(h, launch_zero_u64((unsigned long long*)h->counters.p, 2, h->stream));
  }
  return BEAMS_OK;
}

int phase_mark(beams_renderer* h, int phase, bool end)
{
  auto& f = h->frames[h->cur];
  HIP_TRY(h, hipEventRecord(f.ev[2 * phase + (end ? 1 : 0)], h->stream));
  if (end) f.timed[phase] = true;
  return BEAMS_OK;
}

// Phase 1 (and 3, which equals phase 1 for a single block): build the map.
// Device buffers of the single-block map build (allocated outside any graph capture).
int prepare_map(beams_renderer* h, int32_t mode)
{
  const int64_t nv = (h->map_cells[0] + 1) * (h->map_cells[1] + 1) * (h->map_cells[2] + 1);
  HIP_TRY(h, h->map.ensure((size_t)nv * dtype_size(h->map_dtype)));
  if (mode == BEAMS_SHADOW_LEARNED) {
    const size_t nc = (size_t)(kCoarseCells + 1) * (kCoarseCells + 1) * (kCoarseCells + 1);
    HIP_TRY(h, h->coarse.ensure(2 * nc * sizeof(float) + 6 * (kCoarseCells + 1) * sizeof(int32_t)));
    HIP_TRY(h, h->lerr.ensure(2 * sizeof(unsigned int)));
  }
  if (mode == BEAMS_SHADOW_SWEEP) {
    int axis, a1, a2, step;
    if (!sweep_axis(h, axis, a1, a2, step))
      return h->fail(BEAMS_ERR_UNSUPPORTED, "the sweep needs a light outside the volume's bounds");
    const size_t ns = (size_t)(h->map_cells[a1] + 1) * (size_t)(h->map_cells[a2] + 1);
    HIP_TRY(h, h->sweep_slices.ensure(2 * ns * sizeof(float2)));
    if (h->sweep_block) HIP_TRY(h, h->sweep_sigma.ensure(ns * (size_t)(h->map_cells[axis] + 1) * sizeof(float)));
  }
  return BEAMS_OK;
}

// Enqueues the map build: the reference's per-vertex march, or the exact sweep.
int enqueue_map(beams_renderer* h, const GridParams& g, const MapParams& m, const ShadeParams& s,
                int32_t mode)
{
  if (mode == BEAMS_SHADOW_SWEEP) {
    SweepParams sp{};
    sp.g = g;
    sp.m = m;
    sp.s = s;
    sweep_axis(h, sp.axis, sp.a1, sp.a2, sp.step);
    sp.f = sweep_frame(m, sp.axis, sp.a1, sp.a2);
    sp.h = m.step;
    sp.inv_h = 1.0f / m.step;
    const size_t ns = (size_t)m.mp[sp.a1] * (size_t)m.mp[sp.a2];
    sp.prev = (float2*)h->sweep_slices.p;
    sp.cur = sp.prev + ns;
    const SweepPlan plan = sweep_plan(sp);
    HIP_TRY(h, launch_sweep(sp, plan, h->sweep_block ? (float*)h->sweep_sigma.p : nullptr, h->stream));
  } else {
    HIP_TRY(h, launch_tmap_ref(g, m, s, nullptr, h->stream));
  }
  return BEAMS_OK;
}

LearnedParams learned_params(beams_renderer* h, const GridParams& g, const ShadeParams& s,
                             const MapParams& m)
{
  LearnedParams p{};
  p.g = g;
  p.m = m;
  p.s = s;
  p.cn = kCoarseCells + 1;
  p.coarse_sigma = (float*)h->coarse.p;
  p.coarse_tau = p.coarse_sigma + (size_t)p.cn * p.cn * p.cn;
  p.coarse_ranges = (int32_t*)(p.coarse_tau + (size_t)p.cn * p.cn * p.cn);
  for (int d = 0; d < 3; ++d) {
    p.corg[d] = (float)h->bounds[2 * d];
    p.csp[d] = (float)(h->bounds[2 * d + 1] - h->bounds[2 * d]) / (float)kCoarseCells;
    p.cinv[d] = 1.0f / p.csp[d];
  }
  p.inv_h = 1.0f / m.step;
  p.params = h->net ? mlp_device_params(h->net) : nullptr;
  p.err = (unsigned int*)h->lerr.p;
  p.check_stride = h->check_stride;
  const double bytes = (double)dtype_size(g.dtype) * (double)g.n[0] * (double)g.n[1] * (double)g.n[2];
  p.field_small = bytes < 4294967296.0 ? 1 : 0;
  return p;
}

void learned_built(beams_renderer* h, int32_t kind)
{
  h->learned_kind = kind;
  h->built_budget = h->error_budget;
}

// The learned map's TF-index field (one byte per field point) when the field and colormap
// allow it, else none (the feature patches then read the field).
int prepare_tfidx(beams_renderer* h, const LearnedParams& p)
{
  if (tf_index_field_ok(p)) {
    HIP_TRY(h, h->tfidx.ensure((size_t)(p.g.n[0] * p.g.n[1] * p.g.n[2])));
  } else if (h->tfidx.p) {
    HIP_TRY(h, hipStreamSynchronize(h->stream));  // a replay in flight may read it
    h->tfidx.release();
  }
  return BEAMS_OK;
}

// Device buffers of the learned map build and its fallback (allocated outside any capture).
// With the reference-march fallback, the single block is re-marched tile by tile like a brick:
// a one-brick frame (its descriptor, queue counters and tables uploaded when they change).
int prepare_learned(beams_renderer* h, const GridParams& g, const ShadeParams& s)
{
  int rc = prepare_map(h, BEAMS_SHADOW_LEARNED);
  if (!rc) rc = prepare_map(h, h->fallback_mode);
  if (!rc) rc = prepare_tfidx(h, learned_params(h, g, s, map_params(h, g, BEAMS_SHADOW_LEARNED)));
  if (rc) return rc;
  h->sb_tiles = h->fallback_mode == BEAMS_SHADOW_REF_MARCH;
  if (!h->sb_tiles) return BEAMS_OK;
  BrickDesc d{};
  d.g = g;
  d.m = map_params(h, g, BEAMS_SHADOW_LEARNED);
  d.s = s;
  d.s.counters = nullptr;
  d.map_final = h->map.p;
  d.nverts = d.m.mp[0] * d.m.mp[1] * d.m.mp[2];
  int64_t slots = 1, tiles = 1;
  for (int a = 0; a < 3; ++a) {
    const int64_t c = check_count(h->map_cells[a], h->check_stride);
    slots *= c;
    tiles *= c - 1;
  }
  slots = (slots + 255) / 256 * 256;
  if (slots > INT32_MAX || tiles > INT32_MAX - 2)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "learned map: more than 2^31 check vertices or tiles");
  const int32_t tab[4] = { 0, (int32_t)slots, 0, (int32_t)tiles };
  std::vector<unsigned char> host(sizeof(d) + sizeof(tab));
  std::memcpy(host.data(), &d, sizeof(d));
  std::memcpy(host.data() + sizeof(d), tab, sizeof(tab));
  HIP_TRY(h, h->sb_desc.ensure(sizeof(BrickDesc)));
  HIP_TRY(h, h->sb_tab.ensure(sizeof(tab)));
  HIP_TRY(h, h->sb_queue.ensure(5 * kQueueSlots * sizeof(unsigned int)));
  HIP_TRY(h, h->sb_err.ensure(sizeof(float) * (size_t)slots));
  HIP_TRY(h, h->sb_list.ensure(sizeof(int32_t) * (2 + (size_t)tiles)));
  if (host != h->sb_host) {
    HIP_TRY(h, hipStreamSynchronize(h->stream));  // a replay in flight may read them
    HIP_TRY(h, hipMemcpy(h->sb_desc.p, &d, sizeof(d), hipMemcpyHostToDevice));
    HIP_TRY(h, hipMemcpy(h->sb_tab.p, tab, sizeof(tab), hipMemcpyHostToDevice));
    h->sb_host.swap(host);
  }
  BrickFrame& f = h->sb_f;
  f = BrickFrame{};
  f.bricks = (const BrickDesc*)h->sb_desc.p;
  f.nbricks = 1;
  f.queue = (unsigned int*)h->sb_queue.p;
  f.persistent_groups = h->persistent_groups;
  f.xcd_ranges = h->xcd_ranges;
  // the fallback tiles' march: the schedule's switches (the launch checks the field's size)
  f.map_fast = h->map_fast_fetch;
  f.map_pipeline = h->map_pipeline;
  LearnedBricks& lb = h->sb_lb;
  lb.chk_off = (const int32_t*)h->sb_tab.p;
  lb.tile_off = lb.chk_off + 2;
  lb.verr = (float*)h->sb_err.p;
  lb.list = (int32_t*)h->sb_list.p;
  lb.tile_fail = nullptr;
  lb.n_chk_items = (int32_t)(slots / 256);
  lb.n_tiles = (int32_t)tiles;
  lb.stride = h->check_stride;
  lb.budget = h->error_budget;
  return BEAMS_OK;
}

// The learned map: coarse optical depth, the network's correction per vertex, the sparse exact
// check, and the fallback build gated on the device by the check's largest error (it returns at
// once when the error is within the budget), so the whole sequence is enqueued without a host
// round trip and can be captured.  The fallback writes the learned map's layout (x fastest),
// so the march reads the map the same way whichever build ran.
int enqueue_learned(beams_renderer* h, const GridParams& g, const ShadeParams& s)
{
  const MapParams m = map_params(h, g, BEAMS_SHADOW_LEARNED);
  LearnedParams p = learned_params(h, g, s, m);
  p.verr = h->sb_tiles ? h->sb_lb.verr : nullptr;
  HIP_TRY(h, launch_coarse(p, h->stream));
  if (h->tfidx.p) {
    HIP_TRY(h, launch_tf_index_field(p, (uint8_t*)h->tfidx.p, h->stream));
    p.tfidx = (const uint8_t*)h->tfidx.p;
  }
  HIP_TRY(h, launch_learned_map(p, h->stream));
  HIP_TRY(h, launch_zero_u64((unsigned long long*)h->lerr.p, 1, h->stream));
  HIP_TRY(h, launch_learned_check(p, h->stream));
  if (h->sb_tiles) {
    // the tiles over the budget, re-marched with the reference march (a one-brick frame)
    HIP_TRY(h, launch_zero_u32((unsigned int*)h->sb_lb.list, 2, h->stream));
    HIP_TRY(h, launch_zero_u32(h->sb_f.queue + 4 * kQueueSlots, kQueueSlots, h->stream));
    HIP_TRY(h, launch_learned_tiles(h->sb_f, h->sb_lb, h->stream));
    return launch_bricks_exact_tiles(p, h->sb_f, h->map_dtype, h->sb_lb, h->stream) == hipSuccess
               ? BEAMS_OK
               : h->fail(BEAMS_ERR_HIP, "launch_bricks_exact_tiles failed");
  }
  MapParams fm = map_params(h, g, h->fallback_mode);
  for (int d = 0; d < 3; ++d) fm.ms[d] = m.ms[d];
  fm.gate = (const unsigned int*)h->lerr.p;
  fm.gate_budget = h->error_budget;
  return enqueue_map(h, g, fm, s, h->fallback_mode);
}

int build_map(beams_renderer* h, const GridParams& g, const ShadeParams& s)
{
  if (h->shadow_mode == BEAMS_SHADOW_LEARNED) {
    int rc = prepare_learned(h, g, s);
    if (!rc) rc = enqueue_learned(h, g, s);
    if (rc) return rc;
    learned_built(h, 1);
    h->built_mode = BEAMS_SHADOW_LEARNED;
    h->map_valid = true;
    return BEAMS_OK;
  }
  int rc = prepare_map(h, h->shadow_mode);
  if (rc) return rc;
  MapParams m = map_params(h, g, h->shadow_mode);
  if ((rc = enqueue_map(h, g, m, s, h->shadow_mode))) return rc;
  h->built_mode = h->shadow_mode;
  h->map_valid = true;
  return BEAMS_OK;
}

// ---- bricks ---------------------------------------------------------------------------------

// Front-to-back visibility order of boxes (lo/hi per axis) that form a kd decomposition: an
// axis-aligned plane that no box of `set` straddles (boxes sorted by their low face; the first
// prefix whose highest high face is at or below the next box's low face), the eye's side
// first, then each side recursively.  False when some subset has no such plane.
bool kd_front_to_back(const std::vector<std::array<double, 6>>& boxes, const double eye[3],
                      std::vector<int32_t> set, std::vector<int32_t>& out)
{
  if (set.size() <= 1) {
    out.insert(out.end(), set.begin(), set.end());
    return true;
  }
  for (int a = 0; a < 3; ++a) {
    std::stable_sort(set.begin(), set.end(), [&](int32_t p, int32_t q) {
      return boxes[(size_t)p][(size_t)(2 * a)] < boxes[(size_t)q][(size_t)(2 * a)];
    });
    double hi = -INFINITY;
    for (size_t k = 0; k + 1 < set.size(); ++k) {
      hi = std::max(hi, boxes[(size_t)set[k]][(size_t)(2 * a + 1)]);
      const double plane = boxes[(size_t)set[k + 1]][(size_t)(2 * a)];
      if (hi <= plane) {
        std::vector<int32_t> low(set.begin(), set.begin() + (long)k + 1), high(set.begin() + (long)k + 1, set.end());
        const bool low_first = eye[a] < plane;
        return kd_front_to_back(boxes, eye, low_first ? low : high, out) &&
               kd_front_to_back(boxes, eye, low_first ? high : low, out);
      }
    }
  }
  return false;
}

// Screen rectangle (x0, y0, w, h) that contains every pixel whose camera ray can meet the
// box: the bounding box of the projected corners plus 2 pixels, or the whole image when a
// corner is not in front of the camera.
void footprint(const CameraParams& c, const double* bb, int32_t* fp)
{
  double xmin = 1e300, xmax = -1e300, ymin = 1e300, ymax = -1e300;
  bool front = true;
  for (int k = 0; k < 8; ++k) {
    const double p[3] = { bb[(k & 1) ? 1 : 0], bb[(k & 2) ? 3 : 2], bb[(k & 4) ? 5 : 4] };
    double v[3], z = 0.0, xr = 0.0, yu = 0.0;
    for (int d = 0; d < 3; ++d) v[d] = p[d] - c.eye[d];
    for (int d = 0; d < 3; ++d) {
      z += v[d] * c.fwd[d];
      xr += v[d] * c.right[d];
      yu += v[d] * c.up[d];
    }
    if (!(z > 1e-6)) {
      front = false;
      break;
    }
    const double px = (xr / (z * c.tan_half_fov * c.aspect) + 1.0) * 0.5 * c.width - 0.5;
    const double py = (1.0 - yu / (z * c.tan_half_fov)) * 0.5 * c.height - 0.5;
    xmin = std::min(xmin, px);
    xmax = std::max(xmax, px);
    ymin = std::min(ymin, py);
    ymax = std::max(ymax, py);
  }
  int64_t x0 = 0, x1 = c.width - 1, y0 = 0, y1 = c.height - 1;
  if (front) {
    x0 = std::max<int64_t>(0, (int64_t)std::floor(xmin) - 2);
    x1 = std::min<int64_t>(c.width - 1, (int64_t)std::ceil(xmax) + 2);
    y0 = std::max<int64_t>(0, (int64_t)std::floor(ymin) - 2);
    y1 = std::min<int64_t>(c.height - 1, (int64_t)std::ceil(ymax) + 2);
  }
  fp[0] = (int32_t)x0;
  fp[1] = (int32_t)y0;
  fp[2] = x1 >= x0 ? (int32_t)(x1 - x0 + 1) : 0;
  fp[3] = y1 >= y0 ? (int32_t)(y1 - y0 + 1) : 0;
}

template <typename T>
size_t append(std::vector<unsigned char>& buf, const T* data, size_t n)
{
  const size_t off = (buf.size() + 255) & ~size_t(255);
  buf.resize(off + sizeof(T) * n);
  if (n) std::memcpy(buf.data() + off, data, sizeof(T) * n);
  return off;
}

struct KeyBuilder {
  std::vector<unsigned char> bytes;
  template <typename T>
  KeyBuilder& put(const T& v)
  {
    const unsigned char* p = (const unsigned char*)&v;
    bytes.insert(bytes.end(), p, p + sizeof(T));
    return *this;
  }
};

// The rows (j, k) of a brick's map in light-sheet order, packed j | k << 16.  Seen along the
// light, the rows (lines along x) that lie in one plane through the light's direction share
// their light paths: with the depth axis a (y or z, whichever the light travels along more)
// and the other axis e, row (j, k) at `depth` cells from the face the light enters through
// meets that face at e-index ie - depth (msp_a / msp_e) (d_e / |d_a|).  Rows are grouped by
// that index rounded (a sheet), sheets in order, each from the light's face inward.
void light_sheet_rows(const MapParams& m, int32_t light_kind, const float* light, const double* bb,
                      std::vector<uint32_t>& out)
{
  double d[3];
  for (int a = 0; a < 3; ++a) {
    const double c = 0.5 * (bb[2 * a] + bb[2 * a + 1]);
    d[a] = light_kind == BEAMS_LIGHT_DIRECTIONAL ? (double)light[a] : c - (double)light[a];
  }
  const int a = std::fabs(d[2]) >= std::fabs(d[1]) ? 2 : 1, e = 3 - a;
  const double da = std::fabs(d[a]) > 0.0 ? std::fabs(d[a]) : 1e-300;
  const double shear = ((double)m.msp[a] / (double)m.msp[e]) * (d[e] / da);
  const int64_t n1 = m.mp[1], n2 = m.mp[2], na = m.mp[a];
  struct Row {
    int64_t sheet, depth;
    uint32_t jk;
  };
  std::vector<Row> rows((size_t)(n1 * n2));
  size_t o = 0;
  for (int64_t k = 0; k < n2; ++k)
    for (int64_t j = 0; j < n1; ++j) {
      const int64_t ia = a == 2 ? k : j, ie = a == 2 ? j : k;
      const int64_t depth = d[a] >= 0.0 ? ia : na - 1 - ia;
      const double sh = std::min(std::max((double)ie - (double)depth * shear, -1e15), 1e15);
      rows[o++] = Row{ (int64_t)std::llround(sh), depth, (uint32_t)j | ((uint32_t)k << 16) };
    }
  std::stable_sort(rows.begin(), rows.end(), [](const Row& x, const Row& y) {
    return x.sheet != y.sheet ? x.sheet < y.sheet : x.depth < y.depth;
  });
  out.resize(rows.size());
  for (size_t i = 0; i < rows.size(); ++i) out[i] = rows[i].jk;
}

// Brick descriptors, work lists and depth order for camera c (host), uploaded when changed.
int prepare_bricks(beams_renderer* h, const CameraParams& c, BrickFrame& f, int64_t& total_verts,
                   size_t& partial_pixels)
{
  const GridParams whole = grid_params(h);
  const float user_sd = h->sample_distance;  // each rank defaults to its own |extent| / 200
  ShadeParams s0 = shade_params(h, whole);    // TF, range, counters; per-brick fields below
  h->sample_distance = user_sd;
  const size_t esz = dtype_size(h->dtype);
  const size_t msz = dtype_size(h->map_dtype);
  const int nb = (int)h->bricks.size();
  std::vector<BrickDesc> desc((size_t)nb);
  std::vector<int32_t> item_off((size_t)nb + 1, 0), px_brick, px_tile, order((size_t)nb);
  std::vector<float> depth((size_t)nb);
  int64_t row_total = 0, blk_total = 0;
  bool blk_ok = true;
  total_verts = 0;
  partial_pixels = 0;
  int32_t max_coords = 0;
  const float* xdev = (const float*)h->coords.p;
  const int64_t xoff[3] = { 0, h->dims[0], h->dims[0] + h->dims[1] };
  const int64_t ncoord = h->dims[0] + h->dims[1] + h->dims[2];
  for (int b = 0; b < nb; ++b) {
    const auto& br = h->bricks[(size_t)b];
    BrickDesc& d = desc[(size_t)b];
    GridParams& g = d.g;
    g.dtype = h->dtype;
    g.assoc = 0;
    const int64_t sx = br.sdims[0], sxy = br.sdims[0] * br.sdims[1];
    const size_t foff = esz * (size_t)((br.lo[2] - br.slo[2]) * sxy + (br.lo[1] - br.slo[1]) * sx + (br.lo[0] - br.slo[0]));
    g.field = (const unsigned char*)br.field.p + foff;
    const uint64_t fbytes = (uint64_t)esz * (uint64_t)(br.sdims[0] * br.sdims[1] * br.sdims[2]) - foff;
    g.bytes32 = fbytes < ((uint64_t)1 << 32) ? (uint32_t)fbytes : 0u;
    g.field_stride[0] = 1;
    g.field_stride[1] = sx;
    g.field_stride[2] = sxy;
    for (int k = 0; k < 3; ++k) {
      const std::vector<float>& x = h->hx[k];
      g.n[k] = br.hi[k] - br.lo[k] + 1;
      g.x[k] = xdev + xoff[k] + br.lo[k];
      g.r[k] = g.x[k] + ncoord;
      g.minp[k] = x[(size_t)br.lo[k]];
      g.maxp[k] = x[(size_t)br.hi[k]];
      g.gx0[k] = g.minp[k];
      g.ginv[k] = (float)((double)(g.n[k] - 1) / ((double)g.maxp[k] - (double)g.minp[k]));
      g.bounds[2 * k] = g.minp[k];
      g.bounds[2 * k + 1] = g.maxp[k];
    }
    const float e0 = (float)(g.bounds[1] - g.bounds[0]), e1 = (float)(g.bounds[3] - g.bounds[2]),
                e2 = (float)(g.bounds[5] - g.bounds[4]);
    g.ext_mag = std::sqrt(e0 * e0 + e1 * e1 + e2 * e2);
    max_coords = std::max<int32_t>(max_coords, (int32_t)(g.n[0] + g.n[1] + g.n[2]));
    // map over the brick bounds (CreateDataSetForOpacityMap per rank)
    MapParams& m = d.m;
    m.dtype = h->map_dtype;
    m.light_kind = h->light_kind;
    float size[3];
    for (int k = 0; k < 3; ++k) {
      m.mp[k] = br.cells[k] + 1;
      m.morg[k] = (float)g.bounds[2 * k];
      size[k] = (float)(g.bounds[2 * k + 1] - g.bounds[2 * k]);
      m.msp[k] = size[k] / (float)br.cells[k];
      m.minv[k] = 1.0f / m.msp[k];
      m.light[k] = h->light[k];
      m.light_color[k] = h->light_color[k];
    }
    m.ms[0] = 1;
    m.ms[1] = m.mp[0];
    m.ms[2] = m.mp[0] * m.mp[1];
    m.bytes32 = map_bytes32(m.mp, m.dtype);
    m.step = std::sqrt(size[0] * size[0] + size[1] * size[1] + size[2] * size[2]) / 128.0f;
    m.dir_far = 4.0f * whole.ext_mag;  // one light origin for all bricks
    m.map = (unsigned char*)h->bmap_local.p + msz * (size_t)br.seed_off;
    d.map_final = (unsigned char*)h->bmap_final.p + msz * (size_t)br.seed_off;
    {
      int64_t ybs, zbs;
      blocked_strides(g.n, ybs, zbs);
      d.field_blk = br.field_blk.p;
      d.fblk_ybs = ybs;
      d.fblk_zbs = zbs;
      const uint64_t fb = (uint64_t)esz * (uint64_t)blocked_count(g.n);
      d.fblk_bytes = fb < ((uint64_t)1 << 32) ? (uint32_t)fb : 0u;
      blocked_strides(m.mp, ybs, zbs);
      d.map_blk = (unsigned char*)h->bmap_blk.p + msz * (size_t)blk_total;
      d.mblk_ybs = ybs;
      d.mblk_zbs = zbs;
      const uint64_t mb = (uint64_t)msz * (uint64_t)blocked_count(m.mp);
      d.mblk_bytes = mb < ((uint64_t)1 << 32) ? (uint32_t)mb : 0u;
      blk_total += blocked_count(m.mp);
      blk_ok = blk_ok && d.fblk_bytes != 0 && d.mblk_bytes != 0;
    }
    // per-rank Render() state: its own extent sets the mesh epsilon (and the default step)
    ShadeParams& s = d.s;
    s = s0;
    s.sample_dist = user_sd > 0.0f ? user_sd : g.ext_mag / 200.0f;
    s.mesh_eps = g.ext_mag * 0.0001f;
    s.max_steps = max_ray_steps(g.ext_mag, s.sample_dist);
    d.nverts = br.nverts;
    d.seed_off = br.seed_off;
    for (int k = 0; k < 3; ++k) {
      const float pad = 1e-4f * (float)(g.bounds[2 * k + 1] - g.bounds[2 * k]) + 1e-6f;
      d.box32[2 * k] = (float)g.bounds[2 * k] - pad;
      d.box32[2 * k + 1] = (float)g.bounds[2 * k + 1] + pad;
    }
    footprint(c, g.bounds, d.fp);
    d.img_off = (int64_t)partial_pixels;
    partial_pixels += (size_t)d.fp[2] * (size_t)d.fp[3];
    item_off[(size_t)b + 1] = item_off[(size_t)b] + (int32_t)((br.nverts + kMapItem - 1) / kMapItem);
    d.row_off = row_total;
    row_total += m.mp[1] * m.mp[2];
    const int tiles = ((d.fp[2] + 15) / 16) * ((d.fp[3] + 15) / 16);
    for (int t = 0; t < tiles; ++t) {
      px_brick.push_back(b);
      px_tile.push_back(t);
    }
    total_verts += br.nverts;
    // MapperLightedVolume.cxx:222-236: |camera position - block centre| in FP32
    const float cx = (float)((g.bounds[0] + g.bounds[1]) * 0.5);
    const float cy = (float)((g.bounds[2] + g.bounds[3]) * 0.5);
    const float cz = (float)((g.bounds[4] + g.bounds[5]) * 0.5);
    const float dx = (float)c.eye[0] - cx, dy = (float)c.eye[1] - cy, dz = (float)c.eye[2] - cz;
    depth[(size_t)b] = std::sqrt(dx * dx + dy * dy + dz * dz);
    order[(size_t)b] = b;
  }
  if (h->composite_order == BEAMS_ORDER_KD_FRONT_TO_BACK) {
    std::vector<std::array<double, 6>> boxes((size_t)nb);
    for (int b = 0; b < nb; ++b)
      for (int k = 0; k < 6; ++k) boxes[(size_t)b][(size_t)k] = desc[(size_t)b].g.bounds[k];
    order.clear();
    std::vector<int32_t> all((size_t)nb);
    for (int b = 0; b < nb; ++b) all[(size_t)b] = b;
    if (!kd_front_to_back(boxes, c.eye, all, order))
      return h->fail(BEAMS_ERR_UNSUPPORTED, "kd composite order: the bricks are not a kd decomposition "
                                            "(some set has no axis-aligned plane that no brick straddles)");
  } else {
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return depth[(size_t)a] < depth[(size_t)b]; });
  }
  HIP_TRY(h, h->bpartial.ensure(std::max<size_t>(partial_pixels, 1) * 20));
  // the light-sheet row orders, recomputed when the light or a brick's map changes
  if (h->light_rows) {
    KeyBuilder key;
    key.put(h->light_kind).put(h->light[0]).put(h->light[1]).put(h->light[2]);
    for (const BrickDesc& d : desc) key.put(d.m.mp).put(d.m.msp).put(d.g.bounds);
    if (key.bytes != h->brows_key) {
      std::vector<uint32_t> all((size_t)row_total), one;
      for (int b = 0; b < nb; ++b) {
        light_sheet_rows(desc[(size_t)b].m, h->light_kind, h->light, desc[(size_t)b].g.bounds, one);
        std::copy(one.begin(), one.end(), all.begin() + desc[(size_t)b].row_off);
      }
      HIP_TRY(h, hipStreamSynchronize(h->stream));
      HIP_TRY(h, h->brows.ensure(sizeof(uint32_t) * all.size()));
      HIP_TRY(h, hipMemcpy(h->brows.p, all.data(), sizeof(uint32_t) * all.size(), hipMemcpyHostToDevice));
      h->brows_key.swap(key.bytes);
    }
  }
  std::vector<unsigned char> tab;
  const size_t o_desc = append(tab, desc.data(), desc.size());
  const size_t o_mio = append(tab, item_off.data(), item_off.size());
  const size_t o_pxb = append(tab, px_brick.data(), px_brick.size());
  const size_t o_pxt = append(tab, px_tile.data(), px_tile.size());
  const size_t o_ord = append(tab, order.data(), order.size());
  if (tab != h->btab_host) {
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    HIP_TRY(h, h->btables.ensure(tab.size()));
    HIP_TRY(h, hipMemcpy(h->btables.p, tab.data(), tab.size(), hipMemcpyHostToDevice));
    h->btab_host.swap(tab);
    ++h->btab_version;
  }
  unsigned char* base = (unsigned char*)h->btables.p;
  f = BrickFrame{};
  f.bricks = (const BrickDesc*)(base + o_desc);
  f.nbricks = nb;
  f.map_item_off = (const int32_t*)(base + o_mio);
  f.rows = h->light_rows ? (const uint32_t*)h->brows.p : nullptr;
  f.n_map_items = item_off.back();
  f.px_item_brick = (const int32_t*)(base + o_pxb);
  f.px_item_tile = (const int32_t*)(base + o_pxt);
  f.n_px_items = (int32_t)px_brick.size();
  f.depth_order = (const int32_t*)(base + o_ord);
  f.queue = (unsigned int*)h->bqueue.p;
  f.seeds = (float*)h->bseeds.p;
  f.rm_count = f.queue + 5 * kQueueSlots;
  f.rm_item_off = (int32_t*)(f.rm_count + nb);
  f.rm_list = (uint32_t*)h->brm_list.p;
  f.partial = (float4*)h->bpartial.p;
  f.partial_depth = (float*)((float4*)h->bpartial.p + partial_pixels);
  f.max_coords = max_coords;
  f.persistent_groups = h->persistent_groups;
  f.xcd_ranges = h->xcd_ranges;
  f.light_rows = h->light_rows;
  f.march_xcd_tiles = h->march_xcd_tiles;
  bool fast_ok = true;
  for (const BrickDesc& d : desc)
    fast_ok = fast_ok && d.g.bytes32 != 0 && d.m.bytes32 != 0 && d.g.field_stride[1] < (1 << 24) &&
              d.g.field_stride[2] < (1 << 24) && d.m.ms[1] < (1 << 24) && d.m.ms[2] < (1 << 24);
  f.map_fast = fast_ok && h->map_fast_fetch;
  f.march_fast = fast_ok && h->march_fast_fetch;
  f.map_pipeline = h->map_pipeline;
  f.march_groups_per_cu = h->march_groups_per_cu;
  f.march_blocked = h->march_blocked && blk_ok;
  h->n_map_items = f.n_map_items;
  h->n_px_items = f.n_px_items;
  h->max_coords = max_coords;
  return BEAMS_OK;
}

// Runs the frame's launches `issue(timed)` eagerly, or as a replay of a hipGraph captured
// from them; the graph is re-captured whenever the key (all kernel arguments) changes.
template <typename Issue>
int run_frame(beams_renderer* h, const std::vector<unsigned char>& key, Issue&& issue)
{
  int rc;
  if (!h->use_graph) {
    if ((rc = phase_mark(h, BEAMS_PHASE_FRAME, false))) return rc;
    if ((rc = issue(true))) return rc;
    return phase_mark(h, BEAMS_PHASE_FRAME, true);
  }
  if (!h->graph_exec || key != h->graph_key) {
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    h->graph_exec = nullptr;
    HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    rc = issue(false);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (rc) {
      if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIP_TRY(h, e);
    const hipError_t ei = hipGraphInstantiate(&h->graph_exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_TRY(h, ei);
    h->graph_key = key;
  }
  if ((rc = phase_mark(h, BEAMS_PHASE_FRAME, false))) return rc;
  HIP_TRY(h, hipGraphLaunch(h->graph_exec, h->stream));
  return phase_mark(h, BEAMS_PHASE_FRAME, true);
}

// Learned bricks (phases 1 and 3 of the frame): every brick's final map from the whole volume's
// coarse optical depth and the network, a per-brick sparse check against the whole volume's
// light march (the single-block map at the brick's vertices), the fallback's work list built on
// the device from the checks, and that march for every vertex of the listed bricks.  The whole
// field stays resident beside the bricks' ghosted copies.
int prepare_bricks_learned(beams_renderer* h, const BrickFrame& f)
{
  const int nb = f.nbricks;
  const size_t nc = (size_t)(kCoarseCells + 1) * (kCoarseCells + 1) * (kCoarseCells + 1);
  HIP_TRY(h, h->coarse.ensure(2 * nc * sizeof(float) + 6 * (kCoarseCells + 1) * sizeof(int32_t)));
  std::vector<int32_t> tab(2 * ((size_t)nb + 1), 0);
  int64_t slots = 0, tiles = 0;
  for (int b = 0; b < nb; ++b) {
    const auto& br = h->bricks[(size_t)b];
    int64_t n = 1, t = 1;
    for (int d = 0; d < 3; ++d) {
      const int64_t c = check_count(br.cells[d], h->check_stride);
      n *= c;
      t *= c - 1;
    }
    slots += (n + 255) / 256 * 256;
    tiles += t;
    if (slots > INT32_MAX || tiles > INT32_MAX - 2)
      return h->fail(BEAMS_ERR_UNSUPPORTED, "learned bricks: more than 2^31 check vertices or tiles");
    tab[(size_t)b + 1] = (int32_t)slots;
    tab[(size_t)nb + 2 + (size_t)b] = (int32_t)tiles;
  }
  if (tab != h->lb_host) {
    HIP_TRY(h, h->lb_tab.ensure(sizeof(int32_t) * tab.size()));
    HIP_TRY(h, hipStreamSynchronize(h->stream));  // a replay in flight may read the table
    HIP_TRY(h, hipMemcpy(h->lb_tab.p, tab.data(), sizeof(int32_t) * tab.size(), hipMemcpyHostToDevice));
    h->lb_host = tab;
  }
  HIP_TRY(h, h->lb_err.ensure(sizeof(float) * (size_t)slots));
  HIP_TRY(h, h->lb_list.ensure(sizeof(int32_t) * (2 + (size_t)tiles)));
  HIP_TRY(h, h->lb_fail.ensure((size_t)tiles + 1));
  LearnedBricks& lb = h->lb;
  lb.chk_off = (const int32_t*)h->lb_tab.p;
  lb.tile_off = lb.chk_off + nb + 1;
  lb.verr = (float*)h->lb_err.p;
  lb.list = (int32_t*)h->lb_list.p;
  lb.tile_fail = (uint8_t*)h->lb_fail.p;
  lb.n_chk_items = (int32_t)(slots / 256);
  lb.n_tiles = (int32_t)tiles;
  lb.stride = h->check_stride;
  lb.budget = h->error_budget;
  return BEAMS_OK;
}

LearnedParams bricks_learned_params(beams_renderer* h)
{
  const GridParams g = grid_params(h);
  const float user_sd = h->sample_distance;
  ShadeParams s = shade_params(h, g);
  h->sample_distance = user_sd;
  s.counters = nullptr;
  return learned_params(h, g, s, map_params(h, g, BEAMS_SHADOW_LEARNED));
}

int enqueue_bricks_learned(beams_renderer* h, const BrickFrame& f, const LearnedParams& p,
                           const std::function<int(int, bool)>& mark)
{
  int rc;
  if ((rc = mark(BEAMS_PHASE_MAP, false))) return rc;
  HIP_TRY(h, launch_coarse(p, h->stream));
  LearnedParams pm = p;
  if (h->tfidx.p) {
    HIP_TRY(h, launch_tf_index_field(p, (uint8_t*)h->tfidx.p, h->stream));
    pm.tfidx = (const uint8_t*)h->tfidx.p;
  }
  // the network at the check vertices, the check and the tile list, then the network at the
  // other vertices (waves wholly inside listed tiles skipped), then the fallback
  HIP_TRY(h, launch_bricks_learned_map(pm, f, h->map_dtype, &h->lb, true, h->stream));
  HIP_TRY(h, launch_bricks_learned_check(p, f, h->map_dtype, h->lb, h->stream));
  HIP_TRY(h, launch_bricks_learned_map(pm, f, h->map_dtype, &h->lb, false, h->stream));
  if ((rc = mark(BEAMS_PHASE_MAP, true))) return rc;
  if ((rc = mark(BEAMS_PHASE_REMAP, false))) return rc;
  HIP_TRY(h, launch_bricks_exact_tiles(p, f, h->map_dtype, h->lb, h->stream));
  return mark(BEAMS_PHASE_REMAP, true);
}

int render_bricks(beams_renderer* h, const CameraParams& c, const float* background, float* rgba,
                  float* depth, int32_t mem)
{
  int rc;
  if (h->use_shadow && h->shadow_mode == BEAMS_SHADOW_SWEEP)
    return h->fail(BEAMS_ERR_UNSUPPORTED,
                   "bricks build their maps with the reference march or the learned closure");
  const bool learned = h->use_shadow && h->shadow_mode == BEAMS_SHADOW_LEARNED;
  const size_t msz = dtype_size(h->map_dtype);
  int64_t nv_total = 0;
  for (const auto& br : h->bricks) nv_total += br.nverts;
  HIP_TRY(h, h->bmap_local.ensure((size_t)nv_total * msz));
  HIP_TRY(h, h->bmap_final.ensure((size_t)nv_total * msz));
  int64_t blk_total = 0;
  for (const auto& br : h->bricks) {
    const int64_t mp[3] = { br.cells[0] + 1, br.cells[1] + 1, br.cells[2] + 1 };
    blk_total += blocked_count(mp);
  }
  HIP_TRY(h, h->bmap_blk.ensure((size_t)blk_total * msz));
  HIP_TRY(h, h->bseeds.ensure((size_t)nv_total * sizeof(float)));
  HIP_TRY(h, h->brm_list.ensure((size_t)nv_total * sizeof(uint32_t)));
  // the queue counters, then per brick phase 3's list count and item offsets (and the total)
  const size_t nbr = h->bricks.size();
  HIP_TRY(h, h->bqueue.ensure((5 * kQueueSlots + 2 * nbr + 1) * sizeof(unsigned int)));
  if ((rc = frame_begin(h))) return rc;
  BrickFrame f;
  int64_t total_verts = 0;
  size_t partial_pixels = 0;
  if ((rc = prepare_bricks(h, c, f, total_verts, partial_pixels))) return rc;
  const int32_t ntf = (int32_t)(h->htf.size() / 4);
  const bool shadows = h->use_shadow;
  const size_t npix = (size_t)c.width * (size_t)c.height;
  ImageOut o{};
  if (mem == BEAMS_MEM_DEVICE) {
    o.rgba = rgba;
    o.depth = depth;
  } else {
    HIP_TRY(h, h->image.ensure(sizeof(float) * npix * 5));
    o.rgba = (float*)h->image.p;
    o.depth = depth ? (float*)h->image.p + 4 * npix : nullptr;
  }
  o.use_bg = background ? 1 : 0;
  for (int k = 0; k < 4; ++k) o.bg[k] = background ? background[k] : 0.0f;
  LearnedParams lp{};
  if (learned) {
    if ((rc = prepare_bricks_learned(h, f))) return rc;
    lp = bricks_learned_params(h);
    if ((rc = prepare_tfidx(h, lp))) return rc;
  }
  auto issue = [&](bool timed) -> int {
    int r;
    const std::function<int(int, bool)> mark = [&](int p, bool end) {
      return timed ? phase_mark(h, p, end) : BEAMS_OK;
    };
    HIP_TRY(h, launch_zero_u32(f.queue, 5 * kQueueSlots + f.nbricks, h->stream));
    if (learned && (r = enqueue_bricks_learned(h, f, lp, mark))) return r;
    if (shadows && !learned) {
      if ((r = mark(BEAMS_PHASE_MAP, false))) return r;
      HIP_TRY(h, launch_bricks_map(f, h->dtype, h->map_dtype, ntf, false, h->stream));
      if ((r = mark(BEAMS_PHASE_MAP, true))) return r;
      if ((r = mark(BEAMS_PHASE_EXCHANGE, false))) return r;
      HIP_TRY(h, launch_bricks_exchange(f, h->map_dtype, total_verts, h->stream));
      if ((r = mark(BEAMS_PHASE_EXCHANGE, true))) return r;
      if ((r = mark(BEAMS_PHASE_REMAP, false))) return r;
      HIP_TRY(h, launch_bricks_finalize(f, h->map_dtype, h->stream));
      HIP_TRY(h, launch_bricks_map(f, h->dtype, h->map_dtype, ntf, true, h->stream));
      if ((r = mark(BEAMS_PHASE_REMAP, true))) return r;
    }
    if ((r = mark(BEAMS_PHASE_MARCH, false))) return r;
    if (shadows && f.march_blocked) HIP_TRY(h, launch_block_maps(f, h->map_dtype, h->stream));
    HIP_TRY(h, launch_bricks_march(f, h->dtype, h->map_dtype, ntf, shadows, c, h->stream));
    if ((r = mark(BEAMS_PHASE_MARCH, true))) return r;
    if ((r = mark(BEAMS_PHASE_COMPOSITE, false))) return r;
    HIP_TRY(h, launch_bricks_composite(f, c, o, h->stream));
    return mark(BEAMS_PHASE_COMPOSITE, true);
  };
  KeyBuilder key;
  key.put(1).put(f).put(c).put(o).put(h->btab_version).put(shadows).put(learned).put(h->dtype)
     .put(h->map_dtype).put(ntf);
  if (learned) key.put(lp).put(h->lb).put(h->tfidx.p);
  if ((rc = run_frame(h, key.bytes, issue))) return rc;
  if (learned) learned_built(h, 2);
  if (shadows) h->map_valid = true;
  if (mem == BEAMS_MEM_HOST) {
    HIP_TRY(h, hipMemcpyAsync(rgba, o.rgba, sizeof(float) * 4 * npix, hipMemcpyDeviceToHost, h->stream));
    if (depth) HIP_TRY(h, hipMemcpyAsync(depth, o.depth, sizeof(float) * npix, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
  }
  return BEAMS_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
extern "C" {

int beams_create(beams_renderer** out)
{
  if (!out) return BEAMS_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return BEAMS_ERR_NO_DEVICE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return BEAMS_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return BEAMS_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return BEAMS_ERR_NO_DEVICE;
  void* scr = nullptr;
  if (dev < kMaxDevices) {
    // the scratch array's address on this device, looked up once with the device current
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return BEAMS_ERR_HIP;
      g_scratch[dev] = p;
    }
    scr = g_scratch[dev];
  }
  auto* h = new beams_renderer();
  h->device = dev;
  h->scratch = scr;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return BEAMS_ERR_HIP;
  }
  *out = h;
  return BEAMS_OK;
}

int beams_destroy(beams_renderer* h)
{
  CHECK_HANDLE(h);
  (void)hipStreamSynchronize(h->stream);
  h->field.release();
  h->coords.release();
  h->tf.release();
  h->map.release();
  h->sweep_slices.release();
  h->sweep_sigma.release();
  h->coarse.release();
  h->tfidx.release();
  h->lerr.release();
  h->rays.release();
  h->image.release();
  for (auto& f : h->frames)
    for (auto& e : f.ev)
      if (e) (void)hipEventDestroy(e);
  h->counters.release();
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  for (auto& br : h->bricks) br.release();
  for (auto& br : h->brick_pool) br.release();
  h->bmap_local.release();
  h->bmap_final.release();
  h->bmap_blk.release();
  h->bseeds.release();
  h->btables.release();
  h->bpartial.release();
  h->bqueue.release();
  h->brm_list.release();
  h->brows.release();
  h->lb_tab.release();
  h->lb_err.release();
  h->lb_list.release();
  h->lb_fail.release();
  h->sb_desc.release();
  h->sb_queue.release();
  h->sb_tab.release();
  h->sb_err.release();
  h->sb_list.release();
  (void)hipStreamDestroy(h->stream);
  delete h;
  return BEAMS_OK;
}

const char* beams_last_error(const beams_renderer* h) { return h ? h->err.c_str() : "null handle"; }

int beams_synchronize(beams_renderer* h)
{
  CHECK_HANDLE(h);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return BEAMS_OK;
}

int beams_set_field(beams_renderer* h, const void* values, int32_t dtype, int32_t assoc,
                    const int64_t dims[3], const float* const coords[3], const double origin[3],
                    const double spacing[3], int32_t mem)
{
  CHECK_HANDLE(h);
  if (!values || !dims) return h->fail(BEAMS_ERR_INVALID, "null field values or dims");
  if (dtype != BEAMS_F32 && dtype != BEAMS_F16 && dtype != BEAMS_BF16 && dtype != BEAMS_F64)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "unsupported field dtype");
  if (assoc != BEAMS_ASSOC_POINTS && assoc != BEAMS_ASSOC_CELLS)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "field must be associated with points or cells");
  if (mem != BEAMS_MEM_HOST && mem != BEAMS_MEM_DEVICE) return h->fail(BEAMS_ERR_INVALID, "bad mem kind");
  for (int d = 0; d < 3; ++d)
    if (dims[d] < 2 || dims[d] > (1 << 20)) return h->fail(BEAMS_ERR_INVALID, "point dims must be in [2, 2^20]");
  if (!coords && (!origin || !spacing)) return h->fail(BEAMS_ERR_INVALID, "need coords or origin+spacing");
  std::vector<float> x[3];
  for (int d = 0; d < 3; ++d) {
    x[d].resize((size_t)dims[d]);
    if (coords) {
      if (!coords[d]) return h->fail(BEAMS_ERR_INVALID, "null coordinate array");
      std::memcpy(x[d].data(), coords[d], sizeof(float) * (size_t)dims[d]);
    } else {
      const float o = (float)origin[d], sp = (float)spacing[d];
      for (int64_t i = 0; i < dims[d]; ++i) x[d][(size_t)i] = o + sp * (float)i;
    }
    for (int64_t i = 1; i < dims[d]; ++i)
      if (!(x[d][(size_t)i] > x[d][(size_t)i - 1]))
        return h->fail(BEAMS_ERR_INVALID, "coordinates must be strictly increasing");
    // the kernels divide by a cell width through its reciprocal and one correction step,
    // which gives the IEEE quotient for widths in [2^-60, 2^61] and coordinates below 2^60
    for (int64_t i = 0; i < dims[d]; ++i) {
      const float v = x[d][(size_t)i];
      if (!(std::fabs(v) <= 0x1p60f))
        return h->fail(BEAMS_ERR_INVALID, "coordinates must be finite and below 2^60 in magnitude");
      if (i > 0 && !(v - x[d][(size_t)i - 1] >= 0x1p-60f))
        return h->fail(BEAMS_ERR_INVALID, "cell widths must be at least 2^-60");
    }
  }
  // the kernels stage the coordinates and reciprocal cell widths in LDS
  if (dims[0] + dims[1] + dims[2] > 16384)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "the three point dims must sum to at most 16384");
  // the kernels index cells and strides in 32 bits within an xy plane
  if (dims[0] * dims[1] >= ((int64_t)1 << 31))
    return h->fail(BEAMS_ERR_UNSUPPORTED, "a field's xy plane must hold fewer than 2^31 points");
  const int64_t nvals = (assoc == BEAMS_ASSOC_CELLS) ? (dims[0] - 1) * (dims[1] - 1) * (dims[2] - 1)
                                                     : dims[0] * dims[1] * dims[2];
  const size_t bytes = (size_t)nvals * dtype_size(dtype);
  HIP_TRY(h, h->field.ensure(bytes));
  HIP_TRY(h, hipMemcpyAsync(h->field.p, values, bytes, copy_kind(mem, true), h->stream));
  // coordinates of the three axes, then their reciprocal cell widths (IEEE FP32 quotients,
  // the last entry of each axis 0)
  const size_t nc = (size_t)(dims[0] + dims[1] + dims[2]);
  std::vector<float> flat;
  flat.reserve(2 * nc);
  for (int d = 0; d < 3; ++d) flat.insert(flat.end(), x[d].begin(), x[d].end());
  for (int d = 0; d < 3; ++d) {
    for (size_t i = 0; i + 1 < x[d].size(); ++i) {
      const float w = x[d][i + 1] - x[d][i];
      flat.push_back(1.0f / w);
    }
    flat.push_back(0.0f);
  }
  // Uneven coordinates (some cell not where the uniform-spacing guess puts it): per axis,
  // n-1 uniform bins, each holding the cell its lower edge lies in, so a guess lands within a
  // cell or two of the answer.  Staged in LDS beside the coordinates, so only up to 8192
  // points in all; beyond that the kernels walk from the uniform guess.
  bool uneven = false;
  float binv[3] = { 0.0f, 0.0f, 0.0f };
  for (int d = 0; d < 3; ++d) {
    const double x0 = x[d].front(), span = (double)x[d].back() - x0;
    const float ginv = (float)((double)(dims[d] - 1) / span);
    binv[d] = ginv;
    for (size_t c = 0; c + 1 < x[d].size() && !uneven; ++c) {
      const float mid = (float)(0.5 * ((double)x[d][c] + (double)x[d][c + 1]));
      uneven = (int64_t)((mid - (float)x0) * ginv) != (int64_t)c;
    }
  }
  const bool use_bins = uneven && nc <= 8192;
  if (use_bins)
    for (int d = 0; d < 3; ++d) {
      const int64_t ncell = dims[d] - 1;
      for (int64_t b = 0; b < dims[d]; ++b) {
        const double edge = (double)x[d].front() + (double)std::min(b, ncell - 1) / (double)binv[d];
        int64_t c = (int64_t)(std::upper_bound(x[d].begin(), x[d].end(), (float)edge) - x[d].begin()) - 1;
        c = std::min(std::max<int64_t>(c, 0), ncell - 1);
        const int32_t ci = (int32_t)c;
        float bits;
        std::memcpy(&bits, &ci, 4);
        flat.push_back(bits);
      }
    }
  const size_t words = (use_bins ? 3 : 2) * nc;
  HIP_TRY(h, h->coords.ensure(sizeof(float) * words));
  HIP_TRY(h, hipMemcpyAsync(h->coords.p, flat.data(), sizeof(float) * words, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->use_bins = use_bins;
  for (int d = 0; d < 3; ++d) h->binv[d] = binv[d];
  for (int d = 0; d < 3; ++d) {
    h->hx[d] = std::move(x[d]);
    h->dims[d] = dims[d];
    double lo = h->hx[d].front(), hi = h->hx[d].back();
    h->bounds[2 * d] = lo;
    h->bounds[2 * d + 1] = hi;
  }
  h->dtype = dtype;
  h->assoc = assoc;
  h->has_field = true;
  h->map_valid = false;
  // A new field drops the brick decomposition; its storage waits in the pool for the next
  // beams_set_bricks, which reuses it where the stored regions match.
  for (auto& br : h->brick_pool) br.release();
  h->brick_pool.swap(h->bricks);
  h->bricks.clear();
  return BEAMS_OK;
}

int beams_set_bricks(beams_renderer* h, int32_t nbricks, const int64_t* point_ranges,
                     int32_t ghost, const int64_t* map_cells)
{
  CHECK_HANDLE(h);
  if (nbricks == 0) {
    for (auto& br : h->bricks) br.release();
    h->bricks.clear();
    h->btab_host.clear();
    return BEAMS_OK;
  }
  if (nbricks < 0 || nbricks > 4096 || !point_ranges)
    return h->fail(BEAMS_ERR_INVALID, "bricks: need 1..4096 point ranges");
  if (ghost < 0 || ghost > 8) return h->fail(BEAMS_ERR_INVALID, "bricks: ghost layers must be in [0, 8]");
  if (!h->has_field) return h->fail(BEAMS_ERR_STATE, "bricks: call beams_set_field first");
  if (h->assoc != BEAMS_ASSOC_POINTS || h->dtype == BEAMS_F64)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "bricks need a point field stored as F32, F16 or BF16");
  std::vector<beams_renderer::Brick> bricks((size_t)nbricks);
  int64_t off = 0;
  for (int b = 0; b < nbricks; ++b) {
    auto& br = bricks[(size_t)b];
    for (int d = 0; d < 3; ++d) {
      br.lo[d] = point_ranges[6 * b + 2 * d];
      br.hi[d] = point_ranges[6 * b + 2 * d + 1];
      if (br.lo[d] < 0 || br.hi[d] >= h->dims[d] || br.hi[d] <= br.lo[d])
        return h->fail(BEAMS_ERR_INVALID, "bricks: each range must satisfy 0 <= lo < hi < dims");
      br.slo[d] = std::max<int64_t>(0, br.lo[d] - ghost);
      br.sdims[d] = std::min<int64_t>(h->dims[d] - 1, br.hi[d] + ghost) - br.slo[d] + 1;
      br.cells[d] = map_cells ? map_cells[3 * b + d] : br.hi[d] - br.lo[d];
      if (br.cells[d] < 1 || br.cells[d] > 4096)
        return h->fail(BEAMS_ERR_INVALID, "bricks: map cells must be in [1, 4096]");
    }
    br.nverts = (br.cells[0] + 1) * (br.cells[1] + 1) * (br.cells[2] + 1);
    // the brick kernels index a brick's map vertices in 32 bits
    if (br.nverts > INT32_MAX)
      return h->fail(BEAMS_ERR_UNSUPPORTED, "bricks: a brick's map must have fewer than 2^31 vertices");
    br.seed_off = off;
    off += br.nverts;
  }
  // Storage of the previous decomposition (the current one, or the one a new field replaced)
  // is taken over brick by brick, so that a series of fields on one decomposition keeps every
  // device address, the brick tables and a captured frame graph unchanged.
  std::vector<beams_renderer::Brick>& prev = h->bricks.empty() ? h->brick_pool : h->bricks;
  const size_t esz = dtype_size(h->dtype);
  int rc = BEAMS_OK;
  for (int b = 0; b < nbricks && rc == BEAMS_OK; ++b) {
    auto& br = bricks[(size_t)b];
    if ((size_t)b < prev.size()) {
      std::swap(br.field, prev[(size_t)b].field);
      std::swap(br.field_blk, prev[(size_t)b].field_blk);
    }
    const int64_t own[3] = { br.hi[0] - br.lo[0] + 1, br.hi[1] - br.lo[1] + 1, br.hi[2] - br.lo[2] + 1 };
    hipError_t e = br.field.ensure(esz * (size_t)(br.sdims[0] * br.sdims[1] * br.sdims[2]));
    if (e == hipSuccess) e = br.field_blk.ensure(esz * (size_t)blocked_count(own));
    hipError_t x = e != hipSuccess ? e
                                   : launch_extract_brick(h->field.p, (int32_t)esz, h->dims, br.slo, br.sdims,
                                                          br.field.p, h->stream);
    if (x == hipSuccess) x = launch_block_extract(h->field.p, (int32_t)esz, h->dims, br.lo, own, br.field_blk.p, h->stream);
    if (x != hipSuccess) rc = h->fail(BEAMS_ERR_HIP, std::string("bricks: ") + hipGetErrorString(x));
  }
  if (rc == BEAMS_OK && hipStreamSynchronize(h->stream) != hipSuccess)
    rc = h->fail(BEAMS_ERR_HIP, "bricks: extraction failed");
  for (auto& br : prev) br.release();
  prev.clear();
  if (rc != BEAMS_OK) {
    for (auto& br : bricks) br.release();
    h->bricks.clear();
    h->btab_host.clear();
    h->map_valid = false;
    return rc;
  }
  for (auto& br : h->brick_pool) br.release();
  h->brick_pool.clear();
  // the brick tables are rebuilt every frame and uploaded only when their bytes change
  h->bricks.swap(bricks);
  h->ghost = ghost;
  h->map_valid = false;
  return BEAMS_OK;
}

int beams_get_brick_map(beams_renderer* h, int32_t brick, int32_t which, float* out)
{
  CHECK_HANDLE(h);
  if (!out || brick < 0 || brick >= (int32_t)h->bricks.size() || (which != 0 && which != 1))
    return h->fail(BEAMS_ERR_INVALID, "bad brick index, map selector or output");
  if (!h->map_valid) return h->fail(BEAMS_ERR_STATE, "no brick maps: render with shadows first");
  const auto& br = h->bricks[(size_t)brick];
  const size_t msz = dtype_size(h->map_dtype);
  const unsigned char* src =
    (const unsigned char*)(which ? h->bmap_final.p : h->bmap_local.p) + msz * (size_t)br.seed_off;
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  if (h->map_dtype == BEAMS_F32) {
    HIP_TRY(h, hipMemcpy(out, src, sizeof(float) * (size_t)br.nverts, hipMemcpyDeviceToHost));
  } else {
    std::vector<uint16_t> tmp((size_t)br.nverts);
    HIP_TRY(h, hipMemcpy(tmp.data(), src, 2 * (size_t)br.nverts, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < tmp.size(); ++i) out[i] = f16_to_f32(tmp[i]);
  }
  return BEAMS_OK;
}

int beams_set_persistent_groups(beams_renderer* h, int32_t groups)
{
  CHECK_HANDLE(h);
  if (groups < 1 || groups > 65536) return h->fail(BEAMS_ERR_INVALID, "persistent groups must be in [1, 65536]");
  h->persistent_groups = groups;
  return BEAMS_OK;
}

int beams_set_schedule(beams_renderer* h, const beams_schedule* s)
{
  CHECK_HANDLE(h);
  if (!s) return h->fail(BEAMS_ERR_INVALID, "null schedule");
  if (s->persistent_groups < 1 || s->persistent_groups > 65536)
    return h->fail(BEAMS_ERR_INVALID, "persistent groups must be in [1, 65536]");
  if (s->march_groups_per_cu < 0 || s->march_groups_per_cu > 16)
    return h->fail(BEAMS_ERR_INVALID, "march groups per CU must be in [0, 16]");
  h->persistent_groups = s->persistent_groups;
  h->xcd_ranges = s->xcd_ranges != 0;
  h->light_rows = s->light_rows != 0;
  h->march_xcd_tiles = s->march_xcd_tiles != 0;
  h->map_fast_fetch = s->map_fast_fetch != 0;
  h->march_fast_fetch = s->march_fast_fetch != 0;
  h->map_pipeline = s->map_pipeline != 0;
  h->sweep_block = s->sweep_block != 0;
  h->march_groups_per_cu = s->march_groups_per_cu;
  h->march_blocked = s->march_blocked != 0;
  return BEAMS_OK;
}

int beams_get_schedule(const beams_renderer* h, beams_schedule* s)
{
  if (!h || !s) return BEAMS_ERR_INVALID;
  s->persistent_groups = h->persistent_groups;
  s->xcd_ranges = h->xcd_ranges;
  s->light_rows = h->light_rows;
  s->march_xcd_tiles = h->march_xcd_tiles;
  s->map_fast_fetch = h->map_fast_fetch;
  s->march_fast_fetch = h->march_fast_fetch;
  s->map_pipeline = h->map_pipeline;
  s->sweep_block = h->sweep_block;
  s->march_groups_per_cu = h->march_groups_per_cu;
  s->march_blocked = h->march_blocked;
  return BEAMS_OK;
}

int beams_set_composite_order(beams_renderer* h, int32_t order)
{
  CHECK_HANDLE(h);
  if (order != BEAMS_ORDER_CENTER_DISTANCE && order != BEAMS_ORDER_KD_FRONT_TO_BACK)
    return h->fail(BEAMS_ERR_INVALID, "unknown composite order");
  h->composite_order = order;
  return BEAMS_OK;
}

int beams_set_graph(beams_renderer* h, int32_t enabled)
{
  CHECK_HANDLE(h);
  h->use_graph = enabled != 0;
  if (!h->use_graph && h->graph_exec) {
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    (void)hipGraphExecDestroy(h->graph_exec);
    h->graph_exec = nullptr;
    h->graph_key.clear();
  }
  return BEAMS_OK;
}

int beams_autotune(beams_renderer* h, const beams_camera* cam, int32_t width, int32_t height,
                   int32_t frames, int32_t* best_groups, float* best_ms)
{
  CHECK_HANDLE(h);
  if (!cam || width < 1 || height < 1 || frames < 1 || frames > 1000)
    return h->fail(BEAMS_ERR_INVALID, "autotune: bad camera, size or frame count");
  const size_t npix = (size_t)width * (size_t)height;
  HIP_TRY(h, h->image.ensure(sizeof(float) * npix * 5));
  float* img = (float*)h->image.p;
  hipEvent_t e0, e1;
  HIP_TRY(h, hipEventCreate(&e0));
  HIP_TRY(h, hipEventCreate(&e1));
  int rc = BEAMS_OK;
  // mean frame time of schedule s (one untimed frame first), +inf on failure
  auto time_schedule = [&](const beams_schedule& s) -> float {
    if (rc) return INFINITY;
    beams_set_schedule(h, &s);
    if ((rc = beams_render_image(h, cam, width, height, nullptr, img, img + 4 * npix, BEAMS_MEM_DEVICE)))
      return INFINITY;
    (void)hipEventRecord(e0, h->stream);
    for (int32_t k = 0; k < frames && !rc; ++k)
      rc = beams_render_image(h, cam, width, height, nullptr, img, img + 4 * npix, BEAMS_MEM_DEVICE);
    if (rc) return INFINITY;
    (void)hipEventRecord(e1, h->stream);
    float ms = 0.0f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = h->fail(BEAMS_ERR_HIP, "autotune: event timing failed");
      return INFINITY;
    }
    return ms / (float)frames;
  };
  beams_schedule best{};
  beams_get_schedule(h, &best);
  float best_t = INFINITY;
  if (h->bricks.empty()) {
    best_t = time_schedule(best);  // one block: no schedule knob applies
  } else {
    // coordinate descent: the workgroup count, then each switch flipped in turn
    for (int32_t g : { 512, 1024, 2048, 4096, 8192 }) {
      beams_schedule s = best;
      s.persistent_groups = g;
      const float t = time_schedule(s);
      if (t < best_t) {
        best_t = t;
        best = s;
      }
    }
    // each switch flipped in turn, kept only for a gain above 0.5% (the timing's spread)
    for (int knob = 0; knob < 7 && !rc; ++knob) {
      beams_schedule s = best;
      int32_t* const v[7] = { &s.xcd_ranges, &s.light_rows, &s.march_xcd_tiles, &s.map_fast_fetch,
                              &s.march_fast_fetch, &s.map_pipeline, &s.march_blocked };
      *v[knob] = !*v[knob];
      const float t = time_schedule(s);
      if (t < best_t * 0.995f) {
        best_t = t;
        best = s;
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  beams_set_schedule(h, &best);
  if (rc) return rc;
  if (best_groups) *best_groups = best.persistent_groups;
  if (best_ms) *best_ms = best_t;
  return beams_reset_timings(h);
}

int beams_gen_kd_bricks(uint64_t seed, const int64_t dims[3], int32_t nbricks, int64_t min_cells,
                        int64_t max_cells, int64_t* point_ranges)
{
  if (!dims || !point_ranges || nbricks < 1 || min_cells < 1 || max_cells < 2 * min_cells)
    return BEAMS_ERR_INVALID;
  std::string err;
  return gen_kd_bricks(seed, dims, nbricks, min_cells, max_cells, point_ranges, err) ? BEAMS_ERR_INVALID
                                                                                      : BEAMS_OK;
}

int beams_set_scalar_range(beams_renderer* h, double min_value, double max_value)
{
  CHECK_HANDLE(h);
  if (!std::isfinite(min_value) || !std::isfinite(max_value) || max_value < min_value)
    return h->fail(BEAMS_ERR_INVALID, "scalar range must be finite with min <= max");
  h->smin = min_value;
  h->smax = max_value;
  h->has_range = true;
  return BEAMS_OK;
}

int beams_set_colormap(beams_renderer* h, const float* rgba, int64_t n)
{
  CHECK_HANDLE(h);
  if (!rgba || n < 1 || n > 8192) return h->fail(BEAMS_ERR_INVALID, "colormap needs 1..8192 RGBA entries");
  h->map_valid = false;  // the map's opacities come from the table's alpha
  h->htf.assign(rgba, rgba + 4 * n);
  HIP_TRY(h, h->tf.ensure(sizeof(float) * 4 * (size_t)n));
  HIP_TRY(h, hipMemcpyAsync(h->tf.p, h->htf.data(), sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice, h->stream));
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return BEAMS_OK;
}

int beams_set_light(beams_renderer* h, int32_t kind, const float xyz[3], const float color[3],
                    float intensity)
{
  CHECK_HANDLE(h);
  if (!xyz || !color) return h->fail(BEAMS_ERR_INVALID, "null light vector");
  if (kind != BEAMS_LIGHT_POINT && kind != BEAMS_LIGHT_DIRECTIONAL)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "unknown light kind");
  for (int d = 0; d < 3; ++d)
    if (!std::isfinite(xyz[d]) || !std::isfinite(color[d] * intensity))
      return h->fail(BEAMS_ERR_INVALID, "light position and colour must be finite");
  const float n2 = xyz[0] * xyz[0] + xyz[1] * xyz[1] + xyz[2] * xyz[2];
  if (kind == BEAMS_LIGHT_DIRECTIONAL && !(n2 > 0.0f))
    return h->fail(BEAMS_ERR_INVALID, "directional light needs a nonzero direction");
  const float rs = kind == BEAMS_LIGHT_DIRECTIONAL ? 1.0f / std::sqrt(n2) : 1.0f;
  for (int d = 0; d < 3; ++d) {
    h->light[d] = kind == BEAMS_LIGHT_DIRECTIONAL ? xyz[d] * rs : xyz[d];
    h->light_color[d] = color[d] * intensity;  // Lights::AddLight(radius, pos, Color * Intensity)
  }
  h->light_kind = kind;
  h->has_light = true;
  h->map_valid = false;  // the map belongs to the previous light
  return BEAMS_OK;
}

int beams_set_shadow(beams_renderer* h, int32_t enabled, const int64_t map_cells[3], int32_t mode,
                     int32_t map_dtype)
{
  CHECK_HANDLE(h);
  if (mode != BEAMS_SHADOW_REF_MARCH && mode != BEAMS_SHADOW_SWEEP && mode != BEAMS_SHADOW_LEARNED)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "unknown shadow mode");
  if (map_dtype != BEAMS_F32 && map_dtype != BEAMS_F16)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "map storage must be F32 or F16");
  if (map_cells) {
    for (int d = 0; d < 3; ++d)
      if (map_cells[d] < 1 || map_cells[d] > 4096) return h->fail(BEAMS_ERR_INVALID, "map cells must be in [1, 4096]");
    for (int d = 0; d < 3; ++d) h->map_cells[d] = map_cells[d];
  }
  h->use_shadow = enabled != 0;
  h->shadow_mode = mode;
  h->map_dtype = map_dtype;
  h->map_valid = false;
  return BEAMS_OK;
}

int beams_set_sample_distance(beams_renderer* h, float distance)
{
  CHECK_HANDLE(h);
  if (!(distance > 0.0f)) return h->fail(BEAMS_ERR_INVALID, "Sample distance must be positive.");
  h->sample_distance = distance;
  return BEAMS_OK;
}

int beams_set_alpha_stop(beams_renderer* h, float alpha)
{
  CHECK_HANDLE(h);
  if (!(alpha > 0.0f) || alpha > 1.0f) return h->fail(BEAMS_ERR_INVALID, "alpha stop must be in (0, 1]");
  h->alpha_stop = alpha;
  return BEAMS_OK;
}

int beams_render_rays(beams_renderer* h, int64_t n, const float* const origin[3],
                      const float* const dir[3], float* min_d, float* max_d, float* dist,
                      float* rgba, int32_t mem)
{
  CHECK_HANDLE(h);
  if (n < 0) return h->fail(BEAMS_ERR_INVALID, "negative ray count");
  if (!origin || !dir || !min_d || !max_d || !dist || !rgba)
    return h->fail(BEAMS_ERR_INVALID, "null ray array");
  for (int d = 0; d < 3; ++d)
    if (!origin[d] || !dir[d]) return h->fail(BEAMS_ERR_INVALID, "null ray array");
  if (mem != BEAMS_MEM_HOST && mem != BEAMS_MEM_DEVICE) return h->fail(BEAMS_ERR_INVALID, "bad mem kind");
  int rc = check_ready(h);
  if (rc) return rc;
  if ((rc = frame_begin(h))) return rc;
  GridParams g = grid_params(h);
  ShadeParams s = shade_params(h, g);
  if (h->use_shadow && h->assoc == 0) {
    if ((rc = phase_mark(h, BEAMS_PHASE_MAP, false))) return rc;
    if ((rc = build_map(h, g, s))) return rc;
    if ((rc = phase_mark(h, BEAMS_PHASE_MAP, true))) return rc;
  }
  RaySoA r{};
  r.n = n;
  if (mem == BEAMS_MEM_DEVICE) {
    for (int d = 0; d < 3; ++d) {
      r.org[d] = origin[d];
      r.dir[d] = dir[d];
    }
    r.min_d = min_d;
    r.max_d = max_d;
    r.dist = dist;
    r.rgba = rgba;
  } else {
    const size_t fb = sizeof(float) * (size_t)n;
    HIP_TRY(h, h->rays.ensure(fb * 13 + 64));
    float* base = (float*)h->rays.p;
    float* arr[9];
    for (int k = 0; k < 9; ++k) arr[k] = base + (size_t)k * (size_t)n;
    float* drgba = base + (size_t)9 * (size_t)n;
    for (int d = 0; d < 3; ++d) {
      HIP_TRY(h, hipMemcpyAsync(arr[d], origin[d], fb, hipMemcpyHostToDevice, h->stream));
      HIP_TRY(h, hipMemcpyAsync(arr[3 + d], dir[d], fb, hipMemcpyHostToDevice, h->stream));
      r.org[d] = arr[d];
      r.dir[d] = arr[3 + d];
    }
    HIP_TRY(h, hipMemcpyAsync(arr[6], min_d, fb, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipMemcpyAsync(arr[7], max_d, fb, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipMemcpyAsync(arr[8], dist, fb, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(h, hipMemcpyAsync(drgba, rgba, 4 * fb, hipMemcpyHostToDevice, h->stream));
    r.min_d = arr[6];
    r.max_d = arr[7];
    r.dist = arr[8];
    r.rgba = drgba;
  }
  if ((rc = phase_mark(h, BEAMS_PHASE_MARCH, false))) return rc;
  MapParams m = map_params(h, g, h->built_mode);
  HIP_TRY(h, launch_march_rays(g, m, s, r, h->stream));
  if ((rc = phase_mark(h, BEAMS_PHASE_MARCH, true))) return rc;
  if (mem == BEAMS_MEM_HOST) {
    const size_t fb = sizeof(float) * (size_t)n;
    HIP_TRY(h, hipMemcpyAsync(min_d, r.min_d, fb, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipMemcpyAsync(max_d, r.max_d, fb, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipMemcpyAsync(dist, r.dist, fb, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipMemcpyAsync(rgba, r.rgba, 4 * fb, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
  }
  return BEAMS_OK;
}

int beams_render_rays_noalloc(beams_renderer* h, int64_t n, const float* const origin[3],
                              const float* const dir[3], float* min_d, float* max_d, float* dist,
                              float* rgba)
{
  CHECK_HANDLE(h);
  if (n < 0) return h->fail(BEAMS_ERR_INVALID, "negative ray count");
  if (!origin || !dir || !min_d || !max_d || !dist || !rgba)
    return h->fail(BEAMS_ERR_INVALID, "null ray array");
  for (int d = 0; d < 3; ++d)
    if (!origin[d] || !dir[d]) return h->fail(BEAMS_ERR_INVALID, "null ray array");
  int rc = check_ready(h);
  if (rc) return rc;
  if (!h->bricks.empty())
    return h->fail(BEAMS_ERR_UNSUPPORTED, "the no-allocation render takes a single block, not bricks");
  const bool with_map = h->use_shadow && h->assoc == 0;
  if (with_map && h->shadow_mode != BEAMS_SHADOW_REF_MARCH)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "the no-allocation render builds its map with the reference march");
  if (!h->scratch) return h->fail(BEAMS_ERR_STATE, "no scratch array on this device");
  GridParams g = grid_params(h);
  ShadeParams s = shade_params(h, g);
  s.counters = nullptr;  // the sample counters live in an allocated buffer
  MapParams m = map_params(h, g, BEAMS_SHADOW_REF_MARCH);
  if (with_map) {
    const uint64_t bytes = (uint64_t)m.mp[0] * (uint64_t)m.mp[1] * (uint64_t)m.mp[2] * dtype_size(h->map_dtype);
    if (bytes > kScratchBytes)
      return h->fail(BEAMS_ERR_INVALID, "map of " + std::to_string(bytes) + " B exceeds the " +
                                            std::to_string(kScratchBytes >> 20) +
                                            " MiB scratch of the no-allocation render");
    m.map = h->scratch;
    HIP_TRY(h, launch_tmap_ref(g, m, s, nullptr, h->stream));
  }
  RaySoA r{};
  r.n = n;
  for (int d = 0; d < 3; ++d) {
    r.org[d] = origin[d];
    r.dir[d] = dir[d];
  }
  r.min_d = min_d;
  r.max_d = max_d;
  r.dist = dist;
  r.rgba = rgba;
  HIP_TRY(h, launch_march_rays(g, m, s, r, h->stream));
  return BEAMS_OK;
}

int beams_render_image(beams_renderer* h, const beams_camera* cam, int32_t width, int32_t height,
                       const float* background, float* rgba, float* depth, int32_t mem)
{
  CHECK_HANDLE(h);
  if (!cam || width < 1 || height < 1 || width > 32768 || height > 32768)
    return h->fail(BEAMS_ERR_INVALID, "bad camera or image size");
  if (mem != BEAMS_MEM_HOST && mem != BEAMS_MEM_DEVICE) return h->fail(BEAMS_ERR_INVALID, "bad mem kind");
  if (!rgba) return h->fail(BEAMS_ERR_INVALID, "null rgba output");
  int rc = check_ready(h);
  if (rc) return rc;
  CameraParams c{};
  if (!camera_params(cam, width, height, c))
    return h->fail(BEAMS_ERR_INVALID, "degenerate camera (eye = look_at or up parallel to the view)");
  if (!h->bricks.empty()) return render_bricks(h, c, background, rgba, depth, mem);
  if ((rc = frame_begin(h))) return rc;
  GridParams g = grid_params(h);
  ShadeParams s = shade_params(h, g);
  // The map build (learned, with its check and gated fallback, or exact) is part of the frame's
  // launch sequence and so of its graph.
  const bool with_map = h->use_shadow && h->assoc == 0;
  const bool learned = with_map && h->shadow_mode == BEAMS_SHADOW_LEARNED;
  if (learned && (rc = prepare_learned(h, g, s))) return rc;
  if (with_map && !learned && (rc = prepare_map(h, h->shadow_mode))) return rc;
  const size_t npix = (size_t)width * (size_t)height;
  ImageOut o{};
  if (mem == BEAMS_MEM_DEVICE) {
    o.rgba = rgba;
    o.depth = depth;
  } else {
    HIP_TRY(h, h->image.ensure(sizeof(float) * npix * 5));
    o.rgba = (float*)h->image.p;
    o.depth = depth ? (float*)h->image.p + 4 * npix : nullptr;
  }
  o.use_bg = background ? 1 : 0;
  for (int k = 0; k < 4; ++k) o.bg[k] = background ? background[k] : 0.0f;
  const MapParams m = map_params(h, g, h->shadow_mode);
  auto issue = [&](bool timed) -> int {
    int r;
    auto mark = [&](int p, bool end) { return timed ? phase_mark(h, p, end) : BEAMS_OK; };
    if (with_map) {
      if ((r = mark(BEAMS_PHASE_MAP, false))) return r;
      if ((r = learned ? enqueue_learned(h, g, s) : enqueue_map(h, g, m, s, h->shadow_mode))) return r;
      if ((r = mark(BEAMS_PHASE_MAP, true))) return r;
    }
    if ((r = mark(BEAMS_PHASE_MARCH, false))) return r;
    HIP_TRY(h, launch_march_image(g, m, s, c, o, h->stream));
    return mark(BEAMS_PHASE_MARCH, true);
  };
  KeyBuilder key;
  key.put(0).put(g).put(m).put(s).put(c).put(o).put(with_map).put(h->sweep_block);
  if (learned)
    key.put(h->fallback_mode).put(h->error_budget).put(h->check_stride).put(h->net ? mlp_device_params(h->net) : nullptr)
       .put(h->sb_lb).put(h->sb_f).put(h->tfidx.p);
  if ((rc = run_frame(h, key.bytes, issue))) return rc;
  if (with_map) {
    h->map_valid = true;
    h->built_mode = h->shadow_mode;
    if (learned) learned_built(h, 1);
  }
  if (mem == BEAMS_MEM_HOST) {
    HIP_TRY(h, hipMemcpyAsync(rgba, o.rgba, sizeof(float) * 4 * npix, hipMemcpyDeviceToHost, h->stream));
    if (depth) HIP_TRY(h, hipMemcpyAsync(depth, o.depth, sizeof(float) * npix, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
  }
  return BEAMS_OK;
}

int beams_get_map(beams_renderer* h, float* out, int32_t mem)
{
  CHECK_HANDLE(h);
  if (!out) return h->fail(BEAMS_ERR_INVALID, "null output");
  if (!h->map_valid) return h->fail(BEAMS_ERR_STATE, "no map: render with shadows first");
  const int64_t nv = (h->map_cells[0] + 1) * (h->map_cells[1] + 1) * (h->map_cells[2] + 1);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const MapParams m = map_params(h, grid_params(h), h->built_mode);
  const bool xfast = m.ms[0] == 1 && m.ms[1] == m.mp[0];
  if (h->map_dtype == BEAMS_F32 && xfast) {
    HIP_TRY(h, hipMemcpy(out, h->map.p, sizeof(float) * (size_t)nv, copy_kind(mem, false)));
    return BEAMS_OK;
  }
  // FP16 storage and/or the sweep's slice layout: decode and reorder to x fastest
  std::vector<float> raw((size_t)nv), f((size_t)nv);
  if (h->map_dtype == BEAMS_F32) {
    HIP_TRY(h, hipMemcpy(raw.data(), h->map.p, sizeof(float) * (size_t)nv, hipMemcpyDeviceToHost));
  } else {
    std::vector<uint16_t> tmp((size_t)nv);
    HIP_TRY(h, hipMemcpy(tmp.data(), h->map.p, 2 * (size_t)nv, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < raw.size(); ++i) raw[i] = f16_to_f32(tmp[i]);
  }
  size_t o = 0;
  for (int64_t k = 0; k < m.mp[2]; ++k)
    for (int64_t j = 0; j < m.mp[1]; ++j)
      for (int64_t i = 0; i < m.mp[0]; ++i) f[o++] = raw[(size_t)(i * m.ms[0] + j * m.ms[1] + k * m.ms[2])];
  HIP_TRY(h, hipMemcpy(out, f.data(), sizeof(float) * (size_t)nv,
                       mem == BEAMS_MEM_DEVICE ? hipMemcpyHostToDevice : hipMemcpyHostToHost));
  return BEAMS_OK;
}

int beams_set_learned(beams_renderer* h, const beams_mlp* net, int32_t fallback_mode,
                      float error_budget, int32_t check_stride)
{
  CHECK_HANDLE(h);
  if (fallback_mode != BEAMS_SHADOW_REF_MARCH && fallback_mode != BEAMS_SHADOW_SWEEP)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "the fallback must be an exact mode");
  if (error_budget != error_budget) return h->fail(BEAMS_ERR_INVALID, "error budget is NaN");
  if (check_stride < 1) return h->fail(BEAMS_ERR_INVALID, "check stride must be >= 1");
  h->net = net;
  h->fallback_mode = fallback_mode;
  h->error_budget = error_budget;
  h->check_stride = check_stride;
  return BEAMS_OK;
}

int beams_learned_samples(beams_renderer* h, int32_t mode, int64_t n, uint64_t seed,
                          float* features, float* targets, int32_t mem)
{
  CHECK_HANDLE(h);
  if (mode != BEAMS_SHADOW_REF_MARCH && mode != BEAMS_SHADOW_SWEEP)
    return h->fail(BEAMS_ERR_UNSUPPORTED, "training maps come from an exact mode");
  if (n < 0 || !features || !targets) return h->fail(BEAMS_ERR_INVALID, "bad sample output");
  if (mem != BEAMS_MEM_HOST && mem != BEAMS_MEM_DEVICE) return h->fail(BEAMS_ERR_INVALID, "bad mem kind");
  int rc = check_ready(h);
  if (rc) return rc;
  if (h->assoc != BEAMS_ASSOC_POINTS) return h->fail(BEAMS_ERR_UNSUPPORTED, "learned maps need a point field");
  if (!h->has_light) return h->fail(BEAMS_ERR_STATE, "samples need a light: call beams_set_light");
  if (!h->bricks.empty()) return h->fail(BEAMS_ERR_UNSUPPORTED, "samples come from the single-block map");
  GridParams g = grid_params(h);
  ShadeParams s = shade_params(h, g);
  s.counters = nullptr;
  if ((rc = prepare_map(h, mode))) return rc;
  if ((rc = prepare_map(h, BEAMS_SHADOW_LEARNED))) return rc;
  const MapParams em = map_params(h, g, mode);
  if ((rc = enqueue_map(h, g, em, s, mode))) return rc;
  h->built_mode = mode;
  h->map_valid = true;
  const LearnedParams p = learned_params(h, g, s, map_params(h, g, BEAMS_SHADOW_LEARNED));
  HIP_TRY(h, launch_coarse(p, h->stream));
  float *dx = features, *dt = targets;
  if (mem == BEAMS_MEM_HOST) {
    HIP_TRY(h, h->rays.ensure(sizeof(float) * (size_t)n * (BEAMS_MLP_FEATURES + 1) + 64));
    dx = (float*)h->rays.p;
    dt = dx + (size_t)n * BEAMS_MLP_FEATURES;
  }
  HIP_TRY(h, launch_learned_samples(p, em, n, seed, dx, dt, h->stream));
  if (mem == BEAMS_MEM_HOST) {
    HIP_TRY(h, hipMemcpyAsync(features, dx, sizeof(float) * (size_t)n * BEAMS_MLP_FEATURES, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(h, hipMemcpyAsync(targets, dt, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, h->stream));
  }
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  return BEAMS_OK;
}

int beams_get_learned_stats(beams_renderer* h, float err[2], int32_t* fell_back)
{
  CHECK_HANDLE(h);
  float e[2] = { -1.0f, -1.0f };
  int32_t fb = 0;
  if (h->learned_kind) {
    uint32_t w[2] = { 0, 0 };
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    HIP_TRY(h, hipMemcpy(w, h->learned_kind == 1 ? h->lerr.p : h->lb_list.p, sizeof(w), hipMemcpyDeviceToHost));
    if (h->learned_kind == 1) {
      e[0] = bits_f32(w[0]);
      e[1] = bits_f32(w[1]);
      fb = !(e[0] <= h->built_budget);
      if (h->sb_tiles) {
        uint32_t t = 0;
        HIP_TRY(h, hipMemcpy(&t, h->sb_list.p, sizeof(t), hipMemcpyDeviceToHost));
        fb = (int32_t)t;
      }
    } else {
      e[0] = bits_f32(w[1]);
      fb = (int32_t)w[0];
    }
  }
  if (err) {
    err[0] = e[0];
    err[1] = e[1];
  }
  if (fell_back) *fell_back = fb;
  return BEAMS_OK;
}

int beams_get_timings(beams_renderer* h, float* phase_ms, int32_t n)
{
  CHECK_HANDLE(h);
  if (!phase_ms || n < 1) return h->fail(BEAMS_ERR_INVALID, "null timings output");
  if (h->cur < 0) return h->fail(BEAMS_ERR_STATE, "nothing rendered yet");
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const auto& f = h->frames[h->cur];
  for (int p = 0; p < n && p < BEAMS_NUM_PHASES; ++p) {
    float t = 0.0f;
    if (f.timed[p]) HIP_TRY(h, hipEventElapsedTime(&t, f.ev[2 * p], f.ev[2 * p + 1]));
    phase_ms[p] = t;
  }
  return BEAMS_OK;
}

int beams_get_timing_stats(beams_renderer* h, float* avg_ms, int32_t n, int32_t* nframes)
{
  CHECK_HANDLE(h);
  if (!avg_ms || n < 1) return h->fail(BEAMS_ERR_INVALID, "null timings output");
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  const int64_t used = std::min<int64_t>(h->total_frames, beams_renderer::kMaxFrames);
  for (int p = 0; p < n && p < BEAMS_NUM_PHASES; ++p) {
    double sum = 0.0;
    int cnt = 0;
    for (int64_t i = 0; i < used; ++i) {
      const auto& f = h->frames[(size_t)i];
      if (!f.timed[p]) continue;
      float t = 0.0f;
      HIP_TRY(h, hipEventElapsedTime(&t, f.ev[2 * p], f.ev[2 * p + 1]));
      sum += t;
      ++cnt;
    }
    avg_ms[p] = cnt ? (float)(sum / cnt) : 0.0f;
  }
  if (nframes) *nframes = (int32_t)used;
  return BEAMS_OK;
}

int beams_reset_timings(beams_renderer* h)
{
  CHECK_HANDLE(h);
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  h->total_frames = 0;
  h->cur = -1;
  return BEAMS_OK;
}

int beams_set_sample_counting(beams_renderer* h, int32_t enabled)
{
  CHECK_HANDLE(h);
  h->counting = enabled != 0;
  if (h->counting) HIP_TRY(h, h->counters.ensure(2 * sizeof(unsigned long long)));
  return BEAMS_OK;
}

int beams_get_sample_counts(beams_renderer* h, int64_t* camera_samples, int64_t* light_steps)
{
  CHECK_HANDLE(h);
  if (!h->counting) return h->fail(BEAMS_ERR_STATE, "sample counting is off");
  unsigned long long c[2] = { 0, 0 };
  HIP_TRY(h, hipStreamSynchronize(h->stream));
  HIP_TRY(h, hipMemcpy(c, h->counters.p, sizeof(c), hipMemcpyDeviceToHost));
  if (camera_samples) *camera_samples = (int64_t)c[0];
  if (light_steps) *light_steps = (int64_t)c[1];
  return BEAMS_OK;
}

int beams_gen_field(const beams_field_recipe* r, const int64_t dims[3], int32_t dtype, void* out,
                    int32_t mem)
{
  if (!r || !dims || !out) return BEAMS_ERR_INVALID;
  if (r->nblobs < 0 || r->nblobs > 4096 || r->noise_octaves < 0 || r->noise_octaves > 16)
    return BEAMS_ERR_INVALID;
  for (int d = 0; d < 3; ++d)
    if (dims[d] < 2) return BEAMS_ERR_INVALID;
  if (dtype_size(dtype) == 0) return BEAMS_ERR_UNSUPPORTED;
  FieldRecipe fr{ r->seed, r->nblobs, r->noise_octaves, r->noise_amp, r->noise_freq, r->time,
                  r->drift_speed, r->noise_phase_rate };
  std::string err;
  if (mem == BEAMS_MEM_DEVICE) {
    hipStream_t st;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return BEAMS_ERR_HIP;
    int rc = gen_field(fr, dims, dtype, out, true, st, err);
    (void)hipStreamDestroy(st);
    return rc;
  }
  return gen_field(fr, dims, dtype, out, false, nullptr, err);
}

int beams_gen_colormap(uint64_t seed, int32_t npoints, int64_t n, float* rgba32, double* rgba64)
{
  if (npoints < 2 || n < 1) return BEAMS_ERR_INVALID;
  gen_colormap(seed, npoints, n, rgba32, rgba64);
  return BEAMS_OK;
}

int beams_gen_orbit_camera(const double bounds[6], double azimuth_deg, double elevation_deg,
                           double fov_y_deg, beams_camera* out)
{
  if (!bounds || !out) return BEAMS_ERR_INVALID;
  double c[3], diag2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    c[d] = 0.5 * (bounds[2 * d] + bounds[2 * d + 1]);
    const double e = bounds[2 * d + 1] - bounds[2 * d];
    diag2 += e * e;
  }
  const double dist = std::sqrt(diag2);
  const double az = azimuth_deg * M_PI / 180.0, el = elevation_deg * M_PI / 180.0;
  const double dir[3] = { std::cos(el) * std::sin(az), std::sin(el), std::cos(el) * std::cos(az) };
  for (int d = 0; d < 3; ++d) {
    out->eye[d] = c[d] + dist * dir[d];
    out->look_at[d] = c[d];
  }
  out->up[0] = 0.0;
  out->up[1] = 1.0;
  out->up[2] = 0.0;
  out->fov_y_deg = fov_y_deg;
  out->near_dist = 0.0;
  out->far_dist = 0.0;
  return BEAMS_OK;
}

int beams_gen_camera_rays(const beams_camera* cam, int32_t width, int32_t height,
                          float* const origin[3], float* const dir[3], double* const origin64[3],
                          double* const dir64[3])
{
  if (!cam || width < 1 || height < 1 || !origin || !dir) return BEAMS_ERR_INVALID;
  CameraParams c{};
  if (!camera_params(cam, width, height, c)) return BEAMS_ERR_INVALID;
  for (int32_t py = 0; py < height; ++py)
    for (int32_t px = 0; px < width; ++px) {
      const size_t i = (size_t)py * (size_t)width + (size_t)px;
      float o[3], dd[3];
      double o64[3], d64[3];
      camera_ray(c, px, py, o, dd, o64, d64);
      for (int d = 0; d < 3; ++d) {
        origin[d][i] = o[d];
        dir[d][i] = dd[d];
        if (origin64 && origin64[d]) origin64[d][i] = o64[d];
        if (dir64 && dir64[d]) dir64[d][i] = d64[d];
      }
    }
  return BEAMS_OK;
}

}  // extern "C"
