// Seeded synthetic inputs (the DataSetFactory side): blob + value-noise fields on
// the host or the device, transfer-function tables and orbit cameras.
#include "field_gen.hpp"
#include "gen.hpp"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace beams {

void blob_tables_host(const FieldRecipe& r, const int64_t dims[3], std::vector<double>& amp,
                      std::vector<double> axis[3], NoiseParams& np)
{
  SplitMix64 rng(r.seed);
  const int32_t k = r.nblobs;
  amp.assign((size_t)k, 0.0);
  std::vector<double> c((size_t)k * 3), sig((size_t)k), vel((size_t)k * 3);
  for (int32_t b = 0; b < k; ++b) {
    c[3 * b + 0] = rng.uniform();
    c[3 * b + 1] = rng.uniform();
    c[3 * b + 2] = rng.uniform();
    sig[b] = 0.04 + 0.11 * rng.uniform();
    amp[b] = 0.5 + 0.5 * rng.uniform();
    // drift direction, uniform on the sphere
    const double vz = 2.0 * rng.uniform() - 1.0;
    const double ph = 2.0 * M_PI * rng.uniform();
    const double rr = std::sqrt(std::max(0.0, 1.0 - vz * vz));
    vel[3 * b + 0] = rr * std::cos(ph) * r.drift_speed;
    vel[3 * b + 1] = rr * std::sin(ph) * r.drift_speed;
    vel[3 * b + 2] = vz * r.drift_speed;
  }
  for (int d = 0; d < 3; ++d) {
    const int64_t n = dims[d];
    axis[d].assign((size_t)(k * n), 0.0);
    for (int32_t b = 0; b < k; ++b) {
      const double cc = c[3 * b + d] + vel[3 * b + d] * r.time;
      const double inv2s2 = 1.0 / (2.0 * sig[b] * sig[b]);
      for (int64_t i = 0; i < n; ++i) {
        const double x = (double)i / (double)(n - 1);
        const double dx = x - cc;
        axis[d][(size_t)(b * n + i)] = std::exp(-dx * dx * inv2s2);
      }
    }
  }
  np.octaves = r.noise_octaves;
  np.amp = r.noise_amp;
  np.freq = r.noise_freq;
  np.phase = r.time * r.noise_phase_rate;
  np.salt = SplitMix64(r.seed ^ 0xa0761d6478bd642full).next();
}

namespace {

template <typename F>
void parallel_for(int64_t n, F&& f)
{
  const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const int64_t lo = t * chunk, hi = std::min(n, lo + chunk);
    if (lo >= hi) break;
    th.emplace_back([=, &f] { f(t, lo, hi); });
  }
  for (auto& x : th) x.join();
}

template <typename Store>
void store_all_host(const BlobTables& t, const NoiseParams& np, double mn, double mx,
                    Store&& st)
{
  const int64_t n0 = t.n[0], n1 = t.n[1], total = t.n[0] * t.n[1] * t.n[2];
  const double span = mx - mn;
  parallel_for(total, [&](int, int64_t lo, int64_t hi) {
    for (int64_t id = lo; id < hi; ++id) {
      const double v = field_raw(t, np, id % n0, (id / n0) % n1, id / (n0 * n1));
      st(id, (v - mn) / span);
    }
  });
}

// ---- device generation ------------------------------------------------------------------
struct DevTables {
  BlobTables t;
  NoiseParams np;
};

__global__ void __launch_bounds__(256) k_field_minmax(DevTables dt, double* out, int64_t total)
{
  __shared__ double smin[256], smax[256];
  double mn = __builtin_inf(), mx = -__builtin_inf();
  const int64_t n0 = dt.t.n[0], n1 = dt.t.n[1];
  for (int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * 256) {
    const double v = field_raw(dt.t, dt.np, id % n0, (id / n0) % n1, id / (n0 * n1));
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  smin[threadIdx.x] = mn;
  smax[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = smin[0];
    out[2 * blockIdx.x + 1] = smax[0];
  }
}

__global__ void __launch_bounds__(256) k_field_store(DevTables dt, double mn, double mx,
                                                     int32_t dtype, void* out, int64_t total)
{
  const int64_t n0 = dt.t.n[0], n1 = dt.t.n[1];
  const double span = mx - mn;
  for (int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x; id < total;
       id += (int64_t)gridDim.x * 256) {
    const double v =
      (field_raw(dt.t, dt.np, id % n0, (id / n0) % n1, id / (n0 * n1)) - mn) / span;
    switch (dtype) {
      case kF64: ((double*)out)[id] = v; break;
      case kF32: ((float*)out)[id] = (float)v; break;
      case kF16: ((uint16_t*)out)[id] = f32_to_f16((float)v); break;
      default: ((uint16_t*)out)[id] = f32_to_bf16((float)v); break;
    }
  }
}

}  // namespace

int gen_field(const FieldRecipe& r, const int64_t dims[3], int32_t dtype, void* out, bool device,
              hipStream_t st, std::string& err)
{
  std::vector<double> amp, axis[3];
  NoiseParams np{};
  blob_tables_host(r, dims, amp, axis, np);
  const int64_t total = dims[0] * dims[1] * dims[2];
  if (!device) {
    BlobTables t{ r.nblobs, { dims[0], dims[1], dims[2] }, amp.data(),
                  { axis[0].data(), axis[1].data(), axis[2].data() } };
    std::vector<double> pmn(16, INFINITY), pmx(16, -INFINITY);
    parallel_for(total, [&](int tid, int64_t lo, int64_t hi) {
      double mn = INFINITY, mx = -INFINITY;
      for (int64_t id = lo; id < hi; ++id) {
        const double v = field_raw(t, np, id % dims[0], (id / dims[0]) % dims[1],
                                   id / (dims[0] * dims[1]));
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
      }
      pmn[tid] = mn;
      pmx[tid] = mx;
    });
    const double mn = *std::min_element(pmn.begin(), pmn.end());
    const double mx = *std::max_element(pmx.begin(), pmx.end());
    switch (dtype) {
      case kF64: store_all_host(t, np, mn, mx, [&](int64_t i, double v) { ((double*)out)[i] = v; }); break;
      case kF32: store_all_host(t, np, mn, mx, [&](int64_t i, double v) { ((float*)out)[i] = (float)v; }); break;
      case kF16:
        store_all_host(t, np, mn, mx, [&](int64_t i, double v) { ((uint16_t*)out)[i] = f32_to_f16((float)v); });
        break;
      case kBF16:
        store_all_host(t, np, mn, mx, [&](int64_t i, double v) { ((uint16_t*)out)[i] = f32_to_bf16((float)v); });
        break;
      default: err = "unsupported dtype"; return 2;
    }
    return 0;
  }
  // device: upload the tables, reduce min/max per workgroup, finish on the host
  const int32_t k = r.nblobs;
  const size_t tab_elems = (size_t)k * (size_t)(1 + dims[0] + dims[1] + dims[2]);
  double* dtab = nullptr;
  double* dmm = nullptr;
  const int nb = 2048;
  hipError_t e = hipMalloc(&dtab, sizeof(double) * std::max<size_t>(tab_elems, 1));
  if (e == hipSuccess) e = hipMalloc(&dmm, sizeof(double) * 2 * nb);
  std::vector<double> h((size_t)2 * nb);
  if (e == hipSuccess) {
    std::vector<double> flat;
    flat.reserve(tab_elems);
    flat.insert(flat.end(), amp.begin(), amp.end());
    for (int d = 0; d < 3; ++d) flat.insert(flat.end(), axis[d].begin(), axis[d].end());
    if (!flat.empty())
      e = hipMemcpyAsync(dtab, flat.data(), sizeof(double) * flat.size(), hipMemcpyHostToDevice, st);
    DevTables dt{};
    dt.t.nblobs = k;
    for (int d = 0; d < 3; ++d) dt.t.n[d] = dims[d];
    dt.t.amp = dtab;
    dt.t.axis[0] = dtab + k;
    dt.t.axis[1] = dt.t.axis[0] + k * dims[0];
    dt.t.axis[2] = dt.t.axis[1] + k * dims[1];
    dt.np = np;
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_field_minmax, dim3(nb), dim3(256), 0, st, dt, dmm, total);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dmm, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) {
      double mn = INFINITY, mx = -INFINITY;
      for (int b = 0; b < nb; ++b) {
        mn = std::min(mn, h[2 * b]);
        mx = std::max(mx, h[2 * b + 1]);
      }
      hipLaunchKernelGGL(k_field_store, dim3(8192), dim3(256), 0, st, dt, mn, mx, dtype, out, total);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  if (dtab) (void)hipFree(dtab);
  if (dmm) (void)hipFree(dmm);
  if (e != hipSuccess) {
    err = std::string("field generation: ") + hipGetErrorString(e);
    return 4;
  }
  return 0;
}

int gen_kd_bricks(uint64_t seed, const int64_t dims[3], int32_t nbricks, int64_t min_cells,
                  int64_t max_cells, int64_t* ranges, std::string& err)
{
  struct Box {
    int64_t lo[3], hi[3];  // point index ranges, inclusive; cells = hi - lo
  };
  SplitMix64 rng(seed);
  auto pick = [&](int64_t a, int64_t b) {  // uniform integer in [a, b]
    return a + (int64_t)(rng.uniform() * (double)(b - a + 1));
  };
  auto cells = [](const Box& b, int d) { return b.hi[d] - b.lo[d]; };
  // 1) cut each axis into the fewest slabs no longer than max_cells, at jittered even
  //    positions, and take their product
  std::vector<int64_t> cuts[3];
  for (int d = 0; d < 3; ++d) {
    const int64_t e = dims[d] - 1;
    const int64_t p = (e + max_cells - 1) / max_cells;
    const int64_t base = e / p;
    if (base < min_cells) {
      err = "grid too small for the minimum brick extent";
      return 1;
    }
    const int64_t jit = std::max<int64_t>(0, std::min(base - min_cells, max_cells - base - 1) / 2);
    cuts[d].push_back(0);
    for (int64_t i = 1; i < p; ++i) cuts[d].push_back(i * e / p + pick(-jit, jit));
    cuts[d].push_back(e);
  }
  std::vector<Box> boxes;
  for (size_t k = 0; k + 1 < cuts[2].size(); ++k)
    for (size_t j = 0; j + 1 < cuts[1].size(); ++j)
      for (size_t i = 0; i + 1 < cuts[0].size(); ++i)
        boxes.push_back(Box{ { cuts[0][i], cuts[1][j], cuts[2][k] },
                             { cuts[0][i + 1], cuts[1][j + 1], cuts[2][k + 1] } });
  auto split = [&](size_t i, int d, int64_t at) {
    Box a = boxes[i], b = boxes[i];
    a.hi[d] = a.lo[d] + at;
    b.lo[d] = a.hi[d];  // neighbours share the boundary plane of points
    boxes[i] = a;
    boxes.push_back(b);
  };
  if ((int32_t)boxes.size() > nbricks) {
    err = "too many bricks needed to respect the maximum extent";
    return 1;
  }
  // 2) split the largest-volume splittable box along its longest axis until nbricks
  while ((int32_t)boxes.size() < nbricks) {
    size_t best = boxes.size();
    double bv = -1.0;
    for (size_t i = 0; i < boxes.size(); ++i) {
      int d = 0;
      for (int k = 1; k < 3; ++k)
        if (cells(boxes[i], k) > cells(boxes[i], d)) d = k;
      if (cells(boxes[i], d) < 2 * min_cells) continue;
      const double v = (double)cells(boxes[i], 0) * cells(boxes[i], 1) * cells(boxes[i], 2);
      if (v > bv) {
        bv = v;
        best = i;
      }
    }
    if (best == boxes.size()) {
      err = "no box left to split at the minimum extent";
      return 1;
    }
    int d = 0;
    for (int k = 1; k < 3; ++k)
      if (cells(boxes[best], k) > cells(boxes[best], d)) d = k;
    const int64_t e = cells(boxes[best], d);
    split(best, d, pick(min_cells, e - min_cells));
  }
  for (size_t i = 0; i < boxes.size(); ++i)
    for (int d = 0; d < 3; ++d) {
      ranges[6 * i + 2 * d] = boxes[i].lo[d];
      ranges[6 * i + 2 * d + 1] = boxes[i].hi[d];
    }
  return 0;
}

void gen_colormap(uint64_t seed, int32_t npoints, int64_t n, float* rgba32, double* rgba64)
{
  SplitMix64 rng(seed);
  std::vector<double> xs((size_t)npoints);
  xs[0] = 0.0;
  xs[(size_t)npoints - 1] = 1.0;
  for (int32_t i = 1; i + 1 < npoints; ++i) xs[(size_t)i] = rng.uniform();
  std::sort(xs.begin(), xs.end());
  std::vector<double> cp((size_t)npoints * 4);
  for (auto& v : cp) v = rng.uniform();
  for (int64_t e = 0; e < n; ++e) {
    const double x = (n > 1) ? (double)e / (double)(n - 1) : 0.0;
    int32_t seg = 0;
    while (seg + 2 < npoints && x > xs[(size_t)seg + 1]) ++seg;
    const double x0 = xs[(size_t)seg], x1 = xs[(size_t)seg + 1];
    const double t = (x1 > x0) ? std::min(1.0, std::max(0.0, (x - x0) / (x1 - x0))) : 0.0;
    for (int c = 0; c < 4; ++c) {
      const double a = cp[(size_t)seg * 4 + c], b = cp[(size_t)(seg + 1) * 4 + c];
      const double v = a + t * (b - a);
      if (rgba64) rgba64[4 * e + c] = v;
      if (rgba32) rgba32[4 * e + c] = (float)v;
    }
  }
}

}  // namespace beams
