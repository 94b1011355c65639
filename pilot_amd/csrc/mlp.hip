// Training and inference kernels of the learned shadow closure's network (mlp.hpp) and
// the beams_mlp C-ABI object that owns its parameters and optimiser state.
#include "../../include/beams_mi355x.h"
#include "learned.hpp"
#include "mlp.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace beams {

namespace {

struct TrainLds {
  MlpLds w;
  uint16_t w2[kHid][kLd];    // W2[k][n], k-major: B operand of dH1 = dH2 W2^T
  uint16_t buf[kHid * kLd];  // forward H1, then dH1^T [k][row]
  uint16_t h1t[kHid * kLd];  // H1^T [k][row]: A operand of dW2
  uint16_t dh2[kHid * kLd];  // dH2 [row][n]: A operand of dH1
  uint16_t dh2t[kHid * kLd]; // dH2^T [n][row]: B operand of dW2
  float y[64];
  float dy[64];
};

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Gradients of the mean-squared loss over rows [0, n) of (x, t) added into grads (FP32);
// the summed squared error is added into loss[0].  One wave per 64-row chunk.
__global__ void __launch_bounds__(64) k_mlp_grad(const float* params, const float* x,
                                                 const float* t, int64_t n, float inv_n,
                                                 float* grads, float* loss)
{
  __shared__ TrainLds s;
  mlp_stage(params, s.w);
  for (int i = threadIdx.x; i < kHid * kHid; i += 64) s.w2[i / kHid][i % kHid] = f32_to_bf16(params[kOffW2 + i]);
  __syncthreads();
  const int lane = threadIdx.x, g = lane >> 4, r = lane & 15;
  for (int64_t chunk = blockIdx.x; chunk * 64 < n; chunk += gridDim.x) {
    const int64_t row = chunk * 64 + lane;
    const bool valid = row < n;
    float feat[kFeat];
#pragma unroll
    for (int k = 0; k < kFeat; ++k) feat[k] = valid ? x[row * kFeat + k] : 0.0f;
    const float target = valid ? t[row] : 0.0f;
    f32x4 h1pre[4][4], h2pre[4][4];
    const float y = mlp_forward_wave(s.w, s.buf, s.y, feat, h1pre, h2pre);
    const float err = valid ? (y - target) : 0.0f;
    const float lsum = wave_sum(err * err);
    if (lane == 0) atomicAdd(loss, lsum);
    s.dy[lane] = 2.0f * err * inv_n;
    // H1^T for dW2
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          s.h1t[(16 * ct + r) * kLd + 16 * rt + 4 * g + i] = f32_to_bf16(fmaxf(h1pre[rt][ct][i], 0.0f));
    wave_lds_sync();
    // layer 3 and the ReLU of layer 2: dW3, db2, dH2
    float dw3[4] = { 0, 0, 0, 0 }, db2[4] = { 0, 0, 0, 0 };
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      const float w3 = s.w.w3[16 * ct + r];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int rw = 16 * rt + 4 * g + i;
          const float dyr = s.dy[rw];
          const float pre = h2pre[rt][ct][i];
          dw3[ct] += fmaxf(pre, 0.0f) * dyr;
          const float d = pre > 0.0f ? dyr * w3 : 0.0f;
          db2[ct] += d;
          const uint16_t db = f32_to_bf16(d);
          s.dh2[rw * kLd + 16 * ct + r] = db;
          s.dh2t[(16 * ct + r) * kLd + rw] = db;
        }
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      float a = dw3[ct], b = db2[ct];
      a += __shfl_xor(a, 16, 64);
      a += __shfl_xor(a, 32, 64);
      b += __shfl_xor(b, 16, 64);
      b += __shfl_xor(b, 32, 64);
      if (g == 0) {
        atomicAdd(&grads[kOffW3 + 16 * ct + r], a);
        atomicAdd(&grads[kOffB2 + 16 * ct + r], b);
      }
    }
    const float db3 = wave_sum(s.dy[lane]);
    if (lane == 0) atomicAdd(&grads[kOffB3], db3);
    wave_lds_sync();
    // dW2[k][n] = sum_rows H1[row][k] dH2[row][n]
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
          acc = mfma(lds_frag(s.h1t + (16 * mt + r) * kLd + 32 * ks + 8 * g),
                     lds_frag(s.dh2t + (16 * nt + r) * kLd + 32 * ks + 8 * g), acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) atomicAdd(&grads[kOffW2 + (16 * mt + 4 * g + i) * kHid + 16 * nt + r], acc[i]);
      }
    // dH1[row][k] = sum_n dH2[row][n] W2[k][n], masked by the layer-1 ReLU
    float db1[4] = { 0, 0, 0, 0 };
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) {
        f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
          acc = mfma(lds_frag(s.dh2 + (16 * rt + r) * kLd + 32 * ks + 8 * g),
                     lds_frag(&s.w2[16 * ct + r][32 * ks + 8 * g]), acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float d = h1pre[rt][ct][i] > 0.0f ? acc[i] : 0.0f;
          db1[ct] += d;
          s.buf[(16 * ct + r) * kLd + 16 * rt + 4 * g + i] = f32_to_bf16(d);  // dH1^T
        }
      }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      float b = db1[ct];
      b += __shfl_xor(b, 16, 64);
      b += __shfl_xor(b, 32, 64);
      if (g == 0) atomicAdd(&grads[kOffB1 + 16 * ct + r], b);
    }
    wave_lds_sync();
    // dW1[k][n] = sum_rows X[row][k] dH1[row][n]; X^T fragments come from global memory
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      bf16x8 xa[2];
#pragma unroll
      for (int ks = 0; ks < 2; ++ks) {
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int64_t rr = chunk * 64 + 32 * ks + 8 * g + j;
          v[j] = rr < n ? x[rr * kFeat + 16 * mt + r] : 0.0f;
        }
        xa[ks] = pack_frag(v);
      }
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
          acc = mfma(xa[ks], lds_frag(s.buf + (16 * nt + r) * kLd + 32 * ks + 8 * g), acc);
#pragma unroll
        for (int i = 0; i < 4; ++i) atomicAdd(&grads[kOffW1 + (16 * mt + 4 * g + i) * kHid + 16 * nt + r], acc[i]);
      }
    }
    wave_lds_sync();
  }
}

__global__ void k_adam(float* params, float* m, float* v, float* grads, int32_t n, float lr,
                       float b1, float b2, float c1, float c2, float eps)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gr = grads[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gr;
  const float vi = b2 * v[i] + (1.0f - b2) * gr * gr;
  m[i] = mi;
  v[i] = vi;
  params[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  grads[i] = 0.0f;
}

__global__ void k_zero_f32(float* p, int32_t n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

struct InferLds {
  MlpBias w;
  uint16_t buf[4][64 * kLd];
};

__global__ void __launch_bounds__(256) k_mlp_infer(const float* params, const float* x, int64_t n,
                                                   float* y)
{
  __shared__ InferLds s;
  MlpFrags fr;
  mlp_load_frags(params, fr);
  mlp_stage_bias(params, s.w);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int64_t chunk = (int64_t)blockIdx.x * 4 + wave; chunk * 64 < n; chunk += (int64_t)gridDim.x * 4) {
    const int64_t row = chunk * 64 + lane;
    float feat[kFeat];
#pragma unroll
    for (int k = 0; k < kFeat; ++k) feat[k] = row < n ? x[row * kFeat + k] : 0.0f;
    const float out = mlp_forward_fast(fr, s.w, s.buf[wave], feat);
    if (row < n) y[row] = out;
  }
}

}  // namespace

hipError_t launch_mlp_infer(const float* params, const float* x, int64_t n, float* y, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  const int64_t waves = (n + 63) / 64;
  const unsigned grid = (unsigned)std::min<int64_t>((waves + 3) / 4, 4096);
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(256), 0, st, params, x, n, y);
  return hipGetLastError();
}

hipError_t launch_mlp_grad(const float* params, const float* x, const float* t, int64_t n,
                           float inv_n, float* grads, float* loss, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<int64_t>((n + 63) / 64, 8192);
  hipLaunchKernelGGL(k_mlp_grad, dim3(grid), dim3(64), 0, st, params, x, t, n, inv_n, grads, loss);
  return hipGetLastError();
}

}  // namespace beams

using namespace beams;

struct beams_mlp {
  hipStream_t stream = nullptr;
  float* params = nullptr;  // FP32 master weights
  float* m = nullptr;
  float* v = nullptr;
  float* grads = nullptr;
  float* loss = nullptr;
  int64_t step = 0;
  std::string err;
  int fail(int code, const std::string& msg)
  {
    err = msg;
    return code;
  }
};

const float* beams::mlp_device_params(const beams_mlp* h) { return h ? h->params : nullptr; }

#define MLP_TRY(h, expr)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return (h)->fail(BEAMS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

// Device copies of caller arrays for BEAMS_MEM_HOST calls.
struct Staged {
  void* p = nullptr;
  ~Staged()
  {
    if (p) (void)hipFree(p);
  }
};

int stage_in(beams_mlp* h, const float* src, size_t count, int32_t mem, Staged& st, const float*& out)
{
  if (mem == BEAMS_MEM_DEVICE) {
    out = src;
    return BEAMS_OK;
  }
  MLP_TRY(h, hipMalloc(&st.p, sizeof(float) * std::max<size_t>(count, 1)));
  MLP_TRY(h, hipMemcpy(st.p, src, sizeof(float) * count, hipMemcpyHostToDevice));
  out = (const float*)st.p;
  return BEAMS_OK;
}

}  // namespace

extern "C" {

int beams_mlp_num_params(void) { return kParams; }

int beams_mlp_create(beams_mlp** out, uint64_t seed)
{
  if (!out) return BEAMS_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return BEAMS_ERR_NO_DEVICE;
  auto* h = new beams_mlp();
  bool ok = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) == hipSuccess;
  for (float** p : { &h->params, &h->m, &h->v, &h->grads })
    ok = ok && hipMalloc((void**)p, sizeof(float) * kParams) == hipSuccess;
  ok = ok && hipMalloc((void**)&h->loss, sizeof(float)) == hipSuccess;
  if (!ok) {
    beams_mlp_destroy(h);
    return BEAMS_ERR_HIP;
  }
  // He-uniform initialisation from a SplitMix64 stream; zero biases
  std::vector<float> w(kParams, 0.0f);
  SplitMix64 rng(seed);
  auto fill = [&](int off, int fan_in, int count) {
    const double lim = std::sqrt(6.0 / fan_in);
    for (int i = 0; i < count; ++i) w[off + i] = (float)((2.0 * rng.uniform() - 1.0) * lim);
  };
  fill(kOffW1, 31, kFeat * kHid);
  for (int nn = 0; nn < kHid; ++nn) w[kOffW1 + 31 * kHid + nn] = 0.0f;  // the zero pad's row
  fill(kOffW2, kHid, kHid * kHid);
  fill(kOffW3, kHid, kHid);
  std::vector<float> z(kParams, 0.0f);
  ok = hipMemcpy(h->params, w.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(h->m, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(h->v, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) == hipSuccess &&
       hipMemcpy(h->grads, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) {
    beams_mlp_destroy(h);
    return BEAMS_ERR_HIP;
  }
  *out = h;
  return BEAMS_OK;
}

int beams_mlp_destroy(beams_mlp* h)
{
  if (!h) return BEAMS_ERR_INVALID;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (float* p : { h->params, h->m, h->v, h->grads, h->loss })
    if (p) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return BEAMS_OK;
}

const char* beams_mlp_last_error(const beams_mlp* h) { return h ? h->err.c_str() : "null handle"; }

int beams_mlp_get_params(beams_mlp* h, float* out)
{
  if (!h || !out) return BEAMS_ERR_INVALID;
  MLP_TRY(h, hipStreamSynchronize(h->stream));
  MLP_TRY(h, hipMemcpy(out, h->params, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  return BEAMS_OK;
}

int beams_mlp_set_params(beams_mlp* h, const float* in)
{
  if (!h || !in) return BEAMS_ERR_INVALID;
  for (int i = 0; i < kParams; ++i)
    if (!std::isfinite(in[i])) return h->fail(BEAMS_ERR_INVALID, "parameters must be finite");
  MLP_TRY(h, hipStreamSynchronize(h->stream));
  MLP_TRY(h, hipMemcpy(h->params, in, sizeof(float) * kParams, hipMemcpyHostToDevice));
  std::vector<float> z(kParams, 0.0f);
  MLP_TRY(h, hipMemcpy(h->m, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  MLP_TRY(h, hipMemcpy(h->v, z.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
  h->step = 0;
  return BEAMS_OK;
}

int beams_mlp_forward(beams_mlp* h, const float* x, int64_t n, float* y, int32_t mem)
{
  if (!h || !x || !y || n < 0) return BEAMS_ERR_INVALID;
  Staged sx;
  const float* dx = nullptr;
  int rc = stage_in(h, x, (size_t)n * kFeat, mem, sx, dx);
  if (rc) return rc;
  float* dy = y;
  Staged sy;
  if (mem != BEAMS_MEM_DEVICE) {
    MLP_TRY(h, hipMalloc(&sy.p, sizeof(float) * std::max<int64_t>(n, 1)));
    dy = (float*)sy.p;
  }
  MLP_TRY(h, launch_mlp_infer(h->params, dx, n, dy, h->stream));
  MLP_TRY(h, hipStreamSynchronize(h->stream));
  if (mem != BEAMS_MEM_DEVICE) MLP_TRY(h, hipMemcpy(y, dy, sizeof(float) * n, hipMemcpyDeviceToHost));
  return BEAMS_OK;
}

int beams_mlp_gradients(beams_mlp* h, const float* x, const float* t, int64_t n, float* grads,
                        float* loss, int32_t mem)
{
  if (!h || !x || !t || !grads || n < 1) return BEAMS_ERR_INVALID;
  Staged sx, st;
  const float *dx = nullptr, *dt = nullptr;
  int rc = stage_in(h, x, (size_t)n * kFeat, mem, sx, dx);
  if (!rc) rc = stage_in(h, t, (size_t)n, mem, st, dt);
  if (rc) return rc;
  hipLaunchKernelGGL(k_zero_f32, dim3((kParams + 255) / 256), dim3(256), 0, h->stream, h->grads, kParams);
  hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, h->stream, h->loss, 1);
  MLP_TRY(h, launch_mlp_grad(h->params, dx, dt, n, 1.0f / (float)n, h->grads, h->loss, h->stream));
  MLP_TRY(h, hipStreamSynchronize(h->stream));
  MLP_TRY(h, hipMemcpy(grads, h->grads, sizeof(float) * kParams, hipMemcpyDeviceToHost));
  if (loss) {
    float l = 0.0f;
    MLP_TRY(h, hipMemcpy(&l, h->loss, sizeof(float), hipMemcpyDeviceToHost));
    *loss = l / (float)n;
  }
  hipLaunchKernelGGL(k_zero_f32, dim3((kParams + 255) / 256), dim3(256), 0, h->stream, h->grads, kParams);
  MLP_TRY(h, hipStreamSynchronize(h->stream));
  return BEAMS_OK;
}

int beams_mlp_train(beams_mlp* h, const float* x, const float* t, int64_t n, int32_t epochs,
                    int32_t batch, float lr, float* loss_out, int32_t mem)
{
  if (!h || !x || !t || n < 1 || epochs < 1 || batch < 64 || !(lr > 0.0f)) return BEAMS_ERR_INVALID;
  Staged sx, st;
  const float *dx = nullptr, *dt = nullptr;
  int rc = stage_in(h, x, (size_t)n * kFeat, mem, sx, dx);
  if (!rc) rc = stage_in(h, t, (size_t)n, mem, st, dt);
  if (rc) return rc;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float last = 0.0f;
  for (int32_t e = 0; e < epochs; ++e) {
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, h->stream, h->loss, 1);
    for (int64_t off = 0; off < n; off += batch) {
      const int64_t nb = std::min<int64_t>(batch, n - off);
      MLP_TRY(h, launch_mlp_grad(h->params, dx + off * kFeat, dt + off, nb, 1.0f / (float)nb, h->grads,
                                 h->loss, h->stream));
      ++h->step;
      const float c1 = 1.0f - std::pow(b1, (float)h->step), c2 = 1.0f - std::pow(b2, (float)h->step);
      hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, h->stream, h->params, h->m,
                         h->v, h->grads, kParams, lr, b1, b2, c1, c2, eps);
      MLP_TRY(h, hipGetLastError());
    }
    MLP_TRY(h, hipStreamSynchronize(h->stream));
    MLP_TRY(h, hipMemcpy(&last, h->loss, sizeof(float), hipMemcpyDeviceToHost));
    last /= (float)n;
  }
  if (loss_out) *loss_out = last;
  return BEAMS_OK;
}

}  // extern "C"
