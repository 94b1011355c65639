// The learned shadow closure's network: 32 inputs (31 features + a zero pad) -> 64 -> 64 -> 1,
// ReLU, BF16 weights and activations, FP32 accumulation, evaluated 64 rows per wave on
// the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16).
//
// Fragment layouts of v_mfma_f32_16x16x32_bf16 (lane l, g = l / 16, r = l % 16):
//   A (16 x 32): lane holds A[r][8g .. 8g+7]       B (32 x 16): lane holds B[8g .. 8g+7][r]
//   D (16 x 16): lane holds D[4g .. 4g+3][r]
// Activations move between layers through LDS rows of kLd bf16 (72: 144-byte rows keep the
// 16 rows a fragment load touches in distinct banks).
#pragma once

#include "common.hpp"

namespace beams {

constexpr int kFeat = 32;
constexpr int kHid = 64;
constexpr int kLd = kHid + 8;
// packed FP32 parameters: W1[k][n] (32 x 64, row 31 multiplies the zero pad), b1[64],
// W2[k][n] (64 x 64), b2[64], W3[64], b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kFeat * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid;
constexpr int kParams = kOffB3 + 1;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef __attribute__((__vector_size__(8 * sizeof(uint16_t)))) uint16_t u16x8;

// Weights as the MFMA B operands want them (n-major rows of k), plus the FP32 vectors.
struct MlpLds {
  uint16_t w1t[kHid][kFeat + 8];
  uint16_t w2t[kHid][kLd];  // w2t[n][k] = W2[k][n]
  alignas(16) float b1[kHid];
  alignas(16) float b2[kHid];
  alignas(16) float w3[kHid];
  float b3;
};

__device__ __forceinline__ void mlp_stage(const float* params, MlpLds& w)
{
  for (int i = threadIdx.x; i < kFeat * kHid; i += blockDim.x) {
    const int k = i / kHid, n = i % kHid;
    w.w1t[n][k] = f32_to_bf16(params[kOffW1 + i]);
  }
  for (int i = threadIdx.x; i < kHid * kHid; i += blockDim.x) {
    const int k = i / kHid, n = i % kHid;
    w.w2t[n][k] = f32_to_bf16(params[kOffW2 + i]);
  }
  for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
    w.b1[i] = params[kOffB1 + i];
    w.b2[i] = params[kOffB2 + i];
    w.w3[i] = bf16_to_f32(f32_to_bf16(params[kOffW3 + i]));
  }
  if (threadIdx.x == 0) w.b3 = params[kOffB3];
}

__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ bf16x8 lds_frag(const uint16_t* p)
{
  return __builtin_bit_cast(bf16x8, *(const u16x8*)p);
}

__device__ __forceinline__ bf16x8 pack_frag(const float* v)
{
  u16x8 u;
#pragma unroll
  for (int i = 0; i < 8; ++i) u[i] = f32_to_bf16(v[i]);
  return __builtin_bit_cast(bf16x8, u);
}

// Round to BF16 (nearest even) with the hardware conversion (v_cvt_pk_bf16_f32); equal to
// f32_to_bf16 for every finite value.
__device__ __forceinline__ uint16_t bf16_hw(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }

// ReLU as an integer max on the bits: negative floats (sign bit set) and -0 become +0, every
// other value passes unchanged; one instruction, no NaN canonicalisation.
__device__ __forceinline__ float relu(float x)
{
  return __builtin_bit_cast(float, max(__builtin_bit_cast(int32_t, x), 0));
}

// ReLU of a and b rounded to BF16, packed in one word (a low): both are rounded first (one
// v_cvt_pk_bf16_f32), then a packed 16-bit integer max takes the ReLU on their bits.  Rounding
// keeps the sign (a negative value or -0 gives a BF16 with the sign bit set, NaN keeps its
// sign bit), so the bits equal bf16_hw(relu(a)) and bf16_hw(relu(b)).
__device__ __forceinline__ uint32_t relu_bf16x2(float a, float b)
{
  typedef short s16x2 __attribute__((ext_vector_type(2)));
  typedef float f32x2 __attribute__((ext_vector_type(2)));
  typedef __bf16 b16x2 __attribute__((ext_vector_type(2)));
  const f32x2 v = { a, b };
  const s16x2 w = __builtin_bit_cast(s16x2, __builtin_convertvector(v, b16x2));
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(w, (s16x2)0));
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Forward pass of the whole wave: lane l supplies row l's 32 features; returns row l's
// output.  buf: a per-wave LDS buffer of 64 * kLd bf16; ybuf: 64 floats.  Every lane of
// the wave must call it.  h1pre/h2pre (optional) receive the pre-activations in D layout
// ([tile rt][tile ct][i]: row 16 rt + 4 g + i, column 16 ct + r).
__device__ __forceinline__ float mlp_forward_wave(const MlpLds& w, uint16_t* buf, float* ybuf,
                                                  const float* feat, f32x4 (*h1pre)[4] = nullptr,
                                                  f32x4 (*h2pre)[4] = nullptr)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  // rows -> LDS, then the layer-1 A fragments into registers
  {
    u16x8 a[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int i = 0; i < 8; ++i) a[q][i] = f32_to_bf16(feat[8 * q + i]);
      *(u16x8*)(buf + lane * kLd + 8 * q) = a[q];
    }
  }
  wave_lds_sync();
  bf16x8 xa[4];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) xa[rt] = lds_frag(buf + (16 * rt + r) * kLd + 8 * g);
  wave_lds_sync();
  // layer 1: 4 x 4 output tiles, K = 32
#pragma unroll
  for (int ct = 0; ct < 4; ++ct) {
    const bf16x8 b = lds_frag(&w.w1t[16 * ct + r][8 * g]);
    const float bias = w.b1[16 * ct + r];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
      acc = mfma(xa[rt], b, acc);
      if (h1pre) h1pre[rt][ct] = acc + bias;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        buf[(16 * rt + 4 * g + i) * kLd + 16 * ct + r] = f32_to_bf16(fmaxf(acc[i] + bias, 0.0f));
    }
  }
  wave_lds_sync();
  // layer 2: K = 64 in two steps; layer 3 folded into the epilogue
  float part[4][4];
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int i = 0; i < 4; ++i) part[rt][i] = 0.0f;
#pragma unroll
  for (int ct = 0; ct < 4; ++ct) {
    const bf16x8 b0 = lds_frag(&w.w2t[16 * ct + r][8 * g]);
    const bf16x8 b1 = lds_frag(&w.w2t[16 * ct + r][32 + 8 * g]);
    const float bias = w.b2[16 * ct + r], w3 = w.w3[16 * ct + r];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      f32x4 acc = { 0.0f, 0.0f, 0.0f, 0.0f };
      acc = mfma(lds_frag(buf + (16 * rt + r) * kLd + 8 * g), b0, acc);
      acc = mfma(lds_frag(buf + (16 * rt + r) * kLd + 32 + 8 * g), b1, acc);
      if (h2pre) h2pre[rt][ct] = acc + bias;
#pragma unroll
      for (int i = 0; i < 4; ++i) part[rt][i] += fmaxf(acc[i] + bias, 0.0f) * w3;
    }
  }
  // sum the 16 columns each lane group holds (lanes sharing g)
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float v = part[rt][i];
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 2, 64);
      v += __shfl_xor(v, 4, 64);
      v += __shfl_xor(v, 8, 64);
      part[rt][i] = v;
    }
  if (r == 0)
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int i = 0; i < 4; ++i) ybuf[16 * rt + 4 * g + i] = part[rt][i] + w.b3;
  wave_lds_sync();
  const float y = ybuf[lane];
  wave_lds_sync();
  return y;
}

// Inference weights: each lane keeps its A-operand fragments of W1^T and W2^T in registers
// (48 VGPRs) and the workgroup shares the biases and W3 in LDS, so the LDS per workgroup is
// the per-wave activation tiles and little else.
struct MlpBias {
  alignas(16) float b1[kHid];
  alignas(16) float b2[kHid];
  alignas(16) float w3[kHid];
  float b3;
};

struct MlpFrags {
  bf16x8 w1[4];     // W1^T rows 16 nt + r, features 8g .. 8g+7
  bf16x8 w2[4][2];  // W2^T rows 16 nt + r, units 32 ks + 8g .. +7
};

__device__ __forceinline__ void mlp_load_frags(const float* params, MlpFrags& fr)
{
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    u16x8 u;
#pragma unroll
    for (int j = 0; j < 8; ++j) u[j] = bf16_hw(params[kOffW1 + (8 * g + j) * kHid + 16 * nt + r]);
    fr.w1[nt] = __builtin_bit_cast(bf16x8, u);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
      for (int j = 0; j < 8; ++j) u[j] = bf16_hw(params[kOffW2 + (32 * ks + 8 * g + j) * kHid + 16 * nt + r]);
      fr.w2[nt][ks] = __builtin_bit_cast(bf16x8, u);
    }
  }
}

__device__ __forceinline__ void mlp_stage_bias(const float* params, MlpBias& b)
{
  for (int i = threadIdx.x; i < kHid; i += blockDim.x) {
    b.b1[i] = params[kOffB1 + i];
    b.b2[i] = params[kOffB2 + i];
    b.w3[i] = bf16_to_f32(f32_to_bf16(params[kOffW3 + i]));
  }
  if (threadIdx.x == 0) b.b3 = params[kOffB3];
}

// Inference forward pass, computed transposed (H^T = W^T X^T): the weights are the A
// operands and the rows the B operands, so a lane's accumulator holds four consecutive hidden
// units of one row.  Layer 1 leaves one 8-byte LDS store per tile, and the output needs only
// a reduction over the four lane groups.  buf: a per-wave LDS tile of 64 * kLd bf16; every lane
// of the wave calls it with its row's 32 features and receives that row's output.
__device__ __forceinline__ float mlp_forward_fast(const MlpFrags& fr, const MlpBias& w, uint16_t* buf,
                                                  const float* feat)
{
#pragma clang fp contract(fast)
  const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
  {
    u16x8 a[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
      for (int i = 0; i < 8; ++i) a[q][i] = bf16_hw(feat[8 * q + i]);
      *(u16x8*)(buf + lane * kLd + 8 * q) = a[q];
    }
  }
  wave_lds_sync();
  bf16x8 xb[4];  // X^T fragments: row 16 vt + r, features 8g .. 8g+7
#pragma unroll
  for (int vt = 0; vt < 4; ++vt) xb[vt] = lds_frag(buf + (16 * vt + r) * kLd + 8 * g);
  wave_lds_sync();
  // layer 1: unit tile nt x row tile vt
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const bf16x8 a = fr.w1[nt];
    // the bias enters as the accumulator input (rounded once, after the products)
    const f32x4 bias = *(const f32x4*)&w.b1[16 * nt + 4 * g];
    // the four row tiles' products first, then their epilogues (no wait on each product)
    f32x4 acc[4];
#pragma unroll
    for (int vt = 0; vt < 4; ++vt) acc[vt] = mfma(a, xb[vt], bias);
#pragma unroll
    for (int vt = 0; vt < 4; ++vt) {
      const uint2 h = make_uint2(relu_bf16x2(acc[vt][0], acc[vt][1]), relu_bf16x2(acc[vt][2], acc[vt][3]));
      *(uint2*)(buf + (16 * vt + r) * kLd + 16 * nt + 4 * g) = h;
    }
  }
  wave_lds_sync();
  bf16x8 hb[4][2];  // H1^T fragments: row 16 vt + r, units 32 ks + 8g .. +7
#pragma unroll
  for (int vt = 0; vt < 4; ++vt)
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) hb[vt][ks] = lds_frag(buf + (16 * vt + r) * kLd + 32 * ks + 8 * g);
  wave_lds_sync();
  float part[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const bf16x8 a0 = fr.w2[nt][0];
    const bf16x8 a1 = fr.w2[nt][1];
    const f32x4 bias = *(const f32x4*)&w.b2[16 * nt + 4 * g];
    const f32x4 w3 = *(const f32x4*)&w.w3[16 * nt + 4 * g];
    // the four row tiles' first products, then their second, then the epilogues
    f32x4 acc[4];
#pragma unroll
    for (int vt = 0; vt < 4; ++vt) acc[vt] = mfma(a0, hb[vt][0], bias);
#pragma unroll
    for (int vt = 0; vt < 4; ++vt) acc[vt] = mfma(a1, hb[vt][1], acc[vt]);
#pragma unroll
    for (int vt = 0; vt < 4; ++vt)
#pragma unroll
      for (int i = 0; i < 4; ++i) part[vt] += relu(acc[vt][i]) * w3[i];
  }
  // row 16 vt + r's output: sum over the four lane groups; lane 16 g + r keeps row tile g
#pragma unroll
  for (int vt = 0; vt < 4; ++vt) {
    part[vt] += __shfl_xor(part[vt], 16, 64);
    part[vt] += __shfl_xor(part[vt], 32, 64);
  }
  const float y = g == 0 ? part[0] : (g == 1 ? part[1] : (g == 2 ? part[2] : part[3]));
  return y + w.b3;
}

}  // namespace beams
