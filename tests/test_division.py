"""The corrected-reciprocal quotient the exact kernels use instead of a divide.

A parametric coordinate (p - x_c) / (x_{c+1} - x_c) is computed as q = a*y corrected once,
q + (a - q*b)*y with fused products, from y = RN(1/b) staged per cell; the exchange's FP64
slab quotients use the same step.  Markstein's theorem makes that the correctly rounded
quotient when y is normal and the residual a - q*b is representable; the kernels take the
shortcut only for cell widths in [2^-72, 2^61] and first quotients of at least 2^-28 in FP32
(2^-900 in FP64, and a non-zero direction) and divide otherwise.  This test checks the identity on the CPU, with the
host's IEEE divide and fmaf/fma, over random and grid-shaped operands including all-ones and
power-of-two significands.
"""

import os
import shutil
import subprocess
import tempfile

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 0x2545F4914F6CDD1Dull;
static uint64_t nx(void) {
  s += 0x9e3779b97f4a7c15ull; uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull; z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static float fbits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static int same32(float x, float y) { return memcmp(&x, &y, 4) == 0; }
static int same64(double x, double y) { return memcmp(&x, &y, 8) == 0; }
/* the kernels' div_by_recip / div64_by_recip */
static float q32(float a, float b, float y) {
  const float q = a * y;
  if (fabsf(q) >= 0x1p-28f) return fmaf(fmaf(-q, b, a), y, q);
  return a / b;
}
static double q64(double a, double d, double y) {
  const double q = a * y;
  if (d != 0.0 && fabs(a) >= 0x1p-900 && fabs(q) >= 0x1p-900) return fma(fma(-q, d, a), y, q);
  return a / d;
}
int main(void) {
  long bad = 0, n = 0;
  /* grid cells: x_c = o + c*h in FP32, points in and around each cell */
  const int dims[] = { 2, 20, 33, 64, 65, 257, 513, 1025 };
  for (int di = 0; di < 8; ++di) {
    const int N = dims[di];
    for (int t = 0; t < 3; ++t) {
      const float o = t == 0 ? 0.0f : (t == 1 ? -0.5f : 0.125f), ext = t == 2 ? 2.37f : 1.0f;
      const float h = ext / (float)(N - 1);
      for (int c = 0; c < N - 1; ++c) {
        const float x0 = o + h * (float)c, x1 = o + h * (float)(c + 1), b = x1 - x0, y = 1.0f / b;
        for (int k = 0; k < 64; ++k) {
          float p = x0 + (float)((nx() >> 40) * (1.0 / 16777216.0)) * b * 1.2f - 0.1f * b;
          if (k == 0) p = x0;
          if (k == 1) p = x1;
          if (k == 2) p = nextafterf(x0, x1);
          if (k == 3) p = nextafterf(x1, x0);
          const float a = p - x0;
          ++n;
          if (!same32(q32(a, b, y), a / b)) ++bad;
        }
      }
    }
  }
  /* random operands: widths in [2^-72, 2^61], quotients of magnitude 2^-40 .. 4 */
  for (long i = 0; i < 4000000; ++i) {
    uint32_t ub = ((uint32_t)nx() & 0x007fffffu) | ((uint32_t)(127 - 72 + (int)(nx() % 134)) << 23);
    if ((i & 7) == 0) ub |= 0x7fffffu;
    if ((i & 15) == 1) ub &= 0xff800000u;
    const float b = fbits(ub), y = 1.0f / b;
    float a = b * ldexpf((float)((nx() >> 40) | 1) / 16777216.0f, (int)(nx() % 43) - 39);
    if (nx() & 1) a = -a;
    ++n;
    if (!same32(q32(a, b, y), a / b)) ++bad;
  }
  /* FP64 slab terms: (bound - start) / dir with FP32 inputs */
  for (long i = 0; i < 2000000; ++i) {
    const float st = fbits(((uint32_t)nx() & 0x807fffffu) | ((uint32_t)(127 - 20 + nx() % 24) << 23));
    const float bd = fbits(((uint32_t)nx() & 0x807fffffu) | ((uint32_t)(127 - 20 + nx() % 24) << 23));
    float dv = fbits(((uint32_t)nx() & 0x807fffffu) | ((uint32_t)(127 - 40 + nx() % 41) << 23));
    if ((i & 7) == 0) dv = fbits(((uint32_t)nx() & 0x80000000u) | 0x3f7fffffu);
    const double a = (double)bd - (double)st, d = (double)dv;
    ++n;
    if (!same64(q64(a, d, 1.0 / d), a / d)) ++bad;
  }
  /* zero numerators and zero directions take the divide */
  if (!same32(q32(0.0f, 0.5f, 2.0f), 0.0f / 0.5f)) ++bad;
  if (!same64(q64(1.0, 0.0, 1.0 / 0.0), 1.0 / 0.0)) ++bad;
  printf("%ld %ld\n", n, bad);
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_corrected_reciprocal_quotient_is_the_ieee_quotient():
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "q.c"), os.path.join(d, "q")
        open(src, "w").write(SRC)
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-mfma", "-o", exe, src, "-lm"], check=True)
        n, bad = map(int, subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split())
    assert n > 6_000_000
    assert bad == 0
