// dmsc_device.hpp -- device helpers shared by the gPb kernels.
//
// Numerical contract (identical to the CPU oracle): the library is built with
// -ffp-contract=off; every fused multiply-add below is an explicit fmaf();
// divisions and square roots are IEEE correctly rounded (HIP defaults).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dmsc {
namespace dev {

__host__ __device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Symmetric reflection of index i of a border-padded axis of n samples
// (convert.cu:128-141).
__host__ __device__ inline int mirror(int i, int border, int n) {
  return i < border ? border - i - 1 : (i < border + n ? i - border : border + 2 * n - i - 1);
}

// Cube root from basic IEEE double operations only (exponent split, linear
// initial guess on [1,8), six Newton steps, rounded to float).  Produces the
// same bits on the GPU and on any IEEE host built without contraction.
__device__ inline float cbrt_exact(float xf) {
  if (xf == 0.0f) return xf;
  const double x = (double)xf;
  const double ax = x < 0 ? -x : x;
  const unsigned long long bits = (unsigned long long)__double_as_longlong(ax);
  const int e = (int)((bits >> 52) & 0x7ff) - 1023;
  const int q = (e >= 0) ? e / 3 : -((-e + 2) / 3);
  const int rem = e - 3 * q;
  double m = __longlong_as_double((long long)((bits & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
  for (int i = 0; i < rem; i++) m = m * 2.0;
  double y = 1.0 + (m - 1.0) * (1.0 / 7.0);
#pragma unroll
  for (int it = 0; it < 6; it++) y = y - (y * y * y - m) / (3.0 * y * y);
  const double scale = __longlong_as_double((long long)((unsigned long long)(q + 1023) << 52));
  const double r = y * scale;
  return (float)(x < 0 ? -r : r);
}

// exp of a float argument from basic IEEE double operations: Cody-Waite
// reduction x = n ln2 + r (n the nearest integer to x / ln2, ln2 split in a
// high part with 21 trailing zero bits and a low part, the split of fdlibm's
// e_exp.c), exp(r) by the degree-12 Taylor polynomial (|r| <= ln2 / 2:
// remainder below 2e-16 relative) in Horner form with fused multiply-adds,
// times 2^n, rounded once to float.  The oracle runs the same operations,
// so results are bit-identical to it.
__device__ inline float exp_exact(float xf) {
  if (xf != xf) return xf;
  if (xf > 89.0f) return __int_as_float(0x7f800000);
  if (xf < -104.0f) return 0.0f;
  const double x = (double)xf;
  const double n = rint(x * 1.4426950408889634);
  double r = fma(-n, 6.93147180369123816490e-01, x);
  r = fma(-n, 1.90821492927058770002e-10, r);
  constexpr double f[13] = {1.0,       1.0,       1.0 / 2,    1.0 / 6,     1.0 / 24,      1.0 / 120,      1.0 / 720,
                            1.0 / 5040, 1.0 / 40320, 1.0 / 362880, 1.0 / 3628800, 1.0 / 39916800, 1.0 / 479001600};
  double p = f[12];
#pragma unroll
  for (int k = 11; k >= 0; k--) p = fma(p, r, f[k]);
  const double scale = __longlong_as_double((long long)(((unsigned long long)((long long)n + 1023)) << 52));
  return (float)(p * scale);
}

// In-register 32-slot tree sum in the order of the reference's warp
// reduction (rotate.cu:378-389): (i, i+16), (i, i+8), ..., (0, 1).
template <typename T>
__device__ __forceinline__ float tree32(T* t) {
#pragma unroll
  for (int s = 16; s >= 1; s >>= 1)
#pragma unroll
    for (int i = 0; i < s; i++) t[i] = t[i] + t[i + s];
  return t[0];
}

// tree32 for t[NZ..31] = +0 and t[0..NZ) >= +0: the same sums, without
// the additions of a known +0 (exact for a non-negative left operand).
template <int NZ, typename T>
__device__ __forceinline__ float tree32z(T* t) {
  int nz = NZ;
#pragma unroll
  for (int s = 16; s >= 1; s >>= 1) {
#pragma unroll
    for (int i = 0; i < s; i++)
      if (i + s < nz) t[i] = t[i] + t[i + s];
    nz = nz < s ? nz : s;
  }
  return t[0];
}

// Logical block of a 1-D grid, XCD-aware.  The device deals workgroups to
// its 8 XCDs in turn (workgroup i to XCD i % 8), and each XCD has its own
// L2; this bijection of [0, G) gives XCD x one contiguous range of logical
// blocks, so the neighbour gathers of its rows mostly hit its own L2.
// Correctness never depends on the dealing order.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block(int i, int G) {
  const int x = i % kXcds, j = i / kXcds;
  const int q = G / kXcds, r = G % kXcds;
  return x * q + (x < r ? x : r) + j;
}

}  // namespace dev
}  // namespace dmsc
