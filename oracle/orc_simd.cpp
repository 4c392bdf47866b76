// orc_simd.cpp -- register-tiled SIMD correlations for the CPU baseline.
//
// TEST INFRASTRUCTURE ONLY (the oracle's timing mode; see oracle.h).
//
// The reference's CPU path convolves with cpu_convolve_xmm
// (localcues/cpu_convolve2d.cc:5-156): a 2-row x 16-column register tile of
// SSE accumulators, one broadcast filter tap per step, OpenMP over row
// pairs.  This file is the project's own version of that technique for the
// three correlation stages of the oracle: a 4-row x 32-column tile of
// AVX-512 accumulators (8 x 16 lanes; AVX2: 4 rows x 16 columns) over a
// padded input plane, one broadcast tap per step, OpenMP over row blocks.
//
// Every output lane runs exactly the scalar oracle's per-pixel sequence of
// fused multiply-adds in the scalar loop order (the tap list below), and zero
// taps are skipped as the scalar loops skip them (an exact no-op: the
// accumulators start at +0 and never become -0), so the results equal the
// scalar oracle's bit for bit.  tests/test_oracle.py checks that on all
// three stages.
#include <immintrin.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "oracle.h"

namespace {

// out[y][x] = sum over taps of in[y + ty][x + tx] * f[ty][tx].  The taps are
// listed once per call, in the scalar loop order (filter rows outer for
// order 0, filter columns outer for order 1) and without the zero taps, as
// offsets into `in` (rows in_pitch floats apart) and coefficients.
struct Tap {
  size_t off;
  float c;
};

std::vector<Tap> tap_list(size_t in_pitch, int fh, int fw, int order, const float* f) {
  std::vector<Tap> t;
  for (int a = 0; a < (order ? fw : fh); a++)
    for (int b = 0; b < (order ? fh : fw); b++) {
      const int ty = order ? b : a, tx = order ? a : b;
      const float c = f[ty * fw + tx];
      if (c != 0.0f) t.push_back(Tap{(size_t)ty * in_pitch + (size_t)tx, c});
    }
  return t;
}

// One output pixel, scalar (tile edges).
inline float pixel(const float* in, const std::vector<Tap>& taps) {
  float acc = 0.0f;
  for (const Tap& t : taps) acc = fmaf(in[t.off], t.c, acc);
  return acc;
}

constexpr int kRows = 4;

__attribute__((target("avx512f"))) void tile512(const float* in, size_t in_pitch, float* out, size_t out_pitch,
                                                const Tap* taps, size_t nt) {
  __m512 a[kRows][2];
  for (int r = 0; r < kRows; r++) a[r][0] = a[r][1] = _mm512_setzero_ps();
  for (size_t k = 0; k < nt; k++) {
    const __m512 cv = _mm512_set1_ps(taps[k].c);
    const float* p = in + taps[k].off;
    for (int r = 0; r < kRows; r++) {
      a[r][0] = _mm512_fmadd_ps(_mm512_loadu_ps(p + (size_t)r * in_pitch), cv, a[r][0]);
      a[r][1] = _mm512_fmadd_ps(_mm512_loadu_ps(p + (size_t)r * in_pitch + 16), cv, a[r][1]);
    }
  }
  for (int r = 0; r < kRows; r++) {
    _mm512_storeu_ps(out + (size_t)r * out_pitch, a[r][0]);
    _mm512_storeu_ps(out + (size_t)r * out_pitch + 16, a[r][1]);
  }
}

__attribute__((target("avx2,fma"))) void tile256(const float* in, size_t in_pitch, float* out, size_t out_pitch,
                                                 const Tap* taps, size_t nt) {
  __m256 a[kRows][2];
  for (int r = 0; r < kRows; r++) a[r][0] = a[r][1] = _mm256_setzero_ps();
  for (size_t k = 0; k < nt; k++) {
    const __m256 cv = _mm256_set1_ps(taps[k].c);
    const float* p = in + taps[k].off;
    for (int r = 0; r < kRows; r++) {
      a[r][0] = _mm256_fmadd_ps(_mm256_loadu_ps(p + (size_t)r * in_pitch), cv, a[r][0]);
      a[r][1] = _mm256_fmadd_ps(_mm256_loadu_ps(p + (size_t)r * in_pitch + 8), cv, a[r][1]);
    }
  }
  for (int r = 0; r < kRows; r++) {
    _mm256_storeu_ps(out + (size_t)r * out_pitch, a[r][0]);
    _mm256_storeu_ps(out + (size_t)r * out_pitch + 8, a[r][1]);
  }
}

bool have_avx512() {
  static const bool v = __builtin_cpu_supports("avx512f");
  return v;
}

}  // namespace

extern "C" int orc_simd_width(void) { return have_avx512() ? 32 : 16; }

// Correlation of an h x w output with an fh x fw filter over a padded input
// plane (in: the sample under output (0, 0) and tap (0, 0); in_pitch floats
// per row), OpenMP over blocks of kRows rows.
extern "C" void orc_correlate_simd(const float* in, int in_pitch, float* out, int out_pitch, int w, int h, int fh,
                                   int fw, int order, const float* f) {
  const int cw = orc_simd_width();
  const bool wide = have_avx512();
  const size_t ip = (size_t)in_pitch, op = (size_t)out_pitch;
  const std::vector<Tap> taps = tap_list(ip, fh, fw, order, f);
  const int nblk = (h + kRows - 1) / kRows;
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nblk; b++) {
    const int y0 = b * kRows;
    int x = 0;
    if (y0 + kRows <= h) {
      for (; x + cw <= w; x += cw) {
        if (wide)
          tile512(in + (size_t)y0 * ip + x, ip, out + (size_t)y0 * op + x, op, taps.data(), taps.size());
        else
          tile256(in + (size_t)y0 * ip + x, ip, out + (size_t)y0 * op + x, op, taps.data(), taps.size());
      }
    }
    for (int y = y0; y < y0 + kRows && y < h; y++)
      for (int xx = (y0 + kRows <= h) ? x : 0; xx < w; xx++) out[(size_t)y * op + xx] = pixel(in + (size_t)y * ip + xx, taps);
  }
}

// DIA SpMV y = A x in the scalar oracle's order (diagonals ascending, one
// fused multiply-add each) on rows whose 81 columns all lie in [0, n):
// rows [r0, r1), 16 (AVX-512) or 8 (AVX2) at a time, the rest one by one.
// lin[d] is diagonal d's linear column offset.
namespace {
__attribute__((target("avx512f"))) void spmv_rows512(const float* M, size_t n, const float* x, float* y, int r0,
                                                     int r1, const int* lin, int nd) {
  int r = r0;
  for (; r + 16 <= r1; r += 16) {
    __m512 acc = _mm512_setzero_ps();
    for (int d = 0; d < nd; d++)
      acc = _mm512_fmadd_ps(_mm512_loadu_ps(M + (size_t)d * n + r), _mm512_loadu_ps(x + r + lin[d]), acc);
    _mm512_storeu_ps(y + r, acc);
  }
  for (; r < r1; r++) {
    float acc = 0.0f;
    for (int d = 0; d < nd; d++) acc = fmaf(M[(size_t)d * n + r], x[r + lin[d]], acc);
    y[r] = acc;
  }
}
__attribute__((target("avx2,fma"))) void spmv_rows256(const float* M, size_t n, const float* x, float* y, int r0,
                                                      int r1, const int* lin, int nd) {
  int r = r0;
  for (; r + 8 <= r1; r += 8) {
    __m256 acc = _mm256_setzero_ps();
    for (int d = 0; d < nd; d++)
      acc = _mm256_fmadd_ps(_mm256_loadu_ps(M + (size_t)d * n + r), _mm256_loadu_ps(x + r + lin[d]), acc);
    _mm256_storeu_ps(y + r, acc);
  }
  for (; r < r1; r++) {
    float acc = 0.0f;
    for (int d = 0; d < nd; d++) acc = fmaf(M[(size_t)d * n + r], x[r + lin[d]], acc);
    y[r] = acc;
  }
}
}  // namespace

extern "C" void orc_spmv_interior_simd(const float* M, int n, const float* x, float* y, int r0, int r1,
                                       const int* lin, int nd) {
  const int chunk = 4096;
  const int nc = (r1 - r0 + chunk - 1) / chunk;
  const bool wide = have_avx512();
#pragma omp parallel for schedule(static)
  for (int c = 0; c < nc; c++) {
    const int a = r0 + c * chunk, b = a + chunk < r1 ? a + chunk : r1;
    if (wide)
      spmv_rows512(M, (size_t)n, x, y, a, b, lin, nd);
    else
      spmv_rows256(M, (size_t)n, x, y, a, b, lin, nd);
  }
}
