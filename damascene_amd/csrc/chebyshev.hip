// chebyshev.hip -- block eigensolver for the mixed path: Chebyshev-filtered
// subspace iteration on L_sym for its nvec smallest eigenpairs.
//
// The reference finds them with single-vector Lanczos (lanczos.cu).  The
// wanted eigenvalues sit in a narrow cluster at 0 (481x321: nine values in
// [0, 1.5e-4], neighbours 1e-5 apart) of a spectrum that reaches 2, so one
// Krylov vector needs about 2,000 steps, each a grid-wide reduction.  This
// solver filters a block of kB = 32 vectors with a Chebyshev polynomial that
// damps [a, 2] (a = the block's largest Ritz value): the
// wanted vectors converge at a rate set by their gap to the 33rd eigenvalue
// rather than to the 10th, and a filter step has no reduction.  One outer
// iteration:
//   1. deg filter steps y = alpha (A x - c x) - beta z, one kernel each
//      (the scaled three-term recurrence, so the block stays O(1)); deg is
//      50 at first, then whatever cuts the slowest wanted vector's error
//      tenfold given the current Ritz values (30..300);
//   2. W = A Y; the Gram matrices G = Y'Y, H = Y'W and C = X0'Y in FP64
//      (X0: the previous iteration's Ritz vectors, the filter's input);
//   3. on the host, H s = theta G s in FP64 (Cholesky of G, cyclic Jacobi);
//      the next block X0 = Y S holds G-orthonormal Ritz vectors;
//   4. stop when every wanted Ritz vector moved by less than the bound
//      since the previous iteration: sin of the angle between
//      X0 e_j and Y s_j, from C.  Convergence is linear, so that change
//      bounds the previous iterate's error and exceeds the new one's.  Or
//      stop when every wanted residual |A y - theta y| (from K = W'W) is at
//      most the mixed Lanczos bound, tol * 10^-lanczos_mixed_stop: vectors
//      of a degenerate eigenvalue need not stop moving; or when the
//      residuals stagnate at the FP32 storage floor (reported to the caller).
// Blocks are FP32 in HBM as kB/4 planes of float4 per pixel; the filter
// sums in FP32, the Gram matrices and the rotation in FP64.  Matrix products
// use the reference's stencil semantics: the neighbour at offset (dy, dx)
// of pixel p is the linear index p + dy*w + dx, 0 outside [0, n)
// (stencilMVM.cu:127-180), as in the Lanczos steps.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {

int eigensolver_method = 1;  // mixed path: 0 Lanczos, 1 Chebyshev block, the default (tune key "eigensolver")
// filter tile: 24, 28, 22, 42 register tiles (rows per lane * 10 + float4
// planes per lane); 3RN LDS tiles (R rows per lane, N waves per block);
// 13RN the same with pipelined matrix loads; 23RN those with the LDS rows
// read ahead of their use; 0 (the default) picks 2328 or 2337 per image
// size (chebyshev_shape_for)
int chebyshev_shape = 0;
int chebyshev_reduction = 10;
int chebyshev_first_degree = 50;  // filter degree of the first outer iteration (tune key "chebyshev_first_degree")
int chebyshev_start = 1;  // starting block: 0 hashed uniform values, 1 low-frequency cosines (the default)  // per outer iteration, the filter degree aims to cut the slowest wanted error by this

namespace {

constexpr int kB = 32;             // block vectors
constexpr int kPlanes = kB / 4;    // float4 planes
constexpr int kMinDeg = 30, kMaxDeg = 300;
constexpr int kMaxOuter = 60;
constexpr double kFirstCut = 1e-2;  // damped interval of the first filter: [kFirstCut, 2]
constexpr double kUpper = 2.0;      // spectrum of L_sym lies in [0, 2]
constexpr int kGramPx = 64;         // pixels staged per Gram round
// pinned host block: Gram matrices at [8 KiB, 40 KiB), the rotation at [40 KiB, 48 KiB)
constexpr size_t kPinGram = 8192, kPinRot = 40960, kPinTheta = 49152, kPinRes = 49408;
constexpr int kGrams = 3;  // Y'Y, Y'W, X0'Y
constexpr double kFloorMargin = 4.0;  // residual floor = kFloorMargin * 2 * 2^-24 (about 4.8e-7)

// rows of the radius-5 disc at column offset dx: |dy| <= half_height(dx)
__host__ __device__ constexpr int half_height(int dx) {
  return dx == 0 ? 5 : (dx >= -3 && dx <= 3) ? 4 : (dx == 4 || dx == -4) ? 3 : 0;
}

__device__ __forceinline__ float mad(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double mad(double a, double b, double c) { return fma(a, b, c); }

// Diagonal index of stencil offset (dy, dx): the radius-5 disc in raster
// order (dy outer, dx inner), as make_stencil_offsets enumerates it (the
// host checks the two agree before the first launch).
__host__ __device__ constexpr int disc_start(int dy) {
  return dy == -5 ? 0 : dy == -4 ? 1 : dy == -3 ? 8 : dy == -2 ? 17 : dy == -1 ? 26 : dy == 0 ? 35
       : dy == 1 ? 46 : dy == 2 ? 55 : dy == 3 ? 64 : dy == 4 ? 73 : 80;
}
__host__ __device__ constexpr int disc_index(int dy, int dx) { return disc_start(dy) + dx + half_height(dy); }

// One filter step on a 64 x R pixel tile for VP float4 planes per thread
// (kPlanes / VP waves per block): y = alpha (A x - c x) - beta z.  Each lane
// owns one image column and R rows.  Per column offset dx it walks the
// R + 2 half_height(dx) rows of x the tile reaches, loads each x value once
// and applies it to all R output rows (an offset outside the disc
// multiplies by 0), so the x loads drop from 81 R to 114 (R = 4) per lane.
// The loops over dx and the rows stay rolled and every load is branch-free
// (issued at a clamped index, the value selected afterwards): one row's
// VP + R loads are in flight together and the registers stay few, so many
// waves share each SIMD.  Sums are FP32 in the filter and FP64 (F64) for
// W = A Y: L_sym's rows sum to about 0 against the slowly varying wanted
// vectors, so an FP32 sum of the 81 products carries an error of about
// 1e-7 |y| in a result of about 1e-4 |y|, enough to rotate Ritz vectors
// whose eigenvalues lie 1e-5 apart by 1e-4 through H.  The filter tolerates
// it: its rounding lands in directions the next steps damp or the
// Rayleigh-Ritz step resolves.
template <int R, int VP, bool F64>
__global__ __launch_bounds__(64 * (kPlanes / VP)) void k_cheb_step(
    int w, int h, int ntx, const float* __restrict__ A, size_t pitch, const float4* __restrict__ x,
    const float4* __restrict__ z, float4* __restrict__ y, size_t vs, float alpha, float c, float beta) {
  const int t = dev::xcd_block(blockIdx.x, gridDim.x);
  const int tx = t % ntx, ty = t / ntx;
  const int lane = threadIdx.x & 63;
  const int pl0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * VP;  // wave-uniform
  const int px = tx * 64 + lane, y0 = ty * R;
  const int n = w * h;
  const bool col = px < w;
  using T = typename std::conditional<F64, double, float>::type;
  T acc[R][VP][4];
  int pr[R];
#pragma unroll
  for (int r = 0; r < R; r++) {
    pr[r] = (col && y0 + r < h) ? (y0 + r) * w + px : -1;
#pragma unroll
    for (int v = 0; v < VP; v++)
#pragma unroll
      for (int k = 0; k < 4; k++) acc[r][v][k] = T(0);
  }
#pragma unroll 1
  for (int dx = -5; dx <= 5; dx++) {
    const int hm = half_height(dx);
#pragma unroll 4
    for (int yy = -hm; yy < R + hm; yy++) {
      const int q = (y0 + yy) * w + px + dx;
      const bool in = col && (unsigned)q < (unsigned)n;
      const unsigned qc = in ? (unsigned)q : 0u;
      float4 xv[VP];
#pragma unroll
      for (int v = 0; v < VP; v++) {
        const float4 t4 = x[(size_t)(pl0 + v) * vs + qc];
        xv[v] = in ? t4 : make_float4(0.f, 0.f, 0.f, 0.f);
      }
#pragma unroll
      for (int r = 0; r < R; r++) {
        const int dy = yy - r;
        const bool disc = dy >= -hm && dy <= hm;
        const bool use = disc && pr[r] >= 0;
        const size_t ai = use ? (size_t)disc_index(dy, dx) * pitch + (unsigned)pr[r] : 0;
        const float a0 = A[ai];
        const T a = use ? a0 : 0.f;
#pragma unroll
        for (int v = 0; v < VP; v++) {
          acc[r][v][0] = mad(a, (T)xv[v].x, acc[r][v][0]);
          acc[r][v][1] = mad(a, (T)xv[v].y, acc[r][v][1]);
          acc[r][v][2] = mad(a, (T)xv[v].z, acc[r][v][2]);
          acc[r][v][3] = mad(a, (T)xv[v].w, acc[r][v][3]);
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (!col || y0 + r >= h) continue;
    const int p = (y0 + r) * w + px;
#pragma unroll
    for (int v = 0; v < VP; v++) {
      const size_t i = (size_t)(pl0 + v) * vs + p;
      const float4 xo = x[i];
      const float4 zo = z ? z[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      float4 o;
      o.x = (float)(alpha * mad((T)-c, (T)xo.x, acc[r][v][0]) - beta * zo.x);
      o.y = (float)(alpha * mad((T)-c, (T)xo.y, acc[r][v][1]) - beta * zo.y);
      o.z = (float)(alpha * mad((T)-c, (T)xo.z, acc[r][v][2]) - beta * zo.z);
      o.w = (float)(alpha * mad((T)-c, (T)xo.w, acc[r][v][3]) - beta * zo.w);
      y[i] = o;
    }
  }
}

// The same filter step with the block's input staged in LDS.  A block owns
// a 64 x (R * NW) pixel tile and one group of 8 vectors (two float4
// planes); the four groups of a tile are consecutive blocks, so they run
// on one XCD and read the tile's matrix rows through one L2.  The tile's
// input values plus a 5-pixel frame, ((64 + 10) x (R * NW + 10) slots per
// plane, slot (sy, sx) holding the input at linear index (y0 - 5 + sy) * w
// + (x0 - 5 + sx) or 0 outside [0, n): the reference's linear indexing,
// wrap across image rows included) are loaded once per block.  Each wave
// owns R rows of the tile, each lane one column: per column offset dx it
// loads the R * (2 half_height(dx) + 1) matrix values its rows need, then
// walks the R + 2 half_height(dx) input rows from LDS, applying each to
// every output row whose offset lies in the disc.  The sums run in the
// order of k_cheb_step (dx outer, dy ascending inner), so the results are
// bit-identical to it.  Lanes past the image's right or bottom edge
// compute on clamped matrix addresses and store nothing.
// One column offset dx (HM = half_height(dx)) of k_cheb_lds: the matrix
// values of the lane's R rows, then the R + 2 HM input rows from LDS.
template <int HM, int R, typename T>
__device__ __forceinline__ void cheb_lds_dx(int dx, const float* __restrict__ A, size_t pitch, const int (&prow)[R],
                                            const float4* s0, const float4* s1, int sw, T (&acc)[R][8]) {
  float a[R][2 * HM + 1];
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int dy = -HM; dy <= HM; dy++) a[r][dy + HM] = A[(size_t)disc_index(dy, dx) * pitch + (unsigned)prow[r]];
#pragma unroll
  for (int yy = -HM; yy < R + HM; yy++) {
    const float4 v0 = s0[(yy + 5) * sw + dx];
    const float4 v1 = s1[(yy + 5) * sw + dx];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int dy = yy - r;
      if (dy < -HM || dy > HM) continue;
      const T av = a[r][dy + HM];
      acc[r][0] = mad(av, (T)v0.x, acc[r][0]);
      acc[r][1] = mad(av, (T)v0.y, acc[r][1]);
      acc[r][2] = mad(av, (T)v0.z, acc[r][2]);
      acc[r][3] = mad(av, (T)v0.w, acc[r][3]);
      acc[r][4] = mad(av, (T)v1.x, acc[r][4]);
      acc[r][5] = mad(av, (T)v1.y, acc[r][5]);
      acc[r][6] = mad(av, (T)v1.z, acc[r][6]);
      acc[r][7] = mad(av, (T)v1.w, acc[r][7]);
    }
  }
}

// The same column-offset pass split in two, for software pipelining: the
// matrix values of one offset (into a[r][dy + 5]) and the LDS walk that
// applies them.
// Buffer loads: the lane's row offset (prow * 4 bytes) is the one vector
// offset of all its loads and the diagonal's offset (d * pitch * 4) a scalar,
// so no per-load 64-bit address is formed; the descriptor's range is the
// whole matrix (out-of-range reads would return 0).
template <int HM, int R>
__device__ __forceinline__ void cheb_load_a(int dx, __amdgpu_buffer_rsrc_t rs, unsigned pitch_bytes,
                                            const unsigned (&voff)[R], float (&a)[R][11]) {
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int dy = -HM; dy <= HM; dy++)
      a[r][dy + 5] = __builtin_bit_cast(
          float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff[r], (unsigned)disc_index(dy, dx) * pitch_bytes, 0));
}

// PD > 0: each input row is read from LDS PD rows ahead of the row it
// feeds, so PD rows of reads are in flight while the FMAs run (the sums and
// their order are unchanged).
template <int HM, int R, typename T, int PD = 0>
__device__ __forceinline__ void cheb_apply(int dx, const float (&a)[R][11], const float4* s0, const float4* s1, int sw,
                                           T (&acc)[R][8]) {
  constexpr int NY = R + 2 * HM;
  float4 b0[NY], b1[NY];
#pragma unroll
  for (int k = 0; k < PD && k < NY; k++) {
    b0[k] = s0[(k - HM + 5) * sw + dx];
    b1[k] = s1[(k - HM + 5) * sw + dx];
  }
#pragma unroll
  for (int yy = -HM; yy < R + HM; yy++) {
    const int k = yy + HM;
    if (PD == 0) {
      b0[k] = s0[(yy + 5) * sw + dx];
      b1[k] = s1[(yy + 5) * sw + dx];
    } else if (k + PD < NY) {
      b0[k + PD] = s0[(yy + PD + 5) * sw + dx];
      b1[k + PD] = s1[(yy + PD + 5) * sw + dx];
    }
    const float4 v0 = b0[k], v1 = b1[k];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int dy = yy - r;
      if (dy < -HM || dy > HM) continue;
      const T av = a[r][dy + 5];
      acc[r][0] = mad(av, (T)v0.x, acc[r][0]);
      acc[r][1] = mad(av, (T)v0.y, acc[r][1]);
      acc[r][2] = mad(av, (T)v0.z, acc[r][2]);
      acc[r][3] = mad(av, (T)v0.w, acc[r][3]);
      acc[r][4] = mad(av, (T)v1.x, acc[r][4]);
      acc[r][5] = mad(av, (T)v1.y, acc[r][5]);
      acc[r][6] = mad(av, (T)v1.z, acc[r][6]);
      acc[r][7] = mad(av, (T)v1.w, acc[r][7]);
    }
  }
}

// Empty asm statements that take the accumulators as operands: the sums
// of a pass are complete at this point of the instruction stream, so the
// compiler cannot defer the arithmetic behind later loads.
template <int R, typename T>
__device__ __forceinline__ void cheb_fence(T (&acc)[R][8]) {
#pragma unroll
  for (int r = 0; r < R; r++)
    asm volatile("" : "+v"(acc[r][0]), "+v"(acc[r][1]), "+v"(acc[r][2]), "+v"(acc[r][3]), "+v"(acc[r][4]),
                 "+v"(acc[r][5]), "+v"(acc[r][6]), "+v"(acc[r][7])::"memory");
}

// One stage of the pipelined pass sequence: issue offset dx + 1's matrix
// loads, then apply offset dx's values; the fences keep one stage's loads
// and sums together (left alone, the compiler issues every load of the
// eleven stages first and spills them).
#define CHEB_STAGE(HMC, HMN, DX, CUR, NXT)                                   \
  cheb_load_a<HMN, R>((DX) + 1, rs, pitch_bytes, voff, NXT);                 \
  cheb_apply<HMC, R, T, PD>((DX), CUR, s0, s1, SW, acc);                     \
  cheb_fence<R, T>(acc);

template <int R, int NW, bool F64, bool PF, int PD>
__device__ __forceinline__ void cheb_lds_body(int w, int h, int ntx, const float* __restrict__ A, size_t pitch,
                                              const float4* __restrict__ x, const float4* __restrict__ z,
                                              float4* __restrict__ y, size_t vs, float alpha, float c, float beta) {
  constexpr int TH = R * NW, SW = 64 + 10, SH = TH + 10, NT = 64 * NW;
  constexpr int kSlots = 2 * SH * SW, kIt = (kSlots + NT - 1) / NT;
  __shared__ float4 sx[2][SH][SW];
  const int b = dev::xcd_block(blockIdx.x, gridDim.x);
  const int grp = b & 3, t = b >> 2;
  const int x0 = (t % ntx) * 64, y0 = (t / ntx) * TH;
  const int n = w * h;
  const int pl0 = grp * 2;
  {
    float4 v[kIt];
#pragma unroll
    for (int j = 0; j < kIt; j++) {
      const int i = (int)threadIdx.x + j * NT;
      const int pl = i / (SH * SW), rem = i - pl * (SH * SW);
      const int sy = rem / SW, sxx = rem - sy * SW;
      const int q = (y0 - 5 + sy) * w + (x0 - 5 + sxx);
      const bool in = i < kSlots && (unsigned)q < (unsigned)n;
      const float4 t4 = x[(size_t)(pl0 + (i < kSlots ? pl : 0)) * vs + (in ? (unsigned)q : 0u)];
      v[j] = in ? t4 : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int j = 0; j < kIt; j++) {
      const int i = (int)threadIdx.x + j * NT;
      if (i < kSlots) (&sx[0][0][0])[i] = v[j];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int r0 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)) * R;  // wave-uniform
  const int px = x0 + lane;
  int prow[R];  // matrix row of each output (clamped to 0 past the image: computed, never stored)
#pragma unroll
  for (int r = 0; r < R; r++) prow[r] = (px < w && y0 + r0 + r < h) ? (y0 + r0 + r) * w + px : 0;
  using T = typename std::conditional<F64, double, float>::type;
  T acc[R][8];
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int k = 0; k < 8; k++) acc[r][k] = T(0);
  // input slot of (tile row r0, this lane's column) at dx = 0, dy = -5
  const float4* s0 = &sx[0][r0][lane + 5];
  const float4* s1 = &sx[1][r0][lane + 5];
  if constexpr (PF) {
    // column offsets -5..5 have half heights 0 3 4 4 4 5 4 4 4 3 0
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc((void*)A, (short)0, (int)(kStencilArea * pitch * sizeof(float)), 0x00020000);
    const unsigned pitch_bytes = (unsigned)(pitch * sizeof(float));
    unsigned voff[R];
#pragma unroll
    for (int r = 0; r < R; r++) voff[r] = (unsigned)prow[r] * 4u;
    float a0[R][11], a1[R][11];
    cheb_load_a<0, R>(-5, rs, pitch_bytes, voff, a0);
    CHEB_STAGE(0, 3, -5, a0, a1)
    CHEB_STAGE(3, 4, -4, a1, a0)
    CHEB_STAGE(4, 4, -3, a0, a1)
    CHEB_STAGE(4, 4, -2, a1, a0)
    CHEB_STAGE(4, 5, -1, a0, a1)
    CHEB_STAGE(5, 4, 0, a1, a0)
    CHEB_STAGE(4, 4, 1, a0, a1)
    CHEB_STAGE(4, 4, 2, a1, a0)
    CHEB_STAGE(4, 3, 3, a0, a1)
    CHEB_STAGE(3, 0, 4, a1, a0)
    cheb_apply<0, R, T, PD>(5, a0, s0, s1, SW, acc);
  } else {
#pragma unroll 1
    for (int dx = -5; dx <= 5; dx++) {
      switch (half_height(dx)) {
        case 0: cheb_lds_dx<0, R, T>(dx, A, pitch, prow, s0, s1, SW, acc); break;
        case 3: cheb_lds_dx<3, R, T>(dx, A, pitch, prow, s0, s1, SW, acc); break;
        case 4: cheb_lds_dx<4, R, T>(dx, A, pitch, prow, s0, s1, SW, acc); break;
        default: cheb_lds_dx<5, R, T>(dx, A, pitch, prow, s0, s1, SW, acc); break;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    if (px >= w || y0 + r0 + r >= h) continue;
    const int p = (y0 + r0 + r) * w + px;
#pragma unroll
    for (int v = 0; v < 2; v++) {
      const size_t i = (size_t)(pl0 + v) * vs + p;
      const float4 xo = sx[v][r0 + r + 5][lane + 5];
      const float4 zo = z ? z[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      float4 o;
      o.x = (float)(alpha * mad((T)-c, (T)xo.x, acc[r][4 * v + 0]) - beta * zo.x);
      o.y = (float)(alpha * mad((T)-c, (T)xo.y, acc[r][4 * v + 1]) - beta * zo.y);
      o.z = (float)(alpha * mad((T)-c, (T)xo.z, acc[r][4 * v + 2]) - beta * zo.z);
      o.w = (float)(alpha * mad((T)-c, (T)xo.w, acc[r][4 * v + 3]) - beta * zo.w);
      y[i] = o;
    }
  }
}

template <int R, int NW, bool F64, bool PF = false>
__global__ __launch_bounds__(64 * NW) void k_cheb_lds(int w, int h, int ntx, const float* __restrict__ A,
                                                      size_t pitch, const float4* __restrict__ x,
                                                      const float4* __restrict__ z, float4* __restrict__ y,
                                                      size_t vs, float alpha, float c, float beta) {
  cheb_lds_body<R, NW, F64, PF, 0>(w, h, ntx, A, pitch, x, z, y, vs, alpha, c, beta);
}

// k_cheb_lds (FP32 sums, pipelined matrix loads) with the LDS rows read
// ahead of their FMAs, four rows for R = 2 and three for R = 3 (the depths
// a sweep of 2 to 6 rows found fastest; deeper costs occupancy at R = 3):
// bit-identical results.
template <int R, int NW>
__global__ __launch_bounds__(64 * NW) void k_cheb_ldp(int w, int h, int ntx, const float* __restrict__ A,
                                                      size_t pitch, const float4* __restrict__ x,
                                                      const float4* __restrict__ z, float4* __restrict__ y,
                                                      size_t vs, float alpha, float c, float beta) {
  cheb_lds_body<R, NW, false, true, R == 2 ? 4 : 3>(w, h, ntx, A, pitch, x, z, y, vs, alpha, c, beta);
}

// The FP64-sum product W = AY (alpha 1, c and beta 0) of the read-ahead
// tiles, with pipelined matrix loads: bit-identical to k_cheb_lds<R, NW,
// true>.
template <int R, int NW>
__global__ __launch_bounds__(64 * NW) void k_cheb_ldp64(int w, int h, int ntx, const float* __restrict__ A,
                                                        size_t pitch, const float4* __restrict__ x,
                                                        const float4* __restrict__ z, float4* __restrict__ y,
                                                        size_t vs, float alpha, float c, float beta) {
  cheb_lds_body<R, NW, true, true, 3>(w, h, ntx, A, pitch, x, z, y, vs, alpha, c, beta);
}

// Uniform value in [-1, 1) from an integer hash of element index i.
__device__ __forceinline__ float cheb_hash(size_t i) {
  unsigned u = (unsigned)i * 0x9E3779B1u + 0x7F4A7C15u;
  u ^= u >> 16;
  u *= 0x85EBCA6Bu;
  u ^= u >> 13;
  u *= 0xC2B2AE35u;
  u ^= u >> 16;
  return (float)(u >> 8) * (1.0f / 8388608.0f) - 1.0f;
}

// Starting block: uniform values in [-1, 1) from an integer hash of
// (plane, pixel, component); rows n..vs-1 are 0.
__global__ void k_cheb_init(int n, size_t vs, float4* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)kPlanes * vs) return;
  const int p = (int)(i % vs);
  float e[4];
  for (int k = 0; k < 4; k++) e[k] = p < n ? cheb_hash(i * 4 + k) : 0.0f;
  x[i] = make_float4(e[0], e[1], e[2], e[3]);
}

// The 32 lowest cosine modes of a w x h grid (the eigenvectors of the grid's
// uniform Laplacian): vector j is cos(pi kx_j (x + 1/2) / w) cos(pi ky_j (y +
// 1/2) / h).
struct CosModes {
  int kx[kB], ky[kB];
};

// Starting block of low-frequency cosines plus 2^-6 of the hashed values
// of k_cheb_init, so that every eigenvector has a share of the start.
__global__ void k_cheb_init_cos(int w, int n, size_t vs, CosModes m, float4* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)kPlanes * vs) return;
  const int pl = (int)(i / vs), p = (int)(i % vs);
  const int h = n / w, y = p / w, xx = p - y * w;
  float e[4];
  for (int k = 0; k < 4; k++) {
    const int j = pl * 4 + k;
    e[k] = p < n ? cospif(m.kx[j] * (xx + 0.5f) / (float)w) * cospif(m.ky[j] * (y + 0.5f) / (float)h) +
                       0.015625f * cheb_hash(i * 4 + k)
                 : 0.0f;
  }
  x[i] = make_float4(e[0], e[1], e[2], e[3]);
}

CosModes cos_modes(int w, int h) {
  std::vector<std::pair<double, int>> f;
  for (int kx = 0; kx < std::min(w, kB); kx++)
    for (int ky = 0; ky < std::min(h, kB); ky++)
      f.push_back({(double)kx * kx / ((double)w * w) + (double)ky * ky / ((double)h * h), kx * kB + ky});
  std::sort(f.begin(), f.end());
  CosModes m{};
  for (int j = 0; j < kB; j++) {  // a grid of fewer than 32 modes repeats them (the hashed part differs)
    m.kx[j] = f[j % f.size()].second / kB;
    m.ky[j] = f[j % f.size()].second % kB;
  }
  return m;
}

// Partial Gram matrices over one chunk of pixels, FP64: G = Y'Y, H = Y'W and
// C = X0'Y (entry [i][j] = U_i . V_j), one block per chunk.  The chunk's
// Y, W and X0 are staged in LDS once per round of kGramPx pixels and feed
// all three products.  Thread t owns row i = t / 8 and columns 4 (t % 8) ..
// +3 of each; every entry sums its pixels in order with one fused
// multiply-add each.  (Splitting a round's pixels over two threads per
// entry runs 10 % faster but changes the sums' rounding, and with it the
// iteration at which the stop test passes on some images.)
__global__ __launch_bounds__(256) void k_cheb_gram(int n, size_t vs, int chunk, const float4* __restrict__ Y,
                                                   const float4* __restrict__ Wb, const float4* __restrict__ X0,
                                                   double* __restrict__ part) {
  __shared__ float4 sy[kGramPx][kPlanes], sw[kGramPx][kPlanes], sx[kGramPx][kPlanes];
  const int i = threadIdx.x >> 3, j4 = threadIdx.x & 7;
  double g[4] = {0.0, 0.0, 0.0, 0.0}, hh[4] = {0.0, 0.0, 0.0, 0.0}, cc[4] = {0.0, 0.0, 0.0, 0.0};
  const int p0 = blockIdx.x * chunk, p1 = min(n, p0 + chunk);
  for (int base = p0; base < p1; base += kGramPx) {
    for (int s = threadIdx.x; s < kGramPx * kPlanes; s += 256) {
      const int px = s % kGramPx, pl = s / kGramPx, p = base + px;
      const bool in = p < p1;
      const size_t at = (size_t)pl * vs + p;
      const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
      sy[px][pl] = in ? Y[at] : zero;
      sw[px][pl] = in ? Wb[at] : zero;
      sx[px][pl] = in ? X0[at] : zero;
    }
    __syncthreads();
    const float* fy = &sy[0][0].x;
    const float* fx = &sx[0][0].x;
#pragma unroll 4
    for (int px = 0; px < kGramPx; px++) {
      const double yi = fy[px * kB + i], xi = fx[px * kB + i];
      const float4 yv = sy[px][j4], wv = sw[px][j4];
      g[0] = fma(yi, (double)yv.x, g[0]);
      g[1] = fma(yi, (double)yv.y, g[1]);
      g[2] = fma(yi, (double)yv.z, g[2]);
      g[3] = fma(yi, (double)yv.w, g[3]);
      hh[0] = fma(yi, (double)wv.x, hh[0]);
      hh[1] = fma(yi, (double)wv.y, hh[1]);
      hh[2] = fma(yi, (double)wv.z, hh[2]);
      hh[3] = fma(yi, (double)wv.w, hh[3]);
      cc[0] = fma(xi, (double)yv.x, cc[0]);
      cc[1] = fma(xi, (double)yv.y, cc[1]);
      cc[2] = fma(xi, (double)yv.z, cc[2]);
      cc[3] = fma(xi, (double)yv.w, cc[3]);
    }
    __syncthreads();
  }
  const size_t at = (size_t)blockIdx.x * (kB * kB) + i * kB + j4 * 4, stride = (size_t)gridDim.x * (kB * kB);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    part[at + k] = g[k];
    part[stride + at + k] = hh[k];
    part[2 * stride + at + k] = cc[k];
  }
}

// Residuals of the wanted Ritz pairs, formed directly: r_j = W s_j -
// theta_j Y s_j per pixel in FP64 (S from the rotation buffer, row i =
// input vector i), |r_j|^2 summed per block of 256 pixels into part[blk][j];
// the same pass writes the rotated block Y S (nout < 0: all kB vectors to
// `planes`, the next filter input; nout > 0: vectors 0..nout-1 to out[j * n
// + p], the solver's output), FP64 sums over the input vectors in order.
// The Gram form s'K s - theta^2 cancels: W is stored in FP32, so K carries
// a relative rounding of about 2^-23 and the difference keeps an error of
// about theta * 2^-11.5, which at 61x37 (theta near 1e-2) exceeds the
// residual itself; the direct form errs by about 2^-24 theta |y|.
constexpr int kResidMax = kB / 2;
__global__ __launch_bounds__(256) void k_cheb_resid_rotate(int n, size_t vs, const float4* __restrict__ Y,
                                                           const float4* __restrict__ Wb, const double* __restrict__ S,
                                                           const double* __restrict__ theta, int nvec,
                                                           double* __restrict__ part, int nout,
                                                           float4* __restrict__ planes, float* __restrict__ out) {
  // S is the same for every pixel: its entries are read with scalar loads
  // (wave-uniform addresses) straight into the FMAs, not staged in LDS
  __shared__ double th[kResidMax];
  __shared__ double red[kResidMax][4];
  if (threadIdx.x < kResidMax) th[threadIdx.x] = threadIdx.x < nvec ? theta[threadIdx.x] : 0.0;
  __syncthreads();
  const int p = blockIdx.x * 256 + threadIdx.x;
  double r2[kResidMax];
#pragma unroll
  for (int j = 0; j < kResidMax; j++) r2[j] = 0.0;
  if (p < n) {
    float yv[kB], wv[kB];
#pragma unroll
    for (int pl = 0; pl < kPlanes; pl++) {
      const float4 a = Y[(size_t)pl * vs + p], b = Wb[(size_t)pl * vs + p];
      yv[pl * 4 + 0] = a.x, yv[pl * 4 + 1] = a.y, yv[pl * 4 + 2] = a.z, yv[pl * 4 + 3] = a.w;
      wv[pl * 4 + 0] = b.x, wv[pl * 4 + 1] = b.y, wv[pl * 4 + 2] = b.z, wv[pl * 4 + 3] = b.w;
    }
    // Y s_j for every j (the rotated block), W s_j for the wanted ones;
    // each sum runs over the input vectors in order
    double ys[kResidMax];
    const int nj = nout < 0 ? kB : nout;
#pragma unroll
    for (int j0 = 0; j0 < kB; j0 += 4) {
      if (j0 >= nj && j0 >= nvec) break;
      double o[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
      for (int i = 0; i < kB; i++) {
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = fma((double)yv[i], S[i * kB + j0 + k], o[k]);
      }
      if (j0 < kResidMax) {
#pragma unroll
        for (int k = 0; k < 4; k++) ys[j0 + k] = o[k];
      }
      if (nout < 0) {
        planes[(size_t)(j0 / 4) * vs + p] = make_float4((float)o[0], (float)o[1], (float)o[2], (float)o[3]);
      } else {
        for (int k = 0; k < 4 && j0 + k < nout; k++) out[(size_t)(j0 + k) * n + p] = (float)o[k];
      }
    }
#pragma unroll
    for (int j = 0; j < kResidMax; j++) {
      if (j >= nvec) break;
      double ws = 0.0;
#pragma unroll 8
      for (int i = 0; i < kB; i++) ws = fma((double)wv[i], S[i * kB + j], ws);
      const double r = fma(-th[j], ys[j], ws);
      r2[j] = r * r;
    }
  }
  const int lane = threadIdx.x & 63, wvi = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < kResidMax; j++) {
    double v = r2[j];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[j][wvi] = v;
  }
  __syncthreads();
  if (threadIdx.x < kResidMax)
    part[(size_t)blockIdx.x * kResidMax + threadIdx.x] =
        ((red[threadIdx.x][0] + red[threadIdx.x][1]) + red[threadIdx.x][2]) + red[threadIdx.x][3];
}

// Residual norms from k_cheb_resid_rotate's partials: block j sums column j, each
// thread a strided subset, then the threads' sums in a fixed tree order.
__global__ __launch_bounds__(256) void k_cheb_resid_sum(int nblk, const double* __restrict__ part,
                                                        double* __restrict__ out) {
  __shared__ double red[4];
  const int j = blockIdx.x;
  double v = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 256) v += part[(size_t)b * kResidMax + j];
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) out[j] = sqrt(((red[0] + red[1]) + red[2]) + red[3]);
}

// Sums the partials of each Gram entry: a block covers 64 entries with 16
// threads each; thread j sums the partials of blocks j, j + 16, ... in
// order, then the 16 sums are added in a fixed tree (deterministic).
constexpr int kGramSumSplit = 16;
__global__ __launch_bounds__(1024) void k_cheb_gram_sum(int nblk, const double* __restrict__ part,
                                                        double* __restrict__ gram) {
  __shared__ double red[kGramSumSplit][64];
  const int el = threadIdx.x & 63, j = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + el;
  double sum = 0.0;
  if (e < kGrams * kB * kB) {
    const int which = e / (kB * kB), idx = e % (kB * kB);
    const double* p = part + (size_t)which * nblk * (kB * kB) + idx;
#pragma unroll 4
    for (int b = j; b < nblk; b += kGramSumSplit) sum += p[(size_t)b * (kB * kB)];
  }
  red[j][el] = sum;
  __syncthreads();
  if (j == 0 && e < kGrams * kB * kB) {
    double t[kGramSumSplit];
#pragma unroll
    for (int k = 0; k < kGramSumSplit; k++) t[k] = red[k][el];
#pragma unroll
    for (int w = kGramSumSplit / 2; w >= 1; w >>= 1)
#pragma unroll
      for (int k = 0; k < w; k++) t[k] += t[k + w];
    gram[e] = t[0];
  }
}

template <int R, int VP, bool F64 = false>
int launch_step(hipStream_t s, int w, int h, const float* A, size_t pitch, const float4* x,
                const float4* z, float4* y, size_t vs, double alpha, double c, double beta) {
  const int ntx = div_up(w, 64), nty = div_up(h, R);
  hipLaunchKernelGGL((k_cheb_step<R, VP, F64>), dim3(ntx * nty), dim3(64 * (kPlanes / VP)), 0, s, w, h, ntx, A, pitch,
                     x, z, y, vs, (float)alpha, (float)c, (float)beta);
  return DMSC_OK;
}

template <int R, int NW, bool F64 = false, bool PF = false>
int launch_lds(hipStream_t s, int w, int h, const float* A, size_t pitch, const float4* x, const float4* z,
               float4* y, size_t vs, double alpha, double c, double beta) {
  // the pipelined form addresses the matrix through one buffer descriptor
  // whose range (bytes) is a 31-bit count
  if (PF && (size_t)kStencilArea * pitch * sizeof(float) >= ((size_t)1 << 31))
    return launch_lds<R, NW, F64, false>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  const int ntx = div_up(w, 64), nty = div_up(h, R * NW);
  hipLaunchKernelGGL((k_cheb_lds<R, NW, F64, PF>), dim3(ntx * nty * 4), dim3(64 * NW), 0, s, w, h, ntx, A, pitch, x,
                     z, y, vs, (float)alpha, (float)c, (float)beta);
  return DMSC_OK;
}

template <int R, int NW>
int launch_ldp(hipStream_t s, int w, int h, const float* A, size_t pitch, const float4* x, const float4* z, float4* y,
               size_t vs, double alpha, double c, double beta) {
  if ((size_t)kStencilArea * pitch * sizeof(float) >= ((size_t)1 << 31))  // one buffer descriptor's range
    return launch_lds<R, NW, false, false>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  const int ntx = div_up(w, 64), nty = div_up(h, R * NW);
  hipLaunchKernelGGL((k_cheb_ldp<R, NW>), dim3(ntx * nty * 4), dim3(64 * NW), 0, s, w, h, ntx, A, pitch, x, z, y, vs,
                     (float)alpha, (float)c, (float)beta);
  return DMSC_OK;
}

template <int R, int NW>
int launch_ldp64(hipStream_t s, int w, int h, const float* A, size_t pitch, const float4* x, float4* y, size_t vs) {
  if ((size_t)kStencilArea * pitch * sizeof(float) >= ((size_t)1 << 31))  // one buffer descriptor's range
    return launch_lds<R, NW, true, false>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
  const int ntx = div_up(w, 64), nty = div_up(h, R * NW);
  hipLaunchKernelGGL((k_cheb_ldp64<R, NW>), dim3(ntx * nty * 4), dim3(64 * NW), 0, s, w, h, ntx, A, pitch, x,
                     nullptr, y, vs, 1.0f, 0.0f, 0.0f);
  return DMSC_OK;
}

int step(int shape, hipStream_t s, int w, int h, const float* A, size_t pitch, const float4* x, const float4* z, float4* y, size_t vs, double alpha, double c, double beta) {
  if (shape == 2324) return launch_ldp<2, 4>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 2328) return launch_ldp<2, 8>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 2337) return launch_ldp<3, 7>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 324) return launch_lds<2, 4>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 344) return launch_lds<4, 4>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 342) return launch_lds<4, 2>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 328) return launch_lds<2, 8>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 348) return launch_lds<4, 8>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 1324) return launch_lds<2, 4, false, true>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 1328) return launch_lds<2, 8, false, true>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 1344) return launch_lds<4, 4, false, true>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 1348) return launch_lds<4, 8, false, true>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 1336) return launch_lds<3, 6, false, true>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 1337) return launch_lds<3, 7, false, true>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 42) return launch_step<4, 2>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 24) return launch_step<2, 4>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 28) return launch_step<2, 8>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  return launch_step<2, 2>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
}

// W = A Y with FP64 sums on the filter's tile (bit-identical across tiles:
// every tile sums in the same order).
int product_f64(int shape, hipStream_t s, int w, int h, const float* A, size_t pitch, const float4* x, float4* y,
                size_t vs) {
  if (shape == 324) return launch_lds<2, 4, true>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
  if (shape == 344) return launch_lds<4, 4, true>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
  if (shape == 342) return launch_lds<4, 2, true>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
  if (shape == 328) return launch_lds<2, 8, true>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
  if (shape == 348 || shape == 1348) return launch_lds<4, 8, true>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
  if (shape == 1324) return launch_lds<2, 4, true>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
  if (shape == 1328) return launch_lds<2, 8, true>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
  if (shape == 1344) return launch_lds<4, 4, true>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
  if (shape == 1336) return launch_lds<3, 6, true>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
  if (shape == 1337) return launch_lds<3, 7, true>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
  if (shape == 2337) return launch_ldp64<3, 7>(s, w, h, A, pitch, x, y, vs);
  if (shape == 2324) return launch_ldp64<2, 4>(s, w, h, A, pitch, x, y, vs);
  if (shape == 2328) return launch_ldp64<2, 8>(s, w, h, A, pitch, x, y, vs);
  return launch_step<2, 4, true>(s, w, h, A, pitch, x, nullptr, y, vs, 1.0, 0.0, 0.0);
}

}  // namespace

namespace {
// filter tiles chosen by dmsc_autotune_chebyshev, per image size
std::map<std::pair<int, int>, int>& tuned_filter() {
  static std::map<std::pair<int, int>, int> t;
  return t;
}
std::mutex& tuned_filter_lock() {
  static std::mutex m;
  return m;
}
}  // namespace

namespace {
// Blocks of a filter kernel the device holds at once: CUs x resident blocks
// per CU (LDS bounds both default tiles to two per CU on MI355X).
template <typename K>
int block_slots(int device, K kernel, int threads) {
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess)
    return 0;
  return cus * per_cu;
}
}  // namespace

// The filter tile used at w x h: the autotuner's pick for that size, else the
// tuning key's; with the key at 0, tile 2337 (64 x 21 pixels, 7 waves of 3
// rows, LDS rows read ahead) when its grid fits in one round of the device's
// block slots and the 64 x 16 tile's (8 waves of 2 rows) does not, else
// 2328.  At 481x321 the 64 x 21 tile runs 512 blocks in one round of 512
// slots where the 64 x 16 one runs 672 in 1.31 rounds, the second mostly
// idle (25.1 against 26.9 us per step); where both need several rounds the
// shorter blocks balance better.  Every tile gives bit-identical results.
int chebyshev_shape_for(int device, int w, int h) {
  std::lock_guard<std::mutex> lk(tuned_filter_lock());
  auto it = tuned_filter().find(std::make_pair(w, h));
  if (it != tuned_filter().end()) return it->second;
  if (chebyshev_shape) return chebyshev_shape;
  static std::map<int, std::pair<int, int>> slots;  // device -> (tile 2328, tile 2337)
  auto sl = slots.find(device);
  if (sl == slots.end())
    sl = slots.emplace(device, std::make_pair(block_slots(device, k_cheb_ldp<2, 8>, 64 * 8),
                                              block_slots(device, k_cheb_ldp<3, 7>, 64 * 7))).first;
  const int ntx = div_up(w, 64);
  const long b1328 = 4L * ntx * div_up(h, 16), b1337 = 4L * ntx * div_up(h, 21);
  return (b1337 <= sl->second.second && b1328 > sl->second.first) ? 2337 : 2328;
}

void chebyshev_autotune_clear() {
  std::lock_guard<std::mutex> lk(tuned_filter_lock());
  tuned_filter().clear();
}

int chebyshev_max_vectors() { return kB / 2; }

bool chebyshev_shape_valid(int v) {
  return v == 22 || v == 42 || v == 24 || v == 28 || v == 324 || v == 344 || v == 342 || v == 328 || v == 348 ||
         v == 1324 || v == 1328 || v == 1344 || v == 1348 || v == 1336 || v == 1337 || v == 2324 || v == 2328 ||
         v == 2337;
}

// The nvec smallest eigenpairs of L_sym (A, already converted): unit
// eigenvectors to d_out[j * n + p], eigenvalues to vals; *products = the
// number of block products with A.
int chebyshev_solve(Ctx* ctx, int w, int h, const float* A, int pitch, int nvec, float tol, float* d_out,
                    double* vals, int* products) {
  if (nvec < 1 || nvec > kB / 2) return fail(DMSC_ERR_ARG, "the block eigensolver takes nvec in [1, %d]", kB / 2);
  hipStream_t s = ctx->stream;
  const int n = w * h;
  const size_t vs = (size_t)pitch_floats(n);
  const int chunk = std::max(kGramPx, div_up(div_up(n, 512), kGramPx) * kGramPx);
  const int nblk = div_up(n, chunk);
  const size_t blk = (size_t)kPlanes * vs * sizeof(float4);
  const int nrb = div_up(n, 256);  // k_cheb_resid_rotate blocks
  const size_t o_part = 4 * blk, o_gram = o_part + (size_t)kGrams * nblk * kB * kB * sizeof(double),
               o_rot = o_gram + (size_t)kGrams * kB * kB * sizeof(double),
               o_theta = o_rot + (size_t)kB * kB * sizeof(double),
               o_res = o_theta + (size_t)kResidMax * sizeof(double),
               o_rpart = o_res + (size_t)kResidMax * sizeof(double),
               bytes = o_rpart + (size_t)nrb * kResidMax * sizeof(double);
  void* base;
  DMSC_TRY(ctx->scratch(3, bytes, &base));
  char* Bp = (char*)base;
  float4* X0 = (float4*)Bp;
  float4* ring[3] = {(float4*)(Bp + blk), (float4*)(Bp + 2 * blk), (float4*)(Bp + 3 * blk)};
  double* part = (double*)(Bp + o_part);
  double* gram = (double*)(Bp + o_gram);
  double* rot = (double*)(Bp + o_rot);
  double* d_theta = (double*)(Bp + o_theta);
  double* d_res = (double*)(Bp + o_res);
  double* rpart = (double*)(Bp + o_rpart);
  int* pin;
  DMSC_TRY(ctx->pinned(&pin));
  double* h_gram = (double*)((char*)pin + kPinGram);
  double* h_rot = (double*)((char*)pin + kPinRot);
  double* h_theta = (double*)((char*)pin + kPinTheta);
  double* h_res = (double*)((char*)pin + kPinRes);

  const StencilOffsets so = make_stencil_offsets(w);
  for (int d = 0; d < kStencilArea; d++)  // the filter kernel's closed-form diagonal order
    if (std::abs((int)so.dx[d]) > 5 || std::abs((int)so.dy[d]) > half_height(so.dx[d]) ||
        disc_index(so.dy[d], so.dx[d]) != d)
      return fail(DMSC_ERR_INTERNAL, "block eigensolver: stencil diagonal %d at (%d, %d) is not in raster disc order", d,
                  so.dy[d], so.dx[d]);
  const int shape = chebyshev_shape_for(ctx->device, w, h);
  const size_t pitchz = (size_t)pitch;

  if (chebyshev_start == 1)
    hipLaunchKernelGGL(k_cheb_init_cos, dim3(div_up((long)kPlanes * vs, 256)), dim3(256), 0, s, w, n, vs,
                       cos_modes(w, h), X0);
  else
    hipLaunchKernelGGL(k_cheb_init, dim3(div_up((long)kPlanes * vs, 256)), dim3(256), 0, s, n, vs, X0);
  std::vector<double> theta, S;
  const double bound = 0.03 * (double)tol;  // sin of the largest per-vector move between iterations
  double cut = kFirstCut;
  int deg = chebyshev_first_degree, total = 0, flat = 0;
  double best = HUGE_VAL;
  for (int outer = 0; outer < kMaxOuter; outer++) {
    // scaled Chebyshev filter of degree deg damping [cut, kUpper], normalised at 0
    const double e = (kUpper - cut) / 2, c = (kUpper + cut) / 2;
    double sigma = e / (0.0 - c);
    const double tau = 2 / sigma;
    DMSC_TRY(step(shape, s, w, h, A, pitchz, X0, nullptr, ring[0], vs, sigma / e, c, 0.0));
    int cur = 0;
    const float4* prev = X0;
    for (int i = 2; i <= deg; i++) {
      const double sn = 1 / (tau - sigma);
      const int nx = (cur + 1) % 3;
      DMSC_TRY(step(shape, s, w, h, A, pitchz, ring[cur], prev, ring[nx], vs, 2 * sn / e, c, sigma * sn));
      prev = ring[cur];
      cur = nx;
      sigma = sn;
    }
    float4* Y = ring[cur];
    float4* Wb = ring[(cur + 1) % 3];
    DMSC_TRY(product_f64(shape, s, w, h, A, pitchz, Y, Wb, vs));
    hipLaunchKernelGGL(k_cheb_gram, dim3(nblk), dim3(256), 0, s, n, vs, chunk, Y, Wb, X0, part);
    hipLaunchKernelGGL(k_cheb_gram_sum, dim3(div_up(kGrams * kB * kB, 64)), dim3(1024), 0, s, nblk, part, gram);
    DMSC_HIP(hipMemcpyAsync(h_gram, gram, kGrams * kB * kB * sizeof(double), hipMemcpyDeviceToHost, s));
    // one host round trip per outer iteration: the Gram matrices of this
    // block and the residuals of the previous Ritz pairs (enqueued last
    // iteration, read here)
    DMSC_HIP(hipStreamSynchronize(s));
    const auto t_rr = std::chrono::steady_clock::now();
    if (!rayleigh_ritz(h_gram, h_gram + kB * kB, kB, theta, S))
      return fail(DMSC_ERR_INTERNAL, "block eigensolver: filtered block lost rank at outer iteration %d", outer);
    const double rr_us = std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t_rr).count();
    double resid_prev = HUGE_VAL;  // the previous outer iteration's Ritz pairs
    if (outer > 0) {
      resid_prev = 0.0;
      for (int j = 0; j < nvec; j++) resid_prev = std::max(resid_prev, h_res[j]);
    }
    double move = 0.0;
    if (outer > 0) {
      const double* C = h_gram + 2 * kB * kB;
      for (int j = 0; j < nvec; j++) {
        double cj = 0.0;
        for (int i = 0; i < kB; i++) cj += C[j * kB + i] * S[i * kB + j];
        move = std::max(move, std::sqrt(std::max(0.0, 2.0 * (1.0 - std::fabs(cj)))));
      }
    }
    if (lanczos_trace)
      std::fprintf(stderr,
                   "[chebyshev] outer %d degree %d cut %.4e theta %.6e .. %.6e move %.3e previous residual %.3e "
                   "(Rayleigh-Ritz on the host %.0f us)\n",
                   outer, deg, cut, theta[0], theta[nvec - 1], move, resid_prev, rr_us);
    // done when the wanted vectors stopped moving, or (a degenerate
    // eigenvalue leaves its vectors free to turn) when every residual of the
    // previous Ritz pairs met the mixed Lanczos bound, tol *
    // 10^-lanczos_mixed_stop (this iteration's pairs are at least as good);
    // or when those residuals stagnated (three outer iterations without
    // improving on the best by 10 %) at the floor that FP32 storage of the
    // block sets allows: a stored entry carries a relative rounding of u =
    // 2^-24, so a Ritz vector's residual cannot fall below about
    // u |A - theta I| <= 2u (L_sym's spectrum lies in [0, 2]); kFloorMargin
    // covers the rounding the filter steps accumulate (61x37 image: the
    // residuals flatten at 1.5e-7).  That stop is reported to the caller as
    // DMSC_STOP_FLOOR.
    if (outer > 0) {
      flat = resid_prev < 0.9 * best ? 0 : flat + 1;
      best = std::min(best, resid_prev);
    }
    const double target = (double)tol * std::pow(10.0, -lanczos_mixed_stop);
    const double floor32 = kFloorMargin * 2.0 * std::ldexp(1.0, -24);
    int rule = 0;
    if (outer > 0) {
      if (move <= bound) rule = DMSC_STOP_STEADY;
      else if (resid_prev <= target) rule = DMSC_STOP_RESIDUAL;
      else if (flat >= 3 && resid_prev <= floor32) rule = DMSC_STOP_FLOOR;
    }
    // the rotation S and the wanted Ritz values go to the device either way
    std::memcpy(h_rot, S.data(), sizeof(double) * kB * kB);
    for (int j = 0; j < nvec; j++) h_theta[j] = theta[j];
    DMSC_HIP(hipMemcpyAsync(rot, h_rot, sizeof(double) * kB * kB, hipMemcpyHostToDevice, s));
    DMSC_HIP(hipMemcpyAsync(d_theta, h_theta, sizeof(double) * nvec, hipMemcpyHostToDevice, s));
    // residual |A y_j - theta_j y_j| of each wanted Ritz vector y_j = Y s_j,
    // formed per pixel on the device by the pass that writes the rotated
    // block (the next filter input, or the output when the solve stops
    // here); read back at the next round trip (or below, at a stop)
    hipLaunchKernelGGL(k_cheb_resid_rotate, dim3(nrb), dim3(256), 0, s, n, vs, Y, Wb, rot, d_theta, nvec, rpart,
                       rule ? nvec : -1, rule ? nullptr : X0, rule ? d_out : nullptr);
    hipLaunchKernelGGL(k_cheb_resid_sum, dim3(nvec), dim3(256), 0, s, nrb, rpart, d_res);
    DMSC_HIP(hipMemcpyAsync(h_res, d_res, sizeof(double) * nvec, hipMemcpyDeviceToHost, s));
    if (rule) {
      DMSC_HIP(hipGetLastError());
      DMSC_HIP(hipStreamSynchronize(s));
      double resid = 0.0;
      for (int j = 0; j < nvec; j++) resid = std::max(resid, h_res[j]);
      for (int j = 0; j < nvec; j++) vals[j] = theta[j];
      if (products) *products = total + deg + 1;
      ctx->eig_stop = rule;
      ctx->eig_residual = resid;
      return DMSC_OK;
    }
    DMSC_HIP(hipGetLastError());
    total += deg + 1;
    cut = std::max(theta[kB - 1], 1e-6);
    // next degree: a tenfold reduction of the slowest wanted vector's error,
    // whose components grow by exp(acosh(t)) per degree against the damped
    // interval (t = its distance from the interval's centre in half-widths)
    const double tj = ((kUpper + cut) / 2 - theta[nvec - 1]) / ((kUpper - cut) / 2);
    deg = tj > 1.0 ? (int)std::ceil(std::log((double)chebyshev_reduction) / std::acosh(tj)) : kMaxDeg;
    deg = std::min(kMaxDeg, std::max(kMinDeg, deg));
  }
  return fail(DMSC_ERR_NOCONV, "block eigensolver did not converge in %d outer iterations", kMaxOuter);
}

// Times `products` filter steps (after 5 untimed ones) of the current
// filter tile on the converted matrix A with HIP events on the context's
// stream: the kernel the block eigensolver spends its time in, on a block
// of kB vectors.  Returns milliseconds per step.
int time_chebyshev_products(Ctx* ctx, int w, int h, const float* A, int pitch, int products, float* ms_per_product,
                            int shape) {
  if (shape == 0) shape = chebyshev_shape_for(ctx->device, w, h);
  if (!chebyshev_shape_valid(shape)) return fail(DMSC_ERR_ARG, "unknown filter tile %d", shape);
  if (products < 1 || products > 100000) return fail(DMSC_ERR_ARG, "products out of range");
  hipStream_t s = ctx->stream;
  const int n = w * h;
  const size_t vs = (size_t)pitch_floats(n);
  const size_t blk = (size_t)kPlanes * vs * sizeof(float4);
  void* base;
  DMSC_TRY(ctx->scratch(3, 3 * blk, &base));
  float4* ring[3] = {(float4*)base, (float4*)((char*)base + blk), (float4*)((char*)base + 2 * blk)};
  for (int i = 0; i < 3; i++)
    hipLaunchKernelGGL(k_cheb_init, dim3(div_up((long)kPlanes * vs, 256)), dim3(256), 0, s, n, vs, ring[i]);
  // the solver's first filter (damped interval [kFirstCut, 2]), so the
  // block stays O(1) over any number of steps
  const double e = (kUpper - kFirstCut) / 2, c = (kUpper + kFirstCut) / 2;
  double sigma = e / (0.0 - c);
  const double tau = 2 / sigma;
  hipEvent_t ev[2];
  DMSC_HIP(hipEventCreate(&ev[0]));
  DMSC_HIP(hipEventCreate(&ev[1]));
  int rc = step(shape, s, w, h, A, (size_t)pitch, ring[0], nullptr, ring[1], vs, sigma / e, c, 0.0);
  for (int i = 1; i < products + 6 && rc == DMSC_OK; i++) {
    if (i == 6) rc = hipEventRecord(ev[0], s) == hipSuccess ? DMSC_OK : fail(DMSC_ERR_HIP, "event record failed");
    const double sn = 1 / (tau - sigma);
    if (rc == DMSC_OK)
      rc = step(shape, s, w, h, A, (size_t)pitch, ring[i % 3], ring[(i + 2) % 3], ring[(i + 1) % 3], vs,
                2 * sn / e, c, sigma * sn);
    sigma = sn;
  }
  float ms = 0.f;
  if (rc == DMSC_OK && hipEventRecord(ev[1], s) == hipSuccess && hipEventSynchronize(ev[1]) == hipSuccess &&
      hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) {
    *ms_per_product = ms / (float)products;
  } else if (rc == DMSC_OK) {
    rc = fail(DMSC_ERR_HIP, "filter timing failed");
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  DMSC_HIP(hipGetLastError());
  return rc;
}

// Times every filter tile at w x h (`products` steps each, on the converted
// matrix) and makes the fastest that size's tile for later solves.
int autotune_chebyshev(Ctx* ctx, int w, int h, const float* A, int pitch, int products, int* best, int* shapes,
                       float* us, int max_tried, int* n_tried) {
  static const int kShapes[] = {1328, 1324, 1336, 1337, 2328, 2324, 2337, 1344, 1348, 328, 324, 344, 342, 348,
                                24, 28, 22, 42};
  int nt = 0, top = 0;
  float top_us = 0.f;
  for (int sh : kShapes) {
    float ms = 0.f;
    DMSC_TRY(time_chebyshev_products(ctx, w, h, A, pitch, products, &ms, sh));
    if (shapes && nt < max_tried) {
      shapes[nt] = sh;
      us[nt] = ms * 1e3f;
    }
    nt++;
    if (top == 0 || ms * 1e3f < top_us) {
      top = sh;
      top_us = ms * 1e3f;
    }
  }
  {
    std::lock_guard<std::mutex> lk(tuned_filter_lock());
    tuned_filter()[std::make_pair(w, h)] = top;
  }
  if (best) *best = top;
  if (n_tried) *n_tried = nt;
  return DMSC_OK;
}

}  // namespace dmsc
