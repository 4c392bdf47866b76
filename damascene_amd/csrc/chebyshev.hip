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
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {

int eigensolver_method = 0;  // mixed path: 0 Lanczos, 1 Chebyshev block (tune key "eigensolver")
int chebyshev_shape = 24;    // filter tile: rows per lane * 10 + float4 planes per lane (24, 28, 22, 42)

namespace {

constexpr int kB = 32;             // block vectors
constexpr int kPlanes = kB / 4;    // float4 planes
constexpr int kDeg = 50;           // filter degree of the first outer iteration
constexpr int kMinDeg = 30, kMaxDeg = 300;
constexpr int kMaxOuter = 60;
constexpr double kFirstCut = 1e-2;  // damped interval of the first filter: [kFirstCut, 2]
constexpr double kUpper = 2.0;      // spectrum of L_sym lies in [0, 2]
constexpr int kGramPx = 64;         // pixels staged per Gram round
// pinned host block: Gram matrices at [8 KiB, 40 KiB), the rotation at [40 KiB, 48 KiB)
constexpr size_t kPinGram = 8192, kPinRot = 40960;
constexpr int kGrams = 4;  // Y'Y, Y'W, X0'Y, W'W
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

// Starting block: uniform values in [-1, 1) from an integer hash of
// (plane, pixel, component); rows n..vs-1 are 0.
__global__ void k_cheb_init(int n, size_t vs, float4* __restrict__ x) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)kPlanes * vs) return;
  const int p = (int)(i % vs);
  float e[4];
  for (int k = 0; k < 4; k++) {
    unsigned u = (unsigned)(i * 4 + k) * 0x9E3779B1u + 0x7F4A7C15u;
    u ^= u >> 16;
    u *= 0x85EBCA6Bu;
    u ^= u >> 13;
    u *= 0xC2B2AE35u;
    u ^= u >> 16;
    e[k] = p < n ? (float)(u >> 8) * (1.0f / 8388608.0f) - 1.0f : 0.0f;
  }
  x[i] = make_float4(e[0], e[1], e[2], e[3]);
}

// Partial Gram matrices over one chunk of pixels, FP64: blockIdx.y = 0
// G = Y'Y, 1 H = Y'W, 2 C = X0'Y, 3 K = W'W (entry [i][j] = U_i . V_j).  Thread t
// owns row i = t / 8 and columns 4 (t % 8) .. +3.
__global__ __launch_bounds__(256) void k_cheb_gram(int n, size_t vs, int chunk, const float4* __restrict__ Y,
                                                   const float4* __restrict__ Wb, const float4* __restrict__ X0,
                                                   double* __restrict__ part) {
  __shared__ float su[kGramPx][kB + 1];
  __shared__ float4 sv[kGramPx][kPlanes];
  const int which = blockIdx.y;
  const float4* U = which == 2 ? X0 : which == 3 ? Wb : Y;
  const float4* V = which == 1 || which == 3 ? Wb : Y;
  const int i = threadIdx.x >> 3, j4 = threadIdx.x & 7;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int p0 = blockIdx.x * chunk, p1 = min(n, p0 + chunk);
  for (int base = p0; base < p1; base += kGramPx) {
    for (int s = threadIdx.x; s < kGramPx * kPlanes; s += 256) {
      const int px = s % kGramPx, pl = s / kGramPx, p = base + px;
      const bool in = p < p1;
      const float4 u = in ? U[(size_t)pl * vs + p] : make_float4(0.f, 0.f, 0.f, 0.f);
      su[px][pl * 4 + 0] = u.x;
      su[px][pl * 4 + 1] = u.y;
      su[px][pl * 4 + 2] = u.z;
      su[px][pl * 4 + 3] = u.w;
      sv[px][pl] = in ? V[(size_t)pl * vs + p] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
#pragma unroll 8
    for (int px = 0; px < kGramPx; px++) {
      const double a = su[px][i];
      const float4 b = sv[px][j4];
      acc[0] = fma(a, (double)b.x, acc[0]);
      acc[1] = fma(a, (double)b.y, acc[1]);
      acc[2] = fma(a, (double)b.z, acc[2]);
      acc[3] = fma(a, (double)b.w, acc[3]);
    }
    __syncthreads();
  }
  double* o = part + ((size_t)which * gridDim.x + blockIdx.x) * (kB * kB) + i * kB + j4 * 4;
#pragma unroll
  for (int k = 0; k < 4; k++) o[k] = acc[k];
}

// Sums the partials of each Gram entry in block order (deterministic).
__global__ void k_cheb_gram_sum(int nblk, const double* __restrict__ part, double* __restrict__ gram) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kGrams * kB * kB) return;
  const int which = e / (kB * kB), idx = e % (kB * kB);
  const double* p = part + (size_t)which * nblk * (kB * kB) + idx;
  double s = 0.0;
#pragma unroll 16
  for (int b = 0; b < nblk; b++) s += p[(size_t)b * (kB * kB)];
  gram[e] = s;
}

// Rotation by S (kB x kB, FP64, row i = input vector i), FP64 sums: nout < 0
// writes the whole block to `planes`; nout > 0 writes vectors 0..nout-1 to
// out[j * n + p] (the solver's output layout).
__global__ __launch_bounds__(256) void k_cheb_rotate(int n, size_t vs, const float4* __restrict__ Y,
                                                     const double* __restrict__ S, int nout,
                                                     float4* __restrict__ planes, float* __restrict__ out) {
  __shared__ double ss[kB * kB];
  for (int k = threadIdx.x; k < kB * kB; k += blockDim.x) ss[k] = S[k];
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float yv[kB];
#pragma unroll
  for (int pl = 0; pl < kPlanes; pl++) {
    const float4 v = Y[(size_t)pl * vs + p];
    yv[pl * 4 + 0] = v.x;
    yv[pl * 4 + 1] = v.y;
    yv[pl * 4 + 2] = v.z;
    yv[pl * 4 + 3] = v.w;
  }
  const int nj = nout < 0 ? kB : nout;
  for (int j0 = 0; j0 < nj; j0 += 4) {
    double o[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
    for (int i = 0; i < kB; i++) {
#pragma unroll
      for (int k = 0; k < 4; k++) o[k] = fma((double)yv[i], ss[i * kB + j0 + k], o[k]);
    }
    if (nout < 0) {
      planes[(size_t)(j0 / 4) * vs + p] = make_float4((float)o[0], (float)o[1], (float)o[2], (float)o[3]);
    } else {
      for (int k = 0; k < 4 && j0 + k < nout; k++) out[(size_t)(j0 + k) * n + p] = (float)o[k];
    }
  }
}

// ---------------------------------------------------------------- host FP64
// Cholesky G = L L' (lower, in place); false if G is not positive definite.
bool cholesky(std::vector<double>& L, int m) {
  for (int j = 0; j < m; j++) {
    double d = L[j * m + j];
    for (int k = 0; k < j; k++) d -= L[j * m + k] * L[j * m + k];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    L[j * m + j] = d;
    for (int i = j + 1; i < m; i++) {
      double s = L[i * m + j];
      for (int k = 0; k < j; k++) s -= L[i * m + k] * L[j * m + k];
      L[i * m + j] = s / d;
    }
    for (int k = j + 1; k < m; k++) L[j * m + k] = 0.0;
  }
  return true;
}

// Cyclic Jacobi on symmetric M (m x m): eigenvalues in ev, eigenvectors in
// the columns of Z.
void jacobi(std::vector<double> M, int m, std::vector<double>& ev, std::vector<double>& Z) {
  Z.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; i++) Z[i * m + i] = 1.0;
  for (int sweep = 0; sweep < 60; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++) (i == j ? tot : off) += M[i * m + j] * M[i * m + j];
    if (off <= 1e-30 * (tot + off)) break;
    for (int p = 0; p < m - 1; p++)
      for (int q = p + 1; q < m; q++) {
        const double apq = M[p * m + q];
        if (apq == 0.0) continue;
        const double th = (M[q * m + q] - M[p * m + p]) / (2.0 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < m; k++) {  // M <- J' M J
          const double mkp = M[k * m + p], mkq = M[k * m + q];
          M[k * m + p] = cs * mkp - sn * mkq;
          M[k * m + q] = sn * mkp + cs * mkq;
        }
        for (int k = 0; k < m; k++) {
          const double mpk = M[p * m + k], mqk = M[q * m + k];
          M[p * m + k] = cs * mpk - sn * mqk;
          M[q * m + k] = sn * mpk + cs * mqk;
        }
        for (int k = 0; k < m; k++) {
          const double zkp = Z[k * m + p], zkq = Z[k * m + q];
          Z[k * m + p] = cs * zkp - sn * zkq;
          Z[k * m + q] = sn * zkp + cs * zkq;
        }
      }
  }
  ev.resize(m);
  for (int i = 0; i < m; i++) ev[i] = M[i * m + i];
}

// H s = theta G s: S (row-major kB x kB, column j = s_j, ascending theta)
// with S' G S = I.  False if G is not positive definite.
bool rayleigh_ritz(const double* G, const double* H, std::vector<double>& theta, std::vector<double>& S) {
  const int m = kB;
  std::vector<double> L(G, G + m * m);
  for (int i = 0; i < m; i++)
    for (int j = 0; j < i; j++) L[i * m + j] = L[j * m + i] = 0.5 * (G[i * m + j] + G[j * m + i]);
  if (!cholesky(L, m)) return false;
  // M = L^-1 H L^-T, with H symmetrised
  std::vector<double> T((size_t)m * m), M((size_t)m * m);
  for (int c = 0; c < m; c++)  // T = L^-1 H (forward substitution per column)
    for (int i = 0; i < m; i++) {
      double s = 0.5 * (H[i * m + c] + H[c * m + i]);
      for (int k = 0; k < i; k++) s -= L[i * m + k] * T[k * m + c];
      T[i * m + c] = s / L[i * m + i];
    }
  for (int r = 0; r < m; r++)  // M' = L^-1 T' (M is symmetric)
    for (int i = 0; i < m; i++) {
      double s = T[r * m + i];
      for (int k = 0; k < i; k++) s -= L[i * m + k] * M[r * m + k];
      M[r * m + i] = s / L[i * m + i];
    }
  for (int i = 0; i < m; i++)
    for (int j = 0; j < i; j++) M[i * m + j] = M[j * m + i] = 0.5 * (M[i * m + j] + M[j * m + i]);
  std::vector<double> ev, Z;
  jacobi(M, m, ev, Z);
  std::vector<int> ord(m);
  for (int i = 0; i < m; i++) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return ev[a] < ev[b]; });
  theta.resize(m);
  S.assign((size_t)m * m, 0.0);
  for (int j = 0; j < m; j++) {
    theta[j] = ev[ord[j]];
    for (int i = m - 1; i >= 0; i--) {  // s_j = L^-T z_j (back substitution)
      double s = Z[i * m + ord[j]];
      for (int k = i + 1; k < m; k++) s -= L[k * m + i] * S[k * m + j];
      S[i * m + j] = s / L[i * m + i];
    }
  }
  return true;
}

template <int R, int VP, bool F64 = false>
int launch_step(hipStream_t s, int w, int h, const float* A, size_t pitch, const float4* x,
                const float4* z, float4* y, size_t vs, double alpha, double c, double beta) {
  const int ntx = div_up(w, 64), nty = div_up(h, R);
  hipLaunchKernelGGL((k_cheb_step<R, VP, F64>), dim3(ntx * nty), dim3(64 * (kPlanes / VP)), 0, s, w, h, ntx, A, pitch,
                     x, z, y, vs, (float)alpha, (float)c, (float)beta);
  return DMSC_OK;
}

int step(int shape, hipStream_t s, int w, int h, const float* A, size_t pitch, const float4* x, const float4* z, float4* y, size_t vs, double alpha, double c, double beta) {
  if (shape == 42) return launch_step<4, 2>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 24) return launch_step<2, 4>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  if (shape == 28) return launch_step<2, 8>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
  return launch_step<2, 2>(s, w, h, A, pitch, x, z, y, vs, alpha, c, beta);
}

}  // namespace

bool chebyshev_shape_valid(int v) { return v == 22 || v == 42 || v == 24 || v == 28; }

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
  const size_t o_part = 4 * blk, o_gram = o_part + (size_t)kGrams * nblk * kB * kB * sizeof(double),
               o_rot = o_gram + (size_t)kGrams * kB * kB * sizeof(double),
               bytes = o_rot + (size_t)kB * kB * sizeof(double);
  void* base;
  DMSC_TRY(ctx->scratch(3, bytes, &base));
  char* Bp = (char*)base;
  float4* X0 = (float4*)Bp;
  float4* ring[3] = {(float4*)(Bp + blk), (float4*)(Bp + 2 * blk), (float4*)(Bp + 3 * blk)};
  double* part = (double*)(Bp + o_part);
  double* gram = (double*)(Bp + o_gram);
  double* rot = (double*)(Bp + o_rot);
  int* pin;
  DMSC_TRY(ctx->pinned(&pin));
  double* h_gram = (double*)((char*)pin + kPinGram);
  double* h_rot = (double*)((char*)pin + kPinRot);

  const StencilOffsets so = make_stencil_offsets(w);
  for (int d = 0; d < kStencilArea; d++)  // the filter kernel's closed-form diagonal order
    if (std::abs((int)so.dx[d]) > 5 || std::abs((int)so.dy[d]) > half_height(so.dx[d]) ||
        disc_index(so.dy[d], so.dx[d]) != d)
      return fail(DMSC_ERR_INTERNAL, "block eigensolver: stencil diagonal %d at (%d, %d) is not in raster disc order", d,
                  so.dy[d], so.dx[d]);
  const int shape = chebyshev_shape;
  const size_t pitchz = (size_t)pitch;

  hipLaunchKernelGGL(k_cheb_init, dim3(div_up((long)kPlanes * vs, 256)), dim3(256), 0, s, n, vs, X0);
  std::vector<double> theta, S;
  const double bound = 0.03 * (double)tol;  // sin of the largest per-vector move between iterations
  double cut = kFirstCut;
  int deg = kDeg, total = 0, flat = 0;
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
    DMSC_TRY((launch_step<2, 4, true>(s, w, h, A, pitchz, Y, nullptr, Wb, vs, 1.0, 0.0, 0.0)));
    hipLaunchKernelGGL(k_cheb_gram, dim3(nblk, kGrams), dim3(256), 0, s, n, vs, chunk, Y, Wb, X0, part);
    hipLaunchKernelGGL(k_cheb_gram_sum, dim3(div_up(kGrams * kB * kB, 256)), dim3(256), 0, s, nblk, part, gram);
    DMSC_HIP(hipMemcpyAsync(h_gram, gram, kGrams * kB * kB * sizeof(double), hipMemcpyDeviceToHost, s));
    DMSC_HIP(hipStreamSynchronize(s));
    if (!rayleigh_ritz(h_gram, h_gram + kB * kB, theta, S))
      return fail(DMSC_ERR_INTERNAL, "block eigensolver: filtered block lost rank at outer iteration %d", outer);
    // residual |A y_j - theta_j y_j| of each wanted Ritz vector y_j = Y s_j:
    // s'K s - theta^2 (s'G s = 1, s'H s = theta), exact to FP64 rounding
    // of terms of size theta^2
    double resid = 0.0;
    {
      const double* K = h_gram + 3 * kB * kB;
      for (int j = 0; j < nvec; j++) {
        double q = 0.0;
        for (int a = 0; a < kB; a++) {
          double r = 0.0;
          for (int b = 0; b < kB; b++) r += K[a * kB + b] * S[b * kB + j];
          q += S[a * kB + j] * r;
        }
        resid = std::max(resid, std::sqrt(std::max(0.0, q - theta[j] * theta[j])));
      }
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
      std::fprintf(stderr, "[chebyshev] outer %d degree %d cut %.4e theta %.6e .. %.6e move %.3e residual %.3e\n",
                   outer, deg, cut, theta[0], theta[nvec - 1], move, resid);
    std::memcpy(h_rot, S.data(), sizeof(double) * kB * kB);
    DMSC_HIP(hipMemcpyAsync(rot, h_rot, sizeof(double) * kB * kB, hipMemcpyHostToDevice, s));
    // done when the wanted vectors stopped moving, or (a degenerate
    // eigenvalue leaves its vectors free to turn) when every residual meets
    // the mixed Lanczos bound, tol * 10^-lanczos_mixed_stop; or when the
    // residuals stagnated (three outer iterations without improving on the
    // best by 10 %) at the floor that FP32 storage of the block sets allows:
    // a stored entry carries a relative rounding of u = 2^-24, so a Ritz
    // vector's residual cannot fall below about u |A - theta I| <= 2u
    // (L_sym's spectrum lies in [0, 2]); kFloorMargin covers the rounding
    // the filter steps accumulate (61x37 image: the residuals flatten at
    // 1.5e-7).  That stop is reported to the caller as DMSC_STOP_FLOOR.
    flat = resid < 0.9 * best ? 0 : flat + 1;
    best = std::min(best, resid);
    const double target = (double)tol * std::pow(10.0, -lanczos_mixed_stop);
    const double floor32 = kFloorMargin * 2.0 * std::ldexp(1.0, -24);
    int rule = 0;
    if (outer > 0) {
      if (resid <= target) rule = DMSC_STOP_RESIDUAL;
      else if (move <= bound) rule = DMSC_STOP_STEADY;
      else if (flat >= 3 && resid <= floor32) rule = DMSC_STOP_FLOOR;
    }
    if (rule) {
      hipLaunchKernelGGL(k_cheb_rotate, dim3(div_up(n, 256)), dim3(256), 0, s, n, vs, Y, rot, nvec, nullptr, d_out);
      DMSC_HIP(hipGetLastError());
      for (int j = 0; j < nvec; j++) vals[j] = theta[j];
      if (products) *products = total + deg + 1;
      ctx->eig_stop = rule;
      ctx->eig_residual = resid;
      return DMSC_OK;
    }
    hipLaunchKernelGGL(k_cheb_rotate, dim3(div_up(n, 256)), dim3(256), 0, s, n, vs, Y, rot, -1, X0, nullptr);
    DMSC_HIP(hipGetLastError());
    total += deg + 1;
    cut = std::max(theta[kB - 1], 1e-6);
    // next degree: a tenfold reduction of the slowest wanted vector's error,
    // whose components grow by exp(acosh(t)) per degree against the damped
    // interval (t = its distance from the interval's centre in half-widths)
    const double tj = ((kUpper + cut) / 2 - theta[nvec - 1]) / ((kUpper - cut) / 2);
    deg = tj > 1.0 ? (int)std::ceil(std::log(10.0) / std::acosh(tj)) : kMaxDeg;
    deg = std::min(kMaxDeg, std::max(kMinDeg, deg));
  }
  return fail(DMSC_ERR_NOCONV, "block eigensolver did not converge in %d outer iterations", kMaxOuter);
}

}  // namespace dmsc
