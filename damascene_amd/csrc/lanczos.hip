// lanczos.hip -- GPU-resident generalized eigensolver (reference:
// noReorthog/lanczos.cu, stencilMatrixMultiply/stencilMVM.cu).
//
// (D - W) v = lambda D v is solved as L_sym u = lambda u with
// L_sym = I - D^-1/2 W D^-1/2 and v = D^-1/2 u, by Lanczos without
// reorthogonalisation from q_0 = 1/sqrt(N), with Cullum-Willoughby screening
// of spurious Ritz values (lanczos.cu:279-384).  Two precisions:
//
//  FP32 ("plain", the reference algorithm): FP32 SpMV, vectors, alpha and
//    beta; inner products accumulated in FP64 and rounded to FP32 (one
//    fixed, reproducible choice for the reference's cuBLAS sdot);
//    convergence tested every 500 steps (lanczos.cu:5, 519) with the FP32
//    residual test of lanczos.cu:114-140.
//  MIXED: FP32 matrix and FP32-stored basis, FP64 SpMV accumulation,
//    three-term vector and inner products; tested on an adaptive schedule
//    (multiples of 25 steps) and stopped when every residual
//    |beta_m s_{m,j}| is at most tol * 1e-4 (the "lanczos_mixed_stop"
//    exponent; at 481x321 a bound of tol * 1e-3 left one of three test
//    images at a subspace error of 1.7e-4, tol * 1e-4 kept all three
//    below 2.1e-5).
//
// A plain step is two kernels and no host round trip:
//   K1: beta_{k-1} from the |z|^2 partials (every block reduces them in the
//       same order, so all blocks agree bit for bit), q_k = z_{k-1}/beta_{k-1}
//       formed on the fly for the stencil gather and written to the basis,
//       z_k = A q_k - beta_{k-1} q_{k-1}, partial sums of q_k . z_k;
//   K2: alpha_k from the partials, z_k -= alpha_k q_k, partial |z_k|^2.
// A mixed step is one kernel by default (k_lzmt: pixel tiles whose basis
// values and stencil halo are staged in LDS; see k_lzmf and k_lzmt), with
// the two-kernel form k_lzm1/k_lzm2 kept as a tuning alternative.
// Each node of a chunk's graph carries its step offset and the host sets
// the chunk's first step in device memory, so a whole chunk is built once
// into a HIP graph and replayed.  At a check point one block per wanted
// value finds the FP64 Ritz values of T and of T without its first
// row/column by 256-way multisection over division-free Sturm counts, one
// block screens them, one block per Ritz pair runs inverse iteration with its
// arrays in LDS, and the host reads back a small status record.  Mixed
// checks run on a second stream while further steps are already enqueued.
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <mutex>
#include <cstdio>
#include <tuple>
#include <vector>

#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"
#include "tridiag.hpp"

namespace dmsc {

int lanczos_trace = 0;
int lanczos_k1_threads = 0;
int lanczos_mixed_fused = 2;
int lanczos_tile = 3208;
int lanczos_max_iter = kMaxIter;
int lanczos_check_max = 0;
int lanczos_mixed_stop = 4;
int lanczos_step_loads = 0;

namespace {

constexpr int kGraphSteps = 25;   // steps per graph; periods are multiples of it
constexpr int kMaxEig = 25;
constexpr int kCullumExtra = 20;  // eigCheck = nvec + 20 (lanczos.cu:456)
constexpr int kCullumMax = 1000;  // lanczos.cu:531
constexpr int kWave = 64;
// K1 (the SpMV step kernel): one stencil row per lane, kK1 lanes per block
// (a tuning choice: 64 or 256; larger blocks leave fewer partials for K2).
constexpr int kRed = 256;         // single-block reductions
constexpr int kK2 = 256;          // K2: kK2Rows rows per lane
constexpr int kK2Rows = 8;
constexpr int kBisect = 256;      // shifts per multisection round
constexpr int kInvitIters = 2;
constexpr size_t kInvitBytes = 7 * sizeof(double) + 1;  // LDS bytes per row of T in k_tri_vec
constexpr size_t kInvitLds = 156 * 1024;
constexpr size_t kBisectLdsMax = sizeof(double) * 2 * (kMaxIter + 2);  // (a, b^2) pairs of T
constexpr int kMaxGraphs = 16;
// Mixed path screening: every candidate Ritz pair of the check window gets
// a vector and residual (at most kMaxCand candidates); copies closer than
// kLumpMixed are one eigenvalue, represented by the copy with the smallest
// residual; an isolated value within kSpurMixed (relative) of an eigenvalue
// of T-hat is spurious.
constexpr int kMaxCand = 128;
constexpr int kSpecSteps = 50;  // mixed path: steps enqueued past a check while it runs
constexpr double kLumpMixed = 1e-8;
constexpr double kSpurMixed = 1e-7;

// Device-side solver state.
struct LzState {
  int k;                 // first step of the chunk being replayed
  float* V;              // basis: q_j at V + j * vstride
  int nscreened;         // screened Ritz values found
  int fail;              // 1: fewer than nvec screened values, 2: inverse iteration failed
  int m;                 // order of T at the last check
  double vals[kMaxEig];  // screened Ritz values
  double res[kMaxEig];   // residuals |beta_m s_{m,j}|
  int sel[kMaxEig];      // row of S holding the Ritz vector of vals[j]
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}

// Sum over the block in a fixed order; every thread gets the result.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  double t = red[0];
#pragma unroll
  for (int w = 1; w < NT / kWave; w++) t += red[w];
  __syncthreads();
  return t;
}

// Three block sums with one pair of barriers; each value is summed in
// block_sum's order, so the results are bit-identical to three block_sum
// calls.  red holds 3 * NT / kWave doubles.
template <int NT>
__device__ __forceinline__ void block_sum3(double& a, double& b, double& c, double* red) {
  constexpr int NW = NT / kWave;
  a = wave_sum(a);
  b = wave_sum(b);
  c = wave_sum(c);
  if ((threadIdx.x & (kWave - 1)) == 0) {
    red[threadIdx.x / kWave] = a;
    red[NW + threadIdx.x / kWave] = b;
    red[2 * NW + threadIdx.x / kWave] = c;
  }
  __syncthreads();
  a = red[0];
  b = red[NW];
  c = red[2 * NW];
#pragma unroll
  for (int w = 1; w < NW; w++) {
    a += red[w];
    b += red[NW + w];
    c += red[2 * NW + w];
  }
  __syncthreads();
}

// Sum of nb partials, identical in every block (fixed assignment and order).
template <int NT>
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int nb, double* red) {
  double v = 0;
  for (int i = threadIdx.x; i < nb; i += NT) v += part[i];
  return block_sum<NT>(v, red);
}

__device__ __forceinline__ float fma_t(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return fma(a, b, c); }

// Step index of a kernel inside a chunk's graph: the host sets st->k to
// the chunk's first step before each replay and node t of the chunk carries
// t as an argument, so no kernel updates the index while others read it.

// ---- plain path: FP32 vectors, alpha and beta; two kernels per step.
struct LzBufs {
  const float* A;
  size_t pitch;
  int n;
  size_t vstride;
  float* z0;      // z of even steps
  float* z1;      // z of odd steps (holds q_0 before step 0)
  double* part;   // [nb1] q.z partials of K1's blocks
  double* part2;  // [nb2] |z|^2 partials of K2's blocks
  float* alpha;   // [kMaxIter + 2]
  float* beta;
  int nb1, nb2;
  LzState* st;
};

template <int kK1>
__global__ __launch_bounds__(kK1) void k_lz1(LzBufs b, StencilOffsets so, int t) {
  __shared__ double red[kK1 / kWave];
  const int k = b.st->k + t;
  float* V = b.st->V;
  const float* zp = (k & 1) ? b.z0 : b.z1;
  float* zc = (k & 1) ? b.z1 : b.z0;
  float inv = 1, nbeta = 0;
  if (k > 0) {
    const float beta = (float)sqrt(sum_partials<kK1>(b.part2, b.nb2, red));
    if (blockIdx.x == 0 && threadIdx.x == 0) b.beta[k - 1] = beta;
    inv = 1.0f / beta;
    nbeta = -beta;
  }
  // q_k[c] = z_{k-1}[c] * (1 / beta_{k-1}) in FP32: the value the basis
  // stores (step 0 reads q_0 from z1 with beta 1)
  const int lb = dev::xcd_block(blockIdx.x, gridDim.x);
  const int row = lb * kK1 + threadIdx.x;
  double dot = 0;
  if (row < b.n) {
    const float qv = zp[row] * inv;
    V[(size_t)k * b.vstride + row] = qv;
    // diagonal order, as stencilMVM.cu:127-180; an out-of-range column
    // contributes fma(a, 0, acc) == acc (the reference's bounds test)
    float acc[1];
    spmv_rows<1>(acc, b.A, b.pitch, row, b.n, so, [&](int c) { return zp[c] * inv; });
    float z = acc[0];
    if (k > 0) z = fmaf(nbeta, V[(size_t)(k - 1) * b.vstride + row], z);
    zc[row] = z;
    dot = (double)qv * (double)z;  // exact product
  }
  const double s = block_sum<kK1>(dot, red);
  if (threadIdx.x == 0) b.part[lb] = s;
}

__global__ __launch_bounds__(kK2) void k_lz2(LzBufs b, int t) {
  __shared__ double red[kK2 / kWave];
  const int k = b.st->k + t;
  const float* qk = b.st->V + (size_t)k * b.vstride;
  float* zc = (k & 1) ? b.z1 : b.z0;
  const float alpha = (float)sum_partials<kK2>(b.part, b.nb1, red);
  if (blockIdx.x == 0 && threadIdx.x == 0) b.alpha[k] = alpha;
  double nrm = 0;
  const int base = blockIdx.x * kK2 * kK2Rows + threadIdx.x;
#pragma unroll
  for (int j = 0; j < kK2Rows; j++) {
    const int i = base + j * kK2;
    if (i < b.n) {
      const float zj = fmaf(-alpha, qk[i], zc[i]);
      zc[i] = zj;
      nrm = fma((double)zj, (double)zj, nrm);
    }
  }
  const double s = block_sum<kK2>(nrm, red);
  if (threadIdx.x == 0) b.part2[blockIdx.x] = s;
}

// beta of the last step m-1 (needed by the check's residuals).
__global__ __launch_bounds__(kRed) void k_lz_beta(LzBufs b, int m) {
  __shared__ double red[kRed / kWave];
  const float beta = (float)sqrt(sum_partials<kRed>(b.part2, b.nb2, red));
  if (threadIdx.x == 0) {
    b.beta[m - 1] = beta;
    b.st->m = m;
  }
}

// ---- mixed path: FP32 matrix and basis, FP64 accumulation; two kernels
// per step.  K1m(k) gathers q_k from the basis (FP32), accumulates
// z'_k = A q_k - beta_{k-1} q_{k-1} in FP64 and leaves partial sums of
// q_k.z'_k, |z'_k|^2 and |q_k|^2.  K2m(k) forms alpha_k = sum q.z' and,
// since z_k = z'_k - alpha_k q_k,
//   beta_k^2 = |z_k|^2 = sum z'^2 - 2 alpha_k sum q.z' + alpha_k^2 sum q^2
//            = sum z'^2 + alpha_k^2 (sum q^2 - 2),
// which in FP64 loses at most eps64 |z'|^2 / beta^2 (< 1e-12 here), then
// writes q_{k+1} = (z'_k - alpha_k q_k) / beta_k to the basis in FP32.  The
// basis keeps a zero row before q_0 so step 0 needs no special case.
struct LzmBufs {
  const float* A;
  size_t pitch;
  int n;
  size_t vstride;
  double* z;   // z'_k (two-kernel step); z' of even steps (fused step)
  double* z1;  // fused step: z' of odd steps (holds q_0 before step 0)
  double* p1;  // [2][nb1] q.z' partials (fused step: buffer = step parity)
  double* p2;  // [2][nb1] |z'|^2 partials
  double* p3;  // [2][nb1] |q|^2 partials
  double* alpha;
  double* beta;
  int nb1, nb2;
  LzState* st;
};

template <int kK1>
__global__ __launch_bounds__(kK1) void k_lzm1(LzmBufs b, StencilOffsets so, int t) {
  __shared__ double red[kK1 / kWave];
  const int k = b.st->k + t;
  const float* V = b.st->V;  // row k at V + k * vstride; row -1 is zero
  const float* q = V + (size_t)k * b.vstride;
  const double beta = k > 0 ? b.beta[k - 1] : 0.0;
  const int lb = dev::xcd_block(blockIdx.x, gridDim.x);
  const int row = lb * kK1 + threadIdx.x;
  double d1 = 0, d2 = 0, d3 = 0;
  if (row < b.n) {
    double acc = 0.0;  // FP64 accumulation of exact FP32 products, diagonal order
#pragma unroll 9
    for (int d = 0; d < kStencilArea; d++) {
      const float a = b.A[(size_t)d * b.pitch + row];
      const int col = row + so.off[d];
      const bool in = (unsigned)col < (unsigned)b.n;
      const float xv = q[in ? col : 0];
      acc = fma((double)a, (double)(in ? xv : 0.0f), acc);
    }
    const double qv = (double)q[row];
    const double zj = fma(-beta, (double)V[(long long)(k - 1) * (long long)b.vstride + row], acc);
    b.z[row] = zj;
    d1 = qv * zj;
    d2 = zj * zj;
    d3 = qv * qv;
  }
  d1 = block_sum<kK1>(d1, red);
  d2 = block_sum<kK1>(d2, red);
  d3 = block_sum<kK1>(d3, red);
  if (threadIdx.x == 0) {
    b.p1[lb] = d1;
    b.p2[lb] = d2;
    b.p3[lb] = d3;
  }
}

__global__ __launch_bounds__(kK2) void k_lzm2(LzmBufs b, int t) {
  __shared__ double red[kK2 / kWave];
  const int k = b.st->k + t;
  float* V = b.st->V;
  double v1 = 0, v2 = 0, v3 = 0;
  for (int i = threadIdx.x; i < b.nb1; i += kK2) {
    v1 += b.p1[i];
    v2 += b.p2[i];
    v3 += b.p3[i];
  }
  const double alpha = block_sum<kK2>(v1, red);
  const double s2 = block_sum<kK2>(v2, red), s3 = block_sum<kK2>(v3, red);
  const double beta = sqrt(fmax(fma(alpha * alpha, s3 - 2.0, s2), 0.0));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    b.alpha[k] = alpha;
    b.beta[k] = beta;
  }
  const double inv = 1.0 / beta;
  const float* q = V + (size_t)k * b.vstride;
  float* qn = V + (size_t)(k + 1) * b.vstride;
  const int base = blockIdx.x * kK2 * kK2Rows + threadIdx.x;
#pragma unroll
  for (int j = 0; j < kK2Rows; j++) {
    const int i = base + j * kK2;
    if (i < b.n) qn[i] = (float)(fma(-alpha, (double)q[i], b.z[i]) * inv);
  }
}

// Fused mixed step (one kernel per step).  Kernel k reduces the partials
// kernel k-1 left in buffer (k-1)&1 to alpha_{k-1} and beta_{k-1} (same
// expansion as k_lzm2), stages q_k = (z'_{k-1} - alpha q_{k-1}) / beta_{k-1}
// (rounded to FP32, as k_lzm2 writes it) for its rows plus the stencil halo
// into LDS, writes q_k to the basis for its own rows, and runs the FP64-
// accumulated SpMV from LDS.  Columns outside [0, n) are staged as 0, which
// is the reference's bounds test (fma(a, 0, acc) == acc).  z' and the
// partials alternate between two buffers by step parity, so a block that
// finishes early never overwrites a value another block has yet to read.
// Before step 0 the parity-1 buffers hold z' = q_0 and partials (0, 1, 0)
// (alpha 0, beta 1), and basis row -1 is zero.
template <int NT>
__global__ __launch_bounds__(NT) void k_lzmf(LzmBufs b, StencilOffsets so, int t) {
  extern __shared__ float sq[];
  __shared__ double red[3 * NT / kWave];
  const int k = b.st->k + t;
  float* V = b.st->V;
  const int lb = dev::xcd_block(blockIdx.x, gridDim.x);
  const int row0 = lb * NT, halo = so.halo;
  const size_t po = (size_t)((k - 1) & 1) * b.nb1, pc = (size_t)(k & 1) * b.nb1;
  const double* zp = (k & 1) ? b.z : b.z1;  // z'_{k-1}
  double* zc = (k & 1) ? b.z1 : b.z;        // z'_k
  const float* qm = V + (long long)(k - 1) * (long long)b.vstride;  // q_{k-1}
  const int c0 = row0 - halo, len = NT + 2 * halo;
  // the first kPre staged inputs of this thread are loaded before the
  // reduction of the partials, so the two memory round trips overlap
  constexpr int kPre = 12;
  double zr[kPre];
  float qr[kPre];
#pragma unroll
  for (int j = 0; j < kPre; j++) {
    const int c = c0 + (int)threadIdx.x + j * NT;
    const bool ok = j * NT + (int)threadIdx.x < len && (unsigned)c < (unsigned)b.n;
    zr[j] = ok ? zp[c] : 0.0;
    qr[j] = ok ? qm[c] : 0.0f;
  }
  double v1 = 0, v2 = 0, v3 = 0;
  for (int i = threadIdx.x; i < b.nb1; i += NT) {
    v1 += b.p1[po + i];
    v2 += b.p2[po + i];
    v3 += b.p3[po + i];
  }
  block_sum3<NT>(v1, v2, v3, red);
  const double alpha = v1, s2 = v2, s3 = v3;
  const double beta = sqrt(fmax(fma(alpha * alpha, s3 - 2.0, s2), 0.0));
  if (k > 0 && lb == 0 && threadIdx.x == 0) {
    b.alpha[k - 1] = alpha;
    b.beta[k - 1] = beta;
  }
  const double inv = 1.0 / beta;
#pragma unroll
  for (int j = 0; j < kPre; j++) {
    const int i = (int)threadIdx.x + j * NT;
    if (i < len) sq[i] = (unsigned)(c0 + i) < (unsigned)b.n ? (float)(fma(-alpha, (double)qr[j], zr[j]) * inv) : 0.0f;
  }
  for (int i = (int)threadIdx.x + kPre * NT; i < len; i += NT) {
    const int c = c0 + i;
    float v = 0.0f;
    if ((unsigned)c < (unsigned)b.n) v = (float)(fma(-alpha, (double)qm[c], zp[c]) * inv);
    sq[i] = v;
  }
  __syncthreads();
  const int row = row0 + threadIdx.x;
  double d1 = 0, d2 = 0, d3 = 0;
  if (row < b.n) {
    const float* x = sq + halo + threadIdx.x;
    const float qf = x[0];
    V[(size_t)k * b.vstride + row] = qf;
    double acc = 0.0;  // FP64 accumulation of exact FP32 products, diagonal order
#pragma unroll 9
    for (int d = 0; d < kStencilArea; d++) acc = fma((double)b.A[(size_t)d * b.pitch + row], (double)x[so.off[d]], acc);
    const double qv = (double)qf;
    const double zj = fma(-beta, (double)qm[row], acc);
    zc[row] = zj;
    d1 = qv * zj;
    d2 = zj * zj;
    d3 = qv * qv;
  }
  d1 = block_sum<NT>(d1, red);
  d2 = block_sum<NT>(d2, red);
  d3 = block_sum<NT>(d3, red);
  if (threadIdx.x == 0) {
    b.p1[pc + lb] = d1;
    b.p2[pc + lb] = d2;
    b.p3[pc + lb] = d3;
  }
}

// Tiled fused mixed step: as k_lzmf, but each block owns a TW x TH tile of
// image pixels instead of a run of consecutive rows, so the staged basis
// values (tile plus a 5-pixel frame, (TW + 10) x (TH + 10)) are 3x the rows
// for 32 x 8 tiles instead of 10x for 512 consecutive rows.  A staged slot
// (sy, sx) holds the basis value at linear index (ty0 - 5 + sy) * w +
// (tx0 - 5 + sx) -- the reference's linear indexing, including its wrap
// across image rows at the left and right edges -- or 0 outside [0, n), so
// every product is the one k_lzmf forms.  Blocks are numbered XCD-aware
// over the row-major tile order.
//
// LOADS selects when the 81 matrix loads of a lane are issued: 0 in groups
// of 9 inside the multiply-add loop, 1 all at once after the staging
// barrier, 2 all at once before the partial-sum reduction (so their
// latency overlaps the reduction and the staging).  The FP64 sums run in
// diagonal order in all three, so the results are bit-identical.
template <int TW, int TH, int LOADS>
__global__ __launch_bounds__(TW * TH) void k_lzmt(LzmBufs b, StencilOffsets so, int w, int h, int t) {
  constexpr int NT = TW * TH, SW = TW + 2 * kStencilRadius, SH = TH + 2 * kStencilRadius;
  __shared__ float sq[SW * SH];
  __shared__ double red[3 * NT / kWave];
  const int k = b.st->k + t;
  float* V = b.st->V;
  const int lb = dev::xcd_block(blockIdx.x, gridDim.x);
  const int tiles_x = (w + TW - 1) / TW;
  const int tx0 = (lb % tiles_x) * TW, ty0 = (lb / tiles_x) * TH;
  const size_t po = (size_t)((k - 1) & 1) * b.nb1, pc = (size_t)(k & 1) * b.nb1;
  const double* zp = (k & 1) ? b.z : b.z1;  // z'_{k-1}
  double* zc = (k & 1) ? b.z1 : b.z;        // z'_k
  const float* qm = V + (long long)(k - 1) * (long long)b.vstride;  // q_{k-1}
  const int lx = threadIdx.x % TW, ly = threadIdx.x / TW;
  const int x = tx0 + lx, y = ty0 + ly;
  const bool own = x < w && y < h;
  const int row = own ? y * w + x : 0;
  // Every load of the prologue is issued before the first one is used, with
  // clamped (in-range) indices instead of branches, in the order their
  // values are needed: the previous step's partial sums (the critical
  // path: alpha and beta), then the staged inputs, then (LOADS 2) the
  // matrix.  The hardware returns loads in order, so each use waits only
  // for the loads issued before it.
  constexpr int kPart = 4;  // partials per thread kept in registers; larger grids loop for the rest
  double pv[kPart][3];
#pragma unroll
  for (int j = 0; j < kPart; j++) {
    const int i = (int)threadIdx.x + j * NT;
    const size_t ic = po + (size_t)(i < b.nb1 ? i : 0);
    pv[j][0] = b.p1[ic];
    pv[j][1] = b.p2[ic];
    pv[j][2] = b.p3[ic];
  }
  constexpr int kPre = (SW * SH + NT - 1) / NT;
  double zr[kPre];
  float qr[kPre];
#pragma unroll
  for (int j = 0; j < kPre; j++) {
    const int i = (int)threadIdx.x + j * NT;
    const int sy = i / SW, sx = i - sy * SW;
    const int c = (ty0 - kStencilRadius + sy) * w + (tx0 - kStencilRadius + sx);
    const int cc = (i < SW * SH && (unsigned)c < (unsigned)b.n) ? c : 0;  // selected away below
    zr[j] = zp[cc];
    qr[j] = qm[cc];
  }
  const float qmo = qm[row];  // q_{k-1} at the lane's own pixel
  float am[LOADS ? kStencilArea : 1];
  if constexpr (LOADS == 2) {
#pragma unroll
    for (int d = 0; d < kStencilArea; d++) am[d] = b.A[(size_t)d * b.pitch + row];
  }
  // sums in the order of a plain strided loop from 0: bit-identical to it
  double v1 = 0, v2 = 0, v3 = 0;
#pragma unroll
  for (int j = 0; j < kPart; j++) {
    if ((int)threadIdx.x + j * NT < b.nb1) {
      v1 += pv[j][0];
      v2 += pv[j][1];
      v3 += pv[j][2];
    }
  }
  for (int i = threadIdx.x + kPart * NT; i < b.nb1; i += NT) {
    v1 += b.p1[po + i];
    v2 += b.p2[po + i];
    v3 += b.p3[po + i];
  }
  block_sum3<NT>(v1, v2, v3, red);
  const double alpha = v1, s2 = v2, s3 = v3;
  const double beta = sqrt(fmax(fma(alpha * alpha, s3 - 2.0, s2), 0.0));
  if (k > 0 && lb == 0 && threadIdx.x == 0) {
    b.alpha[k - 1] = alpha;
    b.beta[k - 1] = beta;
  }
  const double inv = 1.0 / beta;
#pragma unroll
  for (int j = 0; j < kPre; j++) {
    const int i = (int)threadIdx.x + j * NT;
    if (i < SW * SH) {
      const int sy = i / SW, sx = i - sy * SW;
      const int c = (ty0 - kStencilRadius + sy) * w + (tx0 - kStencilRadius + sx);
      sq[i] = (unsigned)c < (unsigned)b.n ? (float)(fma(-alpha, (double)qr[j], zr[j]) * inv) : 0.0f;
    }
  }
  __syncthreads();
  double d1 = 0, d2 = 0, d3 = 0;
  if (own) {
    const float* xs = sq + (ly + kStencilRadius) * SW + (lx + kStencilRadius);
    const float qf = xs[0];
    V[(size_t)k * b.vstride + row] = qf;
    double acc = 0.0;  // FP64 accumulation of exact FP32 products, diagonal order
    if constexpr (LOADS == 0) {
#pragma unroll 9
      for (int d = 0; d < kStencilArea; d++)
        acc = fma((double)b.A[(size_t)d * b.pitch + row], (double)xs[so.dy[d] * SW + so.dx[d]], acc);
    } else {
      if constexpr (LOADS == 1) {
#pragma unroll
        for (int d = 0; d < kStencilArea; d++) am[d] = b.A[(size_t)d * b.pitch + row];
      }
#pragma unroll
      for (int d = 0; d < kStencilArea; d++) acc = fma((double)am[d], (double)xs[so.dy[d] * SW + so.dx[d]], acc);
    }
    const double qv = (double)qf;
    const double zj = fma(-beta, (double)qmo, acc);
    zc[row] = zj;
    d1 = qv * zj;
    d2 = zj * zj;
    d3 = qv * qv;
  }
  block_sum3<NT>(d1, d2, d3, red);
  if (threadIdx.x == 0) {
    b.p1[pc + lb] = d1;
    b.p2[pc + lb] = d2;
    b.p3[pc + lb] = d3;
  }
}

// alpha and beta of the last fused step m-1, for the check.
__global__ __launch_bounds__(kRed) void k_lzmf_finish(LzmBufs b, int m) {
  __shared__ double red[kRed / kWave];
  const size_t po = (size_t)((m - 1) & 1) * b.nb1;
  double v1 = 0, v2 = 0, v3 = 0;
  for (int i = threadIdx.x; i < b.nb1; i += kRed) {
    v1 += b.p1[po + i];
    v2 += b.p2[po + i];
    v3 += b.p3[po + i];
  }
  const double alpha = block_sum<kRed>(v1, red);
  const double s2 = block_sum<kRed>(v2, red), s3 = block_sum<kRed>(v3, red);
  if (threadIdx.x == 0) {
    b.alpha[m - 1] = alpha;
    b.beta[m - 1] = sqrt(fmax(fma(alpha * alpha, s3 - 2.0, s2), 0.0));
    b.st->m = m;
  }
}

// Fused-step initial state: the parity-1 partials give alpha 0, beta 1.
__global__ void k_lzmf_init(double* p1, double* p2, double* p3, int nb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nb) {
    p1[nb + i] = 0.0;
    p2[nb + i] = i == 0 ? 1.0 : 0.0;
    p3[nb + i] = 0.0;
  }
}

// FP64 copies of alpha and beta.
template <typename Z>
__global__ void k_widen(const Z* __restrict__ al, const Z* __restrict__ be, double* __restrict__ ad,
                        double* __restrict__ bd, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    ad[i] = (double)al[i];
    bd[i] = (double)be[i];
  }
}

// ---------------------------------------------------------------- check
// Ritz values live in w[which * kWStride + idx] (which 0: T, 1: T-hat), so a
// check can read the previous check's values whatever its window size.
constexpr int kWStride = kCullumMax;

// Block (idx, which): eigenvalue idx+1 of T (which 0, order m) or of T-hat
// (which 1, order m-1: alpha[1..], beta[1..]) by multisection with kBisect
// Sturm counts per round, to relative accuracy 4.4e-16 or absolute
// 2.2e-16 |T|, whichever is larger (the accuracy T's entries carry).  The
// (a_i, b_{i-1}^2) pairs are staged in LDS once and read by every round.
//
// Bracket: T (and T-hat) of the previous check at order prev_m < m is a
// leading principal submatrix of this one, so by Cauchy interlacing its
// value idx+1 bounds this one from above; and T is the Rayleigh quotient
// of L_sym, which is positive semidefinite, so 0 bounds it from below up to
// rounding.  The first round puts its two outer shifts on that bracket and
// falls back to the Gershgorin interval if either side does not hold, so
// the result never depends on the bracket being right.
__global__ __launch_bounds__(kBisect) void k_tri_bisect(const double* __restrict__ alpha,
                                                        const double* __restrict__ beta, const LzState* st,
                                                        double* __restrict__ w, int prev_n) {
  extern __shared__ double2 sab[];
  __shared__ double rlo[kBisect / kWave], rhi[kBisect / kWave], rb2[kBisect / kWave];
  __shared__ unsigned long long masks[kBisect / kWave];
  const int m = st->m;
  const int which = blockIdx.y, idx = blockIdx.x;
  const int n = which ? m - 1 : m;
  if (idx >= n) return;
  const double* a = alpha + which;
  const double* b = beta + which;
  double lo = INFINITY, hi = -INFINITY, bmax = 0;
  for (int i = threadIdx.x; i < n; i += kBisect) {
    const double bl = i > 0 ? b[i - 1] : 0.0, br = i < n - 1 ? b[i] : 0.0;
    sab[i] = make_double2(a[i], bl * bl);
    const double r = fabs(bl) + fabs(br);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
    bmax = fmax(bmax, br * br);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
    bmax = fmax(bmax, __shfl_xor(bmax, o, 64));
  }
  if ((threadIdx.x & (kWave - 1)) == 0) {
    rlo[threadIdx.x / kWave] = lo;
    rhi[threadIdx.x / kWave] = hi;
    rb2[threadIdx.x / kWave] = bmax;
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < kBisect / kWave; v++) {
    lo = fmin(lo, rlo[v]);
    hi = fmax(hi, rhi[v]);
    bmax = fmax(bmax, rb2[v]);
  }
  const double tnorm = fmax(fabs(lo), fabs(hi));
  const double pivmin = fmax(2.2250738585072014e-308 * fmax(1.0, bmax), 1e-300);
  const double abstol = 2.2e-16 * tnorm;
  const double pad = 2.2e-16 * tnorm * n + 2 * pivmin;
  lo -= pad;
  hi += pad;
  const int kk = idx + 1;
  double* slot = w + which * kWStride + idx;
  if (idx < prev_n - which) {  // first round on the interlacing / semidefinite bracket
    const double bl = -64 * abstol - 2 * pivmin;
    const double bh = fmin(hi, *slot + 8.8e-16 * fabs(*slot) + 64 * abstol + 2 * pivmin);
    if (bl < bh) {
      const double step = (bh - bl) / (double)(kBisect - 1);
      const double x = threadIdx.x == kBisect - 1 ? bh : bl + step * threadIdx.x;
      const int c = tri::sturm_pairs(n, sab, x);
      const unsigned long long ge = __ballot(c >= kk);
      if ((threadIdx.x & (kWave - 1)) == 0) masks[threadIdx.x / kWave] = ge;
      __syncthreads();
      int first = -1;
#pragma unroll
      for (int v = 0; v < kBisect / kWave; v++)
        if (first < 0 && masks[v]) first = v * kWave + __ffsll((unsigned long long)masks[v]) - 1;
      __syncthreads();
      if (first > 0) {  // count(bl) < kk <= count(bh): the bracket holds
        lo = bl + step * (first - 1);
        hi = first == kBisect - 1 ? bh : bl + step * first;
      }
    }
  }
  for (int round = 0; round < 24; round++) {
    if (hi - lo <= fmax(4.4e-16 * fmax(fabs(lo), fabs(hi)), abstol) + pivmin) break;
    const double step = (hi - lo) / (double)(kBisect + 1);
    const int c = tri::sturm_pairs(n, sab, lo + step * (threadIdx.x + 1));
    // the first shift with at least kk values below it bounds lambda_kk above
    const unsigned long long ge = __ballot(c >= kk);
    if ((threadIdx.x & (kWave - 1)) == 0) masks[threadIdx.x / kWave] = ge;
    __syncthreads();
    int first = -1;
#pragma unroll
    for (int v = 0; v < kBisect / kWave; v++)
      if (first < 0 && masks[v]) first = v * kWave + __ffsll((unsigned long long)masks[v]) - 1;
    __syncthreads();
    if (first >= 0) {
      const double nhi = lo + step * (first + 1);
      lo = first == 0 ? lo : lo + step * first;
      hi = nhi;
    } else {
      lo = lo + step * kBisect;
    }
  }
  if (threadIdx.x == 0) *slot = 0.5 * (lo + hi);
}

// Screening (lanczos.cu:279-384): lump values closer than 1e-5, drop single
// values that are also eigenvalues of T-hat; the first nvec survivors go to
// st->vals.
__global__ __launch_bounds__(kWave) void k_tri_screen(LzState* st, const double* __restrict__ w, int ncheck,
                                                      int nvec) {
  __shared__ double acc[kCullumMax];
  __shared__ unsigned char dup[kCullumMax];
  if (threadIdx.x != 0) return;
  const int m = st->m;
  const double* wr = w;           // Ritz values of T, ascending
  const double* wc = w + kWStride;  // of T-hat
  const int ec = min(ncheck, m), ech = min(ncheck, m - 1);
  int na = 0;
  for (int j = 0; j < ec; j++) {
    if (j > 0 && wr[j] - wr[j - 1] < 1e-5) {
      dup[na - 1] = 1;
      continue;
    }
    acc[na] = wr[j];
    dup[na] = 0;
    na++;
  }
  int ns = 0;
  for (int j = 0; j < na && ns < nvec; j++) {
    bool accept = true;
    if (!dup[j])
      for (int c = 0; c < ech; c++)
        if (fabs((acc[j] - wc[c]) / acc[j]) <= 1e-10) accept = false;
    if (accept) {
      st->sel[ns] = ns;  // vectors are computed after screening, in this order
      st->vals[ns++] = acc[j];
    }
  }
  st->nscreened = ns;
  st->fail = ns < nvec ? 1 : 0;
}

// Mixed path screening, after k_tri_vec has computed a vector and residual
// for every candidate: values of T closer than kLumpMixed form one group
// (copies of one eigenvalue), represented by its member with the smallest
// residual; a single value within kSpurMixed (relative) of an eigenvalue of
// T-hat is spurious (the Cullum-Willoughby test with tolerances for the
// FP64-accumulated recurrence, whose copies agree to ~1e-10 while distinct
// eigenvalues here lie >= 7e-6 apart).
__global__ __launch_bounds__(kWave) void k_tri_screen_mixed(LzState* st, const double* __restrict__ w,
                                                            const double* __restrict__ res_all, int ncheck,
                                                            int nvec) {
  if (threadIdx.x != 0) return;
  const int m = st->m;
  const double* wr = w;           // Ritz values of T, ascending
  const double* wc = w + kWStride;  // of T-hat
  const int ec = min(ncheck, m), ech = min(ncheck, m - 1);
  int ns = 0;
  for (int j = 0; j < ec && ns < nvec;) {
    int e = j + 1;
    while (e < ec && wr[e] - wr[e - 1] < kLumpMixed) e++;
    bool accept = true;
    if (e == j + 1)
      for (int c = 0; c < ech; c++)
        if (fabs((wr[j] - wc[c]) / wr[j]) <= kSpurMixed) accept = false;
    if (accept) {
      int b = j;
      for (int i = j + 1; i < e; i++)
        if (res_all[i] < res_all[b]) b = i;
      st->vals[ns] = wr[b];
      st->res[ns] = res_all[b];
      st->sel[ns] = b;
      ns++;
    }
    j = e;
  }
  st->nscreened = ns;
  st->fail = ns < nvec ? 1 : 0;
}

// Ritz vector j of T by inverse iteration (block j, one wave: lane 0 runs
// the recurrences, the wave shares the passes without serial dependence)
// and its residual |beta_m s_{m,j}| (lanczos.cu:114-140: FP32
// on the plain path).  When they fit, alpha, beta and the work arrays are
// staged in LDS; otherwise the work arrays live in global memory.
// Plain path: vector j of the screened value st->vals[j], residual to
// st->res[j].  Mixed path (cand != null): vector j of candidate cand[j]
// (every Ritz value of the check window, before screening), residual to
// res_all[j]; a failed solve there only marks that candidate.
__global__ __launch_bounds__(kWave) void k_tri_vec(const double* __restrict__ alpha, const double* __restrict__ beta,
                                                   LzState* st, int mixed, double* __restrict__ S,
                                                   double* __restrict__ gwork, int lds,
                                                   const double* __restrict__ cand, double* __restrict__ res_all) {
  extern __shared__ double dyn[];
  if (!cand && st->fail == 1) return;
  const int j = blockIdx.x, m = st->m;
  double* sj = S + (size_t)j * m;
  const double lam = cand ? cand[j] : st->vals[j];
  // one solve for a candidate (its residual and, if selected, its vector:
  // the other eigenvectors stay below eps |T| / gap), two for the plain path
  const int iters = cand ? 1 : kInvitIters;
  __shared__ double nrm_bc[1];
  int bad;
  if (lds) {
    double* sa = dyn;
    double* sb = dyn + m;
    for (int i = threadIdx.x; i < m; i += kWave) {
      sa[i] = alpha[i];
      sb[i] = beta[i];
    }
    __syncthreads();
    double* wk = dyn + 2 * m;
    bad = tri::invit_wave(m, sa, sb, lam, sj, wk, wk + m, wk + 2 * m, wk + 3 * m, wk + 4 * m,
                          (unsigned char*)(wk + 5 * m), iters, nrm_bc);
  } else {
    double* wk = gwork + (size_t)j * 6 * m;
    bad = tri::invit_wave(m, alpha, beta, lam, sj, wk, wk + m, wk + 2 * m, wk + 3 * m, wk + 4 * m,
                          (unsigned char*)(wk + 5 * m), iters, nrm_bc);
  }
  if (threadIdx.x != 0) return;
  const double bm = beta[m - 1];
  const double r = mixed ? fabs(bm * sj[m - 1]) : (double)fabsf((float)bm * (float)sj[m - 1]);
  if (cand) {
    res_all[j] = bad ? INFINITY : r;
    return;
  }
  st->res[j] = bad ? INFINITY : r;
  if (bad) st->fail = 2;
}

// ---------------------------------------------------------------- finish
// u_j = V s_j (lanczos.cu:249-270), accumulated in FP64 in step order; the
// plain path multiplies by S rounded to FP32 (the reference keeps the Ritz
// coefficients in FP32).
template <int NV, bool S32>
__global__ __launch_bounds__(256) void k_ritz(int n, const LzState* st, size_t vstride, int m,
                                              const double* __restrict__ S, float* __restrict__ out) {
  // S is staged through LDS in chunks of kRitzChunk steps (every lane reads
  // the same coefficient, a broadcast); V is read 8 rows at a time so each
  // lane keeps 8 loads in flight.  The sums run in step order as before.
  constexpr int kRitzChunk = 256;
  __shared__ double sc[NV][kRitzChunk];
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = p < n;
  const float* V = st->V + (live ? p : 0);
  double acc[NV];
#pragma unroll
  for (int j = 0; j < NV; j++) acc[j] = 0;
  for (int k0 = 0; k0 < m; k0 += kRitzChunk) {
    const int len = min(kRitzChunk, m - k0);
    __syncthreads();
    for (int i = threadIdx.x; i < NV * kRitzChunk; i += blockDim.x) {
      const int j = i / kRitzChunk, k = i - j * kRitzChunk;
      if (k < len) {
        const double v = S[(size_t)st->sel[j] * m + k0 + k];
        sc[j][k] = S32 ? (double)(float)v : v;
      }
    }
    __syncthreads();
    int k = 0;
    for (; k + 8 <= len; k += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = (double)V[(size_t)(k0 + k + u) * vstride];
#pragma unroll
      for (int u = 0; u < 8; u++)
#pragma unroll
        for (int j = 0; j < NV; j++) acc[j] = fma(v[u], sc[j][k + u], acc[j]);
    }
    for (; k < len; k++) {
      const double v = (double)V[(size_t)(k0 + k) * vstride];
#pragma unroll
      for (int j = 0; j < NV; j++) acc[j] = fma(v, sc[j][k], acc[j]);
    }
  }
  if (!live) return;
#pragma unroll
  for (int j = 0; j < NV; j++) out[(size_t)j * n + p] = (float)acc[j];
}

// stencilMVM.cu:182-198 scaleEigByD
__global__ void k_scale_rs(int n, int nvec, const float* __restrict__ rs, float* __restrict__ v) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int j = 0; j < nvec; j++) v[(size_t)j * n + p] = v[(size_t)j * n + p] * rs[p];
}

// lanczos.cu:735-885 per-vector minimum and maximum (exact in any order):
// stage 1 per block, stage 2 one block per vector.
__device__ __forceinline__ void block_minmax(float& lo, float& hi, float* smin, float* smax) {
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = 128; s >= 1; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  lo = smin[0];
  hi = smax[0];
}

__global__ __launch_bounds__(256) void k_minmax1(int n, const float* __restrict__ v, float* __restrict__ mm,
                                                 int nblk) {
  __shared__ float smin[256], smax[256];
  const int j = blockIdx.y;
  const float* x = v + (size_t)j * n;
  float lo = INFINITY, hi = -INFINITY;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    lo = fminf(lo, x[i]);
    hi = fmaxf(hi, x[i]);
  }
  block_minmax(lo, hi, smin, smax);
  if (threadIdx.x == 0) {
    mm[((size_t)j * nblk + blockIdx.x) * 2] = lo;
    mm[((size_t)j * nblk + blockIdx.x) * 2 + 1] = hi;
  }
}

__global__ __launch_bounds__(256) void k_minmax2(const float* __restrict__ mm, int nblk, float* __restrict__ fin) {
  __shared__ float smin[256], smax[256];
  const int j = blockIdx.x;
  float lo = INFINITY, hi = -INFINITY;
  for (int b = threadIdx.x; b < nblk; b += 256) {
    lo = fminf(lo, mm[((size_t)j * nblk + b) * 2]);
    hi = fmaxf(hi, mm[((size_t)j * nblk + b) * 2 + 1]);
  }
  block_minmax(lo, hi, smin, smax);
  if (threadIdx.x == 0) {
    fin[2 * j] = lo;
    fin[2 * j + 1] = hi;
  }
}

// lanczos.cu:887-901 NormalizationDev (vectors 1..nvec-1)
__global__ void k_normalize(int n, int nvec, const float* __restrict__ fin, float* __restrict__ v) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int j = 1; j < nvec; j++) {
    const float lo = fin[2 * j], hi = fin[2 * j + 1];
    if ((hi - lo) > 1e-4f) {
      float* x = v + (size_t)j * n + p;
      *x = (*x - lo) / (hi - lo);
    }
  }
}

// Solver state, q_0 in basis row 0 and q_0's source in z1 (the plain
// path's first step reads q_0 from z1 with beta 1).
template <typename Z>
__global__ void k_init(LzState* st, float* V, Z* z1, float q0, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    st->k = 0;
    st->V = V;
    st->nscreened = 0;
    st->fail = 0;
    st->m = 0;
  }
  if (i < n) {
    V[i] = q0;
    z1[i] = (Z)q0;
  }
}

__global__ void k_zero_row(float* row, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) row[i] = 0.0f;
}

__global__ void k_set_basis(LzState* st, float* V) { st->V = V; }
__global__ void k_set_step(LzState* st, int k) { st->k = k; }
__global__ void k_set_m(LzState* st, int m) { st->m = m; }

// ---------------------------------------------------------------- host

// Graphs of kGraphSteps-step chunks, per context; the kernels' arguments
// are baked into the graph, so the key holds every pointer and size they
// depend on (the solver's buffers are carved from its arena at offsets that
// depend only on n, so st and n fix them).
struct GraphKey {
  const void* A;
  const void* st;
  int pitch, n, w, mixed, steps, k1, fused_lds, tile;
  bool operator<(const GraphKey& o) const {
    return std::tie(A, st, pitch, n, w, mixed, steps, k1, fused_lds, tile) <
           std::tie(o.A, o.st, o.pitch, o.n, o.w, o.mixed, o.steps, o.k1, o.fused_lds, o.tile);
  }
};

struct GraphCache {
  std::map<GraphKey, hipGraphExec_t> execs;
  void clear() {
    for (auto& e : execs) (void)hipGraphExecDestroy(e.second);
    execs.clear();
  }
  ~GraphCache() { clear(); }
};

// One cache per context; the map is shared by all contexts (and host
// threads), so it is only touched under graph_lock().  Each cache is then
// used by its own context's thread alone.
std::map<const Ctx*, std::unique_ptr<GraphCache>>& graph_caches() {
  static std::map<const Ctx*, std::unique_ptr<GraphCache>> g;
  return g;
}
std::mutex& graph_lock() {
  static std::mutex m;
  return m;
}
GraphCache* graph_cache(const Ctx* ctx) {
  std::lock_guard<std::mutex> lk(graph_lock());
  auto& gc = graph_caches()[ctx];
  if (!gc) gc.reset(new GraphCache());
  return gc.get();
}

struct Steps {
  bool mixed;
  int k1;            // K1 block size
  size_t fused_lds;  // mixed path: LDS bytes of the linear fused step kernel, 0 = not used
  int tile_w, tile_h;  // mixed path: tile of the tiled fused step, 0 = not used
  int loads;           // tiled fused step: matrix load schedule (k_lzmt's LOADS)
  int w, h;
  LzBufs plain;
  LzmBufs mix;
  StencilOffsets so;
};

// Appends the kernel node fn<<<grid, block, lds>>>(args...) to g after
// *last (the graph copies the arguments when the node is created).  Graphs
// are built node by node rather than by stream capture, so building one on
// a host thread never interacts with other threads' HIP calls.
template <typename... A>
int add_node(hipGraph_t g, hipGraphNode_t* last, const void* fn, dim3 grid, dim3 block, size_t lds, A... args) {
  void* ptrs[] = {(void*)&args...};
  hipKernelNodeParams kp{};
  kp.func = const_cast<void*>(fn);
  kp.gridDim = grid;
  kp.blockDim = block;
  kp.sharedMemBytes = (unsigned)lds;
  kp.kernelParams = ptrs;
  kp.extra = nullptr;
  hipGraphNode_t n = nullptr;
  DMSC_HIP(hipGraphAddKernelNode(&n, g, *last ? last : nullptr, *last ? 1 : 0, &kp));
  *last = n;
  return DMSC_OK;
}

// The kernels of `steps` consecutive Lanczos steps as a chain of graph
// nodes; node t carries its offset t inside the chunk.
int add_steps(hipGraph_t g, const Steps& p, int steps) {
  hipGraphNode_t last = nullptr;
  for (int t = 0; t < steps; t++) {
    if (p.mixed && p.tile_h) {
      const int code = p.tile_w * 100 + p.tile_h;
      const void* fn = nullptr;
#define TILE_FN(TW, TH)                                                        \
  if (code == TW * 100 + TH)                                                   \
    fn = p.loads == 2   ? (const void*)&k_lzmt<TW, TH, 2>                      \
         : p.loads == 1 ? (const void*)&k_lzmt<TW, TH, 1>                      \
                        : (const void*)&k_lzmt<TW, TH, 0>;
      TILE_FN(32, 8) TILE_FN(64, 4) TILE_FN(16, 16) TILE_FN(32, 16) TILE_FN(16, 8) TILE_FN(64, 8)
      TILE_FN(64, 2) TILE_FN(32, 4)
#undef TILE_FN
      if (!fn) return fail(DMSC_ERR_INTERNAL, "no tiled step kernel for tile code %d", code);
      DMSC_TRY(add_node(g, &last, fn, dim3(p.mix.nb1), dim3(p.tile_w * p.tile_h), 0, p.mix, p.so, p.w, p.h, t));
    } else if (p.mixed && p.fused_lds) {
      const void* fn = p.k1 == 512 ? (const void*)&k_lzmf<512> : (const void*)&k_lzmf<256>;
      DMSC_TRY(add_node(g, &last, fn, dim3(p.mix.nb1), dim3(p.k1 == 512 ? 512 : 256), p.fused_lds, p.mix, p.so, t));
    } else if (p.mixed) {
      const void* fn = p.k1 == 64 ? (const void*)&k_lzm1<64> : (const void*)&k_lzm1<256>;
      DMSC_TRY(add_node(g, &last, fn, dim3(p.mix.nb1), dim3(p.k1 == 64 ? 64 : 256), 0, p.mix, p.so, t));
      DMSC_TRY(add_node(g, &last, (const void*)&k_lzm2, dim3(p.mix.nb2), dim3(kK2), 0, p.mix, t));
    } else {
      const void* fn = p.k1 == 64 ? (const void*)&k_lz1<64> : (const void*)&k_lz1<256>;
      DMSC_TRY(add_node(g, &last, fn, dim3(p.plain.nb1), dim3(p.k1 == 64 ? 64 : 256), 0, p.plain, p.so, t));
      DMSC_TRY(add_node(g, &last, (const void*)&k_lz2, dim3(p.plain.nb2), dim3(kK2), 0, p.plain, t));
    }
  }
  return DMSC_OK;
}

// Steps first .. first+steps-1 on ctx->stream through a cached HIP graph.
int run_steps(Ctx* ctx, const Steps& p, LzState* st, int w, int first, int steps) {
  GraphCache* gc = graph_cache(ctx);
  const GraphKey key{p.mixed ? (const void*)p.mix.A : (const void*)p.plain.A, st,
                     (int)(p.mixed ? p.mix.pitch : p.plain.pitch), p.mixed ? p.mix.n : p.plain.n, w, p.mixed ? 1 : 0,
                     steps, p.k1, (int)p.fused_lds, p.loads * 10000 + p.tile_w * 100 + p.tile_h};
  auto it = gc->execs.find(key);
  if (it == gc->execs.end()) {
    if ((int)gc->execs.size() >= kMaxGraphs) gc->clear();
    hipGraph_t g = nullptr;
    DMSC_HIP(hipGraphCreate(&g, 0));
    const int rc = add_steps(g, p, steps);
    hipGraphExec_t exec = nullptr;
    const hipError_t ie = rc == DMSC_OK ? hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) : hipSuccess;
    (void)hipGraphDestroy(g);
    if (rc != DMSC_OK) return rc;
    DMSC_HIP(ie);
    it = gc->execs.emplace(key, exec).first;
  }
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(1), 0, ctx->stream, st, first);
  DMSC_HIP(hipGetLastError());
  DMSC_HIP(hipGraphLaunch(it->second, ctx->stream));
  return DMSC_OK;
}

int launch_ritz(hipStream_t s, bool mixed, int nvec, int n, const LzState* st, size_t vstride, int m,
                const double* S, float* out) {
  const dim3 g(div_up(n, 256)), b(256);
  switch (nvec) {
#define RITZ_CASE(NV)                                                                                   \
  case NV:                                                                                              \
    if (mixed)                                                                                          \
      hipLaunchKernelGGL((k_ritz<NV, false>), g, b, 0, s, n, st, vstride, m, S, out);                   \
    else                                                                                                \
      hipLaunchKernelGGL((k_ritz<NV, true>), g, b, 0, s, n, st, vstride, m, S, out);                    \
    break;
    RITZ_CASE(1) RITZ_CASE(2) RITZ_CASE(3) RITZ_CASE(4) RITZ_CASE(5) RITZ_CASE(6) RITZ_CASE(7) RITZ_CASE(8)
    RITZ_CASE(9) RITZ_CASE(10) RITZ_CASE(11) RITZ_CASE(12) RITZ_CASE(13) RITZ_CASE(14) RITZ_CASE(15)
    RITZ_CASE(16) RITZ_CASE(17) RITZ_CASE(18) RITZ_CASE(19) RITZ_CASE(20) RITZ_CASE(21) RITZ_CASE(22)
    RITZ_CASE(23) RITZ_CASE(24) RITZ_CASE(25)
#undef RITZ_CASE
    default: return fail(DMSC_ERR_ARG, "nvec %d out of range", nvec);
  }
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

// Mixed path check schedule: the largest residual of the wanted Ritz pairs
// falls roughly geometrically with the step count, so the next check goes
// where the trend of the last two checks reaches the tolerance (scaled by
// 0.7), at least kGraphSteps and at most m/4 steps ahead, in whole chunks.
int next_period(int m, const LzState& hs, int nvec, double tol, int& prev_m, double& prev_res) {
  int steps = std::max(kGraphSteps, m / 4);
  if (hs.fail == 0) {
    double r = 0;
    for (int j = 0; j < nvec; j++) r = std::max(r, hs.res[j]);
    if (prev_m > 0 && r > 0 && prev_res > r) {
      const double rate = (std::log(r) - std::log(prev_res)) / (double)(m - prev_m);
      steps = (int)(0.7 * (std::log(tol) - std::log(r)) / rate);
    }
    prev_m = m;
    prev_res = r;
  }
  steps = std::min(std::max(steps, kGraphSteps), std::max(kGraphSteps, m / 4));
  if (lanczos_check_max > 0) steps = std::min(steps, lanczos_check_max);
  return (steps + kGraphSteps - 1) / kGraphSteps * kGraphSteps;
}

// Step-kernel shapes chosen by the autotuner, per (w, h, precision); a
// solve on a size with an entry uses it instead of the process-wide tuning
// values.
std::map<std::tuple<int, int, int>, StepShape>& tuned_shapes() {
  static std::map<std::tuple<int, int, int>, StepShape> t;
  return t;
}
std::mutex& tuned_lock() {
  static std::mutex m;
  return m;
}

StepShape default_shape(int w, int h, bool mixed) {
  {
    std::lock_guard<std::mutex> lk(tuned_lock());
    auto it = tuned_shapes().find(std::make_tuple(w, h, mixed ? 1 : 0));
    if (it != tuned_shapes().end()) return it->second;
  }
  return StepShape{mixed ? lanczos_mixed_fused : 0, lanczos_tile, lanczos_k1_threads, 0.f};
}

// time_steps > 0: run that many steps with no checks between two HIP
// events on the context's stream and return the average ms per step in
// *ms_per_step (the measurement mode of dmsc_lanczos_time_steps).  `force`
// selects the step-kernel shape (the autotuner's trials); otherwise the
// autotuned shape of this size, else the process-wide tuning values.
// convert == false: A already holds L_sym (a later timing run on the same
// matrix).
int solve(Ctx* ctx, int w, int h, float* A, int pitch, int nvec, float tol, bool mixed, float* h_eigvals,
          float* d_eigvecs, int* iterations, int time_steps = 0, float* ms_per_step = nullptr,
          const StepShape* force = nullptr, bool convert = true) {
  const StepShape shape = force ? *force : default_shape(w, h, mixed);
  hipStream_t s = ctx->stream;
  const int n = w * h;
  const size_t vstride = (size_t)pitch_floats(n);
  // dynamic LDS above the 64 KiB default, within the device's limit; set
  // once per process (a kernel attribute), under a lock for host threads
  static size_t invit_lds = 0;
  static std::mutex attr_lock;
  std::unique_lock<std::mutex> attr_lk(attr_lock);
  if (!invit_lds) {
    int lds_max = 0;
    DMSC_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
    if ((size_t)lds_max < kBisectLdsMax + 1024)
      return fail(DMSC_ERR_INTERNAL, "device LDS per block %d below the %zu the eigensolver needs", lds_max,
                  kBisectLdsMax + 1024);
    const size_t vec_lds = std::min(kInvitLds, (size_t)lds_max);
    DMSC_HIP(hipFuncSetAttribute((const void*)k_tri_bisect, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)kBisectLdsMax));
    DMSC_HIP(hipFuncSetAttribute((const void*)k_tri_vec, hipFuncAttributeMaxDynamicSharedMemorySize, (int)vec_lds));
    DMSC_HIP(hipFuncSetAttribute((const void*)k_lzmf<256>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)vec_lds));
    DMSC_HIP(hipFuncSetAttribute((const void*)k_lzmf<512>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)vec_lds));
    invit_lds = vec_lds;
  }
  attr_lk.unlock();
  // Step kernel shape.  The fused mixed step stages its rows plus the
  // stencil halo in LDS (256 or 512 rows per block) and falls back to the
  // two-kernel step when that does not fit; the two-kernel steps take 64 or
  // 256 rows per block.
  const int halo = make_stencil_offsets(w).halo;
  int k1 = shape.k1 == 64 ? 64 : 256;
  size_t fused_lds = 0;
  int tile_w = 0, tile_h = 0;
  if (mixed && shape.fused == 2) {  // pixel tiles, TW x TH
    tile_w = shape.tile / 100;
    tile_h = shape.tile % 100;
    k1 = tile_w * tile_h;
  } else if (mixed && shape.fused == 1) {
    const int kf = (shape.k1 == 0 || shape.k1 == 512) ? 512 : 256;
    const size_t bytes = sizeof(float) * ((size_t)kf + 2 * (size_t)halo);
    if (bytes <= invit_lds) {
      k1 = kf;
      fused_lds = bytes;
    }
  }
  const int nb1 = tile_h ? div_up(w, tile_w) * div_up(h, tile_h) : div_up(n, k1);
  const int nb2 = div_up(n, kK2 * kK2Rows);
  int period = mixed ? kGraphSteps : kCheckIter;  // the plain path checks every 500 (lanczos.cu:519)
  const double stop_tol = mixed ? (double)tol * std::pow(10.0, -lanczos_mixed_stop) : (double)tol;
  const int nblk = std::min(div_up(n, 256), 1024);
  const size_t zb = mixed ? sizeof(double) : sizeof(float);

  // ---- solver state and small vectors (arena 2); offsets depend on n only
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    const size_t o = off;
    off += (bytes + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_st = carve(sizeof(LzState));
  const size_t o_z0 = carve(zb * vstride), o_z1 = carve(zb * vstride);
  const size_t o_rs = carve(sizeof(float) * vstride);
  const size_t o_p1 = carve(sizeof(double) * 2 * nb1), o_p2 = carve(sizeof(double) * 2 * nb1),
               o_p3 = carve(sizeof(double) * 2 * nb1);
  const size_t o_al = carve(zb * (kMaxIter + 2)), o_be = carve(zb * (kMaxIter + 2));
  const size_t o_ad = carve(sizeof(double) * (kMaxIter + 2)), o_bd = carve(sizeof(double) * (kMaxIter + 2));
  const size_t o_w = carve(sizeof(double) * 2 * kWStride);
  const int nvrows = mixed ? kMaxCand : kMaxEig;  // rows of S: candidates (mixed) or screened values
  const size_t o_S = carve(sizeof(double) * nvrows * (kMaxIter + 2));
  const size_t o_work = carve(sizeof(double) * 6 * nvrows * (kMaxIter + 2));
  const size_t o_resall = carve(sizeof(double) * kMaxCand);
  const size_t o_mm = carve(sizeof(float) * 2 * kMaxEig * nblk), o_fin = carve(sizeof(float) * 2 * kMaxEig);
  void* base;
  DMSC_TRY(ctx->scratch(2, off, &base));
  char* B = (char*)base;
  LzState* st = (LzState*)(B + o_st);
  float* rs = (float*)(B + o_rs);
  double* ad = (double*)(B + o_ad);
  double* bd = (double*)(B + o_bd);
  double* wv = (double*)(B + o_w);
  double* S = (double*)(B + o_S);

  // ---- basis (arena 3): a zero row, then q_0, q_1, ...; grown by whole periods
  void* vb;
  DMSC_TRY(ctx->scratch(3, vstride * sizeof(float) * (size_t)(period + 3), &vb));
  size_t cap = ctx->arena_bytes[3] / (vstride * sizeof(float)) - 1;  // rows after the zero row
  float* V = (float*)vb + vstride;

  Steps p;
  p.mixed = mixed;
  p.k1 = k1;
  p.fused_lds = fused_lds;
  p.tile_w = tile_w;
  p.tile_h = tile_h;
  p.loads = lanczos_step_loads;
  p.w = w;
  p.h = h;
  p.so = make_stencil_offsets(w);
  p.plain = LzBufs{A, (size_t)pitch, n, vstride, (float*)(B + o_z0), (float*)(B + o_z1), (double*)(B + o_p1),
                   (double*)(B + o_p2), (float*)(B + o_al), (float*)(B + o_be), nb1, nb2, st};
  p.mix = LzmBufs{A,
                  (size_t)pitch,
                  n,
                  vstride,
                  (double*)(B + o_z0),
                  (double*)(B + o_z1),
                  (double*)(B + o_p1),
                  (double*)(B + o_p2),
                  (double*)(B + o_p3),
                  (double*)(B + o_al),
                  (double*)(B + o_be),
                  nb1,
                  nb2,
                  st};

  if (convert) DMSC_TRY(launch_convert_matrix(s, w, h, A, pitch, nullptr, rs));
  float* mm = (float*)(B + o_mm);
  float* fin = (float*)(B + o_fin);
  // D^-1/2 scaling (stencilMVM.cu:182-198) and min-max normalisation of the unit eigenvectors
  auto finish = [&]() -> int {
    const dim3 g(div_up(n, 256));
    hipLaunchKernelGGL(k_scale_rs, g, dim3(256), 0, s, n, nvec, rs, d_eigvecs);
    hipLaunchKernelGGL(k_minmax1, dim3(nblk, nvec), dim3(256), 0, s, n, d_eigvecs, mm, nblk);
    hipLaunchKernelGGL(k_minmax2, dim3(nvec), dim3(256), 0, s, mm, nblk, fin);
    hipLaunchKernelGGL(k_normalize, g, dim3(256), 0, s, n, nvec, fin, d_eigvecs);
    DMSC_HIP(hipGetLastError());
    return DMSC_OK;
  };
  // the block solver (the mixed default) handles up to half its block; more
  // wanted vectors run on the mixed Lanczos solver
  if (mixed && eigensolver_method == 1 && nvec <= chebyshev_max_vectors() && time_steps == 0 && d_eigvecs) {
    double vals[kMaxEig];
    DMSC_TRY(chebyshev_solve(ctx, w, h, A, pitch, nvec, tol, d_eigvecs, vals, iterations));
    DMSC_TRY(finish());
    for (int j = 0; j < nvec; j++) h_eigvals[j] = (float)vals[j];
    return DMSC_OK;
  }
  const float q0 = (float)(1 / std::sqrt((double)n));  // lanczos.cu:64-72
  const dim3 gi(div_up(std::max(n, nb1), 256));
  hipLaunchKernelGGL(k_zero_row, dim3(div_up(n, 256)), dim3(256), 0, s, V - vstride, n);
  if (mixed) {
    const bool fused = p.fused_lds || p.tile_h;
    hipLaunchKernelGGL(k_init<double>, gi, dim3(256), 0, s, st, V, fused ? p.mix.z1 : p.mix.z, q0, n);
    if (fused)
      hipLaunchKernelGGL(k_lzmf_init, dim3(div_up(nb1, 256)), dim3(256), 0, s, p.mix.p1, p.mix.p2, p.mix.p3, nb1);
  } else {
    hipLaunchKernelGGL(k_init<float>, gi, dim3(256), 0, s, st, V, p.plain.z1, q0, n);
  }
  DMSC_HIP(hipGetLastError());

  if (time_steps > 0) {
    const int steps = std::min(time_steps, kMaxIter - kGraphSteps);
    if ((size_t)steps + kGraphSteps + 2 > cap) {
      void* nbuf;
      const size_t ncap = (size_t)steps + kGraphSteps + 2;
      DMSC_HIP(hipMalloc(&nbuf, vstride * sizeof(float) * (ncap + 1)));
      DMSC_HIP(hipMemcpyAsync(nbuf, V - vstride, vstride * sizeof(float) * 2, hipMemcpyDeviceToDevice, s));
      DMSC_TRY(ctx->adopt(3, nbuf, vstride * sizeof(float) * (ncap + 1)));
      V = (float*)nbuf + vstride;
      cap = ncap;
      hipLaunchKernelGGL(k_set_basis, dim3(1), dim3(1), 0, s, st, V);
    }
    // steps 0..24 build the graph and warm up; steps 25.. are timed
    DMSC_TRY(run_steps(ctx, p, st, w, 0, kGraphSteps));
    hipEvent_t e0, e1;
    DMSC_HIP(hipEventCreate(&e0));
    DMSC_HIP(hipEventCreate(&e1));
    DMSC_HIP(hipEventRecord(e0, s));
    for (int t = 0; t < steps; t += kGraphSteps)
      DMSC_TRY(run_steps(ctx, p, st, w, kGraphSteps + t, std::min(kGraphSteps, steps - t)));
    DMSC_HIP(hipEventRecord(e1, s));
    DMSC_HIP(hipEventSynchronize(e1));
    float ms = 0;
    DMSC_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_step = ms / (float)steps;
    return DMSC_OK;
  }

  int eigCheck = nvec + kCullumExtra;
  int done = 0, prev_m = 0, prev_n = 0;  // prev_n: Ritz values of the last check left in wv
  double prev_res = 0;
  LzState hs_plain{};
  LzState* hsp = &hs_plain;  // the host copy of the status record
  // The mixed path overlaps each check with the next Lanczos steps: the check
  // runs on the context's second stream, after an event recorded behind the
  // step that ends its period, while up to kSpecSteps further steps go to the
  // main stream.  Those steps only write basis rows and coefficients beyond
  // the check's order m, which the check never reads; if the check converges
  // they are discarded.  The status record comes back through pinned memory.
  hipStream_t cs = s;
  hipEvent_t ev_step = nullptr, ev_chk = nullptr;
  if (mixed) {
    DMSC_TRY(ctx->aux_stream(&cs, &ev_step, &ev_chk));
    int* pin;
    DMSC_TRY(ctx->pinned(&pin));
    hsp = reinterpret_cast<LzState*>(reinterpret_cast<char*>(pin) + 1024);
  }
  const LzState& hs = *hsp;
  int ahead = 0;  // steps already enqueued beyond `done`
  bool converged = false;
  // iteration cap: lanczos.cu:4 MAXITER, lowered by the "lanczos_max_iter" tuning value
  const int max_it = std::min(kMaxIter, lanczos_max_iter);
  while (done < max_it) {
    // the next check comes after max(period, ahead) steps
    const int steps = std::min(std::max(period, ahead), max_it - done);
    const int spec = mixed ? kSpecSteps : 0;
    if ((size_t)done + steps + spec + 2 > cap) {
      // grow the basis, keeping the zero row and every row already written:
      // q_0..q_{done+ahead} (the mixed path's K2 writes the row its next K1
      // reads; the plain path's K1 writes its own row, so one extra row is
      // copied harmlessly)
      const size_t ncap =
          std::min<size_t>(std::max(cap * 2, (size_t)done + steps + spec + 2), kMaxIter + kSpecSteps + 2);
      void* nbuf;
      DMSC_HIP(hipMalloc(&nbuf, vstride * sizeof(float) * (ncap + 1)));
      DMSC_HIP(hipMemcpyAsync(nbuf, V - vstride, vstride * sizeof(float) * (done + ahead + 2),
                              hipMemcpyDeviceToDevice, s));
      DMSC_TRY(ctx->adopt(3, nbuf, vstride * sizeof(float) * (ncap + 1)));  // synchronises before freeing
      V = (float*)nbuf + vstride;
      cap = ncap;
      hipLaunchKernelGGL(k_set_basis, dim3(1), dim3(1), 0, s, st, V);
    }
    for (int t = ahead; t < steps; t += kGraphSteps)
      DMSC_TRY(run_steps(ctx, p, st, w, done + t, std::min(kGraphSteps, steps - t)));
    done += steps;
    ahead = 0;
    if (mixed) {
      if (p.fused_lds || p.tile_h)
        hipLaunchKernelGGL(k_lzmf_finish, dim3(1), dim3(kRed), 0, s, p.mix, done);
      else
        hipLaunchKernelGGL(k_set_m, dim3(1), dim3(1), 0, s, st, done);  // alpha, beta[done-1] written by K2m
      hipLaunchKernelGGL(k_widen<double>, dim3(div_up(done, 256)), dim3(256), 0, s, p.mix.alpha, p.mix.beta, ad, bd,
                         done);
      DMSC_HIP(hipEventRecord(ev_step, s));
      DMSC_HIP(hipStreamWaitEvent(cs, ev_step, 0));
    } else {
      hipLaunchKernelGGL(k_lz_beta, dim3(1), dim3(kRed), 0, s, p.plain, done);
      hipLaunchKernelGGL(k_widen<float>, dim3(div_up(done, 256)), dim3(256), 0, s, p.plain.alpha, p.plain.beta, ad,
                         bd, done);
    }
    for (;;) {
      const int ncheck = std::min(eigCheck, done);
      hipLaunchKernelGGL(k_tri_bisect, dim3(ncheck, 2), dim3(kBisect), sizeof(double2) * done, cs, ad, bd, st, wv,
                         prev_n);
      prev_n = ncheck;
      const bool lds = (size_t)done * kInvitBytes <= invit_lds;
      const size_t vlds = lds ? (size_t)done * kInvitBytes : 0;
      if (mixed) {  // vectors and residuals of every candidate, then screening
        double* res_all = (double*)(B + o_resall);
        hipLaunchKernelGGL(k_tri_vec, dim3(ncheck), dim3(kWave), vlds, cs, ad, bd, st, 1, S, (double*)(B + o_work),
                           lds ? 1 : 0, (const double*)wv, res_all);
        hipLaunchKernelGGL(k_tri_screen_mixed, dim3(1), dim3(kWave), 0, cs, st, wv, res_all, ncheck, nvec);
      } else {  // lanczos.cu:279-384 screening, then vectors of the screened values
        hipLaunchKernelGGL(k_tri_screen, dim3(1), dim3(kWave), 0, cs, st, wv, ncheck, nvec);
        hipLaunchKernelGGL(k_tri_vec, dim3(nvec), dim3(kWave), vlds, cs, ad, bd, st, 0, S, (double*)(B + o_work),
                           lds ? 1 : 0, (const double*)nullptr, (double*)nullptr);
      }
      DMSC_HIP(hipGetLastError());
      DMSC_HIP(hipMemcpyAsync(hsp, st, sizeof(LzState), hipMemcpyDeviceToHost, cs));
      if (mixed) {
        DMSC_HIP(hipEventRecord(ev_chk, cs));
        if (ahead == 0) {  // steps past the check, overlapped with it
          ahead = std::min(kSpecSteps, max_it - done);
          for (int t = 0; t < ahead; t += kGraphSteps)
            DMSC_TRY(run_steps(ctx, p, st, w, done + t, std::min(kGraphSteps, ahead - t)));
        }
        DMSC_HIP(hipEventSynchronize(ev_chk));
      } else {
        DMSC_HIP(hipStreamSynchronize(s));
      }
      if (lanczos_trace) {
        double r = 0;
        for (int j = 0; j < std::min(hs.nscreened, nvec); j++) r = std::max(r, hs.res[j]);
        std::fprintf(stderr, "lanczos %s m=%d ncheck=%d screened=%d fail=%d maxres=%.3e vals=[%.6e .. %.6e]\n",
                     mixed ? "mixed" : "plain", done, ncheck, hs.nscreened, hs.fail, r, hs.vals[0],
                     hs.vals[std::max(0, std::min(hs.nscreened, nvec) - 1)]);
        if (lanczos_trace > 1 && hs.fail == 1) {
          std::vector<double> wh(2 * kWStride);
          DMSC_HIP(hipMemcpy(wh.data(), wv, sizeof(double) * 2 * kWStride, hipMemcpyDeviceToHost));
          for (int j = 0; j < ncheck; j++) std::fprintf(stderr, "  T %.9e  That %.9e\n", wh[j], wh[kWStride + j]);
        }
      }
      if (hs.fail != 1 || eigCheck >= done) break;
      eigCheck += kCullumExtra;  // lanczos.cu:528-532: widen the screened window
      if (eigCheck > (mixed ? kMaxCand : kCullumMax))
        return fail(DMSC_ERR_NOCONV, "Cullum screening kept %d of %d values at m=%d with %d candidates",
                    hs.nscreened, nvec, done, ncheck);
    }
    if (hs.fail == 2) return fail(DMSC_ERR_INTERNAL, "inverse iteration failed on the Lanczos tridiagonal");
    if (hs.fail == 0) {
      bool good = true;  // lanczos.cu:114-140 TestForConvergence
      for (int j = 0; j < nvec; j++)
        if (hs.res[j] > stop_tol) good = false;
      if (good) {
        converged = true;
        break;
      }
    }
    if (mixed) period = next_period(done, hs, nvec, stop_tol, prev_m, prev_res);
  }
  // lanczos.cu:519-560: the reference checks once more at MAXITER - 1 and,
  // converged or not, builds the Ritz vectors of that check; so does this
  // solver when the last check screened nvec values (an iteration count at
  // the cap tells the caller the tolerance was not reached)
  if (!converged && !(hs.fail == 0 && hs.nscreened >= nvec && hs.m >= max_it - kGraphSteps))
    return fail(DMSC_ERR_NOCONV, "Lanczos did not converge in %d iterations", max_it);
  const int m = hs.m;
  {
    double r = 0;
    for (int j = 0; j < nvec; j++) r = std::max(r, hs.res[j]);
    ctx->eig_stop = converged ? DMSC_STOP_RESIDUAL : DMSC_STOP_CAP;
    ctx->eig_residual = r;
  }

  // Ritz vectors, D^-1/2 scaling (stencilMVM.cu:182-198), min-max normalisation
  DMSC_TRY(launch_ritz(s, mixed, nvec, n, st, vstride, m, S, d_eigvecs));
  DMSC_TRY(finish());
  for (int j = 0; j < nvec; j++) h_eigvals[j] = (float)hs.vals[j];
  if (iterations) *iterations = m;
  return DMSC_OK;
}

}  // namespace

int run_eigensolve(Ctx* ctx, int w, int h, float* A, int pitch, int nvec, float tol, int precision,
                   float* h_eigvals, float* d_eigvecs, int* iterations) {
  if (nvec < 1 || nvec > kMaxEig) return fail(DMSC_ERR_ARG, "nvec must be in [1, %d]", kMaxEig);
  if (pitch % 4) return fail(DMSC_ERR_ARG, "matrix pitch must be a multiple of 4 floats");
  return solve(ctx, w, h, A, pitch, nvec, tol, precision == DMSC_PREC_MIXED, h_eigvals, d_eigvecs, iterations);
}

int time_lanczos_steps(Ctx* ctx, int w, int h, float* A, int pitch, int precision, int steps, float* ms_per_step) {
  if (pitch % 4) return fail(DMSC_ERR_ARG, "matrix pitch must be a multiple of 4 floats");
  if (steps < kGraphSteps || steps % kGraphSteps)
    return fail(DMSC_ERR_ARG, "steps must be a positive multiple of %d", kGraphSteps);
  return solve(ctx, w, h, A, pitch, 9, 1e-3f, precision == DMSC_PREC_MIXED, nullptr, nullptr, nullptr, steps,
               ms_per_step);
}

// Times every step-kernel shape for this size and precision (25 warm-up
// and `steps` timed steps each, on the matrix converted once in place),
// records the fastest as the size's shape and returns every trial.
int autotune_steps(Ctx* ctx, int w, int h, float* A, int pitch, int precision, int steps, StepShape* best,
                   StepShape* tried, int max_tried, int* n_tried) {
  if (pitch % 4) return fail(DMSC_ERR_ARG, "matrix pitch must be a multiple of 4 floats");
  if (steps < kGraphSteps || steps % kGraphSteps)
    return fail(DMSC_ERR_ARG, "steps must be a positive multiple of %d", kGraphSteps);
  const bool mixed = precision == DMSC_PREC_MIXED;
  std::vector<StepShape> cand;
  if (mixed) {
    for (int tile : kStepTiles) cand.push_back(StepShape{2, tile, 0, 0.f});
    for (int k1 : {512, 256}) cand.push_back(StepShape{1, 0, k1, 0.f});
    for (int k1 : {256, 64}) cand.push_back(StepShape{0, 0, k1, 0.f});
  } else {
    for (int k1 : {256, 64}) cand.push_back(StepShape{0, 0, k1, 0.f});
  }
  int nt = 0;
  StepShape top{};
  top.ms = 0.f;
  for (StepShape c : cand) {
    float ms = 0.f;
    DMSC_TRY(solve(ctx, w, h, A, pitch, 9, 1e-3f, mixed, nullptr, nullptr, nullptr, steps, &ms, &c, nt == 0));
    c.ms = ms;
    if (tried && nt < max_tried) tried[nt] = c;
    nt++;
    if (top.ms == 0.f || ms < top.ms) top = c;
  }
  {
    std::lock_guard<std::mutex> lk(tuned_lock());
    tuned_shapes()[std::make_tuple(w, h, mixed ? 1 : 0)] = top;
  }
  if (best) *best = top;
  if (n_tried) *n_tried = nt;
  return DMSC_OK;
}

void autotune_clear() {
  std::lock_guard<std::mutex> lk(tuned_lock());
  tuned_shapes().clear();
}

void release_eigensolver(const Ctx* ctx) {
  std::unique_ptr<GraphCache> gone;  // destroyed outside the lock
  std::lock_guard<std::mutex> lk(graph_lock());
  auto it = graph_caches().find(ctx);
  if (it == graph_caches().end()) return;
  gone = std::move(it->second);
  graph_caches().erase(it);
}

}  // namespace dmsc
