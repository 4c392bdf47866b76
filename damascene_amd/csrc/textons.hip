// textons.hip -- texton filter bank and k-means (reference: textons/texton.cu,
// textons/kmeans.cu).
//
// Filter bank: exact FP32 correlation in the reference order (x offset
// outer, y offset inner, fused multiply-add), clamp-to-edge borders.  Each
// block stages a clamped image tile with halo in LDS; each thread computes 8
// vertically adjacent outputs so one column of tile values feeds 8
// independent accumulation chains.  Coefficients are wave-uniform (scalar
// loads).
//
// k-means: per iteration one accumulation kernel (int32 wrap-around centroid
// sums via LDS atomics, exact and order independent) and one assignment
// kernel (every block forms the centroids from the sums in LDS while its
// pixels' responses load, then 32 or 64 distances per pixel in the
// reference order, argmin with strict '<').  A device-side done flag lets
// all iterations be enqueued without host round trips.
#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {
namespace {

constexpr int TX = 64, TY = 32, RPT = 8;  // tile 64 x 32, 8 rows per thread

template <int R>
__global__ __launch_bounds__(256) void k_texton_conv(int w, int h, const float* __restrict__ grey,
                                                     const float* __restrict__ coeffs,
                                                     int first_filter, float* __restrict__ out) {
  constexpr int LW = TX + 2 * R, LH = TY + 2 * R, D = 2 * R + 1;
  __shared__ float tile[LH][LW + 1];
  const int f = first_filter + blockIdx.z;
  const int x0 = blockIdx.x * TX, y0 = blockIdx.y * TY;
  for (int i = threadIdx.x; i < LW * LH; i += 256) {
    const int ly = i / LW, lx = i - ly * LW;
    const int gx = dev::clampi(x0 + lx - R, 0, w - 1), gy = dev::clampi(y0 + ly - R, 0, h - 1);
    tile[ly][lx] = grey[(size_t)gy * w + gx];
  }
  __syncthreads();
  const int tx = threadIdx.x & 63, ty = (threadIdx.x >> 6) * RPT;
  const float* c = coeffs + (size_t)(f < 17 ? f * 169 : 17 * 169 + (f - 17) * 361);
  float acc[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++) acc[k] = 0.0f;
#pragma unroll 1
  for (int dx = 0; dx < D; dx++) {
    float col[RPT + 2 * R];
#pragma unroll
    for (int j = 0; j < RPT + 2 * R; j++) col[j] = tile[ty + j][tx + dx];
#pragma unroll
    for (int dy = 0; dy < D; dy++) {
      const float cf = c[dx * D + dy];
#pragma unroll
      for (int k = 0; k < RPT; k++) acc[k] = fmaf(col[k + dy], cf, acc[k]);
    }
  }
  const int gx = x0 + tx;
  if (gx < w) {
    const size_t n = (size_t)w * h;
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const int gy = y0 + ty + k;
      if (gy < h) out[(size_t)f * n + (size_t)gy * w + gx] = acc[k];
    }
  }
}

__global__ void k_zero(float* p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

// ------------------------------------------------------------------ k-means
// Cluster count K is 32 (texton_choice 1) or 64 (texton_choice 2); the
// state is sized for 64 and indexed with stride K.
constexpr int kMaxK = 64;

struct KmState {
  uint32_t mass[2][kNFilters * kMaxK];  // [buffer][filter][cluster], int32 wrap-around sums
  uint32_t count[2][kMaxK];
  int changes[2];
  unsigned int done;        // set when an iteration changed no label
  unsigned int iterations;  // iterations executed
  unsigned int rechecks;    // mixed path: pixels re-labelled by the exact FP32 distances
};

// kmeans.cu:14-35 assignInitialClusters (6x6 grid capped at 31) for K = 32,
// kmeans.cu:37-55 assignInitialClusters_64 (8x8 grid) for K = 64, and
// kmeans.cu:296-307 makeSelfDots over the float responses.
template <int K>
__global__ void k_km_init(int w, int h, int nf, const float* __restrict__ resp, int* __restrict__ labels,
                          float* __restrict__ pdot, KmState* st) {
  const int n = w * h;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {
    for (int b = 0; b < 2; b++) st->changes[b] = 0;
    st->done = 0;
    st->iterations = 0;
    st->rechecks = 0;
  }
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < kNFilters * K; i += blockDim.x) { st->mass[0][i] = 0; st->mass[1][i] = 0; }
    if (threadIdx.x < K) { st->count[0][threadIdx.x] = 0; st->count[1][threadIdx.x] = 0; }
  }
  if (p >= n) return;
  const int y = p / w, x = p - y * w;
  constexpr int G = K == 64 ? 8 : 6;
  const int c = (y / ((h - 1) / G + 1)) * G + x / ((w - 1) / G + 1);
  labels[p] = (K == 32 && c >= 32) ? 31 : c;
  float d = 0.0f;
  for (int f = 0; f < nf; f++) {
    const float r = resp[(size_t)f * n + p];
    d = fmaf(r, r, d);
  }
  pdot[p] = d;
}

// A k-means iteration accumulates the int32 sums of the current labels,
// then assigns: every assignment block forms the centroids from the sums.
// The mixed path accumulates only in its first iteration: its assignment
// also adds each pixel's responses, by the new label, to the next
// iteration's sums, so each later iteration is a small clear kernel and the
// assignment, and reads the responses once.
// The stop test (kmeans.cu:712: no label changed) is taken at the start of
// the next iteration's kernels, from the change count the previous
// assignment left; every block reads the same count, so all agree.
__device__ __forceinline__ bool km_stopped(KmState* st, int it) {
  if (st->done) return true;
  if (it > 0 && st->changes[(it - 1) & 1] <= 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
    return true;
  }
  return false;
}

// kmeans.cu:57-103 findCentroids / 106-153 findCentroids_64: int32 sums of
// (int)(1e5 * response) (INTCONFACTOR, kmeans.h:16), wrap-around
// arithmetic (order independent).
template <int K>
__global__ __launch_bounds__(256) void k_km_accum(int n, int nf, const float* __restrict__ resp,
                                                  const int* __restrict__ labels, KmState* st, int it) {
  if (km_stopped(st, it)) return;
  __shared__ uint32_t smass[kNFilters * K];
  __shared__ uint32_t scount[K];
  for (int i = threadIdx.x; i < nf * K; i += 256) smass[i] = 0;
  if (threadIdx.x < K) scount[threadIdx.x] = 0;
  const int buf = it & 1;
  if (blockIdx.x == 0) {
    // the next iteration's sums (last read by the previous assignment) and
    // this iteration's change count (last read by this kernel's stop test,
    // for iteration it - 2 ... it - 1 of the other parity): cleared here,
    // before the assignment below adds to it
    for (int i = threadIdx.x; i < kNFilters * K; i += 256) st->mass[buf ^ 1][i] = 0;
    if (threadIdx.x < K) st->count[buf ^ 1][threadIdx.x] = 0;
    if (threadIdx.x == 0) st->changes[buf] = 0;
  }
  __syncthreads();
  for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
    const int c = labels[p];
    // all 34 responses are loaded before the first atomic, so their loads
    // are in flight together
    uint32_t v[kNFilters];
#pragma unroll
    for (int f = 0; f < kNFilters; f++) v[f] = (uint32_t)(int)(100000.0f * resp[(size_t)f * n + p]);
    if (c < 0 || c >= K) continue;  // unassigned pixel (all distances >= 1e7)
#pragma unroll
    for (int f = 0; f < kNFilters; f++) atomicAdd(&smass[f * K + c], v[f]);
    atomicAdd(&scount[c], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nf * K; i += 256)
    if (smass[i]) atomicAdd(&st->mass[buf][i], smass[i]);
  if (threadIdx.x < K && scount[threadIdx.x]) atomicAdd(&st->count[buf][threadIdx.x], scount[threadIdx.x]);
}

// Iterations after the first: the next iteration's sums (read last as
// centroids by the previous assignment) and this iteration's change count
// (read last by the previous assignment's stop test) cleared before this
// iteration's assignment adds to them.
template <int K>
__global__ __launch_bounds__(256) void k_km_clear(KmState* st, int it) {
  if (st->done) return;
  const int nb = (it + 1) & 1;
  for (int i = threadIdx.x; i < kNFilters * K; i += 256) st->mass[nb][i] = 0;
  if (threadIdx.x < K) st->count[nb][threadIdx.x] = 0;
  if (threadIdx.x == 0) st->changes[it & 1] = 0;
}

// One pixel's int32 contributions (kmeans.cu findCentroids: (int)(1e5 r))
// to the block's sums of cluster c, and the block's sums added to the next
// iteration's buffer.
__device__ __forceinline__ void km_add(uint32_t* smass, int K, int f, int c, float r) {
  atomicAdd(&smass[f * K + c], (uint32_t)(int)(100000.0f * r));
}
template <int K>
__device__ __forceinline__ void km_flush(const uint32_t* smass, const uint32_t* scount, KmState* st, int nf, int it) {
  const int nb = (it + 1) & 1;
  for (int i = threadIdx.x; i < nf * K; i += 256)
    if (smass[i]) atomicAdd(&st->mass[nb][i], smass[i]);
  if (threadIdx.x < K && scount[threadIdx.x]) atomicAdd(&st->count[nb][threadIdx.x], scount[threadIdx.x]);
}

__device__ __forceinline__ uint16_t to_bf16(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// x ~ hi + lo with hi = bf16(x), lo = bf16(x - hi) (the subtraction is exact)
__device__ __forceinline__ void split_bf16(float x, uint16_t& hi, uint16_t& lo) {
  hi = to_bf16(x);
  lo = to_bf16(x - from_bf16(hi));
}

// Centroids and their norms, computed in LDS by every assignment block.
template <int K>
struct KmCent {
  float cent[kNFilters][K];
  float cdot[K];
  float cnorm[K];  // |c| (mixed path)
};

// kmeans.cu:268-286 finishCentroids(_64) (an empty cluster gives NaN, which
// never wins a comparison) and makeSelfDots for the centroids, from the
// iteration's int32 sums: every block of the assignment computes the same
// values with the same operations, so they agree bit for bit (there is no
// separate centroid kernel between accumulation and assignment).
template <int K>
__device__ __forceinline__ void km_centroids(KmCent<K>& sc, const KmState* st, int nf, int buf) {
  for (int i = threadIdx.x; i < kNFilters * K; i += 256) {
    const int f = i / K, c = i - f * K;
    sc.cent[f][c] = f < nf ? (float)(int)st->mass[buf][i] / (100000.0f * (float)st->count[buf][c]) : 0.0f;
  }
  __syncthreads();
  if (threadIdx.x < K) {
    const int c = threadIdx.x;
    float d = 0.0f;
    for (int f = 0; f < nf; f++) d = fmaf(sc.cent[f][c], sc.cent[f][c], d);
    sc.cdot[c] = d;
    sc.cnorm[c] = sqrtf(d);
  }
}

// Exact FP32 label of one pixel: kmeans.cu:296-338 makeDots + SGEMM(alpha
// -2, beta 1) + 447-472 findDiffLabels: dist_c = fma(-2, x.c, |x|^2 +
// |c|^2), x.c accumulated over the filters in order, strict '<' from 1e7.
template <int K>
__device__ __forceinline__ int km_exact_label(const float (&x)[kNFilters], float pd, const KmCent<K>& sc) {
  float best = 10000000.0f;
  int bl = -1;
  for (int c = 0; c < K; c++) {
    float acc = 0.0f;
#pragma unroll
    for (int f = 0; f < kNFilters; f++) acc = fmaf(x[f], sc.cent[f][c], acc);
    const float d = fmaf(-2.0f, acc, pd + sc.cdot[c]);
    if (d < best) {
      best = d;
      bl = c;
    }
  }
  return bl;
}

// FUSE: also add each pixel's responses to the next iteration's sums (the
// mixed path's scheme); the plain path accumulates in k_km_accum, whose
// many blocks of one pixel per thread measured faster than the atomics
// added to this kernel's tail.
template <int K, bool FUSE>
__global__ __launch_bounds__(256) void k_km_assign(int n, int nf, const float* __restrict__ resp,
                                                   const float* __restrict__ pdot, int* __restrict__ labels,
                                                   KmState* st, int it) {
  if (km_stopped(st, it)) return;
  __shared__ KmCent<K> sc;
  __shared__ uint32_t smass[kNFilters * K], scount[K];
  __shared__ int schanges;
  const int p = blockIdx.x * 256 + threadIdx.x;
  const int pp = p < n ? p : n - 1;
  // the pixel's responses are in flight while the block forms the centroids
  float x[kNFilters];
#pragma unroll
  for (int f = 0; f < kNFilters; f++) x[f] = resp[(size_t)f * n + pp];
  const float pd = pdot[pp];
  const int old = labels[pp];
  if (threadIdx.x == 0) schanges = 0;
  for (int i = threadIdx.x; i < kNFilters * K; i += 256) smass[i] = 0;
  if (threadIdx.x < K) scount[threadIdx.x] = 0;
  km_centroids<K>(sc, st, nf, it & 1);
  __syncthreads();
  int changed = 0;
  if (p < n) {
    const int bl = km_exact_label<K>(x, pd, sc);
    changed = bl != old;
    labels[p] = bl;
    if (FUSE && bl >= 0) {  // -1: no finite distance, in no cluster
#pragma unroll
      for (int f = 0; f < kNFilters; f++) km_add(smass, K, f, bl, x[f]);
      atomicAdd(&scount[bl], 1u);
    }
  }
  const unsigned long long m = __ballot(changed);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&schanges, __popcll(m));
  __syncthreads();
  if (FUSE) km_flush<K>(smass, scount, st, nf, it);
  if (threadIdx.x == 0) {
    if (schanges) atomicAdd(&st->changes[it & 1], schanges);
    if (blockIdx.x == 0) st->iterations = it + 1;
  }
}

// ---------------------------------------------------------------- mixed path
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Distances from v_mfma_f32_16x16x32_bf16 products (clusters as rows, in
// K/16 tiles of 16, pixels as columns, filters 0..31 as K) plus the FP32
// filter-32 term.  Both operands are split, x ~ x_hi + x_lo and c ~ c_hi +
// c_lo (BF16 each, split in registers from the FP32 responses), and x.c is
// taken as x_lo c_hi + x_hi c_lo + x_hi c_hi
// (three MFMAs into one FP32 accumulator).  Per (pixel, cluster) the error
// against the exact FP32 distance is below e = 2^-14 |x| |c| + 2^-20 (|x|^2
// + |c|^2): the split leaves at most 3.01 * 2^-18 |x_f c_f| per product
// (the dropped x_lo c_lo and the BF16 rounding of the low parts), the 96
// FP32 additions of the MFMAs at most 2^-17.4 sum |x_f c_f| and the
// reference-order FP32 dot product at most 2^-19 sum |x_f c_f|, so with
// sum |x_f c_f| <= |x||c| (Cauchy-Schwarz) and the factor 2 of the
// distance below 2^-14.7 |x||c|; FP32 rounding of |x|^2 + |c|^2 - 2 x.c in
// both computations stays inside the second term.  The label of the
// smallest approximate distance is the exact label when its upper bound
// lies below every other cluster's lower bound (no tie is possible then);
// any other pixel is re-labelled with the exact FP32 distances in the
// reference order, so labels are bit-exact.
// Lane layout: A row l%16, B column l%16, K slice 8*(l/16)..+7; result rows
// 4*(l/16)+i of column l%16.
template <int K>
__global__ __launch_bounds__(256) void k_km_assign_mfma(int n, int nf, const float* __restrict__ resp,
                                                        const float* __restrict__ pdot, int* __restrict__ labels,
                                                        KmState* st, int it) {
  constexpr int NT = K / 16;  // cluster tiles
  static_assert(kNFilters == 34, "filters 0..31 on the four lane groups, 32 and 33 on groups 0 and 1");
  if (km_stopped(st, it)) return;
  __shared__ KmCent<K> sc;
  __shared__ uint32_t smass[kNFilters * K], scount[K];
  __shared__ __attribute__((aligned(16))) uint16_t scb[K][32], scbl[K][32];  // BF16 split of filters 0..31
  __shared__ int schanges, srecheck;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  // each wave labels 64 pixels, 16 per MFMA column tile; every tile's
  // operands are in flight while the block forms the centroids
  float xv[4][8];
  float pdv[4], x32v[4];
  int oldl[4];  // the labels being replaced (for the change count)
#pragma unroll
  for (int tile = 0; tile < 4; tile++) {
    const int p = (blockIdx.x * 4 + wave) * 64 + tile * 16 + col;
    const int pp = p < n ? p : n - 1;
#pragma unroll
    for (int j = 0; j < 8; j++) xv[tile][j] = resp[(size_t)(8 * grp + j) * n + pp];
    pdv[tile] = pdot[pp];
    x32v[tile] = resp[(size_t)32 * n + pp];
    oldl[tile] = labels[pp];
  }
  if (threadIdx.x == 0) schanges = srecheck = 0;
  for (int i = threadIdx.x; i < kNFilters * K; i += 256) smass[i] = 0;
  if (threadIdx.x < K) scount[threadIdx.x] = 0;
  km_centroids<K>(sc, st, nf, it & 1);
  for (int i = threadIdx.x; i < K * 32; i += 256) {
    const int c = i / 32, f = i - c * 32;
    split_bf16(sc.cent[f][c], scb[c][f], scbl[c][f]);
  }
  __syncthreads();
  // A operands: centroid rows (clusters) 16*t + col, filters 8*grp..+7
  bf16x8 a[NT], al[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const uint4 v = *reinterpret_cast<const uint4*>(&scb[16 * t + col][8 * grp]);
    __builtin_memcpy(&a[t], &v, 16);
    const uint4 vl = *reinterpret_cast<const uint4*>(&scbl[16 * t + col][8 * grp]);
    __builtin_memcpy(&al[t], &vl, 16);
  }
  int changed = 0, rech = 0;
#pragma unroll
  for (int tile = 0; tile < 4; tile++) {
    const int p = (blockIdx.x * 4 + wave) * 64 + tile * 16 + col;
    const bool live = p < n;
    uint16_t hb[8], lb[8];
#pragma unroll
    for (int j = 0; j < 8; j++) split_bf16(xv[tile][j], hb[j], lb[j]);
    bf16x8 b, bl;
    __builtin_memcpy(&b, hb, 16);
    __builtin_memcpy(&bl, lb, 16);
    f32x4 s[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {  // clusters 16t..16t+15; small terms first
      s[t] = f32x4{0, 0, 0, 0};
      s[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], bl, s[t], 0, 0, 0);
      s[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[t], b, s[t], 0, 0, 0);
      s[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], b, s[t], 0, 0, 0);
    }
    const float pd = pdv[tile], x32 = x32v[tile];
    const float xn = sqrtf(pd);
    // this lane's 4*NT clusters, in increasing order: 16t + 4*grp + i
    float dt[4 * NT];
    int ci[4 * NT];
#pragma unroll
    for (int i = 0; i < 4 * NT; i++) {
      const int c = 16 * (i >> 2) + 4 * grp + (i & 3);
      const float sv = s[i >> 2][i & 3] + x32 * sc.cent[32][c];
      dt[i] = (pd + sc.cdot[c]) - 2.0f * sv;
      ci[i] = c;
    }
    float best = INFINITY;
    int bc = 2 * K;
#pragma unroll
    for (int i = 0; i < 4 * NT; i++)
      if (dt[i] < best) { best = dt[i]; bc = ci[i]; }
#pragma unroll
    for (int o = 16; o <= 32; o <<= 1) {  // the pixel's 4 lanes: smaller distance, then lower index
      const float ob = __shfl_xor(best, o, 64);
      const int oc = __shfl_xor(bc, o, 64);
      if (ob < best || (ob == best && oc < bc)) { best = ob; bc = oc; }
    }
    float lo_other = INFINITY;
#pragma unroll
    for (int i = 0; i < 4 * NT; i++)
      if (ci[i] != bc) {
        const float e = 6.103515625e-05f * xn * sc.cnorm[ci[i]] + 9.5367431640625e-7f * (pd + sc.cdot[ci[i]]);
        lo_other = fminf(lo_other, dt[i] - e);
      }
#pragma unroll
    for (int o = 16; o <= 32; o <<= 1) lo_other = fminf(lo_other, __shfl_xor(lo_other, o, 64));
    bool certain = false;
    if (bc < K) {
      const float e = 6.103515625e-05f * xn * sc.cnorm[bc] + 9.5367431640625e-7f * (pd + sc.cdot[bc]);
      certain = best + e < lo_other && best + e < 10000000.0f;
    }
    int fl = -1;
    if (grp == 0 && live) {
      fl = bc;
      if (!certain) {  // near tie (or no finite distance): exact FP32 distances
        float x[kNFilters];
#pragma unroll
        for (int f = 0; f < kNFilters; f++) x[f] = resp[(size_t)f * n + p];
        fl = km_exact_label<K>(x, pd, sc);
        rech++;
      }
      changed += fl != oldl[tile];
      labels[p] = fl;
    }
    // the pixel's sums by its new label: filters 8 grp .. 8 grp + 7 on each
    // of its four lanes, filter 32 (and the count) on group 0, 33 on group 1
    const int lab = __shfl(fl, col, 64);
    if (live && lab >= 0) {
#pragma unroll
      for (int j = 0; j < 8; j++) km_add(smass, K, 8 * grp + j, lab, xv[tile][j]);
      if (grp == 0) {
        km_add(smass, K, 32, lab, x32);
        atomicAdd(&scount[lab], 1u);
      } else if (grp == 1) {
        km_add(smass, K, 33, lab, resp[(size_t)33 * n + p]);
      }
    }
  }
  if (changed) atomicAdd(&schanges, changed);
  if (rech) atomicAdd(&srecheck, rech);
  __syncthreads();
  km_flush<K>(smass, scount, st, nf, it);
  if (threadIdx.x == 0) {
    if (schanges) atomicAdd(&st->changes[it & 1], schanges);
    if (srecheck) atomicAdd(&st->rechecks, (unsigned)srecheck);
    if (blockIdx.x == 0) st->iterations = it + 1;
  }
}

}  // namespace

int launch_texton_conv(hipStream_t s, int w, int h, const float* grey, const float* d_coeffs, float* resp) {
  dim3 grid(div_up(w, TX), div_up(h, TY), 17);
  hipLaunchKernelGGL(k_texton_conv<6>, grid, dim3(256), 0, s, w, h, grey, d_coeffs, 0, resp);
  grid.z = 16;
  hipLaunchKernelGGL(k_texton_conv<9>, grid, dim3(256), 0, s, w, h, grey, d_coeffs, 17, resp);
  const size_t n = (size_t)w * h;
  hipLaunchKernelGGL(k_zero, dim3(div_up(n, 256)), dim3(256), 0, s, resp + 33 * n, n);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

size_t kmeans_state_bytes() { return sizeof(KmState); }

namespace {
template <int K>
void km_iterations(hipStream_t s, int n, int nf, const float* resp, int* labels, const float* pdot, KmState* st,
                   int max_iter, bool mixed) {
  const int blocks_acc = std::min(div_up(n, 256), 1024);
  for (int it = 0; it < max_iter; it++) {
    if (it == 0 || !mixed)
      hipLaunchKernelGGL(k_km_accum<K>, dim3(blocks_acc), dim3(256), 0, s, n, nf, resp, labels, st, it);
    else
      hipLaunchKernelGGL(k_km_clear<K>, dim3(1), dim3(256), 0, s, st, it);
    if (mixed) {
      hipLaunchKernelGGL(k_km_assign_mfma<K>, dim3(div_up(n, 256)), dim3(256), 0, s, n, nf, resp, pdot, labels, st,
                         it);
    } else
      hipLaunchKernelGGL((k_km_assign<K, false>), dim3(div_up(n, 256)), dim3(256), 0, s, n, nf, resp, pdot, labels, st,
                         it);
  }
}

template <int K>
int run_kmeans(hipStream_t s, int w, int h, int nf, const float* resp, int* labels, float* pdot, KmState* st,
               int max_iter, bool mixed) {
  const int n = w * h;
  hipLaunchKernelGGL(k_km_init<K>, dim3(div_up(n, 256)), dim3(256), 0, s, w, h, nf, resp, labels, pdot, st);
  km_iterations<K>(s, n, nf, resp, labels, pdot, st, max_iter, mixed);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}
}  // namespace

int launch_kmeans_mixed(hipStream_t s, int w, int h, int k, const float* resp, int* labels, float* pdot,
                        void* state, int max_iter) {
  KmState* st = (KmState*)state;
  if ((uintptr_t)state % 16) return fail(DMSC_ERR_INTERNAL, "k-means state misaligned");
  if (k == 64) return run_kmeans<64>(s, w, h, kNFilters, resp, labels, pdot, st, max_iter, true);
  if (k == 32) return run_kmeans<32>(s, w, h, kNFilters, resp, labels, pdot, st, max_iter, true);
  return fail(DMSC_ERR_ARG, "k-means supports 32 or 64 clusters");
}

int launch_kmeans(hipStream_t s, int w, int h, int k, int nf, const float* resp, int* labels,
                  float* pdot, void* state, int max_iter) {
  KmState* st = (KmState*)state;
  if (k == 64) return run_kmeans<64>(s, w, h, nf, resp, labels, pdot, st, max_iter, false);
  if (k == 32) return run_kmeans<32>(s, w, h, nf, resp, labels, pdot, st, max_iter, false);
  return fail(DMSC_ERR_ARG, "k-means supports 32 or 64 clusters");
}

const int* kmeans_iterations_ptr(const void* state) {
  return reinterpret_cast<const int*>(&static_cast<const KmState*>(state)->iterations);
}

int kmeans_read_state(hipStream_t s, void* state, int* changes, int* iterations, int* rechecks) {
  KmState h;
  DMSC_HIP(hipMemcpyAsync(&h, state, sizeof(KmState), hipMemcpyDeviceToHost, s));
  DMSC_HIP(hipStreamSynchronize(s));
  const int last = (int)h.iterations - 1;
  if (iterations) *iterations = (int)h.iterations;
  if (changes) *changes = last >= 0 ? h.changes[last & 1] : 0;
  if (rechecks) *rechecks = (int)h.rechecks;
  return DMSC_OK;
}

}  // namespace dmsc
