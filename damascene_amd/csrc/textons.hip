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
// kernel (centroids rebuilt per block, 32 distances per pixel in the
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
struct KmState {
  uint32_t mass[2][kNFilters * 64];  // [buffer][filter][cluster]
  uint32_t count[2][64];
  int changes[2];
  unsigned int done;        // set when an iteration changed no label
  unsigned int iterations;  // iterations executed
  unsigned int blocks_done[2];
};

// kmeans.cu:14-35 assignInitialClusters (6x6 grid capped at 31) and
// kmeans.cu:296-307 makeSelfDots over the float responses.
__global__ void k_km_init(int w, int h, int nf, const float* __restrict__ resp, int* __restrict__ labels,
                          float* __restrict__ pdot, KmState* st) {
  const int n = w * h;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {
    for (int b = 0; b < 2; b++) { st->changes[b] = 0; st->blocks_done[b] = 0; }
    st->done = 0;
    st->iterations = 0;
  }
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < kNFilters * 64; i += blockDim.x) { st->mass[0][i] = 0; st->mass[1][i] = 0; }
  if (blockIdx.x == 0 && threadIdx.x < 64) { st->count[0][threadIdx.x] = 0; st->count[1][threadIdx.x] = 0; }
  if (p >= n) return;
  const int y = p / w, x = p - y * w;
  int c = (y / ((h - 1) / 6 + 1)) * 6 + x / ((w - 1) / 6 + 1);
  labels[p] = c >= 32 ? 31 : c;
  float d = 0.0f;
  for (int f = 0; f < nf; f++) {
    const float r = resp[(size_t)f * n + p];
    d = fmaf(r, r, d);
  }
  pdot[p] = d;
}

// kmeans.cu:57-103 findCentroids: int32 sums of (int)(1e5 * response)
// (INTCONFACTOR, kmeans.h:16), wrap-around arithmetic.
template <int K>
__global__ __launch_bounds__(256) void k_km_accum(int n, int nf, const float* __restrict__ resp,
                                                  const int* __restrict__ labels, KmState* st, int it) {
  if (st->done) return;
  __shared__ uint32_t smass[kNFilters * K];
  __shared__ uint32_t scount[K];
  for (int i = threadIdx.x; i < nf * K; i += 256) smass[i] = 0;
  if (threadIdx.x < K) scount[threadIdx.x] = 0;
  __syncthreads();
  const int buf = it & 1;
  for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
    const int c = labels[p];
    if (c < 0 || c >= K) continue;  // unassigned pixel (all distances >= 1e7)
#pragma unroll
    for (int f = 0; f < kNFilters; f++)
      atomicAdd(&smass[f * K + c], (uint32_t)(int)(100000.0f * resp[(size_t)f * n + p]));
    atomicAdd(&scount[c], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nf * K; i += 256)
    if (smass[i]) atomicAdd(&st->mass[buf][i], smass[i]);
  if (threadIdx.x < K && scount[threadIdx.x]) atomicAdd(&st->count[buf][threadIdx.x], scount[threadIdx.x]);
}

// kmeans.cu:268-276 finishCentroids + 296-338 makeSelfDots/makeDots +
// SGEMM(alpha=-2, beta=1) + 447-472 findDiffLabels, fused: every block
// rebuilds the 34x32 centroids from the integer sums, then each thread
// labels one pixel: dist_c = fma(-2, x.c, |x|^2 + |c|^2), strict '<' from 1e7.
template <int K>
__global__ __launch_bounds__(256) void k_km_assign(int n, int nf, const float* __restrict__ resp,
                                                   const float* __restrict__ pdot, int* __restrict__ labels,
                                                   KmState* st, int it) {
  if (st->done) return;
  __shared__ float cent[kNFilters][K];
  __shared__ float cdot[K];
  __shared__ int schanges;
  const int buf = it & 1;
  for (int i = threadIdx.x; i < nf * K; i += 256) {
    const int f = i / K, c = i - f * K;
    cent[f][c] = (float)(int)st->mass[buf][i] / (100000.0f * (float)st->count[buf][c]);
  }
  if (threadIdx.x == 0) schanges = 0;
  __syncthreads();
  if (threadIdx.x < K) {
    float d = 0.0f;
    for (int f = 0; f < nf; f++) d = fmaf(cent[f][threadIdx.x], cent[f][threadIdx.x], d);
    cdot[threadIdx.x] = d;
  }
  __syncthreads();
  int local = 0;
  for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
    float x[kNFilters];
#pragma unroll
    for (int f = 0; f < kNFilters; f++) x[f] = resp[(size_t)f * n + p];
    const float pd = pdot[p];
    float best = 10000000.0f;
    int bl = -1;
    for (int c = 0; c < K; c++) {
      float acc = 0.0f;
#pragma unroll
      for (int f = 0; f < kNFilters; f++) acc = fmaf(x[f], cent[f][c], acc);
      const float d = fmaf(-2.0f, acc, pd + cdot[c]);
      if (d < best) { best = d; bl = c; }
    }
    if (bl != labels[p]) local++;
    labels[p] = bl;
  }
  if (local) atomicAdd(&schanges, local);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (schanges) atomicAdd(&st->changes[buf], schanges);
    __threadfence();
    const unsigned int ticket = atomicAdd(&st->blocks_done[buf], 1u);
    if (ticket == gridDim.x - 1) {  // last block: close the iteration
      __threadfence();
      const int ch = atomicAdd(&st->changes[buf], 0);
      st->iterations = it + 1;
      if (ch <= 0) st->done = 1;
      // reset the other buffer for the next iteration
      for (int i = 0; i < kNFilters * 64; i++) st->mass[buf ^ 1][i] = 0;
      for (int i = 0; i < 64; i++) st->count[buf ^ 1][i] = 0;
      st->changes[buf ^ 1] = 0;
      st->blocks_done[buf ^ 1] = 0;
    }
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

size_t kmeans_mixed_work_bytes(int n) { (void)n; return 0; }

int launch_kmeans_mixed(hipStream_t s, int w, int h, const float* resp, int* labels, float* pdot, void* state,
                        int max_iter, void* work) {
  (void)s; (void)w; (void)h; (void)resp; (void)labels; (void)pdot; (void)state; (void)max_iter; (void)work;
  return fail(DMSC_ERR_ARG, "mixed-precision k-means is not available in this build");
}

int launch_kmeans(hipStream_t s, int w, int h, int k, int nf, const float* resp, int* labels,
                  float* pdot, void* state, int max_iter) {
  const int n = w * h;
  KmState* st = (KmState*)state;
  hipLaunchKernelGGL(k_km_init, dim3(div_up(n, 256)), dim3(256), 0, s, w, h, nf, resp, labels, pdot, st);
  const int blocks_acc = std::min(div_up(n, 256), 1024);
  const int blocks_asg = div_up(n, 256);
  for (int it = 0; it < max_iter; it++) {
    hipLaunchKernelGGL(k_km_accum<32>, dim3(blocks_acc), dim3(256), 0, s, n, nf, resp, labels, st, it);
    hipLaunchKernelGGL(k_km_assign<32>, dim3(blocks_asg), dim3(256), 0, s, n, nf, resp, pdot, labels, st, it);
  }
  (void)k;
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int kmeans_read_state(hipStream_t s, void* state, int* changes, int* iterations) {
  KmState h;
  DMSC_HIP(hipMemcpyAsync(&h, state, sizeof(KmState), hipMemcpyDeviceToHost, s));
  DMSC_HIP(hipStreamSynchronize(s));
  const int last = (int)h.iterations - 1;
  if (iterations) *iterations = (int)h.iterations;
  if (changes) *changes = last >= 0 ? h.changes[last & 1] : 0;
  return DMSC_OK;
}

}  // namespace dmsc
