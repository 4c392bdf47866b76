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
  uint32_t mass[2][kNFilters * 32];  // [buffer][filter][cluster], int32 wrap-around sums
  uint32_t count[2][32];
  int changes[2];
  unsigned int done;        // set when an iteration changed no label
  unsigned int iterations;  // iterations executed
  unsigned int rechecks;    // mixed path: pixels re-labelled by the exact FP32 distances
  float cent[kNFilters][32];  // centroids of the current iteration (FP32)
  float cdot[32];             // |c|^2 in the reference order
  alignas(16) uint16_t centb[32][32];  // mixed path: BF16 centroids, filters 0..31, [cluster][filter]
};

// kmeans.cu:14-35 assignInitialClusters (6x6 grid capped at 31) and
// kmeans.cu:296-307 makeSelfDots over the float responses.
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
    for (int i = threadIdx.x; i < kNFilters * 32; i += blockDim.x) { st->mass[0][i] = 0; st->mass[1][i] = 0; }
    if (threadIdx.x < 32) { st->count[0][threadIdx.x] = 0; st->count[1][threadIdx.x] = 0; }
  }
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

// One k-means iteration is three kernels: accumulate, centroids, assign.
// The stop test (kmeans.cu:712: no label changed) is taken at the start of
// the next iteration's accumulation, from the change count the previous
// assignment left; every block reads the same count, so all agree.
__device__ __forceinline__ bool km_stopped(KmState* st, int it) {
  if (st->done) return true;
  if (it > 0 && st->changes[(it - 1) & 1] <= 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 1;
    return true;
  }
  return false;
}

// kmeans.cu:57-103 findCentroids: int32 sums of (int)(1e5 * response)
// (INTCONFACTOR, kmeans.h:16), wrap-around arithmetic (order independent).
__global__ __launch_bounds__(256) void k_km_accum(int n, int nf, const float* __restrict__ resp,
                                                  const int* __restrict__ labels, KmState* st, int it) {
  if (km_stopped(st, it)) return;
  __shared__ uint32_t smass[kNFilters * 32];
  __shared__ uint32_t scount[32];
  for (int i = threadIdx.x; i < nf * 32; i += 256) smass[i] = 0;
  if (threadIdx.x < 32) scount[threadIdx.x] = 0;
  __syncthreads();
  const int buf = it & 1;
  for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
    const int c = labels[p];
    if (c < 0 || c >= 32) continue;  // unassigned pixel (all distances >= 1e7)
#pragma unroll
    for (int f = 0; f < kNFilters; f++)
      atomicAdd(&smass[f * 32 + c], (uint32_t)(int)(100000.0f * resp[(size_t)f * n + p]));
    atomicAdd(&scount[c], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nf * 32; i += 256)
    if (smass[i]) atomicAdd(&st->mass[buf][i], smass[i]);
  if (threadIdx.x < 32 && scount[threadIdx.x]) atomicAdd(&st->count[buf][threadIdx.x], scount[threadIdx.x]);
}

__device__ __forceinline__ uint16_t to_bf16(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// kmeans.cu:268-276 finishCentroids (an empty cluster gives NaN, which never
// wins a comparison) and makeSelfDots for the centroids; one block.  Also
// clears the buffers the next iteration accumulates into.
__global__ __launch_bounds__(256) void k_km_centroids(int nf, KmState* st, int it) {
  if (st->done) return;
  const int buf = it & 1;
  for (int i = threadIdx.x; i < kNFilters * 32; i += 256) {
    const int f = i / 32, c = i - f * 32;
    st->cent[f][c] = f < nf ? (float)(int)st->mass[buf][i] / (100000.0f * (float)st->count[buf][c]) : 0.0f;
  }
  __syncthreads();
  if (threadIdx.x < 32) {
    const int c = threadIdx.x;
    float d = 0.0f;
    for (int f = 0; f < nf; f++) d = fmaf(st->cent[f][c], st->cent[f][c], d);
    st->cdot[c] = d;
  }
  for (int i = threadIdx.x; i < 32 * 32; i += 256) {
    const int c = i / 32, f = i - c * 32;
    st->centb[c][f] = to_bf16(st->cent[f][c]);
  }
  for (int i = threadIdx.x; i < kNFilters * 32; i += 256) st->mass[buf ^ 1][i] = 0;
  if (threadIdx.x < 32) st->count[buf ^ 1][threadIdx.x] = 0;
  if (threadIdx.x == 0) st->changes[buf] = 0;
}

// Exact FP32 label of one pixel: kmeans.cu:296-338 makeDots + SGEMM(alpha
// -2, beta 1) + 447-472 findDiffLabels: dist_c = fma(-2, x.c, |x|^2 +
// |c|^2), x.c accumulated over the filters in order, strict '<' from 1e7.
__device__ __forceinline__ int km_exact_label(const float (&x)[kNFilters], float pd, const float (*cent)[32],
                                              const float* cdot) {
  float best = 10000000.0f;
  int bl = -1;
  for (int c = 0; c < 32; c++) {
    float acc = 0.0f;
#pragma unroll
    for (int f = 0; f < kNFilters; f++) acc = fmaf(x[f], cent[f][c], acc);
    const float d = fmaf(-2.0f, acc, pd + cdot[c]);
    if (d < best) {
      best = d;
      bl = c;
    }
  }
  return bl;
}

__global__ __launch_bounds__(256) void k_km_assign(int n, const float* __restrict__ resp,
                                                   const float* __restrict__ pdot, int* __restrict__ labels,
                                                   KmState* st, int it) {
  if (st->done) return;
  __shared__ float cent[kNFilters][32];
  __shared__ float cdot[32];
  __shared__ int schanges;
  for (int i = threadIdx.x; i < kNFilters * 32; i += 256) (&cent[0][0])[i] = (&st->cent[0][0])[i];
  if (threadIdx.x < 32) cdot[threadIdx.x] = st->cdot[threadIdx.x];
  if (threadIdx.x == 0) schanges = 0;
  __syncthreads();
  const int p = blockIdx.x * 256 + threadIdx.x;
  int changed = 0;
  if (p < n) {
    float x[kNFilters];
#pragma unroll
    for (int f = 0; f < kNFilters; f++) x[f] = resp[(size_t)f * n + p];
    const int bl = km_exact_label(x, pdot[p], cent, cdot);
    changed = bl != labels[p];
    labels[p] = bl;
  }
  const unsigned long long m = __ballot(changed);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&schanges, __popcll(m));
  __syncthreads();
  if (threadIdx.x == 0) {
    if (schanges) atomicAdd(&st->changes[it & 1], schanges);
    if (blockIdx.x == 0) st->iterations = it + 1;
  }
}

// ---------------------------------------------------------------- mixed path
// BF16 copy of filters 0..31 of every pixel, [pixel][filter] (the MFMA B
// operand); filter 32 stays FP32 and filter 33's response is zero.
__global__ void k_km_bf16(int n, const float* __restrict__ resp, uint16_t* __restrict__ xb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  uint32_t pk[16];
#pragma unroll
  for (int f = 0; f < 32; f += 2)
    pk[f / 2] = (uint32_t)to_bf16(resp[(size_t)f * n + p]) | ((uint32_t)to_bf16(resp[(size_t)(f + 1) * n + p]) << 16);
  uint4* dst = reinterpret_cast<uint4*>(xb + (size_t)p * 32);
#pragma unroll
  for (int q = 0; q < 4; q++) dst[q] = make_uint4(pk[4 * q], pk[4 * q + 1], pk[4 * q + 2], pk[4 * q + 3]);
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Distances from v_mfma_f32_16x16x32_bf16 products (clusters as rows,
// pixels as columns, filters 0..31 as K) plus the FP32 filter-32 term.  Per
// (pixel, cluster) the approximation error against the exact FP32 distance
// is below e = 2^-5 |x| |c| + 2^-20 (|x|^2 + |c|^2): BF16 rounding of both
// operands perturbs each product by at most (2^-7 + 2^-16)|x_f c_f|, summing
// to (2^-7 + 2^-16)|x||c| by Cauchy-Schwarz, times 2 for the -2 factor, and
// FP32 rounding of both computations stays inside the second term.  The
// label of the smallest approximate distance is the exact label when its
// upper bound lies below every other cluster's lower bound (no tie is
// possible then); any other pixel is re-labelled with the exact FP32
// distances in the reference order, so labels are bit-exact.
// Lane layout (checked on the device): A row l%16, B column l%16, K slice
// 8*(l/16)..+7; result rows 4*(l/16)+i of column l%16.
__global__ __launch_bounds__(256) void k_km_assign_mfma(int n, const float* __restrict__ resp,
                                                        const uint16_t* __restrict__ xb,
                                                        const float* __restrict__ pdot, int* __restrict__ labels,
                                                        KmState* st, int it) {
  if (st->done) return;
  __shared__ float cent[kNFilters][32];
  __shared__ float cdot[32];
  __shared__ int schanges, srecheck;
  for (int i = threadIdx.x; i < kNFilters * 32; i += 256) (&cent[0][0])[i] = (&st->cent[0][0])[i];
  if (threadIdx.x < 32) cdot[threadIdx.x] = st->cdot[threadIdx.x];
  if (threadIdx.x == 0) schanges = srecheck = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int col = lane & 15, grp = lane >> 4;
  // A operands: centroid rows (clusters) 16*t + col, filters 8*grp..+7
  bf16x8 a0, a1;
  {
    const uint4 v0 = *reinterpret_cast<const uint4*>(&st->centb[col][8 * grp]);
    const uint4 v1 = *reinterpret_cast<const uint4*>(&st->centb[16 + col][8 * grp]);
    __builtin_memcpy(&a0, &v0, 16);
    __builtin_memcpy(&a1, &v1, 16);
  }
  int changed = 0, rech = 0;
  // each wave labels 64 pixels, 16 per MFMA column tile
  for (int tile = 0; tile < 4; tile++) {
    const int p = (blockIdx.x * 4 + wave) * 64 + tile * 16 + col;
    const bool live = p < n;
    const int pp = live ? p : n - 1;
    bf16x8 b;
    {
      const uint4 v = *reinterpret_cast<const uint4*>(xb + (size_t)pp * 32 + 8 * grp);
      __builtin_memcpy(&b, &v, 16);
    }
    f32x4 s0 = {0, 0, 0, 0}, s1 = {0, 0, 0, 0};
    s0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b, s0, 0, 0, 0);  // clusters 0..15
    s1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b, s1, 0, 0, 0);  // clusters 16..31
    const float pd = pdot[pp], x32 = resp[(size_t)32 * n + pp];
    const float xn = sqrtf(pd);
    // this lane's 8 clusters, in increasing order: 4*grp+i, 16+4*grp+i
    float dt[8];
    int ci[8];
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int c = (i < 4 ? 0 : 16) + 4 * grp + (i & 3);
      const float sv = (i < 4 ? s0[i & 3] : s1[i & 3]) + x32 * cent[32][c];
      dt[i] = (pd + cdot[c]) - 2.0f * sv;
      ci[i] = c;
    }
    float best = INFINITY;
    int bc = 64;
#pragma unroll
    for (int i = 0; i < 8; i++)
      if (dt[i] < best) { best = dt[i]; bc = ci[i]; }
#pragma unroll
    for (int o = 16; o <= 32; o <<= 1) {  // the pixel's 4 lanes: smaller distance, then lower index
      const float ob = __shfl_xor(best, o, 64);
      const int oc = __shfl_xor(bc, o, 64);
      if (ob < best || (ob == best && oc < bc)) { best = ob; bc = oc; }
    }
    float lo_other = INFINITY;
#pragma unroll
    for (int i = 0; i < 8; i++)
      if (ci[i] != bc) {
        const float e = 0.03125f * xn * sqrtf(cdot[ci[i]]) + 9.5367431640625e-7f * (pd + cdot[ci[i]]);
        lo_other = fminf(lo_other, dt[i] - e);
      }
#pragma unroll
    for (int o = 16; o <= 32; o <<= 1) lo_other = fminf(lo_other, __shfl_xor(lo_other, o, 64));
    bool certain = false;
    if (bc < 32) {
      const float e = 0.03125f * xn * sqrtf(cdot[bc]) + 9.5367431640625e-7f * (pd + cdot[bc]);
      certain = best + e < lo_other && best + e < 10000000.0f;
    }
    if (grp == 0 && live) {
      int bl = bc;
      if (!certain) {  // near tie (or no finite distance): exact FP32 distances
        float x[kNFilters];
#pragma unroll
        for (int f = 0; f < kNFilters; f++) x[f] = resp[(size_t)f * n + p];
        bl = km_exact_label(x, pd, cent, cdot);
        rech++;
      }
      changed += bl != labels[p];
      labels[p] = bl;
    }
  }
  if (changed) atomicAdd(&schanges, changed);
  if (rech) atomicAdd(&srecheck, rech);
  __syncthreads();
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

// BF16 responses for the mixed path: 32 filters per pixel.
size_t kmeans_mixed_work_bytes(int n) { return (size_t)n * 32 * sizeof(uint16_t) + 256; }

namespace {
int km_iterations(hipStream_t s, int n, int nf, const float* resp, int* labels, const float* pdot, KmState* st,
                  int max_iter, const uint16_t* xb) {
  const int blocks_acc = std::min(div_up(n, 256), 256);
  for (int it = 0; it < max_iter; it++) {
    hipLaunchKernelGGL(k_km_accum, dim3(blocks_acc), dim3(256), 0, s, n, nf, resp, labels, st, it);
    hipLaunchKernelGGL(k_km_centroids, dim3(1), dim3(256), 0, s, nf, st, it);
    if (xb)
      hipLaunchKernelGGL(k_km_assign_mfma, dim3(div_up(n, 256)), dim3(256), 0, s, n, resp, xb, pdot, labels, st,
                         it);
    else
      hipLaunchKernelGGL(k_km_assign, dim3(div_up(n, 256)), dim3(256), 0, s, n, resp, pdot, labels, st, it);
  }
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}
}  // namespace

int launch_kmeans_mixed(hipStream_t s, int w, int h, const float* resp, int* labels, float* pdot, void* state,
                        int max_iter, void* work) {
  const int n = w * h;
  KmState* st = (KmState*)state;
  uint16_t* xb = (uint16_t*)work;
  if ((uintptr_t)xb % 16 || (uintptr_t)state % 16) return fail(DMSC_ERR_INTERNAL, "k-means buffers misaligned");
  hipLaunchKernelGGL(k_km_init, dim3(div_up(n, 256)), dim3(256), 0, s, w, h, kNFilters, resp, labels, pdot, st);
  hipLaunchKernelGGL(k_km_bf16, dim3(div_up(n, 256)), dim3(256), 0, s, n, resp, xb);
  return km_iterations(s, n, kNFilters, resp, labels, pdot, st, max_iter, xb);
}

int launch_kmeans(hipStream_t s, int w, int h, int k, int nf, const float* resp, int* labels,
                  float* pdot, void* state, int max_iter) {
  (void)k;
  const int n = w * h;
  KmState* st = (KmState*)state;
  hipLaunchKernelGGL(k_km_init, dim3(div_up(n, 256)), dim3(256), 0, s, w, h, nf, resp, labels, pdot, st);
  return km_iterations(s, n, nf, resp, labels, pdot, st, max_iter, nullptr);
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
