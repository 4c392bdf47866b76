// stencil.hip -- 81-diagonal stencil matrix: row sums, normalised Laplacian
// transform and SpMV (reference: stencilMatrixMultiply/stencilMVM.cu).
//
// SpMV layout: A is DIA [81][pitch] (diagonal-major), so for a fixed
// diagonal consecutive rows are consecutive floats and a wave's matrix loads
// are contiguous.  Lanes own 1, 2 or 4 consecutive rows (k_spmv<RPL>) or one
// row with all 81 loads in flight (k_spmv_full); x is gathered through the
// caches.  An out-of-range column reads x[0] and contributes
// fma(A, 0, acc) == acc, which makes the reference's column bounds test
// branch-free.  Each row sums its 81 products sequentially in diagonal
// order (fused multiply-add), exactly as the CPU oracle.  Blocks are
// numbered XCD-aware (dev::xcd_block), so each XCD's L2 holds one
// contiguous slice of x.
#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {

namespace {

__global__ void k_rowsum(int n, StencilOffsets so, const float* __restrict__ A, size_t pitch,
                         float* __restrict__ dsum, float* __restrict__ rs) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  float acc = 0.0f;
  for (int d = 0; d < kStencilArea; d++) {
    const int col = row + so.off[d];
    if (col >= 0 && col < n) acc = acc + A[(size_t)d * pitch + row];
  }
  if (dsum) dsum[row] = acc;
  rs[row] = 1.0f / sqrtf(acc);
}

// stencilMVM.cu:98-125 unGeneralizeMatrix
__global__ void k_lsym(int n, StencilOffsets so, float* __restrict__ A, size_t pitch,
                       const float* __restrict__ rs) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  // dii = sum of the row, recomputed in the same order as k_rowsum
  float dii = 0.0f;
  for (int d = 0; d < kStencilArea; d++) {
    const int col = row + so.off[d];
    if (col >= 0 && col < n) dii = dii + A[(size_t)d * pitch + row];
  }
  const float diim5 = rs[row];
  for (int d = 0; d < kStencilArea; d++) {
    const int col = row + so.off[d];
    if (col < 0 || col >= n) continue;
    float* e = &A[(size_t)d * pitch + row];
    if (col == row) *e = (dii - *e) / dii;
    else *e = (-*e * diim5) * rs[col];
  }
}

}  // namespace

int launch_convert_matrix(hipStream_t s, int w, int h, float* A, int pitch, float* dsum, float* rs) {
  const int n = w * h;
  StencilOffsets so = make_stencil_offsets(w);
  hipLaunchKernelGGL(k_rowsum, dim3(div_up(n, 256)), dim3(256), 0, s, n, so, A, (size_t)pitch, dsum, rs);
  hipLaunchKernelGGL(k_lsym, dim3(div_up(n, 256)), dim3(256), 0, s, n, so, A, (size_t)pitch, rs);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

StencilOffsets make_stencil_offsets(int w) {
  StencilOffsets so;
  int delta[121], diag[121];
  const int nd = stencil_geometry(kStencilRadius, w, delta, diag);
  int acc = 0;
  for (int d = 0; d < nd; d++) {
    acc += delta[d];
    so.off[d] = acc;
  }
  // (dy, dx) of each diagonal, in the enumeration order of stencil_geometry
  for (int dy = -kStencilRadius; dy <= kStencilRadius; dy++)
    for (int dx = -kStencilRadius; dx <= kStencilRadius; dx++) {
      const int d = diag[(dy + kStencilRadius) * (2 * kStencilRadius + 1) + dx + kStencilRadius];
      if (d < 0) continue;
      so.dy[d] = (signed char)dy;
      so.dx[d] = (signed char)dx;
    }
  so.halo = kStencilRadius * w + kStencilRadius;
  return so;
}

namespace {

// y = A x (stencilMVM.cu:127-180): one wave per 64*RPL rows.
template <int RPL>
__global__ __launch_bounds__(64) void k_spmv(int n, StencilOffsets so, const float* __restrict__ A, size_t pitch,
                                             const float* __restrict__ x, float* __restrict__ y) {
  const int row = (dev::xcd_block(blockIdx.x, gridDim.x) * 64 + threadIdx.x) * RPL;
  if (row >= n) return;
  float acc[RPL];
  spmv_rows<RPL>(acc, A, pitch, row, n, so, [&](int c) { return x[c]; });
#pragma unroll
  for (int j = 0; j < RPL; j++)
    if (row + j < n) y[row + j] = acc[j];
}

// y = A x with one row per lane in 256-thread blocks: all 81 matrix loads and
// all 81 gathers of a lane are issued before the multiply-add chain, so each
// wave has one memory round trip instead of one per group of diagonals.
__global__ __launch_bounds__(256) void k_spmv_full(int n, StencilOffsets so, const float* __restrict__ A,
                                                   size_t pitch, const float* __restrict__ x,
                                                   float* __restrict__ y) {
  const int row = dev::xcd_block(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (row >= n) return;
  float a[kStencilArea], xv[kStencilArea];
#pragma unroll
  for (int d = 0; d < kStencilArea; d++) a[d] = A[(size_t)d * pitch + row];
#pragma unroll
  for (int d = 0; d < kStencilArea; d++) {
    const int col = row + so.off[d];
    const bool in = (unsigned)col < (unsigned)n;
    const float v = x[in ? col : 0];
    xv[d] = in ? v : 0.0f;  // fma(a, 0, acc) == acc: the reference's bounds test
  }
  float acc = 0.0f;
#pragma unroll
  for (int d = 0; d < kStencilArea; d++) acc = fmaf(a[d], xv[d], acc);
  y[row] = acc;
}

}  // namespace

int spmv_rows_per_lane = 0;  // 0: k_spmv_full; 1, 2, 4: k_spmv<RPL>

int launch_spmv(hipStream_t s, int w, int h, const float* A, int pitch, const float* x, float* y) {
  const int n = w * h;
  if (pitch % 4) return fail(DMSC_ERR_ARG, "spmv: matrix pitch must be a multiple of 4 floats");
  StencilOffsets so = make_stencil_offsets(w);
  switch (spmv_rows_per_lane) {
    case 1:
      hipLaunchKernelGGL(k_spmv<1>, dim3(div_up(n, 64)), dim3(64), 0, s, n, so, A, (size_t)pitch, x, y);
      break;
    case 2:
      hipLaunchKernelGGL(k_spmv<2>, dim3(div_up(n, 128)), dim3(64), 0, s, n, so, A, (size_t)pitch, x, y);
      break;
    case 4:
      hipLaunchKernelGGL(k_spmv<4>, dim3(div_up(n, 256)), dim3(64), 0, s, n, so, A, (size_t)pitch, x, y);
      break;
    default:
      hipLaunchKernelGGL(k_spmv_full, dim3(div_up(n, 256)), dim3(256), 0, s, n, so, A, (size_t)pitch, x, y);
      break;
  }
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

}  // namespace dmsc
