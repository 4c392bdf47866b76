// spectralpb.hip -- oriented spectral Pb (reference: sPb/spectralPb.cu).
//
// For eigenvectors 1..nvec-1: mirror with border 3, correlate with the 8
// first-derivative-of-Gaussian 7x7 filters (u outer, v inner, fused
// multiply-add) and accumulate |response| / sqrt(lambda_i).  One thread per
// pixel keeps the 8 oriented sums for all eigenvectors in registers; the
// block stages each eigenvector tile (mirror-indexed halo) in LDS.
#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {
namespace {

struct SpbParams {
  float K[8 * 49];  // slot o: filter index (3+u) + (3+v)*7 (spectralPb.cu:100)
  float scale[32];  // 1/sqrt(lambda_i)
};

constexpr int STX = 32, STY = 8;

__global__ __launch_bounds__(256) void k_spectral_pb(int w, int h, int nvec, const float* __restrict__ V,
                                                     SpbParams P, float* __restrict__ spb, size_t pitch) {
  __shared__ float tile[STY + 6][STX + 6];
  const int tx = threadIdx.x % STX, ty = threadIdx.x / STX;
  const int x = blockIdx.x * STX + tx, y = blockIdx.y * STY + ty;
  const size_t n = (size_t)w * h;
  float res[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 1; i < nvec; i++) {
    const float* v = V + (size_t)i * n;
    __syncthreads();
    for (int k = threadIdx.x; k < (STY + 6) * (STX + 6); k += 256) {
      const int ly = k / (STX + 6), lx = k - ly * (STX + 6);
      // mirrored coordinates (border 3), clamped to [0, w + 6) before reflecting
      const int gx = dev::mirror(dev::clampi(blockIdx.x * STX + lx, 0, w + 5), 3, w);
      const int gy = dev::mirror(dev::clampi(blockIdx.y * STY + ly, 0, h + 5), 3, h);
      tile[ly][lx] = v[(size_t)gy * w + gx];
    }
    __syncthreads();
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int u = -3; u <= 3; u++)
#pragma unroll
      for (int vv = -3; vv <= 3; vv++) {
        const float val = tile[ty + 3 + vv][tx + 3 + u];
#pragma unroll
        for (int o = 0; o < 8; o++) s[o] = fmaf(val, P.K[o * 49 + (3 + u) + (3 + vv) * 7], s[o]);
      }
#pragma unroll
    for (int o = 0; o < 8; o++) res[o] = fmaf(fabsf(s[o]), P.scale[i], res[o]);
  }
  if (x < w && y < h)
#pragma unroll
    for (int o = 0; o < 8; o++) spb[(size_t)o * pitch + (size_t)y * w + x] = res[o];
}

}  // namespace

// spectralPb.cu:673-795
int launch_spectral_pb(hipStream_t s, const float* h_eigvals, const float* d_eigvecs, int w, int h, int nvec,
                       float* spb, int pitch) {
  if (nvec > 32) return fail(DMSC_ERR_ARG, "spectral_pb: at most 32 vectors");
  SpbParams P;
  spb_filters(P.K);
  for (int i = 0; i < 32; i++) P.scale[i] = 0.f;
  for (int i = 1; i < nvec; i++) P.scale[i] = 1.0f / sqrtf(h_eigvals[i]);
  hipLaunchKernelGGL(k_spectral_pb, dim3(div_up(w, STX), div_up(h, STY)), dim3(256), 0, s, w, h, nvec, d_eigvecs,
                     P, spb, (size_t)pitch);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

}  // namespace dmsc
