// convert.hip -- colour conversion, mirroring and quantisation kernels
// (reference: convert/convert.cu).
#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {

struct GammaLut {
  float v[256];
};

namespace {

__device__ inline void unpack(uint32_t c, float& r, float& g, float& b) {
  r = (float)(c & 0xff) / 255.0f;
  g = (float)((c & 0xff00) >> 8) / 255.0f;
  b = (float)((c & 0xff0000) >> 16) / 255.0f;
}

// convert.cu:11-25
__global__ void k_rgb_to_grey(int n, const uint32_t* __restrict__ rgb, float* __restrict__ grey) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float r, g, b;
  unpack(rgb[i], r, g, b);
  grey[i] = (0.29894f * r + 0.58704f * g) + 0.11402f * b;
}

__device__ inline void lab_of(uint32_t c, const GammaLut& lut, float& L, float& A, float& B) {
  const float r = lut.v[c & 0xff], g = lut.v[(c >> 8) & 0xff], b = lut.v[(c >> 16) & 0xff];
  float x = (0.412453f * r + 0.357580f * g) + 0.180423f * b;
  float y = (0.212671f * r + 0.715160f * g) + 0.072169f * b;
  float z = (0.019334f * r + 0.119193f * g) + 0.950227f * b;
  const float thr = 0.008856f;
  x = x / 0.950456f;
  y = y / 1.000000f;
  z = z / 1.088754f;
  const float cy = dev::cbrt_exact(y);
  const float fx = (x > thr) ? dev::cbrt_exact(x) : (7.787f * x + (16.0f / 116.0f));
  const float fy = (y > thr) ? cy : (7.787f * y + (16.0f / 116.0f));
  const float fz = (z > thr) ? dev::cbrt_exact(z) : (7.787f * z + (16.0f / 116.0f));
  L = (y > thr) ? (116.0f * cy - 16.0f) : (903.3f * y);
  A = 500.0f * (fx - fy);
  B = 200.0f * (fy - fz);
}

__device__ inline void normalise(float& L, float& A, float& B) {
  const float ab_min = -73.0f, ab_range = 95.0f - (-73.0f);
  float l = L / 100.0f, a = (A - ab_min) / ab_range, b = (B - ab_min) / ab_range;
  L = l < 0 ? 0.f : (l > 1 ? 1.f : l);
  A = a < 0 ? 0.f : (a > 1 ? 1.f : a);
  B = b < 0 ? 0.f : (b > 1 ? 1.f : b);
}

// convert.cu:27-64 (gamma by table, cube root by cbrt_exact)
__global__ void k_rgb_to_lab(int n, GammaLut lut, const uint32_t* __restrict__ rgb, float* L,
                             float* A, float* B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float l, a, b;
  lab_of(rgb[i], lut, l, a, b);
  L[i] = l; A[i] = a; B[i] = b;
}

// convert.cu:66-85
__global__ void k_normalize_lab(int n, float* L, float* A, float* B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float l = L[i], a = A[i], b = B[i];
  normalise(l, a, b);
  L[i] = l; A[i] = a; B[i] = b;
}

// Pipeline form: grey and normalised Lab planes in one pass.
__global__ void k_colour_fused(int n, GammaLut lut, const uint32_t* __restrict__ rgb,
                               float* __restrict__ grey, float* __restrict__ L,
                               float* __restrict__ A, float* __restrict__ B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t c = rgb[i];
  float r, g, b;
  unpack(c, r, g, b);
  grey[i] = (0.29894f * r + 0.58704f * g) + 0.11402f * b;
  float l, a, bb;
  lab_of(c, lut, l, a, bb);
  normalise(l, a, bb);
  L[i] = l; A[i] = a; B[i] = bb;
}

template <typename T>
__global__ void k_mirror(int w, int h, int border, const T* __restrict__ in, T* __restrict__ out) {
  const int bw = w + 2 * border, bh = h + 2 * border;
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= bw || y >= bh) return;
  out[(size_t)y * bw + x] = in[(size_t)dev::mirror(y, border, h) * w + dev::mirror(x, border, w)];
}

// convert.cu:205-213
__global__ void k_unmirror(int bw, int border, int ow, int oh, const float* __restrict__ in,
                           float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
  if (x >= ow || y >= oh) return;
  out[(size_t)y * ow + x] = in[(size_t)(y + border) * bw + x + border];
}

// convert.cu:234-246
__global__ void k_quantize(int n, int nbins, const float* __restrict__ in, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int q = (int)floorf(in[i] * (float)nbins);
  out[i] = (q == nbins) ? nbins - 1 : q;
}

GammaLut make_lut(float gamma) {
  GammaLut l;
  gamma_lut(gamma, l.v);
  return l;
}

}  // namespace

int launch_rgb_to_grey(hipStream_t s, int n, const uint32_t* rgb, float* grey) {
  hipLaunchKernelGGL(k_rgb_to_grey, dim3(div_up(n, 256)), dim3(256), 0, s, n, rgb, grey);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int launch_rgb_to_lab(hipStream_t s, int n, float gamma, const uint32_t* rgb, float* L, float* A,
                      float* B) {
  hipLaunchKernelGGL(k_rgb_to_lab, dim3(div_up(n, 256)), dim3(256), 0, s, n, make_lut(gamma), rgb, L, A, B);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int launch_normalize_lab(hipStream_t s, int n, float* L, float* A, float* B) {
  hipLaunchKernelGGL(k_normalize_lab, dim3(div_up(n, 256)), dim3(256), 0, s, n, L, A, B);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int launch_colour_fused(hipStream_t s, int n, float gamma, const uint32_t* rgb, float* grey,
                        float* L, float* A, float* B) {
  hipLaunchKernelGGL(k_colour_fused, dim3(div_up(n, 256)), dim3(256), 0, s, n, make_lut(gamma), rgb, grey, L, A, B);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int launch_mirror_f(hipStream_t s, int w, int h, int border, const float* in, float* out) {
  const int bw = w + 2 * border, bh = h + 2 * border;
  hipLaunchKernelGGL(k_mirror<float>, dim3(div_up(bw, 256), bh), dim3(256), 0, s, w, h, border, in, out);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int launch_mirror_i(hipStream_t s, int w, int h, int border, const int* in, int* out) {
  const int bw = w + 2 * border, bh = h + 2 * border;
  hipLaunchKernelGGL(k_mirror<int>, dim3(div_up(bw, 256), bh), dim3(256), 0, s, w, h, border, in, out);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int launch_unmirror_f(hipStream_t s, int bw, int bh, int border, const float* in, float* out) {
  const int ow = bw - 2 * border, oh = bh - 2 * border;
  hipLaunchKernelGGL(k_unmirror, dim3(div_up(ow, 256), oh), dim3(256), 0, s, bw, border, ow, oh, in, out);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int launch_quantize(hipStream_t s, int n, int nbins, const float* in, int* out) {
  hipLaunchKernelGGL(k_quantize, dim3(div_up(n, 256)), dim3(256), 0, s, n, nbins, in, out);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

}  // namespace dmsc
