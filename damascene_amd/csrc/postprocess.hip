// postprocess.hip -- thinned gPb and sigmoid normalisation (reference:
// postprocess/skeleton.cu).
//
// PostProcess: the thin mask starts as nonmax-mPb > 0.05 (normalizewrtmpb),
// is thinned to a fixed point by eight 3x3 look-up-table passes per
// iteration (applyLut: each pass reads the image the previous pass left and
// clears the pixels whose neighbourhood code the table marks), masks gPb
// (mask), and both gPb and the thin map go through the sigmoid
// normalisation (normalizeSigmoid).  NormalizeGpbAll applies the same
// sigmoid to the eight oriented gPb planes.
//
// The mask lives in u8 ping-pong buffers (each pass is a pure function of
// the previous image); a pass also counts removed pixels, and the host
// stops after the first iteration that removed none, as the reference does.
#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {
namespace {

// One look-up table as a 512-bit set: bit c of word c/32 set when
// neighbourhood code c clears the pixel.
struct LutBits {
  uint32_t w[16];
};

// skeleton.cu:104-129: the neighbourhood codes each of the eight tables
// marks.  The code of a pixel sums 2^(r + 3c) over its non-zero neighbours
// at (row - 1 + r, col - 1 + c), r, c in 0..2 (weights 1 8 64 / 2 16 128 /
// 4 32 256), neighbours outside the image contributing nothing.
constexpr int kLutCodes[8][8] = {
    {89, 91, 217, 219, -1, -1, -1, -1},          {152, 153, 216, 217, 408, 409, 472, 473},
    {23, 31, 55, 63, -1, -1, -1, -1},            {26, 27, 30, 31, 90, 91, 94, 95},
    {464, 472, 496, 504, -1, -1, -1, -1},        {50, 51, 54, 55, 306, 307, 310, 311},
    {308, 310, 436, 438, -1, -1, -1, -1},        {176, 180, 240, 244, 432, 436, 496, 500}};

__global__ void k_pp_seed(int n, const float* __restrict__ nmax, uint8_t* __restrict__ thin) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) thin[p] = nmax[p] > 0.05f ? 1 : 0;
}

__global__ void k_pp_zero(int* counter) { *counter = 0; }

__global__ __launch_bounds__(256) void k_pp_lut(int w, int h, LutBits lut, const uint8_t* __restrict__ in,
                                                uint8_t* __restrict__ out, int* __restrict__ removed) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
  int fired = 0;
  if (x < w && y < h) {
    int code = 0;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const int yy = y + r - 1, xx = x + c - 1;
        if (yy >= 0 && yy < h && xx >= 0 && xx < w && in[yy * w + xx]) code += 1 << (r + 3 * c);
      }
    const bool clear = (lut.w[code >> 5] >> (code & 31)) & 1u;
    const uint8_t v = in[y * w + x];
    out[y * w + x] = clear ? 0 : v;
    fired = clear && v;
  }
  const unsigned long long m = __ballot(fired);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(removed, (int)__popcll(m));
}

// skeleton.cu:84-99 normalizeSigmoid (FP32 arithmetic; the source's double
// literals are single precision on the reference's target).
__device__ __forceinline__ float pp_sigmoid(float val) {
  val = (val <= 0.0f) ? 0.0f : ((val * 1.2f > 1.0f) ? 1.0f : 1.2f * val);
  val = 1.0f / (1.0f + expf(2.6433f - 10.7998f * val));
  return (val - 0.0667f) / 0.9333f;
}

// skeleton.cu:69-82 mask, then normalizeSigmoid on gPb and on the thin map.
__global__ void k_pp_finish(int n, float* __restrict__ gpb, const uint8_t* __restrict__ thin,
                            float* __restrict__ gpb_thin) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float g = gpb[p];
  const float t = thin[p] ? g : 0.0f;
  gpb[p] = pp_sigmoid(g);
  gpb_thin[p] = pp_sigmoid(t);
}

// skeleton.cu:227-245 SigmoidGpbAll.
__global__ void k_pp_sigmoid_all(int npix, int norient, int pitch, float* __restrict__ all) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  for (int o = 0; o < norient; o++) {
    float v = all[(size_t)o * pitch + p];
    v = (v * 1.2f > 1.0f) ? 1.0f : v * 1.2f;
    if (v < 0.0f) v = 0.0f;
    v = 1.0f / (1.0f + expf(2.6433f - 10.7998f * v));
    all[(size_t)o * pitch + p] = (v - 0.0667f) / 0.9333f;
  }
}

}  // namespace

int run_postprocess(Ctx* ctx, int w, int h, float* gpb, const float* nmax, float* gpb_thin, int* iterations) {
  hipStream_t s = ctx->stream;
  const int n = w * h;
  void* base;
  const size_t nb = ((size_t)n + 255) & ~(size_t)255;
  DMSC_TRY(ctx->scratch(5, 2 * nb + 256, &base));
  uint8_t* a = (uint8_t*)base;
  uint8_t* b = a + nb;
  int* removed = (int*)(b + nb);
  LutBits luts[8];
  for (int l = 0; l < 8; l++) {
    for (int i = 0; i < 16; i++) luts[l].w[i] = 0;
    for (int k = 0; k < 8; k++)
      if (kLutCodes[l][k] >= 0) luts[l].w[kLutCodes[l][k] >> 5] |= 1u << (kLutCodes[l][k] & 31);
  }
  const dim3 g1(div_up(n, 256)), g2(div_up(w, 16), div_up(h, 16));
  hipLaunchKernelGGL(k_pp_seed, g1, dim3(256), 0, s, n, nmax, a);
  int it = 0, count = 0;
  // skeleton.cu:143-167: repeat the eight passes until one iteration removes
  // nothing; a pixel can only be removed once, so at most n + 1 iterations
  do {
    hipLaunchKernelGGL(k_pp_zero, dim3(1), dim3(1), 0, s, removed);
    for (int l = 0; l < 8; l++) {
      const uint8_t* in = (l & 1) ? b : a;
      uint8_t* out = (l & 1) ? a : b;
      hipLaunchKernelGGL(k_pp_lut, g2, dim3(256), 0, s, w, h, luts[l], in, out, removed);
    }
    DMSC_HIP(hipGetLastError());
    DMSC_HIP(hipMemcpyAsync(&count, removed, sizeof(int), hipMemcpyDeviceToHost, s));
    DMSC_HIP(hipStreamSynchronize(s));
    it++;
  } while (count > 0 && it <= n);
  hipLaunchKernelGGL(k_pp_finish, g1, dim3(256), 0, s, n, gpb, a, gpb_thin);
  DMSC_HIP(hipGetLastError());
  if (iterations) *iterations = it;
  return DMSC_OK;
}

int launch_normalize_gpb_all(hipStream_t s, int npix, int norient, int pitch, float* all) {
  hipLaunchKernelGGL(k_pp_sigmoid_all, dim3(div_up(npix, 256)), dim3(256), 0, s, npix, norient, pitch, all);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

}  // namespace dmsc
