// combine.hip -- cue combination, oriented non-maximum suppression and the
// global Pb (reference: combine/combine.cu, nonmax/nonmax.cu,
// gPb/globalPb.cu).  One thread per pixel, all orientations in registers.
#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {
namespace {

struct CombineW {
  float coef[12];
  float wts[12];
};

// combine.cu:24-70: mPb and combinedG over the 12 cue x scale planes, in
// the reference order (Bg s0..s2, Cga, Cgb, Tg), fused multiply-add.
__global__ void k_combine(int n, int pitch, const float* __restrict__ bg, const float* __restrict__ cga,
                          const float* __restrict__ cgb, const float* __restrict__ tg, CombineW cw,
                          float* __restrict__ mpb, float* __restrict__ cg) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int o = blockIdx.y;
  if (p >= n) return;
  const float* cues[4] = {bg, cga, cgb, tg};
  float a1 = 0.f, a2 = 0.f;
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const float v = cues[k / 3][(size_t)((k % 3) * 8 + o) * pitch + p];
    a1 = fmaf(v, cw.coef[k], a1);
    a2 = fmaf(v, cw.wts[k], a2);
  }
  mpb[(size_t)o * pitch + p] = a1;
  cg[(size_t)o * pitch + p] = a2;
}

// nonmax.cu:23-94 devFindMax: max over orientations (strict '>', ties to
// the lowest orientation), negative maxima clamped to 0.
__global__ void k_findmax(int n, int pitch, const float* __restrict__ pb, float* __restrict__ mx,
                          int* __restrict__ ori) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float m = -200.0f;
  int mo = 0;
#pragma unroll
  for (int o = 0; o < 8; o++) {
    const float v = pb[(size_t)o * pitch + p];
    if (v > m) { m = v; mo = o; }
  }
  mx[p] = m > 0 ? m : 0.0f;
  ori[p] = mo;
}

// nonmax.cu:147-246 devOriented_2D (x = row, y = column as in the reference).
__global__ void k_oriented_nms(int w, int h, const float* __restrict__ mx, const int* __restrict__ ori,
                               float* __restrict__ nmax) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = w * h;
  if (p >= n) return;
  const int sx = h, sy = w;
  const int x = p / w, y = p - w * x;
  float res = 0.0f;
  const int o = ori[p], th = (o + 4) % 8;
  const float v = mx[p];
  int i0a = 0, i0b = 0, i1a = 0, i1b = 0;
  bool v0, v1;
  if (th == 0) {
    v0 = x > 0; v1 = x < sx - 1;
    i0a = p - sy; i0b = i0a; i1a = p + sy; i1b = i1a;
  } else if (th == 1) {
    v0 = (x > 0) && (y > 0); v1 = (x < sx - 1) && (y < sy - 1);
    i0a = p - sy; i0b = i0a - 1; i1a = p + sy; i1b = i1a + 1;
  } else if (th <= 3) {
    v0 = (x > 0) && (y > 0); v1 = (x < sx - 1) && (y < sy - 1);
    i0a = p - 1; i0b = i0a - sy; i1a = p + 1; i1b = i1a + sy;
  } else if (th <= 5) {
    v0 = (x < sx - 1) && (y > 0); v1 = (x > 0) && (y < sy - 1);
    i0a = p - 1; i0b = i0a + sy; i1a = p + 1; i1b = i1a - sy;
  } else {
    v0 = (x < sx - 1) && (y > 0); v1 = (x > 0) && (y < sy - 1);
    i0a = p + sy; i0b = i0a - 1; i1a = p - sy; i1b = i1a + 1;
  }
  if (v0 && v1) {
    const float tand[8] = {0.f, 0.414214f, 1.f, 0.414214f, 0.f, 0.414214f, 1.f, 0.414214f};
    const float cosori[8] = {1.f, 1.f, 0.923879f, 0.707107f, 0.382684f, 0.707107f, 0.923879f, 1.f};
    const float d = tand[th];
    const int d0a = abs(ori[i0a] - o), d0b = abs(ori[i0b] - o);
    const int d1a = abs(ori[i1a] - o), d1b = abs(ori[i1b] - o);
    const float va = ((1.0f - d) * mx[i0a]) * cosori[d0a] + (d * mx[i0b]) * cosori[d0b];
    const float vb = ((1.0f - d) * mx[i1a]) * cosori[d1a] + (d * mx[i1b]) * cosori[d1b];
    if ((v > va) && (v > vb)) {
      const float t = 1.2f * v;
      res = (t >= 1.0f) ? 1.0f : t;
    }
  }
  nmax[p] = res;
}

// globalPb.cu:5-49 CalcGPb.
__global__ void k_calc_gpb(int n, int pitch, int norient, const float* __restrict__ cg,
                           const float* __restrict__ spb, const float* __restrict__ mpb,
                           float* __restrict__ all, float* __restrict__ gpb) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float m = -200.0f;
  for (int o = 0; o < norient; o++) {
    const float v = fmaf(0.0139f, spb[(size_t)o * pitch + p], cg[(size_t)o * pitch + p]);
    all[(size_t)o * pitch + p] = v;
    if (v > m) m = v;
  }
  gpb[p] = (mpb[p] > 0.05f) ? m : 0.0f;
}

}  // namespace

// combine.cu:16-21: the 32-texton tables (lines 17, 21) and the 64-texton
// ones (lines 16, 20), which differ in the three Tg entries.
static CombineW combine_weights(int ntex) {
  if (ntex == 64)
    return CombineW{{0.0123f, 0.0110f, 0.0117f, 0.0169f, 0.0176f, 0.0198f, 0.0138f, 0.0138f, 0.0145f,
                     0.0104f, 0.0105f, 0.0115f},
                    {0.f, 0.f, 0.0028f, 0.0041f, 0.0042f, 0.0047f, 0.0033f, 0.0033f, 0.0035f, 0.0025f,
                     0.0025f, 0.0137f}};
  return CombineW{{0.0123f, 0.0110f, 0.0117f, 0.0169f, 0.0176f, 0.0198f, 0.0138f, 0.0138f, 0.0145f,
                   0.0144f, 0.0149f, 0.0158f},
                  {0.f, 0.f, 0.0028f, 0.0041f, 0.0042f, 0.0047f, 0.0033f, 0.0033f, 0.0035f, 0.0035f,
                   0.0035f, 0.0188f}};
}

void combine_tables(int ntex, float* coef, float* wts) {
  const CombineW cw = combine_weights(ntex);
  for (int k = 0; k < 12; k++) {
    coef[k] = cw.coef[k];
    wts[k] = cw.wts[k];
  }
}

int launch_combine(hipStream_t s, int n, int pitch, const float* bg, const float* cga,
                   const float* cgb, const float* tg, int ntex, float* mpb, float* cg) {
  if (ntex != 32 && ntex != 64) return fail(DMSC_ERR_ARG, "combine supports 32 or 64 textons");
  hipLaunchKernelGGL(k_combine, dim3(div_up(n, 256), 8), dim3(256), 0, s, n, pitch, bg, cga, cgb, tg,
                     combine_weights(ntex), mpb, cg);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int launch_nonmax(hipStream_t s, int w, int h, const float* pb, int pitch, float* mx, int* ori,
                  float* nmax) {
  const int n = w * h;
  hipLaunchKernelGGL(k_findmax, dim3(div_up(n, 256)), dim3(256), 0, s, n, pitch, pb, mx, ori);
  hipLaunchKernelGGL(k_oriented_nms, dim3(div_up(n, 256)), dim3(256), 0, s, w, h, mx, ori, nmax);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int launch_calc_gpb(hipStream_t s, int n, int pitch, int norient, const float* cg, const float* spb,
                    const float* mpb, float* all, float* gpb) {
  hipLaunchKernelGGL(k_calc_gpb, dim3(div_up(n, 256)), dim3(256), 0, s, n, pitch, norient, cg, spb, mpb,
                     all, gpb);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

}  // namespace dmsc
