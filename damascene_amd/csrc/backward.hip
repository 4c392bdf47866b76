// backward.hip -- differentiable spectral Pb + combine.
//
// The reference has no backward pass; its forward is combine.cu:24-70
// (combinedG_o = sum_k w_k C_{k,o} over the 12 cue x scale planes),
// spectralPb.cu:651-659 (sPb_o = sum_i |K_o * mirror3(v_i)| / sqrt(lambda_i))
// and globalPb.cu:5-49 (gPbAll_o = combinedG_o + 0.0139 sPb_o, gPb = max_o).
// Here the weights are free parameters,
//   params = [w_0 .. w_11, gamma, eta_0 .. eta_{nvec-1}]   (eta_0 unused),
//   gPbAll_o = sum_k w_k C_{k,o} + gamma * sum_{i>=1} eta_i |R_{i,o}|,
//   R_{i,o}  = K_o correlated with v_i mirrored by 3 pixels,
//   gPbPre   = max_o gPbAll_o (strict '>', the lowest orientation wins ties),
// and the backward pass gives, from g = dL/dgPbPre and o* = argmax:
//   dL/dw_k    = sum_p g C_{k,o*}
//   dL/dgamma  = sum_p g sPb_{o*}
//   dL/deta_i  = gamma sum_p g |R_{i,o*}|
//   dL/dv_i    = mirror^T K^T (g gamma eta_i sign(R_{i,o*}) at orientation o*)
// The forward forms every value in the pipeline's operation order, so with
// params = (combine weights, 0.0139, 1/sqrt(lambda_i)) its gPbAll is
// bit-identical to the pipeline's.  Parameter gradients are block partial
// sums in FP64 reduced in a fixed order (deterministic).
#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {
namespace {

constexpr int BTX = 32, BTY = 8, BNT = BTX * BTY;  // pixel tile of the forward and first backward kernel
constexpr int kTaps = 49;

struct SpbK {
  float K[8 * kTaps];  // slot o: filter index (3+u) + (3+v)*7 (spectralPb.cu:100)
};

struct Cues {
  const float* c[4];  // Bg, Cga, Cgb, Tg: [24][cue_pitch], scale-major then orientation
};

// The v_i tile of a 32 x 8 pixel block with its mirrored 3-pixel frame.
__device__ __forceinline__ void stage_tile(float (*tile)[BTX + 6], const float* v, int w, int h) {
  for (int k = threadIdx.x; k < (BTY + 6) * (BTX + 6); k += BNT) {
    const int ly = k / (BTX + 6), lx = k - ly * (BTX + 6);
    const int gx = dev::mirror(dev::clampi(blockIdx.x * BTX + lx, 0, w + 5), 3, w);
    const int gy = dev::mirror(dev::clampi(blockIdx.y * BTY + ly, 0, h + 5), 3, h);
    tile[ly][lx] = v[(size_t)gy * w + gx];
  }
}

__global__ __launch_bounds__(BNT) void k_dspb_forward(int w, int h, int nvec, const float* __restrict__ V, SpbK F,
                                                      const float* __restrict__ prm, Cues C, size_t cue_pitch,
                                                      float* __restrict__ all, size_t pitch,
                                                      float* __restrict__ gpb_pre, int* __restrict__ amax) {
  __shared__ float tile[BTY + 6][BTX + 6];
  const int tx = threadIdx.x % BTX, ty = threadIdx.x / BTX;
  const int x = blockIdx.x * BTX + tx, y = blockIdx.y * BTY + ty;
  const size_t n = (size_t)w * h;
  float res[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 1; i < nvec; i++) {
    __syncthreads();
    stage_tile(tile, V + (size_t)i * n, w, h);
    __syncthreads();
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int u = -3; u <= 3; u++)
#pragma unroll
      for (int vv = -3; vv <= 3; vv++) {
        const float val = tile[ty + 3 + vv][tx + 3 + u];
#pragma unroll
        for (int o = 0; o < 8; o++) s[o] = fmaf(val, F.K[o * kTaps + (3 + u) + (3 + vv) * 7], s[o]);
      }
    const float eta = prm[13 + i];
#pragma unroll
    for (int o = 0; o < 8; o++) res[o] = fmaf(fabsf(s[o]), eta, res[o]);
  }
  if (x >= w || y >= h) return;
  const size_t p = (size_t)y * w + x;
  const float gamma = prm[12];
  float m = -200.0f;
  int mo = 0;
#pragma unroll
  for (int o = 0; o < 8; o++) {
    float a2 = 0.f;
#pragma unroll
    for (int k = 0; k < 12; k++) a2 = fmaf(C.c[k / 3][(size_t)((k % 3) * 8 + o) * cue_pitch + p], prm[k], a2);
    const float v = fmaf(gamma, res[o], a2);
    if (all) all[(size_t)o * pitch + p] = v;
    if (v > m) {
      m = v;
      mo = o;
    }
  }
  gpb_pre[p] = m;
  amax[p] = mo;
}

template <int NT>
__device__ __forceinline__ double bsum(double v, double* red) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  double t = red[0];
#pragma unroll
  for (int k = 1; k < NT / 64; k++) t += red[k];
  __syncthreads();
  return t;
}

// Per pixel: R_{i,o*} for every vector (same order as the forward, so the
// signs are the forward's), the scattered values s_i = g gamma eta_i
// sign(R_{i,o*}) for the second kernel, and block partial sums of the
// parameter gradients (np = 13 + nvec values per block).
__global__ __launch_bounds__(BNT) void k_dspb_backward1(int w, int h, int nvec, const float* __restrict__ V, SpbK F,
                                                        const float* __restrict__ prm, Cues C, size_t cue_pitch,
                                                        const float* __restrict__ grad, const int* __restrict__ amax,
                                                        float* __restrict__ S, double* __restrict__ part) {
  __shared__ float tile[BTY + 6][BTX + 6];
  __shared__ float sk[8 * kTaps];
  __shared__ double red[BNT / 64];
  for (int k = threadIdx.x; k < 8 * kTaps; k += BNT) sk[k] = F.K[k];
  const int tx = threadIdx.x % BTX, ty = threadIdx.x / BTX;
  const int x = blockIdx.x * BTX + tx, y = blockIdx.y * BTY + ty;
  const bool in = x < w && y < h;
  const size_t n = (size_t)w * h, p = in ? (size_t)y * w + x : 0;
  const float g = in ? grad[p] : 0.f;
  const int o = in ? amax[p] : 0;
  const float gamma = prm[12];
  const int np = 13 + nvec;
  const int blk = blockIdx.y * gridDim.x + blockIdx.x;
  double* out = part + (size_t)blk * np;
  float spb = 0.f;
  for (int i = 1; i < nvec; i++) {
    __syncthreads();
    stage_tile(tile, V + (size_t)i * n, w, h);
    __syncthreads();
    float r = 0.f;
#pragma unroll
    for (int u = -3; u <= 3; u++)
#pragma unroll
      for (int vv = -3; vv <= 3; vv++) r = fmaf(tile[ty + 3 + vv][tx + 3 + u], sk[o * kTaps + (3 + u) + (3 + vv) * 7], r);
    const float eta = prm[13 + i];
    spb = fmaf(fabsf(r), eta, spb);
    const float sg = r > 0.f ? 1.f : (r < 0.f ? -1.f : 0.f);
    if (in) S[(size_t)i * n + p] = g * gamma * eta * sg;
    const double de = bsum<BNT>((double)g * (double)gamma * (double)fabsf(r), red);
    if (threadIdx.x == 0) out[13 + i] = de;
  }
#pragma unroll
  for (int k = 0; k < 12; k++) {
    const float c = in ? C.c[k / 3][(size_t)((k % 3) * 8 + o) * cue_pitch + p] : 0.f;
    const double dw = bsum<BNT>((double)g * (double)c, red);
    if (threadIdx.x == 0) out[k] = dw;
  }
  const double dg = bsum<BNT>((double)g * (double)spb, red);
  if (threadIdx.x == 0) {
    out[12] = dg;
    out[13] = 0.0;  // eta_0 (vector 0 is not used)
  }
}

// dL/dv_i at pixel (a, b): every mirrored position (X, Y) that reflects to
// (a, b) -- X = a + 3, and 2 - a near the left edge or 2w + 2 - a near the
// right (dev::mirror with border 3) -- collects K_{o*}(u, v) s_i from the
// pixels (X - 3 - u, Y - 3 - v) whose 7 x 7 window read it.
template <int MAXV>
__global__ __launch_bounds__(256) void k_dspb_backward2(int w, int h, int nvec, SpbK F, const int* __restrict__ amax,
                                                        const float* __restrict__ S, float* __restrict__ dV) {
  __shared__ float sk[8 * kTaps];
  for (int k = threadIdx.x; k < 8 * kTaps; k += 256) sk[k] = F.K[k];
  __syncthreads();
  const size_t n = (size_t)w * h;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= (int)n) return;
  const int a = p % w, b = p / w;
  int px[2], py[2], nx = 0, ny = 0;
  px[nx++] = a + 3;
  if (a <= 2) px[nx++] = 2 - a;
  else if (a >= w - 3) px[nx++] = 2 * w + 2 - a;
  py[ny++] = b + 3;
  if (b <= 2) py[ny++] = 2 - b;
  else if (b >= h - 3) py[ny++] = 2 * h + 2 - b;
  float acc[MAXV];
#pragma unroll
  for (int i = 0; i < MAXV; i++) acc[i] = 0.f;
  for (int jy = 0; jy < ny; jy++)
    for (int jx = 0; jx < nx; jx++)
      for (int u = -3; u <= 3; u++) {
        const int xs = px[jx] - 3 - u;
        if (xs < 0 || xs >= w) continue;
        for (int vv = -3; vv <= 3; vv++) {
          const int ys = py[jy] - 3 - vv;
          if (ys < 0 || ys >= h) continue;
          const size_t q = (size_t)ys * w + xs;
          const float kv = sk[amax[q] * kTaps + (3 + u) + (3 + vv) * 7];
#pragma unroll
          for (int i = 1; i < MAXV; i++)
            if (i < nvec) acc[i] = fmaf(kv, S[(size_t)i * n + q], acc[i]);
        }
      }
  dV[p] = 0.f;  // vector 0 does not enter the forward
#pragma unroll
  for (int i = 1; i < MAXV; i++)
    if (i < nvec) dV[(size_t)i * n + p] = acc[i];
}

// Parameter gradients: the blocks' partials summed in block order.
__global__ __launch_bounds__(256) void k_dspb_reduce(int nblk, int np, const double* __restrict__ part,
                                                     float* __restrict__ gp) {
  __shared__ double red[4];
  for (int j = 0; j < np; j++) {
    double v = 0;
    for (int b = threadIdx.x; b < nblk; b += 256) v += part[(size_t)b * np + j];
    v = bsum<256>(v, red);
    if (threadIdx.x == 0) gp[j] = (float)v;
  }
}

}  // namespace

int spb_combine_forward(Ctx* ctx, int w, int h, int cue_pitch, const float* bg, const float* cga, const float* cgb,
                        const float* tg, int nvec, const float* V, const float* prm, float* all, int pitch,
                        float* gpb_pre, int* amax) {
  if (nvec < 2 || nvec > kMaxDiffVec) return fail(DMSC_ERR_ARG, "nvec must be in [2, %d]", kMaxDiffVec);
  SpbK F;
  spb_filters(F.K);
  hipLaunchKernelGGL(k_dspb_forward, dim3(div_up(w, BTX), div_up(h, BTY)), dim3(BNT), 0, ctx->stream, w, h, nvec, V,
                     F, prm, Cues{{bg, cga, cgb, tg}}, (size_t)cue_pitch, all, (size_t)pitch, gpb_pre, amax);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int spb_combine_backward(Ctx* ctx, int w, int h, int cue_pitch, const float* bg, const float* cga, const float* cgb,
                         const float* tg, int nvec, const float* V, const float* prm, const int* amax,
                         const float* grad, float* grad_prm, float* dV) {
  if (nvec < 2 || nvec > kMaxDiffVec) return fail(DMSC_ERR_ARG, "nvec must be in [2, %d]", kMaxDiffVec);
  const size_t n = (size_t)w * h;
  const dim3 g1(div_up(w, BTX), div_up(h, BTY));
  const int nblk = (int)(g1.x * g1.y), np = 13 + nvec;
  const size_t s_bytes = (sizeof(float) * nvec * n + 255) & ~(size_t)255;
  void* scr;
  DMSC_TRY(ctx->scratch(4, s_bytes + sizeof(double) * (size_t)nblk * np, &scr));
  float* S = (float*)scr;
  double* part = (double*)((char*)scr + s_bytes);
  SpbK F;
  spb_filters(F.K);
  hipLaunchKernelGGL(k_dspb_backward1, g1, dim3(BNT), 0, ctx->stream, w, h, nvec, V, F, prm,
                     Cues{{bg, cga, cgb, tg}}, (size_t)cue_pitch, grad, amax, S, part);
  hipLaunchKernelGGL(k_dspb_backward2<kMaxDiffVec>, dim3(div_up((int)n, 256)), dim3(256), 0, ctx->stream, w, h, nvec,
                     F, amax, S, dV);
  hipLaunchKernelGGL(k_dspb_reduce, dim3(1), dim3(256), 0, ctx->stream, nblk, np, part, grad_prm);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

}  // namespace dmsc
