// intervening.hip -- intervening-contour affinities on the radius-5 disc
// (reference: intervening/intervening.cu).
//
// One thread per pixel walks Bresenham lines from the pixel to every point
// of the boundary of its (clipped) 11x11 window, max-accumulating the
// thinned mPb across crossed pixel edges; points for which the walk is the
// best integer approximant of their direction receive the affinity
// exp(-(1 - C(ic)) / sigma) (later walks overwrite earlier ones, as in the
// reference).  The walks only read mPb, so each block stages its tile plus a
// 5-pixel clamped halo in LDS.  A second kernel symmetrises the matrix over
// forward offsets (each unordered pair is owned by one thread).
#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {
namespace {

struct DiagMap {
  int d[121];  // (dy+5)*11 + dx+5 -> diagonal index (Stencil.cpp:19-20)
};

constexpr int ITX = 32, ITY = 8, HALO = 6;

struct IcCtx {
  const float* tile;  // LDS tile, stride LW, origin at image (x0 - HALO, y0 - HALO)
  int LW, ox, oy;
  int w, h;
  __device__ float pb(int x, int y) const {
    x = dev::clampi(x, 0, w - 1);
    y = dev::clampi(y, 0, h - 1);
    return tile[(y - oy) * LW + (x - ox)];
  }
};

__device__ inline float c_ic_ss(float ic) {  // intervening.cu:32-42 (project exp, as above)
  float val = 1.8682f;
  val += (ic / 0.3130f) * -1.3113f;
  const float post = 1.0f / (1.0f + dev::exp_exact(-val));
  if (!isfinite(post)) return 0;
  if (post < 0) return 0;
  if (post > 1) return 1;
  return post;
}

// intervening.cu:28-30, with the project exp in place of the device expf
__device__ inline float affinity(float c, float rs) { return dev::exp_exact(-(1 - c) * rs); }

// intervening.cu:44-283 ic_walk
__device__ void ic_walk(const IcCtx& C, const DiagMap& dm, float rs, float* M, size_t pitch, int x0, int y0,
                        int x1, int y1, int x2, int y2, int x3, int y3) {
  const int dx1 = x1 - x0, dy1 = y1 - y0, dx2 = x2 - x0, dy2 = y2 - y0, dx3 = x3 - x0, dy3 = y3 - y0;
  const int dot11 = dx1 * dx1 + dy1 * dy1, dot22 = dx2 * dx2 + dy2 * dy2, dot33 = dx3 * dx3 + dy3 * dy3;
  const int adx = abs(dx2), ady = abs(dy2);
  int oct = -1;
  if (dx2 > 0 && dy2 >= 0) oct = adx > ady ? 0 : 1;
  else if (dx2 <= 0 && dy2 > 0) oct = adx < ady ? 2 : 3;
  else if (dy2 <= 0 && dx2 < 0) oct = adx > ady ? 4 : 5;
  else if (dx2 >= 0 && dy2 < 0) oct = adx < ady ? 6 : 7;
  const bool xmaj = (oct == 0 || oct == 3 || oct == 4 || oct == 7);
  const int sx = (oct == 0 || oct == 1 || oct == 6 || oct == 7) ? 1 : -1;
  const int sy = (oct >= 0 && oct <= 3) ? 1 : -1;
  int t = xmaj ? -adx : -ady;
  float maxpb = 0.0f;
  int xi = x0, yi = y0;
  while (xi != x2 || yi != y2) {
    const int ox = xi, oy = yi;
    if (xmaj) {
      xi += sx; t += ady << 1;
      if (t > 0) { yi += sy; t -= adx << 1; }
    } else {
      yi += sy; t += adx << 1;
      if (t > 0) { xi += sx; t -= ady << 1; }
    }
    const int dxi = xi - x0, dyi = yi - y0;
    const int dotii = dxi * dxi + dyi * dyi, doti1 = dxi * dx1 + dyi * dy1;
    const int doti2 = dxi * dx2 + dyi * dy2, doti3 = dxi * dx3 + dyi * dy3;
    const bool good = (doti2 * doti2 == dotii * dot22) ||
                      (dot11 * doti2 * doti2 > dot22 * doti1 * doti1 &&
                       dot33 * doti2 * doti2 >= dot22 * doti3 * doti3);
    float is = 0.0f;
    if (ox == xi) {
      if (yi > oy) is = C.pb(xi, yi - 1);
      else if (yi < oy) is = C.pb(ox, oy - 1);
    } else if (oy == yi) {
      if (xi > ox) is = C.pb(xi - 1, yi);
      else if (xi < ox) is = C.pb(ox - 1, oy);
    } else {
      // diagonal step: two horizontal and two vertical edge crossings
      const int hy = yi > oy ? yi - 1 : oy - 1;        // row of the H edges
      const int vx = xi > ox ? xi - 1 : ox - 1;        // column of the V edges
      is = fmaxf(C.pb(ox, hy), is);
      is = fmaxf(C.pb(xi, hy), is);
      is = fmaxf(C.pb(vx, oy), is);
      is = fmaxf(C.pb(vx, yi), is);
    }
    maxpb = fmaxf(maxpb, is);
    if (!good) continue;
    const int xo = xi - x0, yo = yi - y0;
    if (xo * xo + yo * yo <= 25) {
      const float val = affinity(c_ic_ss(maxpb), rs);
      M[(size_t)dm.d[(yo + 5) * 11 + xo + 5] * pitch + (size_t)y0 * C.w + x0] = val;
    }
  }
}

// intervening.cu:298-475 findAffinities
__global__ __launch_bounds__(256) void k_affinities(int w, int h, float rs, DiagMap dm,
                                                    const float* __restrict__ mpb, float* __restrict__ M,
                                                    size_t pitch) {
  constexpr int LW = ITX + 2 * HALO, LH = ITY + 2 * HALO;
  __shared__ float tile[LH * LW];
  const int bx = blockIdx.x * ITX, by = blockIdx.y * ITY;
  for (int i = threadIdx.x; i < LW * LH; i += 256) {
    const int ly = i / LW, lx = i - ly * LW;
    const int gx = dev::clampi(bx + lx - HALO, 0, w - 1), gy = dev::clampi(by + ly - HALO, 0, h - 1);
    tile[i] = mpb[(size_t)gy * w + gx];
  }
  __syncthreads();
  const int x0 = bx + (threadIdx.x % ITX), y0 = by + (threadIdx.x / ITX);
  if (x0 >= w || y0 >= h) return;
  // The tile holds image pixels (bx-HALO .. bx+ITX+HALO-1) clamped; every read
  // lies within 5 (+1 for edge crossings) of (x0, y0) and is clamped to the
  // image first, so it always falls inside the tile.
  IcCtx C{tile, LW, bx - HALO, by - HALO, w, h};
  const int r = 5;
  M[(size_t)dm.d[r * 11 + r] * pitch + (size_t)y0 * w + x0] = 1.0f;
  const int rxa = max(0, x0 - r), rya = max(0, y0 - r);
  const int rxb = min(x0 + r, w - 1), ryb = min(y0 + r, h - 1);
  if (x0 > rxa) {
    if ((y0 > rya) && (y0 < ryb)) ic_walk(C, dm, rs, M, pitch, x0, y0, rxa, y0 - 1, rxa, y0, rxa, y0 + 1);
    for (int y = y0 - 1; y > rya; y--) ic_walk(C, dm, rs, M, pitch, x0, y0, rxa, y - 1, rxa, y, rxa, y + 1);
  }
  if (x0 > rxa + 1 || y0 > rya + 1 || ((x0 > rxa) && (y0 > rya)))
    ic_walk(C, dm, rs, M, pitch, x0, y0, rxa, rya + 1, rxa, rya, rxa + 1, rya);
  if (y0 > rya)
    for (int x = rxa + 1; x < rxb; x++) ic_walk(C, dm, rs, M, pitch, x0, y0, x - 1, rya, x, rya, x + 1, rya);
  if ((y0 > rya + 1) || (x0 < rxb - 1) || ((y0 > rya) && (x0 < rxb)))
    ic_walk(C, dm, rs, M, pitch, x0, y0, rxb - 1, rya, rxb, rya, rxb, rya + 1);
  if (x0 < rxb)
    for (int y = rya + 1; y < y0; y++) ic_walk(C, dm, rs, M, pitch, x0, y0, rxb, y - 1, rxb, y, rxb, y + 1);
  if (x0 > rxa)
    for (int y = y0 + 1; y < ryb; y++) ic_walk(C, dm, rs, M, pitch, x0, y0, rxa, y - 1, rxa, y, rxa, y + 1);
  if ((x0 > rxa + 1) || (y0 < ryb - 1) || ((x0 > rxa) && (y0 < ryb)))
    ic_walk(C, dm, rs, M, pitch, x0, y0, rxa, ryb - 1, rxa, ryb, rxa + 1, ryb);
  if (y0 < ryb)
    for (int x = rxa + 1; x < rxb; x++) ic_walk(C, dm, rs, M, pitch, x0, y0, x - 1, ryb, x, ryb, x + 1, ryb);
  if ((y0 < ryb - 1) || (x0 < rxb - 1) || ((y0 < ryb) && (x0 < rxb)))
    ic_walk(C, dm, rs, M, pitch, x0, y0, rxb - 1, ryb, rxb, ryb, rxb, ryb - 1);
  if (x0 < rxb) {
    for (int y = ryb - 1; y > y0; y--) ic_walk(C, dm, rs, M, pitch, x0, y0, rxb, y - 1, rxb, y, rxb, y + 1);
    if ((y0 > 0) && (y0 < ryb)) ic_walk(C, dm, rs, M, pitch, x0, y0, rxb, y0 - 1, rxb, y0, rxb, y0 + 1);
  }
  // intervening.cu:459-471 corner fixes (affinity of ic = 0)
  const float fix = affinity(c_ic_ss(0.0f), rs);
  const size_t at = (size_t)y0 * w + x0;
  if ((y0 == rya + 1) && ((x0 == rxa) || (x0 == rxb))) M[(size_t)dm.d[(-1 + 5) * 11 + 5] * pitch + at] = fix;
  if ((x0 == rxa + 1) && ((y0 == rya) || (y0 == ryb))) M[(size_t)dm.d[5 * 11 + (-1 + 5)] * pitch + at] = fix;
  if ((x0 == rxb - 1) && ((y0 == rya) || (y0 == ryb))) M[(size_t)dm.d[5 * 11 + (1 + 5)] * pitch + at] = fix;
  if ((y0 == ryb - 1) && ((x0 == rxa) || (x0 == rxb))) M[(size_t)dm.d[(1 + 5) * 11 + 5] * pitch + at] = fix;
}

// intervening.cu:477-511 symmetrizeMatrix.
__global__ void k_symmetrize(int w, int h, DiagMap dm, float* __restrict__ M, size_t pitch) {
  const int x0 = blockIdx.x * blockDim.x + threadIdx.x, y0 = blockIdx.y;
  if (x0 >= w) return;
  for (int yo = 0; yo <= 5; yo++)
    for (int xo = -5; xo <= 5; xo++) {
      if (xo * xo + yo * yo > 25 || !((yo > 0) || (xo > 0))) continue;
      const int xx = x0 + xo, yy = y0 + yo;
      if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
      float* p1 = &M[(size_t)dm.d[(yo + 5) * 11 + xo + 5] * pitch + (size_t)y0 * w + x0];
      float* p2 = &M[(size_t)dm.d[(5 - yo) * 11 + 5 - xo] * pitch + (size_t)yy * w + xx];
      const float s = 0.5f * (*p1 + *p2);
      *p1 = s;
      *p2 = s;
    }
}

__global__ void k_zero_f(float* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

}  // namespace

int launch_intervene(hipStream_t s, int w, int h, const float* mpb, float* M, int pitch, float sigma) {
  DiagMap dm;
  int offs[121];
  stencil_geometry(kStencilRadius, w, offs, dm.d);
  const size_t total = (size_t)kStencilArea * pitch;
  hipLaunchKernelGGL(k_zero_f, dim3(div_up(total, 256)), dim3(256), 0, s, M, total);
  hipLaunchKernelGGL(k_affinities, dim3(div_up(w, ITX), div_up(h, ITY)), dim3(256), 0, s, w, h, 1.0f / sigma,
                     dm, mpb, M, (size_t)pitch);
  hipLaunchKernelGGL(k_symmetrize, dim3(div_up(w, 64), h), dim3(64), 0, s, w, h, dm, M, (size_t)pitch);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

}  // namespace dmsc
