// localcues.hip -- brightness / colour / texture gradients (reference:
// localcues/localcues.cu, gradient.cu, rotate.cu, parabola.cu).
//
// All 4 channels x 4 rotations (16 frames) are processed together so every
// launch has enough waves to fill 256 CUs:
//   1. labels:    quantised L/a/b (25 bins) or textons (32/64), mirrored (border
//                 30) and scattered through the Bresenham rotation tables into
//                 uint8 rotated frames (255 = outside the rotated image);
//   2. integrals: per-bin 2-D inclusive counts, stored as uint16 modulo
//                 2^16: every half-disc count is a signed sum of four corners
//                 whose true value lies in [0, 2r^2], so modular arithmetic
//                 recovers it exactly at half the bytes of int32.  A frame's
//                 integral is BP/8 planes of 8 bins (16 bytes) per rotated
//                 pixel: a wave's 16-byte gathers of one plane at consecutive
//                 pixels read one contiguous run;
//   3. gradients: per pixel, scale and split: 4+4 corner gathers per bin,
//                 optional bin blur, normalisation to sum 8 and chi-squared
//                 with the reference's 32-slot tree sums (64 textons: two
//                 lanes per pixel, one 32-bin half each);
//   4. parabola:  Savitzky-Golay smoothing of each mirrored gradient plane
//                 over the filter's non-zero taps (v outer, u inner, fma).
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>

#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {

namespace {

constexpr int kFrames = 16;
constexpr int kRadii[4] = {3, 5, 10, 20};

struct FrameGeom {
  int w, h, bw, bh;
  int nw[4], nh[4];
  long long lab_off[kFrames];  // byte offset of frame f = ch*4 + o (label buffer)
  long long int_off[kFrames];  // uint16 element offset of frame f (integral buffer)
  long long band_off[kFrames];  // uint16 element offset of frame f's band sums (k_integral_band)
  int tab_off[4];              // offset of orientation o's tables: bX[bw] bY[bw] gX[bh] gY[bh]
  int rect[4][4][8];           // [orientation][radius index][corner A..H]
  int tbp;                     // bins per integral pixel of the texture frames (32 or 64)
};

// Bins stored per integral-image pixel of frame f: 32 for the L/a/b frames
// (25 used), 32 or 64 for the texton frames 12..15.
__device__ __forceinline__ int bin_pitch(const FrameGeom& g, int f) { return f >= 12 ? g.tbp : 32; }

// Element index (uint16) of bin `bin` at rotated pixel `pix` of frame f:
// plane bin / 8 holds 8 bins per pixel.
__device__ __forceinline__ size_t int_index(const FrameGeom& g, int f, long long pix, int bin) {
  const int o = f & 3;
  const size_t npx = (size_t)g.nw[o] * g.nh[o];
  return (size_t)g.int_off[f] + (size_t)(bin >> 3) * npx * 8 + (size_t)pix * 8 + (bin & 7);
}

struct CueOut {
  float* p[4];  // Bg, Cga, Cgb, Tg; each [24][pitch]
};

__global__ void k_fill_u8(uint8_t* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0xff;
}

// gradient.cu:210-223 (quantizeImage, mirrorImage) + rotate.cu:247-259
// (bresenhamRotate): one mirrored pixel per thread, frames f0, f0 + 1, ...
__global__ void k_cue_labels(FrameGeom g, int f0, const float* __restrict__ L, const float* __restrict__ A,
                             const float* __restrict__ B, const int* __restrict__ tex,
                             const int* __restrict__ tabs, uint8_t* __restrict__ frames) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x, Y = blockIdx.y;
  const int f = f0 + blockIdx.z, ch = f >> 2, o = f & 3;
  if (X >= g.bw) return;
  const size_t si = (size_t)dev::mirror(Y, kBorder, g.h) * g.w + dev::mirror(X, kBorder, g.w);
  int lab;
  if (ch == 3) {
    lab = tex[si];
  } else {
    const float v = (ch == 0 ? L : ch == 1 ? A : B)[si];
    lab = (int)floorf(v * 25.0f);
    if (lab == 25) lab = 24;
  }
  const int* t = tabs + g.tab_off[o];
  const int rx = t[2 * g.bw + Y] + t[X];
  const int ry = t[2 * g.bw + g.bh + Y] + t[g.bw + X];
  frames[g.lab_off[f] + (size_t)ry * g.nw[o] + rx] = (uint8_t)lab;
}

// gradient.cu:66-87: per-bin inclusive prefix counts along each row of a
// rotated frame with BP bins per pixel.  Block = one row; thread = (bin,
// one of 256/BP row segments).  Frames f0 + blockIdx.y.
template <int BP>
__global__ __launch_bounds__(256) void k_integral_rows(FrameGeom g, int f0, const uint8_t* __restrict__ frames,
                                                       uint16_t* __restrict__ I) {
  constexpr int NSEG = 256 / BP;
  extern __shared__ uint8_t row[];
  __shared__ uint32_t segtot[NSEG][BP];
  const int f = f0 + blockIdx.y, o = f & 3, y = blockIdx.x;
  const int nw = g.nw[o];
  if (y >= g.nh[o]) return;
  const uint8_t* src = frames + g.lab_off[f] + (size_t)y * nw;
  for (int i = threadIdx.x; i < nw; i += 256) row[i] = src[i];
  __syncthreads();
  const int bin = threadIdx.x % BP, seg = threadIdx.x / BP;
  const int len = (nw + NSEG - 1) / NSEG, x0 = seg * len, x1 = min(nw, x0 + len);
  uint32_t c = 0;
  for (int x = x0; x < x1; x++) c += (row[x] == bin);
  segtot[seg][bin] = c;
  __syncthreads();
  uint32_t run = 0;
  for (int s = 0; s < seg; s++) run += segtot[s][bin];
  uint16_t* dst = I + int_index(g, f, (long long)y * nw, bin);
  for (int x = x0; x < x1; x++) {
    run += (row[x] == bin);
    dst[(size_t)x * 8] = (uint16_t)run;
  }
}

// gradient.cu:89-103: column pass, two bins per thread, modulo 2^16.
__global__ __launch_bounds__(256) void k_integral_cols(FrameGeom g, int f0, uint16_t* __restrict__ I) {
  const int f = f0 + blockIdx.y, o = f & 3;
  const int nw = g.nw[o], nh = g.nh[o], pairs = bin_pitch(g, f) / 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // (x, bin pair)
  if (t >= nw * pairs) return;
  const int x = t / pairs, pr = t - x * pairs;
  uint32_t* col = reinterpret_cast<uint32_t*>(I + int_index(g, f, x, 2 * pr));
  const size_t stride = (size_t)nw * 4;  // u32 words per plane row
  uint32_t lo = 0, hi = 0;
  int y = 0;
  for (; y + 8 <= nh; y += 8) {
    uint32_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = col[(size_t)(y + k) * stride];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      lo += v[k] & 0xffffu;
      hi += v[k] >> 16;
      col[(size_t)(y + k) * stride] = (lo & 0xffffu) | (hi << 16);
    }
  }
  for (; y < nh; y++) {
    const uint32_t v = col[(size_t)y * stride];
    lo += v & 0xffffu;
    hi += v >> 16;
    col[(size_t)y * stride] = (lo & 0xffffu) | (hi << 16);
  }
}

// Single-write integral images.  A frame is cut into bands of kBand rows;
// I(x, y) for y in band b is the carry C_b(x) = sum over earlier bands of
// their row-prefixed column counts P_b'(x) plus the band's own running sum of
// row prefixes.  Pass 1 (FINAL = false) leaves P_b per band, k_band_carry
// turns them into the exclusive prefix over bands in place, pass 2 (FINAL)
// starts each band's running sums at its carry and writes every integral
// value once.  One block of 1024 threads per (band, frame); thread t owns
// columns x = t + 1024 c (c < NCH) and keeps their running sums for all
// bins in registers (two u16 bins per word, modulo 2^16 like the stored
// integrals).  The row prefix of bin k at column x is the count of k among
// the lanes below x in its wave (ballot + mbcnt) plus the exclusive total
// of the earlier waves (their totals through LDS, one barrier per row).
// Each thread stores its pixel's BP bins as BP/8 16-byte plane entries, so
// a wave's store to one plane is one contiguous kilobyte.
constexpr int kBand = 16;
constexpr int kIntThreads = 1024, kIntWaves = kIntThreads / 64;

template <int BP, int NBU, int NCH, bool FINAL>
__global__ __launch_bounds__(kIntThreads) void k_integral_band(FrameGeom g, int f0,
                                                               const uint8_t* __restrict__ frames,
                                                               uint16_t* __restrict__ bands,
                                                               uint16_t* __restrict__ I) {
  constexpr int NQ = BP / 2;  // words per record (two bins each)
  constexpr int NQU = (NBU + 1) / 2;  // words holding used bins (labels are < NBU; higher bins stay 0)
  constexpr int NS = NCH * kIntWaves;
  __shared__ uint32_t tot[2][NS][NBU];  // per-wave bin totals, by row parity
  const int f = f0 + blockIdx.y, o = f & 3, b = blockIdx.x;
  const int nw = g.nw[o], nh = g.nh[o];
  const int y0 = b * kBand;
  if (y0 >= nh) return;  // the whole block: past this frame's last band
  const int y1 = min(nh, y0 + kBand);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint16_t* bsum = bands + g.band_off[f] + (size_t)b * nw * BP;
  uint32_t S[NCH][NQ];
#pragma unroll
  for (int c = 0; c < NCH; c++) {
    const int x = c * kIntThreads + (int)threadIdx.x;
#pragma unroll
    for (int q = 0; q < NQ; q += 4) {
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (FINAL && x < nw) v = *reinterpret_cast<const uint4*>(bsum + (size_t)x * BP + 2 * q);
      S[c][q] = v.x;
      S[c][q + 1] = v.y;
      S[c][q + 2] = v.z;
      S[c][q + 3] = v.w;
    }
  }
  const uint8_t* lab_base = frames + g.lab_off[f];
  uint32_t labn[NCH];  // the next row's labels, loaded one row ahead
#pragma unroll
  for (int c = 0; c < NCH; c++) {
    const int x = c * kIntThreads + (int)threadIdx.x;
    labn[c] = x < nw ? (uint32_t)lab_base[(size_t)y0 * nw + x] : 255u;
  }
  for (int y = y0; y < y1; y++) {
    const int par = (y - y0) & 1;
    uint32_t lab[NCH];
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      lab[c] = labn[c];
      const int x = c * kIntThreads + (int)threadIdx.x;
      labn[c] = (y + 1 < y1 && x < nw) ? (uint32_t)lab_base[(size_t)(y + 1) * nw + x] : 255u;
    }
    uint32_t Rl[NCH][NQU];
#pragma unroll
    for (int c = 0; c < NCH; c++) {
#pragma unroll
      for (int q = 0; q < NQU; q++) {
        const unsigned long long m0 = __ballot(lab[c] == (uint32_t)(2 * q));
        const uint32_t r0 = __builtin_amdgcn_mbcnt_hi((uint32_t)(m0 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m0, 0u)) +
                            (lab[c] == (uint32_t)(2 * q));
        uint32_t r1 = 0;
        if (2 * q + 1 < NBU) {
          const unsigned long long m1 = __ballot(lab[c] == (uint32_t)(2 * q + 1));
          r1 = __builtin_amdgcn_mbcnt_hi((uint32_t)(m1 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m1, 0u)) +
               (lab[c] == (uint32_t)(2 * q + 1));
          if (lane == 0) tot[par][c * kIntWaves + wv][2 * q + 1] = (uint32_t)__popcll(m1);
        }
        if (lane == 0) tot[par][c * kIntWaves + wv][2 * q] = (uint32_t)__popcll(m0);
        Rl[c][q] = r0 | (r1 << 16);
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      const int x = c * kIntThreads + (int)threadIdx.x;
      // bin `lane`'s count over the row's earlier waves, then broadcast per bin
      uint32_t pv = 0;
      if (lane < NBU)
        for (int k = 0; k < c * kIntWaves + wv; k++) pv += tot[par][k][lane];
#pragma unroll
      for (int q = 0; q < NQU; q++) {
        // row prefix of bins 2q, 2q+1 (each below 2^16: no carry between halves)
        const uint32_t p0 = (uint32_t)__builtin_amdgcn_readlane((int)pv, 2 * q);
        const uint32_t p1 = 2 * q + 1 < NBU ? (uint32_t)__builtin_amdgcn_readlane((int)pv, 2 * q + 1) : 0u;
        const uint32_t r = Rl[c][q] + (p0 | (p1 << 16));
        const uint32_t lo = (S[c][q] + r) & 0xffffu;
        const uint32_t hi = ((S[c][q] >> 16) + (r >> 16)) << 16;
        S[c][q] = lo | hi;
      }
      if (FINAL && x < nw) {
        // plane q / 4 of this pixel: 16 bytes; consecutive lanes, consecutive pixels
        const size_t npx = (size_t)nw * nh;
        uint4* dst = reinterpret_cast<uint4*>(I + g.int_off[f]) + (size_t)y * nw + x;
#pragma unroll
        for (int q = 0; q < NQ; q += 4) dst[(q / 4) * npx] = make_uint4(S[c][q], S[c][q + 1], S[c][q + 2], S[c][q + 3]);
      }
    }
  }
  if (!FINAL) {
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      const int x = c * kIntThreads + (int)threadIdx.x;
      if (x < nw) {
        uint4* dst = reinterpret_cast<uint4*>(bsum + (size_t)x * BP);
#pragma unroll
        for (int q = 0; q < NQ; q += 4) dst[q / 4] = make_uint4(S[c][q], S[c][q + 1], S[c][q + 2], S[c][q + 3]);
      }
    }
  }
}

// Band sums -> exclusive prefix over the bands of each frame (in place).
__global__ __launch_bounds__(256) void k_band_carry(FrameGeom g, int f0, uint16_t* __restrict__ bands) {
  const int f = f0 + blockIdx.y, o = f & 3;
  const int nw = g.nw[o], nb = (g.nh[o] + kBand - 1) / kBand, bp = bin_pitch(g, f);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nw * bp) return;
  uint16_t* p = bands + g.band_off[f] + e;
  const size_t stride = (size_t)nw * bp;
  uint16_t acc = 0;
  for (int b = 0; b < nb; b++) {
    const uint16_t t = p[(size_t)b * stride];
    p[(size_t)b * stride] = acc;
    acc = (uint16_t)(acc + t);
  }
}

struct BlurK {
  float k[16];
};

// rotate.cu:263-475 computeGradient: one (pixel, scale, split) per thread.
// NB bins, KR blur radius (0 = no blur).  Output plane of channel ch:
// [scale][slot] with slot o+4 for the top/bottom split (sense A) and slot o
// for the left/right split (sense B) (gradient.cu:240-241).
// chi^2 of one split of the half-disc pair from the packed corner counts:
// a = c[i0] - c[i1] + c[i2] - c[i3], b = c[j0] - c[j1] + c[j2] - c[j3] per
// bin (uint16 modulo arithmetic recovers the exact counts), optional bin
// blur, normalisation to sum 8 and 0.5 * sum (a-b)^2/(a+b) with the
// reference's 32-slot tree sums (rotate.cu:352-472).
// PAIR: the pixel's 64 bins are split over this lane (32 bins in pk) and
// its neighbour lane (lane ^ 1); each 32-slot tree sum is completed with the
// partner's as lo + hi (rotate.cu:384-389 adds slot 32 to slot 0), and IEEE
// addition being commutative both lanes hold the same total.
template <int NB, int KR, bool PAIR = false>
__device__ __forceinline__ float chi2_split(const uint32_t (&pk)[8][16], int i0, int i1, int i2, int i3, int j0,
                                            int j1, int j2, int j3, const BlurK& bk) {
  static_assert(!PAIR || (NB == 32 && KR == 0), "paired halves hold 32 unblurred bins");
  float a[32], b[32];
  // two bins per 32-bit word: the corner sums in packed 16-bit arithmetic,
  // each half modulo 2^16 on its own
  typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
#pragma unroll
  for (int e = 0; e < 16; e++) {
    if (2 * e < NB) {
      const u16x2 s0 = __builtin_bit_cast(u16x2, pk[i0][e]) - __builtin_bit_cast(u16x2, pk[i1][e]) +
                       __builtin_bit_cast(u16x2, pk[i2][e]) - __builtin_bit_cast(u16x2, pk[i3][e]);
      const u16x2 s1 = __builtin_bit_cast(u16x2, pk[j0][e]) - __builtin_bit_cast(u16x2, pk[j1][e]) +
                       __builtin_bit_cast(u16x2, pk[j2][e]) - __builtin_bit_cast(u16x2, pk[j3][e]);
      a[2 * e] = (float)s0.x;
      b[2 * e] = (float)s1.x;
      a[2 * e + 1] = 2 * e + 1 < NB ? (float)s0.y : 0.f;
      b[2 * e + 1] = 2 * e + 1 < NB ? (float)s1.y : 0.f;
    } else {
      a[2 * e] = a[2 * e + 1] = 0.f;
      b[2 * e] = b[2 * e + 1] = 0.f;
    }
  }
  if (KR > 0) {  // rotate.cu:352-370: zero-padded correlation over bins
    float ca[NB], cb[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) {
      float sa = 0.f, sb = 0.f;
#pragma unroll
      for (int nk = 0; nk < 2 * KR + 1; nk++) {
        // a padding term fma(0, k, s) leaves s unchanged (s >= +0, k
        // finite), so it is not formed
        const int j = i - KR + nk;
        if (j >= 0 && j < NB) {
          sa = fmaf(a[j], bk.k[nk], sa);
          sb = fmaf(b[j], bk.k[nk], sb);
        }
      }
      ca[i] = sa;
      cb[i] = sb;
    }
#pragma unroll
    for (int i = 0; i < NB; i++) {
      a[i] = ca[i];
      b[i] = cb[i];
    }
  }
  // slots NB..31 hold +0 and every other slot is >= +0, so the tree sums
  // skip the additions of those zeros (x + 0 = x exactly) and the chi terms
  // of the empty slots are not formed
  float t32[32];
#pragma unroll
  for (int i = 0; i < 32; i++) t32[i] = a[i];
  float sa = dev::tree32z<NB>(t32);
  if (PAIR) sa = sa + __shfl_xor(sa, 1, 64);
  const float na = (sa != 0.f) ? 8.0f / sa : 0.f;
#pragma unroll
  for (int i = 0; i < 32; i++) t32[i] = b[i];
  float sb = dev::tree32z<NB>(t32);
  if (PAIR) sb = sb + __shfl_xor(sb, 1, 64);
  const float nb = (sb != 0.f) ? 8.0f / sb : 0.f;
#pragma unroll
  for (int i = 0; i < 32; i++) {
    float chi = 0.f;
    if (i < NB) {
      const float ai = a[i] * na, bi = b[i] * nb;
      const float s = ai + bi;
      if (s != 0.f) {
        const float d = ai - bi;
        chi = d * d / s;
      }
    }
    t32[i] = chi;
  }
  float c = dev::tree32z<NB>(t32);
  if (PAIR) c = c + __shfl_xor(c, 1, 64);
  return 0.5f * c;
}

// Both splits of one (pixel, scale, orientation): the eight corners A..H
// are loaded once (uint16 counts, 8 bins per 16-byte load) and feed
//   sense A (top/bottom, slot o+4): a = A-C+E-D, b = H-F+D-E
//   sense B (left/right, slot o):   a = A-F+G-B, b = H-C+B-G
// (rotate.cu:400-445; gradient.cu:240-241).
// Blocks: blockIdx.x runs over (pixel block, scale) with the scale fastest,
// numbered XCD-aware, so each XCD handles a contiguous band of image rows at
// all three scales and its L2 keeps the band's rotated integral rows (plus
// the corner reach above and below) for every gather; blockIdx.y is the
// (channel, orientation).
template <int NB, int KR>
__global__ __launch_bounds__(256) void k_gradients(FrameGeom g, int ch0, int rad0, const int* __restrict__ tabs,
                                                   const uint16_t* __restrict__ I, BlurK bk,
                                                   float* __restrict__ grads, size_t plane) {
  const int n = g.w * g.h;
  const int lb = dev::xcd_block(blockIdx.x, gridDim.x);
  const int p = (lb / 3) * blockDim.x + threadIdx.x;
  const int scale = lb % 3;
  const int ch = ch0 + (blockIdx.y >> 2), o = blockIdx.y & 3, f = ch * 4 + o;
  if (p >= n) return;
  const int y = p / g.w, x = p - y * g.w;
  const int* t = tabs + g.tab_off[o];
  const int ix = x + kBorder, iy = y + kBorder;
  const int rx = t[2 * g.bw + iy] + t[ix];
  const int ry = t[2 * g.bw + g.bh + iy] + t[g.bw + ix];
  const long long pix = (long long)ry * g.nw[o] + rx;
  const int* rc = g.rect[o][rad0 + scale];
  const uint4* base = reinterpret_cast<const uint4*>(I + g.int_off[f]);
  const size_t npx = (size_t)g.nw[o] * g.nh[o];
  constexpr int NQ = (NB + 7) / 8;
  uint32_t pk[8][16];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const uint4* src = base + (size_t)(rc[c] + pix);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      if (q < NQ) {
        const uint4 v = src[q * npx];
        pk[c][4 * q] = v.x;
        pk[c][4 * q + 1] = v.y;
        pk[c][4 * q + 2] = v.z;
        pk[c][4 * q + 3] = v.w;
      } else {
        pk[c][4 * q] = pk[c][4 * q + 1] = pk[c][4 * q + 2] = pk[c][4 * q + 3] = 0u;
      }
    }
  }
  // corner indices: A0 B1 C2 D3 E4 F5 G6 H7
  const float top = chi2_split<NB, KR>(pk, 0, 2, 4, 3, 7, 5, 3, 4, bk);
  const float left = chi2_split<NB, KR>(pk, 0, 5, 6, 1, 7, 2, 1, 6, bk);
  float* dst = grads + (size_t)((ch * 3 + scale) * 8) * plane + p;
  dst[(size_t)(o + 4) * plane] = top;
  dst[(size_t)o * plane] = left;
}

// The texture channel with 64 textons (rotate.cu:780-782,
// computeGradient<64, 64, false, *>): lanes 2p and 2p+1 share pixel p,
// holding bins 0..31 and 32..63 of every corner (integral planes 0..3 and
// 4..7).  Output as k_gradients.
__global__ __launch_bounds__(256) void k_gradients64(FrameGeom g, int rad0, const int* __restrict__ tabs,
                                                     const uint16_t* __restrict__ I, float* __restrict__ grads,
                                                     size_t plane) {
  const int n = g.w * g.h;
  const int lb = dev::xcd_block(blockIdx.x, gridDim.x);  // (pixel block, scale), as k_gradients
  const int gt = (lb / 3) * blockDim.x + threadIdx.x;
  const int half = gt & 1;
  const bool live = (gt >> 1) < n;
  const int p = live ? (gt >> 1) : n - 1;  // both lanes of a pair agree; all lanes reach the shuffles
  const int scale = lb % 3;
  const int ch = 3, o = blockIdx.y, f = ch * 4 + o;
  const int y = p / g.w, x = p - y * g.w;
  const int* t = tabs + g.tab_off[o];
  const int ix = x + kBorder, iy = y + kBorder;
  const int rx = t[2 * g.bw + iy] + t[ix];
  const int ry = t[2 * g.bw + g.bh + iy] + t[g.bw + ix];
  const long long pix = (long long)ry * g.nw[o] + rx;
  const int* rc = g.rect[o][rad0 + scale];
  const size_t npx = (size_t)g.nw[o] * g.nh[o];
  const uint4* base = reinterpret_cast<const uint4*>(I + g.int_off[f]) + (size_t)(4 * half) * npx;  // planes 4h..4h+3
  uint32_t pk[8][16];
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const uint4* src = base + (size_t)(rc[c] + pix);
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint4 v = src[q * npx];
      pk[c][4 * q] = v.x;
      pk[c][4 * q + 1] = v.y;
      pk[c][4 * q + 2] = v.z;
      pk[c][4 * q + 3] = v.w;
    }
  }
  BlurK none{};
  const float top = chi2_split<32, 0, true>(pk, 0, 2, 4, 3, 7, 5, 3, 4, none);
  const float left = chi2_split<32, 0, true>(pk, 0, 5, 6, 1, 7, 2, 1, 6, none);
  if (live && half == 0) {
    float* dst = grads + (size_t)((ch * 3 + scale) * 8) * plane + p;
    dst[(size_t)(o + 4) * plane] = top;
    dst[(size_t)o * plane] = left;
  }
}

// Parabola smoothing (parabola.cu:75-124 / parabola_kernel.cu:1-30) of one
// scale of a channel group, plane z = (channel - ch0) * 8 + orientation.
// Each thread computes QX horizontally adjacent outputs.  For every filter
// row v (the reference's outer loop, ascending) it loads the QX + 2R tile
// values that row needs into registers once, then runs the fully unrolled u
// loop (inner, ascending) over the dense (2R+1)-wide row.  Taps outside the
// ellipse have coefficient 0 and fma(x, 0, acc) == acc for finite x (acc
// starts at +0 and can never become -0), so the sums equal the reference's
// zero-skipping loop bit for bit; rows with no non-zero tap are skipped.
constexpr int QX = 8, QCOLS = 16, QTX = QX * QCOLS, QTY = 16;

template <int R>
constexpr int q_lw() {
  return (QTX + 2 * R + 3) / 4 * 4;  // tile row length, float4-aligned
}

template <int R>
__global__ __launch_bounds__(256) void k_parabola(int w, int h, int scale, int ch0, const float* __restrict__ grads,
                                                  size_t gplane, const float* __restrict__ coef,
                                                  const int2* __restrict__ vrange, CueOut out, int pitch) {
  extern __shared__ float4 tile4[];
  float* tile = reinterpret_cast<float*>(tile4);
  constexpr int LW = q_lw<R>(), LH = QTY + 2 * R, D = 2 * R + 1;
  constexpr int SEG = (QX + 2 * R + 3) / 4 * 4;
  const int ch = ch0 + blockIdx.z / 8, o = blockIdx.z % 8;
  const float* src = grads + (size_t)((ch * 3 + scale) * 8 + o) * gplane;
  const int x0 = blockIdx.x * QTX, y0 = blockIdx.y * QTY;
  for (int i = threadIdx.x; i < LW * LH; i += 256) {
    const int ly = i / LW, lx = i - ly * LW;
    // tile columns past the image only feed discarded outputs: clamp into
    // the mirrored range [0, w + 2*border) before reflecting
    const int gx = dev::mirror(dev::clampi(x0 + lx - R + kBorder, 0, w + 2 * kBorder - 1), kBorder, w);
    const int gy = dev::mirror(dev::clampi(y0 + ly - R + kBorder, 0, h + 2 * kBorder - 1), kBorder, h);
    tile[i] = src[(size_t)gy * w + gx];
  }
  __syncthreads();
  const int tx = (threadIdx.x % QCOLS) * QX, ty = threadIdx.x / QCOLS;
  const float* f = coef + (size_t)o * D * D;
  const int2 vr = vrange[o];
  float acc[QX];
#pragma unroll
  for (int j = 0; j < QX; j++) acc[j] = 0.0f;
  for (int v = vr.x; v <= vr.y; v++) {
    const float4* rowp = reinterpret_cast<const float4*>(tile + (ty + R + v) * LW + tx);
    float seg[SEG];
#pragma unroll
    for (int q = 0; q < SEG / 4; q++) {
      const float4 t = rowp[q];
      seg[4 * q] = t.x;
      seg[4 * q + 1] = t.y;
      seg[4 * q + 2] = t.z;
      seg[4 * q + 3] = t.w;
    }
    const float* fr = f + (v + R) * D;
#pragma unroll
    for (int u = 0; u < D; u++) {
      const float c = fr[u];
#pragma unroll
      for (int j = 0; j < QX; j++) acc[j] = fmaf(seg[j + u], c, acc[j]);
    }
  }
  float* dst = out.p[ch] + (size_t)(scale * 8 + o) * pitch;
  const int gy = y0 + ty;
  if (gy < h)
#pragma unroll
    for (int j = 0; j < QX; j++)
      if (x0 + tx + j < w) dst[(size_t)gy * w + x0 + tx + j] = acc[j];
}

template <int R>
size_t parabola_lds() {
  return sizeof(float) * (size_t)q_lw<R>() * (QTY + 2 * R) + 16;
}

// ---------------------------------------------------------------- host plan

struct CuePlan {
  int w = -1, h = -1;
  FrameGeom g;
  size_t lab_bytes = 0, int_elems = 0, band_elems = 0;
  bool band_ok = false;  // frames narrow enough for the single-write integral pass
  int tap_stride = 0;
  int* d_tabs = nullptr;
  float* d_coef[4] = {nullptr, nullptr, nullptr, nullptr};  // [8][2r+1][2r+1] per radius
  int2* d_vrange[4] = {nullptr, nullptr, nullptr, nullptr};  // rows holding non-zero taps
  BlurK blur_bg, blur_cg;
  int kr_bg = 0, kr_cg = 0;
  ~CuePlan() {
    if (d_tabs) (void)hipFree(d_tabs);
    for (int i = 0; i < 4; i++) {
      if (d_coef[i]) (void)hipFree(d_coef[i]);
      if (d_vrange[i]) (void)hipFree(d_vrange[i]);
    }
  }
};

int build_plan(hipStream_t s, int w, int h, int tbp, CuePlan& P) {
  FrameGeom& g = P.g;
  memset(&g, 0, sizeof(g));
  g.w = w; g.h = h; g.bw = w + 2 * kBorder; g.bh = h + 2 * kBorder;
  g.tbp = tbp;
  std::vector<int> tabs;
  std::vector<int> bX(g.bw), bY(g.bw), gX(g.bh), gY(g.bh);
  for (int o = 0; o < 4; o++) {
    const float thetaPi = -float(o) / float(8);
    if (rotation_tables(g.bw, g.bh, thetaPi, &g.nw[o], &g.nh[o], bX.data(), bY.data(), gX.data(), gY.data()))
      return fail(DMSC_ERR_INTERNAL, "rotation tables failed for %dx%d", w, h);
    g.tab_off[o] = (int)tabs.size();
    tabs.insert(tabs.end(), bX.begin(), bX.end());
    tabs.insert(tabs.end(), bY.begin(), bY.end());
    tabs.insert(tabs.end(), gX.begin(), gX.end());
    tabs.insert(tabs.end(), gY.begin(), gY.end());
    for (int ri = 0; ri < 4; ri++)
      halfdisc_offsets(w, h, thetaPi, kRadii[ri], bX.data(), bY.data(), gX.data(), gY.data(), g.nw[o],
                       g.rect[o][ri]);
  }
  // Every gather below must stay inside its rotated frame: verify on the host
  // (labels scatter over the whole mirrored image; corners around every
  // interior pixel for all radii).
  for (int o = 0; o < 4; o++) {
    const int* t = tabs.data() + g.tab_off[o];
    const int *tbX = t, *tbY = t + g.bw, *tgX = t + 2 * g.bw, *tgY = t + 2 * g.bw + g.bh;
    for (int Y = 0; Y < g.bh; Y++)
      for (int X = 0; X < g.bw; X++) {
        const int rx = tgX[Y] + tbX[X], ry = tgY[Y] + tbY[X];
        if (rx < 0 || ry < 0 || rx >= g.nw[o] || ry >= g.nh[o])
          return fail(DMSC_ERR_INTERNAL, "rotation maps outside the frame (o=%d)", o);
      }
    long long lo_off = 0, hi_off = 0;
    for (int ri = 0; ri < 4; ri++)
      for (int k = 0; k < 8; k++) {
        lo_off = std::min(lo_off, (long long)g.rect[o][ri][k]);
        hi_off = std::max(hi_off, (long long)g.rect[o][ri][k]);
      }
    const long long npx = (long long)g.nw[o] * g.nh[o];
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        const int X = x + kBorder, Y = y + kBorder;
        const long long pix = (long long)(tgY[Y] + tbY[X]) * g.nw[o] + (tgX[Y] + tbX[X]);
        if (pix + lo_off < 0 || pix + hi_off >= npx)
          return fail(DMSC_ERR_INTERNAL, "half-disc corners leave the frame (o=%d)", o);
      }
  }
  size_t lo = 0, io = 0, bo = 0;
  for (int f = 0; f < kFrames; f++) {
    const int o = f & 3;
    g.lab_off[f] = (long long)lo;
    g.int_off[f] = (long long)io;
    g.band_off[f] = (long long)bo;
    lo += ((size_t)g.nw[o] * g.nh[o] + 255) & ~(size_t)255;
    io += (size_t)g.nw[o] * g.nh[o] * (f >= 12 ? tbp : 32);
    bo += (size_t)div_up(g.nh[o], kBand) * g.nw[o] * (f >= 12 ? tbp : 32);
  }
  P.lab_bytes = lo;
  P.int_elems = io;
  P.band_elems = bo;
  // the single-write integral pass covers frames up to 2 * kIntThreads
  // columns; wider ones use the row + column passes
  {
    int maxw = 0;
    for (int o = 0; o < 4; o++) maxw = std::max(maxw, g.nw[o]);
    P.band_ok = maxw <= 2 * kIntThreads;
  }
  DMSC_HIP(hipMalloc(&P.d_tabs, sizeof(int) * tabs.size()));
  DMSC_HIP(hipMemcpy(P.d_tabs, tabs.data(), sizeof(int) * tabs.size(), hipMemcpyHostToDevice));
  for (int ri = 0; ri < 4; ri++) {
    const int r = kRadii[ri], len = 2 * r + 1;
    std::vector<float> filt((size_t)8 * len * len);
    parabola_filters(r, filt.data());
    std::vector<int2> vr(8);
    for (int o = 0; o < 8; o++) {
      int v0 = r + 1, v1 = -r - 1;
      for (int v = -r; v <= r; v++)
        for (int u = -r; u <= r; u++)
          if (filt[(size_t)o * len * len + (v + r) * len + (u + r)] != 0.0f) {
            v0 = std::min(v0, v);
            v1 = std::max(v1, v);
          }
      vr[o] = make_int2(v0, v1);  // empty range (v0 > v1) if every tap is zero
    }
    DMSC_HIP(hipMalloc(&P.d_coef[ri], sizeof(float) * filt.size()));
    DMSC_HIP(hipMemcpy(P.d_coef[ri], filt.data(), sizeof(float) * filt.size(), hipMemcpyHostToDevice));
    DMSC_HIP(hipMalloc(&P.d_vrange[ri], sizeof(int2) * 8));
    DMSC_HIP(hipMemcpy(P.d_vrange[ri], vr.data(), sizeof(int2) * 8, hipMemcpyHostToDevice));
  }
  // gradient.cu:233: blur sigma = (int)(sigma * nbins) -> Bg 2, Cg 1
  memset(&P.blur_bg, 0, sizeof(BlurK));
  memset(&P.blur_cg, 0, sizeof(BlurK));
  blur_kernel((float)(int)(0.1f * 25.0f), P.blur_bg.k, &P.kr_bg);
  blur_kernel((float)(int)(0.05f * 25.0f), P.blur_cg.k, &P.kr_cg);
  if (P.kr_bg != 6 || P.kr_cg != 3) return fail(DMSC_ERR_INTERNAL, "unexpected blur radii");
  P.w = w;
  P.h = h;
  (void)s;
  return DMSC_OK;
}

// Integral images of frames f0 .. f0 + nf - 1 (BP bins) in three launches:
// band sums, carries, final pass (frames up to 2 * kIntThreads wide).
template <int BP, int NBU, int NCH>
int integral_bands_n(hipStream_t s, const FrameGeom& g, int f0, int nf, int maxw, int maxh, const uint8_t* frames,
                     uint16_t* bands, uint16_t* I) {
  const dim3 grid(div_up(maxh, kBand), nf);
  hipLaunchKernelGGL((k_integral_band<BP, NBU, NCH, false>), grid, dim3(kIntThreads), 0, s, g, f0, frames, bands,
                     I);
  hipLaunchKernelGGL(k_band_carry, dim3(div_up(maxw * BP, 256), nf), dim3(256), 0, s, g, f0, bands);
  hipLaunchKernelGGL((k_integral_band<BP, NBU, NCH, true>), grid, dim3(kIntThreads), 0, s, g, f0, frames, bands, I);
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

// BP bins per record, NBU of them used (the L/a/b frames hold 25 quantised
// bins in 32-bin records; the unused bins stay 0)
template <int BP, int NBU>
int integral_bands(hipStream_t s, const FrameGeom& g, int f0, int nf, int maxw, int maxh, const uint8_t* frames,
                   uint16_t* bands, uint16_t* I) {
  if (maxw <= kIntThreads) return integral_bands_n<BP, NBU, 1>(s, g, f0, nf, maxw, maxh, frames, bands, I);
  return integral_bands_n<BP, NBU, 2>(s, g, f0, nf, maxw, maxh, frames, bands, I);
}

struct PlanCache {
  std::map<std::tuple<int, int, int>, std::unique_ptr<CuePlan>> plans;
};

}  // namespace

// Plans are cached per (w, h, texture bin pitch) and context; the map is
// shared by all contexts (and host threads), so it is only touched under
// plan_lock(), and a context's plans are dropped when it is destroyed.
static std::map<const Ctx*, PlanCache>& plan_caches() {
  static std::map<const Ctx*, PlanCache> caches;
  return caches;
}
static std::mutex& plan_lock() {
  static std::mutex m;
  return m;
}

static int get_plan(Ctx* ctx, int w, int h, int tbp, CuePlan** out) {
  const auto key = std::make_tuple(w, h, tbp);
  {
    std::lock_guard<std::mutex> lk(plan_lock());
    PlanCache& pc = plan_caches()[ctx];
    auto it = pc.plans.find(key);
    if (it != pc.plans.end()) {
      *out = it->second.get();
      return DMSC_OK;
    }
  }
  // built outside the lock: only this context's thread builds its plans
  std::unique_ptr<CuePlan> p(new CuePlan());
  DMSC_TRY(build_plan(ctx->stream, w, h, tbp, *p));
  std::lock_guard<std::mutex> lk(plan_lock());
  auto it = plan_caches()[ctx].plans.emplace(key, std::move(p)).first;
  *out = it->second.get();
  return DMSC_OK;
}

void release_local_cues(const Ctx* ctx) {
  std::lock_guard<std::mutex> lk(plan_lock());
  plan_caches().erase(ctx);
}

size_t local_cues_scratch_bytes(int w, int h) {
  const int bw = w + 2 * kBorder, bh = h + 2 * kBorder;
  // generous bound: rotated frames are at most (bw+bh)^2
  const size_t rot = (size_t)(bw + bh) * (bw + bh);
  const size_t lab = 16 * (rot + 256);
  const size_t ints = (12 * 32 + 4 * 64) * rot * sizeof(uint16_t);  // texture frames at up to 64 bins
  const size_t grads = (size_t)4 * 24 * w * h * sizeof(float);
  return lab + ints + grads + 4096;
}

// localcues.cu:257-300 localCues, in two parts that may run on two streams:
// part 1 the L, a and b channels (frames 0..11: labels, integral images,
// Bg and Cg gradients and their parabola fits), which do not need the
// textons; part 2 the texton channel (frames 12..15: Tg); part 3 both.
// local_cues_prepare builds the plan and sizes the scratch arena first, so
// the parts only launch kernels and never reallocate while the other part
// runs.
int local_cues_prepare(Ctx* ctx, int w, int h, int ntex) {
  if (ntex != 32 && ntex != 64) return fail(DMSC_ERR_ARG, "local cues support 32 or 64 textons");
  CuePlan* P;
  DMSC_TRY(get_plan(ctx, w, h, ntex, &P));
  void* scratch;
  const size_t gbytes = (size_t)4 * 24 * w * h * sizeof(float);
  return ctx->scratch(1, P->lab_bytes + P->int_elems * sizeof(uint16_t) + gbytes + P->band_elems * sizeof(uint16_t) + 1024,
                      &scratch);
}

int launch_local_cues_part(Ctx* ctx, hipStream_t s, int part, int w, int h, const float* L, const float* A,
                           const float* B, const int* tex, int ntex, float* bg, float* cga, float* cgb, float* tg,
                           int pitch, float* grads_out, hipEvent_t hist_done) {
  if (ntex != 32 && ntex != 64) return fail(DMSC_ERR_ARG, "local cues support 32 or 64 textons");
  if (part < 1 || part > 3) return fail(DMSC_ERR_INTERNAL, "bad local-cue part %d", part);
  const int tbp = ntex;
  CuePlan* P;
  DMSC_TRY(get_plan(ctx, w, h, tbp, &P));
  const FrameGeom& g = P->g;
  const int n = w * h;
  void* scratch;
  const size_t gbytes = (size_t)4 * 24 * n * sizeof(float);
  const size_t ibytes = P->int_elems * sizeof(uint16_t);
  DMSC_TRY(ctx->scratch(1, P->lab_bytes + ibytes + gbytes + P->band_elems * sizeof(uint16_t) + 1024, &scratch));
  uint8_t* frames = (uint8_t*)scratch;
  uint16_t* I = (uint16_t*)((char*)scratch + P->lab_bytes);
  float* grads = grads_out ? grads_out : (float*)((char*)scratch + P->lab_bytes + ibytes);
  uint16_t* bands = (uint16_t*)((char*)scratch + P->lab_bytes + ibytes + gbytes);
  const bool colour = part & 1, texture = part & 2;
  // frames of channel ch are f = 4 ch + o; the texture frames 12..15 come last
  const size_t split = (size_t)g.lab_off[12];
  const int f0 = colour ? 0 : 12, nf = (colour ? 12 : 0) + (texture ? 4 : 0);
  const size_t fill0 = colour ? 0 : split, fill1 = texture ? P->lab_bytes : split;
  hipLaunchKernelGGL(k_fill_u8, dim3(div_up(fill1 - fill0, 256)), dim3(256), 0, s, frames + fill0, fill1 - fill0);
  hipLaunchKernelGGL(k_cue_labels, dim3(div_up(g.bw, 256), g.bh, nf), dim3(256), 0, s, g, f0, L, A, B, tex,
                     P->d_tabs, frames);
  int maxh = 0, maxw = 0;
  for (int o = 0; o < 4; o++) { maxh = std::max(maxh, g.nh[o]); maxw = std::max(maxw, g.nw[o]); }
  if (colour) {
    if (P->band_ok && integral_pass == 1) {
      DMSC_TRY((integral_bands<32, 25>(s, g, 0, 12, maxw, maxh, frames, bands, I)));
    } else {
      hipLaunchKernelGGL(k_integral_rows<32>, dim3(maxh, 12), dim3(256), (size_t)maxw, s, g, 0, frames, I);
      hipLaunchKernelGGL(k_integral_cols, dim3(div_up(maxw * 32 / 2, 256), 12), dim3(256), 0, s, g, 0, I);
    }
  }
  if (texture) {
    if (P->band_ok && integral_pass == 1) {
      if (tbp == 32) DMSC_TRY((integral_bands<32, 32>(s, g, 12, 4, maxw, maxh, frames, bands, I)));
      else DMSC_TRY((integral_bands<64, 64>(s, g, 12, 4, maxw, maxh, frames, bands, I)));
    } else {
      if (tbp == 32)
        hipLaunchKernelGGL(k_integral_rows<32>, dim3(maxh, 4), dim3(256), (size_t)maxw, s, g, 12, frames, I);
      else
        hipLaunchKernelGGL(k_integral_rows<64>, dim3(maxh, 4), dim3(256), (size_t)maxw, s, g, 12, frames, I);
      hipLaunchKernelGGL(k_integral_cols, dim3(div_up(maxw * tbp / 2, 256), 4), dim3(256), 0, s, g, 12, I);
    }
  }
  const size_t plane = (size_t)n;
  BlurK none;
  memset(&none, 0, sizeof(none));
  // Bg: channel 0, radii {3,5,10}; Cga/Cgb: channels 1-2, radii {5,10,20};
  // Tg: channel 3, radii {5,10,20}.
  if (colour) {
    hipLaunchKernelGGL((k_gradients<25, 6>), dim3(div_up(n, 256) * 3, 4), dim3(256), 0, s, g, 0, 0,
                       P->d_tabs, I, P->blur_bg, grads, plane);
    hipLaunchKernelGGL((k_gradients<25, 3>), dim3(div_up(n, 256) * 3, 8), dim3(256), 0, s, g, 1, 1,
                       P->d_tabs, I, P->blur_cg, grads, plane);
  }
  if (texture) {
    if (ntex == 64)
      hipLaunchKernelGGL(k_gradients64, dim3(div_up(2 * n, 256) * 3, 4), dim3(256), 0, s, g, 1, P->d_tabs, I,
                         grads, plane);
    else
      hipLaunchKernelGGL((k_gradients<32, 0>), dim3(div_up(n, 256) * 3, 4), dim3(256), 0, s, g, 3, 1,
                         P->d_tabs, I, none, grads, plane);
  }
  if (hist_done) DMSC_HIP(hipEventRecord(hist_done, s));
  CueOut out{{bg, cga, cgb, tg}};
  for (int scale = 0; scale < 3; scale++) {
    // (first channel, channels, radius index): Bg; Cga and Cgb; Tg
    const int groups[3][3] = {{0, 1, scale}, {1, 2, scale + 1}, {3, 1, scale + 1}};
    for (int gi = 0; gi < 3; gi++) {
      if ((gi < 2 && !colour) || (gi == 2 && !texture)) continue;
      const int ch0 = groups[gi][0], nch = groups[gi][1], ri = groups[gi][2];
      const dim3 grid(div_up(w, QTX), div_up(h, QTY), nch * 8);
      switch (kRadii[ri]) {
#define PARABOLA_CASE(R)                                                                                   \
  case R:                                                                                                  \
    hipLaunchKernelGGL(k_parabola<R>, grid, dim3(256), parabola_lds<R>(), s, w, h, scale, ch0, grads,       \
                       plane, P->d_coef[ri], P->d_vrange[ri], out, pitch);                                \
    break;
        PARABOLA_CASE(3) PARABOLA_CASE(5) PARABOLA_CASE(10) PARABOLA_CASE(20)
#undef PARABOLA_CASE
        default: return fail(DMSC_ERR_INTERNAL, "no parabola kernel for radius %d", kRadii[ri]);
      }
    }
  }
  DMSC_HIP(hipGetLastError());
  return DMSC_OK;
}

int launch_local_cues(Ctx* ctx, int w, int h, const float* L, const float* A, const float* B,
                      const int* tex, int ntex, float* bg, float* cga, float* cgb, float* tg, int pitch,
                      float* grads_out) {
  DMSC_TRY(local_cues_prepare(ctx, w, h, ntex));
  return launch_local_cues_part(ctx, ctx->stream, 3, w, h, L, A, B, tex, ntex, bg, cga, cgb, tg, pitch, grads_out);
}

}  // namespace dmsc
