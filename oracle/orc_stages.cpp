// orc_stages.cpp -- oracle restatement of the reference's per-pixel stage
// loops (colour conversion, textons, k-means, local cues, combine, nonmax,
// intervening contour, matrix transform, stencil SpMV, spectralPb, gPb).
// TEST INFRASTRUCTURE ONLY.  Loops over pixels are OpenMP-parallel; each
// output element is computed by one iteration in the reference's order, so
// the result does not depend on the thread count.
#include "oracle.h"
#include <omp.h>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
int g_simd = 0;  // orc_set_simd: register-tiled SIMD correlations (same results)
inline int mirror_index(int i, int border, int n) {  // convert.cu:128-141
  if (i < border) return border - i - 1;
  if (i < border + n) return i - border;
  return border + 2 * n - i - 1;
}
}  // namespace

extern "C" {

// OpenMP threads of the oracle's parallel loops (timing of the CPU baseline).
void orc_set_threads(int n) { omp_set_num_threads(n); }
void orc_set_simd(int on) { g_simd = on ? 1 : 0; }
int orc_max_threads(void) { return omp_get_max_threads(); }


// convert.cu:11-25 rgbUtoGreyF_kernel; R in bits 0-7, G 8-15, B 16-23.
void orc_rgb_to_grey(int w, int h, const uint32_t* rgb, float* grey) {
  const int n = w * h;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    uint32_t c = rgb[i];
    float r = (float)(c & 0xff) / 255.0f;
    float g = (float)((c & 0xff00) >> 8) / 255.0f;
    float b = (float)((c & 0xff0000) >> 16) / 255.0f;
    grey[i] = (0.29894f * r + 0.58704f * g) + 0.11402f * b;
  }
}

// convert.cu:27-64 rgbUtoLab3F_kernel (gamma by table, cube root by orc_cbrtf).
void orc_rgb_to_lab(int w, int h, float gamma, const uint32_t* rgb, float* L, float* A, float* B) {
  float lut[256];
  orc_gamma_lut(gamma, lut);
  const int n = w * h;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    uint32_t c = rgb[i];
    float r = lut[c & 0xff], g = lut[(c >> 8) & 0xff], b = lut[(c >> 16) & 0xff];
    float x = (0.412453f * r + 0.357580f * g) + 0.180423f * b;
    float y = (0.212671f * r + 0.715160f * g) + 0.072169f * b;
    float z = (0.019334f * r + 0.119193f * g) + 0.950227f * b;
    const float thr = 0.008856f;
    x = x / 0.950456f;
    y = y / 1.000000f;
    z = z / 1.088754f;
    float fx = (x > thr) ? orc_cbrtf(x) : (7.787f * x + (16.0f / 116.0f));
    float fy = (y > thr) ? orc_cbrtf(y) : (7.787f * y + (16.0f / 116.0f));
    float fz = (z > thr) ? orc_cbrtf(z) : (7.787f * z + (16.0f / 116.0f));
    L[i] = (y > thr) ? (116.0f * orc_cbrtf(y) - 16.0f) : (903.3f * y);
    A[i] = 500.0f * (fx - fy);
    B[i] = 200.0f * (fy - fz);
  }
}

// convert.cu:66-85 normalizeLab_kernel.
void orc_normalize_lab(int w, int h, float* L, float* A, float* B) {
  const int n = w * h;
  for (int i = 0; i < n; i++) {
    const float ab_min = -73.0f, ab_range = 95.0f - (-73.0f);
    float l = L[i] / 100.0f, a = (A[i] - ab_min) / ab_range, b = (B[i] - ab_min) / ab_range;
    l = l < 0 ? 0 : (l > 1 ? 1 : l);
    a = a < 0 ? 0 : (a > 1 ? 1 : a);
    b = b < 0 ? 0 : (b > 1 ? 1 : b);
    L[i] = l; A[i] = a; B[i] = b;
  }
}

// convert.cu:122-144 / 163-185 mirrorImage (symmetric reflection).
void orc_mirror_f(int w, int h, int border, const float* in, float* out) {
  int bw = w + 2 * border, bh = h + 2 * border;
  for (int y = 0; y < bh; y++)
    for (int x = 0; x < bw; x++)
      out[y * bw + x] = in[mirror_index(y, border, h) * w + mirror_index(x, border, w)];
}
void orc_mirror_i(int w, int h, int border, const int* in, int* out) {
  int bw = w + 2 * border, bh = h + 2 * border;
  for (int y = 0; y < bh; y++)
    for (int x = 0; x < bw; x++)
      out[y * bw + x] = in[mirror_index(y, border, h) * w + mirror_index(x, border, w)];
}

// convert.cu:234-246 quantizeImage_kernel.
void orc_quantize(int w, int h, int nbins, const float* in, int* out) {
  for (int i = 0; i < w * h; i++) {
    int q = (int)floorf(in[i] * (float)nbins);
    if (q == nbins) q = nbins - 1;
    out[i] = q;
  }
}

// texton.cu:39-57 convolve: 34 correlations, clamp-to-edge, x offset outer,
// y offset inner, fused multiply-add; filter 33 reads zero coefficients
// (texton.cu:104-109), so its response is exactly +0.
void orc_texton_convolve(int w, int h, const float* grey, float* resp) {
  static float coeffs[9010];
  static int radii[34];
  int nbound = orc_texton_filters(coeffs, radii);
  const int n = w * h;
  if (g_simd) {  // the same sums over a clamp-padded plane, register-tiled
    const int R = 9, pw = w + 2 * R, ph = h + 2 * R;
    std::vector<float> pad((size_t)pw * ph), f(19 * 19);
    for (int yy = 0; yy < ph; yy++)
      for (int xx = 0; xx < pw; xx++)
        pad[(size_t)yy * pw + xx] = grey[clampi(yy - R, 0, h - 1) * w + clampi(xx - R, 0, w - 1)];
    int ci = 0;
    for (int fi = 0; fi < 34; fi++) {
      const int r = radii[fi], L = 2 * r + 1;
      for (int i = 0; i < L; i++)  // coefficients run x offset outer, y offset inner
        for (int j = 0; j < L; j++, ci++) f[j * L + i] = (ci < nbound) ? coeffs[ci] : 0.0f;
      orc_correlate_simd(pad.data() + (size_t)(R - r) * pw + (R - r), pw, resp + (size_t)fi * n, w, w, h, L, L, 1,
                         f.data());
    }
    return;
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      int ci = 0;
      for (int f = 0; f < 34; f++) {
        int r = radii[f];
        float acc = 0.0f;
        for (int cx = x - r; cx <= x + r; cx++) {
          int xx = clampi(cx, 0, w - 1);
          for (int cy = y - r; cy <= y + r; cy++) {
            float c = (ci < nbound) ? coeffs[ci] : 0.0f;
            acc = fmaf(grey[clampi(cy, 0, h - 1) * w + xx], c, acc);
            ci++;
          }
        }
        resp[(size_t)f * n + y * w + x] = acc;
      }
    }
  }
}

// kmeans.cu:602-727 with assignInitialClusters (14-35), findCentroids
// (57-103, int32 wrap-around sums), finishCentroids (268-276), makeSelfDots
// (296-307), makeDots + SGEMM (318-338, 433: dist = (|x|^2+|c|^2) - 2 x.c)
// and findDiffLabels (447-472: strict '<' from 1e7, lowest index wins ties).
int orc_kmeans(int w, int h, int k, int nf, const float* resp, int* labels, int max_iter,
               int* iterations) {
  const int n = w * h;
  std::vector<int> iresp((size_t)nf * n);
  std::vector<float> pdot(n);
  // kmeans.cu:14-35 assignInitialClusters (6x6 grid, capped at 31) for
  // 32 clusters; kmeans.cu:37-55 assignInitialClusters_64 (8x8 grid) for 64
  const int g = (k == 64) ? 8 : 6;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      int c = (y / ((h - 1) / g + 1)) * g + x / ((w - 1) / g + 1);
      if (k != 64 && c >= 32) c = 31;
      labels[y * w + x] = c;
    }
  for (size_t i = 0; i < (size_t)nf * n; i++) iresp[i] = (int)(100000.0f * resp[i]);
  for (int p = 0; p < n; p++) {
    float d = 0;
    for (int f = 0; f < nf; f++) d = fmaf(resp[(size_t)f * n + p], resp[(size_t)f * n + p], d);
    pdot[p] = d;
  }
  std::vector<uint32_t> mass((size_t)nf * k);
  std::vector<uint32_t> count(k);
  std::vector<float> cent((size_t)nf * k), cdot(k);
  std::vector<int> nl(n);
  int it = 0, changes = 0;
  for (it = 0; it < max_iter; it++) {
    std::fill(mass.begin(), mass.end(), 0u);
    std::fill(count.begin(), count.end(), 0u);
    for (int p = 0; p < n; p++) {
      int c = labels[p];
      for (int f = 0; f < nf; f++) mass[(size_t)f * k + c] += (uint32_t)iresp[(size_t)f * n + p];
      count[c]++;
    }
    for (int f = 0; f < nf; f++)
      for (int c = 0; c < k; c++)
        cent[(size_t)f * k + c] = (float)(int32_t)mass[(size_t)f * k + c] / (100000.0f * (float)count[c]);
    for (int c = 0; c < k; c++) {
      float d = 0;
      for (int f = 0; f < nf; f++) d = fmaf(cent[(size_t)f * k + c], cent[(size_t)f * k + c], d);
      cdot[c] = d;
    }
    changes = 0;
#pragma omp parallel for schedule(static) reduction(+ : changes)
    for (int p = 0; p < n; p++) {
      float best = 10000000.0f;
      int bl = -1;
      for (int c = 0; c < k; c++) {
        float acc = 0.0f;
        for (int f = 0; f < nf; f++) acc = fmaf(resp[(size_t)f * n + p], cent[(size_t)f * k + c], acc);
        float d = fmaf(-2.0f, acc, pdot[p] + cdot[c]);
        if (d < best) { best = d; bl = c; }
      }
      nl[p] = bl;
      if (bl != labels[p]) changes++;
    }
    memcpy(labels, nl.data(), sizeof(int) * n);
    if (changes <= 0) { it++; break; }
  }
  if (iterations) *iterations = it;
  return changes;
}

// texton.cu:62-131 findTextons (ntex = 32 or 64 textons, texton.cu:65-67).
int orc_find_textons(int w, int h, const float* grey, int ntex, int* labels, int* iterations) {
  const int n = w * h;
  std::vector<float> resp((size_t)34 * n);
  orc_texton_convolve(w, h, grey, resp.data());
  return orc_kmeans(w, h, ntex, 34, resp.data(), labels, 15, iterations);
}

// gradient.cu:66-118 formIntegralImages (inclusive 2-D counts per bin);
// pixels labelled -1 (outside the rotated image) match no bin.
void orc_integral_images(int W, int H, int nbins, const int* lab, int bp, int32_t* I) {
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nbins; b++) {
    for (int y = 0; y < H; y++) {
      int32_t rowacc = 0;
      for (int x = 0; x < W; x++) {
        rowacc += (lab[y * W + x] == b) ? 1 : 0;
        int32_t above = (y > 0) ? I[((size_t)(y - 1) * W + x) * bp + b] : 0;
        I[((size_t)y * W + x) * bp + b] = rowacc + above;
      }
    }
  }
}

// 32-slot tree reduction of the reference's warp-synchronous sum
// (rotate.cu:378-389): pairs (i, i+16), (i, i+8), ... (0, 1).
static inline float tree32(float* t) {
  for (int s = 16; s >= 1; s >>= 1)
    for (int i = 0; i < s; i++) t[i] = t[i] + t[i + s];
  return t[0];
}

// Bin sum of a histogram padded to 32 or 64 slots: one 32-slot tree per
// 32-thread warp, then slot 0 + slot 32 when there are more than 32 bins
// (rotate.cu:384-389).
static inline float tree_bins(float* t, int nbins) {
  const float lo = tree32(t);
  return nbins > 32 ? lo + tree32(t + 32) : lo;
}

// rotate.cu:263-475 computeGradient for both senses at every pixel.
void orc_gradient_pair(int w, int h, int border, int nbins, int blur, const float* kern,
                       int kr, int rw, const int* bX, const int* bY, const int* gX,
                       const int* gY, const int* rect, const int32_t* I, int bp, float* gA,
                       float* gB) {
  const int kl = 2 * kr + 1;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      int ix = x + border, iy = y + border;
      int rx = gX[iy] + bX[ix], ry = gY[iy] + bY[ix];
      long pix = (long)ry * rw + rx;
      for (int sense = 0; sense < 2; sense++) {
        float a[64], b[64];
        for (int i = 0; i < 64; i++) { a[i] = 0.f; b[i] = 0.f; }
        for (int bin = 0; bin < nbins; bin++) {
          int32_t c[8];
          for (int kk = 0; kk < 8; kk++) c[kk] = I[(size_t)(rect[kk] + pix) * bp + bin];
          float av, bv;
          if (sense == 0) {  // top / bottom halves (A-C+E-D, H-F+D-E)
            av = (float)(c[0] - c[2]); av += (float)(c[4] - c[3]);
            bv = (float)(c[7] - c[5]); bv += (float)(c[3] - c[4]);
          } else {           // left / right halves (A-F+G-B, H-C+B-G)
            av = (float)(c[0] - c[5]); av += (float)(c[6] - c[1]);
            bv = (float)(c[7] - c[2]); bv += (float)(c[1] - c[6]);
          }
          a[bin] = av;
          b[bin] = bv;
        }
        if (blur) {  // rotate.cu:352-370: zero-padded correlation over bins
          float ca[64], cb[64];
          for (int bin = 0; bin < nbins; bin++) {
            float sa = 0.f, sb = 0.f;
            for (int nk = 0; nk < kl; nk++) {
              int j = bin - kr + nk;
              float ha = (j >= 0 && j < nbins) ? a[j] : 0.f;
              float hb = (j >= 0 && j < nbins) ? b[j] : 0.f;
              sa = fmaf(ha, kern[nk], sa);
              sb = fmaf(hb, kern[nk], sb);
            }
            ca[bin] = sa;
            cb[bin] = sb;
          }
          for (int bin = 0; bin < nbins; bin++) { a[bin] = ca[bin]; b[bin] = cb[bin]; }
        }
        const int slots = nbins > 32 ? 64 : 32;
        float t[64];
        memcpy(t, a, sizeof(t));
        float sa = tree_bins(t, nbins);
        float na = (sa != 0.f) ? 8.0f / sa : 0.f;
        memcpy(t, b, sizeof(t));
        float sb = tree_bins(t, nbins);
        float nb = (sb != 0.f) ? 8.0f / sb : 0.f;
        for (int i = 0; i < slots; i++) {
          float ai = a[i] * na, bi = b[i] * nb;
          float s = ai + bi, chi = 0.f;
          if (s != 0.f) { float d = ai - bi; chi = d * d / s; }
          t[i] = chi;
        }
        float out = 0.5f * tree_bins(t, nbins);
        (sense == 0 ? gA : gB)[y * w + x] = out;
      }
    }
  }
}

// gradient.cu:210-281 gradients(): quantised/ textons labels -> mirror
// (border 30) -> 4 rotations -> integral images -> 3 scales x 2 senses.
// Output planes [scale][slot][N]; slot o+4 <- sense A, slot o <- sense B.
void orc_channel_gradients(int w, int h, int nbins, int blur, float sigma, const int* radii3,
                           const int* labels, float* grads) {
  const int border = 30, bw = w + 2 * border, bh = h + 2 * border, n = w * h;
  std::vector<int> mir((size_t)bw * bh);
  orc_mirror_i(w, h, border, labels, mir.data());
  float kern[64];
  int kr = 0;
  if (blur) orc_blur_kernel(sigma, kern, &kr);
  std::vector<int> bX(bw), bY(bw), gX(bh), gY(bh);
  std::vector<float> gA(n), gB(n);
  for (int o = 0; o < 4; o++) {
    float thetaPi = -float(o) / float(8);
    int nw, nh;
    orc_rotation_tables(bw, bh, thetaPi, &nw, &nh, bX.data(), bY.data(), gX.data(), gY.data());
    std::vector<int> rot;
    const int* img = mir.data();
    if (o != 0) {
      rot.assign((size_t)nw * nh, -1);
      for (int yy = 0; yy < bh; yy++)
        for (int xx = 0; xx < bw; xx++) {
          int X = gX[yy] + bX[xx], Y = gY[yy] + bY[xx];
          if (X < 0 || X >= nw || Y < 0 || Y >= nh) abort();
          int& dst = rot[(size_t)Y * nw + X];
          if (dst != -1) abort();  // the Bresenham rotation is injective
          dst = mir[(size_t)yy * bw + xx];
        }
      img = rot.data();
    }
    std::vector<int32_t> I((size_t)nw * nh * nbins);
    orc_integral_images(nw, nh, nbins, img, nbins, I.data());
    for (int s = 0; s < 3; s++) {
      int rect[8];
      orc_halfdisc_offsets(w, h, thetaPi, radii3[s], bX.data(), bY.data(), gX.data(), gY.data(),
                           nw, rect);
      orc_gradient_pair(w, h, border, nbins, blur, kern, kr, nw, bX.data(), bY.data(), gX.data(),
                        gY.data(), rect, I.data(), nbins, gA.data(), gB.data());
      memcpy(grads + (size_t)(s * 8 + o + 4) * n, gA.data(), sizeof(float) * n);
      memcpy(grads + (size_t)(s * 8 + o) * n, gB.data(), sizeof(float) * n);
    }
  }
}

// parabola.cu:75-124 + parabola_kernel.cu:1-30: correlation of the mirrored
// gradient plane with the orientation's Savitzky-Golay filter, v outer, u
// inner, fused multiply-add (zero taps skipped: they leave the sum
// unchanged bit for bit).
void orc_parabola(int w, int h, int border, int r, const float* g8, const float* filt,
                  float* out8) {
  const int len = 2 * r + 1, n = w * h;
  if (g_simd) {  // the same sums over the mirrored plane, register-tiled
    const int pw = w + 2 * r, ph = h + 2 * r;
    std::vector<float> pad((size_t)pw * ph);
    for (int o = 0; o < 8; o++) {
      const float* g = g8 + (size_t)o * n;
#pragma omp parallel for schedule(static)
      for (int yy = 0; yy < ph; yy++) {
        const int sy = mirror_index(yy - r + border, border, h);
        for (int xx = 0; xx < pw; xx++) pad[(size_t)yy * pw + xx] = g[sy * w + mirror_index(xx - r + border, border, w)];
      }
      orc_correlate_simd(pad.data(), pw, out8 + (size_t)o * n, w, w, h, len, len, 0, filt + (size_t)o * len * len);
    }
    return;
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int o = 0; o < 8; o++)
    for (int y = 0; y < h; y++) {
      const float* g = g8 + (size_t)o * n;
      const float* f = filt + (size_t)o * len * len;
      for (int x = 0; x < w; x++) {
        float acc = 0.0f;
        for (int v = -r; v <= r; v++) {
          int yy = mirror_index(y + border + v, border, h);
          for (int u = -r; u <= r; u++) {
            float c = f[(v + r) * len + (u + r)];
            if (c == 0.0f) continue;
            acc = fmaf(g[yy * w + mirror_index(x + border + u, border, w)], c, acc);
          }
        }
        out8[(size_t)o * n + y * w + x] = acc;
      }
    }
}

// localcues.cu:173-300 localCues: Bg (L, 25 bins, blur sigma 0.1, radii
// 3,5,10), Cga/Cgb (25 bins, blur 0.05, radii 5,10,20), Tg (ntex textons, no
// blur, radii 5,10,20); parabola smoothing per scale.  Outputs [24][N].
void orc_local_cues(int w, int h, const float* L, const float* A, const float* B,
                    const int* textons, int ntex, float* bg, float* cga, float* cgb, float* tg) {
  const int n = w * h;
  static const int radii[4] = {3, 5, 10, 20};
  std::vector<float> grads((size_t)24 * n);
  std::vector<int> q(n);
  const float* chans[3] = {L, A, B};
  float* outs[4] = {bg, cga, cgb, tg};
  for (int ch = 0; ch < 4; ch++) {
    const int* rad = (ch == 0) ? &radii[0] : &radii[1];
    if (ch < 3) {
      int nbins = 25;
      float sigma = (ch == 0) ? 0.1f : 0.05f;
      orc_quantize(w, h, nbins, chans[ch], q.data());
      orc_channel_gradients(w, h, nbins, 1, (float)(int)(sigma * (float)nbins), rad, q.data(),
                            grads.data());
    } else {
      orc_channel_gradients(w, h, ntex, 0, 0.0f, rad, textons, grads.data());
    }
    for (int s = 0; s < 3; s++) {
      int r = rad[s], len = 2 * r + 1;
      std::vector<float> f((size_t)8 * len * len);
      orc_parabola_filters(r, f.data());
      orc_parabola(w, h, 30, r, grads.data() + (size_t)s * 8 * n, f.data(),
                   outs[ch] + (size_t)s * 8 * n);
    }
  }
}

// combine.cu:11-93, fused multiply-add in cue order; the 32-texton tables
// (combine.cu:17, 21) or the 64-texton ones (combine.cu:16, 20), which
// differ in the three Tg entries.
void orc_combine(int w, int h, const float* bg, const float* cga, const float* cgb,
                 const float* tg, int ntex, float* mpb, float* cgout) {
  static const float coef32[12] = {0.0123f, 0.0110f, 0.0117f, 0.0169f, 0.0176f, 0.0198f,
                                   0.0138f, 0.0138f, 0.0145f, 0.0144f, 0.0149f, 0.0158f};
  static const float wts32[12] = {0.f,     0.f,     0.0028f, 0.0041f, 0.0042f, 0.0047f,
                                  0.0033f, 0.0033f, 0.0035f, 0.0035f, 0.0035f, 0.0188f};
  static const float coef64[12] = {0.0123f, 0.0110f, 0.0117f, 0.0169f, 0.0176f, 0.0198f,
                                   0.0138f, 0.0138f, 0.0145f, 0.0104f, 0.0105f, 0.0115f};
  static const float wts64[12] = {0.f,     0.f,     0.0028f, 0.0041f, 0.0042f, 0.0047f,
                                  0.0033f, 0.0033f, 0.0035f, 0.0025f, 0.0025f, 0.0137f};
  const float* coef = ntex == 64 ? coef64 : coef32;
  const float* wts = ntex == 64 ? wts64 : wts32;
  const int n = w * h;
  const float* cues[4] = {bg, cga, cgb, tg};
  for (int o = 0; o < 8; o++)
    for (int p = 0; p < n; p++) {
      float a1 = 0.f, a2 = 0.f;
      for (int k = 0; k < 12; k++) {
        float v = cues[k / 3][(size_t)((k % 3) * 8 + o) * n + p];
        a1 = fmaf(v, coef[k], a1);
        a2 = fmaf(v, wts[k], a2);
      }
      mpb[(size_t)o * n + p] = a1;
      cgout[(size_t)o * n + p] = a2;
    }
}

// nonmax.cu:23-94 devFindMax + nonmax.cu:147-246 devOriented_2D
// (GPU variant: output min(1.2 v, 1) at local maxima, 0 elsewhere/borders).
void orc_nonmax(int w, int h, const float* pb, float* nmax) {
  const int n = w * h;
  std::vector<float> mx(n);
  std::vector<int> ori(n);
  for (int p = 0; p < n; p++) {
    float m = -200.0f;
    int mo = 0;
    for (int o = 0; o < 8; o++)
      if (pb[(size_t)o * n + p] > m) { m = pb[(size_t)o * n + p]; mo = o; }
    mx[p] = (m > 0) ? m : 0.0f;
    ori[p] = mo;
  }
  static const float tand[8] = {0.f, 0.414214f, 1.f, 0.414214f, 0.f, 0.414214f, 1.f, 0.414214f};
  static const float cosori[8] = {1.f, 1.f, 0.923879f, 0.707107f, 0.382684f, 0.707107f, 0.923879f, 1.f};
  const int sx = h, sy = w;
  for (int p = 0; p < n; p++) {
    int x = p / w, y = p - w * x;
    nmax[p] = 0.0f;
    int o = ori[p], th = (o + 4) % 8;
    float v = mx[p];
    int i0a = 0, i0b = 0, i1a = 0, i1b = 0;
    bool v0 = false, v1 = false;
    if (th == 0) {
      v0 = (x > 0); v1 = (x < sx - 1);
      if (v0) { i0a = p - sy; i0b = i0a; }
      if (v1) { i1a = p + sy; i1b = i1a; }
    } else if (th == 1) {
      v0 = (x > 0) && (y > 0); v1 = (x < sx - 1) && (y < sy - 1);
      if (v0) { i0a = p - sy; i0b = i0a - 1; }
      if (v1) { i1a = p + sy; i1b = i1a + 1; }
    } else if (th <= 3) {
      v0 = (x > 0) && (y > 0); v1 = (x < sx - 1) && (y < sy - 1);
      if (v0) { i0a = p - 1; i0b = i0a - sy; }
      if (v1) { i1a = p + 1; i1b = i1a + sy; }
    } else if (th <= 5) {
      v0 = (x < sx - 1) && (y > 0); v1 = (x > 0) && (y < sy - 1);
      if (v0) { i0a = p - 1; i0b = i0a + sy; }
      if (v1) { i1a = p + 1; i1b = i1a - sy; }
    } else {
      v0 = (x < sx - 1) && (y > 0); v1 = (x > 0) && (y < sy - 1);
      if (v0) { i0a = p + sy; i0b = i0a - 1; }
      if (v1) { i1a = p - sy; i1b = i1a + 1; }
    }
    if (v0 && v1) {
      float d = tand[th];
      int d0a = std::abs(ori[i0a] - o), d0b = std::abs(ori[i0b] - o);
      int d1a = std::abs(ori[i1a] - o), d1b = std::abs(ori[i1b] - o);
      float va = ((1.0f - d) * mx[i0a]) * cosori[d0a] + (d * mx[i0b]) * cosori[d0b];
      float vb = ((1.0f - d) * mx[i1a]) * cosori[d1a] + (d * mx[i1b]) * cosori[d1b];
      if ((v > va) && (v > vb)) {
        float t = 1.2f * v;
        nmax[p] = (t >= 1.0f) ? 1.0f : t;
      }
    }
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Intervening contour (intervening.cu:28-566)
namespace {
struct IcCtx {
  int w, h, radius, diameter;
  float rsigma;
  const float* mpb;
  const int* diag;
  float* M;
  size_t pitch;
  float pb(int x, int y) const { return mpb[clampi(y, 0, h - 1) * w + clampi(x, 0, w - 1)]; }
};

inline float c_ic_ss(float ic) {  // intervening.cu:32-42
  float val = 1.8682f;
  val += (ic / 0.3130f) * -1.3113f;
  const float post = 1.0f / (1.0f + orc_expf(-val));
  if (!std::isfinite(post)) return 0;
  if (post < 0) return 0;
  if (post > 1) return 1;
  return post;
}
inline float affinity(float c, float rs) { return orc_expf(-(1 - c) * rs); }  // 28-30

void ic_walk(const IcCtx& C, int x0, int y0, int x1, int y1, int x2, int y2, int x3, int y3) {
  const int dx1 = x1 - x0, dy1 = y1 - y0, dx2 = x2 - x0, dy2 = y2 - y0, dx3 = x3 - x0, dy3 = y3 - y0;
  const int dot11 = dx1 * dx1 + dy1 * dy1, dot22 = dx2 * dx2 + dy2 * dy2, dot33 = dx3 * dx3 + dy3 * dy3;
  const int dx = x2 - x0, dy = y2 - y0, adx = std::abs(dx), ady = std::abs(dy);
  int octant = -1;
  if (dx > 0 && dy >= 0) octant = (adx > ady) ? 0 : 1;
  else if (dx <= 0 && dy > 0) octant = (adx < ady) ? 2 : 3;
  else if (dy <= 0 && dx < 0) octant = (adx > ady) ? 4 : 5;
  else if (dx >= 0 && dy < 0) octant = (adx < ady) ? 6 : 7;
  int t = 0;
  switch (octant) {
    case 0: case 3: case 4: case 7: t = -adx; break;
    default: t = -ady; break;
  }
  float maxpb = 0.0f;
  int xi = x0, yi = y0, oldx = xi, oldy = yi;
  while (xi != x2 || yi != y2) {
    switch (octant) {
      case 0: xi++; t += (ady << 1); if (t > 0) { yi++; t -= (adx << 1); } break;
      case 1: yi++; t += (adx << 1); if (t > 0) { xi++; t -= (ady << 1); } break;
      case 2: yi++; t += (adx << 1); if (t > 0) { xi--; t -= (ady << 1); } break;
      case 3: xi--; t += (ady << 1); if (t > 0) { yi++; t -= (adx << 1); } break;
      case 4: xi--; t += (ady << 1); if (t > 0) { yi--; t -= (adx << 1); } break;
      case 5: yi--; t += (adx << 1); if (t > 0) { xi--; t -= (ady << 1); } break;
      case 6: yi--; t += (adx << 1); if (t > 0) { xi++; t -= (ady << 1); } break;
      case 7: xi++; t += (ady << 1); if (t > 0) { yi--; t -= (adx << 1); } break;
    }
    const int dxi = xi - x0, dyi = yi - y0;
    const int dotii = dxi * dxi + dyi * dyi, doti1 = dxi * dx1 + dyi * dy1;
    const int doti2 = dxi * dx2 + dyi * dy2, doti3 = dxi * dx3 + dyi * dy3;
    const bool good = (doti2 * doti2 == dotii * dot22) ||
                      (dot11 * doti2 * doti2 > dot22 * doti1 * doti1 &&
                       dot33 * doti2 * doti2 >= dot22 * doti3 * doti3);
    float is = 0.0f;
    if (oldx == xi) {
      if (yi > oldy) is = C.pb(xi, yi - 1);
      else if (yi < oldy) is = C.pb(oldx, oldy - 1);
    } else if (oldy == yi) {
      if (xi > oldx) is = C.pb(xi - 1, yi);
      else if (xi < oldx) is = C.pb(oldx - 1, oldy);
    } else {
      if ((xi > oldx) && (yi > oldy)) {
        is = fmaxf(C.pb(oldx, yi - 1), is); is = fmaxf(C.pb(xi, yi - 1), is);
        is = fmaxf(C.pb(xi - 1, oldy), is); is = fmaxf(C.pb(xi - 1, yi), is);
      } else if ((xi > oldx) && (yi < oldy)) {
        is = fmaxf(C.pb(oldx, oldy - 1), is); is = fmaxf(C.pb(xi, oldy - 1), is);
        is = fmaxf(C.pb(xi - 1, oldy), is); is = fmaxf(C.pb(xi - 1, yi), is);
      } else if ((xi < oldx) && (yi > oldy)) {
        is = fmaxf(C.pb(oldx, yi - 1), is); is = fmaxf(C.pb(xi, yi - 1), is);
        is = fmaxf(C.pb(oldx - 1, oldy), is); is = fmaxf(C.pb(oldx - 1, yi), is);
      } else if ((xi < oldx) && (yi < oldy)) {
        is = fmaxf(C.pb(oldx, oldy - 1), is); is = fmaxf(C.pb(xi, oldy - 1), is);
        is = fmaxf(C.pb(oldx - 1, oldy), is); is = fmaxf(C.pb(oldx - 1, yi), is);
      }
    }
    maxpb = fmaxf(maxpb, is);
    oldx = xi;
    oldy = yi;
    if (!good) continue;
    float val = affinity(c_ic_ss(maxpb), C.rsigma);
    int xo = xi - x0, yo = yi - y0;
    if (xo * xo + yo * yo <= C.radius * C.radius) {
      int d = C.diag[(yo + C.radius) * C.diameter + xo + C.radius];
      C.M[(size_t)d * C.pitch + (size_t)y0 * C.w + x0] = val;
    }
  }
}

void fix_entry(const IcCtx& C, int x0, int y0, int x, int y) {  // 285-296
  int d = C.diag[(y - y0 + C.radius) * C.diameter + (x - x0) + C.radius];
  C.M[(size_t)d * C.pitch + (size_t)y0 * C.w + x0] = affinity(c_ic_ss(0.0f), C.rsigma);
}

void find_affinities(const IcCtx& C, int x0, int y0) {  // 298-475
  const int r = C.radius, w = C.w, h = C.h;
  C.M[(size_t)C.diag[r * C.diameter + r] * C.pitch + (size_t)y0 * w + x0] = 1.0f;
  const int rxa = std::max(0, x0 - r), rya = std::max(0, y0 - r);
  const int rxb = std::min(x0 + r, w - 1), ryb = std::min(y0 + r, h - 1);
  if (x0 > rxa) {
    if ((y0 > rya) && (y0 < ryb)) ic_walk(C, x0, y0, rxa, y0 - 1, rxa, y0, rxa, y0 + 1);
    for (int y = y0 - 1; y > rya; y--) ic_walk(C, x0, y0, rxa, y - 1, rxa, y, rxa, y + 1);
  }
  if (x0 > rxa + 1 || y0 > rya + 1 || ((x0 > rxa) && (y0 > rya)))
    ic_walk(C, x0, y0, rxa, rya + 1, rxa, rya, rxa + 1, rya);
  if (y0 > rya)
    for (int x = rxa + 1; x < rxb; x++) ic_walk(C, x0, y0, x - 1, rya, x, rya, x + 1, rya);
  if ((y0 > rya + 1) || (x0 < rxb - 1) || ((y0 > rya) && (x0 < rxb)))
    ic_walk(C, x0, y0, rxb - 1, rya, rxb, rya, rxb, rya + 1);
  if (x0 < rxb)
    for (int y = rya + 1; y < y0; y++) ic_walk(C, x0, y0, rxb, y - 1, rxb, y, rxb, y + 1);
  if (x0 > rxa)
    for (int y = y0 + 1; y < ryb; y++) ic_walk(C, x0, y0, rxa, y - 1, rxa, y, rxa, y + 1);
  if ((x0 > rxa + 1) || (y0 < ryb - 1) || ((x0 > rxa) && (y0 < ryb)))
    ic_walk(C, x0, y0, rxa, ryb - 1, rxa, ryb, rxa + 1, ryb);
  if (y0 < ryb)
    for (int x = rxa + 1; x < rxb; x++) ic_walk(C, x0, y0, x - 1, ryb, x, ryb, x + 1, ryb);
  if ((y0 < ryb - 1) || (x0 < rxb - 1) || ((y0 < ryb) && (x0 < rxb)))
    ic_walk(C, x0, y0, rxb - 1, ryb, rxb, ryb, rxb, ryb - 1);
  if (x0 < rxb) {
    for (int y = ryb - 1; y > y0; y--) ic_walk(C, x0, y0, rxb, y - 1, rxb, y, rxb, y + 1);
    if ((y0 > 0) && (y0 < ryb)) ic_walk(C, x0, y0, rxb, y0 - 1, rxb, y0, rxb, y0 + 1);
  }
  if ((y0 == rya + 1) && ((x0 == rxa) || (x0 == rxb))) fix_entry(C, x0, y0, x0, y0 - 1);
  if ((x0 == rxa + 1) && ((y0 == rya) || (y0 == ryb))) fix_entry(C, x0, y0, x0 - 1, y0);
  if ((x0 == rxb - 1) && ((y0 == rya) || (y0 == ryb))) fix_entry(C, x0, y0, x0 + 1, y0);
  if ((y0 == ryb - 1) && ((x0 == rxa) || (x0 == rxb))) fix_entry(C, x0, y0, x0, y0 + 1);
}
}  // namespace

extern "C" {

// intervening.cu:532-566 intervene: affinities (findAffinities) then
// symmetrisation 0.5(a+b) over forward offsets (symmetrizeMatrix 477-511).
void orc_intervene(int w, int h, const float* mpb, float* M, float sigma) {
  const int r = 5, diameter = 11, n = w * h;
  int offs[121], diag[121];
  orc_stencil(r, w, offs, diag);
  memset(M, 0, sizeof(float) * 81 * (size_t)n);
  IcCtx C{w, h, r, diameter, 1.0f / sigma, mpb, diag, M, (size_t)n};
#pragma omp parallel for schedule(dynamic, 2)
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) find_affinities(C, x, y);
  // Each unordered pair is owned by the pixel with the forward offset, so
  // the pairs are disjoint and the order of pixels does not matter.
  for (int y0 = 0; y0 < h; y0++)
    for (int x0 = 0; x0 < w; x0++)
      for (int yo = 0; yo <= r; yo++)
        for (int xo = -r; xo <= r; xo++) {
          if (xo * xo + yo * yo > r * r) continue;
          if (!((yo > 0) || (xo > 0))) continue;
          int xx = x0 + xo, yy = y0 + yo;
          if (xx < 0 || yy < 0 || xx >= w || yy >= h) continue;
          float* p1 = &M[(size_t)diag[(yo + r) * diameter + xo + r] * n + (size_t)y0 * w + x0];
          float* p2 = &M[(size_t)diag[(r - yo) * diameter + r - xo] * n + (size_t)yy * w + xx];
          float s = 0.5f * (*p1 + *p2);
          *p1 = s;
          *p2 = s;
        }
}

// stencilMVM.cu:46-125 stencilSumRows + unGeneralizeMatrix: D = row sums,
// L_sym = I - D^-1/2 W D^-1/2 in place; returns D and D^-1/2 (1/sqrt).
void orc_convert_matrix(int w, int h, float* M, float* dsum, float* rs) {
  const int n = w * h;
  int offs[121], diag[121];
  int nd = orc_stencil(5, w, offs, diag);
  for (int row = 0; row < n; row++) {
    float acc = 0;
    int col = row;
    for (int d = 0; d < nd; d++) {
      col += offs[d];
      if (col >= 0 && col < n) acc = acc + M[(size_t)d * n + row];
    }
    dsum[row] = acc;
    rs[row] = 1.0f / sqrtf(acc);
  }
  for (int row = 0; row < n; row++) {
    int col = row;
    float dii = dsum[row], diim5 = rs[row];
    for (int d = 0; d < nd; d++) {
      col += offs[d];
      if (col >= 0 && col < n) {
        float e = M[(size_t)d * n + row];
        if (col == row) M[(size_t)d * n + row] = (dii - e) / dii;
        else M[(size_t)d * n + row] = (-e * diim5) * rs[col];
      }
    }
  }
}

// stencilMVM.cu:127-180 stencilMVM, sequential over the 81 diagonals, fma.
void orc_spmv(int w, int h, const float* M, const float* x, float* y) {
  const int n = w * h;
  int offs[121], diag[121];
  int nd = orc_stencil(5, w, offs, diag);
  int r0 = 0, r1 = 0;
  if (g_simd) {  // rows whose every column is in range, vectorised in the same order
    int lin[121], c = 0, lo = 0, hi = 0;
    for (int d = 0; d < nd; d++) {
      c += offs[d];
      lin[d] = c;
      lo = std::min(lo, c);
      hi = std::max(hi, c);
    }
    r0 = -lo;
    r1 = std::max(r0, n - hi);
    orc_spmv_interior_simd(M, n, x, y, r0, r1, lin, nd);
  }
#pragma omp parallel for schedule(static)
  for (int row = 0; row < n; row++) {
    if (row >= r0 && row < r1) continue;
    float acc = 0;
    int col = row;
    for (int d = 0; d < nd; d++) {
      col += offs[d];
      if (col >= 0 && col < n) acc = fmaf(M[(size_t)d * n + row], x[col], acc);
    }
    y[row] = acc;
  }
}

void orc_spmv_d(int w, int h, const float* M, const double* x, double* y) {
  const int n = w * h;
  int offs[121], diag[121];
  int nd = orc_stencil(5, w, offs, diag);
#pragma omp parallel for schedule(static)
  for (int row = 0; row < n; row++) {
    double acc = 0;
    int col = row;
    for (int d = 0; d < nd; d++) {
      col += offs[d];
      if (col >= 0 && col < n) acc += (double)M[(size_t)d * n + row] * x[col];
    }
    y[row] = acc;
  }
}

// sPb/spectralPb.cu:673-795: mirror (border 3), 8 oriented first-derivative
// filters (u outer, v inner, index (3+u)+(3+v)*7), accumulate |.|/sqrt(lambda)
// over eigenvectors 1..nvec-1.
void orc_spectral_pb(int w, int h, int nvec, const float* eigvals, const float* eigvecs,
                     float* spb) {
  const int n = w * h, bw = w + 6, bh = h + 6;
  float K[8 * 49];
  orc_spb_filters(K);
  memset(spb, 0, sizeof(float) * 8 * (size_t)n);
  std::vector<float> m((size_t)bw * bh);
  std::vector<float> so(g_simd ? (size_t)8 * n : 0), ko(49);
  for (int i = 1; i < nvec; i++) {
    orc_mirror_f(w, h, 3, eigvecs + (size_t)i * n, m.data());
    float scale = 1.0f / sqrtf(eigvals[i]);
    if (g_simd) {  // per orientation, the same sums (u outer, v inner), register-tiled
      for (int o = 0; o < 8; o++) {
        for (int t = 0; t < 49; t++) ko[t] = K[o * 49 + t];  // [v + 3][u + 3]
        orc_correlate_simd(m.data(), bw, so.data() + (size_t)o * n, w, w, h, 7, 7, 1, ko.data());
      }
#pragma omp parallel for schedule(static)
      for (int p = 0; p < n; p++)
        for (int o = 0; o < 8; o++) {
          float& r = spb[(size_t)o * n + p];
          r = fmaf(fabsf(so[(size_t)o * n + p]), scale, r);
        }
      continue;
    }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int u = -3; u <= 3; u++)
          for (int v = -3; v <= 3; v++) {
            float val = m[(size_t)(y + 3 + v) * bw + (x + 3 + u)];
            for (int o = 0; o < 8; o++) s[o] = fmaf(val, K[o * 49 + (3 + u) + (3 + v) * 7], s[o]);
          }
        for (int o = 0; o < 8; o++) {
          float& r = spb[(size_t)o * n + y * w + x];
          r = fmaf(fabsf(s[o]), scale, r);
        }
      }
  }
}

// gPb/globalPb.cu:5-49 CalcGPb: gPbAll_o = combinedG_o + 0.0139 sPb_o,
// gPb = max_o (strict '>', from -200), zeroed where nonmax-mPb <= 0.05.
void orc_gpb(int w, int h, const float* cg, const float* spb, const float* nmax, float* all,
             float* gpb) {
  const int n = w * h;
  for (int p = 0; p < n; p++) {
    float m = -200.0f;
    for (int o = 0; o < 8; o++) {
      float v = fmaf(0.0139f, spb[(size_t)o * n + p], cg[(size_t)o * n + p]);
      all[(size_t)o * n + p] = v;
      if (v > m) m = v;
    }
    gpb[p] = (nmax[p] > 0.05f) ? m : 0.0f;
  }
}

}  // extern "C"

// ---------------------------------------------------------------- postprocess
// postprocess/skeleton.cu:104-129: neighbourhood codes marked by each table.
static const int kSkelLut[8][8] = {
    {89, 91, 217, 219, -1, -1, -1, -1},   {152, 153, 216, 217, 408, 409, 472, 473},
    {23, 31, 55, 63, -1, -1, -1, -1},     {26, 27, 30, 31, 90, 91, 94, 95},
    {464, 472, 496, 504, -1, -1, -1, -1}, {50, 51, 54, 55, 306, 307, 310, 311},
    {308, 310, 436, 438, -1, -1, -1, -1}, {176, 180, 240, 244, 432, 436, 496, 500}};

// skeleton.cu:8-54 applyLut (reads a snapshot, clears where the table fires)
// with skeleton.cu:189-212 Nhood3's code: weights[r][c] = 2^(r + 3c) for
// the neighbour at (row + r - 1, col + c - 1).
extern "C" int orc_skeleton_pass(int w, int h, int lut, const float* in, float* out) {
  static const int weights[3][3] = {{1, 8, 64}, {2, 16, 128}, {4, 32, 256}};
  bool table[512] = {false};
  for (int k = 0; k < 8; k++)
    if (kSkelLut[lut][k] >= 0) table[kSkelLut[lut][k]] = true;
  int removed = 0;
  for (int row = 0; row < h; row++)
    for (int col = 0; col < w; col++) {
      const int minRow = row == 0 ? 1 : 0, maxRow = row == h - 1 ? 1 : 2;
      const int minCol = col == 0 ? 1 : 0, maxCol = col == w - 1 ? 1 : 2;
      int code = 0;
      for (int r = minRow; r <= maxRow; r++)
        for (int c = minCol; c <= maxCol; c++)
          if (in[(row + r - 1) * w + (col + c - 1)] != 0) code += weights[r][c];
      float v = in[row * w + col];
      if (table[code]) {
        if (v != 0) removed++;
        v = 0;
      }
      out[row * w + col] = v;
    }
  return removed;
}

// skeleton.cu:84-99 normalizeSigmoid in FP32.
static float skel_sigmoid(float val) {
  val = (val <= 0.0f) ? 0.0f : ((val * 1.2f > 1.0f) ? 1.0f : 1.2f * val);
  val = 1.0f / (1.0f + std::exp(2.6433f - 10.7998f * val));
  return (val - 0.0667f) / 0.9333f;
}

// skeleton.cu:172-187 PostProcess; returns the thinning iterations.
extern "C" int orc_postprocess(int w, int h, float* gpb, const float* nmax, float* thin) {
  const int n = w * h;
  std::vector<float> a(n), b(n);
  for (int p = 0; p < n; p++) a[p] = nmax[p] > 0.05f ? 1.0f : 0.0f;
  int it = 0, changed;
  do {
    changed = 0;
    for (int l = 0; l < 8; l++) {
      changed += orc_skeleton_pass(w, h, l, a.data(), b.data());
      a.swap(b);
    }
    it++;
  } while (changed);
  for (int p = 0; p < n; p++) {
    const float t = a[p] != 0 ? gpb[p] : 0.0f;
    gpb[p] = skel_sigmoid(gpb[p]);
    thin[p] = skel_sigmoid(t);
  }
  return it;
}

// skeleton.cu:227-245 SigmoidGpbAll.
extern "C" void orc_normalize_gpb_all(int npix, int norient, int pitch, float* all) {
  for (int o = 0; o < norient; o++)
    for (int p = 0; p < npix; p++) {
      float v = all[(size_t)o * pitch + p];
      v = (v * 1.2f > 1.0f) ? 1.0f : v * 1.2f;
      if (v < 0.0f) v = 0.0f;
      v = 1.0f / (1.0f + std::exp(2.6433f - 10.7998f * v));
      all[(size_t)o * pitch + p] = (v - 0.0667f) / 0.9333f;
    }
}
