// tables.cpp -- host-side tables of the gPb path: texton filter bank,
// Savitzky-Golay parabola filters, spectralPb derivative filters, Bresenham
// rotation tables, half-disc rectangle offsets, histogram blur kernel,
// stencil geometry and the gamma table.
//
// These tables feed integer decisions downstream (k-means truncation,
// ellipse inclusion, histogram corners), so their arithmetic reproduces the
// reference host code's float/double semantics exactly (C++11 overload
// resolution: float sin/cos/tan/exp/round/ceil for float arguments, double
// pow(float,int)); this file is compiled with -ffp-contract=off and
// -fno-builtin so no operation is fused or constant-folded.
#include <cmath>
#include <cstring>
#include <mutex>

#include "tables.hpp"

namespace dmsc {
namespace {

constexpr long double kPiL = 3.1415926535897932384626433832795029L;

// 1-D sampled Gaussian or derivative on [-support, support] (filters.cpp:17-155):
// optional Hilbert transform by direct DFT in single precision, zero mean for
// derivatives, unit L1 norm.
struct Kernel1D {
  std::vector<float> v;

  Kernel1D(double sigma, unsigned deriv, bool hilbert, unsigned long support) {
    const double inv_s2 = double(1) / (sigma * sigma);
    const double g = double(-0.5) * inv_s2;
    const unsigned long len = 2 * support + 1;
    v.resize(len);
    double x = -static_cast<double>(support);
    for (unsigned long i = 0; i < len; ++i, ++x) {
      const double e = exp(x * x * g);
      v[i] = deriv == 0 ? (float)e : deriv == 1 ? (float)(e * (-x)) : (float)(e * (x * x * inv_s2 - 1));
    }
    if (hilbert) analytic_imag(len);
    if (deriv > 0) remove_mean();
    l1_normalise();
  }

  // Imaginary part of the analytic signal: forward DFT, double the positive
  // frequencies, zero the negative ones above len/2+1, inverse DFT.
  void analytic_imag(unsigned long len) {
    std::vector<float> Fr(len, 0.f), Fi(len, 0.f), Gi(len, 0.f);
    for (int k = 0; k < (int)len; ++k)
      for (int j = 0; j < (int)len; ++j) {
        const float c = (float)cos(2 * M_PI * k * j / len);
        const float s = (float)(-sin(2 * M_PI * k * j / len));
        Fr[k] = Fr[k] + v[j] * c;
        Fi[k] = Fi[k] + v[j] * s;
      }
    for (int k = 1; k < (int)len; ++k) {
      if (k < (int)((len + 1) / 2)) { Fr[k] *= 2.0f; Fi[k] *= 2.0f; }
      else if (k > (int)(len / 2 + 1)) { Fr[k] = 0.f; Fi[k] = 0.f; }
    }
    for (int k = 0; k < (int)len; ++k)
      for (int j = 0; j < (int)len; ++j) {
        const float c = (float)(cos(2 * M_PI * k * j / len) / len);
        const float s = (float)(sin(2 * M_PI * k * j / len) / len);
        Gi[k] = Gi[k] + (Fr[j] * s + Fi[j] * c);
      }
    v.assign(Gi.begin(), Gi.end());
  }
  void remove_mean() { remove_mean(v); }
  void l1_normalise() { l1_normalise(v); }
  static void remove_mean(std::vector<float>& a) {
    float s = 0;
    for (float t : a) s += t;
    const float m = s / (float)a.size();
    for (float& t : a) t -= m;
  }
  static void l1_normalise(std::vector<float>& a) {
    float s = 0;
    for (float t : a) s += std::fabs(t);
    for (float& t : a) t /= s;
  }
};

unsigned long rotated_extent(unsigned long sx, unsigned long sy, double ori, bool along_x) {
  const double a = along_x ? sx * cos(ori) : sx * sin(ori);
  const double b = along_x ? sy * sin(ori) : sy * cos(ori);
  const double m0 = std::fabs(a - b), m1 = std::fabs(a + b);
  return (unsigned long)ceil(m0 > m1 ? m0 : m1);
}

// Bilinear resampling of the separable kernel onto the rotated grid
// (filters.cpp:192-258); grid points without four source neighbours stay 0.
void resample_rotated(const std::vector<float>& src, unsigned long nx, unsigned long ny,
                      std::vector<float>& dst, unsigned long mx, unsigned long my, double ori) {
  const double co = cos(ori), si = sin(ori);
  const double cx = static_cast<double>(nx - 1) / 2, cy = static_cast<double>(ny - 1) / 2;
  double u = -(static_cast<double>(mx - 1) / 2);
  size_t o = 0;
  for (unsigned long i = 0; i < mx; ++i, ++u) {
    double v = -(static_cast<double>(my - 1) / 2);
    for (unsigned long j = 0; j < my; ++j, ++v, ++o) {
      const double x = u * co + v * si + cx;
      const double y = v * co - u * si + cy;
      if (!(x >= 0 && y >= 0)) continue;
      const unsigned long x0 = (unsigned long)floor(x), x1 = (unsigned long)ceil(x);
      const unsigned long y0 = (unsigned long)floor(y), y1 = (unsigned long)ceil(y);
      if (x1 >= nx || y1 >= ny) continue;
      const double wx0 = x - x0, wx1 = x1 - x, wy0 = y - y0, wy1 = y1 - y;
      const float a00 = src[x0 * ny + y0], a01 = src[x0 * ny + y1];
      const float a10 = src[x1 * ny + y0], a11 = src[x1 * ny + y1];
      const float r0 = (x0 != x1) ? (float)(wx1 * a00 + wx0 * a10) : a00;
      const float r1 = (x0 != x1) ? (float)(wx1 * a01 + wx0 * a11) : a01;
      dst[o] = (y0 != y1) ? (float)(wy1 * r0 + wy0 * r1) : r0;
    }
  }
}

// 3x3 (generally small) inverse by unblocked LU with partial pivoting and the
// LAPACK-reference inverse sequence, column-major, single precision.  Only
// the first column is consumed by the Savitzky-Golay filter.
void small_inverse(float* A, int n) {
  int piv[8];
  auto at = [&](int i, int j) -> float& { return A[i + j * n]; };
  for (int j = 0; j < n; ++j) {
    int p = j;
    for (int i = j + 1; i < n; ++i)
      if (std::fabs(at(i, j)) > std::fabs(at(p, j))) p = i;
    piv[j] = p;
    if (at(p, j) != 0.f) {
      if (p != j)
        for (int k = 0; k < n; ++k) std::swap(at(j, k), at(p, k));
      const float rcp = 1.0f / at(j, j);
      for (int i = j + 1; i < n; ++i) at(i, j) = at(i, j) * rcp;
    }
    for (int k = j + 1; k < n; ++k) {
      const float t = -at(j, k);
      for (int i = j + 1; i < n; ++i) at(i, k) = at(i, k) + at(i, j) * t;
    }
  }
  for (int j = 0; j < n; ++j) {  // invert U in place
    at(j, j) = 1.0f / at(j, j);
    const float m = -at(j, j);
    for (int c = 0; c < j; ++c) {
      const float t = at(c, j);
      if (t != 0.f) {
        for (int i = 0; i < c; ++i) at(i, j) = at(i, j) + t * at(i, c);
        at(c, j) = at(c, j) * at(c, c);
      }
    }
    for (int i = 0; i < j; ++i) at(i, j) = at(i, j) * m;
  }
  float tmp[8];
  for (int j = n - 1; j >= 0; --j) {  // solve inv(A) L = inv(U)
    for (int i = j + 1; i < n; ++i) { tmp[i] = at(i, j); at(i, j) = 0.f; }
    for (int c = j + 1; c < n; ++c) {
      const float t = -1.0f * tmp[c];
      for (int i = 0; i < n; ++i) at(i, j) = at(i, j) + t * at(i, c);
    }
  }
  for (int j = n - 2; j >= 0; --j)
    if (piv[j] != j)
      for (int i = 0; i < n; ++i) std::swap(at(i, j), at(i, piv[j]));
}

void line(int x0, int y0, int x1, int y1, int& len, int* xs, int* ys) {
  const int dx = x1 - x0, dy = y1 - y0, ax = std::abs(dx), ay = std::abs(dy);
  len = (ax > ay ? ax : ay) + 1;
  int oct;
  if (dx > 0 && dy >= 0) oct = ax > ay ? 0 : 1;
  else if (dx <= 0 && dy > 0) oct = ax < ay ? 2 : 3;
  else if (dy <= 0 && dx < 0) oct = ax > ay ? 4 : 5;
  else if (dx >= 0 && dy < 0) oct = ax < ay ? 6 : 7;
  else oct = -1;
  // per octant: major-axis step, minor-axis step, which axis is major
  static const int major_x[8] = {1, 0, 0, 1, 1, 0, 0, 1};
  static const int sx[8] = {1, 1, -1, -1, -1, -1, 1, 1};
  static const int sy[8] = {1, 1, 1, 1, -1, -1, -1, -1};
  int x = x0, y = y0, k = 0;
  xs[k] = x; ys[k] = y; ++k;
  if (oct < 0) return;
  const bool mx = major_x[oct];
  int t = mx ? -ax : -ay;
  while (x != x1 || y != y1) {
    if (mx) {
      x += sx[oct]; t += ay << 1;
      if (t > 0) { y += sy[oct]; t -= ax << 1; }
    } else {
      y += sy[oct]; t += ax << 1;
      if (t > 0) { x += sx[oct]; t -= ay << 1; }
    }
    xs[k] = x; ys[k] = y; ++k;
  }
}

}  // namespace

// filters.cpp:261-336 gaussian_2D: separable product on the rotated bounding
// grid, bilinear rotation, zero mean for derivatives, unit L1 norm.
void gaussian_2d(float* filter, double sx, double sy, double ori, unsigned deriv, bool hilbert,
                 unsigned long supx, unsigned long supy) {
  const unsigned long rx = rotated_extent(supx, supy, -ori, true);
  const unsigned long ry = rotated_extent(supx, supy, -ori, false);
  Kernel1D kx(sx, 0, false, rx), ky(sy, deriv, hilbert, ry);
  const unsigned long nx = 2 * rx + 1, ny = 2 * ry + 1;
  std::vector<float> sep(nx * ny);
  for (unsigned long i = 0; i < nx; ++i)
    for (unsigned long j = 0; j < ny; ++j) sep[i * ny + j] = kx.v[i] * ky.v[j];
  const unsigned long mx = 2 * supx + 1, my = 2 * supy + 1;
  std::vector<float> out(mx * my, 0.f);
  resample_rotated(sep, nx, ny, out, mx, my, ori);
  if (deriv > 0) Kernel1D::remove_mean(out);
  Kernel1D::l1_normalise(out);
  memcpy(filter, out.data(), sizeof(float) * out.size());
}

const TextonBank& texton_bank() {
  static TextonBank bank;
  static std::once_flag once;
  std::call_once(once, [] {
    // texton.cu:74-82 scales {2, 2*sqrt(2)}, elongation 3; filters.cpp:404-454
    const float scales[2] = {2.0f, (float)(2.0 * M_SQRT2)};
    bank.coeffs.assign(9010, 0.f);
    bank.radii.clear();
    size_t at = 0;
    int last = 0;
    for (float s : scales) {
      const int sup = (int)std::ceil(3 * s);
      const int area = (2 * sup + 1) * (2 * sup + 1);
      for (int hil = 0; hil < 2; ++hil)
        for (int o = 0; o < 8; ++o) {
          gaussian_2d(&bank.coeffs[at], s, s / 3.0f, o * M_PI / 8, 2, hil != 0, sup, sup);
          at += area;
          bank.radii.push_back(sup);
        }
      // centre-surround: surround(sigma) - centre(sigma/sqrt2), zero mean, L1
      std::vector<float> c(area), su(area);
      const double f = (double)M_SQRT2l;
      gaussian_2d(c.data(), s / f, s / f, 0, 0, false, sup, sup);
      gaussian_2d(su.data(), s, s, 0, 0, false, sup, sup);
      std::vector<float> d(area);
      float sum = 0;
      for (int i = 0; i < area; ++i) { d[i] = su[i] - c[i]; sum += d[i]; }
      const float mean = sum / (float)area;
      for (float& t : d) t -= mean;
      Kernel1D::l1_normalise(d);
      memcpy(&bank.coeffs[at], d.data(), sizeof(float) * area);
      at += area;
      bank.radii.push_back(sup);
      last = area;
    }
    bank.nbound = (int)at - last;  // texton.cu:104-109 binds all but the last filter
  });
  return bank;
}

// localcues.cu:21-100 savgol_filter (quadratic fit, value at the centre).
void savgol_filter(float* filt, int d, float inra, float inrb, float theta) {
  const float ra = inra > 1.5f ? inra : 1.5f, rb = inrb > 1.5f ? inrb : 1.5f;
  const float ira2 = (float)(1.0 / (ra * ra)), irb2 = (float)(1.0 / (rb * rb));
  const int wr = int(std::floor(ra > rb ? ra : rb) + 0.5), wd = 2 * wr + 1;
  const float st = std::sin(theta), ct = std::cos(theta);
  memset(filt, 0, sizeof(float) * wd * wd);
  auto inside = [&](int u, int v, float& ai) {
    ai = (float)(-u) * st + (float)v * ct;
    const float bi = (float)u * ct + (float)v * st;
    return !(ai * ai * ira2 + bi * bi * irb2 > 1);
  };
  std::vector<float> mom(2 * d + 1, 0.f);
  for (int u = -wr; u <= wr; ++u)
    for (int v = -wr; v <= wr; ++v) {
      float ai;
      if (!inside(u, v, ai)) continue;
      for (int i = 0; i <= 2 * d; ++i) mom[i] = (float)((double)mom[i] + std::pow((double)ai, i));
    }
  std::vector<float> A((d + 1) * (d + 1));
  for (int i = 0; i <= d; ++i)
    for (int j = 0; j <= d; ++j) A[i * (d + 1) + j] = mom[i + j];
  small_inverse(A.data(), d + 1);
  for (int u = -wr; u <= wr; ++u)
    for (int v = -wr; v <= wr; ++v) {
      float ai;
      if (!inside(u, v, ai)) continue;
      float& f = filt[(v + wr) * wd + u + wr];
      for (int i = 0; i <= d; ++i) f = (float)((double)f + (double)A[i] * std::pow((double)ai, i));
    }
}

void parabola_filters(int radius, float* out) {
  const int len = 2 * radius + 1;
  for (int o = 0; o < kOrient; ++o)
    savgol_filter(out + (size_t)len * len * o, 2, (float)radius, float(radius) / 4.0f,
                  (float)(M_PI / 2 - o * M_PI / 8));
}

void spb_filters(float* out) {
  for (int o = 0; o < kOrient; ++o)
    gaussian_2d(out + (size_t)o * 49, 1, 1, M_PI / 2 + (7 - o) * M_PI / 8, 1, false, 3, 3);
}

// rotate.cu:805-850: baseline/guideline Bresenham tables for thetaPi in
// (-1/2, 0]; the rotated image is (*nw x *nh).
int rotation_tables(int width, int height, float thetaPi, int* nw, int* nh, int* bX, int* bY,
                    int* gX, int* gY) {
  int len = 0;
  if (thetaPi == 0) {
    for (int i = 0; i < width; ++i) { bX[i] = i; bY[i] = 0; }
    for (int i = 0; i < height; ++i) { gX[i] = 0; gY[i] = i; }
    *nw = width;
    *nh = height;
    return 0;
  }
  const bool steep = thetaPi < -.25;
  if (!(thetaPi > -.5 && thetaPi < 0)) return -2;
  const float theta = steep ? (float)((thetaPi + 0.25) * kPiL) : (float)(thetaPi * kPiL);
  const float t = std::tan(theta);
  const float fw1 = (float)(width - 1), fh1 = (float)(height - 1);
  if (!steep) {
    *nw = (int)std::ceil((float)width - fh1 * t);
    *nh = (int)std::ceil((float)height - fw1 * t);
    line(0, 0, width - 1, (int)-std::round(fw1 * t), len, bX, bY);
    if (len != width) return -1;
    line((int)-std::round(fh1 * t), 0, 0, height - 1, len, gX, gY);
  } else {
    *nw = (int)std::ceil((float)height - fw1 * t);
    *nh = (int)std::ceil((float)width - fh1 * t);
    line(0, 0, (int)-std::round(fw1 * t), width - 1, len, bX, bY);
    if (len != width) return -1;
    line(height - 1, 0, 0, (int)-std::round(fh1 * t), len, gX, gY);
  }
  return len == height ? 0 : -1;
}

// rotate.cu:480-614: half-disc as an axis-aligned rectangle in the rotated
// frame; eight corner offsets A..H (linear, rotated image).  The tables
// cover the mirrored frame (w + 2 kBorder by h + 2 kBorder) while the
// reference centres the disc at (w/2, h/2); for images under about 50 px
// that puts its corner lookups outside the tables, so they are clamped to
// the table range (no effect at larger sizes).
void halfdisc_offsets(int w, int h, float thetaPi, int radius, const int* bX, const int* bY,
                      const int* gX, const int* gY, int rw, int* rect) {
  const float theta = (float)(thetaPi * kPiL);
  const float er = (float)((-4.0 + 4.0 * (double)(float)radius * (double)std::sqrt((float)kPiL)) / 8.0);
  const float s = std::sin(theta), c = std::cos(theta);
  const int bw = w + 2 * kBorder, bh = h + 2 * kBorder;
  auto rot = [&](int x, int y, int& rx, int& ry) {
    x = std::min(std::max(x, 0), bw - 1);
    y = std::min(std::max(y, 0), bh - 1);
    rx = gX[y] + bX[x];
    ry = gY[y] + bY[x];
  };
  const int xc = w / 2, yc = h / 2;
  int cx, cy, ax, ay, bx, by;
  rot(xc, yc, cx, cy);
  rot((int)std::round(-er * s - er * c) + xc, (int)std::round(-er * s + er * c) + yc, ax, ay);
  rot((int)std::round(er * s + er * c) + xc, (int)std::round(er * s - er * c) + yc, bx, by);
  const int xd = (bx - cx) - (ax - cx) + 1, yd = (ay - cy) - (by - cy) + 1;
  const int xr = xd / 2, yr = yd / 2;
  const int left = (xd & 1) ? -xr - 1 : -xr, top = (yd & 1) ? -yr - 1 : -yr;
  const int xs[8] = {left, 0, xr, left, xr, left, 0, xr};
  const int ys[8] = {top, top, top, 0, 0, yr, yr, yr};
  for (int i = 0; i < 8; ++i) rect[i] = ys[i] * rw + xs[i];
}

// rotate.cu:29-62 precompute_gaussian.
int blur_kernel(float sigma, float* k, int* radius) {
  const int r = (int)std::ceil(sigma * 3), n = 2 * r + 1;
  const float g = float(-0.5) * (float(1) / sigma / sigma);
  float s = 0;
  for (int i = 0; i < n; ++i) {
    k[i] = std::exp(float(i - r) * float(i - r) * g);
    s += k[i];
  }
  for (int i = 0; i < n; ++i) k[i] /= s;
  *radius = r;
  return n;
}

// Stencil.cpp:6-35.
int stencil_geometry(int radius, int width, int* offsets, int* diagMap) {
  const int dia = 2 * radius + 1;
  int prev = 0, n = 0;
  for (int i = 0; i < dia * dia; ++i) diagMap[i] = -1;
  for (int dy = -radius; dy <= radius; ++dy)
    for (int dx = -radius; dx <= radius; ++dx) {
      if (dy * dy + dx * dx > radius * radius) continue;
      const int lin = dy * width + dx;
      offsets[n] = lin - prev;
      prev = lin;
      diagMap[(dy + radius) * dia + dx + radius] = n++;
    }
  return n;
}

void gamma_lut(float gamma, float* lut) {
  for (int c = 0; c < 256; ++c) lut[c] = (float)std::pow((double)((float)c / 255.0f), (double)gamma);
}

}  // namespace dmsc
