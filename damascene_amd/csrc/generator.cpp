// generator.cpp -- synthetic boundary-detection test images (host only).
//
// Voronoi partition of the image into `ncells` cells (sites uniform in
// [0,w) x [0,h)), each cell a uniform random RGB colour; half of the cells
// (chosen by a seeded shuffle) carry a sinusoidal luminance grating (period
// U[4,16] px, angle U[0,pi), phase U[0,2pi), amplitude 40); Gaussian noise
// sigma 8 per channel; rounded and clipped to uint8.  Deterministic for a
// given (w, h, ncells, seed) on any x86-64 host with the same libm.
#include <cmath>
#include <cstring>

#include "tables.hpp"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)
  int below(int n) { return (int)(uniform() * n); }
  double gauss() {  // Box-Muller, one value per call
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

}  // namespace

extern "C" int dmsc_generate_image(int w, int h, int ncells, uint64_t seed, uint32_t* rgb,
                                   int32_t* cells_out) {
  if (w <= 0 || h <= 0 || ncells <= 0 || !rgb) return dmsc::fail(DMSC_ERR_ARG, "generate_image: bad arguments");
  SplitMix64 rng(seed);
  std::vector<double> sx(ncells), sy(ncells);
  std::vector<int> col(3 * ncells);
  for (int c = 0; c < ncells; ++c) {
    sx[c] = rng.uniform() * w;
    sy[c] = rng.uniform() * h;
    for (int k = 0; k < 3; ++k) col[3 * c + k] = rng.below(256);
  }
  std::vector<int> order(ncells);
  for (int c = 0; c < ncells; ++c) order[c] = c;
  for (int c = ncells - 1; c > 0; --c) std::swap(order[c], order[rng.below(c + 1)]);
  std::vector<char> grating(ncells, 0);
  for (int i = 0; i < ncells / 2; ++i) grating[order[i]] = 1;
  std::vector<double> kx(ncells), ky(ncells), ph(ncells);
  for (int c = 0; c < ncells; ++c) {
    const double period = 4.0 + 12.0 * rng.uniform();
    const double ang = M_PI * rng.uniform();
    kx[c] = 2.0 * M_PI * std::cos(ang) / period;
    ky[c] = 2.0 * M_PI * std::sin(ang) / period;
    ph[c] = 2.0 * M_PI * rng.uniform();
  }
  std::vector<int32_t> cell((size_t)w * h);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      double best = 1e300;
      int bc = 0;
      for (int c = 0; c < ncells; ++c) {
        const double dx = x + 0.5 - sx[c], dy = y + 0.5 - sy[c];
        const double d = dx * dx + dy * dy;
        if (d < best) { best = d; bc = c; }
      }
      cell[(size_t)y * w + x] = bc;
    }
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const int c = cell[(size_t)y * w + x];
      const double g = grating[c] ? 40.0 * std::sin(kx[c] * x + ky[c] * y + ph[c]) : 0.0;
      uint32_t packed = 0;
      for (int k = 0; k < 3; ++k) {
        double v = col[3 * c + k] + g + 8.0 * rng.gauss();
        v = std::floor(v + 0.5);
        const uint32_t q = v < 0 ? 0u : (v > 255 ? 255u : (uint32_t)v);
        packed |= q << (8 * k);
      }
      rgb[(size_t)y * w + x] = packed;
    }
  if (cells_out) memcpy(cells_out, cell.data(), sizeof(int32_t) * (size_t)w * h);
  return DMSC_OK;
}

extern "C" int dmsc_boundary_mask(int w, int h, const int32_t* cells, uint8_t* mask) {
  if (w <= 0 || h <= 0 || !cells || !mask) return dmsc::fail(DMSC_ERR_ARG, "boundary_mask: bad arguments");
  std::vector<uint8_t> edge((size_t)w * h, 0);
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      const int c = cells[(size_t)y * w + x];
      const bool e = (x + 1 < w && cells[(size_t)y * w + x + 1] != c) ||
                     (x > 0 && cells[(size_t)y * w + x - 1] != c) ||
                     (y + 1 < h && cells[(size_t)(y + 1) * w + x] != c) ||
                     (y > 0 && cells[(size_t)(y - 1) * w + x] != c);
      edge[(size_t)y * w + x] = e;
    }
  for (int y = 0; y < h; ++y)
    for (int x = 0; x < w; ++x) {
      uint8_t m = 0;
      for (int dy = -1; dy <= 1 && !m; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int xx = x + dx, yy = y + dy;
          if (xx >= 0 && yy >= 0 && xx < w && yy < h && edge[(size_t)yy * w + xx]) { m = 1; break; }
        }
      mask[(size_t)y * w + x] = m;
    }
  return DMSC_OK;
}
