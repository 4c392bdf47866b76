// io.cpp -- result files of the reference's driver (damascene/damascene.cu:
// 66-128, 480-520): .pb (writeFile), GpbAll.ary (writeArray over the eight
// oriented planes in the orientation order {3,2,1,0,7,6,5,4}, each
// transposed) and an 8-bit PGM of a float map.  Host-only.
#include <cstdio>
#include <vector>

#include "dmsc_internal.hpp"

using namespace dmsc;

namespace {

int write_all(const char* path, const std::vector<std::pair<const void*, size_t>>& parts) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(DMSC_ERR_ARG, "cannot open %s for writing", path);
  for (const auto& p : parts)
    if (std::fwrite(p.first, 1, p.second, f) != p.second) {
      std::fclose(f);
      return fail(DMSC_ERR_ARG, "short write to %s", path);
    }
  if (std::fclose(f) != 0) return fail(DMSC_ERR_ARG, "cannot close %s", path);
  return DMSC_OK;
}

}  // namespace

extern "C" {

// damascene.cu:80-89 writeFile: int width, int height, float[width*height].
int dmsc_write_pb(const char* path, int w, int h, const float* h_data) {
  DMSC_ARG(path && h_data && w > 0 && h > 0, "bad arguments");
  return write_all(path, {{&w, sizeof(int)}, {&h, sizeof(int)}, {h_data, sizeof(float) * (size_t)w * h}});
}

// damascene.cu:109-128 writeArray(ndim 3, dims {8, width, height}) of the
// gPbAll planes, plane i = transpose of orientation {3,2,1,0,7,6,5,4}[i]
// (damascene.cu:501-510).  h_all: [8][pitch] host copy of the planes.
int dmsc_write_gpb_all(const char* path, int w, int h, int pitch, const float* h_all) {
  DMSC_ARG(path && h_all && w > 0 && h > 0 && pitch >= w * h, "bad arguments");
  static const int kOriMap[8] = {3, 2, 1, 0, 7, 6, 5, 4};
  const size_t n = (size_t)w * h;
  std::vector<float> cat(8 * n);
  for (int i = 0; i < 8; i++) {
    const float* in = h_all + (size_t)kOriMap[i] * pitch;
    float* out = cat.data() + i * n;
    for (int row = 0; row < h; row++)
      for (int col = 0; col < w; col++) out[(size_t)col * h + row] = in[(size_t)row * w + col];
  }
  const int ndim = 3, dim[3] = {8, w, h};
  return write_all(path, {{&ndim, sizeof(int)}, {dim, sizeof(dim)}, {cat.data(), sizeof(float) * cat.size()}});
}

// Binary PGM (P5, maxval 255) of a float map: value v -> (unsigned char)(v *
// 255) after clamping v to [0, 1] (the reference's cutSavePGMf is part of
// the unvendored CUDA SDK utilities, so its rounding is not pinned here).
int dmsc_write_pgm(const char* path, int w, int h, const float* h_data) {
  DMSC_ARG(path && h_data && w > 0 && h > 0, "bad arguments");
  char head[64];
  const int hn = std::snprintf(head, sizeof(head), "P5\n%d %d\n255\n", w, h);
  std::vector<unsigned char> px((size_t)w * h);
  for (size_t i = 0; i < px.size(); i++) {
    float v = h_data[i];
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);  // NaN -> 0 below
    px[i] = v == v ? (unsigned char)(v * 255.0f) : 0;
  }
  return write_all(path, {{head, (size_t)hn}, {px.data(), px.size()}});
}

}  // extern "C"
