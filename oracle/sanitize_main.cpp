// Host-code sanitizer driver (test infrastructure): the CPU oracle and the
// library's host table builders (damascene_amd/csrc/tables.cpp, the image
// generator) built with -fsanitize=address,undefined by `make -C oracle
// sanitize` and run by tests/test_sanitize.py.
//
// It generates small synthetic images, builds every host table of the
// library and of the oracle for several sizes and angles (checking the two
// agree bit for bit), and runs the oracle's whole gPb pipeline with 32 and
// 64 textons, so every oracle stage and table builder executes under the
// address and undefined-behaviour checkers.  Exit status 0 = clean run with
// matching tables; the sanitizers abort with their own report otherwise.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "oracle.h"
#include "tables.hpp"

extern "C" int dmsc_generate_image(int w, int h, int ncells, uint64_t seed, uint32_t* rgb, int32_t* cells);

// the library's error hooks (defined by capi.cpp in the shared library)
namespace dmsc {
void set_error(const std::string&) {}
int fail(int code, const char*, ...) { return code; }
}  // namespace dmsc

static int g_bad = 0;
static void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "mismatch: %s\n", what);
    g_bad++;
  }
}

static void tables(int w, int h) {
  using namespace dmsc;
  const int bw = w + 2 * kBorder, bh = h + 2 * kBorder;
  for (float th : {0.0f, -0.125f, -0.25f, -0.375f}) {
    int nw = 0, nh = 0, onw = 0, onh = 0;
    std::vector<int> t(2 * bw + 2 * bh), o(2 * bw + 2 * bh);
    expect(rotation_tables(w + 2 * kBorder, h + 2 * kBorder, th, &nw, &nh, t.data(), t.data() + bw,
                           t.data() + 2 * bw, t.data() + 2 * bw + bh) == 0,
           "rotation_tables rc");
    orc_rotation_tables(w + 2 * kBorder, h + 2 * kBorder, th, &onw, &onh, o.data(), o.data() + bw, o.data() + 2 * bw,
                        o.data() + 2 * bw + bh);
    expect(nw == onw && nh == onh && t == o, "rotation tables");
    for (int r : {3, 5, 10, 20}) {
      int a[8], b[8];
      halfdisc_offsets(w, h, th, r, t.data(), t.data() + bw, t.data() + 2 * bw, t.data() + 2 * bw + bh, nw, a);
      orc_halfdisc_offsets(w, h, th, r, o.data(), o.data() + bw, o.data() + 2 * bw, o.data() + 2 * bw + bh, nw, b);
      expect(std::memcmp(a, b, sizeof(a)) == 0, "half-disc offsets");
    }
  }
  int so[121], sd[121], oo[121], od[121];
  expect(stencil_geometry(5, w, so, sd) == orc_stencil(5, w, oo, od), "stencil size");
  expect(std::memcmp(so, oo, sizeof(int) * 81) == 0 && std::memcmp(sd, od, sizeof(sd)) == 0, "stencil tables");
}

static void fixed_tables() {
  using namespace dmsc;
  const TextonBank& tb = texton_bank();
  std::vector<float> oc(9010);
  std::vector<int> orad(34);
  orc_texton_filters(oc.data(), orad.data());
  expect(std::memcmp(tb.coeffs.data(), oc.data(), sizeof(float) * 9010) == 0 && tb.radii == orad, "texton bank");
  for (int r : {3, 5, 10, 20}) {
    const size_t len = (size_t)8 * (2 * r + 1) * (2 * r + 1);
    std::vector<float> a(len), b(len);
    parabola_filters(r, a.data());
    orc_parabola_filters(r, b.data());
    expect(std::memcmp(a.data(), b.data(), sizeof(float) * len) == 0, "parabola filters");
  }
  float s1[8 * 49], s2[8 * 49];
  spb_filters(s1);
  orc_spb_filters(s2);
  expect(std::memcmp(s1, s2, sizeof(s1)) == 0, "spectral Pb filters");
  for (float sg : {1.0f, 2.0f}) {
    float k1[64], k2[64];
    int r1 = 0, r2 = 0;
    blur_kernel(sg, k1, &r1);
    orc_blur_kernel(sg, k2, &r2);
    expect(r1 == r2 && std::memcmp(k1, k2, sizeof(float) * (2 * r1 + 1)) == 0, "blur kernel");
  }
  float g1[256], g2[256];
  gamma_lut(2.5f, g1);
  orc_gamma_lut(2.5f, g2);
  expect(std::memcmp(g1, g2, sizeof(g1)) == 0, "gamma table");
}

static void pipeline(int w, int h, int cells, uint64_t seed, int ntex) {
  const int n = w * h, nvec = 9;
  std::vector<uint32_t> rgb(n);
  std::vector<int32_t> cell(n);
  expect(dmsc_generate_image(w, h, cells, seed, rgb.data(), cell.data()) == 0, "generator rc");
  std::vector<int> tex(n);
  std::vector<float> mpb(8 * n), cg(8 * n), nmax(n), ev(nvec), evec((size_t)nvec * n), spb(8 * n), all(8 * n),
      gpb(n);
  double ms[8];
  const int rc = orc_gpb_pipeline(w, h, rgb.data(), nvec, 1e-3f, tex.data(), mpb.data(), cg.data(), nmax.data(),
                                  ev.data(), evec.data(), spb.data(), all.data(), gpb.data(), ms, ntex);
  expect(rc == 0, "oracle pipeline rc");
  bool finite = true;
  for (int i = 0; i < n; i++) finite = finite && std::isfinite(gpb[i]);
  expect(finite, "finite gPb");
  std::printf("pipeline %dx%d, %d textons: rc %d\n", w, h, ntex, rc);
}

int main() {
  orc_set_threads(2);
  fixed_tables();
  for (int s : {32, 45, 61}) tables(s + 3, s);
  pipeline(64, 48, 6, 20261202, 32);
  pipeline(48, 40, 5, 20261203, 64);
  std::printf("sanitized run: %d mismatches\n", g_bad);
  return g_bad ? 1 : 0;
}
