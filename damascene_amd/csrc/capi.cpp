// capi.cpp -- extern "C" entry points of libdmsc.so (include/dmsc.h).
#include <cctype>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <iterator>
#include <vector>

#include "dmsc_internal.hpp"
#include "kernels.hpp"

namespace dmsc {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int overlap_stages = 1;
int integral_pass = 1;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

Ctx::~Ctx() {
  for (Ctx* c : lanes) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete static_cast<dmsc_ctx*>(c);
  }
  release_eigensolver(this);
  release_local_cues(this);
  for (int i = 0; i < kArenas; i++)
    if (arena[i]) (void)hipFree(arena[i]);
  if (d_texton_coeffs) (void)hipFree(d_texton_coeffs);
  if (h_pinned) (void)hipHostFree(h_pinned);
  for (hipEvent_t e : aux_ev)
    if (e) (void)hipEventDestroy(e);
  if (aux) (void)hipStreamDestroy(aux);
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

// Pinned block layout: bytes [0, 256) the pipeline's k-means iteration
// count, [1024, 4096) the eigensolver's status record, [8192, 49152) the
// block eigensolver's Gram matrices and rotation.
int Ctx::pinned(int** out) {
  if (!h_pinned) DMSC_HIP(hipHostMalloc((void**)&h_pinned, 65536, hipHostMallocDefault));
  *out = h_pinned;
  return DMSC_OK;
}

int Ctx::aux_stream(hipStream_t* s, hipEvent_t* ev0, hipEvent_t* ev1) {
  if (!aux) {
    DMSC_HIP(hipSetDevice(device));
    DMSC_HIP(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
    DMSC_HIP(hipEventCreateWithFlags(&aux_ev[0], hipEventDisableTiming));
    DMSC_HIP(hipEventCreateWithFlags(&aux_ev[1], hipEventDisableTiming));
  }
  *s = aux;
  *ev0 = aux_ev[0];
  *ev1 = aux_ev[1];
  return DMSC_OK;
}

int Ctx::scratch(int slot, size_t bytes, void** ptr) {
  if (slot < 0 || slot >= kArenas) return fail(DMSC_ERR_INTERNAL, "bad scratch slot %d", slot);
  if (bytes > arena_bytes[slot]) {
    if (arena[slot]) {
      DMSC_HIP(hipStreamSynchronize(stream));
      DMSC_HIP(hipFree(arena[slot]));
      arena[slot] = nullptr;
      arena_bytes[slot] = 0;
    }
    DMSC_HIP(hipMalloc(&arena[slot], bytes));
    arena_bytes[slot] = bytes;
  }
  *ptr = arena[slot];
  return DMSC_OK;
}

int Ctx::adopt(int slot, void* p, size_t bytes) {
  if (arena[slot]) {
    DMSC_HIP(hipStreamSynchronize(stream));
    DMSC_HIP(hipFree(arena[slot]));
  }
  arena[slot] = p;
  arena_bytes[slot] = bytes;
  return DMSC_OK;
}

int Ctx::texton_coeffs(const float** out) {
  if (!d_texton_coeffs) {
    const TextonBank& b = texton_bank();
    DMSC_HIP(hipMalloc(&d_texton_coeffs, sizeof(float) * b.coeffs.size()));
    DMSC_HIP(hipMemcpy(d_texton_coeffs, b.coeffs.data(), sizeof(float) * b.coeffs.size(), hipMemcpyHostToDevice));
  }
  *out = d_texton_coeffs;
  return DMSC_OK;
}

}  // namespace dmsc

using namespace dmsc;

// texton_choice 1 -> 32 textons, 2 -> 64 (texton.cu:65-67)
static int texton_count(int texton_choice) { return texton_choice == DMSC_TEXTON64 ? 64 : 32; }

#define CHECK_CTX(ctx) DMSC_ARG((ctx) != nullptr, "null context")
#define CHECK_SIZE(w, h) \
  DMSC_ARG((w) >= kMinSide && (h) >= kMinSide && (long)(w) * (h) < (1L << 28), "image size %dx%d out of range", (w), (h))

extern "C" {

const char* dmsc_version(void) { return "dmsc 0.1.0 (gfx950)"; }
const char* dmsc_last_error_string(void) { return dmsc::g_last_error.c_str(); }

int dmsc_ctx_create(dmsc_ctx** out, int device, void* stream) {
  DMSC_ARG(out != nullptr, "null output pointer");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(DMSC_ERR_NODEV, "no HIP device available");
  DMSC_ARG(device >= 0 && device < count, "device %d out of range (%d devices)", device, count);
  DMSC_HIP(hipSetDevice(device));
  // the fixed scratch of dmsc_kmeans_noalloc: this device's address, once
  DMSC_TRY(noalloc_scratch_init(device, nullptr));
  dmsc_ctx* c = new dmsc_ctx();
  c->device = device;
  // NULL selects the device's default stream (as in HIP/CUDA); callers that
  // want a private stream create one and pass it in
  c->stream = (hipStream_t)stream;
  c->own_stream = false;
  *out = c;
  return DMSC_OK;
}

int dmsc_ctx_destroy(dmsc_ctx* ctx) {
  if (!ctx) return DMSC_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return DMSC_OK;
}

int dmsc_ctx_synchronize(dmsc_ctx* ctx) {
  CHECK_CTX(ctx);
  DMSC_HIP(hipStreamSynchronize(ctx->stream));
  return DMSC_OK;
}

void* dmsc_ctx_stream(dmsc_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int dmsc_pitch_in_floats(int n) { return pitch_floats(n); }

// ----------------------------------------------------------------- host tables
int dmsc_texton_filters(float* coeffs, int* radii) {
  DMSC_ARG(coeffs && radii, "null output");
  const TextonBank& b = texton_bank();
  memcpy(coeffs, b.coeffs.data(), sizeof(float) * b.coeffs.size());
  memcpy(radii, b.radii.data(), sizeof(int) * b.radii.size());
  return b.nbound;
}
int dmsc_parabola_filters(int radius, float* out) {
  DMSC_ARG(out && radius > 0 && radius <= 64, "bad arguments");
  parabola_filters(radius, out);
  return DMSC_OK;
}
int dmsc_spb_filters(float* out) {
  DMSC_ARG(out, "null output");
  spb_filters(out);
  return DMSC_OK;
}
int dmsc_rotation_tables(int width, int height, float thetaPi, int* nw, int* nh, int* bx, int* by, int* gx,
                         int* gy) {
  DMSC_ARG(width > 0 && height > 0 && nw && nh && bx && by && gx && gy, "bad arguments");
  const int rc = rotation_tables(width, height, thetaPi, nw, nh, bx, by, gx, gy);
  if (rc) return fail(DMSC_ERR_ARG, "rotation_tables: thetaPi %g outside (-1/2, 0]", thetaPi);
  return DMSC_OK;
}
int dmsc_halfdisc_offsets(int w, int h, float thetaPi, int radius, const int* bx, const int* by, const int* gx,
                          const int* gy, int rw, int* rect) {
  DMSC_ARG(bx && by && gx && gy && rect, "null table");
  halfdisc_offsets(w, h, thetaPi, radius, bx, by, gx, gy, rw, rect);
  return DMSC_OK;
}
int dmsc_blur_kernel(float sigma, float* kernel, int* radius) {
  DMSC_ARG(kernel && radius && sigma > 0 && sigma <= 20, "bad arguments");
  return blur_kernel(sigma, kernel, radius);
}
int dmsc_stencil(int radius, int width, int* offsets, int* diag) {
  DMSC_ARG(offsets && diag && radius > 0 && radius <= 10, "bad arguments");
  return stencil_geometry(radius, width, offsets, diag);
}
int dmsc_gamma_lut(float gamma, float* lut) {
  DMSC_ARG(lut, "null output");
  gamma_lut(gamma, lut);
  return DMSC_OK;
}

int dmsc_postprocess(dmsc_ctx* ctx, int w, int h, float* gpb, const float* nmax, float* gpb_thin,
                     int* iterations) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 0 && h > 0 && (long)w * h < (1L << 28) && gpb && nmax && gpb_thin, "bad arguments");
  return run_postprocess(ctx, w, h, gpb, nmax, gpb_thin, iterations);
}
int dmsc_normalize_gpb_all(dmsc_ctx* ctx, int npix, int norient, int pitch, float* all) {
  CHECK_CTX(ctx);
  DMSC_ARG(npix > 0 && norient > 0 && pitch >= npix && all, "bad arguments");
  return launch_normalize_gpb_all(ctx->stream, npix, norient, pitch, all);
}

int dmsc_tune_set(const char* key, int value) {
  DMSC_ARG(key, "null key");
  if (!std::strcmp(key, "spmv_rows_per_lane")) {
    DMSC_ARG(value == 0 || value == 1 || value == 2 || value == 4, "spmv_rows_per_lane must be 0, 1, 2 or 4");
    spmv_rows_per_lane = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_trace")) {
    lanczos_trace = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_k1_threads")) {
    DMSC_ARG(value == 0 || value == 64 || value == 256 || value == 512,
             "lanczos_k1_threads must be 0 (automatic), 64, 256 or 512");
    lanczos_k1_threads = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_mixed_fused")) {
    DMSC_ARG(value >= 0 && value <= 2, "lanczos_mixed_fused must be 0, 1 or 2");
    lanczos_mixed_fused = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_tile")) {
    DMSC_ARG(std::find(std::begin(kStepTiles), std::end(kStepTiles), value) != std::end(kStepTiles),
             "lanczos_tile must be one of 3208, 6404, 1616, 3216, 1608, 6408, 6402, 3204 (width * 100 + height)");
    lanczos_tile = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_mixed_stop")) {
    DMSC_ARG(value >= 0 && value <= 6, "lanczos_mixed_stop must be in [0, 6]");
    lanczos_mixed_stop = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_check_max")) {
    DMSC_ARG(value == 0 || (value >= 25 && value % 25 == 0), "lanczos_check_max must be 0 or a multiple of 25");
    lanczos_check_max = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_max_iter")) {
    DMSC_ARG(value >= 25 && value <= kMaxIter && value % 25 == 0, "lanczos_max_iter must be a multiple of 25 in [25, %d]",
             kMaxIter);
    lanczos_max_iter = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "batch_lanes")) {
    DMSC_ARG(value >= 1 && value <= 16, "batch_lanes must be in [1, 16]");
    batch_lanes = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "eigensolver")) {
    DMSC_ARG(value == 0 || value == 1, "eigensolver must be 0 (Lanczos) or 1 (Chebyshev block, the default)");
    eigensolver_method = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "chebyshev_shape")) {
    DMSC_ARG(value == 0 || chebyshev_shape_valid(value), "chebyshev_shape must be 0 (chosen per size), 24, 28, 22, 42 (register tiles), 324, 344, 342, 328, 348 (LDS tiles), 1324, 1328, 1336, 1337, 1344, 1348 (pipelined LDS tiles) or 2324, 2328, 2337 (pipelined, LDS rows read ahead)");
    chebyshev_shape = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_step_loads")) {
    DMSC_ARG(value >= 0 && value <= 2, "lanczos_step_loads must be 0, 1 or 2");
    lanczos_step_loads = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "overlap_stages")) {
    DMSC_ARG(value == 0 || value == 1, "overlap_stages must be 0 or 1");
    overlap_stages = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "integral_pass")) {
    DMSC_ARG(value == 0 || value == 1, "integral_pass must be 0 (row + column passes) or 1 (banded single write)");
    integral_pass = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "chebyshev_first_degree")) {
    DMSC_ARG(value >= 2 && value <= 300, "chebyshev_first_degree must be in [2, 300]");
    chebyshev_first_degree = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "chebyshev_start")) {
    DMSC_ARG(value == 0 || value == 1, "chebyshev_start must be 0 (hashed uniform values) or 1 (low-frequency cosines)");
    chebyshev_start = value;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "chebyshev_reduction")) {
    DMSC_ARG(value >= 2 && value <= 10000, "chebyshev_reduction must be in [2, 10000]");
    chebyshev_reduction = value;
    return DMSC_OK;
  }
  return fail(DMSC_ERR_ARG, "unknown tuning key %s", key);
}
int dmsc_tune_get(const char* key, int* value) {
  DMSC_ARG(key && value, "null argument");
  if (!std::strcmp(key, "spmv_rows_per_lane")) {
    *value = spmv_rows_per_lane;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_trace")) {
    *value = lanczos_trace;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_k1_threads")) {
    *value = lanczos_k1_threads;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_mixed_fused")) {
    *value = lanczos_mixed_fused;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_tile")) {
    *value = lanczos_tile;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_mixed_stop")) {
    *value = lanczos_mixed_stop;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_check_max")) {
    *value = lanczos_check_max;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_max_iter")) {
    *value = lanczos_max_iter;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "batch_lanes")) {
    *value = batch_lanes;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "eigensolver")) {
    *value = eigensolver_method;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "chebyshev_shape")) {
    *value = chebyshev_shape;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "lanczos_step_loads")) {
    *value = lanczos_step_loads;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "overlap_stages")) {
    *value = overlap_stages;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "integral_pass")) {
    *value = integral_pass;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "chebyshev_first_degree")) {
    *value = chebyshev_first_degree;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "chebyshev_start")) {
    *value = chebyshev_start;
    return DMSC_OK;
  }
  if (!std::strcmp(key, "chebyshev_reduction")) {
    *value = chebyshev_reduction;
    return DMSC_OK;
  }
  return fail(DMSC_ERR_ARG, "unknown tuning key %s", key);
}

// convert.cu:88-95 (cutLoadPPM4ub packing: R in the low byte).
int dmsc_load_ppm(const char* path, int* w, int* h, uint32_t* rgb) {
  DMSC_ARG(path && w && h, "null argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(DMSC_ERR_ARG, "cannot open %s", path);
  auto token = [&](int* out) -> bool {  // next integer, skipping whitespace and # comments
    int c = std::fgetc(f);
    for (;;) {
      while (c != EOF && std::isspace(c)) c = std::fgetc(f);
      if (c != '#') break;
      while (c != EOF && c != '\n') c = std::fgetc(f);
    }
    if (c == EOF || !std::isdigit(c)) return false;
    long v = 0;
    while (c != EOF && std::isdigit(c)) {
      v = v * 10 + (c - '0');
      if (v > (1L << 20)) return false;
      c = std::fgetc(f);
    }
    *out = (int)v;
    return true;  // the single whitespace after the number has been consumed
  };
  char magic[2] = {0, 0};
  int pw = 0, ph = 0, maxval = 0;
  const bool ok = std::fread(magic, 1, 2, f) == 2 && magic[0] == 'P' && magic[1] == '6' && token(&pw) &&
                  token(&ph) && token(&maxval);
  if (!ok || pw <= 0 || ph <= 0 || maxval != 255) {
    std::fclose(f);
    return fail(DMSC_ERR_ARG, "%s: not a binary PPM with maxval 255", path);
  }
  *w = pw;
  *h = ph;
  if (rgb) {
    std::vector<unsigned char> px((size_t)pw * ph * 3);
    const size_t got = std::fread(px.data(), 1, px.size(), f);
    if (got != px.size()) {
      std::fclose(f);
      return fail(DMSC_ERR_ARG, "%s: truncated pixel data", path);
    }
    for (size_t i = 0; i < (size_t)pw * ph; i++)
      rgb[i] = (uint32_t)px[3 * i] | ((uint32_t)px[3 * i + 1] << 8) | ((uint32_t)px[3 * i + 2] << 16);
  }
  std::fclose(f);
  return DMSC_OK;
}

// ----------------------------------------------------------------- colour
int dmsc_rgb_to_grey(dmsc_ctx* ctx, int w, int h, const uint32_t* rgb, float* grey) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 0 && h > 0 && rgb && grey, "bad arguments");
  return launch_rgb_to_grey(ctx->stream, w * h, rgb, grey);
}
int dmsc_rgb_to_lab(dmsc_ctx* ctx, int w, int h, float gamma, const uint32_t* rgb, float* L, float* a, float* b) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 0 && h > 0 && rgb && L && a && b, "bad arguments");
  return launch_rgb_to_lab(ctx->stream, w * h, gamma, rgb, L, a, b);
}
int dmsc_normalize_lab(dmsc_ctx* ctx, int w, int h, float* L, float* a, float* b) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 0 && h > 0 && L && a && b, "bad arguments");
  return launch_normalize_lab(ctx->stream, w * h, L, a, b);
}
int dmsc_mirror_image_f(dmsc_ctx* ctx, int w, int h, int border, const float* in, float* out) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 0 && h > 0 && border >= 0 && border <= w && border <= h && in && out, "bad arguments");
  return launch_mirror_f(ctx->stream, w, h, border, in, out);
}
int dmsc_mirror_image_i(dmsc_ctx* ctx, int w, int h, int border, const int* in, int* out) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 0 && h > 0 && border >= 0 && border <= w && border <= h && in && out, "bad arguments");
  return launch_mirror_i(ctx->stream, w, h, border, in, out);
}
int dmsc_unmirror_image_f(dmsc_ctx* ctx, int bw, int bh, int border, const float* in, float* out) {
  CHECK_CTX(ctx);
  DMSC_ARG(bw > 2 * border && bh > 2 * border && border >= 0 && in && out, "bad arguments");
  return launch_unmirror_f(ctx->stream, bw, bh, border, in, out);
}
int dmsc_quantize_image(dmsc_ctx* ctx, int w, int h, int nbins, const float* in, int* out) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 0 && h > 0 && nbins > 0 && in && out, "bad arguments");
  return launch_quantize(ctx->stream, w * h, nbins, in, out);
}

// ----------------------------------------------------------------- textons
int dmsc_texton_responses(dmsc_ctx* ctx, int w, int h, const float* grey, float* resp) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 0 && h > 0 && grey && resp, "bad arguments");
  const float* c;
  DMSC_TRY(ctx->texton_coeffs(&c));
  return launch_texton_conv(ctx->stream, w, h, grey, c, resp);
}

int dmsc_kmeans(dmsc_ctx* ctx, int w, int h, int k, int nfilters, const float* resp, int* labels, int max_iter,
                int conv_thresh, int precision, int* changes, int* iterations) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 0 && h > 0 && resp && labels, "bad arguments");
  DMSC_ARG((k == 32 || k == 64) && nfilters == kNFilters, "kmeans supports k=32 or 64 over 34 filters");
  DMSC_ARG(max_iter >= 1 && max_iter <= 1000, "max_iter out of range");
  DMSC_ARG(conv_thresh == 0, "only conv_thresh=0 (the reference default) is supported");
  DMSC_ARG(precision == DMSC_PREC_FP32 || precision == DMSC_PREC_MIXED, "bad precision");
  const int n = w * h;
  void* scr;
  const size_t st = (kmeans_state_bytes() + 255) & ~(size_t)255;
  DMSC_TRY(ctx->scratch(2, st + sizeof(float) * (size_t)pitch_floats(n), &scr));
  float* pdot = (float*)((char*)scr + st);
  if (precision == DMSC_PREC_MIXED) {
    DMSC_TRY(launch_kmeans_mixed(ctx->stream, w, h, k, resp, labels, pdot, scr, max_iter));
  } else {
    DMSC_TRY(launch_kmeans(ctx->stream, w, h, k, nfilters, resp, labels, pdot, scr, max_iter));
  }
  return kmeans_read_state(ctx->stream, scr, changes, iterations);
}

int dmsc_kmeans_rechecks(dmsc_ctx* ctx, int* rechecks) {
  CHECK_CTX(ctx);
  DMSC_ARG(rechecks, "null pointer");
  DMSC_ARG(ctx->arena[2] && ctx->arena_bytes[2] >= kmeans_state_bytes(), "no k-means run on this context");
  return kmeans_read_state(ctx->stream, ctx->arena[2], nullptr, nullptr, rechecks);
}

int dmsc_eigensolve_report(dmsc_ctx* ctx, int* stop_rule, double* max_residual) {
  CHECK_CTX(ctx);
  DMSC_ARG(stop_rule && max_residual, "null pointer");
  *stop_rule = ctx->eig_stop;
  *max_residual = ctx->eig_residual;
  return DMSC_OK;
}

int dmsc_find_textons(dmsc_ctx* ctx, int w, int h, const float* grey, int* textons, int texton_choice,
                      int precision, int* iterations) {
  CHECK_CTX(ctx);
  CHECK_SIZE(w, h);
  DMSC_ARG(grey && textons, "bad arguments");
  DMSC_ARG(texton_choice == DMSC_TEXTON32 || texton_choice == DMSC_TEXTON64,
           "texton_choice must be 1 (32 textons) or 2 (64 textons)");
  const int n = w * h;
  void* r;
  DMSC_TRY(ctx->scratch(0, sizeof(float) * (size_t)kNFilters * n, &r));
  DMSC_TRY(dmsc_texton_responses(ctx, w, h, grey, (float*)r));
  int changes;
  return dmsc_kmeans(ctx, w, h, texton_count(texton_choice), kNFilters, (float*)r, textons, 15, 0, precision,
                     &changes, iterations);
}

// ----------------------------------------------------------------- cues
int dmsc_local_cues(dmsc_ctx* ctx, int w, int h, const float* L, const float* a, const float* b, const int* tex,
                    float* bg, float* cga, float* cgb, float* tg, int cue_pitch, int texton_choice) {
  CHECK_CTX(ctx);
  CHECK_SIZE(w, h);
  DMSC_ARG(L && a && b && tex && bg && cga && cgb && tg, "null pointer");
  DMSC_ARG(cue_pitch >= w * h, "cue pitch smaller than the image");
  DMSC_ARG(texton_choice == DMSC_TEXTON32 || texton_choice == DMSC_TEXTON64,
           "texton_choice must be 1 (32 textons) or 2 (64 textons)");
  return launch_local_cues(ctx, w, h, L, a, b, tex, texton_count(texton_choice), bg, cga, cgb, tg, cue_pitch,
                           nullptr);
}

int dmsc_local_cues_raw(dmsc_ctx* ctx, int w, int h, const float* L, const float* a, const float* b,
                        const int* tex, float* bg, float* cga, float* cgb, float* tg, int cue_pitch,
                        float* raw, int texton_choice) {
  CHECK_CTX(ctx);
  CHECK_SIZE(w, h);
  DMSC_ARG(L && a && b && tex && bg && cga && cgb && tg && raw, "null pointer");
  DMSC_ARG(cue_pitch >= w * h, "cue pitch smaller than the image");
  DMSC_ARG(texton_choice == DMSC_TEXTON32 || texton_choice == DMSC_TEXTON64,
           "texton_choice must be 1 (32 textons) or 2 (64 textons)");
  return launch_local_cues(ctx, w, h, L, a, b, tex, texton_count(texton_choice), bg, cga, cgb, tg, cue_pitch,
                           raw);
}

// ----------------------------------------------------------------- combine / nonmax / gPb
int dmsc_combine(dmsc_ctx* ctx, int w, int h, int cue_pitch, const float* bg, const float* cga, const float* cgb,
                 const float* tg, float* mpb, float* cg, int texton_choice) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 0 && h > 0 && cue_pitch >= w * h && bg && cga && cgb && tg && mpb && cg, "bad arguments");
  DMSC_ARG(texton_choice == DMSC_TEXTON32 || texton_choice == DMSC_TEXTON64,
           "texton_choice must be 1 (32 textons) or 2 (64 textons)");
  return launch_combine(ctx->stream, w * h, cue_pitch, bg, cga, cgb, tg, texton_count(texton_choice), mpb, cg);
}

int dmsc_nonmax(dmsc_ctx* ctx, int w, int h, const float* pb, int pitch, float* nmax) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 0 && h > 0 && pitch >= w * h && pb && nmax, "bad arguments");
  const int n = w * h;
  void* scr;
  DMSC_TRY(ctx->scratch(2, sizeof(float) * (size_t)n * 2 + 512, &scr));
  float* mx = (float*)scr;
  int* ori = (int*)((char*)scr + (((size_t)n * sizeof(float) + 255) & ~(size_t)255));
  return launch_nonmax(ctx->stream, w, h, pb, pitch, mx, ori, nmax);
}

int dmsc_calc_gpb(dmsc_ctx* ctx, int npix, int pitch, int norient, const float* cg, const float* spb,
                  const float* mpb, float* all, float* gpb) {
  CHECK_CTX(ctx);
  DMSC_ARG(npix > 0 && pitch >= npix && norient > 0 && norient <= 8 && cg && spb && mpb && all && gpb,
           "bad arguments");
  return launch_calc_gpb(ctx->stream, npix, pitch, norient, cg, spb, mpb, all, gpb);
}

// ----------------------------------------------------------------- affinity / eigen
int dmsc_intervene(dmsc_ctx* ctx, int w, int h, const float* mpb, float* M, int pitch, float sigma) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 5 && h > 5 && pitch >= w * h && mpb && M && sigma > 0, "bad arguments");
  return launch_intervene(ctx->stream, w, h, mpb, M, pitch, sigma);
}

int dmsc_convert_matrix(dmsc_ctx* ctx, int w, int h, float* M, int pitch, float* dsum, float* rs) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 5 && h > 5 && pitch >= w * h && M && rs, "bad arguments");
  return launch_convert_matrix(ctx->stream, w, h, M, pitch, dsum, rs);
}

int dmsc_stencil_spmv(dmsc_ctx* ctx, int w, int h, const float* M, int pitch, const float* x, float* y) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 5 && h > 5 && pitch >= w * h && M && x && y, "bad arguments");
  DMSC_ARG(((uintptr_t)M % 16) == 0 && ((uintptr_t)y % 16) == 0, "matrix and y must be 16-byte aligned");
  return launch_spmv(ctx->stream, w, h, M, pitch, x, y);
}

int dmsc_generalized_eigensolve(dmsc_ctx* ctx, int w, int h, float* M, int pitch, int nvec, float tol,
                                int precision, float* eigvals, float* eigvecs, int* iterations) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 5 && h > 5 && pitch >= w * h && M && eigvals && eigvecs, "bad arguments");
  DMSC_ARG(nvec >= 2 && nvec <= 25, "nvec must be in [2, 25] (damascene.cu:130-141)");
  DMSC_ARG(tol >= 1e-5f && tol <= 1e-1f, "tol must be in [1e-5, 1e-1] (damascene.cu:142-151)");
  DMSC_ARG(precision == DMSC_PREC_FP32 || precision == DMSC_PREC_MIXED, "bad precision");
  DMSC_ARG((pitch % 4) == 0 && ((uintptr_t)M % 16) == 0, "matrix must be 16-byte aligned with pitch % 4 == 0");
  return run_eigensolve(ctx, w, h, M, pitch, nvec, tol, precision, eigvals, eigvecs, iterations);
}

int dmsc_lanczos_time_steps(dmsc_ctx* ctx, int w, int h, float* M, int pitch, int precision, int steps,
                            float* ms_per_step) {
  CHECK_CTX(ctx);
  CHECK_SIZE(w, h);
  DMSC_ARG(M && ms_per_step && pitch >= w * h, "bad arguments");
  DMSC_ARG(precision == DMSC_PREC_FP32 || precision == DMSC_PREC_MIXED, "bad precision");
  return time_lanczos_steps(ctx, w, h, M, pitch, precision, steps, ms_per_step);
}

int dmsc_chebyshev_time_products(dmsc_ctx* ctx, int w, int h, const float* M, int pitch, int products,
                                 float* ms_per_product) {
  CHECK_CTX(ctx);
  CHECK_SIZE(w, h);
  DMSC_ARG(M && ms_per_product && pitch >= w * h, "bad arguments");
  return time_chebyshev_products(ctx, w, h, M, pitch, products, ms_per_product);
}

int dmsc_autotune_eigensolver(dmsc_ctx* ctx, int w, int h, float* M, int pitch, int precision, int steps,
                              dmsc_step_shape* best, dmsc_step_shape* tried, int max_tried, int* n_tried) {
  CHECK_CTX(ctx);
  CHECK_SIZE(w, h);
  DMSC_ARG(M && pitch >= w * h && max_tried >= 0 && (tried || max_tried == 0), "bad arguments");
  DMSC_ARG(precision == DMSC_PREC_FP32 || precision == DMSC_PREC_MIXED, "bad precision");
  std::vector<StepShape> t(std::max(max_tried, 0));
  StepShape b{};
  int nt = 0;
  DMSC_TRY(autotune_steps(ctx, w, h, M, pitch, precision, steps, &b, t.data(), max_tried, &nt));
  auto put = [](const StepShape& x, dmsc_step_shape* o) {
    o->mixed_fused = x.fused;
    o->tile = x.tile;
    o->k1_threads = x.k1;
    o->ms_per_step = x.ms;
  };
  if (best) put(b, best);
  for (int i = 0; i < std::min(nt, max_tried); i++) put(t[i], &tried[i]);
  if (n_tried) *n_tried = nt;
  return DMSC_OK;
}

int dmsc_autotune_spmv(dmsc_ctx* ctx, int w, int h, const float* M, int pitch, const float* x, float* y, int reps,
                       int* best_rows_per_lane, float* us_per_launch) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 5 && h > 5 && pitch >= w * h && M && x && y && reps > 0, "bad arguments");
  DMSC_ARG(((uintptr_t)M % 16) == 0 && ((uintptr_t)y % 16) == 0, "matrix and y must be 16-byte aligned");
  const int shapes[4] = {0, 1, 2, 4};
  hipEvent_t e0, e1;
  DMSC_HIP(hipEventCreate(&e0));
  DMSC_HIP(hipEventCreate(&e1));
  const int saved = spmv_rows_per_lane;
  int best = saved, rc = DMSC_OK;
  float best_us = 0.f;
  for (int i = 0; i < 4 && rc == DMSC_OK; i++) {
    spmv_rows_per_lane = shapes[i];
    rc = launch_spmv(ctx->stream, w, h, M, pitch, x, y);  // warm-up
    if (rc == DMSC_OK && hipEventRecord(e0, ctx->stream) != hipSuccess) rc = fail(DMSC_ERR_HIP, "event record");
    for (int r = 0; r < reps && rc == DMSC_OK; r++) rc = launch_spmv(ctx->stream, w, h, M, pitch, x, y);
    float ms = 0.f;
    if (rc == DMSC_OK && (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                          hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
      rc = fail(DMSC_ERR_HIP, "SpMV timing events failed");
    const float us = 1e3f * ms / reps;
    if (us_per_launch) us_per_launch[i] = us;
    if (rc == DMSC_OK && (best_us == 0.f || us < best_us)) {
      best_us = us;
      best = shapes[i];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  spmv_rows_per_lane = rc == DMSC_OK ? best : saved;
  if (rc == DMSC_OK && best_rows_per_lane) *best_rows_per_lane = best;
  return rc;
}

int dmsc_combine_weights(int texton_choice, float* h_coef, float* h_wts) {
  DMSC_ARG(h_coef && h_wts, "null pointer");
  DMSC_ARG(texton_choice == DMSC_TEXTON32 || texton_choice == DMSC_TEXTON64, "texton_choice must be 1 or 2");
  combine_tables(texton_count(texton_choice), h_coef, h_wts);
  return DMSC_OK;
}

int dmsc_spb_combine_forward(dmsc_ctx* ctx, int w, int h, int cue_pitch, const float* bg, const float* cga,
                             const float* cgb, const float* tg, int nvec, const float* eigvecs, const float* params,
                             float* gpb_all, int pitch, float* gpb_pre, int* argmax) {
  CHECK_CTX(ctx);
  CHECK_SIZE(w, h);
  DMSC_ARG(bg && cga && cgb && tg && eigvecs && params && gpb_pre && argmax, "null pointer");
  DMSC_ARG(cue_pitch >= w * h && (!gpb_all || pitch >= w * h), "bad pitch");
  return spb_combine_forward(ctx, w, h, cue_pitch, bg, cga, cgb, tg, nvec, eigvecs, params, gpb_all, pitch, gpb_pre,
                             argmax);
}

int dmsc_spb_combine_backward(dmsc_ctx* ctx, int w, int h, int cue_pitch, const float* bg, const float* cga,
                              const float* cgb, const float* tg, int nvec, const float* eigvecs, const float* params,
                              const int* argmax, const float* grad_gpb_pre, float* grad_params,
                              float* grad_eigvecs) {
  CHECK_CTX(ctx);
  CHECK_SIZE(w, h);
  DMSC_ARG(bg && cga && cgb && tg && eigvecs && params && argmax && grad_gpb_pre && grad_params && grad_eigvecs,
           "null pointer");
  DMSC_ARG(cue_pitch >= w * h, "bad pitch");
  return spb_combine_backward(ctx, w, h, cue_pitch, bg, cga, cgb, tg, nvec, eigvecs, params, argmax, grad_gpb_pre,
                              grad_params, grad_eigvecs);
}

int dmsc_autotune_chebyshev(dmsc_ctx* ctx, int w, int h, const float* M, int pitch, int products, int* best_shape,
                            int* shapes_tried, float* us_per_product, int max_tried, int* n_tried) {
  CHECK_CTX(ctx);
  CHECK_SIZE(w, h);
  DMSC_ARG(M && pitch >= w * h && products >= 1 && products <= 10000, "bad arguments");
  DMSC_ARG(max_tried >= 0 && ((shapes_tried && us_per_product) || max_tried == 0), "bad arguments");
  return autotune_chebyshev(ctx, w, h, M, pitch, products, best_shape, shapes_tried, us_per_product, max_tried,
                            n_tried);
}

int dmsc_chebyshev_tile(dmsc_ctx* ctx, int w, int h, int* shape) {
  CHECK_CTX(ctx);
  CHECK_SIZE(w, h);
  DMSC_ARG(shape != nullptr, "null pointer");
  *shape = chebyshev_shape_for(ctx->device, w, h);
  return DMSC_OK;
}

int dmsc_autotune_clear(void) {
  autotune_clear();
  chebyshev_autotune_clear();
  return DMSC_OK;
}

int dmsc_spectral_pb(dmsc_ctx* ctx, const float* eigvals, const float* eigvecs, int w, int h, int nvec, float* spb,
                     int pitch) {
  CHECK_CTX(ctx);
  DMSC_ARG(w > 3 && h > 3 && pitch >= w * h && eigvals && eigvecs && spb && nvec >= 2 && nvec <= 32,
           "bad arguments");
  return launch_spectral_pb(ctx->stream, eigvals, eigvecs, w, h, nvec, spb, pitch);
}

// ----------------------------------------------------------------- pipeline
void dmsc_gpb_default_params(dmsc_gpb_params* p) {
  if (!p) return;
  p->n_eig = 9;
  p->tol = 1e-3f;
  p->texton_choice = DMSC_TEXTON32;
  p->precision = DMSC_PREC_FP32;
}

int dmsc_gpb(dmsc_ctx* ctx, int w, int h, const uint32_t* rgb, const dmsc_gpb_params* prm, dmsc_gpb_outputs* out,
             int timing) {
  CHECK_CTX(ctx);
  CHECK_SIZE(w, h);
  DMSC_ARG(rgb && out, "null pointer");
  dmsc_gpb_params p;
  if (prm) p = *prm; else dmsc_gpb_default_params(&p);
  DMSC_ARG(p.n_eig >= 2 && p.n_eig <= 25, "n_eig must be in [2, 25]");
  DMSC_ARG(p.texton_choice == DMSC_TEXTON32 || p.texton_choice == DMSC_TEXTON64,
           "texton_choice must be 1 (32 textons) or 2 (64 textons)");
  const int ntex = texton_count(p.texton_choice);
  hipStream_t s = ctx->stream;
  const int n = w * h;
  const int pitch = out->pitch ? out->pitch : pitch_floats(n);
  DMSC_ARG(pitch >= n && pitch % 4 == 0, "bad pitch");
  const size_t P = (size_t)pitch;
  // arena 0 layout (floats unless noted)
  size_t off = 0;
  auto carve = [&](size_t bytes) { size_t o = off; off += (bytes + 255) & ~(size_t)255; return o; };
  const size_t o_grey = carve(4 * P), o_L = carve(4 * P), o_A = carve(4 * P), o_B = carve(4 * P);
  const size_t o_resp = carve(4 * (size_t)kNFilters * n);
  const size_t o_tex = carve(4 * P);
  const size_t o_cues = carve(4 * 4 * 24 * P);
  const size_t o_mpb = carve(4 * 8 * P), o_cg = carve(4 * 8 * P), o_nmax = carve(4 * P);
  const size_t o_mat = carve(4 * (size_t)kStencilArea * P);
  const size_t o_vec = carve(4 * (size_t)p.n_eig * n);
  const size_t o_spb = carve(4 * 8 * P), o_all = carve(4 * 8 * P), o_gpb = carve(4 * P);
  void* base;
  DMSC_TRY(ctx->scratch(0, off, &base));
  char* B0 = (char*)base;
  float* grey = (float*)(B0 + o_grey);
  float *L = (float*)(B0 + o_L), *A = (float*)(B0 + o_A), *Bc = (float*)(B0 + o_B);
  float* resp = (float*)(B0 + o_resp);
  int* tex = out->d_textons ? out->d_textons : (int*)(B0 + o_tex);
  float* cues = (float*)(B0 + o_cues);
  float* mpb = out->d_mpb ? out->d_mpb : (float*)(B0 + o_mpb);
  float* cg = out->d_combined_g ? out->d_combined_g : (float*)(B0 + o_cg);
  float* nmax = out->d_nmax ? out->d_nmax : (float*)(B0 + o_nmax);
  float* mat = (float*)(B0 + o_mat);
  float* vec = out->d_eigvecs ? out->d_eigvecs : (float*)(B0 + o_vec);
  float* spb = out->d_spb ? out->d_spb : (float*)(B0 + o_spb);
  float* all = out->d_gpb_all ? out->d_gpb_all : (float*)(B0 + o_all);
  float* gpb = out->d_gpb ? out->d_gpb : (float*)(B0 + o_gpb);

  int* kit = nullptr;
  // ev[0..6]: stage boundaries; ev[7]: after the texton filter bank, ev[8]:
  // after the chi-squared gradients (finer splits, timed when the stages run
  // in order on one stream)
  hipEvent_t ev[9] = {};
  const bool overlap = ctx->overlap_stages && overlap_stages;
  if (timing) {
    for (int i = 0; i < 9; i++) DMSC_HIP(hipEventCreate(&ev[i]));
    DMSC_HIP(hipEventRecord(ev[0], s));
  }
  // damascene.cu:269-285 grey + textons; 309-317 Lab (gamma 2.5), normalised
  DMSC_TRY(launch_colour_fused(s, n, 2.5f, rgb, grey, L, A, Bc));
  // damascene.cu:352 localCues, L/a/b channels: they need no textons, so they
  // run on the context's second stream while the textons are computed
  // (on the main stream after the textons when the context does not overlap)
  hipStream_t aux = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  DMSC_TRY(local_cues_prepare(ctx, w, h, ntex));
  if (overlap) {
    DMSC_TRY(ctx->aux_stream(&aux, &fork, &join));
    DMSC_HIP(hipEventRecord(fork, s));
    DMSC_HIP(hipStreamWaitEvent(aux, fork, 0));
    DMSC_TRY(launch_local_cues_part(ctx, aux, 1, w, h, L, A, Bc, tex, ntex, cues, cues + 24 * P, cues + 48 * P,
                                    cues + 72 * P, pitch, nullptr));
    DMSC_HIP(hipEventRecord(join, aux));
  }
  const float* coeffs;
  DMSC_TRY(ctx->texton_coeffs(&coeffs));
  DMSC_TRY(launch_texton_conv(s, w, h, grey, coeffs, resp));
  if (timing) DMSC_HIP(hipEventRecord(ev[7], s));
  {
    void* scr;
    const size_t st = (kmeans_state_bytes() + 255) & ~(size_t)255;
    DMSC_TRY(ctx->scratch(2, st + sizeof(float) * P, &scr));
    float* pdot = (float*)((char*)scr + st);
    if (p.precision == DMSC_PREC_MIXED)
      DMSC_TRY(launch_kmeans_mixed(s, w, h, ntex, resp, tex, pdot, scr, 15));
    else
      DMSC_TRY(launch_kmeans(s, w, h, ntex, kNFilters, resp, tex, pdot, scr, 15));
    // the iteration count reaches the host without a stream synchronisation
    // here: the eigensolver synchronises before it returns
    DMSC_TRY(ctx->pinned(&kit));
    DMSC_HIP(hipMemcpyAsync(kit, kmeans_iterations_ptr(scr), sizeof(int), hipMemcpyDeviceToHost, s));
  }
  if (timing) DMSC_HIP(hipEventRecord(ev[1], s));
  // damascene.cu:352 localCues, texton channel; then join the L/a/b part
  DMSC_TRY(launch_local_cues_part(ctx, s, overlap ? 2 : 3, w, h, L, A, Bc, tex, ntex, cues, cues + 24 * P,
                                  cues + 48 * P, cues + 72 * P, pitch, nullptr, timing ? ev[8] : nullptr));
  if (overlap) DMSC_HIP(hipStreamWaitEvent(s, join, 0));
  if (timing) DMSC_HIP(hipEventRecord(ev[2], s));
  // damascene.cu:388-404 combine + nonMaxSuppression
  DMSC_TRY(launch_combine(s, n, pitch, cues, cues + 24 * P, cues + 48 * P, cues + 72 * P, ntex, mpb, cg));
  DMSC_TRY(launch_nonmax(s, w, h, mpb, pitch, spb, (int*)all, nmax));  // spb/all as scratch here
  if (timing) DMSC_HIP(hipEventRecord(ev[3], s));
  // damascene.cu:417-420 intervene
  DMSC_TRY(launch_intervene(s, w, h, nmax, mat, pitch, 0.1f));
  if (timing) DMSC_HIP(hipEventRecord(ev[4], s));
  // damascene.cu:433 generalizedEigensolve
  DMSC_TRY(run_eigensolve(ctx, w, h, mat, pitch, p.n_eig, p.tol, p.precision, out->h_eigvals, vec,
                          &out->lanczos_iterations));
  out->kmeans_iterations = *kit;  // copied before the eigensolver's own synchronisations
  if (timing) DMSC_HIP(hipEventRecord(ev[5], s));
  // damascene.cu:441-459 spectralPb + StartCalcGPb
  DMSC_TRY(launch_spectral_pb(s, out->h_eigvals, vec, w, h, p.n_eig, spb, pitch));
  DMSC_TRY(launch_calc_gpb(s, n, pitch, 8, cg, spb, nmax, all, gpb));
  if (timing) {
    DMSC_HIP(hipEventRecord(ev[6], s));
    DMSC_HIP(hipEventSynchronize(ev[6]));
    for (int i = 0; i < 6; i++) DMSC_HIP(hipEventElapsedTime(&out->stage_ms[i], ev[i], ev[i + 1]));
    out->stage_ms[6] = out->stage_ms[7] = 0.f;
    if (!overlap) {
      DMSC_HIP(hipEventElapsedTime(&out->stage_ms[6], ev[0], ev[7]));
      DMSC_HIP(hipEventElapsedTime(&out->stage_ms[7], ev[1], ev[8]));
    }
    for (int i = 0; i < 9; i++) (void)hipEventDestroy(ev[i]);
  }
  return DMSC_OK;
}

}  // extern "C"
