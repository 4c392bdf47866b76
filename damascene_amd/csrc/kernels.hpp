// kernels.hpp -- launch wrappers of the gPb kernels and the device helpers
// shared between translation units.
#pragma once
#include "dmsc_internal.hpp"

namespace dmsc {

// ------------------------------------------------------------- colour
int launch_rgb_to_grey(hipStream_t s, int n, const uint32_t* rgb, float* grey);
int launch_rgb_to_lab(hipStream_t s, int n, float gamma, const uint32_t* rgb, float* L, float* A, float* B);
int launch_normalize_lab(hipStream_t s, int n, float* L, float* A, float* B);
int launch_colour_fused(hipStream_t s, int n, float gamma, const uint32_t* rgb, float* grey, float* L,
                        float* A, float* B);
int launch_mirror_f(hipStream_t s, int w, int h, int border, const float* in, float* out);
int launch_mirror_i(hipStream_t s, int w, int h, int border, const int* in, int* out);
int launch_unmirror_f(hipStream_t s, int bw, int bh, int border, const float* in, float* out);
int launch_quantize(hipStream_t s, int n, int nbins, const float* in, int* out);

// ------------------------------------------------------------- textons
int launch_texton_conv(hipStream_t s, int w, int h, const float* grey, const float* d_coeffs, float* resp);
size_t kmeans_state_bytes();
int launch_kmeans(hipStream_t s, int w, int h, int k, int nf, const float* resp, int* labels, float* pdot,
                  void* state, int max_iter);
int launch_kmeans_mixed(hipStream_t s, int w, int h, const float* resp, int* labels, float* pdot, void* state,
                        int max_iter, void* work);
size_t kmeans_mixed_work_bytes(int n);
int kmeans_read_state(hipStream_t s, void* state, int* changes, int* iterations);

// ------------------------------------------------------------- local cues
size_t local_cues_scratch_bytes(int w, int h);
int launch_local_cues(Ctx* ctx, int w, int h, const float* L, const float* A, const float* B, const int* tex,
                      float* bg, float* cga, float* cgb, float* tg, int pitch, float* grads_out);

// ------------------------------------------------------------- combine / nonmax / gPb
int launch_combine(hipStream_t s, int n, int pitch, const float* bg, const float* cga, const float* cgb,
                   const float* tg, float* mpb, float* cg);
int launch_nonmax(hipStream_t s, int w, int h, const float* pb, int pitch, float* mx, int* ori, float* nmax);
int launch_calc_gpb(hipStream_t s, int n, int pitch, int norient, const float* cg, const float* spb,
                    const float* mpb, float* all, float* gpb);

// ------------------------------------------------------------- affinity / stencil
int launch_intervene(hipStream_t s, int w, int h, const float* mpb, float* M, int pitch, float sigma);

struct StencilOffsets {
  int off[kStencilArea];  // cumulative linear offset of each diagonal
  int halo;               // 5*w + 5
};
StencilOffsets make_stencil_offsets(int w);
int launch_convert_matrix(hipStream_t s, int w, int h, float* A, int pitch, float* dsum, float* rs);
int launch_spmv(hipStream_t s, int w, int h, const float* A, int pitch, const float* x, float* y);

constexpr int kSpmvThreads = 256;
constexpr int kSpmvRows = kSpmvThreads * 4;

// Stage x[r0 - halo, r0 + kSpmvRows + halo) * scale into LDS, zero outside
// [0, n).
__device__ __forceinline__ void spmv_stage_x(float* xs, const float* __restrict__ x, int n, int r0, int halo,
                                             float scale) {
  const int len = kSpmvRows + 2 * halo;
  for (int i = threadIdx.x; i < len; i += blockDim.x) {
    const int g = r0 - halo + i;
    xs[i] = (g >= 0 && g < n) ? x[g] * scale : 0.0f;
  }
}

// Four consecutive rows starting at `row`; xs index of row is `xb`.
__device__ __forceinline__ void spmv_rows4(float4& acc, const float* __restrict__ A, size_t pitch,
                                           const float* xs, int xb, const StencilOffsets& so) {
  acc = make_float4(0.f, 0.f, 0.f, 0.f);
  const int row_off = 0;
  (void)row_off;
#pragma unroll 9
  for (int d = 0; d < kStencilArea; d++) {
    const float4 a = *reinterpret_cast<const float4*>(A + (size_t)d * pitch);
    const float* xp = xs + xb + so.off[d];
    acc.x = fmaf(a.x, xp[0], acc.x);
    acc.y = fmaf(a.y, xp[1], acc.y);
    acc.z = fmaf(a.z, xp[2], acc.z);
    acc.w = fmaf(a.w, xp[3], acc.w);
  }
}

// ------------------------------------------------------------- eigensolver
int run_eigensolve(Ctx* ctx, int w, int h, float* A, int pitch, int nvec, float tol, int precision,
                   float* h_eigvals, float* d_eigvecs, int* iterations);

// ------------------------------------------------------------- spectral Pb
int launch_spectral_pb(hipStream_t s, const float* h_eigvals, const float* d_eigvecs, int w, int h, int nvec,
                       float* spb, int pitch);

}  // namespace dmsc
