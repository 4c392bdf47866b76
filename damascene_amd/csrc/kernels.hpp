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
int launch_kmeans_mixed(hipStream_t s, int w, int h, int k, const float* resp, int* labels, float* pdot,
                        void* state, int max_iter);
int kmeans_read_state(hipStream_t s, void* state, int* changes, int* iterations, int* rechecks = nullptr);
const int* kmeans_iterations_ptr(const void* state);  // device address of the iteration count
// Device address of the current device's copy of the fixed k-means scratch
// of dmsc_kmeans_noalloc (looked up once per device; `device` is current).
int noalloc_scratch_init(int device, void** out);

// ------------------------------------------------------------- local cues
size_t local_cues_scratch_bytes(int w, int h);
int launch_local_cues(Ctx* ctx, int w, int h, const float* L, const float* A, const float* B, const int* tex,
                      int ntex, float* bg, float* cga, float* cgb, float* tg, int pitch, float* grads_out);
// The local cues in two parts for two streams: part 1 the L, a, b channels
// (needs no textons), part 2 the texton channel, 3 both; prepare first.
int local_cues_prepare(Ctx* ctx, int w, int h, int ntex);
// hist_done (may be null) is recorded on s between the chi-squared
// gradients and the parabola fits.
int launch_local_cues_part(Ctx* ctx, hipStream_t s, int part, int w, int h, const float* L, const float* A,
                           const float* B, const int* tex, int ntex, float* bg, float* cga, float* cgb, float* tg,
                           int pitch, float* grads_out, hipEvent_t hist_done = nullptr);
extern int overlap_stages;  // dmsc_gpb: L/a/b local cues beside the textons (tuning key "overlap_stages")
extern int integral_pass;   // local cues: 1 single-write banded integral images, 0 row + column passes

// ------------------------------------------------------------- combine / nonmax / gPb
int launch_combine(hipStream_t s, int n, int pitch, const float* bg, const float* cga, const float* cgb,
                   const float* tg, int ntex, float* mpb, float* cg);
int launch_nonmax(hipStream_t s, int w, int h, const float* pb, int pitch, float* mx, int* ori, float* nmax);
int launch_calc_gpb(hipStream_t s, int n, int pitch, int norient, const float* cg, const float* spb,
                    const float* mpb, float* all, float* gpb);

// ------------------------------------------------------------- affinity / stencil
int launch_intervene(hipStream_t s, int w, int h, const float* mpb, float* M, int pitch, float sigma);

struct StencilOffsets {
  int off[kStencilArea];          // cumulative linear offset of each diagonal
  signed char dy[kStencilArea];   // the diagonal's image-row offset (off = dy * w + dx)
  signed char dx[kStencilArea];   // and column offset
  int halo;                       // 5*w + 5
};
StencilOffsets make_stencil_offsets(int w);
int launch_convert_matrix(hipStream_t s, int w, int h, float* A, int pitch, float* dsum, float* rs);
int launch_spmv(hipStream_t s, int w, int h, const float* A, int pitch, const float* x, float* y);

// Stencil SpMV building block: RPL consecutive rows per lane starting at
// `row` (RPL * 4 bytes aligned), one vector load of A per diagonal, x
// gathered through the caches; a column outside [0, n) contributes
// fma(A, 0, acc) == acc (acc is never -0), which is the reference's bounds
// test without a branch.  XF(col) supplies x[col] (any float-valued source).
template <int RPL, typename XF>
__device__ __forceinline__ void spmv_rows(float (&acc)[RPL], const float* __restrict__ A, size_t pitch, int row,
                                          int n, const StencilOffsets& so, XF xf) {
#pragma unroll
  for (int j = 0; j < RPL; j++) acc[j] = 0.0f;
#pragma unroll 9
  for (int d = 0; d < kStencilArea; d++) {
    float a[RPL];
    const float* ap = A + (size_t)d * pitch + row;
    if constexpr (RPL == 4) {
      const float4 v = *reinterpret_cast<const float4*>(ap);
      a[0] = v.x; a[1] = v.y; a[2] = v.z; a[3] = v.w;
    } else if constexpr (RPL == 2) {
      const float2 v = *reinterpret_cast<const float2*>(ap);
      a[0] = v.x; a[1] = v.y;
    } else {
#pragma unroll
      for (int j = 0; j < RPL; j++) a[j] = ap[j];
    }
#pragma unroll
    for (int j = 0; j < RPL; j++) {
      const int col = row + j + so.off[d];
      const bool in = (unsigned)col < (unsigned)n;
      const float xv = xf(in ? col : 0);
      acc[j] = fmaf(a[j], in ? xv : 0.0f, acc[j]);
    }
  }
}

// SpMV launch shape chosen by the tuner (0: k_spmv_full; 1, 2, 4 rows per lane).
extern int spmv_rows_per_lane;
// Non-zero: the eigensolver prints one line per check to stderr.
extern int lanczos_trace;
// Rows per block of the Lanczos SpMV step kernel: 0 = automatic (256 for the
// two-kernel steps, 512 for the fused mixed step), or 64, 256, 512.
extern int lanczos_k1_threads;
// Mixed eigensolver step: 2 = one fused kernel per step over pixel tiles
// (default), 1 = one fused kernel over runs of consecutive rows (when
// the stencil halo fits in LDS), 0 = two kernels per step.
extern int lanczos_mixed_fused;
// Tile of the tiled fused mixed step, TW * 100 + TH: one of kStepTiles
// (3208, the default, first).
extern int lanczos_tile;
// Tiled fused mixed step: when a lane issues its 81 matrix loads (0 in
// groups of 9, 1 all after the staging barrier, 2 all before the partial
// reduction); results are identical.
extern int lanczos_step_loads;
extern int lanczos_check_max;  // mixed path: most steps between checks (0 = a quarter of the steps so far)
extern int lanczos_mixed_stop;  // mixed path: stop when every residual is <= tol * 10^-this (default 4)
extern int lanczos_max_iter;  // iteration cap, at most kMaxIter (multiple of 25)
extern int batch_lanes;
// Mixed path eigensolver: 0 Lanczos, 1 Chebyshev-filtered block subspace
// iteration (chebyshev.hip); its filter tile (chebyshev_shape_valid; 0 =
// chosen per size by chebyshev_shape_for).
extern int eigensolver_method;
extern int chebyshev_shape;
extern int chebyshev_reduction;  // tune key "chebyshev_reduction"
extern int chebyshev_start;      // tune key "chebyshev_start"
extern int chebyshev_first_degree;  // tune key "chebyshev_first_degree"
bool chebyshev_shape_valid(int v);
int chebyshev_max_vectors();  // most eigenvectors the block solver returns
// shape 0: the tile chebyshev_shape_for(w, h) picks
int time_chebyshev_products(Ctx* ctx, int w, int h, const float* A, int pitch, int products, float* ms_per_product,
                            int shape = 0);
int chebyshev_shape_for(int device, int w, int h);  // the filter tile a solve at w x h uses
int autotune_chebyshev(Ctx* ctx, int w, int h, const float* A, int pitch, int products, int* best, int* shapes,
                       float* us, int max_tried, int* n_tried);
void chebyshev_autotune_clear();
int chebyshev_solve(Ctx* ctx, int w, int h, const float* A, int pitch, int nvec, float tol, float* d_out,
                    double* vals, int* products);

// ------------------------------------------------------------- eigensolver
int run_eigensolve(Ctx* ctx, int w, int h, float* A, int pitch, int nvec, float tol, int precision,
                   float* h_eigvals, float* d_eigvecs, int* iterations);
void release_eigensolver(const Ctx* ctx);  // drops the context's captured graphs
void release_local_cues(const Ctx* ctx);   // drops the context's cached local-cue plans
int time_lanczos_steps(Ctx* ctx, int w, int h, float* A, int pitch, int precision, int steps, float* ms_per_step);
// Step-kernel shape of the Lanczos solver: mixed step form (0 two kernels,
// 1 linear fused, 2 tiled fused), tile code (width * 100 + height) and
// block size (0 = default); ms is the measured time per step.
struct StepShape {
  int fused, tile, k1;
  float ms;
};
// Tiles of the tiled fused mixed step (width * 100 + height); the first is
// the default.
constexpr int kStepTiles[] = {3208, 6404, 1616, 3216, 1608, 6408, 6402, 3204};
int autotune_steps(Ctx* ctx, int w, int h, float* A, int pitch, int precision, int steps, StepShape* best,
                   StepShape* tried, int max_tried, int* n_tried);
void autotune_clear();

// ------------------------------------------------------------- post-processing
int run_postprocess(Ctx* ctx, int w, int h, float* gpb, const float* nmax, float* gpb_thin, int* iterations);
int launch_normalize_gpb_all(hipStream_t s, int npix, int norient, int pitch, float* all);

// ------------------------------------------------------------- spectral Pb
int launch_spectral_pb(hipStream_t s, const float* h_eigvals, const float* d_eigvecs, int w, int h, int nvec,
                       float* spb, int pitch);

// combine.cu:16-21: mPb coefficients and combinedG weights (12 each) for
// 32 or 64 textons.
void combine_tables(int ntex, float* coef, float* wts);

// Differentiable spectral Pb + combine (params [12 cue weights, gamma,
// eta_0..eta_{nvec-1}] in device memory; see backward.hip).
constexpr int kMaxDiffVec = 16;
int spb_combine_forward(Ctx* ctx, int w, int h, int cue_pitch, const float* bg, const float* cga, const float* cgb,
                        const float* tg, int nvec, const float* V, const float* prm, float* all, int pitch,
                        float* gpb_pre, int* amax);
int spb_combine_backward(Ctx* ctx, int w, int h, int cue_pitch, const float* bg, const float* cga, const float* cgb,
                         const float* tg, int nvec, const float* V, const float* prm, const int* amax,
                         const float* grad, float* grad_prm, float* dV);

}  // namespace dmsc
