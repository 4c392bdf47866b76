/*
 * dmsc.h -- C ABI of the MI355X-native gPb library (libdmsc.so).
 *
 * Plain pointers and sizes only.  "d_" arguments are device pointers
 * (hipMalloc'd memory on the context's device), "h_" arguments are host
 * pointers.  Every entry point returns DMSC_OK (0) or an error code and
 * never exits the process; dmsc_last_error_string() describes the last
 * failure on the calling thread.  All work is enqueued on the context's
 * stream; entry points that return host results synchronise that stream.
 *
 * Each stage entry point replaces one function of the reference's stage
 * API (the reference's include/ headers); the replaced interface is
 * cited beside it.  The reference-named C++ functions (texton.h, kmeans.h,
 * localcues.h, ...) in this directory are thin wrappers over these.
 *
 * Layouts: images are row-major [h][w] (index y*w + x), N = w*h.  Oriented
 * stacks are [planes][pitch] floats with pitch >= N (cue stacks: 24 planes,
 * scale-major then orientation; mPb/combinedG/sPb/gPbAll: 8 planes).  The
 * affinity matrix is DIA: [81][pitch] floats, diagonal-major, diagonal d
 * being the d-th point of the radius-5 disc in scan order.
 */
#ifndef DMSC_H
#define DMSC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DMSC_OK = 0,
  DMSC_ERR_ARG = 1,      /* invalid argument (sizes, null pointers, ranges) */
  DMSC_ERR_HIP = 2,      /* HIP runtime error */
  DMSC_ERR_NOCONV = 3,   /* eigensolver screening failed, or reached the cap without nvec screened values */
  DMSC_ERR_NODEV = 4,    /* no usable GPU */
  DMSC_ERR_INTERNAL = 5
};

enum { DMSC_PREC_FP32 = 0, DMSC_PREC_MIXED = 1 };
enum { DMSC_TEXTON32 = 1, DMSC_TEXTON64 = 2 };

typedef struct dmsc_ctx dmsc_ctx;

const char* dmsc_version(void);
const char* dmsc_last_error_string(void);

/* Context: device, stream (NULL = the device's default stream; the stream
 * is borrowed, never destroyed) and grow-only scratch arenas.  Not
 * thread-safe; use one context per thread. */
int dmsc_ctx_create(dmsc_ctx** ctx, int device, void* stream);
int dmsc_ctx_destroy(dmsc_ctx* ctx);
int dmsc_ctx_synchronize(dmsc_ctx* ctx);
void* dmsc_ctx_stream(dmsc_ctx* ctx);

/* stencilMVM.cu:215-221 findPitchInFloats: row pitch (floats) of a pitched
 * plane of n floats (256-byte aligned). */
int dmsc_pitch_in_floats(int n);

/* ------------------------------------------------------------------ host utilities
 * (CPU only; safe without a GPU) */

/* Synthetic test image: Voronoi cells (uniform sites), uniform random cell
 * colours, gratings on half the cells (period U[4,16], angle U[0,pi),
 * amplitude 40), Gaussian noise sigma 8, clipped to uint8.  Packed RGB8 in a
 * uint32 per pixel (R bits 0-7, G 8-15, B 16-23; convert.h:6).  cells may be
 * NULL; otherwise receives the cell index per pixel. */
int dmsc_generate_image(int w, int h, int ncells, uint64_t seed, uint32_t* h_rgb, int32_t* h_cells);
/* Ground-truth boundary mask: pixels with a 4-neighbour in another cell,
 * dilated by one pixel (3x3). */
int dmsc_boundary_mask(int w, int h, const int32_t* h_cells, uint8_t* h_mask);

/* Host tables (bit-identical to what the kernels use). */
int dmsc_texton_filters(float* h_coeffs /*9010*/, int* h_radii /*34*/); /* filters.cpp:404 */
int dmsc_parabola_filters(int radius, float* h_out /*8*(2r+1)^2*/);     /* localcues.cu:102 */
int dmsc_spb_filters(float* h_out /*8*49*/);                            /* spectralPb.cu:708 */
int dmsc_rotation_tables(int width, int height, float thetaPi, int* new_w, int* new_h,
                         int* h_bx, int* h_by, int* h_gx, int* h_gy);   /* rotate.cu:805 */
int dmsc_halfdisc_offsets(int w, int h, float thetaPi, int radius, const int* h_bx,
                          const int* h_by, const int* h_gx, const int* h_gy, int rotated_w,
                          int* h_rect /*8*/);                            /* rotate.cu:480 */
int dmsc_blur_kernel(float sigma, float* h_kernel /*<=64*/, int* radius); /* rotate.cu:29 */
int dmsc_stencil(int radius, int width, int* h_offsets, int* h_diag_map); /* Stencil.cpp:6 */
int dmsc_gamma_lut(float gamma, float* h_lut /*256*/);
/* combine.cu:16-21: the 12 mPb coefficients and 12 combinedG weights for
 * texton_choice 1 (32 textons) or 2 (64). */
int dmsc_combine_weights(int texton_choice, float* h_coef /*12*/, float* h_wts /*12*/);
/* No reference counterpart: eigenvalues (unsorted) and eigenvectors (column
 * j of the row-major n x n output) of a symmetric FP64 matrix, by
 * Householder tridiagonalisation and implicit symmetric QR: the block
 * eigensolver's Rayleigh-Ritz step, exported for testing. */
int dmsc_symmetric_eigen(int n, const double* h_a, double* h_eigvals, double* h_eigvecs);
/* convert.cu:88-95 loadPPM_rgbU (cutLoadPPM4ub packing): binary PPM (P6,
 * maxval 255) -> host uint32 per pixel, R in bits 0-7, G 8-15, B 16-23.
 * h_rgb may be NULL to query the size. */
int dmsc_load_ppm(const char* path, int* w, int* h, uint32_t* h_rgb);

/* ------------------------------------------------------------------ colour (convert.h) */
int dmsc_rgb_to_grey(dmsc_ctx* ctx, int w, int h, const uint32_t* d_rgb, float* d_grey);          /* convert.h:53 */
int dmsc_rgb_to_lab(dmsc_ctx* ctx, int w, int h, float gamma, const uint32_t* d_rgb,
                    float* d_L, float* d_a, float* d_b);                                          /* convert.h:27 */
int dmsc_normalize_lab(dmsc_ctx* ctx, int w, int h, float* d_L, float* d_a, float* d_b);          /* convert.h:40 */
int dmsc_mirror_image_f(dmsc_ctx* ctx, int w, int h, int border, const float* d_in, float* d_out); /* convert.h:75 */
int dmsc_mirror_image_i(dmsc_ctx* ctx, int w, int h, int border, const int* d_in, int* d_out);     /* convert.h:97 */
int dmsc_unmirror_image_f(dmsc_ctx* ctx, int bw, int bh, int border, const float* d_in,
                          float* d_out);                                                          /* convert.h:120 */
int dmsc_quantize_image(dmsc_ctx* ctx, int w, int h, int nbins, const float* d_in, int* d_out);   /* convert.h:140 */

/* ------------------------------------------------------------------ textons */
/* texton.cu:39-57: 34 filter responses [34][N] (filter 33 is zero). */
int dmsc_texton_responses(dmsc_ctx* ctx, int w, int h, const float* d_grey, float* d_resp);
/* kmeans.h:18 kmeans: k=32 or 64 clusters over [nfilters][N] responses; labels
 * out; *changes = label changes of the last iteration; *iterations = count.
 * precision: DMSC_PREC_FP32 (VALU distances, reference order) or
 * DMSC_PREC_MIXED (BF16 MFMA distances + FP32 re-check of near ties; same
 * labels). */
int dmsc_kmeans(dmsc_ctx* ctx, int w, int h, int k, int nfilters, const float* d_resp,
                int* d_labels, int max_iter, int conv_thresh, int precision, int* changes,
                int* iterations);
/* No reference counterpart: after a DMSC_PREC_MIXED dmsc_kmeans call on this
 * context, the number of pixel assignments (summed over iterations) whose
 * BF16 MFMA distances could not certify the winner and were recomputed with
 * the exact FP32 distances. */
int dmsc_kmeans_rechecks(dmsc_ctx* ctx, int* rechecks);
/* No reference counterpart: dmsc_kmeans for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and labels as
 * dmsc_kmeans; its scratch (the k-means state and |x|^2 per pixel) lives in
 * a fixed device array of DMSC_KMEANS_NOALLOC_SCRATCH_BYTES per GPU, held
 * from library load to process exit, so the call allocates nothing.  Inputs
 * of more than dmsc_kmeans_noalloc_max_pixels() pixels (2,092,608: a
 * 1920x1080 image fits) return DMSC_ERR_ARG.  With changes and iterations
 * both NULL the call only enqueues work on the context's stream (no
 * read-back, no synchronisation).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other contexts,
 * other threads) race. */
#define DMSC_KMEANS_NOALLOC_SCRATCH_BYTES (8u << 20)
size_t dmsc_kmeans_noalloc_max_pixels(void);
int dmsc_kmeans_noalloc(dmsc_ctx* ctx, int w, int h, int k, int nfilters, const float* d_resp,
                        int* d_labels, int max_iter, int conv_thresh, int precision, int* changes,
                        int* iterations);
/* texton.h:16 findTextons (texton_choice 1 = 32 textons, 2 = 64). */
int dmsc_find_textons(dmsc_ctx* ctx, int w, int h, const float* d_grey, int* d_textons,
                      int texton_choice, int precision, int* iterations);

/* ------------------------------------------------------------------ local cues */
/* localcues.h:20 localCues: Bg/Cga/Cgb/Tg, each [24][cue_pitch]. */
int dmsc_local_cues(dmsc_ctx* ctx, int w, int h, const float* d_L, const float* d_a,
                    const float* d_b, const int* d_textons, float* d_bg, float* d_cga,
                    float* d_cgb, float* d_tg, int cue_pitch, int texton_choice);
/* As dmsc_local_cues, also exporting the raw chi-squared planes before the
 * parabola smoothing (gradient.cu:240-241): d_raw [4 cues][3 scales][8][N]. */
int dmsc_local_cues_raw(dmsc_ctx* ctx, int w, int h, const float* d_L, const float* d_a,
                        const float* d_b, const int* d_textons, float* d_bg, float* d_cga,
                        float* d_cgb, float* d_tg, int cue_pitch, float* d_raw, int texton_choice);

/* ------------------------------------------------------------------ combine / nonmax / gPb */
/* combine.h:15 combine -> mPb [8][pitch], combinedG [8][pitch]. */
int dmsc_combine(dmsc_ctx* ctx, int w, int h, int cue_pitch, const float* d_bg,
                 const float* d_cga, const float* d_cgb, const float* d_tg, float* d_mpb,
                 float* d_combined_g, int texton_choice);
/* nonmax.h:16 nonMaxSuppression: pb [8][pitch] -> nmax [N]. */
int dmsc_nonmax(dmsc_ctx* ctx, int w, int h, const float* d_pb, int pitch, float* d_nmax);
/* globalPb.h:24 StartCalcGPb. */
int dmsc_calc_gpb(dmsc_ctx* ctx, int npix, int pitch, int norient, const float* d_combined_g,
                  const float* d_spb, const float* d_mpb, float* d_gpb_all, float* d_gpb);

/* ------------------------------------------------------------------ affinity / eigensolve */
/* intervening.h:15 intervene: DIA affinity matrix [81][matrix_pitch]. */
int dmsc_intervene(dmsc_ctx* ctx, int w, int h, const float* d_mpb, float* d_matrix,
                   int matrix_pitch, float sigma);
/* stencilMVM.h:34 convertMatrix: W -> L_sym = I - D^-1/2 W D^-1/2 in place;
 * d_dsum (may be NULL) = row sums, d_rsqrt = D^-1/2. */
int dmsc_convert_matrix(dmsc_ctx* ctx, int w, int h, float* d_matrix, int matrix_pitch,
                        float* d_dsum, float* d_rsqrt);
/* stencilMVM.h:26 stencilMVM: y = A x. */
int dmsc_stencil_spmv(dmsc_ctx* ctx, int w, int h, const float* d_matrix, int matrix_pitch,
                      const float* d_x, float* d_y);
/* lanczos.h:17 generalizedEigensolve: (D-W) v = lambda D v from the affinity
 * matrix W (overwritten by L_sym).  h_eigvals[nvec] (host), d_eigvecs
 * [nvec][N] (device, vectors 1.. min-max normalised as the reference).
 * precision: DMSC_PREC_FP32 = reference algorithm (FP32 Lanczos without
 * reorthogonalisation, Cullum screening every 500 steps, residual test at
 * tol); DMSC_PREC_MIXED = FP32 matrix and FP32-stored basis with FP64 SpMV
 * accumulation, three-term vector and inner products, Cullum screening on
 * an adaptive schedule, residuals at tol * 1e-4 (the "lanczos_mixed_stop"
 * tuning value).  The FP64 tridiagonal
 * eigensolve runs on the GPU in both. */
int dmsc_generalized_eigensolve(dmsc_ctx* ctx, int w, int h, float* d_matrix, int matrix_pitch,
                                int nvec, float tol, int precision, float* h_eigvals,
                                float* d_eigvecs, int* iterations);
/* No reference counterpart: how the last dmsc_generalized_eigensolve (or
 * dmsc_gpb) on this context stopped, and the largest residual
 * |L_sym u - lambda u| of its wanted pairs at that point (the Lanczos
 * estimate |beta_m s_mj|, or the block solver's FP64 Gram residual).
 *   DMSC_STOP_RESIDUAL: every residual met the tolerance (plain: tol; mixed:
 *                       tol * 10^-lanczos_mixed_stop);
 *   DMSC_STOP_STEADY:   block solver: no wanted Ritz vector moved by more
 *                       than 0.03 tol since the previous outer iteration;
 *   DMSC_STOP_FLOOR:    block solver: the residuals stagnated above the
 *                       tolerance at the FP32 storage floor (about 4.8e-7);
 *   DMSC_STOP_CAP:      Lanczos reached its iteration cap and returned the
 *                       last check's Ritz pairs (lanczos.cu:519-560). */
enum { DMSC_STOP_NONE = 0, DMSC_STOP_RESIDUAL = 1, DMSC_STOP_STEADY = 2, DMSC_STOP_FLOOR = 3, DMSC_STOP_CAP = 4 };
int dmsc_eigensolve_report(dmsc_ctx* ctx, int* stop_rule, double* max_residual);
/* spectralPb.h:19 spectralPb -> [8][pitch] (accumulates into d_spb after
 * zeroing it). */
int dmsc_spectral_pb(dmsc_ctx* ctx, const float* h_eigvals, const float* d_eigvecs, int w,
                     int h, int nvec, float* d_spb, int pitch);

/* ------------------------------------------------------------------ post-processing */
/* skeleton.cu:172-187 PostProcess: thin mask = nmax > 0.05 thinned by the
 * eight look-up tables to a fixed point; d_gpb_thin = sigmoid(gPb on the
 * mask); d_gpb is replaced by sigmoid(gPb) (in place, as the reference).
 * *iterations (may be NULL) = thinning iterations run. */
int dmsc_postprocess(dmsc_ctx* ctx, int w, int h, float* d_gpb, const float* d_nmax, float* d_gpb_thin,
                     int* iterations);
/* skeleton.cu:247-252 NormalizeGpbAll: sigmoid of the oriented gPb planes,
 * in place. */
int dmsc_normalize_gpb_all(dmsc_ctx* ctx, int npix, int norient, int pitch, float* d_gpb_all);

/* ------------------------------------------------------------------ result files (host) */
/* damascene.cu:80-89 writeFile (.pb): int width, int height, float[w*h]. */
int dmsc_write_pb(const char* path, int w, int h, const float* h_data);
/* damascene.cu:109-128 writeArray of GpbAll.ary: int 3, int {8, w, h}, then
 * the planes of orientations {3,2,1,0,7,6,5,4} (damascene.cu:501-510), each
 * transposed; h_all is the host copy of the [8][pitch] gPbAll planes. */
int dmsc_write_gpb_all(const char* path, int w, int h, int pitch, const float* h_all);
/* 8-bit binary PGM of a float map in [0, 1] (Pb.pgm). */
int dmsc_write_pgm(const char* path, int w, int h, const float* h_data);

/* ------------------------------------------------------------------ tuning */
/* Launch-shape knobs read by the kernels at launch (process-wide; no
 * reference counterpart: they back the project's tuner).  Keys:
 *   "spmv_rows_per_lane": 0 = one row per lane in 256-thread blocks with all
 *                         81 loads in flight (default); 1, 2, 4 = one wave
 *                         per 64*RPL rows.
 *   "lanczos_k1_threads": rows per block of the eigensolver's SpMV step
 *                         kernel: 0 = automatic (default: 256 for two-kernel
 *                         steps, 512 for the fused mixed step), 64, 256 or
 *                         512 (512 applies to the fused step only).
 *   "lanczos_mixed_fused": mixed eigensolver step as one kernel staging the
 *                         basis vector in LDS, over pixel tiles (2, default)
 *                         or runs of consecutive rows (1, when the stencil
 *                         halo fits), or as two kernels (0).
 *   "lanczos_tile":       tile of the tiled fused step, width * 100 + height:
 *                         3208 (default), 6404, 1616, 3216, 1608, 6408,
 *                         6402, 3204.
 *   "batch_lanes":        default lane count of dmsc_gpb_batch (1-16, 4).
 *   "lanczos_mixed_stop": the mixed eigensolver stops when every Ritz
 *                         residual is at most tol * 10^-value (0-6, default
 *                         4).
 *   "lanczos_check_max":  the mixed eigensolver's most steps between two
 *                         checks (0 = a quarter of the steps so far, the
 *                         default; else a multiple of 25).
 *   "lanczos_max_iter":   Lanczos iteration cap, a multiple of 25 up to 6000
 *                         (lanczos.cu:4 MAXITER, the default).  At the cap
 *                         the eigensolver returns the Ritz pairs of its last
 *                         check, as the reference does, with the iteration
 *                         count equal to the cap.
 *   "lanczos_trace":      non-zero prints one line per eigensolver check to
 *                         stderr (2: also the Ritz values of a failed
 *                         screening).
 *   "eigensolver":        mixed precision only: 0 Lanczos (the default), 1
 *                         Chebyshev-filtered subspace iteration on a block of
 *                         32 vectors with an FP64 Rayleigh-Ritz step (no
 *                         reference counterpart; nvec <= 16).  It stops when
 *                         every wanted residual is at most
 *                         tol * 10^-lanczos_mixed_stop, when no wanted Ritz
 *                         vector moved by more than 0.03 tol (sine of the
 *                         angle) since the previous outer iteration, or when
 *                         the residuals stagnate at the FP32 floor
 *                         (dmsc_eigensolve_report tells which); the iteration
 *                         count it reports is the number of block products
 *                         with the matrix.
 *   "chebyshev_shape":    its filter kernel's tile, rows per lane * 10 +
 *                         float4 planes per lane: 24 (the default), 28, 22
 *                         or 42.
 *   "overlap_stages":     dmsc_gpb runs the L, a and b local cues on the
 *                         context's second stream beside the textons (1, the
 *                         default) or every stage in order on one stream (0;
 *                         results are bit-identical, and per-stage timing
 *                         then splits the textons and the local cues).
 * Returns DMSC_ERR_ARG for an unknown key or value. */
/* Measurement (no reference counterpart; it backs the roofline figure of
 * the benchmark): converts d_matrix in place as the eigensolver does, runs
 * `steps` (a multiple of 25) Lanczos steps of the given precision with no
 * convergence checks between two HIP events on the context's stream, after
 * 25 untimed steps, and returns the average milliseconds per step.  The step
 * kernels are the ones dmsc_generalized_eigensolve uses. */
int dmsc_lanczos_time_steps(dmsc_ctx* ctx, int w, int h, float* d_matrix, int matrix_pitch, int precision,
                            int steps, float* ms_per_step);
/* No reference counterpart: times `products` filter steps of the block
 * eigensolver (eigensolver = 1) on a converted matrix (L_sym, as
 * dmsc_convert_matrix leaves it) with the tile dmsc_chebyshev_tile names, a
 * block of 32 vectors, after 5 untimed steps, with HIP events on the
 * context's stream; returns the average milliseconds per step. */
int dmsc_chebyshev_time_products(dmsc_ctx* ctx, int w, int h, const float* d_matrix, int matrix_pitch,
                                 int products, float* ms_per_product);
/* No reference counterpart: the filter tile (a chebyshev_shape value) that a
 * block solve at w x h uses on this context's device: the autotuner's pick
 * for that size, else the tuning key, else (key 0, the default) 2337 where
 * its grid fits in one round of the device's block slots and 2328's does
 * not, else 2328. */
int dmsc_chebyshev_tile(dmsc_ctx* ctx, int w, int h, int* shape);
/* No reference counterpart: times every filter tile of the block eigensolver
 * (the 18 chebyshev_shape values) at w x h on a converted matrix, `products`
 * steps each, and makes the fastest the tile of later block solves at that
 * size (tiles give bit-identical results).  shapes_tried/us_per_product
 * receive up to max_tried (tile, microseconds per step) pairs;
 * dmsc_autotune_clear forgets the choices. */
int dmsc_autotune_chebyshev(dmsc_ctx* ctx, int w, int h, const float* d_matrix, int matrix_pitch, int products,
                            int* best_shape, int* shapes_tried, float* us_per_product, int max_tried,
                            int* n_tried);
/* Autotuner (no reference counterpart; the reference's launch shapes are
 * fixed).  dmsc_autotune_eigensolver converts d_matrix in place once, then
 * times every step-kernel shape of the given precision -- mixed: tiled
 * fused steps with tiles 32x8, 64x4, 16x16, 32x16, 16x8, 64x8, 64x2, 32x4,
 * linear fused steps
 * of 512 and 256 rows per block, two-kernel steps of 256 and 64; plain:
 * two-kernel steps of 256 and 64 -- with `steps` (a multiple of 25) timed
 * steps each, and records the fastest as the shape every later eigensolve
 * of this size and precision uses (in this process).  Shapes change results
 * only through the rounding of block partial sums.  `tried` receives up to
 * max_tried trials, *n_tried their number. */
typedef struct {
  int mixed_fused;    /* 0 two kernels, 1 linear fused, 2 tiled fused (mixed only) */
  int tile;           /* tiled step: tile width * 100 + height */
  int k1_threads;     /* block size of the (first) step kernel, 0 = default */
  float ms_per_step;  /* measured */
} dmsc_step_shape;
int dmsc_autotune_eigensolver(dmsc_ctx* ctx, int w, int h, float* d_matrix, int matrix_pitch, int precision,
                              int steps, dmsc_step_shape* best, dmsc_step_shape* tried, int max_tried,
                              int* n_tried);
/* Times the four SpMV launch shapes (rows per lane 0 = one row per lane
 * over all rows, 1, 2, 4; `reps` launches each on d_x -> d_y), leaves the
 * fastest selected process-wide ("spmv_rows_per_lane") and returns it with
 * the four average launch durations in microseconds. */
int dmsc_autotune_spmv(dmsc_ctx* ctx, int w, int h, const float* d_matrix, int matrix_pitch, const float* d_x,
                       float* d_y, int reps, int* best_rows_per_lane, float* us_per_launch /* [4] */);
/* Forgets every autotuned eigensolver shape. */
int dmsc_autotune_clear(void);
int dmsc_tune_set(const char* key, int value);
int dmsc_tune_get(const char* key, int* value);

/* ------------------------------------------------------------------ whole pipeline */
typedef struct {
  int n_eig;          /* number of eigenvectors (2..25), default 9 */
  float tol;          /* Lanczos residual tolerance, default 1e-3 */
  int texton_choice;  /* DMSC_TEXTON32 (default) or DMSC_TEXTON64 */
  int precision;      /* DMSC_PREC_FP32 or DMSC_PREC_MIXED */
} dmsc_gpb_params;

typedef struct {
  /* caller-allocated device buffers; any may be NULL */
  int* d_textons;        /* [N] */
  float* d_mpb;          /* [8][pitch] */
  float* d_combined_g;   /* [8][pitch] */
  float* d_nmax;         /* [N] */
  float* d_eigvecs;      /* [n_eig][N] */
  float* d_spb;          /* [8][pitch] */
  float* d_gpb_all;      /* [8][pitch] */
  float* d_gpb;          /* [N] */
  int pitch;             /* pitch of the oriented stacks (0 = dmsc_pitch_in_floats(N)) */
  /* host results */
  float h_eigvals[32];
  int kmeans_iterations;
  int lanczos_iterations;
  float stage_ms[8];     /* when timing is on: [0] colour + textons (+ the
                            L/a/b local cues when they overlap), [1] local
                            cues (the rest), [2] combine + nonmax, [3]
                            intervene, [4] eigensolve, [5] spectralPb + gPb;
                            when the stages do not overlap ("overlap_stages"
                            0): [6] colour + texton filter bank (part of
                            [0]) and [7] rotated labels, integral images and
                            chi-squared gradients (part of [1]); else 0 */
} dmsc_gpb_outputs;

void dmsc_gpb_default_params(dmsc_gpb_params* p);
/* damascene.cu:269-459: RGB8 (device) -> gPb.  timing != 0 records
 * per-stage milliseconds with HIP events (adds synchronisation). */
int dmsc_gpb(dmsc_ctx* ctx, int w, int h, const uint32_t* d_rgb, const dmsc_gpb_params* params,
             dmsc_gpb_outputs* out, int timing);

/* Differentiable spectral Pb + combine (no reference counterpart: the
 * reference has no backward pass).  The forward generalises
 * combine.cu:24-70, spectralPb.cu:651-659 and globalPb.cu:5-49 with free
 * weights, d_params = [w_0..w_11 (cue x scale weights, combine.cu:21 order),
 * gamma (the sPb weight, 0.0139 in globalPb.cu:13), eta_0..eta_{nvec-1}
 * (eigenvector weights, 1/sqrt(lambda_i) in spectralPb.cu:759; eta_0
 * unused)], a device array of 13 + nvec floats:
 *   gPbAll_o = sum_k w_k C_{k,o} + gamma sum_{i>=1} eta_i |K_o * mirror3(v_i)|,
 *   gPbPre   = max_o gPbAll_o, argmax = the first maximising orientation.
 * With the reference's weights gPbAll is bit-identical to dmsc_gpb's.
 * d_gpb_all ([8][pitch]) may be NULL.  nvec in [2, 16].
 * The backward takes dL/dgPbPre and the forward's argmax and writes
 * dL/dparams (13 + nvec floats, device; eta_0's is 0) and dL/deigvecs
 * ([nvec][N], device; vector 0's is 0).  Parameter gradients are reduced in
 * a fixed order (deterministic). */
int dmsc_spb_combine_forward(dmsc_ctx* ctx, int w, int h, int cue_pitch, const float* d_bg, const float* d_cga,
                             const float* d_cgb, const float* d_tg, int nvec, const float* d_eigvecs,
                             const float* d_params, float* d_gpb_all, int pitch, float* d_gpb_pre, int* d_argmax);
int dmsc_spb_combine_backward(dmsc_ctx* ctx, int w, int h, int cue_pitch, const float* d_bg, const float* d_cga,
                              const float* d_cgb, const float* d_tg, int nvec, const float* d_eigvecs,
                              const float* d_params, const int* d_argmax, const float* d_grad_gpb_pre,
                              float* d_grad_params, float* d_grad_eigvecs);

/* One image of a batch: its size, its RGB8 pixels (device) and the
 * outputs, exactly as the arguments of dmsc_gpb. */
typedef struct {
  int w, h;
  const uint32_t* d_rgb;
  dmsc_gpb_outputs* out;
} dmsc_batch_image;

/* No counterpart in the reference, which runs one image per process
 * (damascene.cu:130-207).  Runs every image of the batch through the
 * dmsc_gpb pipeline with the same parameters, spread over `lanes` child
 * contexts of ctx (0 = the "batch_lanes" tuning value, default 4), each with
 * its own stream and host thread, largest images first.  The lanes start
 * after the work already enqueued on ctx's stream; the call returns when
 * every image is finished.  Each image's results are bit-identical to a
 * dmsc_gpb call on it.  Images may differ in size; their outputs must not
 * overlap.  The first failing image stops the batch and its error is
 * returned. */
int dmsc_gpb_batch(dmsc_ctx* ctx, int count, const dmsc_batch_image* images, const dmsc_gpb_params* params,
                   int lanes);

#ifdef __cplusplus
}
#endif
#endif
