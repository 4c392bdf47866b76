/*
 * oracle.h -- CPU restatement of the Damascene gPb stage loops.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (damascene_amd/libdmsc.so) never links or calls it.
 *
 * Every function restates the reference's per-element operation order
 * (file:line cited at each definition in the .cpp files).  Floating-point
 * contract shared with the HIP kernels:
 *   - all code is compiled with -ffp-contract=off;
 *   - multiply-adds inside accumulations are explicit fmaf() where the
 *     definition says "fma", plain a*b+c otherwise;
 *   - reductions follow the reference's order (sequential, or the 32-slot
 *     tree of the histogram kernel);
 *   - transcendentals on the exact (integer-decision) path are either host
 *     tables built with glibc in double, or the project-defined cube root
 *     orc_cbrtf (basic IEEE operations only, bit-identical on CPU and GPU).
 *
 * Layouts: images are row-major float[h][w] (index y*w + x); cue stacks are
 * [24][N] planes (scale-major, then orientation); the affinity matrix is
 * DIA float[81][N] (diagonal-major); eigenvectors are [nvec][N].
 */
#ifndef DMSC_ORACLE_H
#define DMSC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tables (host code of the reference) ---- */
int  orc_texton_filters(float* coeffs /*9010*/, int* radii /*34*/);
void orc_gaussian_2d(float* filter, double sigma_x, double sigma_y, double ori,
                     unsigned deriv, int hilbert, unsigned long sx, unsigned long sy);
void orc_savgol_filter(float* filt, int d, float ra, float rb, float theta);
void orc_parabola_filters(int radius, float* out /* 8*(2r+1)^2 */);
void orc_spb_filters(float* out /* 8*49 */);
int  orc_rotation_tables(int width, int height, float thetaPi, int* newW, int* newH,
                         int* baseX, int* baseY, int* guideX, int* guideY);
void orc_halfdisc_offsets(int width, int height, float thetaPi, int radius,
                          const int* baseX, const int* baseY,
                          const int* guideX, const int* guideY, int rotatedWidth,
                          int* rectOffsets /*8*/);
int  orc_blur_kernel(float sigma, float* kernel /*<=64*/, int* radius);
int  orc_stencil(int radius, int width, int* offsets /*delta-coded*/, int* diagMap);
void orc_gamma_lut(float gamma, float* lut /*256*/);

/* ---- deterministic math shared by contract with the GPU ---- */
float orc_cbrtf(float x);
float orc_expf(float x);  /* project exp (basic IEEE operations), bit-identical on CPU and GPU */

/* ---- stages ---- */
void orc_rgb_to_grey(int w, int h, const uint32_t* rgb, float* grey);
void orc_rgb_to_lab(int w, int h, float gamma, const uint32_t* rgb, float* L, float* a, float* b);
void orc_normalize_lab(int w, int h, float* L, float* a, float* b);
void orc_mirror_f(int w, int h, int border, const float* in, float* out);
void orc_mirror_i(int w, int h, int border, const int* in, int* out);
void orc_quantize(int w, int h, int nbins, const float* in, int* out);

void orc_texton_convolve(int w, int h, const float* grey, float* responses /*[34][N]*/);
int  orc_kmeans(int w, int h, int k, int nfilters, const float* responses, int* labels,
                int max_iter, int* iterations);
int  orc_find_textons(int w, int h, const float* grey, int ntex /*32|64*/, int* labels, int* iterations);

void orc_integral_images(int W, int H, int nbins, const int* labels, int binPitch, int32_t* integrals);
void orc_gradient_pair(int w, int h, int border, int nbins, int blur, const float* blurKernel,
                       int kernelRadius, int rotatedWidth,
                       const int* baseX, const int* baseY, const int* guideX, const int* guideY,
                       const int* rectOffsets, const int32_t* integrals, int binPitch,
                       float* gradA, float* gradB);
void orc_channel_gradients(int w, int h, int nbins, int blur, float sigma, const int* radii3,
                           const int* labels /*unmirrored w*h*/, float* grads /*[3][8][N]*/);
void orc_parabola(int w, int h, int border, int radius, const float* grads8 /*[8][N]*/,
                  const float* filters /*8*(2r+1)^2*/, float* out8 /*[8][N]*/);
void orc_local_cues(int w, int h, const float* L, const float* a, const float* b,
                    const int* textons, int ntex, float* bg, float* cga, float* cgb, float* tg);
void orc_combine(int w, int h, const float* bg, const float* cga, const float* cgb,
                 const float* tg, int ntex, float* mpb /*[8][N]*/, float* combinedG /*[8][N]*/);
void orc_nonmax(int w, int h, const float* pb8, float* nmax);

void orc_intervene(int w, int h, const float* mpb, float* matrix /*[81][N]*/, float sigma);
void orc_convert_matrix(int w, int h, float* matrix, float* dsum, float* rsqrt_dsum);
void orc_spmv(int w, int h, const float* matrix, const float* x, float* y);
void orc_spmv_d(int w, int h, const float* matrix, const double* x, double* y);

/* tridiagonal helpers (FP64) */
int  orc_tridiag_eigvals(int n, const double* alpha, const double* beta, int il, int iu,
                         double* w);
int  orc_tridiag_eigvec(int n, const double* alpha, const double* beta, double lambda,
                        double* v, const double* prev, int nprev);

/* eigensolver: reference algorithm in FP32 (no reorthogonalisation, Cullum screening) */
int  orc_lanczos_fp32(int w, int h, const float* lsym, int nvec, float tol,
                      float* eigvals, float* eigvecs /*[nvec][N], u = D^1/2 v basis*/,
                      int* iterations);
/* the Cullum-Willoughby screening step of the above on a given T (size i+1) */
int  orc_cullum_screen(int i, const float* alpha, const float* beta, int eigCheck, double* vals);
/* FP64 oracle: Lanczos with full reorthogonalisation on the same L_sym */
int  orc_lanczos_fp64(int w, int h, const float* lsym, int nvec, double tol, int maxiter,
                      double* eigvals, double* eigvecs, int* iterations);
void orc_finish_eigvecs(int w, int h, int nvec, const float* rsqrt_dsum, float* eigvecs);
int  orc_generalized_eigensolve(int w, int h, const float* W /*[81][N]*/, int nvec, float tol,
                                float* eigvals, float* eigvecs, int* iterations);

void orc_spectral_pb(int w, int h, int nvec, const float* eigvals, const float* eigvecs,
                     float* spb /*[8][N]*/);
void orc_gpb(int w, int h, const float* combinedG, const float* spb, const float* nmax,
             float* gpbAll, float* gpb);

/* postprocess/skeleton.cu: one look-up-table pass (lut 0..7) over a 0/1
 * image, returns the number of cleared pixels; the whole PostProcess
 * (gpb replaced by its sigmoid, thin = sigmoid of gPb on the thinned mask);
 * NormalizeGpbAll in place. */
int  orc_skeleton_pass(int w, int h, int lut, const float* in, float* out);
int  orc_postprocess(int w, int h, float* gpb, const float* nmax, float* thin);
void orc_normalize_gpb_all(int npix, int norient, int pitch, float* all);

/* whole pipeline (plain FP32 reference order); returns 0 on success */
void orc_set_threads(int n);  /* OpenMP threads of the parallel loops */
/* Timing mode of the CPU baseline: the texton bank, parabola and spectral Pb
 * correlations run as register-tiled SIMD loops (orc_simd.cpp, the
 * technique of the reference's cpu_convolve2d.cc) with every output's
 * operation sequence unchanged, so results are identical to the default. */
void orc_set_simd(int on);
int  orc_simd_width(void);  /* columns per register tile: 32 (AVX-512) or 16 (AVX2) */
void orc_spmv_interior_simd(const float* M, int n, const float* x, float* y, int r0, int r1,
                            const int* lin, int nd);
void orc_correlate_simd(const float* in, int in_pitch, float* out, int out_pitch, int w, int h, int fh,
                        int fw, int order /* 0: filter rows outer, 1: filter columns outer */,
                        const float* f /* fh x fw, row-major */);
int  orc_max_threads(void);
int  orc_gpb_pipeline(int w, int h, const uint32_t* rgb, int nvec, float tol,
                      int* textons, float* mpb8, float* cg8, float* nmax, float* eigvals,
                      float* eigvecs, float* spb8, float* gpbAll, float* gpb, double* stage_ms,
                      int ntex);

#ifdef __cplusplus
}
#endif
#endif
