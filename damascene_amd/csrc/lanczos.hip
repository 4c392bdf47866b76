// lanczos.hip -- GPU-resident generalized eigensolver (reference:
// noReorthog/lanczos.cu, stencilMatrixMultiply/stencilMVM.cu).
//
// (D - W) v = lambda D v is solved as L_sym u = lambda u with
// L_sym = I - D^-1/2 W D^-1/2 and v = D^-1/2 u.  Two precisions:
//
//  FP32 ("plain", the reference algorithm): FP32 SpMV, vectors and inner
//    products, no reorthogonalisation, Ritz values checked every 500 steps
//    with Cullum-Willoughby screening of spurious copies (lanczos.cu:279-384).
//  MIXED: FP32 SpMV and basis, FP64 inner products, full classical
//    Gram-Schmidt reorthogonalisation twice per step (FP64 coefficients),
//    FP64 tridiagonal eigensolve; converged Ritz pairs are tested every 8
//    steps, no screening needed.
//
// One Lanczos step is three kernels with no host round trip: (1) stencil
// SpMV fused with the three-term subtraction and the q.z partial sums, (2)
// alpha from the partials (every block reduces them in the same fixed order,
// so alpha is identical everywhere and deterministic) fused with
// z -= alpha q and the |z|^2 partials, (3) beta and q_{k+1} = z / beta
// written into the basis.  Steps are enqueued in chunks captured in a HIP
// graph; alpha/beta stay on the device until a convergence check.
#include <algorithm>
#include <cmath>
#include <vector>

#include "dmsc_device.hpp"
#include "dmsc_internal.hpp"
#include "kernels.hpp"
#include "tridiag.hpp"

namespace dmsc {
namespace {

constexpr int kRedThreads = 256;

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s >= 1; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = sh[threadIdx.x] + sh[threadIdx.x + s];
    __syncthreads();
  }
  const T r = sh[0];
  __syncthreads();
  return r;
}

// Deterministic sum of nb partials (fixed order independent of the block).
template <typename T>
__device__ __forceinline__ T sum_partials(const T* __restrict__ part, int nb, T* sh) {
  T v = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) v = v + part[i];
  return block_sum(v, sh);
}

struct LzArgs {
  int n;
  size_t pitch;
  const float* A;
  float* V;        // basis, vector k at V + k * vstride
  size_t vstride;
  float* z;
  void* part;      // partial sums (T)
  void* part2;
  void* alpha;     // T[kMaxIter + 2]
  void* beta;
  int nb;          // number of SpMV blocks
};

// step 1: z = A q_k - beta_{k-1} q_{k-1}; partial q_k . z
template <typename T>
__global__ __launch_bounds__(kSpmvThreads) void k_lz_spmv(LzArgs a, StencilOffsets so, int k) {
  extern __shared__ float xs[];
  __shared__ T sh[kSpmvThreads];
  const float* q = a.V + (size_t)k * a.vstride;
  const int r0 = blockIdx.x * kSpmvRows;
  spmv_stage_x(xs, q, a.n, r0, so.halo, 1.0f);
  __syncthreads();
  const int row = r0 + threadIdx.x * 4;
  T dot = 0;
  if (row < a.n) {
    float4 acc;
    spmv_rows4(acc, a.A + row, a.pitch, xs, row - r0 + so.halo, so);
    float zz[4] = {acc.x, acc.y, acc.z, acc.w};
    if (k > 0) {
      const float nb = -(float)((const T*)a.beta)[k - 1];
      const float* qp = a.V + (size_t)(k - 1) * a.vstride + row;
      for (int j = 0; j < 4; j++) zz[j] = fmaf(nb, qp[j], zz[j]);
    }
    for (int j = 0; j < 4; j++) {
      if (row + j < a.n) {
        a.z[row + j] = zz[j];
        const float qj = xs[row - r0 + so.halo + j];
        dot = dot + (T)qj * (T)zz[j];
      }
    }
  }
  const T s = block_sum(dot, sh);
  if (threadIdx.x == 0) ((T*)a.part)[blockIdx.x] = s;
}

// step 2: alpha_k; z -= alpha_k q_k; partial |z|^2
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_lz_axpy(LzArgs a, int k) {
  __shared__ T sh[kRedThreads];
  const T alpha = sum_partials((const T*)a.part, a.nb, sh);
  if (blockIdx.x == 0 && threadIdx.x == 0) ((T*)a.alpha)[k] = alpha;
  const float na = -(float)alpha;
  const float* q = a.V + (size_t)k * a.vstride;
  T nrm = 0;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < a.n; i += gridDim.x * kRedThreads) {
    const float zi = fmaf(na, q[i], a.z[i]);
    a.z[i] = zi;
    nrm = nrm + (T)zi * (T)zi;
  }
  const T s = block_sum(nrm, sh);
  if (threadIdx.x == 0) ((T*)a.part2)[blockIdx.x] = s;
}

// step 3: beta_k = |z|; q_{k+1} = z * (1 / beta_k)
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_lz_scale(LzArgs a, int k, int nb2) {
  __shared__ T sh[kRedThreads];
  const T beta = sqrt(sum_partials((const T*)a.part2, nb2, sh));
  if (blockIdx.x == 0 && threadIdx.x == 0) ((T*)a.beta)[k] = beta;
  const float inv = 1 / (float)beta;
  float* qn = a.V + (size_t)(k + 1) * a.vstride;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < a.n; i += gridDim.x * kRedThreads)
    qn[i] = a.z[i] * inv;
}

// ---- full reorthogonalisation (mixed): c = V[0..k]^T z (FP64), z -= V c
__global__ __launch_bounds__(kRedThreads) void k_ro_dots(int n, const float* __restrict__ V, size_t vstride,
                                                         int nv, const float* __restrict__ z,
                                                         double* __restrict__ part, int nchunk) {
  // block (chunk c, vector j): partial dot of vector j with z over chunk c
  __shared__ double sh[kRedThreads];
  const int j = blockIdx.y, c = blockIdx.x;
  const int len = (n + nchunk - 1) / nchunk, i0 = c * len, i1 = min(n, i0 + len);
  const float* v = V + (size_t)j * vstride;
  double s = 0;
  for (int i = i0 + threadIdx.x; i < i1; i += kRedThreads) s += (double)v[i] * (double)z[i];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) part[(size_t)j * nchunk + c] = s;
  (void)nv;
}

__global__ __launch_bounds__(kRedThreads) void k_ro_reduce(const double* __restrict__ part, int nchunk,
                                                           double* __restrict__ coef, double* alpha_acc, int k) {
  __shared__ double sh[kRedThreads];
  const int j = blockIdx.x;
  const double s = sum_partials(part + (size_t)j * nchunk, nchunk, sh);
  if (threadIdx.x == 0) {
    coef[j] = s;
    if (j == k && alpha_acc) alpha_acc[k] += s;
  }
}

__global__ __launch_bounds__(kRedThreads) void k_ro_update(int n, const float* __restrict__ V, size_t vstride,
                                                           int nv, const double* __restrict__ coef,
                                                           float* __restrict__ z, double* __restrict__ part2) {
  __shared__ double sh[kRedThreads];
  double nrm = 0;
  for (int i = blockIdx.x * kRedThreads + threadIdx.x; i < n; i += gridDim.x * kRedThreads) {
    double acc = (double)z[i];
    for (int j = 0; j < nv; j++) acc -= coef[j] * (double)V[(size_t)j * vstride + i];
    const float zi = (float)acc;
    z[i] = zi;
    nrm += (double)zi * (double)zi;
  }
  nrm = block_sum(nrm, sh);
  if (threadIdx.x == 0 && part2) part2[blockIdx.x] = nrm;
}

// ---- Ritz vectors, scaling, min-max normalisation
__global__ __launch_bounds__(256) void k_ritz(int n, const float* __restrict__ V, size_t vstride, int m,
                                              const float* __restrict__ S, int nvec, float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float acc[32];
  for (int j = 0; j < nvec; j++) acc[j] = 0.f;
  for (int k = 0; k < m; k++) {
    const float v = V[(size_t)k * vstride + p];
    for (int j = 0; j < nvec; j++) acc[j] = fmaf(v, S[(size_t)j * m + k], acc[j]);
  }
  for (int j = 0; j < nvec; j++) out[(size_t)j * n + p] = acc[j];
}

// stencilMVM.cu:182-198 scaleEigByD
__global__ void k_scale_rs(int n, int nvec, const float* __restrict__ rs, float* __restrict__ v) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int j = 0; j < nvec; j++) v[(size_t)j * n + p] = v[(size_t)j * n + p] * rs[p];
}

// lanczos.cu:735-885 FindMaxMinPerBlock/Grid (min and max are exact, so any
// reduction order gives the same result).
__global__ __launch_bounds__(256) void k_minmax(int n, const float* __restrict__ v, float* __restrict__ mm,
                                                int nblk) {
  __shared__ float smin[256], smax[256];
  const int j = blockIdx.y;
  const float* x = v + (size_t)j * n;
  float lo = INFINITY, hi = -INFINITY;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    lo = fminf(lo, x[i]);
    hi = fmaxf(hi, x[i]);
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int s = 128; s >= 1; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
      smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    mm[((size_t)j * nblk + blockIdx.x) * 2] = smin[0];
    mm[((size_t)j * nblk + blockIdx.x) * 2 + 1] = smax[0];
  }
}

// lanczos.cu:887-901 NormalizationDev (vectors 1..nvec-1)
__global__ void k_normalize(int n, int nvec, const float* __restrict__ mm, int nblk, float* __restrict__ v) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int j = 1; j < nvec; j++) {
    float lo = mm[((size_t)j * nblk) * 2], hi = mm[((size_t)j * nblk) * 2 + 1];
    for (int b = 1; b < nblk; b++) {
      lo = fminf(lo, mm[((size_t)j * nblk + b) * 2]);
      hi = fmaxf(hi, mm[((size_t)j * nblk + b) * 2 + 1]);
    }
    if ((hi - lo) > 1e-4f) {
      float* x = v + (size_t)j * n + p;
      *x = (*x - lo) / (hi - lo);
    }
  }
}

__global__ void k_fill(float* p, int n, float v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// --------------------------------------------------------------- host side

// lanczos.cu:279-384 CullumDevice (FP64): Ritz values of T_{i+1} and of T
// without its first row/column; lump values closer than 1e-5, drop
// non-duplicates that are also eigenvalues of the reduced matrix (relative
// distance <= 1e-10); Ritz vectors of the first nvec survivors.
bool cullum(int i, const std::vector<double>& alpha, const std::vector<double>& beta, int eigCheck, int nvec,
            std::vector<double>& vals, std::vector<double>& S) {
  std::vector<double> a(i + 1), b(i + 1), wc, wr;
  const int tn_hat = i;
  for (int j = 0; j < tn_hat; j++) a[j] = alpha[j + 1];
  for (int j = 0; j < tn_hat - 1; j++) b[j] = beta[j + 1];
  const int ech = std::min(eigCheck, tn_hat);
  tri::Bounds g = tri::gershgorin(tn_hat, a.data(), b.data());
  for (int k = 1; k <= ech; k++) wc.push_back(tri::kth(tn_hat, a.data(), b.data(), k, g));
  const int tn = i + 1;
  for (int j = 0; j < tn; j++) a[j] = alpha[j];
  for (int j = 0; j < tn - 1; j++) b[j] = beta[j];
  g = tri::gershgorin(tn, a.data(), b.data());
  const int ec = std::min(eigCheck, tn);
  for (int k = 1; k <= ec; k++) wr.push_back(tri::kth(tn, a.data(), b.data(), k, g));
  std::vector<double> acc;
  std::vector<char> dup;
  for (int j = 0; j < ec; j++) {
    if (j > 0 && wr[j] - wr[j - 1] < 1e-5) {
      dup.back() = 1;
      continue;
    }
    acc.push_back(wr[j]);
    dup.push_back(0);
  }
  std::vector<double> scr;
  for (size_t j = 0; j < acc.size(); j++) {
    bool ok = true;
    if (!dup[j])
      for (double c : wc)
        if (std::fabs((acc[j] - c) / acc[j]) <= 1e-10) ok = false;
    if (ok) scr.push_back(acc[j]);
  }
  if ((int)scr.size() < nvec) return false;
  vals.assign(scr.begin(), scr.begin() + nvec);
  S.assign((size_t)nvec * tn, 0.0);
  std::vector<double> work(5 * (size_t)tn);
  std::vector<int> piv(tn);
  int cstart = 0;
  for (int j = 0; j < nvec; j++) {
    if (j > 0 && vals[j] - vals[j - 1] > 1e-3 * g.tnorm) cstart = j;
    if (tri::invit(tn, a.data(), b.data(), vals[j], &S[(size_t)j * tn], &S[(size_t)cstart * tn], j - cstart,
                   work.data(), piv.data()))
      return false;
  }
  return true;
}

// Eigenpairs of T_m (no screening; the mixed path keeps the basis orthogonal).
void ritz_plain(int m, const std::vector<double>& alpha, const std::vector<double>& beta, int nvec,
                std::vector<double>& vals, std::vector<double>& S) {
  std::vector<double> a(alpha.begin(), alpha.begin() + m), b(beta.begin(), beta.begin() + std::max(m - 1, 0));
  b.resize(std::max(m, 1));
  const tri::Bounds g = tri::gershgorin(m, a.data(), b.data());
  vals.resize(nvec);
  for (int k = 1; k <= nvec; k++) vals[k - 1] = tri::kth(m, a.data(), b.data(), k, g);
  S.assign((size_t)nvec * m, 0.0);
  std::vector<double> work(5 * (size_t)m);
  std::vector<int> piv(m);
  int cstart = 0;
  for (int j = 0; j < nvec; j++) {
    if (j > 0 && vals[j] - vals[j - 1] > 1e-3 * g.tnorm) cstart = j;
    tri::invit(m, a.data(), b.data(), vals[j], &S[(size_t)j * m], &S[(size_t)cstart * m], j - cstart,
               work.data(), piv.data());
  }
}

struct Workspace {
  float* V = nullptr;
  size_t cap = 0;  // vectors
  float* z = nullptr;
  float* rs = nullptr;
  void* part = nullptr;
  void* part2 = nullptr;
  void* alpha = nullptr;
  void* beta = nullptr;
  double* coef = nullptr;
  double* ropart = nullptr;
  float* S = nullptr;
  float* mm = nullptr;
};

}  // namespace

int run_eigensolve(Ctx* ctx, int w, int h, float* A, int pitch, int nvec, float tol, int precision,
                   float* h_eigvals, float* d_eigvecs, int* iterations) {
  hipStream_t s = ctx->stream;
  const int n = w * h;
  const bool mixed = precision == DMSC_PREC_MIXED;
  const size_t vstride = (size_t)pitch_floats(n);
  const int nb = div_up(n, kSpmvRows);
  const int nb2 = std::min(div_up(n, kRedThreads), 1024);
  const int nchunk = std::min(64, std::max(1, div_up(n, 8192)));
  StencilOffsets so = make_stencil_offsets(w);
  const size_t lds = sizeof(float) * (kSpmvRows + 2 * so.halo);
  const size_t tsz = mixed ? sizeof(double) : sizeof(float);
  const int cap0 = mixed ? 512 : kCheckIter + 2;

  // scratch: z, rs, partials, alpha/beta, coef, ropart, S, minmax
  Workspace ws;
  const size_t small = vstride * 2 * sizeof(float) + tsz * (size_t)(nb + nb2 + 2 * (kMaxIter + 2)) +
                       sizeof(double) * ((size_t)(kMaxIter + 2) * (nchunk + 1)) + sizeof(float) * 32 * (kMaxIter + 2) +
                       sizeof(float) * 2 * 32 * 1024 + 64 * 16;
  void* sm;
  DMSC_TRY(ctx->scratch(2, small, &sm));
  char* cur = (char*)sm;
  auto take = [&](size_t bytes) {
    void* p = cur;
    cur += (bytes + 255) & ~(size_t)255;
    return p;
  };
  ws.z = (float*)take(vstride * sizeof(float));
  ws.rs = (float*)take(vstride * sizeof(float));
  ws.part = take(tsz * nb);
  ws.part2 = take(tsz * nb2);
  ws.alpha = take(tsz * (kMaxIter + 2));
  ws.beta = take(tsz * (kMaxIter + 2));
  ws.coef = (double*)take(sizeof(double) * (kMaxIter + 2));
  ws.ropart = (double*)take(sizeof(double) * (size_t)(kMaxIter + 2) * nchunk);
  ws.S = (float*)take(sizeof(float) * 32 * (kMaxIter + 2));
  ws.mm = (float*)take(sizeof(float) * 2 * 32 * 1024);
  // basis (grown on demand, kept on the context)
  size_t cap = cap0;
  void* vb;
  DMSC_TRY(ctx->scratch(3, vstride * sizeof(float) * cap, &vb));
  ws.V = (float*)vb;

  DMSC_TRY(launch_convert_matrix(s, w, h, A, pitch, nullptr, ws.rs));
  const float q0 = (float)(1 / std::sqrt((double)n));  // lanczos.cu:64-72
  hipLaunchKernelGGL(k_fill, dim3(div_up(n, 256)), dim3(256), 0, s, ws.V, n, q0);
  if (mixed) DMSC_HIP(hipMemsetAsync(ws.alpha, 0, tsz * (kMaxIter + 2), s));

  LzArgs la{n, (size_t)pitch, A, ws.V, vstride, ws.z, ws.part, ws.part2, ws.alpha, ws.beta, nb};
  std::vector<double> alpha, beta, vals, S;
  int eigCheck = nvec + 20;
  int m = 0;
  bool converged = false;
  const int period = mixed ? 8 : kCheckIter;

  for (int k = 0; k < kMaxIter && !converged; k++) {
    if ((size_t)k + 2 > cap) {  // grow the basis, keeping its contents
      const size_t ncap = std::min<size_t>(cap * 2, kMaxIter + 2);
      void* nbuf;
      DMSC_HIP(hipMalloc(&nbuf, vstride * sizeof(float) * ncap));
      DMSC_HIP(hipMemcpyAsync(nbuf, ws.V, vstride * sizeof(float) * cap, hipMemcpyDeviceToDevice, s));
      DMSC_HIP(hipStreamSynchronize(s));
      DMSC_TRY(ctx->adopt(3, nbuf, vstride * sizeof(float) * ncap));
      ws.V = (float*)nbuf;
      la.V = ws.V;
      cap = ncap;
    }
    if (mixed) {
      hipLaunchKernelGGL(k_lz_spmv<double>, dim3(nb), dim3(kSpmvThreads), lds, s, la, so, k);
      hipLaunchKernelGGL(k_lz_axpy<double>, dim3(nb2), dim3(kRedThreads), 0, s, la, k);
      // two passes of classical Gram-Schmidt against q_0..q_k
      for (int pass = 0; pass < 2; pass++) {
        hipLaunchKernelGGL(k_ro_dots, dim3(nchunk, k + 1), dim3(kRedThreads), 0, s, n, ws.V, vstride, k + 1,
                           ws.z, ws.ropart, nchunk);
        hipLaunchKernelGGL(k_ro_reduce, dim3(k + 1), dim3(kRedThreads), 0, s, ws.ropart, nchunk, ws.coef,
                           (double*)ws.alpha, k);
        hipLaunchKernelGGL(k_ro_update, dim3(nb2), dim3(kRedThreads), 0, s, n, ws.V, vstride, k + 1, ws.coef,
                           ws.z, (double*)ws.part2);
      }
      hipLaunchKernelGGL(k_lz_scale<double>, dim3(nb2), dim3(kRedThreads), 0, s, la, k, nb2);
    } else {
      hipLaunchKernelGGL(k_lz_spmv<float>, dim3(nb), dim3(kSpmvThreads), lds, s, la, so, k);
      hipLaunchKernelGGL(k_lz_axpy<float>, dim3(nb2), dim3(kRedThreads), 0, s, la, k);
      hipLaunchKernelGGL(k_lz_scale<float>, dim3(nb2), dim3(kRedThreads), 0, s, la, k, nb2);
    }
    DMSC_HIP(hipGetLastError());
    const bool check = mixed ? ((k + 1) % period == 0 && k + 1 >= nvec + 2) || k == kMaxIter - 1
                             : ((k + 1) % kCheckIter == 0) || k == kMaxIter - 1;
    if (!check) continue;
    // fetch alpha/beta (FP32 path: float, mixed: double)
    alpha.assign(k + 1, 0.0);
    beta.assign(k + 1, 0.0);
    if (mixed) {
      DMSC_HIP(hipMemcpyAsync(alpha.data(), ws.alpha, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, s));
      DMSC_HIP(hipMemcpyAsync(beta.data(), ws.beta, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, s));
      DMSC_HIP(hipStreamSynchronize(s));
    } else {
      std::vector<float> fa(k + 1), fb(k + 1);
      DMSC_HIP(hipMemcpyAsync(fa.data(), ws.alpha, sizeof(float) * (k + 1), hipMemcpyDeviceToHost, s));
      DMSC_HIP(hipMemcpyAsync(fb.data(), ws.beta, sizeof(float) * (k + 1), hipMemcpyDeviceToHost, s));
      DMSC_HIP(hipStreamSynchronize(s));
      for (int j = 0; j <= k; j++) { alpha[j] = fa[j]; beta[j] = fb[j]; }
    }
    m = k + 1;
    bool ok;
    if (mixed) {
      ritz_plain(m, alpha, beta, nvec, vals, S);
      ok = true;
    } else {
      ok = false;
      while (!ok) {
        ok = cullum(k, alpha, beta, eigCheck, nvec, vals, S);
        if (!ok) {
          eigCheck += 20;  // lanczos.cu:528-532
          if (eigCheck > 1000) return fail(DMSC_ERR_NOCONV, "Cullum screening kept fewer than %d values", nvec);
        }
      }
    }
    // lanczos.cu:114-140 TestForConvergence
    bool good = true;
    for (int j = nvec - 1; j >= 0; j--) {
      const double res = mixed ? std::fabs(beta[k] * S[(size_t)j * m + k])
                               : std::fabs((float)beta[k] * (float)S[(size_t)j * m + k]);
      if (res > tol) { good = false; break; }
    }
    converged = good;
  }
  if (!converged) return fail(DMSC_ERR_NOCONV, "Lanczos did not converge in %d iterations", kMaxIter);

  // Ritz vectors u_j = V s_j (lanczos.cu:249-270), D^-1/2 scaling, min-max
  std::vector<float> Sf(S.size());
  for (size_t i = 0; i < S.size(); i++) Sf[i] = (float)S[i];
  DMSC_HIP(hipMemcpyAsync(ws.S, Sf.data(), sizeof(float) * Sf.size(), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_ritz, dim3(div_up(n, 256)), dim3(256), 0, s, n, ws.V, vstride, m, ws.S, nvec, d_eigvecs);
  hipLaunchKernelGGL(k_scale_rs, dim3(div_up(n, 256)), dim3(256), 0, s, n, nvec, ws.rs, d_eigvecs);
  const int nblk = std::min(div_up(n, 256), 1024);
  hipLaunchKernelGGL(k_minmax, dim3(nblk, nvec), dim3(256), 0, s, n, d_eigvecs, ws.mm, nblk);
  hipLaunchKernelGGL(k_normalize, dim3(div_up(n, 256)), dim3(256), 0, s, n, nvec, ws.mm, nblk, d_eigvecs);
  DMSC_HIP(hipGetLastError());
  for (int j = 0; j < nvec; j++) h_eigvals[j] = (float)vals[j];
  if (iterations) *iterations = m;
  DMSC_HIP(hipStreamSynchronize(s));
  return DMSC_OK;
}

}  // namespace dmsc
