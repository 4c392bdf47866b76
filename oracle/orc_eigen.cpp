// orc_eigen.cpp -- oracle restatement of the generalized eigensolve
// (noReorthog/lanczos.cu) plus an FP64 reference eigensolver.
// TEST INFRASTRUCTURE ONLY.
//
// The tridiagonal eigenproblems that the reference hands to ACML
// (dstebz: bisection on Sturm counts; dstein: inverse iteration) are solved
// here by the textbook routines behind those LAPACK calls.
#include "oracle.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

// Number of eigenvalues of T strictly less than x (Sturm sequence).
int sturm_count(int n, const double* a, const double* b, double x, double pivmin) {
  int cnt = 0;
  double q = a[0] - x;
  if (std::fabs(q) < pivmin) q = -pivmin;
  if (q < 0) cnt++;
  for (int i = 1; i < n; i++) {
    q = a[i] - x - (b[i - 1] * b[i - 1]) / q;
    if (std::fabs(q) < pivmin) q = -pivmin;
    if (q < 0) cnt++;
  }
  return cnt;
}

}  // namespace

extern "C" {

// Eigenvalues il..iu (1-based, ascending) of the symmetric tridiagonal
// matrix (alpha, beta) by bisection to full double precision (the algorithm
// of LAPACK dstebz with RANGE='I', ORDER='E', ABSTOL=0).
int orc_tridiag_eigvals(int n, const double* a, const double* b, int il, int iu, double* w) {
  if (n <= 0 || il < 1 || iu > n || il > iu) return -1;
  double lo = a[0], hi = a[0], bmax = 0;
  for (int i = 0; i < n; i++) {
    double r = 0;
    if (i > 0) r += std::fabs(b[i - 1]);
    if (i < n - 1) r += std::fabs(b[i]);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
    if (i < n - 1) bmax = std::max(bmax, b[i] * b[i]);
  }
  double tnorm = std::max(std::fabs(lo), std::fabs(hi));
  double pivmin = std::max(2.2250738585072014e-308 * std::max(1.0, bmax), 1e-300);
  lo -= 2.2e-16 * tnorm * n + 2 * pivmin;
  hi += 2.2e-16 * tnorm * n + 2 * pivmin;
  for (int k = il; k <= iu; k++) {
    double l = lo, h = hi;
    for (int it = 0; it < 200; it++) {
      double mid = 0.5 * (l + h);
      if (mid <= l || mid >= h) break;
      if (h - l <= 2.2e-16 * 2 * std::max(std::fabs(l), std::fabs(h)) + pivmin) break;
      if (sturm_count(n, a, b, mid, pivmin) >= k) h = mid;
      else l = mid;
    }
    w[k - il] = 0.5 * (l + h);
  }
  return 0;
}

// Eigenvector for eigenvalue lambda by inverse iteration (the method of
// LAPACK dstein): LU of T - lambda I with partial pivoting, deterministic
// start, re-orthogonalised against `prev` (vectors of nearby eigenvalues).
// Normalised to unit length with its largest-magnitude entry positive.
int orc_tridiag_eigvec(int n, const double* a, const double* b, double lambda, double* v,
                       const double* prev, int nprev) {
  std::vector<double> dl(n), d(n), du(n), du2(n), x(n);
  std::vector<int> piv(n);
  double tnorm = 0;
  for (int i = 0; i < n; i++)
    tnorm = std::max(tnorm, std::fabs(a[i]) + (i > 0 ? std::fabs(b[i - 1]) : 0) +
                                (i < n - 1 ? std::fabs(b[i]) : 0));
  double eps = 2.2e-16 * std::max(tnorm, 1e-300);
  // factor T - lambda I (dgttrf-style)
  for (int i = 0; i < n; i++) d[i] = a[i] - lambda;
  for (int i = 0; i < n - 1; i++) { dl[i] = b[i]; du[i] = b[i]; }
  for (int i = 0; i < n - 1; i++) du2[i] = 0;
  for (int i = 0; i < n - 1; i++) {
    if (std::fabs(d[i]) >= std::fabs(dl[i])) {
      piv[i] = i;
      if (d[i] == 0) d[i] = eps;
      double f = dl[i] / d[i];
      dl[i] = f;
      d[i + 1] -= f * du[i];
    } else {
      piv[i] = i + 1;
      double f = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = f;
      double t = du[i];
      du[i] = d[i + 1];
      d[i + 1] = t - f * d[i + 1];
      if (i < n - 2) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
    }
  }
  piv[n - 1] = n - 1;
  if (d[n - 1] == 0) d[n - 1] = eps;
  for (int i = 0; i < n; i++) if (std::fabs(d[i]) < eps) d[i] = (d[i] < 0 ? -eps : eps);
  // deterministic start vector
  uint64_t s = 0x9E3779B97F4A7C15ULL;
  for (int i = 0; i < n; i++) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    x[i] = ((double)(s >> 11) / 9007199254740992.0) - 0.5;
  }
  for (int iter = 0; iter < 6; iter++) {
    for (int j = 0; j < nprev; j++) {
      const double* p = prev + (size_t)j * n;
      double dot = 0;
      for (int i = 0; i < n; i++) dot += p[i] * x[i];
      for (int i = 0; i < n; i++) x[i] -= dot * p[i];
    }
    // solve L U y = P x
    for (int i = 0; i < n - 1; i++) {
      if (piv[i] == i) {
        x[i + 1] -= dl[i] * x[i];
      } else {
        double t = x[i];
        x[i] = x[i + 1];
        x[i + 1] = t - dl[i] * x[i];
      }
    }
    x[n - 1] /= d[n - 1];
    if (n > 1) x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; i--) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / d[i];
    double nrm = 0;
    for (int i = 0; i < n; i++) nrm += x[i] * x[i];
    nrm = std::sqrt(nrm);
    if (!(nrm > 0) || !std::isfinite(nrm)) return -1;
    for (int i = 0; i < n; i++) x[i] /= nrm;
  }
  for (int j = 0; j < nprev; j++) {
    const double* p = prev + (size_t)j * n;
    double dot = 0;
    for (int i = 0; i < n; i++) dot += p[i] * x[i];
    for (int i = 0; i < n; i++) x[i] -= dot * p[i];
  }
  double nrm = 0, big = 0;
  int jmax = 0;
  for (int i = 0; i < n; i++) {
    nrm += x[i] * x[i];
    if (std::fabs(x[i]) > big) { big = std::fabs(x[i]); jmax = i; }
  }
  nrm = std::sqrt(nrm);
  double scl = (x[jmax] < 0 ? -1.0 : 1.0) / nrm;
  for (int i = 0; i < n; i++) v[i] = x[i] * scl;
  return 0;
}

}  // extern "C"

namespace {

// Eigenvectors for a sorted list of eigenvalues; vectors whose eigenvalues
// lie within 1e-3*|T| of each other are orthogonalised against each other.
int tridiag_vectors(int n, const double* a, const double* b, const double* lam, int m, double* V) {
  double tnorm = 0;
  for (int i = 0; i < n; i++) tnorm = std::max(tnorm, std::fabs(a[i]) + (i ? std::fabs(b[i - 1]) : 0));
  int cstart = 0;
  for (int j = 0; j < m; j++) {
    if (j > 0 && lam[j] - lam[j - 1] > 1e-3 * tnorm) cstart = j;
    int rc = orc_tridiag_eigvec(n, a, b, lam[j], V + (size_t)j * n, V + (size_t)cstart * n, j - cstart);
    if (rc) return rc;
  }
  return 0;
}

// lanczos.cu:279-384 CullumDevice: Ritz values of T (size i+1) and of T-hat
// (first row/column removed); merge values closer than LUMPTOL=1e-5, drop
// non-duplicate values that are also eigenvalues of T-hat (relative distance
// <= 1e-10); take the first nEig survivors and their Ritz vectors.
bool screen(int i, const float* alpha, const float* beta, int eigCheck, std::vector<double>& scr) {
  int tn = i;
  std::vector<double> ta(i + 1), tb(i + 1), wc(eigCheck), wr(eigCheck);
  for (int j = 0; j < tn; j++) ta[j] = alpha[j + 1];
  for (int j = 0; j < tn - 1; j++) tb[j] = beta[j + 1];
  if (orc_tridiag_eigvals(tn, ta.data(), tb.data(), 1, std::min(eigCheck, tn), wc.data())) return false;
  tn = i + 1;
  for (int j = 0; j < tn; j++) ta[j] = alpha[j];
  for (int j = 0; j < tn - 1; j++) tb[j] = beta[j];
  int ec = std::min(eigCheck, tn);
  if (orc_tridiag_eigvals(tn, ta.data(), tb.data(), 1, ec, wr.data())) return false;
  std::vector<double> acc;
  std::vector<char> dup;
  for (int j = 0; j < ec; j++) {
    bool accept = true;
    if (j > 0 && wr[j] - wr[j - 1] < 1e-5) {
      accept = false;
      dup.back() = 1;
    }
    if (accept) { acc.push_back(wr[j]); dup.push_back(0); }
  }
  scr.clear();
  for (size_t j = 0; j < acc.size(); j++) {
    bool accept = true;
    if (!dup[j])
      for (int k = 0; k < std::min(eigCheck, i); k++)
        if (std::fabs((acc[j] - wc[k]) / acc[j]) <= 1e-10) accept = false;
    if (accept) scr.push_back(acc[j]);
  }
  return true;
}

bool cullum(int i, const float* alpha, const float* beta, int eigCheck, int nEig, double* vals,
            std::vector<double>& S) {
  std::vector<double> scr;
  if (!screen(i, alpha, beta, eigCheck, scr)) return false;
  if ((int)scr.size() < nEig) return false;
  for (int j = 0; j < nEig; j++) vals[j] = scr[j];
  const int tn = i + 1;
  std::vector<double> ta(tn), tb(tn, 0.0);
  for (int j = 0; j < tn; j++) ta[j] = alpha[j];
  for (int j = 0; j < tn - 1; j++) tb[j] = beta[j];
  S.assign((size_t)nEig * tn, 0.0);
  return tridiag_vectors(tn, ta.data(), tb.data(), vals, nEig, S.data()) == 0;
}

}  // namespace

extern "C" {

// lanczos.cu:279-384 CullumDevice's screening alone: T = tridiag(alpha[0..i],
// beta[0..i-1]) and T-hat = T without its first row and column; the
// screened Ritz values of the first eigCheck values of T, ascending, into
// vals (at most eigCheck); returns their number, or -1 if a tridiagonal
// eigenvalue solve failed.
int orc_cullum_screen(int i, const float* alpha, const float* beta, int eigCheck, double* vals) {
  std::vector<double> scr;
  if (!screen(i, alpha, beta, eigCheck, scr)) return -1;
  for (size_t j = 0; j < scr.size(); j++) vals[j] = scr[j];
  return (int)scr.size();
}

// lanczos.cu:386-733 lanczos (no reorthogonalisation) on L_sym.  FP32
// vectors and SpMV; inner products accumulate in double and are rounded to
// float (a more accurate stand-in for the cuBLAS Sdot/Snrm2 of lanczos.cu:
// 234, 241).  Convergence is tested every 500 iterations (CHECKITER) with
// |beta_i s_{i,j}| <= tol for the first nvec Ritz pairs.  Output: Ritz
// values and Ritz vectors u_j = V s_j (before the D^-1/2 scaling).
int orc_lanczos_fp32(int w, int h, const float* A, int nvec, float tol, float* eigvals,
                     float* eigvecs, int* iterations) {
  const int n = w * h, MAXITER = 6000;
  std::vector<std::vector<float>> V;
  std::vector<float> alpha(MAXITER + 1), beta(MAXITER + 1), q(n), z(n);
  float q0 = (float)(1 / std::sqrt((double)n));
  for (int k = 0; k < n; k++) q[k] = q0;
  V.push_back(q);
  int eigCheck = nvec + 20, i = 0;
  std::vector<double> vals(nvec), S;
  bool converged = false;
  for (int iter = 0; iter < MAXITER; iter++) {
    i = iter;
    orc_spmv(w, h, A, q.data(), z.data());
    if (iter > 0) {
      const float nb = -beta[iter - 1];
      const float* vp = V[iter - 1].data();
      for (int k = 0; k < n; k++) z[k] = fmaf(nb, vp[k], z[k]);
    }
    double dot = 0;
    for (int k = 0; k < n; k++) dot += (double)q[k] * z[k];
    alpha[iter] = (float)dot;
    const float na = -alpha[iter];
    for (int k = 0; k < n; k++) z[k] = fmaf(na, q[k], z[k]);
    double nr = 0;
    for (int k = 0; k < n; k++) nr += (double)z[k] * z[k];
    beta[iter] = (float)std::sqrt(nr);
    float inv = 1 / beta[iter];
    for (int k = 0; k < n; k++) q[k] = z[k] * inv;
    V.push_back(q);
    if (((i + 1) % 500 == 0) || (i == MAXITER - 1)) {
      bool ok = false;
      while (!ok) {
        ok = cullum(i, alpha.data(), beta.data(), eigCheck, nvec, vals.data(), S);
        if (!ok) {
          eigCheck += 20;
          if (eigCheck > 1000) return -2;
        }
      }
      bool good = true;
      for (int j = nvec - 1; j >= 0; j--)
        if (std::fabs(beta[i] * (float)S[(size_t)j * (i + 1) + i]) > tol) { good = false; break; }
      if (good) { converged = true; break; }
    }
  }
  const int m = i + 1;
  for (int j = 0; j < nvec; j++) eigvals[j] = (float)vals[j];
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++)
    for (int j = 0; j < nvec; j++) {
      double acc = 0;
      for (int t = 0; t < m; t++) acc += (double)V[t][k] * (float)S[(size_t)j * m + t];
      eigvecs[(size_t)j * n + k] = (float)acc;
    }
  if (iterations) *iterations = m;
  return converged ? 0 : 1;
}

// FP64 reference: Lanczos with full (two-pass classical Gram-Schmidt)
// reorthogonalisation, double vectors, on the float L_sym entries; stops
// when the first nvec Ritz residuals |beta_m s_{m,j}| are below tol.
int orc_lanczos_fp64(int w, int h, const float* A, int nvec, double tol, int maxiter,
                     double* eigvals, double* eigvecs, int* iterations) {
  const int n = w * h;
  std::vector<double> Q;  // m x n
  std::vector<double> alpha, beta, q(n, 1.0 / std::sqrt((double)n)), z(n), coef;
  Q.insert(Q.end(), q.begin(), q.end());
  int m = 0;
  std::vector<double> vals(nvec), S;
  bool conv = false;
  for (int it = 0; it < maxiter; it++) {
    orc_spmv_d(w, h, A, q.data(), z.data());
    int k = it + 1;  // number of basis vectors
    for (int pass = 0; pass < 2; pass++) {
      coef.assign(k, 0.0);
#pragma omp parallel for schedule(static)
      for (int j = 0; j < k; j++) {
        const double* qj = &Q[(size_t)j * n];
        double d = 0;
        for (int t = 0; t < n; t++) d += qj[t] * z[t];
        coef[j] = d;
      }
      if (pass == 0) alpha.push_back(coef[it]);
      else alpha.back() += coef[it];
#pragma omp parallel for schedule(static)
      for (int t = 0; t < n; t++) {
        double s = z[t];
        for (int j = 0; j < k; j++) s -= coef[j] * Q[(size_t)j * n + t];
        z[t] = s;
      }
    }
    double nr = 0;
    for (int t = 0; t < n; t++) nr += z[t] * z[t];
    nr = std::sqrt(nr);
    beta.push_back(nr);
    m = it + 1;
    if (m >= nvec + 2 && (m % 25 == 0 || it == maxiter - 1 || nr < 1e-14)) {
      std::vector<double> b(beta.begin(), beta.end() - 1);
      orc_tridiag_eigvals(m, alpha.data(), b.data(), 1, nvec, vals.data());
      S.assign((size_t)nvec * m, 0.0);
      tridiag_vectors(m, alpha.data(), b.data(), vals.data(), nvec, S.data());
      bool good = true;
      for (int j = 0; j < nvec; j++)
        if (std::fabs(nr * S[(size_t)j * m + m - 1]) > tol) good = false;
      if (good) { conv = true; break; }
    }
    if (nr < 1e-300) break;
    for (int t = 0; t < n; t++) q[t] = z[t] / nr;
    Q.insert(Q.end(), q.begin(), q.end());
  }
  for (int j = 0; j < nvec; j++) eigvals[j] = vals[j];
#pragma omp parallel for schedule(static)
  for (int t = 0; t < n; t++)
    for (int j = 0; j < nvec; j++) {
      double acc = 0;
      for (int k = 0; k < m; k++) acc += Q[(size_t)k * n + t] * S[(size_t)j * m + k];
      eigvecs[(size_t)j * n + t] = acc;
    }
  if (iterations) *iterations = m;
  return conv ? 0 : 1;
}

// stencilMVM.cu:182-198 scaleEigByD (v = D^-1/2 u) then lanczos.cu:735-968
// NormalizeEigVecDev: vectors 1..nvec-1 min-max normalised when the range
// exceeds 1e-4.
void orc_finish_eigvecs(int w, int h, int nvec, const float* rs, float* V) {
  const int n = w * h;
  for (int j = 0; j < nvec; j++)
    for (int k = 0; k < n; k++) V[(size_t)j * n + k] = V[(size_t)j * n + k] * rs[k];
  for (int j = 1; j < nvec; j++) {
    float* v = V + (size_t)j * n;
    float mn = v[0], mx = v[0];
    for (int k = 1; k < n; k++) { mn = std::min(mn, v[k]); mx = std::max(mx, v[k]); }
    if ((mx - mn) > 1e-4f)
      for (int k = 0; k < n; k++) v[k] = (v[k] - mn) / (mx - mn);
  }
}

// lanczos.cu:1079-1101 generalizedEigensolve.
int orc_generalized_eigensolve(int w, int h, const float* W, int nvec, float tol,
                               float* eigvals, float* eigvecs, int* iterations) {
  const int n = w * h;
  std::vector<float> M(W, W + (size_t)81 * n), d(n), rs(n);
  orc_convert_matrix(w, h, M.data(), d.data(), rs.data());
  int rc = orc_lanczos_fp32(w, h, M.data(), nvec, tol, eigvals, eigvecs, iterations);
  orc_finish_eigvecs(w, h, nvec, rs.data(), eigvecs);
  return rc;
}

// damascene.cu:269-459 stage order for the plain FP32 path.
int orc_gpb_pipeline(int w, int h, const uint32_t* rgb, int nvec, float tol, int* textons,
                     float* mpb8, float* cg8, float* nmax, float* eigvals, float* eigvecs,
                     float* spb8, float* gpbAll, float* gpb, double* ms, int ntex) {
  using clk = std::chrono::steady_clock;
  const int n = w * h;
  auto t0 = clk::now();
  auto lap = [&](int k) {
    auto t1 = clk::now();
    if (ms) ms[k] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    t0 = t1;
  };
  std::vector<float> grey(n), L(n), A(n), B(n);
  orc_rgb_to_grey(w, h, rgb, grey.data());
  int it = 0;
  orc_find_textons(w, h, grey.data(), ntex, textons, &it);
  lap(0);
  orc_rgb_to_lab(w, h, 2.5f, rgb, L.data(), A.data(), B.data());
  orc_normalize_lab(w, h, L.data(), A.data(), B.data());
  std::vector<float> bg((size_t)24 * n), ca((size_t)24 * n), cb((size_t)24 * n), tg((size_t)24 * n);
  orc_local_cues(w, h, L.data(), A.data(), B.data(), textons, ntex, bg.data(), ca.data(), cb.data(), tg.data());
  lap(1);
  orc_combine(w, h, bg.data(), ca.data(), cb.data(), tg.data(), ntex, mpb8, cg8);
  orc_nonmax(w, h, mpb8, nmax);
  lap(2);
  std::vector<float> Wm((size_t)81 * n);
  orc_intervene(w, h, nmax, Wm.data(), 0.1f);
  lap(3);
  int iters = 0;
  int rc = orc_generalized_eigensolve(w, h, Wm.data(), nvec, tol, eigvals, eigvecs, &iters);
  lap(4);
  orc_spectral_pb(w, h, nvec, eigvals, eigvecs, spb8);
  orc_gpb(w, h, cg8, spb8, nmax, gpbAll, gpb);
  lap(5);
  return rc < 0 ? rc : 0;
}

}  // extern "C"
