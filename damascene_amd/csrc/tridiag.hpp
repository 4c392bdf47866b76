// tridiag.hpp -- symmetric tridiagonal eigen-solver in FP64 (bisection on
// Sturm counts for eigenvalues, inverse iteration for eigenvectors), usable
// from host and device code.  It stands in for the LAPACK dstebz/dstein
// calls of noReorthog/lanczos.cu:292, 307, 373.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace dmsc {
namespace tri {

// Number of eigenvalues of T (diagonal a[0..n), off-diagonal b[0..n-1))
// strictly below x.
__host__ __device__ inline int sturm(int n, const double* a, const double* b, double x, double pivmin) {
  int cnt = 0;
  double q = a[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  cnt += q < 0;
  for (int i = 1; i < n; i++) {
    q = (a[i] - x) - (b[i - 1] * b[i - 1]) / q;
    if (fabs(q) < pivmin) q = -pivmin;
    cnt += q < 0;
  }
  return cnt;
}

struct Bounds {
  double lo, hi, pivmin, tnorm;
};

__host__ __device__ inline Bounds gershgorin(int n, const double* a, const double* b) {
  Bounds g;
  g.lo = a[0];
  g.hi = a[0];
  double bmax = 0;
  for (int i = 0; i < n; i++) {
    const double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < n - 1 ? fabs(b[i]) : 0.0);
    g.lo = fmin(g.lo, a[i] - r);
    g.hi = fmax(g.hi, a[i] + r);
    if (i < n - 1) bmax = fmax(bmax, b[i] * b[i]);
  }
  g.tnorm = fmax(fabs(g.lo), fabs(g.hi));
  g.pivmin = fmax(2.2250738585072014e-308 * fmax(1.0, bmax), 1e-300);
  const double pad = 2.2e-16 * g.tnorm * n + 2 * g.pivmin;
  g.lo -= pad;
  g.hi += pad;
  return g;
}

// k-th smallest eigenvalue (1-based) by bisection to full precision.
__host__ __device__ inline double kth(int n, const double* a, const double* b, int k, const Bounds& g) {
  double l = g.lo, h = g.hi;
  for (int it = 0; it < 200; it++) {
    const double m = 0.5 * (l + h);
    if (m <= l || m >= h) break;
    if (h - l <= 4.4e-16 * fmax(fabs(l), fabs(h)) + g.pivmin) break;
    if (sturm(n, a, b, m, g.pivmin) >= k) h = m;
    else l = m;
  }
  return 0.5 * (l + h);
}

// Inverse iteration for eigenvalue lam; work needs 5n doubles and n ints.
// prev/np: vectors (stride n) to orthogonalise against (nearby eigenvalues).
// Result normalised, largest-magnitude entry positive.
__host__ __device__ inline int invit(int n, const double* a, const double* b, double lam, double* v,
                                     const double* prev, int np, double* work, int* piv) {
  double *dl = work, *d = work + n, *du = work + 2 * n, *du2 = work + 3 * n, *x = work + 4 * n;
  double tnorm = 0;
  for (int i = 0; i < n; i++)
    tnorm = fmax(tnorm, fabs(a[i]) + (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < n - 1 ? fabs(b[i]) : 0.0));
  const double eps = 2.2e-16 * fmax(tnorm, 1e-300);
  for (int i = 0; i < n; i++) d[i] = a[i] - lam;
  for (int i = 0; i < n - 1; i++) { dl[i] = b[i]; du[i] = b[i]; du2[i] = 0; }
  for (int i = 0; i < n - 1; i++) {  // LU with partial pivoting (dgttrf)
    if (fabs(d[i]) >= fabs(dl[i])) {
      piv[i] = i;
      if (d[i] == 0) d[i] = eps;
      const double f = dl[i] / d[i];
      dl[i] = f;
      d[i + 1] -= f * du[i];
    } else {
      piv[i] = i + 1;
      const double f = d[i] / dl[i];
      d[i] = dl[i];
      dl[i] = f;
      const double t = du[i];
      du[i] = d[i + 1];
      d[i + 1] = t - f * d[i + 1];
      if (i < n - 2) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
    }
  }
  piv[n - 1] = n - 1;
  for (int i = 0; i < n; i++)
    if (fabs(d[i]) < eps) d[i] = d[i] < 0 ? -eps : eps;
  unsigned long long s = 0x9E3779B97F4A7C15ULL;
  for (int i = 0; i < n; i++) {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    x[i] = ((double)(s >> 11) / 9007199254740992.0) - 0.5;
  }
  for (int iter = 0; iter < 6; iter++) {
    for (int j = 0; j < np; j++) {
      const double* p = prev + (size_t)j * n;
      double dot = 0;
      for (int i = 0; i < n; i++) dot += p[i] * x[i];
      for (int i = 0; i < n; i++) x[i] -= dot * p[i];
    }
    for (int i = 0; i < n - 1; i++) {
      if (piv[i] == i) {
        x[i + 1] -= dl[i] * x[i];
      } else {
        const double t = x[i];
        x[i] = x[i + 1];
        x[i + 1] = t - dl[i] * x[i];
      }
    }
    x[n - 1] /= d[n - 1];
    if (n > 1) x[n - 2] = (x[n - 2] - du[n - 2] * x[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; i--) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / d[i];
    double nrm = 0;
    for (int i = 0; i < n; i++) nrm += x[i] * x[i];
    nrm = sqrt(nrm);
    if (!(nrm > 0) || !isfinite(nrm)) return -1;
    for (int i = 0; i < n; i++) x[i] /= nrm;
  }
  for (int j = 0; j < np; j++) {
    const double* p = prev + (size_t)j * n;
    double dot = 0;
    for (int i = 0; i < n; i++) dot += p[i] * x[i];
    for (int i = 0; i < n; i++) x[i] -= dot * p[i];
  }
  double nrm = 0, big = 0;
  int jm = 0;
  for (int i = 0; i < n; i++) {
    nrm += x[i] * x[i];
    if (fabs(x[i]) > big) { big = fabs(x[i]); jm = i; }
  }
  const double scl = (x[jm] < 0 ? -1.0 : 1.0) / sqrt(nrm);
  for (int i = 0; i < n; i++) v[i] = x[i] * scl;
  return 0;
}

}  // namespace tri
}  // namespace dmsc
