// symeig.cpp -- dense symmetric eigendecomposition on the host (FP64), for
// the block eigensolver's Rayleigh-Ritz step (a 32 x 32 problem per outer
// iteration).
//
// Householder reduction to tridiagonal form T = Q' A Q, then implicit
// symmetric QR steps with Wilkinson shifts on T, the rotations accumulated
// into Q (Golub and Van Loan, Matrix Computations, 4th ed., Algorithms 5.1.1,
// 8.3.1 and 8.3.3).  About 9 n^3 flops in all, against roughly 6 n^3 per
// sweep for cyclic Jacobi.
#include <cmath>
#include <vector>

#include "dmsc.h"
#include "tables.hpp"

namespace dmsc {

namespace {

// Householder vector of x (length m): v[0] = 1 and beta with
// (I - beta v v') x = +-|x| e_1 (Algorithm 5.1.1).
void house(const double* x, int m, double* v, double* beta) {
  double sigma = 0.0;
  for (int i = 1; i < m; i++) sigma += x[i] * x[i];
  v[0] = 1.0;
  for (int i = 1; i < m; i++) v[i] = x[i];
  if (sigma == 0.0) {
    *beta = 0.0;  // already a multiple of e_1 (its sign is kept)
    return;
  }
  const double mu = std::sqrt(x[0] * x[0] + sigma);
  const double v0 = x[0] <= 0.0 ? x[0] - mu : -sigma / (x[0] + mu);
  *beta = 2.0 * v0 * v0 / (sigma + v0 * v0);
  for (int i = 1; i < m; i++) v[i] /= v0;
}

}  // namespace

bool sym_eigen(const double* A_in, int n, double* ev, double* Z) {
  if (n <= 0) return false;
  // work on A / 2^ceil(log2 max|a_ij|) (undone on the eigenvalues), so the squares in the
  // Givens rotations neither overflow nor underflow
  double amax = 0.0;
  for (size_t i = 0; i < (size_t)n * n; i++) amax = std::fmax(amax, std::fabs(A_in[i]));
  int ex = 0;
  if (amax > 0.0) (void)std::frexp(amax, &ex);
  const double scale = std::ldexp(1.0, ex);  // a power of two: the scaling is exact
  std::vector<double> A((size_t)n * n);
  for (size_t i = 0; i < (size_t)n * n; i++) A[i] = A_in[i] / scale;
  // Q' is accumulated as T (row j = column j of Q), so every update walks
  // contiguous rows; Z = T' at the end
  std::vector<double> T((size_t)n * n, 0.0), srow(n);
  for (int i = 0; i < n; i++) T[(size_t)i * n + i] = 1.0;
  std::vector<double> v(n), p(n), w(n), x(n);
  // 1. A <- Q' A Q tridiagonal, Q accumulated in Z
  for (int k = 0; k + 2 < n; k++) {
    const int m = n - k - 1;
    for (int i = 0; i < m; i++) x[i] = A[(k + 1 + i) * n + k];
    double beta;
    house(x.data(), m, v.data(), &beta);
    if (beta == 0.0) continue;
    // p = beta A22 v, w = p - (beta p'v / 2) v, A22 <- A22 - v w' - w v'
    for (int i = 0; i < m; i++) {
      double s = 0.0;
      for (int j = 0; j < m; j++) s += A[(k + 1 + i) * n + (k + 1 + j)] * v[j];
      p[i] = beta * s;
    }
    double pv = 0.0;
    for (int i = 0; i < m; i++) pv += p[i] * v[i];
    for (int i = 0; i < m; i++) w[i] = p[i] - 0.5 * beta * pv * v[i];
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++) A[(k + 1 + i) * n + (k + 1 + j)] -= v[i] * w[j] + w[i] * v[j];
    // the reflected column: x -> (x_0 - beta (v'x)) e_1
    double vx = 0.0;
    for (int i = 0; i < m; i++) vx += v[i] * x[i];
    const double top = x[0] - beta * vx;
    A[(k + 1) * n + k] = A[k * n + k + 1] = top;
    for (int i = 1; i < m; i++) A[(k + 1 + i) * n + k] = A[k * n + k + 1 + i] = 0.0;
    // Q <- Q (I - beta v v') on columns k+1 ..: rows k+1 .. of T
    for (int r = 0; r < n; r++) srow[r] = 0.0;
    for (int j = 0; j < m; j++) {
      const double* t = &T[(size_t)(k + 1 + j) * n];
      for (int r = 0; r < n; r++) srow[r] += t[r] * v[j];
    }
    for (int j = 0; j < m; j++) {
      double* t = &T[(size_t)(k + 1 + j) * n];
      const double bv = beta * v[j];
      for (int r = 0; r < n; r++) t[r] -= srow[r] * bv;
    }
  }
  std::vector<double> d(n), e(n, 0.0);
  for (int i = 0; i < n; i++) d[i] = A[i * n + i];
  for (int i = 0; i + 1 < n; i++) e[i] = A[(i + 1) * n + i];
  // 2. implicit symmetric QR with Wilkinson shifts on (d, e)
  const double eps = 2.220446049250313e-16;
  for (int iter = 0; iter < 40 * n; iter++) {
    for (int i = 0; i + 1 < n; i++)
      if (std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1]))) e[i] = 0.0;
    int hi = n - 1;
    while (hi > 0 && e[hi - 1] == 0.0) hi--;
    if (hi == 0) {
      for (int i = 0; i < n; i++) ev[i] = d[i] * scale;
      for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Z[i * n + j] = T[(size_t)j * n + i];
      return true;
    }
    int lo = hi - 1;
    while (lo > 0 && e[lo - 1] != 0.0) lo--;
    // Wilkinson shift: the eigenvalue of the trailing 2 x 2 block nearer d[hi]
    const double dd = 0.5 * (d[hi - 1] - d[hi]), bb = e[hi - 1];
    const double den = dd + std::copysign(std::hypot(dd, bb), dd == 0.0 ? 1.0 : dd);
    const double mu = d[hi] - bb * bb / den;
    double x = d[lo] - mu, z = e[lo], bulge = 0.0;
    for (int k = lo; k < hi; k++) {
      const double r = std::sqrt(x * x + z * z);  // |x|, |z| = O(1) after the scaling
      const double c = r == 0.0 ? 1.0 : x / r, s = r == 0.0 ? 0.0 : -z / r;
      if (k > lo) e[k - 1] = c * e[k - 1] - s * bulge;
      const double a = d[k], b = e[k], ap = d[k + 1];
      d[k] = c * c * a - 2.0 * c * s * b + s * s * ap;
      d[k + 1] = s * s * a + 2.0 * c * s * b + c * c * ap;
      e[k] = c * s * (a - ap) + (c * c - s * s) * b;
      if (k + 1 < hi) {
        bulge = -s * e[k + 1];
        e[k + 1] = c * e[k + 1];
        x = e[k];
        z = bulge;
      }
      double* tk = &T[(size_t)k * n];
      double* tk1 = &T[(size_t)(k + 1) * n];
      for (int col = 0; col < n; col++) {
        const double zk = tk[col], zk1 = tk1[col];
        tk[col] = c * zk - s * zk1;
        tk1[col] = s * zk + c * zk1;
      }
    }
  }
  return false;
}

}  // namespace dmsc

extern "C" int dmsc_symmetric_eigen(int n, const double* h_a, double* h_eigvals, double* h_eigvecs) {
  if (n < 1 || n > 4096 || !h_a || !h_eigvals || !h_eigvecs) return dmsc::fail(DMSC_ERR_ARG, "bad arguments");
  if (!dmsc::sym_eigen(h_a, n, h_eigvals, h_eigvecs))
    return dmsc::fail(DMSC_ERR_NOCONV, "symmetric QR iteration did not converge");
  return DMSC_OK;
}
