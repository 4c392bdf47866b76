// symeig.cpp -- the block eigensolver's Rayleigh-Ritz step on the host
// (FP64): Cholesky of the Gram matrix and a dense symmetric
// eigendecomposition (a 32 x 32 problem per outer iteration).
//
// Householder reduction to tridiagonal form T = Q' A Q, then implicit
// symmetric QR steps with Wilkinson shifts on T, the rotations accumulated
// into Q (Golub and Van Loan, Matrix Computations, 4th ed., Algorithms 5.1.1,
// 8.3.1 and 8.3.3).  About 9 n^3 flops in all, against roughly 6 n^3 per
// sweep for cyclic Jacobi.
#include <algorithm>
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

// Implicit symmetric QR with Wilkinson shifts (Algorithm 8.3.3) on the
// tridiagonal (d, e), |entries| = O(1): d becomes the eigenvalues and the
// rotations are applied to T (n x n, row k = column k of the current Q).
// The rotations of a sweep form a dependent chain (about 100 cycles per
// rotation at n = 32, where the square root and the divisions sit); their
// updates of T are vector work that runs beside it.  (Eigenvalues alone,
// then inverse iteration for the vectors, measured slower at n = 32: the
// value-only sweeps keep the same chain, and the solves add chains of their
// own.  Applying a sweep's rotations to T afterwards, in registers, did not
// shorten it either.)  False if 40 n sweeps do not converge.
bool tridiag_qr(double* d, double* e, int n, double* T) {
  const double eps = 2.220446049250313e-16;
  for (int iter = 0; iter < 40 * n; iter++) {
    for (int i = 0; i + 1 < n; i++)
      if (std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1]))) e[i] = 0.0;
    int hi = n - 1;
    while (hi > 0 && e[hi - 1] == 0.0) hi--;
    if (hi == 0) return true;
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

}  // namespace

// Built at -O3 with floating-point contraction, in two clones picked at load
// time by the CPU: AVX2 + FMA (every host this library runs on) and
// baseline x86-64.  A host always runs the same clone, so its results are
// reproducible there.
__attribute__((target_clones("arch=haswell", "default"))) bool sym_eigen(const double* A_in, int n, double* ev,
                                                                         double* Z) {
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
    // (four rows' sums at a time: independent chains, each in column order)
    int i4 = 0;
    for (; i4 + 4 <= m; i4 += 4) {
      const double* a0 = &A[(k + 1 + i4) * n + k + 1];
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      for (int j = 0; j < m; j++) {
        s0 += a0[j] * v[j];
        s1 += a0[n + j] * v[j];
        s2 += a0[2 * n + j] * v[j];
        s3 += a0[3 * n + j] * v[j];
      }
      p[i4] = beta * s0;
      p[i4 + 1] = beta * s1;
      p[i4 + 2] = beta * s2;
      p[i4 + 3] = beta * s3;
    }
    for (int i = i4; i < m; i++) {
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
  // 2. implicit symmetric QR on (d, e), the rotations accumulated into Q
  if (!tridiag_qr(d.data(), e.data(), n, T.data())) return false;
  for (int i = 0; i < n; i++) ev[i] = d[i] * scale;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) Z[i * n + j] = T[(size_t)j * n + i];
  return true;
}


namespace {

// Cholesky G = L L' (lower, in place); false if G is not positive definite.
bool cholesky(std::vector<double>& L, int m) {
  for (int j = 0; j < m; j++) {
    double d = L[j * m + j];
    for (int k = 0; k < j; k++) d -= L[j * m + k] * L[j * m + k];
    if (!(d > 0.0)) return false;
    d = std::sqrt(d);
    L[j * m + j] = d;
    // the rows below, four at a time (independent chains, each in k order)
    const double* lj = &L[j * m];
    int i = j + 1;
    for (; i + 4 <= m; i += 4) {
      double* l0 = &L[i * m];
      double s0 = l0[j], s1 = l0[m + j], s2 = l0[2 * m + j], s3 = l0[3 * m + j];
      for (int k = 0; k < j; k++) {
        s0 -= l0[k] * lj[k];
        s1 -= l0[m + k] * lj[k];
        s2 -= l0[2 * m + k] * lj[k];
        s3 -= l0[3 * m + k] * lj[k];
      }
      l0[j] = s0 / d;
      l0[m + j] = s1 / d;
      l0[2 * m + j] = s2 / d;
      l0[3 * m + j] = s3 / d;
    }
    for (; i < m; i++) {
      double s = L[i * m + j];
      for (int k = 0; k < j; k++) s -= L[i * m + k] * L[j * m + k];
      L[i * m + j] = s / d;
    }
    for (int k = j + 1; k < m; k++) L[j * m + k] = 0.0;
  }
  return true;
}

}  // namespace

// H s = theta G s: S (row-major m x m, column j = s_j, ascending theta)
// with S' G S = I.  False if G is not positive definite.
__attribute__((target_clones("arch=haswell", "default"))) bool rayleigh_ritz(const double* G, const double* H, int m,
                                                                             std::vector<double>& theta,
                                                                             std::vector<double>& S) {
  std::vector<double> L(G, G + m * m);
  for (int i = 0; i < m; i++)
    for (int j = 0; j < i; j++) L[i * m + j] = L[j * m + i] = 0.5 * (G[i * m + j] + G[j * m + i]);
  if (!cholesky(L, m)) return false;
  // M = L^-1 H L^-T, with H symmetrised.  The triangular solves run row by
  // row with the columns side by side (each entry's sum still in k order).
  std::vector<double> T((size_t)m * m), U((size_t)m * m), M((size_t)m * m), acc(m);
  for (int i = 0; i < m; i++) {  // T = L^-1 H
    for (int c = 0; c < m; c++) acc[c] = 0.5 * (H[i * m + c] + H[c * m + i]);
    for (int k = 0; k < i; k++) {
      const double l = L[i * m + k];
      for (int c = 0; c < m; c++) acc[c] -= l * T[k * m + c];
    }
    for (int c = 0; c < m; c++) T[i * m + c] = acc[c] / L[i * m + i];
  }
  for (int i = 0; i < m; i++) {  // U = L^-1 T' = M' (M is symmetric)
    for (int r = 0; r < m; r++) acc[r] = T[r * m + i];
    for (int k = 0; k < i; k++) {
      const double l = L[i * m + k];
      for (int r = 0; r < m; r++) acc[r] -= l * U[k * m + r];
    }
    for (int r = 0; r < m; r++) U[i * m + r] = acc[r] / L[i * m + i];
  }
  for (int i = 0; i < m; i++) {
    M[i * m + i] = U[i * m + i];
    for (int j = 0; j < i; j++) M[i * m + j] = M[j * m + i] = 0.5 * (U[j * m + i] + U[i * m + j]);
  }
  std::vector<double> ev(m), Z((size_t)m * m);
  if (!sym_eigen(M.data(), m, ev.data(), Z.data())) return false;
  std::vector<int> ord(m);
  for (int i = 0; i < m; i++) ord[i] = i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return ev[a] < ev[b]; });
  theta.resize(m);
  S.assign((size_t)m * m, 0.0);
  for (int j = 0; j < m; j++) theta[j] = ev[ord[j]];
  for (int i = m - 1; i >= 0; i--) {  // s_j = L^-T z_j (back substitution, all j side by side)
    for (int j = 0; j < m; j++) acc[j] = Z[i * m + ord[j]];
    for (int k = i + 1; k < m; k++) {
      const double l = L[k * m + i];
      for (int j = 0; j < m; j++) acc[j] -= l * S[k * m + j];
    }
    for (int j = 0; j < m; j++) S[i * m + j] = acc[j] / L[i * m + i];
  }
  return true;
}

}  // namespace dmsc

extern "C" int dmsc_symmetric_eigen(int n, const double* h_a, double* h_eigvals, double* h_eigvecs) {
  if (n < 1 || n > 4096 || !h_a || !h_eigvals || !h_eigvecs) return dmsc::fail(DMSC_ERR_ARG, "bad arguments");
  if (!dmsc::sym_eigen(h_a, n, h_eigvals, h_eigvecs))
    return dmsc::fail(DMSC_ERR_NOCONV, "symmetric QR iteration did not converge");
  return DMSC_OK;
}
