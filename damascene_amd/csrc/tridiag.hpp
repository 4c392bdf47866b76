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

// Sturm count by the three-term recurrence of the leading principal minors,
// p_i = (a_i - x) p_{i-1} - b_{i-1}^2 p_{i-2}: the number of sign changes in
// p_0 = 1, p_1, ..., p_n equals the number of eigenvalues below x (a zero
// p_i counts as a change, the convention of the pivot form above).  One
// dependent multiply-add per row instead of a division; the pair is
// rescaled by an exact power of two every 4 rows to stay in range.
// ab(i) returns (a_i, b_{i-1}^2).
template <typename AB>
__host__ __device__ inline int sturm_count(int n, AB ab, double x) {
  const double kBig = 0x1p+400, kSmall = 0x1p-400;
  double pp = 1.0, p = ab(0).x - x;  // p_0, p_1
  if (p == 0) p = -0x1p-1000;
  int cnt = p < 0;
  int i = 1;
  for (; i + 4 <= n; i += 4) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const double2 e = ab(i + j);
      double pn = fma(e.x - x, p, -e.y * pp);
      if (pn == 0) pn = p < 0 ? 0x1p-1000 : -0x1p-1000;
      cnt += (pn < 0) != (p < 0);
      pp = p;
      p = pn;
    }
    const double m = fmax(fabs(p), fabs(pp));
    const double sc = m > kBig ? kSmall : (m < kSmall ? kBig : 1.0);
    p *= sc;
    pp *= sc;
  }
  for (; i < n; i++) {
    const double2 e = ab(i);
    double pn = fma(e.x - x, p, -e.y * pp);
    if (pn == 0) pn = p < 0 ? 0x1p-1000 : -0x1p-1000;
    cnt += (pn < 0) != (p < 0);
    pp = p;
    p = pn;
  }
  return cnt;
}

// Diagonal a[0..n) and squared off-diagonal b2[0..n-1).
__host__ __device__ inline int sturm_poly(int n, const double* a, const double* b2, double x) {
  return sturm_count(n, [&](int i) { return make_double2(a[i], i > 0 ? b2[i - 1] : 0.0); }, x);
}

// Pairs (a_i, b_{i-1}^2), e.g. an LDS copy.
__host__ __device__ inline int sturm_pairs(int n, const double2* ab, double x) {
  return sturm_count(n, [&](int i) { return ab[i]; }, x);
}

// 1/x: on the device v_rcp_f64 refined by two Newton steps (a few ulp,
// enough for pivots and multipliers), on the host a division.
__host__ __device__ inline double recip(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  r = fma(r, fma(-x, r, 1.0), r);
  r = fma(r, fma(-x, r, 1.0), r);
  return r;
#else
  return 1.0 / x;
#endif
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

// Inverse iteration (LU with partial pivoting of T - lam I, as LAPACK
// dgttrf/dstein) for the eigenvalue lam, `iters` solves from a fixed
// pseudo-random start.  dl, d, du, du2, x: n doubles each, piv: n bytes
// (device callers place them in LDS).  No orthogonalisation against
// neighbouring vectors: a value separated from the rest of the spectrum of T
// by g is separated by a factor g / (|T| eps) per solve (1e5 for the 1e-10
// copies the mixed path keeps apart, far more for distinct eigenvalues).
// The result is normalised with its largest-magnitude entry positive.
// Returns -1 on a non-finite iterate.
//
// Every loop reads its inputs two rows ahead into registers, so on the
// device the memory latency overlaps the serial recurrence instead of
// adding to it.
__host__ __device__ inline int invit(int n, const double* __restrict__ a, const double* __restrict__ b, double lam,
                                     double* __restrict__ v, double* __restrict__ dl, double* __restrict__ d,
                                     double* __restrict__ du, double* __restrict__ du2, double* __restrict__ x,
                                     unsigned char* __restrict__ piv, int iters) {
  double tnorm = 0;
  for (int i = 0; i < n; i++)
    tnorm = fmax(tnorm, fabs(a[i]) + (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < n - 1 ? fabs(b[i]) : 0.0));
  const double eps = 2.2e-16 * fmax(tnorm, 1e-300);
  // factorisation (dgttrf), carrying the running diagonal entry di and the
  // running upper entry ui (modified by a preceding row interchange); row i
  // uses li = b[i], dn = a[i+1] - lam, un = b[i+1] (0 past the end)
  auto rowin = [&](int i, double& li, double& dn, double& un) {
    li = b[i];
    dn = a[i + 1] - lam;
    un = i < n - 2 ? b[i + 1] : 0.0;
  };
  double di = a[0] - lam, ui = n > 1 ? b[0] : 0.0;
  double l1 = 0, d1 = 0, u1 = 0, l2 = 0, d2 = 0, u2 = 0;
  if (n > 1) rowin(0, l1, d1, u1);
  if (n > 2) rowin(1, l2, d2, u2);
  for (int i = 0; i < n - 1; i++) {
    const double li = l1, dn = d1, un = u1;
    l1 = l2; d1 = d2; u1 = u2;
    if (i + 2 < n - 1) rowin(i + 2, l2, d2, u2);
    if (fabs(di) >= fabs(li)) {
      piv[i] = 0;
      if (di == 0) di = eps;
      const double f = li * recip(di);
      dl[i] = f;
      d[i] = di;
      du[i] = ui;
      du2[i] = 0;
      di = dn - f * ui;
      ui = un;
    } else {
      piv[i] = 1;
      const double f = di * recip(li);
      dl[i] = f;
      d[i] = li;
      du[i] = dn;
      du2[i] = un;
      di = ui - f * dn;
      ui = -f * un;
    }
  }
  d[n - 1] = di;
  piv[n - 1] = 0;
  // reciprocal pivots and the start vector (a hash of the row index: no
  // serial dependence between rows)
  for (int i = 0; i < n; i++) {
    double t = d[i];
    if (fabs(t) < eps) t = t < 0 ? -eps : eps;
    d[i] = recip(t);
    unsigned long long h = (unsigned long long)(i + 1) * 0x9E3779B97F4A7C15ULL;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ULL;
    h ^= h >> 27;
    x[i] = ((double)(h >> 11) / 9007199254740992.0) - 0.5;
  }
  for (int iter = 0; iter < iters; iter++) {
    double xi = x[0];
    {  // forward: L with interchanges (row i reads x[i+1], dl[i], piv[i])
      double xa = n > 1 ? x[1] : 0, fa = dl[0], xb = n > 2 ? x[2] : 0, fb = n > 1 ? dl[1] : 0;
      int pa = piv[0], pb = n > 1 ? piv[1] : 0;
      for (int i = 0; i < n - 1; i++) {
        const double xn = xa, f = fa;
        const int pv = pa;
        xa = xb; fa = fb; pa = pb;
        if (i + 2 < n - 1) {
          xb = x[i + 3];
          fb = dl[i + 2];
          pb = piv[i + 2];
        }
        if (pv) {
          x[i] = xn;
          xi = xi - f * xn;
        } else {
          x[i] = xi;
          xi = xn - f * xi;
        }
      }
    }
    double x1 = xi * d[n - 1], x2 = 0;  // backward: U (three bands)
    x[n - 1] = x1;
    if (n > 1) {
      x2 = x1;
      x1 = (x[n - 2] - du[n - 2] * x2) * d[n - 2];
      x[n - 2] = x1;
    }
    double nrm = x1 * x1 + x2 * x2;
    if (n > 2) {  // row i reads x[i], du[i], du2[i], d[i]
      double ra[4], rb[4] = {0, 0, 0, 0};
      ra[0] = x[n - 3]; ra[1] = du[n - 3]; ra[2] = du2[n - 3]; ra[3] = d[n - 3];
      if (n > 3) { rb[0] = x[n - 4]; rb[1] = du[n - 4]; rb[2] = du2[n - 4]; rb[3] = d[n - 4]; }
      for (int i = n - 3; i >= 0; i--) {
        const double xv = ra[0], u = ra[1], u2 = ra[2], dv = ra[3];
        ra[0] = rb[0]; ra[1] = rb[1]; ra[2] = rb[2]; ra[3] = rb[3];
        if (i - 2 >= 0) { rb[0] = x[i - 2]; rb[1] = du[i - 2]; rb[2] = du2[i - 2]; rb[3] = d[i - 2]; }
        const double xi0 = (xv - u * x1 - u2 * x2) * dv;
        x[i] = xi0;
        x2 = x1;
        x1 = xi0;
        nrm += xi0 * xi0;
      }
    }
    nrm = sqrt(nrm);
    if (!(nrm > 0) || !isfinite(nrm)) return -1;
    const double inv = 1.0 / nrm;
    for (int i = 0; i < n; i++) x[i] = x[i] * inv;
  }
  double big = 0;
  int jm = 0;
  for (int i = 0; i < n; i++)
    if (fabs(x[i]) > big) { big = fabs(x[i]); jm = i; }
  const double sg = x[jm] < 0 ? -1.0 : 1.0;
  for (int i = 0; i < n; i++) v[i] = x[i] * sg;
  return 0;
}

#ifdef __HIPCC__
// invit on one 64-lane wave (blockDim.x == 64, every lane calls it): lane 0
// runs the three recurrences that carry a serial dependence (factorisation,
// forward and backward substitution); the passes with none (|T| row sums,
// reciprocal pivots and start vector, normalisation, largest entry and
// sign) are shared by the 64 lanes.  Every value is formed by the same
// operations as in invit (the row-sum maximum and the first largest entry
// do not depend on the order of the comparisons), so the results are
// bit-identical to it.  All arrays must be visible to every lane (LDS or
// global memory of this block).
__device__ inline int invit_wave(int n, const double* __restrict__ a, const double* __restrict__ b, double lam,
                                 double* __restrict__ v, double* __restrict__ dl, double* __restrict__ d,
                                 double* __restrict__ du, double* __restrict__ du2, double* __restrict__ x,
                                 unsigned char* __restrict__ piv, int iters, double* __restrict__ shared4) {
  const int lane = threadIdx.x;
  double tnorm = 0;
  for (int i = lane; i < n; i += 64)
    tnorm = fmax(tnorm, fabs(a[i]) + (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < n - 1 ? fabs(b[i]) : 0.0));
  for (int o = 32; o >= 1; o >>= 1) tnorm = fmax(tnorm, __shfl_xor(tnorm, o, 64));
  const double eps = 2.2e-16 * fmax(tnorm, 1e-300);
  if (lane == 0) {  // factorisation (dgttrf), as in invit
    auto rowin = [&](int i, double& li, double& dn, double& un) {
      li = b[i];
      dn = a[i + 1] - lam;
      un = i < n - 2 ? b[i + 1] : 0.0;
    };
    double di = a[0] - lam, ui = n > 1 ? b[0] : 0.0;
    double l1 = 0, d1 = 0, u1 = 0, l2 = 0, d2 = 0, u2 = 0;
    if (n > 1) rowin(0, l1, d1, u1);
    if (n > 2) rowin(1, l2, d2, u2);
    for (int i = 0; i < n - 1; i++) {
      const double li = l1, dn = d1, un = u1;
      l1 = l2; d1 = d2; u1 = u2;
      if (i + 2 < n - 1) rowin(i + 2, l2, d2, u2);
      if (fabs(di) >= fabs(li)) {
        piv[i] = 0;
        if (di == 0) di = eps;
        const double f = li * recip(di);
        dl[i] = f;
        d[i] = di;
        du[i] = ui;
        du2[i] = 0;
        di = dn - f * ui;
        ui = un;
      } else {
        piv[i] = 1;
        const double f = di * recip(li);
        dl[i] = f;
        d[i] = li;
        du[i] = dn;
        du2[i] = un;
        di = ui - f * dn;
        ui = -f * un;
      }
    }
    d[n - 1] = di;
    piv[n - 1] = 0;
  }
  __syncthreads();
  for (int i = lane; i < n; i += 64) {  // reciprocal pivots and the hashed start vector
    double t = d[i];
    if (fabs(t) < eps) t = t < 0 ? -eps : eps;
    d[i] = recip(t);
    unsigned long long h = (unsigned long long)(i + 1) * 0x9E3779B97F4A7C15ULL;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ULL;
    h ^= h >> 27;
    x[i] = ((double)(h >> 11) / 9007199254740992.0) - 0.5;
  }
  __syncthreads();
  for (int iter = 0; iter < iters; iter++) {
    if (lane == 0) {
      double xi = x[0];
      {  // forward: L with interchanges
        double xa = n > 1 ? x[1] : 0, fa = dl[0], xb = n > 2 ? x[2] : 0, fb = n > 1 ? dl[1] : 0;
        int pa = piv[0], pb = n > 1 ? piv[1] : 0;
        for (int i = 0; i < n - 1; i++) {
          const double xn = xa, f = fa;
          const int pv = pa;
          xa = xb; fa = fb; pa = pb;
          if (i + 2 < n - 1) {
            xb = x[i + 3];
            fb = dl[i + 2];
            pb = piv[i + 2];
          }
          if (pv) {
            x[i] = xn;
            xi = xi - f * xn;
          } else {
            x[i] = xi;
            xi = xn - f * xi;
          }
        }
      }
      double x1 = xi * d[n - 1], x2 = 0;  // backward: U (three bands)
      x[n - 1] = x1;
      if (n > 1) {
        x2 = x1;
        x1 = (x[n - 2] - du[n - 2] * x2) * d[n - 2];
        x[n - 2] = x1;
      }
      double nrm = x1 * x1 + x2 * x2;
      if (n > 2) {
        double ra[4], rb[4] = {0, 0, 0, 0};
        ra[0] = x[n - 3]; ra[1] = du[n - 3]; ra[2] = du2[n - 3]; ra[3] = d[n - 3];
        if (n > 3) { rb[0] = x[n - 4]; rb[1] = du[n - 4]; rb[2] = du2[n - 4]; rb[3] = d[n - 4]; }
        for (int i = n - 3; i >= 0; i--) {
          const double xv = ra[0], u = ra[1], u2 = ra[2], dv = ra[3];
          ra[0] = rb[0]; ra[1] = rb[1]; ra[2] = rb[2]; ra[3] = rb[3];
          if (i - 2 >= 0) { rb[0] = x[i - 2]; rb[1] = du[i - 2]; rb[2] = du2[i - 2]; rb[3] = d[i - 2]; }
          const double xi0 = (xv - u * x1 - u2 * x2) * dv;
          x[i] = xi0;
          x2 = x1;
          x1 = xi0;
          nrm += xi0 * xi0;
        }
      }
      shared4[0] = sqrt(nrm);
    }
    __syncthreads();
    const double nrm = shared4[0];
    if (!(nrm > 0) || !isfinite(nrm)) return -1;
    const double inv = 1.0 / nrm;
    for (int i = lane; i < n; i += 64) x[i] = x[i] * inv;
    __syncthreads();
  }
  double big = 0;
  int jm = n;  // first index of the largest |x|
  for (int i = lane; i < n; i += 64)
    if (fabs(x[i]) > big) { big = fabs(x[i]); jm = i; }
  for (int o = 32; o >= 1; o >>= 1) {
    const double ob = __shfl_xor(big, o, 64);
    const int oj = __shfl_xor(jm, o, 64);
    if (ob > big || (ob == big && oj < jm)) { big = ob; jm = oj; }
  }
  if (jm >= n) jm = 0;
  const double sg = x[jm] < 0 ? -1.0 : 1.0;
  for (int i = lane; i < n; i += 64) v[i] = x[i] * sg;
  __syncthreads();
  return 0;
}
#endif

}  // namespace tri
}  // namespace dmsc
