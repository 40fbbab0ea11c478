// Numerical primitives for the host-side operator build: Bessel J_n / Y_n,
// Gauss-Legendre-10, Chebyshev polynomials, the 6-significant-digit table
// rounding, and a complex Householder QR with column pivoting.
//
// Reference call sites these replace:
//   besselJ / besselY (Boost.Math cyl_bessel_j / cyl_neumann)  GMRES/G_FMM2DTree.hpp:73-100
//   Gauss_Legendre_Nodes_and_Weights(10, ...)                   GMRES/Gauss_Legendre_Nodes_and_Weights.hpp:17
//   get_ChebPoly (cos(n acos x))                                 GMRES/G_FMM2DTree.hpp:670-672
//   Eigen text round trip of tables (stream precision 6)         GMRES/G_FMM2DTree.hpp:1252-1315
//   Eigen ColPivHouseholderQR factor + solve                     GMRES/G_FMM2DTree.hpp:2135,2346,3379,3536
#pragma once
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <algorithm>
#include <limits>

#include "bessel.hpp"

namespace lsg {

using cplx = std::complex<double>;

// ---------------------------------------------------------------------------
// Gauss-Legendre 10-point rule on [-1,1], regenerated by Newton iteration on
// P_10 in extended precision (the reference ships the same rule as a literal
// table, GMRES/Gauss_Legendre_Nodes_and_Weights.hpp case 10).
// Nodes ascending.
// ---------------------------------------------------------------------------
inline void gauss_legendre(int n, double* nodes, double* weights) {
  for (int i = 0; i < n; ++i) {
    long double z = std::cos(kPi * (i + 0.75L) / (n + 0.5L));
    long double pp = 0;
    for (int it = 0; it < 100; ++it) {
      long double p1 = 1, p2 = 0;
      for (int j = 1; j <= n; ++j) { long double p3 = p2; p2 = p1; p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j; }
      pp = n * (z * p1 - p2) / (z * z - 1);
      long double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs((double)(z - z1)) == 0.0) break;
    }
    {
      long double p1 = 1, p2 = 0;
      for (int j = 1; j <= n; ++j) { long double p3 = p2; p2 = p1; p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j; }
      pp = n * (z * p1 - p2) / (z * z - 1);
    }
    nodes[n - 1 - i] = (double)z;  // cos(...) runs from +1 down; store ascending
    weights[n - 1 - i] = (double)(2.0L / ((1 - z * z) * pp * pp));
  }
}

inline double cheb_poly(double x, int n) { return std::cos(n * std::acos(x)); }

// Value after an Eigen/iostream text round trip at the default stream precision
// (6 significant digits, %g format), which the reference applies to the M and
// neighbour tables before using them (writeMToFile / readMFromFile).
inline double round6(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.6g", v);
  return std::strtod(buf, nullptr);
}
inline cplx round6(cplx v) { return cplx(round6(v.real()), round6(v.imag())); }

// ---------------------------------------------------------------------------
// Complex Householder QR with column pivoting and the rank-truncated basic
// solve, following Eigen's ColPivHouseholderQR semantics (pivot = largest
// remaining column norm with norm downdating; pivots below
// eps*max_colnorm*sqrt(rows-k) are treated as zero; solve applies Q^H,
// back-substitutes on the leading nonzero-pivot block and zeroes the rest).
// Column-major storage.
// ---------------------------------------------------------------------------
class ColPivQR {
 public:
  int rows = 0, cols = 0, size = 0, nonzero = 0;
  std::vector<cplx> qr, tau;
  std::vector<int> perm;  // perm[i] = original column now at position i

  void factor(const cplx* A, int m, int n) {
    rows = m; cols = n; size = std::min(m, n);
    qr.assign(A, A + size_t(m) * n);
    tau.assign(size, cplx(0));
    std::vector<double> nu(n), nd(n);
    std::vector<int> transp(n);
    for (int k = 0; k < n; ++k) { nd[k] = colnorm(k, 0); nu[k] = nd[k]; }
    double maxn = 0;
    for (int k = 0; k < n; ++k) maxn = std::max(maxn, nu[k]);
    const double eps = std::numeric_limits<double>::epsilon();
    const double thr_helper = (maxn * eps) * (maxn * eps) / double(m);
    const double down_thr = std::sqrt(eps);
    nonzero = size;
    std::vector<cplx> tmp(n);
    for (int k = 0; k < size; ++k) {
      int big = k;
      double bv = nu[k];
      for (int j = k + 1; j < n; ++j) if (nu[j] > bv) { bv = nu[j]; big = j; }
      if (nonzero == size && bv * bv < thr_helper * double(m - k)) nonzero = k;
      transp[k] = big;
      if (big != k) {
        for (int i = 0; i < m; ++i) std::swap(at(i, k), at(i, big));
        std::swap(nu[k], nu[big]); std::swap(nd[k], nd[big]);
      }
      // Householder on column k, rows k..m-1
      cplx c0 = at(k, k);
      double tail = 0;
      for (int i = k + 1; i < m; ++i) tail += std::norm(at(i, k));
      const double tol = std::numeric_limits<double>::min();
      double beta;
      cplx t;
      if (tail <= tol && std::imag(c0) * std::imag(c0) <= tol) {
        t = 0; beta = std::real(c0);
        for (int i = k + 1; i < m; ++i) at(i, k) = 0;
      } else {
        beta = std::sqrt(std::norm(c0) + tail);
        if (std::real(c0) >= 0) beta = -beta;
        cplx d = c0 - beta;
        for (int i = k + 1; i < m; ++i) at(i, k) /= d;
        t = std::conj((beta - c0) / beta);
      }
      tau[k] = t;
      at(k, k) = beta;
      // apply H_k = I - t v v^H to the trailing columns (v = [1; essential])
      if (k + 1 < n) {
        if (m - k == 1) {
          for (int j = k + 1; j < n; ++j) at(k, j) *= (1.0 - t);
        } else if (t != cplx(0)) {
          for (int j = k + 1; j < n; ++j) {
            cplx s = at(k, j);
            for (int i = k + 1; i < m; ++i) s += std::conj(at(i, k)) * at(i, j);
            tmp[j] = s;
          }
          for (int j = k + 1; j < n; ++j) {
            cplx ts = t * tmp[j];
            at(k, j) -= ts;
            for (int i = k + 1; i < m; ++i) at(i, j) -= at(i, k) * ts;
          }
        }
      }
      for (int j = k + 1; j < n; ++j) {
        if (nu[j] != 0.0) {
          double tt = std::abs(at(k, j)) / nu[j];
          tt = (1.0 + tt) * (1.0 - tt);
          if (tt < 0) tt = 0;
          double r = nu[j] / nd[j];
          double t2 = tt * r * r;
          if (t2 <= down_thr) { nd[j] = colnorm(j, k + 1); nu[j] = nd[j]; }
          else nu[j] *= std::sqrt(tt);
        }
      }
    }
    perm.resize(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < size; ++k) std::swap(perm[k], perm[transp[k]]);
  }

  // X (cols x nrhs) = basic solution of A X = B (B: rows x nrhs), column-major.
  void solve(const cplx* B, int nrhs, cplx* X) const {
    std::vector<cplx> c(rows);
    for (int r = 0; r < nrhs; ++r) {
      for (int i = 0; i < cols; ++i) X[size_t(r) * cols + i] = 0;
      if (nonzero == 0) continue;
      std::copy(B + size_t(r) * rows, B + size_t(r + 1) * rows, c.begin());
      for (int k = 0; k < nonzero; ++k) {
        // R = H_{n-1}..H_0 A with H_k = I - tau_k v v^H (H_k a_k = beta e_1),
        // so Q^H b applies H_0, H_1, ... in order with tau_k itself.
        cplx t = tau[k];
        if (rows - k == 1) { c[k] *= (1.0 - t); continue; }
        if (t == cplx(0)) continue;
        cplx s = c[k];
        for (int i = k + 1; i < rows; ++i) s += std::conj(at(i, k)) * c[i];
        cplx ts = t * s;
        c[k] -= ts;
        for (int i = k + 1; i < rows; ++i) c[i] -= at(i, k) * ts;
      }
      for (int i = nonzero - 1; i >= 0; --i) {
        cplx s = c[i];
        for (int j = i + 1; j < nonzero; ++j) s -= at(i, j) * c[j];
        c[i] = s / at(i, i);
      }
      for (int i = 0; i < nonzero; ++i) X[size_t(r) * cols + perm[i]] = c[i];
    }
  }

 private:
  cplx& at(int i, int j) { return qr[size_t(j) * rows + i]; }
  const cplx& at(int i, int j) const { return qr[size_t(j) * rows + i]; }
  double colnorm(int j, int from) const {
    // scaled 2-norm, like Eigen's stableNorm-free norm(): sqrt(sum |a|^2)
    double s = 0;
    for (int i = from; i < rows; ++i) s += std::norm(at(i, j));
    return std::sqrt(s);
  }
};

// SplitMix64 stream used for every seeded input (contrast bumps, angles, vectors).
struct SplitMix64 {
  unsigned long long s;
  explicit SplitMix64(unsigned long long seed) : s(seed) {}
  unsigned long long next() {
    unsigned long long z = (s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * 0x1.0p-53; }
};

}  // namespace lsg
