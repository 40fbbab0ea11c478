// Bessel functions J_n, Y_n of integer order for real x > 0, shared by the
// host operator build and the device entry kernels (the same operations in
// the same order on both sides).
//
// Reference call sites these replace:
//   besselJ / besselY (Boost.Math cyl_bessel_j / cyl_neumann)  GMRES/G_FMM2DTree.hpp:73-100
#pragma once
#include <algorithm>
#include <cmath>
#include <limits>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LSG_HD __host__ __device__
#else
#define LSG_HD
#endif

namespace lsg {

constexpr double kPi = 3.1415926535897932384;   // same literal as GMRES/ACA.hpp:14
constexpr double kEulerGamma = 0.57721566490153286061;

// ---------------------------------------------------------------------------
// Bessel functions of the first and second kind, integer orders 0..nmax, x > 0.
// Method (Zhang & Jin, "Computation of Special Functions", 1996, JYNB), split
// into three ranges:
//   x < 1      : power series for every J_n, series for Y_0 and Y_1 (A&S 9.1.10-11)
//   1 <= x < 25: Miller backward recurrence normalised by J_0 + 2 sum J_2k = 1;
//                Y_0, Y_1 from the Neumann series of the same J's
//   x >= 25    : Hankel asymptotic expansion for J_0, J_1, Y_0, Y_1, then forward
//                recurrence (stable for n < x)
// Y_n for n >= 2 always by forward recurrence.
// ---------------------------------------------------------------------------
LSG_HD inline void hankel_asym01(double x, double& j0, double& j1, double& y0, double& y1) {
  // P_nu, Q_nu asymptotic series; a_k(nu) = prod_{i=1..k}(4nu^2-(2i-1)^2) / (k! 8^k)
  double s = std::sin(x), c = std::cos(x);
  double pre = std::sqrt(2.0 / (kPi * x));
  for (int nu = 0; nu <= 1; ++nu) {
    double mu = 4.0 * nu * nu;
    double P = 0.0, Q = 0.0;
    double term = 1.0;  // a_k / x^k with sign bookkeeping handled below
    for (int k = 0; k < 60; ++k) {
      // term = a_k(nu) / x^k
      double contrib = term;
      int r = k & 3;  // P: k even, sign (-1)^(k/2); Q: k odd, sign (-1)^((k-1)/2)
      if (r == 0) P += contrib;
      else if (r == 1) Q += contrib;
      else if (r == 2) P -= contrib;
      else Q -= contrib;
      double odd = 2.0 * k + 1.0;
      double next = term * (mu - odd * odd) / ((k + 1.0) * 8.0 * x);
      if (std::fabs(next) < 1e-18 * std::fabs(P) || std::fabs(next) > std::fabs(term)) break;
      term = next;
    }
    // chi = x - (nu/2 + 1/4) pi ; expand cos/sin(chi) in cos x, sin x to avoid
    // the argument-reduction error of forming x - const.
    double cchi, schi;
    const double r2 = 0.70710678118654752440;
    if (nu == 0) { cchi = r2 * (c + s); schi = r2 * (s - c); }
    else { cchi = r2 * (s - c); schi = -r2 * (c + s); }
    double jv = pre * (P * cchi - Q * schi);
    double yv = pre * (P * schi + Q * cchi);
    if (nu == 0) { j0 = jv; y0 = yv; } else { j1 = jv; y1 = yv; }
  }
}

LSG_HD inline void bessel_jy(double x, int nmax, double* J, double* Y) {
  if (x <= 0.0) {
    for (int n = 0; n <= nmax; ++n) { J[n] = (n == 0) ? 1.0 : 0.0; Y[n] = -std::numeric_limits<double>::infinity(); }
    return;
  }
  const int nm = std::max(nmax, 1);
  double Jl[64], Yl[2];
  if (x < 1.0) {
    const double h = 0.5 * x, h2 = h * h;
    double fact_n = 1.0, hpow = 1.0;
    for (int n = 0; n <= nm; ++n) {
      if (n > 0) { fact_n *= n; hpow *= h; }
      double term = hpow / fact_n, sum = term;
      for (int k = 1; k < 40; ++k) {
        term *= -h2 / (double(k) * double(n + k));
        sum += term;
        if (std::fabs(term) < 1e-18 * std::fabs(sum)) break;
      }
      Jl[n] = sum;
    }
    // Y_0 = (2/pi)[(ln(x/2)+gamma) J_0 + sum_{k>=1} (-1)^{k+1} H_k (x^2/4)^k/(k!)^2]
    const double lg = std::log(h) + kEulerGamma;
    double t = 1.0, Hk = 0.0, s0 = 0.0;
    for (int k = 1; k < 40; ++k) {
      t *= -h2 / (double(k) * double(k));
      Hk += 1.0 / k;
      double c = -t * Hk;
      s0 += c;
      if (std::fabs(c) < 1e-18 * std::fabs(s0)) break;
    }
    Yl[0] = (2.0 / kPi) * (lg * Jl[0] + s0);
    // Y_1 = -2/(pi x) + (2/pi) ln(x/2) J_1 - (1/pi)(x/2) sum_k [psi(k+1)+psi(k+2)] (-x^2/4)^k/(k!(k+1)!)
    double tk = 1.0, psi1 = -kEulerGamma, psi2 = 1.0 - kEulerGamma, s1 = 0.0;
    for (int k = 0; k < 40; ++k) {
      if (k > 0) { tk *= -h2 / (double(k) * double(k + 1)); psi1 += 1.0 / k; psi2 += 1.0 / (k + 1); }
      double c = (psi1 + psi2) * tk;
      s1 += c;
      if (k > 2 && std::fabs(c) < 1e-18 * std::fabs(s1)) break;
    }
    Yl[1] = -2.0 / (kPi * x) + (2.0 / kPi) * std::log(h) * Jl[1] - (1.0 / kPi) * h * s1;
  } else if (x < 25.0) {
    const double xm = std::max(x, double(nm));
    const int M = int(xm + 30.0 + 10.0 * std::cbrt(xm));
    double f2 = 0.0, f1 = 1e-100, f = 0.0, bs = 0.0, su = 0.0, sv = 0.0;
    for (int k = M; k >= 0; --k) {
      f = 2.0 * (k + 1.0) / x * f1 - f2;
      if (k <= nm) Jl[k] = f;
      if ((k & 1) == 0 && k != 0) {
        bs += 2.0 * f;
        su += (((k / 2) & 1) ? -1.0 : 1.0) * f / k;
      } else if (k > 1) {
        sv += (((k / 2) & 1) ? -1.0 : 1.0) * k / (double(k) * k - 1.0) * f;
      }
      f2 = f1; f1 = f;
    }
    const double s0 = bs + f;
    for (int k = 0; k <= nm; ++k) Jl[k] /= s0;
    const double ec = std::log(0.5 * x) + kEulerGamma;
    Yl[0] = (2.0 / kPi) * (ec * Jl[0] - 4.0 * su / s0);
    Yl[1] = (2.0 / kPi) * ((ec - 1.0) * Jl[1] - Jl[0] / x - 4.0 * sv / s0);
  } else {
    double j0 = 0.0, j1 = 0.0, y0 = 0.0, y1 = 0.0;
    hankel_asym01(x, j0, j1, y0, y1);
    Jl[0] = j0; Jl[1] = j1; Yl[0] = y0; Yl[1] = y1;
    for (int n = 1; n < nm; ++n) Jl[n + 1] = 2.0 * n / x * Jl[n] - Jl[n - 1];
  }
  for (int n = 0; n <= nmax; ++n) J[n] = Jl[n];
  Y[0] = Yl[0];
  if (nmax >= 1) Y[1] = Yl[1];
  for (int n = 1; n < nmax; ++n) Y[n + 1] = 2.0 * n / x * Y[n] - Y[n - 1];
}

}  // namespace lsg
