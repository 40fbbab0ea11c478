// Problem definition: contrast q(x), incident plane wave, right-hand side.
// Restates GMRES/GMRES.cpp:5-125 (contrast functions and userkernel) and adds
// the seeded Gaussian-bump contrast and an incidence angle.
#pragma once
#include <cmath>
#include <complex>
#include "numerics.hpp"
#include "../../include/lsg.h"

namespace lsg {

// GMRES/GMRES.cpp:5-8
inline double q_gaussian(double x, double y) {
  double R2 = x * x + y * y;
  return 1.5 * std::exp(-160.0 * R2);
}

// GMRES/GMRES.cpp:10-59 (20 fixed centres, a = 0.0013)
inline double q_multiple_gaussians(double x, double y) {
  static const double cxs[20] = {-1.17182014706864, -0.861066625441127, -1.14533939124057, -0.856639124568525,
                                 -0.620066350133170, -0.457064031294315, -0.482782006538245, -0.635070123186999,
                                 0.108594058866164, -0.0315722590552472, -0.177274727610547, 0.0342987980936911,
                                 0.369662097773613, 0.591444264226524, 0.401225715100565, 0.666822699314234,
                                 1.10327781188617, 0.827519268441487, 0.873411278211136, 1.09482906183464};
  static const double cys[20] = {-0.361066625441127, 0.0719520938231629, 0.354660608759432, 1.14336087543147,
                                 -1.12006635013317, -0.457064031294315, 0.364929876813001, 0.802261227250738,
                                 -0.394084977497864, -0.0315722590552472, 0.322725272389453, 1.03429879809369,
                                 -1.13033790222639, 0.0137164199764117, 0.401225715100565, 1.16682269931423,
                                 -0.345187707108795, -0.172480731558513, 0.373411278211136, 1.09482906183464};
  double q = 0.0;
  const double a = 0.0013;
  for (int i = 0; i < 20; ++i) {
    double R2 = (x - cxs[i]) * (x - cxs[i]) + (y - cys[i]) * (y - cys[i]);
    q += 1.5 * std::exp(-R2 / a);
  }
  return q;
}

// GMRES/GMRES.cpp:66-83
inline double q_plasma(double x, double y) {
  const double C = 0.4987;
  double psi = 1.0 - (x - 0.15 * (1 - x * x)) * (x - 0.15 * (1 - x * x)) - C * (1.0 + 0.3 * x) * (1.0 + 0.3 * x) * y * y;
  static const double a[5] = {0.45, 0.195, 0.51, 0.195, 0.63};
  static const double xc[5] = {0.8, 0.54, -0.14, -0.5, 0.18};
  static const double yc[5] = {0.0, -0.28, 0.7, -0.01, 0.8};
  if (psi <= 0.05) return 0.0;
  double g = 0.0;
  for (int i = 0; i < 5; ++i) {
    double t = (x - xc[i]) * (x - xc[i]) + (y - yc[i]) * (y - yc[i]);
    g += a[i] * std::exp(-t / 0.01);
  }
  return -1.5 * (psi - 0.05) - g * std::cos(0.9 * y);
}

// GMRES/GMRES.cpp:61-64
inline double q_flat_bump(double x, double y) {
  double R2 = x * x + y * y;
  return 0.5 * std::erfc(5.0 * (R2 - 1.0));
}

// GMRES/GMRES.cpp:85-91
inline double q_cavity(double x, double y) {
  double R = std::sqrt(x * x + y * y);
  double theta = std::atan2(y, x);
  return (1.0 - std::pow(std::sin(0.5 * theta), 500)) * std::exp(-2000 * (0.1 - R * R) * (0.1 - R * R));
}

// GMRES/GMRES.cpp:93-98
inline double q_lens(double x, double y) {
  double R = std::sqrt(x * x + y * y);
  return 4 * (y - 0.1) * (1.0 - std::erf(25.0 * (R - 0.3)));
}

struct Problem {
  double kappa = 0;
  double inc_angle = 0;
  lsg_contrast contrast{};

  // userkernel::ContrastFunction, GMRES/GMRES.cpp:104-117
  double q(double x, double y) const {
    switch (contrast.kind) {
      case 0: return q_gaussian(x, y);
      case 1: return q_multiple_gaussians(x, y);
      case 2: return q_plasma(x, y);
      case 3: return q_flat_bump(x, y);
      case 4: return q_cavity(x, y);
      case 5: return q_lens(x, y);
      default: {
        double s = 0.0;
        for (int i = 0; i < contrast.n_bumps; ++i) {
          double dx = x - contrast.cx[i], dy = y - contrast.cy[i];
          double w = contrast.width[i];
          s += contrast.amp[i] * std::exp(-(dx * dx + dy * dy) / (w * w));
        }
        return s;
      }
    }
  }
  // IncidenceFunction (GMRES/GMRES.cpp:118-121) generalised to a direction.
  // At inc_angle == 0 the phase is kappa*x exactly, as in the reference.
  cplx u_inc(double x, double y) const {
    double ph = (inc_angle == 0.0) ? kappa * x : kappa * (x * std::cos(inc_angle) + y * std::sin(inc_angle));
    return std::exp(cplx(0.0, 1.0) * ph);
  }
  // RHSFunction, GMRES/GMRES.cpp:122-125
  cplx rhs(double x, double y) const { return -1.0 * kappa * kappa * q(x, y) * u_inc(x, y); }
};

}  // namespace lsg
