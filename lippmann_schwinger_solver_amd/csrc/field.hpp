// Re-interpolation of a leaf-node field onto arbitrary points in the domain
// (the Phi -> Psi step of the reference's field validation).
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include "tables.hpp"
#include "tree.hpp"

namespace lsg {

struct FieldInterp {
  int nleaves = 0, rank = 0, nb = 0, side = 0;
  double r = 0;
  double* qpinv = nullptr;   // nb x rank
  int* ab = nullptr;         // basis degrees (a, b) per basis function
  int* leaf_at = nullptr;    // leaf index of box (column, row)
  double *cx = nullptr, *cy = nullptr;  // leaf centres per column / row
  double2* coef = nullptr;   // nleaves x nb basis coefficients
  int* bad = nullptr;        // set when a point lies outside [-L, L)^2
  std::string setup(const Tree& t, const BasisSet& bs);
  // out[i] = field at (x[i], y[i]); device arrays; *outside when any point is
  // outside the domain (its value is 0)
  std::string run(const double2* f, const double* x, const double* y, long long n, double2* out, hipStream_t s,
                  bool* outside);
  void release();
};

}  // namespace lsg
