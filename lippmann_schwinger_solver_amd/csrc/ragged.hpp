// Ragged batch: independent scatterers (uniform trees of different depth,
// wavenumber and contrast) fused into one plan and solved together by a
// batched GMRES with one Krylov basis, Hessenberg and stopping point per
// system over segments of one concatenated vector.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "device.hpp"

namespace lsg {

// The fused plan of independent plans P_0..P_{n-1} (uniform trees, the same
// nodes per leaf): one state vector [x_all | y_all | charge and potential
// slots of P_0, P_1, ...], where x_all and y_all concatenate the systems'
// vectors in order; the k-th upward pass of every plan becomes part of the
// k-th fused upward pass (likewise downward, leaf M2L and leaf L2L), which
// keeps each plan's own pass order (the plans share no data). The near-field
// tables become one table set per plan. offsets (n + 1) receives the first
// unknown of each system.
std::string fuse_plans(const std::vector<const Plan*>& plans, Plan& fused, std::vector<int>& offsets);

struct RaggedWork {
  int m = -1;
  double2 *V = nullptr, *w = nullptr, *partial = nullptr, *d1 = nullptr, *d2 = nullptr, *y = nullptr;
  double2 *H = nullptr, *sn = nullptr, *g = nullptr;
  double *cs = nullptr, *hn = nullptr, *beta = nullptr, *bnorm = nullptr, *tolq = nullptr;
  int* kq = nullptr;
  double* host_rel = nullptr;   // pinned, n
  double* host_est = nullptr;   // pinned, m x n
  int* host_active = nullptr;   // pinned, n
  int* host_kq = nullptr;       // pinned, n
  std::string reserve(int Ntot, int nrb, int n, int m);
  void release();
};

struct RaggedOperator {
  DeviceOperator op;                 // the fused plan
  int n = 0;                         // systems
  std::vector<int> off;              // n + 1 offsets into the concatenated vectors
  // row blocks of at most 512 rows, none straddling two systems
  int nrb = 0;
  int *rb_row0 = nullptr, *rb_nrow = nullptr, *rb_sys = nullptr;
  int* sys_rb0 = nullptr;            // n + 1: first row block of each system
  // per-system activity flags read by the fused operator's kernels (the
  // matvec skips inactive systems); all 1 outside the solver
  int* active = nullptr;
  RaggedWork work;
  std::string setup(const Plan& fused, const std::vector<int>& offsets);
  void all_active();                 // every system active (enqueued on the operator's stream)
  void release();
};

struct RaggedStats {
  int iterations = 0, restarts = 0, converged = 0;
  double rel_residual = 0;
};
struct RaggedResult {
  std::vector<RaggedStats> sys;
  int status = 0;
  std::string error;
  double seconds = 0;
};
// GMRES(m) on every system at once (B, X: concatenated device vectors, X the
// initial guesses): per system classical Gram-Schmidt with one
// reorthogonalisation, complex Givens, and the stopping rules of the
// single-system solver (estimate at 0.5 tol, breakdown below 1e-8); every
// cycle opens with each system's true FP64 residual, which decides convergence.
RaggedResult ragged_gmres(RaggedOperator& R, const double2* B, double2* X, int m, int max_restarts, double tol);

}  // namespace lsg
