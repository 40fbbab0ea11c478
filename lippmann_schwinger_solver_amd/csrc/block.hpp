// Multi right-hand-side (block) matvec and batched GMRES on the device.
//
// Block vectors use a row-split layout: row n of an N x nb block holds
// [Re v(n, 0..nb) | Im v(n, 0..nb)] (2*nb doubles), so one gathered source
// row feeds all right-hand sides with coalesced loads. Every pass of the
// single-vector plan (nca.cpp) becomes, per destination, a real GEMM
//   [Yr; Yi] (2R x nb) = [[Ar, -Ai], [Ai, Ar]] (2R x 2C) [Xr; Xi] (2C x nb)
// on MFMA tiles (v_mfma_f64_16x16x4f64, or v_mfma_f32_16x16x4f32 for the
// FP32 operator), one wavefront per destination. On a uniform tree one
// kernel per leaf computes the near field (the 9 colleague tables side by
// side, rank x 9 rank, shared by all leaves), the leaf M2L and leaf L2L
// products and the x term, and writes the leaf's output rows once.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include <vector>
#include "device.hpp"

namespace lsg {

constexpr int kBlockRhs = 64;  // right-hand sides per block (4 MFMA column tiles)

// Device workspace of the batched GMRES, kept between solves of the same
// restart length (64-system Krylov basis, Hessenbergs, rotations, partials).
struct BlockWork {
  int N = 0, m = -1;
  double *V = nullptr, *w = nullptr;
  _Float16* V16 = nullptr;  // FP16 basis of the FP32 inner solves
  float* w32 = nullptr;   // FP32 product A v_j of the FP32 inner solves (when the leaf kernel writes every output row)
  float* x32 = nullptr;   // FP32 copy of the current FP16 basis vector: the next block matvec's input
  double* Z = nullptr;    // flexible GMRES: the m preconditioned directions (FP64), allocated by flex()
  double2 *partial = nullptr, *h = nullptr, *y = nullptr, *H = nullptr, *sn = nullptr, *g = nullptr;
  double *cs = nullptr, *hn = nullptr, *bnorm = nullptr, *tolq = nullptr;
  int *active = nullptr, *kq = nullptr;
  double *host_rel = nullptr, *host_est = nullptr;  // pinned
  int* host_active = nullptr;                        // pinned
  std::string reserve(int N, int m);   // everything but the basis
  std::string basis(bool f16);         // the FP64 or FP16 basis, on first use
  std::string flex();                  // Z, on first use
  void release();
};

struct BlockOperator {
  DeviceOperator* op = nullptr;
  int nb = kBlockRhs;
  double* bstate = nullptr;   // (n_state + 1) rows x 2nb; row n_state stays zero
  // uniform tree: per leaf, its destination in the leaf M2L and leaf L2L
  // passes (2 per leaf, -1 for none) when the leaf kernel folds them in
  int* leaf_dest = nullptr;
  bool leaf_fused = false;
  int* near_src = nullptr;    // n_leaves x 9 rank near-field source rows (zero row for an absent colleague)
  // per right-hand-side contrast (N x 64, row-major [n][q]) for a volume-potential
  // operator: Y[:, q] = X[:, q] + qk2_cols[:, q] .* (V X)[:, q]
  double* qk2_cols = nullptr;
  float* bstate32 = nullptr;  // FP32 pass state for matvec32
  // right-hand-side tiles of 16 per wavefront in the FP32 block passes (1 or
  // 2), chosen by timing when the FP32 operator is enabled
  int nt32 = 1;
  // split-BF16 operator of the inner solves (a = ah + al in bf16, products
  // (ah + al)(bh + bl) as four BF16 MFMA terms with FP32 sums; 16 significant bits),
  // built by enable_sb: per destination, chunks of 16 columns of 128-byte rows
  char* ops_sb = nullptr;
  char* T_sb = nullptr;            // the near-field tables in the same layout
  long long* d_opsb = nullptr;    // per destination: its first 128-byte row in ops_sb
  int nt_sb = 1;                   // right-hand-side tiles per wavefront (chosen by timing)
  bool sb_split = false;         // near field on the second stream beside the passes (chosen by timing)
  std::vector<std::string> notes;  // timed choices (tiles per wavefront, near field beside the passes)
  bool inner_sb = true;          // the FP32 inner solves use the split-BF16 kernels
  bool sb_active = false;        // set while they run: FP32 block matvecs go to those kernels
  std::vector<int> pass_max_rows;  // largest destination of each pass (selects the row-tile count)
  std::vector<int> pass_nseg;      // column segments per destination (1: not split)
  void* split_part = nullptr;      // partial products of split passes
  BlockWork work;
  BlockWork outer;    // the FP64 outer iteration of block_fgmres
  std::string setup(DeviceOperator& o);
  std::string enable_fp32();  // FP32 operators (DeviceOperator::enable_fp32) and pass state
  std::string enable_sb();    // the split-BF16 operator (after enable_fp32)
  // the inner solves' arithmetic: enable_sb or enable_fp32 (inner_sb)
  std::string enable_inner() { return inner_sb ? enable_sb() : enable_fp32(); }
  void release();
  // Y = A X for one block (row-split, nb columns), on op->stream
  void matvec(const double* X, double* Y);
  // same with FP32 operators, FP32 MFMA and FP32 pass state (Y FP64; X FP64,
  // rounded into the pass state, or FP32 and read in place)
  void matvec32(const double* X, double* Y);
  void matvec32(const float* X, double* Y);
  // same with the product stored in FP32; only where the leaf kernel writes
  // every output row (can_store_f32)
  void matvec32(const float* X, float* Y);
  // FP16 input (the inner solves' basis), converted into the FP32 pass state
  void matvec32(const _Float16* X, double* Y);
  void matvec32(const _Float16* X, float* Y);
  bool can_store_f32() const { return op && leaf_fused && !qk2_cols; }
};

// sb_active set to `on` for the lifetime of the object, then restored
struct SbScope {
  BlockOperator& bo;
  bool prev;
  SbScope(BlockOperator& b, bool on) : bo(b), prev(b.sb_active) { bo.sb_active = on; }
  ~SbScope() { bo.sb_active = prev; }
};

// Batched restarted GMRES: nb independent systems A x_q = b_q advanced in
// lockstep with the block matvec; classical Gram-Schmidt (one projection,
// ||w - V h||^2 = ||w||^2 - |h|^2), complex Givens per system; a system that
// reaches 0.5 tol in its Givens estimate, or breaks down, is frozen for the
// rest of the cycle. Every cycle opens with the true FP64 residual of each
// system, which decides convergence (rel_residual <= tol for all nrhs).
struct BatchResult {
  int iterations = 0, restarts = 0, converged = 0;  // converged: systems meeting tol
  int refinements = 0;
  double max_rel_residual = 0, seconds = 0;
  int status = 0;
  std::string error;
};
// B, X: N x nb row-split device blocks (X holds the initial guesses); only the
// first nrhs columns are solved, the others must be zero.
// fp32: the block matvecs use the FP32 pass state with split-BF16 MFMA
// operators (bo.inner_sb) or FP32 operators and FP32 MFMA, and the Krylov basis
// is stored in FP16; tol_per (host, nrhs):
// per-system tolerances instead of tol; allow_zero: zero right-hand sides are
// accepted (their systems are converged with x unchanged).
// x_zero: X is zero on entry (the first residual is B, no matvec).
BatchResult block_gmres(BlockOperator& bo, const double* B, double* X, int nrhs, int m, int max_restarts, double tol,
                        double* rel_residuals /* host, nrhs */, bool fp32 = false, const double* tol_per = nullptr,
                        bool allow_zero = false, bool x_zero = false);
// FP64 iterative refinement of the batch around FP32 batched GMRES:
// R = B - A X (FP64 block matvec); systems above tol solve A D = R with the FP32
// block operator to max(inner_tol, tol ||b_q|| / (2 ||r_q||)); X += D.
BatchResult block_gmres_ir(BlockOperator& bo, const double* B, double* X, int nrhs, int m, int max_restarts,
                           double tol, double inner_tol, int max_outer, double* rel_residuals);
// FP64 flexible GMRES of the batch, outer Krylov dimension K per cycle (at most
// max_cycles cycles), preconditioned by the FP32 batched GMRES: outer step k
// solves A z_k = v_k with the FP32 block operator (m steps per cycle,
// max_restarts cycles, to inner_tol), w = A z_k in FP64 is orthogonalised
// against the FP64 outer basis (classical Gram-Schmidt, Givens per system), and
// X += Z y closes the cycle. Unlike refinement, every correction is re-weighted
// by the outer least-squares problem. Each cycle opens with the true FP64
// residual of every system, which decides convergence.
BatchResult block_fgmres(BlockOperator& bo, const double* B, double* X, int nrhs, int m, int max_restarts,
                         double tol, double inner_tol, int K, int max_cycles, double* rel_residuals);

// column-major complex (x + 2*ld*j) <-> row-split block, nrhs <= nb columns
void to_split(const double2* x, long long ld, int nrhs, int N, int nb, double* out, hipStream_t s);
void from_split(const double* in, int N, int nb, int nrhs, double2* x, long long ld, hipStream_t s);

}  // namespace lsg
