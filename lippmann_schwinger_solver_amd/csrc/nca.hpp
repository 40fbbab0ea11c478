// Nested cross approximation (NCA) skeletons and the fused far-field operators
// of the DAFMM, flattened into destination-major passes for the device.
//
// Reference: GMRES/ACA.hpp:253-481 (ACA_only_nodes),
//            GMRES/G_FMM2DTree.hpp:603-658 (getMatrixEntry),
//            :2013-2385 (getNodes_LFR), :3232-3571 (getNodes_HFR),
//            :3146-3230 (getUserCheckPoints),
//            :2387-2500 (LFR_M2M), :3573-3827 (HFR_M2M),
//            :2502-2651 (LFR_M2L), :3829-3942 (HFR_M2L),
//            :2653-2841 (LFR_L2L), :3944-4236 (HFR_L2L),
//            :2884-2960 (evaluate_NearField_Using_Precomputations_LS).
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "numerics.hpp"
#include "problem.hpp"
#include "tables.hpp"
#include "tree.hpp"

namespace lsg {

// Far-field matrix entry of the (optionally contrast-scaled) volume operator
// between a target node i and a source node j in a different, well separated
// leaf: the Graf series of G = (i/4) H0 against the projected source basis.
struct EntryKernel {
  const Tree* tree = nullptr;
  int rank = 0;
  double kappa = 0;
  bool find_phi = true;
  std::vector<double> qk2;                // kappa*kappa*q(x_i)
  std::vector<std::vector<cplx>> M;       // per level: (2*kSeries+1) x rank row-major
  std::vector<int> leaf_level;            // per leaf
  std::vector<double> leaf_cx, leaf_cy;   // per leaf

  // h[n+kSeries] = (J_n(R) + i Y_n(R)) e^{i n theta}, R = kappa |x_i - c_leaf|
  void hvec(int i, int leaf, cplx* h) const;
  cplx entry(const cplx* h, int i, int j) const;
  // out (col-major, nr x nc) = K(rows, cols)
  void block(const int* rows, int nr, const int* cols, int nc, cplx* out) const;
};

struct AcaResult {
  std::vector<int> row_bases, col_bases;
  int rank = 0;
  std::vector<cplx> Ar;  // rank x ncols, row-major: K(row pivots, all cols)
};
// Partially pivoted ACA of K(rows, cols) with the reference's stopping rules.
// abs_exit: the absolute exit on the new cross vectors' norms (the reference's 1e-8)
void aca_only_nodes(const EntryKernel& K, const std::vector<int>& rows, const std::vector<int>& cols, double eps,
                    AcaResult& out, double abs_exit = 1e-8);

// One destination of a device pass: rows x ncol column-major operator,
// ncol source state indices, contiguous destination [dst, dst+rows).
struct PassHost {
  std::string name;
  int level = 0;
  int kind = 0;  // 0 = store (y = A x), 1 = accumulate (y += A x)
  std::vector<int> rows, ncol, dst;
  std::vector<int64_t> op_off, src_off;
  int64_t entries = 0, n_src = 0, n_dst = 0;
};

struct Plan {
  int N = 0, rank = 0, n_leaves = 0;
  int64_t n_state = 0;      // complex slots: [x | y | outgoing charges | incoming potentials]
  std::vector<cplx> ops;    // operator entries of every pass
  std::vector<int> src;     // source state indices of every pass
  std::vector<PassHost> passes;   // execution order: upward, downward, leaf (after near field)
  // near field (leaf level of a uniform tree): y = x + qk2 .* sum_s T_s x_nbr(s)
  std::vector<cplx> T;      // 9 x rank x rank, row-major per slot (per table set)
  // several near-field table sets (a fused plan of independent trees): T holds
  // the sets back to back and near_group_set names the set of each group of
  // 16 consecutive leaves; empty for one set
  std::vector<int> near_group_set;
  // fused plans: the system of every destination (passes in order); empty
  // for one system
  std::vector<int> dest_sys;
  std::vector<int> nbr;     // n_leaves x 9 leaf indices, -1 where absent
  std::vector<double> qk2;  // N
  bool find_phi = true;
  int64_t E_far = 0, E_near = 0;
  int max_rank = 0;
  std::vector<int> level_max_rank;
  // Skeleton dump, owners in the order: level nLevels..2, LFR boxes by index
  // then HFR (box, cone); per owner: level, box, cone (-1 = LFR), out_active,
  // in_active, |out_check|, |out_charge|, |in_check|, |in_charge|, |ucp|, then
  // the five lists (node ids).
  std::vector<int> skeleton;
  // Per owner, in the skeleton dump's order: level, box, cone (-1 = LFR), then
  // the outgoing and the incoming ACA's rows, columns and rank (-1 when an
  // operand list was empty and no ACA ran).
  std::vector<int> aca;
  std::vector<cplx> far_leaf;  // far-field table of the leaf level (after rounding)
  // build time of: far-field tables, NCA skeletons, operator fill, near-field tables
  double phase_seconds[4] = {0, 0, 0, 0};
};

// One ACA of K(rows, cols) (rows and cols are node lists).
struct AcaJob {
  const std::vector<int>* rows = nullptr;
  const std::vector<int>* cols = nullptr;
  AcaResult* out = nullptr;
};
// One entry block K(rows, cols), written column-major (nr x nc) to out.
struct BlockJob {
  const int* rows = nullptr;
  int nr = 0;
  const int* cols = nullptr;
  int nc = 0;
  cplx* out = nullptr;
};

// X (n x nrhs) = A^-1 B for a square A (n x n), all column-major: the factor
// and basic solve of ColPivQR (column-pivoted Householder QR).
struct QrSolveJob {
  const cplx* A = nullptr;
  int n = 0;
  const cplx* B = nullptr;
  int nrhs = 0;
  cplx* X = nullptr;
};

// Entry generation and ACA on the GPU: the entry kernel, the pivot order and
// the stopping rules of EntryKernel and aca_only_nodes, evaluated in device
// arithmetic (results agree with the host to rounding; pivots may differ
// where two candidates tie to rounding).
class DeviceNca {
 public:
  virtual ~DeviceNca() {}
  virtual std::string aca(const std::vector<AcaJob>& jobs, double eps, double abs_exit = 1e-8) = 0;
  virtual std::string blocks(const std::vector<BlockJob>& jobs) = 0;
  // Batched column-pivoted QR factor + solve (ColPivQR semantics in device
  // arithmetic); jobs whose n exceeds the device limit are solved on the host.
  virtual std::string qr_solve(const std::vector<QrSolveJob>& jobs) = 0;
  // The integrals of near_integrals for every table in `tabs` (each rank x nb,
  // row-major), by the same adaptive quadrature on the GPU.
  virtual std::string near_integrals(const BasisSet& bs, double kappa, const std::vector<NearGeom>& tabs,
                                     std::vector<std::vector<cplx>>& V1) = 0;
};
typedef DeviceNca* (*DeviceNcaFactory)(const EntryKernel& K, std::string& err);

struct BuildOptions {
  int TOL_POW = 10;
  double aca_abs_exit = 1e-8;   // the ACA's absolute exit (the reference's 1e-8)
  bool find_phi = true;
  // when set, NCA skeletons and operator entries are computed on the GPU
  DeviceNcaFactory device = nullptr;
  // when not empty, the far-field and colleague tables are read from the
  // reference's cache files under <table_dir>/precomputations
  std::string table_dir;
};

// Factory of the HIP implementation (csrc/nca_device.hip).
DeviceNca* make_device_nca(const EntryKernel& K, std::string& err);

// The entry kernel of a built tree (far-field tables of the leaf levels, k^2 q
// at the nodes). With table_dir set the tables are read from the reference's
// cache files under it instead of computed. Returns "" on success.
std::string make_entry_kernel(const Tree& tree, const BasisSet& bs, const Problem& prob, bool find_phi,
                              const std::string& table_dir, EntryKernel& K);

// Build tables, NCA skeletons and the plan. Returns "" on success.
std::string build_plan(const Tree& tree, const BasisSet& bs, const Problem& prob, const BuildOptions& opt, Plan& plan);

}  // namespace lsg
