// Device-resident operator and solver state (HIP, gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "nca.hpp"

namespace lsg {

struct DevPass {
  std::string name;
  int kind = 0;        // 0 store, 1 accumulate
  int role = 0;        // 1: leaf M2L
  bool wave = false;   // one wavefront (not one workgroup) per destination
  bool lane = false;   // one lane per output row (overrides wave)
  bool lane_ok = false;  // the row map holds this pass (rows < 256 and < 2^23 destinations)
  int ndest = 0;
  int grid_cap = 0;       // one-workgroup kernel: at most this many workgroups (grid-stride); 0 = one per destination
  int64_t dest_base = 0;  // index into the concatenated destination arrays
  int64_t row_base = 0;   // index into the concatenated row map
  int64_t entries = 0, n_src = 0, n_dst = 0;
  int64_t bytes() const { return 16 * entries + 4 * n_src + 16 * n_src + 16 * n_dst; }
};

// Passes run as one launch of one workgroup per subtree: `first` is the
// position of the first fused pass in its list (up or down), n the number of
// passes (stages); task t's destinations of stage s are
// d_fdest[d_task_stage[t (n + 1) + s] ... d_task_stage[t (n + 1) + s + 1]).
struct FusedPasses {
  int first = 0, n = 0, tasks = 0;
  bool on = false;
  int* d_task_stage = nullptr;
  int* d_fdest = nullptr;
};

// Device workspace of the GMRES solver, kept between solves of the same
// restart length (Krylov basis, Hessenberg, rotations, reduction partials).
struct GmresWork {
  int N = 0, m = -1;
  double2 *V = nullptr, *w = nullptr, *partial = nullptr, *d1 = nullptr, *d2 = nullptr, *y = nullptr;
  double2* Z = nullptr;       // flexible GMRES: preconditioned directions, m x N
  double2 *H = nullptr, *sn = nullptr, *g = nullptr;
  double *cs = nullptr, *hn = nullptr, *beta = nullptr;
  double* hostbuf = nullptr;  // pinned: per-step residual estimates and norms
  // Set when the workspace lives in device memory the caller provides (carve):
  // hostbuf is then a device buffer, copied here after each synchronisation.
  double* host_mirror = nullptr;
  std::string reserve(int N, int m, bool flexible = false);
  // point every buffer into [base, base + cap_bytes) (device memory, not owned);
  // host_mirror must hold 2 (m + 2) doubles; fails if the workspace does not fit
  std::string carve(void* base, size_t cap_bytes, int N, int m, double* host_mirror);
  void release();
};
// bytes carve() needs for GMRES(m) on N unknowns
size_t gmres_work_bytes(int N, int m);

// The library's allocation-free workspace on `device`: the device address of a
// file-scope array of kScratchBytes, looked up once per device (the device
// must be current at the first call).
constexpr size_t kScratchBytes = size_t(128) << 20;
std::string scratch_workspace(int device, void** ptr);

// Right preconditioner of flexible GMRES: z = M(v), device vectors of N
// complex128, enqueued on stream s (M may be nonlinear).
struct Precond {
  virtual ~Precond() = default;
  virtual void apply(const double2* v, double2* z, hipStream_t s) = 0;
};

// One matvec = upward passes -> downward passes -> leaf L2L on the main
// stream, while the near field (from the start) and the leaf M2L (after the
// upward passes that write the charges it reads: the first one, the leaf
// level, on a uniform tree) run on a second stream and join before the leaf
// L2L. x is read and y written in place; the
// state's own x and y regions are used only by profile() and in-place calls.
struct DeviceOperator {
  hipStream_t stream = nullptr;   // main stream (all public calls are ordered on it)
  hipStream_t aux = nullptr;      // second stream, same default priority
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_leaf = nullptr;
  int device = 0;
  int max_lds = 65536;            // LDS bytes per workgroup on this device
  int N = 0, rank = 0, n_leaves = 0;
  bool find_phi = true;
  int64_t n_state = 0;
  double2* state = nullptr;    // [x | y | charges | potentials]
  double2* ops = nullptr;
  int* src = nullptr;
  int* d_rows = nullptr;
  int* d_ncol = nullptr;
  int* d_dst = nullptr;
  long long* d_op = nullptr;
  long long* d_srcoff = nullptr;
  // per pass, one entry per output row: (destination within the pass) << 8 | row,
  // destinations ordered by decreasing column count, rows of one destination adjacent
  unsigned* d_rowmap = nullptr;
  // Consecutive passes fused per subtree (pass_fused_kernel: one workgroup per
  // task, a barrier between the passes' stages; used when `on`): the leaf
  // level's upward pass and the two above it (fup), and the last two
  // wavefront-sized downward passes before the first large one (fdown).
  FusedPasses fup, fdown;
  double2* T = nullptr;        // 9 x rank x rank, column-major per slot
  int* nbr = nullptr;          // n_leaves x 9
  double* qk2 = nullptr;       // N
  // contrast supplied by the caller for a volume-potential operator
  // (find_phi = false): matvec then returns x + qk2_user .* (V x)
  double* qk2_user = nullptr;
  int64_t n_entries = 0;        // operator entries of all passes
  // FP32 copies (operators, near-field tables, pass state), made by enable_fp32
  float2* ops32 = nullptr;
  float2* T32 = nullptr;
  float2* state32 = nullptr;
  GmresWork work;
  int near_leaves_per_thread = 2;     // near-field variant of the FMA kernel (autotuned)
  // FP64 near field on MFMA (autotuned against the FMA kernel): the tables as
  // real-ified 16 x 4 fragments per (slot, row tile, k-step, lane)
  double* Tm = nullptr;
  bool near_mfma = false;
  // several table sets (fused plans): set of each 16-leaf group, and the
  // fragment stride of one set; the near field then runs on MFMA only
  int n_table_sets = 1;
  int* near_set = nullptr;
  long long tm_set_stride = 0;
  // fused plans: system of every destination, and the per-system activity
  // flags (device, owned by the caller) that make the passes skip the
  // destinations of inactive systems; nullptr: every destination runs
  int* d_dest_sys = nullptr;
  const int* sys_active = nullptr;
  hipGraphExec_t graph_exec = nullptr;  // captured FP64 matvec (capture_graph)
  const double2** xy_dev = nullptr;     // its input/output addresses
  std::vector<std::string> tuning;      // autotuner decisions
  std::vector<DevPass> passes;
  std::vector<int> up, down, leaf_m2l, leaf_l2l;  // pass indices by role
  // near field of an adaptive tree as passes (stored operators); empty for a
  // uniform tree, whose near field is the 9-slot table kernel
  std::vector<int> near;
  // the leaf M2L starts once upward passes 0..leaf_after (in `up` order) are
  // done: the last one writing a charge it reads
  int leaf_after = 0;
  // when the leaf M2L starts on the second stream (autotuned on the whole
  // matvec): 0 after upward pass leaf_after, 1 after every upward pass, 2
  // before the last downward pass
  int m2l_start = 0;
  int64_t near_bytes = 0;

  std::string upload(const Plan& plan);
  void release();
  // y = A x on the main stream; x, y device pointers of N complex128 (may alias)
  void matvec(const double2* x, double2* y);
  // same with FP32 operators and FP32 accumulation inside each pass (x, y FP64);
  // enable_fp32 must have succeeded
  void matvec32(const double2* x, double2* y);
  std::string enable_fp32();
  // capture the FP64 matvec into a hipGraph; matvec() then sets the addresses
  // and replays it (one launch)
  std::string capture_graph();
  void drop_graph();
  // time the kernel variants of every pass and of the near field, keep the faster
  std::string autotune(int reps, int* changed);
  // event-timed per-pass averages over reps matvecs (passes run serialised)
  std::string profile(int reps, std::vector<double>& ms, std::vector<int64_t>& bytes, std::vector<std::string>& names);
  // average time of the full two-stream matvec (ms)
  double time_matvec(int reps, bool fp32 = false);
};

// Restarted GMRES (classical Gram-Schmidt with one reorthogonalisation,
// complex Givens) with device-resident Krylov basis, Hessenberg and rotations.
struct GmresResult {
  int iterations = 0, restarts = 0, converged = 0, refinements = 0;
  double rel_residual = 0, est_residual = 0, seconds = 0;
  int status = 0;  // 0 ok, else LSG_ERR_*
  std::string error;
};
// fp32: every matvec of the Krylov process uses the FP32 operators (FP64 vectors,
// Gram-Schmidt and Givens)
// ext: a workspace carved from caller-provided memory (nothing is allocated);
// otherwise op.work, reserved on first use
GmresResult device_gmres(DeviceOperator& op, const double2* b_dev, double2* x_dev, int m, int max_restarts, double tol,
                         bool fp32 = false, Precond* prec = nullptr, GmresWork* ext = nullptr);
// FP64 iterative refinement around FP32-operator GMRES: r = b - A x with the FP64
// operator, d ~ A32^{-1} r by GMRES(m) to relative residual max(inner_tol, tol/2 ||b||/||r||),
// x += d; stops when ||b - A x|| <= tol ||b|| (FP64) or after max_outer refinements.
GmresResult device_gmres_ir(DeviceOperator& op, const double2* b_dev, double2* x_dev, int m, int max_restarts,
                            double tol, double inner_tol, int max_outer);

}  // namespace lsg
