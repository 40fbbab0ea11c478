// C ABI of the solver (include/lsg.h, include/lsg_host.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lsg.h"
#include "../../include/lsg_host.h"
#include "../../include/lsg_learn.h"
#include "block.hpp"
#include "device.hpp"
#include "field.hpp"
#include "learn.hpp"
#include "nca.hpp"
#include "precomp.hpp"
#include "problem.hpp"
#include "ragged.hpp"
#include "tables.hpp"
#include "tree.hpp"

using namespace lsg;

namespace {

thread_local std::string g_create_error;
thread_local std::string g_host_error;

struct HostBuild {
  lsg_inputs in{};
  BasisSet bs;
  Tree tree;
  Problem prob;
  Plan plan;
  double setup_seconds = 0;
  std::string table_dir;  // copy of lsg_inputs.precomputations_dir ("" = compute the tables)
  std::vector<cplx> far;  // leaf-level M table
};

std::string validate(const lsg_inputs* in) {
  if (!in) return "inputs is NULL";
  if (in->nChebNodes < 2 || in->nChebNodes > 8) return "nChebNodes must be in [2, 8]";
  if (in->degreeOfBases < 1 || in->degreeOfBases > in->nChebNodes) return "degreeOfBases must be in [1, nChebNodes]";
  if (!(in->L > 0) || !std::isfinite(in->L)) return "L must be positive";
  if (!(in->kappa > 0) || !std::isfinite(in->kappa)) return "kappa must be positive";
  if (in->nCones_LFR < 1) return "nCones_LFR must be >= 1";
  if (in->TOL_POW < 1 || in->TOL_POW > 16) return "TOL_POW must be in [1, 16]";
  if (in->tree_mode != 0 && in->tree_mode != 1) return "tree_mode must be 0 (uniform) or 1 (adaptive)";
  if (in->tree_mode == 0 && (in->nLevelsUniform < 2 || in->nLevelsUniform > 9))
    return "nLevelsUniform must be in [2, 9]";
  if (in->tree_mode == 1 && (in->treeAdaptivity < 1 || in->treeAdaptivity > 12))
    return "treeAdaptivity must be in [1, 12]";
  if (in->Qchoice != in->contrast.kind) return "Qchoice must equal contrast.kind";
  if (in->contrast.kind < 0 || in->contrast.kind > 6) return "contrast.kind must be in [0, 6]";
  if (in->contrast.kind == 6 && (in->contrast.n_bumps < 0 || in->contrast.n_bumps > LSG_MAX_BUMPS))
    return "contrast.n_bumps out of range";
  if (in->setup_device != 0 && in->setup_device != 1) return "setup_device must be 0 or 1";
  if (in->aca_exit_pow != 0 && (in->aca_exit_pow < 8 || in->aca_exit_pow > 20))
    return "aca_exit_pow must be 0 (the reference's 1e-8) or in [8, 20]";
  return "";
}

// No entry of v[0, n) is NaN or Inf: an IEEE double is not finite exactly
// when its exponent bits are all ones; the integer test vectorises.
bool all_finite(const double* v, int64_t n) {
  uint64_t bad = 0;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t u;
    std::memcpy(&u, v + i, sizeof(u));
    bad |= uint64_t((u & 0x7ff0000000000000ull) == 0x7ff0000000000000ull);
  }
  return bad == 0;
}

// Returns LSG_OK, LSG_ERR_ARG (invalid inputs) or LSG_ERR_SETUP; message in err.
int host_build(const lsg_inputs* in, HostBuild& hb, std::string& err, bool tree_only = false) {
  err = validate(in);
  if (!err.empty()) return LSG_ERR_ARG;
  std::string e;
  auto t0 = std::chrono::steady_clock::now();
  hb.in = *in;
  hb.table_dir = in->precomputations_dir ? in->precomputations_dir : "";
  hb.in.precomputations_dir = nullptr;  // the caller's string need not outlive the call
  hb.bs.init(in->nChebNodes, in->degreeOfBases);
  TreeParams tp;
  tp.nCones_LFR = in->nCones_LFR;
  tp.nChebNodes = in->nChebNodes;
  tp.L = in->L;
  tp.kappa = in->kappa;
  tp.yes2DFMM = in->yes2DFMM;
  tp.nLevelsUniform = in->nLevelsUniform;
  tp.tree_mode = in->tree_mode;
  tp.treeAdaptivity = in->treeAdaptivity;
  hb.prob.kappa = in->kappa;
  hb.prob.inc_angle = in->inc_angle;
  hb.prob.contrast = in->contrast;
  const Problem* prob = &hb.prob;
  tp.contrast = [prob](double x, double y) { return prob->q(x, y); };
  e = hb.tree.build(tp, hb.bs);
  hb.tree.p.contrast = nullptr;  // refers to hb.prob, used only while the tree grows
  if (!e.empty()) { err = e; return LSG_ERR_SETUP; }
  if (tree_only) return LSG_OK;
  BuildOptions bo;
  bo.TOL_POW = in->TOL_POW;
  bo.aca_abs_exit = in->aca_exit_pow ? std::pow(10.0, -in->aca_exit_pow) : 1e-8;
  bo.find_phi = in->find_phi != 0;
  if (in->setup_device) bo.device = make_device_nca;
  bo.table_dir = hb.table_dir;
  e = build_plan(hb.tree, hb.bs, hb.prob, bo, hb.plan);
  if (!e.empty()) { err = e; return LSG_ERR_SETUP; }
  hb.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return LSG_OK;
}

void fill_info(const HostBuild& hb, lsg_info* info) {
  std::memset(info, 0, sizeof(*info));
  info->N = hb.plan.N;
  info->nLevels = hb.tree.nLevels;
  info->level_LFR = hb.tree.level_LFR;
  info->n_leaves = hb.plan.n_leaves;
  info->E_far = hb.plan.E_far;
  info->E_near = hb.plan.E_near;
  int64_t bytes = 0, nsrc = 0;
  for (const PassHost& p : hb.plan.passes) {
    bytes += 16 * p.entries + 4 * p.n_src + 16 * p.n_dst;
    nsrc += p.n_src;
  }
  // operators + index lists + destinations + x in + y out + qk2 + near tables
  info->bytes_matvec = bytes + 16 * int64_t(hb.plan.N) * 2 + 8 * int64_t(hb.plan.N) +
                       16 * int64_t(hb.plan.T.size());
  info->flops_matvec = 8 * (hb.plan.E_far + hb.plan.E_near) + 4 * int64_t(hb.plan.N);
  info->n_state = hb.plan.n_state;
  info->max_rank = hb.plan.max_rank;
  info->setup_seconds = hb.setup_seconds;
  for (int i = 0; i < 4; ++i) info->setup_phase_seconds[i] = hb.plan.phase_seconds[i];
}

}  // namespace

struct lsg_ctx {
  HostBuild hb;
  DeviceOperator dev;
  BlockOperator blk;          // multi right-hand-side operator, set up on first use
  GridMaps maps;              // node <-> grid maps of the attached learned preconditioner
  FieldInterp field;          // re-interpolation tables, set up on first use
  LearnedPrecond prec;
  std::vector<double> qk2_user_host;  // column 0 of lsg_set_contrast
  mutable std::string tuning;         // last autotuner report and the block operator's choices
  std::string err;
  double2* dx = nullptr;
  double2* dy = nullptr;
  // allocation-free GMRES (lsg_gmres_noalloc): the device's scratch array,
  // looked up at creation, and the host copy of its per-step estimates
  void* scratch = nullptr;
  std::string scratch_err;
  std::vector<double> noalloc_host;
  double* bx = nullptr;       // two row-split blocks (N x 64) for block calls
  double* by = nullptr;
  double2* bcm = nullptr;     // column-major staging for up to 64 host vectors
  // pinned host staging of two vectors for host-buffer calls: the caller's
  // (pageable) data is copied here by the CPU and moved by DMA at full rate
  double2* pin = nullptr;
  double2* pinned() {
    if (!pin && hipHostMalloc((void**)&pin, 2 * sizeof(double2) * size_t(hb.plan.N), hipHostMallocDefault) != hipSuccess)
      pin = nullptr;
    return pin;
  }
  // host -> device through the pinned staging (slot 0 or 1), on stream s
  bool to_device(double2* d, const double* h, int slot, hipStream_t s) {
    const size_t nb = sizeof(double2) * size_t(hb.plan.N);
    double2* p = pinned();
    if (!p) return hipMemcpyAsync(d, h, nb, hipMemcpyHostToDevice, s) == hipSuccess;
    double2* ps = p + size_t(slot) * size_t(hb.plan.N);
    // the previous use of this slot has finished (every host-buffer call ends synchronised)
    std::memcpy(ps, h, nb);
    return hipMemcpyAsync(d, ps, nb, hipMemcpyHostToDevice, s) == hipSuccess;
  }
  // device -> host through the pinned staging (slot 1); synchronises s
  hipError_t to_host(double* h, const double2* d, hipStream_t s) {
    const size_t nb = sizeof(double2) * size_t(hb.plan.N);
    double2* p = pinned();
    if (!p) {
      hipError_t e = hipMemcpyAsync(h, d, nb, hipMemcpyDeviceToHost, s);
      return e != hipSuccess ? e : hipStreamSynchronize(s);
    }
    double2* ps = p + size_t(hb.plan.N);
    hipError_t e = hipMemcpyAsync(ps, d, nb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) std::memcpy(h, ps, nb);
    return e;
  }
  std::string block_ready() {
    std::string e = blk.setup(dev);
    if (!e.empty()) return e;
    const size_t nbk = sizeof(double) * size_t(hb.plan.N) * 2 * kBlockRhs;
    if (!bx && hipMalloc((void**)&bx, nbk) != hipSuccess) return "hipMalloc of block buffers failed";
    if (!by && hipMalloc((void**)&by, nbk) != hipSuccess) return "hipMalloc of block buffers failed";
    if (!bcm && hipMalloc((void**)&bcm, sizeof(double2) * size_t(hb.plan.N) * kBlockRhs) != hipSuccess)
      return "hipMalloc of block buffers failed";
    return "";
  }
  int set(int code, const std::string& m) {
    err = m;
    return code;
  }
};

struct lsg_host {
  HostBuild hb;
};

struct lsg_ragged {
  RaggedOperator R;
  std::string err;
  std::string tuning;
  double2 *db = nullptr, *dx = nullptr;   // concatenated device vectors
  double2* pin = nullptr;                  // pinned staging of one concatenated vector
  int set(int code, const std::string& m) {
    err = m;
    return code;
  }
  // host vectors of every system -> concatenated device vector d (through the staging)
  bool gather(double2* d, const double* const* v) {
    const int n = R.n;
    for (int i = 0; i < n; ++i)
      std::memcpy(pin + R.off[size_t(i)], v[i], sizeof(double2) * size_t(R.off[size_t(i) + 1] - R.off[size_t(i)]));
    return hipMemcpyAsync(d, pin, sizeof(double2) * size_t(R.off[size_t(n)]), hipMemcpyHostToDevice, R.op.stream) ==
               hipSuccess &&
           hipStreamSynchronize(R.op.stream) == hipSuccess;
  }
  bool scatter(double* const* v, const double2* d) {
    const int n = R.n;
    if (hipMemcpyAsync(pin, d, sizeof(double2) * size_t(R.off[size_t(n)]), hipMemcpyDeviceToHost, R.op.stream) !=
            hipSuccess ||
        hipStreamSynchronize(R.op.stream) != hipSuccess)
      return false;
    for (int i = 0; i < n; ++i)
      std::memcpy(v[i], pin + R.off[size_t(i)], sizeof(double2) * size_t(R.off[size_t(i) + 1] - R.off[size_t(i)]));
    return true;
  }
};

extern "C" {

int lsg_bumps_from_seed(uint64_t seed, int n, double cmin, double cmax, double wmin, double wmax, double amin,
                        double amax, lsg_contrast* out) {
  if (!out || n < 0 || n > LSG_MAX_BUMPS) return LSG_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  out->kind = 6;
  out->n_bumps = n;
  SplitMix64 rng(seed);
  for (int i = 0; i < n; ++i) {
    out->cx[i] = cmin + (cmax - cmin) * rng.uniform();
    out->cy[i] = cmin + (cmax - cmin) * rng.uniform();
    out->width[i] = wmin + (wmax - wmin) * rng.uniform();
    out->amp[i] = amin + (amax - amin) * rng.uniform();
  }
  return LSG_OK;
}

static int finish_create(std::unique_ptr<lsg_ctx> ctx, lsg_ctx** out);

static int device_present() {
  int ndev = 0;
  hipError_t he = hipGetDeviceCount(&ndev);
  if (he != hipSuccess || ndev == 0) {
    g_create_error = std::string("no HIP device available (") + (he != hipSuccess ? hipGetErrorString(he) : "0 devices") +
                     "); this library has no CPU execution path";
    return LSG_ERR_DEVICE;
  }
  return LSG_OK;
}

int lsg_create(const lsg_inputs* in, lsg_ctx** out) {
  g_create_error.clear();
  if (!out) { g_create_error = "out is NULL"; return LSG_ERR_ARG; }
  *out = nullptr;
  if (int rc = device_present(); rc != LSG_OK) return rc;
  auto ctx = std::make_unique<lsg_ctx>();
  std::string e;
  int rc = host_build(in, ctx->hb, e);
  if (rc != LSG_OK) {
    g_create_error = e;
    return rc;
  }
  return finish_create(std::move(ctx), out);
}

// upload a context's host build and allocate its matvec buffers
static int finish_create(std::unique_ptr<lsg_ctx> ctx, lsg_ctx** out) {
  std::string e = ctx->dev.upload(ctx->hb.plan);
  if (!e.empty()) {
    ctx->dev.release();
    g_create_error = e;
    return LSG_ERR_DEVICE;
  }
  // the host copy of the operators is no longer needed
  std::vector<cplx>().swap(ctx->hb.plan.ops);
  std::vector<int>().swap(ctx->hb.plan.src);
  const size_t nb = sizeof(double2) * size_t(ctx->hb.plan.N);
  if (hipMalloc((void**)&ctx->dx, nb) != hipSuccess || hipMalloc((void**)&ctx->dy, nb) != hipSuccess) {
    ctx->dev.release();
    g_create_error = "hipMalloc of matvec buffers failed";
    return LSG_ERR_DEVICE;
  }
  // the allocation-free entry point's workspace: the scratch array's address
  // on this device (once per device) and a host buffer for its estimates
  ctx->scratch_err = scratch_workspace(ctx->dev.device, &ctx->scratch);
  ctx->noalloc_host.assign(2 * (500 + 2), 0.0);
  *out = ctx.release();
  return LSG_OK;
}

void lsg_destroy(lsg_ctx* ctx) {
  if (!ctx) return;
  ctx->maps.release();
  ctx->field.release();
  ctx->blk.release();
  ctx->dev.release();
  if (ctx->dx) hipFree(ctx->dx);
  if (ctx->dy) hipFree(ctx->dy);
  if (ctx->bx) hipFree(ctx->bx);
  if (ctx->by) hipFree(ctx->by);
  if (ctx->bcm) hipFree(ctx->bcm);
  if (ctx->pin) hipHostFree(ctx->pin);
  delete ctx;
}

const char* lsg_last_error(const lsg_ctx* ctx) { return ctx ? ctx->err.c_str() : "context is NULL"; }
const char* lsg_create_error(void) { return g_create_error.c_str(); }

int lsg_get_info(const lsg_ctx* ctx, lsg_info* info) {
  if (!ctx || !info) return LSG_ERR_ARG;
  fill_info(ctx->hb, info);
  return LSG_OK;
}

int lsg_grid_points(const lsg_ctx* ctx, double* x, double* y) {
  if (!ctx || !x || !y) return LSG_ERR_ARG;
  std::memcpy(x, ctx->hb.tree.gx.data(), sizeof(double) * ctx->hb.tree.N);
  std::memcpy(y, ctx->hb.tree.gy.data(), sizeof(double) * ctx->hb.tree.N);
  return LSG_OK;
}

int lsg_leaf_boxes(const lsg_ctx* ctx, double* cx, double* cy, double* radius) {
  if (!ctx || !cx || !cy || !radius) return LSG_ERR_ARG;
  const Tree& t = ctx->hb.tree;
  for (size_t i = 0; i < t.leaves.size(); ++i) {
    const Box& b = t.box(t.leaves[i]);
    cx[i] = b.cx;
    cy[i] = b.cy;
    radius[i] = b.radius;
  }
  return LSG_OK;
}

int lsg_contrast_at_nodes(const lsg_ctx* ctx, double* qk2) {
  if (!ctx || !qk2) return LSG_ERR_ARG;
  const Tree& t = ctx->hb.tree;
  const double k = ctx->hb.prob.kappa;
  for (int i = 0; i < t.N; ++i) qk2[i] = k * k * ctx->hb.prob.q(t.gx[i], t.gy[i]);
  return LSG_OK;
}

int lsg_incident(const lsg_ctx* ctx, double* u) {
  if (!ctx || !u) return LSG_ERR_ARG;
  const Tree& t = ctx->hb.tree;
  for (int i = 0; i < t.N; ++i) {
    cplx v = ctx->hb.prob.u_inc(t.gx[i], t.gy[i]);
    u[2 * i] = v.real();
    u[2 * i + 1] = v.imag();
  }
  return LSG_OK;
}

int lsg_interpolate(lsg_ctx* ctx, const double* f, const double* x, const double* y, int64_t npts, double* out) {
  if (!ctx || !f || (npts > 0 && (!x || !y || !out)) || npts < 0) return LSG_ERR_ARG;
  if (npts == 0) return LSG_OK;
  if (ctx->hb.in.tree_mode != 0) return ctx->set(LSG_ERR_ARG, "lsg_interpolate needs a uniform tree (tree_mode 0)");
  std::string e = ctx->field.setup(ctx->hb.tree, ctx->hb.bs);
  if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
  const size_t N = size_t(ctx->hb.plan.N), n = size_t(npts);
  double2* df = nullptr;
  double *dx = nullptr, *dy = nullptr;
  double2* dout = nullptr;
  bool outside = false;
  auto cleanup = [&]() {
    for (void* p : {(void*)df, (void*)dx, (void*)dy, (void*)dout})
      if (p) (void)hipFree(p);
  };
  if (hipMalloc((void**)&df, sizeof(double2) * N) != hipSuccess || hipMalloc((void**)&dx, sizeof(double) * n) != hipSuccess ||
      hipMalloc((void**)&dy, sizeof(double) * n) != hipSuccess || hipMalloc((void**)&dout, sizeof(double2) * n) != hipSuccess) {
    cleanup();
    return ctx->set(LSG_ERR_DEVICE, "hipMalloc failed in lsg_interpolate");
  }
  hipStream_t s = ctx->dev.stream;
  if (hipMemcpyAsync(df, f, sizeof(double2) * N, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(dy, y, sizeof(double) * n, hipMemcpyHostToDevice, s) != hipSuccess) {
    cleanup();
    return ctx->set(LSG_ERR_DEVICE, "copy to the device failed in lsg_interpolate");
  }
  e = ctx->field.run(df, dx, dy, (long long)npts, dout, s, &outside);
  if (e.empty() && hipMemcpy(out, dout, sizeof(double2) * n, hipMemcpyDeviceToHost) != hipSuccess)
    e = "copy to the host failed in lsg_interpolate";
  cleanup();
  if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
  if (outside) return ctx->set(LSG_ERR_ARG, "a point lies outside [-L, L)^2");
  return LSG_OK;
}

int lsg_rhs(const lsg_ctx* ctx, double* b) {
  if (!ctx || !b) return LSG_ERR_ARG;
  const Tree& t = ctx->hb.tree;
  for (int i = 0; i < t.N; ++i) {
    cplx v = ctx->hb.prob.rhs(t.gx[i], t.gy[i]);
    b[2 * i] = v.real();
    b[2 * i + 1] = v.imag();
  }
  return LSG_OK;
}

int lsg_matvec(lsg_ctx* ctx, const double* x, double* y, int nrhs, int64_t ld) {
  if (!ctx) return LSG_ERR_ARG;
  const int64_t N = ctx->hb.plan.N;
  if (!x || !y || nrhs < 1 || ld < N) return ctx->set(LSG_ERR_ARG, "bad matvec arguments");
  const size_t nb = sizeof(double2) * size_t(N);
  if (nrhs >= 2) {
    // block path: 64 right-hand sides per block matvec (FP64 MFMA)
    for (int r = 0; r < nrhs; ++r)
      if (!all_finite(x + 2 * ld * r, 2 * N)) return ctx->set(LSG_ERR_NUMERIC, "input vector holds NaN/Inf");
    std::string e = ctx->block_ready();
    if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
    hipStream_t s = ctx->dev.stream;
    for (int r0 = 0; r0 < nrhs; r0 += kBlockRhs) {
      const int nr = std::min(kBlockRhs, nrhs - r0);
      for (int r = 0; r < nr; ++r)
        hipMemcpyAsync(ctx->bcm + size_t(r) * N, x + 2 * ld * (r0 + r), nb, hipMemcpyHostToDevice, s);
      to_split(ctx->bcm, N, nr, int(N), kBlockRhs, ctx->bx, s);
      ctx->blk.matvec(ctx->bx, ctx->by);
      from_split(ctx->by, int(N), kBlockRhs, nr, ctx->bcm, N, s);
      for (int r = 0; r < nr; ++r)
        hipMemcpyAsync(y + 2 * ld * (r0 + r), ctx->bcm + size_t(r) * N, nb, hipMemcpyDeviceToHost, s);
      hipError_t he = hipStreamSynchronize(s);
      if (he != hipSuccess) return ctx->set(LSG_ERR_DEVICE, std::string("block matvec: ") + hipGetErrorString(he));
    }
    return LSG_OK;
  }
  for (int r = 0; r < nrhs; ++r) {
    const double* xr = x + 2 * ld * r;
    if (!all_finite(xr, 2 * N)) return ctx->set(LSG_ERR_NUMERIC, "input vector holds NaN/Inf");
    hipStream_t s = ctx->dev.stream;
    if (!ctx->to_device(ctx->dx, xr, 0, s)) return ctx->set(LSG_ERR_DEVICE, "copy in failed");
    ctx->dev.matvec(ctx->dx, ctx->dy);
    hipError_t e = ctx->to_host(y + 2 * ld * r, ctx->dy, s);
    if (e != hipSuccess) return ctx->set(LSG_ERR_DEVICE, std::string("matvec: ") + hipGetErrorString(e));
  }
  return LSG_OK;
}

int lsg_matvec_device(lsg_ctx* ctx, const void* x_dev, void* y_dev) {
  if (!ctx || !x_dev || !y_dev) return LSG_ERR_ARG;
  ctx->dev.matvec(static_cast<const double2*>(x_dev), static_cast<double2*>(y_dev));
  return LSG_OK;
}

int lsg_sync(lsg_ctx* ctx) {
  if (!ctx) return LSG_ERR_ARG;
  hipError_t e = hipStreamSynchronize(ctx->dev.stream);
  if (e != hipSuccess) return ctx->set(LSG_ERR_DEVICE, hipGetErrorString(e));
  return LSG_OK;
}

int lsg_profile_passes(lsg_ctx* ctx, int reps, double* ms, int64_t* bytes, int max_pass) {
  if (!ctx || reps < 1 || !ms || !bytes) return LSG_ERR_ARG;
  std::vector<double> m;
  std::vector<int64_t> b;
  std::vector<std::string> names;
  std::string e = ctx->dev.profile(reps, m, b, names);
  if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
  const int n = std::min<int>(max_pass, int(m.size()));
  for (int i = 0; i < n; ++i) { ms[i] = m[i]; bytes[i] = b[i]; }
  std::string all;
  for (auto& s : names) all += s + ",";
  ctx->err = all;  // pass names are reported through lsg_last_error
  return int(m.size());
}

int lsg_time_matvec(lsg_ctx* ctx, int reps, double* ms) {
  if (!ctx || reps < 1 || !ms) return LSG_ERR_ARG;
  *ms = ctx->dev.time_matvec(reps);
  if (*ms < 0) return ctx->set(LSG_ERR_DEVICE, "time_matvec allocation failed");
  return LSG_OK;
}

int lsg_set_contrast(lsg_ctx* ctx, const double* qk2, int ncol) {
  if (!ctx) return LSG_ERR_ARG;
  if (ctx->dev.find_phi) return ctx->set(LSG_ERR_ARG, "lsg_set_contrast needs a volume-potential context (find_phi = 0)");
  const int64_t N = ctx->hb.plan.N;
  if (qk2) {
    if (ncol < 1 || ncol > kBlockRhs) return ctx->set(LSG_ERR_ARG, "ncol must be in [1, 64]");
    for (int64_t i = 0; i < N * ncol; ++i)
      if (!std::isfinite(qk2[i])) return ctx->set(LSG_ERR_NUMERIC, "contrast holds NaN/Inf");
  }
  // A captured matvec graph decided at capture time whether it applies the
  // contrast and which buffer it reads: drop it (after in-flight replays)
  // and capture again once the new contrast is in place.
  const bool had_graph = ctx->dev.graph_exec != nullptr;
  ctx->dev.drop_graph();
  hipError_t se = hipStreamSynchronize(ctx->dev.stream);
  if (se != hipSuccess) return ctx->set(LSG_ERR_DEVICE, std::string("lsg_set_contrast: ") + hipGetErrorString(se));
  auto recapture = [&]() {
    if (!had_graph) return LSG_OK;
    std::string ge = ctx->dev.capture_graph();
    return ge.empty() ? LSG_OK : ctx->set(LSG_ERR_DEVICE, ge);
  };
  if (!qk2) {
    if (ctx->dev.qk2_user) hipFree(ctx->dev.qk2_user);
    if (ctx->blk.qk2_cols) hipFree(ctx->blk.qk2_cols);
    ctx->dev.qk2_user = nullptr;
    ctx->blk.qk2_cols = nullptr;
    ctx->qk2_user_host.clear();
    return recapture();
  }
  std::string e = ctx->block_ready();
  if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
  if (!ctx->dev.qk2_user && hipMalloc((void**)&ctx->dev.qk2_user, sizeof(double) * N) != hipSuccess)
    return ctx->set(LSG_ERR_DEVICE, "hipMalloc failed");
  if (!ctx->blk.qk2_cols && hipMalloc((void**)&ctx->blk.qk2_cols, sizeof(double) * N * kBlockRhs) != hipSuccess)
    return ctx->set(LSG_ERR_DEVICE, "hipMalloc failed");
  std::vector<double> rowmajor(size_t(N) * kBlockRhs);
  for (int64_t n = 0; n < N; ++n)
    for (int q = 0; q < kBlockRhs; ++q) rowmajor[size_t(n) * kBlockRhs + q] = qk2[size_t(q < ncol ? q : 0) * N + n];
  if (hipMemcpy(ctx->dev.qk2_user, qk2, sizeof(double) * N, hipMemcpyHostToDevice) != hipSuccess)
    return ctx->set(LSG_ERR_DEVICE, "copy of the contrast failed");
  ctx->qk2_user_host.assign(qk2, qk2 + N);
  if (hipMemcpy(ctx->blk.qk2_cols, rowmajor.data(), sizeof(double) * rowmajor.size(), hipMemcpyHostToDevice) !=
      hipSuccess)
    return ctx->set(LSG_ERR_DEVICE, "copy of the contrast failed");
  return recapture();
}

int lsg_attach_preconditioner(lsg_ctx* ctx, lsgn_net* net) {
  if (!ctx) return LSG_ERR_ARG;
  ctx->maps.release();
  ctx->prec.net = nullptr;
  ctx->prec.maps = nullptr;
  if (!net) return LSG_OK;
  if (ctx->hb.in.tree_mode != 0)
    return ctx->set(LSG_ERR_ARG, "the learned preconditioner's grid maps need a uniform tree (tree_mode 0)");
  const Tree& t = ctx->hb.tree;
  std::vector<double> cx(t.leaves.size()), cy(t.leaves.size());
  double radius = 0;
  for (size_t i = 0; i < t.leaves.size(); ++i) {
    const Box& b = t.box(t.leaves[i]);
    cx[i] = b.cx;
    cy[i] = b.cy;
    radius = b.radius;
  }
  std::vector<double> qk2;
  if (!ctx->qk2_user_host.empty()) {
    qk2 = ctx->qk2_user_host;
  } else {
    qk2.resize(size_t(t.N));
    lsg_contrast_at_nodes(ctx, qk2.data());
  }
  std::string e = ctx->maps.build(cx, cy, radius, ctx->hb.in.nChebNodes, ctx->hb.in.L, net->net.G, qk2,
                                  ctx->hb.in.kappa);
  if (!e.empty()) {
    ctx->maps.release();
    return ctx->set(LSG_ERR_ARG, e);
  }
  ctx->prec.net = &net->net;
  ctx->prec.maps = &ctx->maps;
  return LSG_OK;
}

int lsg_apply_preconditioner(lsg_ctx* ctx, const double* v, double* z) {
  if (!ctx || !v || !z) return LSG_ERR_ARG;
  if (!ctx->prec.net) return ctx->set(LSG_ERR_ARG, "no preconditioner attached");
  const size_t nb = sizeof(double2) * size_t(ctx->hb.plan.N);
  hipStream_t s = ctx->dev.stream;
  hipMemcpyAsync(ctx->dx, v, nb, hipMemcpyHostToDevice, s);
  ctx->prec.apply(ctx->dx, ctx->dy, s);
  hipMemcpyAsync(z, ctx->dy, nb, hipMemcpyDeviceToHost, s);
  hipError_t he = hipStreamSynchronize(s);
  return he == hipSuccess ? LSG_OK : ctx->set(LSG_ERR_DEVICE, hipGetErrorString(he));
}

int lsg_grid_resample(lsg_ctx* ctx, const double* in, double* out, int to_grid) {
  if (!ctx || !in || !out) return LSG_ERR_ARG;
  if (!ctx->prec.net) return ctx->set(LSG_ERR_ARG, "no preconditioner attached (the maps belong to it)");
  const int N = int(ctx->hb.plan.N), GG = ctx->maps.G * ctx->maps.G;
  const size_t nin = sizeof(double2) * size_t(to_grid ? N : GG), nout = sizeof(double2) * size_t(to_grid ? GG : N);
  double2 *a = nullptr, *b = nullptr;
  if (hipMalloc((void**)&a, nin) != hipSuccess || hipMalloc((void**)&b, nout) != hipSuccess) {
    if (a) hipFree(a);
    return ctx->set(LSG_ERR_DEVICE, "hipMalloc failed");
  }
  hipStream_t s = ctx->dev.stream;
  hipMemcpyAsync(a, in, nin, hipMemcpyHostToDevice, s);
  if (to_grid) ctx->maps.to_grid(a, b, s);
  else ctx->maps.to_nodes(a, b, s);
  hipMemcpyAsync(out, b, nout, hipMemcpyDeviceToHost, s);
  hipError_t he = hipStreamSynchronize(s);
  hipFree(a);
  hipFree(b);
  return he == hipSuccess ? LSG_OK : ctx->set(LSG_ERR_DEVICE, hipGetErrorString(he));
}

int lsg_gmres_learned(lsg_ctx* ctx, const double* b, double* x, const lsg_solve_opts* o, int flags, lsg_stats* stats) {
  if (!ctx) return LSG_ERR_ARG;
  if (!o || !b || !x || o->mode != LSG_SOLVE_FP64 || o->restart_len < 1 || o->restart_len > 500 ||
      o->max_restarts < 1 || !(o->tol > 0) || (flags & ~(LSG_PREC_GUESS | LSG_PREC_RIGHT)))
    return ctx->set(LSG_ERR_ARG, "bad learned gmres arguments (mode FP64, tol > 0, restart_len in [1,500])");
  if (flags && !ctx->prec.net) return ctx->set(LSG_ERR_ARG, "no preconditioner attached");
  const int64_t N = ctx->hb.plan.N;
  const size_t nb = sizeof(double2) * size_t(N);
  hipStream_t s = ctx->dev.stream;
  double2 *db = nullptr, *dxs = nullptr;
  if (o->device_ptrs) {
    db = const_cast<double2*>(reinterpret_cast<const double2*>(b));
    dxs = reinterpret_cast<double2*>(x);
  } else {
    if (!all_finite(b, 2 * N) || !all_finite(x, 2 * N)) return ctx->set(LSG_ERR_NUMERIC, "b or x holds NaN/Inf");
    // the context's own vectors (one call at a time per context)
    db = ctx->dx;
    dxs = ctx->dy;
    if (!ctx->to_device(db, b, 0, s) || !ctx->to_device(dxs, x, 1, s)) return ctx->set(LSG_ERR_DEVICE, "copy in failed");
  }
  auto t0 = std::chrono::steady_clock::now();
  if (flags & LSG_PREC_GUESS) ctx->prec.apply(db, dxs, s);
  // The network preconditions the first restart cycle only; later cycles are
  // plain FP64 GMRES from the current iterate, so convergence does not depend
  // on the network (a nonlinear M can stall flexible GMRES near its own noise).
  GmresResult r;
  if (flags & LSG_PREC_RIGHT) {
    r = device_gmres(ctx->dev, db, dxs, o->restart_len, 1, o->tol, false, &ctx->prec);
    if ((r.status == 0 || r.status == LSG_ERR_NOCONV) && !r.converged && o->max_restarts > 1) {
      GmresResult r2 = device_gmres(ctx->dev, db, dxs, o->restart_len, o->max_restarts - 1, o->tol, false, nullptr);
      r2.iterations += r.iterations;
      r2.restarts += r.restarts;
      r = r2;
    }
  } else {
    r = device_gmres(ctx->dev, db, dxs, o->restart_len, o->max_restarts, o->tol, false, nullptr);
  }
  r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (!o->device_ptrs) {
    hipError_t ce = ctx->to_host(x, dxs, s);
    if (ce != hipSuccess && r.status == 0) {
      r.status = LSG_ERR_DEVICE;
      r.error = std::string("copy out failed: ") + hipGetErrorString(ce);
    }
  }
  if (stats) {
    stats->iterations = r.iterations;
    stats->restarts = r.restarts;
    stats->converged = r.converged;
    stats->rel_residual = r.rel_residual;
    stats->est_residual = r.est_residual;
    stats->seconds = r.seconds;
    stats->refinements = 0;
  }
  if (r.status != 0) return ctx->set(r.status, r.error);
  return LSG_OK;
}

static int time_block(lsg_ctx* ctx, int reps, double* ms, bool fp32) {
  if (!ctx || reps < 1 || !ms) return LSG_ERR_ARG;
  std::string e = ctx->block_ready();
  if (e.empty() && fp32) e = ctx->blk.enable_inner();
  if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
  SbScope sb_scope(ctx->blk, fp32 && ctx->blk.inner_sb);
  auto mv = [&]() {
    if (fp32) ctx->blk.matvec32(ctx->bx, ctx->by);
    else ctx->blk.matvec(ctx->bx, ctx->by);
  };
  hipStream_t s = ctx->dev.stream;
  const size_t nbk = sizeof(double) * size_t(ctx->hb.plan.N) * 2 * kBlockRhs;
  hipMemsetAsync(ctx->bx, 0, nbk, s);
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  mv();
  hipEventRecord(a, s);
  for (int r = 0; r < reps; ++r) mv();
  hipEventRecord(b, s);
  hipError_t he = hipEventSynchronize(b);
  float t = 0;
  hipEventElapsedTime(&t, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  if (he != hipSuccess) return ctx->set(LSG_ERR_DEVICE, hipGetErrorString(he));
  *ms = double(t) / reps;
  return LSG_OK;
}

int lsg_time_matvec_block(lsg_ctx* ctx, int reps, double* ms) { return time_block(ctx, reps, ms, false); }
int lsg_time_matvec_block_fp32(lsg_ctx* ctx, int reps, double* ms) { return time_block(ctx, reps, ms, true); }

int lsg_time_matvec_fp32(lsg_ctx* ctx, int reps, double* ms) {
  if (!ctx || reps < 1 || !ms) return LSG_ERR_ARG;
  std::string e = ctx->dev.enable_fp32();
  if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
  *ms = ctx->dev.time_matvec(reps, true);
  if (*ms < 0) return ctx->set(LSG_ERR_DEVICE, "time_matvec allocation failed");
  return LSG_OK;
}

int lsg_matvec_fp32(lsg_ctx* ctx, const double* x, double* y, int nrhs, int64_t ld) {
  if (!ctx) return LSG_ERR_ARG;
  const int64_t N = ctx->hb.plan.N;
  if (!x || !y || nrhs < 1 || ld < N) return ctx->set(LSG_ERR_ARG, "bad matvec arguments");
  std::string e = ctx->dev.enable_fp32();
  if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
  const size_t nb = sizeof(double2) * size_t(N);
  if (nrhs >= 2) {
    // block path: 64 right-hand sides per block matvec in the inner solves'
    // arithmetic (split-BF16 or FP32 MFMA, FP32 pass state)
    for (int r = 0; r < nrhs; ++r)
      if (!all_finite(x + 2 * ld * r, 2 * N)) return ctx->set(LSG_ERR_NUMERIC, "input vector holds NaN/Inf");
    e = ctx->block_ready();
    if (e.empty()) e = ctx->blk.enable_inner();
    if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
    SbScope sb_scope(ctx->blk, ctx->blk.inner_sb);
    hipStream_t s = ctx->dev.stream;
    for (int r0 = 0; r0 < nrhs; r0 += kBlockRhs) {
      const int nr = std::min(kBlockRhs, nrhs - r0);
      for (int r = 0; r < nr; ++r)
        hipMemcpyAsync(ctx->bcm + size_t(r) * N, x + 2 * ld * (r0 + r), nb, hipMemcpyHostToDevice, s);
      to_split(ctx->bcm, N, nr, int(N), kBlockRhs, ctx->bx, s);
      ctx->blk.matvec32(static_cast<const double*>(ctx->bx), ctx->by);
      from_split(ctx->by, int(N), kBlockRhs, nr, ctx->bcm, N, s);
      for (int r = 0; r < nr; ++r)
        hipMemcpyAsync(y + 2 * ld * (r0 + r), ctx->bcm + size_t(r) * N, nb, hipMemcpyDeviceToHost, s);
      hipError_t he = hipStreamSynchronize(s);
      if (he != hipSuccess) return ctx->set(LSG_ERR_DEVICE, std::string("block matvec: ") + hipGetErrorString(he));
    }
    return LSG_OK;
  }
  for (int r = 0; r < nrhs; ++r) {
    const double* xr = x + 2 * ld * r;
    if (!all_finite(xr, 2 * N)) return ctx->set(LSG_ERR_NUMERIC, "input vector holds NaN/Inf");
    hipStream_t s = ctx->dev.stream;
    if (hipMemcpyAsync(ctx->dx, xr, nb, hipMemcpyHostToDevice, s) != hipSuccess)
      return ctx->set(LSG_ERR_DEVICE, "copy in failed");
    ctx->dev.matvec32(ctx->dx, ctx->dy);
    if (hipMemcpyAsync(y + 2 * ld * r, ctx->dy, nb, hipMemcpyDeviceToHost, s) != hipSuccess)
      return ctx->set(LSG_ERR_DEVICE, "copy out failed");
    hipError_t he = hipStreamSynchronize(s);
    if (he != hipSuccess) return ctx->set(LSG_ERR_DEVICE, std::string("matvec: ") + hipGetErrorString(he));
  }
  return LSG_OK;
}

int lsg_gmres_ex(lsg_ctx* ctx, const double* b, double* x, const lsg_solve_opts* o, lsg_stats* stats) {
  if (!ctx) return LSG_ERR_ARG;
  if (!o || !b || !x || o->restart_len < 1 || o->restart_len > 500 || o->max_restarts < 1)
    return ctx->set(LSG_ERR_ARG, "bad gmres arguments (restart_len in [1,500], max_restarts >= 1)");
  if (o->mode != LSG_SOLVE_FP64 && o->mode != LSG_SOLVE_FP32_IR)
    return ctx->set(LSG_ERR_ARG, "unknown solver mode (FP32_FGMRES is batched only: lsg_gmres_batch)");
  // tol <= 0 is the reference's fixed-cycle mode; NaN or Inf is an error
  if (!std::isfinite(o->tol)) return ctx->set(LSG_ERR_ARG, "tol must be finite");
  if (o->mode == LSG_SOLVE_FP32_IR && (!(o->tol > 0) || !(o->inner_tol > 0) || o->max_outer < 1))
    return ctx->set(LSG_ERR_ARG, "FP32_IR needs tol > 0, inner_tol > 0, max_outer >= 1");
  const int64_t N = ctx->hb.plan.N;
  const size_t nb = sizeof(double2) * size_t(N);
  hipStream_t s = ctx->dev.stream;
  double2 *db = nullptr, *dxs = nullptr;
  if (o->device_ptrs) {
    db = const_cast<double2*>(reinterpret_cast<const double2*>(b));
    dxs = reinterpret_cast<double2*>(x);
  } else {
    if (!all_finite(b, 2 * N) || !all_finite(x, 2 * N)) return ctx->set(LSG_ERR_NUMERIC, "b or x holds NaN/Inf");
    // the context's own vectors (one call at a time per context)
    db = ctx->dx;
    dxs = ctx->dy;
    if (!ctx->to_device(db, b, 0, s) || !ctx->to_device(dxs, x, 1, s)) return ctx->set(LSG_ERR_DEVICE, "copy in failed");
  }
  GmresResult r = o->mode == LSG_SOLVE_FP64
                      ? device_gmres(ctx->dev, db, dxs, o->restart_len, o->max_restarts, o->tol)
                      : device_gmres_ir(ctx->dev, db, dxs, o->restart_len, o->max_restarts, o->tol, o->inner_tol,
                                        o->max_outer);
  if (!o->device_ptrs) {
    hipError_t ce = ctx->to_host(x, dxs, s);
    if (ce != hipSuccess && r.status == 0) {
      r.status = LSG_ERR_DEVICE;
      r.error = std::string("copy out failed: ") + hipGetErrorString(ce);
    }
  }
  if (stats) {
    stats->iterations = r.iterations;
    stats->restarts = r.restarts;
    stats->converged = r.converged;
    stats->rel_residual = r.rel_residual;
    stats->est_residual = r.est_residual;
    stats->seconds = r.seconds;
    stats->refinements = r.refinements;
  }
  if (r.status != 0) return ctx->set(r.status, r.error);
  return LSG_OK;
}

int lsg_gmres_noalloc(lsg_ctx* ctx, const void* b_dev, void* x_dev, int restart_len, int max_restarts, double tol,
                      lsg_stats* stats) {
  if (!ctx) return LSG_ERR_ARG;
  if (!b_dev || !x_dev || restart_len < 1 || restart_len > 500 || max_restarts < 1 || !std::isfinite(tol))
    return ctx->set(LSG_ERR_ARG, "bad gmres arguments (device b and x, restart_len in [1,500], max_restarts >= 1, "
                                 "finite tol)");
  if (!ctx->scratch) return ctx->set(LSG_ERR_DEVICE, "no scratch workspace on this device: " + ctx->scratch_err);
  const int N = int(ctx->hb.plan.N);
  const size_t need = gmres_work_bytes(N, restart_len);
  if (need > kScratchBytes)
    return ctx->set(LSG_ERR_ARG, "GMRES(" + std::to_string(restart_len) + ") on " + std::to_string(N) +
                                     " unknowns needs " + std::to_string(need) + " bytes of scratch; the maximum is " +
                                     std::to_string(kScratchBytes) + " (use lsg_gmres_ex)");
  GmresWork W;
  std::string e = W.carve(ctx->scratch, kScratchBytes, N, restart_len, ctx->noalloc_host.data());
  if (!e.empty()) return ctx->set(LSG_ERR_ARG, e);
  GmresResult r = device_gmres(ctx->dev, static_cast<const double2*>(b_dev), static_cast<double2*>(x_dev), restart_len,
                               max_restarts, tol, false, nullptr, &W);
  if (stats) {
    stats->iterations = r.iterations;
    stats->restarts = r.restarts;
    stats->converged = r.converged;
    stats->rel_residual = r.rel_residual;
    stats->est_residual = r.est_residual;
    stats->seconds = r.seconds;
    stats->refinements = 0;
  }
  if (r.status != 0) return ctx->set(r.status, r.error);
  return LSG_OK;
}

// outer cycles of the batched flexible GMRES (each opens with the true residual)
constexpr int kFgmresCycles = 20;

int lsg_gmres_batch(lsg_ctx* ctx, const double* b, double* x, int nrhs, int64_t ld, const lsg_solve_opts* o,
                    double* rel_residuals, lsg_stats* stats) {
  if (!ctx) return LSG_ERR_ARG;
  const int64_t N = ctx->hb.plan.N;
  if (!o || !b || !x || nrhs < 1 || nrhs > kBlockRhs || ld < N || !(o->tol > 0) || o->restart_len < 1 ||
      o->restart_len > 500 || o->max_restarts < 1 ||
      (o->mode != LSG_SOLVE_FP64 && o->mode != LSG_SOLVE_FP32_IR && o->mode != LSG_SOLVE_FP32_FGMRES) ||
      (o->mode != LSG_SOLVE_FP64 && (!(o->inner_tol > 0) || o->max_outer < 1)) ||
      (o->mode == LSG_SOLVE_FP32_FGMRES && o->max_outer > 500))
    return ctx->set(LSG_ERR_ARG,
                    "bad batched gmres arguments (1 <= nrhs <= 64, ld >= N, tol > 0, restart_len in [1,500]; "
                    "FP32_IR, FP32_FGMRES: inner_tol > 0, max_outer >= 1 (<= 500 for FP32_FGMRES))");
  std::string e = ctx->block_ready();
  if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
  hipStream_t s = ctx->dev.stream;
  const size_t nb = sizeof(double2) * size_t(N);
  // B -> bx, X -> by (row-split)
  auto load = [&](const double* src, double* dst) {
    if (o->device_ptrs) {
      to_split(reinterpret_cast<const double2*>(src), ld, nrhs, int(N), kBlockRhs, dst, s);
    } else {
      for (int r = 0; r < nrhs; ++r) hipMemcpyAsync(ctx->bcm + size_t(r) * N, src + 2 * ld * r, nb, hipMemcpyHostToDevice, s);
      to_split(ctx->bcm, N, nrhs, int(N), kBlockRhs, dst, s);
    }
  };
  if (!o->device_ptrs)
    for (int r = 0; r < nrhs; ++r)
      if (!all_finite(b + 2 * ld * r, 2 * N) || !all_finite(x + 2 * ld * r, 2 * N))
        return ctx->set(LSG_ERR_NUMERIC, "b or x holds NaN/Inf");
  load(b, ctx->bx);
  load(x, ctx->by);
  std::vector<double> rel(kBlockRhs, 0.0);
  BatchResult r = o->mode == LSG_SOLVE_FP64
                      ? block_gmres(ctx->blk, ctx->bx, ctx->by, nrhs, o->restart_len, o->max_restarts, o->tol,
                                    rel.data())
                  : o->mode == LSG_SOLVE_FP32_IR
                      ? block_gmres_ir(ctx->blk, ctx->bx, ctx->by, nrhs, o->restart_len, o->max_restarts, o->tol,
                                       o->inner_tol, o->max_outer, rel.data())
                      : block_fgmres(ctx->blk, ctx->bx, ctx->by, nrhs, o->restart_len, o->max_restarts, o->tol,
                                     o->inner_tol, o->max_outer, kFgmresCycles, rel.data());
  if (o->device_ptrs) {
    from_split(ctx->by, int(N), kBlockRhs, nrhs, reinterpret_cast<double2*>(x), ld, s);
  } else {
    from_split(ctx->by, int(N), kBlockRhs, nrhs, ctx->bcm, N, s);
    for (int q = 0; q < nrhs; ++q) hipMemcpyAsync(x + 2 * ld * q, ctx->bcm + size_t(q) * N, nb, hipMemcpyDeviceToHost, s);
  }
  hipError_t se = hipStreamSynchronize(s);
  if (se != hipSuccess && r.status == 0) {
    r.status = LSG_ERR_DEVICE;
    r.error = std::string("batched GMRES: copy out failed: ") + hipGetErrorString(se);
  }
  if (rel_residuals)
    for (int q = 0; q < nrhs; ++q) rel_residuals[q] = rel[q];
  if (stats) {
    stats->iterations = r.iterations;
    stats->restarts = r.restarts;
    stats->converged = r.converged;
    stats->rel_residual = r.max_rel_residual;
    stats->est_residual = 0.0;
    stats->seconds = r.seconds;
    stats->refinements = r.refinements;
  }
  if (r.status != 0) return ctx->set(r.status, r.error);
  return LSG_OK;
}

int lsg_autotune(lsg_ctx* ctx, int reps, int* n_changed) {
  if (!ctx || reps < 1) return LSG_ERR_ARG;
  const bool had_graph = ctx->dev.graph_exec != nullptr;
  std::string e = ctx->dev.autotune(reps, n_changed);
  if (e.empty() && had_graph) e = ctx->dev.capture_graph();
  if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
  return LSG_OK;
}

const char* lsg_tuning(const lsg_ctx* ctx) {
  if (!ctx) return "";
  ctx->tuning.clear();
  for (const std::string& t : ctx->dev.tuning) ctx->tuning += t + "\n";
  for (const std::string& t : ctx->blk.notes) ctx->tuning += t + "\n";
  return ctx->tuning.c_str();
}

int lsg_set_block_tiles(lsg_ctx* ctx, int tiles) {
  if (!ctx || tiles < 1 || tiles > 4) return LSG_ERR_ARG;
  std::string e = ctx->block_ready();
  if (e.empty()) e = ctx->blk.enable_inner();
  if (!e.empty()) return ctx->set(LSG_ERR_DEVICE, e);
  if (tiles > 2 && !(ctx->blk.inner_sb && ctx->dev.T && ctx->blk.leaf_fused))
    return ctx->set(LSG_ERR_ARG, "the near field runs beside the passes only in split BF16 on a uniform tree");
  if (ctx->blk.inner_sb) {
    ctx->blk.nt_sb = 2 - (tiles & 1);
    ctx->blk.sb_split = tiles > 2;
  } else {
    ctx->blk.nt32 = tiles;
  }
  return LSG_OK;
}

int lsg_set_inner_arith(lsg_ctx* ctx, int arith) {
  if (!ctx || (arith != LSG_ARITH_FP32 && arith != LSG_ARITH_SPLIT_BF16)) return LSG_ERR_ARG;
  ctx->blk.inner_sb = arith == LSG_ARITH_SPLIT_BF16;
  return LSG_OK;
}

int lsg_set_pass_kernel(lsg_ctx* ctx, int kernel) {
  if (!ctx || kernel < 0 || kernel > 5) return LSG_ERR_ARG;
  const bool had_graph = ctx->dev.graph_exec != nullptr;
  ctx->dev.drop_graph();
  const int k = kernel % 3;
  for (DevPass& p : ctx->dev.passes) {
    p.lane = k == 2 && p.lane_ok;
    p.wave = k == 1;
  }
  ctx->dev.fup.on = kernel >= 3 && ctx->dev.fup.tasks > 0;
  ctx->dev.fdown.on = kernel >= 3 && ctx->dev.fdown.tasks > 0;
  std::string e = had_graph ? ctx->dev.capture_graph() : std::string();
  return e.empty() ? LSG_OK : ctx->set(LSG_ERR_DEVICE, e);
}

int lsg_use_graph(lsg_ctx* ctx, int on) {
  if (!ctx) return LSG_ERR_ARG;
  if (!on) {
    hipStreamSynchronize(ctx->dev.stream);
    ctx->dev.drop_graph();
    return LSG_OK;
  }
  std::string e = ctx->dev.capture_graph();
  return e.empty() ? LSG_OK : ctx->set(LSG_ERR_DEVICE, e);
}

int lsg_gmres_ragged(lsg_ctx* const* ctxs, int n, const double* const* b, double* const* x, const lsg_solve_opts* o,
                     lsg_stats* stats) {
  if (!ctxs || n < 1 || !b || !x || !o) return LSG_ERR_ARG;
  for (int i = 0; i < n; ++i)
    if (!ctxs[i] || !b[i] || !x[i]) return LSG_ERR_ARG;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      if (ctxs[i] == ctxs[j]) return ctxs[i]->set(LSG_ERR_ARG, "a context appears twice in the ragged batch");
  std::vector<int> rc(size_t(n), LSG_OK);
  std::vector<std::thread> th;
  th.reserve(size_t(n));
  for (int i = 0; i < n; ++i)
    th.emplace_back([&, i]() { rc[size_t(i)] = lsg_gmres_ex(ctxs[i], b[i], x[i], o, stats ? &stats[i] : nullptr); });
  for (auto& t : th) t.join();
  int worst = LSG_OK;
  for (int i = 0; i < n; ++i)
    if (rc[size_t(i)] != LSG_OK && (worst == LSG_OK || rc[size_t(i)] != LSG_ERR_NOCONV)) worst = rc[size_t(i)];
  return worst;
}

int lsg_gmres(lsg_ctx* ctx, const double* b, double* x, int restart_len, int max_restarts, double tol,
              lsg_stats* stats) {
  lsg_solve_opts o{};
  o.mode = LSG_SOLVE_FP64;
  o.restart_len = restart_len;
  o.max_restarts = max_restarts;
  o.tol = tol;
  return lsg_gmres_ex(ctx, b, x, &o, stats);
}

// ----------------------------------------------------------- host-only API
int lsg_host_build(const lsg_inputs* in, lsg_host** out) {
  g_host_error.clear();
  if (!out) { g_host_error = "out is NULL"; return LSG_ERR_ARG; }
  *out = nullptr;
  auto h = std::make_unique<lsg_host>();
  std::string e;
  int rc = host_build(in, h->hb, e);
  if (rc != LSG_OK) {
    g_host_error = e;
    return rc;
  }
  h->hb.far = h->hb.plan.far_leaf;
  *out = h.release();
  return LSG_OK;
}

int lsg_host_build_tree(const lsg_inputs* in, lsg_host** out) {
  g_host_error.clear();
  if (!out) { g_host_error = "out is NULL"; return LSG_ERR_ARG; }
  *out = nullptr;
  auto h = std::make_unique<lsg_host>();
  std::string e;
  int rc = host_build(in, h->hb, e, true);
  if (rc != LSG_OK) {
    g_host_error = e;
    return rc;
  }
  h->hb.plan.N = h->hb.tree.N;
  *out = h.release();
  return LSG_OK;
}

int lsg_create_from_host(const lsg_host* h, lsg_ctx** out) {
  g_create_error.clear();
  if (!out) { g_create_error = "out is NULL"; return LSG_ERR_ARG; }
  *out = nullptr;
  if (!h) { g_create_error = "host build is NULL"; return LSG_ERR_ARG; }
  if (h->hb.plan.passes.empty() && h->hb.plan.T.empty())
    { g_create_error = "the host build has no operators (lsg_host_build_tree builds only the tree)"; return LSG_ERR_ARG; }
  int64_t entries = 0;
  for (const PassHost& ph : h->hb.plan.passes) entries += ph.entries;
  if (int64_t(h->hb.plan.ops.size()) != entries)
    { g_create_error = "the host build's operators are incomplete"; return LSG_ERR_ARG; }
  if (int rc = device_present(); rc != LSG_OK) return rc;
  auto ctx = std::make_unique<lsg_ctx>();
  ctx->hb = h->hb;
  std::vector<cplx>().swap(ctx->hb.far);
  return finish_create(std::move(ctx), out);
}

void lsg_host_destroy(lsg_host* h) { delete h; }

int lsg_ragged_create(const lsg_host* const* hosts, int n, lsg_ragged** out) {
  g_create_error.clear();
  if (!out) { g_create_error = "out is NULL"; return LSG_ERR_ARG; }
  *out = nullptr;
  if (!hosts || n < 1) { g_create_error = "need at least one host build"; return LSG_ERR_ARG; }
  std::vector<const Plan*> plans;
  for (int i = 0; i < n; ++i) {
    if (!hosts[i]) { g_create_error = "host build " + std::to_string(i) + " is NULL"; return LSG_ERR_ARG; }
    for (int j = 0; j < i; ++j)
      if (hosts[j] == hosts[i]) { g_create_error = "a host build appears twice"; return LSG_ERR_ARG; }
    if (hosts[i]->hb.in.tree_mode != 0) { g_create_error = "the fused batch needs uniform trees"; return LSG_ERR_ARG; }
    plans.push_back(&hosts[i]->hb.plan);
  }
  if (int rc = device_present(); rc != LSG_OK) return rc;
  Plan fused;
  std::vector<int> off;
  std::string e = fuse_plans(plans, fused, off);
  if (!e.empty()) { g_create_error = e; return LSG_ERR_ARG; }
  auto r = std::make_unique<lsg_ragged>();
  e = r->R.setup(fused, off);
  const size_t nb = sizeof(double2) * size_t(off.back());
  if (e.empty() && (hipMalloc((void**)&r->db, nb) != hipSuccess || hipMalloc((void**)&r->dx, nb) != hipSuccess ||
                    hipHostMalloc((void**)&r->pin, nb, hipHostMallocDefault) != hipSuccess))
    e = "allocation of the fused vectors failed";
  if (!e.empty()) {
    lsg_ragged_destroy(r.release());
    g_create_error = e;
    return LSG_ERR_DEVICE;
  }
  *out = r.release();
  return LSG_OK;
}

void lsg_ragged_destroy(lsg_ragged* r) {
  if (!r) return;
  r->R.release();
  if (r->db) (void)hipFree(r->db);
  if (r->dx) (void)hipFree(r->dx);
  if (r->pin) (void)hipHostFree(r->pin);
  delete r;
}

const char* lsg_ragged_error(const lsg_ragged* r) { return r ? r->err.c_str() : "ragged batch is NULL"; }

int64_t lsg_ragged_size(const lsg_ragged* r, int i) {
  if (!r || i < -1 || i >= r->R.n) return LSG_ERR_ARG;
  if (i < 0) return r->R.off.back();
  return r->R.off[size_t(i) + 1] - r->R.off[size_t(i)];
}

int lsg_ragged_matvec(lsg_ragged* r, const double* const* x, double* const* y) {
  if (!r) return LSG_ERR_ARG;
  if (!x || !y) return r->set(LSG_ERR_ARG, "x and y must hold one vector per system");
  for (int i = 0; i < r->R.n; ++i) {
    if (!x[i] || !y[i]) return r->set(LSG_ERR_ARG, "a vector is NULL");
    if (!all_finite(x[i], 2 * lsg_ragged_size(r, i))) return r->set(LSG_ERR_NUMERIC, "input vector holds NaN/Inf");
  }
  if (!r->gather(r->dx, x)) return r->set(LSG_ERR_DEVICE, "copy in failed");
  r->R.all_active();
  r->R.op.matvec(r->dx, r->db);
  if (!r->scatter(y, r->db)) return r->set(LSG_ERR_DEVICE, "matvec or copy out failed");
  return LSG_OK;
}

int lsg_ragged_gmres(lsg_ragged* r, const double* const* b, double* const* x, int restart_len, int max_restarts,
                     double tol, lsg_stats* stats) {
  if (!r) return LSG_ERR_ARG;
  if (!b || !x) return r->set(LSG_ERR_ARG, "b and x must hold one vector per system");
  for (int i = 0; i < r->R.n; ++i) {
    if (!b[i] || !x[i]) return r->set(LSG_ERR_ARG, "a vector is NULL");
    const int64_t Ni = lsg_ragged_size(r, i);
    if (!all_finite(b[i], 2 * Ni) || !all_finite(x[i], 2 * Ni)) return r->set(LSG_ERR_NUMERIC, "b or x holds NaN/Inf");
  }
  if (!r->gather(r->db, b) || !r->gather(r->dx, x)) return r->set(LSG_ERR_DEVICE, "copy in failed");
  RaggedResult res = ragged_gmres(r->R, r->db, r->dx, restart_len, max_restarts, tol);
  if (res.status != LSG_ERR_ARG && !r->scatter(x, r->dx) && (res.status == 0 || res.status == LSG_ERR_NOCONV)) {
    res.status = LSG_ERR_DEVICE;
    res.error = "copy out failed";
  }
  if (stats)
    for (int i = 0; i < r->R.n; ++i) {
      const RaggedStats& q = res.sys[size_t(i)];
      stats[i] = lsg_stats{};
      stats[i].iterations = q.iterations;
      stats[i].restarts = q.restarts;
      stats[i].converged = q.converged;
      stats[i].rel_residual = q.rel_residual;
      stats[i].seconds = res.seconds;
    }
  if (res.status != 0) return r->set(res.status, res.error);
  return LSG_OK;
}

int lsg_ragged_autotune(lsg_ragged* r, int reps, int* n_changed) {
  if (!r || reps < 1) return LSG_ERR_ARG;
  const bool had_graph = r->R.op.graph_exec != nullptr;
  r->R.all_active();
  std::string e = r->R.op.autotune(reps, n_changed);
  if (e.empty() && had_graph) e = r->R.op.capture_graph();
  if (!e.empty()) return r->set(LSG_ERR_DEVICE, e);
  return LSG_OK;
}

int lsg_ragged_use_graph(lsg_ragged* r, int on) {
  if (!r) return LSG_ERR_ARG;
  if (!on) {
    r->R.op.drop_graph();
    return LSG_OK;
  }
  std::string e = r->R.op.capture_graph();
  return e.empty() ? LSG_OK : r->set(LSG_ERR_DEVICE, e);
}

int lsg_ragged_time_matvec(lsg_ragged* r, int reps, double* ms) {
  if (!r || reps < 1 || !ms) return LSG_ERR_ARG;
  r->R.all_active();
  *ms = r->R.op.time_matvec(reps);
  if (*ms < 0) return r->set(LSG_ERR_DEVICE, "time_matvec allocation failed");
  return LSG_OK;
}
const char* lsg_host_error(void) { return g_host_error.c_str(); }

int lsg_host_info(const lsg_host* h, lsg_info* info) {
  if (!h || !info) return LSG_ERR_ARG;
  fill_info(h->hb, info);
  return LSG_OK;
}

int64_t lsg_host_far_table(const lsg_host* h, double* out) {
  if (!h) return LSG_ERR_ARG;
  if (out) std::memcpy(out, h->hb.far.data(), sizeof(cplx) * h->hb.far.size());
  return int64_t(h->hb.far.size());
}

int64_t lsg_host_near_tables(const lsg_host* h, double* out) {
  if (!h) return LSG_ERR_ARG;
  if (out) std::memcpy(out, h->hb.plan.T.data(), sizeof(cplx) * h->hb.plan.T.size());
  return int64_t(h->hb.plan.T.size());
}

int64_t lsg_host_skeleton(const lsg_host* h, int* out) {
  if (!h) return LSG_ERR_ARG;
  if (out) std::memcpy(out, h->hb.plan.skeleton.data(), sizeof(int) * h->hb.plan.skeleton.size());
  return int64_t(h->hb.plan.skeleton.size());
}

int lsg_host_il_pairs(const lsg_host* h, int64_t* out, int max_levels) {
  if (!h || !out) return LSG_ERR_ARG;
  const Tree& t = h->hb.tree;
  const int n = std::min<int>(max_levels, t.nLevels + 1);
  for (int l = 0; l < n; ++l) {
    int64_t s = 0;
    for (const Box& b : t.lv[l]) {
      s += int64_t(b.il.size());
      for (const auto& c : b.cone_il) s += int64_t(c.size());
    }
    out[l] = s;
  }
  return n;
}

int64_t lsg_host_qpinv(const lsg_host* h, double* out) {
  if (!h) return LSG_ERR_ARG;
  if (out) std::memcpy(out, h->hb.bs.qpinv.data(), sizeof(double) * h->hb.bs.qpinv.size());
  return int64_t(h->hb.bs.qpinv.size());
}

int64_t lsg_host_plan_ops(const lsg_host* h, double* out) {
  if (!h) return LSG_ERR_ARG;
  if (out) std::memcpy(out, h->hb.plan.ops.data(), sizeof(cplx) * h->hb.plan.ops.size());
  return int64_t(h->hb.plan.ops.size());
}

int64_t lsg_host_plan_src(const lsg_host* h, int* out) {
  if (!h) return LSG_ERR_ARG;
  if (out) std::memcpy(out, h->hb.plan.src.data(), sizeof(int) * h->hb.plan.src.size());
  return int64_t(h->hb.plan.src.size());
}

int64_t lsg_host_plan_dests(const lsg_host* h, int64_t* out) {
  if (!h) return LSG_ERR_ARG;
  int64_t n = 0;
  const auto& ps = h->hb.plan.passes;
  for (size_t p = 0; p < ps.size(); ++p)
    for (size_t d = 0; d < ps[p].rows.size(); ++d, ++n)
      if (out) {
        int64_t* o = out + 7 * n;
        o[0] = int64_t(p); o[1] = ps[p].kind; o[2] = ps[p].rows[d]; o[3] = ps[p].ncol[d];
        o[4] = ps[p].dst[d]; o[5] = ps[p].op_off[d]; o[6] = ps[p].src_off[d];
      }
  return n;
}

namespace {
// Entry kernel of a host build, and its GPU counterpart when asked for.
int probe_kernel(const lsg_host* h, int device, EntryKernel& K, std::unique_ptr<DeviceNca>& dev) {
  if (!make_entry_kernel(h->hb.tree, h->hb.bs, h->hb.prob, h->hb.in.find_phi != 0, h->hb.table_dir, K).empty())
    return LSG_ERR_SETUP;
  if (device) {
    std::string e;
    dev.reset(make_device_nca(K, e));
    if (!dev) return LSG_ERR_DEVICE;
  }
  return LSG_OK;
}
}  // namespace

int lsg_write_precomputations(const lsg_inputs* in, const char* dir) {
  if (!in || !dir) return LSG_ERR_ARG;
  std::string e = validate(in);
  if (!e.empty()) {
    g_create_error = e;
    return LSG_ERR_ARG;
  }
  BasisSet bs;
  bs.init(in->nChebNodes, in->degreeOfBases);
  // the far-field tables cover the tree's levels (writeMToFile runs after the
  // tree is built, G_FMM2DTree.hpp:1252-1269): an adaptive tree is grown first
  int n_levels = in->nLevelsUniform;
  if (in->tree_mode == 1) {
    HostBuild hb;
    if (host_build(in, hb, e, true) != LSG_OK) {
      g_create_error = e;
      return LSG_ERR_SETUP;
    }
    n_levels = hb.tree.nLevels;
  }
  e = write_precomputations(dir, bs, in->kappa, in->L, n_levels, 10);
  if (!e.empty()) {
    g_create_error = e;
    return LSG_ERR_SETUP;
  }
  return LSG_OK;
}

int lsg_host_entries(const lsg_host* h, const int* rows, int nr, const int* cols, int nc, int device, double* out) {
  if (!h || !rows || !cols || !out || nr < 1 || nc < 1) return LSG_ERR_ARG;
  for (int i = 0; i < nr; ++i)
    if (rows[i] < 0 || rows[i] >= h->hb.tree.N) return LSG_ERR_ARG;
  for (int i = 0; i < nc; ++i)
    if (cols[i] < 0 || cols[i] >= h->hb.tree.N) return LSG_ERR_ARG;
  EntryKernel K;
  std::unique_ptr<DeviceNca> dev;
  int rc = probe_kernel(h, device, K, dev);
  if (rc != LSG_OK) return rc;
  cplx* o = reinterpret_cast<cplx*>(out);
  if (!dev) {
    K.block(rows, nr, cols, nc, o);
    return LSG_OK;
  }
  std::vector<BlockJob> jobs{BlockJob{rows, nr, cols, nc, o}};
  return dev->blocks(jobs).empty() ? LSG_OK : LSG_ERR_DEVICE;
}

int lsg_host_aca(const lsg_host* h, const int* rows, int n1, const int* cols, int n2, int device, int* rank,
                 int* row_piv, int* col_piv) {
  if (!h || !rows || !cols || !rank || !row_piv || !col_piv || n1 < 1 || n2 < 1) return LSG_ERR_ARG;
  for (int i = 0; i < n1; ++i)
    if (rows[i] < 0 || rows[i] >= h->hb.tree.N) return LSG_ERR_ARG;
  for (int i = 0; i < n2; ++i)
    if (cols[i] < 0 || cols[i] >= h->hb.tree.N) return LSG_ERR_ARG;
  EntryKernel K;
  std::unique_ptr<DeviceNca> dev;
  int rc = probe_kernel(h, device, K, dev);
  if (rc != LSG_OK) return rc;
  const std::vector<int> r(rows, rows + n1), c(cols, cols + n2);
  const double eps = std::pow(10.0, -1.0 * h->hb.in.TOL_POW);
  const double abs_exit = h->hb.in.aca_exit_pow ? std::pow(10.0, -h->hb.in.aca_exit_pow) : 1e-8;
  AcaResult res;
  if (dev) {
    std::vector<AcaJob> jobs{AcaJob{&r, &c, &res}};
    if (!dev->aca(jobs, eps, abs_exit).empty()) return LSG_ERR_DEVICE;
  } else {
    aca_only_nodes(K, r, c, eps, res, abs_exit);
  }
  *rank = res.rank;
  std::copy(res.row_bases.begin(), res.row_bases.end(), row_piv);
  std::copy(res.col_bases.begin(), res.col_bases.end(), col_piv);
  return LSG_OK;
}

int64_t lsg_host_plan_nbr(const lsg_host* h, int* out) {
  if (!h) return LSG_ERR_ARG;
  if (out) std::memcpy(out, h->hb.plan.nbr.data(), sizeof(int) * h->hb.plan.nbr.size());
  return int64_t(h->hb.plan.nbr.size());
}

int64_t lsg_host_plan_qk2(const lsg_host* h, double* out) {
  if (!h) return LSG_ERR_ARG;
  if (out) std::memcpy(out, h->hb.plan.qk2.data(), sizeof(double) * h->hb.plan.qk2.size());
  return int64_t(h->hb.plan.qk2.size());
}

int lsg_host_level_boxes(const lsg_host* h, int* boxes, int* leaves, int max_levels) {
  if (!h || !boxes || !leaves) return LSG_ERR_ARG;
  const Tree& t = h->hb.tree;
  const int n = std::min<int>(max_levels, t.nLevels + 1);
  for (int l = 0; l < n; ++l) {
    boxes[l] = int(t.lv[l].size());
    int nl = 0;
    for (const Box& b : t.lv[l]) nl += b.leaf ? 1 : 0;
    leaves[l] = nl;
  }
  return n;
}

int64_t lsg_host_leaf_boxes(const lsg_host* h, int* out) {
  if (!h) return LSG_ERR_ARG;
  const Tree& t = h->hb.tree;
  if (out)
    for (size_t i = 0; i < t.leaves.size(); ++i) {
      out[2 * i] = t.leaves[i].l;
      out[2 * i + 1] = t.leaves[i].k;
    }
  return int64_t(t.leaves.size());
}

int64_t lsg_host_aca_shapes(const lsg_host* h, int* out) {
  if (!h) return LSG_ERR_ARG;
  if (out) std::memcpy(out, h->hb.plan.aca.data(), sizeof(int) * h->hb.plan.aca.size());
  return int64_t(h->hb.plan.aca.size());
}

int lsg_host_level_ranks(const lsg_host* h, int* out, int max_levels) {
  if (!h || !out) return LSG_ERR_ARG;
  const auto& v = h->hb.plan.level_max_rank;
  const int n = std::min<int>(max_levels, int(v.size()));
  for (int i = 0; i < n; ++i) out[i] = v[i];
  return n;
}

}  // extern "C"
