// Ragged batch on one device (gfx950): fused plan of independent scatterers
// and a batched GMRES with per-system Krylov bases over segments of one vector.
#include "ragged.hpp"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>

#include "../../include/lsg.h"

namespace lsg {

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return std::string(#expr ": ") + hipGetErrorString(e_); \
  } while (0)

// ------------------------------------------------------------ fused plan ----
namespace {
enum Role { kUp = 0, kDown = 1, kLeafM2L = 2, kLeafL2L = 3 };
int role_of(const std::string& name) {
  if (name.rfind("up_", 0) == 0) return kUp;
  if (name == "leaf_m2l") return kLeafM2L;
  if (name == "leaf_l2l") return kLeafL2L;
  if (name == "near") return -1;
  return kDown;
}
}  // namespace

std::string fuse_plans(const std::vector<const Plan*>& P, Plan& F, std::vector<int>& off) {
  const int n = int(P.size());
  if (n < 1) return "no plans to fuse";
  F = Plan{};
  const int rank = P[0]->rank;
  F.rank = rank;
  F.find_phi = P[0]->find_phi;
  const size_t set_sz = size_t(9) * rank * rank;
  off.assign(size_t(n) + 1, 0);
  int64_t slots = 0, leaves = 0;
  for (int i = 0; i < n; ++i) {
    const Plan& p = *P[i];
    if (p.rank != rank) return "fused systems must have the same nodes per leaf";
    if (p.find_phi != F.find_phi) return "fused systems must all be solve operators (find_phi) or all volume potentials";
    if (p.T.size() != set_sz || p.nbr.size() != size_t(p.n_leaves) * 9) return "fused systems must be uniform trees";
    if (p.n_leaves % 16) return "fused systems need a multiple of 16 leaves (uniform depth >= 2)";
    int64_t entries = 0;
    for (const PassHost& ph : p.passes) {
      if (role_of(ph.name) < 0) return "fused systems must be uniform trees (no near-field pass)";
      entries += ph.entries;
    }
    if (int64_t(p.ops.size()) != entries) return "a host build's operators are incomplete";
    off[size_t(i) + 1] = off[size_t(i)] + p.N;
    slots += p.n_state - 2 * int64_t(p.N);
    leaves += p.n_leaves;
  }
  const int64_t Ntot = off[size_t(n)];
  if (2 * Ntot + slots >= int64_t(INT_MAX)) return "fused state exceeds 32-bit indices";
  F.N = int(Ntot);
  F.n_leaves = int(leaves);
  F.n_state = 2 * Ntot + slots;
  // state index of plan i -> fused state index
  std::vector<int64_t> soff(size_t(n), 0);
  for (int i = 1; i < n; ++i) soff[size_t(i)] = soff[size_t(i) - 1] + P[i - 1]->n_state - 2 * int64_t(P[i - 1]->N);
  auto remap = [&](int i, int64_t s) -> int64_t {
    const int64_t Ni = P[i]->N;
    if (s < Ni) return off[size_t(i)] + s;
    if (s < 2 * Ni) return Ntot + off[size_t(i)] + (s - Ni);
    return 2 * Ntot + soff[size_t(i)] + (s - 2 * Ni);
  };
  // operators and source lists back to back, plan by plan
  std::vector<int64_t> op_base(static_cast<size_t>(n)), src_base(static_cast<size_t>(n));
  size_t n_ops = 0, n_src = 0;
  for (int i = 0; i < n; ++i) {
    op_base[size_t(i)] = int64_t(n_ops);
    src_base[size_t(i)] = int64_t(n_src);
    n_ops += P[i]->ops.size();
    n_src += P[i]->src.size();
  }
  F.ops.reserve(n_ops);
  F.src.reserve(n_src);
  for (int i = 0; i < n; ++i) {
    F.ops.insert(F.ops.end(), P[i]->ops.begin(), P[i]->ops.end());
    for (int s : P[i]->src) F.src.push_back(int(remap(i, s)));
  }
  // passes by role and position within the role
  std::vector<std::vector<std::vector<int>>> by_role(4, std::vector<std::vector<int>>(size_t(n)));
  for (int i = 0; i < n; ++i)
    for (size_t pi = 0; pi < P[i]->passes.size(); ++pi) by_role[size_t(role_of(P[i]->passes[pi].name))][size_t(i)].push_back(int(pi));
  static const char* kName[4] = {"up_f", "down_f", "leaf_m2l", "leaf_l2l"};
  for (int r = 0; r < 4; ++r) {
    size_t K = 0;
    for (int i = 0; i < n; ++i) K = std::max(K, by_role[size_t(r)][size_t(i)].size());
    if (r >= kLeafM2L && K > 1) return "a plan holds two leaf passes of one kind";
    for (size_t k = 0; k < K; ++k) {
      PassHost m;
      m.name = r <= kDown ? std::string(kName[r]) + std::to_string(k) : std::string(kName[r]);
      m.level = -1;
      m.kind = -1;
      for (int i = 0; i < n; ++i) {
        if (k >= by_role[size_t(r)][size_t(i)].size()) continue;
        const PassHost& ph = P[i]->passes[size_t(by_role[size_t(r)][size_t(i)][k])];
        if (m.kind < 0) m.kind = ph.kind;
        if (m.kind != ph.kind) return "passes of one role differ in kind";
        for (size_t q = 0; q < ph.rows.size(); ++q) {
          F.dest_sys.push_back(i);
          m.rows.push_back(ph.rows[q]);
          m.ncol.push_back(ph.ncol[q]);
          m.dst.push_back(int(remap(i, ph.dst[q])));
          m.op_off.push_back(op_base[size_t(i)] + ph.op_off[q]);
          m.src_off.push_back(src_base[size_t(i)] + ph.src_off[q]);
        }
        m.entries += ph.entries;
        m.n_src += ph.n_src;
        m.n_dst += ph.n_dst;
      }
      F.passes.push_back(std::move(m));
    }
  }
  // near field: one table set per plan, one set index per 16 leaves
  int64_t loff = 0;
  for (int i = 0; i < n; ++i) {
    const Plan& p = *P[i];
    F.T.insert(F.T.end(), p.T.begin(), p.T.end());
    for (int g = 0; g < p.n_leaves / 16; ++g) F.near_group_set.push_back(i);
    for (int v : p.nbr) F.nbr.push_back(v >= 0 ? int(v + loff) : -1);
    F.qk2.insert(F.qk2.end(), p.qk2.begin(), p.qk2.end());
    F.E_far += p.E_far;
    F.E_near += p.E_near;
    F.max_rank = std::max(F.max_rank, p.max_rank);
    loff += p.n_leaves;
  }
  return "";
}

// --------------------------------------------------------------- kernels ----
namespace {
constexpr int kT = 256;
constexpr int kRB = 2 * kT;   // rows per row block (two per thread)
constexpr int kCols = 8;      // Krylov columns per workgroup in the dot kernel
constexpr int kMaxM = 500;

__device__ __forceinline__ double2 wave_sum(double2 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o);
    v.y += __shfl_xor(v.y, o);
  }
  return v;
}
// acc += conj(v) w
__device__ __forceinline__ void cdot(double2& acc, double2 v, double2 w) {
  acc.x = fma(v.x, w.x, fma(v.y, w.y, acc.x));
  acc.y = fma(v.x, w.y, fma(-v.y, w.x, acc.y));
}
// acc += v h
__device__ __forceinline__ void cmac(double2& acc, double2 v, double2 h) {
  acc.x = fma(v.x, h.x, fma(-v.y, h.y, acc.x));
  acc.y = fma(v.x, h.y, fma(v.y, h.x, acc.y));
}

// workgroup sums of acc[0..kCols) into partial[blk*ncx + g0 + u] (u < ncx - g0)
__device__ void store_partials(double2 (*ws)[kCols], const double2 (&acc)[kCols], int g0, int ncx, double2* partial,
                               long long blk) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
  for (int u = 0; u < kCols; ++u) {
    const double2 s = wave_sum(acc[u]);
    if (lane == 0) ws[wv][u] = s;
  }
  __syncthreads();
  if (t < kCols && g0 + t < ncx) {
    double2 s = ws[0][t];
    for (int q = 1; q < kT / 64; ++q) {
      s.x += ws[q][t].x;
      s.y += ws[q][t].y;
    }
    partial[blk * ncx + g0 + t] = s;
  }
  __syncthreads();
}

// partial[rb*ncx + col]: conj(V_col) . w over row block rb (col < nc), |w|^2 (col == nc)
__global__ __launch_bounds__(kT) void rg_dot_kernel(const double2* __restrict__ V, long long ldv, int nc, int ncx,
                                                    const double2* __restrict__ w, const int* __restrict__ rb_row0,
                                                    const int* __restrict__ rb_nrow, const int* __restrict__ rb_sys,
                                                    const int* __restrict__ active, double2* __restrict__ partial) {
  __shared__ double2 ws[kT / 64][kCols];
  const int t = threadIdx.x, rb = blockIdx.x, g0 = blockIdx.y * kCols;
  // the partials of inactive systems are never read (active == nullptr: all systems)
  if (active && !active[rb_sys[rb]]) return;
  const int row0 = rb_row0[rb], nr = rb_nrow[rb];
  const double2 z = make_double2(0.0, 0.0);
  long long n[2];
  bool ok[2];
  double2 wv[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int i = t + r * kT;
    ok[r] = i < nr;
    n[r] = (long long)row0 + i;
    wv[r] = ok[r] ? w[n[r]] : z;
  }
  double2 acc[kCols];
#pragma unroll
  for (int u = 0; u < kCols; ++u) {
    const int col = g0 + u;
    acc[u] = z;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      double2 v = z;
      if (col < nc) {
        if (ok[r]) v = V[col * ldv + n[r]];
      } else if (col == nc) {
        v = wv[r];
      }
      cdot(acc[u], v, wv[r]);
    }
  }
  store_partials(ws, acc, g0, ncx, partial, rb);
}

// sums of columns [0, ncx) of system blockIdx.x's row blocks, into LDS (and out)
__device__ void reduce_system(const double2* __restrict__ partial, int ncx, int rb0, int rb1, double2* out,
                              double2* lds, double2* red) {
  const int t = threadIdx.x;
  for (int c0 = 0; c0 < ncx; c0 += kT) {
    const int ncc = min(kT, ncx - c0);
    const int S = kT / ncc;
    const int c = t % ncc, sg = t / ncc;
    double2 a = make_double2(0.0, 0.0);
    if (sg < S)
      for (int q = rb0 + sg; q < rb1; q += S) {
        const double2 v = partial[(long long)q * ncx + c0 + c];
        a.x += v.x;
        a.y += v.y;
      }
    red[t] = a;
    __syncthreads();
    if (t < ncc) {
      double2 s = red[t];
      for (int q = 1; q < S; ++q) {
        s.x += red[q * ncc + t].x;
        s.y += red[q * ncc + t].y;
      }
      if (out) out[c0 + t] = s;
      if (lds) lds[c0 + t] = s;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kT) void rg_reduce_kernel(const double2* __restrict__ partial, int ncx,
                                                       const int* __restrict__ sys_rb0, double2* __restrict__ out,
                                                       int ostride) {
  __shared__ double2 red[kT];
  const int s = blockIdx.x;
  reduce_system(partial, ncx, sys_rb0[s], sys_rb0[s + 1], out + (long long)s * ostride, nullptr, red);
}

// ||v_s|| of every system (one column of partials)
__global__ __launch_bounds__(kT) void rg_norm_kernel(const double2* __restrict__ partial,
                                                     const int* __restrict__ sys_rb0, double* __restrict__ out) {
  __shared__ double2 red[kT];
  __shared__ double2 r2[1];
  const int s = blockIdx.x;
  reduce_system(partial, 1, sys_rb0[s], sys_rb0[s + 1], nullptr, r2, red);
  if (threadIdx.x == 0) out[s] = sqrt(fmax(r2[0].x, 0.0));
}

// w <- w - V h1_s per row; partials of V^H w' (nc columns) and |w'|^2 (column nc)
__global__ __launch_bounds__(kT) void rg_proj_dot_kernel(const double2* __restrict__ V, long long ldv, int nc,
                                                         const double2* __restrict__ h1, int ostride,
                                                         double2* __restrict__ w, const int* __restrict__ rb_row0,
                                                         const int* __restrict__ rb_nrow, const int* __restrict__ rb_sys,
                                                         const int* __restrict__ active, double2* __restrict__ partial) {
  __shared__ double2 hs[kMaxM + 1];
  __shared__ double2 ws[kT / 64][kCols];
  const int t = threadIdx.x, rb = blockIdx.x;
  if (!active[rb_sys[rb]]) return;  // the rows of inactive systems are not used
  const double2* hsys = h1 + (long long)rb_sys[rb] * ostride;
  for (int i = t; i < nc; i += kT) hs[i] = hsys[i];
  __syncthreads();
  const int row0 = rb_row0[rb], nr = rb_nrow[rb];
  const double2 z = make_double2(0.0, 0.0);
  long long n[2];
  bool ok[2];
  double2 wv[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int i = t + r * kT;
    ok[r] = i < nr;
    n[r] = (long long)row0 + i;
    wv[r] = ok[r] ? w[n[r]] : z;
  }
  {
    double2 a[2] = {z, z};
    for (int i = 0; i < nc; ++i)
#pragma unroll
      for (int r = 0; r < 2; ++r)
        if (ok[r]) cmac(a[r], V[i * ldv + n[r]], hs[i]);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      wv[r].x -= a[r].x;
      wv[r].y -= a[r].y;
      if (ok[r]) w[n[r]] = wv[r];
    }
  }
  const int ncx = nc + 1;
  for (int g0 = 0; g0 < ncx; g0 += kCols) {
    double2 acc[kCols];
#pragma unroll
    for (int u = 0; u < kCols; ++u) {
      const int col = g0 + u;
      acc[u] = z;
#pragma unroll
      for (int r = 0; r < 2; ++r) {
        double2 v = z;
        if (col < nc) {
          if (ok[r]) v = V[col * ldv + n[r]];
        } else if (col == nc) {
          v = wv[r];
        }
        cdot(acc[u], v, wv[r]);
      }
    }
    store_partials(ws, acc, g0, ncx, partial, rb);
  }
}

struct RgGivens {
  double2* H;      // per system (m+1) x m, column-major
  double* cs;      // per system m+1
  double2* sn;     // per system m+1
  double2* g;      // per system m+2
  double* hn;      // per system
  double* beta;    // per system
  double* bnorm;   // per system
  double* tolq;    // per system
  int* active;     // per system
  int* kq;         // per system: Arnoldi steps of this cycle
};

// Cycle start of system blockIdx.x: beta = ||r_s||, g = beta e1; active if
// beta / ||b_s|| > tol.
__global__ __launch_bounds__(kT) void rg_cycle_init_kernel(const double2* __restrict__ partial,
                                                           const int* __restrict__ sys_rb0, RgGivens st, int m,
                                                           double* host_rel, int* host_active) {
  __shared__ double2 red[kT];
  __shared__ double2 r2[1];
  const int s = blockIdx.x, t = threadIdx.x;
  reduce_system(partial, 1, sys_rb0[s], sys_rb0[s + 1], nullptr, r2, red);
  const double b = sqrt(fmax(r2[0].x, 0.0));
  const double rel = b / st.bnorm[s];
  const int act = (rel > st.tolq[s] && b > 0.0) ? 1 : 0;
  double2* g = st.g + (long long)s * (m + 2);
  for (int i = t; i <= m; i += kT) g[i] = make_double2(i == 0 ? b : 0.0, 0.0);
  if (t == 0) {
    st.beta[s] = b;
    st.active[s] = act;
    st.kq[s] = 0;
    st.hn[s] = act ? b : 0.0;
    host_rel[s] = rel;
    host_active[s] = act;
  }
}

// v = w / beta_s for active systems, 0 otherwise
__global__ __launch_bounds__(kT) void rg_scale_kernel(const double2* __restrict__ w, RgGivens st,
                                                      double2* __restrict__ v, const int* __restrict__ rb_row0,
                                                      const int* __restrict__ rb_nrow, const int* __restrict__ rb_sys) {
  const int rb = blockIdx.x, s = rb_sys[rb];
  const double d = st.active[s] ? st.beta[s] : 0.0;
  const double inv = d > 1e-300 ? 1.0 / d : 0.0;
  for (int i = threadIdx.x; i < rb_nrow[rb]; i += kT) {
    const long long n = (long long)rb_row0[rb] + i;
    v[n] = make_double2(w[n].x * inv, w[n].y * inv);
  }
}

// Arnoldi step j of system blockIdx.x after both Gram-Schmidt passes: h2 from
// the second pass's partials (into d2), H column j = h1 + h2, ||w''||, the
// previous rotations, the new one; freeze at 0.5 tol or below 1e-8 (breakdown).
__global__ __launch_bounds__(kT) void rg_tail_kernel(const double2* __restrict__ partial,
                                                     const int* __restrict__ sys_rb0, const double2* __restrict__ d1,
                                                     double2* __restrict__ d2, int ostride, RgGivens st, int m, int j,
                                                     int n_sys, double* host_est, int* host_active) {
  __shared__ double2 red[kT];
  __shared__ double2 h2[kMaxM + 2];
  __shared__ double2 col[kMaxM + 2];
  const int s = blockIdx.x, t = threadIdx.x;
  reduce_system(partial, j + 2, sys_rb0[s], sys_rb0[s + 1], d2 + (long long)s * ostride, h2, red);
  if (t != 0) return;
  if (!st.active[s]) {
    st.hn[s] = 0.0;
    host_est[(long long)j * n_sys + s] = -1.0;
    host_active[s] = 0;
    return;
  }
  const double2* h1 = d1 + (long long)s * ostride;
  double s2 = 0.0;
  for (int i = 0; i <= j; ++i) {
    col[i] = make_double2(h1[i].x + h2[i].x, h1[i].y + h2[i].y);
    s2 += h2[i].x * h2[i].x + h2[i].y * h2[i].y;
  }
  const double hn = sqrt(fmax(h2[j + 1].x - s2, 0.0));
  double* cs = st.cs + (long long)s * (m + 1);
  double2* sn = st.sn + (long long)s * (m + 1);
  double2* g = st.g + (long long)s * (m + 2);
  col[j + 1] = make_double2(hn, 0.0);
  for (int i = 0; i < j; ++i) {
    const double c = cs[i];
    const double2 sv = sn[i];
    const double2 a = col[i], b = col[i + 1];
    col[i] = make_double2(c * a.x + (sv.x * b.x - sv.y * b.y), c * a.y + (sv.x * b.y + sv.y * b.x));
    col[i + 1] = make_double2(c * b.x - (sv.x * a.x + sv.y * a.y), c * b.y - (sv.x * a.y - sv.y * a.x));
  }
  const double2 a = col[j];
  const double an = hypot(a.x, a.y);
  double c;
  double2 sv;
  if (an == 0.0) {
    c = 0.0;
    sv = make_double2(1.0, 0.0);
    col[j] = make_double2(hn, 0.0);
  } else {
    const double r = hypot(an, hn);
    c = an / r;
    const double2 ph = make_double2(a.x / an, a.y / an);
    sv = make_double2(ph.x * hn / r, ph.y * hn / r);
    col[j] = make_double2(ph.x * r, ph.y * r);
  }
  col[j + 1] = make_double2(0.0, 0.0);
  double2* Hc = st.H + (long long)s * (m + 1) * m + (long long)j * (m + 1);
  for (int i = 0; i <= j + 1; ++i) Hc[i] = col[i];
  cs[j] = c;
  sn[j] = sv;
  const double2 gj = g[j];
  g[j] = make_double2(c * gj.x, c * gj.y);
  const double2 g1 = make_double2(-(sv.x * gj.x + sv.y * gj.y), -(sv.x * gj.y - sv.y * gj.x));
  g[j + 1] = g1;
  const double est = hypot(g1.x, g1.y) / st.bnorm[s];
  st.kq[s] = j + 1;
  const bool stop = est <= 0.5 * st.tolq[s] || fabs(hn) < 1e-8;
  if (stop) st.active[s] = 0;
  // a frozen system's next basis vector is zero (it is never used)
  st.hn[s] = stop ? 0.0 : hn;
  host_est[(long long)j * n_sys + s] = est;
  host_active[s] = stop ? 0 : 1;
}

// v_{j+1} = (w' - V h2_s) / hn_s per row (zero where hn_s = 0)
__global__ __launch_bounds__(kT) void rg_update_kernel(const double2* __restrict__ V, long long ldv, int nc,
                                                       const double2* __restrict__ d2, int ostride,
                                                       const double2* __restrict__ w, const double* __restrict__ hn,
                                                       double2* __restrict__ vout, const int* __restrict__ rb_row0,
                                                       const int* __restrict__ rb_nrow, const int* __restrict__ rb_sys) {
  __shared__ double2 hs[kMaxM + 1];
  const int t = threadIdx.x, rb = blockIdx.x, s = rb_sys[rb];
  const double d = hn[s];
  // a system frozen at this step or before has hn = 0: its next basis vector
  // is never read (the dots, the matvec and the solution update skip it)
  if (!(d > 1e-300)) return;
  for (int i = t; i < nc; i += kT) hs[i] = d2[(long long)s * ostride + i];
  __syncthreads();
  const double inv = 1.0 / d;
  for (int r = t; r < rb_nrow[rb]; r += kT) {
    const long long n = (long long)rb_row0[rb] + r;
    double2 a = make_double2(0.0, 0.0);
    for (int i = 0; i < nc; ++i) cmac(a, V[i * ldv + n], hs[i]);
    const double2 wv = w[n];
    vout[n] = make_double2((wv.x - a.x) * inv, (wv.y - a.y) * inv);
  }
}

// y_s = R_s^{-1} g_s over the kq[s] steps of system blockIdx.x (zero beyond)
__global__ __launch_bounds__(64) void rg_trisolve_kernel(RgGivens st, int m, double2* __restrict__ y) {
  const int s = blockIdx.x;
  if (threadIdx.x != 0) return;
  const int k = st.kq[s];
  double2* ys = y + (long long)s * (m + 1);
  const double2* g = st.g + (long long)s * (m + 2);
  const double2* H = st.H + (long long)s * (m + 1) * m;
  for (int i = k; i <= m; ++i) ys[i] = make_double2(0.0, 0.0);
  for (int i = k - 1; i >= 0; --i) {
    double2 v = g[i];
    for (int c = i + 1; c < k; ++c) {
      const double2 r = H[(long long)c * (m + 1) + i], yc = ys[c];
      v.x -= r.x * yc.x - r.y * yc.y;
      v.y -= r.x * yc.y + r.y * yc.x;
    }
    const double2 dg = H[(long long)i * (m + 1) + i];
    const double den = dg.x * dg.x + dg.y * dg.y;
    ys[i] = den > 0.0 ? make_double2((v.x * dg.x + v.y * dg.y) / den, (v.y * dg.x - v.x * dg.y) / den)
                      : make_double2(0.0, 0.0);
  }
}

// x += V y_s over the kq[s] steps of the system's cycle (the only columns it
// wrote; systems without steps are left unchanged)
__global__ __launch_bounds__(kT) void rg_axpy_kernel(const double2* __restrict__ V, long long ldv,
                                                     const int* __restrict__ kq, const double2* __restrict__ y, int m,
                                                     double2* __restrict__ x, const int* __restrict__ rb_row0,
                                                     const int* __restrict__ rb_nrow, const int* __restrict__ rb_sys) {
  __shared__ double2 ys[kMaxM + 1];
  const int t = threadIdx.x, rb = blockIdx.x, s = rb_sys[rb];
  const int kmax = kq[s];
  if (kmax == 0) return;
  for (int i = t; i < kmax; i += kT) ys[i] = y[(long long)s * (m + 1) + i];
  __syncthreads();
  for (int r = t; r < rb_nrow[rb]; r += kT) {
    const long long n = (long long)rb_row0[rb] + r;
    double2 a = make_double2(0.0, 0.0);
    for (int i = 0; i < kmax; ++i) cmac(a, V[i * ldv + n], ys[i]);
    double2 o = x[n];
    o.x += a.x;
    o.y += a.y;
    x[n] = o;
  }
}

__global__ void rg_fill_kernel(int* __restrict__ a, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void rg_sub_kernel(const double2* __restrict__ b, double2* __restrict__ w, long long N) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) w[n] = make_double2(b[n].x - w[n].x, b[n].y - w[n].y);
}
}  // namespace

// ------------------------------------------------------------- operator ----
std::string RaggedOperator::setup(const Plan& F, const std::vector<int>& offsets) {
  release();
  std::string e = op.upload(F);
  if (!e.empty()) return e;
  off = offsets;
  n = int(off.size()) - 1;
  std::vector<int> r0, nr, sy, first;
  for (int s = 0; s < n; ++s) {
    first.push_back(int(r0.size()));
    for (int a = off[size_t(s)]; a < off[size_t(s) + 1]; a += kRB) {
      r0.push_back(a);
      nr.push_back(std::min(kRB, off[size_t(s) + 1] - a));
      sy.push_back(s);
    }
  }
  first.push_back(int(r0.size()));
  nrb = int(r0.size());
  auto up = [](const std::vector<int>& h, int** d) -> std::string {
    HIP_OK(hipMalloc((void**)d, sizeof(int) * h.size()));
    HIP_OK(hipMemcpy(*d, h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice));
    return "";
  };
  if (!(e = up(r0, &rb_row0)).empty() || !(e = up(nr, &rb_nrow)).empty() || !(e = up(sy, &rb_sys)).empty() ||
      !(e = up(first, &sys_rb0)).empty() || !(e = up(std::vector<int>(size_t(n), 1), &active)).empty())
    return e;
  // the passes skip the destinations of systems whose flag is 0
  op.sys_active = active;
  return "";
}

void RaggedOperator::all_active() {
  hipLaunchKernelGGL(rg_fill_kernel, dim3((n + 255) / 256), dim3(256), 0, op.stream, active, n, 1);
}

void RaggedOperator::release() {
  work.release();
  op.release();
  for (int* p : {rb_row0, rb_nrow, rb_sys, sys_rb0, active})
    if (p) (void)hipFree(p);
  rb_row0 = rb_nrow = rb_sys = sys_rb0 = active = nullptr;
}

std::string RaggedWork::reserve(int Ntot, int nrb, int n, int m_) {
  if (m == m_) return "";
  release();
  HIP_OK(hipMalloc((void**)&V, sizeof(double2) * size_t(Ntot) * (m_ + 1)));
  HIP_OK(hipMalloc((void**)&w, sizeof(double2) * size_t(Ntot)));
  HIP_OK(hipMalloc((void**)&partial, sizeof(double2) * size_t(nrb) * (m_ + 2)));
  HIP_OK(hipMalloc((void**)&d1, sizeof(double2) * size_t(n) * (m_ + 2)));
  HIP_OK(hipMalloc((void**)&d2, sizeof(double2) * size_t(n) * (m_ + 2)));
  HIP_OK(hipMalloc((void**)&y, sizeof(double2) * size_t(n) * (m_ + 1)));
  HIP_OK(hipMalloc((void**)&H, sizeof(double2) * size_t(n) * (m_ + 1) * m_));
  HIP_OK(hipMalloc((void**)&cs, sizeof(double) * size_t(n) * (m_ + 1)));
  HIP_OK(hipMalloc((void**)&sn, sizeof(double2) * size_t(n) * (m_ + 1)));
  HIP_OK(hipMalloc((void**)&g, sizeof(double2) * size_t(n) * (m_ + 2)));
  for (double** p : {&hn, &beta, &bnorm, &tolq}) HIP_OK(hipMalloc((void**)p, sizeof(double) * size_t(n)));
  HIP_OK(hipMalloc((void**)&kq, sizeof(int) * size_t(n)));
  HIP_OK(hipHostMalloc((void**)&host_rel, sizeof(double) * size_t(n), hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&host_est, sizeof(double) * size_t(n) * m_, hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&host_active, sizeof(int) * size_t(n), hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&host_kq, sizeof(int) * size_t(n), hipHostMallocDefault));
  m = m_;
  return "";
}

void RaggedWork::release() {
  void* dev[] = {V, w, partial, d1, d2, y, H, cs, sn, g, hn, beta, bnorm, tolq, kq};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  void* host[] = {host_rel, host_est, host_active, host_kq};
  for (void* p : host)
    if (p) (void)hipHostFree(p);
  *this = RaggedWork{};
}

// ---------------------------------------------------------------- solver ----
RaggedResult ragged_gmres(RaggedOperator& R, const double2* B, double2* X, int m, int max_restarts, double tol) {
  RaggedResult res;
  auto t0 = std::chrono::steady_clock::now();
  const int n = R.n, nrb = R.nrb;
  DeviceOperator& op = R.op;
  hipStream_t s = op.stream;
  const long long Ntot = op.N;
  res.sys.assign(size_t(n), RaggedStats{});
  if (m < 1 || m > kMaxM || max_restarts < 1 || !(tol > 0) || !std::isfinite(tol)) {
    res.status = LSG_ERR_ARG;
    res.error = "ragged GMRES: 1 <= restart_len <= 500, max_restarts >= 1, finite tol > 0";
    return res;
  }
  {
    std::string e = R.work.reserve(int(Ntot), nrb, n, m);
    if (!e.empty()) {
      R.work.release();
      res.status = LSG_ERR_DEVICE;
      res.error = "ragged GMRES workspace allocation failed: " + e;
      return res;
    }
  }
  RaggedWork& W = R.work;
  RgGivens st{W.H, W.cs, W.sn, W.g, W.hn, W.beta, W.bnorm, W.tolq, R.active, W.kq};
  const int ostride = m + 2;
  const unsigned gv = unsigned((Ntot + kT - 1) / kT);
  auto dot = [&](int nc, const double2* vec, bool norm, bool masked) {
    const int ncx = nc + (norm ? 1 : 0);
    hipLaunchKernelGGL(rg_dot_kernel, dim3(nrb, (ncx + kCols - 1) / kCols), dim3(kT), 0, s, W.V, Ntot, nc, ncx, vec,
                       R.rb_row0, R.rb_nrow, R.rb_sys, masked ? R.active : nullptr, W.partial);
  };
  auto fail = [&](int code, const std::string& msg) {
    (void)hipStreamSynchronize(s);
    res.status = code;
    res.error = msg;
  };
  // ||b_s|| of every system
  dot(0, B, true, false);
  hipLaunchKernelGGL(rg_norm_kernel, dim3(n), dim3(kT), 0, s, W.partial, R.sys_rb0, W.bnorm);
  std::vector<double> bn(static_cast<size_t>(n)), tolv(static_cast<size_t>(n), tol);
  if (hipMemcpyAsync(bn.data(), W.bnorm, sizeof(double) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    fail(LSG_ERR_DEVICE, "ragged GMRES: norm of b failed");
    return res;
  }
  for (int q = 0; q < n; ++q)
    if (!(bn[size_t(q)] > 0) || !std::isfinite(bn[size_t(q)])) {
      fail(LSG_ERR_NUMERIC, "ragged GMRES: right-hand side " + std::to_string(q) + " is zero or holds NaN/Inf");
      return res;
    }
  (void)hipMemcpyAsync(W.tolq, tolv.data(), sizeof(double) * n, hipMemcpyHostToDevice, s);
  std::vector<int> in_cycle(size_t(n), 0);
  for (int cyc = 0; cyc <= max_restarts; ++cyc) {
    // the true residual of every system (the cycle start then decides which
    // systems stay active; the Arnoldi matvecs skip the others)
    R.all_active();
    op.matvec(X, W.w);
    hipLaunchKernelGGL(rg_sub_kernel, dim3(gv), dim3(kT), 0, s, B, W.w, Ntot);
    dot(0, W.w, true, false);
    hipLaunchKernelGGL(rg_cycle_init_kernel, dim3(n), dim3(kT), 0, s, W.partial, R.sys_rb0, st, m, W.host_rel,
                       W.host_active);
    if (hipStreamSynchronize(s) != hipSuccess) {
      fail(LSG_ERR_DEVICE, "ragged GMRES: device error");
      break;
    }
    int n_act = 0;
    bool finite = true;
    for (int q = 0; q < n; ++q) {
      RaggedStats& rs = res.sys[size_t(q)];
      rs.rel_residual = W.host_rel[q];
      rs.converged = W.host_rel[q] <= tol ? 1 : 0;
      finite = finite && std::isfinite(W.host_rel[q]);
      in_cycle[size_t(q)] = W.host_active[q];
      n_act += W.host_active[q];
    }
    if (!finite) {
      fail(LSG_ERR_NUMERIC, "ragged GMRES: a residual is NaN/Inf");
      break;
    }
    if (n_act == 0 || cyc == max_restarts) break;
    hipLaunchKernelGGL(rg_scale_kernel, dim3(nrb), dim3(kT), 0, s, W.w, st, W.V, R.rb_row0, R.rb_nrow, R.rb_sys);
    int steps = 0;
    bool dev_err = false;
    for (int j = 0; j < m;) {
      const int batch = std::min(m - j, 4);
      for (int jj = j; jj < j + batch; ++jj) {
        op.matvec(W.V + size_t(jj) * Ntot, W.w);
        dot(jj + 1, W.w, false, true);
        hipLaunchKernelGGL(rg_reduce_kernel, dim3(n), dim3(kT), 0, s, W.partial, jj + 1, R.sys_rb0, W.d1, ostride);
        hipLaunchKernelGGL(rg_proj_dot_kernel, dim3(nrb), dim3(kT), 0, s, W.V, Ntot, jj + 1, W.d1, ostride, W.w,
                           R.rb_row0, R.rb_nrow, R.rb_sys, R.active, W.partial);
        hipLaunchKernelGGL(rg_tail_kernel, dim3(n), dim3(kT), 0, s, W.partial, R.sys_rb0, W.d1, W.d2, ostride, st, m,
                           jj, n, W.host_est, W.host_active);
        hipLaunchKernelGGL(rg_update_kernel, dim3(nrb), dim3(kT), 0, s, W.V, Ntot, jj + 1, W.d2, ostride, W.w, W.hn,
                           W.V + size_t(jj + 1) * Ntot, R.rb_row0, R.rb_nrow, R.rb_sys);
      }
      if (hipStreamSynchronize(s) != hipSuccess) {
        dev_err = true;
        break;
      }
      j += batch;
      steps = j;
      int still = 0;
      for (int q = 0; q < n; ++q) still += W.host_active[q];
      if (still == 0) break;
    }
    if (dev_err) {
      fail(LSG_ERR_DEVICE, "ragged GMRES: device error");
      break;
    }
    hipLaunchKernelGGL(rg_trisolve_kernel, dim3(n), dim3(64), 0, s, st, m, W.y);
    hipLaunchKernelGGL(rg_axpy_kernel, dim3(nrb), dim3(kT), 0, s, W.V, Ntot, W.kq, W.y, m, X, R.rb_row0, R.rb_nrow,
                       R.rb_sys);
    if (hipMemcpyAsync(W.host_kq, W.kq, sizeof(int) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      fail(LSG_ERR_DEVICE, "ragged GMRES: device error");
      break;
    }
    for (int q = 0; q < n; ++q)
      if (in_cycle[size_t(q)]) {
        res.sys[size_t(q)].iterations += W.host_kq[q];
        res.sys[size_t(q)].restarts += 1;
      }
  }
  if (res.status == 0)
    for (const RaggedStats& rs : res.sys)
      if (!rs.converged) {
        res.status = LSG_ERR_NOCONV;
        res.error = "ragged GMRES: not every system reached tol";
        break;
      }
  R.all_active();
  (void)hipStreamSynchronize(s);
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

}  // namespace lsg
