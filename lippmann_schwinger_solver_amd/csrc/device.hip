// HIP kernels of the DAFMM matvec and the device GMRES, gfx950 (CDNA4).
//
// Matvec = upward passes (LFR_M2M / HFR_M2M, GMRES/G_FMM2DTree.hpp:2472, 3772),
// near field (:2884), downward passes (LFR_M2L / HFR_M2L / HFR_L2L / LFR_L2L,
// :2577, 3882, 4208, 2800) and the two leaf passes, all as destination-major
// gather-GEMV passes over fused operators (nca.cpp builds them).
#include "device.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>

namespace lsg {

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 1024;   // source columns staged in LDS per step (16 KB)
constexpr int kNearLeaves = 16;  // leaves per workgroup in the near-field kernel (rank <= 48)

typedef double dvec2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void cmac(double2& acc, const dvec2 a, const double2 x) {
  acc.x += a.x * x.x - a.y * x.y;
  acc.y += a.x * x.y + a.y * x.x;
}
__device__ __forceinline__ void cmac(double2& acc, const double2 a, const double2 x) {
  acc.x += a.x * x.x - a.y * x.y;
  acc.y += a.x * x.y + a.y * x.x;
}

// One workgroup per destination: Y[dst : dst+R] (=|+=) A (R x C, col-major) * state[src[0..C)].
// Threads are split into (row, phase); phase p walks columns p, p+nph, ... so each
// step of the workgroup reads nph*R consecutive complex entries. The operator is
// streamed once per matvec, so it is read with non-temporal loads.
template <bool ACC>
__global__ __launch_bounds__(kThreads) void pass_kernel(const dvec2* __restrict__ ops, const int* __restrict__ src,
                                                        double2* __restrict__ state, const int* __restrict__ rows,
                                                        const int* __restrict__ ncol, const int* __restrict__ dst,
                                                        const long long* __restrict__ op_off,
                                                        const long long* __restrict__ src_off) {
  __shared__ double2 xs[kChunk];
  __shared__ double2 red[kThreads];
  const int d = blockIdx.x;
  const int R = rows[d], C = ncol[d];
  const dvec2* A = ops + op_off[d];
  const int* S = src + src_off[d];
  double2* Y = state + dst[d];
  const int t = threadIdx.x;
  for (int r0 = 0; r0 < R; r0 += kThreads) {
    const int Rt = min(kThreads, R - r0);
    const int nph = kThreads / Rt;
    const int row = t % Rt, ph = t / Rt;
    const bool active = ph < nph;
    double2 acc = make_double2(0.0, 0.0);
    for (int c0 = 0; c0 < C; c0 += kChunk) {
      const int Cc = min(kChunk, C - c0);
      __syncthreads();
      for (int i = t; i < Cc; i += kThreads) xs[i] = state[S[c0 + i]];
      __syncthreads();
      if (active) {
        const dvec2* Ap = A + (size_t)c0 * R + r0 + row;
        int c = ph;
        for (; c + 7 * nph < Cc; c += 8 * nph) {
          dvec2 a[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) a[u] = __builtin_nontemporal_load(Ap + (size_t)(c + u * nph) * R);
#pragma unroll
          for (int u = 0; u < 8; ++u) cmac(acc, a[u], xs[c + u * nph]);
        }
        for (; c < Cc; c += nph) cmac(acc, __builtin_nontemporal_load(Ap + (size_t)c * R), xs[c]);
      }
    }
    red[t] = acc;
    __syncthreads();
    if (t < Rt) {
      double2 s = red[t];
      for (int p = 1; p < nph; ++p) { s.x += red[p * Rt + t].x; s.y += red[p * Rt + t].y; }
      if (ACC) { s.x += Y[r0 + t].x; s.y += Y[r0 + t].y; }
      Y[r0 + t] = s;
    }
    __syncthreads();
  }
}

// Near field for `lpw` leaves per workgroup: each colleague table T_s
// (rank x rank, column-major) is staged once in LDS and applied to the slot-s
// neighbours of all leaves of the group; y_t = [x_t +] qk2_t * sum_s T_s x_nbr(s).
// Dynamic LDS: rank*rank + lpw*rank complex values; lpw*rank <= 3*kThreads.
__global__ __launch_bounds__(kThreads) void near_kernel(const double2* __restrict__ T, const int* __restrict__ nbr,
                                                        const double* __restrict__ qk2, double2* __restrict__ state,
                                                        int N, int n_leaves, int rank, int lpw, int find_phi) {
  extern __shared__ double2 smem[];
  double2* Ts = smem;
  double2* Xs = smem + rank * rank;
  const int t = threadIdx.x;
  const int leaf0 = blockIdx.x * lpw;
  const int nl = min(lpw, n_leaves - leaf0);
  const int nout = nl * rank;
  const double2* x = state;
  double2* y = state + N;
  double2 acc[3] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  for (int s = 0; s < 9; ++s) {
    __syncthreads();
    for (int i = t; i < rank * rank; i += kThreads) Ts[i] = T[(size_t)s * rank * rank + i];
    for (int i = t; i < nout; i += kThreads) {
      const int l = i / rank, c = i % rank;
      const int nb = nbr[(leaf0 + l) * 9 + s];
      Xs[i] = nb >= 0 ? x[(size_t)nb * rank + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int o = t + q * kThreads;
      if (o < nout) {
        const int l = o / rank, i = o % rank;
        const double2* xv = Xs + l * rank;
        for (int c = 0; c < rank; ++c) cmac(acc[q], Ts[c * rank + i], xv[c]);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const int o = t + q * kThreads;
    if (o < nout) {
      const size_t g = (size_t)leaf0 * rank + o;
      double2 s = acc[q];
      if (find_phi) {
        const double qq = qk2[g];
        s.x = x[g].x + qq * s.x;
        s.y = x[g].y + qq * s.y;
      }
      y[g] = s;
    }
  }
}

// ---------------------------------------------------------------- GMRES ----
constexpr int kDotRows = 2;   // rows per thread in the block dot
constexpr int kDotCols = 8;   // columns reduced together

__device__ __forceinline__ double2 wave_sum(double2 v) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, 64);
    v.y += __shfl_xor(v.y, o, 64);
  }
  return v;
}

// partial[wave*(ncols+1) + i] = sum over the wave's rows of conj(V[n,i]) * w[n]
// for i < ncols, and of |w[n]|^2 for i = ncols. Skipped when *flag == 0.
__global__ __launch_bounds__(kThreads) void dot_partial_kernel(const double2* __restrict__ V, long long ldv, int ncols,
                                                               const double2* __restrict__ w, int N,
                                                               double2* __restrict__ partial,
                                                               const int* __restrict__ flag) {
  if (flag && *flag == 0) return;
  const int t = threadIdx.x, lane = t & 63;
  const long long wave = (long long)blockIdx.x * (kThreads / 64) + (t >> 6);
  const long long base = (long long)blockIdx.x * kThreads * kDotRows;
  const int ncx = ncols + 1;
  double2 wv[kDotRows];
  long long n[kDotRows];
  for (int r = 0; r < kDotRows; ++r) {
    n[r] = base + (long long)r * kThreads + t;
    wv[r] = n[r] < N ? w[n[r]] : make_double2(0.0, 0.0);
  }
  for (int g = 0; g < ncx; g += kDotCols) {
    double2 acc[kDotCols];
#pragma unroll
    for (int u = 0; u < kDotCols; ++u) {
      acc[u] = make_double2(0.0, 0.0);
      const int i = g + u;
      if (i < ncx) {
#pragma unroll
        for (int r = 0; r < kDotRows; ++r) {
          if (n[r] < N) {
            const double2 v = i < ncols ? V[(long long)i * ldv + n[r]] : wv[r];
            acc[u].x += v.x * wv[r].x + v.y * wv[r].y;
            acc[u].y += v.x * wv[r].y - v.y * wv[r].x;
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kDotCols; ++u) {
      const double2 s = wave_sum(acc[u]);
      if (lane == 0 && g + u < ncx) partial[wave * ncx + g + u] = s;
    }
  }
}

// out[i] = sum over parts of partial[p*ncx + i]; one block per column.
__global__ __launch_bounds__(kThreads) void dot_reduce_kernel(const double2* __restrict__ partial, int nparts, int ncx,
                                                              double2* __restrict__ out, const int* __restrict__ flag) {
  if (flag && *flag == 0) return;
  __shared__ double2 wsum[kThreads / 64];
  const int i = blockIdx.x, t = threadIdx.x;
  double2 acc = make_double2(0.0, 0.0);
  for (int b = t; b < nparts; b += kThreads) {
    const double2 p = partial[(long long)b * ncx + i];
    acc.x += p.x; acc.y += p.y;
  }
  acc = wave_sum(acc);
  if ((t & 63) == 0) wsum[t >> 6] = acc;
  __syncthreads();
  if (t == 0) {
    double2 s = wsum[0];
    for (int q = 1; q < kThreads / 64; ++q) { s.x += wsum[q].x; s.y += wsum[q].y; }
    out[i] = s;
  }
}

// w[n] += sgn * sum_i V[n,i] coef[i]; skipped when *flag == 0
__global__ __launch_bounds__(kThreads) void axpy_cols_kernel(const double2* __restrict__ V, long long ldv, int ncols,
                                                             const double2* __restrict__ coef, double sgn,
                                                             double2* __restrict__ w, int N,
                                                             const int* __restrict__ flag) {
  if (flag && *flag == 0) return;
  __shared__ double2 cs[512];
  const int t = threadIdx.x;
  for (int i = t; i < ncols; i += kThreads) cs[i] = coef[i];
  __syncthreads();
  const long long n = (long long)blockIdx.x * kThreads + t;
  if (n >= N) return;
  double2 acc = make_double2(0.0, 0.0);
  for (int i = 0; i < ncols; ++i) cmac(acc, V[(long long)i * ldv + n], cs[i]);
  double2 o = w[n];
  o.x += sgn * acc.x;
  o.y += sgn * acc.y;
  w[n] = o;
}

// out = b - a (residual); out may alias a
__global__ void sub_kernel(const double2* __restrict__ b, const double2* a, double2* out, int N) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) { double2 v = a[n]; out[n] = make_double2(b[n].x - v.x, b[n].y - v.y); }
}

// v = w / s[0] (zero when s[0] vanishes: breakdown, the column is never used)
__global__ void scale_kernel(const double2* __restrict__ w, const double* __restrict__ s, double2* __restrict__ v,
                             int N) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double d = s[0];
  const double inv = d > 1e-300 ? 1.0 / d : 0.0;
  if (n < N) v[n] = make_double2(w[n].x * inv, w[n].y * inv);
}

struct GivensState {
  double2* H;      // (m+1) x m, column-major, rotated in place (R factor)
  double* cs;      // rotation cosines
  double2* sn;     // rotation sines
  double2* g;      // rotated beta e1
  double* hn;      // ||w|| of the current Arnoldi step
  double* beta;    // ||r0||
  int* flag;       // 1: the current step needs a second Gram-Schmidt pass
};

// Cycle start: beta = ||r|| from the fused dot (|w|^2 in column 0), g = beta e1.
__global__ void cycle_init_kernel(const double2* nrm2, GivensState st, int m, double* host_out) {
  const int t = threadIdx.x;
  const double b = sqrt(fmax(nrm2[0].x, 0.0));
  for (int i = t; i <= m; i += blockDim.x) st.g[i] = make_double2(i == 0 ? b : 0.0, 0.0);
  if (t == 0) {
    st.beta[0] = b;
    host_out[0] = b;
  }
}

// Arnoldi step j, after the first (phase 1) or second (phase 2) Gram-Schmidt
// pass. Phase 1: h1 = V^H w and |w|^2 (before projection) are in d1[0..j+1].
// If |w - V h1|^2 = |w|^2 - |h1|^2 > |w|^2 / 2 (DGKS test, eta = 1/sqrt 2),
// the step is complete; otherwise flag = 1 and phase 2 adds h2 from d2 with
// |w'|^2 measured directly. The completed column is rotated with the previous
// Givens rotations, a new rotation zeroes H(j+1,j), and |g(j+1)| (the residual
// estimate) and ||w|| are written to host_out[2j], host_out[2j+1].
__global__ void arnoldi_tail_kernel(const double2* d1, const double2* d2, double2* hsum, GivensState st, int m, int j,
                                    int phase, double* host_out) {
  __shared__ double2 col[512 + 2];
  __shared__ double csv[512];
  __shared__ double2 snv[512];
  __shared__ int done;
  const int t = threadIdx.x;
  if (phase == 2 && st.flag[0] == 0) return;
  if (t == 0) {
    double hn2;
    if (phase == 1) {
      double s = 0.0;
      for (int i = 0; i <= j; ++i) s += d1[i].x * d1[i].x + d1[i].y * d1[i].y;
      const double n1 = d1[j + 1].x;
      hn2 = n1 - s;
      done = (hn2 > 0.5 * n1) ? 1 : 0;
      st.flag[0] = done ? 0 : 1;
    } else {
      double s = 0.0;
      for (int i = 0; i <= j; ++i) s += d2[i].x * d2[i].x + d2[i].y * d2[i].y;
      hn2 = d2[j + 1].x - s;
      done = 1;
    }
    st.hn[0] = sqrt(fmax(hn2, 0.0));
  }
  __syncthreads();
  if (!done) return;
  // H column j = h1 (+ h2), loaded in parallel with the rotations
  for (int i = t; i <= j; i += blockDim.x) {
    double2 v = d1[i];
    if (phase == 2) { v.x += d2[i].x; v.y += d2[i].y; }
    col[i] = v;
    hsum[i] = v;
    if (i < j) { csv[i] = st.cs[i]; snv[i] = st.sn[i]; }
  }
  __syncthreads();
  if (t != 0) return;
  const double hn = st.hn[0];
  col[j + 1] = make_double2(hn, 0.0);
  for (int i = 0; i < j; ++i) {
    const double c = csv[i];
    const double2 s = snv[i];
    const double2 a = col[i], b = col[i + 1];
    // [a; b] <- [c a + s b; -conj(s) a + c b]
    col[i] = make_double2(c * a.x + (s.x * b.x - s.y * b.y), c * a.y + (s.x * b.y + s.y * b.x));
    col[i + 1] = make_double2(c * b.x - (s.x * a.x + s.y * a.y), c * b.y - (s.x * a.y - s.y * a.x));
  }
  const double2 a = col[j];
  const double an = hypot(a.x, a.y);
  double c;
  double2 s;
  if (an == 0.0) {
    c = 0.0;
    s = make_double2(1.0, 0.0);
    col[j] = make_double2(hn, 0.0);
  } else {
    const double r = hypot(an, hn);
    c = an / r;
    const double2 ph = make_double2(a.x / an, a.y / an);
    s = make_double2(ph.x * hn / r, ph.y * hn / r);
    col[j] = make_double2(ph.x * r, ph.y * r);
  }
  col[j + 1] = make_double2(0.0, 0.0);
  double2* Hc = st.H + (long long)j * (m + 1);
  for (int i = 0; i <= j + 1; ++i) Hc[i] = col[i];
  st.cs[j] = c;
  st.sn[j] = s;
  const double2 gj = st.g[j];
  st.g[j] = make_double2(c * gj.x, c * gj.y);
  st.g[j + 1] = make_double2(-(s.x * gj.x + s.y * gj.y), -(s.x * gj.y - s.y * gj.x));
  host_out[2 * j] = hypot(st.g[j + 1].x, st.g[j + 1].y);
  host_out[2 * j + 1] = hn;
}

// y = R^{-1} g (upper triangular k x k of the rotated H)
__global__ void trisolve_kernel(GivensState st, int m, int k, double2* y) {
  if (threadIdx.x != 0) return;
  for (int i = k - 1; i >= 0; --i) {
    double2 s = st.g[i];
    for (int q = i + 1; q < k; ++q) {
      const double2 r = st.H[(long long)q * (m + 1) + i];
      s.x -= r.x * y[q].x - r.y * y[q].y;
      s.y -= r.x * y[q].y + r.y * y[q].x;
    }
    const double2 d = st.H[(long long)i * (m + 1) + i];
    const double den = d.x * d.x + d.y * d.y;
    y[i] = make_double2((s.x * d.x + s.y * d.y) / den, (s.y * d.x - s.x * d.y) / den);
  }
}

#define HIP_OK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return std::string(#expr ": ") + hipGetErrorString(e_);      \
  } while (0)

template <class T>
std::string upload_vec(const std::vector<T>& h, T** d) {
  *d = nullptr;
  if (h.empty()) return "";
  HIP_OK(hipMalloc((void**)d, h.size() * sizeof(T)));
  HIP_OK(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return "";
}

}  // namespace

std::string DeviceOperator::upload(const Plan& plan) {
  N = plan.N;
  rank = plan.rank;
  n_leaves = plan.n_leaves;
  find_phi = plan.find_phi;
  n_state = plan.n_state;
  if (rank > 64) return "nChebNodes^2 > 64 is not supported by the near-field kernel";
  HIP_OK(hipGetDevice(&device));
  HIP_OK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  // the near kernel stages one rank x rank table plus at least one leaf
  if (int64_t(sizeof(double2)) * (rank * rank + rank) > max_lds)
    return "near-field table does not fit in LDS: nChebNodes too large for this device";
  if (int64_t(sizeof(double2)) * (rank * rank + std::min(kNearLeaves, 3 * kThreads / rank) * rank) > 65536)
    HIP_OK(hipFuncSetAttribute((const void*)near_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, max_lds));
  HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIP_OK(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
  HIP_OK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  HIP_OK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  HIP_OK(hipMalloc((void**)&state, size_t(n_state) * sizeof(double2)));
  HIP_OK(hipMemset(state, 0, size_t(n_state) * sizeof(double2)));
  static_assert(sizeof(cplx) == sizeof(double2), "layout");
  std::string e;
  if (!(e = upload_vec(reinterpret_cast<const std::vector<double2>&>(plan.ops), &ops)).empty()) return e;
  if (!(e = upload_vec(plan.src, &src)).empty()) return e;
  std::vector<int> rows, ncol, dst;
  std::vector<long long> opo, srco;
  passes.clear();
  up.clear(); down.clear(); leaf_m2l.clear(); leaf_l2l.clear();
  for (const PassHost& ph : plan.passes) {
    DevPass dp;
    dp.name = ph.name;
    dp.kind = ph.kind;
    dp.ndest = int(ph.rows.size());
    dp.dest_base = int64_t(rows.size());
    dp.entries = ph.entries;
    dp.n_src = ph.n_src;
    dp.n_dst = ph.n_dst;
    rows.insert(rows.end(), ph.rows.begin(), ph.rows.end());
    ncol.insert(ncol.end(), ph.ncol.begin(), ph.ncol.end());
    dst.insert(dst.end(), ph.dst.begin(), ph.dst.end());
    opo.insert(opo.end(), ph.op_off.begin(), ph.op_off.end());
    srco.insert(srco.end(), ph.src_off.begin(), ph.src_off.end());
    const int idx = int(passes.size());
    if (ph.name.rfind("up_", 0) == 0) up.push_back(idx);
    else if (ph.name == "leaf_m2l") leaf_m2l.push_back(idx);
    else if (ph.name == "leaf_l2l") leaf_l2l.push_back(idx);
    else down.push_back(idx);
    passes.push_back(dp);
  }
  if (!(e = upload_vec(rows, &d_rows)).empty()) return e;
  if (!(e = upload_vec(ncol, &d_ncol)).empty()) return e;
  if (!(e = upload_vec(dst, &d_dst)).empty()) return e;
  if (!(e = upload_vec(opo, &d_op)).empty()) return e;
  if (!(e = upload_vec(srco, &d_srcoff)).empty()) return e;
  // near-field tables transposed to column-major per slot
  std::vector<double2> Tc(plan.T.size());
  for (int s = 0; s < 9; ++s)
    for (int i = 0; i < rank; ++i)
      for (int c = 0; c < rank; ++c) {
        const cplx v = plan.T[size_t(s) * rank * rank + size_t(i) * rank + c];
        Tc[size_t(s) * rank * rank + size_t(c) * rank + i] = make_double2(v.real(), v.imag());
      }
  if (!(e = upload_vec(Tc, &T)).empty()) return e;
  if (!(e = upload_vec(plan.nbr, &nbr)).empty()) return e;
  if (!(e = upload_vec(plan.qk2, &qk2)).empty()) return e;
  near_bytes = 16 * int64_t(9) * rank * rank + 16 * 10 * int64_t(N) + 8 * int64_t(N);
  HIP_OK(hipDeviceSynchronize());
  return "";
}

void DeviceOperator::release() {
  if (stream) hipStreamSynchronize(stream);
  if (aux) hipStreamSynchronize(aux);
  void* ptrs[] = {state, ops, src, d_rows, d_ncol, d_dst, d_op, d_srcoff, T, nbr, qk2};
  for (void* p : ptrs)
    if (p) hipFree(p);
  state = ops = nullptr;
  src = d_rows = d_ncol = d_dst = nbr = nullptr;
  d_op = d_srcoff = nullptr;
  T = nullptr;
  qk2 = nullptr;
  if (ev_fork) hipEventDestroy(ev_fork);
  if (ev_join) hipEventDestroy(ev_join);
  ev_fork = ev_join = nullptr;
  if (aux) hipStreamDestroy(aux);
  if (stream) hipStreamDestroy(stream);
  aux = stream = nullptr;
}

static void launch_pass(const DeviceOperator& o, const DevPass& p, hipStream_t s) {
  const int64_t b = p.dest_base;
  const dvec2* ops = reinterpret_cast<const dvec2*>(o.ops);
  if (p.kind == 0)
    hipLaunchKernelGGL(pass_kernel<false>, dim3(p.ndest), dim3(kThreads), 0, s, ops, o.src, o.state,
                       o.d_rows + b, o.d_ncol + b, o.d_dst + b, o.d_op + b, o.d_srcoff + b);
  else
    hipLaunchKernelGGL(pass_kernel<true>, dim3(p.ndest), dim3(kThreads), 0, s, ops, o.src, o.state,
                       o.d_rows + b, o.d_ncol + b, o.d_dst + b, o.d_op + b, o.d_srcoff + b);
}

static void launch_near(const DeviceOperator& o, hipStream_t s) {
  const int fit = int(o.max_lds / int(sizeof(double2))) - o.rank * o.rank;
  const int lpw = std::max(1, std::min({kNearLeaves, 3 * kThreads / o.rank, fit / o.rank}));
  const size_t lds = sizeof(double2) * size_t(o.rank * o.rank + lpw * o.rank);
  hipLaunchKernelGGL(near_kernel, dim3((o.n_leaves + lpw - 1) / lpw), dim3(kThreads), lds, s, o.T, o.nbr, o.qk2,
                     o.state, o.N, o.n_leaves, o.rank, lpw, o.find_phi ? 1 : 0);
}

void DeviceOperator::matvec(const double2* x, double2* y) {
  hipMemcpyAsync(state, x, size_t(N) * sizeof(double2), hipMemcpyDeviceToDevice, stream);
  // the first upward pass is the leaf level (the only one the leaf M2L reads)
  size_t u = 0;
  if (!up.empty()) launch_pass(*this, passes[up[0]], stream), u = 1;
  hipEventRecord(ev_fork, stream);
  hipStreamWaitEvent(aux, ev_fork, 0);
  launch_near(*this, aux);
  for (int p : leaf_m2l) launch_pass(*this, passes[p], aux);
  hipEventRecord(ev_join, aux);
  for (; u < up.size(); ++u) launch_pass(*this, passes[up[u]], stream);
  for (int p : down) launch_pass(*this, passes[p], stream);
  hipStreamWaitEvent(stream, ev_join, 0);
  for (int p : leaf_l2l) launch_pass(*this, passes[p], stream);
  hipMemcpyAsync(y, state + N, size_t(N) * sizeof(double2), hipMemcpyDeviceToDevice, stream);
}

double DeviceOperator::time_matvec(int reps) {
  double2 *x = nullptr, *y = nullptr;
  if (hipMalloc((void**)&x, sizeof(double2) * size_t(N)) != hipSuccess) return -1;
  if (hipMalloc((void**)&y, sizeof(double2) * size_t(N)) != hipSuccess) { hipFree(x); return -1; }
  hipMemcpyAsync(x, state, sizeof(double2) * size_t(N), hipMemcpyDeviceToDevice, stream);
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  matvec(x, y);
  hipEventRecord(a, stream);
  for (int r = 0; r < reps; ++r) matvec(x, y);
  hipEventRecord(b, stream);
  hipEventSynchronize(b);
  float ms = 0;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  hipFree(x);
  hipFree(y);
  return double(ms) / reps;
}

std::string DeviceOperator::profile(int reps, std::vector<double>& ms, std::vector<int64_t>& bytes,
                                    std::vector<std::string>& names) {
  // serialised on the main stream: upward, near field, downward, leaf M2L, leaf L2L
  std::vector<int> order(up.begin(), up.end());
  order.push_back(-1);
  order.insert(order.end(), down.begin(), down.end());
  order.insert(order.end(), leaf_m2l.begin(), leaf_m2l.end());
  order.insert(order.end(), leaf_l2l.begin(), leaf_l2l.end());
  const int np = int(order.size());
  std::vector<hipEvent_t> ev(np + 1);
  for (auto& e : ev) HIP_OK(hipEventCreate(&e));
  ms.assign(np, 0.0);
  bytes.assign(np, 0);
  names.assign(np, "");
  for (int r = 0; r < reps; ++r) {
    HIP_OK(hipEventRecord(ev[0], stream));
    for (int q = 0; q < np; ++q) {
      if (order[q] < 0) launch_near(*this, stream);
      else launch_pass(*this, passes[order[q]], stream);
      HIP_OK(hipEventRecord(ev[q + 1], stream));
    }
    HIP_OK(hipEventSynchronize(ev[np]));
    for (int q = 0; q < np; ++q) {
      float t = 0;
      HIP_OK(hipEventElapsedTime(&t, ev[q], ev[q + 1]));
      ms[q] += t / reps;
    }
  }
  for (int q = 0; q < np; ++q) {
    if (order[q] < 0) { names[q] = "near_field"; bytes[q] = near_bytes; }
    else { names[q] = passes[order[q]].name; bytes[q] = passes[order[q]].bytes(); }
  }
  for (auto& e : ev) hipEventDestroy(e);
  return "";
}

// ------------------------------------------------------------------ GMRES ---
GmresResult device_gmres(DeviceOperator& op, const double2* b_dev, double2* x_dev, int m, int max_restarts,
                         double tol) {
  GmresResult res;
  auto t0 = std::chrono::steady_clock::now();
  const int N = op.N;
  hipStream_t s = op.stream;
  const int nb_dot = (N + kThreads * kDotRows - 1) / (kThreads * kDotRows);
  const int nparts = nb_dot * (kThreads / 64);
  const int nb_vec = (N + kThreads - 1) / kThreads;
  const long long ldv = N;
  double2 *V = nullptr, *w = nullptr, *partial = nullptr, *d1 = nullptr, *d2 = nullptr, *hsum = nullptr,
          *y = nullptr;
  GivensState st{};
  double* hostbuf = nullptr;
  std::vector<void*> owned;
  auto alloc = [&](void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e == hipSuccess) owned.push_back(*p);
    return e == hipSuccess;
  };
  bool ok = alloc((void**)&V, sizeof(double2) * size_t(N) * (m + 1)) && alloc((void**)&w, sizeof(double2) * size_t(N)) &&
            alloc((void**)&partial, sizeof(double2) * size_t(nparts) * (m + 2)) &&
            alloc((void**)&d1, sizeof(double2) * (m + 2)) && alloc((void**)&d2, sizeof(double2) * (m + 2)) &&
            alloc((void**)&hsum, sizeof(double2) * (m + 2)) && alloc((void**)&y, sizeof(double2) * (m + 1)) &&
            alloc((void**)&st.H, sizeof(double2) * size_t(m + 1) * m) && alloc((void**)&st.cs, sizeof(double) * (m + 1)) &&
            alloc((void**)&st.sn, sizeof(double2) * (m + 1)) && alloc((void**)&st.g, sizeof(double2) * (m + 2)) &&
            alloc((void**)&st.hn, sizeof(double)) && alloc((void**)&st.beta, sizeof(double)) &&
            alloc((void**)&st.flag, sizeof(int));
  if (!ok || hipHostMalloc((void**)&hostbuf, sizeof(double) * 2 * (m + 2), hipHostMallocDefault) != hipSuccess) {
    for (void* p : owned) hipFree(p);
    res.status = LSG_ERR_DEVICE;
    res.error = "GMRES workspace allocation failed";
    return res;
  }
  auto dotw = [&](const double2* Vb, int ncols, const double2* vec, double2* out, const int* flag) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nb_dot), dim3(kThreads), 0, s, Vb, ldv, ncols, vec, N, partial, flag);
    hipLaunchKernelGGL(dot_reduce_kernel, dim3(ncols + 1), dim3(kThreads), 0, s, partial, nparts, ncols + 1, out, flag);
  };
  // ||b||
  dotw(V, 0, b_dev, d1, nullptr);
  double2 bn2;
  hipMemcpyAsync(&bn2, d1, sizeof(double2), hipMemcpyDeviceToHost, s);
  hipError_t e = hipStreamSynchronize(s);
  const double bnorm = std::sqrt(bn2.x);
  if (e != hipSuccess) {
    res.status = LSG_ERR_DEVICE;
    res.error = hipGetErrorString(e);
  } else if (!(bnorm > 0) || !std::isfinite(bnorm)) {
    res.status = LSG_ERR_NUMERIC;
    res.error = !std::isfinite(bnorm) ? "right-hand side holds NaN/Inf" : "right-hand side is zero";
  }
  int total_it = 0;
  bool converged = false;
  const bool ref_mode = !(tol > 0);
  // Each cycle opens with the true FP64 residual r = b - A x, which is also the
  // convergence test; the loop runs at most max_restarts Arnoldi cycles and
  // always ends on a residual evaluation.
  for (int cyc = 0; res.status == 0 && cyc <= max_restarts; ++cyc) {
    op.matvec(x_dev, w);
    hipLaunchKernelGGL(sub_kernel, dim3(nb_vec), dim3(kThreads), 0, s, b_dev, w, w, N);
    dotw(V, 0, w, d1, nullptr);
    hipLaunchKernelGGL(cycle_init_kernel, dim3(1), dim3(64), 0, s, d1, st, m, hostbuf);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) { res.status = LSG_ERR_DEVICE; res.error = hipGetErrorString(e); break; }
    const double beta = hostbuf[0];
    res.rel_residual = beta / bnorm;
    if (!std::isfinite(beta)) { res.status = LSG_ERR_NUMERIC; res.error = "residual is NaN/Inf"; break; }
    if (!ref_mode && beta / bnorm <= tol) { converged = true; break; }
    if (cyc == max_restarts || beta == 0.0) break;
    hipLaunchKernelGGL(scale_kernel, dim3(nb_vec), dim3(kThreads), 0, s, w, st.beta, V, N);
    int k = 0;
    bool stop = false;
    double last_est = 1.0;
    for (int j = 0; j < m && !stop;) {
      // enqueue a batch of Arnoldi steps before reading their estimates back;
      // steps past an exit found inside the batch are discarded
      const int batch = std::min(m - j, (!ref_mode && last_est > 100.0 * tol) || ref_mode ? 4 : 1);
      for (int jj = j; jj < j + batch; ++jj) {
        op.matvec(V + (size_t)jj * N, w);
        const int nc = jj + 1;
        dotw(V, nc, w, d1, nullptr);
        hipLaunchKernelGGL(axpy_cols_kernel, dim3(nb_vec), dim3(kThreads), 0, s, V, ldv, nc, d1, -1.0, w, N,
                           (const int*)nullptr);
        hipLaunchKernelGGL(arnoldi_tail_kernel, dim3(1), dim3(kThreads), 0, s, d1, d2, hsum, st, m, jj, 1, hostbuf);
        dotw(V, nc, w, d2, st.flag);
        hipLaunchKernelGGL(axpy_cols_kernel, dim3(nb_vec), dim3(kThreads), 0, s, V, ldv, nc, d2, -1.0, w, N,
                           (const int*)st.flag);
        hipLaunchKernelGGL(arnoldi_tail_kernel, dim3(1), dim3(kThreads), 0, s, d1, d2, hsum, st, m, jj, 2, hostbuf);
        hipLaunchKernelGGL(scale_kernel, dim3(nb_vec), dim3(kThreads), 0, s, w, st.hn, V + (size_t)(jj + 1) * N, N);
      }
      if ((e = hipStreamSynchronize(s)) != hipSuccess) { res.status = LSG_ERR_DEVICE; res.error = hipGetErrorString(e); break; }
      for (int jj = j; jj < j + batch; ++jj) {
        const double est = hostbuf[2 * jj], hn = hostbuf[2 * jj + 1];
        ++total_it;
        k = jj + 1;
        res.est_residual = est / bnorm;
        last_est = est / bnorm;
        if (!std::isfinite(est)) { res.status = LSG_ERR_NUMERIC; res.error = "Arnoldi produced NaN/Inf"; stop = true; break; }
        if (std::fabs(hn) < 1e-8) { stop = true; break; }   // breakdown exit, GMRES/GMRES.cpp:284-288
        if (!ref_mode && est / bnorm <= 0.5 * tol) { stop = true; break; }
      }
      j += batch;
    }
    if (res.status != 0) break;
    // x += V y, y = R^-1 g
    hipLaunchKernelGGL(trisolve_kernel, dim3(1), dim3(64), 0, s, st, m, k, y);
    hipLaunchKernelGGL(axpy_cols_kernel, dim3(nb_vec), dim3(kThreads), 0, s, V, ldv, k, y, 1.0, x_dev, N,
                       (const int*)nullptr);
    res.restarts = cyc + 1;
  }
  if (res.status == 0) {
    converged = !ref_mode ? (res.rel_residual <= tol) : true;
    if (!ref_mode && !converged) { res.status = LSG_ERR_NOCONV; res.error = "GMRES did not reach tol"; }
  }
  res.iterations = total_it;
  res.converged = converged ? 1 : 0;
  hipStreamSynchronize(s);
  for (void* p : owned) hipFree(p);
  hipHostFree(hostbuf);
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

}  // namespace lsg
