// HIP kernels of the DAFMM matvec and the device GMRES, gfx950 (CDNA4).
//
// Matvec = upward passes (LFR_M2M / HFR_M2M, GMRES/G_FMM2DTree.hpp:2472, 3772),
// near field (:2884), downward passes (LFR_M2L / HFR_M2L / HFR_L2L / LFR_L2L,
// :2577, 3882, 4208, 2800) and the leaf far field, all as destination-major
// gather-GEMV passes over fused operators (see nca.cpp).
#include "device.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>

namespace lsg {

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 1024;  // source columns staged in LDS per step (16 KB)

__device__ __forceinline__ void cmac(double2& acc, const double2 a, const double2 x) {
  acc.x += a.x * x.x - a.y * x.y;
  acc.y += a.x * x.y + a.y * x.x;
}

// One workgroup per destination: y[dst : dst+R] (=|+=) A (R x C, col-major) * state[src[0..C)].
// Threads are split into (row, phase); phase p walks columns p, p+nph, ... so a
// workgroup step reads nph*R consecutive complex entries (coalesced stream).
template <bool ACC>
__global__ __launch_bounds__(kThreads) void pass_kernel(const double2* __restrict__ ops, const int* __restrict__ src,
                                                        double2* __restrict__ state, const int* __restrict__ rows,
                                                        const int* __restrict__ ncol, const int* __restrict__ dst,
                                                        const long long* __restrict__ op_off,
                                                        const long long* __restrict__ src_off) {
  __shared__ double2 xs[kChunk];
  __shared__ double2 red[kThreads];
  const int d = blockIdx.x;
  const int R = rows[d], C = ncol[d];
  const double2* A = ops + op_off[d];
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
        const double2* Ap = A + (size_t)c0 * R + r0 + row;
        int c = ph;
#pragma unroll 1
        for (; c + 3 * nph < Cc; c += 4 * nph) {
          const double2 a0 = Ap[(size_t)c * R];
          const double2 a1 = Ap[(size_t)(c + nph) * R];
          const double2 a2 = Ap[(size_t)(c + 2 * nph) * R];
          const double2 a3 = Ap[(size_t)(c + 3 * nph) * R];
          cmac(acc, a0, xs[c]);
          cmac(acc, a1, xs[c + nph]);
          cmac(acc, a2, xs[c + 2 * nph]);
          cmac(acc, a3, xs[c + 3 * nph]);
        }
        for (; c < Cc; c += nph) cmac(acc, Ap[(size_t)c * R], xs[c]);
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

// Near field of one leaf per workgroup: y_t = [x_t +] qk2_t * sum_s T_s x_nbr(s).
// T column-major per slot so that row-threads read consecutive addresses.
__global__ __launch_bounds__(kThreads) void near_kernel(const double2* __restrict__ T, const int* __restrict__ nbr,
                                                        const double* __restrict__ qk2, double2* __restrict__ state,
                                                        int N, int rank, int find_phi) {
  __shared__ double2 xs[9 * 64];
  __shared__ double2 red[kThreads];
  const int leaf = blockIdx.x, t = threadIdx.x;
  const double2* x = state;
  double2* y = state + N;
  for (int i = t; i < 9 * rank; i += kThreads) {
    const int s = i / rank, c = i % rank;
    const int nb = nbr[leaf * 9 + s];
    xs[i] = nb >= 0 ? x[(size_t)nb * rank + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  const int nph = kThreads / rank;
  const int row = t % rank, ph = t / rank;
  double2 acc = make_double2(0.0, 0.0);
  if (ph < nph) {
    for (int s = ph; s < 9; s += nph) {
      if (nbr[leaf * 9 + s] < 0) continue;
      const double2* Ts = T + (size_t)s * rank * rank + row;
      const double2* xv = xs + s * rank;
      for (int c = 0; c < rank; ++c) cmac(acc, Ts[(size_t)c * rank], xv[c]);
    }
  }
  red[t] = acc;
  __syncthreads();
  if (t < rank) {
    double2 s = red[t];
    for (int p = 1; p < nph; ++p) { s.x += red[p * rank + t].x; s.y += red[p * rank + t].y; }
    const size_t g = (size_t)leaf * rank + t;
    if (find_phi) {
      const double q = qk2[g];
      s.x = x[g].x + q * s.x;
      s.y = x[g].y + q * s.y;
    }
    y[g] = s;
  }
}

// ---------------------------------------------------------------- GMRES ----
constexpr int kDotRows = 2;  // rows per thread in the block dot

__device__ __forceinline__ double2 wave_sum(double2 v) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, 64);
    v.y += __shfl_xor(v.y, o, 64);
  }
  return v;
}

// partial[b*ncols + i] = sum over the block's rows of conj(V[n,i]) * w[n]
__global__ __launch_bounds__(kThreads) void dot_partial_kernel(const double2* __restrict__ V, long long ldv, int ncols,
                                                               const double2* __restrict__ w, int N,
                                                               double2* __restrict__ partial) {
  __shared__ double2 wsum[kThreads / 64];
  const int t = threadIdx.x;
  const long long base = (long long)blockIdx.x * kThreads * kDotRows;
  double2 wv[kDotRows];
  long long n[kDotRows];
  for (int r = 0; r < kDotRows; ++r) {
    n[r] = base + (long long)r * kThreads + t;
    wv[r] = n[r] < N ? w[n[r]] : make_double2(0.0, 0.0);
  }
  for (int i = 0; i < ncols; ++i) {
    double2 acc = make_double2(0.0, 0.0);
    for (int r = 0; r < kDotRows; ++r) {
      if (n[r] < N) {
        const double2 v = V[(long long)i * ldv + n[r]];
        acc.x += v.x * wv[r].x + v.y * wv[r].y;
        acc.y += v.x * wv[r].y - v.y * wv[r].x;
      }
    }
    acc = wave_sum(acc);
    if ((t & 63) == 0) wsum[t >> 6] = acc;
    __syncthreads();
    if (t == 0) {
      double2 s = wsum[0];
      for (int q = 1; q < kThreads / 64; ++q) { s.x += wsum[q].x; s.y += wsum[q].y; }
      partial[(long long)blockIdx.x * ncols + i] = s;
    }
    __syncthreads();
  }
}

// out[i] (=|+=) sum_b partial[b*ncols + i]
__global__ __launch_bounds__(kThreads) void dot_reduce_kernel(const double2* __restrict__ partial, int nblocks,
                                                              int ncols, double2* __restrict__ out, int accumulate) {
  __shared__ double2 wsum[kThreads / 64];
  const int i = blockIdx.x, t = threadIdx.x;
  double2 acc = make_double2(0.0, 0.0);
  for (int b = t; b < nblocks; b += kThreads) {
    const double2 p = partial[(long long)b * ncols + i];
    acc.x += p.x; acc.y += p.y;
  }
  acc = wave_sum(acc);
  if ((t & 63) == 0) wsum[t >> 6] = acc;
  __syncthreads();
  if (t == 0) {
    double2 s = wsum[0];
    for (int q = 1; q < kThreads / 64; ++q) { s.x += wsum[q].x; s.y += wsum[q].y; }
    if (accumulate) { s.x += out[i].x; s.y += out[i].y; }
    out[i] = s;
  }
}

// w[n] += sgn * sum_i V[n,i] coef[i]
__global__ __launch_bounds__(kThreads) void axpy_cols_kernel(const double2* __restrict__ V, long long ldv, int ncols,
                                                             const double2* __restrict__ coef, double sgn,
                                                             double2* __restrict__ w, int N) {
  __shared__ double2 cs[512];
  const int t = threadIdx.x;
  for (int i = t; i < ncols && i < 512; i += kThreads) cs[i] = coef[i];
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

// out = b - a (residual), out may alias a
__global__ void sub_kernel(const double2* __restrict__ b, const double2* a, double2* out, int N) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) { double2 v = a[n]; out[n] = make_double2(b[n].x - v.x, b[n].y - v.y); }
}

// v = w / s where s = sqrt(real(nrm2[0])) (or scal[0] if given)
__global__ void scale_kernel(const double2* __restrict__ w, const double* __restrict__ s, double2* __restrict__ v,
                             int N) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double inv = 1.0 / s[0];
  if (n < N) v[n] = make_double2(w[n].x * inv, w[n].y * inv);
}

struct GivensState {
  // H column-major (m+1) x m, rotations, rhs g, misc scalars
  double2* H;
  double* cs;
  double2* sn;
  double2* g;
  double* hn;     // ||w|| of the current step (device)
  double* beta;   // ||r0|| (device)
};

// Cycle start: beta = sqrt(re(nrm2)), g = beta e1, out[0] = beta
__global__ void cycle_init_kernel(const double2* nrm2, GivensState st, int m, double* host_out) {
  if (threadIdx.x != 0) return;
  const double b = sqrt(nrm2[0].x);
  st.beta[0] = b;
  for (int i = 0; i <= m; ++i) st.g[i] = make_double2(0.0, 0.0);
  st.g[0] = make_double2(b, 0.0);
  host_out[0] = b;
}

// Column j: H[0..j, j] = h (already summed), H[j+1, j] = sqrt(re(nrm2)).
// Apply previous rotations, build rotation j, update g, write |g[j+1]| and hn.
__global__ void givens_kernel(const double2* h, const double2* nrm2, GivensState st, int m, int j, double* host_out) {
  if (threadIdx.x != 0) return;
  double2* col = st.H + (long long)j * (m + 1);
  for (int i = 0; i <= j; ++i) col[i] = h[i];
  const double hn = sqrt(fmax(nrm2[0].x, 0.0));
  st.hn[0] = hn;
  col[j + 1] = make_double2(hn, 0.0);
  for (int i = 0; i < j; ++i) {
    const double c = st.cs[i];
    const double2 s = st.sn[i];
    const double2 a = col[i], b = col[i + 1];
    // [a; b] <- [c a + s b; -conj(s) a + c b]
    col[i] = make_double2(c * a.x + (s.x * b.x - s.y * b.y), c * a.y + (s.x * b.y + s.y * b.x));
    col[i + 1] = make_double2(c * b.x - (s.x * a.x + s.y * a.y), c * b.y - (s.x * a.y - s.y * a.x));
  }
  const double2 a = col[j];
  const double bb = hn;
  const double an = hypot(a.x, a.y);
  double c;
  double2 s;
  if (an == 0.0) {
    c = 0.0;
    s = make_double2(1.0, 0.0);
    col[j] = make_double2(bb, 0.0);
  } else {
    const double r = hypot(an, bb);
    c = an / r;
    const double2 ph = make_double2(a.x / an, a.y / an);
    s = make_double2(ph.x * bb / r, ph.y * bb / r);
    col[j] = make_double2(ph.x * r, ph.y * r);
  }
  col[j + 1] = make_double2(0.0, 0.0);
  st.cs[j] = c;
  st.sn[j] = s;
  const double2 gj = st.g[j];
  st.g[j] = make_double2(c * gj.x, c * gj.y);
  // g[j+1] = -conj(s) g_j
  st.g[j + 1] = make_double2(-(s.x * gj.x + s.y * gj.y), -(s.x * gj.y - s.y * gj.x));
  host_out[0] = hypot(st.g[j + 1].x, st.g[j + 1].y);
  host_out[1] = hn;
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
  HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIP_OK(hipMalloc((void**)&state, size_t(n_state) * sizeof(double2)));
  HIP_OK(hipMemset(state, 0, size_t(n_state) * sizeof(double2)));
  static_assert(sizeof(cplx) == sizeof(double2), "layout");
  std::string e;
  if (!(e = upload_vec(reinterpret_cast<const std::vector<double2>&>(plan.ops), &ops)).empty()) return e;
  if (!(e = upload_vec(plan.src, &src)).empty()) return e;
  std::vector<int> rows, ncol, dst;
  std::vector<long long> opo, srco;
  passes.clear();
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
  near_bytes = 16 * int64_t(9) * rank * rank + 16 * 3 * int64_t(N) + 8 * int64_t(N);
  HIP_OK(hipDeviceSynchronize());
  return "";
}

void DeviceOperator::release() {
  if (stream) hipStreamSynchronize(stream);
  void* ptrs[] = {state, ops, src, d_rows, d_ncol, d_dst, d_op, d_srcoff, T, nbr, qk2};
  for (void* p : ptrs)
    if (p) hipFree(p);
  state = ops = nullptr;
  src = d_rows = d_ncol = d_dst = nbr = nullptr;
  d_op = d_srcoff = nullptr;
  T = nullptr;
  qk2 = nullptr;
  if (stream) hipStreamDestroy(stream);
  stream = nullptr;
}

static void launch_pass(const DeviceOperator& o, const DevPass& p) {
  const int64_t b = p.dest_base;
  if (p.kind == 0)
    hipLaunchKernelGGL(pass_kernel<false>, dim3(p.ndest), dim3(kThreads), 0, o.stream, o.ops, o.src, o.state,
                       o.d_rows + b, o.d_ncol + b, o.d_dst + b, o.d_op + b, o.d_srcoff + b);
  else
    hipLaunchKernelGGL(pass_kernel<true>, dim3(p.ndest), dim3(kThreads), 0, o.stream, o.ops, o.src, o.state,
                       o.d_rows + b, o.d_ncol + b, o.d_dst + b, o.d_op + b, o.d_srcoff + b);
}

void DeviceOperator::matvec(const double2* x, double2* y) {
  hipMemcpyAsync(state, x, size_t(N) * sizeof(double2), hipMemcpyDeviceToDevice, stream);
  for (const DevPass& p : passes)
    if (p.kind == 0 && p.name.rfind("up_", 0) == 0) launch_pass(*this, p);
  hipLaunchKernelGGL(near_kernel, dim3(n_leaves), dim3(kThreads), 0, stream, T, nbr, qk2, state, N, rank,
                     find_phi ? 1 : 0);
  for (const DevPass& p : passes)
    if (!(p.kind == 0 && p.name.rfind("up_", 0) == 0)) launch_pass(*this, p);
  hipMemcpyAsync(y, state + N, size_t(N) * sizeof(double2), hipMemcpyDeviceToDevice, stream);
}

std::string DeviceOperator::profile(int reps, std::vector<double>& ms, std::vector<int64_t>& bytes,
                                    std::vector<std::string>& names) {
  // order of execution: upward passes, near field, the rest
  std::vector<int> order;
  for (size_t i = 0; i < passes.size(); ++i)
    if (passes[i].kind == 0 && passes[i].name.rfind("up_", 0) == 0) order.push_back(int(i));
  order.push_back(-1);
  for (size_t i = 0; i < passes.size(); ++i)
    if (!(passes[i].kind == 0 && passes[i].name.rfind("up_", 0) == 0)) order.push_back(int(i));
  const int np = int(order.size());
  std::vector<hipEvent_t> ev(np + 1);
  for (auto& e : ev) HIP_OK(hipEventCreate(&e));
  ms.assign(np, 0.0);
  bytes.assign(np, 0);
  names.assign(np, "");
  for (int r = 0; r < reps; ++r) {
    HIP_OK(hipEventRecord(ev[0], stream));
    for (int q = 0; q < np; ++q) {
      if (order[q] < 0)
        hipLaunchKernelGGL(near_kernel, dim3(n_leaves), dim3(kThreads), 0, stream, T, nbr, qk2, state, N, rank,
                           find_phi ? 1 : 0);
      else
        launch_pass(*this, passes[order[q]]);
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
  const int nb_vec = (N + kThreads - 1) / kThreads;
  double2 *V = nullptr, *w = nullptr, *partial = nullptr, *h = nullptr, *h2 = nullptr, *nrm = nullptr, *y = nullptr;
  GivensState st{};
  double* hostbuf = nullptr;
  auto fail = [&](const char* what, hipError_t e) {
    res.status = LSG_ERR_DEVICE;
    res.error = std::string(what) + ": " + hipGetErrorString(e);
  };
  hipError_t e;
  const long long ldv = N;
  if ((e = hipMalloc((void**)&V, sizeof(double2) * size_t(N) * (m + 1))) != hipSuccess) { fail("V", e); return res; }
  hipMalloc((void**)&w, sizeof(double2) * size_t(N));
  hipMalloc((void**)&partial, sizeof(double2) * size_t(nb_dot) * (m + 1));
  hipMalloc((void**)&h, sizeof(double2) * (m + 2));
  hipMalloc((void**)&h2, sizeof(double2) * (m + 2));
  hipMalloc((void**)&nrm, sizeof(double2) * 2);
  hipMalloc((void**)&y, sizeof(double2) * (m + 1));
  hipMalloc((void**)&st.H, sizeof(double2) * size_t(m + 1) * m);
  hipMalloc((void**)&st.cs, sizeof(double) * (m + 1));
  hipMalloc((void**)&st.sn, sizeof(double2) * (m + 1));
  hipMalloc((void**)&st.g, sizeof(double2) * (m + 2));
  hipMalloc((void**)&st.hn, sizeof(double));
  hipMalloc((void**)&st.beta, sizeof(double));
  if ((e = hipHostMalloc((void**)&hostbuf, sizeof(double) * 4, hipHostMallocDefault)) != hipSuccess) {
    fail("hostbuf", e);
    return res;
  }
  auto norm2 = [&](const double2* v, double2* out) {
    hipLaunchKernelGGL(dot_partial_kernel, dim3(nb_dot), dim3(kThreads), 0, s, v, ldv, 1, v, N, partial);
    hipLaunchKernelGGL(dot_reduce_kernel, dim3(1), dim3(kThreads), 0, s, partial, nb_dot, 1, out, 0);
  };
  // ||b||
  norm2(b_dev, nrm);
  double2 bn2;
  hipMemcpyAsync(&bn2, nrm, sizeof(double2), hipMemcpyDeviceToHost, s);
  hipStreamSynchronize(s);
  const double bnorm = std::sqrt(bn2.x);
  if (!(bnorm > 0) || !std::isfinite(bnorm)) {
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
    norm2(w, nrm);
    hipLaunchKernelGGL(cycle_init_kernel, dim3(1), dim3(64), 0, s, nrm, st, m, hostbuf);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) { fail("cycle", e); break; }
    const double beta = hostbuf[0];
    res.rel_residual = beta / bnorm;
    if (!std::isfinite(beta)) { res.status = LSG_ERR_NUMERIC; res.error = "residual is NaN/Inf"; break; }
    if (!ref_mode && beta / bnorm <= tol) { converged = true; break; }
    if (cyc == max_restarts || beta == 0.0) break;
    hipLaunchKernelGGL(scale_kernel, dim3(nb_vec), dim3(kThreads), 0, s, w, st.beta, V, N);
    int k = 0;
    for (int j = 0; j < m; ++j) {
      double2* vj = V + (size_t)j * N;
      op.matvec(vj, w);
      const int nc = j + 1;
      // classical Gram-Schmidt, twice
      hipLaunchKernelGGL(dot_partial_kernel, dim3(nb_dot), dim3(kThreads), 0, s, V, ldv, nc, w, N, partial);
      hipLaunchKernelGGL(dot_reduce_kernel, dim3(nc), dim3(kThreads), 0, s, partial, nb_dot, nc, h, 0);
      hipLaunchKernelGGL(axpy_cols_kernel, dim3(nb_vec), dim3(kThreads), 0, s, V, ldv, nc, h, -1.0, w, N);
      hipLaunchKernelGGL(dot_partial_kernel, dim3(nb_dot), dim3(kThreads), 0, s, V, ldv, nc, w, N, partial);
      hipLaunchKernelGGL(dot_reduce_kernel, dim3(nc), dim3(kThreads), 0, s, partial, nb_dot, nc, h2, 0);
      hipLaunchKernelGGL(axpy_cols_kernel, dim3(nb_vec), dim3(kThreads), 0, s, V, ldv, nc, h2, -1.0, w, N);
      hipLaunchKernelGGL(dot_reduce_kernel, dim3(nc), dim3(kThreads), 0, s, partial, nb_dot, nc, h, 1);
      norm2(w, nrm);
      hipLaunchKernelGGL(givens_kernel, dim3(1), dim3(64), 0, s, h, nrm, st, m, j, hostbuf);
      if ((e = hipStreamSynchronize(s)) != hipSuccess) { fail("arnoldi", e); break; }
      ++total_it;
      k = j + 1;
      const double est = hostbuf[0], hn = hostbuf[1];
      res.est_residual = est / bnorm;
      if (!std::isfinite(est)) { res.status = LSG_ERR_NUMERIC; res.error = "Arnoldi produced NaN/Inf"; break; }
      // breakdown exit as GMRES/GMRES.cpp:284-288
      if (std::fabs(hn) < 1e-8) break;
      hipLaunchKernelGGL(scale_kernel, dim3(nb_vec), dim3(kThreads), 0, s, w, st.hn, V + (size_t)(j + 1) * N, N);
      if (!ref_mode && est / bnorm <= 0.5 * tol) break;
    }
    if (res.status != 0) break;
    // x += V y, y = R^-1 g
    hipLaunchKernelGGL(trisolve_kernel, dim3(1), dim3(64), 0, s, st, m, k, y);
    hipLaunchKernelGGL(axpy_cols_kernel, dim3(nb_vec), dim3(kThreads), 0, s, V, ldv, k, y, 1.0, x_dev, N);
    res.restarts = cyc + 1;
  }
  if (res.status == 0) {
    converged = !ref_mode ? (res.rel_residual <= tol) : true;
    if (!ref_mode && !converged) { res.status = LSG_ERR_NOCONV; res.error = "GMRES did not reach tol"; }
  }
  res.iterations = total_it;
  res.converged = converged ? 1 : 0;
  hipFree(V); hipFree(w); hipFree(partial); hipFree(h); hipFree(h2); hipFree(nrm); hipFree(y);
  hipFree(st.H); hipFree(st.cs); hipFree(st.sn); hipFree(st.g); hipFree(st.hn); hipFree(st.beta);
  hipHostFree(hostbuf);
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

}  // namespace lsg
