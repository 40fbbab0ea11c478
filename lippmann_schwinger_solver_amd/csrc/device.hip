// HIP kernels of the DAFMM matvec and the device GMRES, gfx950 (CDNA4).
//
// Matvec = upward passes (LFR_M2M / HFR_M2M, GMRES/G_FMM2DTree.hpp:2472, 3772),
// near field (:2884), downward passes (LFR_M2L / HFR_M2L / HFR_L2L / LFR_L2L,
// :2577, 3882, 4208, 2800) and the two leaf passes, all as destination-major
// gather-GEMV passes over fused operators (nca.cpp builds them).
#include "device.hpp"

#include <algorithm>
#include <functional>
#include <chrono>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <type_traits>

namespace lsg {

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 1024;    // source columns staged in LDS per step (16 KB)
constexpr int kMaxRestart = 500;
constexpr int kWaveEntries = 4096;  // passes with fewer entries per destination use one wave per destination
constexpr int kWaveCols = 16;       // columns per lane per step in the one-wave pass kernel

template <class R>
struct cx;
template <>
struct cx<double> {
  typedef double2 t;
  typedef double v __attribute__((ext_vector_type(2)));
};
template <>
struct cx<float> {
  typedef float2 t;
  typedef float v __attribute__((ext_vector_type(2)));
};

// acc += a * x as four fused multiply-adds
template <class Acc, class A, class X>
__device__ __forceinline__ void cmac(Acc& acc, const A a, const X x) {
  acc.x = fma(a.x, x.x, acc.x);
  acc.x = fma(-a.y, x.y, acc.x);
  acc.y = fma(a.x, x.y, acc.y);
  acc.y = fma(a.y, x.x, acc.y);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Source value of a pass: state indices below N are the input vector x (FP64,
// read in place and rounded to the pass precision; the state's x region is not
// filled).
template <class C, class XC>
__device__ __forceinline__ C load_src(const XC* __restrict__ x, const C* __restrict__ state, int i, int N) {
  C v;
  if (i < N) {
    const XC a = x[i];
    v.x = a.x;
    v.y = a.y;
  } else {
    v = state[i];
  }
  return v;
}

// Write (or add) s to element r of destination d: [N, 2N) is the output
// vector y (FP64, written in place), everything else the pass-precision state.
template <bool ACC, class C, class XC>
__device__ __forceinline__ void put_dst(C* state, XC* y, int d, int N, int r, C s) {
  if (d >= N && d < 2 * N) {
    XC* Y = y + (d - N) + r;
    XC o;
    if (ACC) {
      o = *Y;
      o.x += s.x;
      o.y += s.y;
    } else {
      o.x = s.x;
      o.y = s.y;
    }
    *Y = o;
  } else {
    C* Y = state + d + r;
    if (ACC) {
      s.x += Y->x;
      s.y += Y->y;
    }
    *Y = s;
  }
}

// Sum over the nph phases of red[ph * rw + row] into red[row]: a halving tree
// when there are many phases (few rows), else left to the caller's serial
// loop. nph and rw are the same for the whole workgroup, so every thread
// reaches the same barriers. Returns whether the tree ran.
template <class C>
__device__ __forceinline__ bool phase_tree(C* red, int nph, int rw, int ph, int row) {
  if (nph <= 8) return false;
  for (int n = nph; n > 1;) {
    const int h = (n + 1) / 2;
    if (ph < n - h) {
      C& a = red[ph * rw + row];
      const C b = red[(ph + h) * rw + row];
      a.x += b.x;
      a.y += b.y;
    }
    __syncthreads();
    n = h;
  }
  return true;
}

// Destinations of inactive systems skipped (a fused plan of several systems):
// destination d belongs to system dsys[d], active while act[system] != 0.
// dsys == nullptr: every destination runs.
struct DestMask {
  const int* dsys;
  const int* act;
  __device__ __forceinline__ bool skip(int d) const { return dsys && !act[dsys[d]]; }
};

// One workgroup per destination: Y[0:R] (=|+=) A (R x C, col-major) * src[0..C).
// Threads are split into (row, phase); phase p walks columns p, p+nph, ... so each
// step of the workgroup reads nph*R consecutive complex entries. The operator is
// streamed once per matvec, so it is read with non-temporal loads. ROLE only
// names the leaf M2L instantiation apart in kernel traces. Workgroups walk the
// destinations with a stride of gridDim.x (one destination each when the grid
// covers them all).
template <bool ACC, class R, int ROLE>
__global__ __launch_bounds__(kThreads) void pass_block_kernel(
    const typename cx<R>::v* __restrict__ ops, const int* __restrict__ src, typename cx<R>::t* state,
    const double2* __restrict__ x_arg, double2* y_arg, int N, const int* __restrict__ rows,
    const int* __restrict__ ncol, const int* __restrict__ dst, const long long* __restrict__ op_off,
    const long long* __restrict__ src_off, int ndest, const double2* const* __restrict__ xyp, DestMask mk) {
  // pointers from device memory when the matvec runs as a captured graph
  const double2* __restrict__ x = xyp ? xyp[0] : x_arg;
  double2* y = xyp ? const_cast<double2*>(xyp[1]) : y_arg;
  typedef typename cx<R>::t C;
  typedef typename cx<R>::v V;
  __shared__ C xs[kChunk];
  __shared__ C red[kThreads];
  for (int d = blockIdx.x; d < ndest; d += gridDim.x) {
  if (mk.skip(d)) continue;  // uniform over the workgroup
  const int Rw = rows[d], Cn = ncol[d];
  const V* A = ops + op_off[d];
  const int* S = src + src_off[d];
  const int dd = dst[d];
  const int t = threadIdx.x;
  for (int r0 = 0; r0 < Rw; r0 += kThreads) {
    const int Rt = min(kThreads, Rw - r0);
    const int nph = kThreads / Rt;
    const int row = t % Rt, ph = t / Rt;
    const bool active = ph < nph;
    C acc;
    acc.x = 0;
    acc.y = 0;
    for (int c0 = 0; c0 < Cn; c0 += kChunk) {
      const int Cc = min(kChunk, Cn - c0);
      const V* Ap = A + (size_t)c0 * Rw + r0 + row;
      // the first 8 operator columns of this thread are requested before the
      // source gather, so the operator stream does not wait for it
      V a[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int c = ph + u * nph;
        if (active && c < Cc) a[u] = __builtin_nontemporal_load(Ap + (size_t)c * Rw);
        else a[u] = V{0, 0};
      }
      __syncthreads();
      for (int i = t; i < Cc; i += kThreads) xs[i] = load_src(x, state, S[c0 + i], N);
      __syncthreads();
      if (active) {
        int c = ph;
        for (; c + 7 * nph < Cc; c += 8 * nph) {
          // next group in flight while this one is multiplied
          V nx[8];
          const int cn = c + 8 * nph;
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int cc = cn + u * nph;
            if (cc < Cc) nx[u] = __builtin_nontemporal_load(Ap + (size_t)cc * Rw);
            else nx[u] = V{0, 0};
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) cmac(acc, a[u], xs[c + u * nph]);
#pragma unroll
          for (int u = 0; u < 8; ++u) a[u] = nx[u];
        }
        // tail: fewer than 8 columns left, already in a[]
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (c + u * nph < Cc) cmac(acc, a[u], xs[c + u * nph]);
      }
    }
    red[t] = acc;
    __syncthreads();
    const bool tree = phase_tree(red, nph, Rt, ph, row);
    if (t < Rt) {
      C s = red[t];
      if (!tree)
        for (int p = 1; p < nph; ++p) {
          s.x += red[p * Rt + t].x;
          s.y += red[p * Rt + t].y;
        }
      put_dst<ACC>(state, y, dd, N, r0 + t, s);
    }
    __syncthreads();
  }
  }
}

// Destination d of a pass computed by one wavefront: lanes split into (row,
// phase), phase p walks columns p, p + nph, ... WC at a time (indices, then
// operator entries, then sources in flight together); the phases are summed
// through the wave's own LDS slice rw (64 entries), so no workgroup barrier
// is needed.
template <bool ACC, class R, int WC>
__device__ __forceinline__ void wave_dest(const typename cx<R>::v* __restrict__ ops, const int* __restrict__ src,
                                          typename cx<R>::t* state, const double2* __restrict__ x, double2* y, int N,
                                          const int* __restrict__ rows, const int* __restrict__ ncol,
                                          const int* __restrict__ dst, const long long* __restrict__ op_off,
                                          const long long* __restrict__ src_off, int d, typename cx<R>::t* rw,
                                          int lane) {
  typedef typename cx<R>::t C;
  typedef typename cx<R>::v V;
  const int Rw = rows[d], Cn = ncol[d];
  const V* A = ops + op_off[d];
  const int* S = src + src_off[d];
  const int dd = dst[d];
  for (int r0 = 0; r0 < Rw; r0 += 64) {
    const int Rt = min(64, Rw - r0);
    const int nph = 64 / Rt;
    const int row = lane % Rt, ph = lane / Rt;
    C acc;
    acc.x = 0;
    acc.y = 0;
    if (ph < nph) {
      const V* Ap = A + r0 + row;
      for (int c = ph; c < Cn; c += WC * nph) {
        int si[WC];
        V a[WC];
#pragma unroll
        for (int u = 0; u < WC; ++u) {
          const int cc = c + u * nph;
          si[u] = cc < Cn ? S[cc] : -1;
        }
#pragma unroll
        for (int u = 0; u < WC; ++u) {
          const int cc = c + u * nph;
          if (cc < Cn) a[u] = __builtin_nontemporal_load(Ap + (size_t)cc * Rw);
          else a[u] = V{0, 0};
        }
#pragma unroll
        for (int u = 0; u < WC; ++u) {
          C xv;
          if (si[u] >= 0) xv = load_src(x, state, si[u], N);
          else xv.x = xv.y = 0;
          cmac(acc, a[u], xv);
        }
      }
    }
    rw[lane] = acc;
    wave_sync();
    if (lane < Rt) {
      C s = rw[lane];
      for (int p = 1; p < nph; ++p) {
        s.x += rw[p * Rt + lane].x;
        s.y += rw[p * Rt + lane].y;
      }
      put_dst<ACC>(state, y, dd, N, r0 + lane, s);
    }
    wave_sync();
  }
}

// One wavefront per destination (four per workgroup) for passes of small
// operators (wave_dest); waves retire independently.
template <bool ACC, class R>
__global__ __launch_bounds__(kThreads) void pass_wave_kernel(
    const typename cx<R>::v* __restrict__ ops, const int* __restrict__ src, typename cx<R>::t* state,
    const double2* __restrict__ x_arg, double2* y_arg, int N, const int* __restrict__ rows,
    const int* __restrict__ ncol, const int* __restrict__ dst, const long long* __restrict__ op_off,
    const long long* __restrict__ src_off, int ndest, const double2* const* __restrict__ xyp, DestMask mk) {
  // pointers from device memory when the matvec runs as a captured graph
  const double2* __restrict__ x = xyp ? xyp[0] : x_arg;
  double2* y = xyp ? const_cast<double2*>(xyp[1]) : y_arg;
  typedef typename cx<R>::t C;
  __shared__ C red[kThreads];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int d = blockIdx.x * (kThreads / 64) + wv;
  if (d >= ndest || mk.skip(d)) return;
  wave_dest<ACC, R, kWaveCols>(ops, src, state, x, y, N, rows, ncol, dst, op_off, src_off, d, red + wv * 64, lane);
}

// Upward passes fused per subtree. Workgroup t runs the destinations of task
// t stage by stage (stage s: its destinations of the s-th fused upward pass,
// fdest[task_stage[t (nstage + 1) + s] ...]), one wavefront per destination
// round robin over the 16 wavefronts, with a workgroup barrier between
// stages. Every source a destination reads is x or a charge written by an
// earlier stage of the same task (checked when the tasks are built), so no
// other workgroup is waited for; the barrier makes the charges written by one
// wavefront visible to the others.
constexpr int kFusedThreads = 1024;
template <class R>
__global__ __launch_bounds__(kFusedThreads) void pass_fused_kernel(
    const typename cx<R>::v* __restrict__ ops, const int* __restrict__ src, typename cx<R>::t* state,
    const double2* __restrict__ x_arg, double2* y_arg, int N, const int* __restrict__ rows,
    const int* __restrict__ ncol, const int* __restrict__ dst, const long long* __restrict__ op_off,
    const long long* __restrict__ src_off, const int* __restrict__ task_stage, const int* __restrict__ fdest,
    int nstage, const double2* const* __restrict__ xyp, DestMask mk) {
  const double2* __restrict__ x = xyp ? xyp[0] : x_arg;
  double2* y = xyp ? const_cast<double2*>(xyp[1]) : y_arg;
  typedef typename cx<R>::t C;
  __shared__ C red[kFusedThreads];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int* ts = task_stage + (size_t)blockIdx.x * (nstage + 1);
  for (int s = 0; s < nstage; ++s) {
    const int e = ts[s + 1];
    for (int i = ts[s] + wv; i < e; i += kFusedThreads / 64) {
      const int d = fdest[i];
      if (!mk.skip(d))
        wave_dest<false, R, 8>(ops, src, state, x, y, N, rows, ncol, dst, op_off, src_off, d, red + wv * 64, lane);
    }
    __syncthreads();
  }
}

// One lane per output row, for passes of short rows: lane k takes row
// rowmap[k] & 255 of destination rowmap[k] >> 8 and walks all its columns
// alone (8 columns in flight: indices, then operator entries, then sources),
// so there is no phase reduction and no barrier; lanes of one destination
// read each operator column as one contiguous run and share its source loads.
// The row map orders destinations by decreasing column count, so the lanes of
// a wavefront run loops of similar length.
template <bool ACC, class R>
__global__ __launch_bounds__(kThreads) void pass_lane_kernel(
    const typename cx<R>::v* __restrict__ ops, const int* __restrict__ src, typename cx<R>::t* state,
    const double2* __restrict__ x_arg, double2* y_arg, int N, const int* __restrict__ rows,
    const int* __restrict__ ncol, const int* __restrict__ dst, const long long* __restrict__ op_off,
    const long long* __restrict__ src_off, const unsigned* __restrict__ rowmap, int nrows,
    const double2* const* __restrict__ xyp, DestMask mk) {
  const double2* __restrict__ x = xyp ? xyp[0] : x_arg;
  double2* y = xyp ? const_cast<double2*>(xyp[1]) : y_arg;
  typedef typename cx<R>::t C;
  typedef typename cx<R>::v V;
  const int k = blockIdx.x * kThreads + threadIdx.x;
  if (k >= nrows) return;
  const unsigned pm = rowmap[k];
  const int d = int(pm >> 8), r = int(pm & 255u);
  if (mk.skip(d)) return;
  const int Rw = rows[d], Cn = ncol[d];
  const V* A = ops + op_off[d] + r;
  const int* S = src + src_off[d];
  C acc;
  acc.x = 0;
  acc.y = 0;
  int c = 0;
  for (; c + 8 <= Cn; c += 8) {
    int si[8];
    V a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) si[u] = S[c + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = __builtin_nontemporal_load(A + (size_t)(c + u) * Rw);
#pragma unroll
    for (int u = 0; u < 8; ++u) cmac(acc, a[u], load_src(x, state, si[u], N));
  }
  for (; c < Cn; ++c) cmac(acc, __builtin_nontemporal_load(A + (size_t)c * Rw), load_src(x, state, S[c], N));
  put_dst<ACC>(state, y, dst[d], N, r, acc);
}

// Near field of a uniform tree as one GEMM over the 9 colleague slots:
// y_l = [x_l +] qk2_l * sum_s T_s x_nbr(l,s). A workgroup owns LB = NL*(256/RK)
// leaves; thread (row i, group p) accumulates row i of leaves NL p .. NL p + NL-1,
// so each table entry read from LDS feeds NL complex multiply-adds (NL = 1 or
// 2, chosen per operator by the autotuner). The next
// slot's table and neighbour vectors are fetched into registers while the
// current slot is being applied.
template <class R, int RK, int NL>
__global__ __launch_bounds__(kThreads) void near_kernel(const typename cx<R>::t* __restrict__ T,
                                                        const int* __restrict__ nbr, const double* __restrict__ qk2,
                                                        const double2* __restrict__ x_arg, double2* __restrict__ y_arg,
                                                        int n_leaves, int find_phi, const double2* const* __restrict__ xyp) {
  const double2* __restrict__ x = xyp ? xyp[0] : x_arg;
  double2* __restrict__ y = xyp ? const_cast<double2*>(xyp[1]) : y_arg;
  typedef typename cx<R>::t C;
  constexpr int LP = kThreads / RK;   // leaf groups per workgroup
  constexpr int LB = NL * LP;         // leaves per workgroup (NL per thread)
  constexpr int NT = (RK * RK + kThreads - 1) / kThreads;
  constexpr int NX = (LB * RK + kThreads - 1) / kThreads;
  __shared__ C Ts[RK * RK];
  __shared__ C Xs[LB * RK];
  const int t = threadIdx.x;
  const int i = t % RK, p = t / RK;
  const int leaf0 = blockIdx.x * LB;
  // registers holding the next slot's table entries and neighbour values
  C tr[NT], xr[NX];
#define LSG_NEAR_FETCH(slot)                                                  \
  {                                                                           \
    _Pragma("unroll") for (int k = 0; k < NT; ++k) {                          \
      const int idx = t + k * kThreads;                                       \
      tr[k] = idx < RK * RK ? T[(slot) * RK * RK + idx] : C{};                \
    }                                                                         \
    _Pragma("unroll") for (int k = 0; k < NX; ++k) {                          \
      const int idx = t + k * kThreads;                                       \
      const int l = idx / RK, c = idx - l * RK;                               \
      const int leaf = leaf0 + l;                                             \
      const int nb = (idx < LB * RK && leaf < n_leaves) ? nbr[leaf * 9 + (slot)] : -1; \
      if (nb >= 0) {                                                          \
        const double2 xv = x[(size_t)nb * RK + c];                            \
        xr[k].x = xv.x;                                                       \
        xr[k].y = xv.y;                                                       \
      } else {                                                                \
        xr[k] = C{};                                                          \
      }                                                                       \
    }                                                                         \
  }
  LSG_NEAR_FETCH(0)
  C a[NL];
#pragma unroll
  for (int u = 0; u < NL; ++u) a[u].x = a[u].y = 0;
  for (int s = 0; s < 9; ++s) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      const int idx = t + k * kThreads;
      if (idx < RK * RK) Ts[idx] = tr[k];
    }
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      const int idx = t + k * kThreads;
      if (idx < LB * RK) Xs[idx] = xr[k];
    }
    __syncthreads();
    if (s + 1 < 9) LSG_NEAR_FETCH(s + 1)
    if (p < LP) {
      const C* xa = Xs + NL * p * RK;
#pragma unroll 6
      for (int c = 0; c < RK; ++c) {
        const C tv = Ts[c * RK + i];
#pragma unroll
        for (int u = 0; u < NL; ++u) cmac(a[u], tv, xa[u * RK + c]);
      }
    }
  }
#undef LSG_NEAR_FETCH
  if (p < LP) {
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int leaf = leaf0 + NL * p + u;
      if (leaf < n_leaves) {
        const size_t g = (size_t)leaf * RK + i;
        double2 s = make_double2(a[u].x, a[u].y);
        if (find_phi) {
          const double qq = qk2[g];
          const double2 xv = x[g];
          s.x = xv.x + qq * s.x;
          s.y = xv.y + qq * s.y;
        }
        y[g] = s;
      }
    }
  }
}

// ---------------------------------------------------------------- GMRES ----
// Classical Gram-Schmidt with one full reorthogonalisation (CGS2) per Arnoldi
// step, in five launches:
//   gs_dot      partials of V^H w                  (reads V once)
//   gs_reduce   h1 = sum of the partials
//   gs_proj_dot w' = w - V h1, partials of V^H w' and |w'|^2
//   gs_tail     h2, h = h1 + h2, ||w''||^2 = |w'|^2 - |h2|^2, Givens update
//   gs_update   v_{j+1} = (w' - V h2) / ||w''||
constexpr int kGsRpt = 2;                  // rows per thread in the dot kernels
constexpr int kGsRows = kGsRpt * kThreads;  // rows per workgroup in the dot kernels
constexpr int kGsCols = 8;             // columns per workgroup in gs_dot

__device__ __forceinline__ double2 wave_sum(double2 v) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_xor(v.x, o, 64);
    v.y += __shfl_xor(v.y, o, 64);
  }
  return v;
}

// acc += conj(v) * w
__device__ __forceinline__ void cdot(double2& acc, const double2 v, const double2 w) {
  acc.x = fma(v.x, w.x, acc.x);
  acc.x = fma(v.y, w.y, acc.x);
  acc.y = fma(v.x, w.y, acc.y);
  acc.y = fma(-v.y, w.x, acc.y);
}

// Sum of four per-wave values of column u (ws[wave][u]) into partial[blk*ncx + g0 + u].
__device__ __forceinline__ void store_block_partials(double2 (*ws)[kGsCols], const double2 (&acc)[kGsCols], int g0,
                                                     int ncx, double2* partial, long long blk) {
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
#pragma unroll
  for (int u = 0; u < kGsCols; ++u) {
    const double2 s = wave_sum(acc[u]);
    if (lane == 0) ws[wv][u] = s;
  }
  __syncthreads();
  if (t < kGsCols && g0 + t < ncx) {
    double2 s = ws[0][t];
    for (int q = 1; q < kThreads / 64; ++q) {
      s.x += ws[q][t].x;
      s.y += ws[q][t].y;
    }
    partial[blk * ncx + g0 + t] = s;
  }
  __syncthreads();
}

// partial[blockIdx.x*ncx + col] for col in [blockIdx.y*8, +8): conj(V[:,col]).w over
// the workgroup's kGsRows rows for col < nc, |w|^2 for col == nc (when ncx = nc+1).
__global__ __launch_bounds__(kThreads) void gs_dot_kernel(const double2* __restrict__ V, long long ldv, int nc,
                                                          int ncx, const double2* __restrict__ w, int N,
                                                          double2* __restrict__ partial) {
  __shared__ double2 ws[kThreads / 64][kGsCols];
  const int t = threadIdx.x;
  const int g0 = blockIdx.y * kGsCols;
  const double2 z = make_double2(0.0, 0.0);
  long long n[kGsRpt];
  double2 wv[kGsRpt];
#pragma unroll
  for (int r = 0; r < kGsRpt; ++r) {
    n[r] = (long long)blockIdx.x * kGsRows + t + r * kThreads;
    wv[r] = n[r] < N ? w[n[r]] : z;
  }
  double2 acc[kGsCols];
#pragma unroll
  for (int u = 0; u < kGsCols; ++u) {
    const int col = g0 + u;
    acc[u] = z;
#pragma unroll
    for (int r = 0; r < kGsRpt; ++r) {
      double2 v = z;
      if (col < nc) {
        if (n[r] < N) v = V[col * ldv + n[r]];
      } else if (col == nc) {
        v = wv[r];
      }
      cdot(acc[u], v, wv[r]);
    }
  }
  store_block_partials(ws, acc, g0, ncx, partial, blockIdx.x);
}

// out[c] = sum_p partial[p*ncx + c] for c < ncx, by one workgroup; also kept in lds_out.
__device__ void block_reduce_cols(const double2* __restrict__ partial, int nparts, int ncx, double2* out,
                                  double2* lds_out, double2* red) {
  const int t = threadIdx.x;
  for (int c0 = 0; c0 < ncx; c0 += kThreads) {
    const int ncc = min(kThreads, ncx - c0);
    const int S = kThreads / ncc;
    const int c = t % ncc, sg = t / ncc;
    double2 a = make_double2(0.0, 0.0);
    if (sg < S) {
#pragma unroll 4
      for (int q = sg; q < nparts; q += S) {
        const double2 v = partial[(long long)q * ncx + c0 + c];
        a.x += v.x;
        a.y += v.y;
      }
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
      if (lds_out) lds_out[c0 + t] = s;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kThreads) void gs_reduce_kernel(const double2* __restrict__ partial, int nparts,
                                                             int ncx, double2* __restrict__ out) {
  __shared__ double2 red[kThreads];
  block_reduce_cols(partial, nparts, ncx, out, nullptr, red);
}

// w <- w - V h (h = first nc entries of hg); then partials of conj(V[:,c]).w for c < nc and
// |w|^2 (column nc) per workgroup of kGsRows rows.
__global__ __launch_bounds__(kThreads) void gs_proj_dot_kernel(const double2* __restrict__ V, long long ldv, int nc,
                                                               const double2* __restrict__ hg, double2* __restrict__ w,
                                                               int N, double2* __restrict__ partial) {
  __shared__ double2 hs[kMaxRestart + 1];
  __shared__ double2 ws[kThreads / 64][kGsCols];
  const int t = threadIdx.x;
  for (int i = t; i < nc; i += kThreads) hs[i] = hg[i];
  __syncthreads();
  const double2 z = make_double2(0.0, 0.0);
  long long n[kGsRpt];
  bool ok[kGsRpt];
  double2 wv[kGsRpt];
#pragma unroll
  for (int r = 0; r < kGsRpt; ++r) {
    n[r] = (long long)blockIdx.x * kGsRows + t + r * kThreads;
    ok[r] = n[r] < N;
    wv[r] = ok[r] ? w[n[r]] : z;
  }
  {
    double2 a[kGsRpt];
#pragma unroll
    for (int r = 0; r < kGsRpt; ++r) a[r] = z;
    int i = 0;
    for (; i + 4 <= nc; i += 4) {
      double2 v[kGsRpt][4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < kGsRpt; ++r) v[r][u] = ok[r] ? V[(i + u) * ldv + n[r]] : z;
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int r = 0; r < kGsRpt; ++r) cmac(a[r], v[r][u], hs[i + u]);
    }
    for (; i < nc; ++i)
#pragma unroll
      for (int r = 0; r < kGsRpt; ++r)
        if (ok[r]) cmac(a[r], V[i * ldv + n[r]], hs[i]);
#pragma unroll
    for (int r = 0; r < kGsRpt; ++r) {
      wv[r].x -= a[r].x;
      wv[r].y -= a[r].y;
      if (ok[r]) w[n[r]] = wv[r];
    }
  }
  const int ncx = nc + 1;
  for (int g0 = 0; g0 < ncx; g0 += kGsCols) {
    double2 acc[kGsCols];
#pragma unroll
    for (int u = 0; u < kGsCols; ++u) {
      const int col = g0 + u;
      acc[u] = z;
#pragma unroll
      for (int r = 0; r < kGsRpt; ++r) {
        double2 v = z;
        if (col < nc) {
          if (ok[r]) v = V[col * ldv + n[r]];
        } else if (col == nc) {
          v = wv[r];
        }
        cdot(acc[u], v, wv[r]);
      }
    }
    store_block_partials(ws, acc, g0, ncx, partial, blockIdx.x);
  }
}

struct GivensState {
  double2* H;      // (m+1) x m, column-major, rotated in place (R factor)
  double* cs;      // rotation cosines
  double2* sn;     // rotation sines
  double2* g;      // rotated beta e1
  double* hn;      // ||w|| of the current Arnoldi step
  double* beta;    // ||r0||
};

// Cycle start: beta = ||r|| from the partials of |r|^2 (one column), g = beta e1.
__global__ __launch_bounds__(kThreads) void gs_cycle_init_kernel(const double2* __restrict__ partial, int nparts,
                                                                 GivensState st, int m, double* host_out) {
  __shared__ double2 red[kThreads];
  __shared__ double2 nrm2[1];
  block_reduce_cols(partial, nparts, 1, nullptr, nrm2, red);
  const int t = threadIdx.x;
  const double b = sqrt(fmax(nrm2[0].x, 0.0));
  for (int i = t; i <= m; i += blockDim.x) st.g[i] = make_double2(i == 0 ? b : 0.0, 0.0);
  if (t == 0) {
    st.beta[0] = b;
    host_out[0] = b;
  }
}

// Arnoldi step j after both Gram-Schmidt passes: d2 = reduced second-pass
// partials (h2 in [0, j], |w'|^2 at j+1), h1 in d1. Writes H column j = h1 + h2,
// ||w''|| = sqrt(|w'|^2 - |h2|^2), rotates the column by the previous Givens
// rotations, forms the rotation that zeroes H(j+1, j), and writes |g(j+1)|
// (the residual estimate) and ||w''|| to host_out[2j], host_out[2j+1].
__global__ __launch_bounds__(kThreads) void gs_tail_kernel(const double2* __restrict__ partial, int nparts,
                                                           const double2* __restrict__ d1, double2* __restrict__ d2,
                                                           GivensState st, int m, int j, double* host_out) {
  __shared__ double2 red[kThreads];
  __shared__ double2 h2[kMaxRestart + 2];
  __shared__ double2 col[kMaxRestart + 2];
  __shared__ double csv[kMaxRestart];
  __shared__ double2 snv[kMaxRestart];
  block_reduce_cols(partial, nparts, j + 2, d2, h2, red);
  const int t = threadIdx.x;
  for (int i = t; i <= j; i += blockDim.x) {
    const double2 a = d1[i];
    col[i] = make_double2(a.x + h2[i].x, a.y + h2[i].y);
    if (i < j) {
      csv[i] = st.cs[i];
      snv[i] = st.sn[i];
    }
  }
  __syncthreads();
  if (t != 0) return;
  double s2 = 0.0;
  for (int i = 0; i <= j; ++i) s2 += h2[i].x * h2[i].x + h2[i].y * h2[i].y;
  const double hn = sqrt(fmax(h2[j + 1].x - s2, 0.0));
  st.hn[0] = hn;
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

// vout = (w - V h) * (1/s[0]) with h the first nc entries of hg; vout = 0 when
// s[0] vanishes (breakdown: the column is never used).
__global__ __launch_bounds__(kThreads) void gs_update_kernel(const double2* __restrict__ V, long long ldv, int nc,
                                                             const double2* __restrict__ hg,
                                                             const double2* __restrict__ w,
                                                             const double* __restrict__ s, double2* __restrict__ vout,
                                                             int N) {
  __shared__ double2 hs[kMaxRestart + 1];
  const int t = threadIdx.x;
  for (int i = t; i < nc; i += kThreads) hs[i] = hg[i];
  __syncthreads();
  const long long n = (long long)blockIdx.x * kThreads + t;
  if (n >= N) return;
  double2 a = make_double2(0.0, 0.0);
  int i = 0;
  for (; i + 8 <= nc; i += 8) {
    double2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = V[(i + u) * ldv + n];
#pragma unroll
    for (int u = 0; u < 8; ++u) cmac(a, v[u], hs[i + u]);
  }
  for (; i < nc; ++i) cmac(a, V[i * ldv + n], hs[i]);
  const double d = s ? s[0] : 1.0;
  const double inv = d > 1e-300 ? 1.0 / d : 0.0;
  const double2 wv = w[n];
  vout[n] = make_double2((wv.x - a.x) * inv, (wv.y - a.y) * inv);
}

// x += V y (k columns)
__global__ __launch_bounds__(kThreads) void axpy_cols_kernel(const double2* __restrict__ V, long long ldv, int nc,
                                                             const double2* __restrict__ coef,
                                                             double2* __restrict__ x, int N) {
  __shared__ double2 cs[kMaxRestart + 1];
  const int t = threadIdx.x;
  for (int i = t; i < nc; i += kThreads) cs[i] = coef[i];
  __syncthreads();
  const long long n = (long long)blockIdx.x * kThreads + t;
  if (n >= N) return;
  double2 a = make_double2(0.0, 0.0);
  int i = 0;
  for (; i + 8 <= nc; i += 8) {
    double2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = V[(i + u) * ldv + n];
#pragma unroll
    for (int u = 0; u < 8; ++u) cmac(a, v[u], cs[i + u]);
  }
  for (; i < nc; ++i) cmac(a, V[i * ldv + n], cs[i]);
  double2 o = x[n];
  o.x += a.x;
  o.y += a.y;
  x[n] = o;
}

// out = b - a (residual); out may alias a
__global__ void sub_kernel(const double2* __restrict__ b, const double2* a, double2* out, int N) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) {
    double2 v = a[n];
    out[n] = make_double2(b[n].x - v.x, b[n].y - v.y);
  }
}

// y = R^{-1} g (upper triangular k x k of the rotated H), column-oriented back
// substitution by one workgroup.
__global__ __launch_bounds__(kThreads) void trisolve_kernel(GivensState st, int m, int k, double2* y) {
  __shared__ double2 s[kMaxRestart];
  __shared__ double2 yi;
  const int t = threadIdx.x;
  for (int i = t; i < k; i += blockDim.x) s[i] = st.g[i];
  __syncthreads();
  for (int i = k - 1; i >= 0; --i) {
    const double2* Rc = st.H + (long long)i * (m + 1);
    if (t == 0) {
      const double2 d = Rc[i];
      const double den = d.x * d.x + d.y * d.y;
      const double2 v = s[i];
      const double2 r = make_double2((v.x * d.x + v.y * d.y) / den, (v.y * d.x - v.x * d.y) / den);
      y[i] = r;
      yi = r;
    }
    __syncthreads();
    const double2 r = yi;
    for (int q = t; q < i; q += blockDim.x) {
      const double2 a = Rc[q];
      s[q].x -= a.x * r.x - a.y * r.y;
      s[q].y -= a.x * r.y + a.y * r.x;
    }
    __syncthreads();
  }
}

__global__ void scale_kernel(const double2* __restrict__ w, const double* __restrict__ s, double2* __restrict__ v,
                             int N) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double d = s[0];
  const double inv = d > 1e-300 ? 1.0 / d : 0.0;
  if (n < N) v[n] = make_double2(w[n].x * inv, w[n].y * inv);
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

static std::vector<double> near_fragments(const std::vector<cplx>& T, int rk);

// Tasks of a fused group: the destinations of the nf passes plist[0..nf)
// (stage s = pass plist[s]) grouped into the connected components of "reads a
// slot written by", i.e. one box's subtree over nf levels, HFR cones of one
// box included (upward: children's charges; downward: the parent's incoming
// potential). Each component is one task; its stage-s destinations are listed
// in fdest from task_stage[t (nf + 1) + s]. Returns an empty string and no
// tasks when a destination reads a slot of its own or a later stage, a pass
// accumulates, or a task grows beyond a few subtrees' worth of destinations
// (one workgroup would serialise them).
static std::string fused_tasks(int nf, const int* plist, const std::vector<DevPass>& passes,
                                  const std::vector<int>& rows, const std::vector<int>& ncol,
                                  const std::vector<int>& dst, const std::vector<long long>& srco,
                                  const std::vector<int>& src, int64_t n_state, std::vector<int>& task_stage,
                                  std::vector<int>& fdest) {
  task_stage.clear();
  fdest.clear();
  std::vector<int> gl, stage;
  for (int s = 0; s < nf; ++s) {
    const DevPass& P = passes[size_t(plist[s])];
    if (P.kind != 0) return "";
    for (int q = 0; q < P.ndest; ++q) {
      gl.push_back(int(P.dest_base) + q);
      stage.push_back(s);
    }
  }
  const int n = int(gl.size());
  if (n == 0) return "";
  std::vector<int> wr(static_cast<size_t>(n_state), -1), parent(static_cast<size_t>(n));
  for (int i = 0; i < n; ++i) {
    parent[size_t(i)] = i;
    const int g = gl[size_t(i)];
    for (int r = 0; r < rows[size_t(g)]; ++r) wr[size_t(dst[size_t(g)]) + r] = i;
  }
  std::function<int(int)> find = [&](int a) {
    while (parent[size_t(a)] != a) a = parent[size_t(a)] = parent[size_t(parent[size_t(a)])];
    return a;
  };
  for (int i = 0; i < n; ++i) {
    const int g = gl[size_t(i)];
    for (int c = 0; c < ncol[size_t(g)]; ++c) {
      const int si = src[size_t(srco[size_t(g)]) + c];
      if (si < 0 || si >= n_state) return "fused upward passes: source index out of range";
      const int j = wr[size_t(si)];
      if (j < 0) continue;                             // x, or a slot written before the group
      if (stage[size_t(j)] >= stage[size_t(i)]) return "";
      const int a = find(i), b = find(j);
      if (a != b) parent[size_t(std::max(a, b))] = std::min(a, b);
    }
  }
  // components in order of their first destination; destinations by stage
  std::vector<int> task_of(size_t(n), -1), root_task(size_t(n), -1);
  int ntask = 0;
  for (int i = 0; i < n; ++i) {
    const int r = find(i);
    if (root_task[size_t(r)] < 0) root_task[size_t(r)] = ntask++;
    task_of[size_t(i)] = root_task[size_t(r)];
  }
  {
    // consecutive components share a task while every stage keeps at most
    // one destination per wavefront of the workgroup
    constexpr int kPack = kFusedThreads / 64;
    std::vector<int> ccnt(size_t(ntask) * nf, 0), packed(static_cast<size_t>(ntask));
    for (int i = 0; i < n; ++i) ++ccnt[size_t(task_of[size_t(i)]) * nf + stage[size_t(i)]];
    std::vector<int> cur(size_t(nf), 0);
    int np = 0;
    for (int c = 0; c < ntask; ++c) {
      bool fits = c > 0;
      for (int s = 0; s < nf && fits; ++s) fits = cur[size_t(s)] + ccnt[size_t(c) * nf + s] <= kPack;
      if (!fits) {
        std::fill(cur.begin(), cur.end(), 0);
        ++np;
      }
      for (int s = 0; s < nf; ++s) cur[size_t(s)] += ccnt[size_t(c) * nf + s];
      packed[size_t(c)] = np - 1;
    }
    for (int i = 0; i < n; ++i) task_of[size_t(i)] = packed[size_t(task_of[size_t(i)])];
    ntask = np;
  }
  std::vector<int> cnt(size_t(ntask) * nf, 0);
  for (int i = 0; i < n; ++i) ++cnt[size_t(task_of[size_t(i)]) * nf + stage[size_t(i)]];
  constexpr int kMaxStage = 256;
  for (int c : cnt)
    if (c > kMaxStage) return "";
  task_stage.assign(size_t(ntask) * (nf + 1), 0);
  std::vector<int> pos(size_t(ntask) * nf);
  int off = 0;
  for (int t = 0; t < ntask; ++t) {
    for (int s = 0; s < nf; ++s) {
      task_stage[size_t(t) * (nf + 1) + s] = off;
      pos[size_t(t) * nf + s] = off;
      off += cnt[size_t(t) * nf + s];
    }
    task_stage[size_t(t) * (nf + 1) + nf] = off;
  }
  fdest.assign(size_t(n), -1);
  for (int i = 0; i < n; ++i) fdest[size_t(pos[size_t(task_of[size_t(i)]) * nf + stage[size_t(i)]]++)] = gl[size_t(i)];
  return "";
}

std::string DeviceOperator::upload(const Plan& plan) {
  N = plan.N;
  rank = plan.rank;
  n_leaves = plan.n_leaves;
  find_phi = plan.find_phi;
  n_state = plan.n_state;
  if (rank > 64) return "nChebNodes^2 > 64 is not supported by the near-field kernel";
  HIP_OK(hipGetDevice(&device));
  HIP_OK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  if (!plan.T.empty()) {
    const int nc = int(std::lround(std::sqrt(double(rank))));
    if (nc * nc != rank || nc < 2 || nc > 8) return "rank must be nChebNodes^2 with nChebNodes in [2, 8]";
    // the near kernel holds one rank x rank table and 2*(256/rank) leaf vectors in LDS
    if (int64_t(sizeof(double2)) * (rank * rank + 2 * (kThreads / rank) * rank) > max_lds)
      return "near-field table does not fit in LDS: nChebNodes too large for this device";
  }
  HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  HIP_OK(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking));
  HIP_OK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  HIP_OK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  HIP_OK(hipEventCreateWithFlags(&ev_leaf, hipEventDisableTiming));
  HIP_OK(hipMalloc((void**)&state, size_t(n_state) * sizeof(double2)));
  HIP_OK(hipMemset(state, 0, size_t(n_state) * sizeof(double2)));
  static_assert(sizeof(cplx) == sizeof(double2), "layout");
  std::string e;
  n_entries = int64_t(plan.ops.size());
  if (!(e = upload_vec(reinterpret_cast<const std::vector<double2>&>(plan.ops), &ops)).empty()) return e;
  if (!(e = upload_vec(plan.src, &src)).empty()) return e;
  std::vector<int> rows, ncol, dst;
  std::vector<long long> opo, srco;
  std::vector<unsigned> rowmap;
  passes.clear();
  up.clear(); down.clear(); leaf_m2l.clear(); leaf_l2l.clear(); near.clear();
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
    else if (ph.name == "leaf_m2l") leaf_m2l.push_back(idx), dp.role = 1;
    else if (ph.name == "leaf_l2l") leaf_l2l.push_back(idx);
    else if (ph.name == "near") near.push_back(idx);
    else down.push_back(idx);
    dp.wave = dp.ndest > 0 && ph.entries <= int64_t(kWaveEntries) * dp.ndest;
    {
      // row map of the one-lane-per-row variant (pass_lane_kernel)
      bool ok = dp.ndest > 0 && dp.ndest < (1 << 24) && ph.n_dst < (int64_t(1) << 31);
      for (int r : ph.rows) ok = ok && r >= 0 && r <= 256;
      dp.lane_ok = ok;
      dp.row_base = int64_t(rowmap.size());
      if (ok) {
        std::vector<int> order(size_t(dp.ndest));
        for (int q = 0; q < dp.ndest; ++q) order[size_t(q)] = q;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ph.ncol[size_t(a)] > ph.ncol[size_t(b)]; });
        for (int q : order)
          for (int r = 0; r < ph.rows[size_t(q)]; ++r) rowmap.push_back((unsigned(q) << 8) | unsigned(r));
        // the kernel runs one lane per entry: n_dst of them
        if (int64_t(rowmap.size()) - dp.row_base != dp.n_dst) {
          rowmap.resize(size_t(dp.row_base));
          dp.lane_ok = false;
        }
      }
    }
    for (size_t q = 0; q < ph.rows.size(); ++q) {
      // a destination lies wholly in the output vector or wholly in the charge/potential slots
      const int64_t d0 = ph.dst[q], d1 = d0 + ph.rows[q];
      if (!((d0 >= N && d1 <= 2 * int64_t(N)) || d0 >= 2 * int64_t(N)) || d1 > n_state)
        return "plan destination straddles the state regions";
    }
    for (int64_t q = 0; q < int64_t(ph.src_off.size()); ++q)
      for (int c = 0; c < ph.ncol[q]; ++c) {
        const int si = plan.src[ph.src_off[q] + c];
        if (si < 0 || si >= n_state || (si >= N && si < 2 * N)) return "plan source index out of range";
      }
    passes.push_back(dp);
  }
  {
    // the leaf M2L waits for the last upward pass that writes one of its sources
    std::vector<int> writer(size_t(n_state), -1);
    for (size_t u = 0; u < up.size(); ++u) {
      const PassHost& ph = plan.passes[size_t(up[u])];
      for (size_t q = 0; q < ph.rows.size(); ++q)
        for (int r = 0; r < ph.rows[q]; ++r) writer[size_t(ph.dst[q]) + r] = int(u);
    }
    leaf_after = 0;
    for (int pi : leaf_m2l) {
      const PassHost& ph = plan.passes[size_t(pi)];
      for (int64_t q = 0; q < int64_t(ph.src_off.size()); ++q)
        for (int c = 0; c < ph.ncol[q]; ++c) leaf_after = std::max(leaf_after, writer[size_t(plan.src[ph.src_off[q] + c])]);
    }
    // the near field runs beside the upward passes: it may read only x
    for (int pi : near) {
      const PassHost& ph = plan.passes[size_t(pi)];
      for (int64_t q = 0; q < int64_t(ph.src_off.size()); ++q)
        for (int c = 0; c < ph.ncol[q]; ++c)
          if (plan.src[ph.src_off[q] + c] >= N) return "near-field pass reads beyond the input vector";
    }
  }
  if (!(e = upload_vec(rows, &d_rows)).empty()) return e;
  if (!(e = upload_vec(ncol, &d_ncol)).empty()) return e;
  if (!(e = upload_vec(dst, &d_dst)).empty()) return e;
  if (!(e = upload_vec(opo, &d_op)).empty()) return e;
  if (!(e = upload_vec(srco, &d_srcoff)).empty()) return e;
  if (!(e = upload_vec(rowmap, &d_rowmap)).empty()) return e;
  {
    // fused groups (each kept or not by the autotuner): the leaf level's
    // upward pass and the two above it; the last two consecutive downward
    // passes of wavefront size (at most kWaveEntries entries per destination
    // on average) before the first larger one
    auto make_group = [&](const std::vector<int>& list, int first, int nf, FusedPasses& g) -> std::string {
      g = FusedPasses();
      if (nf < 2) return "";
      std::vector<int> ts, fd;
      std::string e2 = fused_tasks(nf, list.data() + first, passes, rows, ncol, dst, srco, plan.src, n_state, ts, fd);
      if (!e2.empty() || ts.empty()) return e2;
      if (!(e2 = upload_vec(ts, &g.d_task_stage)).empty()) return e2;
      if (!(e2 = upload_vec(fd, &g.d_fdest)).empty()) return e2;
      g.first = first;
      g.n = nf;
      g.tasks = int(ts.size()) / (nf + 1);
      return "";
    };
    if (!(e = make_group(up, 0, int(std::min<size_t>(3, up.size())), fup)).empty()) return e;
    int last = -1;
    for (size_t i = 0; i < down.size(); ++i) {
      const DevPass& P = passes[size_t(down[i])];
      if (P.ndest > 0 && P.entries > int64_t(kWaveEntries) * P.ndest) break;
      last = int(i);
    }
    if (last >= 1 && !(e = make_group(down, last - 1, 2, fdown)).empty()) return e;
  }
  // near-field tables transposed to column-major per slot (uniform tree only:
  // an adaptive tree's near field is a pass)
  const size_t set_sz = size_t(9) * rank * rank;
  n_table_sets = plan.T.empty() ? 1 : int(plan.T.size() / set_sz);
  if (!plan.T.empty() && plan.T.size() != set_sz * size_t(n_table_sets)) return "near-field tables: expected 9 slots per set";
  if (n_table_sets > 1) {
    // one set per group of 16 leaves (the MFMA kernel's workgroup)
    if (n_leaves % 16 || plan.near_group_set.size() != size_t(n_leaves / 16))
      return "near-field table sets: one set index per 16 leaves expected";
    for (int g : plan.near_group_set)
      if (g < 0 || g >= n_table_sets) return "near-field table set index out of range";
    if (!(rank == 16 || rank == 25 || rank == 36)) return "several near-field table sets need rank 16, 25 or 36";
  }
  std::vector<double2> Tc(plan.T.size());
  for (int ts = 0; ts < (plan.T.empty() ? 0 : n_table_sets); ++ts)
    for (int s = 0; s < 9; ++s)
      for (int i = 0; i < rank; ++i)
        for (int c = 0; c < rank; ++c) {
          const size_t b = ts * set_sz + size_t(s) * rank * rank;
          const cplx v = plan.T[b + size_t(i) * rank + c];
          Tc[b + size_t(c) * rank + i] = make_double2(v.real(), v.imag());
        }
  if (!(e = upload_vec(Tc, &T)).empty()) return e;
  if (!plan.T.empty() && (rank == 16 || rank == 25 || rank == 36)) {
    std::vector<double> frag;
    for (int ts = 0; ts < n_table_sets; ++ts) {
      const std::vector<cplx> one(plan.T.begin() + ts * set_sz, plan.T.begin() + (ts + 1) * set_sz);
      std::vector<double> f = near_fragments(one, rank);
      tm_set_stride = (long long)f.size();
      frag.insert(frag.end(), f.begin(), f.end());
    }
    if (!(e = upload_vec(frag, &Tm)).empty()) return e;
    near_mfma = true;
  }
  if (n_table_sets > 1 && !(e = upload_vec(plan.near_group_set, &near_set)).empty()) return e;
  if (!plan.dest_sys.empty()) {
    if (plan.dest_sys.size() != rows.size()) return "plan: one system index per destination expected";
    if (!(e = upload_vec(plan.dest_sys, &d_dest_sys)).empty()) return e;
  }
  if (!(e = upload_vec(plan.nbr, &nbr)).empty()) return e;
  if (!(e = upload_vec(plan.qk2, &qk2)).empty()) return e;
  near_bytes = plan.T.empty() ? 0 : 16 * int64_t(9) * rank * rank + 16 * 10 * int64_t(N) + 8 * int64_t(N);
  HIP_OK(hipDeviceSynchronize());
  return "";
}

void DeviceOperator::release() {
  if (stream) hipStreamSynchronize(stream);
  if (aux) hipStreamSynchronize(aux);
  work.release();
  drop_graph();
  if (xy_dev) hipFree(xy_dev);
  xy_dev = nullptr;
  void* ptrs[] = {state, ops, src, d_rows, d_ncol, d_dst, d_op, d_srcoff, T, nbr, qk2, ops32, T32, state32, qk2_user,
                  Tm, near_set, d_dest_sys, d_rowmap, fup.d_task_stage, fup.d_fdest, fdown.d_task_stage,
                  fdown.d_fdest};
  for (void* p : ptrs)
    if (p) hipFree(p);
  state = ops = nullptr;
  src = d_rows = d_ncol = d_dst = nbr = nullptr;
  d_op = d_srcoff = nullptr;
  d_rowmap = nullptr;
  fup = FusedPasses();
  fdown = FusedPasses();
  T = nullptr;
  qk2 = nullptr;
  ops32 = T32 = state32 = nullptr;
  qk2_user = nullptr;
  Tm = nullptr;
  near_set = d_dest_sys = nullptr;
  sys_active = nullptr;
  n_table_sets = 1;
  near_mfma = false;
  if (ev_fork) hipEventDestroy(ev_fork);
  if (ev_join) hipEventDestroy(ev_join);
  if (ev_leaf) hipEventDestroy(ev_leaf);
  ev_fork = ev_join = ev_leaf = nullptr;
  if (aux) hipStreamDestroy(aux);
  if (stream) hipStreamDestroy(stream);
  aux = stream = nullptr;
}

template <class R>
static void launch_block(const DeviceOperator& o, const DevPass& p, const typename cx<R>::v* ops,
                         typename cx<R>::t* st, const double2* x, double2* y, hipStream_t s,
                         const double2* const* xyp) {
  const int64_t b = p.dest_base;
  const dim3 grid(p.grid_cap > 0 ? std::min(p.ndest, p.grid_cap) : p.ndest);
  const DestMask mk{o.d_dest_sys && o.sys_active ? o.d_dest_sys + b : nullptr, o.sys_active};
  if (p.kind == 0)
    hipLaunchKernelGGL((pass_block_kernel<false, R, 0>), grid, dim3(kThreads), 0, s, ops, o.src, st, x, y, o.N,
                       o.d_rows + b, o.d_ncol + b, o.d_dst + b, o.d_op + b, o.d_srcoff + b, p.ndest, xyp, mk);
  else if (p.role == 1)
    hipLaunchKernelGGL((pass_block_kernel<true, R, 1>), grid, dim3(kThreads), 0, s, ops, o.src, st, x, y, o.N,
                       o.d_rows + b, o.d_ncol + b, o.d_dst + b, o.d_op + b, o.d_srcoff + b, p.ndest, xyp, mk);
  else
    hipLaunchKernelGGL((pass_block_kernel<true, R, 0>), grid, dim3(kThreads), 0, s, ops, o.src, st, x, y, o.N,
                       o.d_rows + b, o.d_ncol + b, o.d_dst + b, o.d_op + b, o.d_srcoff + b, p.ndest, xyp, mk);
}

template <class R>
static void launch_pass(const DeviceOperator& o, const DevPass& p, const double2* x, double2* y, hipStream_t s,
                        const double2* const* xyp = nullptr) {
  const int64_t b = p.dest_base;
  typedef typename cx<R>::v V;
  typedef typename cx<R>::t C;
  const V* ops = reinterpret_cast<const V*>(std::is_same<R, double>::value ? (const void*)o.ops : (const void*)o.ops32);
  C* st = reinterpret_cast<C*>(std::is_same<R, double>::value ? (void*)o.state : (void*)o.state32);
  if (p.lane && p.lane_ok) {
    const int nrows = int(p.n_dst);
    const dim3 grid((nrows + kThreads - 1) / kThreads);
    const DestMask mk{o.d_dest_sys && o.sys_active ? o.d_dest_sys + b : nullptr, o.sys_active};
    const unsigned* rm = o.d_rowmap + p.row_base;
    if (p.kind == 0)
      hipLaunchKernelGGL((pass_lane_kernel<false, R>), grid, dim3(kThreads), 0, s, ops, o.src, st, x, y, o.N,
                         o.d_rows + b, o.d_ncol + b, o.d_dst + b, o.d_op + b, o.d_srcoff + b, rm, nrows, xyp, mk);
    else
      hipLaunchKernelGGL((pass_lane_kernel<true, R>), grid, dim3(kThreads), 0, s, ops, o.src, st, x, y, o.N,
                         o.d_rows + b, o.d_ncol + b, o.d_dst + b, o.d_op + b, o.d_srcoff + b, rm, nrows, xyp, mk);
    return;
  }
  if (p.wave) {
    const dim3 grid((p.ndest + 3) / 4);
    const DestMask mk{o.d_dest_sys && o.sys_active ? o.d_dest_sys + b : nullptr, o.sys_active};
    if (p.kind == 0)
      hipLaunchKernelGGL((pass_wave_kernel<false, R>), grid, dim3(kThreads), 0, s, ops, o.src, st, x, y, o.N,
                         o.d_rows + b, o.d_ncol + b, o.d_dst + b, o.d_op + b, o.d_srcoff + b, p.ndest, xyp, mk);
    else
      hipLaunchKernelGGL((pass_wave_kernel<true, R>), grid, dim3(kThreads), 0, s, ops, o.src, st, x, y, o.N,
                         o.d_rows + b, o.d_ncol + b, o.d_dst + b, o.d_op + b, o.d_srcoff + b, p.ndest, xyp, mk);
    return;
  }
  launch_block<R>(o, p, ops, st, x, y, s, xyp);
}

// a fused group of passes as one launch of pass_fused_kernel
template <class R>
static void launch_fused(const DeviceOperator& o, const FusedPasses& g, const double2* x, double2* y, hipStream_t s,
                         const double2* const* xyp = nullptr) {
  typedef typename cx<R>::v V;
  typedef typename cx<R>::t C;
  const V* ops = reinterpret_cast<const V*>(std::is_same<R, double>::value ? (const void*)o.ops : (const void*)o.ops32);
  C* st = reinterpret_cast<C*>(std::is_same<R, double>::value ? (void*)o.state : (void*)o.state32);
  const DestMask mk{o.d_dest_sys && o.sys_active ? o.d_dest_sys : nullptr, o.sys_active};
  hipLaunchKernelGGL((pass_fused_kernel<R>), dim3(g.tasks), dim3(kFusedThreads), 0, s, ops, o.src, st, x, y, o.N,
                     o.d_rows, o.d_ncol, o.d_dst, o.d_op, o.d_srcoff, g.d_task_stage, g.d_fdest, g.n, xyp, mk);
}

template <class R, int RK, int NL>
static void launch_near_rk(const DeviceOperator& o, const double2* x, double2* y, hipStream_t s,
                           const double2* const* xyp) {
  constexpr int LB = NL * (kThreads / RK);
  typedef typename cx<R>::t C;
  const C* T = reinterpret_cast<const C*>(std::is_same<R, double>::value ? (const void*)o.T : (const void*)o.T32);
  hipLaunchKernelGGL((near_kernel<R, RK, NL>), dim3((o.n_leaves + LB - 1) / LB), dim3(kThreads), 0, s, T, o.nbr,
                     o.qk2, x, y, o.n_leaves, o.find_phi ? 1 : 0, xyp);
}

template <class R, int NL>
static void launch_near_nl(const DeviceOperator& o, const double2* x, double2* y, hipStream_t s,
                           const double2* const* xyp) {
  switch (o.rank) {
    case 4: launch_near_rk<R, 4, NL>(o, x, y, s, xyp); break;
    case 9: launch_near_rk<R, 9, NL>(o, x, y, s, xyp); break;
    case 16: launch_near_rk<R, 16, NL>(o, x, y, s, xyp); break;
    case 25: launch_near_rk<R, 25, NL>(o, x, y, s, xyp); break;
    case 36: launch_near_rk<R, 36, NL>(o, x, y, s, xyp); break;
    case 49: launch_near_rk<R, 49, NL>(o, x, y, s, xyp); break;
    case 64: launch_near_rk<R, 64, NL>(o, x, y, s, xyp); break;
    default: break;  // rejected at upload
  }
}

static bool launch_near_mfma(const DeviceOperator& o, const double2* x, double2* y, hipStream_t s,
                             const double2* const* xyp);

template <class R>
static void launch_near(const DeviceOperator& o, const double2* x, double2* y, hipStream_t s,
                        const double2* const* xyp = nullptr) {
  if (std::is_same<R, double>::value && (o.near_mfma || o.n_table_sets > 1) && o.Tm && launch_near_mfma(o, x, y, s, xyp))
    return;
  if (o.n_table_sets > 1) return;  // rejected at upload (FP32 and FMA variants hold one table set)
  if (o.near_leaves_per_thread == 1) launch_near_nl<R, 1>(o, x, y, s, xyp);
  else launch_near_nl<R, 2>(o, x, y, s, xyp);
}

// Near field of a uniform tree on FP64 MFMA (16 x 16 x 4): per slot s the
// product T_s [x_nbr(l,s)] over all leaves l is a GEMM with the leaves as
// columns. Complex arithmetic is real-ified: real row 2r + a (a = 0 Re, 1 Im)
// of table row r, real column 2c + b of table column c, entries
// [[Re T, -Im T], [Im T, Re T]]. A workgroup owns 16 leaves (the MFMA
// columns); wave w owns real rows [16w, 16w + 16). Lane l holds the operand
// fragments A(row l%16, k l/16), read from Tm (pre-arranged per slot, row
// tile, k-step and lane: one coalesced load per k-step), and B(k l/16, leaf
// l%16), read from the slot's neighbour vectors staged in LDS; the result
// rows l/16 + 4v of leaf l%16. The next slot's neighbour vectors and
// fragments are loaded into registers while the current slot's products run.
template <int RK>
__global__ __launch_bounds__(64 * ((2 * RK + 15) / 16)) void near_mfma_kernel(
    const double* __restrict__ Tm_all, const int* __restrict__ nbr, const double* __restrict__ qk2,
    const double2* __restrict__ x_arg, double2* __restrict__ y_arg, int n_leaves, int find_phi,
    const double2* const* __restrict__ xyp, const int* __restrict__ tset, long long set_stride,
    const int* __restrict__ set_active) {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  constexpr int MT = (2 * RK + 15) / 16;  // real row tiles = waves
  constexpr int KS = (2 * RK + 3) / 4;    // k-steps of 4 real columns
  constexpr int NT = 64 * MT;
  constexpr int NE = 16 * 2 * RK;         // reals of one slot's 16 neighbour vectors
  constexpr int NG = (NE + NT - 1) / NT;  // of them per thread
  constexpr int LS = 2 * RK + 2;          // LDS row stride (reals) of one leaf
  const double* __restrict__ x = reinterpret_cast<const double*>(xyp ? xyp[0] : x_arg);
  double* __restrict__ y = reinterpret_cast<double*>(xyp ? const_cast<double2*>(xyp[1]) : y_arg);
  __shared__ double Xs[2][16 * LS];
  __shared__ int nbs[16 * 9];
  const int t = threadIdx.x, lane = t & 63, mt = t >> 6;
  const int j = lane & 15, k4 = lane >> 4;
  const int leaf0 = blockIdx.x * 16;
  // the workgroup's 16 leaves share one table set, one per system of a fused
  // plan; the leaves of inactive systems are skipped
  if (tset && set_active && !set_active[tset[blockIdx.x]]) return;
  const double* __restrict__ Tm = Tm_all + (tset ? tset[blockIdx.x] * set_stride : 0);
  for (int e = t; e < 16 * 9; e += NT) nbs[e] = leaf0 + e / 9 < n_leaves ? nbr[leaf0 * 9 + e] : -1;
  __syncthreads();
  double g[NG], a_cur[KS], a_nxt[KS];
  auto gather = [&](int s) {
#pragma unroll
    for (int q = 0; q < NG; ++q) {
      const int e = t + q * NT;
      const int l = e / (2 * RK), d = e - l * (2 * RK);
      const int nb = e < NE ? nbs[l * 9 + s] : -1;
      g[q] = nb >= 0 ? x[size_t(nb) * 2 * RK + d] : 0.0;
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int q = 0; q < NG; ++q) {
      const int e = t + q * NT;
      const int l = e / (2 * RK), d = e - l * (2 * RK);
      if (e < NE) Xs[buf][l * LS + d] = g[q];
    }
  };
  auto frags = [&](int s, double* a) {
    const double* A = Tm + (size_t(s) * MT + mt) * KS * 64 + lane;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) a[ks] = A[ks * 64];
  };
  gather(0);
  stage(0);
  frags(0, a_cur);
  acc_t acc = {0, 0, 0, 0};
  for (int s = 0; s < 9; ++s) {
    __syncthreads();  // slot s staged; the other buffer is no longer read
    if (s + 1 < 9) {
      gather(s + 1);
      frags(s + 1, a_nxt);
    }
    const double* B = Xs[s & 1] + j * LS + k4;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const double bv = 4 * ks + k4 < 2 * RK ? B[4 * ks] : 0.0;
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a_cur[ks], bv, acc, 0, 0, 0);
    }
    if (s + 1 < 9) {
      stage((s + 1) & 1);
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) a_cur[ks] = a_nxt[ks];
    }
  }
  const int leaf = leaf0 + j;
  if (leaf >= n_leaves) return;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int ri = 16 * mt + k4 + 4 * v;
    if (ri < 2 * RK) {
      const size_t g2 = size_t(leaf) * RK + (ri >> 1);
      double val = acc[v];
      if (find_phi) val = x[2 * g2 + (ri & 1)] + qk2[g2] * val;
      y[2 * g2 + (ri & 1)] = val;
    }
  }
}

template <int RK>
static void launch_near_mfma_rk(const DeviceOperator& o, const double2* x, double2* y, hipStream_t s,
                                const double2* const* xyp) {
  constexpr int MT = (2 * RK + 15) / 16;
  hipLaunchKernelGGL((near_mfma_kernel<RK>), dim3((o.n_leaves + 15) / 16), dim3(64 * MT), 0, s, o.Tm, o.nbr, o.qk2,
                     x, y, o.n_leaves, o.find_phi ? 1 : 0, xyp, o.near_set, o.tm_set_stride, o.sys_active);
}

static bool launch_near_mfma(const DeviceOperator& o, const double2* x, double2* y, hipStream_t s,
                             const double2* const* xyp) {
  switch (o.rank) {
    case 16: launch_near_mfma_rk<16>(o, x, y, s, xyp); return true;
    case 25: launch_near_mfma_rk<25>(o, x, y, s, xyp); return true;
    case 36: launch_near_mfma_rk<36>(o, x, y, s, xyp); return true;
    default: return false;
  }
}

// Tm from the 9 tables (T: row-major rank x rank per slot): fragment value of
// lane l at (slot s, row tile mt, k-step ks) is the real-ified entry at real
// row 16 mt + l%16, real column 4 ks + l/16 (zero beyond 2 rank)
static std::vector<double> near_fragments(const std::vector<cplx>& T, int rk) {
  const int MT = (2 * rk + 15) / 16, KS = (2 * rk + 3) / 4;
  std::vector<double> out(size_t(9) * MT * KS * 64, 0.0);
  for (int s = 0; s < 9; ++s)
    for (int mt = 0; mt < MT; ++mt)
      for (int ks = 0; ks < KS; ++ks)
        for (int l = 0; l < 64; ++l) {
          const int ri = 16 * mt + (l & 15), rc = 4 * ks + (l >> 4);
          if (ri >= 2 * rk || rc >= 2 * rk) continue;
          const int r = ri >> 1, a = ri & 1, c = rc >> 1, b = rc & 1;
          const cplx v = T[size_t(s) * rk * rk + size_t(r) * rk + c];
          const double e = a == b ? v.real() : (a == 1 ? v.imag() : -v.imag());
          out[((size_t(s) * MT + mt) * KS + ks) * 64 + l] = e;
        }
  return out;
}

// y = x + qk2 .* y: the Lippmann-Schwinger operator from the volume potential
__global__ void diag_kernel(const double2* __restrict__ x_arg, double2* __restrict__ y_arg,
                            const double* __restrict__ qk2, int N, const double2* const* __restrict__ xyp) {
  const double2* __restrict__ x = xyp ? xyp[0] : x_arg;
  double2* __restrict__ y = xyp ? const_cast<double2*>(xyp[1]) : y_arg;
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) {
    const double q = qk2[n];
    const double2 a = x[n], v = y[n];
    y[n] = make_double2(a.x + q * v.x, a.y + q * v.y);
  }
}

template <class R>
static void matvec_impl(DeviceOperator& o, const double2* x, double2* y, const double2* const* xyp = nullptr) {
  if (!xyp && x == y) {
    // in-place call: the passes read x while the output is written
    hipMemcpyAsync(o.state, x, size_t(o.N) * sizeof(double2), hipMemcpyDeviceToDevice, o.stream);
    x = o.state;
  }
  // The near field needs only x: it starts on the second stream at once. The
  // leaf M2L there waits for the first upward pass (the leaf level, the only
  // charges it reads).
  hipEventRecord(o.ev_fork, o.stream);
  hipStreamWaitEvent(o.aux, o.ev_fork, 0);
  if (o.T) launch_near<R>(o, x, y, o.aux, xyp);
  for (int p : o.near) launch_pass<R>(o, o.passes[p], x, y, o.aux, xyp);
  // the leaf M2L on the second stream, once the charges it reads are written
  // (and, by the schedule the autotuner picked, later)
  bool m2l_started = false;
  auto start_m2l = [&]() {
    if (m2l_started) return;
    m2l_started = true;
    hipEventRecord(o.ev_leaf, o.stream);
    hipStreamWaitEvent(o.aux, o.ev_leaf, 0);
    for (int p : o.leaf_m2l) launch_pass<R>(o, o.passes[p], x, y, o.aux, xyp);
    hipEventRecord(o.ev_join, o.aux);
  };
  size_t u = 0;
  if (o.fup.on) {
    launch_fused<R>(o, o.fup, x, y, o.stream, xyp);
    u = size_t(o.fup.n);
  }
  for (; u < o.up.size() && int(u) <= o.leaf_after; ++u) launch_pass<R>(o, o.passes[o.up[u]], x, y, o.stream, xyp);
  if (o.m2l_start == 0) start_m2l();
  for (; u < o.up.size(); ++u) launch_pass<R>(o, o.passes[o.up[u]], x, y, o.stream, xyp);
  if (o.m2l_start == 1) start_m2l();
  for (size_t i = 0; i < o.down.size(); ++i) {
    const bool fused = o.fdown.on && int(i) == o.fdown.first;
    const size_t last = i + (fused ? size_t(o.fdown.n) - 1 : 0);
    if (o.m2l_start == 2 && last + 1 == o.down.size()) start_m2l();
    if (fused) launch_fused<R>(o, o.fdown, x, y, o.stream, xyp);
    else launch_pass<R>(o, o.passes[o.down[i]], x, y, o.stream, xyp);
    i = last;
  }
  start_m2l();
  hipStreamWaitEvent(o.stream, o.ev_join, 0);
  for (int p : o.leaf_l2l) launch_pass<R>(o, o.passes[p], x, y, o.stream, xyp);
  if (o.qk2_user)
    hipLaunchKernelGGL(diag_kernel, dim3((o.N + kThreads - 1) / kThreads), dim3(kThreads), 0, o.stream, x, y,
                       o.qk2_user, o.N, xyp);
}

__global__ void set_xy_kernel(const double2** xy, const double2* x, double2* y) {
  xy[0] = x;
  xy[1] = y;
}

void DeviceOperator::matvec(const double2* x, double2* y) {
  if (graph_exec && x != y) {
    // the captured sequence reads its input and output addresses from xy_dev
    hipLaunchKernelGGL(set_xy_kernel, dim3(1), dim3(1), 0, stream, xy_dev, x, y);
    hipGraphLaunch(graph_exec, stream);
    return;
  }
  matvec_impl<double>(*this, x, y);
}

std::string DeviceOperator::capture_graph() {
  if (graph_exec) return "";
  if (!xy_dev) HIP_OK(hipMalloc((void**)&xy_dev, 2 * sizeof(double2*)));
  HIP_OK(hipStreamSynchronize(stream));
  HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  matvec_impl<double>(*this, nullptr, nullptr, const_cast<const double2* const*>(xy_dev));
  hipGraph_t g = nullptr;
  HIP_OK(hipStreamEndCapture(stream, &g));
  hipError_t e = hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    graph_exec = nullptr;
    return std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
  }
  return "";
}

void DeviceOperator::drop_graph() {
  if (!graph_exec) return;
  // a replay may still be in flight after an asynchronous call
  (void)hipStreamSynchronize(stream);
  (void)hipStreamSynchronize(aux);
  hipGraphExecDestroy(graph_exec);
  graph_exec = nullptr;
}

// Per-pass choice between one wavefront and one workgroup per destination, and
// 1 or 2 leaves per thread in the near field: each candidate is timed with HIP
// events on this operator (best of `reps`) and the faster one kept.
std::string DeviceOperator::autotune(int reps, int* changed) {
  drop_graph();
  const double2* x = state;
  double2* y = state + N;
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  auto time_it = [&](auto&& launch) {
    float best = 1e30f;
    launch();   // warm-up
    for (int r = 0; r < reps; ++r) {
      hipEventRecord(a, stream);
      launch();
      hipEventRecord(b, stream);
      hipEventSynchronize(b);
      float ms = 0;
      hipEventElapsedTime(&ms, a, b);
      best = std::min(best, ms);
    }
    return best;
  };
  int nchg = 0;
  tuning.clear();
  for (DevPass& p : passes) {
    const bool wave_before = p.wave, lane_before = p.lane;
    p.lane = false;
    p.wave = false;
    const float tb = time_it([&]() { launch_pass<double>(*this, p, x, y, stream); });
    p.wave = true;
    const float tw = time_it([&]() { launch_pass<double>(*this, p, x, y, stream); });
    float tl = 1e30f;
    if (p.lane_ok) {
      p.lane = true;
      tl = time_it([&]() { launch_pass<double>(*this, p, x, y, stream); });
    }
    p.lane = tl < std::min(tb, tw);
    p.wave = !p.lane && tw < tb;
    nchg += (p.wave != wave_before) + (p.lane != lane_before);
    tuning.push_back(p.name + (p.lane ? ": lane" : p.wave ? ": wave" : ": block") + " (" + std::to_string(tb * 1e3f) +
                     " us block, " + std::to_string(tw * 1e3f) + " us wave, " +
                     (p.lane_ok ? std::to_string(tl * 1e3f) + " us lane)" : std::string("no lane variant)")));
  }
  auto tune_group = [&](FusedPasses& g, const std::vector<int>& list, const char* what) {
    if (g.tasks == 0) return;
    const bool before = g.on;
    const float tf = time_it([&]() { launch_fused<double>(*this, g, x, y, stream); });
    const float ts = time_it([&]() {
      for (int u = g.first; u < g.first + g.n; ++u) launch_pass<double>(*this, passes[size_t(list[size_t(u)])], x, y, stream);
    });
    g.on = tf < ts;
    nchg += g.on != before;
    tuning.push_back(std::string(what) + " subtrees: " + (g.on ? "fused" : "separate") + " (" + std::to_string(tf * 1e3f) +
                     " us fused over " + std::to_string(g.tasks) + " subtrees, " + std::to_string(ts * 1e3f) + " us as " +
                     std::to_string(g.n) + " passes from " + passes[size_t(list[size_t(g.first)])].name + ")");
  };
  tune_group(fup, up, "upward");
  tune_group(fdown, down, "downward");
  if (!leaf_m2l.empty()) {
    // the two-stream schedule, timed on the whole matvec: when the leaf M2L
    // starts, and how many workgroups it may hold (grid-stride), so that the
    // dependent chain of small passes on the main stream still finds CUs
    DevPass& lp = passes[size_t(leaf_m2l.front())];
    const int start_before = m2l_start, cap_before = lp.grid_cap;
    const int ncu = [&]() {
      int n = 0;
      return hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && n > 0 ? n : 256;
    }();
    const int caps[4] = {0, 4 * ncu, 2 * ncu, ncu};
    float best = 1e30f;
    int bs = 0, bc = 0;
    std::string rep = "leaf M2L schedule:";
    for (int st = 0; st < 3; ++st)
      for (int ci = 0; ci < 4; ++ci) {
        if (lp.wave || lp.lane) {
          if (ci > 0) continue;  // the cap applies to the one-workgroup kernel
        }
        m2l_start = st;
        lp.grid_cap = caps[ci];
        const float t = time_it([&]() { matvec_impl<double>(*this, x, y); });
        rep += " start" + std::to_string(st) + "/cap" + std::to_string(caps[ci]) + " " + std::to_string(t * 1e3f) + " us;";
        if (t < best) {
          best = t;
          bs = st;
          bc = caps[ci];
        }
      }
    m2l_start = bs;
    lp.grid_cap = bc;
    nchg += (m2l_start != start_before) + (lp.grid_cap != cap_before);
    tuning.push_back(rep + " kept start" + std::to_string(bs) + "/cap" + std::to_string(bc));
  }
  const int nl_before = near_leaves_per_thread;
  if (!T) {
    tuning.push_back("near_field: passes (adaptive tree)");
  } else if (n_table_sets > 1) {
    tuning.push_back("near_field: MFMA (" + std::to_string(n_table_sets) + " table sets)");
  } else {
  const bool mfma_before = near_mfma;
  near_mfma = false;
  near_leaves_per_thread = 1;
  const float t1 = time_it([&]() { launch_near<double>(*this, x, y, stream); });
  near_leaves_per_thread = 2;
  const float t2 = time_it([&]() { launch_near<double>(*this, x, y, stream); });
  near_leaves_per_thread = t1 < t2 ? 1 : 2;
  float tm = 1e30f;
  if (Tm) {
    near_mfma = true;
    tm = time_it([&]() { launch_near<double>(*this, x, y, stream); });
    near_mfma = tm < std::min(t1, t2);
  }
  nchg += (near_leaves_per_thread != nl_before) + (near_mfma != mfma_before);
  tuning.push_back("near_field: " + std::string(near_mfma ? "MFMA" : "FMA") + "; FMA kernel " +
                   std::to_string(near_leaves_per_thread) + " leaves per thread (" + std::to_string(t1 * 1e3f) +
                   " us with 1, " + std::to_string(t2 * 1e3f) + " us with 2, " +
                   (Tm ? std::to_string(tm * 1e3f) + " us on MFMA)" : std::string("no MFMA variant)")));
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  HIP_OK(hipStreamSynchronize(stream));
  if (changed) *changed = nchg;
  return "";
}
void DeviceOperator::matvec32(const double2* x, double2* y) { matvec_impl<float>(*this, x, y); }

__global__ void to_f32_kernel(const double2* __restrict__ a, float2* __restrict__ b, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = make_float2(float(a[i].x), float(a[i].y));
}

std::string DeviceOperator::enable_fp32() {
  if (ops32) return "";
  if (n_table_sets > 1) return "the FP32 operator holds one near-field table set (not a fused plan)";
  const long long n_ops = n_entries, n_T = T ? 9LL * rank * rank : 0;
  HIP_OK(hipMalloc((void**)&ops32, sizeof(float2) * size_t(std::max(n_ops, 1LL))));
  if (n_T) HIP_OK(hipMalloc((void**)&T32, sizeof(float2) * size_t(n_T)));
  HIP_OK(hipMalloc((void**)&state32, sizeof(float2) * size_t(n_state)));
  HIP_OK(hipMemsetAsync(state32, 0, sizeof(float2) * size_t(n_state), stream));
  if (n_ops > 0)
    hipLaunchKernelGGL(to_f32_kernel, dim3(unsigned((n_ops + 255) / 256)), dim3(256), 0, stream, ops, ops32, n_ops);
  if (n_T)
    hipLaunchKernelGGL(to_f32_kernel, dim3(unsigned((n_T + 255) / 256)), dim3(256), 0, stream, T, T32, n_T);
  HIP_OK(hipStreamSynchronize(stream));
  return "";
}

double DeviceOperator::time_matvec(int reps, bool fp32) {
  double2 *x = nullptr, *y = nullptr;
  if (hipMalloc((void**)&x, sizeof(double2) * size_t(N)) != hipSuccess) return -1;
  if (hipMalloc((void**)&y, sizeof(double2) * size_t(N)) != hipSuccess) {
    hipFree(x);
    return -1;
  }
  hipMemsetAsync(x, 0, sizeof(double2) * size_t(N), stream);
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  auto mv = [&]() {
    if (fp32) matvec32(x, y);
    else matvec(x, y);
  };
  mv();
  hipEventRecord(a, stream);
  for (int r = 0; r < reps; ++r) mv();
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
  // serialised on the main stream: upward, near field, downward, leaf M2L, leaf L2L;
  // input and output are the state's own x and y regions
  const double2* x = state;
  double2* y = state + N;
  std::vector<int> order(up.begin(), up.end());
  if (T) order.push_back(-1);
  order.insert(order.end(), near.begin(), near.end());
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
      if (order[q] < 0) launch_near<double>(*this, x, y, stream);
      else launch_pass<double>(*this, passes[order[q]], x, y, stream);
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
    if (order[q] < 0) {
      names[q] = "near_field";
      bytes[q] = near_bytes;
    } else {
      names[q] = passes[order[q]].name;
      bytes[q] = passes[order[q]].bytes();
    }
  }
  for (auto& e : ev) hipEventDestroy(e);
  return "";
}

// ------------------------------------------------------------------ GMRES ---
void GmresWork::release() {
  void* ptrs[] = {V, w, partial, d1, d2, y, H, sn, g, cs, hn, beta, Z};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (hostbuf) hipHostFree(hostbuf);
  *this = GmresWork{};
}

std::string GmresWork::reserve(int N_, int m_, bool flexible) {
  if (m == m_ && N == N_ && (!flexible || Z)) return "";
  release();
  if (flexible) HIP_OK(hipMalloc((void**)&Z, sizeof(double2) * size_t(N_) * m_));
  const int nrb = (N_ + kGsRows - 1) / kGsRows;
  HIP_OK(hipMalloc((void**)&V, sizeof(double2) * size_t(N_) * (m_ + 1)));
  HIP_OK(hipMalloc((void**)&w, sizeof(double2) * size_t(N_)));
  HIP_OK(hipMalloc((void**)&partial, sizeof(double2) * size_t(nrb) * (m_ + 2)));
  HIP_OK(hipMalloc((void**)&d1, sizeof(double2) * (m_ + 2)));
  HIP_OK(hipMalloc((void**)&d2, sizeof(double2) * (m_ + 2)));
  HIP_OK(hipMalloc((void**)&y, sizeof(double2) * (m_ + 1)));
  HIP_OK(hipMalloc((void**)&H, sizeof(double2) * size_t(m_ + 1) * m_));
  HIP_OK(hipMalloc((void**)&cs, sizeof(double) * (m_ + 1)));
  HIP_OK(hipMalloc((void**)&sn, sizeof(double2) * (m_ + 1)));
  HIP_OK(hipMalloc((void**)&g, sizeof(double2) * (m_ + 2)));
  HIP_OK(hipMalloc((void**)&hn, sizeof(double)));
  HIP_OK(hipMalloc((void**)&beta, sizeof(double)));
  HIP_OK(hipHostMalloc((void**)&hostbuf, sizeof(double) * 2 * (m_ + 2), hipHostMallocDefault));
  m = m_;
  N = N_;
  return "";
}

// ------------------------------------------------ allocation-free workspace
// One array per device for the whole process (the library's scratch): calls
// that use it on one device must not overlap. Its host-side name stands for
// no device address; scratch_workspace() returns the address on a device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::string scratch_workspace(int device, void** ptr) {
  static std::mutex mu;
  static void* addr[64] = {};
  if (device < 0 || device >= 64) return "device index out of range";
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    int cur = -1;
    HIP_OK(hipGetDevice(&cur));
    if (cur != device) return "scratch_workspace: the device must be current";
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return "hipGetSymbolAddress returned no address for the scratch array";
    addr[device] = p;
  }
  *ptr = addr[device];
  return "";
}

namespace {
// sub-buffer sizes of a GMRES(m) workspace, each rounded to 256 bytes
struct WorkLayout {
  size_t off[13] = {};
  size_t total = 0;
  WorkLayout(int N, int m) {
    const int nrb = (N + kGsRows - 1) / kGsRows;
    const size_t sz[13] = {sizeof(double2) * size_t(N) * (m + 1),   // V
                           sizeof(double2) * size_t(N),             // w
                           sizeof(double2) * size_t(nrb) * (m + 2), // partial
                           sizeof(double2) * (m + 2),               // d1
                           sizeof(double2) * (m + 2),               // d2
                           sizeof(double2) * (m + 1),               // y
                           sizeof(double2) * size_t(m + 1) * m,     // H
                           sizeof(double) * (m + 1),                // cs
                           sizeof(double2) * (m + 1),               // sn
                           sizeof(double2) * (m + 2),               // g
                           sizeof(double),                          // hn
                           sizeof(double),                          // beta
                           sizeof(double) * 2 * (m + 2)};           // hostbuf (device side)
    for (int i = 0; i < 13; ++i) {
      off[i] = total;
      total += (sz[i] + 255) / 256 * 256;
    }
  }
};
}  // namespace

size_t gmres_work_bytes(int N, int m) { return WorkLayout(N, m).total; }

std::string GmresWork::carve(void* base, size_t cap_bytes, int N_, int m_, double* mirror) {
  const WorkLayout L(N_, m_);
  if (L.total > cap_bytes) return "workspace does not fit";
  char* b = static_cast<char*>(base);
  V = reinterpret_cast<double2*>(b + L.off[0]);
  w = reinterpret_cast<double2*>(b + L.off[1]);
  partial = reinterpret_cast<double2*>(b + L.off[2]);
  d1 = reinterpret_cast<double2*>(b + L.off[3]);
  d2 = reinterpret_cast<double2*>(b + L.off[4]);
  y = reinterpret_cast<double2*>(b + L.off[5]);
  H = reinterpret_cast<double2*>(b + L.off[6]);
  cs = reinterpret_cast<double*>(b + L.off[7]);
  sn = reinterpret_cast<double2*>(b + L.off[8]);
  g = reinterpret_cast<double2*>(b + L.off[9]);
  hn = reinterpret_cast<double*>(b + L.off[10]);
  beta = reinterpret_cast<double*>(b + L.off[11]);
  hostbuf = reinterpret_cast<double*>(b + L.off[12]);
  host_mirror = mirror;
  Z = nullptr;
  N = N_;
  m = m_;
  return "";
}

GmresResult device_gmres(DeviceOperator& op, const double2* b_dev, double2* x_dev, int m, int max_restarts,
                         double tol, bool fp32, Precond* prec, GmresWork* ext) {
  GmresResult res;
  auto t0 = std::chrono::steady_clock::now();
  const int N = op.N;
  hipStream_t s = op.stream;
  if (m < 1 || m > kMaxRestart) {
    res.status = LSG_ERR_ARG;
    res.error = "restart length must be in [1, 500]";
    return res;
  }
  if (fp32 && !op.ops32) {
    res.status = LSG_ERR_ARG;
    res.error = "FP32 operators are not enabled";
    return res;
  }
  const int nrb = (N + kGsRows - 1) / kGsRows;   // row blocks of the dot kernels
  const int nb_vec = (N + kThreads - 1) / kThreads;
  const long long ldv = N;
  if (ext && (ext->N != N || ext->m != m || prec)) {
    res.status = LSG_ERR_ARG;
    res.error = "external GMRES workspace does not match (N, m), or has no room for a preconditioner";
    return res;
  }
  if (!ext) {
    std::string e = op.work.reserve(N, m, prec != nullptr);
    if (!e.empty()) {
      op.work.release();
      res.status = LSG_ERR_DEVICE;
      res.error = "GMRES workspace allocation failed: " + e;
      return res;
    }
  }
  GmresWork& W = ext ? *ext : op.work;
  double2 *V = W.V, *w = W.w, *partial = W.partial, *d1 = W.d1, *d2 = W.d2, *y = W.y;
  GivensState st{W.H, W.cs, W.sn, W.g, W.hn, W.beta};
  // the kernels write hostbuf; the host reads hb (the same pinned buffer, or
  // a host copy of a device buffer made after each synchronisation)
  double* hostbuf = W.hostbuf;
  const double* hb = W.host_mirror ? W.host_mirror : W.hostbuf;
  auto pull = [&](int n) -> hipError_t {
    if (!W.host_mirror) return hipSuccess;
    hipError_t pe = hipMemcpyAsync(W.host_mirror, W.hostbuf, sizeof(double) * size_t(n), hipMemcpyDeviceToHost, s);
    return pe != hipSuccess ? pe : hipStreamSynchronize(s);
  };
  auto matvec = [&](const double2* xin, double2* yout) {
    if (fp32) op.matvec32(xin, yout);
    else op.matvec(xin, yout);
  };
  // partials of V[:, 0:nc]^H vec (+ |vec|^2 as column nc when norm)
  auto dot = [&](int nc, const double2* vec, bool norm) {
    const int ncx = nc + (norm ? 1 : 0);
    hipLaunchKernelGGL(gs_dot_kernel, dim3(nrb, (ncx + kGsCols - 1) / kGsCols), dim3(kThreads), 0, s, V, ldv, nc,
                       ncx, vec, N, partial);
  };
  // ||b||
  dot(0, b_dev, true);
  hipLaunchKernelGGL(gs_reduce_kernel, dim3(1), dim3(kThreads), 0, s, partial, nrb, 1, d1);
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
    matvec(x_dev, w);
    hipLaunchKernelGGL(sub_kernel, dim3(nb_vec), dim3(kThreads), 0, s, b_dev, w, w, N);
    dot(0, w, true);
    hipLaunchKernelGGL(gs_cycle_init_kernel, dim3(1), dim3(kThreads), 0, s, partial, nrb, st, m, hostbuf);
    if ((e = hipStreamSynchronize(s)) != hipSuccess || (e = pull(1)) != hipSuccess) {
      res.status = LSG_ERR_DEVICE;
      res.error = hipGetErrorString(e);
      break;
    }
    const double beta = hb[0];
    res.rel_residual = beta / bnorm;
    if (!std::isfinite(beta)) {
      res.status = LSG_ERR_NUMERIC;
      res.error = "residual is NaN/Inf";
      break;
    }
    if (!ref_mode && beta / bnorm <= tol) {
      converged = true;
      break;
    }
    if (cyc == max_restarts || beta == 0.0) break;
    hipLaunchKernelGGL(scale_kernel, dim3(nb_vec), dim3(kThreads), 0, s, w, st.beta, V, N);
    int k = 0;
    bool stop = false;
    double last_est = beta / bnorm, rate = 0.0;
    for (int j = 0; j < m && !stop;) {
      // Enqueue a batch of Arnoldi steps before reading their estimates back
      // (each read is a host synchronisation, about 40 us of idle device);
      // steps past an exit found inside the batch are discarded. In reference
      // mode only a breakdown ends a cycle early: batches of 16. With a
      // tolerance, half the steps the last batch's rate of decrease still
      // predicts before tol (at most 16), and 4 before a rate is known.
      int batch = 4;
      if (ref_mode) {
        batch = 16;
      } else if (rate > 0.0 && rate < 1.0 && last_est > 0.5 * tol) {
        const double left = std::log(0.5 * tol / last_est) / std::log(rate);
        batch = int(std::max(1.0, std::min(16.0, std::floor(0.5 * left))));
      } else if (rate >= 1.0) {
        batch = 1;
      }
      batch = std::min(batch, m - j);
      for (int jj = j; jj < j + batch; ++jj) {
        if (prec) {
          prec->apply(V + (size_t)jj * N, W.Z + (size_t)jj * N, s);
          matvec(W.Z + (size_t)jj * N, w);
        } else {
          matvec(V + (size_t)jj * N, w);
        }
        const int nc = jj + 1;
        dot(nc, w, false);
        hipLaunchKernelGGL(gs_reduce_kernel, dim3(1), dim3(kThreads), 0, s, partial, nrb, nc, d1);
        hipLaunchKernelGGL(gs_proj_dot_kernel, dim3(nrb), dim3(kThreads), 0, s, V, ldv, nc, d1, w, N, partial);
        hipLaunchKernelGGL(gs_tail_kernel, dim3(1), dim3(kThreads), 0, s, partial, nrb, d1, d2, st, m, jj, hostbuf);
        hipLaunchKernelGGL(gs_update_kernel, dim3(nb_vec), dim3(kThreads), 0, s, V, ldv, nc, d2, w, st.hn,
                           V + (size_t)(jj + 1) * N, N);
      }
      if ((e = hipStreamSynchronize(s)) != hipSuccess || (e = pull(2 * (j + batch))) != hipSuccess) {
        res.status = LSG_ERR_DEVICE;
        res.error = hipGetErrorString(e);
        break;
      }
      const double est_before = last_est;
      int n_read = 0;
      for (int jj = j; jj < j + batch; ++jj) {
        const double est = hb[2 * jj], hn = hb[2 * jj + 1];
        ++total_it;
        ++n_read;
        k = jj + 1;
        res.est_residual = est / bnorm;
        last_est = est / bnorm;
        if (!std::isfinite(est)) {
          res.status = LSG_ERR_NUMERIC;
          res.error = "Arnoldi produced NaN/Inf";
          stop = true;
          break;
        }
        if (std::fabs(hn) < 1e-8) {  // breakdown exit, GMRES/GMRES.cpp:284-288
          stop = true;
          break;
        }
        if (!ref_mode && est / bnorm <= 0.5 * tol) {
          stop = true;
          break;
        }
      }
      // geometric mean decrease per step over this batch (the first batch of a
      // cycle starts from the cycle's residual, 1 relative to itself)
      if (n_read > 0 && est_before > 0.0 && last_est > 0.0 && std::isfinite(last_est))
        rate = std::pow(last_est / est_before, 1.0 / n_read);
      j += batch;
    }
    if (res.status != 0) break;
    // x += V y, y = R^-1 g
    hipLaunchKernelGGL(trisolve_kernel, dim3(1), dim3(kThreads), 0, s, st, m, k, y);
    hipLaunchKernelGGL(axpy_cols_kernel, dim3(nb_vec), dim3(kThreads), 0, s, prec ? W.Z : V, ldv, k, y, x_dev, N);
    res.restarts = cyc + 1;
  }
  if (res.status == 0) {
    converged = !ref_mode ? (res.rel_residual <= tol) : true;
    if (!ref_mode && !converged) {
      res.status = LSG_ERR_NOCONV;
      res.error = "GMRES did not reach tol";
    }
  }
  res.iterations = total_it;
  res.converged = converged ? 1 : 0;
  hipStreamSynchronize(s);
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

// ||v|| by the Gram-Schmidt dot kernels (norm column only), synchronous.
static double device_norm(DeviceOperator& op, const double2* v, double2* partial, double2* out) {
  const int N = op.N;
  const int nrb = (N + kGsRows - 1) / kGsRows;
  hipLaunchKernelGGL(gs_dot_kernel, dim3(nrb, 1), dim3(kThreads), 0, op.stream, (const double2*)nullptr, (long long)N,
                     0, 1, v, N, partial);
  hipLaunchKernelGGL(gs_reduce_kernel, dim3(1), dim3(kThreads), 0, op.stream, partial, nrb, 1, out);
  double2 h;
  hipMemcpyAsync(&h, out, sizeof(double2), hipMemcpyDeviceToHost, op.stream);
  if (hipStreamSynchronize(op.stream) != hipSuccess) return std::nan("");
  return std::sqrt(std::fmax(h.x, 0.0));
}

__global__ void add_kernel(double2* __restrict__ x, const double2* __restrict__ d, int N) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n < N) {
    x[n].x += d[n].x;
    x[n].y += d[n].y;
  }
}

GmresResult device_gmres_ir(DeviceOperator& op, const double2* b_dev, double2* x_dev, int m, int max_restarts,
                            double tol, double inner_tol, int max_outer) {
  GmresResult res;
  auto t0 = std::chrono::steady_clock::now();
  const int N = op.N;
  hipStream_t s = op.stream;
  if (!(tol > 0) || !(inner_tol > 0) || max_outer < 1) {
    res.status = LSG_ERR_ARG;
    res.error = "iterative refinement needs tol > 0, inner_tol > 0 and max_outer >= 1";
    return res;
  }
  std::string e = op.enable_fp32();
  if (!e.empty()) {
    res.status = LSG_ERR_DEVICE;
    res.error = e;
    return res;
  }
  const int nrb = (N + kGsRows - 1) / kGsRows;
  const int nb_vec = (N + kThreads - 1) / kThreads;
  double2 *r = nullptr, *d = nullptr, *partial = nullptr, *out = nullptr;
  bool ok = hipMalloc((void**)&r, sizeof(double2) * size_t(N)) == hipSuccess &&
            hipMalloc((void**)&d, sizeof(double2) * size_t(N)) == hipSuccess &&
            hipMalloc((void**)&partial, sizeof(double2) * size_t(nrb)) == hipSuccess &&
            hipMalloc((void**)&out, sizeof(double2)) == hipSuccess;
  const double bnorm = ok ? device_norm(op, b_dev, partial, out) : 0.0;
  if (!ok) {
    res.status = LSG_ERR_DEVICE;
    res.error = "refinement workspace allocation failed";
  } else if (!(bnorm > 0) || !std::isfinite(bnorm)) {
    res.status = LSG_ERR_NUMERIC;
    res.error = !std::isfinite(bnorm) ? "right-hand side holds NaN/Inf" : "right-hand side is zero";
  }
  // x <- x + A32^{-1} (b - A64 x), the FP64 residual deciding convergence
  for (int outer = 0; res.status == 0; ++outer) {
    op.matvec(x_dev, r);
    hipLaunchKernelGGL(sub_kernel, dim3(nb_vec), dim3(kThreads), 0, s, b_dev, r, r, N);
    const double rn = device_norm(op, r, partial, out);
    res.rel_residual = rn / bnorm;
    if (!std::isfinite(rn)) {
      res.status = LSG_ERR_NUMERIC;
      res.error = "residual is NaN/Inf";
      break;
    }
    if (rn <= tol * bnorm) {
      res.converged = 1;
      break;
    }
    if (outer == max_outer) break;
    // inner target relative to ||r||: enough for this refinement to reach tol,
    // but not below what the FP32 operator can resolve
    const double eta = std::max(inner_tol, 0.5 * tol * bnorm / rn);
    hipMemsetAsync(d, 0, sizeof(double2) * size_t(N), s);
    GmresResult g = device_gmres(op, r, d, m, max_restarts, eta, true);
    res.iterations += g.iterations;
    res.restarts += g.restarts;
    res.est_residual = g.est_residual;
    if (g.status != 0 && g.status != LSG_ERR_NOCONV) {
      res.status = g.status;
      res.error = "inner FP32 solve: " + g.error;
      break;
    }
    hipLaunchKernelGGL(add_kernel, dim3(nb_vec), dim3(kThreads), 0, s, x_dev, d, N);
    res.refinements = outer + 1;
  }
  if (res.status == 0 && !res.converged) {
    res.status = LSG_ERR_NOCONV;
    res.error = "iterative refinement did not reach tol";
  }
  hipStreamSynchronize(s);
  if (r) hipFree(r);
  if (d) hipFree(d);
  if (partial) hipFree(partial);
  if (out) hipFree(out);
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

}  // namespace lsg
