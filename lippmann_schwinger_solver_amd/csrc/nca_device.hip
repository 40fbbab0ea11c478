// NCA skeletons and far-field operator entries on the GPU (gfx950).
//
// Four pieces, all driven by the host build in nca.cpp:
//   - the entry kernel: the Graf series of (i/4) H0 against the projected
//     source basis, with the Bessel recurrences of bessel.hpp evaluated on the
//     device (GMRES/G_FMM2DTree.hpp:603-658, getMatrixEntry);
//   - a batched partially pivoted ACA, one workgroup per skeleton, with the
//     pivot order and stopping rules of GMRES/ACA.hpp:253-481 (ACA_only_nodes);
//   - batched entry blocks K(check points, charges) for the fused M2L and L2L
//     operators (G_FMM2DTree.hpp:2502-2575, 3829-3880);
//   - the adaptive quadrature of the near-field tables' integrals
//     (NearFieldInteraction1.hpp:98-128 over GMRES/domain2D.hpp:46-138).
// The arithmetic follows the host code operation by operation (this file is
// compiled without multiply-add contraction); the device's sin, cos, atan2,
// log and hypot may differ from the host's in the last bit, so a pivot can
// differ where two candidates tie to rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

#include "nca.hpp"

namespace lsg {

namespace {

constexpr int kH = 2 * kSeries + 1;  // Graf series terms n = -15..15
constexpr int kAT = 256;             // threads per ACA workgroup (4 waves)
constexpr int kBT = 64;              // threads per entry-block workgroup
constexpr int kCapFirst = 64;        // rank capacity of the first ACA attempt
constexpr size_t kChunkBytes = size_t(4) << 30;  // device workspace per launch

#define NCA_HIP(expr)                                                             \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return std::string(#expr ": ") + hipGetErrorString(e_); \
  } while (0)

struct DevEntry {
  const double* gx;
  const double* gy;
  const double* qk2;
  const double* lcx;
  const double* lcy;
  const int* llev;
  const double2* M;         // far-field table of level l at M + moff[l], kH x rank row-major
  const long long* moff;
  int rank;
  double kappa;
  int find_phi;
};

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ inline double2 cjmul(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
// a / b by Smith's method (the libgcc complex division the host build links)
__device__ inline double2 cdiv(double2 a, double2 b) {
  if (fabs(b.x) < fabs(b.y)) {
    const double r = b.x / b.y, den = b.x * r + b.y;
    return make_double2((a.x * r + a.y) / den, (a.y * r - a.x) / den);
  }
  const double r = b.y / b.x, den = b.y * r + b.x;
  return make_double2((a.y * r + a.x) / den, (a.y - a.x * r) / den);
}
__device__ inline double cabs2(double2 a) { return hypot(a.x, a.y); }

// h[n + kSeries] = (J_n(R) + i Y_n(R)) e^{i n theta}, R = kappa |x_i - c_leaf|
// (EntryKernel::hvec)
__device__ void hvec_dev(const DevEntry& K, int i, int leaf, double2* h) {
  const double xi = K.gx[i], yi = K.gy[i];
  const double cx = K.lcx[leaf], cy = K.lcy[leaf];
  const double th = atan2(yi - cy, xi - cx);
  const double R2 = (yi - cy) * (yi - cy) + (xi - cx) * (xi - cx);
  const double R = K.kappa * sqrt(R2);
  double J[kSeries + 1], Y[kSeries + 1];
  bessel_jy(R, kSeries, J, Y);
  for (int n = -kSeries; n <= kSeries; ++n) {
    const int an = n < 0 ? -n : n;
    double jn = J[an], yn = Y[an];
    if (n < 0 && (an & 1)) {
      jn = -jn;
      yn = -yn;
    }
    const double a = double(n) * th;
    const double c = cos(a), s = sin(a);
    h[n + kSeries] = make_double2(jn * c - yn * s, jn * s + yn * c);
  }
}

// K(i, j) from the series vector of (i, leaf of j) (EntryKernel::entry)
__device__ double2 entry_dev(const DevEntry& K, const double2* h, int i, int j) {
  const int leaf = j / K.rank, idx = j - leaf * K.rank;
  const double2* Mr = K.M + K.moff[K.llev[leaf]];
  double2 s = make_double2(0.0, 0.0);
  for (int t = 0; t < kH; ++t) {
    const double2 p = cmul(Mr[t * K.rank + idx], h[t]);
    s.x += p.x;
    s.y += p.y;
  }
  s = make_double2(s.x * 0.0 - s.y * 0.25, s.x * 0.25 + s.y * 0.0);  // times i/4
  if (K.find_phi) {
    const double q = K.qk2[i];
    s.x *= q;
    s.y *= q;
    if (i == j) s.x += 1.0;
  }
  return s;
}

// ------------------------------------------------------------------ ACA ----
struct AcaDesc {
  int n1, n2, cap, nruns;
  long long rows, cols;     // node lists in idx
  long long colrun, runs;   // per column: its run of equal source leaf; per run: the leaf
  long long ws;             // double2 workspace
  long long fl;             // remaining-row and remaining-column flags (n1 + n2 bytes)
  long long piv;            // row pivots [cap] then column pivots [cap]
};

__device__ inline bool better(double ov, int oi, double bv, int bi) {
  return oi >= 0 && (bi < 0 || ov > bv || (ov == bv && oi < bi));
}

// Index of the largest |v[c]| over allowed c, the smallest index on ties
// (maxAbsVector, ACA.hpp:153-166, over an ascending set); -1 if none allowed.
__device__ int block_argmax(const double2* v, const unsigned char* allowed, int n, double* shv, int* shi) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double bv = 0.0;
  int bi = -1;
  for (int c = t; c < n; c += kAT)
    if (allowed[c]) {
      const double a = cabs2(v[c]);
      if (bi < 0 || a > bv) {
        bv = a;
        bi = c;
      }
    }
  for (int o = 32; o > 0; o >>= 1) {
    const double ov = __shfl_xor(bv, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (better(ov, oi, bv, bi)) {
      bv = ov;
      bi = oi;
    }
  }
  __syncthreads();
  if (lane == 0) {
    shv[w] = bv;
    shi[w] = bi;
  }
  __syncthreads();
  bv = shv[0];
  bi = shi[0];
  for (int q = 1; q < kAT / 64; ++q)
    if (better(shv[q], shi[q], bv, bi)) {
      bv = shv[q];
      bi = shi[q];
    }
  return bi;
}

__device__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double s = sh[0];
  for (int q = 1; q < kAT / 64; ++q) s += sh[q];
  return s;
}

__device__ double block_norm(const double2* v, int n, double* sh) {
  double a = 0.0;
  for (int c = threadIdx.x; c < n; c += kAT) a += v[c].x * v[c].x + v[c].y * v[c].y;
  return sqrt(block_sum(a, sh));
}

// One workgroup per skeleton: partially pivoted ACA of K(rows, cols).
// Workspace: row (n2), col (n1), run series (nruns x kH), U (cap x LU),
// V (cap x LV), raw rows A (cap x n2), dots (2 cap); U holds the residual
// columns and V the scaled residual rows when n1 < n2, the other way round
// otherwise, as in the reference. res[2b] = rank, res[2b+1] = 1 when the
// rank capacity was reached before the stopping rules ended the loop.
__global__ __launch_bounds__(kAT) void aca_kernel(DevEntry K, const AcaDesc* __restrict__ desc,
                                                  const int* __restrict__ idx, const int* __restrict__ runs,
                                                  double2* __restrict__ wsp, unsigned char* __restrict__ flg,
                                                  int* __restrict__ piv, int* __restrict__ res, double eps,
                                                  double abs_exit) {
  __shared__ double shv[kAT / 64];
  __shared__ int shi[kAT / 64];
  __shared__ double shs[kAT / 64];
  __shared__ double sh_norm[2];
  const AcaDesc d = desc[blockIdx.x];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int N1 = d.n1, N2 = d.n2, cap = d.cap;
  const int* rows = idx + d.rows;
  const int* cols = idx + d.cols;
  const int* colrun = runs + d.colrun;
  const int* runleaf = runs + d.runs;
  const bool rowfirst = N1 < N2;
  const int LU = rowfirst ? N1 : N2, LV = rowfirst ? N2 : N1;
  double2* row = wsp + d.ws;
  double2* col = row + N2;
  double2* hb = col + N1;
  double2* U = hb + (size_t)d.nruns * kH;
  double2* V = U + (size_t)cap * LU;
  double2* A = V + (size_t)cap * LV;
  double2* dU = A + (size_t)cap * N2;
  double2* dV = dU + cap;
  unsigned char* rrow = flg + d.fl;
  unsigned char* rcol = rrow + N1;
  int* rbase = piv + d.piv;
  int* cbase = rbase + cap;
  for (int i = t; i < N1; i += kAT) rrow[i] = 1;
  for (int i = t; i < N2; i += kAT) rcol[i] = 1;
  __syncthreads();

  // K(rows[i], cols[:]) with one series vector per run of equal source leaf.
  // Both entry routines open with a barrier: every thread has finished reading
  // the previous contents (loop conditions read row[ci] and col[ri]).
  auto get_row = [&](int i, double2* dst) {
    __syncthreads();
    for (int r = t; r < d.nruns; r += kAT) hvec_dev(K, rows[i], runleaf[r], hb + (size_t)r * kH);
    __syncthreads();
    for (int c = t; c < N2; c += kAT) dst[c] = entry_dev(K, hb + (size_t)colrun[c] * kH, rows[i], cols[c]);
    __syncthreads();
  };
  // K(rows[:], cols[j])
  auto get_col = [&](int j, double2* dst) {
    __syncthreads();
    const int leaf = cols[j] / K.rank;
    double2 h[kH];
    for (int r = t; r < N1; r += kAT) {
      hvec_dev(K, rows[r], leaf, h);
      dst[r] = entry_dev(K, h, rows[r], cols[j]);
    }
    __syncthreads();
  };
  // normS update after a rank-one term (u, v) was appended as term rank-1:
  // normS^2 + (|u||v|)^2 + sum_l 2 |<U_l, u><V_l, v>| (the sum includes the new term)
  auto norm_update = [&](double normS, int rank, double nu, double nv) {
    const double2* u = U + (size_t)(rank - 1) * LU;
    const double2* v = V + (size_t)(rank - 1) * LV;
    for (int l = wv; l < rank; l += kAT / 64) {
      double2 a = make_double2(0.0, 0.0), b = make_double2(0.0, 0.0);
      for (int s = lane; s < LU; s += 64) {
        const double2 p = cjmul(U[(size_t)l * LU + s], u[s]);
        a.x += p.x;
        a.y += p.y;
      }
      for (int s = lane; s < LV; s += 64) {
        const double2 p = cjmul(V[(size_t)l * LV + s], v[s]);
        b.x += p.x;
        b.y += p.y;
      }
      for (int o = 32; o > 0; o >>= 1) {
        a.x += __shfl_xor(a.x, o, 64);
        a.y += __shfl_xor(a.y, o, 64);
        b.x += __shfl_xor(b.x, o, 64);
        b.y += __shfl_xor(b.y, o, 64);
      }
      if (lane == 0) {
        dU[l] = a;
        dV[l] = b;
      }
    }
    __syncthreads();
    if (t == 0) {
      double s = normS * normS + (nu * nv) * (nu * nv);
      for (int l = 0; l < rank; ++l) s += 2.0 * cabs2(cmul(dU[l], dV[l]));
      sh_norm[0] = sqrt(s);
    }
    __syncthreads();
    const double r = sh_norm[0];
    __syncthreads();
    return r;
  };

  const int mn = min(N1, N2);
  int rank = 0, flag = 0;
  if (rowfirst) {
    int l_local, ci = -1;
    for (l_local = 0; l_local < N1; ++l_local) {
      get_row(l_local, row);
      ci = block_argmax(row, rcol, N2, shv, shi);
      if (cabs2(row[ci]) > eps) break;
    }
    if (l_local < N1) {
      int ri = l_local;
      const double2 pv = row[ci];
      for (int s = t; s < N2; s += kAT) {
        V[s] = cdiv(row[s], pv);
        A[s] = row[s];
      }
      get_col(ci, col);
      for (int s = t; s < N1; s += kAT) U[s] = col[s];
      if (t == 0) {
        rbase[0] = ri;
        cbase[0] = ci;
        rcol[ci] = 0;
        rrow[ri] = 0;
      }
      __syncthreads();
      rank = 1;
      double normS = 0.0, prev = 1000.0;
      ri = block_argmax(col, rrow, N1, shv, shi);
      while (ri >= 0 && cabs2(row[ci]) > eps && cabs2(col[ri]) > eps && fabs(prev - normS) >= eps * prev &&
             rank < mn) {
        if (rank == cap) {
          flag = 1;
          break;
        }
        get_row(ri, row);
        for (int s = t; s < N2; s += kAT) {
          double2 r = row[s];
          A[(size_t)rank * N2 + s] = r;
          for (int l = 0; l < rank; ++l) {
            const double2 p = cmul(U[(size_t)l * N1 + ri], V[(size_t)l * N2 + s]);
            r.x -= p.x;
            r.y -= p.y;
          }
          row[s] = r;
        }
        __syncthreads();
        ci = block_argmax(row, rcol, N2, shv, shi);
        const double2 pr = row[ci];
        for (int s = t; s < N2; s += kAT) V[(size_t)rank * N2 + s] = cdiv(row[s], pr);
        get_col(ci, col);
        for (int s = t; s < N1; s += kAT) {
          double2 c = col[s];
          for (int l = 0; l < rank; ++l) {
            const double2 p = cmul(V[(size_t)l * N2 + ci], U[(size_t)l * N1 + s]);
            c.x -= p.x;
            c.y -= p.y;
          }
          col[s] = c;
          U[(size_t)rank * N1 + s] = c;
        }
        __syncthreads();
        const double nu = block_norm(U + (size_t)rank * N1, N1, shs);
        const double nv = block_norm(V + (size_t)rank * N2, N2, shs);
        if (nu < abs_exit || nv < abs_exit) break;
        if (t == 0) {
          rbase[rank] = ri;
          cbase[rank] = ci;
          rcol[ci] = 0;
          rrow[ri] = 0;
        }
        __syncthreads();
        ++rank;
        if (rank != 2) prev = normS;
        normS = norm_update(normS, rank, nu, nv);
        ri = block_argmax(col, rrow, N1, shv, shi);
      }
    }
  } else {
    int l_local, ri = -1;
    for (l_local = 0; l_local < N2; ++l_local) {
      get_col(l_local, col);
      ri = block_argmax(col, rrow, N1, shv, shi);
      if (cabs2(col[ri]) > eps) break;
    }
    if (l_local < N2) {
      int ci = l_local;
      const double2 pv = col[ri];
      for (int s = t; s < N1; s += kAT) V[s] = cdiv(col[s], pv);
      get_row(ri, row);
      for (int s = t; s < N2; s += kAT) {
        U[s] = row[s];
        A[s] = row[s];
      }
      if (t == 0) {
        rbase[0] = ri;
        cbase[0] = ci;
        rrow[ri] = 0;
        rcol[ci] = 0;
      }
      __syncthreads();
      rank = 1;
      double normS = 0.0, prev = 1000.0;
      ci = block_argmax(row, rcol, N2, shv, shi);
      while (ci >= 0 && cabs2(row[ci]) > eps && cabs2(col[ri]) > eps && fabs(prev - normS) >= eps * prev &&
             rank < mn) {
        if (rank == cap) {
          flag = 1;
          break;
        }
        get_col(ci, col);
        for (int s = t; s < N1; s += kAT) {
          double2 c = col[s];
          for (int l = 0; l < rank; ++l) {
            const double2 p = cmul(U[(size_t)l * N2 + ci], V[(size_t)l * N1 + s]);
            c.x -= p.x;
            c.y -= p.y;
          }
          col[s] = c;
        }
        __syncthreads();
        ri = block_argmax(col, rrow, N1, shv, shi);
        const double2 pc = col[ri];
        for (int s = t; s < N1; s += kAT) V[(size_t)rank * N1 + s] = cdiv(col[s], pc);
        get_row(ri, row);
        for (int s = t; s < N2; s += kAT) {
          double2 r = row[s];
          A[(size_t)rank * N2 + s] = r;
          for (int l = 0; l < rank; ++l) {
            const double2 p = cmul(V[(size_t)l * N1 + ri], U[(size_t)l * N2 + s]);
            r.x -= p.x;
            r.y -= p.y;
          }
          row[s] = r;
          U[(size_t)rank * N2 + s] = r;
        }
        __syncthreads();
        const double nu = block_norm(U + (size_t)rank * N2, N2, shs);
        const double nv = block_norm(V + (size_t)rank * N1, N1, shs);
        if (nu < abs_exit || nv < abs_exit) break;
        if (t == 0) {
          cbase[rank] = ci;
          rbase[rank] = ri;
          rrow[ri] = 0;
          rcol[ci] = 0;
        }
        __syncthreads();
        ++rank;
        if (rank != 2) prev = normS;
        normS = norm_update(normS, rank, nu, nv);
        ci = block_argmax(row, rcol, N2, shv, shi);
      }
    }
  }
  if (t == 0) {
    res[2 * blockIdx.x] = rank;
    res[2 * blockIdx.x + 1] = flag;
  }
}

// Packs the raw pivot rows A[0, rank) of every job: out + off[b] holds rank x n2.
__global__ void aca_pack_kernel(const AcaDesc* __restrict__ desc, const int* __restrict__ res,
                                const long long* __restrict__ off, const double2* __restrict__ wsp,
                                double2* __restrict__ out) {
  const AcaDesc d = desc[blockIdx.x];
  const int rank = res[2 * blockIdx.x];
  const bool rowfirst = d.n1 < d.n2;
  const int LU = rowfirst ? d.n1 : d.n2, LV = rowfirst ? d.n2 : d.n1;
  const double2* A = wsp + d.ws + d.n2 + d.n1 + (size_t)d.nruns * kH + (size_t)d.cap * (LU + LV);
  const long long n = (long long)rank * d.n2;
  for (long long e = threadIdx.x; e < n; e += blockDim.x) out[off[blockIdx.x] + e] = A[e];
}

// ------------------------------------------------------------- blocks ----
struct BlockDesc {
  int nr, nc;
  long long rows, cols, out;
};

// out (column-major nr x nc) = K(rows, cols); one series vector per row and
// run of equal source leaf (EntryKernel::block).
__global__ __launch_bounds__(kBT) void block_kernel(DevEntry K, const BlockDesc* __restrict__ desc,
                                                    const int* __restrict__ idx, double2* __restrict__ out) {
  const BlockDesc d = desc[blockIdx.x];
  const int* rows = idx + d.rows;
  const int* cols = idx + d.cols;
  double2* o = out + d.out;
  double2 h[kH];
  for (int r = threadIdx.x; r < d.nr; r += kBT) {
    int prev = -1;
    const int i = rows[r];
    for (int c = 0; c < d.nc; ++c) {
      const int j = cols[c], leaf = j / K.rank;
      if (leaf != prev) {
        hvec_dev(K, i, leaf, h);
        prev = leaf;
      }
      o[(size_t)c * d.nr + r] = entry_dev(K, h, i, j);
    }
  }
}

// ------------------------------------------- near-field table quadrature ---
// V1(i, l) = int_[-1,1]^2 H0(kappa rB |X_i - P|) T_a(P_x) T_b(P_y) dP by the
// adaptive tensor Gauss-Legendre-10 rule of adaptive_quad (tables.cpp, after
// GMRES/domain2D.hpp:46-138): a panel is accepted when its own value differs
// from the sum of its four children's by less than eps, accepted panels are
// summed in depth-first order (children SW, SE, NE, NW). One wavefront per
// integral: the 400 points of a refinement are spread over the lanes and
// summed by shuffles; the depth-first panel stack lives in LDS (every lane
// writes the same, wave-uniform values). A panel is also accepted at depth
// kQMaxDepth or after kQMaxSteps refinements, so every wavefront finishes;
// such a cap is reported and fails the build.
constexpr int kQT = 256;                 // threads per quadrature workgroup (4 integrals)
constexpr int kQMaxDepth = 50;
constexpr int kQStack = 3 * kQMaxDepth + 4;
constexpr int kQMaxSteps = 20000;       // a log-singular integral needs a few hundred

struct QuadJob {
  double Xx, Xy, krb;  // target node in source-box units, kappa rB
  int a, b;            // T_a(P_x) T_b(P_y)
};

__device__ __forceinline__ double q_cheb(double x, int n) { return cos(n * acos(x)); }

__device__ __forceinline__ void q_f(const QuadJob& q, double px, double py, double& re, double& im) {
  const double R = sqrt((q.Xx - px) * (q.Xx - px) + (q.Xy - py) * (q.Xy - py));
  double J[2], Y[2];
  bessel_jy(q.krb * R, 0, J, Y);
  const double ta = q_cheb(px, q.a), tb = q_cheb(py, q.b);
  re = J[0] * ta * tb;
  im = Y[0] * ta * tb;
}

__device__ __forceinline__ double q_wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(kQT) void quad_kernel(const QuadJob* __restrict__ jobs, int njobs,
                                                   const double* __restrict__ glx, const double* __restrict__ glw,
                                                   double eps, double2* __restrict__ out, int* __restrict__ capped) {
  __shared__ double stk[kQT / 64][kQStack][6];  // cx, cy, half width, re, im, depth
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int jid = blockIdx.x * (kQT / 64) + wv;
  if (jid >= njobs) return;   // whole wavefront: no workgroup barrier follows
  const QuadJob q = jobs[jid];
  double (*st)[6] = stk[wv];
  // the root panel [-1,1]^2
  double rr = 0.0, ri = 0.0;
  for (int p = lane; p < 100; p += 64) {
    const int m = p / 10, n = p % 10;
    double fr, fi;
    q_f(q, glx[m], glx[n], fr, fi);
    const double w = glw[m] * glw[n];
    rr += w * fr;
    ri += w * fi;
  }
  rr = q_wsum(rr);
  ri = q_wsum(ri);
  int sp = 0;
  st[0][0] = 0.0; st[0][1] = 0.0; st[0][2] = 1.0; st[0][3] = rr; st[0][4] = ri; st[0][5] = 0.0;
  sp = 1;
  double ar = 0.0, ai = 0.0;
  int steps = 0, cap = 0;
  while (sp > 0) {
    --sp;
    const double cx = st[sp][0], cy = st[sp][1], h = st[sp][2], pr = st[sp][3], pi = st[sp][4];
    const int d = int(st[sp][5]);
    if (d >= kQMaxDepth || steps >= kQMaxSteps) {
      ar += pr;
      ai += pi;
      cap = 1;
      continue;
    }
    ++steps;
    const double r = 0.5 * h;
    double cr[4] = {0.0, 0.0, 0.0, 0.0}, ci[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const double ccx = (c == 1 || c == 2) ? cx + r : cx - r;
      const double ccy = (c >= 2) ? cy + r : cy - r;
      double sr = 0.0, si = 0.0;
      for (int p = lane; p < 100; p += 64) {
        const int m = p / 10, n = p % 10;
        double fr, fi;
        q_f(q, ccx + r * glx[m], ccy + r * glx[n], fr, fi);
        const double w = glw[m] * glw[n];
        sr += w * fr;
        si += w * fi;
      }
      cr[c] = q_wsum(sr) * (r * r);
      ci[c] = q_wsum(si) * (r * r);
    }
    double er = pr, ei = pi;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      er -= cr[c];
      ei -= ci[c];
    }
    if (hypot(er, ei) < eps) {
      ar += pr;
      ai += pi;
    } else {
      // children pushed NW, NE, SE, SW so that SW is refined first
#pragma unroll
      for (int c = 3; c >= 0; --c) {
        st[sp][0] = (c == 1 || c == 2) ? cx + r : cx - r;
        st[sp][1] = (c >= 2) ? cy + r : cy - r;
        st[sp][2] = r;
        st[sp][3] = cr[c];
        st[sp][4] = ci[c];
        st[sp][5] = double(d + 1);
        ++sp;
      }
    }
  }
  if (lane == 0) {
    out[jid] = make_double2(ar, ai);
    capped[jid] = cap;
  }
}

// ---------------------------------------------------------------------------
// Batched column-pivoted Householder QR and basic solve (ColPivQR of the host
// build, step by step: the pivot is the first largest downdated column norm,
// pivots below eps max_norm sqrt(n - k) end the nonzero block, H_k = I - tau
// v v^H with v = [1; essential], Q^H b applies H_0, H_1, ... with tau itself,
// back substitution on the nonzero block, the rest zero). One wavefront per
// system: the matrix lives in LDS; thread j owns column j in the trailing
// update and the norm downdates, thread 0 the sequential reductions, so every
// sum is formed in the host's order. Thread r solves right-hand sides r,
// r + 64, ... in the global workspace W.
constexpr int kQrMax = 60;     // largest system solved on the device (A in 64 KB of LDS)
struct QrDesc {
  int n, nrhs;
  long long a, b;   // offsets of A (n x n) and of B / X / W (n x nrhs), complex
};

__global__ __launch_bounds__(64) void qr_solve_kernel(const QrDesc* __restrict__ desc, const double2* __restrict__ Aall,
                                                     const double2* __restrict__ Ball, double2* __restrict__ Xall,
                                                     double2* __restrict__ Wall) {
  extern __shared__ double2 qr[];   // n x n, column-major
  __shared__ double nu[kQrMax], nd[kQrMax];
  __shared__ double2 tau[kQrMax];
  __shared__ int transp[kQrMax], perm[kQrMax];
  __shared__ int nonzero_s, big_s, zero_s;
  __shared__ double2 d_s, t_s;
  __shared__ double beta_s;
  const QrDesc D = desc[blockIdx.x];
  const int n = D.n, t = threadIdx.x;
  const double2* A = Aall + D.a;
  for (int e = t; e < n * n; e += 64) qr[e] = A[e];
  __syncthreads();
  auto at = [&](int i, int j) -> double2& { return qr[j * n + i]; };
  if (t < n) {
    double sq = 0.0;
    for (int i = 0; i < n; ++i) sq += at(i, t).x * at(i, t).x + at(i, t).y * at(i, t).y;
    nd[t] = nu[t] = sqrt(sq);
  }
  __syncthreads();
  const double eps = 2.220446049250313e-16;
  double maxn = 0.0;
  for (int k = 0; k < n; ++k) maxn = fmax(maxn, nu[k]);
  const double thr_helper = (maxn * eps) * (maxn * eps) / double(n);
  const double down_thr = sqrt(eps);
  if (t == 0) nonzero_s = n;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    if (t == 0) {
      int big = k;
      double bv = nu[k];
      for (int j = k + 1; j < n; ++j)
        if (nu[j] > bv) {
          bv = nu[j];
          big = j;
        }
      if (nonzero_s == n && bv * bv < thr_helper * double(n - k)) nonzero_s = k;
      transp[k] = big;
      big_s = big;
    }
    __syncthreads();
    const int big = big_s;
    if (big != k) {
      for (int i = t; i < n; i += 64) {
        const double2 a = at(i, k);
        at(i, k) = at(i, big);
        at(i, big) = a;
      }
      if (t == 0) {
        double a = nu[k]; nu[k] = nu[big]; nu[big] = a;
        a = nd[k]; nd[k] = nd[big]; nd[big] = a;
      }
    }
    __syncthreads();
    if (t == 0) {
      const double2 c0 = at(k, k);
      double tail = 0.0;
      for (int i = k + 1; i < n; ++i) tail += at(i, k).x * at(i, k).x + at(i, k).y * at(i, k).y;
      const double tiny = 2.2250738585072014e-308;
      if (tail <= tiny && c0.y * c0.y <= tiny) {
        zero_s = 1;
        t_s = make_double2(0.0, 0.0);
        beta_s = c0.x;
      } else {
        zero_s = 0;
        double beta = sqrt(c0.x * c0.x + c0.y * c0.y + tail);
        if (c0.x >= 0) beta = -beta;
        d_s = make_double2(c0.x - beta, c0.y);
        const double2 q = make_double2((beta - c0.x) / beta, -c0.y / beta);
        t_s = make_double2(q.x, -q.y);   // conj((beta - c0) / beta)
        beta_s = beta;
      }
    }
    __syncthreads();
    for (int i = k + 1 + t; i < n; i += 64) at(i, k) = zero_s ? make_double2(0.0, 0.0) : cdiv(at(i, k), d_s);
    __syncthreads();
    const double2 tk = t_s;
    if (t == 0) {
      tau[k] = tk;
      at(k, k) = make_double2(beta_s, 0.0);
    }
    __syncthreads();
    // H_k applied to the trailing columns, one column per thread
    const int j = k + 1 + t;
    if (j < n) {
      if (n - k == 1) {
        at(k, j) = cmul(at(k, j), make_double2(1.0 - tk.x, -tk.y));
      } else if (tk.x != 0.0 || tk.y != 0.0) {
        double2 sv = at(k, j);
        for (int i = k + 1; i < n; ++i) {
          const double2 p = cjmul(at(i, k), at(i, j));
          sv.x += p.x;
          sv.y += p.y;
        }
        const double2 ts = cmul(tk, sv);
        at(k, j).x -= ts.x;
        at(k, j).y -= ts.y;
        for (int i = k + 1; i < n; ++i) {
          const double2 p = cmul(at(i, k), ts);
          at(i, j).x -= p.x;
          at(i, j).y -= p.y;
        }
      }
      if (nu[j] != 0.0) {
        double tt = cabs2(at(k, j)) / nu[j];
        tt = (1.0 + tt) * (1.0 - tt);
        if (tt < 0) tt = 0;
        const double r = nu[j] / nd[j];
        const double t2 = tt * r * r;
        if (t2 <= down_thr) {
          double sq = 0.0;
          for (int i = k + 1; i < n; ++i) sq += at(i, j).x * at(i, j).x + at(i, j).y * at(i, j).y;
          nd[j] = nu[j] = sqrt(sq);
        } else {
          nu[j] *= sqrt(tt);
        }
      }
    }
    __syncthreads();
  }
  if (t == 0) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
      const int a = perm[k];
      perm[k] = perm[transp[k]];
      perm[transp[k]] = a;
    }
  }
  __syncthreads();
  const int nz = nonzero_s;
  for (int r = t; r < D.nrhs; r += 64) {
    double2* c = Wall + D.b + (long long)r * n;
    const double2* bcol = Ball + D.b + (long long)r * n;
    double2* xcol = Xall + D.b + (long long)r * n;
    for (int i = 0; i < n; ++i) {
      c[i] = bcol[i];
      xcol[i] = make_double2(0.0, 0.0);
    }
    if (nz == 0) continue;
    for (int k = 0; k < nz; ++k) {
      const double2 tk = tau[k];
      if (n - k == 1) {
        c[k] = cmul(c[k], make_double2(1.0 - tk.x, -tk.y));
        continue;
      }
      if (tk.x == 0.0 && tk.y == 0.0) continue;
      double2 sv = c[k];
      for (int i = k + 1; i < n; ++i) {
        const double2 p = cjmul(at(i, k), c[i]);
        sv.x += p.x;
        sv.y += p.y;
      }
      const double2 ts = cmul(tk, sv);
      c[k].x -= ts.x;
      c[k].y -= ts.y;
      for (int i = k + 1; i < n; ++i) {
        const double2 p = cmul(at(i, k), ts);
        c[i].x -= p.x;
        c[i].y -= p.y;
      }
    }
    for (int i = nz - 1; i >= 0; --i) {
      double2 sv = c[i];
      for (int jj = i + 1; jj < nz; ++jj) {
        const double2 p = cmul(at(i, jj), c[jj]);
        sv.x -= p.x;
        sv.y -= p.y;
      }
      c[i] = cdiv(sv, at(i, i));
    }
    for (int i = 0; i < nz; ++i) xcol[perm[i]] = c[i];
  }
}

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  std::string reserve(size_t m) {
    if (m <= n) return "";
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    NCA_HIP(hipMalloc((void**)&p, std::max<size_t>(m, 1) * sizeof(T)));
    n = m;
    return "";
  }
  std::string put(const std::vector<T>& h) {
    std::string e = reserve(h.size());
    if (!e.empty()) return e;
    if (!h.empty()) NCA_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return "";
  }
};

class HipNca : public DeviceNca {
 public:
  std::string init(const EntryKernel& K) {
    const Tree& t = *K.tree;
    rank_ = K.rank;
    std::vector<double2> M;
    std::vector<long long> moff(K.M.size(), 0);
    for (size_t l = 0; l < K.M.size(); ++l) {
      moff[l] = (long long)M.size();
      for (const cplx& z : K.M[l]) M.push_back(make_double2(z.real(), z.imag()));
    }
    std::string e;
    if (!(e = gx_.put(t.gx)).empty() || !(e = gy_.put(t.gy)).empty() || !(e = qk2_.put(K.qk2)).empty() ||
        !(e = lcx_.put(K.leaf_cx)).empty() || !(e = lcy_.put(K.leaf_cy)).empty() ||
        !(e = llev_.put(K.leaf_level)).empty() || !(e = M_.put(M)).empty() || !(e = moff_.put(moff)).empty())
      return e;
    dk_.gx = gx_.p;
    dk_.gy = gy_.p;
    dk_.qk2 = qk2_.p;
    dk_.lcx = lcx_.p;
    dk_.lcy = lcy_.p;
    dk_.llev = llev_.p;
    dk_.M = M_.p;
    dk_.moff = moff_.p;
    dk_.rank = K.rank;
    dk_.kappa = K.kappa;
    dk_.find_phi = K.find_phi ? 1 : 0;
    return "";
  }

  std::string aca(const std::vector<AcaJob>& jobs, double eps, double abs_exit) override {
    std::vector<int> caps(jobs.size());
    for (size_t q = 0; q < jobs.size(); ++q)
      caps[q] = std::min<int>(kCapFirst, int(std::min(jobs[q].rows->size(), jobs[q].cols->size())));
    std::vector<size_t> todo(jobs.size());
    for (size_t q = 0; q < jobs.size(); ++q) todo[q] = q;
    // first attempt at the default capacity; skeletons that reach it are redone
    // with capacity min(n1, n2)
    for (int attempt = 0; attempt < 2 && !todo.empty(); ++attempt) {
      std::vector<size_t> again;
      size_t q0 = 0;
      while (q0 < todo.size()) {
        size_t q1 = q0, bytes = 0;
        while (q1 < todo.size()) {
          const size_t b = ws_elems(jobs[todo[q1]], caps[todo[q1]]) * sizeof(double2);
          if (q1 > q0 && bytes + b > kChunkBytes) break;
          bytes += b;
          ++q1;
        }
        std::string e = aca_chunk(jobs, todo, q0, q1, caps, eps, abs_exit, again);
        if (!e.empty()) return e;
        q0 = q1;
      }
      for (size_t q : again) caps[q] = int(std::min(jobs[q].rows->size(), jobs[q].cols->size()));
      todo.swap(again);
    }
    if (!todo.empty()) return "ACA rank capacity exceeded";
    return "";
  }

  std::string near_integrals(const BasisSet& bs, double kappa, const std::vector<NearGeom>& tabs,
                             std::vector<std::vector<cplx>>& V1) override {
    V1.assign(tabs.size(), std::vector<cplx>());
    if (tabs.empty()) return "";
    std::vector<QuadJob> jobs;
    jobs.reserve(tabs.size() * size_t(bs.rank) * bs.nb);
    for (const NearGeom& g : tabs)
      for (int i = 0; i < bs.rank; ++i)
        for (int l = 0; l < bs.nb; ++l) {
          QuadJob q;
          q.Xx = (g.rA * bs.sx[i] + 0.0 - g.bx) / g.rB;
          q.Xy = (g.rA * bs.sy[i] + 0.0 - g.by) / g.rB;
          q.krb = kappa * g.rB;
          q.a = bs.ab[l][0];
          q.b = bs.ab[l][1];
          jobs.push_back(q);
        }
    std::vector<double> gx(kQuadNodes), gw(kQuadNodes);
    gauss_legendre(kQuadNodes, gx.data(), gw.data());
    DevBuf<QuadJob> dj;
    DevBuf<double> dx, dw;
    DevBuf<double2> dout;
    DevBuf<int> dcap;
    std::string e;
    if (!(e = dj.put(jobs)).empty() || !(e = dx.put(gx)).empty() || !(e = dw.put(gw)).empty() ||
        !(e = dout.reserve(jobs.size())).empty() || !(e = dcap.reserve(jobs.size())).empty())
      return e;
    const int n = int(jobs.size());
    hipLaunchKernelGGL(quad_kernel, dim3(unsigned((n + kQT / 64 - 1) / (kQT / 64))), dim3(kQT), 0, 0, dj.p, n, dx.p,
                       dw.p, kQuadEps, dout.p, dcap.p);
    NCA_HIP(hipGetLastError());
    std::vector<double2> h(jobs.size());
    std::vector<int> cap(jobs.size());
    NCA_HIP(hipMemcpy(h.data(), dout.p, h.size() * sizeof(double2), hipMemcpyDeviceToHost));
    NCA_HIP(hipMemcpy(cap.data(), dcap.p, cap.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int c : cap)
      if (c) return "table quadrature reached its depth or step limit";
    const size_t per = size_t(bs.rank) * bs.nb;
    for (size_t t = 0; t < tabs.size(); ++t) {
      V1[t].resize(per);
      for (size_t k = 0; k < per; ++k) V1[t][k] = cplx(h[t * per + k].x, h[t * per + k].y);
    }
    return "";
  }

  std::string qr_solve(const std::vector<QrSolveJob>& jobs) override {
    // systems above the device limit: the host ColPivQR
    std::vector<size_t> host_jobs, dev_jobs;
    for (size_t q = 0; q < jobs.size(); ++q) (jobs[q].n > kQrMax || jobs[q].n < 1 ? host_jobs : dev_jobs).push_back(q);
#pragma omp parallel for schedule(dynamic, 4)
    for (int h = 0; h < int(host_jobs.size()); ++h) {
      const QrSolveJob& j = jobs[host_jobs[size_t(h)]];
      if (j.n < 1) continue;
      ColPivQR qr;
      qr.factor(j.A, j.n, j.n);
      qr.solve(j.B, j.nrhs, j.X);
    }
    size_t q0 = 0;
    while (q0 < dev_jobs.size()) {
      size_t q1 = q0, ea = 0, eb = 0;
      while (q1 < dev_jobs.size()) {
        const QrSolveJob& j = jobs[dev_jobs[q1]];
        const size_t na = size_t(j.n) * j.n, nb = size_t(j.n) * j.nrhs;
        if (q1 > q0 && (ea + na + 3 * (eb + nb)) * sizeof(double2) > kChunkBytes / 4) break;
        ea += na;
        eb += nb;
        ++q1;
      }
      std::vector<QrDesc> dd(q1 - q0);
      std::vector<double2> Ah(ea), Bh(eb);
      long long oa = 0, ob = 0;
      for (size_t q = q0; q < q1; ++q) {
        const QrSolveJob& j = jobs[dev_jobs[q]];
        dd[q - q0] = QrDesc{j.n, j.nrhs, oa, ob};
        std::memcpy(Ah.data() + oa, j.A, sizeof(double2) * size_t(j.n) * j.n);
        std::memcpy(Bh.data() + ob, j.B, sizeof(double2) * size_t(j.n) * j.nrhs);
        oa += (long long)j.n * j.n;
        ob += (long long)j.n * j.nrhs;
      }
      std::string e;
      if (!(e = qdesc_.put(dd)).empty() || !(e = qa_.put(Ah)).empty() || !(e = qb_.put(Bh)).empty() ||
          !(e = qx_.reserve(eb)).empty() || !(e = qw_.reserve(eb)).empty())
        return e;
      int nmax = 1;
      for (const QrDesc& d : dd) nmax = std::max(nmax, d.n);
      hipLaunchKernelGGL(qr_solve_kernel, dim3(unsigned(dd.size())), dim3(64), sizeof(double2) * nmax * nmax, 0,
                         qdesc_.p, qa_.p, qb_.p, qx_.p, qw_.p);
      NCA_HIP(hipGetLastError());
      NCA_HIP(hipMemcpy(Bh.data(), qx_.p, sizeof(double2) * eb, hipMemcpyDeviceToHost));
      for (size_t q = q0; q < q1; ++q) {
        const QrSolveJob& j = jobs[dev_jobs[q]];
        std::memcpy(j.X, Bh.data() + dd[q - q0].b, sizeof(double2) * size_t(j.n) * j.nrhs);
      }
      q0 = q1;
    }
    return "";
  }

  std::string blocks(const std::vector<BlockJob>& jobs) override {
    size_t q0 = 0;
    while (q0 < jobs.size()) {
      size_t q1 = q0, elems = 0;
      while (q1 < jobs.size()) {
        const size_t n = size_t(jobs[q1].nr) * jobs[q1].nc;
        if (q1 > q0 && (elems + n) * sizeof(double2) > kChunkBytes / 4) break;
        elems += n;
        ++q1;
      }
      std::vector<BlockDesc> dd;
      std::vector<int> ids;
      dd.reserve(q1 - q0);
      long long off = 0;
      for (size_t q = q0; q < q1; ++q) {
        const BlockJob& j = jobs[q];
        BlockDesc b;
        b.nr = j.nr;
        b.nc = j.nc;
        b.rows = (long long)ids.size();
        ids.insert(ids.end(), j.rows, j.rows + j.nr);
        b.cols = (long long)ids.size();
        ids.insert(ids.end(), j.cols, j.cols + j.nc);
        b.out = off;
        off += (long long)j.nr * j.nc;
        dd.push_back(b);
      }
      std::string e;
      if (!(e = bdesc_.put(dd)).empty() || !(e = idx_.put(ids)).empty() || !(e = out_.reserve(size_t(off))).empty())
        return e;
      if (!dd.empty()) {
        hipLaunchKernelGGL(block_kernel, dim3(unsigned(dd.size())), dim3(kBT), 0, 0, dk_, bdesc_.p, idx_.p, out_.p);
        NCA_HIP(hipGetLastError());
      }
      // blocks that follow each other in the caller's memory come back in one copy
      bool contiguous = true;
      for (size_t q = q0 + 1; q < q1 && contiguous; ++q)
        contiguous = jobs[q].out == jobs[q - 1].out + size_t(jobs[q - 1].nr) * jobs[q - 1].nc;
      if (contiguous) {
        if (off) NCA_HIP(hipMemcpy(reinterpret_cast<void*>(jobs[q0].out), out_.p, size_t(off) * sizeof(double2),
                                   hipMemcpyDeviceToHost));
      } else {
        host_.resize(size_t(off));
        if (off) NCA_HIP(hipMemcpy(host_.data(), out_.p, size_t(off) * sizeof(double2), hipMemcpyDeviceToHost));
        for (size_t q = q0; q < q1; ++q) {
          const BlockJob& j = jobs[q];
          std::memcpy(reinterpret_cast<void*>(j.out), host_.data() + dd[q - q0].out,
                      size_t(j.nr) * j.nc * sizeof(double2));
        }
      }
      q0 = q1;
    }
    return "";
  }

 private:
  static size_t ws_elems(const AcaJob& j, int cap) {
    const size_t n1 = j.rows->size(), n2 = j.cols->size();
    return n1 + n2 + size_t(cap) * (n1 + 2 * n2 + 2) + kH * n2;  // runs <= n2
  }

  std::string aca_chunk(const std::vector<AcaJob>& jobs, const std::vector<size_t>& todo, size_t q0, size_t q1,
                        const std::vector<int>& caps, double eps, double abs_exit, std::vector<size_t>& again) {
    const size_t nj = q1 - q0;
    std::vector<AcaDesc> dd(nj);
    std::vector<int> ids, runs;
    long long ws = 0, fl = 0, pv = 0;
    for (size_t q = 0; q < nj; ++q) {
      const AcaJob& j = jobs[todo[q0 + q]];
      AcaDesc& d = dd[q];
      d.n1 = int(j.rows->size());
      d.n2 = int(j.cols->size());
      d.cap = caps[todo[q0 + q]];
      d.rows = (long long)ids.size();
      ids.insert(ids.end(), j.rows->begin(), j.rows->end());
      d.cols = (long long)ids.size();
      ids.insert(ids.end(), j.cols->begin(), j.cols->end());
      // runs of consecutive columns with the same source leaf
      d.colrun = (long long)runs.size();
      std::vector<int> leaves;
      int prev = -1;
      for (int c = 0; c < d.n2; ++c) {
        const int leaf = (*j.cols)[size_t(c)] / rank_;
        if (leaf != prev) {
          leaves.push_back(leaf);
          prev = leaf;
        }
        runs.push_back(int(leaves.size()) - 1);
      }
      d.nruns = int(leaves.size());
      d.runs = (long long)runs.size();
      runs.insert(runs.end(), leaves.begin(), leaves.end());
      d.ws = ws;
      const int LU = d.n1 < d.n2 ? d.n1 : d.n2, LV = d.n1 < d.n2 ? d.n2 : d.n1;
      ws += d.n2 + d.n1 + (long long)d.nruns * kH + (long long)d.cap * (LU + LV + d.n2) + 2LL * d.cap;
      d.fl = fl;
      fl += d.n1 + d.n2;
      d.piv = pv;
      pv += 2LL * d.cap;
    }
    std::string e;
    if (!(e = adesc_.put(dd)).empty() || !(e = idx_.put(ids)).empty() || !(e = runs_.put(runs)).empty() ||
        !(e = ws_.reserve(size_t(ws))).empty() || !(e = fl_.reserve(size_t(fl))).empty() ||
        !(e = piv_.reserve(size_t(pv))).empty() || !(e = res_.reserve(2 * nj)).empty())
      return e;
    hipLaunchKernelGGL(aca_kernel, dim3(unsigned(nj)), dim3(kAT), 0, 0, dk_, adesc_.p, idx_.p, runs_.p, ws_.p,
                       fl_.p, piv_.p, res_.p, eps, abs_exit);
    NCA_HIP(hipGetLastError());
    std::vector<int> res(2 * nj), pivh(static_cast<size_t>(pv));
    NCA_HIP(hipMemcpy(res.data(), res_.p, res.size() * sizeof(int), hipMemcpyDeviceToHost));
    NCA_HIP(hipMemcpy(pivh.data(), piv_.p, pivh.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<long long> off(nj);
    long long tot = 0;
    for (size_t q = 0; q < nj; ++q) {
      off[q] = tot;
      tot += (long long)res[2 * q] * dd[q].n2;
    }
    if (!(e = off_.put(off)).empty() || !(e = out_.reserve(size_t(tot))).empty()) return e;
    hipLaunchKernelGGL(aca_pack_kernel, dim3(unsigned(nj)), dim3(kAT), 0, 0, adesc_.p, res_.p, off_.p, ws_.p,
                       out_.p);
    NCA_HIP(hipGetLastError());
    host_.resize(size_t(tot));
    if (tot) NCA_HIP(hipMemcpy(host_.data(), out_.p, size_t(tot) * sizeof(double2), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < nj; ++q) {
      if (res[2 * q + 1]) {
        again.push_back(todo[q0 + q]);
        continue;
      }
      AcaResult& r = *jobs[todo[q0 + q]].out;
      const int rk = res[2 * q];
      r.rank = rk;
      r.row_bases.assign(pivh.begin() + dd[q].piv, pivh.begin() + dd[q].piv + rk);
      r.col_bases.assign(pivh.begin() + dd[q].piv + dd[q].cap, pivh.begin() + dd[q].piv + dd[q].cap + rk);
      r.Ar.resize(size_t(rk) * dd[q].n2);
      std::memcpy(reinterpret_cast<void*>(r.Ar.data()), host_.data() + off[q], r.Ar.size() * sizeof(double2));
    }
    return "";
  }

  int rank_ = 0;
  DevEntry dk_{};
  DevBuf<double> gx_, gy_, qk2_, lcx_, lcy_;
  DevBuf<int> llev_;
  DevBuf<double2> M_;
  DevBuf<long long> moff_;
  DevBuf<AcaDesc> adesc_;
  DevBuf<BlockDesc> bdesc_;
  DevBuf<int> idx_, runs_, piv_, res_;
  DevBuf<unsigned char> fl_;
  DevBuf<double2> ws_, out_;
  DevBuf<long long> off_;
  DevBuf<QrDesc> qdesc_;
  DevBuf<double2> qa_, qb_, qx_, qw_;
  std::vector<double2> host_;
};

}  // namespace

DeviceNca* make_device_nca(const EntryKernel& K, std::string& err) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    err = "no GPU";
    return nullptr;
  }
  std::unique_ptr<HipNca> d(new HipNca());
  err = d->init(K);
  if (!err.empty()) return nullptr;
  return d.release();
}

}  // namespace lsg
