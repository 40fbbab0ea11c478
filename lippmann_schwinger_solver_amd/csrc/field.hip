// Re-interpolation of a leaf-node field onto arbitrary points (gfx950).
//
// Reference: GMRES_Hybrid/LippmannSchwinger/main/hybridTrueValidation_gmres.cpp:186-209
// (Phi at the nodes of one tree -> Psi at the nodes of another): per leaf the
// coefficients c_l = sum_r Q_pinv(l, r) f(r) of the total-degree Chebyshev
// basis T_a(x) T_b(y), a + b < degreeOfBases, evaluated at each point of the
// leaf box [c - r, c + r) that contains it, with T_n(x) = cos(n acos x)
// (get_ChebPoly, G_FMM2DTree.hpp:670-672).
#include "field.hpp"

#include <cmath>

namespace lsg {

namespace {

constexpr int kFT = 256;

// c[leaf * nb + l] = sum_r qpinv[l * rank + r] f[leaf * rank + r]
__global__ __launch_bounds__(kFT) void interp_coef_kernel(const double* __restrict__ qpinv,
                                                          const double2* __restrict__ f, int nleaves, int rank,
                                                          int nb, double2* __restrict__ c) {
  const long long e = (long long)blockIdx.x * kFT + threadIdx.x;
  if (e >= (long long)nleaves * nb) return;
  const int leaf = int(e / nb), l = int(e - (long long)leaf * nb);
  const double2* fl = f + (size_t)leaf * rank;
  const double* q = qpinv + (size_t)l * rank;
  double2 s = make_double2(0.0, 0.0);
  for (int r = 0; r < rank; ++r) {
    s.x += q[r] * fl[r].x;
    s.y += q[r] * fl[r].y;
  }
  c[e] = s;
}

// Index of the leaf box along one axis holding v, tested as the reference
// tests it (centre - r <= v < centre + r); -1 outside the domain.
__device__ inline int locate(double v, const double* __restrict__ centre, int side, double r) {
  if (!(v >= centre[0] - r && v < centre[side - 1] + r)) return -1;  // outside, or NaN
  int k = int(floor((v - (centre[0] - r)) / (2.0 * r)));
  k = k < 0 ? 0 : (k >= side ? side - 1 : k);
  if (v < centre[k] - r && k > 0) --k;
  else if (!(v < centre[k] + r) && k < side - 1) ++k;
  return (centre[k] - r <= v && v < centre[k] + r) ? k : -1;
}

__global__ __launch_bounds__(kFT) void interp_eval_kernel(const double2* __restrict__ c, const int* __restrict__ ab,
                                                          const int* __restrict__ leaf_at,
                                                          const double* __restrict__ cx, const double* __restrict__ cy,
                                                          int side, double r, int nb, const double* __restrict__ x,
                                                          const double* __restrict__ y, long long n,
                                                          double2* __restrict__ out, int* __restrict__ bad) {
  const long long i = (long long)blockIdx.x * kFT + threadIdx.x;
  if (i >= n) return;
  const int ix = locate(x[i], cx, side, r), iy = locate(y[i], cy, side, r);
  if (ix < 0 || iy < 0) {
    out[i] = make_double2(0.0, 0.0);
    atomicOr(bad, 1);
    return;
  }
  const int leaf = leaf_at[iy * side + ix];
  const double px = (x[i] - cx[ix]) / r, py = (y[i] - cy[iy]) / r;
  const double ax = acos(px), ay = acos(py);
  const double2* cl = c + (size_t)leaf * nb;
  double2 s = make_double2(0.0, 0.0);
  for (int l = 0; l < nb; ++l) {
    const double ta = cos(ab[2 * l] * ax), tb = cos(ab[2 * l + 1] * ay);
    s.x += (cl[l].x * ta) * tb;
    s.y += (cl[l].y * ta) * tb;
  }
  out[i] = s;
}

}  // namespace

#define FIELD_HIP(expr)                                                           \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return std::string(#expr ": ") + hipGetErrorString(e_); \
  } while (0)

std::string FieldInterp::setup(const Tree& t, const BasisSet& bs) {
  if (qpinv) return "";
  const int lvl = t.nLevels;
  side = 1 << lvl;
  nleaves = int(t.leaves.size());
  if (nleaves != side * side) return "re-interpolation needs a uniform tree";
  rank = bs.rank;
  nb = bs.nb;
  r = t.boxRadius[size_t(lvl)];
  // leaf centres per axis and the leaf index of each (column, row) box
  std::vector<double> cxs(static_cast<size_t>(side)), cys(static_cast<size_t>(side));
  std::vector<int> at(size_t(side) * side, -1);
  for (int k = 0; k < side; ++k) {
    cxs[size_t(k)] = -t.p.L + r * (2 * k + 1);
    cys[size_t(k)] = cxs[size_t(k)];
  }
  for (int li = 0; li < nleaves; ++li) {
    const Box& b = t.box(t.leaves[size_t(li)]);
    const int ix = int(std::lround((b.cx + t.p.L) / (2 * r) - 0.5)), iy = int(std::lround((b.cy + t.p.L) / (2 * r) - 0.5));
    if (ix < 0 || ix >= side || iy < 0 || iy >= side) return "leaf outside the domain";
    at[size_t(iy) * side + ix] = li;
    cxs[size_t(ix)] = b.cx;
    cys[size_t(iy)] = b.cy;
  }
  std::vector<int> abv(size_t(2) * nb);
  for (int l = 0; l < nb; ++l) {
    abv[size_t(2 * l)] = bs.ab[size_t(l)][0];
    abv[size_t(2 * l + 1)] = bs.ab[size_t(l)][1];
  }
  FIELD_HIP(hipMalloc((void**)&qpinv, sizeof(double) * bs.qpinv.size()));
  FIELD_HIP(hipMemcpy(qpinv, bs.qpinv.data(), sizeof(double) * bs.qpinv.size(), hipMemcpyHostToDevice));
  FIELD_HIP(hipMalloc((void**)&ab, sizeof(int) * abv.size()));
  FIELD_HIP(hipMemcpy(ab, abv.data(), sizeof(int) * abv.size(), hipMemcpyHostToDevice));
  FIELD_HIP(hipMalloc((void**)&leaf_at, sizeof(int) * at.size()));
  FIELD_HIP(hipMemcpy(leaf_at, at.data(), sizeof(int) * at.size(), hipMemcpyHostToDevice));
  FIELD_HIP(hipMalloc((void**)&cx, sizeof(double) * side));
  FIELD_HIP(hipMemcpy(cx, cxs.data(), sizeof(double) * side, hipMemcpyHostToDevice));
  FIELD_HIP(hipMalloc((void**)&cy, sizeof(double) * side));
  FIELD_HIP(hipMemcpy(cy, cys.data(), sizeof(double) * side, hipMemcpyHostToDevice));
  FIELD_HIP(hipMalloc((void**)&coef, sizeof(double2) * size_t(nleaves) * nb));
  FIELD_HIP(hipMalloc((void**)&bad, sizeof(int)));
  return "";
}

std::string FieldInterp::run(const double2* f, const double* x, const double* y, long long n, double2* out,
                             hipStream_t s, bool* outside) {
  FIELD_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  const long long ne = (long long)nleaves * nb;
  hipLaunchKernelGGL(interp_coef_kernel, dim3(unsigned((ne + kFT - 1) / kFT)), dim3(kFT), 0, s, qpinv, f, nleaves,
                     rank, nb, coef);
  if (n > 0)
    hipLaunchKernelGGL(interp_eval_kernel, dim3(unsigned((n + kFT - 1) / kFT)), dim3(kFT), 0, s, coef, ab, leaf_at,
                       cx, cy, side, r, nb, x, y, n, out, bad);
  int hb = 0;
  FIELD_HIP(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  FIELD_HIP(hipStreamSynchronize(s));
  *outside = hb != 0;
  return "";
}

void FieldInterp::release() {
  void* ptrs[] = {qpinv, ab, leaf_at, cx, cy, coef, bad};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  qpinv = cx = cy = nullptr;
  ab = leaf_at = bad = nullptr;
  coef = nullptr;
}

}  // namespace lsg
