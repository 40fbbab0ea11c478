#include "tables.hpp"

#include <mutex>

namespace lsg {

const double kColleagueCenters[9][2] = {{-2, -2}, {0, -2}, {2, -2}, {2, 0}, {2, 2}, {0, 2}, {-2, 2}, {-2, 0}, {0, 0}};
const double kFineCenters[12][2] = {{-1.5, -1.5}, {-0.5, -1.5}, {0.5, -1.5}, {1.5, -1.5}, {1.5, -0.5}, {1.5, 0.5},
                                    {1.5, 1.5},   {0.5, 1.5},   {-0.5, 1.5}, {-1.5, 1.5}, {-1.5, 0.5}, {-1.5, -0.5}};
const double kSepFineCenters[20][2] = {{-2.5, -2.5}, {-1.5, -2.5}, {-0.5, -2.5}, {0.5, -2.5}, {1.5, -2.5},
                                       {2.5, -2.5},  {2.5, -1.5},  {2.5, -0.5},  {2.5, 0.5},  {2.5, 1.5},
                                       {2.5, 2.5},   {1.5, 2.5},   {0.5, 2.5},   {-0.5, 2.5}, {-1.5, 2.5},
                                       {-2.5, 2.5},  {-2.5, 1.5},  {-2.5, 0.5},  {-2.5, -0.5}, {-2.5, -1.5}};
const double kCoarseCenters[12][2] = {{-3, -3}, {-1, -3}, {1, -3}, {3, -3}, {3, -1}, {3, 1},
                                      {3, 3},   {1, 3},   {-1, 3}, {-3, 3}, {-3, 1}, {-3, -1}};

namespace {
struct GL10 {
  double x[kQuadNodes], w[kQuadNodes];
  GL10() { gauss_legendre(kQuadNodes, x, w); }
};
const GL10& gl10() {
  static GL10 g;
  return g;
}

cplx panel_integral(const std::function<cplx(double, double)>& f, double cx, double cy, double hx, double hy) {
  const GL10& g = gl10();
  double xs[kQuadNodes], ys[kQuadNodes];
  for (int l = 0; l < kQuadNodes; ++l) { xs[l] = cx + hx * g.x[l]; ys[l] = cy + hy * g.x[l]; }
  cplx s(0.0, 0.0);
  for (int m = 0; m < kQuadNodes; ++m)
    for (int n = 0; n < kQuadNodes; ++n) s += g.w[m] * g.w[n] * f(xs[m], ys[n]);
  s *= (hx * hy);
  return s;
}

// create_Tree(j,k) of GMRES/domain2D.hpp:69-120: children 0..3 = SW, SE, NE, NW.
void refine(const std::function<cplx(double, double)>& f, double cx, double cy, double hx, double hy, cplx parent,
            double eps, cplx& acc) {
  double rx = 0.5 * hx, ry = 0.5 * hy;
  double ccx[4] = {cx - rx, cx + rx, cx + rx, cx - rx};
  double ccy[4] = {cy - ry, cy - ry, cy + ry, cy + ry};
  cplx ch[4];
  for (int c = 0; c < 4; ++c) ch[c] = panel_integral(f, ccx[c], ccy[c], rx, ry);
  cplx err = parent;
  for (int c = 0; c < 4; ++c) err -= ch[c];
  if (std::abs(err) < eps) {
    acc += parent;
  } else {
    for (int c = 0; c < 4; ++c) refine(f, ccx[c], ccy[c], rx, ry, ch[c], eps, acc);
  }
}
}  // namespace

cplx adaptive_quad(const std::function<cplx(double, double)>& f, double eps) {
  cplx root = panel_integral(f, 0.0, 0.0, 1.0, 1.0);
  cplx acc(0.0, 0.0);
  refine(f, 0.0, 0.0, 1.0, 1.0, root, eps, acc);
  return acc;
}

void BasisSet::init(int nCheb, int degree) {
  n = nCheb;
  rank = n * n;
  deg = degree;
  nb = deg * (deg + 1) / 2;
  node1d.resize(n);
  for (int k = 0; k < n; ++k) node1d[k] = -std::cos((k + 0.5) / n * kPi);
  sx.resize(rank);
  sy.resize(rank);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) { sx[j * n + k] = node1d[k]; sy[j * n + k] = node1d[j]; }
  ab.clear();
  for (int d = 0; d < deg; ++d)
    for (int a = 0; a < d + 1; ++a) ab.push_back({a, d - a});
  // Q (rank x nb), Q(i,l) = T_a(x_i) T_b(y_i). For a, b < n the columns are
  // orthogonal at the tensor Chebyshev nodes (discrete orthogonality of T_k),
  // with Q^T Q = diag(d), d = n^2 (a?1/2:1)(b?1/2:1), so the pseudo-inverse
  // (FarFieldInteraction.hpp:148) is exactly diag(1/d) Q^T. The divisor is a
  // power of two times n^2, so every entry is correctly rounded.
  qpinv.assign(size_t(nb) * rank, 0.0);
  for (int l = 0; l < nb; ++l) {
    const double d = double(n) * n * (ab[l][0] ? 0.5 : 1.0) * (ab[l][1] ? 0.5 : 1.0);
    for (int i = 0; i < rank; ++i) qpinv[size_t(l) * rank + i] = cheb_poly(sx[i], ab[l][0]) * cheb_poly(sy[i], ab[l][1]) / d;
  }
}

// C(n,l) = r^2 * int_{[-1,1]^2} J_n(kappa r |P|) e^{-i n theta_P} T_a(P_x) T_b(P_y) dP;
// M = C Q_pinv.  FarFieldInteraction.hpp:62-74, 115-160.
std::vector<cplx> far_table(const BasisSet& bs, double kappa, double radius) {
  const int ns = 2 * kSeries + 1;
  std::vector<cplx> C(size_t(ns) * bs.nb);
#pragma omp parallel for schedule(dynamic)
  for (int t = 0; t < ns * bs.nb; ++t) {
    const int n = t / bs.nb - kSeries, l = t % bs.nb;
    const int a = bs.ab[l][0], b = bs.ab[l][1];
    auto f = [&](double px, double py) -> cplx {
      double normX = std::sqrt(px * px + py * py);
      double arg = kappa * radius * normX;
      double th = std::atan2(py, px);
      double J[kSeries + 1], Y[kSeries + 1];
      int an = n < 0 ? -n : n;
      bessel_jy(arg, an, J, Y);
      double jn = J[an];
      if (n < 0 && (an & 1)) jn = -jn;
      return jn * std::exp(-cplx(0.0, 1.0) * double(n) * th) * cheb_poly(px, a) * cheb_poly(py, b);
    };
    C[t] = radius * radius * adaptive_quad(f, kQuadEps);
  }
  std::vector<cplx> M(size_t(ns) * bs.rank);
  for (int r = 0; r < ns; ++r)
    for (int j = 0; j < bs.rank; ++j) {
      cplx s(0.0, 0.0);
      for (int l = 0; l < bs.nb; ++l) s += C[size_t(r) * bs.nb + l] * bs.qpinv[size_t(l) * bs.rank + j];
      M[size_t(r) * bs.rank + j] = round6(s);
    }
  return M;
}

// T = (i rB^2 / 4) V1 Q_pinv, V1(i,l) = int H0(kappa rB |X_i - P|) T_a T_b dP with
// X_i = (node_i of A - centre_B)/rB.  NearFieldInteraction1.hpp:98-128,
// G_FMM2DTree.hpp:1773-1805.
std::vector<cplx> near_integrals(const BasisSet& bs, double kappa, double rA, double rB, double bx, double by) {
  std::vector<cplx> V(size_t(bs.rank) * bs.nb);
#pragma omp parallel for schedule(dynamic)
  for (int t = 0; t < bs.rank * bs.nb; ++t) {
    const int i = t / bs.nb, l = t % bs.nb;
    const int a = bs.ab[l][0], b = bs.ab[l][1];
    const double Xx = (rA * bs.sx[i] + 0.0 - bx) / rB;
    const double Xy = (rA * bs.sy[i] + 0.0 - by) / rB;
    auto f = [&](double px, double py) -> cplx {
      double R = std::sqrt((Xx - px) * (Xx - px) + (Xy - py) * (Xy - py));
      double arg = kappa * rB * R;
      double J[2], Y[2];
      bessel_jy(arg, 0, J, Y);
      return (J[0] + cplx(0.0, 1.0) * Y[0]) * cheb_poly(px, a) * cheb_poly(py, b);
    };
    V[t] = adaptive_quad(f, kQuadEps);
  }
  return V;
}

std::vector<cplx> near_table_from_integrals(const BasisSet& bs, double rB, const std::vector<cplx>& V) {
  std::vector<cplx> T(size_t(bs.rank) * bs.rank);
  for (int i = 0; i < bs.rank; ++i)
    for (int j = 0; j < bs.rank; ++j) {
      cplx s(0.0, 0.0);
      for (int l = 0; l < bs.nb; ++l) {
        cplx vt = V[size_t(i) * bs.nb + l] * cplx(0.0, 1.0) * rB * rB / 4.0;
        s += vt * bs.qpinv[size_t(l) * bs.rank + j];
      }
      T[size_t(i) * bs.rank + j] = round6(s);
    }
  return T;
}

std::vector<cplx> near_table(const BasisSet& bs, double kappa, double rA, double rB, double bx, double by) {
  return near_table_from_integrals(bs, rB, near_integrals(bs, kappa, rA, rB, bx, by));
}

}  // namespace lsg
