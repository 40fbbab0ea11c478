// Level tables shared by all boxes of a level:
//   Q_pinv  - pseudo-inverse of the Chebyshev Vandermonde of the leaf nodes
//   M[l]    - far-field Graf coefficient table (2*15+1 x n^2), M = C Q_pinv
//   T[l][s] - near-field operators (n^2 x n^2) for each neighbour slot s
// All M and T values pass through the reference's 6-significant-digit text
// round trip before use.
//
// Reference: GMRES/FarFieldInteraction.hpp:115-160 (evaluateQ/C/M),
//            GMRES/NearFieldInteraction1.hpp:98-128 (evaluateV, integrand),
//            GMRES/G_FMM2DTree.hpp:1773-1805 (getOperator),
//            GMRES/G_FMM2DTree.hpp:1815-1911 (neighbour table geometry),
//            GMRES/domain2D.hpp:46-138 (adaptive Gauss-Legendre quadrature).
#pragma once
#include <array>
#include <functional>
#include <vector>
#include "numerics.hpp"

namespace lsg {

constexpr int kSeries = 15;           // Graf series half-length, FarFieldInteraction.hpp:55
constexpr double kQuadEps = 1e-7;     // quadrature tolerance, FarFieldInteraction.hpp:37 / NearFieldInteraction1.hpp:112
constexpr int kQuadNodes = 10;        // GL points per panel side

// Neighbour slot geometry in units of the target box radius
// (GMRES/G_FMM2DTree.hpp:19-72): colleagues 9 (slot 8 = self), fine 12,
// separated fine 20, coarse 12.
extern const double kColleagueCenters[9][2];
extern const double kFineCenters[12][2];
extern const double kSepFineCenters[20][2];
extern const double kCoarseCenters[12][2];

struct BasisSet {
  int n = 0;        // Chebyshev nodes per side
  int rank = 0;     // n*n
  int deg = 0;      // degreeOfBases
  int nb = 0;       // deg*(deg+1)/2
  std::vector<double> node1d;            // -cos((k+0.5)pi/n)
  std::vector<double> sx, sy;            // standard nodes, index j*n+k -> (node1d[k], node1d[j])
  std::vector<std::array<int, 2>> ab;    // basis (a,b) in the reference order
  std::vector<double> qpinv;             // nb x rank, row-major
  void init(int nCheb, int degree);
};

// Adaptive tensor Gauss-Legendre quadrature on [cx-hx,cx+hx]x[cy-hy,cy+hy]:
// a panel is accepted when |I_panel - sum I_children| < eps, and the accepted
// panel's own value is summed (depth-first order).
cplx adaptive_quad(const std::function<cplx(double, double)>& f, double eps);

// M table at a level: (2*kSeries+1) x rank, row-major, rounded to 6 digits.
std::vector<cplx> far_table(const BasisSet& bs, double kappa, double radius);

// Near-field operator (rank x rank, row-major, rounded to 6 digits) from a
// source box B of radius rB centred at (bx,by) to the nodes of a target box of
// radius rA centred at the origin.
std::vector<cplx> near_table(const BasisSet& bs, double kappa, double rA, double rB, double bx, double by);
// Its two halves: the integrals V1(i, l) = int H0(kappa rB |X_i - P|) T_a T_b dP
// (rank x nb, row-major, by adaptive_quad), and T = round6((i rB^2/4) V1 Q_pinv).
std::vector<cplx> near_integrals(const BasisSet& bs, double kappa, double rA, double rB, double bx, double by);
std::vector<cplx> near_table_from_integrals(const BasisSet& bs, double rB, const std::vector<cplx>& V1);

// Geometry of one near-field table: target box radius rA at the origin, source
// box radius rB centred at (bx, by).
struct NearGeom {
  double rA = 0, rB = 0, bx = 0, by = 0;
};

}  // namespace lsg
