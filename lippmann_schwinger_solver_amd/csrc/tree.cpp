#include "tree.hpp"

#include <functional>

namespace lsg {

namespace {
const double kChildDx[4] = {-0.5, 0.5, 0.5, -0.5};  // children 0..3 = SW, SE, NE, NW (G_FMM2DTree.hpp:808-823)
const double kChildDy[4] = {-0.5, -0.5, 0.5, 0.5};
}  // namespace

int Tree::cone_toward(const Box& from, const Box& to, int level) const {
  // G_FMM2DTree.hpp:2524-2526 / 3335-3337: cone of `from` that contains `to`
  // as seen by the M2L/NCA passes (argument = direction from `from` to `to`).
  double arg = std::atan2(to.cy - from.cy, to.cx - from.cx);
  double a = std::fmod(arg + 2 * kPi, 2 * kPi);
  return int(a / aperture[level]);
}

std::string Tree::build(const TreeParams& params, const BasisSet& bs) {
  p = params;
  rank = bs.rank;
  if (p.tree_mode != 0) return "only the uniform tree (createUniformTree) is supported";
  if (p.nLevelsUniform < 2 || p.nLevelsUniform > 12) return "nLevelsUniform must be in [2, 12]";
  // level_LFR (G_FMM2DTree.hpp:316-326)
  if (p.yes2DFMM == 1) {
    if (p.kappa == 0) level_LFR = 2;
    else level_LFR = int(std::floor(std::log(p.kappa * p.L / kLfrB) / std::log(2.0)));
    if (level_LFR < 2) level_LFR = 2;
  } else {
    level_LFR = 2;
  }
  lv.assign(1, std::vector<Box>(1));
  Box& root = lv[0][0];
  root.level = 0; root.number = 0; root.radius = p.L; root.cx = 0.0; root.cy = 0.0;
  leaves.clear();
  // createTreeU (G_FMM2DTree.hpp:833-889), depth first
  std::string err;
  std::function<void(int, int)> grow = [&](int j, int k) {
    if (!err.empty()) return;
    bool cond3 = p.kappa * lv[j][k].radius <= 2 * kPi * 3.0;
    if (cond3 && j == p.nLevelsUniform) {
      lv[j][k].leaf = true;
      leaves.push_back({j, k});
      return;
    }
    if (j >= p.nLevelsUniform) { err = "leaf boxes hold more than 3 wavelengths (kappa*r > 6 pi)"; return; }
    if (int(lv.size()) == j + 1) lv.emplace_back();
    int n = int(lv[j + 1].size());
    for (int c = 0; c < 4; ++c) {
      Box b;
      b.level = j + 1;
      b.number = lv[j][k].number * 4 + c;
      b.parent = k;
      b.radius = 0.5 * lv[j][k].radius;
      b.cx = lv[j][k].cx + kChildDx[c] * lv[j][k].radius;
      b.cy = lv[j][k].cy + kChildDy[c] * lv[j][k].radius;
      lv[j][k].child[c] = n + c;
      lv[j + 1].push_back(b);
    }
    for (int c = 0; c < 4; ++c) grow(j + 1, n + c);
  };
  grow(0, 0);
  if (!err.empty()) return err;
  nLevels = int(lv.size()) - 1;
  boxRadius.assign(1, p.L);
  for (int j = 1; j <= std::max(10, nLevels + 1); ++j) boxRadius.push_back(boxRadius[j - 1] / 2.0);
  if (level_LFR >= nLevels) level_LFR = nLevels - 1;
  if (level_LFR < 1) level_LFR = 1;

  // createCones (G_FMM2DTree.hpp:1952-1974)
  nCones.clear(); aperture.clear();
  nCones.push_back(int(p.nCones_LFR * std::pow(2.0, level_LFR - 1)));
  aperture.push_back(2 * kPi / nCones[0]);
  for (int j = 1; j <= level_LFR - 1; ++j) {
    nCones.push_back(int(nCones[j - 1] / 2.0));
    aperture.push_back(aperture[j - 1] * 2.0);
  }
  for (int j = 0; j <= level_LFR - 1 && j <= nLevels; ++j)
    for (auto& b : lv[j])
      if (!b.leaf) b.cone_il.assign(nCones[j], {});

  // assign_Tree_Interactions (G_FMM2DTree.hpp:1356-1498)
  for (int c = 0; c < 4; ++c)
    for (int n = 0; n < 4; ++n) lv[1][c].nbrs.push_back({1, n});
  const double kap = p.kappa;
  auto sep_lfr = [&](const Box& A, int ja, const Box& B, int jb) {
    return std::fabs(A.cx - B.cx) >= 3.0 * boxRadius[ja] + boxRadius[jb] - kRoundOff ||
           std::fabs(A.cy - B.cy) >= 3.0 * boxRadius[ja] + boxRadius[jb] - kRoundOff;
  };
  std::function<void(int, int, int, int, std::vector<BoxRef>&)> future_lfr =
      [&](int bj, int bk, int pnj, int pni, std::vector<BoxRef>& out) {
        const Box& P = lv[pnj][pni];
        if (P.leaf) {
          if (!sep_lfr(lv[bj][bk], bj, P, pnj)) out.push_back({pnj, pni});
        } else {
          for (int nc = 0; nc < 4; ++nc) future_lfr(bj, bk, pnj + 1, P.child[nc], out);
        }
      };
  for (int j = 1; j <= nLevels - 1; ++j) {
    for (int k = 0; k < int(lv[j].size()); ++k) {
      if (lv[j][k].leaf) continue;
      for (int c = 0; c < 4; ++c) {
        const int ja = j + 1, ka = lv[j][k].child[c];
        const std::vector<BoxRef> parent_nbrs = lv[j][k].nbrs;
        for (const BoxRef& pn : parent_nbrs) {
          Box& A = lv[ja][ka];
          const Box& P = lv[pn.l][pn.k];
          if (ja >= level_LFR || (ja < level_LFR && A.leaf)) {
            if (P.leaf) {
              if (sep_lfr(A, ja, P, pn.l)) A.il.push_back(pn);
              else A.nbrs.push_back(pn);
            } else {
              for (int nc = 0; nc < 4; ++nc) {
                BoxRef bref{pn.l + 1, P.child[nc]};
                const Box& Bb = lv[bref.l][bref.k];
                if (sep_lfr(A, ja, Bb, bref.l)) {
                  A.il.push_back(bref);
                } else if (!A.leaf) {
                  A.nbrs.push_back(bref);
                } else {
                  std::vector<BoxRef> fg;
                  future_lfr(ja, ka, bref.l, bref.k, fg);
                  for (auto& f : fg) A.nbrs.push_back(f);
                }
              }
            }
          } else {
            const double r = boxRadius[ja];
            if (P.leaf) {
              double thr = kap * r * r + r + boxRadius[pn.l] - kRoundOff;
              if (std::fabs(A.cx - P.cx) >= thr || std::fabs(A.cy - P.cy) >= thr) {
                double arg = std::atan2(A.cy - P.cy, A.cx - P.cx);
                arg = std::fmod(arg + 2 * kPi + kPi, 2 * kPi);
                A.cone_il[int(arg / aperture[ja])].push_back(pn);
              } else {
                A.nbrs.push_back(pn);
              }
            } else {
              for (int nc = 0; nc < 4; ++nc) {
                BoxRef bref{pn.l + 1, P.child[nc]};
                const Box& Bb = lv[bref.l][bref.k];
                // x uses kappa r^2 + r + r_B, y uses kappa r^2 + r_B (G_FMM2DTree.hpp:1441-1442)
                if (std::fabs(A.cx - Bb.cx) >= kap * r * r + r + boxRadius[bref.l] - kRoundOff ||
                    std::fabs(A.cy - Bb.cy) >= kap * r * r + boxRadius[bref.l] - kRoundOff) {
                  double arg = std::atan2(A.cy - Bb.cy, A.cx - Bb.cx);
                  arg = std::fmod(arg + 2 * kPi + kPi, 2 * kPi);
                  A.cone_il[int(arg / aperture[ja])].push_back(bref);
                } else {
                  A.nbrs.push_back(bref);
                }
              }
            }
          }
        }
      }
    }
  }

  // getNeighbors colleague slots (G_FMM2DTree.hpp:1500-1601)
  double nang[8];
  for (int i = 0; i < 8; ++i) nang[i] = std::atan2(kColleagueCenters[i][1], kColleagueCenters[i][0]);
  for (int j = 1; j <= nLevels; ++j)
    for (auto& A : lv[j])
      for (const BoxRef& nb : A.nbrs) {
        const Box& B = lv[nb.l][nb.k];
        if (A.radius != B.radius) continue;  // fine / coarse slots: adaptive trees only
        double ang = std::atan2(B.cy - A.cy, B.cx - A.cx);
        if (std::fabs(B.cy - A.cy) < kRoundOff && std::fabs(B.cx - A.cx) < kRoundOff) {
          A.colleague[8] = nb.k;
        } else {
          for (int i = 0; i < 8; ++i)
            if (std::fabs(nang[i] - ang) < kRoundOff) { A.colleague[i] = nb.k; break; }
        }
      }

  // assignLeafChargeLocations (G_FMM2DTree.hpp:1150-1163)
  N = int(leaves.size()) * rank;
  gx.resize(N); gy.resize(N);
  for (size_t t = 0; t < leaves.size(); ++t) {
    Box& b = box(leaves[t]);
    b.leaf_index = int(t);
    const double r = boxRadius[leaves[t].l];
    for (int i = 0; i < rank; ++i) {
      gx[t * rank + i] = r * bs.sx[i] + b.cx;
      gy[t * rank + i] = r * bs.sy[i] + b.cy;
    }
  }
  return "";
}

}  // namespace lsg
