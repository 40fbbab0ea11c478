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

// get_S (G_FMM2DTree.hpp:674-680): the 1-D Chebyshev interpolation kernel
// sum of T_k(x) T_k(y) at n first-kind nodes
static double cheb_s(double x, double y, int n) {
  double S = 0.5;
  for (int k = 1; k < n; ++k) S += std::cos(k * std::acos(x)) * std::cos(k * std::acos(y));
  return 2.0 / n * S;
}

std::string Tree::build(const TreeParams& params, const BasisSet& bs) {
  p = params;
  rank = bs.rank;
  if (p.tree_mode != 0 && p.tree_mode != 1) return "tree_mode must be 0 (uniform) or 1 (adaptive)";
  if (p.tree_mode == 1 && !p.contrast) return "adaptive tree without a contrast function";
  if (p.tree_mode == 0 && (p.nLevelsUniform < 2 || p.nLevelsUniform > 12)) return "nLevelsUniform must be in [2, 12]";
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
  auto make_children = [&](int j, int k) {
    if (int(lv.size()) == j + 1) lv.emplace_back();
    const int n = int(lv[j + 1].size());
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
    return n;
  };
  if (p.tree_mode == 0) {
    grow(0, 0);
  } else {
    // createTree (G_FMM2DTree.hpp:776-831): a box is a leaf when the contrast
    // at its children's nodes is reproduced to 10^-treeAdaptivity (max abs)
    // by interpolation from its own nodes (L2L of get_Transfer_Matrix,
    // :720-733) and it spans at most one wavelength (kappa r / 2 pi <= 1).
    // The reference's right-hand-side criterion is always true (:777-779).
    const int n1 = bs.n;
    std::vector<double> l2l(size_t(4) * rank * rank);
    for (int c = 0; c < 4; ++c)
      for (int j = 0; j < rank; ++j) {
        // standardChebNodesChild (:696-719): 0.5 x -/+ 0.5
        const double chx = 0.5 * bs.sx[j] + kChildDx[c], chy = 0.5 * bs.sy[j] + kChildDy[c];
        for (int k = 0; k < rank; ++k)
          l2l[(size_t(c) * rank + j) * rank + k] = cheb_s(bs.sx[k], chx, n1) * cheb_s(bs.sy[k], chy, n1);
      }
    const double eps_tree = std::pow(10.0, -p.treeAdaptivity);
    std::vector<double> qn(rank);
    auto resolve_contrast = [&](const Box& b) {
      for (int k = 0; k < rank; ++k) qn[k] = p.contrast(b.radius * bs.sx[k] + b.cx, b.radius * bs.sy[k] + b.cy);
      double err = 0.0;
      for (int c = 0; c < 4; ++c)
        for (int j = 0; j < rank; ++j) {
          double a = 0.0;
          for (int k = 0; k < rank; ++k) a += l2l[(size_t(c) * rank + j) * rank + k] * qn[k];
          const double chx = 0.5 * bs.sx[j] + kChildDx[c], chy = 0.5 * bs.sy[j] + kChildDy[c];
          const double tv = p.contrast(b.radius * chx + b.cx, b.radius * chy + b.cy);
          err = std::max(err, std::fabs(tv - a));
        }
      return err;
    };
    std::function<void(int, int)> grow_adaptive = [&](int j, int k) {
      if (!err.empty()) return;
      const bool c1 = resolve_contrast(lv[j][k]) < eps_tree;
      const bool c3 = p.kappa * lv[j][k].radius / 2 / kPi <= 1.0;
      if (c1 && c3) {
        lv[j][k].leaf = true;
        leaves.push_back({j, k});
        return;
      }
      if (j >= 12) { err = "adaptive tree deeper than 12 levels: contrast not resolved"; return; }
      const int n = make_children(j, k);
      for (int c = 0; c < 4; ++c) grow_adaptive(j + 1, n + c);
    };
    grow_adaptive(0, 0);
  }
  if (!err.empty()) return err;
  nLevels = int(lv.size()) - 1;
  boxRadius.assign(1, p.L);
  for (int j = 1; j <= std::max(10, nLevels + 1); ++j) boxRadius.push_back(boxRadius[j - 1] / 2.0);
  if (level_LFR >= nLevels) level_LFR = nLevels - 1;
  if (level_LFR < 1) level_LFR = 1;
  if (nLevels < 2) return "tree has fewer than 2 levels";

  auto sep_lfr = [&](const Box& A, int ja, const Box& B, int jb) {
    return std::fabs(A.cx - B.cx) >= 3.0 * boxRadius[ja] + boxRadius[jb] - kRoundOff ||
           std::fabs(A.cy - B.cy) >= 3.0 * boxRadius[ja] + boxRadius[jb] - kRoundOff;
  };
  // getFutureGeneration_LFR (:943-964): the leaves under P near box (bj, bk)
  std::function<void(int, int, int, int, std::vector<BoxRef>&)> future_lfr =
      [&](int bj, int bk, int pnj, int pni, std::vector<BoxRef>& out) {
        const Box& P = lv[pnj][pni];
        if (P.leaf) {
          if (!sep_lfr(lv[bj][bk], bj, P, pnj)) out.push_back({pnj, pni});
        } else {
          for (int nc = 0; nc < 4; ++nc) future_lfr(bj, bk, pnj + 1, P.child[nc], out);
        }
      };

  if (p.tree_mode == 1) {
    // makeLevelRestriction_Tree (:1019-1137): nLevels sweeps; each sweep lists
    // for every box the leaves near it (assign_Tree_Interactions_LFR,
    // :966-1017) and, for every leaf of the sweep's start, refines a near leaf
    // two or more levels coarser, or the leaf itself when a near leaf is two
    // or more levels finer. Refined boxes leave the leaf order; their children
    // are appended to it.
    for (int sweep = 0; sweep < nLevels; ++sweep) {
      std::vector<std::vector<std::vector<BoxRef>>> near(nLevels + 1);
      for (int j = 0; j <= nLevels; ++j) near[j].resize(lv[j].size());
      for (int c = 0; c < 4; ++c)
        for (int n = 0; n < 4; ++n) near[1][c].push_back({1, n});
      for (int j = 1; j <= nLevels - 1; ++j)
        for (int k = 0; k < int(lv[j].size()); ++k) {
          if (lv[j][k].leaf) continue;
          for (int c = 0; c < 4; ++c) {
            const int ka = lv[j][k].child[c];
            for (const BoxRef& pn : near[j][k]) future_lfr(j + 1, ka, pn.l, pn.k, near[j + 1][ka]);
          }
        }
      const std::vector<BoxRef> old = leaves;
      std::vector<char> killed(old.size(), 0);
      std::vector<std::vector<int>> pos(nLevels + 1);
      for (int j = 0; j <= nLevels; ++j) pos[j].assign(lv[j].size(), -1);
      for (int l = int(old.size()) - 1; l >= 0; --l) pos[old[l].l][old[l].k] = l;
      auto refine = [&](int l, BoxRef b) {
        killed[l] = 1;
        if (b.l + 1 > nLevels) { err = "2:1 restriction would refine below the finest level"; return; }
        lv[b.l][b.k].leaf = false;
        const int n = make_children(b.l, b.k);
        for (int c = 0; c < 4; ++c) {
          lv[b.l + 1][n + c].leaf = true;
          leaves.push_back({b.l + 1, n + c});
        }
      };
      for (int l = 0; l < int(old.size()) && err.empty(); ++l) {
        const BoxRef lf = old[l];
        for (const BoxRef& nb : near[lf.l][lf.k]) {
          if (lf.l - nb.l >= 2) {
            const int idx = pos[nb.l][nb.k];
            if (idx >= 0 && !killed[idx]) refine(idx, nb);
          } else if (nb.l - lf.l >= 2) {
            if (!killed[l]) refine(l, lf);
          }
        }
      }
      if (!err.empty()) return err;
      std::vector<BoxRef> kept;
      for (size_t l = 0; l < leaves.size(); ++l)
        if (l >= old.size() || !killed[l]) kept.push_back(leaves[l]);
      leaves.swap(kept);
    }
  }

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

  // getNeighbors (G_FMM2DTree.hpp:1500-1745): the slot of a neighbour is the
  // first slot centre whose polar angle matches the direction to it within
  // 1e-8; a coarse or fine neighbour matching no slot is dropped, a separated
  // fine one (farther than 2.5 r) falls into the last slot, as there.
  double nang[8], cang[12], fang[12], sang[20];
  for (int i = 0; i < 8; ++i) nang[i] = std::atan2(kColleagueCenters[i][1], kColleagueCenters[i][0]);
  for (int i = 0; i < 12; ++i) cang[i] = std::atan2(kCoarseCenters[i][1], kCoarseCenters[i][0]);
  for (int i = 0; i < 12; ++i) fang[i] = std::atan2(kFineCenters[i][1], kFineCenters[i][0]);
  for (int i = 0; i < 20; ++i) sang[i] = std::atan2(kSepFineCenters[i][1], kSepFineCenters[i][0]);
  for (int j = 1; j <= nLevels; ++j)
    for (auto& A : lv[j])
      for (const BoxRef& nb : A.nbrs) {
        const Box& B = lv[nb.l][nb.k];
        const double ang = std::atan2(B.cy - A.cy, B.cx - A.cx);
        if (A.radius == B.radius) {
          if (std::fabs(B.cy - A.cy) < kRoundOff && std::fabs(B.cx - A.cx) < kRoundOff) {
            A.colleague[8] = nb.k;
          } else {
            for (int i = 0; i < 8; ++i)
              if (std::fabs(nang[i] - ang) < kRoundOff) { A.colleague[i] = nb.k; break; }
          }
        } else if (A.radius < B.radius) {
          for (int i = 0; i < 12; ++i)
            if (std::fabs(cang[i] - ang) < kRoundOff) { A.coarse[i] = nb.k; break; }
        } else {
          const double dist = std::sqrt((A.cx - B.cx) * (A.cx - B.cx) + (A.cy - B.cy) * (A.cy - B.cy));
          if (dist > 2.5 * A.radius) {
            int slot = 19;
            for (int i = 0; i < 19; ++i)
              if (std::fabs(sang[i] - ang) < kRoundOff) { slot = i; break; }
            A.sepfine[slot] = nb.k;
          } else {
            for (int i = 0; i < 12; ++i)
              if (std::fabs(fang[i] - ang) < kRoundOff) { A.fine[i] = nb.k; break; }
          }
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
