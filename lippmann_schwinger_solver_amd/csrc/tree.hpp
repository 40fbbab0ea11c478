// Quadtree of the DAFMM: boxes, leaf order, interaction lists (plain for the
// low-frequency regime, per angular cone for the high-frequency regime) and
// the colleague slots of the near field.
//
// Reference: GMRES/G_FMM2DTree.hpp
//   FMM2DTree ctor / level_LFR            :293-328
//   get_Transfer_Matrix / resolveContrast :669-757
//   createTree / createAdaptiveTree       :776-831, 891-915
//   createTreeU / createUniformTree       :833-889, 917-941
//   makeLevelRestriction_Tree (2:1)       :943-1137
//   assign_Child_Interaction (IL rules)   :1356-1461
//   getNeighbors (slot numbering)         :1500-1745
//   createCones                           :1952-1974
//   assignLeafChargeLocations (node order):1150-1163
#pragma once
#include <functional>
#include <string>
#include <vector>
#include "numerics.hpp"
#include "tables.hpp"

namespace lsg {

constexpr double kRoundOff = 1e-8;   // epsilonRoundOff, G_FMM2DTree.hpp:116
constexpr double kLfrB = 0.35;       // B in level_LFR = floor(log2(kappa L / B)), :243

struct BoxRef {
  int l = -1, k = -1;
};

struct Box {
  int level = 0, number = 0, parent = -1;
  int child[4] = {-1, -1, -1, -1};
  bool leaf = false;
  double cx = 0, cy = 0, radius = 0;
  std::vector<BoxRef> nbrs;                  // neighborNumbers (incl. self)
  std::vector<BoxRef> il;                    // LFR interaction list
  std::vector<std::vector<BoxRef>> cone_il;  // HFR: interaction list per cone
  // near-field slots (box index at the slot's level, -1 = empty): colleagues
  // (same level, 8 = self), coarse (level - 1), fine and separated fine
  // (level + 1); only adaptive trees fill the last three
  int colleague[9] = {-1, -1, -1, -1, -1, -1, -1, -1, -1};
  int coarse[12] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  int fine[12] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  int sepfine[20] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  int leaf_index = -1;                       // position in leaf order
};

struct TreeParams {
  int nCones_LFR = 16;
  int nChebNodes = 6;
  double L = 0.5;
  double kappa = 1.0;
  int yes2DFMM = 1;
  int nLevelsUniform = 4;
  int tree_mode = 0;       // 0 uniform (createUniformTree), 1 adaptive (createAdaptiveTree + 2:1 restriction)
  int treeAdaptivity = 6;  // adaptive: contrast resolved to 10^-treeAdaptivity
  std::function<double(double, double)> contrast;  // adaptive: q(x, y)
};

struct Tree {
  TreeParams p;
  int rank = 0;
  int nLevels = 0;
  int level_LFR = 2;
  int N = 0;
  std::vector<double> boxRadius;
  std::vector<int> nCones;
  std::vector<double> aperture;
  std::vector<std::vector<Box>> lv;
  std::vector<BoxRef> leaves;
  std::vector<double> gx, gy;  // node coordinates in leaf-major order

  // returns empty string on success
  std::string build(const TreeParams& params, const BasisSet& bs);

  Box& box(BoxRef r) { return lv[r.l][r.k]; }
  const Box& box(BoxRef r) const { return lv[r.l][r.k]; }
  bool hfr(int l, int k) const { return l < level_LFR && !lv[l][k].leaf; }
  int cone_toward(const Box& from, const Box& to, int level) const;  // M2L/NCA cone of `from` facing `to`
};

}  // namespace lsg
