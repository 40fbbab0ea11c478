// The reference's table cache on disk: precomputations/M and
// precomputations/Neighbor in its directory layout and text format (one matrix
// row per line, "(re,im)" entries at the stream's default 6 significant
// digits, right-aligned to the widest entry as Eigen prints a matrix).
//
// Reference: GMRES/G_FMM2DTree.hpp:1252-1269 (writeMToFile), :1271-1299
// (ReadFromTextFile), :1301-1313 (readMFromFile), :1806-1909
// (getNeighborInteractions, writeNeighborInteractionsForLevel), :1911-1950
// (readNeighborInteractions).
#pragma once
#include <string>
#include <vector>
#include "tables.hpp"

namespace lsg {

// "<root>/precomputations/M/M_<n>_<int kappa>_<deg>" and the Neighbor twin.
std::string far_table_dir(const std::string& root, int n, double kappa, int deg);
std::string near_table_dir(const std::string& root, int n, double kappa, int deg);

// One matrix in the reference's text format; rows x cols, row-major values.
std::string write_matrix(const std::string& path, const std::vector<cplx>& A, int rows, int cols);
std::string read_matrix(const std::string& path, std::vector<cplx>& A, int& rows, int& cols);

// Far-field tables of levels 0..n_levels and the neighbour tables (9
// colleague, 12 fine, 20 separated-fine, 12 coarse) of levels 2..max_level,
// as the reference writes them. Returns "" on success.
std::string write_precomputations(const std::string& root, const BasisSet& bs, double kappa, double L, int n_levels,
                                  int max_level);

// The leaf level's far-field table ((2 kSeries + 1) x rank) and colleague
// tables (9 x rank x rank, slot-major) read back from the cache.
std::string read_far_table(const std::string& root, const BasisSet& bs, double kappa, int level,
                           std::vector<cplx>& M);
std::string read_colleague_tables(const std::string& root, const BasisSet& bs, double kappa, int level,
                                  std::vector<cplx>& T);
// One neighbour table of a level: kind 0 colleague, 1 fine, 2 separated fine,
// 3 coarse (readNeighborInteractions, G_FMM2DTree.hpp:1911-1950).
std::string read_neighbor_table(const std::string& root, const BasisSet& bs, double kappa, int level, int kind,
                                int slot, std::vector<cplx>& T);

}  // namespace lsg
