#include "precomp.hpp"

#include <sys/stat.h>

#include <cerrno>
#include <cstring>
#include <fstream>
#include <sstream>

namespace lsg {

namespace {

std::string dir_name(const std::string& root, const char* kind, int n, double kappa, int deg) {
  return root + "/precomputations/" + kind + "/" + kind + "_" + std::to_string(n) + "_" +
         std::to_string(int(kappa)) + "_" + std::to_string(deg);
}

// mkdir -p of <root>/precomputations/<kind>/<leaf>
std::string make_dirs(const std::string& root, const char* kind, const std::string& leaf) {
  for (const std::string& d : {root, root + "/precomputations", root + "/precomputations/" + kind, leaf})
    if (mkdir(d.c_str(), 0775) != 0 && errno != EEXIST) return "cannot create " + d + ": " + std::strerror(errno);
  return "";
}

// std::complex<double> inserted into a default-formatted stream: "(re,im)"
// with 6 significant digits in %g style
std::string entry_text(const cplx& z) {
  std::ostringstream o;
  o << z;
  return o.str();
}

const char* const kNeighborKinds[4] = {"colleagueNeighborInteraction", "fineNeighborInteraction",
                                       "separatedFineNeighborInteraction", "coarseNeighborInteraction"};

}  // namespace

std::string far_table_dir(const std::string& root, int n, double kappa, int deg) {
  return dir_name(root, "M", n, kappa, deg);
}
std::string near_table_dir(const std::string& root, int n, double kappa, int deg) {
  return dir_name(root, "Neighbor", n, kappa, deg);
}

std::string write_matrix(const std::string& path, const std::vector<cplx>& A, int rows, int cols) {
  if (size_t(rows) * cols != A.size()) return "matrix size mismatch for " + path;
  std::vector<std::string> txt(A.size());
  size_t width = 0;
  for (size_t e = 0; e < A.size(); ++e) {
    txt[e] = entry_text(A[e]);
    width = std::max(width, txt[e].size());
  }
  std::ofstream f(path);
  if (!f) return "cannot write " + path;
  for (int i = 0; i < rows; ++i) {
    if (i) f << '\n';
    for (int j = 0; j < cols; ++j) {
      if (j) f << ' ';
      const std::string& s = txt[size_t(i) * cols + j];
      f << std::string(width - s.size(), ' ') << s;
    }
  }
  f << std::endl;
  return f ? "" : "write failed for " + path;
}

std::string read_matrix(const std::string& path, std::vector<cplx>& A, int& rows, int& cols) {
  std::ifstream f(path);
  if (!f) return "cannot open " + path;
  std::string line;
  A.clear();
  rows = 0;
  cols = -1;
  while (std::getline(f, line)) {
    std::istringstream ls(line);
    int c = 0;
    cplx z;
    while (ls >> z) {
      A.push_back(z);
      ++c;
    }
    if (!ls.eof()) return "malformed entry in " + path;
    if (c == 0) continue;  // the blank line after the last row
    if (cols >= 0 && c != cols) return path + " does not form a matrix";
    cols = c;
    ++rows;
  }
  if (rows == 0) return path + " is empty";
  return "";
}

std::string write_precomputations(const std::string& root, const BasisSet& bs, double kappa, double L, int n_levels,
                                  int max_level) {
  const int n = bs.n, deg = bs.deg;
  auto radius = [&](int l) { return L / double(1 << l); };
  const std::string md = far_table_dir(root, n, kappa, deg), nd = near_table_dir(root, n, kappa, deg);
  std::string e = make_dirs(root, "M", md);
  if (!e.empty()) return e;
  for (int l = 0; l <= n_levels; ++l) {
    e = write_matrix(md + "/M_" + std::to_string(n) + "_" + std::to_string(l), far_table(bs, kappa, radius(l)),
                     2 * kSeries + 1, bs.rank);
    if (!e.empty()) return e;
  }
  e = make_dirs(root, "Neighbor", nd);
  if (!e.empty()) return e;
  const double (*centres[4])[2] = {kColleagueCenters, kFineCenters, kSepFineCenters, kCoarseCenters};
  const int counts[4] = {9, 12, 20, 12};
  const int dlevel[4] = {0, 1, 1, -1};  // source box level relative to the target's
  for (int la = 2; la <= max_level; ++la) {
    const double rA = radius(la);
    for (int kind = 0; kind < 4; ++kind) {
      const double rB = radius(la + dlevel[kind]);
      for (int s = 0; s < counts[kind]; ++s) {
        const std::vector<cplx> T =
            near_table(bs, kappa, rA, rB, centres[kind][s][0] * rA, centres[kind][s][1] * rA);
        e = write_matrix(nd + "/" + kNeighborKinds[kind] + "_" + std::to_string(la) + "_" + std::to_string(s), T,
                         bs.rank, bs.rank);
        if (!e.empty()) return e;
      }
    }
  }
  return "";
}

std::string read_far_table(const std::string& root, const BasisSet& bs, double kappa, int level,
                           std::vector<cplx>& M) {
  int r = 0, c = 0;
  const std::string p = far_table_dir(root, bs.n, kappa, bs.deg) + "/M_" + std::to_string(bs.n) + "_" +
                        std::to_string(level);
  std::string e = read_matrix(p, M, r, c);
  if (!e.empty()) return e;
  if (r != 2 * kSeries + 1 || c != bs.rank) return p + ": expected a " + std::to_string(2 * kSeries + 1) + " x " +
                                                   std::to_string(bs.rank) + " table";
  return "";
}

std::string read_colleague_tables(const std::string& root, const BasisSet& bs, double kappa, int level,
                                  std::vector<cplx>& T) {
  T.clear();
  const std::string nd = near_table_dir(root, bs.n, kappa, bs.deg);
  for (int s = 0; s < 9; ++s) {
    std::vector<cplx> A;
    int r = 0, c = 0;
    const std::string p = nd + "/" + kNeighborKinds[0] + "_" + std::to_string(level) + "_" + std::to_string(s);
    std::string e = read_matrix(p, A, r, c);
    if (!e.empty()) return e;
    if (r != bs.rank || c != bs.rank) return p + ": expected a " + std::to_string(bs.rank) + " x " +
                                             std::to_string(bs.rank) + " table";
    T.insert(T.end(), A.begin(), A.end());
  }
  return "";
}

std::string read_neighbor_table(const std::string& root, const BasisSet& bs, double kappa, int level, int kind,
                                int slot, std::vector<cplx>& T) {
  if (kind < 0 || kind > 3) return "neighbour table kind must be in [0, 3]";
  int r = 0, c = 0;
  const std::string p = near_table_dir(root, bs.n, kappa, bs.deg) + "/" + kNeighborKinds[kind] + "_" +
                        std::to_string(level) + "_" + std::to_string(slot);
  std::string e = read_matrix(p, T, r, c);
  if (!e.empty()) return e;
  if (r != bs.rank || c != bs.rank)
    return p + ": expected a " + std::to_string(bs.rank) + " x " + std::to_string(bs.rank) + " table";
  return "";
}

}  // namespace lsg
