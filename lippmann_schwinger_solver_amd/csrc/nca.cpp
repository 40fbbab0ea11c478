#include "nca.hpp"

#include "precomp.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <set>

namespace lsg {

// ---------------------------------------------------------------------------
// Entry kernel: G_FMM2DTree.hpp:630-657
// ---------------------------------------------------------------------------
void EntryKernel::hvec(int i, int leaf, cplx* h) const {
  const double xi = tree->gx[i], yi = tree->gy[i];
  const double cx = leaf_cx[leaf], cy = leaf_cy[leaf];
  const double th = std::atan2(yi - cy, xi - cx);
  const double R2 = (yi - cy) * (yi - cy) + (xi - cx) * (xi - cx);
  const double R = kappa * std::sqrt(R2);
  double J[kSeries + 1], Y[kSeries + 1];
  bessel_jy(R, kSeries, J, Y);
  const cplx I(0.0, 1.0);
  for (int n = -kSeries; n <= kSeries; ++n) {
    int an = n < 0 ? -n : n;
    double jn = J[an], yn = Y[an];
    if (n < 0 && (an & 1)) { jn = -jn; yn = -yn; }
    h[n + kSeries] = (jn + I * yn) * std::exp(I * double(n) * th);
  }
}

cplx EntryKernel::entry(const cplx* h, int i, int j) const {
  const int leaf = j / rank, idx = j % rank;
  const cplx* Mr = M[leaf_level[leaf]].data();
  cplx s(0.0, 0.0);
  for (int t = 0; t < 2 * kSeries + 1; ++t) s += Mr[size_t(t) * rank + idx] * h[t];
  s *= cplx(0.0, 1.0) / 4.0;
  if (find_phi) {
    s *= qk2[i];
    if (i == j) s += 1.0;
  }
  return s;
}

void EntryKernel::block(const int* rows, int nr, const int* cols, int nc, cplx* out) const {
  cplx h[2 * kSeries + 1];
  for (int r = 0; r < nr; ++r) {
    int prev = -1;
    for (int c = 0; c < nc; ++c) {
      int leaf = cols[c] / rank;
      if (leaf != prev) { hvec(rows[r], leaf, h); prev = leaf; }
      out[size_t(c) * nr + r] = entry(h, rows[r], cols[c]);
    }
  }
}

// ---------------------------------------------------------------------------
// ACA_only_nodes, GMRES/ACA.hpp:253-481
// ---------------------------------------------------------------------------
namespace {
// maxAbsVector (ACA.hpp:153-166): first index of the largest |v| over `allowed`
// (ascending); -1 if `allowed` is empty.
int max_abs(const std::vector<cplx>& v, const std::set<int>& allowed) {
  if (allowed.empty()) return -1;
  auto it = allowed.begin();
  int index = *it;
  double mx = std::abs(v[index]);
  for (; it != allowed.end(); ++it)
    if (std::abs(v[*it]) > mx) { index = *it; mx = std::abs(v[index]); }
  return index;
}
double vnorm(const std::vector<cplx>& v) {
  double s = 0;
  for (const cplx& z : v) s += std::norm(z);
  return std::sqrt(s);
}
cplx vdot(const std::vector<cplx>& a, const std::vector<cplx>& b) {  // Eigen a.dot(b) = a^H b
  cplx s(0.0, 0.0);
  for (size_t i = 0; i < a.size(); ++i) s += std::conj(a[i]) * b[i];
  return s;
}
}  // namespace

void aca_only_nodes(const EntryKernel& K, const std::vector<int>& rows, const std::vector<int>& cols, double eps,
                    AcaResult& out, double abs_exit) {
  const int N1 = int(rows.size()), N2 = int(cols.size());
  out.row_bases.clear(); out.col_bases.clear(); out.Ar.clear(); out.rank = 0;
  const int mn = std::min(N1, N2);
  std::vector<cplx> row(N2), col(N1), u, v, row_temp, col_temp;
  std::vector<std::vector<cplx>> U, V;
  std::vector<std::vector<cplx>> Arows;
  std::set<int> rem_row, rem_col;
  for (int l = 0; l < N1; ++l) rem_row.insert(l);
  for (int l = 0; l < N2; ++l) rem_col.insert(l);
  auto getRow = [&](int i, std::vector<cplx>& dst) {
    dst.resize(N2);
    K.block(&rows[i], 1, cols.data(), N2, dst.data());
  };
  // Column j needs h(rows[r], leaf of cols[j]) for every row. Pivot columns
  // often share a source leaf (a leaf box's own 36 nodes, a parent's four
  // children), so these series vectors are kept per source leaf for the
  // duration of the call (first in, first out, about 16 MB at most). The
  // entries are the same arithmetic as EntryKernel::block.
  constexpr int kH = 2 * kSeries + 1;
  const size_t hbytes = size_t(N1) * kH * sizeof(cplx);
  const int hcap = int(std::max<size_t>(1, std::min<size_t>(64, (size_t(16) << 20) / std::max<size_t>(hbytes, 1))));
  std::vector<int> hleaf;
  std::vector<std::vector<cplx>> hval;
  int hnext = 0;
  auto getCol = [&](int j, std::vector<cplx>& dst) {
    dst.resize(N1);
    const int leaf = cols[j] / K.rank;
    int slot = -1;
    for (size_t q = 0; q < hleaf.size(); ++q)
      if (hleaf[q] == leaf) slot = int(q);
    if (slot < 0) {
      if (int(hleaf.size()) < hcap) {
        hleaf.push_back(leaf);
        hval.emplace_back(size_t(N1) * kH);
        slot = int(hleaf.size()) - 1;
      } else {
        slot = hnext;
        hnext = (hnext + 1) % hcap;
        hleaf[size_t(slot)] = leaf;
      }
      cplx* h = hval[size_t(slot)].data();
      for (int r = 0; r < N1; ++r) K.hvec(rows[r], leaf, h + size_t(r) * kH);
    }
    const cplx* h = hval[size_t(slot)].data();
    for (int r = 0; r < N1; ++r) dst[r] = K.entry(h + size_t(r) * kH, rows[r], cols[j]);
  };
  int rank = 0;
  int row_index = 0, col_index = 0;
  if (N1 < N2) {
    int l_local;
    for (l_local = 0; l_local < N1; ++l_local) {
      row_index = l_local;
      getRow(row_index, row);
      col_index = max_abs(row, rem_col);
      if (std::abs(row[col_index]) > eps) break;
    }
    if (l_local == N1) return;
    v.resize(N2);
    for (int l = 0; l < N2; ++l) v[l] = row[l] / row[col_index];
    out.row_bases.push_back(row_index);
    out.col_bases.push_back(col_index);
    getCol(col_index, col);
    u = col;
    U.push_back(u); V.push_back(v);
    Arows.push_back(row);
    rem_col.erase(col_index); rem_row.erase(row_index);
    rank = 1;
    double normS = 0.0, prev_normS = 1000.0;
    row_index = max_abs(col, rem_row);
    while (row_index >= 0 && std::abs(row[col_index]) > eps && std::abs(col[row_index]) > eps &&
           std::fabs(prev_normS - normS) >= eps * prev_normS && rank < mn) {
      out.row_bases.push_back(row_index);
      getRow(row_index, row);
      row_temp = row;
      for (int l = 0; l <= rank - 1; ++l)
        for (size_t s = 0; s < V[l].size(); ++s) row[s] -= U[l][row_index] * V[l][s];
      col_index = max_abs(row, rem_col);
      out.col_bases.push_back(col_index);
      for (size_t l = 0; l < row.size(); ++l) v[l] = row[l] / row[col_index];
      getCol(col_index, col);
      col_temp = col;
      for (int l = 0; l <= rank - 1; ++l)
        for (size_t s = 0; s < U[l].size(); ++s) col[s] -= V[l][col_index] * U[l][s];
      u = col;
      if (vnorm(u) < abs_exit || vnorm(v) < abs_exit) {
        out.row_bases.pop_back(); out.col_bases.pop_back();
        break;
      }
      U.push_back(u); V.push_back(v);
      Arows.push_back(row_temp);
      ++rank;
      rem_col.erase(col_index); rem_row.erase(row_index);
      if (rank != 2) prev_normS = normS;
      normS = std::pow(normS, 2) + std::pow(vnorm(u) * vnorm(v), 2);
      for (int l = 0; l <= rank - 1; ++l) normS += 2 * std::abs(vdot(U[l], u) * vdot(V[l], v));
      normS = std::sqrt(normS);
      row_index = max_abs(col, rem_row);
    }
  } else {
    int l_local;
    for (l_local = 0; l_local < N2; ++l_local) {
      col_index = l_local;
      getCol(col_index, col);
      row_index = max_abs(col, rem_row);
      if (std::abs(col[row_index]) > eps) break;
    }
    if (l_local == N2) return;
    out.col_bases.push_back(col_index);
    v.resize(N1);
    for (int l = 0; l < N1; ++l) v[l] = col[l] / col[row_index];
    out.row_bases.push_back(row_index);
    getRow(row_index, row);
    u = row;
    U.push_back(u); V.push_back(v);
    Arows.push_back(row);
    rank = 1;
    rem_row.erase(row_index); rem_col.erase(col_index);
    double normS = 0.0, prev_normS = 1000.0;
    col_index = max_abs(row, rem_col);
    while (col_index >= 0 && std::abs(row[col_index]) > eps && std::abs(col[row_index]) > eps &&
           std::fabs(prev_normS - normS) >= eps * prev_normS && rank < mn) {
      out.col_bases.push_back(col_index);
      getCol(col_index, col);
      col_temp = col;
      for (int l = 0; l <= rank - 1; ++l)
        for (size_t s = 0; s < V[l].size(); ++s) col[s] -= U[l][col_index] * V[l][s];
      row_index = max_abs(col, rem_row);
      out.row_bases.push_back(row_index);
      for (size_t l = 0; l < col.size(); ++l) v[l] = col[l] / col[row_index];
      getRow(row_index, row);
      row_temp = row;
      for (int l = 0; l <= rank - 1; ++l)
        for (size_t s = 0; s < U[l].size(); ++s) row[s] -= V[l][row_index] * U[l][s];
      u = row;
      if (vnorm(u) < abs_exit || vnorm(v) < abs_exit) {
        out.col_bases.pop_back(); out.row_bases.pop_back();
        break;
      }
      U.push_back(u); V.push_back(v);
      Arows.push_back(row_temp);
      ++rank;
      rem_row.erase(row_index); rem_col.erase(col_index);
      if (rank != 2) prev_normS = normS;
      normS = std::pow(normS, 2) + std::pow(vnorm(u) * vnorm(v), 2);
      for (int l = 0; l <= rank - 1; ++l) normS += 2 * std::abs(vdot(U[l], u) * vdot(V[l], v));
      normS = std::sqrt(normS);
      col_index = max_abs(row, rem_col);
    }
  }
  out.rank = rank;
  out.Ar.resize(size_t(rank) * N2);
  for (int r = 0; r < rank; ++r) std::copy(Arows[r].begin(), Arows[r].end(), out.Ar.begin() + size_t(r) * N2);
}

// ---------------------------------------------------------------------------
// NCA skeletons and plan
// ---------------------------------------------------------------------------
namespace {

struct Contrib {
  std::vector<int> rows;                 // target nodes
  const std::vector<int>* cols = nullptr;  // source charge nodes
  const ColPivQR* Zt = nullptr;            // L2L only: QR of Atilde_in^T; operator = K Atilde_in^-1
  const std::vector<cplx>* At = nullptr;   // L2L, device build: Atilde_in^T itself (factored on the GPU)
  const std::vector<int>* src_list = nullptr;  // source state indices (M2L)
  int64_t src_base = -1;                       // contiguous source (L2L)
};

struct Owner {
  bool out_active = false, in_active = false;
  // ACA operand sizes and rank (-1: not run, an operand was empty), outgoing
  // then incoming: {rows, cols, rank}
  int aca_out[3] = {0, 0, -1}, aca_in[3] = {0, 0, -1};
  std::vector<int> out_check, out_charge, in_check, in_charge, ucp;
  std::vector<int> src_nodes;  // outgoing ACA columns
  std::vector<cplx> Ar;        // rank x src, row-major
  std::vector<cplx> P;         // rank x src, col-major (Atilde^-1 Ar)
  ColPivQR in_qrT;             // QR of Atilde_in^T (incoming check x charge points)
  std::vector<cplx> in_At;     // device build: Atilde_in^T (r x r, col-major), factored with each solve
  std::vector<int> fout;       // state indices aligned with out_charge
  int64_t uin = -1;            // state offset of the incoming potential (size ucp)
  std::vector<Contrib> contribs;
  std::vector<Contrib> leaf_m2l, leaf_l2l;  // leaves only
};

struct Builder {
  const Tree& t;
  const BasisSet& bs;
  const EntryKernel& K;
  double eps;
  double abs_exit = 1e-8;
  std::vector<std::vector<Owner>> L;  // L[l][k]: LFR owners (and leaves)
  std::vector<std::vector<Owner>> H;  // H[l][k*nc+c]: HFR cones
  int rank;

  Builder(const Tree& tr, const BasisSet& b, const EntryKernel& k, double e) : t(tr), bs(b), K(k), eps(e) {
    rank = bs.rank;
    L.resize(t.nLevels + 1);
    H.resize(t.nLevels + 1);
    for (int l = 0; l <= t.nLevels; ++l) {
      L[l].resize(t.lv[l].size());
      if (l < t.level_LFR) H[l].resize(t.lv[l].size() * size_t(t.nCones[l]));
    }
  }
  bool is_lfr(int l, int k) const { return l >= t.level_LFR || t.lv[l][k].leaf; }
  Owner& hc(int l, int k, int c) { return H[l][size_t(k) * t.nCones[l] + c]; }
  void leaf_nodes(int l, int k, std::vector<int>& out) const {
    int li = t.lv[l][k].leaf_index;
    for (int i = 0; i < rank; ++i) out.push_back(li * rank + i);
  }
  static void app(std::vector<int>& dst, const std::vector<int>& src) { dst.insert(dst.end(), src.begin(), src.end()); }

  // getParticlesFromChildrenLFR_* (G_FMM2DTree.hpp:2013-2060, 2242-2292)
  // which: 0 = outgoing charges, 1 = incoming checks
  void lfr_children(int j, int k, int which, std::vector<int>& out) {
    if (t.lv[j][k].leaf) { leaf_nodes(j, k, out); return; }
    const int J = j + 1;
    for (int c = 0; c < 4; ++c) {
      int Kc = t.lv[j][k].child[c];
      if (is_lfr(J, Kc)) app(out, which == 0 ? L[J][Kc].out_charge : L[J][Kc].in_check);
      else
        for (int cone = 0; cone < t.nCones[J]; ++cone) app(out, which == 0 ? hc(J, Kc, cone).out_charge : hc(J, Kc, cone).in_check);
    }
  }
  // getParticlesFromChildrenHFR_* (G_FMM2DTree.hpp:3258-3320, 3421-3483)
  void hfr_children(int j, int k, int cone, int which, std::vector<int>& out) {
    if (t.lv[j][k].leaf) { leaf_nodes(j, k, out); return; }
    const int J = j + 1;
    for (int c = 0; c < 4; ++c) {
      int Kc = t.lv[j][k].child[c];
      if (j == t.level_LFR - 1 || t.lv[J][Kc].leaf) app(out, which == 0 ? L[J][Kc].out_charge : L[J][Kc].in_check);
      else app(out, which == 0 ? hc(J, Kc, cone / 2).out_charge : hc(J, Kc, cone / 2).in_check);
    }
  }

  // K(row pivots, column pivots), column-major, gathered from the ACA's raw
  // pivot rows Ar = K(row pivots, all columns): the same entries K.block of the
  // skeleton lists computes (same series vector per row and source leaf).
  static std::vector<cplx> pivot_block(const AcaResult& res, int ncols) {
    const int r = res.rank;
    std::vector<cplx> A(size_t(r) * r);
    for (int j = 0; j < r; ++j)
      for (int i = 0; i < r; ++i) A[size_t(j) * r + i] = res.Ar[size_t(i) * ncols + res.col_bases[size_t(j)]];
    return A;
  }
  void finish_out(Owner& o, const std::vector<int>& rows, const std::vector<int>& cols, const AcaResult& res) {
    for (int r : res.row_bases) o.out_check.push_back(rows[r]);
    for (int c : res.col_bases) o.out_charge.push_back(cols[c]);
    if (dev) {
      // P = Atilde^-1 Ar is solved on the GPU for the whole level (qr_level)
      o.P.resize(size_t(res.rank) * cols.size());
      o.src_nodes = cols;
      return;
    }
    ColPivQR qr;
    {
      const std::vector<cplx> A = pivot_block(res, int(cols.size()));
      qr.factor(A.data(), res.rank, res.rank);
    }
    const int r = res.rank, ns = int(cols.size());
    std::vector<cplx> B(size_t(r) * ns);
    for (int i = 0; i < r; ++i)
      for (int j = 0; j < ns; ++j) B[size_t(j) * r + i] = res.Ar[size_t(i) * ns + j];
    o.P.resize(size_t(qr.cols) * ns);
    qr.solve(B.data(), ns, o.P.data());
    o.src_nodes = cols;
  }
  void finish_in(Owner& o, const std::vector<int>& rows, const std::vector<int>& cols, const AcaResult& res) {
    for (int r : res.row_bases) o.in_check.push_back(rows[r]);
    for (int c : res.col_bases) o.in_charge.push_back(cols[c]);
    // The downward operator K(child, in_charge) Atilde_in^-1 is formed row by
    // row from solves with Atilde_in^T (backward stable), never through an
    // explicit inverse, whose entries grow like cond(Atilde_in).
    const std::vector<cplx> A = pivot_block(res, int(cols.size()));
    const int r = res.rank;
    std::vector<cplx> At(size_t(r) * r);
    for (int i = 0; i < r; ++i)
      for (int j = 0; j < r; ++j) At[size_t(i) * r + j] = A[size_t(j) * r + i];
    if (dev) {
      o.in_At = std::move(At);   // factored on the GPU with every L2L solve
      return;
    }
    o.in_qrT.factor(At.data(), r, r);
  }

  // One skeleton construction: an owner (LFR box or HFR cone), the direction
  // (outgoing or incoming) and its ACA operands.
  struct Job {
    Owner* o = nullptr;
    bool out = true;
    std::vector<int> rows, cols;
    AcaResult res;
  };
  // getNodes_LFR_outgoing_box / _incoming_box (G_FMM2DTree.hpp:2083-2212, 2294-2358):
  // outgoing: columns = the box's charges (children's outgoing charges or leaf
  // nodes), rows = the incoming check points of its interaction list; incoming:
  // rows = the box's check points, columns = the interaction list's charges.
  void lfr_job(int j, int k, bool out, Job& jb) {
    jb.o = &L[j][k];
    jb.out = out;
    lfr_children(j, k, out ? 0 : 1, out ? jb.cols : jb.rows);
    for (const BoxRef& b : t.lv[j][k].il) lfr_children(b.l, b.k, out ? 1 : 0, out ? jb.rows : jb.cols);
  }
  // getNodes_HFR_outgoing_box / _incoming_box (G_FMM2DTree.hpp:3322-3392, 3485-3546)
  void hfr_job(int j, int k, int cone, bool out, Job& jb) {
    jb.o = &hc(j, k, cone);
    jb.out = out;
    hfr_children(j, k, cone, out ? 0 : 1, out ? jb.cols : jb.rows);
    for (const BoxRef& b : t.lv[j][k].cone_il[cone]) {
      int coneB = t.cone_toward(t.lv[b.l][b.k], t.lv[j][k], b.l);
      hfr_children(b.l, b.k, coneB, out ? 1 : 0, out ? jb.rows : jb.cols);
    }
  }
  static bool job_active(const Job& jb) { return !jb.rows.empty() && !jb.cols.empty(); }
  void complete(Job& jb) {
    Owner& o = *jb.o;
    int* shape = jb.out ? o.aca_out : o.aca_in;
    shape[0] = int(jb.rows.size());
    shape[1] = int(jb.cols.size());
    shape[2] = job_active(jb) ? jb.res.rank : -1;
    if (jb.out) {
      if (job_active(jb)) {
        o.out_active = true;
        if (jb.res.rank > 0) finish_out(o, jb.rows, jb.cols, jb.res);
      } else {
        o.out_active = false;
        o.out_charge = jb.cols;
      }
    } else {
      if (job_active(jb)) {
        o.in_active = true;
        if (jb.res.rank > 0) finish_in(o, jb.rows, jb.cols, jb.res);
      } else {
        o.in_active = false;
        o.in_check = jb.rows;
      }
    }
  }
  // Device build: P = Atilde_out^-1 Ar of every outgoing skeleton of a level,
  // one batched QR factor + solve on the GPU.
  std::string qr_level(std::vector<Job>& jobs) {
    std::vector<QrSolveJob> qj;
    std::vector<std::vector<cplx>> As(jobs.size()), Bs(jobs.size());
    for (size_t q = 0; q < jobs.size(); ++q) {
      Job& jb = jobs[q];
      if (!jb.out || !job_active(jb) || jb.res.rank == 0) continue;
      const int r = jb.res.rank, ns = int(jb.cols.size());
      As[q] = pivot_block(jb.res, ns);
      Bs[q].resize(size_t(r) * ns);
      for (int i = 0; i < r; ++i)
        for (int j = 0; j < ns; ++j) Bs[q][size_t(j) * r + i] = jb.res.Ar[size_t(i) * ns + j];
      qj.push_back(QrSolveJob{As[q].data(), r, Bs[q].data(), ns, jb.o->P.data()});
    }
    return qj.empty() ? "" : dev->qr_solve(qj);
  }
  // Skeletons of one level: operand lists (from the level below), the ACAs
  // (host threads or one batched device launch sequence), then the QR finish.
  std::string run_level(std::vector<Job>& jobs) {
    const int nj = int(jobs.size());
    if (dev) {
      std::vector<AcaJob> aj;
      for (Job& jb : jobs)
        if (job_active(jb)) aj.push_back(AcaJob{&jb.rows, &jb.cols, &jb.res});
      std::string e = dev->aca(aj, eps, abs_exit);
      if (!e.empty()) return e;
    } else {
#pragma omp parallel for schedule(dynamic, 1)
      for (int q = 0; q < nj; ++q)
        if (job_active(jobs[q])) aca_only_nodes(K, jobs[q].rows, jobs[q].cols, eps, jobs[q].res, abs_exit);
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < nj; ++q) complete(jobs[q]);
    return dev ? qr_level(jobs) : "";
  }

  DeviceNca* dev = nullptr;

  std::string run_nca() {
    // getNodes_LFR (G_FMM2DTree.hpp:2062-2081)
    for (int j = t.nLevels; j >= t.level_LFR; --j) {
      const int nb = int(t.lv[j].size());
      std::vector<Job> jobs(size_t(2) * nb);
#pragma omp parallel for schedule(dynamic, 8)
      for (int q = 0; q < 2 * nb; ++q) lfr_job(j, q >> 1, (q & 1) == 0, jobs[size_t(q)]);
      std::string e = run_level(jobs);
      if (!e.empty()) return e;
    }
    // getNodes_HFR (G_FMM2DTree.hpp:3232-3255); leaf boxes above level_LFR keep LFR skeletons
    for (int j = t.level_LFR - 1; j >= 2; --j) {
      const int nb = int(t.lv[j].size()), nc = t.nCones[j];
      std::vector<Job> jobs(size_t(2) * nb * nc);
      std::vector<char> used(jobs.size(), 0);
#pragma omp parallel for schedule(dynamic, 8)
      for (int q = 0; q < 2 * nb * nc; ++q) {
        const int w = q >> 1, k = w / nc, c = w % nc;
        const bool out = (q & 1) == 0;
        if (t.lv[j][k].leaf) {
          if (c == 0) { lfr_job(j, k, out, jobs[size_t(q)]); used[size_t(q)] = 1; }
        } else {
          hfr_job(j, k, c, out, jobs[size_t(q)]);
          used[size_t(q)] = 1;
        }
      }
      std::vector<Job> live;
      live.reserve(jobs.size());
      for (size_t q = 0; q < jobs.size(); ++q)
        if (used[q]) live.push_back(std::move(jobs[q]));
      std::string e = run_level(live);
      if (!e.empty()) return e;
    }
    // getUserCheckPoints (G_FMM2DTree.hpp:3146-3230)
    for (int j = t.nLevels; j >= t.level_LFR; --j)
      for (int k = 0; k < int(t.lv[j].size()); ++k) {
        Owner& o = L[j][k];
        if (t.lv[j][k].leaf) leaf_nodes(j, k, o.ucp);
        else if (o.in_active) o.ucp = o.in_check;
        else
          for (int c = 0; c < 4; ++c) app(o.ucp, L[j + 1][t.lv[j][k].child[c]].ucp);
      }
    for (int j = t.level_LFR - 1; j >= 2; --j)
      for (int k = 0; k < int(t.lv[j].size()); ++k) {
        if (t.lv[j][k].leaf) { leaf_nodes(j, k, L[j][k].ucp); continue; }
        for (int cone = 0; cone < t.nCones[j]; ++cone) {
          Owner& o = hc(j, k, cone);
          if (o.in_active) { o.ucp = o.in_check; continue; }
          for (int c = 0; c < 4; ++c) {
            int J = j + 1, Kc = t.lv[j][k].child[c];
            if (t.lv[J][Kc].leaf || J == t.level_LFR) app(o.ucp, L[J][Kc].ucp);
            else app(o.ucp, hc(J, Kc, cone / 2).ucp);
          }
        }
      }
    return "";
  }

  // ------------------------------------------------------------------ plan
  int64_t next_slot = 0;

  // outgoing charge state indices, aligned with out_charge (reference:
  // LFR_M2M_ILActive_False / HFR_M2M_ILActiveFalse concatenations)
  void assign_fout() {
    for (int j = t.nLevels; j >= 2; --j) {
      for (int k = 0; k < int(t.lv[j].size()); ++k) {
        if (is_lfr(j, k)) {
          Owner& o = L[j][k];
          if (o.out_active) {
            for (size_t i = 0; i < o.out_charge.size(); ++i) o.fout.push_back(int(next_slot++));
          } else if (t.lv[j][k].leaf) {
            leaf_nodes(j, k, o.fout);
          } else {
            for (int c = 0; c < 4; ++c) {
              int J = j + 1, Kc = t.lv[j][k].child[c];
              if (is_lfr(J, Kc)) app(o.fout, L[J][Kc].fout);
              else
                for (int cone = 0; cone < t.nCones[J]; ++cone) app(o.fout, hc(J, Kc, cone).fout);
            }
          }
        } else {
          for (int cone = 0; cone < t.nCones[j]; ++cone) {
            Owner& o = hc(j, k, cone);
            if (o.out_active) {
              for (size_t i = 0; i < o.out_charge.size(); ++i) o.fout.push_back(int(next_slot++));
            } else {
              for (int c = 0; c < 4; ++c) {
                int J = j + 1, Kc = t.lv[j][k].child[c];
                if (j == t.level_LFR - 1 || t.lv[J][Kc].leaf) app(o.fout, L[J][Kc].fout);
                else app(o.fout, hc(J, Kc, cone / 2).fout);
              }
            }
          }
        }
      }
    }
  }
  // source state indices of an outgoing operator, aligned with src_nodes
  void out_src(int j, int k, int cone, std::vector<int>& s) {
    if (t.lv[j][k].leaf) { leaf_nodes(j, k, s); return; }
    const int J = j + 1;
    for (int c = 0; c < 4; ++c) {
      int Kc = t.lv[j][k].child[c];
      if (cone < 0) {
        if (is_lfr(J, Kc)) app(s, L[J][Kc].fout);
        else
          for (int cc = 0; cc < t.nCones[J]; ++cc) app(s, hc(J, Kc, cc).fout);
      } else {
        if (j == t.level_LFR - 1 || t.lv[J][Kc].leaf) app(s, L[J][Kc].fout);
        else app(s, hc(J, Kc, cone / 2).fout);
      }
    }
  }

  struct Dest {
    int64_t dst;
    int rows;
    const std::vector<cplx>* P = nullptr;      // upward: single operator
    const std::vector<int>* P_src = nullptr;
    std::vector<int> up_src;
    std::vector<Contrib>* contribs = nullptr;  // downward
  };

  // Children that receive a share of an owner's incoming potential, with the
  // row segment of the owner's ucp each one holds (getUserCheckPoints order).
  struct ChildTarget { Owner* o; int level; };
  void child_targets(int j, int k, int cone, std::vector<ChildTarget>& out) {
    const int J = j + 1;
    for (int c = 0; c < 4; ++c) {
      int Kc = t.lv[j][k].child[c];
      if (cone < 0) out.push_back({&L[J][Kc], J});
      else if (t.lv[J][Kc].leaf || J == t.level_LFR) out.push_back({&L[J][Kc], J});
      else out.push_back({&hc(J, Kc, cone / 2), J});
    }
  }

  void add_m2l(int j, int k, int cone, Owner& o) {
    const Box& A = t.lv[j][k];
    const std::vector<BoxRef>& il = cone < 0 ? A.il : A.cone_il[cone];
    if (!o.in_active || o.ucp.empty()) return;
    for (const BoxRef& b : il) {
      Owner* s;
      if (b.l < t.level_LFR && !t.lv[b.l][b.k].leaf) s = &hc(b.l, b.k, t.cone_toward(t.lv[b.l][b.k], A, b.l));
      else s = &L[b.l][b.k];
      if (s->out_charge.empty()) continue;
      Contrib c;
      c.rows = o.ucp;
      c.cols = &s->out_charge;
      c.src_list = &s->fout;
      o.contribs.push_back(std::move(c));
    }
  }

  // resolve an owner at level j after all its contributions are known
  // Leaves get two destinations: the M2L blocks (sources are outgoing charges
  // of same-level boxes, ready right after the leaf-level upward pass) and the
  // L2L blocks (sources are incoming potentials of ancestors).
  void resolve(int j, int k, int cone, Owner& o, std::vector<Dest>& dests, std::vector<Dest>& leaf_m2l,
               std::vector<Dest>& leaf_l2l) {
    const bool leaf = t.lv[j][k].leaf;
    if (leaf) {
      for (Contrib& c : o.contribs) (c.src_list ? o.leaf_m2l : o.leaf_l2l).push_back(std::move(c));
      o.contribs.clear();
      for (int part = 0; part < 2; ++part) {
        std::vector<Contrib>& cs = part == 0 ? o.leaf_m2l : o.leaf_l2l;
        if (cs.empty()) continue;
        Dest d;
        d.dst = t.N + int64_t(t.lv[j][k].leaf_index) * rank;
        d.rows = rank;
        d.contribs = &cs;
        (part == 0 ? leaf_m2l : leaf_l2l).push_back(std::move(d));
      }
      return;
    }
    if (!o.in_active) {
      // pass-through: hand each contribution's row segments to the children
      std::vector<ChildTarget> ch;
      child_targets(j, k, cone, ch);
      for (Contrib& c : o.contribs) {
        size_t off = 0;
        for (ChildTarget& tg : ch) {
          size_t n = tg.o->ucp.size();
          if (n > 0) {
            Contrib s;
            s.rows.assign(c.rows.begin() + off, c.rows.begin() + off + n);
            s.cols = c.cols; s.Zt = c.Zt; s.At = c.At; s.src_list = c.src_list; s.src_base = c.src_base;
            tg.o->contribs.push_back(std::move(s));
          }
          off += n;
        }
      }
      o.contribs.clear();
      return;
    }
    if (o.ucp.empty() || o.contribs.empty()) return;
    o.uin = next_slot;
    next_slot += int64_t(o.ucp.size());
    Dest d;
    d.dst = o.uin;
    d.rows = int(o.ucp.size());
    d.contribs = &o.contribs;
    dests.push_back(std::move(d));
    if (o.in_charge.empty()) return;
    // L2L (fused with Atilde_in^-1) into the children
    std::vector<ChildTarget> ch;
    child_targets(j, k, cone, ch);
    for (ChildTarget& tg : ch) {
      if (tg.o->ucp.empty()) continue;
      Contrib c;
      c.rows = tg.o->ucp;
      c.cols = &o.in_charge;
      if (dev) c.At = &o.in_At;
      else c.Zt = &o.in_qrT;
      c.src_base = o.uin;
      tg.o->contribs.push_back(std::move(c));
    }
  }
};

int64_t contrib_cols(const Contrib& c) { return int64_t(c.cols->size()); }

}  // namespace

std::string make_entry_kernel(const Tree& t, const BasisSet& bs, const Problem& prob, bool find_phi,
                              const std::string& table_dir, EntryKernel& K) {
  K = EntryKernel();
  K.tree = &t;
  K.rank = bs.rank;
  K.kappa = prob.kappa;
  K.find_phi = find_phi;
  K.qk2.resize(t.N);
  for (int i = 0; i < t.N; ++i) K.qk2[i] = prob.kappa * prob.kappa * prob.q(t.gx[i], t.gy[i]);
  K.leaf_level.resize(t.leaves.size());
  K.leaf_cx.resize(t.leaves.size());
  K.leaf_cy.resize(t.leaves.size());
  K.M.resize(t.nLevels + 1);
  for (size_t i = 0; i < t.leaves.size(); ++i) {
    const Box& b = t.box(t.leaves[i]);
    K.leaf_level[i] = t.leaves[i].l;
    K.leaf_cx[i] = b.cx;
    K.leaf_cy[i] = b.cy;
    const int l = t.leaves[i].l;
    if (!K.M[l].empty()) continue;
    if (table_dir.empty()) {
      K.M[l] = far_table(bs, prob.kappa, t.boxRadius[l]);
    } else {
      std::string e = read_far_table(table_dir, bs, prob.kappa, l, K.M[l]);
      if (!e.empty()) return e;
    }
  }
  return "";
}

// Near field of an adaptive tree (evaluate_NearField_Using_Precomputations_LS,
// G_FMM2DTree.hpp:2884-2960) as one store pass named "near": leaf l's
// destination is its rank rows, its columns the nodes of its neighbour leaves
// slot by slot (colleagues, fine, separated fine, coarse), and its operator the
// side-by-side blocks diag(k^2 q) T (+ I on the self slot; T alone for the
// volume potential). T of slot s at leaf level j is getOperator for a source
// box of level j, j+1 or j-1 centred at the slot centre times r_j (the
// writeNeighborInteractionsForLevel geometry, :1812-1857); each table in use
// is computed once (or read from the table cache).
static std::string adaptive_near_pass(const Tree& t, const BasisSet& bs, const Problem& prob, const BuildOptions& opt,
                                      const std::vector<double>& qk2, DeviceNca* dev, Plan& plan) {
  const int rk = bs.rank, nL = int(t.leaves.size());
  // kinds in the table cache's numbering: 0 colleague, 1 fine, 2 separated fine, 3 coarse
  static const int kSlots[4] = {9, 12, 20, 12};
  static const int kDl[4] = {0, 1, 1, -1};
  auto centres = [](int kind, int s) -> const double* {
    switch (kind) {
      case 0: return kColleagueCenters[s];
      case 1: return kFineCenters[s];
      case 2: return kSepFineCenters[s];
      default: return kCoarseCenters[s];
    }
  };
  auto slot_box = [](const Box& b, int kind, int s) {
    switch (kind) {
      case 0: return b.colleague[s];
      case 1: return b.fine[s];
      case 2: return b.sepfine[s];
      default: return b.coarse[s];
    }
  };
  // tables in use, keyed by (level, kind, slot)
  auto key = [](int l, int kind, int s) { return (l * 4 + kind) * 32 + s; };
  std::vector<int> need;
  for (int li = 0; li < nL; ++li) {
    const Box& b = t.box(t.leaves[li]);
    const int l = t.leaves[li].l;
    for (int kind = 0; kind < 4; ++kind)
      for (int s = 0; s < kSlots[kind]; ++s) {
        const int nk = slot_box(b, kind, s);
        if (nk < 0) continue;
        const int nl = l + kDl[kind];
        if (nl < 0 || nl > t.nLevels || nk >= int(t.lv[nl].size()) || !t.lv[nl][nk].leaf)
          return "near field: a neighbour slot of leaf " + std::to_string(li) + " is not a leaf";
        need.push_back(key(l, kind, s));
      }
  }
  std::sort(need.begin(), need.end());
  need.erase(std::unique(need.begin(), need.end()), need.end());
  std::vector<std::vector<cplx>> tabs(need.size());
  std::vector<std::string> errs(need.size());
  std::vector<NearGeom> geom(need.size());
  for (size_t q = 0; q < need.size(); ++q) {
    const int l = need[q] / 128, kind = (need[q] / 32) % 4, s = need[q] % 32;
    const double rA = t.boxRadius[l], rB = t.boxRadius[l + kDl[kind]];
    const double* c = centres(kind, s);
    geom[q] = NearGeom{rA, rB, c[0] * rA, c[1] * rA};
  }
  if (dev && opt.table_dir.empty()) {
    // every table's integrals in one device launch, assembled and rounded here
    std::vector<std::vector<cplx>> V1;
    std::string e = dev->near_integrals(bs, prob.kappa, geom, V1);
    if (!e.empty()) return "device tables: " + e;
    for (size_t q = 0; q < need.size(); ++q) tabs[q] = near_table_from_integrals(bs, geom[q].rB, V1[q]);
  } else {
#pragma omp parallel for schedule(dynamic, 1)
    for (int q = 0; q < int(need.size()); ++q) {
      const int l = need[q] / 128, kind = (need[q] / 32) % 4, s = need[q] % 32;
      if (opt.table_dir.empty())
        tabs[q] = near_table(bs, prob.kappa, geom[q].rA, geom[q].rB, geom[q].bx, geom[q].by);
      else
        errs[q] = read_neighbor_table(opt.table_dir, bs, prob.kappa, l, kind, s, tabs[q]);
    }
  }
  for (const std::string& e : errs)
    if (!e.empty()) return "tables: " + e;
  auto table = [&](int l, int kind, int s) -> const std::vector<cplx>& {
    return tabs[std::lower_bound(need.begin(), need.end(), key(l, kind, s)) - need.begin()];
  };

  PassHost ph;
  ph.name = "near";
  ph.level = t.nLevels;
  ph.kind = 0;
  int64_t op_total = int64_t(plan.ops.size()), src_total = int64_t(plan.src.size());
  for (int li = 0; li < nL; ++li) {
    const Box& b = t.box(t.leaves[li]);
    const int l = t.leaves[li].l;
    int ns = 0;
    for (int kind = 0; kind < 4; ++kind)
      for (int s = 0; s < kSlots[kind]; ++s) ns += slot_box(b, kind, s) >= 0;
    if (ns == 0) continue;
    const int nc = ns * rk;
    ph.rows.push_back(rk);
    ph.ncol.push_back(nc);
    ph.dst.push_back(t.N + li * rk);
    ph.op_off.push_back(op_total);
    ph.src_off.push_back(src_total);
    plan.ops.resize(size_t(op_total) + size_t(nc) * rk);
    plan.src.resize(size_t(src_total) + size_t(nc));
    cplx* op = plan.ops.data() + op_total;
    int* sp = plan.src.data() + src_total;
    int col = 0;
    // slot order of the reference: colleagues, fine, separated fine, coarse
    for (int kind = 0; kind < 4; ++kind)
      for (int s = 0; s < kSlots[kind]; ++s) {
        const int nk = slot_box(b, kind, s);
        if (nk < 0) continue;
        const int nleaf = t.lv[l + kDl[kind]][nk].leaf_index;
        const std::vector<cplx>& T = table(l, kind, s);   // row-major rk x rk
        for (int c = 0; c < rk; ++c) {
          sp[col + c] = nleaf * rk + c;
          for (int i = 0; i < rk; ++i) {
            cplx v = T[size_t(i) * rk + c];
            if (opt.find_phi) {
              v *= qk2[size_t(li) * rk + i];
              if (kind == 0 && s == 8 && i == c) v += 1.0;
            }
            op[size_t(col + c) * rk + i] = v;
          }
        }
        col += rk;
      }
    op_total += int64_t(nc) * rk;
    src_total += nc;
    ph.entries += int64_t(nc) * rk;
    ph.n_src += nc;
    ph.n_dst += rk;
  }
  plan.E_near = ph.entries;
  if (!ph.rows.empty()) plan.passes.push_back(std::move(ph));
  return "";
}

std::string build_plan(const Tree& t, const BasisSet& bs, const Problem& prob, const BuildOptions& opt, Plan& plan) {
  typedef std::chrono::steady_clock clk;
  auto since = [](clk::time_point a) { return std::chrono::duration<double>(clk::now() - a).count(); };
  double phase[4] = {0, 0, 0, 0};
  auto tp = clk::now();
  EntryKernel K;
  {
    std::string e = make_entry_kernel(t, bs, prob, opt.find_phi, opt.table_dir, K);
    if (!e.empty()) return "tables: " + e;
  }
  phase[0] = since(tp);
  tp = clk::now();
  Builder B(t, bs, K, std::pow(10, -1.0 * opt.TOL_POW));
  B.abs_exit = opt.aca_abs_exit;
  std::unique_ptr<DeviceNca> dev;
  if (opt.device) {
    std::string e;
    dev.reset(opt.device(K, e));
    if (!dev) return "device NCA: " + e;
    B.dev = dev.get();
  }
  {
    std::string e = B.run_nca();
    if (!e.empty()) return "NCA: " + e;
  }
  phase[1] = since(tp);
  tp = clk::now();

  plan = Plan();
  plan.far_leaf = K.M[t.nLevels];
  plan.N = t.N;
  plan.rank = bs.rank;
  plan.n_leaves = int(t.leaves.size());
  plan.find_phi = opt.find_phi;
  plan.qk2 = opt.find_phi ? K.qk2 : std::vector<double>(t.N, 0.0);
  B.next_slot = 2 * int64_t(t.N);
  B.assign_fout();
  auto dump = [&](int l, int k, int cone, const Owner& o) {
    std::vector<int>& d = plan.skeleton;
    d.insert(d.end(), {l, k, cone, o.out_active ? 1 : 0, o.in_active ? 1 : 0, int(o.out_check.size()),
                       int(o.out_charge.size()), int(o.in_check.size()), int(o.in_charge.size()), int(o.ucp.size())});
    for (const std::vector<int>* v : {&o.out_check, &o.out_charge, &o.in_check, &o.in_charge, &o.ucp})
      d.insert(d.end(), v->begin(), v->end());
    plan.aca.insert(plan.aca.end(), {l, k, cone, o.aca_out[0], o.aca_out[1], o.aca_out[2], o.aca_in[0], o.aca_in[1],
                                     o.aca_in[2]});
  };
  for (int l = t.nLevels; l >= 2; --l) {
    for (int k = 0; k < int(t.lv[l].size()); ++k)
      if (B.is_lfr(l, k)) dump(l, k, -1, B.L[l][k]);
    for (int k = 0; k < int(t.lv[l].size()); ++k)
      if (!B.is_lfr(l, k))
        for (int c = 0; c < t.nCones[l]; ++c) dump(l, k, c, B.hc(l, k, c));
  }

  plan.level_max_rank.assign(t.nLevels + 1, 0);
  auto note_rank = [&](int l, int r) {
    plan.level_max_rank[l] = std::max(plan.level_max_rank[l], r);
    plan.max_rank = std::max(plan.max_rank, r);
  };

  using Dest = Builder::Dest;
  std::vector<std::pair<PassHost, std::vector<Dest>>> passes;

  // upward passes: leaf level to level 2 (LFR_M2M then HFR_M2M)
  for (int j = t.nLevels; j >= 2; --j) {
    std::vector<Dest> ds;
    for (int k = 0; k < int(t.lv[j].size()); ++k) {
      if (B.is_lfr(j, k)) {
        Owner& o = B.L[j][k];
        note_rank(j, int(o.out_charge.size()) * (o.out_active ? 1 : 0));
        note_rank(j, int(o.in_charge.size()));
        if (o.out_active && !o.out_charge.empty()) {
          Dest d;
          d.dst = o.fout[0];
          d.rows = int(o.out_charge.size());
          d.P = &o.P;
          B.out_src(j, k, -1, d.up_src);
          ds.push_back(std::move(d));
        }
      } else {
        for (int c = 0; c < t.nCones[j]; ++c) {
          Owner& o = B.hc(j, k, c);
          note_rank(j, int(o.out_charge.size()) * (o.out_active ? 1 : 0));
          note_rank(j, int(o.in_charge.size()));
          if (o.out_active && !o.out_charge.empty()) {
            Dest d;
            d.dst = o.fout[0];
            d.rows = int(o.out_charge.size());
            d.P = &o.P;
            B.out_src(j, k, c, d.up_src);
            ds.push_back(std::move(d));
          }
        }
      }
    }
    PassHost ph;
    ph.name = "up_L" + std::to_string(j);
    ph.level = j;
    ph.kind = 0;
    if (!ds.empty()) passes.push_back({ph, std::move(ds)});
  }

  // downward passes: level 2 to leaves (M2L + L2L grouped by destination)
  std::vector<Dest> leaf_m2l, leaf_l2l;
  for (int j = 2; j <= t.nLevels; ++j) {
    std::vector<Dest> ds;
    for (int k = 0; k < int(t.lv[j].size()); ++k) {
      if (B.is_lfr(j, k)) B.add_m2l(j, k, -1, B.L[j][k]);
      else
        for (int c = 0; c < t.nCones[j]; ++c) B.add_m2l(j, k, c, B.hc(j, k, c));
    }
    for (int k = 0; k < int(t.lv[j].size()); ++k) {
      if (B.is_lfr(j, k)) B.resolve(j, k, -1, B.L[j][k], ds, leaf_m2l, leaf_l2l);
      else
        for (int c = 0; c < t.nCones[j]; ++c) B.resolve(j, k, c, B.hc(j, k, c), ds, leaf_m2l, leaf_l2l);
    }
    PassHost ph;
    ph.name = "down_L" + std::to_string(j);
    ph.level = j;
    ph.kind = 0;
    if (!ds.empty()) passes.push_back({ph, std::move(ds)});
  }
  {
    // leaf rows: M2L from the interaction lists, then L2L from the parents
    PassHost ph;
    ph.name = "leaf_m2l";
    ph.level = t.nLevels;
    ph.kind = 1;
    if (!leaf_m2l.empty()) passes.push_back({ph, std::move(leaf_m2l)});
    ph.name = "leaf_l2l";
    if (!leaf_l2l.empty()) passes.push_back({ph, std::move(leaf_l2l)});
  }
  plan.n_state = B.next_slot;

  // sizes and offsets
  int64_t op_total = 0, src_total = 0;
  for (auto& pr : passes) {
    PassHost& ph = pr.first;
    for (Dest& d : pr.second) {
      int64_t nc = 0;
      if (d.P) nc = int64_t(d.up_src.size());
      else
        for (const Contrib& c : *d.contribs) nc += contrib_cols(c);
      ph.rows.push_back(d.rows);
      ph.ncol.push_back(int(nc));
      ph.dst.push_back(int(d.dst));
      ph.op_off.push_back(op_total);
      ph.src_off.push_back(src_total);
      op_total += nc * d.rows;
      src_total += nc;
      ph.entries += nc * d.rows;
      ph.n_src += nc;
      ph.n_dst += d.rows;
    }
    plan.E_far += ph.entries;
  }
  plan.ops.resize(op_total);
  plan.src.resize(src_total);

  // fill operators: M2L blocks K(check points, charges) and L2L blocks
  // K Atilde_in^-1 (row solves with Atilde_in^T)
  const bool on_dev = B.dev != nullptr;
  for (auto& pr : passes) {
    PassHost& ph = pr.first;
    std::vector<Dest>& ds = pr.second;
    const int nd = int(ds.size());
    // on the device every entry block of the pass is generated in one batch,
    // in place; the L2L blocks are turned into K Atilde^-1 by the loop below
    if (on_dev) {
      std::vector<BlockJob> jobs;
      for (int di = 0; di < nd; ++di) {
        Dest& d = ds[di];
        if (d.P) continue;
        cplx* op = plan.ops.data() + ph.op_off[di];
        for (const Contrib& c : *d.contribs) {
          const int nc = int(c.cols->size());
          if (int(c.rows.size()) != d.rows) continue;
          jobs.push_back(BlockJob{c.rows.data(), d.rows, c.cols->data(), nc, op});
          op += size_t(nc) * d.rows;
        }
      }
      std::string e = B.dev->blocks(jobs);
      if (!e.empty()) return "device NCA: " + e;
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int di = 0; di < nd; ++di) {
      Dest& d = ds[di];
      cplx* op = plan.ops.data() + ph.op_off[di];
      int* sp = plan.src.data() + ph.src_off[di];
      const int nr = d.rows;
      if (d.P) {
        std::copy(d.P->begin(), d.P->end(), op);
        std::copy(d.up_src.begin(), d.up_src.end(), sp);
        continue;
      }
      std::vector<cplx> kb;
      for (const Contrib& c : *d.contribs) {
        const int nc = int(c.cols->size());
        if (int(c.rows.size()) != nr) continue;  // cannot happen: rows are the destination's ucp
        const cplx* kblk;
        if (on_dev) {
          kblk = op;  // read into Bm below before op is overwritten
        } else {
          kb.resize(size_t(nr) * nc);
          K.block(c.rows.data(), nr, c.cols->data(), nc, kb.data());
          kblk = kb.data();
        }
        if (c.At) {
          // device build: solved below for the whole pass in one GPU batch
          for (int q = 0; q < nc; ++q) sp[q] = int(c.src_base + q);
        } else if (c.Zt) {
          // op = K Atilde^-1: rows of op solve Atilde^T op_row^T = K_row^T
          std::vector<cplx> Bm(size_t(nc) * nr), X(size_t(nc) * nr);
          for (int r = 0; r < nr; ++r)
            for (int p = 0; p < nc; ++p) Bm[size_t(r) * nc + p] = kblk[size_t(p) * nr + r];
          c.Zt->solve(Bm.data(), nr, X.data());
          for (int q = 0; q < nc; ++q)
            for (int r = 0; r < nr; ++r) op[size_t(q) * nr + r] = X[size_t(r) * nc + q];
          for (int q = 0; q < nc; ++q) sp[q] = int(c.src_base + q);
        } else {
          if (!on_dev) std::copy(kb.begin(), kb.end(), op);
          std::copy(c.src_list->begin(), c.src_list->end(), sp);
        }
        op += size_t(nc) * nr;
        sp += nc;
      }
    }
    if (on_dev) {
      // L2L blocks of the device build: op = K Atilde_in^-1, i.e. X = Atilde_in^-T K^T
      // for the K block in op, one batched QR factor + solve on the GPU
      struct L2L { cplx* op; int nr, nc; const std::vector<cplx>* At; };
      std::vector<L2L> items;
      for (int di = 0; di < nd; ++di) {
        Dest& d = ds[di];
        if (d.P) continue;
        cplx* op = plan.ops.data() + ph.op_off[di];
        for (const Contrib& c : *d.contribs) {
          const int nc = int(c.cols->size());
          if (int(c.rows.size()) != d.rows) continue;
          if (c.At) items.push_back(L2L{op, d.rows, nc, c.At});
          op += size_t(nc) * d.rows;
        }
      }
      if (!items.empty()) {
        std::vector<std::vector<cplx>> Bm(items.size()), X(items.size());
#pragma omp parallel for schedule(dynamic, 16)
        for (int q = 0; q < int(items.size()); ++q) {
          const L2L& it = items[size_t(q)];
          Bm[size_t(q)].resize(size_t(it.nc) * it.nr);
          X[size_t(q)].resize(size_t(it.nc) * it.nr);
          for (int r = 0; r < it.nr; ++r)
            for (int p = 0; p < it.nc; ++p) Bm[size_t(q)][size_t(r) * it.nc + p] = it.op[size_t(p) * it.nr + r];
        }
        std::vector<QrSolveJob> qj(items.size());
        for (size_t q = 0; q < items.size(); ++q)
          qj[q] = QrSolveJob{items[q].At->data(), items[q].nc, Bm[q].data(), items[q].nr, X[q].data()};
        std::string e = B.dev->qr_solve(qj);
        if (!e.empty()) return "device NCA: " + e;
#pragma omp parallel for schedule(dynamic, 16)
        for (int q = 0; q < int(items.size()); ++q) {
          const L2L& it = items[size_t(q)];
          for (int c = 0; c < it.nc; ++c)
            for (int r = 0; r < it.nr; ++r) it.op[size_t(c) * it.nr + r] = X[size_t(q)][size_t(r) * it.nc + c];
        }
      }
    }
    plan.passes.push_back(std::move(ph));
  }

  phase[2] = since(tp);
  tp = clk::now();
  if (t.p.tree_mode == 1) {
    std::string e = adaptive_near_pass(t, bs, prob, opt, K.qk2, B.dev, plan);
    if (!e.empty()) return e;
    phase[3] = since(tp);
    for (int i = 0; i < 4; ++i) plan.phase_seconds[i] = phase[i];
    return "";
  }
  // near field: colleague tables of the leaf level (uniform tree)
  const int lvl = t.nLevels;
  const double rA = t.boxRadius[lvl];
  if (opt.table_dir.empty()) {
    plan.T.resize(size_t(9) * bs.rank * bs.rank);
    std::vector<NearGeom> geom(9);
    for (int s = 0; s < 9; ++s) geom[s] = NearGeom{rA, rA, kColleagueCenters[s][0] * rA, kColleagueCenters[s][1] * rA};
    std::vector<std::vector<cplx>> V1;
    if (B.dev) {
      std::string e = B.dev->near_integrals(bs, prob.kappa, geom, V1);
      if (!e.empty()) return "device tables: " + e;
    }
    for (int s = 0; s < 9; ++s) {
      std::vector<cplx> Ts = B.dev ? near_table_from_integrals(bs, rA, V1[s])
                                   : near_table(bs, prob.kappa, rA, rA, geom[s].bx, geom[s].by);
      std::copy(Ts.begin(), Ts.end(), plan.T.begin() + size_t(s) * bs.rank * bs.rank);
    }
  } else {
    std::string e = read_colleague_tables(opt.table_dir, bs, prob.kappa, lvl, plan.T);
    if (!e.empty()) return "tables: " + e;
  }
  plan.nbr.assign(size_t(plan.n_leaves) * 9, -1);
  for (int li = 0; li < plan.n_leaves; ++li) {
    const Box& b = t.box(t.leaves[li]);
    for (int s = 0; s < 9; ++s)
      if (b.colleague[s] >= 0) {
        plan.nbr[size_t(li) * 9 + s] = t.lv[t.leaves[li].l][b.colleague[s]].leaf_index;
        plan.E_near += int64_t(bs.rank) * bs.rank;
      }
  }
  phase[3] = since(tp);
  for (int i = 0; i < 4; ++i) plan.phase_seconds[i] = phase[i];
  return "";
}

}  // namespace lsg
