// Block (multi right-hand-side) matvec on FP64 MFMA and batched GMRES, gfx950.
#include "block.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <type_traits>
#include <vector>

namespace lsg {

namespace {

constexpr int kT = 256;
constexpr int kNB = kBlockRhs;      // 64 right-hand sides = 4 column tiles of 16
constexpr int kLD = 2 * kNB;        // doubles per block row
constexpr int kBMaxRows = 64;       // complex rows per destination (8 row tiles of 16 reals)
#ifndef LSG_SPLIT_COLS
#define LSG_SPLIT_COLS 256
#endif
#ifndef LSG_SPLIT_MAX_DEST
#define LSG_SPLIT_MAX_DEST 2048
#endif
constexpr int kSplitCols = LSG_SPLIT_COLS;        // columns per workgroup in split passes
constexpr int kSplitMaxDest = LSG_SPLIT_MAX_DEST; // passes with fewer destinations (and longer rows) are split

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return std::string(#expr ": ") + hipGetErrorString(e_); \
  } while (0)

// MFMA tile traits per precision: 16 x 16 x 4 products; the result lane l holds
// column l%16 and rows row(l/16, v), v < 4 (FP64: l/16 + 4v; FP32: 4(l/16) + v).
template <class R>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double acc __attribute__((ext_vector_type(4)));
  typedef double2 c2;
  static __device__ __forceinline__ acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int k4, int v) { return k4 + 4 * v; }
};
template <>
struct Mfma<float> {
  typedef float acc __attribute__((ext_vector_type(4)));
  typedef float2 c2;
  static __device__ __forceinline__ acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int k4, int v) { return 4 * k4 + v; }
};

// Workgroups are dispatched round-robin over the 8 XCDs. With kXcdGroup = G
// > 0, position b of nb is mapped so that each XCD takes runs of G
// consecutive positions (neighbouring destinations share source rows, which
// then meet in one XCD's L2) while the runs still interleave over the XCDs
// (destinations of unequal work stay spread over the chip); G = 0 keeps the
// dispatch order. A bijection on [0, nb).
#ifndef LSG_XCD_GROUP
#define LSG_XCD_GROUP 0
#endif
constexpr int kXcdGroup = LSG_XCD_GROUP;
__device__ __forceinline__ int xcd_order(int b, int nb) {
  if (kXcdGroup <= 0) return b;
  constexpr int span = 8 * (kXcdGroup > 0 ? kXcdGroup : 1);
  const int full = nb - nb % span;   // positions in complete spans
  if (b >= full) return b;
  const int x = b & 7, i = b >> 3;    // XCD, its i-th workgroup
  return (i / kXcdGroup) * span + x * kXcdGroup + i % kXcdGroup;
}

// The same with the run length G as an argument (G <= 0: dispatch order).
__device__ __forceinline__ int xcd_order_g(int b, int nb, int G) {
  if (G <= 0) return b;
  const int span = 8 * G;
  const int full = nb - nb % span;
  if (b >= full) return b;
  const int x = b & 7, i = b >> 3;
  return (i / G) * span + x * G + i % G;
}
// Leaves per XCD run in the leaf kernels: a leaf gathers the x rows of its 9
// colleagues, so leaves that share colleagues should meet in one XCD's L2.
#ifndef LSG_XCD_GROUP_LEAF
#define LSG_XCD_GROUP_LEAF 0
#endif
constexpr int kXcdGroupLeaf = LSG_XCD_GROUP_LEAF;

// complex operator columns per LDS chunk
#ifndef LSG_BLOCK_KC
#define LSG_BLOCK_KC 16
#endif
constexpr int kKC = LSG_BLOCK_KC;

// Right-hand-side tiles of 16 per wavefront in the block passes, NT: 1 (four
// wavefronts per workgroup) or 2 (two wavefronts, each operator value read
// from LDS feeds two MFMAs; chosen by timing for the FP32 operator, where it
// pays at C3 but not at C2). Threads per workgroup:
template <int NT>
constexpr int block_threads() { return 64 * (kNB / 16 / NT); }

typedef __attribute__((address_space(3))) void* lds_void_ptr;

// gfx950 LDS-direct loads: lane l of the wavefront copies 4 (16) bytes from g
// to wave_base + 4 l (16 l), wave_base the same for every lane; completion is
// counted by vmcnt.
__device__ __forceinline__ void lds_copy4(const void* g, void* wave_base) {
  __builtin_amdgcn_global_load_lds(g, (lds_void_ptr)wave_base, 4, 0, 0);
}
__device__ __forceinline__ void lds_copy16(const void* g, void* wave_base) {
  __builtin_amdgcn_global_load_lds(g, (lds_void_ptr)wave_base, 16, 0, 0);
}

// Operand stream of one workgroup: up to three segments, each a complex
// column-major block A_k (Rw rows, C_k columns) whose column c reads source
// row S_k[c]; walked in chunks of kKC columns.
template <class R>
struct Segments {
  const typename Mfma<R>::c2 *A0, *A1, *A2;
  const int *S0, *S1, *S2;
  int C0, C1, C2;
  int n0, n01, nch;  // chunk counts: segment 0, segments 0-1, all
};
// a segment as the workgroup reads it from LDS: operator, indices, columns, first chunk
struct SegDesc {
  const void* A;
  const int* S;
  int C, first;
};

// Block product of one workgroup on 64 right-hand sides: acc (wave w: RHS
// 16 w .. 16 w + 15) += sum over the segments of A_k X(rows S_k).
// Real-ified product [Yr; Yi] = [[Ar, -Ai], [Ai, Ar]] [Xr; Xi] with
// interleaved real rows 2r + a and real columns 2c + b; MFMA 16 x 16 x 4:
// lane l feeds A(real row 16 mt + l%16, real column 4 ks + l/16), i.e. the
// entry (row 8 mt + (l%16)/2, column 2 ks + (l/16)/2), Re when
// l%2 == (l/16)%2, else Im, times rm[mt] (sign -1 for (a, b) = (0, 1), 0 for
// rows past Rw, segment 0's row scale where asked), and X(real row
// 4 ks + l/16, RHS 16 w + l%16). Pipeline per chunk i: the operator columns
// of chunk i + 1 are copied into LDS by LDS-direct loads (double buffered,
// shared by the 4 waves), the column indices of chunk i + 2 into a second
// pair of LDS buffers, and each wave loads its source values of chunk i + 1
// into registers, while it runs the MFMAs of chunk i; one wait and one
// barrier per chunk. Columns past C_k read zeros (operator: the zero row of
// the pass state; source: row zrow).
template <class R, int MT, int MTU, int NT>
__device__ __forceinline__ void block_product(typename Mfma<R>::acc (&acc)[MT][NT], const Segments<R>& sg, int Rw,
                                              const R (&rm0)[MT], const R (&rm1)[MT], const R* __restrict__ X,
                                              const R* __restrict__ bst, int nx, int zrow, R* As0, R* As1,
                                              int* Sr, const SegDesc* sd) {
  constexpr int KS = kKC / 2;                        // MFMA k-steps per chunk
  constexpr int PB = sizeof(R) == 8 ? 16 : 4;        // bytes per operator copy (one complex FP64, or one float)
  constexpr int kBT = block_threads<NT>();
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int i16 = lane & 15, k4 = lane >> 4, kb = k4 & 1, kc = k4 >> 1;
  const int sel = (i16 & 1) == kb ? 0 : 1;
  int roff[MTU];
#pragma unroll
  for (int mt = 0; mt < MTU; ++mt) roff[mt] = 2 * min(8 * mt + (i16 >> 1), Rw - 1) + sel;
  const R* zero = bst + (size_t)zrow * kLD;
  const int per_col = 2 * Rw * int(sizeof(R)) / PB;  // copies per operator column
  const int n_cp = kKC * per_col;
  // segment of chunk i (descriptors in LDS, picked by a uniform index)
  struct Pick {
    const char* A;
    const int* S;
    int C, c0;
    bool seg0;
  };
  auto seg_of = [=](int i) __attribute__((always_inline)) {
    const SegDesc& d = sd[(i >= sg.n0 ? 1 : 0) + (i >= sg.n01 ? 1 : 0)];
    Pick p;
    p.A = reinterpret_cast<const char*>(d.A);
    p.S = d.S;
    p.C = d.C;
    p.c0 = (i - d.first) * kKC;
    p.seg0 = i < sg.n0;
    return p;
  };
  // operator columns of chunk i into As (every thread), indices into Sb (wave 0)
  auto copy_ops = [=](int i, R* As) __attribute__((always_inline)) {
    const Pick p = seg_of(i);
    const int c0 = p.c0, Ck = p.C;
    const char* src = p.A + (size_t)c0 * per_col * PB;
    const char* z = reinterpret_cast<const char*>(zero);
    const int valid = min(kKC, Ck - c0) * per_col;
    for (int q = 0; q * kBT < n_cp; ++q) {
      const int e = q * kBT + t;
      const char* g = e < valid ? src + (size_t)e * PB : z + (e & (kLD * int(sizeof(R)) / PB - 1)) * PB;
      char* wb = reinterpret_cast<char*>(As) + (size_t)(q * kBT + 64 * w) * PB;
      if (e < n_cp) {
        if (PB == 16) lds_copy16(g, wb);
        else lds_copy4(g, wb);
      }
    }
  };
  auto copy_idx = [=](int i, int* Sb) __attribute__((always_inline)) {
    if (w != 0 || i >= sg.nch) return;
    const Pick p = seg_of(i);
    const int c0 = p.c0, Ck = p.C;
    const int* Sk = p.S;
    if (lane < kKC) lds_copy4(Sk + min(c0 + lane, Ck - 1), Sb);
  };
  // this lane's source values of chunk i (indices in Sb)
  auto load_src = [=](int i, const int* Sb, R (&bv)[KS][NT]) __attribute__((always_inline)) {
    const Pick p = seg_of(i);
    const int c0 = p.c0, Ck = p.C;
    int sv[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) sv[ks] = Sb[2 * ks + kc];  // every index of the chunk was copied
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int r = c0 + 2 * ks + kc < Ck ? sv[ks] : zrow;
      const R* row = (r < nx ? X : bst) + (size_t)r * kLD + kb * kNB + 16 * NT * w + i16;
#pragma unroll
      for (int c = 0; c < NT; ++c) bv[ks][c] = row[16 * c];
    }
  };
  // index buffers: Sc holds chunk i + 1's indices, Sf receives chunk i + 2's
  int* Sc = Sr + kKC;
  int* Sf = Sr;
  R bcur[KS][NT], bnext[KS][NT];
  copy_idx(0, Sf);
  copy_idx(1, Sc);
  copy_ops(0, As0);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  __syncthreads();
  load_src(0, Sf, bcur);
  // nothing is in flight when the loop starts, so the first MFMAs wait only for LDS
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __syncthreads();
  for (int i = 0; i < sg.nch; ++i) {
    const bool odd = i & 1;
    const R* As = odd ? As1 : As0;
    const bool seg0 = i < sg.n0;
    if (i + 1 < sg.nch) copy_ops(i + 1, odd ? As0 : As1);
    copy_idx(i + 2, Sf);
    if (i + 1 < sg.nch) load_src(i + 1, Sc, bnext);
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      const int cl = 2 * ks + kc;
#pragma unroll
      for (int mt = 0; mt < MTU; ++mt) {
        const R a = As[cl * 2 * Rw + roff[mt]] * (seg0 ? rm0[mt] : rm1[mt]);
#pragma unroll
        for (int c = 0; c < NT; ++c) acc[mt][c] = Mfma<R>::op(a, bcur[ks][c], acc[mt][c]);
      }
    }
    // the MFMAs stay ahead of the wait for the next chunk's loads
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
#pragma unroll
      for (int c = 0; c < NT; ++c) bcur[ks][c] = bnext[ks][c];
    int* tmp = Sc;
    Sc = Sf;
    Sf = tmp;
  }
}

// The lane's operator multipliers: the real-ification sign, 0 for rows past
// Rw, times the row's scale (scale: per complex row, or null).
template <class R, int MT>
__device__ __forceinline__ void lane_rmul(R (&rmul)[MT], int Rw, int lane, const double* scale) {
  const int i16 = lane & 15, kb = (lane >> 4) & 1;
  const R sgn = ((i16 & 1) == 0 && kb == 1) ? R(-1) : R(1);
#pragma unroll
  for (int mt = 0; mt < MT; ++mt) {
    const int r = 8 * mt + (i16 >> 1);
    rmul[mt] = r < Rw ? (scale ? R(double(sgn) * scale[r]) : sgn) : R(0);
  }
}

// block_product over the row tiles that hold rows of the destinations (a
// workgroup-uniform choice); the LDS stage buffers are sized for MT tiles.
template <class R, int MT, int NT>
__device__ __forceinline__ void block_product_rows(typename Mfma<R>::acc (&acc)[MT][NT], const Segments<R>& sg, int Rw,
                                                   const R (&rm0)[MT], const R (&rm1)[MT], const R* __restrict__ X,
                                                   const R* __restrict__ bst, int nx, int zrow) {
  __shared__ R As0[2 * 8 * MT * kKC], As1[2 * 8 * MT * kKC];
  __shared__ int Sr[2 * kKC];
  __shared__ SegDesc sd[3];
  if (sg.nch <= 0) return;
  if (threadIdx.x == 0) {
    sd[0] = SegDesc{sg.A0, sg.S0, sg.C0, 0};
    sd[1] = SegDesc{sg.A1, sg.S1, sg.C1, sg.n0};
    sd[2] = SegDesc{sg.A2, sg.S2, sg.C2, sg.n01};
  }
  __syncthreads();
  const int mtu = (2 * Rw + 15) >> 4;
  if (mtu >= MT) {
    block_product<R, MT, MT, NT>(acc, sg, Rw, rm0, rm1, X, bst, nx, zrow, As0, As1, Sr, sd);
  } else if (MT > 1 && mtu == MT - 1) {
    block_product<R, MT, (MT > 1 ? MT - 1 : 1), NT>(acc, sg, Rw, rm0, rm1, X, bst, nx, zrow, As0, As1, Sr, sd);
  } else if (MT > 2 && mtu == MT - 2) {
    block_product<R, MT, (MT > 2 ? MT - 2 : 1), NT>(acc, sg, Rw, rm0, rm1, X, bst, nx, zrow, As0, As1, Sr, sd);
  } else {
    block_product<R, MT, (MT > 3 ? MT - 3 : 1), NT>(acc, sg, Rw, rm0, rm1, X, bst, nx, zrow, As0, As1, Sr, sd);
  }
}

// One pass of the plan on 64 right-hand sides: one workgroup per destination
// (or per column segment of a destination in split passes: kseg > 0, item =
// d * nseg + segment, partial products to part, summed by
// bsplit_reduce_kernel). R is the pass precision (operators, pass state,
// accumulation); Y is FP64. MT row tiles of 16 reals: 2 * (rows of every
// destination of the pass) <= 16 MT.
template <bool ACC, class R, int MT, int NT>
__global__ __launch_bounds__(block_threads<NT>()) void bpass_kernel(const typename Mfma<R>::c2* __restrict__ ops,
                                                   const int* __restrict__ src, R* bst, const R* __restrict__ X,
                                                   double* Y, int N, const int* __restrict__ rows,
                                                   const int* __restrict__ ncol, const int* __restrict__ dst,
                                                   const long long* __restrict__ op_off,
                                                   const long long* __restrict__ src_off, int nx, int zrow, int kseg,
                                                   int nseg, R* __restrict__ part) {
  typedef typename Mfma<R>::acc Acc;
  const int item = xcd_order(blockIdx.x, gridDim.x);
  const int d = kseg > 0 ? item / nseg : item;
  const int sgi = kseg > 0 ? item % nseg : 0;
  const int Rw = rows[d];
  const int Cfull = ncol[d];
  const int cbeg = kseg > 0 ? min(Cfull, sgi * kseg) : 0;
  const int C = kseg > 0 ? min(Cfull, cbeg + kseg) : Cfull;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  Segments<R> sg;
  sg.A0 = sg.A1 = sg.A2 = ops + op_off[d] + (size_t)cbeg * Rw;
  sg.S0 = sg.S1 = sg.S2 = src + src_off[d] + cbeg;
  sg.C0 = C - cbeg;
  sg.C1 = sg.C2 = 0;
  sg.n0 = sg.n01 = sg.nch = (sg.C0 + kKC - 1) / kKC;
  Acc acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int c = 0; c < NT; ++c) acc[mt][c] = Acc{0, 0, 0, 0};
  R rm[MT];
  lane_rmul<R, MT>(rm, Rw, lane, nullptr);
  block_product_rows<R, MT, NT>(acc, sg, Rw, rm, rm, X, bst, nx, zrow);
  const int i16 = lane & 15, k4 = lane >> 4;
  if (kseg > 0) {
    R* P = part + (size_t)item * (MT * 16) * kNB;
#pragma unroll
    for (int c = 0; c < NT; ++c) {
      const int q = 16 * (NT * w + c) + i16;
#pragma unroll
      for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int v = 0; v < 4; ++v) P[(size_t)((mt << 4) + Mfma<R>::row(k4, v)) * kNB + q] = acc[mt][c][v];
    }
    return;
  }
  const int dd = dst[d];
#pragma unroll
  for (int c = 0; c < NT; ++c) {
    const int q = 16 * (NT * w + c) + i16;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ri = (mt << 4) + Mfma<R>::row(k4, v);
        const int r = ri >> 1, a = ri & 1;
        if (r < Rw) {
          const int row = dd + r;
          if (row >= N && row < 2 * N) {
            double* p = Y + (size_t)(row - N) * kLD + a * kNB + q;
            double val = double(acc[mt][c][v]);
            if (ACC) val += *p;
            *p = val;
          } else {
            R* p = bst + (size_t)row * kLD + a * kNB + q;
            R val = acc[mt][c][v];
            if (ACC) val += *p;
            *p = val;
          }
        }
      }
    }
  }
}

// Output rows of the leaves of a uniform tree, one workgroup per leaf:
//   near field: the 9 colleague tables T (rank x 9 rank, shared by all leaves)
//   applied to the x rows of the colleagues (near_src: 9 rank rows per leaf,
//   the zero row for an absent colleague); find_phi: scaled row-wise by qk2
//   (applied to the table rows as they are loaded);
// then the leaf's destinations in the fused leaf passes (leaf_dest[2 leaf + k],
// -1 for none; rank rows at N + leaf rank) accumulated on top, and the rows
// stored once: Y = (find_phi ? X : 0) + qk2 .* (T x) + leaf passes (in FP64,
// or rounded to FP32 for the FP32 inner solves' Krylov product).
// Waves per SIMD asked of the FP64/FP32 leaf kernel with one right-hand-side
// tile per wavefront: 4 keeps it in 120 registers without spills (three waves
// without the request; C3 FP64 block matvec 5.68 ms against 5.99 ms).
#ifndef LSG_LEAF_WAVES
#define LSG_LEAF_WAVES 4
#endif
template <class R, class XT, class YT, int MT, int NT>
__global__ __attribute__((amdgpu_waves_per_eu(NT == 1 ? LSG_LEAF_WAVES : 1))) __launch_bounds__(block_threads<NT>())
void bleaf_kernel(const typename Mfma<R>::c2* __restrict__ T,
                                                   const int* __restrict__ near_src, int rank,
                                                   const typename Mfma<R>::c2* __restrict__ ops,
                                                   const int* __restrict__ src, const int* __restrict__ ncol,
                                                   const long long* __restrict__ op_off,
                                                   const long long* __restrict__ src_off,
                                                   const int* __restrict__ leaf_dest, const R* __restrict__ X,
                                                   const R* __restrict__ bst, int nx, const XT* __restrict__ Xin,
                                                   const double* __restrict__ qk2, int find_phi, YT* __restrict__ Y,
                                                   int zrow) {
  typedef typename Mfma<R>::acc Acc;
  const int leaf = xcd_order_g(blockIdx.x, gridDim.x, kXcdGroupLeaf);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  Segments<R> sg;
  sg.A0 = T;
  sg.S0 = near_src + (size_t)leaf * 9 * rank;
  sg.C0 = 9 * rank;
  const int d1 = leaf_dest[2 * leaf], d2 = leaf_dest[2 * leaf + 1];
  sg.A1 = d1 >= 0 ? ops + op_off[d1] : T;
  sg.S1 = d1 >= 0 ? src + src_off[d1] : sg.S0;
  sg.C1 = d1 >= 0 ? ncol[d1] : 0;
  sg.A2 = d2 >= 0 ? ops + op_off[d2] : T;
  sg.S2 = d2 >= 0 ? src + src_off[d2] : sg.S0;
  sg.C2 = d2 >= 0 ? ncol[d2] : 0;
  sg.n0 = (sg.C0 + kKC - 1) / kKC;
  sg.n01 = sg.n0 + (sg.C1 + kKC - 1) / kKC;
  sg.nch = sg.n01 + (sg.C2 + kKC - 1) / kKC;
  Acc acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int c = 0; c < NT; ++c) acc[mt][c] = Acc{0, 0, 0, 0};
  R rm0[MT], rm1[MT];
  lane_rmul<R, MT>(rm0, rank, lane, find_phi ? qk2 + (size_t)leaf * rank : nullptr);
  lane_rmul<R, MT>(rm1, rank, lane, nullptr);
  block_product_rows<R, MT, NT>(acc, sg, rank, rm0, rm1, X, bst, nx, zrow);
  const int i16 = lane & 15, k4 = lane >> 4;
#pragma unroll
  for (int c = 0; c < NT; ++c) {
    const int q = 16 * (NT * w + c) + i16;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int ri = (mt << 4) + Mfma<R>::row(k4, v);
        const int r = ri >> 1, a = ri & 1;
        if (r >= rank) continue;
        const size_t o = ((size_t)leaf * rank + r) * kLD + a * kNB + q;
        double val = double(acc[mt][c][v]);
        if (find_phi) val += double(Xin[o]);
        Y[o] = YT(val);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Split-BF16 block product (the inner solves' operator arithmetic): every FP32
// value a is carried as a = ah + al, ah = bf16(a), al = bf16(a - ah) (16
// significant bits, the pair within 2^-17 |a| of a), and a product as
// ah bh + ah bl + al bh + al bl on v_mfma_f32_16x16x32_bf16 with FP32
// accumulation: the exact product of the two pairs, summed in FP32. One MFMA
// takes K = 32 real columns, 16 complex ones: a sub-chunk of the operator.
//
// Real-ification with separate real and imaginary row tiles: for 16 complex
// destination rows (row tile p) and complex columns c,
//   Re y = sum_c [Re A, Im A] . [Re x, -Im x],  Im y = sum_c [Re A, Im A] . [Im x, Re x],
// so both tiles take the operator as stored (a word = bf16 Re in the low half,
// bf16 Im in the high half) and differ only in the source fragment. Lane l
// feeds A(row 16 p + l % 16, columns 4 (l / 16) .. + 3 of the chunk: eight bf16
// = four words) and B(the same four columns, right-hand side l % 16 of its
// tile); the result lane l holds rows 16 p + 4 (l / 16) + v, v < 4.
//
// Operator layout (built by to_sb_kernel): per destination, chunks of 16
// columns; a chunk is Rw rows of 128 bytes, the 16 columns' high words in
// bytes 0..63 and their low words in 64..127, columns past C zero. In LDS the
// eight 16-byte slots of row r are stored XOR-swizzled by r & 7, so the 16
// rows a wavefront reads per slot fall in distinct banks.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// two FP32 values rounded to bf16 (nearest even), packed: a in the low half
__device__ __forceinline__ unsigned pk_bf16(float a, float b) {
  typedef float f2 __attribute__((ext_vector_type(2)));
  typedef __bf16 b2 __attribute__((ext_vector_type(2)));
  return __builtin_bit_cast(unsigned, __builtin_convertvector(f2{a, b}, b2));
}
// high and low words of the pair (a, b)
__device__ __forceinline__ void split_pair(float a, float b, unsigned& hi, unsigned& lo) {
  hi = pk_bf16(a, b);
  lo = pk_bf16(a - __uint_as_float(hi << 16), b - __uint_as_float(hi & 0xffff0000u));
}
__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

__global__ void to_sb_kernel(const double2* __restrict__ ops, const long long* __restrict__ op_off,
                             const int* __restrict__ rows, const int* __restrict__ ncol,
                             const long long* __restrict__ opsb_off, unsigned* __restrict__ out) {
  const int d = blockIdx.x;
  const int Rw = rows[d], C = ncol[d], nk = (C + 15) / 16;
  const double2* A = ops + op_off[d];
  unsigned* O = out + opsb_off[d] * 32;
  for (int e = threadIdx.x; e < nk * Rw * 16; e += blockDim.x) {
    const int c16 = e & 15, rr = e >> 4;  // rr = chunk * Rw + row
    const int k = rr / Rw, r = rr - k * Rw, c = 16 * k + c16;
    const double2 a = c < C ? A[(size_t)c * Rw + r] : make_double2(0.0, 0.0);
    unsigned hi, lo;
    split_pair(float(a.x), float(a.y), hi, lo);
    O[(size_t)rr * 32 + c16] = hi;
    O[(size_t)rr * 32 + 16 + c16] = lo;
  }
}
// the near-field tables as one block (rank rows, 9 rank columns)
__global__ void to_sb_table_kernel(const double2* __restrict__ T, int Rw, int C, unsigned* __restrict__ out) {
  const int nk = (C + 15) / 16;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < (long long)nk * Rw * 16;
       e += (long long)gridDim.x * blockDim.x) {
    const int c16 = int(e & 15);
    const long long rr = e >> 4;
    const int k = int(rr / Rw), r = int(rr - (long long)k * Rw), c = 16 * k + c16;
    const double2 a = c < C ? T[(size_t)c * Rw + r] : make_double2(0.0, 0.0);
    unsigned hi, lo;
    split_pair(float(a.x), float(a.y), hi, lo);
    out[rr * 32 + c16] = hi;
    out[rr * 32 + 16 + c16] = lo;
  }
}

// Operand stream of one workgroup in the split-BF16 layout: up to three
// segments (operator block A_k, column indices S_k, C_k columns).
struct SegSb {
  const char *A0, *A1, *A2;
  const int *S0, *S1, *S2;
  int C0, C1, C2;
  int n0, n01, nch;  // chunk counts: segment 0, segments 0-1, all
};

// Sub-chunks of 16 columns (one MFMA k-step) per LDS chunk of the split-BF16
// product: more columns per chunk put more loads in flight per barrier.
#ifndef LSG_SB_SUB
#define LSG_SB_SUB 2
#endif
constexpr int kSbSub = LSG_SB_SUB;
// products per term: 4 (the exact product of the pairs) or 3 (al bl dropped:
// 2^-16 more error per product, 24 more inner steps of 456 on C3 x 64 angles)
#ifndef LSG_SB_TERMS
#define LSG_SB_TERMS 4
#endif
constexpr int kSbTerms = LSG_SB_TERMS;
constexpr int kSbCols = 16 * kSbSub;  // columns per chunk
__host__ __device__ constexpr int sb_chunks(int C) { return (C + kSbCols - 1) / kSbCols; }

// Block product of one workgroup on 64 right-hand sides in split BF16:
// are/aim[p][c] (wave w: right-hand sides 16 (NT w + c) .. + 15) += the real
// and imaginary rows of row tile p over all segments. With scale0, the rows are
// multiplied by scale0[row] once segment 0 is done (the leaf kernel's qk2 on
// the near field). Pipeline as in block_product: chunk i + 1's operator rows
// into LDS by LDS-direct loads (double buffered), chunk i + 2's indices into a
// second pair of buffers, chunk i + 1's source values into registers, while the
// MFMAs of chunk i run; one wait and one barrier per chunk. A chunk holds up
// to kSbSub sub-chunks of 16 columns, stored in LDS as consecutive blocks of
// Rw rows (row rr of the chunk: sub-chunk rr / Rw).
template <int MTP, int NT, class XS>
__device__ __forceinline__ void block_product_sb(f32x4 (&are)[MTP][NT], f32x4 (&aim)[MTP][NT], const SegSb& sg,
                                                 int Rw, const XS* __restrict__ X, const float* __restrict__ bst,
                                                 int nx, const float (&scale0)[MTP][4], bool scaled, char* As0,
                                                 char* As1, int* Sr, const SegDesc* sd) {
  constexpr int kBT = block_threads<NT>();
  constexpr int KS = kSbSub;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int i16 = lane & 15, g = lane >> 4;
  struct Pick {
    const char* A;
    const int* S;
    int C, c0, nsub;
  };
  auto seg_of = [=](int i) __attribute__((always_inline)) {
    const SegDesc& d = sd[(i >= sg.n0 ? 1 : 0) + (i >= sg.n01 ? 1 : 0)];
    Pick p;
    p.A = reinterpret_cast<const char*>(d.A);
    p.S = d.S;
    p.C = d.C;
    p.c0 = (i - d.first) * kSbCols;
    p.nsub = min(KS, (d.C - p.c0 + 15) >> 4);
    return p;
  };
  auto copy_ops = [=](int i, char* As) __attribute__((always_inline)) {
    const Pick p = seg_of(i);
    const char* src = p.A + (size_t)(p.c0 >> 4) * Rw * 128;
    const int n_cp = 8 * Rw * p.nsub;  // 16-byte slots of the chunk's sub-chunks
    for (int q = 0; q * kBT < 8 * Rw * KS; ++q) {
      const int e = q * kBT + t;
      const int r = e >> 3, sl = (e & 7) ^ (r & 7);
      char* wb = As + (size_t)(q * kBT + 64 * w) * 16;
      if (e < n_cp) lds_copy16(src + (size_t)(8 * r + sl) * 16, wb);
    }
  };
  auto copy_idx = [=](int i, int* Sb) __attribute__((always_inline)) {
    if (w != 0 || i >= sg.nch) return;
    const Pick p = seg_of(i);
    if (lane < kSbCols) lds_copy4(p.S + min(p.c0 + lane, p.C - 1), Sb);
  };
  // this lane's source values of chunk i: columns 16 s + 4 g .. + 3 of each
  // sub-chunk s, right-hand sides of its NT tiles (real and imaginary parts)
  auto load_src = [=](int i, const int* Sb, float (&xv)[NT][4 * KS][2]) __attribute__((always_inline)) {
    const Pick p = seg_of(i);
#pragma unroll
    for (int sb = 0; sb < KS; ++sb) {
      if (sb >= p.nsub) break;
      const int4 sv = *reinterpret_cast<const int4*>(Sb + 16 * sb + 4 * g);
      const int s4[4] = {sv.x, sv.y, sv.z, sv.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const size_t o = (size_t)s4[j] * kLD + 16 * NT * w + i16;
        if (s4[j] < nx) {
#pragma unroll
          for (int c = 0; c < NT; ++c) {
            xv[c][4 * sb + j][0] = float(X[o + 16 * c]);
            xv[c][4 * sb + j][1] = float(X[o + kNB + 16 * c]);
          }
        } else {
#pragma unroll
          for (int c = 0; c < NT; ++c) {
            xv[c][4 * sb + j][0] = bst[o + 16 * c];
            xv[c][4 * sb + j][1] = bst[o + kNB + 16 * c];
          }
        }
      }
    }
  };
  // chunk c's indices live in index buffer c & 1, its operator rows in As0/As1 by the same parity
  typedef float XBuf[NT][4 * KS][2];
  XBuf xa, xb;
  copy_idx(0, Sr);
  copy_idx(1, Sr + kSbCols);
  copy_ops(0, As0);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  __syncthreads();
  load_src(0, Sr, xa);
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __syncthreads();
  // chunk i: its source values in xc; chunk i + 1's are loaded into xn
  auto step = [&](int i, XBuf& xc, XBuf& xn) __attribute__((always_inline)) {
    const char* As = (i & 1) ? As1 : As0;
    const int nsub = seg_of(i).nsub;
    if (i + 1 < sg.nch) copy_ops(i + 1, (i & 1) ? As0 : As1);
    copy_idx(i + 2, Sr + kSbCols * (i & 1));
    if (i + 1 < sg.nch) load_src(i + 1, Sr + kSbCols * ((i + 1) & 1), xn);
#pragma unroll
    for (int sb = 0; sb < KS; ++sb) {
      if (sb >= nsub) break;
      // source fragments: [Re x, -Im x] for the real rows, [Im x, Re x] for the imaginary rows
      u32x4 b1h[NT], b1l[NT], b2h[NT], b2l[NT];
#pragma unroll
      for (int c = 0; c < NT; ++c)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          unsigned hi, lo;
          split_pair(xc[c][4 * sb + j][0], -xc[c][4 * sb + j][1], hi, lo);
          b1h[c][j] = hi;
          b1l[c][j] = lo;
          // [Im x, Re x] = the halves of [Re x, -Im x] swapped, the new low half negated
          b2h[c][j] = __builtin_amdgcn_alignbit(hi, hi, 16) ^ 0x8000u;
          b2l[c][j] = __builtin_amdgcn_alignbit(lo, lo, 16) ^ 0x8000u;
        }
#pragma unroll
      for (int p = 0; p < MTP; ++p) {
        const int r = sb * Rw + min(16 * p + i16, Rw - 1);  // rows past Rw compute unused results
        const char* rowp = As + r * 128;
        const int sw = r & 7;
        const bf16x8 ah = as_bf16x8(*reinterpret_cast<const u32x4*>(rowp + ((g ^ sw) << 4)));
        const bf16x8 al = as_bf16x8(*reinterpret_cast<const u32x4*>(rowp + (((4 + g) ^ sw) << 4)));
#pragma unroll
        for (int c = 0; c < NT; ++c) {
          are[p][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, as_bf16x8(b1h[c]), are[p][c], 0, 0, 0);
          aim[p][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, as_bf16x8(b2h[c]), aim[p][c], 0, 0, 0);
          are[p][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, as_bf16x8(b1l[c]), are[p][c], 0, 0, 0);
          aim[p][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, as_bf16x8(b2l[c]), aim[p][c], 0, 0, 0);
          are[p][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, as_bf16x8(b1h[c]), are[p][c], 0, 0, 0);
          aim[p][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, as_bf16x8(b2h[c]), aim[p][c], 0, 0, 0);
          if (kSbTerms == 4) {
            are[p][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, as_bf16x8(b1l[c]), are[p][c], 0, 0, 0);
            aim[p][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, as_bf16x8(b2l[c]), aim[p][c], 0, 0, 0);
          }
        }
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
  };
  // chunks [i0, i1) two at a time (the buffers swap roles without copies);
  // on return xa holds chunk i1's source values
  auto run = [&](int i0, int i1) __attribute__((always_inline)) {
    int i = i0;
    for (; i + 1 < i1; i += 2) {
      step(i, xa, xb);
      step(i + 1, xb, xa);
    }
    if (i < i1) {
      step(i, xa, xb);
#pragma unroll
      for (int c = 0; c < NT; ++c)
#pragma unroll
        for (int k = 0; k < 4 * KS; ++k) {
          xa[c][k][0] = xb[c][k][0];
          xa[c][k][1] = xb[c][k][1];
        }
    }
  };
  // segment 0, its row scaling (outside the loops: the accumulators stay in
  // the MFMA registers), then the other segments
  run(0, sg.n0);
  if (scaled) {
#pragma unroll
    for (int p = 0; p < MTP; ++p)
#pragma unroll
      for (int c = 0; c < NT; ++c)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          are[p][c][v] *= scale0[p][v];
          aim[p][c][v] *= scale0[p][v];
        }
  }
  run(sg.n0, sg.nch);
}

// LDS of one split-BF16 workgroup: two chunk buffers of up to 64 rows per
// sub-chunk, the index buffers and the segment descriptors
template <int MTP, int NT, class XS>
__device__ __forceinline__ void block_product_sb_lds(f32x4 (&are)[MTP][NT], f32x4 (&aim)[MTP][NT], const SegSb& sg,
                                                     int Rw, const XS* __restrict__ X,
                                                     const float* __restrict__ bst, int nx,
                                                     const float (&scale0)[MTP][4], bool scaled) {
  __shared__ __attribute__((aligned(16))) char As0[kSbSub * 16 * MTP * 128], As1[kSbSub * 16 * MTP * 128];
  __shared__ __attribute__((aligned(16))) int Sr[2 * kSbCols];
  __shared__ SegDesc sd[3];
  if (sg.nch <= 0) return;
  if (threadIdx.x == 0) {
    sd[0] = SegDesc{sg.A0, sg.S0, sg.C0, 0};
    sd[1] = SegDesc{sg.A1, sg.S1, sg.C1, sg.n0};
    sd[2] = SegDesc{sg.A2, sg.S2, sg.C2, sg.n01};
  }
  __syncthreads();
  block_product_sb<MTP, NT, XS>(are, aim, sg, Rw, X, bst, nx, scale0, scaled, As0, As1, Sr, sd);
}

// One pass on 64 right-hand sides in split BF16 (as bpass_kernel<ACC, float,
// ...>): one workgroup per destination or per column segment of one; MTP row
// tiles of 16 complex rows (rows of every destination <= 16 MTP); opsb_off in
// 128-byte rows of the split-BF16 operator; XS: type of the input block's
// rows (the first nx source rows), FP32 or the inner solves' FP16 basis.
// Waves per SIMD asked of the split-BF16 kernels with one right-hand-side tile
// per wavefront: the compiler then keeps the accumulators in VGPRs and fits the
// leaf kernel in 128 registers without spills, four waves per SIMD (two
// without the request; C3 block matvec 1.75 ms against 2.20 ms).
#ifndef LSG_SB_WAVES
#define LSG_SB_WAVES 4
#endif
#define LSG_SB_ATTR __attribute__((amdgpu_waves_per_eu(NT == 1 ? LSG_SB_WAVES : 1)))
template <bool ACC, int MTP, int NT, class XS>
__global__ LSG_SB_ATTR __launch_bounds__(block_threads<NT>()) void bpass_sb_kernel(
    const char* __restrict__ ops_sb, const int* __restrict__ src, float* bst, const XS* __restrict__ X, double* Y,
    int N, const int* __restrict__ rows, const int* __restrict__ ncol, const int* __restrict__ dst,
    const long long* __restrict__ opsb_off, const long long* __restrict__ src_off, int nx, int kseg, int nseg,
    int rt, float* __restrict__ part) {
  const int item = xcd_order(blockIdx.x, gridDim.x);
  const int d = kseg > 0 ? item / nseg : item;
  const int sgi = kseg > 0 ? item % nseg : 0;
  const int Rw = rows[d];
  const int Cfull = ncol[d];
  const int cbeg = kseg > 0 ? min(Cfull, sgi * kseg) : 0;
  const int C = kseg > 0 ? min(Cfull, cbeg + kseg) : Cfull;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  SegSb sg;
  sg.A0 = sg.A1 = sg.A2 = ops_sb + (opsb_off[d] + (long long)(cbeg >> 4) * Rw) * 128;
  sg.S0 = sg.S1 = sg.S2 = src + src_off[d] + cbeg;
  sg.C0 = C - cbeg;
  sg.C1 = sg.C2 = 0;
  sg.n0 = sg.n01 = sg.nch = sb_chunks(sg.C0);
  f32x4 are[MTP][NT], aim[MTP][NT];
#pragma unroll
  for (int p = 0; p < MTP; ++p)
#pragma unroll
    for (int c = 0; c < NT; ++c) are[p][c] = aim[p][c] = f32x4{0, 0, 0, 0};
  const float none[MTP][4] = {};
  block_product_sb_lds<MTP, NT, XS>(are, aim, sg, Rw, X, bst, nx, none, false);
  const int i16 = lane & 15, g = lane >> 4;
#pragma unroll
  for (int c = 0; c < NT; ++c) {
    const int q = 16 * (NT * w + c) + i16;
#pragma unroll
    for (int p = 0; p < MTP; ++p)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * p + 4 * g + v;
        if (r >= Rw) continue;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const float val = a ? aim[p][c][v] : are[p][c][v];
          if (kseg > 0) {
            part[((size_t)item * rt + 2 * r + a) * kNB + q] = val;
            continue;
          }
          const int row = dst[d] + r;
          if (row >= N && row < 2 * N) {
            double* yp = Y + (size_t)(row - N) * kLD + a * kNB + q;
            double y = double(val);
            if (ACC) y += *yp;
            *yp = y;
          } else {
            float* sp = bst + (size_t)row * kLD + a * kNB + q;
            float y = val;
            if (ACC) y += *sp;
            *sp = y;
          }
        }
      }
  }
}

// The leaf kernel in split BF16 (as bleaf_kernel<float, XT, YT, ...>): one
// workgroup per leaf, near field (scaled by qk2 when find_phi), leaf M2L and
// leaf L2L, x term, rows stored once (part 0). Part 1 is the near field and x
// term alone (stored), part 2 the leaf passes alone (added to the stored rows),
// so that the near field can run beside the upward and downward passes.
template <class XS, class XT, class YT, int MTP, int NT>
__global__ LSG_SB_ATTR __launch_bounds__(block_threads<NT>()) void bleaf_sb_kernel(
    const char* __restrict__ T_sb, const int* __restrict__ near_src, int rank, const char* __restrict__ ops_sb,
    const int* __restrict__ src, const int* __restrict__ ncol, const long long* __restrict__ opsb_off,
    const long long* __restrict__ src_off, const int* __restrict__ leaf_dest, const XS* __restrict__ X,
    const float* __restrict__ bst, int nx, const XT* __restrict__ Xin, const double* __restrict__ qk2, int find_phi,
    YT* __restrict__ Y, int part) {
  const int leaf = xcd_order_g(blockIdx.x, gridDim.x, kXcdGroupLeaf);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i16 = lane & 15, g = lane >> 4;
  SegSb sg;
  sg.A0 = T_sb;
  sg.S0 = near_src + (size_t)leaf * 9 * rank;
  sg.C0 = part == 2 ? 0 : 9 * rank;
  const int d1 = part == 1 ? -1 : leaf_dest[2 * leaf], d2 = part == 1 ? -1 : leaf_dest[2 * leaf + 1];
  sg.A1 = d1 >= 0 ? ops_sb + opsb_off[d1] * 128 : T_sb;
  sg.S1 = d1 >= 0 ? src + src_off[d1] : sg.S0;
  sg.C1 = d1 >= 0 ? ncol[d1] : 0;
  sg.A2 = d2 >= 0 ? ops_sb + opsb_off[d2] * 128 : T_sb;
  sg.S2 = d2 >= 0 ? src + src_off[d2] : sg.S0;
  sg.C2 = d2 >= 0 ? ncol[d2] : 0;
  if (part == 2 && sg.C1 + sg.C2 == 0) return;  // nothing to add
  sg.n0 = sb_chunks(sg.C0);
  sg.n01 = sg.n0 + sb_chunks(sg.C1);
  sg.nch = sg.n01 + sb_chunks(sg.C2);
  f32x4 are[MTP][NT], aim[MTP][NT];
#pragma unroll
  for (int p = 0; p < MTP; ++p)
#pragma unroll
    for (int c = 0; c < NT; ++c) are[p][c] = aim[p][c] = f32x4{0, 0, 0, 0};
  float sc[MTP][4];
#pragma unroll
  for (int p = 0; p < MTP; ++p)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = 16 * p + 4 * g + v;
      sc[p][v] = (find_phi && r < rank) ? float(qk2[(size_t)leaf * rank + r]) : 1.0f;
    }
  block_product_sb_lds<MTP, NT, XS>(are, aim, sg, rank, X, bst, nx, sc, find_phi != 0 && part != 2);
#pragma unroll
  for (int c = 0; c < NT; ++c) {
    const int q = 16 * (NT * w + c) + i16;
#pragma unroll
    for (int p = 0; p < MTP; ++p)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = 16 * p + 4 * g + v;
        if (r >= rank) continue;
        const size_t o = ((size_t)leaf * rank + r) * kLD + q;
        double vr = double(are[p][c][v]), vi = double(aim[p][c][v]);
        if (part == 2) {
          vr += double(Y[o]);
          vi += double(Y[o + kNB]);
        } else if (find_phi) {
          vr += double(Xin[o]);
          vi += double(Xin[o + kNB]);
        }
        Y[o] = YT(vr);
        Y[o + kNB] = YT(vi);
      }
  }
}

// Sum of the nseg partial products of destination blockIdx.x (tiles of rt real
// rows x 64), stored or accumulated into its rows (Y or the pass state).
template <bool ACC, class R>
__global__ __launch_bounds__(kT) void bsplit_reduce_kernel(const R* __restrict__ part, int nseg, int rt,
                                                           const int* __restrict__ rows, const int* __restrict__ dst,
                                                           R* bst, double* Y, int N) {
  const int d = blockIdx.x;
  const int Rw = rows[d], dd = dst[d];
  const R* P = part + (size_t)d * nseg * rt * kNB;
  for (int e = threadIdx.x; e < 2 * Rw * kNB; e += kT) {
    const int ri = e / kNB, q = e - ri * kNB;
    R s = 0;
    for (int sg = 0; sg < nseg; ++sg) s += P[((size_t)sg * rt + ri) * kNB + q];
    const int r = ri >> 1, a = ri & 1, row = dd + r;
    if (row >= N && row < 2 * N) {
      double* p = Y + (size_t)(row - N) * kLD + a * kNB + q;
      double val = double(s);
      if (ACC) val += *p;
      *p = val;
    } else {
      R* p = bst + (size_t)row * kLD + a * kNB + q;
      R val = s;
      if (ACC) val += *p;
      *p = val;
    }
  }
}

// Y = X + qk2c .* Y with a contrast per right-hand side (qk2c[n * 64 + q])
template <class XT>
__global__ void bdiag_kernel(const XT* __restrict__ X, double* __restrict__ Y, const double* __restrict__ qk2c,
                             long long n_total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_total) Y[i] = double(X[i]) + qk2c[(i / kLD) * kNB + (i % kNB)] * Y[i];
}

// ------------------------------------------------------- batched GMRES ----
constexpr int kBRows = 256;   // rows per workgroup in the block dot
constexpr int kBCols = 8;     // Krylov columns per workgroup in the block dot
constexpr int kRS = 16;       // first-stage segments of the row-block sums
constexpr int kUpdRows = 8;   // rows per thread (and system group) in the basis updates

// Basis loads of the Gram-Schmidt kernels: non-temporal by default (the
// streamed basis then leaves the caches to the product, which the dot reads
// once per group of columns; C3 batch 44.6 against 44.1 solves/s with the FP16
// kernels), or plain (LSG_GS_NT 0).
#ifndef LSG_GS_NT
#define LSG_GS_NT 1
#endif
template <class T>
__device__ __forceinline__ T gs_load(const T* p) {
  if (LSG_GS_NT) return __builtin_nontemporal_load(p);
  return *p;
}


// partial[((rb * ncx) + col) * 64 + q] = sum over the workgroup's rows of
// conj(V_col[n, q]) w[n, q] (col < nc) or |w[n, q]|^2 (col == nc < ncx).
template <class TV, class TW>
__global__ __launch_bounds__(kT) void bdot_kernel(const TV* __restrict__ V, long long vs, int nc, int ncx,
                                                  const TW* __restrict__ w, int N, double2* __restrict__ partial) {
  __shared__ double2 red[kT / 64][kBCols][kNB];
  const int t = threadIdx.x, q = t & 63, rl = t >> 6;
  const int g0 = blockIdx.y * kBCols;
  const long long n0 = (long long)blockIdx.x * kBRows;
  double2 acc[kBCols];
#pragma unroll
  for (int u = 0; u < kBCols; ++u) acc[u] = make_double2(0.0, 0.0);
  for (int r = rl; r < kBRows; r += kT / 64) {
    const long long n = n0 + r;
    if (n >= N) break;
    const double wr = double(w[n * kLD + q]), wi = double(w[n * kLD + kNB + q]);
#pragma unroll
    for (int u = 0; u < kBCols; ++u) {
      const int col = g0 + u;
      double vr = 0.0, vi = 0.0;
      if (col < nc) {
        vr = double(gs_load(V + col * vs + n * kLD + q));
        vi = double(gs_load(V + col * vs + n * kLD + kNB + q));
      } else if (col == nc) {
        vr = wr;
        vi = wi;
      }
      acc[u].x = fma(vr, wr, fma(vi, wi, acc[u].x));
      acc[u].y = fma(vr, wi, fma(-vi, wr, acc[u].y));
    }
  }
#pragma unroll
  for (int u = 0; u < kBCols; ++u) red[rl][u][q] = acc[u];
  __syncthreads();
  for (int e = t; e < kBCols * kNB; e += kT) {
    const int u = e / kNB, qq = e % kNB;
    if (g0 + u < ncx) {
      double2 s = red[0][u][qq];
      for (int k = 1; k < kT / 64; ++k) {
        s.x += red[k][u][qq].x;
        s.y += red[k][u][qq].y;
      }
      partial[((long long)blockIdx.x * ncx + g0 + u) * kNB + qq] = s;
    }
  }
}

// The block dot for the inner solves' FP16 basis and FP32 product: lane l of a
// wavefront covers the systems 4 (l % 16) .. + 3 with one 8-byte load per
// basis value (16 for the product) and the quarter-wave l / 16 takes one of
// four consecutive rows, so one load instruction moves four rows' 512 bytes of
// the basis (its rows are 256 bytes: 64 real parts, then 64 imaginary parts).
// Each thread sums its 16 rows in FP32 (the basis is FP16: these sums are far
// below its rounding), the quarter-waves, wavefronts and row blocks in FP64.
// Same partial layout as bdot_kernel.
// (a request of 4 waves spills 24 registers and made the C3 batch 10% slower)
#ifndef LSG_BDOT_WAVES
#define LSG_BDOT_WAVES 1
#endif
__global__ __attribute__((amdgpu_waves_per_eu(LSG_BDOT_WAVES))) __launch_bounds__(kT) void bdot_h_kernel(const _Float16* __restrict__ V, long long vs, int nc, int ncx,
                                                    const float* __restrict__ w, int N, double2* __restrict__ partial) {
  typedef _Float16 h4 __attribute__((ext_vector_type(4)));
  __shared__ double2 red[kT / 64][kBCols][kNB];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6, rs = lane >> 4, sg = lane & 15;
  const int g0 = blockIdx.y * kBCols;
  const long long n0 = (long long)blockIdx.x * kBRows;
  float2 acc[kBCols][4];
#pragma unroll
  for (int u = 0; u < kBCols; ++u)
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[u][e] = make_float2(0.f, 0.f);
#ifndef LSG_BDOT_UNROLL
#define LSG_BDOT_UNROLL 1
#endif
#if LSG_BDOT_UNROLL == 2
#pragma unroll 2
#endif
  for (int k = 0; k < kBRows / 16; ++k) {
    const long long n = n0 + 16 * k + 4 * wv + rs;
    if (n >= N) break;
    const float4 wr = *reinterpret_cast<const float4*>(w + n * kLD + 4 * sg);
    const float4 wi = *reinterpret_cast<const float4*>(w + n * kLD + kNB + 4 * sg);
    const float wre[4] = {wr.x, wr.y, wr.z, wr.w}, wim[4] = {wi.x, wi.y, wi.z, wi.w};
#pragma unroll
    for (int u = 0; u < kBCols; ++u) {
      const int col = g0 + u;
      float vre[4] = {0.f, 0.f, 0.f, 0.f}, vim[4] = {0.f, 0.f, 0.f, 0.f};
      if (col < nc) {
        const h4 a = gs_load(reinterpret_cast<const h4*>(V + col * vs + n * kLD + 4 * sg));
        const h4 b = gs_load(reinterpret_cast<const h4*>(V + col * vs + n * kLD + kNB + 4 * sg));
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          vre[e] = float(a[e]);
          vim[e] = float(b[e]);
        }
      } else if (col == nc) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          vre[e] = wre[e];
          vim[e] = wim[e];
        }
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        acc[u][e].x = fmaf(vre[e], wre[e], fmaf(vim[e], wim[e], acc[u][e].x));
        acc[u][e].y = fmaf(vre[e], wim[e], fmaf(-vim[e], wre[e], acc[u][e].y));
      }
    }
  }
#pragma unroll
  for (int u = 0; u < kBCols; ++u)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double x = double(acc[u][e].x), y = double(acc[u][e].y);
      x += __shfl_xor(x, 16);
      y += __shfl_xor(y, 16);
      x += __shfl_xor(x, 32);
      y += __shfl_xor(y, 32);
      if (rs == 0) red[wv][u][4 * sg + e] = make_double2(x, y);
    }
  __syncthreads();
  for (int e = t; e < kBCols * kNB; e += kT) {
    const int u = e / kNB, qq = e % kNB;
    if (g0 + u < ncx) {
      double2 sm = red[0][u][qq];
      for (int k = 1; k < kT / 64; ++k) {
        sm.x += red[k][u][qq].x;
        sm.y += red[k][u][qq].y;
      }
      partial[((long long)blockIdx.x * ncx + g0 + u) * kNB + qq] = sm;
    }
  }
}

// The update for the FP16 basis and FP32 product: v_out = (w - sum_i V_i h_i) /
// hn; lane l covers the systems 4 (l % 16) .. + 3 (8-byte basis loads) of rows
// n0 + 4 r + l / 16, r < kUpdRows, so one load instruction moves four rows.
// The projection sums run in FP32 (the result is rounded to FP16); xout32
// receives the FP32 value of every stored entry (the next matvec's input).
__global__ __launch_bounds__(kT) void bupdate_h_kernel(const _Float16* __restrict__ V, long long vs, int nc,
                                                       const double2* __restrict__ h, const float* __restrict__ w,
                                                       const double* __restrict__ hn, _Float16* __restrict__ vout,
                                                       float* __restrict__ xout32, long long n_rows) {
  typedef _Float16 h4 __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x & 63, sg = lane & 15, rs = lane >> 4;
  const long long n0 = ((long long)blockIdx.x * (kT / 64) + (threadIdx.x >> 6)) * (4 * kUpdRows);
  float2 acc[kUpdRows][4];
#pragma unroll
  for (int r = 0; r < kUpdRows; ++r)
#pragma unroll
    for (int e = 0; e < 4; ++e) acc[r][e] = make_float2(0.f, 0.f);
  for (int i = 0; i < nc; ++i) {
    float hr[4], hi[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double2 hh = h[i * kNB + 4 * sg + e];
      hr[e] = float(hh.x);
      hi[e] = float(hh.y);
    }
    const _Float16* Vi = V + i * vs + 4 * sg;
#pragma unroll
    for (int r = 0; r < kUpdRows; ++r) {
      const long long n = n0 + 4 * r + rs;
      if (n < n_rows) {
        const h4 a = gs_load(reinterpret_cast<const h4*>(Vi + n * kLD));
        const h4 b = gs_load(reinterpret_cast<const h4*>(Vi + n * kLD + kNB));
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float vr = float(a[e]), vi = float(b[e]);
          acc[r][e].x = fmaf(vr, hr[e], fmaf(-vi, hi[e], acc[r][e].x));
          acc[r][e].y = fmaf(vr, hi[e], fmaf(vi, hr[e], acc[r][e].y));
        }
      }
    }
  }
  float inv[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const double d = hn[4 * sg + e];
    inv[e] = d > 0.0 ? float(1.0 / d) : 0.0f;
  }
#pragma unroll
  for (int r = 0; r < kUpdRows; ++r) {
    const long long n = n0 + 4 * r + rs;
    if (n < n_rows) {
      const long long o = n * kLD + 4 * sg;
      const float4 wr = *reinterpret_cast<const float4*>(w + o), wi = *reinterpret_cast<const float4*>(w + o + kNB);
      const float wre[4] = {wr.x, wr.y, wr.z, wr.w}, wim[4] = {wi.x, wi.y, wi.z, wi.w};
      h4 a, b;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        a[e] = _Float16((wre[e] - acc[r][e].x) * inv[e]);
        b[e] = _Float16((wim[e] - acc[r][e].y) * inv[e]);
      }
      *reinterpret_cast<h4*>(vout + o) = a;
      *reinterpret_cast<h4*>(vout + o + kNB) = b;
      *reinterpret_cast<float4*>(xout32 + o) = make_float4(float(a[0]), float(a[1]), float(a[2]), float(a[3]));
      *reinterpret_cast<float4*>(xout32 + o + kNB) = make_float4(float(b[0]), float(b[1]), float(b[2]), float(b[3]));
    }
  }
}

// First stage of the row-block sums: workgroup (col, sg) adds the row blocks
// rb = sg, sg + kRS, ... into part2[(sg * ncx + col) * 64 + q].
__global__ __launch_bounds__(kT) void breduce1_kernel(const double2* __restrict__ partial, int nrb, int ncx,
                                                      double2* __restrict__ part2) {
  __shared__ double2 red[kT];
  const int col = blockIdx.x, sg = blockIdx.y, t = threadIdx.x, q = t & 63, ln = t >> 6;
  double2 a = make_double2(0.0, 0.0);
  for (int rb = sg + kRS * ln; rb < nrb; rb += kRS * (kT / 64)) {
    const double2 v = partial[((long long)rb * ncx + col) * kNB + q];
    a.x += v.x;
    a.y += v.y;
  }
  red[t] = a;
  __syncthreads();
  if (t < kNB) {
    double2 s = red[t];
    for (int k = 1; k < kT / 64; ++k) {
      s.x += red[k * kNB + t].x;
      s.y += red[k * kNB + t].y;
    }
    part2[((long long)sg * ncx + col) * kNB + t] = s;
  }
}

// Second stage: out[col * 64 + q] = sum over the kRS segments (fixed order).
__global__ void breduce2_kernel(const double2* __restrict__ part2, int ncx, double2* __restrict__ out) {
  const int col = blockIdx.x, q = threadIdx.x;
  double2 s = part2[(long long)col * kNB + q];
  for (int sg = 1; sg < kRS; ++sg) {
    const double2 v = part2[((long long)sg * ncx + col) * kNB + q];
    s.x += v.x;
    s.y += v.y;
  }
  out[col * kNB + q] = s;
}

// out = sums over the nrb row blocks of partial (ncx columns); part2 holds kRS x ncx x 64
void block_reduce(const double2* partial, int nrb, int ncx, double2* part2, double2* out, hipStream_t s) {
  hipLaunchKernelGGL(breduce1_kernel, dim3(ncx, kRS), dim3(kT), 0, s, partial, nrb, ncx, part2);
  hipLaunchKernelGGL(breduce2_kernel, dim3(ncx), dim3(kNB), 0, s, part2, ncx, out);
}

struct BGivens {
  double2* H;     // [(j*(m+1) + i) * 64 + q]
  double* cs;     // [j * 64 + q]
  double2* sn;    // [j * 64 + q]
  double2* g;     // [i * 64 + q], i <= m
  double* hn;     // [q]
  double* bnorm;  // [q]
  int* active;    // [q]
  int* kq;        // [q] Arnoldi steps applied in this cycle
  double* tolq;   // [q] relative tolerance of system q
};

// Cycle start per system q: beta = ||r_q|| (norm column of h), relative
// residual to host_rel[q]; systems already at tol (or padding) are inactive.
__global__ void bcycle_init_kernel(const double2* __restrict__ h, BGivens st, int m, int nrhs,
                                   double* host_rel, int* host_active) {
  const int q = threadIdx.x;
  if (q >= kNB) return;
  const double beta = sqrt(fmax(h[q].x, 0.0));
  const double rel = beta / st.bnorm[q];
  const int act = (q < nrhs && rel > st.tolq[q] && beta > 0.0) ? 1 : 0;
  st.active[q] = act;
  st.kq[q] = 0;
  st.hn[q] = act ? beta : 0.0;
  for (int i = 0; i <= m; ++i) st.g[i * kNB + q] = make_double2(i == 0 ? beta : 0.0, 0.0);
  host_rel[q] = rel;
  host_active[q] = act;
}

// V0[n, q] = w[n, q] / beta_q for active q, 0 otherwise (beta in st.hn)
// (v32: when given, the FP32 value of each stored entry as well)
template <class TV>
__global__ void bscale_kernel(const double* __restrict__ w, BGivens st, TV* __restrict__ v, long long n_total,
                              float* __restrict__ v32) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_total) return;
  const int q = int(i % kNB);
  const TV x = TV(st.active[q] ? w[i] / st.hn[q] : 0.0);
  v[i] = x;
  if (v32) v32[i] = float(x);
}

// Arnoldi step j for every active system: h in [0, j], ||w||^2 at j+1 (from the
// block dot); ||w - V h|| by Pythagoras, or from hn2 (||w - V h||^2 per
// system, when the caller orthogonalised twice); Givens rotation of H column
// j; the relative estimate |g_{j+1}| / ||b_q|| goes to host_est[j*64 + q]. A
// system whose estimate reaches 0.5 tol, or whose new vector vanishes, is
// frozen.
__global__ void btail_kernel(const double2* __restrict__ h, BGivens st, int m, int j, double* host_est,
                             int* host_active, const double* __restrict__ hn2) {
  const int q = threadIdx.x;
  if (q >= kNB) return;
  if (!st.active[q]) {
    st.hn[q] = 0.0;
    host_est[j * kNB + q] = -1.0;
    host_active[q] = 0;
    return;
  }
  double s2 = 0.0;
  double2* Hc = st.H + (long long)j * (m + 1) * kNB;
  for (int i = 0; i <= j; ++i) {
    const double2 v = h[i * kNB + q];
    s2 += v.x * v.x + v.y * v.y;
    Hc[i * kNB + q] = v;
  }
  const double wn2 = h[(j + 1) * kNB + q].x;
  const double hn = sqrt(fmax(hn2 ? hn2[q] : wn2 - s2, 0.0));
  for (int i = 0; i < j; ++i) {
    const double c = st.cs[i * kNB + q];
    const double2 s = st.sn[i * kNB + q];
    const double2 a = Hc[i * kNB + q], b = Hc[(i + 1) * kNB + q];
    Hc[i * kNB + q] = make_double2(c * a.x + (s.x * b.x - s.y * b.y), c * a.y + (s.x * b.y + s.y * b.x));
    Hc[(i + 1) * kNB + q] = make_double2(c * b.x - (s.x * a.x + s.y * a.y), c * b.y - (s.x * a.y - s.y * a.x));
  }
  const double2 a = Hc[j * kNB + q];
  const double an = hypot(a.x, a.y);
  double c;
  double2 s;
  if (an == 0.0) {
    c = 0.0;
    s = make_double2(1.0, 0.0);
    Hc[j * kNB + q] = make_double2(hn, 0.0);
  } else {
    const double r = hypot(an, hn);
    c = an / r;
    const double2 ph = make_double2(a.x / an, a.y / an);
    s = make_double2(ph.x * hn / r, ph.y * hn / r);
    Hc[j * kNB + q] = make_double2(ph.x * r, ph.y * r);
  }
  Hc[(j + 1) * kNB + q] = make_double2(0.0, 0.0);
  st.cs[j * kNB + q] = c;
  st.sn[j * kNB + q] = s;
  const double2 gj = st.g[j * kNB + q];
  st.g[j * kNB + q] = make_double2(c * gj.x, c * gj.y);
  const double2 g1 = make_double2(-(s.x * gj.x + s.y * gj.y), -(s.x * gj.y - s.y * gj.x));
  st.g[(j + 1) * kNB + q] = g1;
  const double est = hypot(g1.x, g1.y) / st.bnorm[q];
  st.kq[q] = j + 1;
  st.hn[q] = hn;
  const bool stop = est <= 0.5 * st.tolq[q] || !(hn > 1e-14 * sqrt(fmax(wn2, 0.0)));
  if (stop) st.active[q] = 0;
  host_est[j * kNB + q] = est;
  host_active[q] = stop ? 0 : 1;
}

// Second pass of a twice-orthogonalised step: h2 holds the projections of the
// once-projected vector w' and ||w'||^2 at nc; h1 += h2 gives the step's
// Hessenberg column, hn2[q] = ||w'||^2 - sum |h2|^2 its subdiagonal squared
// (no cancellation left: w' is already nearly orthogonal to V).
__global__ void bcgs2_combine_kernel(double2* __restrict__ h1, const double2* __restrict__ h2, int nc,
                                     double* __restrict__ hn2) {
  const int q = threadIdx.x;
  if (q >= kNB) return;
  double s2 = 0.0;
  for (int i = 0; i < nc; ++i) {
    const double2 a = h1[i * kNB + q], b = h2[i * kNB + q];
    s2 += b.x * b.x + b.y * b.y;
    h1[i * kNB + q] = make_double2(a.x + b.x, a.y + b.y);
  }
  hn2[q] = h2[nc * kNB + q].x - s2;
}

__global__ void bfill_kernel(double* __restrict__ a, int n, double v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// v[n, q] = (w[n, q] - sum_i V_i[n, q] h[i, q]) / hn_q, zero for systems with hn_q = 0
// (frozen before this step) — hn is the step's norm, set by btail. Thread
// (q, row group) owns kUpdRows consecutive rows of system q, so each
// coefficient h[i, q] is loaded once per kUpdRows rows.
template <class TV, class TW>
__global__ __launch_bounds__(kT) void bupdate_kernel(const TV* __restrict__ V, long long vs, int nc,
                                                     const double2* __restrict__ h, const TW* __restrict__ w,
                                                     const double* __restrict__ hn, TV* __restrict__ vout,
                                                     long long n_rows) {
  const int q = threadIdx.x & 63;
  const long long n0 = ((long long)blockIdx.x * (kT / 64) + (threadIdx.x >> 6)) * kUpdRows;
  if (n0 >= n_rows) return;
  const int nr = int(min((long long)kUpdRows, n_rows - n0));
  double ar[kUpdRows], ai[kUpdRows];
#pragma unroll
  for (int r = 0; r < kUpdRows; ++r) ar[r] = ai[r] = 0.0;
  const long long base = n0 * kLD + q;
#pragma unroll 2
  for (int i = 0; i < nc; ++i) {
    const double2 hh = h[i * kNB + q];
    const TV* Vi = V + i * vs + base;
#pragma unroll
    for (int r = 0; r < kUpdRows; ++r) {
      if (r < nr) {
        const double vr = double(gs_load(Vi + r * kLD)), vi = double(gs_load(Vi + r * kLD + kNB));
        ar[r] = fma(vr, hh.x, fma(-vi, hh.y, ar[r]));
        ai[r] = fma(vr, hh.y, fma(vi, hh.x, ai[r]));
      }
    }
  }
  const double d = hn[q];
  const double inv = d > 0.0 ? 1.0 / d : 0.0;
#pragma unroll
  for (int r = 0; r < kUpdRows; ++r) {
    if (r < nr) {
      const long long o = base + r * kLD;
      vout[o] = TV((double(w[o]) - ar[r]) * inv);
      vout[o + kNB] = TV((double(w[o + kNB]) - ai[r]) * inv);
    }
  }
}

// y_q = R_q^{-1} g_q over the kq[q] steps of system q (zero beyond)
__global__ void btrisolve_kernel(BGivens st, int m, double2* __restrict__ y) {
  const int q = threadIdx.x;
  if (q >= kNB) return;
  const int k = st.kq[q];
  for (int i = k; i <= m; ++i) y[i * kNB + q] = make_double2(0.0, 0.0);
  for (int i = k - 1; i >= 0; --i) {
    double2 s = st.g[i * kNB + q];
    for (int c = i + 1; c < k; ++c) {
      const double2 r = st.H[((long long)c * (m + 1) + i) * kNB + q];
      const double2 yc = y[c * kNB + q];
      s.x -= r.x * yc.x - r.y * yc.y;
      s.y -= r.x * yc.y + r.y * yc.x;
    }
    const double2 dg = st.H[((long long)i * (m + 1) + i) * kNB + q];
    const double den = dg.x * dg.x + dg.y * dg.y;
    y[i * kNB + q] = den > 0.0 ? make_double2((s.x * dg.x + s.y * dg.y) / den, (s.y * dg.x - s.x * dg.y) / den)
                               : make_double2(0.0, 0.0);
  }
}

// X[n, q] += sum_i V_i[n, q] y[i, q], i < nc
template <class TV>
__global__ __launch_bounds__(kT) void baxpy_kernel(const TV* __restrict__ V, long long vs, int nc,
                                                   const double2* __restrict__ y, double* __restrict__ X,
                                                   long long n_rows) {
  const long long e = (long long)blockIdx.x * kT + threadIdx.x;
  const long long n = e / kNB;
  const int q = int(e % kNB);
  if (n >= n_rows) return;
  const long long base = n * kLD + q;
  double ar = 0.0, ai = 0.0;
  for (int i = 0; i < nc; ++i) {
    const double vr = double(V[i * vs + base]), vi = double(V[i * vs + base + kNB]);
    const double2 yy = y[i * kNB + q];
    ar = fma(vr, yy.x, fma(-vi, yy.y, ar));
    ai = fma(vr, yy.y, fma(vi, yy.x, ai));
  }
  X[base] += ar;
  X[base + kNB] += ai;
}

// X += V y for the FP16 basis: lane l covers the systems 4 (l % 16) .. + 3 of
// one row (8-byte basis loads, as bupdate_h_kernel), the quarter-wave l / 16
// one of four consecutive rows; sums in FP64 (X is the FP64 solution).
__global__ __launch_bounds__(kT) void baxpy_h_kernel(const _Float16* __restrict__ V, long long vs, int nc,
                                                     const double2* __restrict__ y, double* __restrict__ X,
                                                     long long n_rows) {
  typedef _Float16 h4 __attribute__((ext_vector_type(4)));
  const int lane = threadIdx.x & 63, sg = lane & 15, rs = lane >> 4;
  const long long n = ((long long)blockIdx.x * (kT / 64) + (threadIdx.x >> 6)) * 4 + rs;
  if (n >= n_rows) return;
  double ar[4] = {0.0, 0.0, 0.0, 0.0}, ai[4] = {0.0, 0.0, 0.0, 0.0};
  const long long base = n * kLD + 4 * sg;
  for (int i = 0; i < nc; ++i) {
    const h4 a = *reinterpret_cast<const h4*>(V + i * vs + base);
    const h4 b = *reinterpret_cast<const h4*>(V + i * vs + base + kNB);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const double2 yy = y[i * kNB + 4 * sg + e];
      const double vr = double(a[e]), vi = double(b[e]);
      ar[e] = fma(vr, yy.x, fma(-vi, yy.y, ar[e]));
      ai[e] = fma(vr, yy.y, fma(vi, yy.x, ai[e]));
    }
  }
  double4* xr = reinterpret_cast<double4*>(X + base);
  double4* xi = reinterpret_cast<double4*>(X + base + kNB);
  double4 u = *xr, v = *xi;
  u.x += ar[0]; u.y += ar[1]; u.z += ar[2]; u.w += ar[3];
  v.x += ai[0]; v.y += ai[1]; v.z += ai[2]; v.w += ai[3];
  *xr = u;
  *xi = v;
}

// zero the columns q of a block whose mask[q] is 0
__global__ void bmask_kernel(double* __restrict__ V, const int* __restrict__ mask, long long n_total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_total && !mask[i % kNB]) V[i] = 0.0;
}

// X += D
__global__ void badd_kernel(double* __restrict__ X, const double* __restrict__ D, long long n_total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_total) X[i] += D[i];
}

// R = B - R
__global__ void bsub_kernel(const double* __restrict__ B, double* __restrict__ R, long long n_total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_total) R[i] = B[i] - R[i];
}

__global__ void to_split_kernel(const double2* __restrict__ x, long long ld, int nrhs, int N, double* __restrict__ out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = e / kNB;
  const int q = int(e % kNB);
  if (n >= N) return;
  double2 v = make_double2(0.0, 0.0);
  if (q < nrhs) v = x[q * ld + n];
  out[n * kLD + q] = v.x;
  out[n * kLD + kNB + q] = v.y;
}

__global__ void from_split_kernel(const double* __restrict__ in, int N, int nrhs, double2* __restrict__ x, long long ld) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = e / kNB;
  const int q = int(e % kNB);
  if (n >= N || q >= nrhs) return;
  x[q * ld + n] = make_double2(in[n * kLD + q], in[n * kLD + kNB + q]);
}

inline unsigned blocks_for(long long n, int per) { return unsigned((n + per - 1) / per); }

}  // namespace

void to_split(const double2* x, long long ld, int nrhs, int N, int nb, double* out, hipStream_t s) {
  (void)nb;
  hipLaunchKernelGGL(to_split_kernel, dim3(blocks_for((long long)N * kNB, kT)), dim3(kT), 0, s, x, ld, nrhs, N, out);
}

void from_split(const double* in, int N, int nb, int nrhs, double2* x, long long ld, hipStream_t s) {
  (void)nb;
  hipLaunchKernelGGL(from_split_kernel, dim3(blocks_for((long long)N * kNB, kT)), dim3(kT), 0, s, in, N, nrhs, x, ld);
}

std::string BlockOperator::setup(DeviceOperator& o) {
  if (op == &o && bstate) return "";
  if (o.n_table_sets > 1) return "block matvec: the operator holds several near-field table sets (a fused plan)";
  release();
  op = &o;
  nb = kNB;
  // every destination must fit the kernel's row tiles
  {
    std::vector<int> rows_h(size_t(o.passes.empty() ? 0 : o.passes.back().dest_base + o.passes.back().ndest));
    if (!rows_h.empty())
      HIP_OK(hipMemcpy(rows_h.data(), o.d_rows, rows_h.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int r : rows_h)
      if (r > kBMaxRows) return "block matvec: a destination has more than 64 rows";
    if (o.rank > kBMaxRows) return "block matvec: rank above 64";
    pass_max_rows.assign(o.passes.size(), 1);
    for (size_t pi = 0; pi < o.passes.size(); ++pi)
      for (int d = 0; d < o.passes[pi].ndest; ++d)
        pass_max_rows[pi] = std::max(pass_max_rows[pi], rows_h[size_t(o.passes[pi].dest_base + d)]);
    // passes with few destinations and long operators are split over columns
    std::vector<int> ncol_h(rows_h.size());
    if (!ncol_h.empty())
      HIP_OK(hipMemcpy(ncol_h.data(), o.d_ncol, ncol_h.size() * sizeof(int), hipMemcpyDeviceToHost));
    pass_nseg.assign(o.passes.size(), 1);
    size_t part_elems = 0;
    for (size_t pi = 0; pi < o.passes.size(); ++pi) {
      const DevPass& p = o.passes[pi];
      int64_t cols = 0;
      int cmax = 0;
      for (int d = 0; d < p.ndest; ++d) {
        cols += ncol_h[size_t(p.dest_base + d)];
        cmax = std::max(cmax, ncol_h[size_t(p.dest_base + d)]);
      }
      if (p.ndest > 0 && p.ndest < kSplitMaxDest && cols > int64_t(kSplitCols) * p.ndest) {
        pass_nseg[pi] = (cmax + kSplitCols - 1) / kSplitCols;
        const int rt = 16 * std::max(1, (2 * pass_max_rows[pi] + 15) / 16);
        part_elems = std::max(part_elems, size_t(p.ndest) * pass_nseg[pi] * rt * kNB);
      }
    }
    if (part_elems) HIP_OK(hipMalloc((void**)&split_part, part_elems * sizeof(double)));
  }
  HIP_OK(hipMalloc((void**)&bstate, sizeof(double) * size_t(o.n_state + 1) * kLD));
  HIP_OK(hipMemset(bstate, 0, sizeof(double) * size_t(o.n_state + 1) * kLD));
  if (!o.T) return "";
  // Uniform tree: the leaf kernel writes every y row (near field, qk2, x). The
  // leaf M2L and leaf L2L passes are folded into it when each of them is one
  // accumulating pass whose destinations are whole leaves (rank rows at
  // N + leaf rank, at most one per leaf) and no other pass writes y rows;
  // otherwise they run as passes after it.
  const int L = o.n_leaves, rk = o.rank;
  if (int64_t(L) * rk != o.N) return "block matvec: a uniform tree must have N = leaves x rank";
  std::vector<int> ld(size_t(L) * 2, -1);
  bool fuse = o.near.empty() && o.leaf_m2l.size() <= 1 && o.leaf_l2l.size() <= 1;
  {
    const int64_t nd = o.passes.empty() ? 0 : o.passes.back().dest_base + o.passes.back().ndest;
    std::vector<int> rows_h(static_cast<size_t>(nd)), dst_h(static_cast<size_t>(nd));
    if (nd) {
      HIP_OK(hipMemcpy(rows_h.data(), o.d_rows, rows_h.size() * sizeof(int), hipMemcpyDeviceToHost));
      HIP_OK(hipMemcpy(dst_h.data(), o.d_dst, dst_h.size() * sizeof(int), hipMemcpyDeviceToHost));
    }
    auto writes_y = [&](const DevPass& p) {
      for (int d = 0; d < p.ndest && fuse; ++d) {
        const int64_t a = dst_h[size_t(p.dest_base + d)], b = a + rows_h[size_t(p.dest_base + d)];
        if (a < 2 * int64_t(o.N) && b > o.N) return true;
      }
      return false;
    };
    for (int pi : o.up) fuse = fuse && !writes_y(o.passes[size_t(pi)]);
    for (int pi : o.down) fuse = fuse && !writes_y(o.passes[size_t(pi)]);
    const std::vector<int>* lists[2] = {&o.leaf_m2l, &o.leaf_l2l};
    for (int k = 0; k < 2 && fuse; ++k) {
      if (lists[k]->empty()) continue;
      const DevPass& p = o.passes[size_t(lists[k]->front())];
      fuse = p.kind == 1;
      for (int d = 0; d < p.ndest && fuse; ++d) {
        const int64_t g = p.dest_base + d;
        const int64_t off = int64_t(dst_h[size_t(g)]) - o.N;
        const int64_t leaf = off / rk;
        fuse = rows_h[size_t(g)] == rk && off >= 0 && off % rk == 0 && leaf < L && ld[size_t(2 * leaf + k)] < 0;
        if (fuse) ld[size_t(2 * leaf + k)] = int(g);
      }
    }
  }
  leaf_fused = fuse;
  if (!fuse) std::fill(ld.begin(), ld.end(), -1);
  HIP_OK(hipMalloc((void**)&leaf_dest, ld.size() * sizeof(int)));
  HIP_OK(hipMemcpy(leaf_dest, ld.data(), ld.size() * sizeof(int), hipMemcpyHostToDevice));
  // near-field source rows: leaf l reads colleague s's x rows (nbl rank + c),
  // the state's zero row where the colleague is absent
  std::vector<int> nbr_h(size_t(L) * 9);
  HIP_OK(hipMemcpy(nbr_h.data(), o.nbr, nbr_h.size() * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> ns(size_t(L) * 9 * rk);
  for (int l = 0; l < L; ++l)
    for (int sl = 0; sl < 9; ++sl) {
      const int nbl = nbr_h[size_t(l) * 9 + sl];
      for (int c = 0; c < rk; ++c) ns[(size_t(l) * 9 + sl) * rk + c] = nbl >= 0 ? nbl * rk + c : int(o.n_state);
    }
  HIP_OK(hipMalloc((void**)&near_src, ns.size() * sizeof(int)));
  HIP_OK(hipMemcpy(near_src, ns.data(), ns.size() * sizeof(int), hipMemcpyHostToDevice));
  return "";
}

void BlockOperator::release() {
  notes.clear();
  work.release();
  outer.release();
  void* ptrs[] = {bstate, leaf_dest, near_src, qk2_cols, bstate32, split_part, ops_sb, T_sb, d_opsb};
  ops_sb = nullptr;
  T_sb = nullptr;
  d_opsb = nullptr;
  sb_active = false;
  qk2_cols = nullptr;
  bstate32 = nullptr;
  split_part = nullptr;
  leaf_dest = nullptr;
  near_src = nullptr;
  leaf_fused = false;
  for (void* p : ptrs)
    if (p) hipFree(p);
  bstate = nullptr;
  op = nullptr;
}

template <bool ACC, class R, int MT, int NT>
static void launch_bpass_nt(const BlockOperator& bo, int ndest, const typename Mfma<R>::c2* ops, const int* src,
                            R* bst, const R* X, int nx, double* Y, const int* rows, const int* ncol, const int* dst,
                            const long long* op_off, const long long* src_off, int nseg) {
  hipStream_t s = bo.op->stream;
  const int zrow = int(bo.op->n_state);  // the pass state's all-zero row
  if (nseg <= 1) {
    hipLaunchKernelGGL((bpass_kernel<ACC, R, MT, NT>), dim3(ndest), dim3(block_threads<NT>()), 0, s, ops, src, bst, X,
                       Y, bo.op->N, rows, ncol, dst, op_off, src_off, nx, zrow, 0, 1, (R*)nullptr);
    return;
  }
  R* part = reinterpret_cast<R*>(bo.split_part);
  hipLaunchKernelGGL((bpass_kernel<ACC, R, MT, NT>), dim3(ndest * nseg), dim3(block_threads<NT>()), 0, s, ops, src, bst,
                     X, Y, bo.op->N, rows, ncol, dst, op_off, src_off, nx, zrow, kSplitCols, nseg, part);
  hipLaunchKernelGGL((bsplit_reduce_kernel<ACC, R>), dim3(ndest), dim3(kT), 0, s, part, nseg, MT * 16, rows, dst, bst,
                     Y, bo.op->N);
}

// right-hand-side tiles per wavefront: the operator's choice for FP32 (bo.nt32), 1 for FP64
template <bool ACC, class R, int MT>
static void launch_bpass_mt(const BlockOperator& bo, int ndest, const typename Mfma<R>::c2* ops, const int* src,
                            R* bst, const R* X, int nx, double* Y, const int* rows, const int* ncol, const int* dst,
                            const long long* op_off, const long long* src_off, int nseg) {
  if (std::is_same<R, float>::value && bo.nt32 == 2)
    launch_bpass_nt<ACC, R, MT, 2>(bo, ndest, ops, src, bst, X, nx, Y, rows, ncol, dst, op_off, src_off, nseg);
  else
    launch_bpass_nt<ACC, R, MT, 1>(bo, ndest, ops, src, bst, X, nx, Y, rows, ncol, dst, op_off, src_off, nseg);
}

// row tiles from the largest destination of the pass
template <bool ACC, class R>
static void launch_bpass_any(const BlockOperator& bo, int max_rows, int ndest, const typename Mfma<R>::c2* ops,
                             const int* src, R* bst, const R* X, int nx, double* Y, const int* rows, const int* ncol,
                             const int* dst, const long long* op_off, const long long* src_off, int nseg = 1) {
  const int mt = std::max(1, (2 * max_rows + 15) / 16);
#define LSG_BPASS_CASE(M)                                                                                       \
  case M:                                                                                                       \
    launch_bpass_mt<ACC, R, M>(bo, ndest, ops, src, bst, X, nx, Y, rows, ncol, dst, op_off, src_off, nseg);    \
    break;
  switch (mt) {
    LSG_BPASS_CASE(1)
    LSG_BPASS_CASE(2)
    LSG_BPASS_CASE(3)
    LSG_BPASS_CASE(4)
    LSG_BPASS_CASE(5)
    LSG_BPASS_CASE(6)
    LSG_BPASS_CASE(7)
    default:
      launch_bpass_mt<ACC, R, 8>(bo, ndest, ops, src, bst, X, nx, Y, rows, ncol, dst, op_off, src_off, nseg);
  }
#undef LSG_BPASS_CASE
}

template <class R>
static void launch_bpass(BlockOperator& bo, int pidx, const R* X, int nx, double* Y) {
  DeviceOperator& o = *bo.op;
  const DevPass& p = o.passes[size_t(pidx)];
  const int64_t b = p.dest_base;
  typedef typename Mfma<R>::c2 C2;
  const C2* ops = reinterpret_cast<const C2*>(std::is_same<R, double>::value ? (const void*)o.ops : (const void*)o.ops32);
  R* bst = reinterpret_cast<R*>(std::is_same<R, double>::value ? (void*)bo.bstate : (void*)bo.bstate32);
  const int mr = bo.pass_max_rows[size_t(pidx)];
  const int nseg = bo.pass_nseg[size_t(pidx)];
  if (p.kind == 0)
    launch_bpass_any<false, R>(bo, mr, p.ndest, ops, o.src, bst, X, nx, Y, o.d_rows + b, o.d_ncol + b, o.d_dst + b,
                               o.d_op + b, o.d_srcoff + b, nseg);
  else
    launch_bpass_any<true, R>(bo, mr, p.ndest, ops, o.src, bst, X, nx, Y, o.d_rows + b, o.d_ncol + b, o.d_dst + b,
                              o.d_op + b, o.d_srcoff + b, nseg);
}

// split-BF16 pass: MTP row tiles of 16 complex rows from the largest destination
template <bool ACC, int NT, class XS>
static void launch_bpass_sb_nt(BlockOperator& bo, const DevPass& p, int max_rows, int nseg, const XS* X, int nx,
                               double* Y) {
  DeviceOperator& o = *bo.op;
  hipStream_t s = o.stream;
  const int64_t b = p.dest_base;
  const int mtp = (max_rows + 15) / 16;
  const int rt = 16 * std::max(1, (2 * max_rows + 15) / 16);  // partial rows (the FP32 layout)
  const int kseg = nseg > 1 ? kSplitCols : 0;
  const unsigned grid = unsigned(p.ndest) * unsigned(nseg > 1 ? nseg : 1);
  float* part = reinterpret_cast<float*>(bo.split_part);
#define LSG_SB_CASE(M)                                                                                         \
  case M:                                                                                                      \
    hipLaunchKernelGGL((bpass_sb_kernel<ACC, M, NT, XS>), dim3(grid), dim3(block_threads<NT>()), 0, s, bo.ops_sb, o.src, \
                       bo.bstate32, X, Y, o.N, o.d_rows + b, o.d_ncol + b, o.d_dst + b, bo.d_opsb + b,            \
                       o.d_srcoff + b, nx, kseg, nseg > 1 ? nseg : 1, rt, part);                                \
    break;
  switch (mtp) {
    LSG_SB_CASE(1)
    LSG_SB_CASE(2)
    LSG_SB_CASE(3)
    default:
      LSG_SB_CASE(4)
  }
#undef LSG_SB_CASE
  if (nseg > 1)
    hipLaunchKernelGGL((bsplit_reduce_kernel<ACC, float>), dim3(p.ndest), dim3(kT), 0, s, part, nseg, rt,
                       o.d_rows + b, o.d_dst + b, bo.bstate32, Y, o.N);
}

template <class XS>
static void launch_bpass_sb(BlockOperator& bo, int pidx, const XS* X, int nx, double* Y) {
  const DevPass& p = bo.op->passes[size_t(pidx)];
  const int mr = bo.pass_max_rows[size_t(pidx)];
  const int nseg = bo.pass_nseg[size_t(pidx)];
  if (p.kind == 0) {
    if (bo.nt_sb == 2) launch_bpass_sb_nt<false, 2>(bo, p, mr, nseg, X, nx, Y);
    else launch_bpass_sb_nt<false, 1>(bo, p, mr, nseg, X, nx, Y);
  } else {
    if (bo.nt_sb == 2) launch_bpass_sb_nt<true, 2>(bo, p, mr, nseg, X, nx, Y);
    else launch_bpass_sb_nt<true, 1>(bo, p, mr, nseg, X, nx, Y);
  }
}

// the split-BF16 leaf kernel
template <class XS, class XT, class YT, int NT>
static void launch_bleaf_sb_nt(BlockOperator& bo, const XS* Xr, int nx, const XT* X, YT* Y, int part,
                               hipStream_t s) {
  DeviceOperator& o = *bo.op;
  const int mtp = (o.rank + 15) / 16;
  const dim3 grid(o.n_leaves);
#define LSG_SBL_CASE(M)                                                                                        \
  case M:                                                                                                      \
    hipLaunchKernelGGL((bleaf_sb_kernel<XS, XT, YT, M, NT>), grid, dim3(block_threads<NT>()), 0, s, bo.T_sb,     \
                       bo.near_src, o.rank, bo.ops_sb, o.src, o.d_ncol, bo.d_opsb, o.d_srcoff, bo.leaf_dest, Xr,     \
                       bo.bstate32, nx, X, o.qk2, o.find_phi ? 1 : 0, Y, part);                                \
    break;
  switch (mtp) {
    LSG_SBL_CASE(1)
    LSG_SBL_CASE(2)
    LSG_SBL_CASE(3)
    default:
      LSG_SBL_CASE(4)
  }
#undef LSG_SBL_CASE
}

// FP32 copy of the input block (FP64 or FP16) into the first N rows of the FP32 pass state
template <class XT>
__global__ void bround_kernel(const XT* __restrict__ X, float* __restrict__ bst32, long long n_total) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_total) bst32[i] = float(X[i]);
}

// The block matvec in split BF16 (the inner solves' arithmetic): the same
// passes and leaf kernel as the FP32 path; the input block's rows are read in
// place (XS = XT, FP32 or FP16) or from its FP32 copy in the pass state (nx = 0).
template <class XS, class XT, class YT>
static void block_matvec_sb(BlockOperator& bo, const XS* Xs, int nx, const XT* X, YT* Y) {
  constexpr bool y64 = std::is_same<YT, double>::value;
  double* Yd = y64 ? reinterpret_cast<double*>(Y) : nullptr;
  DeviceOperator& o = *bo.op;
  if (!y64 && !(o.T && bo.can_store_f32())) return;  // not reached: callers check can_store_f32
  auto leaf = [&](int part, hipStream_t s) {
    if (bo.nt_sb == 2) launch_bleaf_sb_nt<XS, XT, YT, 2>(bo, Xs, nx, X, Y, part, s);
    else launch_bleaf_sb_nt<XS, XT, YT, 1>(bo, Xs, nx, X, Y, part, s);
  };
  if (o.T && bo.leaf_fused && bo.sb_split) {
    // the near field on the second stream beside the upward and downward
    // passes, the leaf passes added after both
    (void)hipEventRecord(o.ev_fork, o.stream);
    (void)hipStreamWaitEvent(o.aux, o.ev_fork, 0);
    leaf(1, o.aux);
    for (int p : o.up) launch_bpass_sb(bo, p, Xs, nx, Yd);
    for (int p : o.down) launch_bpass_sb(bo, p, Xs, nx, Yd);
    (void)hipEventRecord(o.ev_join, o.aux);
    (void)hipStreamWaitEvent(o.stream, o.ev_join, 0);
    leaf(2, o.stream);
  } else {
    for (int p : o.up) launch_bpass_sb(bo, p, Xs, nx, Yd);
    if (o.T && bo.leaf_fused) {
      for (int p : o.down) launch_bpass_sb(bo, p, Xs, nx, Yd);
      leaf(0, o.stream);
    } else if constexpr (y64) {
      if (o.T) leaf(0, o.stream);
      for (int p : o.near) launch_bpass_sb(bo, p, Xs, nx, Yd);
      for (int p : o.down) launch_bpass_sb(bo, p, Xs, nx, Yd);
      for (int p : o.leaf_m2l) launch_bpass_sb(bo, p, Xs, nx, Yd);
      for (int p : o.leaf_l2l) launch_bpass_sb(bo, p, Xs, nx, Yd);
    }
  }
  // a contrast per right-hand side (volume-potential context): Y = X + qk2 .* Y
  if constexpr (y64) {
    if (bo.qk2_cols)
      hipLaunchKernelGGL(bdiag_kernel<XT>, dim3(blocks_for((long long)o.N * kLD, kT)), dim3(kT), 0, o.stream, X, Yd,
                         bo.qk2_cols, (long long)o.N * kLD);
  }
}

// R: pass precision; XT: type of the input block. An FP64 or FP16 input to
// the FP32 passes is converted into the pass state's x rows first; an input
// already in the pass precision is read in place.
template <class R, class XT, class YT = double>
static void block_matvec_impl(BlockOperator& bo, const XT* X, YT* Y) {
  // an FP32 product is stored by the leaf kernel alone (can_store_f32): the
  // other passes then write only the pass state, and get no output pointer
  constexpr bool y64 = std::is_same<YT, double>::value;
  double* Yd = y64 ? reinterpret_cast<double*>(Y) : nullptr;
  DeviceOperator& o = *bo.op;
  typedef typename Mfma<R>::c2 C2;
  const C2* T = reinterpret_cast<const C2*>(std::is_same<R, double>::value ? (const void*)o.T : (const void*)o.T32);
  R* bst = reinterpret_cast<R*>(std::is_same<R, double>::value ? (void*)bo.bstate : (void*)bo.bstate32);
  if constexpr (std::is_same<R, float>::value) {
    if (bo.sb_active) {
      if constexpr (std::is_same<XT, double>::value) {
        hipLaunchKernelGGL(bround_kernel<XT>, dim3(blocks_for((long long)o.N * kLD, kT)), dim3(kT), 0, o.stream, X,
                           bo.bstate32, (long long)o.N * kLD);
        block_matvec_sb<float, XT, YT>(bo, bo.bstate32, 0, X, Y);
      } else {
        block_matvec_sb<XT, XT, YT>(bo, X, o.N, X, Y);
      }
      return;
    }
  }
  const R* Xr;
  int nx;
  if (std::is_same<R, XT>::value) {
    Xr = reinterpret_cast<const R*>(X);
    nx = o.N;
  } else {
    hipLaunchKernelGGL(bround_kernel<XT>, dim3(blocks_for((long long)o.N * kLD, kT)), dim3(kT), 0, o.stream, X,
                       bo.bstate32, (long long)o.N * kLD);
    Xr = bst;
    nx = 0;
  }
  // the leaf kernel: near field, qk2 and x, and the leaf passes when folded in
  auto leaf = [&]() {
    const C2* ops = reinterpret_cast<const C2*>(std::is_same<R, double>::value ? (const void*)o.ops
                                                                                : (const void*)o.ops32);
    const int mt = std::max(1, (2 * o.rank + 15) / 16);
    const int nt = std::is_same<R, float>::value ? bo.nt32 : 1;
    const dim3 grid(o.n_leaves);
#define LSG_LEAF_CASE(M)                                                                                        \
  case M:                                                                                                       \
    if (nt == 2)                                                                                                \
      hipLaunchKernelGGL((bleaf_kernel<R, XT, YT, M, 2>), grid, dim3(block_threads<2>()), 0, o.stream, T,        \
                         bo.near_src, o.rank, ops, o.src, o.d_ncol, o.d_op, o.d_srcoff, bo.leaf_dest, Xr, bst, nx, X, \
                         o.qk2, o.find_phi ? 1 : 0, Y, int(o.n_state));                                        \
    else                                                                                                        \
      hipLaunchKernelGGL((bleaf_kernel<R, XT, YT, M, 1>), grid, dim3(block_threads<1>()), 0, o.stream, T,        \
                         bo.near_src, o.rank, ops, o.src, o.d_ncol, o.d_op, o.d_srcoff, bo.leaf_dest, Xr, bst, nx, X, \
                         o.qk2, o.find_phi ? 1 : 0, Y, int(o.n_state));                                        \
    break;
    switch (mt) {
      LSG_LEAF_CASE(1)
      LSG_LEAF_CASE(2)
      LSG_LEAF_CASE(3)
      LSG_LEAF_CASE(4)
      LSG_LEAF_CASE(5)
      LSG_LEAF_CASE(6)
      LSG_LEAF_CASE(7)
      default:
        hipLaunchKernelGGL((bleaf_kernel<R, XT, YT, 8, 1>), grid, dim3(block_threads<1>()), 0, o.stream, T,
                           bo.near_src, o.rank, ops, o.src, o.d_ncol, o.d_op, o.d_srcoff, bo.leaf_dest, Xr, bst, nx, X,
                           o.qk2, o.find_phi ? 1 : 0, Y, int(o.n_state));
    }
#undef LSG_LEAF_CASE
  };
  if (!y64 && !(o.T && bo.can_store_f32())) return;  // not reached: callers check can_store_f32
  for (int p : o.up) launch_bpass<R>(bo, p, Xr, nx, Yd);
  if (o.T && bo.leaf_fused) {
    for (int p : o.down) launch_bpass<R>(bo, p, Xr, nx, Yd);
    leaf();
  } else if constexpr (y64) {
    if (o.T) leaf();
    for (int p : o.near) launch_bpass<R>(bo, p, Xr, nx, Yd);
    for (int p : o.down) launch_bpass<R>(bo, p, Xr, nx, Yd);
    for (int p : o.leaf_m2l) launch_bpass<R>(bo, p, Xr, nx, Yd);
    for (int p : o.leaf_l2l) launch_bpass<R>(bo, p, Xr, nx, Yd);
  }
  if constexpr (y64) {
    if (bo.qk2_cols)
      hipLaunchKernelGGL(bdiag_kernel<XT>, dim3(blocks_for((long long)o.N * kLD, kT)), dim3(kT), 0, o.stream, X, Yd,
                         bo.qk2_cols, (long long)o.N * kLD);
  }
}

void BlockOperator::matvec(const double* X, double* Y) { block_matvec_impl<double, double>(*this, X, Y); }
void BlockOperator::matvec32(const double* X, double* Y) { block_matvec_impl<float, double>(*this, X, Y); }
void BlockOperator::matvec32(const float* X, double* Y) { block_matvec_impl<float, float>(*this, X, Y); }
void BlockOperator::matvec32(const float* X, float* Y) { block_matvec_impl<float, float, float>(*this, X, Y); }
void BlockOperator::matvec32(const _Float16* X, double* Y) { block_matvec_impl<float, _Float16>(*this, X, Y); }
void BlockOperator::matvec32(const _Float16* X, float* Y) { block_matvec_impl<float, _Float16, float>(*this, X, Y); }

std::string BlockOperator::enable_fp32() {
  std::string e = op->enable_fp32();
  if (!e.empty()) return e;
  if (bstate32) return "";
  HIP_OK(hipMalloc((void**)&bstate32, sizeof(float) * size_t(op->n_state + 1) * kLD));
  HIP_OK(hipMemset(bstate32, 0, sizeof(float) * size_t(op->n_state + 1) * kLD));
  // one or two right-hand-side tiles per wavefront in the FP32 passes: the
  // faster on this operator (a zero block in, best of three after a warm-up)
  const size_t nb = size_t(op->N) * kLD;
  float* x = nullptr;
  double* y = nullptr;
  HIP_OK(hipMalloc((void**)&x, sizeof(float) * nb));
  if (hipMalloc((void**)&y, sizeof(double) * nb) != hipSuccess) {
    (void)hipFree(x);
    return "FP32 block operator: allocation failed";
  }
  hipStream_t s = op->stream;
  (void)hipMemsetAsync(x, 0, sizeof(float) * nb, s);
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  float best[2] = {1e30f, 1e30f};
  for (int nt = 1; nt <= 2; ++nt) {
    nt32 = nt;
    matvec32(x, y);
    for (int r = 0; r < 3; ++r) {
      (void)hipEventRecord(a, s);
      matvec32(x, y);
      (void)hipEventRecord(b, s);
      (void)hipEventSynchronize(b);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, a, b);
      best[nt - 1] = std::min(best[nt - 1], ms);
    }
  }
  nt32 = best[1] < best[0] ? 2 : 1;
  notes.push_back("FP32 block matvec: tiles per wavefront " + std::to_string(nt32) + " (1: " +
                  std::to_string(best[0]) + " ms, 2: " + std::to_string(best[1]) + " ms)");
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(x);
  (void)hipFree(y);
  return "";
}

std::string BlockOperator::enable_sb() {
  std::string e = enable_fp32();
  if (!e.empty()) return e;
  if (ops_sb) return "";
  DeviceOperator& o = *op;
  const int64_t nd = o.passes.empty() ? 0 : o.passes.back().dest_base + o.passes.back().ndest;
  std::vector<int> rows_h(static_cast<size_t>(nd)), ncol_h(static_cast<size_t>(nd));
  if (nd) {
    HIP_OK(hipMemcpy(rows_h.data(), o.d_rows, rows_h.size() * sizeof(int), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(ncol_h.data(), o.d_ncol, ncol_h.size() * sizeof(int), hipMemcpyDeviceToHost));
  }
  std::vector<long long> off(static_cast<size_t>(nd));
  long long total = 0;
  for (int64_t d = 0; d < nd; ++d) {
    off[size_t(d)] = total;
    total += (long long)rows_h[size_t(d)] * ((ncol_h[size_t(d)] + 15) / 16);
  }
  HIP_OK(hipMalloc((void**)&ops_sb, size_t(std::max(total, 1LL)) * 128));
  HIP_OK(hipMalloc((void**)&d_opsb, sizeof(long long) * size_t(std::max<int64_t>(nd, 1))));
  if (nd) {
    HIP_OK(hipMemcpy(d_opsb, off.data(), off.size() * sizeof(long long), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(to_sb_kernel, dim3(unsigned(nd)), dim3(kT), 0, o.stream, o.ops, o.d_op, o.d_rows, o.d_ncol,
                       d_opsb, reinterpret_cast<unsigned*>(ops_sb));
  }
  if (o.T) {
    const long long trows = (long long)o.rank * ((9 * o.rank + 15) / 16);
    HIP_OK(hipMalloc((void**)&T_sb, size_t(trows) * 128));
    hipLaunchKernelGGL(to_sb_table_kernel, dim3(blocks_for(trows * 16, kT)), dim3(kT), 0, o.stream, o.T, o.rank,
                       9 * o.rank, reinterpret_cast<unsigned*>(T_sb));
  }
  HIP_OK(hipStreamSynchronize(o.stream));
  // one or two right-hand-side tiles per wavefront: the faster on this operator
  const size_t nb = size_t(o.N) * kLD;
  float* x = nullptr;
  double* y = nullptr;
  HIP_OK(hipMalloc((void**)&x, sizeof(float) * nb));
  if (hipMalloc((void**)&y, sizeof(double) * nb) != hipSuccess) {
    (void)hipFree(x);
    return "split-BF16 block operator: allocation failed";
  }
  hipStream_t s = o.stream;
  (void)hipMemsetAsync(x, 0, sizeof(float) * nb, s);
  hipEvent_t a, b;
  (void)hipEventCreate(&a);
  (void)hipEventCreate(&b);
  // and the near field beside the passes (sb_split) or after them
  float best[4] = {1e30f, 1e30f, 1e30f, 1e30f};
  SbScope sb_scope(*this, true);
  for (int v = 0; v < 4; ++v) {
    nt_sb = 1 + (v & 1);
    sb_split = v >= 2;
    if (sb_split && !(o.T && leaf_fused)) continue;
    matvec32(x, y);
    for (int r = 0; r < 3; ++r) {
      (void)hipEventRecord(a, s);
      matvec32(x, y);
      (void)hipEventRecord(b, s);
      (void)hipEventSynchronize(b);
      float ms = 0;
      (void)hipEventElapsedTime(&ms, a, b);
      best[v] = std::min(best[v], ms);
    }
  }
  const int vb = int(std::min_element(best, best + 4) - best);
  nt_sb = 1 + (vb & 1);
  sb_split = vb >= 2;
  {
    std::string n = "split-BF16 block matvec: tiles per wavefront " + std::to_string(nt_sb) +
                    (sb_split ? ", near field beside the passes" : ", near field after the passes") + " (";
    const char* lab[4] = {"1/after", "2/after", "1/beside", "2/beside"};
    for (int v = 0; v < 4; ++v)
      if (best[v] < 1e29f) n += std::string(v ? ", " : "") + lab[v] + ": " + std::to_string(best[v]) + " ms";
    notes.push_back(n + ")");
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  (void)hipFree(x);
  (void)hipFree(y);
  HIP_OK(hipGetLastError());
  return "";
}

std::string BlockWork::reserve(int N_, int m_) {
  const long long vs = (long long)N_ * kLD;
  const int nrb = (N_ + kBRows - 1) / kBRows;
  HIP_OK(hipMalloc((void**)&w, sizeof(double) * size_t(vs)));
  HIP_OK(hipMalloc((void**)&partial, sizeof(double2) * size_t(nrb + kRS) * (m_ + 2) * kNB));
  HIP_OK(hipMalloc((void**)&h, sizeof(double2) * size_t(m_ + 2) * kNB));
  HIP_OK(hipMalloc((void**)&y, sizeof(double2) * size_t(m_ + 1) * kNB));
  HIP_OK(hipMalloc((void**)&H, sizeof(double2) * size_t(m_ + 1) * m_ * kNB));
  HIP_OK(hipMalloc((void**)&cs, sizeof(double) * size_t(m_) * kNB));
  HIP_OK(hipMalloc((void**)&sn, sizeof(double2) * size_t(m_) * kNB));
  HIP_OK(hipMalloc((void**)&g, sizeof(double2) * size_t(m_ + 1) * kNB));
  HIP_OK(hipMalloc((void**)&hn, sizeof(double) * kNB));
  HIP_OK(hipMalloc((void**)&bnorm, sizeof(double) * kNB));
  HIP_OK(hipMalloc((void**)&active, sizeof(int) * kNB));
  HIP_OK(hipMalloc((void**)&kq, sizeof(int) * kNB));
  HIP_OK(hipMalloc((void**)&tolq, sizeof(double) * kNB));
  HIP_OK(hipHostMalloc((void**)&host_rel, sizeof(double) * kNB, hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&host_est, sizeof(double) * size_t(m_) * kNB, hipHostMallocDefault));
  HIP_OK(hipHostMalloc((void**)&host_active, sizeof(int) * kNB, hipHostMallocDefault));
  m = m_;
  N = N_;
  return "";
}

std::string BlockWork::flex() {
  if (!Z) HIP_OK(hipMalloc((void**)&Z, sizeof(double) * size_t(N) * kLD * size_t(m)));
  return "";
}

std::string BlockWork::basis(bool f32) {
  const size_t n = size_t(N) * kLD * size_t(m + 1);
  if (f32 && !V16) HIP_OK(hipMalloc((void**)&V16, sizeof(_Float16) * n));
  if (f32 && !w32) HIP_OK(hipMalloc((void**)&w32, sizeof(float) * size_t(N) * kLD));
  if (f32 && !x32) HIP_OK(hipMalloc((void**)&x32, sizeof(float) * size_t(N) * kLD));
  if (!f32 && !V) HIP_OK(hipMalloc((void**)&V, sizeof(double) * n));
  return "";
}

void BlockWork::release() {
  void* ptrs[] = {V, V16, w32, x32, Z, w, partial, h, y, H, cs, sn, g, hn, bnorm, active, kq, tolq};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (host_rel) (void)hipHostFree(host_rel);
  if (host_est) (void)hipHostFree(host_est);
  if (host_active) (void)hipHostFree(host_active);
  *this = BlockWork{};
}

// TV: storage of the Krylov basis (double; _Float16 for the FP32 inner solves,
// whose accuracy target is far above the basis rounding: the basis vectors have
// unit norm per system, so their entries sit in FP16's normal range for any N
// this library builds, and FP16 storage leaves the C3 inner and outer step
// counts unchanged against FP32 storage). Dots, norms, Givens and the solution
// update accumulate in FP64 either way.
template <class TV>
static BatchResult block_gmres_t(BlockOperator& bo, const double* B, double* X, int nrhs, int m, int max_restarts,
                                 double tol, double* rel_residuals, bool fp32, const double* tol_per, bool allow_zero,
                                 bool x_zero) {
  constexpr bool lowv = !std::is_same<TV, double>::value;   // the FP16 basis of the FP32 inner solves
  BatchResult res;
  auto t0 = std::chrono::steady_clock::now();
  DeviceOperator& o = *bo.op;
  hipStream_t s = o.stream;
  const int N = o.N;
  if (nrhs < 1 || nrhs > kNB || m < 1 || m > 500 || max_restarts < 1 || (!(tol > 0) && !tol_per)) {
    res.status = LSG_ERR_ARG;
    res.error = "batched GMRES: 1 <= nrhs <= 64, 1 <= m <= 500, max_restarts >= 1, tol > 0";
    return res;
  }
  const long long vs = (long long)N * kLD;  // doubles per block vector
  const int nrb = (N + kBRows - 1) / kBRows;
  // workspace kept on the block operator between solves of the same restart length
  BlockWork& W = bo.work;
  if (W.m != m || W.N != N) {
    W.release();
    std::string e = W.reserve(N, m);
    if (!e.empty()) {
      W.release();
      res.status = LSG_ERR_DEVICE;
      res.error = "batched GMRES workspace allocation failed: " + e;
      return res;
    }
  }
  {
    std::string e = W.basis(lowv);
    if (!e.empty()) {
      W.release();
      res.status = LSG_ERR_DEVICE;
      res.error = "batched GMRES basis allocation failed: " + e;
      return res;
    }
  }
  TV* V = reinterpret_cast<TV*>(lowv ? (void*)W.V16 : (void*)W.V);
  double* w = W.w;
  // the FP32 inner solves keep the Arnoldi product A v_j in FP32 too (the
  // FP32 operator computed it in FP32) where the leaf kernel stores it
  float* w32 = lowv && fp32 && bo.can_store_f32() ? W.w32 : nullptr;
  double2 *partial = W.partial, *h = W.h, *y = W.y;
  BGivens st{W.H, W.cs, W.sn, W.g, W.hn, W.bnorm, W.active, W.kq, W.tolq};
  double *host_rel = W.host_rel, *host_est = W.host_est;
  int* host_active = W.host_active;
  auto cleanup = [&]() { hipStreamSynchronize(s); };
  auto dot = [&](int nc, const auto* vec, bool norm) {
    typedef std::remove_cv_t<std::remove_pointer_t<decltype(vec)>> TW;
    const int ncx = nc + (norm ? 1 : 0);
    if constexpr (lowv && std::is_same<TW, float>::value)
      hipLaunchKernelGGL(bdot_h_kernel, dim3(nrb, (ncx + kBCols - 1) / kBCols), dim3(kT), 0, s, V, vs, nc, ncx, vec,
                         N, partial);
    else
      hipLaunchKernelGGL((bdot_kernel<TV, TW>), dim3(nrb, (ncx + kBCols - 1) / kBCols), dim3(kT), 0, s, V, vs, nc,
                         ncx, vec, N, partial);
    block_reduce(partial, nrb, ncx, partial + size_t(nrb) * (m + 2) * kNB, h, s);
  };
  // ||b_q||
  dot(0, B, true);
  std::vector<double2> bn2(kNB);
  hipMemcpyAsync(bn2.data(), h, sizeof(double2) * kNB, hipMemcpyDeviceToHost, s);
  if (hipStreamSynchronize(s) != hipSuccess) {
    cleanup();
    res.status = LSG_ERR_DEVICE;
    res.error = "batched GMRES: norm of b failed";
    return res;
  }
  std::vector<double> bnorm(kNB, 1.0), tolv(kNB, tol);
  for (int q = 0; q < nrhs; ++q) {
    bnorm[q] = std::sqrt(std::max(bn2[q].x, 0.0));
    if (tol_per) tolv[q] = tol_per[q];
    if (allow_zero && bnorm[q] == 0.0) {
      bnorm[q] = 1.0;   // a zero system: x stays as given, counted as converged
      continue;
    }
    if (!(bnorm[q] > 0) || !std::isfinite(bnorm[q])) {
      cleanup();
      res.status = LSG_ERR_NUMERIC;
      res.error = "batched GMRES: a right-hand side is zero or holds NaN/Inf";
      return res;
    }
  }
  // the FP32 block matvecs of this solve run the split-BF16 kernels when asked
  SbScope sb_scope(bo, bo.sb_active);
  if (fp32) {
    std::string e = bo.enable_inner();
    if (!e.empty()) {
      cleanup();
      res.status = LSG_ERR_DEVICE;
      res.error = e;
      return res;
    }
    bo.sb_active = bo.inner_sb;
  }
  auto amv = [&](const auto* xin, double* yout) {
    typedef std::remove_cv_t<std::remove_pointer_t<decltype(xin)>> XT;
    if constexpr (!std::is_same<XT, double>::value) {
      bo.matvec32(xin, yout);
    } else {
      if (fp32) bo.matvec32(xin, yout);
      else bo.matvec(xin, yout);
    }
  };
  hipMemcpyAsync(st.bnorm, bnorm.data(), sizeof(double) * kNB, hipMemcpyHostToDevice, s);
  hipMemcpyAsync(st.tolq, tolv.data(), sizeof(double) * kNB, hipMemcpyHostToDevice, s);
  const long long ntot = vs;                                  // doubles per block vector
  const unsigned gv = blocks_for(ntot, kT);                   // one thread per double
  const unsigned gq = blocks_for((long long)N * kNB, kT);     // one thread per (row, system)
  const unsigned gu = blocks_for((N + kUpdRows - 1) / kUpdRows, kT / 64);   // kUpdRows rows per (thread, system)
  int total_it = 0;
  for (int cyc = 0; cyc <= max_restarts; ++cyc) {
    // the cycle's first residual: b itself when x = 0 (its squared norms are
    // still in h from the check of b above), else b - A x
    const double* r0 = w;
    if (cyc == 0 && x_zero) {
      r0 = B;
    } else {
      amv(X, w);
      hipLaunchKernelGGL(bsub_kernel, dim3(gv), dim3(kT), 0, s, B, w, ntot);
      dot(0, w, true);
    }
    hipLaunchKernelGGL(bcycle_init_kernel, dim3(1), dim3(kNB), 0, s, h, st, m, nrhs, host_rel, host_active);
    if (hipStreamSynchronize(s) != hipSuccess) {
      res.status = LSG_ERR_DEVICE;
      res.error = "batched GMRES: device error";
      break;
    }
    int n_act = 0, n_conv = 0;
    double worst = 0.0;
    bool finite = true;
    for (int q = 0; q < nrhs; ++q) {
      n_act += host_active[q];
      n_conv += host_rel[q] <= tolv[q] ? 1 : 0;
      worst = std::max(worst, host_rel[q]);
      finite = finite && std::isfinite(host_rel[q]);
      if (rel_residuals) rel_residuals[q] = host_rel[q];
    }
    res.converged = n_conv;
    res.max_rel_residual = worst;
    if (!finite) {
      res.status = LSG_ERR_NUMERIC;
      res.error = "batched GMRES: residual is NaN/Inf";
      break;
    }
    if (n_act == 0 || cyc == max_restarts) break;
    hipLaunchKernelGGL(bscale_kernel<TV>, dim3(gv), dim3(kT), 0, s, r0, st, V, ntot, w32 ? W.x32 : (float*)nullptr);
    int steps = 0;
    for (int j = 0; j < m;) {
      const int batch = std::min(m - j, 4);
      for (int jj = j; jj < j + batch; ++jj) {
        if constexpr (lowv) {
          if (w32) {
            // the FP32 copy of v_jj (written with it) is the matvec's input
            bo.matvec32((const float*)W.x32, w32);
            dot(jj + 1, (const float*)w32, true);
            hipLaunchKernelGGL(btail_kernel, dim3(1), dim3(kNB), 0, s, h, st, m, jj, host_est, host_active,
                               (const double*)nullptr);
            hipLaunchKernelGGL(bupdate_h_kernel, dim3(blocks_for((N + 4 * kUpdRows - 1) / (4 * kUpdRows), kT / 64)),
                               dim3(kT), 0, s, (const TV*)V, vs, jj + 1, h, (const float*)w32, st.hn,
                               V + size_t(jj + 1) * vs, W.x32, (long long)N);
            continue;
          }
        }
        amv((const TV*)(V + size_t(jj) * vs), w);
        dot(jj + 1, (const double*)w, true);
        hipLaunchKernelGGL(btail_kernel, dim3(1), dim3(kNB), 0, s, h, st, m, jj, host_est, host_active,
                           (const double*)nullptr);
        hipLaunchKernelGGL((bupdate_kernel<TV, double>), dim3(gu), dim3(kT), 0, s, V, vs, jj + 1, h,
                           (const double*)w, st.hn, V + size_t(jj + 1) * vs, (long long)N);
      }
      if (hipStreamSynchronize(s) != hipSuccess) {
        res.status = LSG_ERR_DEVICE;
        res.error = "batched GMRES: device error";
        break;
      }
      j += batch;
      steps = j;
      int still = 0;
      for (int q = 0; q < nrhs; ++q) still += host_active[q];
      if (still == 0) break;
    }
    if (res.status != 0) break;
    total_it += steps;
    hipLaunchKernelGGL(btrisolve_kernel, dim3(1), dim3(kNB), 0, s, st, m, y);
    if constexpr (lowv)
      hipLaunchKernelGGL(baxpy_h_kernel, dim3(blocks_for((N + 3) / 4, kT / 64)), dim3(kT), 0, s, (const _Float16*)V,
                         vs, steps, y, X, (long long)N);
    else
      hipLaunchKernelGGL(baxpy_kernel<TV>, dim3(gq), dim3(kT), 0, s, V, vs, steps, y, X, (long long)N);
    res.restarts = cyc + 1;
  }
  res.iterations = total_it;
  if (res.status == 0 && res.converged < nrhs) {
    res.status = LSG_ERR_NOCONV;
    res.error = "batched GMRES: not every system reached tol";
  }
  cleanup();
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

BatchResult block_gmres(BlockOperator& bo, const double* B, double* X, int nrhs, int m, int max_restarts, double tol,
                        double* rel_residuals, bool fp32, const double* tol_per, bool allow_zero, bool x_zero) {
  // the FP32 inner solves keep their Krylov basis in FP16 (a quarter of the
  // FP64 Gram-Schmidt traffic); the FP64 solver keeps it in FP64
  if (fp32)
    return block_gmres_t<_Float16>(bo, B, X, nrhs, m, max_restarts, tol, rel_residuals, true, tol_per, allow_zero,
                                   x_zero);
  return block_gmres_t<double>(bo, B, X, nrhs, m, max_restarts, tol, rel_residuals, false, tol_per, allow_zero,
                               x_zero);
}

BatchResult block_gmres_ir(BlockOperator& bo, const double* B, double* X, int nrhs, int m, int max_restarts,
                           double tol, double inner_tol, int max_outer, double* rel_residuals) {
  BatchResult res;
  auto t0 = std::chrono::steady_clock::now();
  DeviceOperator& o = *bo.op;
  hipStream_t s = o.stream;
  const int N = o.N;
  if (nrhs < 1 || nrhs > kNB || !(tol > 0) || !(inner_tol > 0) || max_outer < 1) {
    res.status = LSG_ERR_ARG;
    res.error = "batched refinement: 1 <= nrhs <= 64, tol > 0, inner_tol > 0, max_outer >= 1";
    return res;
  }
  std::string e = bo.enable_fp32();
  if (!e.empty()) {
    res.status = LSG_ERR_DEVICE;
    res.error = e;
    return res;
  }
  const long long vs = (long long)N * kLD;
  const int nrb = (N + kBRows - 1) / kBRows;
  double *R = nullptr, *D = nullptr;
  double2 *partial = nullptr, *h = nullptr;
  int* mask = nullptr;
  bool ok = hipMalloc((void**)&R, sizeof(double) * size_t(vs)) == hipSuccess &&
            hipMalloc((void**)&D, sizeof(double) * size_t(vs)) == hipSuccess &&
            hipMalloc((void**)&partial, sizeof(double2) * size_t(nrb + kRS) * kNB) == hipSuccess &&
            hipMalloc((void**)&h, sizeof(double2) * kNB) == hipSuccess &&
            hipMalloc((void**)&mask, sizeof(int) * kNB) == hipSuccess;
  auto cleanup = [&]() {
    hipStreamSynchronize(s);
    void* ptrs[] = {R, D, partial, h, mask};
    for (void* p : ptrs)
      if (p) hipFree(p);
  };
  if (!ok) {
    cleanup();
    res.status = LSG_ERR_DEVICE;
    res.error = "batched refinement workspace allocation failed";
    return res;
  }
  const unsigned gv = blocks_for(vs, kT);
  // per-system 2-norms of a block
  auto norms = [&](const double* v, std::vector<double>& out) {
    hipLaunchKernelGGL((bdot_kernel<double, double>), dim3(nrb, 1), dim3(kT), 0, s, (const double*)nullptr, vs, 0, 1, v, N,
                       partial);
    block_reduce(partial, nrb, 1, partial + size_t(nrb) * kNB, h, s);
    std::vector<double2> hh(kNB);
    hipMemcpyAsync(hh.data(), h, sizeof(double2) * kNB, hipMemcpyDeviceToHost, s);
    const bool good = hipStreamSynchronize(s) == hipSuccess;
    out.assign(kNB, 0.0);
    for (int q = 0; q < kNB; ++q) out[q] = std::sqrt(std::max(hh[q].x, 0.0));
    return good;
  };
  std::vector<double> bn, rn;
  if (!norms(B, bn)) {
    cleanup();
    res.status = LSG_ERR_DEVICE;
    res.error = "batched refinement: device error";
    return res;
  }
  for (int q = 0; q < nrhs; ++q)
    if (!(bn[q] > 0) || !std::isfinite(bn[q])) {
      cleanup();
      res.status = LSG_ERR_NUMERIC;
      res.error = "batched refinement: a right-hand side is zero or holds NaN/Inf";
      return res;
    }
  std::vector<int> maskh(kNB, 0);
  std::vector<double> tol_in(kNB, 1.0);
  for (int outer = 0; res.status == 0; ++outer) {
    bo.matvec(X, R);   // FP64 residual
    hipLaunchKernelGGL(bsub_kernel, dim3(gv), dim3(kT), 0, s, B, R, vs);
    if (!norms(R, rn)) {
      res.status = LSG_ERR_DEVICE;
      res.error = "batched refinement: device error";
      break;
    }
    int n_conv = 0;
    double worst = 0.0;
    bool finite = true;
    for (int q = 0; q < nrhs; ++q) {
      const double rel = rn[q] / bn[q];
      finite = finite && std::isfinite(rel);
      worst = std::max(worst, rel);
      if (rel_residuals) rel_residuals[q] = rel;
      maskh[q] = rel > tol ? 1 : 0;
      n_conv += 1 - maskh[q];
      tol_in[q] = maskh[q] ? std::max(inner_tol, 0.5 * tol * bn[q] / rn[q]) : 1.0;
    }
    res.converged = n_conv;
    res.max_rel_residual = worst;
    if (!finite) {
      res.status = LSG_ERR_NUMERIC;
      res.error = "batched refinement: residual is NaN/Inf";
      break;
    }
    if (n_conv == nrhs || outer == max_outer) break;
    // inner FP32 solves of A d = r for the unconverged systems (converged columns zeroed)
    hipMemcpyAsync(mask, maskh.data(), sizeof(int) * kNB, hipMemcpyHostToDevice, s);
    hipLaunchKernelGGL(bmask_kernel, dim3(gv), dim3(kT), 0, s, R, mask, vs);
    hipMemsetAsync(D, 0, sizeof(double) * size_t(vs), s);
    BatchResult g = block_gmres(bo, R, D, nrhs, m, max_restarts, 0.0, nullptr, true, tol_in.data(), true, true);
    res.iterations += g.iterations;
    res.restarts += g.restarts;
    if (g.status != 0 && g.status != LSG_ERR_NOCONV) {
      res.status = g.status;
      res.error = "inner FP32 batched solve: " + g.error;
      break;
    }
    hipLaunchKernelGGL(badd_kernel, dim3(gv), dim3(kT), 0, s, X, D, vs);
    res.refinements = outer + 1;
  }
  if (res.status == 0 && res.converged < nrhs) {
    res.status = LSG_ERR_NOCONV;
    res.error = "batched refinement: not every system reached tol";
  }
  cleanup();
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

BatchResult block_fgmres(BlockOperator& bo, const double* B, double* X, int nrhs, int m, int max_restarts,
                         double tol, double inner_tol, int K, int max_cycles, double* rel_residuals) {
  BatchResult res;
  auto t0 = std::chrono::steady_clock::now();
  DeviceOperator& o = *bo.op;
  hipStream_t s = o.stream;
  const int N = o.N;
  if (nrhs < 1 || nrhs > kNB || !(tol > 0) || !(inner_tol > 0) || K < 1 || K > 500 || m < 1 || m > 500 ||
      max_restarts < 1 || max_cycles < 1) {
    res.status = LSG_ERR_ARG;
    res.error = "batched flexible GMRES: 1 <= nrhs <= 64, tol > 0, inner_tol > 0, 1 <= K, m <= 500";
    return res;
  }
  const long long vs = (long long)N * kLD;
  const int nrb = (N + kBRows - 1) / kBRows;
  BlockWork& W = bo.outer;
  if (W.m != K || W.N != N) {
    W.release();
    std::string e = W.reserve(N, K);
    if (!e.empty()) {
      W.release();
      res.status = LSG_ERR_DEVICE;
      res.error = "batched flexible GMRES workspace allocation failed: " + e;
      return res;
    }
  }
  {
    std::string e = W.basis(false);
    if (e.empty()) e = W.flex();
    if (!e.empty()) {
      W.release();
      res.status = LSG_ERR_DEVICE;
      res.error = "batched flexible GMRES basis allocation failed: " + e;
      return res;
    }
  }
  double *V = W.V, *Z = W.Z, *w = W.w;
  double2 *partial = W.partial, *h = W.h, *y = W.y;
  // per system: ones (the unnormalised first projection) and ||w'||^2 - |h2|^2
  double* small = nullptr;
  if (hipMalloc((void**)&small, sizeof(double) * 2 * kNB) != hipSuccess) {
    res.status = LSG_ERR_DEVICE;
    res.error = "batched flexible GMRES: allocation failed";
    return res;
  }
  struct Release {
    hipStream_t s;
    void* p;
    ~Release() {
      (void)hipStreamSynchronize(s);
      (void)hipFree(p);
    }
  } release_small{s, small};
  double *ones = small, *hn2 = small + kNB;
  hipLaunchKernelGGL(bfill_kernel, dim3(1), dim3(kNB), 0, s, ones, kNB, 1.0);
  BGivens st{W.H, W.cs, W.sn, W.g, W.hn, W.bnorm, W.active, W.kq, W.tolq};
  double *host_rel = W.host_rel, *host_est = W.host_est;
  int* host_active = W.host_active;
  auto dot = [&](int nc, const double* vec) {
    const int ncx = nc + 1;
    hipLaunchKernelGGL((bdot_kernel<double, double>), dim3(nrb, (ncx + kBCols - 1) / kBCols), dim3(kT), 0, s, V, vs,
                       nc, ncx, vec, N, partial);
    block_reduce(partial, nrb, ncx, partial + size_t(nrb) * (K + 2) * kNB, h, s);
  };
  dot(0, B);
  std::vector<double2> bn2(kNB);
  hipMemcpyAsync(bn2.data(), h, sizeof(double2) * kNB, hipMemcpyDeviceToHost, s);
  if (hipStreamSynchronize(s) != hipSuccess) {
    res.status = LSG_ERR_DEVICE;
    res.error = "batched flexible GMRES: norm of b failed";
    return res;
  }
  std::vector<double> bnorm(kNB, 1.0), tolv(kNB, tol), tol_in(kNB, inner_tol);
  for (int q = 0; q < nrhs; ++q) {
    bnorm[q] = std::sqrt(std::max(bn2[q].x, 0.0));
    if (!(bnorm[q] > 0) || !std::isfinite(bnorm[q])) {
      (void)hipStreamSynchronize(s);
      res.status = LSG_ERR_NUMERIC;
      res.error = "batched flexible GMRES: a right-hand side is zero or holds NaN/Inf";
      return res;
    }
  }
  hipMemcpyAsync(st.bnorm, bnorm.data(), sizeof(double) * kNB, hipMemcpyHostToDevice, s);
  hipMemcpyAsync(st.tolq, tolv.data(), sizeof(double) * kNB, hipMemcpyHostToDevice, s);
  const unsigned gv = blocks_for(vs, kT);
  const unsigned gq = blocks_for((long long)N * kNB, kT);
  const unsigned gu = blocks_for((N + kUpdRows - 1) / kUpdRows, kT / 64);
  for (int cyc = 0; cyc <= max_cycles; ++cyc) {
    bo.matvec(X, w);   // true FP64 residual
    hipLaunchKernelGGL(bsub_kernel, dim3(gv), dim3(kT), 0, s, B, w, vs);
    dot(0, w);
    hipLaunchKernelGGL(bcycle_init_kernel, dim3(1), dim3(kNB), 0, s, h, st, K, nrhs, host_rel, host_active);
    if (hipStreamSynchronize(s) != hipSuccess) {
      res.status = LSG_ERR_DEVICE;
      res.error = "batched flexible GMRES: device error";
      break;
    }
    int n_act = 0, n_conv = 0;
    double worst = 0.0;
    bool finite = true;
    for (int q = 0; q < nrhs; ++q) {
      n_act += host_active[q];
      n_conv += host_rel[q] <= tol ? 1 : 0;
      worst = std::max(worst, host_rel[q]);
      finite = finite && std::isfinite(host_rel[q]);
      if (rel_residuals) rel_residuals[q] = host_rel[q];
    }
    res.converged = n_conv;
    res.max_rel_residual = worst;
    if (!finite) {
      res.status = LSG_ERR_NUMERIC;
      res.error = "batched flexible GMRES: residual is NaN/Inf";
      break;
    }
    if (n_act == 0 || cyc == max_cycles) break;
    hipLaunchKernelGGL(bscale_kernel<double>, dim3(gv), dim3(kT), 0, s, w, st, V, vs, (float*)nullptr);
    int steps = 0;
    for (int k = 0; k < K; ++k) {
      // z_k ~ A^-1 v_k by the FP32 batched GMRES (columns of frozen systems are
      // zero: converged at once)
      double* zk = Z + size_t(k) * vs;
      hipMemsetAsync(zk, 0, sizeof(double) * size_t(vs), s);
      BatchResult g = block_gmres(bo, V + size_t(k) * vs, zk, nrhs, m, max_restarts, 0.0, nullptr, true, tol_in.data(),
                                  true, true);
      res.iterations += g.iterations;
      if (g.status != 0 && g.status != LSG_ERR_NOCONV) {
        res.status = g.status;
        res.error = "inner FP32 batched solve: " + g.error;
        break;
      }
      bo.matvec(zk, w);
      // classical Gram-Schmidt twice: A z_k lies almost in span(V) (z_k solves
      // A z = v_k to inner_tol), so the new direction is a small remainder
      // whose norm cannot come from ||w||^2 - |h|^2 (that difference cancels)
      dot(k + 1, w);
      hipMemcpyAsync(y, h, sizeof(double2) * size_t(k + 1) * kNB, hipMemcpyDeviceToDevice, s);
      hipLaunchKernelGGL((bupdate_kernel<double, double>), dim3(gu), dim3(kT), 0, s, V, vs, k + 1, y,
                         (const double*)w, ones, w, (long long)N);
      dot(k + 1, w);
      hipLaunchKernelGGL(bcgs2_combine_kernel, dim3(1), dim3(kNB), 0, s, y, h, k + 1, hn2);
      hipLaunchKernelGGL(btail_kernel, dim3(1), dim3(kNB), 0, s, y, st, K, k, host_est, host_active,
                         (const double*)hn2);
      hipLaunchKernelGGL((bupdate_kernel<double, double>), dim3(gu), dim3(kT), 0, s, V, vs, k + 1, h,
                         (const double*)w, st.hn, V + size_t(k + 1) * vs, (long long)N);
      ++res.refinements;
      steps = k + 1;
      if (hipStreamSynchronize(s) != hipSuccess) {
        res.status = LSG_ERR_DEVICE;
        res.error = "batched flexible GMRES: device error";
        break;
      }
      int still = 0;
      for (int q = 0; q < nrhs; ++q) still += host_active[q];
      if (still == 0) break;
    }
    if (res.status != 0) break;
    hipLaunchKernelGGL(btrisolve_kernel, dim3(1), dim3(kNB), 0, s, st, K, y);
    hipLaunchKernelGGL(baxpy_kernel<double>, dim3(gq), dim3(kT), 0, s, Z, vs, steps, y, X, (long long)N);
    res.restarts = cyc + 1;
  }
  if (res.status == 0 && res.converged < nrhs) {
    res.status = LSG_ERR_NOCONV;
    res.error = "batched flexible GMRES: not every system reached tol";
  }
  (void)hipStreamSynchronize(s);
  res.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return res;
}

}  // namespace lsg
