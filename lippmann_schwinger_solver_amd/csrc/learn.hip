// Learned preconditioner: convolutional network training and inference
// kernels (BF16 storage, FP32 accumulation), Adam, and the node <-> grid maps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lsg_learn.h"
#include "learn.hpp"
#include "numerics.hpp"

namespace lsg {
namespace {

constexpr int kT = 256;
constexpr int kWidth = 32;     // hidden channels (the kernels are instantiated for it)
constexpr int kWPix = 32;      // pixels per LDS tile in the weight-gradient kernel
constexpr int kWChunk = 1024;  // pixels per workgroup in the weight-gradient kernel

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (finite inputs)
  const uint32_t u = __float_as_uint(f);
  return uint16_t((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// out[b, y, x, co] = act(bias[co] + sum_{ky,kx,ci} w[co, ky, kx, ci] in[b, y + (ky-1) d, x + (kx-1) d, ci])
template <int CIN, int COUT>
__global__ __launch_bounds__(kT) void conv_fwd_kernel(const uint16_t* __restrict__ in, const uint16_t* __restrict__ w,
                                                      const float* __restrict__ bias, int G, int dil, int relu,
                                                      uint16_t* __restrict__ out_bf, float* __restrict__ out_f) {
  __shared__ float ws[COUT * 9 * CIN];
  __shared__ float bs[COUT];
  const int t = threadIdx.x;
  for (int e = t; e < COUT * 9 * CIN; e += kT) ws[e] = bf2f(w[e]);
  if (t < COUT) bs[t] = bias[t];
  __syncthreads();
  const int pix = blockIdx.x * kT + t, b = blockIdx.y;
  if (pix >= G * G) return;
  const int y = pix / G, x = pix - y * G;
  float acc[COUT];
#pragma unroll
  for (int co = 0; co < COUT; ++co) acc[co] = bs[co];
  for (int k = 0; k < 9; ++k) {
    const int yy = y + (k / 3 - 1) * dil, xx = x + (k % 3 - 1) * dil;
    if (yy < 0 || yy >= G || xx < 0 || xx >= G) continue;
    const uint16_t* ip = in + ((size_t(b) * G + yy) * G + xx) * CIN;
    float v[CIN];
#pragma unroll
    for (int ci = 0; ci < CIN; ++ci) v[ci] = bf2f(ip[ci]);
#pragma unroll
    for (int co = 0; co < COUT; ++co) {
      const float* wr = ws + (co * 9 + k) * CIN;
#pragma unroll
      for (int ci = 0; ci < CIN; ++ci) acc[co] = fmaf(wr[ci], v[ci], acc[co]);
    }
  }
  const size_t o = (size_t(b) * G * G + pix) * COUT;
#pragma unroll
  for (int co = 0; co < COUT; ++co) {
    const float a = relu ? fmaxf(acc[co], 0.0f) : acc[co];
    if (out_bf) out_bf[o + co] = f2bf(a);
    if (out_f) out_f[o + co] = a;
  }
}

// gin[b, q, ci] = sum_{k, co} w[co, k, ci] g[b, q - off_k, co]
template <int CIN, int COUT>
__global__ __launch_bounds__(kT) void conv_bwd_data_kernel(const float* __restrict__ g, const uint16_t* __restrict__ w,
                                                           int G, int dil, float* __restrict__ gin) {
  __shared__ float ws[COUT * 9 * CIN];
  const int t = threadIdx.x;
  for (int e = t; e < COUT * 9 * CIN; e += kT) ws[e] = bf2f(w[e]);
  __syncthreads();
  const int pix = blockIdx.x * kT + t, b = blockIdx.y;
  if (pix >= G * G) return;
  const int y = pix / G, x = pix - y * G;
  float acc[CIN];
#pragma unroll
  for (int ci = 0; ci < CIN; ++ci) acc[ci] = 0.0f;
  for (int k = 0; k < 9; ++k) {
    const int yy = y - (k / 3 - 1) * dil, xx = x - (k % 3 - 1) * dil;
    if (yy < 0 || yy >= G || xx < 0 || xx >= G) continue;
    const float* gp = g + ((size_t(b) * G + yy) * G + xx) * COUT;
#pragma unroll
    for (int co = 0; co < COUT; ++co) {
      const float gv = gp[co];
      const float* wr = ws + (co * 9 + k) * CIN;
#pragma unroll
      for (int ci = 0; ci < CIN; ++ci) acc[ci] = fmaf(wr[ci], gv, acc[ci]);
    }
  }
  float* op = gin + (size_t(b) * G * G + pix) * CIN;
#pragma unroll
  for (int ci = 0; ci < CIN; ++ci) op[ci] = acc[ci];
}

// ---------------------------------------------------------------------------
// Hidden 32 -> 32 layers on the matrix cores. A wave owns a tile of 16
// consecutive pixels of one image row (G is a multiple of 16) and forms each
// 3x3 convolution as 9 shifted GEMMs, one per tap.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kTilesPerWg = kT / 64;

__device__ __forceinline__ bf16x8 load_bf16x8(const uint16_t* p) {
  bf16x8 v;
  __builtin_memcpy(&v, p, 16);
  return v;
}

// Forward on v_mfma_f32_16x16x32_bf16: per tap, C[16 pixels x 16 cout] +=
// A[16 pixels x 32 cin] B[32 cin x 16 cout]; lane l supplies A(pixel l%16,
// cin 8(l/16)..+8) from the NHWC input (one 16-byte load) and B(cin
// 8(l/16)..+8, cout l%16) from the weights [cout][tap][cin] (one 16-byte
// load, kept in registers for all tiles: 9 taps x 2 cout tiles); the result
// lane l holds pixels 4(l/16)+i, cout l%16. FP32 accumulation, bias and ReLU
// in FP32, output rounded to BF16 (and kept in FP32 when out_f is given).
__global__ __launch_bounds__(kT) void conv_fwd_mfma_kernel(const uint16_t* __restrict__ in,
                                                           const uint16_t* __restrict__ w,
                                                           const float* __restrict__ bias, int G, int dil, int relu,
                                                           long long ntiles, uint16_t* __restrict__ out_bf,
                                                           float* __restrict__ out_f) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, kg = lane >> 4;
  bf16x8 wf[9][2];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) wf[t][ct] = load_bf16x8(w + ((ct * 16 + r16) * 9 + t) * kWidth + 8 * kg);
  const float b0 = bias[r16], b1 = bias[16 + r16];
  const int GG = G * G;
  const bf16x8 zero = {};
  for (long long tile = (long long)blockIdx.x * kTilesPerWg + wave; tile < ntiles;
       tile += (long long)gridDim.x * kTilesPerWg) {
    const long long p0 = tile * 16;
    const int b = int(p0 / GG), pix0 = int(p0 - (long long)b * GG);
    const int y = pix0 / G, x = pix0 - y * G + r16;
    const uint16_t* img = in + (size_t)b * GG * kWidth;
    f32x4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int yy = y + (t / 3 - 1) * dil, xx = x + (t % 3 - 1) * dil;
      const bool ok = yy >= 0 && yy < G && xx >= 0 && xx < G;
      const bf16x8 a = ok ? load_bf16x8(img + ((size_t)yy * G + xx) * kWidth + 8 * kg) : zero;
      acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, wf[t][0], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, wf[t][1], acc1, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const size_t o = (size_t)(p0 + 4 * kg + i) * kWidth;
      float v0 = acc0[i] + b0, v1 = acc1[i] + b1;
      if (relu) {
        v0 = fmaxf(v0, 0.0f);
        v1 = fmaxf(v1, 0.0f);
      }
      if (out_bf) {
        out_bf[o + r16] = f2bf(v0);
        out_bf[o + 16 + r16] = f2bf(v1);
      }
      if (out_f) {
        out_f[o + r16] = v0;
        out_f[o + 16 + r16] = v1;
      }
    }
  }
}

// Data gradient on v_mfma_f32_16x16x4f32: gin[p, ci] = sum_{tap, co}
// g[p - off_tap, co] w[co, tap, ci]. Per tap, 8 k-steps of 4 cout: in k-step
// s, lane l supplies A(pixel l%16, k l/16) = g(pixel - off, cout 8(l/16) + s)
// (its 8 cout values are one contiguous 32-byte read) and B(k l/16, cin l%16)
// = w(cout 8(l/16) + s, tap, cin) from the weights staged in LDS as FP32;
// the result lane l holds pixels 4(l/16)+v, cin l%16 (+16 for the second tile).
__global__ __launch_bounds__(kT) void conv_bwd_data_mfma_kernel(const float* __restrict__ g,
                                                                const uint16_t* __restrict__ w, int G, int dil,
                                                                long long ntiles, float* __restrict__ gin) {
  __shared__ float ws[kWidth * 9 * kWidth];
  for (int e = threadIdx.x; e < kWidth * 9 * kWidth; e += kT) ws[e] = bf2f(w[e]);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, kg = lane >> 4;
  const int GG = G * G;
  for (long long tile = (long long)blockIdx.x * kTilesPerWg + wave; tile < ntiles;
       tile += (long long)gridDim.x * kTilesPerWg) {
    const long long p0 = tile * 16;
    const int b = int(p0 / GG), pix0 = int(p0 - (long long)b * GG);
    const int y = pix0 / G, x = pix0 - y * G + r16;
    const float* img = g + (size_t)b * GG * kWidth;
    f32x4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
    for (int t = 0; t < 9; ++t) {
      const int yy = y - (t / 3 - 1) * dil, xx = x - (t % 3 - 1) * dil;
      float a[8];
      if (yy >= 0 && yy < G && xx >= 0 && xx < G) {
        const float4* gp = reinterpret_cast<const float4*>(img + ((size_t)yy * G + xx) * kWidth + 8 * kg);
        const float4 u = gp[0], v = gp[1];
        a[0] = u.x; a[1] = u.y; a[2] = u.z; a[3] = u.w;
        a[4] = v.x; a[5] = v.y; a[6] = v.z; a[7] = v.w;
      } else {
#pragma unroll
        for (int s = 0; s < 8; ++s) a[s] = 0.0f;
      }
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const float* wr = ws + ((8 * kg + s) * 9 + t) * kWidth;
        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], wr[r16], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a[s], wr[16 + r16], acc1, 0, 0, 0);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float* op = gin + (size_t)(p0 + 4 * kg + v) * kWidth;
      op[r16] = acc0[v];
      op[16 + r16] = acc1[v];
    }
  }
}

// Weight and bias gradients on v_mfma_f32_16x16x4f32, per workgroup of
// kWChunk pixels of one image (the partial layout of conv_bwd_w_kernel):
// dW[co, tap, ci] = sum_p g[p, co] in[p + off_tap, ci]. In a k-step of 4
// pixels, lane l supplies A(cout l%16 (+16), k l/16) = g(p0 + l/16, cout) and
// B(k l/16, cin l%16 (+16)) = in(p0 + l/16 + off, cin); wave w owns taps w,
// w+4, w+8 (2 x 2 cout/cin tiles each); wave 0 also sums the bias gradient.
__global__ __launch_bounds__(kT) void conv_bwd_w_mfma_kernel(const uint16_t* __restrict__ in,
                                                             const float* __restrict__ g, int G, int dil,
                                                             float* __restrict__ part) {
  constexpr int NW = kWidth * 9 * kWidth;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r16 = lane & 15, kg = lane >> 4;
  const int GG = G * G;
  const int chunks = (GG + kWChunk - 1) / kWChunk;
  const int b = blockIdx.x / chunks, c0 = (blockIdx.x % chunks) * kWChunk;
  const int c1 = min(GG, c0 + kWChunk);
  const float* gi = g + (size_t)b * GG * kWidth;
  const uint16_t* ii = in + (size_t)b * GG * kWidth;
  constexpr int kMaxTaps = 3;
  int taps[kMaxTaps];
  int ntap = 0;
  for (int t = wave; t < 9; t += 4) taps[ntap++] = t;
  f32x4 acc[kMaxTaps][2][2];
#pragma unroll
  for (int q = 0; q < kMaxTaps; ++q)
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) acc[q][mt][ct] = f32x4{0, 0, 0, 0};
  float bsum0 = 0.0f, bsum1 = 0.0f;
  for (int p0 = c0; p0 < c1; p0 += 4) {
    const int p = p0 + kg;
    float a0 = 0.0f, a1 = 0.0f;
    int y = 0, x = 0;
    if (p < c1) {
      a0 = gi[(size_t)p * kWidth + r16];
      a1 = gi[(size_t)p * kWidth + 16 + r16];
      y = p / G;
      x = p - y * G;
    }
    bsum0 += a0;
    bsum1 += a1;
#pragma unroll
    for (int q = 0; q < kMaxTaps; ++q) {
      if (q < ntap) {
        const int t = taps[q];
        const int yy = y + (t / 3 - 1) * dil, xx = x + (t % 3 - 1) * dil;
        float v0 = 0.0f, v1 = 0.0f;
        if (p < c1 && yy >= 0 && yy < G && xx >= 0 && xx < G) {
          const uint16_t* ip = ii + ((size_t)yy * G + xx) * kWidth;
          v0 = bf2f(ip[r16]);
          v1 = bf2f(ip[16 + r16]);
        }
        acc[q][0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, v0, acc[q][0][0], 0, 0, 0);
        acc[q][0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, v1, acc[q][0][1], 0, 0, 0);
        acc[q][1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, v0, acc[q][1][0], 0, 0, 0);
        acc[q][1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, v1, acc[q][1][1], 0, 0, 0);
      }
    }
  }
  float* P = part + (size_t)blockIdx.x * (NW + kWidth);
#pragma unroll
  for (int q = 0; q < kMaxTaps; ++q) {
    if (q < ntap) {
      const int t = taps[q];
#pragma unroll
      for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int co = 16 * mt + 4 * kg + v, ci = 16 * ct + r16;
            P[(co * 9 + t) * kWidth + ci] = acc[q][mt][ct][v];
          }
    }
  }
  if (wave == 0) {
    // bias: sum over the 4 lanes sharing r16 (one per pixel of a k-step)
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
      bsum0 += __shfl_xor(bsum0, o);
      bsum1 += __shfl_xor(bsum1, o);
    }
    if (kg == 0) {
      P[NW + r16] = bsum0;
      P[NW + 16 + r16] = bsum1;
    }
  }
}

// g *= (act > 0): gradient through ReLU (act is the layer's stored output)
__global__ void relu_bwd_kernel(float* __restrict__ g, const uint16_t* __restrict__ act, long long n) {
  const long long i = (long long)blockIdx.x * kT + threadIdx.x;
  if (i < n && !(bf2f(act[i]) > 0.0f)) g[i] = 0.0f;
}

// Partial weight and bias gradients of one workgroup (kWChunk pixels of one
// sample): part[wg, (co*9 + k)*CIN + ci] = sum_p g[p, co] in[p + off_k, ci],
// part[wg, COUT*9*CIN + co] = sum_p g[p, co].
template <int CIN, int COUT>
__global__ __launch_bounds__(kT) void conv_bwd_w_kernel(const uint16_t* __restrict__ in, const float* __restrict__ g,
                                                        int G, int dil, float* __restrict__ part) {
  constexpr int NW = COUT * 9 * CIN;
  constexpr int PER = (NW + COUT + kT - 1) / kT;
  __shared__ float Gt[kWPix][COUT];
  __shared__ float It[kWPix][9 * CIN];
  const int t = threadIdx.x;
  const int chunks = (G * G + kWChunk - 1) / kWChunk;
  const int b = blockIdx.x / chunks, c0 = (blockIdx.x % chunks) * kWChunk;
  const int c1 = min(G * G, c0 + kWChunk);
  // parameter u of this thread: gradient column og[u] and LDS input offset oi[u]
  // (oi < 0: bias, og < 0: none)
  float acc[PER];
  int og[PER], oi[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    acc[u] = 0.0f;
    const int e = t + u * kT;
    if (e < NW) {
      const int co = e / (9 * CIN), r = e - co * 9 * CIN;
      og[u] = co;
      oi[u] = r;
    } else if (e < NW + COUT) {
      og[u] = e - NW;
      oi[u] = -1;
    } else {
      og[u] = -1;
      oi[u] = -1;
    }
  }
  for (int p0 = c0; p0 < c1; p0 += kWPix) {
    __syncthreads();
    for (int e = t; e < kWPix * COUT; e += kT) {
      const int pl = e / COUT, co = e - pl * COUT, p = p0 + pl;
      Gt[pl][co] = p < c1 ? g[(size_t(b) * G * G + p) * COUT + co] : 0.0f;
    }
    for (int e = t; e < kWPix * 9 * CIN; e += kT) {
      const int pl = e / (9 * CIN), r = e - pl * 9 * CIN, k = r / CIN, ci = r - k * CIN, p = p0 + pl;
      float v = 0.0f;
      if (p < c1) {
        const int y = p / G, x = p - y * G;
        const int yy = y + (k / 3 - 1) * dil, xx = x + (k % 3 - 1) * dil;
        if (yy >= 0 && yy < G && xx >= 0 && xx < G) v = bf2f(in[((size_t(b) * G + yy) * G + xx) * CIN + ci]);
      }
      It[pl][r] = v;
    }
    __syncthreads();
    for (int pl = 0; pl < kWPix; ++pl) {
#pragma unroll
      for (int u = 0; u < PER; ++u) {
        if (og[u] >= 0) acc[u] = fmaf(Gt[pl][og[u]], oi[u] >= 0 ? It[pl][oi[u]] : 1.0f, acc[u]);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int e = t + u * kT;
    if (e < NW + COUT) part[(size_t)blockIdx.x * (NW + COUT) + e] = acc[u];
  }
}

// out[e] = sum_p part[p * n + e]
__global__ void reduce_parts_kernel(const float* __restrict__ part, int nparts, int n, float* __restrict__ out) {
  const int e = blockIdx.x * kT + threadIdx.x;
  if (e >= n) return;
  float s = 0.0f;
  for (int p = 0; p < nparts; ++p) s += part[(size_t)p * n + e];
  out[e] = s;
}

// grad = 2 (out - tgt) / n; per-block sums of (out - tgt)^2
__global__ __launch_bounds__(kT) void mse_kernel(const float* __restrict__ out, const float* __restrict__ tgt,
                                                 long long n, float* __restrict__ grad, float* __restrict__ lpart) {
  __shared__ float red[kT];
  const long long i = (long long)blockIdx.x * kT + threadIdx.x;
  float d = 0.0f;
  if (i < n) {
    d = out[i] - tgt[i];
    grad[i] = 2.0f * d / float(n);
  }
  red[threadIdx.x] = d * d;
  __syncthreads();
  for (int s = kT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long long n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
  const long long i = (long long)blockIdx.x * kT + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void to_bf16_kernel(const float* __restrict__ a, uint16_t* __restrict__ b, long long n) {
  const long long i = (long long)blockIdx.x * kT + threadIdx.x;
  if (i < n) b[i] = f2bf(a[i]);
}

__global__ void gather_batch_kernel(const uint16_t* __restrict__ din, const float* __restrict__ dtgt,
                                    const int* __restrict__ idx, int GG, uint16_t* __restrict__ act0,
                                    float* __restrict__ tgt) {
  const long long e = (long long)blockIdx.x * kT + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= GG) return;
  const size_t s = size_t(idx[b]);
  for (int c = 0; c < 3; ++c) act0[(size_t(b) * GG + e) * 3 + c] = din[(s * GG + e) * 3 + c];
  for (int c = 0; c < 2; ++c) tgt[(size_t(b) * GG + e) * 2 + c] = dtgt[(s * GG + e) * 2 + c];
}

// y[r] = sum_j val[j] x[col[j]], j in [rowptr[r], rowptr[r+1]); complex x, y, real weights
// y[r] = base[r] + scale * sum_j val[j] x[col[j]] (base optional)
__global__ void spmv_c_kernel(const int* __restrict__ rowptr, const int* __restrict__ col,
                              const double* __restrict__ val, const double2* __restrict__ x, double2* __restrict__ y,
                              int nrows, const double* __restrict__ scale, const double2* __restrict__ base) {
  const int r = blockIdx.x * kT + threadIdx.x;
  if (r >= nrows) return;
  double2 a = make_double2(0.0, 0.0);
  for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
    const double w = val[j];
    const double2 v = x[col[j]];
    a.x += w * v.x;
    a.y += w * v.y;
  }
  const double s = scale ? scale[0] : 1.0;
  double2 o = make_double2(a.x * s, a.y * s);
  if (base) {
    o.x += base[r].x;
    o.y += base[r].y;
  }
  y[r] = o;
}

// sum of |v|^2 into out[0] (one workgroup)
__global__ __launch_bounds__(kT) void norm2_kernel(const double2* __restrict__ v, int n, double* __restrict__ out) {
  __shared__ double red[kT];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kT) s += v[i].x * v[i].x + v[i].y * v[i].y;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = kT / 2; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// network input from a grid vector: [q, Re g / sigma, Im g / sigma]; sigma = sqrt(nrm2 / n) (RMS);
// sigma is also written to sig_out for the output scaling
__global__ void prec_input_kernel(const double* __restrict__ qgrid, const double2* __restrict__ g,
                                  const double* __restrict__ nrm2, int n_nodes, int GG, uint16_t* __restrict__ act0,
                                  double* __restrict__ sig_out) {
  const int p = blockIdx.x * kT + threadIdx.x;
  const double sigma = sqrt(fmax(nrm2[0], 0.0) / double(n_nodes));
  const double inv = sigma > 0.0 ? 1.0 / sigma : 0.0;
  if (p == 0) sig_out[0] = sigma;
  if (p >= GG) return;
  act0[size_t(p) * 3 + 0] = f2bf(float(qgrid[p]));
  act0[size_t(p) * 3 + 1] = f2bf(float(g[p].x * inv));
  act0[size_t(p) * 3 + 2] = f2bf(float(g[p].y * inv));
}

__global__ void out_to_complex_kernel(const float* __restrict__ o, double2* __restrict__ g, int GG) {
  const int p = blockIdx.x * kT + threadIdx.x;
  if (p < GG) g[p] = make_double2(double(o[2 * p]), double(o[2 * p + 1]));
}

inline unsigned nblk(long long n) { return unsigned((n + kT - 1) / kT); }

#define HIP_OK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return std::string(#expr ": ") + hipGetErrorString(e_); \
  } while (0)

}  // namespace

// ----------------------------------------------------------------- network --
std::string Net::create(int n_layers, int width, const int* dil, int grid, int max_batch, uint64_t seed) {
  if (n_layers < 2 || n_layers > LSGN_MAX_LAYERS) return "n_layers must be in [2, 8]";
  if (width != kWidth) return "width must be 32";
  if (grid < 16 || grid > 512 || grid % 16) return "grid must be a multiple of 16 in [16, 512]";
  if (max_batch < 1 || max_batch > 256) return "max_batch must be in [1, 256]";
  G = grid;
  maxB = max_batch;
  layers.clear();
  int64_t off = 0;
  for (int l = 0; l < n_layers; ++l) {
    Layer L;
    L.cin = l == 0 ? 3 : width;
    L.cout = l == n_layers - 1 ? 2 : width;
    L.dil = dil ? dil[l] : 1;
    if (L.dil < 1 || L.dil > G / 2) return "dilation out of range";
    L.relu = l < n_layers - 1;
    L.w_off = off;
    off += int64_t(L.cout) * 9 * L.cin;
    L.b_off = off;
    off += L.cout;
    layers.push_back(L);
  }
  nparams = off;
  std::vector<float> h(size_t(nparams), 0.0f);
  SplitMix64 rng(seed);
  for (size_t l = 0; l + 1 < layers.size(); ++l) {
    // He-normal, Box-Muller on the SplitMix64 stream; the last layer starts at
    // zero, so an untrained network is the identity preconditioner
    const Layer& L = layers[l];
    const double sd = std::sqrt(2.0 / (9.0 * L.cin));
    for (int64_t i = 0; i < int64_t(L.cout) * 9 * L.cin; ++i) {
      const double u1 = std::max(rng.uniform(), 1e-300), u2 = rng.uniform();
      h[size_t(L.w_off + i)] = float(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2));
    }
  }
  HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  const size_t np = size_t(nparams);
  HIP_OK(hipMalloc((void**)&p, np * sizeof(float)));
  HIP_OK(hipMalloc((void**)&pw, np * sizeof(uint16_t)));
  HIP_OK(hipMalloc((void**)&grad, np * sizeof(float)));
  HIP_OK(hipMalloc((void**)&m, np * sizeof(float)));
  HIP_OK(hipMalloc((void**)&v, np * sizeof(float)));
  HIP_OK(hipMemcpy(p, h.data(), np * sizeof(float), hipMemcpyHostToDevice));
  HIP_OK(hipMemset(m, 0, np * sizeof(float)));
  HIP_OK(hipMemset(v, 0, np * sizeof(float)));
  const size_t GG = size_t(G) * G;
  for (int l = 0; l <= n_layers; ++l) {
    const int c = l == 0 ? 3 : layers[l - 1].cout;
    uint16_t* a = nullptr;
    HIP_OK(hipMalloc((void**)&a, size_t(maxB) * GG * c * sizeof(uint16_t)));
    act.push_back(a);
  }
  HIP_OK(hipMalloc((void**)&out, size_t(maxB) * GG * 2 * sizeof(float)));
  HIP_OK(hipMalloc((void**)&tgt, size_t(maxB) * GG * 2 * sizeof(float)));
  HIP_OK(hipMalloc((void**)&g0, size_t(maxB) * GG * width * sizeof(float)));
  HIP_OK(hipMalloc((void**)&g1, size_t(maxB) * GG * width * sizeof(float)));
  const int chunks = int((GG + kWChunk - 1) / kWChunk);
  HIP_OK(hipMalloc((void**)&wpart, size_t(maxB) * chunks * (size_t(width) * 9 * width + width) * sizeof(float)));
  HIP_OK(hipMalloc((void**)&lpart, size_t(nblk((long long)maxB * GG * 2)) * sizeof(float)));
  HIP_OK(hipMalloc((void**)&didx, size_t(maxB) * sizeof(int)));
  return sync_weights();
}

std::string Net::sync_weights() {
  hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk(nparams)), dim3(kT), 0, stream, p, pw, (long long)nparams);
  HIP_OK(hipStreamSynchronize(stream));
  return "";
}

void Net::release() {
  if (stream) (void)hipStreamSynchronize(stream);
  void* ptrs[] = {p, pw, grad, m, v, out, tgt, g0, g1, wpart, lpart, didx, din, dtgt};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (uint16_t* a : act)
    if (a) (void)hipFree(a);
  act.clear();
  p = grad = m = v = out = tgt = g0 = g1 = wpart = lpart = dtgt = nullptr;
  pw = din = nullptr;
  didx = nullptr;
  if (stream) (void)hipStreamDestroy(stream);
  stream = nullptr;
}

// the 32 -> 32 layers run on the matrix cores when their weights are 16-byte
// aligned (the 16-byte fragment loads); G is a multiple of 16 by construction
static bool mfma_ok(const Net::Layer& L) { return L.cin == kWidth && L.cout == kWidth && L.w_off % 8 == 0; }
static unsigned mfma_grid(long long ntiles) {
  return unsigned(std::max(1LL, std::min(4096LL, (ntiles + kTilesPerWg - 1) / kTilesPerWg)));
}

void Net::forward(int B, hipStream_t s) {
  const size_t GG = size_t(G) * G;
  const dim3 grid(nblk((long long)GG), B);
  const long long ntiles = (long long)B * (long long)GG / 16;
  for (size_t l = 0; l < layers.size(); ++l) {
    const Layer& L = layers[l];
    const bool last = l + 1 == layers.size();
    uint16_t* o_bf = act[l + 1];
    float* o_f = last ? out : nullptr;
    if (L.cin == 3)
      hipLaunchKernelGGL((conv_fwd_kernel<3, kWidth>), grid, dim3(kT), 0, s, act[l], pw + L.w_off, p + L.b_off, G,
                         L.dil, L.relu ? 1 : 0, o_bf, o_f);
    else if (L.cout == 2)
      hipLaunchKernelGGL((conv_fwd_kernel<kWidth, 2>), grid, dim3(kT), 0, s, act[l], pw + L.w_off, p + L.b_off, G,
                         L.dil, L.relu ? 1 : 0, o_bf, o_f);
    else if (mfma_ok(L))
      hipLaunchKernelGGL(conv_fwd_mfma_kernel, dim3(mfma_grid(ntiles)), dim3(kT), 0, s, act[l], pw + L.w_off,
                         p + L.b_off, G, L.dil, L.relu ? 1 : 0, ntiles, o_bf, o_f);
    else
      hipLaunchKernelGGL((conv_fwd_kernel<kWidth, kWidth>), grid, dim3(kT), 0, s, act[l], pw + L.w_off,
                         p + L.b_off, G, L.dil, L.relu ? 1 : 0, o_bf, o_f);
  }
}

// loss of out vs tgt into lpart (block sums), gradients of all parameters into grad
void Net::backward(int B, hipStream_t s) {
  const size_t GG = size_t(G) * G;
  const long long nout = (long long)B * GG * 2;
  const int nlb = int(nblk(nout));
  hipLaunchKernelGGL(mse_kernel, dim3(nlb), dim3(kT), 0, s, out, tgt, nout, g0, lpart);
  const int chunks = int((GG + kWChunk - 1) / kWChunk);
  const dim3 grid(nblk((long long)GG), B);
  const long long ntiles = (long long)B * (long long)GG / 16;
  float* gcur = g0;
  float* gnext = g1;
  for (int l = int(layers.size()) - 1; l >= 0; --l) {
    const Layer& L = layers[size_t(l)];
    if (L.relu)
      hipLaunchKernelGGL(relu_bwd_kernel, dim3(nblk((long long)B * GG * L.cout)), dim3(kT), 0, s, gcur, act[l + 1],
                         (long long)B * GG * L.cout);
    const int nw = L.cout * 9 * L.cin + L.cout;
    if (L.cin == 3)
      hipLaunchKernelGGL((conv_bwd_w_kernel<3, kWidth>), dim3(B * chunks), dim3(kT), 0, s, act[l], gcur, G, L.dil,
                         wpart);
    else if (L.cout == 2)
      hipLaunchKernelGGL((conv_bwd_w_kernel<kWidth, 2>), dim3(B * chunks), dim3(kT), 0, s, act[l], gcur, G, L.dil,
                         wpart);
    else if (mfma_ok(L))
      hipLaunchKernelGGL(conv_bwd_w_mfma_kernel, dim3(B * chunks), dim3(kT), 0, s, act[l], gcur, G, L.dil, wpart);
    else
      hipLaunchKernelGGL((conv_bwd_w_kernel<kWidth, kWidth>), dim3(B * chunks), dim3(kT), 0, s, act[l], gcur, G,
                         L.dil, wpart);
    hipLaunchKernelGGL(reduce_parts_kernel, dim3(nblk(nw)), dim3(kT), 0, s, wpart, B * chunks, nw, grad + L.w_off);
    if (l == 0) break;
    if (L.cout == 2)
      hipLaunchKernelGGL((conv_bwd_data_kernel<kWidth, 2>), grid, dim3(kT), 0, s, gcur, pw + L.w_off, G, L.dil,
                         gnext);
    else if (mfma_ok(L))
      hipLaunchKernelGGL(conv_bwd_data_mfma_kernel, dim3(mfma_grid(ntiles)), dim3(kT), 0, s, gcur, pw + L.w_off, G,
                         L.dil, ntiles, gnext);
    else
      hipLaunchKernelGGL((conv_bwd_data_kernel<kWidth, kWidth>), grid, dim3(kT), 0, s, gcur, pw + L.w_off, G, L.dil,
                         gnext);
    std::swap(gcur, gnext);
  }
}

std::string Net::loss_value(int B, float* loss) {
  const long long nout = (long long)B * G * G * 2;
  const int nlb = int(nblk(nout));
  std::vector<float> h(static_cast<size_t>(nlb), 0.0f);
  HIP_OK(hipMemcpyAsync(h.data(), lpart, h.size() * sizeof(float), hipMemcpyDeviceToHost, stream));
  HIP_OK(hipStreamSynchronize(stream));
  double sacc = 0.0;
  for (float x : h) sacc += x;
  *loss = float(sacc / double(nout));
  return "";
}

std::string Net::train_step(const int* idx, int B, float lr, float* loss) {
  if (!din) return "no training data set";
  if (B < 1 || B > maxB) return "batch out of range";
  for (int b = 0; b < B; ++b)
    if (idx[b] < 0 || idx[b] >= n_data) return "sample index out of range";
  HIP_OK(hipMemcpyAsync(didx, idx, sizeof(int) * B, hipMemcpyHostToDevice, stream));
  const int GG = G * G;
  hipLaunchKernelGGL(gather_batch_kernel, dim3(nblk(GG), B), dim3(kT), 0, stream, din, dtgt, didx, GG, act[0], tgt);
  forward(B, stream);
  backward(B, stream);
  ++adam_t;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, float(adam_t)), c2 = 1.0f - std::pow(b2, float(adam_t));
  hipLaunchKernelGGL(adam_kernel, dim3(nblk(nparams)), dim3(kT), 0, stream, p, grad, m, v, (long long)nparams, lr, b1,
                     b2, 1e-8f, c1, c2);
  hipLaunchKernelGGL(to_bf16_kernel, dim3(nblk(nparams)), dim3(kT), 0, stream, p, pw, (long long)nparams);
  return loss_value(B, loss);
}

std::string Net::upload_batch(const float* in, const float* target, int B) {
  if (B < 1 || B > maxB) return "batch out of range";
  const size_t GG = size_t(G) * G;
  std::vector<uint16_t> hb(size_t(B) * GG * 3);
  for (size_t i = 0; i < hb.size(); ++i) {
    uint32_t u;
    std::memcpy(&u, &in[i], 4);
    hb[i] = uint16_t((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
  }
  HIP_OK(hipMemcpyAsync(act[0], hb.data(), hb.size() * sizeof(uint16_t), hipMemcpyHostToDevice, stream));
  if (target) HIP_OK(hipMemcpyAsync(tgt, target, size_t(B) * GG * 2 * sizeof(float), hipMemcpyHostToDevice, stream));
  HIP_OK(hipStreamSynchronize(stream));
  return "";
}

std::string Net::set_data(const float* in, const float* target, int n) {
  if (n < 1) return "empty training set";
  if (din) (void)hipFree(din);
  if (dtgt) (void)hipFree(dtgt);
  din = nullptr;
  dtgt = nullptr;
  const size_t GG = size_t(G) * G;
  std::vector<uint16_t> hb(size_t(n) * GG * 3);
  for (size_t i = 0; i < hb.size(); ++i) {
    uint32_t u;
    std::memcpy(&u, &in[i], 4);
    hb[i] = uint16_t((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
  }
  HIP_OK(hipMalloc((void**)&din, hb.size() * sizeof(uint16_t)));
  HIP_OK(hipMalloc((void**)&dtgt, size_t(n) * GG * 2 * sizeof(float)));
  HIP_OK(hipMemcpy(din, hb.data(), hb.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(dtgt, target, size_t(n) * GG * 2 * sizeof(float), hipMemcpyHostToDevice));
  n_data = n;
  return "";
}

// ----------------------------------------------------------- grid maps -----
std::string GridMaps::build(const std::vector<double>& cx, const std::vector<double>& cy, double radius, int nc,
                            double L, int grid, const std::vector<double>& qk2_nodes, double kappa) {
  release();
  G = grid;
  const int n_leaves = int(cx.size());
  const int rank = nc * nc;
  N = n_leaves * rank;
  const int side = int(std::lround(L / radius));  // leaves per side (uniform tree), radius = half leaf side
  if (side < 1 || side * side != n_leaves) return "grid maps need a uniform tree";
  // leaf index of each (column, row) of leaves
  std::vector<int> leaf_at(size_t(side) * side, -1);
  for (int l = 0; l < n_leaves; ++l) {
    const int ix = int(std::floor((cx[size_t(l)] + L) / (2.0 * radius)));
    const int iy = int(std::floor((cy[size_t(l)] + L) / (2.0 * radius)));
    if (ix < 0 || iy < 0 || ix >= side || iy >= side) return "leaf outside the domain";
    leaf_at[size_t(iy) * side + ix] = l;
  }
  std::vector<double> tn(static_cast<size_t>(nc));
  for (int k = 0; k < nc; ++k) tn[size_t(k)] = -std::cos((k + 0.5) * M_PI / nc);
  auto lagrange = [&](double t, double* wgt) {
    for (int a = 0; a < nc; ++a) {
      double num = 1.0, den = 1.0;
      for (int b = 0; b < nc; ++b)
        if (b != a) {
          num *= t - tn[size_t(b)];
          den *= tn[size_t(a)] - tn[size_t(b)];
        }
      wgt[a] = num / den;
    }
  };
  const double h = 2.0 * L / G;
  // nodes -> grid: Chebyshev (Lagrange) interpolation inside the leaf holding the grid point
  std::vector<int> rp1(size_t(G) * G + 1, 0), c1;
  std::vector<double> v1;
  std::vector<double> wx(static_cast<size_t>(nc)), wy(static_cast<size_t>(nc));
  for (int gy = 0; gy < G; ++gy)
    for (int gx = 0; gx < G; ++gx) {
      const double x = -L + (gx + 0.5) * h, y = -L + (gy + 0.5) * h;
      const int ix = std::min(side - 1, int(std::floor((x + L) / (2.0 * radius))));
      const int iy = std::min(side - 1, int(std::floor((y + L) / (2.0 * radius))));
      const int l = leaf_at[size_t(iy) * side + ix];
      if (l < 0) return "missing leaf";
      lagrange((x - cx[size_t(l)]) / radius, wx.data());
      lagrange((y - cy[size_t(l)]) / radius, wy.data());
      for (int b = 0; b < nc; ++b)
        for (int a = 0; a < nc; ++a) {
          c1.push_back(l * rank + b * nc + a);
          v1.push_back(wx[size_t(a)] * wy[size_t(b)]);
        }
      rp1[size_t(gy) * G + gx + 1] = int(c1.size());
    }
  // grid -> nodes: bilinear interpolation of the cell-centred grid
  std::vector<int> rp2(size_t(N) + 1, 0), c2;
  std::vector<double> v2;
  for (int l = 0; l < n_leaves; ++l)
    for (int b = 0; b < nc; ++b)
      for (int a = 0; a < nc; ++a) {
        const double x = cx[size_t(l)] + radius * tn[size_t(a)], y = cy[size_t(l)] + radius * tn[size_t(b)];
        const double fx = (x + L) / h - 0.5, fy = (y + L) / h - 0.5;
        const int i0 = std::min(G - 2, std::max(0, int(std::floor(fx))));
        const int j0 = std::min(G - 2, std::max(0, int(std::floor(fy))));
        const double tx = fx - i0, ty = fy - j0;
        const int gi[4] = {j0 * G + i0, j0 * G + i0 + 1, (j0 + 1) * G + i0, (j0 + 1) * G + i0 + 1};
        const double gw[4] = {(1 - tx) * (1 - ty), tx * (1 - ty), (1 - tx) * ty, tx * ty};
        for (int k = 0; k < 4; ++k) {
          c2.push_back(gi[k]);
          v2.push_back(gw[k]);
        }
        rp2[size_t(l) * rank + b * nc + a + 1] = int(c2.size());
      }
  // q channel on the grid: (k^2 q / k^2) interpolated from the nodes
  std::vector<double> qg(size_t(G) * G, 0.0);
  for (int r = 0; r < G * G; ++r)
    for (int j = rp1[size_t(r)]; j < rp1[size_t(r) + 1]; ++j) qg[size_t(r)] += v1[size_t(j)] * qk2_nodes[size_t(c1[size_t(j)])];
  for (double& q : qg) q /= kappa * kappa;
  HIP_OK(hipMalloc((void**)&rp_ng, rp1.size() * sizeof(int)));
  HIP_OK(hipMalloc((void**)&col_ng, c1.size() * sizeof(int)));
  HIP_OK(hipMalloc((void**)&val_ng, v1.size() * sizeof(double)));
  HIP_OK(hipMalloc((void**)&rp_gn, rp2.size() * sizeof(int)));
  HIP_OK(hipMalloc((void**)&col_gn, c2.size() * sizeof(int)));
  HIP_OK(hipMalloc((void**)&val_gn, v2.size() * sizeof(double)));
  HIP_OK(hipMalloc((void**)&qgrid, qg.size() * sizeof(double)));
  HIP_OK(hipMalloc((void**)&gbuf, size_t(G) * G * sizeof(double2)));
  HIP_OK(hipMalloc((void**)&scal, 2 * sizeof(double)));
  HIP_OK(hipMemcpy(rp_ng, rp1.data(), rp1.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(col_ng, c1.data(), c1.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(val_ng, v1.data(), v1.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(rp_gn, rp2.data(), rp2.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(col_gn, c2.data(), c2.size() * sizeof(int), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(val_gn, v2.data(), v2.size() * sizeof(double), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(qgrid, qg.data(), qg.size() * sizeof(double), hipMemcpyHostToDevice));
  return "";
}

void GridMaps::release() {
  void* ptrs[] = {rp_ng, col_ng, val_ng, rp_gn, col_gn, val_gn, qgrid, gbuf, scal};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  rp_ng = col_ng = rp_gn = col_gn = nullptr;
  val_ng = val_gn = qgrid = scal = nullptr;
  gbuf = nullptr;
}

void GridMaps::to_grid(const double2* v, double2* g, hipStream_t s) const {
  hipLaunchKernelGGL(spmv_c_kernel, dim3(nblk((long long)G * G)), dim3(kT), 0, s, rp_ng, col_ng, val_ng, v, g, G * G,
                     (const double*)nullptr, (const double2*)nullptr);
}

void GridMaps::to_nodes(const double2* g, double2* v, hipStream_t s, const double* scale, const double2* base) const {
  hipLaunchKernelGGL(spmv_c_kernel, dim3(nblk(N)), dim3(kT), 0, s, rp_gn, col_gn, val_gn, g, v, N, scale, base);
}

void LearnedPrecond::apply(const double2* vin, double2* z, hipStream_t s) {
  const int GG = maps->G * maps->G;
  hipLaunchKernelGGL(norm2_kernel, dim3(1), dim3(kT), 0, s, vin, maps->N, maps->scal);
  maps->to_grid(vin, maps->gbuf, s);
  hipLaunchKernelGGL(prec_input_kernel, dim3(nblk(GG)), dim3(kT), 0, s, maps->qgrid, maps->gbuf, maps->scal, maps->N,
                     GG, net->act[0], maps->scal + 1);
  net->forward(1, s);
  hipLaunchKernelGGL(out_to_complex_kernel, dim3(nblk(GG)), dim3(kT), 0, s, net->out, maps->gbuf, GG);
  maps->to_nodes(maps->gbuf, z, s, maps->scal + 1, vin);   // z = v + sigma * correction
}

}  // namespace lsg

// --------------------------------------------------------------- C ABI ------
extern "C" {

int lsgn_create(int n_layers, int width, const int* dilations, int grid, int max_batch, uint64_t seed,
                lsgn_net** out) {
  if (!out) return LSG_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return LSG_ERR_DEVICE;
  lsgn_net* n = new lsgn_net();
  std::string e = n->net.create(n_layers, width, dilations, grid, max_batch, seed);
  if (!e.empty()) {
    const bool arg = e.find("must be") != std::string::npos || e.find("out of range") != std::string::npos;
    n->net.release();
    delete n;
    return arg ? LSG_ERR_ARG : LSG_ERR_DEVICE;
  }
  *out = n;
  return LSG_OK;
}

void lsgn_destroy(lsgn_net* net) {
  if (!net) return;
  net->net.release();
  delete net;
}

const char* lsgn_last_error(const lsgn_net* net) { return net ? net->err.c_str() : "net is NULL"; }

int64_t lsgn_num_params(const lsgn_net* net) { return net ? net->net.nparams : LSG_ERR_ARG; }

int lsgn_get_params(const lsgn_net* net, float* out) {
  if (!net || !out) return LSG_ERR_ARG;
  if (hipMemcpy(out, net->net.p, sizeof(float) * size_t(net->net.nparams), hipMemcpyDeviceToHost) != hipSuccess)
    return LSG_ERR_DEVICE;
  return LSG_OK;
}

int lsgn_set_params(lsgn_net* net, const float* in) {
  if (!net || !in) return LSG_ERR_ARG;
  if (hipMemcpy(net->net.p, in, sizeof(float) * size_t(net->net.nparams), hipMemcpyHostToDevice) != hipSuccess)
    return net->set(LSG_ERR_DEVICE, "copy failed");
  std::string e = net->net.sync_weights();
  return e.empty() ? LSG_OK : net->set(LSG_ERR_DEVICE, e);
}

int lsgn_forward(lsgn_net* net, const float* in, int batch, float* out) {
  if (!net || !in || !out) return LSG_ERR_ARG;
  std::string e = net->net.upload_batch(in, nullptr, batch);
  if (!e.empty()) return net->set(LSG_ERR_ARG, e);
  net->net.forward(batch, net->net.stream);
  const size_t n = size_t(batch) * net->net.G * net->net.G * 2;
  (void)hipMemcpyAsync(out, net->net.out, n * sizeof(float), hipMemcpyDeviceToHost, net->net.stream);
  if (hipStreamSynchronize(net->net.stream) != hipSuccess) return net->set(LSG_ERR_DEVICE, "forward failed");
  return LSG_OK;
}

int lsgn_loss_grad(lsgn_net* net, const float* in, const float* target, int batch, float* loss, float* grad) {
  if (!net || !in || !target || !loss || !grad) return LSG_ERR_ARG;
  std::string e = net->net.upload_batch(in, target, batch);
  if (!e.empty()) return net->set(LSG_ERR_ARG, e);
  net->net.forward(batch, net->net.stream);
  net->net.backward(batch, net->net.stream);
  e = net->net.loss_value(batch, loss);
  if (!e.empty()) return net->set(LSG_ERR_DEVICE, e);
  if (hipMemcpy(grad, net->net.grad, sizeof(float) * size_t(net->net.nparams), hipMemcpyDeviceToHost) != hipSuccess)
    return net->set(LSG_ERR_DEVICE, "copy failed");
  return LSG_OK;
}

int lsgn_set_training_data(lsgn_net* net, const float* in, const float* target, int n) {
  if (!net || !in || !target || n < 1) return LSG_ERR_ARG;
  std::string e = net->net.set_data(in, target, n);
  return e.empty() ? LSG_OK : net->set(LSG_ERR_DEVICE, e);
}

int lsgn_train_step(lsgn_net* net, const int* idx, int batch, float lr, float* loss) {
  if (!net || !idx || !loss) return LSG_ERR_ARG;
  std::string e = net->net.train_step(idx, batch, lr, loss);
  if (!e.empty()) {
    const bool arg = e.find("range") != std::string::npos || e.find("no training") != std::string::npos;
    return net->set(arg ? LSG_ERR_ARG : LSG_ERR_DEVICE, e);
  }
  return LSG_OK;
}

}  // extern "C"
