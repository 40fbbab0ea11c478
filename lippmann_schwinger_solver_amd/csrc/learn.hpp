// Learned preconditioner: network state on the device, node <-> grid maps,
// and the flexible-GMRES preconditioner built from them.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include "device.hpp"

namespace lsg {

struct Net {
  struct Layer {
    int cin = 0, cout = 0, dil = 1;
    bool relu = true;
    int64_t w_off = 0, b_off = 0;  // offsets in the parameter vector
  };
  std::vector<Layer> layers;
  int G = 0, maxB = 0;
  int64_t nparams = 0;
  hipStream_t stream = nullptr;
  float* p = nullptr;       // FP32 master parameters
  uint16_t* pw = nullptr;   // BF16 copy used by the kernels (weights; biases are read from p)
  float *grad = nullptr, *m = nullptr, *v = nullptr;  // gradient, Adam moments
  int adam_t = 0;
  std::vector<uint16_t*> act;   // act[0] input (3 ch), act[l+1] output of layer l (BF16, NHWC)
  float* out = nullptr;         // last layer output, FP32 (B G G 2)
  float* tgt = nullptr;         // targets of the current batch
  float *g0 = nullptr, *g1 = nullptr;  // activation gradients (FP32)
  float* wpart = nullptr;       // weight-gradient partials
  float* lpart = nullptr;       // loss partials
  int* didx = nullptr;
  uint16_t* din = nullptr;      // resident training inputs (BF16)
  float* dtgt = nullptr;        // resident training targets (FP32)
  int n_data = 0;

  std::string create(int n_layers, int width, const int* dil, int grid, int max_batch, uint64_t seed);
  void release();
  std::string sync_weights();
  void forward(int B, hipStream_t s);     // act[0] -> out
  void backward(int B, hipStream_t s);    // MSE(out, tgt) -> grad
  std::string loss_value(int B, float* loss);
  std::string upload_batch(const float* in, const float* target, int B);
  std::string set_data(const float* in, const float* target, int n);
  std::string train_step(const int* idx, int B, float lr, float* loss);
};

// Linear maps between the N leaf nodes (leaf-major Chebyshev nodes of a
// uniform tree) and a G x G cell-centred grid on [-L, L]^2.
struct GridMaps {
  int G = 0, N = 0;
  int *rp_ng = nullptr, *col_ng = nullptr;  // nodes -> grid (CSR, Lagrange weights)
  double* val_ng = nullptr;
  int *rp_gn = nullptr, *col_gn = nullptr;  // grid -> nodes (CSR, bilinear weights)
  double* val_gn = nullptr;
  double* qgrid = nullptr;    // q = k^2 q / k^2 on the grid
  double2* gbuf = nullptr;    // G G complex scratch
  double* scal = nullptr;     // [||v||^2, sigma]
  std::string build(const std::vector<double>& cx, const std::vector<double>& cy, double radius, int nc, double L,
                    int grid, const std::vector<double>& qk2_nodes, double kappa);
  void release();
  void to_grid(const double2* v, double2* g, hipStream_t s) const;
  void to_nodes(const double2* g, double2* v, hipStream_t s, const double* scale = nullptr,
                const double2* base = nullptr) const;
};

// z = v + sigma * P_nodes(net(q, P_grid(v) / sigma)), sigma = RMS of v: the
// network predicts the correction A^-1 v - v of a unit-RMS right-hand side
struct LearnedPrecond : Precond {
  Net* net = nullptr;
  GridMaps* maps = nullptr;
  void apply(const double2* v, double2* z, hipStream_t s) override;
};

}  // namespace lsg

// handle behind the C ABI of include/lsg_learn.h
struct lsgn_net {
  lsg::Net net;
  std::string err;
  int set(int code, const std::string& m) {
    err = m;
    return code;
  }
};
