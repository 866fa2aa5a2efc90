// Learned convolution functions (see learned_cf.cpp and kernels_mlp.hip).
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "ygrid_internal.h"

namespace ygrid {

class Gridder;

constexpr int kLcfIn = 5;  // network inputs: w, fracu, fracv, dx, dy
enum LcfEpilogue { kEpiNone = 0, kEpiBias = 1, kEpiBiasSilu = 2, kEpiSiluGrad = 3 };

// kernel launchers (kernels_mlp.hip)
void mlp_gemm(const Ctx& c, bool ta, bool tb, const float* A, const void* B, bool b_half, float* C, int M, int N, int K,
              int lda, int ldb, int ldc, int epi, const float* bias, float* Zout, const float* Zin);
void mlp_colsum(const Ctx& c, const float* D, int M, int N, int ldd, float* db);
void mlp_mse(const Ctx& c, const float* Y, const float* T, int M, float* dY, double* loss);
void mlp_adam(const Ctx& c, float* p, const float* g, float* m, float* v, void* p16, int n, float lr, float b1, float b2,
              float eps, int step);
void mlp_to_half(const Ctx& c, const float* p, void* p16, int n);
void mlp_gather(const Ctx& c, const float* X, const float* T, const int* idx, int M, float* Xb, float* Tb);
void lcf_inputs(const Ctx& c, const float4* sin, int s0, int ns, int S, float inv_os, float inv_wnorm, float* X);
void lcf_scatter(const Ctx& c, const float4* sin, const float* Y, int s0, int ns, int S, int P, long long plane_elems,
                 float inv_tscale, void* table, int prec);

struct LcfTrainOptions {
  int epochs = 20;
  int batch = 4096;
  float lr = 2e-3f;
  uint64_t seed = 1;
  int holdout_every = 8;  // every n-th w plane is held out of training and judges the error
};

struct LcfReport {
  int epochs = 0, steps = 0, examples = 0, cfs = 0, holdout_cfs = 0;
  double first_loss = 0, last_loss = 0;
  double holdout_max_err = 0, holdout_rms_err = 0;  // relative to the table's peak |CF|
  double train_s = 0;
};

// A coordinate network x = (w / wnorm, fracu / os, fracv / os, dx / S, dy / S) -> CF tap * tscale,
// dense layers [5, H x layers, 2], SiLU between them.  Master weights FP32 (training, Adam),
// an FP16 copy for inference (products accumulated in FP32).
class LearnedCF {
 public:
  LearnedCF(std::shared_ptr<Ctx> ctx, int hidden, int layers, uint64_t seed);
  int nparams() const { return nparams_; }
  const std::vector<int>& dims() const { return dims_; }
  void get_params(float* host) const;
  void set_params(const float* host);
  // loss and gradients (host arrays) of one forward + backward pass over n examples (host arrays)
  double loss_and_grads_host(const float* x, const float* t, int n, float* grad);
  // inference over n rows (host arrays), FP16 or FP32 weights
  void forward_host(const float* x, float* y, int n, bool half);
  LcfReport train_wproject(const Gridder& wp, const LcfTrainOptions& o);
  // per-sample CF planes from the samples' (w, fracu, fracv, valid) (device), in stream order
  void reserve_eval(int S);
  void eval_samples(const float4* sin, int nsamp, void* table, int P, long long plane_elems, int prec);

  int support = 0, oversample = 1;
  double wnorm = 1.0;
  float tscale = 1.f;
  bool trained = false;
  double holdout_max_err = 1e30;

 private:
  void forward_dev(const float* X, int M, bool half, bool keep, float* Y);  // keep: Z and X per layer
  double backward_dev(const float* T, int M);                               // after forward_dev(keep)
  void ensure_train_buffers(int M);
  const float* W(int l) const { return p_.as<float>() + woff_[l]; }
  const float* B(int l) const { return p_.as<float>() + boff_[l]; }

  std::shared_ptr<Ctx> ctx_;
  int H_, L_, nparams_ = 0;
  std::vector<int> dims_;
  std::vector<size_t> woff_, boff_;
  DevBuf p_, p16_, g_, m_, v_;
  DevBuf xs_, zs_, dz_, y_, dy_, loss_;  // training activations (per layer, batch rows)
  int train_rows_ = 0;
  DevBuf ex_, ey_, ea_, eb_;  // inference buffers
  int eval_rows_ = 0;
  int adam_step_ = 0;
};

}  // namespace ygrid
