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
// fused FP16 matrix-core inference (width 128, every layer's weights in LDS): whether a network
// of this size has it, the bytes of its packed weights, the packing (from the FP16 and FP32
// parameters), and inference over n input rows or over the taps of ns samples (outputs into their
// CF planes, as lcf_inputs + forward + lcf_scatter)
bool lcf_fused_supported(int hidden, int layers);
size_t lcf_fused_pack_bytes(int layers);
void lcf_fused_pack(const Ctx& c, int layers, const long long* woff, const long long* boff, const void* p16,
                    const float* p32, void* img);
void lcf_fused_rows(const Ctx& c, const void* img, int layers, const float* x, float* y, long long n);
// same (device flag, may be null): 0 means the samples equal those whose planes are at src (null:
// already in table): the kernel copies them instead of evaluating
void lcf_fused_samples(const Ctx& c, const void* img, int layers, const float4* sin, int ns, int S, float inv_os,
                       float inv_wnorm, int P, long long plane_elems, float inv_tscale, void* table, int prec,
                       const int* same, const void* src);
// flag := 1 if !compare, else whether a[0..n) and b[0..n) differ bitwise; keep: b := a
void lcf_differ(const Ctx& c, const float4* a, const float4* b, int n, int* flag, bool compare);
void lcf_keep(const Ctx& c, const float4* a, float4* b, int n);

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
  // inference over n rows (host arrays): mode 0 FP32, 1 FP16 weights (FP32 activations), 2 the
  // fused FP16 matrix-core form (FP16 weights and activations, FP32 sums; what gridding uses
  // whenever the network has it)
  void forward_host(const float* x, float* y, int n, int mode);
  bool fused() const;
  void clear_cache() {
    last_planes_.reset();
    last_key_ = EvalKey{};
  }
  LcfReport train_wproject(const Gridder& wp, const LcfTrainOptions& o);
  // per-sample CF planes from the samples' (w, fracu, fracv, valid) (device), in stream order
  void reserve_eval(int S);
  // table: the caller's planes buffer.  The network's last evaluation is remembered (its inputs
  // and buffer): a pass over the same samples copies those planes (or, in the same buffer, keeps
  // them) instead of evaluating again -- decided on the device, so this stays kernels only.
  void eval_samples(const float4* sin, int nsamp, const std::shared_ptr<DevBuf>& table, int P, long long plane_elems,
                    int prec);

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
  const void* packed();          // the fused kernel's weight image (current weights)
  // the last evaluation: its planes buffer, inputs and parameters
  std::shared_ptr<DevBuf> last_planes_;
  DevBuf last_in_, same_;
  struct EvalKey {
    int nsamp = -1, S = -1, P = -1, prec = -1;
    long long plane_elems = -1, version = -1;
    bool operator==(const EvalKey& o) const {
      return nsamp == o.nsamp && S == o.S && P == o.P && prec == o.prec && plane_elems == o.plane_elems &&
             version == o.version;
    }
  } last_key_;
  DevBuf fused_img_;
  long long version_ = 0, packed_version_ = -1;
};

}  // namespace ygrid
