// Learned convolution functions: training and inference drivers of the coordinate network
// (kernels in kernels_mlp.hip).
//
// Training data are the W-projection gridder's own CF table (WProjectVisGridder::generate,
// WProjectVisGridder.cc:435-522: one padded plane per (w plane, fracv, fracu)): every tap of
// every plane is an example (w of the plane, fracu, fracv, dx, dy) -> tap, taps outside a plane's
// support are zero.  Every holdout_every-th w plane is left out and judges the error afterwards
// (the table is the judge); a gridder uses the network only while that error is within the
// caller's bound, and the table otherwise.
#include "learned_cf.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <random>

#include "gridders.h"

namespace ygrid {

LearnedCF::LearnedCF(std::shared_ptr<Ctx> ctx, int hidden, int layers, uint64_t seed)
    : ctx_(std::move(ctx)), H_(hidden), L_(layers) {
  YG_CHECK(hidden >= 4 && hidden <= 1024 && layers >= 1 && layers <= 16, 1, "network size out of range");
  dims_.push_back(kLcfIn);
  for (int l = 0; l < L_; ++l) dims_.push_back(H_);
  dims_.push_back(2);
  size_t o = 0;
  for (size_t l = 0; l + 1 < dims_.size(); ++l) {
    woff_.push_back(o);
    o += (size_t)dims_[l] * dims_[l + 1];
    boff_.push_back(o);
    o += dims_[l + 1];
  }
  nparams_ = (int)o;
  // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases
  std::vector<float> h(nparams_);
  std::mt19937_64 rng(seed);
  for (size_t l = 0; l + 1 < dims_.size(); ++l) {
    const float a = 1.f / std::sqrt((float)dims_[l]);
    std::uniform_real_distribution<float> U(-a, a);
    for (size_t i = woff_[l]; i < boff_[l] + dims_[l + 1]; ++i) h[i] = U(rng);
  }
  p_.resize(nparams_ * sizeof(float));
  p16_.resize(nparams_ * 2);
  g_.resize(nparams_ * sizeof(float));
  m_.resize(nparams_ * sizeof(float));
  v_.resize(nparams_ * sizeof(float));
  set_params(h.data());
}

void LearnedCF::get_params(float* host) const {
  YG_HIP(hipStreamSynchronize(ctx_->stream));
  YG_HIP(hipMemcpy(host, p_.get(), nparams_ * sizeof(float), hipMemcpyDeviceToHost));
}

void LearnedCF::set_params(const float* host) {
  YG_HIP(hipStreamSynchronize(ctx_->stream));
  YG_HIP(hipMemcpy(p_.get(), host, nparams_ * sizeof(float), hipMemcpyHostToDevice));
  mlp_to_half(*ctx_, p_.as<float>(), p16_.get(), nparams_);
  YG_HIP(hipMemsetAsync(m_.get(), 0, nparams_ * sizeof(float), ctx_->stream));
  YG_HIP(hipMemsetAsync(v_.get(), 0, nparams_ * sizeof(float), ctx_->stream));
  adam_step_ = 0;
  ++version_;
  YG_HIP(hipStreamSynchronize(ctx_->stream));
}

bool LearnedCF::fused() const { return lcf_fused_supported(H_, L_); }

// The weights in the layout of the fused kernel's LDS image, rebuilt after the weights change.
const void* LearnedCF::packed() {
  if (packed_version_ != version_) {
    fused_img_.resize(lcf_fused_pack_bytes(L_));
    std::vector<long long> wo(woff_.begin(), woff_.end()), bo(boff_.begin(), boff_.end());
    lcf_fused_pack(*ctx_, L_, wo.data(), bo.data(), p16_.get(), p_.as<float>(), fused_img_.get());
    packed_version_ = version_;
  }
  return fused_img_.get();
}

void LearnedCF::ensure_train_buffers(int M) {
  if (M <= train_rows_) return;
  size_t xs = 0, zs = 0;
  for (int l = 0; l <= L_; ++l) xs += (size_t)M * dims_[l];  // X_0 .. X_L
  for (int l = 1; l <= L_; ++l) zs += (size_t)M * dims_[l];  // Z_1 .. Z_L
  xs_.resize(xs * sizeof(float));
  zs_.resize(zs * sizeof(float));
  dz_.resize((size_t)2 * M * H_ * sizeof(float));
  y_.resize((size_t)M * 2 * sizeof(float));
  dy_.resize((size_t)M * 2 * sizeof(float));
  loss_.resize(sizeof(double));
  train_rows_ = M;
}

// X (M x 5) in xs_ slot 0 when keep; layer activations X_l / pre-activations Z_l kept for the
// backward pass, or two ping-pong buffers for inference.
void LearnedCF::forward_dev(const float* X, int M, bool half, bool keep, float* Y) {
  const Ctx& c = *ctx_;
  auto wp = [&](int l) -> const void* {
    return half ? (const void*)(p16_.as<uint16_t>() + woff_[l]) : (const void*)(p_.as<float>() + woff_[l]);
  };
  const float* in = X;
  size_t xo = (size_t)M * dims_[0], zo = 0;
  for (int l = 0; l < L_; ++l) {
    float *out, *z = nullptr;
    if (keep) {
      out = xs_.as<float>() + xo;
      z = zs_.as<float>() + zo;
      xo += (size_t)M * dims_[l + 1];
      zo += (size_t)M * dims_[l + 1];
    } else {
      out = (l % 2 == 0) ? ea_.as<float>() : eb_.as<float>();
    }
    mlp_gemm(c, false, false, in, wp(l), half, out, M, dims_[l + 1], dims_[l], dims_[l], dims_[l + 1], dims_[l + 1],
             kEpiBiasSilu, B(l), z, nullptr);
    in = out;
  }
  mlp_gemm(c, false, false, in, wp(L_), half, Y, M, 2, dims_[L_], dims_[L_], 2, 2, kEpiBias, B(L_), nullptr, nullptr);
}

// gradients into g_ (and the loss) of the examples whose forward pass forward_dev(keep) ran
double LearnedCF::backward_dev(const float* T, int M) {
  const Ctx& c = *ctx_;
  float* g = g_.as<float>();
  YG_HIP(hipMemsetAsync(loss_.get(), 0, sizeof(double), c.stream));
  mlp_mse(c, y_.as<float>(), T, M, dy_.as<float>(), loss_.as<double>());
  // X_l and Z_l offsets
  std::vector<size_t> xoff(L_ + 1), zoff(L_ + 1);
  size_t xo = 0, zo = 0;
  for (int l = 0; l <= L_; ++l) {
    xoff[l] = xo;
    xo += (size_t)M * dims_[l];
  }
  for (int l = 1; l <= L_; ++l) {
    zoff[l] = zo;
    zo += (size_t)M * dims_[l];
  }
  const float* X = xs_.as<float>();
  const float* Z = zs_.as<float>();
  float* dzbuf[2] = {dz_.as<float>(), dz_.as<float>() + (size_t)M * H_};
  // output layer L_: dZ = dY
  const float* dZ = dy_.as<float>();
  for (int l = L_; l >= 0; --l) {
    const int din = dims_[l], dout = dims_[l + 1];
    // dW_l = X_l^T dZ  (din x dout), db_l = column sums of dZ
    mlp_gemm(c, true, false, X + xoff[l], dZ, false, g + woff_[l], din, dout, M, din, dout, dout, kEpiNone, nullptr,
             nullptr, nullptr);
    mlp_colsum(c, dZ, M, dout, dout, g + boff_[l]);
    if (l == 0) break;
    // dZ_{l} (layer l's input pre-activation) = (dZ W_l^T) * silu'(Z_l)
    float* next = dzbuf[l % 2];
    mlp_gemm(c, false, true, dZ, W(l), false, next, M, din, dout, dout, dout, din, kEpiSiluGrad, nullptr, nullptr,
             Z + zoff[l]);
    dZ = next;
  }
  double loss = 0;
  YG_HIP(hipMemcpyAsync(&loss, loss_.get(), sizeof(double), hipMemcpyDeviceToHost, c.stream));
  YG_HIP(hipStreamSynchronize(c.stream));
  return loss;
}

double LearnedCF::loss_and_grads_host(const float* x, const float* t, int n, float* grad) {
  YG_CHECK(n > 0, 1, "no examples");
  ensure_train_buffers(n);
  DevBuf T((size_t)n * 2 * sizeof(float));
  YG_HIP(hipStreamSynchronize(ctx_->stream));  // xs_ may still be read by earlier work on the stream
  YG_HIP(hipMemcpy(xs_.get(), x, (size_t)n * kLcfIn * sizeof(float), hipMemcpyHostToDevice));
  YG_HIP(hipMemcpy(T.get(), t, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice));
  forward_dev(xs_.as<float>(), n, false, true, y_.as<float>());
  const double loss = backward_dev(T.as<float>(), n);
  if (grad) YG_HIP(hipMemcpy(grad, g_.get(), nparams_ * sizeof(float), hipMemcpyDeviceToHost));
  return loss;
}

void LearnedCF::forward_host(const float* x, float* y, int n, int mode) {
  YG_CHECK(n > 0, 1, "no rows");
  YG_CHECK(mode >= 0 && mode <= 2, 1, "inference mode is 0 (FP32), 1 (FP16 weights) or 2 (FP16 matrix cores)");
  YG_CHECK(mode != 2 || fused(), 1, "the FP16 matrix-core inference needs hidden width 128 and at most 5 layers");
  DevBuf X((size_t)n * kLcfIn * sizeof(float)), Y((size_t)n * 2 * sizeof(float));
  YG_HIP(hipMemcpy(X.get(), x, (size_t)n * kLcfIn * sizeof(float), hipMemcpyHostToDevice));
  if (mode == 2) {
    lcf_fused_rows(*ctx_, packed(), L_, X.as<float>(), Y.as<float>(), n);
  } else {
    ea_.resize((size_t)n * H_ * sizeof(float));
    eb_.resize((size_t)n * H_ * sizeof(float));
    forward_dev(X.as<float>(), n, mode == 1, false, Y.as<float>());
  }
  YG_HIP(hipStreamSynchronize(ctx_->stream));
  YG_HIP(hipMemcpy(y, Y.get(), (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost));
}

LcfReport LearnedCF::train_wproject(const Gridder& wp, const LcfTrainOptions& o) {
  const auto t0 = std::chrono::steady_clock::now();
  const CFTable* cf = wp.cfTable();
  std::vector<double> wterm;
  double wmax = 0;
  YG_CHECK(wp.wTerms(wterm, wmax), 1, "learned CFs are trained on a W-projection gridder's table");
  YG_CHECK(cf && cf->n_elems > 0, 2, "the gridder has no CF table yet: run a pass first");
  const int os = cf->oversample, nz = cf->nz;
  YG_CHECK(nz == (int)wterm.size(), 2, "CF table and w sampling disagree");
  int S = 0;
  for (int s : cf->supp) S = std::max(S, s);
  const int W = 2 * S + 1;
  std::vector<float2> table(cf->n_elems);
  YG_HIP(hipStreamSynchronize(ctx_->stream));
  YG_HIP(hipMemcpy(table.data(), cf->table.get(), cf->n_elems * sizeof(float2), hipMemcpyDeviceToHost));
  float peak = 0.f;
  for (const float2& v : table) peak = std::max(peak, std::hypot(v.x, v.y));
  YG_CHECK(peak > 0.f, 2, "empty CF table");
  support = S;
  oversample = os;
  wnorm = wmax > 0 ? wmax : 1.0;
  tscale = 1.f / peak;
  const int hold = std::max(2, o.holdout_every);
  auto held = [&](int z) { return nz > 2 && z % hold == hold / 2; };
  // examples of one plane (z, fracu, fracv): W*W taps
  auto plane_examples = [&](int z, int fu, int fv, float* x, float* t) {
    const int k = fu + os * (fv + os * z);
    const int Sk = cf->supp[k], Pk = 2 * Sk + 1 + 2 * kGuardU;
    for (int iy = 0; iy < W; ++iy)
      for (int ix = 0; ix < W; ++ix) {
        const int dx = ix - S, dy = iy - S;
        float* xr = x + (size_t)(iy * W + ix) * kLcfIn;
        xr[0] = (float)(wterm[z] / wnorm);
        xr[1] = (float)fu / os;
        xr[2] = (float)fv / os;
        xr[3] = S > 0 ? (float)dx / S : 0.f;
        xr[4] = S > 0 ? (float)dy / S : 0.f;
        float2 v = make_float2(0.f, 0.f);
        if (std::abs(dx) <= Sk && std::abs(dy) <= Sk)
          v = table[cf->off[k] + (size_t)(dy + Sk + kGuardV) * Pk + (dx + Sk + kGuardU)];
        t[2 * (iy * W + ix)] = v.x * tscale;
        t[2 * (iy * W + ix) + 1] = v.y * tscale;
      }
  };
  std::vector<float> X, T, HX, HT;
  int ntrain_cf = 0, nhold_cf = 0;
  for (int z = 0; z < nz; ++z)
    for (int fv = 0; fv < os; ++fv)
      for (int fu = 0; fu < os; ++fu) {
        std::vector<float>& xv = held(z) ? HX : X;
        std::vector<float>& tv = held(z) ? HT : T;
        const size_t b = xv.size() / kLcfIn;
        xv.resize((b + (size_t)W * W) * kLcfIn);
        tv.resize((b + (size_t)W * W) * 2);
        plane_examples(z, fu, fv, xv.data() + b * kLcfIn, tv.data() + b * 2);
        (held(z) ? nhold_cf : ntrain_cf) += 1;
      }
  const int N = (int)(X.size() / kLcfIn);
  YG_CHECK(N > 0, 2, "no training examples");
  DevBuf dX(X.size() * sizeof(float)), dT(T.size() * sizeof(float)), didx((size_t)N * sizeof(int));
  YG_HIP(hipMemcpy(dX.get(), X.data(), X.size() * sizeof(float), hipMemcpyHostToDevice));
  YG_HIP(hipMemcpy(dT.get(), T.data(), T.size() * sizeof(float), hipMemcpyHostToDevice));
  const int M = std::min(o.batch, N);
  ensure_train_buffers(M);
  DevBuf Tb((size_t)M * 2 * sizeof(float));
  std::vector<int> idx(N);
  std::iota(idx.begin(), idx.end(), 0);
  std::mt19937_64 rng(o.seed);
  LcfReport rep;
  rep.examples = N;
  rep.cfs = ntrain_cf;
  rep.holdout_cfs = nhold_cf;
  const int spe = N / M;  // steps per epoch (the remainder of each shuffle is skipped)
  const long long total = (long long)o.epochs * spe;
  long long step = 0;
  for (int e = 0; e < o.epochs; ++e) {
    std::shuffle(idx.begin(), idx.end(), rng);
    YG_HIP(hipMemcpy(didx.get(), idx.data(), (size_t)N * sizeof(int), hipMemcpyHostToDevice));
    double eloss = 0;
    for (int sI = 0; sI < spe; ++sI, ++step) {
      mlp_gather(*ctx_, dX.as<float>(), dT.as<float>(), didx.as<int>() + (size_t)sI * M, M, xs_.as<float>(),
                 Tb.as<float>());
      forward_dev(xs_.as<float>(), M, false, true, y_.as<float>());
      // loss only read back at the end of an epoch: accumulate on the device
      const Ctx& c = *ctx_;
      if (sI == 0) YG_HIP(hipMemsetAsync(loss_.get(), 0, sizeof(double), c.stream));
      mlp_mse(c, y_.as<float>(), Tb.as<float>(), M, dy_.as<float>(), loss_.as<double>());
      // backward (as backward_dev, without its loss reset and readback)
      std::vector<size_t> xoff(L_ + 1), zoff(L_ + 1);
      size_t xo = 0, zo = 0;
      for (int l = 0; l <= L_; ++l) { xoff[l] = xo; xo += (size_t)M * dims_[l]; }
      for (int l = 1; l <= L_; ++l) { zoff[l] = zo; zo += (size_t)M * dims_[l]; }
      float* g = g_.as<float>();
      float* dzbuf[2] = {dz_.as<float>(), dz_.as<float>() + (size_t)M * H_};
      const float* dZ = dy_.as<float>();
      for (int l = L_; l >= 0; --l) {
        const int din = dims_[l], dout = dims_[l + 1];
        mlp_gemm(c, true, false, xs_.as<float>() + xoff[l], dZ, false, g + woff_[l], din, dout, M, din, dout, dout,
                 kEpiNone, nullptr, nullptr, nullptr);
        mlp_colsum(c, dZ, M, dout, dout, g + boff_[l]);
        if (l == 0) break;
        float* next = dzbuf[l % 2];
        mlp_gemm(c, false, true, dZ, p_.as<float>() + woff_[l], false, next, M, din, dout, dout, dout, din,
                 kEpiSiluGrad, nullptr, nullptr, zs_.as<float>() + zoff[l]);
        dZ = next;
      }
      // cosine decay of the learning rate to lr/20
      const double fr = total > 1 ? (double)step / (double)(total - 1) : 1.0;
      const float lr = (float)(o.lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(M_PI * fr))));
      mlp_adam(c, p_.as<float>(), g, m_.as<float>(), v_.as<float>(), p16_.get(), nparams_, lr, 0.9f, 0.999f, 1e-8f,
               ++adam_step_);
    }
    YG_HIP(hipMemcpyAsync(&eloss, loss_.get(), sizeof(double), hipMemcpyDeviceToHost, ctx_->stream));
    YG_HIP(hipStreamSynchronize(ctx_->stream));
    eloss /= std::max(1, spe);
    if (e == 0) rep.first_loss = eloss;
    rep.last_loss = eloss;
  }
  rep.epochs = o.epochs;
  rep.steps = (int)step;
  // the table judges the held-out planes (FP16 weights, as used for gridding)
  const int NH = (int)(HX.size() / kLcfIn);
  if (NH > 0) {
    std::vector<float> y((size_t)NH * 2);
    ++version_;  // new weights: the packed copy is stale
    forward_host(HX.data(), y.data(), NH, fused() ? 2 : 1);
    double mx = 0, ss = 0;
    for (int i = 0; i < NH; ++i) {
      // targets are the taps divided by the table's peak: errors are in units of the peak
      const double er = y[2 * i] - HT[2 * i], ei = y[2 * i + 1] - HT[2 * i + 1];
      const double e2 = er * er + ei * ei;
      mx = std::max(mx, std::sqrt(e2));
      ss += e2;
    }
    rep.holdout_max_err = mx;
    rep.holdout_rms_err = std::sqrt(ss / NH);
  } else {
    rep.holdout_max_err = rep.holdout_rms_err = 1e30;
  }
  holdout_max_err = rep.holdout_max_err;
  trained = true;
  rep.train_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rep;
}

void LearnedCF::reserve_eval(int S) {
  const int W2 = (2 * S + 1) * (2 * S + 1);
  const int ns = std::max(1, 65536 / W2);
  const int rows = ns * W2;
  if (rows <= eval_rows_) return;
  ex_.resize((size_t)rows * kLcfIn * sizeof(float));
  ey_.resize((size_t)rows * 2 * sizeof(float));
  ea_.resize((size_t)rows * H_ * sizeof(float));
  eb_.resize((size_t)rows * H_ * sizeof(float));
  eval_rows_ = rows;
}

// Kernels only (capturable): chunks of samples through the network (FP16 weights), outputs
// into the samples' padded planes.
void LearnedCF::eval_samples(const float4* sin, int nsamp, const std::shared_ptr<DevBuf>& tbuf, int P,
                             long long plane_elems, int prec) {
  const int S = support;
  void* table = tbuf->get();
  if (fused()) {
    const Ctx& c = *ctx_;
    EvalKey key;
    key.nsamp = nsamp;
    key.S = S;
    key.P = P;
    key.prec = prec;
    key.plane_elems = plane_elems;
    key.version = version_;
    const bool comparable = last_planes_ && key == last_key_;
    same_.resize(sizeof(int));
    last_in_.resize((size_t)std::max(nsamp, 1) * sizeof(float4));
    lcf_differ(c, sin, last_in_.as<float4>(), nsamp, same_.as<int>(), comparable);
    const void* src = comparable && last_planes_ != tbuf ? last_planes_->get() : nullptr;
    lcf_fused_samples(c, packed(), L_, sin, nsamp, S, 1.f / oversample, (float)(1.0 / wnorm), P, plane_elems,
                      1.f / tscale, table, prec, same_.as<int>(), src);
    lcf_keep(c, sin, last_in_.as<float4>(), nsamp);
    last_planes_ = tbuf;
    last_key_ = key;
    return;
  }
  const int W2 = (2 * S + 1) * (2 * S + 1);
  reserve_eval(S);
  const int ns = std::max(1, eval_rows_ / W2);
  for (int s0 = 0; s0 < nsamp; s0 += ns) {
    const int n = std::min(ns, nsamp - s0);
    lcf_inputs(*ctx_, sin, s0, n, S, 1.f / oversample, (float)(1.0 / wnorm), ex_.as<float>());
    forward_dev(ex_.as<float>(), n * W2, true, false, ey_.as<float>());
    lcf_scatter(*ctx_, sin, ey_.as<float>(), s0, n, S, P, plane_elems, 1.f / tscale, table, prec);
  }
}

}  // namespace ygrid
