// extern "C" boundary of libygrid (declared in include/ygrid.h).  Every entry point catches
// C++ exceptions and turns them into a status code plus a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "batch.h"
#include "gridders.h"
#include "snapshot.h"

struct ygrid_ctx {
  std::shared_ptr<ygrid::Ctx> c;
};
struct ygrid_gridder {
  std::unique_ptr<ygrid::Gridder> g;
  std::unique_ptr<ygrid::Snapshot> s;  // snapshot imaging around g (gridder.snapshotimaging)
};
struct ygrid_batch {
  std::unique_ptr<ygrid::Batch> b;
};
struct ygrid_lcf {
  std::shared_ptr<ygrid::LearnedCF> l;
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guard(F&& f) {
  try {
    f();
    return YGRID_OK;
  } catch (const ygrid::Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return YGRID_ERR_ARG;
  } catch (...) {
    g_last_error = "unknown error";
    return YGRID_ERR_ARG;
  }
}

#define NEED(p) YG_CHECK((p) != nullptr, YGRID_ERR_ARG, #p " is null")
}  // namespace

extern "C" {

const char* ygrid_last_error(void) { return g_last_error.c_str(); }
int ygrid_version(void) { return YGRID_VERSION; }

int ygrid_ctx_create(int device, ygrid_ctx** out) {
  return guard([&] {
    NEED(out);
    int n = 0;
    YG_HIP(hipGetDeviceCount(&n));
    YG_CHECK(device >= 0 && device < n, YGRID_ERR_ARG, "device index out of range");
    YG_HIP(hipSetDevice(device));
    auto c = std::make_shared<ygrid::Ctx>();
    c->device = device;
    hipDeviceProp_t prop;
    YG_HIP(hipGetDeviceProperties(&prop, device));
    c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    // tile-kernel variants "waves,pipeline[,matrix]" (e.g. "3,0" or "2,1,0"): performance
    // tuning only; matrix selects the matrix-core degridding (on unless given as 0)
    const auto variant = [](const char* env, ygrid::TileVariant& v) {
      if (const char* s = std::getenv(env)) {
        int w = 0, p = 0, m = v.matrix;
        if (std::sscanf(s, "%d,%d,%d", &w, &p, &m) >= 2 && (w == 2 || w == 3))
          v = ygrid::TileVariant{w, p != 0, m != 0};
      }
    };
    variant("YGRID_GRID_VARIANT", c->grid_variant);
    variant("YGRID_DEGRID_VARIANT", c->degrid_variant);
    if (const char* s = std::getenv("YGRID_ITEM_PAIRS")) c->item_pairs = std::max(16, std::atoi(s));
    if (const char* s = std::getenv("YGRID_SMALL_SUPPORT"))
      c->small_support_max = std::min(std::max(0, std::atoi(s)), ygrid::kSmallSupportMax);
    // the allocation-free entry point's workspace on this device (its device is current here)
    ygrid::noalloc_setup(device);
    YG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
    *out = new ygrid_ctx{c};
  });
}

int ygrid_ctx_destroy(ygrid_ctx* ctx) {
  return guard([&] {
    if (!ctx) return;
    if (ctx->c.use_count() == 1 && ctx->c->own_stream && ctx->c->stream) {
      (void)hipStreamSynchronize(ctx->c->stream);
      (void)hipStreamDestroy(ctx->c->stream);
      ctx->c->stream = nullptr;
    }
    delete ctx;
  });
}

int ygrid_ctx_set_stream(ygrid_ctx* ctx, void* stream) {
  return guard([&] {
    NEED(ctx);
    if (ctx->c->own_stream && ctx->c->stream) {
      YG_HIP(hipStreamSynchronize(ctx->c->stream));
      YG_HIP(hipStreamDestroy(ctx->c->stream));
    }
    ctx->c->stream = static_cast<hipStream_t>(stream);
    ctx->c->own_stream = false;
  });
}

int ygrid_ctx_synchronize(ygrid_ctx* ctx) {
  return guard([&] {
    NEED(ctx);
    YG_HIP(hipStreamSynchronize(ctx->c->stream));
  });
}

int ygrid_ctx_clear_cache(ygrid_ctx* ctx) {
  return guard([&] {
    NEED(ctx);
    ctx->c->bins.valid = false;
  });
}

int ygrid_ctx_set_small_support(ygrid_ctx* ctx, int max_support) {
  return guard([&] {
    NEED(ctx);
    YG_CHECK(max_support >= 0 && max_support <= ygrid::kSmallSupportMax, YGRID_ERR_ARG,
             "small-support limit must lie in 0..8");
    ctx->c->small_support_max = max_support;
  });
}

int ygrid_ctx_set_small_band_rows(ygrid_ctx* ctx, int rows) {
  return guard([&] {
    NEED(ctx);
    YG_CHECK(rows == 0 || rows == 4 || rows == 8 || rows == 16, YGRID_ERR_ARG,
             "small-support band rows must be 0 (by support), 4, 8 or 16");
    ctx->c->small_band_rows = rows;
  });
}

int ygrid_gridder_autotune_small(ygrid_gridder* g, const ygrid_image_desc* img, const ygrid_chunk* acc, int trials,
                                 int* best_rows, double* best_ms) {
  return guard([&] {
    NEED(g);
    NEED(img);
    NEED(acc);
    NEED(best_rows);
    NEED(best_ms);
    YG_HIP(hipSetDevice(g->g->context()->device));
    ygrid::Ctx& c = *g->g->context();
    const int r0 = c.small_band_rows;
    const int cand[] = {4, 8, 16};
    int best = r0;
    double bms = -1.0;
    try {
      for (int rows : cand) {
        c.small_band_rows = rows;
        double t = 1e300;
        for (int k = 0; k < std::max(1, trials); ++k) {
          auto t1 = g->g->clone();
          t1->initialiseGrid(*img, false, false);
          YG_HIP(hipStreamSynchronize(c.stream));
          const auto a = std::chrono::steady_clock::now();
          t1->grid(*acc);
          YG_HIP(hipStreamSynchronize(c.stream));
          t = std::min(t, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - a).count());
        }
        if (bms < 0 || t < bms) {
          bms = t;
          best = rows;
        }
      }
    } catch (...) {
      c.small_band_rows = r0;
      throw;
    }
    c.small_band_rows = best;
    *best_rows = best;
    *best_ms = bms;
  });
}

int ygrid_gridder_create(ygrid_ctx* ctx, const char* name, const char* parset, ygrid_gridder** out) {
  return guard([&] {
    NEED(ctx);
    NEED(name);
    NEED(out);
    YG_HIP(hipSetDevice(ctx->c->device));
    const ygrid::ParamMap m = ygrid::parse_parset(parset);
    auto g = ygrid::make_gridder(ctx->c, name, m);
    auto h = std::unique_ptr<ygrid_gridder>(new ygrid_gridder{std::move(g)});
    if (ygrid::Params{m}.flag("snapshotimaging", false)) h->s.reset(new ygrid::Snapshot(h->g.get(), ygrid::Params{m}));
    *out = h.release();
  });
}

int ygrid_gridder_configure(ygrid_gridder* g, const char* options) {
  return guard([&] {
    NEED(g);
    const ygrid::ParamMap m = ygrid::parse_parset(options);
    for (const auto& kv : m) {
      static const char* known[] = {"padding", "alldatapsf", "oversampleweight", "cleargrids", "MaxPointingSeparation",
                                    "parotation", "parotation.angle", "swappols", "snapshotimaging",
                                    "snapshotimaging.wtolerance", "snapshotimaging.clipping",
                                    "snapshotimaging.weightsclipping", "snapshotimaging.reprojectpsf",
                                    "snapshotimaging.interpmethod", "snapshotimaging.coorddecimation",
                                    "snapshotimaging.longtrack"};
      bool ok = false;
      for (const char* k : known) ok = ok || kv.first == k;
      YG_CHECK(ok, YGRID_ERR_ARG, "ygrid_gridder_configure: unknown option " + kv.first);
    }
    YG_CHECK(!ygrid::Params{m}.flag("swappols", false), YGRID_ERR_ARG,
             "polarisation swapping (gridder.swappols) needs parallactic angle rotation, which the GPU gridders do not do");
    g->g->configure_globals(ygrid::Params{m});
    // VisGridderFactory.cc:248-270: gridder.snapshotimaging wraps the gridder in the snapshot
    // imaging adapter (once; later configurations leave an existing wrapper as it is)
    const ygrid::Params p{m};
    if (p.has("snapshotimaging")) {
      if (!p.flag("snapshotimaging", false))
        g->s.reset();
      else if (!g->s)
        g->s.reset(new ygrid::Snapshot(g->g.get(), p));
    }
  });
}

int ygrid_gridder_clone(const ygrid_gridder* g, ygrid_gridder** out) {
  return guard([&] {
    NEED(g);
    NEED(out);
    auto c = new ygrid_gridder{g->g->clone()};
    if (g->s) c->s.reset(new ygrid::Snapshot(*g->s, c->g.get()));
    *out = c;
  });
}

int ygrid_gridder_destroy(ygrid_gridder* g) {
  return guard([&] { delete g; });
}

int ygrid_initialise_grid(ygrid_gridder* g, const ygrid_image_desc* img, int dopsf, int dopcf) {
  return guard([&] {
    NEED(g);
    NEED(img);
    if (g->s) g->s->initialiseGrid(*img, dopsf != 0, dopcf != 0);
    else g->g->initialiseGrid(*img, dopsf != 0, dopcf != 0);
  });
}

int ygrid_grid(ygrid_gridder* g, const ygrid_chunk* acc) {
  return guard([&] {
    NEED(g);
    NEED(acc);
    if (g->s) g->s->grid(*acc);
    else g->g->grid(*acc);
  });
}

int ygrid_finalise_grid(ygrid_gridder* g, float* out, int mem) {
  return guard([&] {
    NEED(g);
    NEED(out);
    if (g->s) g->s->finaliseGrid(out, mem);
    else g->g->finaliseGrid(out, mem);
  });
}

int ygrid_finalise_weights(ygrid_gridder* g, float* out, int mem) {
  return guard([&] {
    NEED(g);
    NEED(out);
    if (g->s) g->s->finaliseWeights(out, mem);
    else g->g->finaliseWeights(out, mem);
  });
}

int ygrid_initialise_degrid(ygrid_gridder* g, const ygrid_image_desc* img, const float* model, int mem) {
  return guard([&] {
    NEED(g);
    NEED(img);
    NEED(model);
    if (g->s) g->s->initialiseDegrid(*img, model, mem);
    else g->g->initialiseDegrid(*img, model, mem);
  });
}

int ygrid_degrid(ygrid_gridder* g, const ygrid_chunk* acc) {
  return guard([&] {
    NEED(g);
    NEED(acc);
    if (g->s) g->s->degrid(*acc);
    else g->g->degrid(*acc);
  });
}

int ygrid_degrid_noalloc(ygrid_gridder* g, const ygrid_chunk* acc) {
  return guard([&] {
    NEED(g);
    NEED(acc);
    YG_CHECK(!g->s, YGRID_ERR_ARG, "ygrid_degrid_noalloc: not for snapshot imaging gridders");
    const int dev = g->g->context()->device;
    YG_HIP(hipSetDevice(dev));
    g->g->degridNoAlloc(*acc, ygrid::noalloc_scratch(dev), ygrid::kNoAllocScratchBytes);
  });
}

int ygrid_finalise_degrid(ygrid_gridder* g) {
  return guard([&] {
    NEED(g);
    if (g->s) g->s->finaliseDegrid();
    else g->g->finaliseDegrid();
  });
}

int ygrid_is_model_empty(const ygrid_gridder* g, int* out) {
  return guard([&] {
    NEED(g);
    NEED(out);
    *out = (g->s ? g->s->isModelEmpty() : g->g->isModelEmpty()) ? 1 : 0;
  });
}

int ygrid_set_mfs_weight(ygrid_gridder* g, double ref_freq, int order) {
  return guard([&] {
    NEED(g);
    YG_CHECK(order < 0 || ref_freq > 0, YGRID_ERR_ARG, "reference frequency must be positive");
    g->g->setMFSWeight(ref_freq, order);
  });
}

int ygrid_get_stats(const ygrid_gridder* g, ygrid_stats* out) {
  return guard([&] {
    NEED(g);
    NEED(out);
    std::memset(out, 0, sizeof(*out));
    g->g->stats(out);
    if (g->s) g->s->stats(out);
  });
}

int ygrid_set_profiling(ygrid_gridder* g, int enable) {
  return guard([&] {
    NEED(g);
    g->g->setProfiling(enable != 0);
  });
}

int ygrid_get_cf(const ygrid_gridder* g, int k, float* out, int* support) {
  return guard([&] {
    NEED(g);
    g->g->getCF(k, out, support);
  });
}

int ygrid_get_grid(const ygrid_gridder* g, float* out, size_t* n) {
  return guard([&] {
    NEED(g);
    g->g->getGrid(out, n);
  });
}

int ygrid_imaging_chunk(ygrid_gridder* const* degridders, int ndegrid, ygrid_gridder* const* gridders, int ngrid,
                        const ygrid_chunk* chunk) {
  return guard([&] {
    NEED(chunk);
    YG_CHECK(ndegrid >= 0 && ngrid >= 0, YGRID_ERR_ARG, "negative gridder count");
    YG_CHECK(ndegrid == 0 || degridders, YGRID_ERR_ARG, "degridders is null");
    YG_CHECK(ngrid == 0 || gridders, YGRID_ERR_ARG, "gridders is null");
    std::vector<ygrid::Gridder*> d, g;
    for (int i = 0; i < ndegrid; ++i) {
      NEED(degridders[i]);
      YG_CHECK(!degridders[i]->s, YGRID_ERR_ARG, "ygrid_imaging_chunk: not for snapshot imaging gridders");
      d.push_back(degridders[i]->g.get());
    }
    for (int i = 0; i < ngrid; ++i) {
      NEED(gridders[i]);
      YG_CHECK(!gridders[i]->s, YGRID_ERR_ARG, "ygrid_imaging_chunk: not for snapshot imaging gridders");
      g.push_back(gridders[i]->g.get());
    }
    ygrid::imaging_chunk(d, g, *chunk);
  });
}

int ygrid_uvw_rotate(ygrid_ctx* ctx, const ygrid_uvw_rotation* r) {
  return guard([&] {
    NEED(ctx);
    NEED(r);
    YG_CHECK(r->nrow >= 0, YGRID_ERR_ARG, "negative row count");
    if (r->nrow == 0) return;
    NEED(r->uvw);
    NEED(r->phase_centre);
    NEED(r->uvw_out);
    YG_HIP(hipSetDevice(ctx->c->device));
    const ygrid::Ctx& c = *ctx->c;
    const ygrid::UvwTargets T = ygrid::uvw_targets(r->tangent_lon, r->tangent_lat, r->centre_lon, r->centre_lat);
    const size_t n = (size_t)r->nrow;
    const size_t npc = r->phase_centre_per_row ? 2 * n : 2;
    if (r->mem == YGRID_MEM_DEVICE) {
      ygrid::launch_uvw_rotate(c, r->uvw, r->phase_centre, r->phase_centre_per_row ? 1 : 0, T, r->nrow, r->uvw_out,
                               r->delay_out);
      return;
    }
    YG_CHECK(r->mem == YGRID_MEM_HOST, YGRID_ERR_ARG, "mem must be YGRID_MEM_HOST or YGRID_MEM_DEVICE");
    ygrid::DevBuf in(n * 3 * sizeof(double)), pc(npc * sizeof(double)), out(n * 3 * sizeof(double)),
        dl(r->delay_out ? n * sizeof(double) : 0);
    YG_HIP(hipMemcpyAsync(in.get(), r->uvw, n * 3 * sizeof(double), hipMemcpyHostToDevice, c.stream));
    YG_HIP(hipMemcpyAsync(pc.get(), r->phase_centre, npc * sizeof(double), hipMemcpyHostToDevice, c.stream));
    ygrid::launch_uvw_rotate(c, in.as<double>(), pc.as<double>(), r->phase_centre_per_row ? 1 : 0, T, r->nrow,
                             out.as<double>(), r->delay_out ? dl.as<double>() : nullptr);
    YG_HIP(hipMemcpyAsync(r->uvw_out, out.get(), n * 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    if (r->delay_out)
      YG_HIP(hipMemcpyAsync(r->delay_out, dl.get(), n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    YG_HIP(hipStreamSynchronize(c.stream));
  });
}

int ygrid_batch_create(ygrid_ctx* ctx, ygrid_batch** out) {
  return guard([&] {
    NEED(ctx);
    NEED(out);
    YG_HIP(hipSetDevice(ctx->c->device));
    *out = new ygrid_batch{std::make_unique<ygrid::Batch>(ctx->c)};
  });
}

int ygrid_batch_destroy(ygrid_batch* b) {
  return guard([&] { delete b; });
}

int ygrid_batch_add(ygrid_batch* b, ygrid_gridder* g, const ygrid_chunk* acc) {
  return guard([&] {
    NEED(b);
    NEED(g);
    NEED(acc);
    YG_CHECK(!g->s, YGRID_ERR_ARG, "ygrid_batch_add: not for snapshot imaging gridders");
    b->b->add(g->g.get(), *acc);
  });
}

int ygrid_batch_set_graph(ygrid_batch* b, int enable) {
  return guard([&] {
    NEED(b);
    b->b->set_graph(enable != 0);
  });
}

int ygrid_batch_run(ygrid_batch* b) {
  return guard([&] {
    NEED(b);
    b->b->run();
  });
}

int ygrid_batch_autotune(ygrid_batch* b, int trials, ygrid_tuning* best) {
  return guard([&] {
    NEED(b);
    const ygrid::BatchTuning t = b->b->autotune(trials);
    if (best) {
      best->grid_waves = t.grid_waves;
      best->grid_pipeline = t.grid_pipeline;
      best->degrid_waves = t.degrid_waves;
      best->degrid_pipeline = t.degrid_pipeline;
      best->item_pairs = t.item_pairs;
      best->ms = t.ms;
      best->trials = t.trials;
    }
  });
}

int ygrid_lcf_create(ygrid_ctx* ctx, int hidden, int layers, unsigned long long seed, ygrid_lcf** out) {
  return guard([&] {
    NEED(ctx);
    NEED(out);
    YG_HIP(hipSetDevice(ctx->c->device));
    *out = new ygrid_lcf{std::make_shared<ygrid::LearnedCF>(ctx->c, hidden, layers, seed)};
  });
}

int ygrid_lcf_destroy(ygrid_lcf* l) {
  return guard([&] { delete l; });
}

int ygrid_lcf_num_params(const ygrid_lcf* l, int* n) {
  return guard([&] {
    NEED(l);
    NEED(n);
    *n = l->l->nparams();
  });
}

int ygrid_lcf_get_params(const ygrid_lcf* l, float* out) {
  return guard([&] {
    NEED(l);
    NEED(out);
    l->l->get_params(out);
  });
}

int ygrid_lcf_set_params(ygrid_lcf* l, const float* in) {
  return guard([&] {
    NEED(l);
    NEED(in);
    l->l->set_params(in);
  });
}

int ygrid_lcf_clear_cache(ygrid_lcf* l) {
  return guard([&] {
    NEED(l);
    l->l->clear_cache();
  });
}

int ygrid_lcf_loss_grad(ygrid_lcf* l, const float* x, const float* t, int n, double* loss, float* grad) {
  return guard([&] {
    NEED(l);
    NEED(x);
    NEED(t);
    const double v = l->l->loss_and_grads_host(x, t, n, grad);
    if (loss) *loss = v;
  });
}

int ygrid_lcf_forward(ygrid_lcf* l, const float* x, float* y, int n, int half) {
  return guard([&] {
    NEED(l);
    NEED(x);
    NEED(y);
    l->l->forward_host(x, y, n, half);
  });
}

int ygrid_lcf_train_wproject(ygrid_lcf* l, const ygrid_gridder* wp, const ygrid_lcf_train_opts* opts,
                             ygrid_lcf_report* report) {
  return guard([&] {
    NEED(l);
    NEED(wp);
    ygrid::LcfTrainOptions o;
    if (opts) {
      o.epochs = opts->epochs;
      o.batch = opts->batch;
      o.lr = opts->lr;
      o.seed = opts->seed;
      o.holdout_every = opts->holdout_every;
    }
    YG_CHECK(o.epochs > 0 && o.batch > 0 && o.lr > 0.f, YGRID_ERR_ARG, "bad training options");
    const ygrid::LcfReport r = l->l->train_wproject(*wp->g, o);
    if (report) {
      report->epochs = r.epochs;
      report->steps = r.steps;
      report->examples = r.examples;
      report->cfs = r.cfs;
      report->holdout_cfs = r.holdout_cfs;
      report->first_loss = r.first_loss;
      report->last_loss = r.last_loss;
      report->holdout_max_err = r.holdout_max_err;
      report->holdout_rms_err = r.holdout_rms_err;
      report->train_s = r.train_s;
      report->support = l->l->support;
    }
  });
}

int ygrid_gridder_set_learned_cf(ygrid_gridder* g, ygrid_lcf* lcf, double max_err) {
  return guard([&] {
    NEED(g);
    g->g->setLearnedCF(lcf ? lcf->l : nullptr, max_err);
  });
}

int ygrid_batch_info(const ygrid_batch* b, double* last_ms, int* captures) {
  return guard([&] {
    NEED(b);
    if (last_ms) *last_ms = b->b->last_ms();
    if (captures) *captures = b->b->captures();
  });
}

}  // extern "C"
