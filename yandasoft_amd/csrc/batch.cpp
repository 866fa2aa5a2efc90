// Batched passes: many (gridder, chunk) passes run as one sequence, with one persistent
// tile-kernel launch per kind (grid / degrid, polarisations, CF precision) over the work items
// of every pass, optionally captured once in a HIP graph and replayed.
//
// The reference runs its gridders one after the other (ImageFFTEquation.cc:362-387 loops over
// the images, each TableVisGridder::generic over its accessor); a batch gives each gridder the
// same result as its own ygrid_grid / ygrid_degrid on its chunk.  What the batch adds is the
// launch structure: per pass the prep kernel and a binning whose buffers are sized by upper
// bounds (no host readback, no allocation, no memset nodes), so that the whole sequence can be
// captured in a graph; then the tile kernels take work items from all passes at once.
#include "batch.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <map>
#include <tuple>

namespace ygrid {

Batch::Batch(std::shared_ptr<Ctx> ctx) : ctx_(std::move(ctx)) {}

Batch::~Batch() { drop_graph(); }

void Batch::drop_graph() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

void Batch::invalidate() {
  drop_graph();
  prepared_ = false;
}

void Batch::add(Gridder* g, const ygrid_chunk& a) {
  YG_CHECK(g != nullptr, 1, "null gridder");
  YG_CHECK(a.mem == YGRID_MEM_DEVICE, 1, "batched passes need chunks in device memory");
  YG_CHECK(g->nGridPlanes() == 1, 1, "batched passes take single-grid gridders (not w stacking)");
  YG_CHECK(!g->needsPolConversion(a), 1,
           "batched passes take visibilities in the image's polarisation frame (convert with a single pass)");
  for (const Pass& p : passes_) YG_CHECK(p.g != g, 1, "a gridder can take part in one pass of a batch only");
  Pass p;
  p.g = g;
  p.a = a;
  p.forward = g->degrid_mode();
  passes_.push_back(std::move(p));
  invalidate();
}

// Eager part: every pass's setup (staging, CF generation, buffer sizes), upper-bound binning
// buffers, the job tables and the combined item buffers.  All device memory the sequence
// touches exists after this and keeps its address until the batch changes.
void Batch::prepare() {
  drop_graph();
  groups_.clear();
  std::map<std::tuple<int, int, int>, int> gidx;
  const int chunk = ctx_->item_pairs > 0 ? ctx_->item_pairs : kDefaultItemPairs;
  for (Pass& p : passes_) {
    p.active = p.g->setup_pass(p.a, p.forward, p.ps);
    if (!p.active) continue;
    const Gridder::PassSetup& ps = p.ps;
    const int nkz = ps.pp.nzg * p.g->shape_[3];
    size_static_bins(*ctx_, p.sb, ps.d.nrow, ps.d.nchan, p.g->shape_[0], p.g->shape_[1], nkz, ps.cfv, chunk);
    const auto key = std::make_tuple((int)p.forward, p.g->shape_[2], tile_prec(ps.cfv));
    auto it = gidx.find(key);
    if (it == gidx.end()) {
      it = gidx.emplace(key, (int)groups_.size()).first;
      Group gr;
      gr.forward = p.forward;
      gr.npol = p.g->shape_[2];
      gr.prec = tile_prec(ps.cfv);
      groups_.push_back(std::move(gr));
    }
    Group& gr = groups_[it->second];
    p.group = it->second;
    p.job = (int)gr.passes.size();
    gr.passes.push_back((int)(&p - passes_.data()));
    gr.items_cap += p.sb.v.nitems_cap;
  }
  for (Group& gr : groups_) {
    YG_CHECK(gr.items_cap < (size_t)INT_MAX, 1, "too many work items in one batched launch");
    gr.items.resize(gr.items_cap * sizeof(int4));
    gr.base.resize((gr.passes.size() + 1) * sizeof(int));
    gr.counter.resize(kQueueInts * sizeof(int));
    std::vector<TileJob> jobs;
    for (int pi : gr.passes) {
      Pass& p = passes_[pi];
      Gridder& g = *p.g;
      jobs.push_back(make_tile_job(p.sb.v.pairs, p.ps.out.smp, p.ps.cfv, g.grid_.as<float>(),
                                   p.forward ? p.ps.acc : nullptr, g.shape_[0], g.shape_[1], g.shape_[2],
                                   p.ps.pp.nzg));
    }
    gr.jobs.resize(jobs.size() * sizeof(TileJob));
    YG_HIP(hipMemcpy(gr.jobs.get(), jobs.data(), jobs.size() * sizeof(TileJob), hipMemcpyHostToDevice));
  }
  prepared_ = true;
}

// The device work of one run, in stream order.  Capturable: kernels only.
void Batch::enqueue() {
  const Ctx& c = *ctx_;
  for (Group& gr : groups_) launch_fill_u32(c, gr.base.get(), gr.passes.size() + 1, 0u);
  for (Pass& p : passes_) {
    if (!p.active) continue;
    Gridder& g = *p.g;
    Group& gr = groups_[p.group];
    launch_fill_u32(c, p.ps.out.counters, 8 * 2, 0u);
    launch_fill_u32(c, p.ps.out.error, 1, 0u);
    if (p.forward) launch_fill_u32(c, p.ps.acc, (size_t)p.ps.nsamp * 8, 0u);
    launch_prep(c, p.ps.pp, p.ps.in, p.ps.out);
    g.afterPrep(p.ps);
    int* base = gr.base.as<int>();
    enqueue_static_binning(c, p.sb.v, p.ps.out.smp, p.ps.cfv, gr.items.as<int4>(), base + p.job, base + p.job + 1,
                           p.job);
  }
  for (Group& gr : groups_) {
    launch_tile_jobs(c, gr.forward ? c.degrid_variant : c.grid_variant, gr.forward, gr.npol, gr.prec,
                     gr.jobs.as<TileJob>(), gr.items.as<int4>(), gr.base.as<int>() + gr.passes.size(),
                     gr.counter.as<int>());
  }
  for (Pass& p : passes_) {
    if (!p.active || !p.forward) continue;
    const ChunkDev& d = p.ps.d;
    launch_degrid_finish(c, p.ps.out.smp, p.ps.acc, d.nrow, d.nchan, d.npol, d.vis_out, d.vis_out_ld);
  }
}

void Batch::run() {
  YG_CHECK(!passes_.empty(), 2, "empty batch");
  if (!prepared_) prepare();
  const auto t0 = std::chrono::steady_clock::now();
  hipStream_t s = ctx_->stream;
  if (graph_on_) {
    YG_CHECK(s != nullptr, 2, "graph capture needs the context's own stream (not the null stream)");
    if (!exec_) {
      YG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        enqueue();
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(s, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
      }
      YG_HIP(hipStreamEndCapture(s, &graph_));
      YG_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
      ++n_captures_;
    }
    YG_HIP(hipGraphLaunch(exec_, s));
  } else {
    enqueue();
  }
  YG_HIP(hipStreamSynchronize(s));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  last_ms_ = 1e3 * secs;
  ++n_runs_;
  int nact = 0;
  for (const Pass& p : passes_) nact += p.active ? 1 : 0;
  for (Pass& p : passes_) {
    if (!p.active) continue;
    unsigned long long cnt[5] = {0, 0, 0, 0, 0};
    int err = 0;
    YG_HIP(hipMemcpy(cnt, p.ps.out.counters, sizeof(cnt), hipMemcpyDeviceToHost));
    YG_HIP(hipMemcpy(&err, p.ps.out.error, sizeof(int), hipMemcpyDeviceToHost));
    p.g->account_pass(p.forward, cnt, err, secs / nact);
  }
}

void Batch::set_graph(bool on) {
  if (on != graph_on_) drop_graph();
  graph_on_ = on;
}

// Times every tile-kernel variant and work-item size on this batch (min over `trials` eager
// runs each) and keeps the fastest on the context.  Every trial is a full run: grids and
// degridded visibilities accumulate it, so re-initialise them before the runs that count.
BatchTuning Batch::autotune(int trials) {
  YG_CHECK(!passes_.empty(), 2, "empty batch");
  trials = std::max(1, trials);
  const bool graph = graph_on_;
  graph_on_ = false;
  drop_graph();
  bool f32 = true;
  if (!prepared_) prepare();
  for (const Group& gr : groups_) f32 = f32 && gr.prec == kPrecF32;
  std::vector<TileVariant> vars = {{2, 1}, {2, 0}, {3, 1}, {3, 0}};
  if (!f32) vars = {{2, 1}};  // 16-bit CF tables have the one variant
  const int chunks[] = {256, 512, 2048};
  BatchTuning best{};
  best.ms = -1.0;
  const TileVariant g0 = ctx_->grid_variant, d0 = ctx_->degrid_variant;
  const int c0 = ctx_->item_pairs;
  n_tuned_ = 0;
  try {
    for (int ch : chunks) {
      ctx_->item_pairs = ch;
      invalidate();
      for (const TileVariant& v : vars) {
        ctx_->grid_variant = v;
        ctx_->degrid_variant = TileVariant{v.waves, v.pipeline, d0.matrix};
        double t = 1e300;
        for (int r = 0; r < trials; ++r) {
          run();
          t = std::min(t, last_ms_);
        }
        ++n_tuned_;
        if (best.ms < 0 || t < best.ms) {
          best.ms = t;
          best.grid_waves = best.degrid_waves = v.waves;
          best.grid_pipeline = best.degrid_pipeline = v.pipeline;
          best.item_pairs = ch;
        }
      }
    }
  } catch (...) {
    ctx_->grid_variant = g0;
    ctx_->degrid_variant = d0;
    ctx_->item_pairs = c0;
    graph_on_ = graph;
    invalidate();
    throw;
  }
  ctx_->grid_variant = TileVariant{best.grid_waves, best.grid_pipeline, 0};
  ctx_->degrid_variant = TileVariant{best.degrid_waves, best.degrid_pipeline, d0.matrix};
  ctx_->item_pairs = best.item_pairs;
  best.trials = n_tuned_;
  graph_on_ = graph;
  invalidate();
  return best;
}

}  // namespace ygrid
