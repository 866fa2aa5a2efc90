// Batched passes over several gridders (see batch.cpp).
#pragma once

#include <memory>
#include <vector>

#include "gridders.h"

namespace ygrid {

struct BatchTuning {
  int grid_waves, grid_pipeline, degrid_waves, degrid_pipeline, item_pairs;
  double ms;   // best run time (ms) of the batch
  int trials;  // configurations timed
};

class Batch {
 public:
  explicit Batch(std::shared_ptr<Ctx> ctx);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  // Adds the pass of gridder g over chunk a (device memory): gridding or degridding as g was
  // initialised.  The chunk's arrays must keep their addresses and shapes while the batch is used.
  void add(Gridder* g, const ygrid_chunk& a);
  void run();
  void set_graph(bool on);
  BatchTuning autotune(int trials);
  double last_ms() const { return last_ms_; }
  int captures() const { return n_captures_; }

 private:
  struct Pass {
    Gridder* g = nullptr;
    ygrid_chunk a{};
    bool forward = false, active = false;
    Gridder::PassSetup ps;
    StaticBins sb;
    int group = -1, job = -1;
  };
  struct Group {
    bool forward = false;
    int npol = 1, prec = kPrecF32;
    std::vector<int> passes;
    DevBuf jobs, items, base, counter;
    size_t items_cap = 0;
  };
  void prepare();
  void enqueue();
  void drop_graph();
  void invalidate();

  std::shared_ptr<Ctx> ctx_;
  std::vector<Pass> passes_;
  std::vector<Group> groups_;
  bool prepared_ = false, graph_on_ = false;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  double last_ms_ = 0;
  int n_runs_ = 0, n_captures_ = 0, n_tuned_ = 0;
};

}  // namespace ygrid
