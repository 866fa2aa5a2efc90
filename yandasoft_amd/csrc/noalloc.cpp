// The allocation-free degridding pass (ygrid_degrid_noalloc): no counterpart in the reference.
// Every device buffer of the pass beyond the caller's chunk and the gridder's own state (grid,
// CF cache) is carved from `scratch`, a file-scope device array of kNoAllocScratchBytes that each
// GPU holds once the library's code is loaded on it, until the process ends.
#include <hip/hip_runtime.h>

#include <chrono>
#include <mutex>

#include "gridders.h"

namespace ygrid {

static_assert(kNoAllocScratchBytes == YGRID_NOALLOC_WORKSPACE_BYTES, "workspace size of the header");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
// per device: the device address of that GPU's copy of `scratch`, and a lock that serialises the
// allocation-free passes on it (they share the array)
void* g_scratch[kMaxDevices] = {};
std::mutex g_pass_mu[kMaxDevices];
}  // namespace

// Resolves `scratch` on `device`, which must be current.  Called when a context is created on the
// device (ygrid_ctx_create); later calls for the same device return at once.
void noalloc_setup(int device) {
  YG_CHECK(device >= 0 && device < kMaxDevices, 1, "device index beyond the workspace table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return;
  int cur = -1;
  YG_HIP(hipGetDevice(&cur));
  YG_CHECK(cur == device, 2, "the workspace of a device is resolved with that device current");
  void* p = nullptr;
  YG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  YG_CHECK(p != nullptr, 3, "hipGetSymbolAddress returned a null workspace address");
  g_scratch[device] = p;
}

void* noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  YG_CHECK(device >= 0 && device < kMaxDevices && g_scratch[device], 2,
           "the device workspace is not resolved: create a context on the device first");
  return g_scratch[device];
}

// TableVisGridder::generic(acc, true) (TableVisGridder.cc:451-884) with the pass's buffers in the
// workspace: the prep kernel, the sync-free binning and one tile-kernel launch, as ygrid_degrid.
void Gridder::degridNoAlloc(const ygrid_chunk& a, void* ws, size_t bytes) {
  const auto t0 = std::chrono::steady_clock::now();
  YG_CHECK(initialised_ && degrid_mode_, 2, "degrid called on a gridder not initialised for degridding");
  if (model_empty_) return;  // TableVisGridder.cc:453-454
  YG_CHECK(a.mem == YGRID_MEM_DEVICE, 1, "the allocation-free degridding takes a chunk in device memory");
  YG_CHECK(!needsPolConversion(a), 1,
           "the allocation-free degridding takes visibilities in the image's polarisation frame");
  YG_CHECK(cf_ && cf_->n_elems > 0, 2,
           "the allocation-free degridding needs the gridder's convolution functions: run one ordinary degrid "
           "pass first (or clone a gridder that has)");
  YG_CHECK(!cfPerChunk() && !learnedActive(), 1,
           "the allocation-free degridding needs convolution functions that do not change per chunk "
           "(not A-projection, not learned CFs)");
  YG_CHECK(cf_prec_ == kPrecF32 || (cf_->table16_prec == cf_prec_ && cf_->table16.size() == cf_->n_elems * 4), 2,
           "the 16-bit copy of the CF table does not exist yet: run one ordinary degrid pass first");
  YG_CHECK(ws != nullptr, 2, "no device workspace");
  std::lock_guard<std::mutex> lk(g_pass_mu[ctx_->device]);
  PassArena arena{static_cast<char*>(ws), bytes};
  PassSetup ps;
  if (!setup_pass(a, true, ps, &arena)) return;
  const ChunkDev& d = ps.d;
  const int nsamp = ps.nsamp;
  const Ctx& c = *ctx_;
  StaticBinsView v;
  const int chunk = c.item_pairs > 0 ? c.item_pairs : kDefaultItemPairs;
  static_bins_shape(c, v, d.nrow, d.nchan, shape_[0], shape_[1], ps.pp.nzg * shape_[3], ps.cfv, chunk);
  const size_t nk1 = v.nkeys + 1;
  v.counts = arena.take_n<int>(nk1);
  v.offsets = arena.take_n<int>(nk1);
  v.cursor = arena.take_n<int>(nk1);
  v.icount = arena.take_n<int>(nk1);
  v.ioffsets = arena.take_n<int>(nk1);
  v.pairs = arena.take_n<int4>(std::max<size_t>(v.npairs_cap, 1));
  v.temp = arena.take(v.temp_bytes);
  v.wide = v.kstride > 1 ? arena.take_n<int>(nk1 * v.kstride) : nullptr;
  int4* items = arena.take_n<int4>(v.nitems_cap);
  int* base = arena.take_n<int>(2);  // [0] = 0: the item base; [1]: the item count
  int* counter = arena.take_n<int>(kQueueInts);
  TileJob* job = arena.take_n<TileJob>(1);
  // the workspace keeps the previous call's contents: clear what must start at zero (fill
  // kernels, no memset nodes)
  launch_fill_u32(c, ps.out.counters, 16, 0u);
  launch_fill_u32(c, ps.out.error, 2, 0u);
  launch_fill_u32(c, ps.acc, (size_t)nsamp * 8, 0u);
  launch_fill_u32(c, base, 2, 0u);
  if (profiling_) YG_HIP(hipEventRecord(ev_[0], c.stream));
  launch_prep(c, ps.pp, ps.in, ps.out);
  if (profiling_) YG_HIP(hipEventRecord(ev_[1], c.stream));
  enqueue_static_binning(c, v, ps.out.smp, ps.cfv, items, base, base + 1, 0);
  noalloc_job_ = make_tile_job(v.pairs, Samples{}, ps.cfv, grid_.as<float>(), ps.acc, shape_[0], shape_[1], shape_[2],
                               ps.pp.nzg, nGridPlanes() > 1 ? gridPlaneElems() : 0);
  YG_HIP(hipMemcpyAsync(job, &noalloc_job_, sizeof(TileJob), hipMemcpyHostToDevice, c.stream));
  if (profiling_) YG_HIP(hipEventRecord(ev_[2], c.stream));
  launch_tile_jobs(c, c.degrid_variant, true, shape_[2], tile_prec(ps.cfv), job, items, base + 1, counter);
  if (profiling_) YG_HIP(hipEventRecord(ev_[3], c.stream));
  launch_degrid_finish(c, ps.out.smp, ps.acc, d.nrow, d.nchan, d.npol, d.vis_out, d.vis_out_ld);
  unsigned long long cnt[5] = {0, 0, 0, 0, 0};
  int err = 0;
  YG_HIP(hipMemcpyAsync(cnt, ps.out.counters, sizeof(cnt), hipMemcpyDeviceToHost, c.stream));
  YG_HIP(hipMemcpyAsync(&err, ps.out.error, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  YG_HIP(hipStreamSynchronize(c.stream));
  if (profiling_) {
    float ta = 0, tb = 0, tk = 0;
    YG_HIP(hipEventElapsedTime(&ta, ev_[0], ev_[1]));
    YG_HIP(hipEventElapsedTime(&tb, ev_[1], ev_[2]));
    YG_HIP(hipEventElapsedTime(&tk, ev_[2], ev_[3]));
    ms_prep_ += ta;
    ms_bin_ += tb;
    ms_dk_ += tk;
    n_dk_ += 1;
  }
  account_pass(true, cnt, err, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
}

}  // namespace ygrid
