// C++ host side of libygrid: a casacore-free mirror of askap::synthesis::TableVisGridder
// (askap/gridding/TableVisGridder.h) and its Box / SphFunc / WProject / AWProject subclasses,
// driving the HIP kernels.  The C ABI in ygrid_api.cpp wraps these classes.
#pragma once

#include <cmath>
#include <map>
#include <memory>
#include <string>
#include <complex>
#include <vector>

#include "../../include/ygrid.h"
#include "learned_cf.h"
#include "ygrid_internal.h"

namespace ygrid {

using ParamMap = std::map<std::string, std::string>;
ParamMap parse_parset(const char* text);
// SynthesisParamsHelper::convertQuantity for angle (rad, deg, arcmin, arcsec) and length (m, cm,
// mm, km) units; a bare number is taken in `unit`
double quantity(const std::string& s, const std::string& unit);
// MVDirection::separation: great-circle angle between two (longitude, latitude) directions
double separation(double l1, double b1, double l2, double b2);

struct Params {
  const ParamMap& m;
  bool has(const std::string& k) const { return m.count(k) != 0; }
  std::string str(const std::string& k, const std::string& d) const;
  int i32(const std::string& k, int d) const;
  double f64(const std::string& k, double d) const;
  bool flag(const std::string& k, bool d) const;
};

// w sampling of the W-dependent gridders (WDependentGridderBase.cc:53-245, PowerWSampling.cc,
// GaussianWSampling.cc)
struct WSamplingDesc {
  double wmax = 35000;
  int nw = 65;
  int kind = 0;  // 0 linear, 1 powerlaw, 2 gaussian
  double exponent = 1, ts = 1, amp = 1;
  bool wmaxclip = false;
  double wscale() const { return nw > 1 ? wmax / double((nw - 1) / 2) : 1.0; }
  double map(double plane) const {
    if (kind == 1) return plane > 0 ? std::pow(plane, exponent) : (plane < 0 ? -std::pow(-plane, exponent) : 0.);
    if (kind == 2) {
      const double a = amp * (1 - std::exp(-plane * plane / ts));
      return plane < 0 ? -a : (plane > 0 ? a : 0.);
    }
    return plane;
  }
  // WDependentGridderBase::getWTerm (WDependentGridderBase.cc:157-167)
  double wterm(int plane) const {
    const int half = (nw - 1) / 2;
    if (kind != 0 && nw > 1) return map(double(plane - half) / double(half)) * wmax;
    return double(plane - half) * wscale();
  }
  // GaussianWSampling::calculateDistributionParameters (GaussianWSampling.cc:127-166)
  void gaussian(double w50) {
    auto target = [](double t) { return std::sqrt(-t * std::log(0.5 + 0.5 * std::exp(-1. / t))); };
    YG_CHECK(w50 > 0 && w50 * std::sqrt(2.) < 1., 1, "wsampling.nwplanes50 out of range");
    double lo = w50 * w50 / std::log(2.) / 2., hi = 0.2 / (1 / std::sqrt(2.) - w50);
    if (hi < lo) { lo = hi; hi = 5.; }
    double t = lo;
    for (int it = 0; it < 100; ++it) {
      t = (lo + hi) / 2.;
      if (target(t) < w50) lo = t; else hi = t;
      if (hi - lo < 1e-5) break;
    }
    ts = t;
    amp = 1. / (1. - std::exp(-1. / ts));
  }
};

// WDependentGridderBase::configureWSampling (WDependentGridderBase.cc:192-245)
void configure_wsampling(WSamplingDesc& ws, const Params& p);

// Device-resident CF cache (TableVisGridder::itsConvFunc, TableVisGridder.h:356), shared by
// clones once generated (WProjectVisGridder's shared cache, WProjectVisGridder.cc:51-52).
struct CFTable {
  int oversample = 1;
  int nz = 0;            // planes before oversampling
  int support = 0;       // largest support
  std::vector<int32_t> off, supp;      // per CF plane (nz*os*os)
  std::vector<int32_t> ofs_u, ofs_v;   // per z plane (empty: no offsets)
  DevBuf table, d_off, d_supp, d_ofs_u, d_ofs_v;
  size_t n_elems = 0;
  // the table rounded to 16-bit precision table16_prec (built on first use at that precision)
  DevBuf table16;
  int table16_prec = 0;
  // tables of 2 GiB or more: per key group of zg_zdiv planes, the element offset of its first
  // plane (device; built on first use for that grouping)
  DevBuf d_zg;
  int zg_zdiv = 0, zg_n = 0;
  void upload(const Ctx& c);
};

// Device copies of one accessor chunk (staged from host when needed).
struct ChunkDev {
  int nrow = 0, nchan = 0, npol = 0;
  const double* uvw = nullptr;
  const double* delay = nullptr;
  const double* freq = nullptr;
  const float2* vis = nullptr;
  float2* vis_out = nullptr;
  const float2* noise = nullptr;
  const uint8_t* flag = nullptr;
  const int32_t* feed1 = nullptr;
  const double* pointing1 = nullptr;  // nrow x 2 (lon, lat) or null
  const double* dish1 = nullptr;      // nrow x 2 or null
  int vis_out_ld = 0;                 // rows of the cube vis_out points into
  std::vector<double> freq_host;
  int feed0 = 0;
  // host copies of per-row metadata for gridders whose CF cache depends on it (AWProject)
  std::vector<int32_t> feed_h;
  std::vector<double> point_h;  // nrow x 2, empty: every row points at the image centre
  std::vector<float> pa_h;      // nrow parallactic angles (rad), empty: all 0
};

struct SupportResult {
  int size, offu, offv;
};
// SupportSearcher + WProjectVisGridder::extractSupport on one device plane (nx x ny).
class Gridder;
// ImageFFTEquation::calcImagingEquations' data loop body for one chunk (ImageFFTEquation.cc:437-480).
void imaging_chunk(const std::vector<Gridder*>& degridders, const std::vector<Gridder*>& gridders,
                   const ygrid_chunk& a);

// A device copy of a host chunk's input arrays (vis_out is not copied: NULL in the copy); the
// arrays of a device chunk are used as they are.
struct DeviceChunk {
  ygrid_chunk c{};
  DevBuf store;
  DeviceChunk(const Ctx& ctx, const ygrid_chunk& a);
};

// Polarisation conversion matrix out = M in between two lists of casacore Stokes codes
// (PolConverter semantics, polconv.cpp); row-major nout x nin.  determined: throw when an output
// polarisation cannot be formed from the inputs (imaging); else the Stokes parameters the inputs
// do not hold are taken as zero (predicting).
std::vector<std::complex<double>> pol_matrix(const std::vector<int>& from, const std::vector<int>& to,
                                             bool determined = true);

SupportResult search_support(const Ctx& c, const float2* plane, int nx, int ny, double cutoff, bool absolute,
                             bool offset, int os);

class Gridder {
 public:
  explicit Gridder(std::shared_ptr<Ctx> ctx) : ctx_(std::move(ctx)) {}
  virtual ~Gridder();
  virtual std::unique_ptr<Gridder> clone() const = 0;
  virtual std::string name() const = 0;

  void configure_common(const Params& p);
  // the options VisGridderFactory::make sets on a created gridder (VisGridderFactory.cc:137-232)
  void configure_globals(const Params& p);
  void initialiseGrid(const ygrid_image_desc& img, bool dopsf, bool dopcf);
  void grid(const ygrid_chunk& acc);
  void finaliseGrid(float* out, int mem);
  virtual void finaliseWeights(float* out, int mem);
  void initialiseDegrid(const ygrid_image_desc& img, const float* model, int mem);
  void degrid(const ygrid_chunk& acc);
  // degrid() with every device buffer of the pass carved from `bytes` of caller-held device
  // memory at `scratch` (no allocation; the CF cache must already exist)
  void degridNoAlloc(const ygrid_chunk& acc, void* scratch, size_t bytes);
  void finaliseDegrid();
  bool isModelEmpty() const { return model_empty_; }
  bool initialisedFor(bool degrid) const { return initialised_ && degrid_mode_ == degrid; }
  void setMFSWeight(double ref, int order) { vw_ref_ = ref; vw_order_ = order; }
  void setProfiling(bool on);
  void stats(ygrid_stats* out) const;
  void getCF(int k, float* out, int* support) const;
  void getGrid(float* out, size_t* n) const;
  const CFTable* cfTable() const { return cf_.get(); }
  const std::shared_ptr<Ctx>& context() const { return ctx_; }
  // W-projection: the w (wavelengths) of every CF plane and wmax; false for other gridders
  virtual bool wTerms(std::vector<double>& w, double& wmax) const {
    (void)w;
    (void)wmax;
    return false;
  }
  // Learned CFs (W-projection only): the network gives every sample its CF at its exact w while
  // its held-out error is <= max_err; the table otherwise (and with lcf null).
  virtual void setLearnedCF(std::shared_ptr<LearnedCF> lcf, double max_err) {
    (void)lcf;
    (void)max_err;
    throw Error(1, "learned CFs need a W-projection gridder");
  }
  virtual bool learnedActive() const { return false; }

 protected:
  Gridder(const Gridder& o);
  // TableVisGridder::initConvolutionFunction: build cf_ if not built yet.
  virtual void initConvolutionFunction(const ChunkDev& acc) = 0;
  // TableVisGridder::cIndex as a device-side descriptor.
  virtual CIndexDesc cindex(const ChunkDev& acc) const { (void)acc; return CIndexDesc{}; }
  // number of CF planes before oversampling (rows of the sum of weights)
  virtual int nZPlanes() const { return 1; }
  // grids per image plane (w stacking: one per w plane; the tile kernels select it by the
  // sample's w plane) and the transforms between them and the image
  virtual int nGridPlanes() const { return 1; }
  virtual void gridsToImage(const double* ccfx, const double* ccfy, float* out);
  virtual void modelToGrids(const float* model, const double* ccfx, const double* ccfy);
  size_t gridPlaneElems() const { return (size_t)shape_[0] * shape_[1] * shape_[2] * shape_[3]; }
  // itsOverSample as configured (the CF cache may not exist yet)
  virtual int oversampling() const { return 1; }
  // 1-D factors of correctConvolution (empty: no correction, e.g. Box)
  virtual void gridCorrection(std::vector<double>& cx, std::vector<double>& cy) const;
  virtual bool doesCorrection() const { return true; }
  // whether initConvolutionFunction needs host copies of feed1/pointing1
  virtual bool needsRowMeta() const { return false; }
  // whether a chunk can need CF planes the cache does not hold yet (A-projection: new feeds or
  // fields); such gridders build CFs inside their passes
  virtual bool cfPerChunk() const { return false; }

  void initialiseCommon(const ygrid_image_desc& img);
  void generic(const ygrid_chunk& acc, bool forward);
  // polarisation conversion around generic (polconv.cpp): true when the chunk's Stokes codes
  // differ from the image's
  bool needsPolConversion(const ygrid_chunk& acc) const;
  void convertedPass(const ygrid_chunk& acc, bool forward);
  friend class Batch;
  struct PassSetup {
    ChunkDev d;
    PrepParams pp{};
    PrepInputs in{};
    PrepOutputs out{};
    CFView cfv{};
    int nsamp = 0;
    float* acc = nullptr;  // degridding: per-sample accumulators (nsamp x 8)
    bool small = false;    // the CF support allows the small-support kernels
  };
  // arena: carve the pass's buffers from it (the allocation-free entry point); null: the
  // gridder's own buffers, grown as needed
  bool setup_pass(const ygrid_chunk& acc, bool forward, PassSetup& ps, PassArena* arena = nullptr);
  // cnt: samples, taps, flagged, w-flagged, rows rejected by pointing
  void account_pass(bool forward, const unsigned long long cnt[5], int err, double seconds);
  // hooks of a pass: after its setup (may replace the CF view and prep arguments), and after the
  // prep kernel (device work only: also runs inside captured graphs)
  virtual void customizePass(PassSetup& ps) { (void)ps; }
  virtual void afterPrep(PassSetup& ps) { (void)ps; }
 public:
  bool degrid_mode() const { return degrid_mode_; }
  int npol() const { return shape_[2]; }
 protected:
  void stage(const ygrid_chunk& acc, ChunkDev& d, bool forward);
  // the representative feed and field of a PSF pass: those of the first row the pointing test
  // keeps (TableVisGridder.cc:571-593); false when the chunk has no such row
  bool choosePSFField(const ygrid_chunk& acc);
  void uploadCorrection();

  std::shared_ptr<Ctx> ctx_;
  std::shared_ptr<CFTable> cf_;
  int shape_[4] = {0, 0, 1, 1};  // padded nx, ny, npol, nchan (itsShape)
  int img_[4] = {0, 0, 1, 1};    // unpadded image
  double uvcell_[2] = {0, 0};    // itsUVCellSize
  float padding_ = 1.f;
  bool dopsf_ = false, dopcf_ = false;
  bool model_empty_ = false;
  bool track_per_os_ = false, use_all_psf_ = false, clear_grid_ = false;
  bool first_psf_ = true;
  int psf_feed_ = -1;
  int psf_field_ = 0;           // the representative field's pointing is known (psf_dir_)
  double psf_dir_[2] = {0, 0};  // itsPointingUsedForPSF (dish pointing of the first used row)
  double max_sep_ = -1.0;       // itsMaxPointingSeparation (radians; <= 0: every row is used)
  double st_rejected_ = 0;      // itsRowsRejectedDueToMaxPointingSeparation
  bool has_freq_axis_ = false;
  double fstart_ = 0, fend_ = 0;
  double centre_[2] = {0, 0};  // image centre (lon, lat)
  std::vector<int> img_stokes_;  // Stokes codes of the image polarisations (empty: the data's frame)
  double vw_ref_ = 0;
  int vw_order_ = -1;
  int nsw_rows_ = 1;
  int alpha_int_ = 1;
  bool initialised_ = false, degrid_mode_ = false;

  DevBuf grid_, sumw_, scratch_, ccfx_, ccfy_;
  DevBuf smp_hdr_, smp_key_, smp_coef_, acc_, counters_, err_, chanmap_, visweight_, staging_;
  // the CF table at the configured storage precision (building the 16-bit copy if needed)
  CFView cf_view();
  int cf_prec_ = kPrecF32, vis_prec_ = kPrecF32;  // storage precisions (cfprecision, visprecision)
  DevBuf smp_coef16_;
  DevBuf gscale_;  // degridding, FP16 tables: max |model grid| bits and its power-of-two scale
  double n_bin_reused_ = 0;  // passes that reused the context's cached binning
  double n_small_ = 0;       // passes that ran the small-support kernels
  // A-projection CF cache statistics (itsNumberOfIterations, itsNumberOfCFGenerations[DueToPA/Freq])
  double cf_iter_ = 0, cf_gen_ = 0, cf_gen_pa_ = 0, cf_gen_freq_ = 0;
  TileJob noalloc_job_{};    // host copy of the allocation-free pass's job (read by its H2D copy)
  PlanPtr plan_;
  double st_samples_g_ = 0, st_taps_g_ = 0, st_samples_d_ = 0, st_taps_d_ = 0, st_flag_ = 0, st_wflag_ = 0;
  double t_grid_ = 0, t_degrid_ = 0, t_cf_ = 0;
  // profiling events: 0 start, 1 after prep, 2 after binning, 3 after the main kernel
  bool profiling_ = false;
  hipEvent_t ev_[4] = {nullptr, nullptr, nullptr, nullptr};
  double ms_prep_ = 0, ms_bin_ = 0, ms_gk_ = 0, ms_dk_ = 0, n_gk_ = 0, n_dk_ = 0;
};

std::unique_ptr<Gridder> make_gridder(std::shared_ptr<Ctx> ctx, const std::string& name, const ParamMap& parset);
// "fp32" | "fp16" | "bf16" (case-insensitive; also float32, float16, half, bfloat16) -> Prec
int parse_precision(const std::string& v, const char* key);

}  // namespace ygrid
