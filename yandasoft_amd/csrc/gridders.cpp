// Gridder classes: TableVisGridder's host logic (askap/gridding/TableVisGridder.cc) around the
// HIP kernels, plus the Box, SphFunc and WProject convolution functions.
#include "gridders.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <sstream>

namespace ygrid {

// helpers declared in the kernel units
void launch_gather_cross(const Ctx& c, const float2* plane, int nx, int ny, int px, int py, float2* out);


// ---------------------------------------------------------------------------------------
// parset
// ---------------------------------------------------------------------------------------
static std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

ParamMap parse_parset(const char* text) {
  ParamMap m;
  if (!text) return m;
  std::string all(text), line;
  for (char& ch : all)
    if (ch == ';') ch = '\n';
  std::istringstream is(all);
  while (std::getline(is, line)) {
    line = trim(line);
    if (line.empty() || line[0] == '#') continue;
    const size_t eq = line.find('=');
    YG_CHECK(eq != std::string::npos, 1, "parset line without '=': " + line);
    m[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
  }
  return m;
}

double quantity(const std::string& s, const std::string& unit) {
  size_t i = 0;
  double v = 0;
  try { v = std::stod(s, &i); } catch (...) { throw Error(1, "bad quantity " + s); }
  std::string u = s.substr(i);
  u.erase(0, u.find_first_not_of(" \t"));
  if (u.empty() || u == unit) return v;
  if (unit == "rad") {
    if (u == "deg") return v * kPi / 180.;
    if (u == "arcmin") return v * kPi / 180. / 60.;
    if (u == "arcsec") return v * kPi / 180. / 3600.;
  }
  if (unit == "m") {
    if (u == "cm") return v * 0.01;
    if (u == "mm") return v * 1e-3;
    if (u == "km") return v * 1e3;
  }
  throw Error(1, "cannot convert " + s + " to " + unit);
}

double separation(double l1, double b1, double l2, double b2) {
  const double dl = l2 - l1;
  const double a = std::cos(b2) * std::sin(dl);
  const double b = std::cos(b1) * std::sin(b2) - std::sin(b1) * std::cos(b2) * std::cos(dl);
  const double c = std::sin(b1) * std::sin(b2) + std::cos(b1) * std::cos(b2) * std::cos(dl);
  return std::atan2(std::sqrt(a * a + b * b), c);
}

std::string Params::str(const std::string& k, const std::string& d) const {
  auto it = m.find(k);
  return it == m.end() ? d : it->second;
}
int Params::i32(const std::string& k, int d) const {
  auto it = m.find(k);
  if (it == m.end()) return d;
  try { return std::stoi(it->second); } catch (...) { throw Error(1, "bad integer for " + k); }
}
double Params::f64(const std::string& k, double d) const {
  auto it = m.find(k);
  if (it == m.end()) return d;
  try { return std::stod(it->second); } catch (...) { throw Error(1, "bad number for " + k); }
}
bool Params::flag(const std::string& k, bool d) const {
  auto it = m.find(k);
  if (it == m.end()) return d;
  std::string v = it->second;
  std::transform(v.begin(), v.end(), v.begin(), ::tolower);
  if (v == "true" || v == "1" || v == "yes" || v == "t") return true;
  if (v == "false" || v == "0" || v == "no" || v == "f") return false;
  throw Error(1, "bad boolean for " + k);
}

void CFTable::upload(const Ctx& c) {
  d_off.resize(off.size() * sizeof(int32_t));
  d_supp.resize(supp.size() * sizeof(int32_t));
  YG_HIP(hipMemcpyAsync(d_off.get(), off.data(), off.size() * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
  YG_HIP(hipMemcpyAsync(d_supp.get(), supp.data(), supp.size() * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
  if (!ofs_u.empty()) {
    d_ofs_u.resize(ofs_u.size() * sizeof(int32_t));
    d_ofs_v.resize(ofs_v.size() * sizeof(int32_t));
    YG_HIP(hipMemcpyAsync(d_ofs_u.get(), ofs_u.data(), ofs_u.size() * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
    YG_HIP(hipMemcpyAsync(d_ofs_v.get(), ofs_v.data(), ofs_v.size() * sizeof(int32_t), hipMemcpyHostToDevice, c.stream));
  }
  YG_HIP(hipStreamSynchronize(c.stream));
}

// Layout helper: offsets of the padded planes for per-z supports.
static void layout_table(CFTable& t, const std::vector<int>& supp_z, int os) {
  t.nz = (int)supp_z.size();
  t.oversample = os;
  t.off.assign((size_t)t.nz * os * os, 0);
  t.supp.assign((size_t)t.nz * os * os, 0);
  size_t total = 0;
  t.support = 0;
  for (int z = 0; z < t.nz; ++z) {
    const int S = supp_z[z];
    t.support = std::max(t.support, S);
    const size_t blk = (size_t)cf_pitch(S) * cf_rows(S);
    for (int f = 0; f < os * os; ++f) {
      YG_CHECK(total < (size_t)INT_MAX, 1, "convolution function table exceeds 2^31 elements");
      t.off[(size_t)z * os * os + f] = (int32_t)total;
      t.supp[(size_t)z * os * os + f] = S;
      total += blk;
    }
  }
  t.n_elems = total;
}

// ---------------------------------------------------------------------------------------
// Gridder (TableVisGridder)
// ---------------------------------------------------------------------------------------
Gridder::Gridder(const Gridder& o)
    : ctx_(o.ctx_), cf_(o.cf_), padding_(o.padding_), track_per_os_(o.track_per_os_), use_all_psf_(o.use_all_psf_),
      clear_grid_(o.clear_grid_), max_sep_(o.max_sep_), vw_ref_(o.vw_ref_), vw_order_(o.vw_order_),
      alpha_int_(o.alpha_int_), cf_prec_(o.cf_prec_), vis_prec_(o.vis_prec_) {
  std::memcpy(shape_, o.shape_, sizeof(shape_));
  std::memcpy(img_, o.img_, sizeof(img_));
  uvcell_[0] = o.uvcell_[0];
  uvcell_[1] = o.uvcell_[1];
}

Gridder::~Gridder() {
  for (auto& e : ev_)
    if (e) (void)hipEventDestroy(e);
}

void Gridder::setProfiling(bool on) {
  profiling_ = on;
  if (on)
    for (auto& e : ev_)
      if (!e) YG_HIP(hipEventCreate(&e));
}

// Global gridder options (VisGridderFactory.cc:137-232); keys a parset does not give keep their
// values.
void Gridder::configure_globals(const Params& p) {
  padding_ = (float)p.f64("padding", padding_);
  YG_CHECK(padding_ >= 1.f, 1, "padding must be >= 1");
  use_all_psf_ = p.flag("alldatapsf", use_all_psf_);
  track_per_os_ = p.flag("oversampleweight", track_per_os_);
  clear_grid_ = p.flag("cleargrids", clear_grid_);
  if (p.has("MaxPointingSeparation")) max_sep_ = quantity(p.str("MaxPointingSeparation", "-1rad"), "rad");
  YG_CHECK(!p.flag("parotation", false), 1,
           "parallactic angle rotation (gridder.parotation) is not supported by the GPU gridders");
}

void Gridder::configure_common(const Params& p) {
  configure_globals(p);
  // Storage precision of the CF taps and of the gridding coefficients (weighted visibilities);
  // no counterpart in the reference, whose tables and accessors are FP32 throughout.
  cf_prec_ = parse_precision(p.str("cfprecision", "fp32"), "cfprecision");
  vis_prec_ = parse_precision(p.str("visprecision", "fp32"), "visprecision");
}

int parse_precision(const std::string& v, const char* key) {
  std::string s;
  for (char ch : v) s += (char)std::tolower((unsigned char)ch);
  if (s == "fp32" || s == "float32" || s == "float") return kPrecF32;
  if (s == "fp16" || s == "float16" || s == "half") return kPrecF16;
  if (s == "bf16" || s == "bfloat16") return kPrecBF16;
  throw Error(1, std::string(key) + " must be fp32, fp16 or bf16, not " + v);
}

CFView Gridder::cf_view() {
  const int smax = cf_->supp.empty() ? 0 : *std::max_element(cf_->supp.begin(), cf_->supp.end());
  if (cf_prec_ == kPrecF32) return CFView{cf_->table.get(), cf_->table.size(), smax, kPrecF32};
  if (cf_->table16_prec != cf_prec_ || cf_->table16.size() != cf_->n_elems * sizeof(uint32_t)) {
    cf_->table16.resize(cf_->n_elems * sizeof(uint32_t));
    launch_cf_round(*ctx_, cf_->table.as<float2>(), cf_->table16.as<uint32_t>(), cf_->n_elems, cf_prec_);
    cf_->table16_prec = cf_prec_;
  }
  CFView v{cf_->table16.get(), cf_->table16.size(), smax, cf_prec_, vis_prec_ == cf_prec_ ? 1 : 0};
  if (degrid_mode_ && cf_prec_ == kPrecF16 && gscale_.get()) v.gscale = gscale_.as<float>() + 1;
  return v;
}

void Gridder::initialiseCommon(const ygrid_image_desc& img) {
  YG_CHECK(img.nx > 0 && img.ny > 0 && img.npol > 0 && img.nchan > 0, 1, "bad image shape");
  YG_CHECK(img.npol <= 4, 1, "at most 4 image polarisations are supported");
  YG_CHECK(img.cellx != 0 && img.celly != 0, 1, "zero cell size");
  img_[0] = img.nx; img_[1] = img.ny; img_[2] = img.npol; img_[3] = img.nchan;
  // PaddingUtils::paddedShape: both direction axes scaled by the padding factor
  shape_[0] = (int)(img.nx * padding_);
  shape_[1] = (int)(img.ny * padding_);
  shape_[2] = img.npol;
  shape_[3] = img.nchan;
  YG_CHECK(shape_[0] % 2 == 0 && shape_[1] % 2 == 0, 1, "padded grid sizes must be even");
  // TableVisGridder::initialiseCellSize (TableVisGridder.cc:1038-1049)
  uvcell_[0] = 1. / (std::fabs(img.cellx) * double(shape_[0]));
  uvcell_[1] = 1. / (std::fabs(img.celly) * double(shape_[1]));
  has_freq_axis_ = img.has_freq_axis != 0;
  fstart_ = img.freq_start;
  fend_ = img.freq_end;
  centre_[0] = img.centre_lon;
  centre_[1] = img.centre_lat;
  img_stokes_.clear();
  if (img.stokes) {
    img_stokes_.assign(img.stokes, img.stokes + img.npol);
    pol_matrix(img_stokes_, img_stokes_);  // the codes must be known and independent
  }
  initialised_ = true;
}

void Gridder::gridCorrection(std::vector<double>& cx, std::vector<double>& cy) const {
  cx.clear();
  cy.clear();
}

void Gridder::uploadCorrection() {
  std::vector<double> cx, cy;
  gridCorrection(cx, cy);
  if (cx.empty()) {
    ccfx_.release();
    ccfy_.release();
    return;
  }
  ccfx_.resize(cx.size() * sizeof(double));
  ccfy_.resize(cy.size() * sizeof(double));
  YG_HIP(hipMemcpyAsync(ccfx_.get(), cx.data(), cx.size() * sizeof(double), hipMemcpyHostToDevice, ctx_->stream));
  YG_HIP(hipMemcpyAsync(ccfy_.get(), cy.data(), cy.size() * sizeof(double), hipMemcpyHostToDevice, ctx_->stream));
  YG_HIP(hipStreamSynchronize(ctx_->stream));
}

// TableVisGridder::initialiseGrid (TableVisGridder.cc:997-1030)
void Gridder::initialiseGrid(const ygrid_image_desc& img, bool dopsf, bool dopcf) {
  initialiseCommon(img);
  YG_CHECK(!(dopsf && dopcf), 1, "a gridder is either a PSF or a PCF gridder");
  dopsf_ = dopsf;
  dopcf_ = dopcf;
  degrid_mode_ = false;
  const size_t n = gridPlaneElems() * nGridPlanes();
  grid_.resize(n * sizeof(float2));
  YG_HIP(hipMemsetAsync(grid_.get(), 0, n * sizeof(float2), ctx_->stream));
  first_psf_ = true;
  psf_feed_ = -1;
  // sum of weights (resizeSumOfWeights, TableVisGridder.cc:1067-1079)
  const int os = oversampling();
  nsw_rows_ = nZPlanes() * (track_per_os_ ? os * os : 1);
  const size_t nsw = (size_t)nsw_rows_ * shape_[2] * shape_[3];
  sumw_.resize(nsw * sizeof(double));
  YG_HIP(hipMemsetAsync(sumw_.get(), 0, nsw * sizeof(double), ctx_->stream));
  YG_HIP(hipStreamSynchronize(ctx_->stream));
}

void Gridder::stage(const ygrid_chunk& a, ChunkDev& d, bool forward) {
  YG_CHECK(a.nrow >= 0 && a.nchan > 0 && a.npol > 0, 1, "bad chunk shape");
  YG_CHECK(a.uvw && a.freq, 1, "chunk needs uvw and frequencies");
  YG_CHECK(forward ? (a.vis_out != nullptr) : (a.vis != nullptr || dopsf_ || dopcf_), 1,
           forward ? "degridding needs vis_out" : "gridding needs visibilities");
  d.nrow = a.nrow;
  d.nchan = a.nchan;
  d.npol = a.npol;
  const size_t nr = a.nrow, nc = a.nchan, np = a.npol;
  const size_t ncube = nr * nc * np;
  d.freq_host.resize(nc);
  if (needsRowMeta()) {
    d.feed_h.assign(nr, 0);
    d.point_h.clear();
    const hipMemcpyKind k = a.mem == YGRID_MEM_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
    if (a.feed1 && nr) YG_HIP(hipMemcpy(d.feed_h.data(), a.feed1, nr * sizeof(int32_t), k));
    if (a.pointing1 && nr) {
      d.point_h.resize(nr * 2);
      YG_HIP(hipMemcpy(d.point_h.data(), a.pointing1, nr * 2 * sizeof(double), k));
    }
    d.pa_h.clear();
    if (a.feed1_pa && nr) {
      d.pa_h.resize(nr);
      YG_HIP(hipMemcpy(d.pa_h.data(), a.feed1_pa, nr * sizeof(float), k));
    }
  }
  d.vis_out_ld = a.vis_out_rows > 0 ? a.vis_out_rows : a.nrow;
  YG_CHECK(d.vis_out_ld >= a.nrow, 1, "vis_out_rows must be 0 or at least nrow");
  if (a.mem == YGRID_MEM_DEVICE) {
    d.uvw = a.uvw;
    d.delay = a.delay;
    d.freq = a.freq;
    d.vis = reinterpret_cast<const float2*>(a.vis);
    d.vis_out = reinterpret_cast<float2*>(a.vis_out);
    d.noise = reinterpret_cast<const float2*>(a.noise);
    d.flag = a.flag;
    d.feed1 = a.feed1;
    d.pointing1 = a.pointing1;
    d.dish1 = a.dish_pointing1;
    YG_HIP(hipMemcpyAsync(d.freq_host.data(), a.freq, nc * sizeof(double), hipMemcpyDeviceToHost, ctx_->stream));
    d.feed0 = 0;
    if (a.feed1 && a.nrow > 0)
      YG_HIP(hipMemcpyAsync(&d.feed0, a.feed1, sizeof(int32_t), hipMemcpyDeviceToHost, ctx_->stream));
    YG_HIP(hipStreamSynchronize(ctx_->stream));
    return;
  }
  YG_CHECK(a.mem == YGRID_MEM_HOST, 1, "unknown memory kind");
  std::memcpy(d.freq_host.data(), a.freq, nc * sizeof(double));
  d.feed0 = (a.feed1 && a.nrow > 0) ? a.feed1[0] : 0;
  // one staging allocation, 256-byte aligned sub-buffers
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const bool want_vis = !forward && a.vis && !(dopsf_ || dopcf_);
  // the pointings are read by the pointing test and by a PSF pass's field test
  const bool want_point = max_sep_ > 0 || (dopsf_ && !use_all_psf_);
  // degridding: the span of the (ld-row) cube this chunk's rows occupy
  const size_t vspan = ncube ? nr + (size_t)d.vis_out_ld * (nc * np - 1) : 0;
  size_t bytes = al(nr * 3 * 8) + al(nr * 8) + al(nc * 8) + (want_vis ? al(ncube * 8) : 0) + (forward ? al(vspan * 8) : 0) +
                 (a.noise ? al(ncube * 8) : 0) + (a.flag ? al(ncube) : 0) + (a.feed1 ? al(nr * 4) : 0) +
                 (want_point && a.pointing1 ? al(nr * 16) : 0) + (want_point && a.dish_pointing1 ? al(nr * 16) : 0);
  staging_.resize(bytes);
  char* p = staging_.as<char>();
  auto put = [&](const void* src, size_t b) -> void* {
    void* dst = p;
    if (b) YG_HIP(hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, ctx_->stream));
    p += al(b);
    return dst;
  };
  d.uvw = static_cast<const double*>(put(a.uvw, nr * 3 * 8));
  d.delay = a.delay ? static_cast<const double*>(put(a.delay, nr * 8)) : nullptr;
  if (!a.delay) p += al(nr * 8);
  d.freq = static_cast<const double*>(put(a.freq, nc * 8));
  d.vis = want_vis ? static_cast<const float2*>(put(a.vis, ncube * 8)) : nullptr;
  d.vis_out = forward ? static_cast<float2*>(put(a.vis_out, vspan * 8)) : nullptr;
  d.noise = a.noise ? static_cast<const float2*>(put(a.noise, ncube * 8)) : nullptr;
  d.flag = a.flag ? static_cast<const uint8_t*>(put(a.flag, ncube)) : nullptr;
  d.feed1 = a.feed1 ? static_cast<const int32_t*>(put(a.feed1, nr * 4)) : nullptr;
  d.pointing1 = want_point && a.pointing1 ? static_cast<const double*>(put(a.pointing1, nr * 16)) : nullptr;
  d.dish1 = want_point && a.dish_pointing1 ? static_cast<const double*>(put(a.dish_pointing1, nr * 16)) : nullptr;
}

bool Gridder::choosePSFField(const ygrid_chunk& a) {
  const size_t nr = a.nrow;
  if (nr == 0) return false;
  const hipMemcpyKind k = a.mem == YGRID_MEM_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost;
  std::vector<double> point;
  if (a.pointing1) {
    point.resize(nr * 2);
    YG_HIP(hipMemcpy(point.data(), a.pointing1, nr * 2 * sizeof(double), k));
  }
  size_t i0 = 0;
  if (max_sep_ > 0 && !point.empty()) {
    while (i0 < nr && separation(centre_[0], centre_[1], point[2 * i0], point[2 * i0 + 1]) > max_sep_) ++i0;
    if (i0 == nr) return false;
  }
  int32_t feed = 0;
  if (a.feed1) YG_HIP(hipMemcpy(&feed, a.feed1 + i0, sizeof(int32_t), k));
  psf_feed_ = feed;
  psf_field_ = 0;
  if (a.dish_pointing1) {
    YG_HIP(hipMemcpy(psf_dir_, a.dish_pointing1 + 2 * i0, 2 * sizeof(double), k));
    psf_field_ = 1;
  } else if (!point.empty()) {
    psf_dir_[0] = point[2 * i0];
    psf_dir_[1] = point[2 * i0 + 1];
    psf_field_ = 1;
  }
  return true;
}

// TableVisGridder::generic (TableVisGridder.cc:451-884)
// The part of TableVisGridder::generic before the prep kernel: stage the chunk, channel map,
// MFS weights, CF generation, buffer sizes, and the prep kernel's arguments.  Returns false
// when the pass has nothing to do (an empty model when degridding, no rows).
// A buffer of the pass: from the arena when one is given, else the gridder's own (grown).
static void* pass_buf(PassArena* arena, DevBuf& own, size_t bytes) {
  if (arena) return arena->take(bytes);
  own.resize(bytes);
  return own.get();
}

bool Gridder::setup_pass(const ygrid_chunk& a, bool forward, PassSetup& ps, PassArena* arena) {
  YG_CHECK(initialised_, 2, "gridder not initialised");
  YG_CHECK(forward == degrid_mode_, 2,
           forward ? "degrid called on a gridder initialised for gridding" : "grid called on a degridder");
  if (forward && model_empty_) return false;
  ChunkDev& d = ps.d;
  stage(a, d, forward);
  YG_CHECK(d.npol == shape_[2], 1,
           "chunk npol must equal image npol unless both polarisation frames are given (stokes)");
  if (d.nrow == 0) return false;
  const double r0 = d.freq_host[0] / kC;
  YG_CHECK(r0 > 0.1 && r0 < 30000, 4,
           "Check frequencies in the input file as the order of magnitude is likely to be wrong");
  // frequency mapping (FrequencyMapper::setupMapping, FrequencyMapper.cc:63-97)
  std::vector<int32_t> cmap(d.nchan, 0);
  if (has_freq_axis_ && shape_[3] > 1) {
    const double inc = (fend_ - fstart_) / double(shape_[3] - 1);
    YG_CHECK(std::fabs(inc) > 0, 1, "Frequency axis in the image has the same start and end frequency");
    for (int c = 0; c < d.nchan; ++c) {
      double x = (d.freq_host[c] - fstart_) / inc;
      x += (x < 0.) ? -0.5 : 0.5;
      const int ic = int(x);
      cmap[c] = (ic < 0 || ic >= shape_[3]) ? -1 : ic;
    }
  }
  int32_t* dchan = static_cast<int32_t*>(pass_buf(arena, chanmap_, cmap.size() * sizeof(int32_t)));
  YG_HIP(hipMemcpyAsync(dchan, cmap.data(), cmap.size() * 4, hipMemcpyHostToDevice, ctx_->stream));
  // MFS visibility weights (VisWeightsMultiFrequency::getWeight, VisWeightsMultiFrequency.cc:59-75)
  const float* vw = nullptr;
  if (vw_order_ >= 0 && !dopcf_) {
    std::vector<float> w(d.nchan);
    for (int c = 0; c < d.nchan; ++c) {
      if (vw_order_ == 0) { w[c] = 1.f; continue; }
      const double ratio = (d.freq_host[c] - vw_ref_) / vw_ref_;
      w[c] = (float)(vw_order_ == 1 ? ratio : (vw_order_ == 2 ? ratio * ratio : std::pow(ratio, vw_order_)));
    }
    float* dw = static_cast<float*>(pass_buf(arena, visweight_, w.size() * sizeof(float)));
    YG_HIP(hipMemcpyAsync(dw, w.data(), w.size() * 4, hipMemcpyHostToDevice, ctx_->stream));
    vw = dw;
  }
  // the copies read host vectors of this frame: complete them before it returns
  if (arena) YG_HIP(hipStreamSynchronize(ctx_->stream));
  if (!forward && dopsf_ && first_psf_) {
    // the first row the pointing test keeps fixes the representative feed and field
    // (TableVisGridder.cc:581-593), once per initialiseGrid
    if (use_all_psf_) {
      psf_feed_ = -1;
      first_psf_ = false;
    } else if (choosePSFField(a)) {
      first_psf_ = false;
    }
  }
  const auto tcf = std::chrono::steady_clock::now();
  // with an arena the caller has checked that the CF cache exists and holds every plane
  if (!arena) initConvolutionFunction(d);
  t_cf_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - tcf).count();
  YG_CHECK(cf_ && cf_->n_elems > 0, 2, "convolution function not initialised");
  if (!forward) {
    // the sum-of-weights cube follows the CF cache (resizeSumOfWeights when the cache is
    // first built, AWProjectVisGridder.cc:274-276)
    const int rows = nZPlanes() * (track_per_os_ ? cf_->oversample * cf_->oversample : 1);
    YG_CHECK(!arena || rows == nsw_rows_, 2, "the sum-of-weights cube must exist before an allocation-free pass");
    if (rows != nsw_rows_) {
      nsw_rows_ = rows;
      const size_t nsw = (size_t)nsw_rows_ * shape_[2] * shape_[3];
      sumw_.release();
      sumw_.resize(nsw * sizeof(double));
      YG_HIP(hipMemsetAsync(sumw_.get(), 0, nsw * sizeof(double), ctx_->stream));
    }
  }

  const int nsamp = d.nrow * d.nchan;
  Samples smp{};
  smp.hdr = static_cast<SampleHdr*>(pass_buf(arena, smp_hdr_, (size_t)nsamp * sizeof(SampleHdr)));
  smp.key = static_cast<int32_t*>(pass_buf(arena, smp_key_, (size_t)nsamp * sizeof(int32_t)));
  smp.coef = static_cast<float*>(pass_buf(arena, smp_coef_, (size_t)nsamp * kCoef * sizeof(float)));
  smp.coef16 = vis_prec_ != kPrecF32
                   ? static_cast<uint16_t*>(pass_buf(arena, smp_coef16_, (size_t)nsamp * kCoef * sizeof(uint16_t)))
                   : nullptr;
  smp.coef_prec = vis_prec_;
  auto* counters = static_cast<unsigned long long*>(pass_buf(arena, counters_, 8 * sizeof(unsigned long long)));
  int* errp = static_cast<int*>(pass_buf(arena, err_, sizeof(int) * 2));
  ps.acc = forward ? static_cast<float*>(pass_buf(arena, acc_, (size_t)nsamp * 8 * sizeof(float))) : nullptr;

  PrepParams& pp = ps.pp;
  pp.nrow = d.nrow;
  pp.nchan = d.nchan;
  pp.npol = d.npol;
  pp.nx = shape_[0];
  pp.ny = shape_[1];
  pp.nchan_img = shape_[3];
  pp.uvcell_u = uvcell_[0];
  pp.uvcell_v = uvcell_[1];
  pp.oversample = cf_->oversample;
  pp.mode = forward ? kModeDegrid : (dopsf_ ? kModePSF : (dopcf_ ? kModePCF : kModeGrid));
  pp.track_per_os = track_per_os_ ? 1 : 0;
  pp.nsw_rows = nsw_rows_;
  pp.psf_feed = psf_feed_;
  pp.psf_field = psf_field_;
  pp.psf_lon = psf_dir_[0];
  pp.psf_lat = psf_dir_[1];
  pp.max_sep = max_sep_;
  pp.centre_lon = centre_[0];
  pp.centre_lat = centre_[1];
  pp.ci = cindex(d);
  const int ntiles = ((shape_[0] + kTile - 1) / kTile) * ((shape_[1] + kTile - 1) / kTile);
  const long long budget = 1ll << 24;
  const int nz = std::max(1, cf_->nz);
  long long zdiv = ((long long)nz * ntiles * shape_[3] + budget - 1) / budget;
  if (zdiv < 1) zdiv = 1;
  pp.zdiv = (int)zdiv;
  pp.nzg = (int)((nz + zdiv - 1) / zdiv);
  if (nGridPlanes() > 1) {
    // w stacking: the key names the grid, one group per w plane
    YG_CHECK((long long)nGridPlanes() * ntiles * shape_[3] <= budget, 1, "too many w planes x tiles x channels");
    pp.zdiv = 1;
    pp.nzg = nGridPlanes();
  }
  pp.vis_prec = vis_prec_;

  PrepInputs& in = ps.in;
  in.uvw = d.uvw;
  in.delay = d.delay;
  in.freq = d.freq;
  in.vis = d.vis;
  in.noise = d.noise;
  in.flag = d.flag;
  in.feed1 = d.feed1;
  in.pointing1 = d.pointing1;
  in.dish1 = d.dish1;
  in.chanmap = dchan;
  in.visweight = vw;
  in.cf_off = cf_->d_off.as<int32_t>();
  in.cf_supp = cf_->d_supp.as<int32_t>();
  in.cf_ofs_u = cf_->ofs_u.empty() ? nullptr : cf_->d_ofs_u.as<int32_t>();
  in.cf_ofs_v = cf_->ofs_v.empty() ? nullptr : cf_->d_ofs_v.as<int32_t>();
  in.n_cf = (int)cf_->off.size();
  PrepOutputs& out = ps.out;
  out.smp = smp;
  out.sumw = forward ? nullptr : sumw_.as<double>();
  out.counters = counters;
  out.error = errp;
  ps.nsamp = nsamp;
  ps.cfv = cf_view();
  if (!arena) customizePass(ps);
  // small supports: the small-support kernels, and one CF-plane group, so that a binning key
  // names its own region of the grid (the small-support gridding kernel owns a key's region;
  // with w stacking the key already names a grid)
  ps.small = ps.cfv.max_support <= std::min(ctx_->small_support_max, kSmallSupportMax);
  if (ps.small && nGridPlanes() == 1) {
    pp.zdiv = nz;
    pp.nzg = 1;
  }
  // the tile kernels address a table through 32-bit buffer offsets: one of 2 GiB or more is
  // addressed per key group (the small-support kernels use 64-bit element offsets)
  const bool own_table = ps.cfv.ptr == cf_->table.get() || ps.cfv.ptr == cf_->table16.get();
  if (!ps.small && own_table && nGridPlanes() == 1 &&
      ps.cfv.bytes + cf_resource_pad(ps.cfv.max_support, ps.cfv.prec) > kCFResourceMax) {
    const size_t os2 = (size_t)cf_->oversample * cf_->oversample;
    const size_t esz = cf_elem_bytes(ps.cfv.prec);
    if (cf_->zg_zdiv != pp.zdiv || cf_->zg_n != pp.nzg) {
      YG_CHECK(!arena, 1, "a CF table of 2 GiB or more needs a pass that may allocate before the allocation-free one");
      std::vector<long long> base(pp.nzg);
      for (int g = 0; g < pp.nzg; ++g) {
        const size_t z0 = (size_t)g * pp.zdiv, z1 = std::min((size_t)nz, (size_t)(g + 1) * pp.zdiv);
        const size_t b = (size_t)cf_->off[z0 * os2];
        const size_t e = z1 < (size_t)nz ? (size_t)cf_->off[z1 * os2] : cf_->n_elems;
        YG_CHECK((e - b) * esz + cf_resource_pad(ps.cfv.max_support, ps.cfv.prec) <= kCFResourceMax, 1,
                 "the CF planes of one w-plane group span 2 GiB or more");
        base[g] = (long long)b;
      }
      cf_->d_zg.release();
      cf_->d_zg.resize(base.size() * sizeof(long long));
      YG_HIP(hipMemcpy(cf_->d_zg.get(), base.data(), base.size() * sizeof(long long), hipMemcpyHostToDevice));
      cf_->zg_zdiv = pp.zdiv;
      cf_->zg_n = pp.nzg;
    }
    ps.cfv.zg_base = cf_->d_zg.as<long long>();
    ps.cfv.nzg = pp.nzg;
  }
  return true;
}

// Error flags and counters of a finished pass (TableVisGridder.cc:182-317 statistics).
void Gridder::account_pass(bool forward, const unsigned long long cnt[5], int err, double seconds) {
  YG_CHECK(!(err & 1), 4, "Weight is supposed to be a positive number (zero noise)");
  YG_CHECK(!(err & 2), 4, "W scaling error: recommend allowing larger range of w");
  YG_CHECK(!(err & 4), 4, "Index into convolution functions exceeds number of planes");
  if (forward) {
    st_samples_d_ += (double)cnt[0];
    st_taps_d_ += (double)cnt[1];
    t_degrid_ += seconds;
  } else {
    st_samples_g_ += (double)cnt[0];
    st_taps_g_ += (double)cnt[1];
    st_flag_ += (double)cnt[2];
    t_grid_ += seconds;
  }
  st_wflag_ += (double)cnt[3];
  st_rejected_ += (double)cnt[4];
}

void Gridder::generic(const ygrid_chunk& a, bool forward) {
  const auto t0 = std::chrono::steady_clock::now();
  PassSetup ps;
  if (!setup_pass(a, forward, ps)) return;
  const ChunkDev& d = ps.d;
  const int nsamp = ps.nsamp;
  YG_HIP(hipMemsetAsync(ps.out.counters, 0, 8 * sizeof(unsigned long long), ctx_->stream));
  YG_HIP(hipMemsetAsync(ps.out.error, 0, sizeof(int), ctx_->stream));
  if (profiling_) YG_HIP(hipEventRecord(ev_[0], ctx_->stream));
  launch_prep(*ctx_, ps.pp, ps.in, ps.out);
  afterPrep(ps);
  if (profiling_) YG_HIP(hipEventRecord(ev_[1], ctx_->stream));
  const Samples smp = ps.out.smp;
  const CFView& cfv = ps.cfv;
  const size_t wstride = nGridPlanes() > 1 ? gridPlaneElems() : 0;
  // w stacking grids on the tile kernels: its keys (w grid x tile) hold few samples each, which
  // the small-support kernel's per-band work units serve worse
  if (ps.small && (forward || nGridPlanes() == 1)) {
    // small supports: degridding reads each sample's footprint directly (no binning);
    // gridding bins with one work item per key, without a host round trip
    ++n_small_;
    if (forward) {
      if (profiling_) YG_HIP(hipEventRecord(ev_[2], ctx_->stream));
      SmallDegridArgs da{smp, cfv.ptr, grid_.as<float>(), ps.acc, d.nrow, d.nchan, shape_[0],
                         shape_[1], ps.pp.nzg, wstride};
      da.round16 = cfv.mma16;
      da.gscale = cfv.gscale;
      launch_degrid_small(*ctx_, da, shape_[2], cfv.prec, cfv.max_support);
      if (profiling_) YG_HIP(hipEventRecord(ev_[3], ctx_->stream));
      launch_degrid_finish(*ctx_, smp, ps.acc, d.nrow, d.nchan, d.npol, d.vis_out, d.vis_out_ld);
    } else {
      SmallGridBufs& sb = ctx_->small;
      constexpr int kOneItemPerKey = (1 << 30) - 1;
      size_static_bins(*ctx_, sb.bins, d.nrow, d.nchan, shape_[0], shape_[1], ps.pp.nzg * shape_[3], cfv,
                       kOneItemPerKey);
      const StaticBinsView& v = sb.bins.v;
      YG_CHECK(v.npairs_cap < (size_t)kOneItemPerKey && v.nitems_cap < (size_t)INT_MAX, 1,
               "too many (sample, tile) pairs for one pass: split the chunk");
      sb.items.resize(v.nitems_cap * sizeof(int4));
      sb.base.resize(2 * sizeof(int));
      sb.counter.resize(kQueueInts * sizeof(int));
      int* base = sb.base.as<int>();
      launch_fill_u32(*ctx_, base, 2, 0u);
      launch_fill_u32(*ctx_, sb.counter.get(), kQueueInts, 0u);
      enqueue_static_binning(*ctx_, v, smp, cfv, sb.items.as<int4>(), base, base + 1, 0);
      if (profiling_) YG_HIP(hipEventRecord(ev_[2], ctx_->stream));
      const int ntu = (shape_[0] + kTile - 1) / kTile, ntv = (shape_[1] + kTile - 1) / kTile;
      launch_grid_small(*ctx_, sb.items.as<int4>(), base + 1, sb.counter.as<int>(), (int)v.nitems_cap,
                        SmallGridArgs{v.pairs, smp, cfv.ptr, grid_.as<float>(), shape_[0], shape_[1], ntu, ntu * ntv,
                                      ps.pp.nzg, wstride},
                        shape_[2], cfv.prec, cfv.max_support);
      if (profiling_) YG_HIP(hipEventRecord(ev_[3], ctx_->stream));
    }
  } else {
    if (bin_samples(*ctx_, smp, d.nrow, d.nchan, shape_[0], shape_[1], ps.pp.nzg * shape_[3], cfv)) ++n_bin_reused_;
    Binning& bins = ctx_->bins.bin;
    if (forward) {
      YG_HIP(hipMemsetAsync(ps.acc, 0, (size_t)nsamp * 8 * sizeof(float), ctx_->stream));
      if (profiling_) YG_HIP(hipEventRecord(ev_[2], ctx_->stream));
      launch_degrid(*ctx_, bins, cfv, grid_.as<float2>(), shape_[0], shape_[1], shape_[2], ps.pp.nzg, ps.acc, wstride);
      if (profiling_) YG_HIP(hipEventRecord(ev_[3], ctx_->stream));
      launch_degrid_finish(*ctx_, smp, ps.acc, d.nrow, d.nchan, d.npol, d.vis_out, d.vis_out_ld);
    } else {
      if (profiling_) YG_HIP(hipEventRecord(ev_[2], ctx_->stream));
      launch_grid(*ctx_, bins, smp, cfv, grid_.as<float2>(), shape_[0], shape_[1], shape_[2], ps.pp.nzg, wstride);
      if (profiling_) YG_HIP(hipEventRecord(ev_[3], ctx_->stream));
    }
  }
  unsigned long long cnt[5] = {0, 0, 0, 0, 0};
  int err = 0;
  YG_HIP(hipMemcpyAsync(cnt, ps.out.counters, sizeof(cnt), hipMemcpyDeviceToHost, ctx_->stream));
  YG_HIP(hipMemcpyAsync(&err, ps.out.error, sizeof(int), hipMemcpyDeviceToHost, ctx_->stream));
  if (forward && a.mem == YGRID_MEM_HOST && nsamp) {
    const size_t span = (size_t)d.nrow + (size_t)d.vis_out_ld * ((size_t)d.nchan * d.npol - 1);
    YG_HIP(hipMemcpyAsync(a.vis_out, d.vis_out, span * sizeof(float2), hipMemcpyDeviceToHost, ctx_->stream));
  }
  YG_HIP(hipStreamSynchronize(ctx_->stream));
  if (profiling_) {
    float ta = 0, tb = 0, tk = 0;
    YG_HIP(hipEventElapsedTime(&ta, ev_[0], ev_[1]));
    YG_HIP(hipEventElapsedTime(&tb, ev_[1], ev_[2]));
    YG_HIP(hipEventElapsedTime(&tk, ev_[2], ev_[3]));
    ms_prep_ += ta;
    ms_bin_ += tb;
    if (forward) { ms_dk_ += tk; n_dk_ += 1; } else { ms_gk_ += tk; n_gk_ += 1; }
  }
  account_pass(forward, cnt, err, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
}

void Gridder::grid(const ygrid_chunk& a) {
  if (needsPolConversion(a)) convertedPass(a, false);
  else generic(a, false);
}
void Gridder::degrid(const ygrid_chunk& a) {
  if (needsPolConversion(a)) convertedPass(a, true);
  else generic(a, true);
}

// TableVisGridder::finaliseGrid (TableVisGridder.cc:1109-1236)
void Gridder::finaliseGrid(float* out, int mem) {
  YG_CHECK(initialised_ && !degrid_mode_ && grid_.get(), 2, "finaliseGrid without an initialised grid");
  const int nxp = shape_[0], nyp = shape_[1], np = shape_[2] * shape_[3];
  const size_t ng = (size_t)nxp * nyp * np;
  scratch_.resize(ng * sizeof(double2));
  if (doesCorrection() && !dopcf_) uploadCorrection();
  const double* cx = (doesCorrection() && !dopcf_) ? ccfx_.as<double>() : nullptr;
  const double* cy = (doesCorrection() && !dopcf_) ? ccfy_.as<double>() : nullptr;
  const size_t nout = (size_t)img_[0] * img_[1] * np;
  float* dout = out;
  DevBuf tmp;
  if (mem != YGRID_MEM_DEVICE) {
    tmp.resize(nout * sizeof(float));
    dout = tmp.as<float>();
  }
  // FP64 transform: the grid correction amplifies the image edges by 1/grdsf (~1e2), which
  // would lift an FP32 transform's rounding there to ~1e-4 of the image peak
  gridsToImage(cx, cy, dout);
  if (mem != YGRID_MEM_DEVICE)
    YG_HIP(hipMemcpyAsync(out, dout, nout * sizeof(float), hipMemcpyDeviceToHost, ctx_->stream));
  YG_HIP(hipStreamSynchronize(ctx_->stream));
  if (clear_grid_) grid_.release();
}

void Gridder::gridsToImage(const double* cx, const double* cy, float* dout) {
  const int nxp = shape_[0], nyp = shape_[1], np = shape_[2] * shape_[3];
  grid_to_image_z(*ctx_, grid_.as<float2>(), nxp, nyp, np, cx, cy, img_[0], img_[1], scratch_.as<double2>(), dout,
                  plan_);
}

void Gridder::modelToGrids(const float* dmodel, const double* cx, const double* cy) {
  const int nxp = shape_[0], nyp = shape_[1], np = shape_[2] * shape_[3];
  scratch_.resize(gridPlaneElems() * sizeof(double2));
  model_to_grid_z(*ctx_, dmodel, img_[0], img_[1], np, cx, cy, nxp, nyp, scratch_.as<double2>(), grid_.as<float2>(),
                  plan_);
}

// TableVisGridder::finaliseWeights (TableVisGridder.cc:1255-1284)
void Gridder::finaliseWeights(float* out, int mem) {
  YG_CHECK(initialised_ && !degrid_mode_, 2, "finaliseWeights without an initialised grid");
  const int npol = shape_[2], nchan = shape_[3];
  std::vector<double> sw((size_t)nsw_rows_ * npol * nchan);
  YG_HIP(hipMemcpyAsync(sw.data(), sumw_.get(), sw.size() * sizeof(double), hipMemcpyDeviceToHost, ctx_->stream));
  YG_HIP(hipStreamSynchronize(ctx_->stream));
  const size_t plane = (size_t)img_[0] * img_[1];
  std::vector<float> host(plane * npol * nchan);
  for (int c = 0; c < nchan; ++c)
    for (int p = 0; p < npol; ++p) {
      double s = 0.0;
      for (int z = 0; z < nsw_rows_; ++z) s += sw[z + (size_t)nsw_rows_ * (p + (size_t)npol * c)];
      std::fill(host.begin() + (size_t)(c * npol + p) * plane, host.begin() + (size_t)(c * npol + p + 1) * plane, (float)s);
    }
  YG_HIP(hipMemcpy(out, host.data(), host.size() * sizeof(float),
                   mem == YGRID_MEM_DEVICE ? hipMemcpyHostToDevice : hipMemcpyHostToHost));
}

// TableVisGridder::initialiseDegrid (TableVisGridder.cc:1286-1319)
void Gridder::initialiseDegrid(const ygrid_image_desc& img, const float* model, int mem) {
  initialiseCommon(img);
  dopsf_ = dopcf_ = false;
  degrid_mode_ = true;
  YG_CHECK(model != nullptr, 1, "model image is null");
  const int nxp = shape_[0], nyp = shape_[1], np = shape_[2] * shape_[3];
  const size_t nin = (size_t)img_[0] * img_[1] * np;
  const float* dmodel = model;
  DevBuf tmp;
  if (mem != YGRID_MEM_DEVICE) {
    bool any = false;
    for (size_t i = 0; i < nin && !any; ++i) any = model[i] != 0.f;
    model_empty_ = !any;
    if (!model_empty_) {
      tmp.resize(nin * sizeof(float));
      YG_HIP(hipMemcpyAsync(tmp.get(), model, nin * sizeof(float), hipMemcpyHostToDevice, ctx_->stream));
      dmodel = tmp.as<float>();
    }
  } else {
    DevBuf flag(sizeof(int));
    int h = 0;
    YG_HIP(hipMemsetAsync(flag.get(), 0, sizeof(int), ctx_->stream));
    launch_any_nonzero(*ctx_, model, nin, flag.as<int>());
    YG_HIP(hipMemcpyAsync(&h, flag.get(), sizeof(int), hipMemcpyDeviceToHost, ctx_->stream));
    YG_HIP(hipStreamSynchronize(ctx_->stream));
    model_empty_ = h == 0;
  }
  const size_t ng = (size_t)nxp * nyp * np * nGridPlanes();
  grid_.resize(ng * sizeof(float2));
  if (model_empty_) {
    YG_HIP(hipMemsetAsync(grid_.get(), 0, ng * sizeof(float2), ctx_->stream));
    YG_HIP(hipStreamSynchronize(ctx_->stream));
    return;
  }
  if (doesCorrection()) uploadCorrection();
  const double* cx = doesCorrection() ? ccfx_.as<double>() : nullptr;
  const double* cy = doesCorrection() ? ccfy_.as<double>() : nullptr;
  modelToGrids(dmodel, cx, cy);
  if (cf_prec_ == kPrecF16) {
    // the power-of-two scale that keeps the model grid inside FP16's range when the 16-bit
    // matrix-core degridding rounds it
    gscale_.resize(2 * sizeof(uint32_t));
    launch_grid_scale16(*ctx_, grid_.as<float>(), 2 * ng, gscale_.as<uint32_t>(), gscale_.as<float>() + 1);
  }
  YG_HIP(hipStreamSynchronize(ctx_->stream));
  scratch_.release();
}

void Gridder::finaliseDegrid() {
  if (clear_grid_) grid_.release();
}

void Gridder::stats(ygrid_stats* o) const {
  o->samples_gridded = st_samples_g_;
  o->taps_gridded = st_taps_g_;
  o->samples_degridded = st_samples_d_;
  o->taps_degridded = st_taps_d_;
  o->vectors_flagged = st_flag_;
  o->vectors_wflagged = st_wflag_;
  o->time_grid_s = t_grid_;
  o->time_degrid_s = t_degrid_;
  o->time_cf_s = t_cf_;
  o->support = cf_ ? cf_->support : 0;
  o->oversample = cf_ ? cf_->oversample : 0;
  o->n_cf = cf_ ? (int)cf_->off.size() : 0;
  o->ms_prep = ms_prep_;
  o->ms_bin = ms_bin_;
  o->ms_grid_kernel = ms_gk_;
  o->ms_degrid_kernel = ms_dk_;
  o->n_grid_launches = n_gk_;
  o->n_degrid_launches = n_dk_;
  o->n_bin_reused = n_bin_reused_;
  o->small_support_passes = n_small_;
  o->grid_waves = ctx_->grid_variant.waves;
  o->grid_pipeline = ctx_->grid_variant.pipeline;
  o->degrid_waves = ctx_->degrid_variant.waves;
  o->degrid_pipeline = ctx_->degrid_variant.pipeline;
  o->cf_precision = cf_prec_;
  o->vis_precision = vis_prec_;
  o->learned_cf_active = learnedActive() ? 1 : 0;
  o->degrid_matrix = ctx_->degrid_variant.matrix && shape_[2] >= 2 ? 1 : 0;
  o->rows_rejected_pointing = st_rejected_;
  o->cf_iterations = cf_iter_;
  o->cf_generations = cf_gen_;
  o->cf_generations_pa = cf_gen_pa_;
  o->cf_generations_freq = cf_gen_freq_;
}

void Gridder::getCF(int k, float* out, int* support) const {
  YG_CHECK(cf_ && cf_->n_elems > 0, 2, "convolution function not generated yet");
  YG_CHECK(k >= 0 && k < (int)cf_->off.size(), 1, "CF plane index out of range");
  const int S = cf_->supp[k], W = 2 * S + 1, P = cf_pitch(S);
  if (support) *support = S;
  if (!out) return;
  std::vector<float2> host((size_t)P * cf_rows(S));
  YG_HIP(hipStreamSynchronize(ctx_->stream));  // the table may still be filling on the context stream
  YG_HIP(hipMemcpy(host.data(), cf_->table.as<float2>() + cf_->off[k], host.size() * sizeof(float2), hipMemcpyDeviceToHost));
  for (int iy = 0; iy < W; ++iy)
    for (int ix = 0; ix < W; ++ix) {
      const float2 v = host[(size_t)(iy + kGuardV) * P + ix + kGuardU];
      out[2 * ((size_t)iy * W + ix)] = v.x;
      out[2 * ((size_t)iy * W + ix) + 1] = v.y;
    }
}

void Gridder::getGrid(float* out, size_t* n) const {
  const size_t ng = gridPlaneElems() * nGridPlanes();
  if (n) *n = 2 * ng;
  if (!out) return;
  YG_CHECK(grid_.get(), 2, "no grid");
  YG_HIP(hipStreamSynchronize(ctx_->stream));  // passes still in flight on the context stream
  YG_HIP(hipMemcpy(out, grid_.get(), ng * sizeof(float2), hipMemcpyDeviceToHost));
}

// ---------------------------------------------------------------------------------------
// Box (BoxVisGridder.cc:62-76)
// ---------------------------------------------------------------------------------------
class BoxGridder : public Gridder {
 public:
  using Gridder::Gridder;
  std::unique_ptr<Gridder> clone() const override { return std::unique_ptr<Gridder>(new BoxGridder(*this)); }
  std::string name() const override { return "GPUBox"; }

 protected:
  bool doesCorrection() const override { return false; }
  void initConvolutionFunction(const ChunkDev&) override {
    if (cf_) return;
    auto t = std::make_shared<CFTable>();
    layout_table(*t, {1}, 1);
    std::vector<float2> host((size_t)cf_pitch(1) * cf_rows(1), make_float2(0.f, 0.f));
    host[(size_t)(1 + kGuardV) * cf_pitch(1) + 1 + kGuardU] = make_float2(1.f, 0.f);
    t->table.resize(host.size() * sizeof(float2));
    YG_HIP(hipMemcpy(t->table.get(), host.data(), host.size() * sizeof(float2), hipMemcpyHostToDevice));
    t->upload(*ctx_);
    cf_ = t;
  }
};

// ---------------------------------------------------------------------------------------
// SphFunc (SphFuncVisGridder.cc:43-266)
// ---------------------------------------------------------------------------------------
class SphFuncGridder : public Gridder {
 public:
  SphFuncGridder(std::shared_ptr<Ctx> c, double alpha, int support, int oversample)
      : Gridder(std::move(c)), alpha_(alpha), support_(support), os_(oversample) {
    YG_CHECK(support >= 3, 1, "SphFunc support must be >= 3");
    YG_CHECK(oversample >= 1, 1, "oversample must be >= 1");
    YG_CHECK(alpha >= 0. && alpha <= 2., 1, "alpha must be in [0, 2]");
  }
  std::unique_ptr<Gridder> clone() const override { return std::unique_ptr<Gridder>(new SphFuncGridder(*this)); }
  std::string name() const override { return "GPUSphFunc"; }

 protected:
  int oversampling() const override { return os_; }
  void gridCorrection(std::vector<double>& cx, std::vector<double>& cy) const override {
    sph_grid_correction_xy(shape_[0], shape_[1], alpha_, 3, cx, cy);
  }
  void initConvolutionFunction(const ChunkDev&) override {
    if (cf_) return;
    auto t = std::make_shared<CFTable>();
    if (dopcf_) {
      // nearest-neighbour (1, 3i) kernel (SphFuncVisGridder.cc:99-125), normalised by the sum
      // of |C| like every SphFunc plane (SphFuncVisGridder.cc:159-167): (1, 3)/sqrt(10)
      layout_table(*t, {1}, os_);
      std::vector<float2> host(t->n_elems, make_float2(0.f, 0.f));
      const double nrm = std::sqrt(10.0);
      for (int fu = 0; fu < os_; ++fu)
        for (int fv = 0; fv < os_; ++fv) {
          const int k = fu + os_ * fv;
          const int ix = (int)(-float(fu) / float(os_)), iy = (int)(-float(fv) / float(os_));
          host[t->off[k] + (size_t)(iy + 1 + kGuardV) * cf_pitch(1) + ix + 1 + kGuardU] =
              make_float2((float)(1.0 / nrm), (float)(3.0 / nrm));
        }
      t->table.resize(host.size() * sizeof(float2));
      YG_HIP(hipMemcpy(t->table.get(), host.data(), host.size() * sizeof(float2), hipMemcpyHostToDevice));
      t->upload(*ctx_);
      cf_ = t;
      return;
    }
    const int S = support_, W = 2 * S + 1;
    layout_table(*t, {S}, os_);
    const Spheroidal sph(kPi * S, alpha_);
    std::vector<double> b((size_t)os_ * W);
    for (int f = 0; f < os_; ++f)
      for (int i = 0; i < W; ++i) {
        const double nu = std::fabs(double(os_ * (i - S) + f)) / double(S * os_);
        b[(size_t)f * W + i] = sph(nu) * std::pow(1.0 - nu * nu, alpha_);
      }
    DevBuf db(b.size() * sizeof(double));
    YG_HIP(hipMemcpy(db.get(), b.data(), b.size() * sizeof(double), hipMemcpyHostToDevice));
    t->table.resize(t->n_elems * sizeof(float2));
    t->upload(*ctx_);
    launch_sph_cf(*ctx_, db.as<double>(), db.as<double>(), os_, S, t->d_off.as<int32_t>(), t->table.as<float2>());
    YG_HIP(hipStreamSynchronize(ctx_->stream));
    cf_ = t;
  }

 private:
  double alpha_;
  int support_, os_;
};

// ---------------------------------------------------------------------------------------
// WProject (WProjectVisGridder.cc, WDependentGridderBase.cc)
// ---------------------------------------------------------------------------------------
// WDependentGridderBase::configureWSampling (WDependentGridderBase.cc:192-245)
void configure_wsampling(WSamplingDesc& ws, const Params& p) {
  const std::string smp = p.str("wsampling", "linear");
  if (smp == "powerlaw") {
    ws.kind = 1;
    ws.exponent = p.f64("wsampling.exponent", 1.0);
    YG_CHECK(ws.exponent != 0., 1, "Zero exponent in power law doesn't make sense");
  } else if (smp == "gaussian") {
    ws.kind = 2;
    YG_CHECK(p.has("wsampling.nwplanes50"), 1, "wsampling.nwplanes50 is required");
    const double n50 = p.f64("wsampling.nwplanes50", 1.0);
    YG_CHECK(n50 > 0. && n50 < double(ws.nw) / std::sqrt(2.), 1, "wsampling.nwplanes50 out of range");
    ws.gaussian(n50 / ws.nw);
  } else {
    YG_CHECK(smp == "linear", 1, "W-sampling " + smp + " is not implemented");
  }
  ws.wmaxclip = p.flag("wmaxclip", false);
}

class WProjectGridder : public Gridder {
 public:
  WProjectGridder(std::shared_ptr<Ctx> c, const Params& p) : Gridder(std::move(c)) {
    // WProjectVisGridder::createGridder / configureGridder (WProjectVisGridder.cc:303-363)
    ws_.wmax = p.f64("wmax", 35000.0);
    ws_.nw = p.i32("nwplanes", 65);
    cutoff_ = p.f64("cutoff", 1e-3);
    os_ = p.i32("oversample", 8);
    maxsupport_ = p.i32("maxsupport", 256);
    limitsupport_ = p.i32("limitsupport", 0);
    alpha_ = p.f64("alpha", 1.0);
    variable_ = p.flag("variablesupport", false);
    offset_ = p.flag("offsetsupport", false);
    abs_cutoff_ = p.flag("cutoff.absolute", false);
    YG_CHECK(ws_.wmax > 0.0, 1, "Baseline length must be greater than zero");
    YG_CHECK(ws_.nw > 0 && ws_.nw % 2 == 1, 1, "Number of w planes must be a positive odd number");
    YG_CHECK(os_ > 0, 1, "Oversampling must be greater than 0");
    YG_CHECK(maxsupport_ > 0, 1, "Maximum support must be greater than 0");
    YG_CHECK(!offset_ || variable_, 1, "offsetsupport option of the gridder should only be used together with variablesupport option");
    if (!abs_cutoff_) YG_CHECK(cutoff_ > 0.0 && cutoff_ < 1.0, 1, "Cutoff must be in (0, 1)");
    configure_wsampling(ws_, p);
  }
  std::unique_ptr<Gridder> clone() const override { return std::unique_ptr<Gridder>(new WProjectGridder(*this)); }
  std::string name() const override { return "GPUWProject"; }
  bool wTerms(std::vector<double>& w, double& wmax) const override {
    w.resize(ws_.nw);
    for (int z = 0; z < ws_.nw; ++z) w[z] = ws_.wterm(z);
    wmax = ws_.wmax;
    return true;
  }
  void setLearnedCF(std::shared_ptr<LearnedCF> lcf, double max_err) override {
    lcf_ = std::move(lcf);
    lcf_max_err_ = max_err;
  }
  bool learnedActive() const override {
    return lcf_ && lcf_->trained && lcf_->holdout_max_err <= lcf_max_err_;
  }

 protected:
  int oversampling() const override { return os_; }
  int nZPlanes() const override { return ws_.nw; }
  void gridCorrection(std::vector<double>& cx, std::vector<double>& cy) const override {
    sph_grid_correction_xy(shape_[0], shape_[1], alpha_, 3, cx, cy);
  }
  CIndexDesc cindex(const ChunkDev&) const override {
    CIndexDesc d;
    d.kind = 1;
    d.nw = ws_.nw;
    d.wscale = ws_.wscale();
    d.wmax = ws_.wmax;
    d.sampling = ws_.kind;
    d.exponent = ws_.exponent;
    d.g_ts = ws_.ts;
    d.g_amp = ws_.amp;
    d.wmaxclip = ws_.wmaxclip ? 1 : 0;
    return d;
  }
  void initConvolutionFunction(const ChunkDev&) override;
  void initPCFKernel();

  // Learned CFs: every sample its own padded plane of the network's support, at the gridder's
  // CF storage precision, filled by the network after the prep kernel.
  void customizePass(PassSetup& ps) override {
    if (!learnedActive() || dopcf_) return;  // the preconditioner pass keeps its 3x3 kernels
    const int S = lcf_->support;
    const int P = 2 * S + 1 + 2 * kGuardU, R = 2 * S + 1 + 2 * kGuardV;
    const long long plane = (long long)R * P;
    const size_t bytes = (size_t)ps.nsamp * plane * cf_elem_bytes(cf_prec_);
    YG_CHECK(bytes < ((size_t)1 << 31) - ((size_t)1 << 24), 1,
             "learned CF planes of this chunk exceed 2 GiB: use fewer samples per chunk or cfprecision fp16");
    if (!lcf_buf_.table || lcf_buf_.table->size() < bytes || lcf_buf_.S != S || lcf_buf_.prec != cf_prec_) {
      // a new buffer (the network may still hold the old one as its last evaluation)
      lcf_buf_.table = std::make_shared<DevBuf>(bytes);
      YG_HIP(hipMemsetAsync(lcf_buf_.table->get(), 0, lcf_buf_.table->size(), ctx_->stream));  // guard bands
      lcf_buf_.S = S;
      lcf_buf_.prec = cf_prec_;
    }
    lcf_buf_.in.resize((size_t)ps.nsamp * sizeof(float4));
    lcf_->reserve_eval(S);
    ps.pp.lcf_S = S;
    ps.pp.lcf_plane_elems = plane;
    ps.out.lcf_in = lcf_buf_.in.as<float4>();
    ps.cfv = CFView{lcf_buf_.table->get(), lcf_buf_.table->size(), S, cf_prec_,
                    cf_prec_ != kPrecF32 && vis_prec_ == cf_prec_ ? 1 : 0};
    if (degrid_mode_ && cf_prec_ == kPrecF16 && gscale_.get()) ps.cfv.gscale = gscale_.as<float>() + 1;
  }
  void afterPrep(PassSetup& ps) override {
    if (!ps.out.lcf_in) return;
    const int S = ps.pp.lcf_S;
    lcf_->eval_samples(ps.out.lcf_in, ps.nsamp, lcf_buf_.table, 2 * S + 1 + 2 * kGuardU, ps.pp.lcf_plane_elems,
                       ps.cfv.prec);
  }

 private:
  // per-gridder buffers of the learned planes (clones start without them)
  struct LcfBuffers {
    std::shared_ptr<DevBuf> table;
    DevBuf in;
    int S = -1, prec = -1;
    LcfBuffers() = default;
    LcfBuffers(const LcfBuffers&) {}
    LcfBuffers& operator=(const LcfBuffers&) { return *this; }
  };
  std::shared_ptr<LearnedCF> lcf_;
  double lcf_max_err_ = 0.0;
  LcfBuffers lcf_buf_;
  WSamplingDesc ws_;
  double cutoff_ = 1e-3, alpha_ = 1.0;
  int os_ = 8, maxsupport_ = 256, limitsupport_ = 0;
  bool variable_ = false, offset_ = false, abs_cutoff_ = false;
};

// SupportSearcher on the row/column through the peak (SupportSearcher.tcc:161-257) and
// WProjectVisGridder::extractSupport (WProjectVisGridder.cc:198-263).
SupportResult search_support(const Ctx& c, const float2* plane, int nx, int ny, double cutoff, bool absolute,
                                    bool offset, int os) {
  DevBuf am(sizeof(unsigned long long)), cross((size_t)(nx + ny) * sizeof(float2));
  launch_argmax(c, plane, nx * ny, am.as<unsigned long long>());
  unsigned long long pk = 0;
  YG_HIP(hipMemcpyAsync(&pk, am.get(), sizeof(pk), hipMemcpyDeviceToHost, c.stream));
  YG_HIP(hipStreamSynchronize(c.stream));
  const unsigned idx = 0xffffffffu - (unsigned)(pk & 0xffffffffu);
  const int px = (int)(idx % nx), py = (int)(idx / nx);
  launch_gather_cross(c, plane, nx, ny, px, py, cross.as<float2>());
  std::vector<float2> h((size_t)nx + ny);
  YG_HIP(hipMemcpyAsync(h.data(), cross.get(), h.size() * sizeof(float2), hipMemcpyDeviceToHost, c.stream));
  YG_HIP(hipStreamSynchronize(c.stream));
  auto amp = [](float2 v) { return (double)v.x * v.x + (double)v.y * v.y; };
  double peak = std::sqrt(amp(h[px]));
  YG_CHECK(peak > 0.0, 4, "Unable to find peak in the support searcher");
  if (absolute) peak = 1.0;
  YG_CHECK(px > 0 && py > 0 && px + 1 < nx && py + 1 < ny, 4,
           "Peak position of the convolution function is too close to the edge, increase maxsupport");
  const double thr = std::fabs(cutoff * peak) * std::fabs(cutoff * peak);
  int blc0 = -1, blc1 = -1, trc0 = -1, trc1 = -1;
  for (int ix = 0; ix <= px; ++ix) if (amp(h[ix]) > thr) { blc0 = ix; break; }
  for (int iy = 0; iy <= py; ++iy) if (amp(h[nx + iy]) > thr) { blc1 = iy; break; }
  for (int ix = nx - 1; ix >= px; --ix) if (amp(h[ix]) > thr) { trc0 = ix; break; }
  for (int iy = ny - 1; iy >= py; --iy) if (amp(h[nx + iy]) > thr) { trc1 = iy; break; }
  YG_CHECK(blc0 >= 0 && blc1 >= 0 && trc0 >= 0 && trc1 >= 0, 4,
           "Unable to find the support on one of the coordinates (try decreasing the value of cutoff)");
  SupportResult r{0, 0, 0};
  if (offset) {
    r.size = std::max(trc0 - blc0, trc1 - blc1);
    r.offu = (px - nx / 2) / os;
    r.offv = (py - ny / 2) / os;
  } else {
    const int c0 = nx / 2, c1 = ny / 2;
    const int xm = std::max(std::abs(trc0 - c0), std::abs(c0 - blc0));
    const int ym = std::max(std::abs(trc1 - c1), std::abs(c1 - blc1));
    r.size = 2 * std::max(xm, ym);
    YG_CHECK(r.size > 0, 4, "Unable to determine support of convolution function");
  }
  r.size /= 2 * os;
  if (r.size < 3) r.size = 3;
  return r;
}

// Nearest-neighbour preconditioner kernel (WProjectVisGridder::doPCFGridder,
// WProjectVisGridder.cc:385-434): support 1, per w plane and oversampling offset a 3x3 plane whose
// centre is (1, wKernelPix), the fitted width of the plane's W-projection kernel in pixels:
// wThetaPix = |w| / du^2; 3 below one pixel, else 6 + 1.14 wThetaPix (cutoff < 0.01) or
// sqrt(49 + wThetaPix^2).  Any offset support is left at zero (WProjectVisGridder.cc:164-171).
void WProjectGridder::initPCFKernel() {
  auto t = std::make_shared<CFTable>();
  const int nw = ws_.nw;
  layout_table(*t, std::vector<int>(nw, 1), os_);
  std::vector<float2> host(t->n_elems, make_float2(0.f, 0.f));
  for (int iw = 0; iw < nw; ++iw) {
    const float wThetaPix = (float)(std::fabs(ws_.wterm(iw)) / (uvcell_[0] * uvcell_[0]));
    float wKernelPix;
    if (wThetaPix < 1) {
      wKernelPix = 3;
    } else if (cutoff_ < 0.01) {
      wKernelPix = (float)(6 + 1.14 * wThetaPix);
    } else {
      wKernelPix = std::sqrt(49.f + wThetaPix * wThetaPix);
    }
    for (int fu = 0; fu < os_; ++fu)
      for (int fv = 0; fv < os_; ++fv) {
        const int k = fu + os_ * (fv + os_ * iw);
        // the reference's tap index int(-frac/os) truncates to 0: the centre of every plane
        const int ix = (int)(-float(fu) / float(os_)), iy = (int)(-float(fv) / float(os_));
        host[t->off[k] + (size_t)(iy + 1 + kGuardV) * cf_pitch(1) + ix + 1 + kGuardU] = make_float2(1.f, wKernelPix);
      }
  }
  t->table.resize(host.size() * sizeof(float2));
  YG_HIP(hipMemcpy(t->table.get(), host.data(), host.size() * sizeof(float2), hipMemcpyHostToDevice));
  t->upload(*ctx_);
  cf_ = t;
}

// WProjectVisGridder::generate (WProjectVisGridder.cc:435-522) with setup (:661-703)
void WProjectGridder::initConvolutionFunction(const ChunkDev&) {
  if (cf_) return;
  if (dopcf_) {
    initPCFKernel();
    return;
  }
  const Ctx& c = *ctx_;
  const int nx = maxsupport_, ny = maxsupport_;
  YG_CHECK(nx % 2 == 0, 1, "maxsupport must be even");
  const int qnx = nx / os_, qny = ny / os_;
  YG_CHECK(qnx > 3 && qny > 3, 1, "maxsupport/oversample too small");
  const double cellx = 1.0 / (double(shape_[0]) * uvcell_[0]);
  const double celly = 1.0 / (double(shape_[1]) * uvcell_[1]);
  const double ccellx = double(shape_[0]) * cellx / double(qnx);
  const double ccelly = double(shape_[1]) * celly / double(qny);
  const Spheroidal sph(kPi * 3, alpha_);
  auto taper = [&](int q) {
    std::vector<float> v(q);
    for (int i = 0; i < q; ++i) {
      const float nu = std::fabs(float(i - q / 2)) / float(q / 2);
      v[i] = (float)(sph((double)nu) / (double)float(q));
    }
    v[0] = 3.f * (v[1] - v[2]) + v[3];
    if (q % 2 == 1) v[q - 1] = v[0];
    return v;
  };
  const std::vector<float> hx = taper(qnx), hy = taper(qny);
  DevBuf dx(hx.size() * 4), dy(hy.size() * 4);
  YG_HIP(hipMemcpy(dx.get(), hx.data(), hx.size() * 4, hipMemcpyHostToDevice));
  YG_HIP(hipMemcpy(dy.get(), hy.data(), hy.size() * 4, hipMemcpyHostToDevice));

  const int nw = ws_.nw;
  std::vector<double> tpw(nw);
  for (int iw = 0; iw < nw; ++iw) tpw[iw] = 2.0f * kPi * ws_.wterm(iw);
  DevBuf dtpw(nw * sizeof(double));
  YG_HIP(hipMemcpy(dtpw.get(), tpw.data(), nw * sizeof(double), hipMemcpyHostToDevice));

  const size_t plane_bytes = (size_t)nx * ny * sizeof(float2);
  int nb = (int)std::max<size_t>(1, ((size_t)1 << 30) / plane_bytes);
  nb = std::min(nb, nw);
  DevBuf buf(plane_bytes * nb);
  PlanPtr plan;
  std::vector<int> supp_z(nw, 0), offu(nw, 0), offv(nw, 0);
  int common = 0;
  // pieces: per batch, the cut-out planes (layout is known once the batch's supports are)
  std::vector<std::pair<int, DevBuf>> pieces;
  for (int z0 = 0; z0 < nw; z0 += nb) {
    const int cnt = std::min(nb, nw - z0);
    launch_wproj_screen(c, buf.as<float2>(), nx, ny, qnx, qny, ccellx, ccelly, dx.as<float>(), dy.as<float>(),
                        dtpw.as<double>() + z0, cnt);
    fft_plain_c2c(c, buf.as<float2>(), nx, ny, cnt, true, plan);
    launch_checkerboard_sign(c, buf.as<float2>(), nx, ny, cnt);
    for (int b = 0; b < cnt; ++b) {
      const int iw = z0 + b;
      int size = common;
      if (variable_ || common == 0) {
        const SupportResult r = search_support(c, buf.as<float2>() + (size_t)b * nx * ny, nx, ny, cutoff_, abs_cutoff_,
                                               offset_, os_);
        YG_CHECK((r.size + 1) * os_ < nx / 2, 4,
                 "Overflowing convolution function - increase maxSupport or cutoff or decrease overSample");
        size = r.size;
        if (limitsupport_ > 0 && size > limitsupport_) size = limitsupport_;
        if (common == 0) common = size;
        if (offset_) { offu[iw] = r.offu; offv[iw] = r.offv; }
      }
      const int S = variable_ ? size : common;
      supp_z[iw] = S;
      // calcConvFuncOverflow (WProjectVisGridder.cc:704-726)
      const int kxmin = (-S + offu[iw]) * os_ + nx / 2, kxmax = (S + offu[iw]) * os_ + os_ - 1 + nx / 2;
      const int kymin = (-S + offv[iw]) * os_ + ny / 2, kymax = (S + offv[iw]) * os_ + os_ - 1 + ny / 2;
      YG_CHECK(kxmin >= 0 && kymin >= 0 && kxmax < nx && kymax < ny, 4,
               "Convolution function overflowing - increase maxsupport or cutoff or decrease oversample");
    }
    // cut out this batch into its own piece
    CFTable part;
    std::vector<int> sb(supp_z.begin() + z0, supp_z.begin() + z0 + cnt);
    layout_table(part, sb, os_);
    std::vector<int32_t> first(cnt);
    for (int b = 0; b < cnt; ++b) first[b] = part.off[(size_t)b * os_ * os_];
    DevBuf dfirst(cnt * 4), dsup(cnt * 4), dou(cnt * 4), dov(cnt * 4);
    YG_HIP(hipMemcpy(dfirst.get(), first.data(), cnt * 4, hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(dsup.get(), sb.data(), cnt * 4, hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(dou.get(), offu.data() + z0, cnt * 4, hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(dov.get(), offv.data() + z0, cnt * 4, hipMemcpyHostToDevice));
    DevBuf piece(part.n_elems * sizeof(float2));
    launch_cf_cutout(c, buf.as<float2>(), nx, ny, cnt, os_, dsup.as<int>(), dou.as<int>(), dov.as<int>(),
                     dfirst.as<int32_t>(), 0, 1.0, piece.as<float2>());
    YG_HIP(hipStreamSynchronize(c.stream));
    pieces.emplace_back((int)part.n_elems, std::move(piece));
  }
  auto t = std::make_shared<CFTable>();
  layout_table(*t, supp_z, os_);
  if (offset_) {
    t->ofs_u.assign(offu.begin(), offu.end());
    t->ofs_v.assign(offv.begin(), offv.end());
  }
  t->table.resize(t->n_elems * sizeof(float2));
  size_t pos = 0;
  for (auto& pc : pieces) {
    YG_HIP(hipMemcpyAsync(t->table.as<float2>() + pos, pc.second.get(), (size_t)pc.first * sizeof(float2),
                          hipMemcpyDeviceToDevice, c.stream));
    pos += (size_t)pc.first;
  }
  t->upload(c);
  cf_ = t;
}

// ---------------------------------------------------------------------------------------
// factory (VisGridderFactory::make, VisGridderFactory.cc:112-286)
// ---------------------------------------------------------------------------------------
std::unique_ptr<Gridder> make_awproject(std::shared_ptr<Ctx> ctx, const Params& p);
std::unique_ptr<Gridder> make_aprojectwstack(std::shared_ptr<Ctx> ctx, const Params& p);

// ---------------------------------------------------------------------------------------
// WStack (WStackVisGridder.cc): the SphFunc kernel (WDependentGridderBase derives from
// SphFuncVisGridder with its defaults, alpha 1, support 3, oversample 128) gridding each sample
// into the grid of its w plane (gIndex, :339-344); finaliseGrid transforms every plane, applies
// its w screen and sums the real parts (:255-289), initialiseDegrid fills every plane from the
// corrected model times the conjugate screen (:314-327).
// ---------------------------------------------------------------------------------------
class WStackGridder : public SphFuncGridder {
 public:
  WStackGridder(std::shared_ptr<Ctx> c, const Params& p) : SphFuncGridder(std::move(c), 1.0, 3, 128) {
    // WStackVisGridder::createGridder (WStackVisGridder.cc:355-363)
    ws_.wmax = p.f64("wmax", 35000.0);
    ws_.nw = p.i32("nwplanes", 65);
    YG_CHECK(ws_.wmax > 0.0, 1, "Baseline length must be greater than zero");
    YG_CHECK(ws_.nw > 0 && ws_.nw % 2 == 1, 1, "Number of w planes must be a positive odd number");
    configure_wsampling(ws_, p);
  }
  std::unique_ptr<Gridder> clone() const override { return std::unique_ptr<Gridder>(new WStackGridder(*this)); }
  std::string name() const override { return "GPUWStack"; }

 protected:
  int nGridPlanes() const override { return ws_.nw; }
  CIndexDesc cindex(const ChunkDev&) const override {
    CIndexDesc d;
    d.kind = 3;
    d.nw = ws_.nw;
    d.wscale = ws_.wscale();
    d.wmax = ws_.wmax;
    d.sampling = ws_.kind;
    d.exponent = ws_.exponent;
    d.g_ts = ws_.ts;
    d.g_amp = ws_.amp;
    d.wmaxclip = ws_.wmaxclip ? 1 : 0;
    return d;
  }
  // the screen's constants as WStackVisGridder::multiply computes them (:137-144)
  void screen(int iw, float& w2pi, float& cellx, float& celly) const {
    w2pi = (float)(2.0f * kPi * ws_.wterm(iw));
    cellx = (float)(1.0 / (float(shape_[0]) * uvcell_[0]));
    celly = (float)(1.0 / (float(shape_[1]) * uvcell_[1]));
  }
  // the planes that received samples: only those are transformed (WStackVisGridder.cc:236,272
  // skip planes whose largest amplitude is 0)
  std::vector<int> usedPlanes() {
    const size_t ng = gridPlaneElems();
    DevBuf flags((size_t)ws_.nw * sizeof(int));
    launch_fill_u32(*ctx_, flags.get(), ws_.nw, 0u);
    for (int iw = 0; iw < ws_.nw; ++iw)
      launch_any_nonzero(*ctx_, reinterpret_cast<const float*>(grid_.as<float2>() + ng * iw), 2 * ng,
                         flags.as<int>() + iw);
    std::vector<int> f(ws_.nw, 0), used;
    YG_HIP(hipMemcpyAsync(f.data(), flags.get(), f.size() * sizeof(int), hipMemcpyDeviceToHost, ctx_->stream));
    YG_HIP(hipStreamSynchronize(ctx_->stream));
    for (int iw = 0; iw < ws_.nw; ++iw)
      if (f[iw]) used.push_back(iw);
    return used;
  }
  void gridsToImage(const double* cx, const double* cy, float* dout) override {
    const int nxp = shape_[0], nyp = shape_[1], np = shape_[2] * shape_[3];
    const size_t ng = gridPlaneElems();
    const std::vector<int> used = usedPlanes();
    if (dopcf_) {
      // WStackVisGridder::finaliseGrid's preconditioner form (:222-262): the planes' uv grids,
      // their imaginary parts (the w support of the preconditioner kernel, gridded as 1 +- 3i)
      // scaled by the would-be W-projection kernel size over 3 (updatew, :173-209), are summed and
      // transformed once; the image is the real part of the normalised inverse transform, with
      // no correction and no scaling
      scratch_.resize(ng * sizeof(double2));
      launch_fill_u32(*ctx_, scratch_.get(), 4 * ng, 0u);
      const double norm = 1.0 / ((double)nxp * (double)nyp);
      for (int iw : used) {
        const float wThetaPix = (float)(std::fabs(ws_.wterm(iw)) / (uvcell_[0] * uvcell_[0]));
        float wKernelPix = wThetaPix < 1 ? 3.f : (float)(6 + 1.14 * wThetaPix);
        wKernelPix /= 3.f;
        wstack_pcf_stack(*ctx_, grid_.as<float2>() + ng * iw, nxp, nyp, np, (double)wKernelPix, norm,
                         scratch_.as<double2>());
      }
      stack_to_image_z(*ctx_, scratch_.as<double2>(), nxp, nyp, np, img_[0], img_[1], dout, plan_);
      YG_HIP(hipStreamSynchronize(ctx_->stream));
      return;
    }
    DevBuf acc(ng * sizeof(double));
    launch_fill_u32(*ctx_, acc.get(), 2 * ng, 0u);
    for (int iw : used) {
      float w2pi, cellx, celly;
      screen(iw, w2pi, cellx, celly);
      wstack_accumulate(*ctx_, grid_.as<float2>() + ng * iw, nxp, nyp, np, w2pi, cellx, celly, scratch_.as<double2>(),
                        acc.as<double>(), plan_);
    }
    wstack_finalise(*ctx_, acc.as<double>(), nxp, nyp, np, cx, cy, img_[0], img_[1], dout);
    YG_HIP(hipStreamSynchronize(ctx_->stream));
  }
  void modelToGrids(const float* dmodel, const double* cx, const double* cy) override {
    const int nxp = shape_[0], nyp = shape_[1], np = shape_[2] * shape_[3];
    const size_t ng = gridPlaneElems();
    scratch_.resize(ng * sizeof(double2));
    for (int iw = 0; iw < ws_.nw; ++iw) {
      float w2pi, cellx, celly;
      screen(iw, w2pi, cellx, celly);
      wstack_model_to_grid(*ctx_, dmodel, img_[0], img_[1], np, cx, cy, nxp, nyp, w2pi, cellx, celly,
                           scratch_.as<double2>(), grid_.as<float2>() + ng * iw, plan_);
    }
  }

 private:
  WSamplingDesc ws_;
};

std::unique_ptr<Gridder> make_gridder(std::shared_ptr<Ctx> ctx, const std::string& name, const ParamMap& parset) {
  const Params p{parset};
  std::unique_ptr<Gridder> g;
  if (name == "GPUBox") {
    g.reset(new BoxGridder(ctx));
  } else if (name == "GPUSphFunc") {
    // SphFuncVisGridder::createGridder (SphFuncVisGridder.cc:67-76)
    g.reset(new SphFuncGridder(ctx, p.f64("alpha", 1.0), p.i32("support", 3), p.i32("oversample", 128)));
  } else if (name == "GPUWProject") {
    g.reset(new WProjectGridder(ctx, p));
  } else if (name == "GPUWStack") {
    g.reset(new WStackGridder(ctx, p));
  } else if (name == "GPUAWProject") {
    g = make_awproject(ctx, p);
  } else if (name == "GPUAProjectWStack") {
    g = make_aprojectwstack(ctx, p);
  } else {
    throw Error(5, "Unknown gridder " + name);
  }
  g->configure_common(p);
  return g;
}

}  // namespace ygrid
