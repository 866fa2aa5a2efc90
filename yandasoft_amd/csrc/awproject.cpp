// AWProject gridder: A-projection (disk or ATCA illumination) x W-projection convolution functions
// (AWProjectVisGridder.cc, AProjectGridderBase.cc/.tcc, DiskIllumination.cc, ATCAIllumination.cc),
// built on the device; and its w-stacking form, AProjectWStack (AProjectWStackVisGridder.cc):
// A-projection CFs without the w term, one grid per w plane.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

#include "gridders.h"
#include "illumination.h"

namespace ygrid {

class AWProjectGridder : public Gridder {
 public:
  // wstack: AProjectWStackVisGridder (its createGridder is createAProjectGridder alone:
  // AProjectWStackVisGridder.cc:490-493, AProjectGridderBase.tcc:53-115)
  AWProjectGridder(std::shared_ptr<Ctx> c, const Params& p, bool wstack = false) : Gridder(std::move(c)), wstack_(wstack) {
    // AProjectGridderBase::createAProjectGridder (AProjectGridderBase.tcc:53-115)
    pointing_tol_ = quantity(p.str("pointingtolerance", "0.0001rad"), "rad");
    pa_tol_ = quantity(p.str("patolerance", "0.1rad"), "rad");
    std::string ft = p.str("freqtolerance", "undefined");
    std::transform(ft.begin(), ft.end(), ft.begin(), ::tolower);
    freq_tol_ = -1.;
    if (ft != "infinite") {
      freq_tol_ = p.f64("freqtolerance", 1e-6);
      YG_CHECK(freq_tol_ >= 0., 1, "Frequency tolerance parameter is supposed to be either a non-negative number or infinite");
    }
    ws_.wmax = p.f64("wmax", 10000.0);
    ws_.nw = p.i32("nwplanes", 65);
    cutoff_ = p.f64("cutoff", 1e-3);
    os_ = p.i32("oversample", 8);
    maxsupport_ = p.i32("maxsupport", 128);
    limitsupport_ = p.i32("limitsupport", 0);
    maxfeeds_ = p.i32("maxfeeds", 1);
    maxfields_ = p.i32("maxfields", 1);
    freqdep_ = p.flag("frequencydependent", true);
    sph_taper_ = p.flag("spheroidaltaper", false);
    sph_wcut_ = p.f64("spheroidalweightscutoff", 0.0);
    alpha_ = p.f64("alpha", 1.0);
    abs_cutoff_ = p.flag("cutoff.absolute", false);
    // WProjectVisGridder::configureGridder (WProjectVisGridder.cc:331-363), which
    // AWProjectVisGridder::createGridder calls (AWProjectVisGridder.cc:883-888)
    variable_ = !wstack && p.flag("variablesupport", false);
    offset_ = !wstack && p.flag("offsetsupport", false);
    YG_CHECK(!offset_ || variable_, 1,
             "offsetsupport option of the gridder should only be used together with variablesupport option");
    YG_CHECK(ws_.wmax > 0.0, 1, "Baseline length must be greater than zero");
    YG_CHECK(ws_.nw > 0 && ws_.nw % 2 == 1, 1, "Number of w planes must be a positive odd number");
    YG_CHECK(os_ > 0 && maxsupport_ > 0 && maxsupport_ % 2 == 0, 1, "bad oversample / maxsupport");
    YG_CHECK(maxfeeds_ > 0 && maxfields_ > 0, 1, "maxfeeds and maxfields must be positive");
    if (!abs_cutoff_ && !wstack) YG_CHECK(cutoff_ > 0.0 && cutoff_ < 1.0, 1, "Cutoff must be in (0, 1)");
    YG_CHECK(!(wstack && sph_taper_), 1, "Spheroidal taper not currently availble for AProjectWStack");
    // makeIllumination (AProjectGridderBase.cc:290-348): the disk and the ATCA models
    const std::string ill = p.str("illumination", "disk");
    YG_CHECK(ill == "disk" || ill == "ATCA", 1,
             "GPUAWProject has the disk and ATCA illuminations, not " + ill + " (Unknown illumination type)");
    YG_CHECK(p.has("diameter") && p.has("blockage"), 1, ill + " illumination needs diameter and blockage");
    diameter_ = quantity(p.str("diameter", "12m"), "m");
    blockage_ = quantity(p.str("blockage", "0m"), "m");
    YG_CHECK(diameter_ > 0 && blockage_ >= 0 && diameter_ > blockage_, 1, "bad disk diameter/blockage");
    if (ill == "ATCA") {
      atca_ = true;
      // ATCAIllumination::simulateTapering / simulateFeedLegShadows / simulateFeedLegWedges and
      // their parset defaults (AProjectGridderBase.cc:305-345, ATCAIllumination.cc:37-74)
      a_taper_ = p.flag("illumination.tapering", true);
      if (a_taper_) a_defocus_ = quantity(p.str("illumination.tapering.defocusing", "0rad"), "rad");
      a_legs_ = p.flag("illumination.feedlegs", true);
      if (a_legs_) {
        a_leg_half_ = quantity(p.str("illumination.feedlegs.width", "1.8m"), "m") / 2.;
        a_leg_rot_ = quantity(p.str("illumination.feedlegs.rotation", "45deg"), "rad");
        a_leg_shadow_ = p.f64("illumination.feedlegs.shadowing", 0.75);
        YG_CHECK(a_leg_shadow_ <= 1. && a_leg_shadow_ >= 0., 1,
                 "shadowingFactor is supposed to be from [0,1] interval");
        a_wedges_ = p.flag("illumination.feedlegs.wedges", true);
        if (a_wedges_) {
          std::vector<double> ws = doubles(p.str("illumination.feedlegs.wedges.shadowing", "[0.6,0.5]"));
          YG_CHECK(!ws.empty() && ws.size() < 3, 1,
                   "illumination.feedlegs.wedges.shadowing can have either 1 or 2 elements only");
          if (ws.size() == 1) ws.push_back(ws[0]);
          a_wedge_s1_ = ws[0];
          a_wedge_s2_ = ws[1];
          a_wedge_angle_ = quantity(p.str("illumination.feedlegs.wedges.angle", "15deg"), "rad");
          a_wedge_r0_ = quantity(p.str("illumination.feedlegs.wedges.startradius", "3.5m"), "m");
          YG_CHECK(a_wedge_s1_ <= 1. && a_wedge_s1_ >= 0. && a_wedge_s2_ <= 1. && a_wedge_s2_ >= 0., 1,
                   "wedge shadowing factors are supposed to be from [0,1] interval");
          YG_CHECK(a_wedge_angle_ < kPi / 2. && a_wedge_angle_ > 0, 1,
                   "Opening angle is supposed to be from the (0, pi/2) interval");
          YG_CHECK(a_wedge_r0_ < diameter_ / 2., 1, "the wedge starting radius must lie inside the dish");
        }
        YG_CHECK(a_leg_half_ < diameter_ / 2., 1, "the feed-leg width must be less than the dish diameter");
      }
    }
    // configureWSampling (AProjectGridderBase.tcc:112): linear, power-law or gaussian planes
    configure_wsampling(ws_, p);
    done_.assign((size_t)maxfeeds_ * maxfields_, 0);
    pointings_.assign((size_t)maxfeeds_ * maxfields_ * 2, 0.0);
  }
  std::unique_ptr<Gridder> clone() const override {
    auto g = std::unique_ptr<AWProjectGridder>(new AWProjectGridder(*this));
    return g;
  }
  std::string name() const override { return wstack_ ? "GPUAProjectWStack" : "GPUAWProject"; }

 protected:
  int oversampling() const override { return os_; }
  // CF planes before oversampling: AWProject w x feeds x fields x channels; AProjectWStack
  // feeds x fields x channels (its w term is in the grids)
  int nZPlanes() const override { return (wstack_ ? 1 : ws_.nw) * maxfeeds_ * maxfields_ * std::max(1, nchan_cf_); }
  int nGridPlanes() const override { return wstack_ ? ws_.nw : 1; }
  bool needsRowMeta() const override { return true; }
  bool cfPerChunk() const override { return true; }
  // AProjectWStackVisGridder::correctConvolution is empty (AProjectWStackVisGridder.cc:479-480)
  bool doesCorrection() const override { return !wstack_ && sph_taper_; }
  void gridsToImage(const double* cx, const double* cy, float* out) override;
  void modelToGrids(const float* model, const double* cx, const double* cy) override;
  void gridCorrection(std::vector<double>& cx, std::vector<double>& cy) const override {
    if (sph_taper_) sph_grid_correction_xy(shape_[0], shape_[1], alpha_, 3, cx, cy);
  }
  CIndexDesc cindex(const ChunkDev&) const override {
    CIndexDesc d;
    d.kind = wstack_ ? 4 : 2;
    d.nw = ws_.nw;
    d.wscale = ws_.wscale();
    d.wmax = ws_.wmax;
    d.sampling = ws_.kind;
    d.exponent = ws_.exponent;
    d.g_ts = ws_.ts;
    d.g_amp = ws_.amp;
    d.wmaxclip = ws_.wmaxclip ? 1 : 0;
    d.freqdep = freqdep_ ? 1 : 0;
    d.nchan_cf = std::max(1, nchan_cf_);
    d.maxfeeds = maxfeeds_;
    d.field = field_;
    return d;
  }
  void initConvolutionFunction(const ChunkDev& acc) override;
  void finaliseWeights(float* out, int mem) override;

 private:
  // WDependentGridderBase::getWTerm (WDependentGridderBase.cc:157-167)
  double wterm(int plane) const { return ws_.wterm(plane); }
  // ATCAIllumination::isSymmetric: without feed legs the pattern does not rotate with the
  // parallactic angle; the disk never does
  bool symmetric() const { return !atca_ || !a_legs_; }
  // a parset vector "[a, b, ...]" (or "a, b") of numbers
  static std::vector<double> doubles(std::string s) {
    for (char& ch : s)
      if (ch == '[' || ch == ']' || ch == ',') ch = ' ';
    std::istringstream is(s);
    std::vector<double> v;
    double x;
    while (is >> x) v.push_back(x);
    YG_CHECK(is.eof(), 1, "bad number list: " + s);
    return v;
  }
  void indexField(const ChunkDev& acc);
  void validateCache(const ChunkDev& acc);
  void generateFeed(const ChunkDev& acc, int feed, double lon, double lat, double pa);
  // the aperture pattern of one channel (DiskIllumination::getPattern / ATCAIllumination::getPattern
  // parameters) on an nx x ny buffer
  IllumParams illumParams(double freq, double l, double m, double pa, int nx, int ny) const;
  // AProjectWStackVisGridder::initConvolutionFunction for one (feed, field) (:258-370)
  void generateWStackFeed(const ChunkDev& acc, int feed, double lon, double lat, double pa);
  std::vector<int> usedPlanes();
  void wscreen(int iw, float& w2pi, float& cellx, float& celly) const;
  bool wstack_ = false;
  void ensureTable(int S);
  void writePCFPlanes(const Ctx& c, DevBuf& buf, int nx, int ny, int cnt, int z0, int chan, int nchan, int feed);
  int pcf_support_ = 0;  // preconditioner form: the support found on the first plane
  // Plane-dependent support (variablesupport, offsetsupport).  Each plane's support (and offset)
  // is found on the plane itself; a (feed, field)'s planes are cut out into pieces of their own
  // sizes and the table is rebuilt around them (VarPiece: the first z, the z count, the piece).
  struct VarPiece {
    int z0, cnt;
    size_t elems;
    DevBuf data;
  };
  // search the support (and offset) of the cnt planes in buf, as extractSupport +
  // limitSupportIfNecessary + calcConvFuncOverflow (AWProjectVisGridder.cc:444-496)
  void planeSupports(const Ctx& c, const DevBuf& buf, int nx, int ny, int cnt, int zfirst, std::vector<int>& sz,
                     std::vector<int>& ou, std::vector<int>& ov);
  void cutVariable(const Ctx& c, const DevBuf& buf, int nx, int ny, int cnt, int zfirst, std::vector<VarPiece>& out);
  void commitVariable(int feed, std::vector<VarPiece>& pieces);
  bool variable_ = false, offset_ = false;
  int first_support_ = 0;                  // itsSupport: the first plane's support
  std::vector<int> zsupp_, zou_, zov_;     // per z plane before oversampling (0: not generated)

  double pointing_tol_ = 1e-4, pa_tol_ = 0.1, freq_tol_ = 1e-6;
  WSamplingDesc ws_{10000.0, 65};
  double cutoff_ = 1e-3, alpha_ = 1.0, diameter_ = 12, blockage_ = 0, sph_wcut_ = 0;
  int os_ = 8, maxsupport_ = 128, limitsupport_ = 0, maxfeeds_ = 1, maxfields_ = 1;
  bool freqdep_ = true, sph_taper_ = false, abs_cutoff_ = false;
  int nchan_cf_ = 0, field_ = 0, last_field_ = -1;
  // ATCA illumination (atca_ false: the disk)
  bool atca_ = false, a_taper_ = false, a_legs_ = false, a_wedges_ = false;
  double a_defocus_ = 0, a_leg_half_ = 0, a_leg_rot_ = 0, a_leg_shadow_ = 1, a_wedge_s1_ = 1, a_wedge_s2_ = 1,
         a_wedge_angle_ = 0, a_wedge_r0_ = 0;
  float cf_pa_ = 0.f;  // itsCFParallacticAngle (float): the angle the cache was last built for
  bool invalid_pa_ = false, invalid_freq_ = false;  // itsCFInvalidDueToPA / DueToFreq of this chunk
  std::vector<char> done_;
  std::vector<double> pointings_;
  std::vector<double> cached_freq_;
};

// AProjectGridderBase::indexField (AProjectGridderBase.cc:166-194): the field of the first row.
void AWProjectGridder::indexField(const ChunkDev& acc) {
  const int feed = acc.feed_h.empty() ? 0 : acc.feed_h[0];
  YG_CHECK(feed >= 0 && feed < maxfeeds_, 1, "Too many feeds: increase maxfeeds");
  const double lon = acc.point_h.empty() ? centre_[0] : acc.point_h[0];
  const double lat = acc.point_h.empty() ? centre_[1] : acc.point_h[1];
  for (int f = last_field_; f > -1; --f) {
    const double* pt = &pointings_[((size_t)feed * maxfields_ + f) * 2];
    if (separation(lon, lat, pt[0], pt[1]) < pointing_tol_) {
      field_ = f;
      return;
    }
  }
  ++last_field_;
  field_ = last_field_;
  YG_CHECK(field_ < maxfields_, 1, "Too many fields: increase maxfields");
  double* pt = &pointings_[((size_t)feed * maxfields_ + field_) * 2];
  pt[0] = lon;
  pt[1] = lat;
}

// AProjectGridderBase::validateCFCache (AProjectGridderBase.cc:202-253): a pattern that rotates
// with the parallactic angle invalidates the whole cache at the first row whose angle differs from
// the cached one by more than the tolerance (and caches that row's angle); otherwise a change of
// the frequencies beyond the relative tolerance does.
void AWProjectGridder::validateCache(const ChunkDev& acc) {
  invalid_pa_ = invalid_freq_ = false;
  bool& invalid_pa = invalid_pa_;
  if (!symmetric()) {
    for (size_t row = 0; row < acc.pa_h.size(); ++row)
      if ((double)std::fabs(acc.pa_h[row] - cf_pa_) > pa_tol_) {  // float difference, as the reference
        invalid_pa = true;
        cf_pa_ = acc.pa_h[row];
        std::fill(done_.begin(), done_.end(), 0);
        break;
      }
  }
  bool& invalid_freq = invalid_freq_;
  if (!invalid_pa && freq_tol_ >= 0.) {
    invalid_freq = acc.freq_host.size() != cached_freq_.size();
    for (size_t c = 0; !invalid_freq && c < acc.freq_host.size(); ++c)
      invalid_freq = std::fabs(cached_freq_[c] - acc.freq_host[c]) / acc.freq_host[c] > freq_tol_;
    if (invalid_freq) std::fill(done_.begin(), done_.end(), 0);
  }
  if ((invalid_pa || invalid_freq) && freq_tol_ >= 0.) cached_freq_ = acc.freq_host;
}

void AWProjectGridder::ensureTable(int S) {
  if (cf_ && cf_.use_count() == 1 && cf_->support == S) return;
  auto t = std::make_shared<CFTable>();
  const int nz = nZPlanes();
  t->nz = nz;
  t->oversample = os_;
  t->support = S;
  const size_t blk = (size_t)cf_pitch(S) * cf_rows(S);
  const size_t n = (size_t)nz * os_ * os_;
  YG_CHECK(n * blk < (size_t)INT_MAX, 1, "convolution function table exceeds 2^31 elements");
  t->off.resize(n);
  t->supp.assign(n, S);
  for (size_t k = 0; k < n; ++k) t->off[k] = (int32_t)(k * blk);
  t->n_elems = n * blk;
  t->table.resize(t->n_elems * sizeof(float2));
  if (cf_ && cf_->support == S && cf_->n_elems == t->n_elems) {
    // copy on write: a clone shares the cache until it has to add planes
    YG_HIP(hipMemcpyAsync(t->table.get(), cf_->table.get(), t->n_elems * sizeof(float2), hipMemcpyDeviceToDevice,
                          ctx_->stream));
  } else {
    YG_HIP(hipMemsetAsync(t->table.get(), 0, t->n_elems * sizeof(float2), ctx_->stream));
  }
  t->upload(*ctx_);
  cf_ = t;
}

// The body of AWProjectVisGridder::initConvolutionFunction for one (feed, field)
// (AWProjectVisGridder.cc:366-573): per channel, the disk pattern is transformed to the image
// plane; per w plane, |A|^2 x taper x phase screen is transformed back, the support is found on
// the first plane ever generated, and the os^2 cut-outs are scaled by os^2.
// DiskIllumination::getPattern (DiskIllumination.cc:85-177), ATCAIllumination::getPattern
// (ATCAIllumination.cc:159-331): the parameters of one channel's pattern on an nx x ny buffer of
// cells uvcell / os, with the normalisation of the pattern
IllumParams AWProjectGridder::illumParams(double freq, double l, double m, double pa, int nx, int ny) const {
  const double cellU = uvcell_[0] / os_, cellV = uvcell_[1] / os_;
  YG_CHECK(std::fabs(std::fabs(cellU / cellV) - 1.) < 1e-7, 1, "Rectangular cells are not supported");
  const double cell = std::fabs(cellU * (kC / freq));
  const double rmax2 = (diameter_ / (2.0 * cell)) * (diameter_ / (2.0 * cell));
  double rmin2 = (blockage_ / (2.0 * cell)) * (blockage_ / (2.0 * cell));
  IllumParams ip{};
  ip.lscaled = 2. * kPi * cellU * l;
  ip.mscaled = 2. * kPi * cellV * m;
  if (!atca_) {
    YG_CHECK(rmax2 > rmin2, 1, "Disk hole is supposed to be less than the disk size");
    YG_CHECK(double(nx) * nx > rmax2 && double(ny) * ny > rmax2, 1,
             "The pattern buffer is too small for the dish: increase maxsupport");
    if (rmin2 + 1 >= rmax2) --rmin2;
    ip.kind = kIllumDisk;
  } else {
    YG_CHECK(double(nx) * nx > rmax2 || double(ny) * ny > rmax2, 1,
             "The pattern buffer is too small for the dish: increase maxsupport");
    ip.kind = kIllumATCA;
    ip.taper = a_taper_ ? 1 : 0;
    ip.legs = a_legs_ ? 1 : 0;
    ip.wedges = a_legs_ && a_wedges_ ? 1 : 0;
    ip.defocus = a_defocus_;
    ip.leg_half = a_leg_half_ / cell;
    ip.leg_angle = a_leg_rot_ + pa;
    ip.leg_shadow = a_leg_shadow_;
    ip.wedge_tan = std::tan(a_wedge_angle_);
    ip.wedge_r0 = a_wedge_r0_ / cell;
    ip.wedge_s1 = a_wedge_s1_;
    ip.wedge_s2 = a_wedge_s2_;
  }
  ip.rmin2 = rmin2;
  ip.rmax2 = rmax2;
  // the normalisation: the disk counts its cells (in float), ATCA sums |value|^2 (in double)
  double sum = 0;
  for (int iu = 0; iu < nx; ++iu) {
    const double ou = double(iu) - double(nx) / 2.;
    for (int iv = 0; iv < ny; ++iv) {
      const double ov = double(iv) - double(ny) / 2.;
      const double r2 = ou * ou + ov * ov;
      if (r2 >= rmin2 && r2 <= rmax2) {
        double vr, vi;
        illum_value(ip, ou, ov, r2, vr, vi);
        sum += atca_ ? vr * vr + vi * vi : 1.;
      }
    }
  }
  if (!atca_) {
    YG_CHECK(sum > 0., 4, "Integral of the disk should be non-zero");
    ip.scale = (float)(float(nx) * float(ny) / float(sum));
  } else {
    YG_CHECK(sum > 0., 4, "Integral of the pattern should be non-zero");
    ip.scale = (float)(double(nx) * double(ny) / std::sqrt(sum));
  }
  return ip;
}

void AWProjectGridder::generateFeed(const ChunkDev& acc, int feed, double lon, double lat, double pa) {
  if (wstack_) {
    generateWStackFeed(acc, feed, lon, lat, pa);
    return;
  }
  const Ctx& c = *ctx_;
  const int nx = maxsupport_, ny = maxsupport_;
  const int qnx = nx / os_, qny = ny / os_;
  const double ccellx = 1.0 / (double(qnx) * uvcell_[0]);
  const double ccelly = 1.0 / (double(qny) * uvcell_[1]);
  const bool psf = dopsf_ || dopcf_;
  double l = 0., m = 0.;
  if (!psf) {  // DiskIllumination.cc:111-116
    l = std::sin(lon - centre_[0]) * std::cos(lat);
    m = std::sin(lat) * std::cos(centre_[1]) - std::cos(lat) * std::sin(centre_[1]) * std::cos(lon - centre_[0]);
  }
  DevBuf dtx, dty;
  if (sph_taper_) {
    const Spheroidal sph(kPi * 3, alpha_);
    auto axis = [&](int q) {
      std::vector<double> v(q);
      for (int i = 0; i < q; ++i) v[i] = sph(std::fabs(double(i) - double(q / 2)) / double(q / 2));
      v[0] = 3. * (v[1] - v[2]) + v[3];
      if (q % 2 == 1) v[q - 1] = v[0];
      return v;
    };
    const auto hx = axis(qnx), hy = axis(qny);
    dtx.resize(hx.size() * 8);
    dty.resize(hy.size() * 8);
    YG_HIP(hipMemcpy(dtx.get(), hx.data(), hx.size() * 8, hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(dty.get(), hy.data(), hy.size() * 8, hipMemcpyHostToDevice));
  }
  std::vector<double> tpw(ws_.nw);
  for (int iw = 0; iw < ws_.nw; ++iw) tpw[iw] = 2.0f * kPi * wterm(iw);
  DevBuf dtpw(ws_.nw * sizeof(double));
  YG_HIP(hipMemcpy(dtpw.get(), tpw.data(), ws_.nw * sizeof(double), hipMemcpyHostToDevice));
  const size_t plane_bytes = (size_t)nx * ny * sizeof(float2);
  int nb = (int)std::max<size_t>(1, ((size_t)1 << 29) / plane_bytes);
  nb = std::min(nb, ws_.nw);
  DevBuf pat(plane_bytes), buf(plane_bytes * nb), dprm(sizeof(IllumParams));
  PlanPtr plan1, planb;
  const int nchan = nchan_cf_;
  std::vector<VarPiece> pieces;  // variablesupport: this (feed, field)'s planes
  for (int chan = 0; chan < nchan; ++chan) {
    const IllumParams ip = illumParams(acc.freq_host[chan], l, m, pa, nx, ny);
    YG_HIP(hipMemcpyAsync(dprm.get(), &ip, sizeof(ip), hipMemcpyHostToDevice, c.stream));
    launch_illum_pattern(c, pat.as<float2>(), nx, ny, dprm.as<IllumParams>(), 1);
    // fft2d(pattern, false): backward, normalised by 1/(nx ny) -> folded into the screen kernel
    fft_plain_c2c(c, pat.as<float2>(), nx, ny, 1, false, plan1);
    for (int z0 = 0; z0 < ws_.nw; z0 += nb) {
      const int cnt = std::min(nb, ws_.nw - z0);
      launch_aw_screen(c, pat.as<float2>(), nx, ny, qnx, qny, ccellx, ccelly, (float)(1.0 / (double(nx) * ny)),
                       sph_taper_ ? dtx.as<double>() : nullptr, sph_taper_ ? dty.as<double>() : nullptr,
                       dtpw.as<double>() + z0, cnt, buf.as<float2>());
      fft_plain_c2c(c, buf.as<float2>(), nx, ny, cnt, true, planb);
      launch_scale_sign(c, buf.as<float2>(), nx, ny, cnt, (float)(1.0 / (double(nx) * ny)));
      if (dopcf_) {
        writePCFPlanes(c, buf, nx, ny, cnt, z0, chan, nchan, feed);
        continue;
      }
      if (variable_) {
        cutVariable(c, buf, nx, ny, cnt, z0 + ws_.nw * (chan + nchan * (feed + maxfeeds_ * field_)), pieces);
        continue;
      }
      if (!cf_ || cf_->support == 0) {
        // support from the first plane generated (AWProjectVisGridder.cc:436-466)
        const SupportResult r = search_support(c, buf.as<float2>(), nx, ny, cutoff_, abs_cutoff_, false, os_);
        YG_CHECK((r.size + 1) * os_ < nx / 2, 4,
                 "Overflowing convolution function - increase maxSupport or cutoff or decrease overSample");
        int S = r.size;
        if (limitsupport_ > 0 && S > limitsupport_) S = limitsupport_;
        ensureTable(S);
      } else {
        ensureTable(cf_->support);
      }
      const int S = cf_->support;
      const int kmin = -S * os_ + nx / 2, kmax = S * os_ + os_ - 1 + nx / 2;
      YG_CHECK(kmin >= 0 && kmax < nx && kmax < ny, 4,
               "Convolution function overflowing - increase maxsupport or cutoff or decrease oversample");
      std::vector<int32_t> first(cnt);
      std::vector<int> sup(cnt, S), zero(cnt, 0);
      for (int b = 0; b < cnt; ++b) {
        const int z = (z0 + b) + ws_.nw * (chan + nchan * (feed + maxfeeds_ * field_));
        first[b] = cf_->off[(size_t)z * os_ * os_];
      }
      DevBuf dfirst(cnt * 4), dsup(cnt * 4), dzero(cnt * 4);
      YG_HIP(hipMemcpy(dfirst.get(), first.data(), cnt * 4, hipMemcpyHostToDevice));
      YG_HIP(hipMemcpy(dsup.get(), sup.data(), cnt * 4, hipMemcpyHostToDevice));
      YG_HIP(hipMemcpy(dzero.get(), zero.data(), cnt * 4, hipMemcpyHostToDevice));
      launch_cf_cutout(c, buf.as<float2>(), nx, ny, cnt, os_, dsup.as<int>(), dzero.as<int>(), dzero.as<int>(),
                       dfirst.as<int32_t>(), 1, double(os_) * os_, cf_->table.as<float2>());
      YG_HIP(hipStreamSynchronize(c.stream));
    }
  }
  if (variable_ && !dopcf_) commitVariable(feed, pieces);
}

void AWProjectGridder::planeSupports(const Ctx& c, const DevBuf& buf, int nx, int ny, int cnt, int zfirst,
                                     std::vector<int>& sz, std::vector<int>& ou, std::vector<int>& ov) {
  sz.assign(cnt, 0);
  ou.assign(cnt, 0);
  ov.assign(cnt, 0);
  for (int b = 0; b < cnt; ++b) {
    const SupportResult r =
        search_support(c, buf.as<float2>() + (size_t)b * nx * ny, nx, ny, cutoff_, abs_cutoff_, offset_, os_);
    YG_CHECK((r.size + 1) * os_ < nx / 2, 4,
             "Overflowing convolution function - increase maxSupport or cutoff or decrease overSample");
    int S = r.size;
    if (limitsupport_ > 0 && S > limitsupport_) S = limitsupport_;
    if (first_support_ == 0) first_support_ = S;
    if (offset_) {
      ou[b] = r.offu;
      ov[b] = r.offv;
    }
    sz[b] = S;
    const int kxmin = (-S + ou[b]) * os_ + nx / 2, kxmax = (S + ou[b]) * os_ + os_ - 1 + nx / 2;
    const int kymin = (-S + ov[b]) * os_ + ny / 2, kymax = (S + ov[b]) * os_ + os_ - 1 + ny / 2;
    YG_CHECK(kxmin >= 0 && kymin >= 0 && kxmax < nx && kymax < ny, 4,
             "Convolution function overflowing - increase maxsupport or cutoff or decrease oversample");
    const int z = zfirst + b;
    zsupp_[z] = S;
    zou_[z] = ou[b];
    zov_[z] = ov[b];
  }
}

// The cnt planes of z = zfirst .. zfirst + cnt - 1 cut out at their own supports and offsets (each
// z's os^2 planes consecutive, as the table holds them) into one piece.
void AWProjectGridder::cutVariable(const Ctx& c, const DevBuf& buf, int nx, int ny, int cnt, int zfirst,
                                   std::vector<VarPiece>& out) {
  std::vector<int> sz, ou, ov;
  planeSupports(c, buf, nx, ny, cnt, zfirst, sz, ou, ov);
  std::vector<int32_t> first(cnt);
  size_t total = 0;
  for (int b = 0; b < cnt; ++b) {
    first[b] = (int32_t)total;
    total += (size_t)os_ * os_ * cf_pitch(sz[b]) * cf_rows(sz[b]);
    YG_CHECK(total < (size_t)INT_MAX, 1, "convolution function table exceeds 2^31 elements");
  }
  DevBuf dfirst(cnt * 4), dsup(cnt * 4), dou(cnt * 4), dov(cnt * 4);
  YG_HIP(hipMemcpy(dfirst.get(), first.data(), cnt * 4, hipMemcpyHostToDevice));
  YG_HIP(hipMemcpy(dsup.get(), sz.data(), cnt * 4, hipMemcpyHostToDevice));
  YG_HIP(hipMemcpy(dou.get(), ou.data(), cnt * 4, hipMemcpyHostToDevice));
  YG_HIP(hipMemcpy(dov.get(), ov.data(), cnt * 4, hipMemcpyHostToDevice));
  VarPiece pc{zfirst, cnt, total, DevBuf(total * sizeof(float2))};
  launch_cf_cutout(c, buf.as<float2>(), nx, ny, cnt, os_, dsup.as<int>(), dou.as<int>(), dov.as<int>(),
                   dfirst.as<int32_t>(), 1, double(os_) * os_, pc.data.as<float2>());
  YG_HIP(hipStreamSynchronize(c.stream));
  out.push_back(std::move(pc));
}

// A new table holding every plane generated so far at its own support: the planes of the other
// (feed, field)s copied from the current table, those of `feed` (this field) from the pieces.
// Planes not generated yet take no space (support 0, never selected by a sample).  The current
// table is left untouched for the clones that share it.
void AWProjectGridder::commitVariable(int feed, std::vector<VarPiece>& pieces) {
  const Ctx& c = *ctx_;
  const int nz = ws_.nw * maxfeeds_ * maxfields_ * nchan_cf_;
  const size_t os2 = (size_t)os_ * os_;
  const int zb = ws_.nw * nchan_cf_ * (feed + maxfeeds_ * field_), ze = zb + ws_.nw * nchan_cf_;
  auto t = std::make_shared<CFTable>();
  t->nz = nz;
  t->oversample = os_;
  t->off.assign((size_t)nz * os2, 0);
  t->supp.assign((size_t)nz * os2, 0);
  size_t total = 0, new_zb = 0, new_ze = 0;
  t->support = 0;
  for (int z = 0; z < nz; ++z) {
    if (z == zb) new_zb = total;
    if (z == ze) new_ze = total;
    const int S = zsupp_[z];
    const size_t blk = S > 0 ? (size_t)cf_pitch(S) * cf_rows(S) : 0;
    for (size_t f = 0; f < os2; ++f) {
      YG_CHECK(total < (size_t)INT_MAX, 1, "convolution function table exceeds 2^31 elements");
      t->off[(size_t)z * os2 + f] = (int32_t)total;
      t->supp[(size_t)z * os2 + f] = S;
      total += blk;
    }
    t->support = std::max(t->support, S);
  }
  if (ze == nz) new_ze = total;
  t->n_elems = total;
  if (offset_) {
    t->ofs_u.assign(zou_.begin(), zou_.end());
    t->ofs_v.assign(zov_.begin(), zov_.end());
  }
  t->table.resize(std::max<size_t>(total, 1) * sizeof(float2));
  float2* dst = t->table.as<float2>();
  if (cf_ && cf_->n_elems > 0) {
    // the planes before and after the block keep their sizes: two contiguous runs
    const size_t old_zb = cf_->off[(size_t)zb * os2];
    const size_t old_ze = ze < nz ? (size_t)cf_->off[(size_t)ze * os2] : cf_->n_elems;
    const float2* src = cf_->table.as<float2>();
    if (old_zb > 0)
      YG_HIP(hipMemcpyAsync(dst, src, old_zb * sizeof(float2), hipMemcpyDeviceToDevice, c.stream));
    if (cf_->n_elems > old_ze)
      YG_HIP(hipMemcpyAsync(dst + new_ze, src + old_ze, (cf_->n_elems - old_ze) * sizeof(float2),
                            hipMemcpyDeviceToDevice, c.stream));
  }
  for (const VarPiece& pc : pieces)
    YG_HIP(hipMemcpyAsync(dst + t->off[(size_t)pc.z0 * os2], pc.data.get(), pc.elems * sizeof(float2),
                          hipMemcpyDeviceToDevice, c.stream));
  (void)new_zb;
  t->upload(c);
  cf_ = t;
  pieces.clear();
}

// AProjectWStackVisGridder::initConvolutionFunction for one (feed, field) (:258-370): per channel
// the pattern (on min(maxsupport, grid) cells of uvcell / os) is transformed to the image plane,
// its power |A|^2 normalised to a peak of 1 is transformed back (/ nx ny), and cut out at
// k = t os + frac about the centre for t in [-S, S-1] (the last row and column of each plane stay
// zero), times os^2.  S is set once, from the first channel's pattern: the dish diameter in
// pattern cells, 1 + 2 int(R) / os (DiskIllumination.cc:139, ATCAIllumination.cc:245).
void AWProjectGridder::generateWStackFeed(const ChunkDev& acc, int feed, double lon, double lat, double pa) {
  const Ctx& c = *ctx_;
  const int nx = std::min(maxsupport_, shape_[0]), ny = std::min(maxsupport_, shape_[1]);
  const bool psf = dopsf_ || dopcf_;
  double l = 0., m = 0.;
  if (!psf) {
    l = std::sin(lon - centre_[0]) * std::cos(lat);
    m = std::sin(lat) * std::cos(centre_[1]) - std::cos(lat) * std::sin(centre_[1]) * std::cos(lon - centre_[0]);
  }
  const size_t plane_bytes = (size_t)nx * ny * sizeof(float2);
  DevBuf pat(plane_bytes), buf(plane_bytes), dprm(sizeof(IllumParams)), tmp(sizeof(uint32_t));
  DevBuf dfirst(4), dsup(4);
  PlanPtr plan1, plan2;
  for (int chan = 0; chan < nchan_cf_; ++chan) {
    const double freq = acc.freq_host[chan];
    const IllumParams ip = illumParams(freq, l, m, pa, nx, ny);
    YG_HIP(hipMemcpyAsync(dprm.get(), &ip, sizeof(ip), hipMemcpyHostToDevice, c.stream));
    launch_illum_pattern(c, pat.as<float2>(), nx, ny, dprm.as<IllumParams>(), 1);
    fft_plain_c2c(c, pat.as<float2>(), nx, ny, 1, false, plan1);  // fft2d(pattern, false)
    launch_aws_cf_power(c, pat.as<float2>(), nx, ny, (float)(1.0 / (double(nx) * ny)), buf.as<float2>(),
                        tmp.as<uint32_t>());
    fft_plain_c2c(c, buf.as<float2>(), nx, ny, 1, true, plan2);  // fft2d(pattern, true) / (nx ny)
    launch_scale_sign(c, buf.as<float2>(), nx, ny, 1, (float)(1.0 / (double(nx) * ny)));
    if (!cf_ || cf_->support == 0) {
      const double cell = std::fabs(uvcell_[0] / os_ * (kC / freq));
      int S = 1 + (int)(2u * (unsigned)(diameter_ / (2.0 * cell)) / (unsigned)os_);
      YG_CHECK(S > 0, 4, "Unable to determine support of convolution function");
      YG_CHECK(S * os_ < nx / 2, 4, "Overflowing convolution function - increase maxSupport or decrease overSample");
      if (limitsupport_ > 0 && S > limitsupport_) S = limitsupport_;
      ensureTable(S);
    } else {
      ensureTable(cf_->support);
    }
    const int S = cf_->support;
    const int z = chan + nchan_cf_ * (feed + maxfeeds_ * field_);
    const int32_t first = cf_->off[(size_t)z * os_ * os_];
    YG_HIP(hipMemcpy(dfirst.get(), &first, 4, hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(dsup.get(), &S, 4, hipMemcpyHostToDevice));
    launch_cf_cutout(c, buf.as<float2>(), nx, ny, 1, os_, dsup.as<int>(), nullptr, nullptr, dfirst.as<int32_t>(), 2,
                     double(os_) * os_, cf_->table.as<float2>());
    YG_HIP(hipStreamSynchronize(c.stream));
  }
}

// the w planes whose grids received samples (WStackVisGridder.cc:236,272 skip planes whose
// largest amplitude is 0)
std::vector<int> AWProjectGridder::usedPlanes() {
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

// WStackVisGridder::multiply's screen constants (WStackVisGridder.cc:137-144)
void AWProjectGridder::wscreen(int iw, float& w2pi, float& cellx, float& celly) const {
  w2pi = (float)(2.0f * kPi * ws_.wterm(iw));
  cellx = (float)(1.0 / (float(shape_[0]) * uvcell_[0]));
  celly = (float)(1.0 / (float(shape_[1]) * uvcell_[1]));
}

// AProjectWStack: WStackVisGridder::finaliseGrid (WStackVisGridder.cc:255-289) with the empty
// correction: per used w plane the inverse transform times its screen, real parts summed.
void AWProjectGridder::gridsToImage(const double* cx, const double* cy, float* dout) {
  if (!wstack_) {
    Gridder::gridsToImage(cx, cy, dout);
    return;
  }
  const int nxp = shape_[0], nyp = shape_[1], np = shape_[2] * shape_[3];
  const size_t ng = gridPlaneElems();
  const std::vector<int> used = usedPlanes();
  if (dopcf_) {
    // the preconditioner pass: AProjectWStackVisGridder has no form of its own, so its A-term CFs
    // are gridded by the generic preconditioner branch and WStackVisGridder::finaliseGrid's
    // preconditioner form (:222-262) stacks the used planes, imaginary parts scaled by the would-be
    // W-projection kernel size over 3 (updatew, :173-209), into one normalised inverse transform
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
  scratch_.resize(ng * sizeof(double2));
  for (int iw : used) {
    float w2pi, cellx, celly;
    wscreen(iw, w2pi, cellx, celly);
    wstack_accumulate(*ctx_, grid_.as<float2>() + ng * iw, nxp, nyp, np, w2pi, cellx, celly, scratch_.as<double2>(),
                      acc.as<double>(), plan_);
  }
  wstack_finalise(*ctx_, acc.as<double>(), nxp, nyp, np, nullptr, nullptr, img_[0], img_[1], dout);
  YG_HIP(hipStreamSynchronize(ctx_->stream));
}

// AProjectWStack: WStackVisGridder::initialiseDegrid (WStackVisGridder.cc:314-327) without the
// correction: each w plane's grid is the forward transform of conj(screen x model).
void AWProjectGridder::modelToGrids(const float* dmodel, const double* cx, const double* cy) {
  if (!wstack_) {
    Gridder::modelToGrids(dmodel, cx, cy);
    return;
  }
  const int nxp = shape_[0], nyp = shape_[1], np = shape_[2] * shape_[3];
  const size_t ng = gridPlaneElems();
  scratch_.resize(ng * sizeof(double2));
  for (int iw = 0; iw < ws_.nw; ++iw) {
    float w2pi, cellx, celly;
    wscreen(iw, w2pi, cellx, celly);
    wstack_model_to_grid(*ctx_, dmodel, img_[0], img_[1], np, nullptr, nullptr, nxp, nyp, w2pi, cellx, celly,
                         scratch_.as<double2>(), grid_.as<float2>() + ng * iw, plan_);
  }
}

// Preconditioner form (AWProjectVisGridder.cc:436-466,511-516): the support is found on the first
// plane generated as for the imaging kernels, and every plane becomes 3x3 with the centre tap
// (1, support); the table itself then has support 1.
void AWProjectGridder::writePCFPlanes(const Ctx& c, DevBuf& buf, int nx, int ny, int cnt, int z0, int chan, int nchan,
                                      int feed) {
  const int zfirst = z0 + ws_.nw * (chan + nchan * (feed + maxfeeds_ * field_));
  if (variable_) {
    // each plane's centre carries its own support, and its offset applies when it is gridded
    // (AWProjectVisGridder.cc:444-466,499,511-516)
    std::vector<int> sz, ou, ov;
    planeSupports(c, buf, nx, ny, cnt, zfirst, sz, ou, ov);
    if (pcf_support_ == 0) pcf_support_ = sz[0];
    ensureTable(1);
    const size_t blk = (size_t)cf_pitch(1) * cf_rows(1);
    const size_t k0 = (size_t)zfirst * os_ * os_, nk = (size_t)cnt * os_ * os_;
    std::vector<float2> host(nk * blk, make_float2(0.f, 0.f));
    for (size_t k = 0; k < nk; ++k)
      host[k * blk + (size_t)(1 + kGuardV) * cf_pitch(1) + 1 + kGuardU] =
          make_float2(1.f, (float)sz[k / ((size_t)os_ * os_)]);
    YG_HIP(hipMemcpyAsync(cf_->table.as<float2>() + cf_->off[k0], host.data(), host.size() * sizeof(float2),
                          hipMemcpyHostToDevice, c.stream));
    if (offset_) {
      cf_->ofs_u.assign(zou_.begin(), zou_.end());
      cf_->ofs_v.assign(zov_.begin(), zov_.end());
      cf_->upload(c);
    }
    YG_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  if (pcf_support_ == 0) {
    const SupportResult r = search_support(c, buf.as<float2>(), nx, ny, cutoff_, abs_cutoff_, false, os_);
    YG_CHECK((r.size + 1) * os_ < nx / 2, 4,
             "Overflowing convolution function - increase maxSupport or cutoff or decrease overSample");
    int S = r.size;
    if (limitsupport_ > 0 && S > limitsupport_) S = limitsupport_;
    pcf_support_ = S;
  }
  const int S = pcf_support_;
  const int kmin = -S * os_ + nx / 2, kmax = S * os_ + os_ - 1 + nx / 2;
  YG_CHECK(kmin >= 0 && kmax < nx && kmax < ny, 4,
           "Convolution function overflowing - increase maxsupport or cutoff or decrease oversample");
  ensureTable(1);
  const size_t blk = (size_t)cf_pitch(1) * cf_rows(1);
  const size_t k0 = (size_t)zfirst * os_ * os_, nk = (size_t)cnt * os_ * os_;
  std::vector<float2> host(nk * blk, make_float2(0.f, 0.f));
  for (size_t k = 0; k < nk; ++k)
    host[k * blk + (size_t)(1 + kGuardV) * cf_pitch(1) + 1 + kGuardU] = make_float2(1.f, (float)S);
  // on the context's stream, after the table's clearing
  YG_HIP(hipMemcpyAsync(cf_->table.as<float2>() + cf_->off[k0], host.data(), host.size() * sizeof(float2),
                        hipMemcpyHostToDevice, c.stream));
  YG_HIP(hipStreamSynchronize(c.stream));
}

void AWProjectGridder::initConvolutionFunction(const ChunkDev& acc) {
  YG_CHECK(!acc.freq_host.empty(), 1, "no channels");
  if (nchan_cf_ == 0) nchan_cf_ = freqdep_ ? acc.nchan : 1;
  YG_CHECK(!freqdep_ || acc.nchan == nchan_cf_, 1, "the number of channels changed between chunks");
  if (zsupp_.empty()) {
    const size_t nz = (size_t)ws_.nw * maxfeeds_ * maxfields_ * nchan_cf_;
    zsupp_.assign(nz, 0);
    zou_.assign(nz, 0);
    zov_.assign(nz, 0);
  }
  indexField(acc);
  validateCache(acc);
  int ndone = 0;
  for (int row = 0; row < acc.nrow; ++row) {
    const int feed = acc.feed_h.empty() ? 0 : acc.feed_h[row];
    YG_CHECK(feed >= 0 && feed < maxfeeds_, 1, "Exceeded specified maximum number of feeds");
    char& d = done_[(size_t)feed * maxfields_ + field_];
    if (d) continue;
    d = 1;
    ++ndone;
    const double lon = acc.point_h.empty() ? centre_[0] : acc.point_h[2 * (size_t)row];
    const double lat = acc.point_h.empty() ? centre_[1] : acc.point_h[2 * (size_t)row + 1];
    // the row's parallactic angle where the pattern rotates with it (AWProjectVisGridder.cc:369)
    const double pa = symmetric() || acc.pa_h.empty() ? 0.0 : (double)acc.pa_h[row];
    generateFeed(acc, feed, lon, lat, pa);
  }
  // AProjectGridderBase::updateStats (AProjectGridderBase.cc:272-282)
  ++cf_iter_;
  cf_gen_ += ndone;
  if (invalid_pa_) cf_gen_pa_ += ndone;
  if (invalid_freq_) cf_gen_freq_ += ndone;
}

// AWProjectVisGridder::finaliseWeights (AWProjectVisGridder.cc:618-770): the weights image is
// sum_z wt_z |iFFT(CF_z)|^2 on a min(maxsupport, nx) grid, Fourier-interpolated to the image.
void AWProjectGridder::finaliseWeights(float* out, int mem) {
  YG_CHECK(initialised_ && !degrid_mode_, 2, "finaliseWeights without an initialised grid");
  const Ctx& c = *ctx_;
  const int nx = shape_[0], ny = shape_[1], npol = shape_[2], nchan = shape_[3], ncp = npol * nchan;
  const int cnx = std::min(maxsupport_, nx), cny = std::min(maxsupport_, ny);
  // AProjectWStackVisGridder::finaliseWeights (:373-473): taps -S .. S-1 placed, each transform
  // normalised to its peak before squaring
  std::vector<double> sw((size_t)nsw_rows_ * ncp);
  YG_HIP(hipMemcpyAsync(sw.data(), sumw_.get(), sw.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  YG_HIP(hipStreamSynchronize(c.stream));
  DevBuf cout((size_t)cnx * cny * ncp * sizeof(double));
  YG_HIP(hipMemsetAsync(cout.get(), 0, (size_t)cnx * cny * ncp * sizeof(double), c.stream));
  DevBuf dtx, dty;
  if (sph_taper_) {
    const Spheroidal sph(kPi * 3, alpha_);
    auto axis = [&](int n) {
      std::vector<double> v(n);
      for (int i = 0; i < n; ++i) {
        const double val = sph(std::fabs(double(i) - double(n / 2)) / double(n / 2));
        v[i] = val > sph_wcut_ ? val : 0.;
      }
      v[0] = 3. * (v[1] - v[2]) + v[3];
      if (n % 2 == 1) v[n - 1] = v[0];
      return v;
    };
    const auto hx = axis(cnx), hy = axis(cny);
    dtx.resize(hx.size() * 8);
    dty.resize(hy.size() * 8);
    YG_HIP(hipMemcpy(dtx.get(), hx.data(), hx.size() * 8, hipMemcpyHostToDevice));
    YG_HIP(hipMemcpy(dty.get(), hy.data(), hy.size() * 8, hipMemcpyHostToDevice));
  }
  // rows with data
  std::vector<int> rows;
  for (int z = 0; z < nsw_rows_; ++z) {
    bool has = false;
    for (int cp = 0; cp < ncp; ++cp) has |= sw[z + (size_t)nsw_rows_ * cp] > 0.0;
    if (has) rows.push_back(z);
  }
  if (cf_ && !rows.empty()) {
    const int nb = std::max(1, std::min((int)rows.size(), (int)(((size_t)1 << 28) / ((size_t)cnx * cny * 8))));
    DevBuf planes((size_t)nb * cnx * cny * sizeof(double2)), dw((size_t)nb * ncp * sizeof(double));
    DevBuf doff(nb * 4), dsup(nb * 4), dou(nb * 4), dov(nb * 4);
    DevBuf max2(wstack_ ? (size_t)nb * sizeof(unsigned long long) : 0);
    PlanPtr plan;
    for (size_t r0 = 0; r0 < rows.size(); r0 += nb) {
      const int cnt = (int)std::min<size_t>(nb, rows.size() - r0);
      std::vector<int32_t> off(cnt);
      std::vector<int> sup(cnt), ou(cnt, 0), ov(cnt, 0);
      std::vector<double> w((size_t)cnt * ncp);
      for (int b = 0; b < cnt; ++b) {
        const int z = rows[r0 + b];
        const int k = track_per_os_ ? z : os_ * os_ * z;  // cfIndexFromSumOfWeightsRow
        off[b] = cf_->off[k];
        sup[b] = cf_->supp[k];
        // getConvFuncOffset(iz) takes the sum-of-weights row (AWProjectVisGridder.cc:706), which
        // is the plane before oversampling unless weights are tracked per oversampling plane;
        // rows beyond the offsets give (0, 0) (TableVisGridder.cc:1428-1435)
        if (!cf_->ofs_u.empty() && z < (int)cf_->ofs_u.size()) {
          ou[b] = cf_->ofs_u[z];
          ov[b] = cf_->ofs_v[z];
        }
        for (int cp = 0; cp < ncp; ++cp) w[(size_t)b * ncp + cp] = sw[z + (size_t)nsw_rows_ * cp];
      }
      // the previous batch's kernels read doff/dsup/dw: wait for them before overwriting
      if (r0 > 0) YG_HIP(hipStreamSynchronize(c.stream));
      YG_HIP(hipMemcpy(doff.get(), off.data(), cnt * 4, hipMemcpyHostToDevice));
      YG_HIP(hipMemcpy(dsup.get(), sup.data(), cnt * 4, hipMemcpyHostToDevice));
      YG_HIP(hipMemcpy(dou.get(), ou.data(), cnt * 4, hipMemcpyHostToDevice));
      YG_HIP(hipMemcpy(dov.get(), ov.data(), cnt * 4, hipMemcpyHostToDevice));
      YG_HIP(hipMemcpy(dw.get(), w.data(), w.size() * 8, hipMemcpyHostToDevice));
      launch_place_cf(c, cf_->table.as<float2>(), doff.as<int32_t>(), dsup.as<int>(), dou.as<int>(), dov.as<int>(),
                      cnx, cny, cnt, planes.as<double2>(), wstack_ ? 1 : 0);
      // fft2d(thisPlane, false) * (cnx*cny): the unnormalised backward transform
      fft_plain_z2z(c, planes.as<double2>(), cnx, cny, cnt, false, plan);
      launch_aw_weights_acc(c, planes.as<double2>(), cnt, cnx, cny, dw.as<double>(), ncp,
                            sph_taper_ ? dtx.as<double>() : nullptr, sph_taper_ ? dty.as<double>() : nullptr,
                            cout.as<double>(), wstack_ ? max2.as<unsigned long long>() : nullptr);
    }
  }
  // PaddingUtils::fftPad onto the padded image, then the unpadded centre
  // (the centred backward transform is unnormalised: its 1/(cnx cny) is applied on conversion)
  DevBuf small((size_t)cnx * cny * ncp * sizeof(double2)), big((size_t)nx * ny * ncp * sizeof(double2));
  launch_d2c(c, cout.as<double>(), (size_t)cnx * cny * ncp, 1.0 / (double(cnx) * cny), small.as<double2>());
  PlanPtr p1, p2;
  fft_centred_z2z(c, small.as<double2>(), cnx, cny, ncp, false, p1);
  launch_pad_centre(c, small.as<double2>(), cnx, cny, nx, ny, ncp, big.as<double2>());
  fft_centred_z2z(c, big.as<double2>(), nx, ny, ncp, true, p2);
  const size_t nout = (size_t)img_[0] * img_[1] * ncp;
  DevBuf tmp;
  float* dout = out;
  if (mem != YGRID_MEM_DEVICE) {
    tmp.resize(nout * sizeof(float));
    dout = tmp.as<float>();
  }
  launch_extract_real(c, big.as<double2>(), nx, ny, img_[0], img_[1], ncp, dout);
  if (mem != YGRID_MEM_DEVICE) YG_HIP(hipMemcpyAsync(out, dout, nout * sizeof(float), hipMemcpyDeviceToHost, c.stream));
  YG_HIP(hipStreamSynchronize(c.stream));
}

std::unique_ptr<Gridder> make_awproject(std::shared_ptr<Ctx> ctx, const Params& p) {
  return std::unique_ptr<Gridder>(new AWProjectGridder(std::move(ctx), p));
}

std::unique_ptr<Gridder> make_aprojectwstack(std::shared_ptr<Ctx> ctx, const Params& p) {
  return std::unique_ptr<Gridder>(new AWProjectGridder(std::move(ctx), p, true));
}

}  // namespace ygrid
