// Snapshot imaging (SnapShotImagingGridderAdapter.cc) around a wrapped gridder.
//
// Every chunk is gridded with w taken relative to the current plane w = A u + B v.  The plane is
// refitted (least squares over the chunk's rows) when a row departs from it by more than
// wtolerance; then the snapshot gridded so far is finalised by the wrapped gridder, regridded
// from the plane's slanted frame to the target frame on the device and added to the image and
// weight buffers (:222-258,416-443), and the wrapped gridder starts a new snapshot.  Degridding
// regrids the model into each plane's frame before the wrapped gridder's initialiseDegrid
// (:347-395).  The PSF and preconditioner passes go straight to the wrapped gridder (with the
// plane-relative w) unless reprojectpsf (:195-211).  BestWPlaneDataAccessor (base-accessors) and
// casacore's ImageRegrid are absent: the fit is least squares, the regrid bilinear or Keys cubic
// with exact per-pixel coordinates (kernels_snapshot.hip).
#include "snapshot.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstring>

namespace ygrid {

Snapshot::Snapshot(Gridder* inner, const Params& p) : g_(inner), ctx_(inner->context()) {
  YG_CHECK(p.has("snapshotimaging.wtolerance"), 1, "gridder.snapshotimaging.wtolerance is required");
  tol_ = p.f64("snapshotimaging.wtolerance", 0.);
  YG_CHECK(tol_ >= 0., 1, "gridder.snapshotimaging.wtolerance must be non-negative");
  clip_ = p.f64("snapshotimaging.clipping", 0.);
  wclip_ = p.f64("snapshotimaging.weightsclipping", 0.);
  // setClippingFactor / setWeightsClippingFactor (:615-640)
  YG_CHECK(clip_ >= 0. && clip_ < 1., 1, "Clipping factor should be a non-negative number less than 1");
  YG_CHECK(wclip_ >= 0. && wclip_ < 1., 1, "Clipping factor should be a non-negative number less than 1");
  reproject_psf_ = p.flag("snapshotimaging.reprojectpsf", false);
  std::string m = p.str("snapshotimaging.interpmethod", "cubic");
  std::transform(m.begin(), m.end(), m.begin(), ::tolower);
  YG_CHECK(m == "cubic" || m == "linear", 1,
           "snapshotimaging.interpmethod: cubic or linear (got " + m + ")");
  cubic_ = m == "cubic";
  // coorddecimation: the reference computes coordinates on a decimated grid; here every pixel's
  // coordinates are computed exactly, so the key is accepted and has no effect
  YG_CHECK(p.i32("snapshotimaging.coorddecimation", 3) >= 0, 1, "coorddecimation must be non-negative");
  YG_CHECK(!p.flag("snapshotimaging.longtrack", false), 1,
           "snapshotimaging.longtrack (w-plane prediction) is not supported by the GPU snapshot imaging");
}

Snapshot::Snapshot(const Snapshot& o, Gridder* inner)
    : g_(inner), ctx_(inner->context()), tol_(o.tol_), clip_(o.clip_), wclip_(o.wclip_),
      reproject_psf_(o.reproject_psf_), cubic_(o.cubic_) {}

void Snapshot::copyOut(const DevBuf& b, float* out, int mem) const {
  YG_HIP(hipMemcpyAsync(out, b.get(), imageElems() * sizeof(float),
                        mem == YGRID_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, ctx_->stream));
  YG_HIP(hipStreamSynchronize(ctx_->stream));
}

void Snapshot::initialiseGrid(const ygrid_image_desc& img, bool dopsf, bool dopcf) {
  dopsf_ = dopsf;
  dopcf_ = dopcf;
  img_ = img;
  stokes_.assign(img.stokes, img.stokes ? img.stokes + img.npol : img.stokes);
  img_.stokes = stokes_.empty() ? nullptr : stokes_.data();
  if (passthrough()) {
    g_->initialiseGrid(img_, dopsf, dopcf);
    return;
  }
  YG_CHECK(!dopcf, 1, "the preconditioner pass with snapshotimaging.reprojectpsf is not supported");
  const size_t n = imageElems();
  image_.resize(n * sizeof(float));
  weights_.resize(n * sizeof(float));
  launch_fill_u32(*ctx_, image_.get(), n, 0u);
  launch_fill_u32(*ctx_, weights_.get(), n, 0u);
  first_ = true;
  done_ = true;
}

bool Snapshot::updatePlane(const ygrid_chunk& acc, double& A, double& B) {
  A = A_;
  B = B_;
  const int n = acc.nrow;
  if (n <= 0) return false;
  const double* duvw = acc.uvw;
  if (acc.mem != YGRID_MEM_DEVICE) {
    uvw_in_.resize((size_t)n * 3 * sizeof(double));
    YG_HIP(hipMemcpyAsync(uvw_in_.get(), acc.uvw, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice,
                          ctx_->stream));
    duvw = uvw_in_.as<double>();
  }
  sums_.resize(6 * sizeof(double));
  launch_wplane_sums(*ctx_, duvw, n, A_, B_, sums_.as<double>());
  double s[6];
  YG_HIP(hipMemcpyAsync(s, sums_.get(), sizeof(s), hipMemcpyDeviceToHost, ctx_->stream));
  // the tolerance is in wavelengths: the largest deviation (metres) at the chunk's highest
  // frequency, as the reference's statistics convert the adapter's residual w
  // (VisMetaDataStats.cc:408,451-452)
  std::vector<double> f(acc.nchan > 0 ? acc.nchan : 0);
  if (acc.nchan > 0)
    YG_HIP(hipMemcpyAsync(f.data(), acc.freq, f.size() * sizeof(double),
                          acc.mem == YGRID_MEM_DEVICE ? hipMemcpyDeviceToHost : hipMemcpyHostToHost, ctx_->stream));
  YG_HIP(hipStreamSynchronize(ctx_->stream));
  double fmax = 0.;
  for (double x : f) fmax = std::max(fmax, x);
  if (!(s[5] * fmax / kC > tol_)) return false;
  // least-squares plane: [uu uv; uv vv] [A B]' = [uw vw]'
  const double det = s[0] * s[2] - s[1] * s[1];
  if (det > 0.) {
    A = (s[3] * s[2] - s[4] * s[1]) / det;
    B = (s[4] * s[0] - s[3] * s[1]) / det;
  } else {
    A = B = 0.;
  }
  n_fits_ += 1;
  return true;
}

ygrid_chunk Snapshot::relative(const ygrid_chunk& acc) {
  ygrid_chunk r = acc;
  const int n = acc.nrow;
  if (n <= 0) return r;
  const double* duvw = acc.mem == YGRID_MEM_DEVICE ? acc.uvw : uvw_in_.as<double>();  // staged by updatePlane
  uvw_rel_.resize((size_t)n * 3 * sizeof(double));
  launch_wplane_apply(*ctx_, duvw, n, A_, B_, uvw_rel_.as<double>());
  if (acc.mem == YGRID_MEM_DEVICE) {
    r.uvw = uvw_rel_.as<double>();
  } else {
    uvw_host_.resize((size_t)n * 3);
    YG_HIP(hipMemcpyAsync(uvw_host_.data(), uvw_rel_.get(), (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx_->stream));
    YG_HIP(hipStreamSynchronize(ctx_->stream));
    r.uvw = uvw_host_.data();
  }
  return r;
}

void Snapshot::regrid(const float* src, float* dst, bool to_target, bool add, double clip) {
  SnapshotRegridArgs a;
  a.A = A_;
  a.B = B_;
  a.cellx = std::fabs(img_.cellx);
  a.celly = std::fabs(img_.celly);
  a.to_target = to_target ? 1 : 0;
  a.cubic = cubic_ ? 1 : 0;
  a.add = add ? 1 : 0;
  // imageClip (:640-652): keep the centred int(n (1 - factor)) pixels of each axis
  a.clip_kx = clip > 0. ? (int)(img_.nx * (1. - clip)) : img_.nx;
  a.clip_ky = clip > 0. ? (int)(img_.ny * (1. - clip)) : img_.ny;
  a.clip_x0 = (img_.nx - a.clip_kx) / 2;
  a.clip_y0 = (img_.ny - a.clip_ky) / 2;
  launch_snapshot_regrid(*ctx_, src, dst, img_.nx, img_.ny, img_.npol * img_.nchan, a);
  n_regrids_ += 1;
}

// finaliseGriddingOfCurrentPlane (:416-443)
void Snapshot::finaliseCurrent() {
  scratch_.resize(imageElems() * sizeof(float));
  g_->finaliseGrid(scratch_.as<float>(), YGRID_MEM_DEVICE);
  regrid(scratch_.as<float>(), image_.as<float>(), true, true, clip_);
  if (!dopcf_) {
    g_->finaliseWeights(scratch_.as<float>(), YGRID_MEM_DEVICE);
    regrid(scratch_.as<float>(), weights_.as<float>(), true, true, wclip_ != 0. ? wclip_ : clip_);
  }
  YG_HIP(hipStreamSynchronize(ctx_->stream));
  done_ = true;
}

void Snapshot::grid(const ygrid_chunk& acc) {
  double A, B;
  const bool changed = updatePlane(acc, A, B);
  if (passthrough()) {
    A_ = A;
    B_ = B;
    g_->grid(relative(acc));
    return;
  }
  if (changed || first_) {
    if (!first_) {
      finaliseCurrent();  // in the frame of the plane it was gridded in
      first_ = true;
    }
    A_ = A;
    B_ = B;
  }
  if (first_) {
    g_->initialiseGrid(img_, dopsf_, dopcf_);
    first_ = false;
  }
  g_->grid(relative(acc));
  done_ = false;
}

void Snapshot::finaliseGrid(float* out, int mem) {
  if (passthrough()) {
    g_->finaliseGrid(out, mem);
    return;
  }
  if (!done_) finaliseCurrent();
  copyOut(image_, out, mem);
}

void Snapshot::finaliseWeights(float* out, int mem) {
  if (passthrough()) {
    g_->finaliseWeights(out, mem);
    return;
  }
  if (!done_) finaliseCurrent();
  copyOut(weights_, out, mem);
}

void Snapshot::initialiseDegrid(const ygrid_image_desc& img, const float* model, int mem) {
  img_ = img;
  stokes_.assign(img.stokes, img.stokes ? img.stokes + img.npol : img.stokes);
  img_.stokes = stokes_.empty() ? nullptr : stokes_.data();
  dopsf_ = dopcf_ = false;
  const size_t n = imageElems();
  image_.resize(n * sizeof(float));
  YG_HIP(hipMemcpyAsync(image_.get(), model, n * sizeof(float),
                        mem == YGRID_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx_->stream));
  DevBuf flag(sizeof(int));
  launch_fill_u32(*ctx_, flag.get(), 1, 0u);
  launch_any_nonzero(*ctx_, image_.as<float>(), n, flag.as<int>());
  int h = 0;
  YG_HIP(hipMemcpyAsync(&h, flag.get(), sizeof(int), hipMemcpyDeviceToHost, ctx_->stream));
  YG_HIP(hipStreamSynchronize(ctx_->stream));
  model_empty_ = h == 0;
  first_ = true;
}

void Snapshot::degrid(const ygrid_chunk& acc) {
  if (model_empty_) return;
  double A, B;
  const bool changed = updatePlane(acc, A, B);
  if (changed || first_) {
    if (!first_) {
      g_->finaliseDegrid();
      first_ = true;
    }
    A_ = A;
    B_ = B;
  }
  if (first_) {
    scratch_.resize(imageElems() * sizeof(float));
    regrid(image_.as<float>(), scratch_.as<float>(), false, false, clip_);
    g_->initialiseDegrid(img_, scratch_.as<float>(), YGRID_MEM_DEVICE);
    first_ = false;
  }
  g_->degrid(relative(acc));
}

void Snapshot::finaliseDegrid() {
  if (model_empty_) return;
  YG_CHECK(!first_, 1, "finaliseDegrid is called while the itsFirstAccessor flag is true. This is not supposed to happen");
  g_->finaliseDegrid();
}

void Snapshot::stats(ygrid_stats* out) const {
  out->snapshot_fits = n_fits_;
  out->snapshot_regrids = n_regrids_;
  out->snapshot_a = A_;
  out->snapshot_b = B_;
}

}  // namespace ygrid
