// Snapshot imaging around a gridder: a casacore-free mirror of
// askap::synthesis::SnapShotImagingGridderAdapter (askap/gridding/SnapShotImagingGridderAdapter.h),
// its w-plane fit (base-accessors' BestWPlaneDataAccessor) and image-plane regrid on the device.
#pragma once

#include <memory>
#include <vector>

#include "gridders.h"

namespace ygrid {

class Snapshot {
 public:
  // VisGridderFactory.cc:248-270: gridder.snapshotimaging.{wtolerance (wavelengths at a chunk's
  // highest frequency), clipping, weightsclipping,
  // reprojectpsf, interpmethod, coorddecimation, longtrack}, keys here without the prefix
  // "gridder." ("snapshotimaging.wtolerance", ...)
  Snapshot(Gridder* inner, const Params& p);
  Snapshot(const Snapshot& o, Gridder* inner);
  void initialiseGrid(const ygrid_image_desc& img, bool dopsf, bool dopcf);
  void grid(const ygrid_chunk& acc);
  void finaliseGrid(float* out, int mem);
  void finaliseWeights(float* out, int mem);
  void initialiseDegrid(const ygrid_image_desc& img, const float* model, int mem);
  void degrid(const ygrid_chunk& acc);
  void finaliseDegrid();
  bool isModelEmpty() const { return model_empty_; }
  void stats(ygrid_stats* out) const;

 private:
  bool passthrough() const { return (dopsf_ || dopcf_) && !reproject_psf_; }
  // the plane for this chunk: refit (returns true) when a row departs from the current plane by
  // more than the tolerance; the chunk with the plane-relative w in `rel`
  bool updatePlane(const ygrid_chunk& acc, double& A, double& B);
  ygrid_chunk relative(const ygrid_chunk& acc);
  void finaliseCurrent();
  void regrid(const float* src, float* dst, bool to_target, bool add, double clip);
  size_t imageElems() const { return (size_t)img_.nx * img_.ny * img_.npol * img_.nchan; }
  void copyOut(const DevBuf& b, float* out, int mem) const;

  Gridder* g_;
  std::shared_ptr<Ctx> ctx_;
  double tol_ = 0, clip_ = 0, wclip_ = 0;
  bool reproject_psf_ = false, cubic_ = true;
  double A_ = 0, B_ = 0;
  bool dopsf_ = false, dopcf_ = false, first_ = true, done_ = true, model_empty_ = false;
  ygrid_image_desc img_{};
  std::vector<int32_t> stokes_;
  DevBuf image_, weights_, scratch_, sums_, uvw_in_, uvw_rel_;
  std::vector<double> uvw_host_;
  double n_fits_ = 0, n_regrids_ = 0;
};

}  // namespace ygrid
