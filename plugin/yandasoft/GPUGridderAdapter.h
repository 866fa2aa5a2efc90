/// @file GPUGridderAdapter.h
/// @brief yandasoft-side adapter that forwards IVisGridder calls to libygrid (MI355X).
///
/// This file is compiled inside a yandasoft build (it needs casacore and yandasoft's own
/// headers), not in this repository.  The adapter is a TableVisGridder
/// (askap/gridding/TableVisGridder.h), so that everything yandasoft applies to gridders through
/// that class works on it: the factory's options (VisGridderFactory.cc:137-232: padding,
/// MaxPointingSeparation, alldatapsf, oversampleweight, parotation, swappols, cleargrids) and
/// ImageFFTEquation's source index for direction-dependent calibration
/// (ImageFFTEquation.cc:283-287).  It overrides every IVisGridder call (IVisGridder.h:64-140) and
/// forwards it to the C ABI of include/ygrid.h one to one; TableVisGridder's own loop is never
/// run.  Six gridder names are registered: GPUBox, GPUSphFunc, GPUWProject, GPUAWProject,
/// GPUWStack and GPUAProjectWStack; a parset selects one with `gridder = GPUWProject` and keeps every
/// `gridder.WProject.*` key under `gridder.GPUWProject.*`.
#ifndef YGRID_GPU_GRIDDER_ADAPTER_H
#define YGRID_GPU_GRIDDER_ADAPTER_H

#include <askap/gridding/IVisWeights.h>
#include <askap/gridding/TableVisGridder.h>
#include <Common/ParameterSet.h>

#include <memory>
#include <string>
#include <vector>

#include "ygrid.h"

namespace askap {
namespace synthesis {

class GPUGridderAdapter : public TableVisGridder {
 public:
  /// @param kind one of "GPUBox", "GPUSphFunc", "GPUWProject", "GPUAWProject", "GPUWStack",
  ///             "GPUAProjectWStack"
  /// @param parset gridder subset (gridder name already stripped), as the factory passes it
  GPUGridderAdapter(const std::string& kind, const LOFAR::ParameterSet& parset);
  ~GPUGridderAdapter() override;

  IVisGridder::ShPtr clone() override;
  void initialiseGrid(const scimath::Axes& axes, const casacore::IPosition& shape,
                      const bool dopsf = true, const bool dopcf = false) override;
  void grid(accessors::IConstDataAccessor& acc) override;
  void finaliseGrid(casacore::Array<imtype>& out) override;
  void finaliseWeights(casacore::Array<imtype>& out) override;
  void initialiseDegrid(const scimath::Axes& axes, const casacore::Array<imtype>& image) override;
  void customiseForContext(const std::string& context) override;
  void initVisWeights(const IVisWeights::ShPtr& viswt) override;
  void degrid(accessors::IDataAccessor& acc) override;
  void finaliseDegrid() override;
  bool isModelEmpty() const override;

  /// factory hooks (VisGridderFactory.h:64-93)
  static IVisGridder::ShPtr createGPUBox(const LOFAR::ParameterSet& parset);
  static IVisGridder::ShPtr createGPUSphFunc(const LOFAR::ParameterSet& parset);
  static IVisGridder::ShPtr createGPUWProject(const LOFAR::ParameterSet& parset);
  static IVisGridder::ShPtr createGPUAWProject(const LOFAR::ParameterSet& parset);
  static IVisGridder::ShPtr createGPUWStack(const LOFAR::ParameterSet& parset);
  static IVisGridder::ShPtr createGPUAProjectWStack(const LOFAR::ParameterSet& parset);

 protected:
  /// TableVisGridder's per-gridder hooks: only its own loop calls them, and every entry point
  /// that would reach that loop is overridden here, so they throw if ever called.
  void initConvolutionFunction(const accessors::IConstDataAccessor& acc) override;
  void initIndices(const accessors::IConstDataAccessor& acc) override;
  void correctConvolution(casacore::Array<imtype>& image) override;

 private:
  GPUGridderAdapter(const GPUGridderAdapter& other);
  /// the options the factory set on this TableVisGridder, forwarded to the library
  void applyFactoryOptions();
  /// TableVisGridder::initialiseGrid's geometry (itsAxes, padded itsShape), so that
  /// getImageCentre / getTangentPoint answer as they do for the reference's gridders
  void setGeometry(const scimath::Axes& axes, const casacore::IPosition& shape);
  ygrid_image_desc describe(const scimath::Axes& axes, const casacore::IPosition& shape);
  ygrid_chunk stage(const accessors::IConstDataAccessor& acc, bool forward);
  static void check(int status);

  std::string itsKind;
  std::shared_ptr<ygrid_ctx> itsCtx;  ///< one context (device + stream) per process
  ygrid_gridder* itsGridder = nullptr;
  ygrid_image_desc itsImage{};
  double itsMFSRefFreq = 0.0;  ///< > 0 once MFS visibility weights are set
  // host staging of one accessor (reused across calls)
  std::vector<double> itsUVW, itsDelay, itsPointing, itsDish;
  std::vector<int32_t> itsFeed, itsImageStokes, itsVisStokes;
  std::vector<float> itsPA;
  std::vector<uint8_t> itsFlag;
};

}  // namespace synthesis
}  // namespace askap

#endif
