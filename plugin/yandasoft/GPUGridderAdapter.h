/// @file GPUGridderAdapter.h
/// @brief yandasoft-side adapter that forwards IVisGridder calls to libygrid (MI355X).
///
/// This file is compiled inside a yandasoft build (it needs casacore and yandasoft's own
/// headers), not in this repository.  It implements askap::synthesis::IVisGridder
/// (askap/gridding/IVisGridder.h:64-140) by converting each accessor into a ygrid_chunk and
/// calling the C ABI of include/ygrid.h one to one.  Four gridder names are registered:
/// GPUBox, GPUSphFunc, GPUWProject and GPUAWProject; a parset selects one with
/// `gridder = GPUWProject` and keeps every `gridder.WProject.*` key under
/// `gridder.GPUWProject.*`.
#ifndef YGRID_GPU_GRIDDER_ADAPTER_H
#define YGRID_GPU_GRIDDER_ADAPTER_H

#include <askap/gridding/IVisGridder.h>
#include <askap/gridding/IVisWeights.h>
#include <Common/ParameterSet.h>

#include <memory>
#include <string>
#include <vector>

#include "ygrid.h"

namespace askap {
namespace synthesis {

class GPUGridderAdapter : public IVisGridder {
 public:
  /// @param kind one of "GPUBox", "GPUSphFunc", "GPUWProject", "GPUAWProject"
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

 private:
  GPUGridderAdapter(const GPUGridderAdapter& other);
  ygrid_image_desc describe(const scimath::Axes& axes, const casacore::IPosition& shape);
  ygrid_chunk stage(const accessors::IConstDataAccessor& acc, bool forward);
  static void check(int status);

  std::string itsKind;
  std::shared_ptr<ygrid_ctx> itsCtx;  ///< one context (device + stream) per process
  ygrid_gridder* itsGridder = nullptr;
  ygrid_image_desc itsImage{};
  double itsMFSRefFreq = 0.0;  ///< > 0 once MFS visibility weights are set
  // host staging of one accessor (reused across calls)
  std::vector<double> itsUVW, itsDelay, itsPointing;
  std::vector<int32_t> itsFeed, itsImageStokes, itsVisStokes;
  std::vector<uint8_t> itsFlag;
};

}  // namespace synthesis
}  // namespace askap

#endif
