/// @file GPUGridderAdapter.cc
/// @brief IVisGridder → libygrid forwarding (compiled inside a yandasoft build).
///
/// Every call maps to one entry point of include/ygrid.h.  The accessor's casacore cubes are
/// column-major (nRow, nChan, nPol), which is the layout the C ABI takes, so visibilities and
/// noise are passed by pointer without reordering; uvw, delays, feeds, pointings and flags are
/// staged once per accessor.
#include "GPUGridderAdapter.h"

#include <askap/askap/AskapError.h>
#include <askap/gridding/VisGridderFactory.h>
#include <askap/measurementequation/ImageParamsHelper.h>
#include <askap/measurementequation/SynthesisParamsHelper.h>
#include <casacore/coordinates/Coordinates/DirectionCoordinate.h>

#include <sstream>

namespace askap {
namespace synthesis {

namespace {
std::shared_ptr<ygrid_ctx> shared_context() {
  static std::weak_ptr<ygrid_ctx> cache;
  auto ctx = cache.lock();
  if (!ctx) {
    ygrid_ctx* raw = nullptr;
    if (ygrid_ctx_create(0, &raw) != YGRID_OK) ASKAPTHROW(AskapError, "libygrid: " << ygrid_last_error());
    ctx.reset(raw, [](ygrid_ctx* c) { ygrid_ctx_destroy(c); });
    cache = ctx;
  }
  return ctx;
}

/// the gridder subset as "key = value" lines (the form ygrid_gridder_create parses)
std::string to_text(const LOFAR::ParameterSet& parset) {
  std::ostringstream os;
  for (auto it = parset.begin(); it != parset.end(); ++it) os << it->first << " = " << it->second.get() << "\n";
  return os.str();
}
}  // namespace

void GPUGridderAdapter::check(int status) {
  if (status != YGRID_OK) ASKAPTHROW(AskapError, "libygrid: " << ygrid_last_error());
}

GPUGridderAdapter::GPUGridderAdapter(const std::string& kind, const LOFAR::ParameterSet& parset)
    : itsKind(kind), itsCtx(shared_context()) {
  check(ygrid_gridder_create(itsCtx.get(), kind.c_str(), to_text(parset).c_str(), &itsGridder));
}

GPUGridderAdapter::GPUGridderAdapter(const GPUGridderAdapter& other)
    : IVisGridder(other), itsKind(other.itsKind), itsCtx(other.itsCtx), itsImage(other.itsImage) {
  check(ygrid_gridder_clone(other.itsGridder, &itsGridder));
}

GPUGridderAdapter::~GPUGridderAdapter() { ygrid_gridder_destroy(itsGridder); }

IVisGridder::ShPtr GPUGridderAdapter::clone() { return IVisGridder::ShPtr(new GPUGridderAdapter(*this)); }

ygrid_image_desc GPUGridderAdapter::describe(const scimath::Axes& axes, const casacore::IPosition& shape) {
  ygrid_image_desc d{};
  d.nx = shape(0);
  d.ny = shape(1);
  d.npol = shape.nelements() > 2 ? shape(2) : 1;
  d.nchan = shape.nelements() > 3 ? shape(3) : 1;
  // TableVisGridder::initialiseCellSize (TableVisGridder.cc:1003-1021)
  const casacore::DirectionCoordinate dc = axes.directionAxis();
  d.cellx = dc.increment()(0);
  d.celly = dc.increment()(1);
  const casacore::MVDirection centre = SynthesisParamsHelper::directionAxisCentre(axes);  // getImageCentre
  d.centre_lon = centre.getLong();
  d.centre_lat = centre.getLat();
  d.has_freq_axis = axes.has("FREQUENCY") ? 1 : 0;
  if (d.has_freq_axis) {
    d.freq_start = axes.start("FREQUENCY");
    d.freq_end = axes.end("FREQUENCY");
  }
  // TableVisGridder::initStokes (TableVisGridder.cc:982-994): the stokes axis, or I
  itsImageStokes.clear();
  if (axes.has("STOKES")) {
    const casacore::Vector<casacore::Stokes::StokesTypes> st = axes.stokesAxis();
    for (casacore::uInt p = 0; p < st.nelements(); ++p) itsImageStokes.push_back(int32_t(st[p]));
  } else {
    itsImageStokes.push_back(int32_t(casacore::Stokes::I));
  }
  d.stokes = itsImageStokes.data();
  return d;
}

ygrid_chunk GPUGridderAdapter::stage(const accessors::IConstDataAccessor& acc, bool forward) {
  const casacore::uInt nRow = acc.nRow(), nChan = acc.nChannel(), nPol = acc.nPol();
  const casacore::MVDirection centre(itsImage.centre_lon, itsImage.centre_lat);
  // TableVisGridder.cc:487-491: rotated uvw and the rotation delay for the tangent point
  const auto& uvw = acc.rotatedUVW(centre);
  const auto& delay = acc.uvwRotationDelay(centre, centre);
  itsUVW.resize(3 * nRow);
  itsDelay.resize(nRow);
  itsFeed.resize(nRow);
  itsPointing.resize(2 * nRow);
  for (casacore::uInt i = 0; i < nRow; ++i) {
    for (int k = 0; k < 3; ++k) itsUVW[3 * i + k] = uvw(i)(k);
    itsDelay[i] = delay(i);
    itsFeed[i] = acc.feed1()(i);
    itsPointing[2 * i] = acc.pointingDir1()(i).getLong();
    itsPointing[2 * i + 1] = acc.pointingDir1()(i).getLat();
  }
  const casacore::Cube<casacore::Bool>& flags = acc.flag();
  itsFlag.assign(flags.begin(), flags.end());

  ygrid_chunk c{};
  c.nrow = nRow;
  c.nchan = nChan;
  c.npol = nPol;
  c.mem = YGRID_MEM_HOST;
  c.uvw = itsUVW.data();
  c.delay = itsDelay.data();
  c.freq = acc.frequency().data();
  c.feed1 = itsFeed.data();
  c.pointing1 = itsPointing.data();
  c.flag = itsFlag.data();
  // acc.stokes(): the visibilities' polarisation frame (TableVisGridder.cc:517-524)
  itsVisStokes.clear();
  for (casacore::uInt p = 0; p < nPol; ++p) itsVisStokes.push_back(int32_t(acc.stokes()[p]));
  c.stokes = itsVisStokes.data();
  if (forward) {
    ASKAPCHECK(acc.visibility().contiguousStorage() && acc.noise().contiguousStorage(),
               "GPU gridder needs contiguous visibility and noise cubes");
    c.vis = reinterpret_cast<const float*>(acc.visibility().data());
    c.noise = reinterpret_cast<const float*>(acc.noise().data());
  }
  return c;
}

void GPUGridderAdapter::initialiseGrid(const scimath::Axes& axes, const casacore::IPosition& shape,
                                       const bool dopsf, const bool dopcf) {
  itsImage = describe(axes, shape);
  check(ygrid_initialise_grid(itsGridder, &itsImage, dopsf ? 1 : 0, dopcf ? 1 : 0));
}

void GPUGridderAdapter::grid(accessors::IConstDataAccessor& acc) {
  const ygrid_chunk c = stage(acc, true);
  check(ygrid_grid(itsGridder, &c));
}

void GPUGridderAdapter::finaliseGrid(casacore::Array<imtype>& out) {
  casacore::Array<float> buf(out.shape());
  check(ygrid_finalise_grid(itsGridder, buf.data(), YGRID_MEM_HOST));
  casacore::convertArray(out, buf);
}

void GPUGridderAdapter::finaliseWeights(casacore::Array<imtype>& out) {
  casacore::Array<float> buf(out.shape());
  check(ygrid_finalise_weights(itsGridder, buf.data(), YGRID_MEM_HOST));
  casacore::convertArray(out, buf);
}

void GPUGridderAdapter::initialiseDegrid(const scimath::Axes& axes, const casacore::Array<imtype>& image) {
  itsImage = describe(axes, image.shape());
  casacore::Array<float> buf(image.shape());
  casacore::convertArray(buf, image);
  check(ygrid_initialise_degrid(itsGridder, &itsImage, buf.data(), YGRID_MEM_HOST));
}

void GPUGridderAdapter::degrid(accessors::IDataAccessor& acc) {
  ygrid_chunk c = stage(acc, false);
  casacore::Cube<casacore::Complex>& vis = acc.rwVisibility();
  ASKAPCHECK(vis.contiguousStorage(), "GPU degridder needs a contiguous visibility cube");
  c.vis_out = reinterpret_cast<float*>(vis.data());
  check(ygrid_degrid(itsGridder, &c));
}

void GPUGridderAdapter::finaliseDegrid() { check(ygrid_finalise_degrid(itsGridder)); }

bool GPUGridderAdapter::isModelEmpty() const {
  int empty = 1;
  check(ygrid_is_model_empty(itsGridder, &empty));
  return empty != 0;
}

void GPUGridderAdapter::initVisWeights(const IVisWeights::ShPtr& viswt) {
  // Only the MFS weights exist (VisGridderFactory.cc:247-255).  Their reference frequency
  // is private; with order 1 the weight is (f - f0)/f0, so f0 = f/(w(f) + 1).
  if (!viswt) {
    check(ygrid_set_mfs_weight(itsGridder, 0.0, -1));
    return;
  }
  IVisWeights::ShPtr probe = viswt->clone();
  probe->setParameters(1);
  const double f = 1e9;
  itsMFSRefFreq = f / (double(probe->getWeight(f)) + 1.0);
  check(ygrid_set_mfs_weight(itsGridder, itsMFSRefFreq, 0));
}

void GPUGridderAdapter::customiseForContext(const std::string& context) {
  // TableVisGridder.cc:1382-1405: Taylor term order from the image name
  if (itsMFSRefFreq > 0) {
    const ImageParamsHelper iph(context);
    const int order = iph.isTaylorTerm() ? iph.order() : 1;
    check(ygrid_set_mfs_weight(itsGridder, itsMFSRefFreq, order));
  }
}

IVisGridder::ShPtr GPUGridderAdapter::createGPUBox(const LOFAR::ParameterSet& p) {
  return IVisGridder::ShPtr(new GPUGridderAdapter("GPUBox", p));
}
IVisGridder::ShPtr GPUGridderAdapter::createGPUSphFunc(const LOFAR::ParameterSet& p) {
  return IVisGridder::ShPtr(new GPUGridderAdapter("GPUSphFunc", p));
}
IVisGridder::ShPtr GPUGridderAdapter::createGPUWProject(const LOFAR::ParameterSet& p) {
  return IVisGridder::ShPtr(new GPUGridderAdapter("GPUWProject", p));
}
IVisGridder::ShPtr GPUGridderAdapter::createGPUAWProject(const LOFAR::ParameterSet& p) {
  return IVisGridder::ShPtr(new GPUGridderAdapter("GPUAWProject", p));
}
IVisGridder::ShPtr GPUGridderAdapter::createGPUWStack(const LOFAR::ParameterSet& p) {
  return IVisGridder::ShPtr(new GPUGridderAdapter("GPUWStack", p));
}

}  // namespace synthesis
}  // namespace askap

// VisGridderFactory::createGridder (VisGridderFactory.cc:68-98) dlopens libaskap_<name>.so
// for an unknown gridder name and calls register_<name>(); each of the five names gets a
// symlink to this one library so that any of them loads it.
extern "C" {
static void ygrid_register_all() {
  using askap::synthesis::GPUGridderAdapter;
  using askap::synthesis::VisGridderFactory;
  VisGridderFactory::registerGridder("GPUBox", GPUGridderAdapter::createGPUBox);
  VisGridderFactory::registerGridder("GPUSphFunc", GPUGridderAdapter::createGPUSphFunc);
  VisGridderFactory::registerGridder("GPUWProject", GPUGridderAdapter::createGPUWProject);
  VisGridderFactory::registerGridder("GPUAWProject", GPUGridderAdapter::createGPUAWProject);
  VisGridderFactory::registerGridder("GPUWStack", GPUGridderAdapter::createGPUWStack);
}
void register_gpubox() { ygrid_register_all(); }
void register_gpusphfunc() { ygrid_register_all(); }
void register_gpuwproject() { ygrid_register_all(); }
void register_gpuawproject() { ygrid_register_all(); }
void register_gpuwstack() { ygrid_register_all(); }
}
