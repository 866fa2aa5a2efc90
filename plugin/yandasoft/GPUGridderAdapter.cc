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
#include <askap/scimath/utils/PaddingUtils.h>
#include <casacore/coordinates/Coordinates/DirectionCoordinate.h>

#include <iomanip>
#include <sstream>

namespace askap {
namespace synthesis {

namespace {
// VisGridderFactory::make applies its global options after creating a gridder, through
// TableVisGridder's setters (VisGridderFactory.cc:137-232); the fields they set are private and
// have no getters, and setSourceIndex (ImageFFTEquation.cc:283-287) is the same.  The usual
// access checks do not apply to the names in an explicit template instantiation, so a pointer to
// each field can be formed there, which lets the adapter read the values without a change to the
// reference's headers.  (With a getter added upstream, read that instead.)
template <class Tag, typename Tag::type M>
struct MemberOf {
  friend typename Tag::type field(Tag) { return M; }
};
#define YGRID_TVG_FIELD(TAG, TYPE, NAME)                  \
  struct TAG {                                            \
    typedef TYPE TableVisGridder::*type;                  \
    friend type field(TAG);                               \
  };                                                      \
  template struct MemberOf<TAG, &TableVisGridder::NAME>;
YGRID_TVG_FIELD(UseAllPSF, bool, itsUseAllDataForPSF)
YGRID_TVG_FIELD(MaxPointingSep, double, itsMaxPointingSeparation)
YGRID_TVG_FIELD(TrackPerPlane, bool, itsTrackWeightPerOversamplePlane)
YGRID_TVG_FIELD(PARotation, bool, itsPARotation)
YGRID_TVG_FIELD(SwapPols, bool, itsSwapPols)
YGRID_TVG_FIELD(ClearGrid, bool, itsClearGrid)
YGRID_TVG_FIELD(SourceIndex, int, itsSourceIndex)
#undef YGRID_TVG_FIELD

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
    : TableVisGridder(1, 1), itsKind(kind), itsCtx(shared_context()) {
  check(ygrid_gridder_create(itsCtx.get(), kind.c_str(), to_text(parset).c_str(), &itsGridder));
}

// TableVisGridder's copy constructor copies the factory's options (TableVisGridder.cc:126-168);
// the library clone shares the CF cache (WProjectVisGridder.cc:51-52)
GPUGridderAdapter::GPUGridderAdapter(const GPUGridderAdapter& other)
    : TableVisGridder(other), itsKind(other.itsKind), itsCtx(other.itsCtx), itsImage(other.itsImage),
      itsMFSRefFreq(other.itsMFSRefFreq) {
  check(ygrid_gridder_clone(other.itsGridder, &itsGridder));
}

void GPUGridderAdapter::applyFactoryOptions() {
  ASKAPCHECK(!(this->*field(PARotation{})) && !(this->*field(SwapPols{})),
             "gridder.parotation / gridder.swappols: parallactic angle rotation is not done by the GPU gridders");
  std::ostringstream os;
  os << std::setprecision(17) << "padding = " << paddingFactor() << "\n"
     << "alldatapsf = " << ((this->*field(UseAllPSF{})) ? "true" : "false") << "\n"
     << "oversampleweight = " << ((this->*field(TrackPerPlane{})) ? "true" : "false") << "\n"
     << "cleargrids = " << ((this->*field(ClearGrid{})) ? "true" : "false") << "\n"
     << "MaxPointingSeparation = " << (this->*field(MaxPointingSep{})) << "\n";
  check(ygrid_gridder_configure(itsGridder, os.str().c_str()));
}

void GPUGridderAdapter::setGeometry(const scimath::Axes& axes, const casacore::IPosition& shape) {
  // TableVisGridder::initialiseGrid (TableVisGridder.cc:1002-1003)
  itsAxes = axes;
  itsShape = scimath::PaddingUtils::paddedShape(shape, paddingFactor());
}

void GPUGridderAdapter::initConvolutionFunction(const accessors::IConstDataAccessor&) {
  ASKAPTHROW(AskapError, "GPUGridderAdapter: the convolution functions live in libygrid");
}
void GPUGridderAdapter::initIndices(const accessors::IConstDataAccessor&) {
  ASKAPTHROW(AskapError, "GPUGridderAdapter: the CF indices are computed in libygrid");
}
void GPUGridderAdapter::correctConvolution(casacore::Array<imtype>&) {
  ASKAPTHROW(AskapError, "GPUGridderAdapter: the grid correction is applied in libygrid");
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
  // TableVisGridder::initStokes (TableVisGridder.cc:982-994): the stokes axis, or I, one code per
  // polarisation plane of the grid
  itsImageStokes.clear();
  if (axes.has("STOKES")) {
    const casacore::Vector<casacore::Stokes::StokesTypes> st = axes.stokesAxis();
    for (casacore::uInt p = 0; p < st.nelements(); ++p) itsImageStokes.push_back(int32_t(st[p]));
  } else {
    itsImageStokes.push_back(int32_t(casacore::Stokes::I));
  }
  ASKAPCHECK(int(itsImageStokes.size()) == d.npol, "Stokes axis is not consistent with the shape of the grid. There are "
             << d.npol << " planes in the grid and " << itsImageStokes.size()
             << " polarisation descriptors defined by the STOKES axis");
  d.stokes = itsImageStokes.data();
  return d;
}

ygrid_chunk GPUGridderAdapter::stage(const accessors::IConstDataAccessor& acc, bool forward) {
  const casacore::uInt nRow = acc.nRow(), nChan = acc.nChannel(), nPol = acc.nPol();
  // TableVisGridder.cc:481-495: uvw rotated to the tangent point and the delay that moves the
  // phase centre from the tangent point to the image centre
  const casacore::MVDirection imageCentre = getImageCentre();
  const casacore::MVDirection tangentPoint = getTangentPoint();
  const auto& uvw = acc.rotatedUVW(tangentPoint);
  const auto& delay = acc.uvwRotationDelay(tangentPoint, imageCentre);
  itsUVW.resize(3 * nRow);
  itsDelay.resize(nRow);
  itsFeed.resize(nRow);
  itsPointing.resize(2 * nRow);
  itsDish.resize(2 * nRow);
  itsPA.resize(nRow);
  for (casacore::uInt i = 0; i < nRow; ++i) {
    for (int k = 0; k < 3; ++k) itsUVW[3 * i + k] = uvw(i)(k);
    itsDelay[i] = delay(i);
    itsFeed[i] = acc.feed1()(i);
    itsPA[i] = acc.feed1PA()(i);
    itsPointing[2 * i] = acc.pointingDir1()(i).getLong();
    itsPointing[2 * i + 1] = acc.pointingDir1()(i).getLat();
    itsDish[2 * i] = acc.dishPointing1()(i).getLong();
    itsDish[2 * i + 1] = acc.dishPointing1()(i).getLat();
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
  c.dish_pointing1 = itsDish.data();
  // AProjectGridderBase.cc:212-219: the A-projection CF cache follows the feeds' parallactic angle
  c.feed1_pa = itsPA.data();
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
  applyFactoryOptions();
  setGeometry(axes, shape);
  configureForPSF(dopsf);
  configureForPCF(dopcf);
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
  applyFactoryOptions();
  setGeometry(axes, image.shape());
  configureForPSF(false);
  configureForPCF(false);
  itsImage = describe(axes, image.shape());
  casacore::Array<float> buf(image.shape());
  casacore::convertArray(buf, image);
  check(ygrid_initialise_degrid(itsGridder, &itsImage, buf.data(), YGRID_MEM_HOST));
}

void GPUGridderAdapter::degrid(accessors::IDataAccessor& acc) {
  ygrid_chunk c = stage(acc, false);
  casacore::Cube<casacore::Complex>& vis = acc.rwVisibility();
  ASKAPCHECK(vis.contiguousStorage(), "GPU degridder needs a contiguous visibility cube");
  // direction-dependent calibration: source s predicts into rows s*nRow .. of the cube
  // (TableVisGridder.cc:567,861)
  const int source = this->*field(SourceIndex{});
  const size_t offset = size_t(source) * c.nrow;
  ASKAPCHECK(offset + c.nrow <= vis.nrow(), "source index " << source << " beyond the visibility cube's rows");
  c.vis_out = reinterpret_cast<float*>(vis.data() + offset);
  c.vis_out_rows = int(vis.nrow());
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
IVisGridder::ShPtr GPUGridderAdapter::createGPUAProjectWStack(const LOFAR::ParameterSet& p) {
  return IVisGridder::ShPtr(new GPUGridderAdapter("GPUAProjectWStack", p));
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
  VisGridderFactory::registerGridder("GPUAProjectWStack", GPUGridderAdapter::createGPUAProjectWStack);
}
void register_gpubox() { ygrid_register_all(); }
void register_gpusphfunc() { ygrid_register_all(); }
void register_gpuwproject() { ygrid_register_all(); }
void register_gpuawproject() { ygrid_register_all(); }
void register_gpuwstack() { ygrid_register_all(); }
}
