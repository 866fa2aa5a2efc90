// One chunk of the imaging equation's data loop on the device.
//
// ImageFFTEquation::calcImagingEquations (ImageFFTEquation.cc:437-480) runs, per accessor chunk:
// the model visibilities = sum over the model degridders whose model is not empty (`:445-455`),
// then residual = -(model - data) (`:464-465`), then every residual, PSF and preconditioner
// gridder grids that residual (`:467-476`).  Here the chunk is made device-resident once (host
// arrays are copied in, device arrays are used as they are), the model cube lives in device
// memory, the subtraction is one kernel, and the gridders read the residual in place: nothing
// returns to the host between the degridding and the gridding.
#include <algorithm>
#include <vector>

#include "gridders.h"

namespace ygrid {

// A host chunk's input arrays copied to the device (every pointer of the copy a device pointer,
// vis_out NULL); a device chunk is used as it is.
DeviceChunk::DeviceChunk(const Ctx& ctx, const ygrid_chunk& a) : c(a) {
  if (a.mem == YGRID_MEM_DEVICE) return;
  YG_CHECK(a.mem == YGRID_MEM_HOST, 1, "unknown memory kind");
  const size_t nr = a.nrow, nc = a.nchan, ncube = nr * nc * (size_t)a.npol;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t sizes[] = {a.uvw ? nr * 24 : 0,       a.delay ? nr * 8 : 0,      a.freq ? nc * 8 : 0,
                          a.vis ? ncube * 8 : 0,     a.noise ? ncube * 8 : 0,   a.flag ? ncube : 0,
                          a.feed1 ? nr * 4 : 0,      a.pointing1 ? nr * 16 : 0, a.dish_pointing1 ? nr * 16 : 0,
                          a.feed1_pa ? nr * 4 : 0};
  size_t bytes = 0;
  for (size_t b : sizes) bytes += al(b);
  store.resize(std::max<size_t>(bytes, 256));
  char* p = store.as<char>();
  auto put = [&](const void* src, size_t b) -> void* {
    if (!src) return nullptr;
    void* dst = p;
    if (b) YG_HIP(hipMemcpyAsync(dst, src, b, hipMemcpyHostToDevice, ctx.stream));
    p += al(b);
    return dst;
  };
  c.uvw = static_cast<const double*>(put(a.uvw, sizes[0]));
  c.delay = static_cast<const double*>(put(a.delay, sizes[1]));
  c.freq = static_cast<const double*>(put(a.freq, sizes[2]));
  c.vis = static_cast<const float*>(put(a.vis, sizes[3]));
  c.noise = static_cast<const float*>(put(a.noise, sizes[4]));
  c.flag = static_cast<const uint8_t*>(put(a.flag, sizes[5]));
  c.feed1 = static_cast<const int32_t*>(put(a.feed1, sizes[6]));
  c.pointing1 = static_cast<const double*>(put(a.pointing1, sizes[7]));
  c.dish_pointing1 = static_cast<const double*>(put(a.dish_pointing1, sizes[8]));
  c.feed1_pa = static_cast<const float*>(put(a.feed1_pa, sizes[9]));
  c.vis_out = nullptr;
  c.vis_out_rows = 0;
  c.mem = YGRID_MEM_DEVICE;
  // the host arrays may be reused by the caller as soon as this returns
  YG_HIP(hipStreamSynchronize(ctx.stream));
}

void imaging_chunk(const std::vector<Gridder*>& degridders, const std::vector<Gridder*>& gridders,
                   const ygrid_chunk& a) {
  YG_CHECK(!degridders.empty() || !gridders.empty(), 1, "no gridders given");
  const std::shared_ptr<Ctx>& cp = degridders.empty() ? gridders[0]->context() : degridders[0]->context();
  for (const Gridder* g : degridders) {
    YG_CHECK(g->context() == cp, 1, "all gridders must share one context");
    YG_CHECK(g->initialisedFor(true), 2, "a model degridder is not initialised for degridding");
  }
  for (const Gridder* g : gridders) {
    YG_CHECK(g->context() == cp, 1, "all gridders must share one context");
    YG_CHECK(g->initialisedFor(false), 2, "a residual gridder is not initialised for gridding");
  }
  YG_CHECK(a.nrow >= 0 && a.nchan > 0 && a.npol > 0, 1, "bad chunk shape");
  YG_CHECK(a.vis != nullptr, 1, "the imaging pass needs the data visibilities (chunk vis)");
  YG_CHECK(a.vis_out_rows == 0 || a.vis_out_rows == a.nrow, 1, "the imaging pass returns a residual cube of nrow rows");
  const Ctx& c = *cp;
  const size_t ncube = (size_t)a.nrow * a.nchan * a.npol;
  DeviceChunk dc(c, a);
  // accBuffer.rwVisibility().set(0.0), then every non-empty model degridder adds its prediction
  DevBuf resid(std::max<size_t>(ncube, 1) * sizeof(float2));
  launch_fill_u32(c, resid.get(), 2 * ncube, 0u);
  ygrid_chunk dg = dc.c;
  dg.vis_out = resid.as<float>();
  for (Gridder* g : degridders)
    if (!g->isModelEmpty()) g->degrid(dg);
  // rwVisibility() -= data; rwVisibility() *= -1: residual = data - model
  launch_residual(c, reinterpret_cast<const float2*>(dc.c.vis), resid.as<float2>(), ncube);
  ygrid_chunk gr = dc.c;
  gr.vis = resid.as<float>();
  gr.vis_out = nullptr;
  for (Gridder* g : gridders) g->grid(gr);
  if (a.vis_out && ncube)
    YG_HIP(hipMemcpyAsync(a.vis_out, resid.get(), ncube * sizeof(float2),
                          a.mem == YGRID_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c.stream));
  YG_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace ygrid
