// The accessor's coordinate step on the device: rotatedUVW and uvwRotationDelay
// (IConstDataAccessor, as called by TableVisGridder::generic, TableVisGridder.cc:481-495).
//
// A row's uvw (metres) is given in the frame of its phase centre p: u toward the east, v toward
// the north, w toward p.  With e_u, e_v, e_w the frame's unit vectors in equatorial xyz at
// direction (lon, lat),
//   e_u = (-sin lon, cos lon, 0)
//   e_v = (-sin lat cos lon, -sin lat sin lon, cos lat)
//   e_w = (cos lat cos lon, cos lat sin lon, sin lat),
// the baseline is b = u e_u(p) + v e_v(p) + w e_w(p).  rotatedUVW(t) is b in the frame of the
// tangent point t, (b.e_u(t), b.e_v(t), b.e_w(t)).  uvwRotationDelay(t, c) is the path
// difference that rephases the visibilities from p to the image centre c, b.e_w(c) - w: the
// gridder multiplies each visibility's conjugate by exp(2 pi i f delay / c) (TableVisGridder.cc:
// 652-655), so a source at c grids as a source at the centre of the uv grid while the uvw keep
// the tangent point's projection.  The rotation is the plain change of frame; casacore's
// UVWMachine (base-accessors' choice of its projection mode) is not part of the reference.
//
// One thread per row, FP64 throughout: 24 bytes in, 32 out per row (HBM bound; a 189,000-row
// chunk takes microseconds).
#include <hip/hip_runtime.h>

#include "ygrid_internal.h"

namespace ygrid {

namespace {

struct Frame {
  double u[3], v[3], w[3];
};

__host__ __device__ inline Frame frame_of(double lon, double lat) {
  const double sl = sin(lon), cl = cos(lon), sb = sin(lat), cb = cos(lat);
  Frame f;
  f.u[0] = -sl; f.u[1] = cl; f.u[2] = 0.;
  f.v[0] = -sb * cl; f.v[1] = -sb * sl; f.v[2] = cb;
  f.w[0] = cb * cl; f.w[1] = cb * sl; f.w[2] = sb;
  return f;
}

__device__ inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

__global__ void k_uvw_rotate(const double* __restrict__ uvw, const double* __restrict__ pc, int pc_per_row,
                             UvwTargets T, int nrow, double* __restrict__ out, double* __restrict__ delay) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  const size_t k = pc_per_row ? (size_t)r : 0;
  const Frame p = frame_of(pc[2 * k], pc[2 * k + 1]);
  const double u = uvw[3 * (size_t)r], v = uvw[3 * (size_t)r + 1], w = uvw[3 * (size_t)r + 2];
  double b[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) b[i] = u * p.u[i] + v * p.v[i] + w * p.w[i];
  out[3 * (size_t)r] = dot3(b, T.tu);
  out[3 * (size_t)r + 1] = dot3(b, T.tv);
  out[3 * (size_t)r + 2] = dot3(b, T.tw);
  if (delay) delay[r] = dot3(b, T.cw) - w;
}

}  // namespace

UvwTargets uvw_targets(double tangent_lon, double tangent_lat, double centre_lon, double centre_lat) {
  const Frame t = frame_of(tangent_lon, tangent_lat), c = frame_of(centre_lon, centre_lat);
  UvwTargets T;
  for (int i = 0; i < 3; ++i) {
    T.tu[i] = t.u[i];
    T.tv[i] = t.v[i];
    T.tw[i] = t.w[i];
    T.cw[i] = c.w[i];
  }
  return T;
}

void launch_uvw_rotate(const Ctx& c, const double* uvw, const double* pc, int pc_per_row, const UvwTargets& T,
                       int nrow, double* out, double* delay) {
  if (nrow <= 0) return;
  hipLaunchKernelGGL(k_uvw_rotate, dim3((nrow + 255) / 256), dim3(256), 0, c.stream, uvw, pc, pc_per_row, T, nrow,
                     out, delay);
  YG_HIP(hipGetLastError());
}

}  // namespace ygrid
