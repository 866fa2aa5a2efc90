// Snapshot imaging on the device (SnapShotImagingGridderAdapter.cc): the w-plane fit over a
// chunk's rows, the plane-relative w handed to the wrapped gridder, and the image-plane regrid
// between the snapshot's slanted SIN frame and the target frame.
//
// Fit: one pass over the rows accumulates u.u, u.v, v.v, u.w, v.w (FP64) and the largest
// |w - A u - B v| for the current plane; the 2 x 2 normal equations are solved on the host.
// Regrid: one thread per output pixel computes, in FP64, where that pixel's direction falls in
// the input frame (a source at (l, m) images at (l - A (1 - n), m - B (1 - n)) in the frame of
// the plane w = A u + B v, so the target -> snapshot direction is that closed form and the
// snapshot -> target direction its inverse, the root near 1 of a quadratic in n), then
// interpolates every image plane there (bilinear, or Keys' cubic convolution with a = -1/2),
// 0 where the stencil leaves the image, and adds to or replaces the output; pixels outside the
// clipping box are zeroed.  Images are [plane][y][x] float, pixel (x, y) at
// ((x - nx/2) cellx, (y - ny/2) celly).  HBM bound: a cubic regrid reads 16 input pixels per
// output pixel and plane (mostly from L2: neighbouring threads share rows).
#include <hip/hip_runtime.h>

#include "ygrid_internal.h"

namespace ygrid {

namespace {

__global__ void k_wplane_sums(const double* __restrict__ uvw, int nrow, double A, double B,
                              double* __restrict__ sums, unsigned long long* __restrict__ maxdev_bits) {
  __shared__ double s[5][256];
  __shared__ double smax[256];
  double a[5] = {0., 0., 0., 0., 0.}, mx = 0.;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrow; r += gridDim.x * blockDim.x) {
    const double u = uvw[3 * (size_t)r], v = uvw[3 * (size_t)r + 1], w = uvw[3 * (size_t)r + 2];
    a[0] += u * u;
    a[1] += u * v;
    a[2] += v * v;
    a[3] += u * w;
    a[4] += v * w;
    mx = fmax(mx, fabs(w - A * u - B * v));
  }
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < 5; ++k) s[k][t] = a[k];
  smax[t] = mx;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if (t < h) {
#pragma unroll
      for (int k = 0; k < 5; ++k) s[k][t] += s[k][t + h];
      smax[t] = fmax(smax[t], smax[t + h]);
    }
    __syncthreads();
  }
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < 5; ++k) atomicAdd(sums + k, s[k][0]);
    // non-negative doubles order as their bit patterns
    atomicMax(maxdev_bits, (unsigned long long)__double_as_longlong(smax[0]));
  }
}

__global__ void k_wplane_apply(const double* __restrict__ uvw, int nrow, double A, double B, double* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrow) return;
  const double u = uvw[3 * (size_t)r], v = uvw[3 * (size_t)r + 1], w = uvw[3 * (size_t)r + 2];
  out[3 * (size_t)r] = u;
  out[3 * (size_t)r + 1] = v;
  out[3 * (size_t)r + 2] = w - A * u - B * v;
}

__device__ inline double keys(double t) {
  t = fabs(t);
  if (t <= 1.) return (1.5 * t - 2.5) * t * t + 1.;
  if (t < 2.) return ((-0.5 * t + 2.5) * t - 4.) * t + 2.;
  return 0.;
}

__global__ void k_snapshot_regrid(const float* __restrict__ src, float* __restrict__ dst, int nx, int ny, int nplanes,
                                  SnapshotRegridArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  const size_t plane = (size_t)nx * ny;
  const size_t o = (size_t)y * nx + x;
  const bool keep = x >= a.clip_x0 && x < a.clip_x0 + a.clip_kx && y >= a.clip_y0 && y < a.clip_y0 + a.clip_ky;
  if (!keep) {
    for (int p = 0; p < nplanes; ++p) dst[p * plane + o] = 0.f;
    return;
  }
  // this pixel's projection-plane coordinates and where they fall in the input frame
  const double l = double(x - nx / 2) * a.cellx, m = double(y - ny / 2) * a.celly;
  double sl, sm;
  if (a.to_target) {
    const double n = sqrt(fmax(0., 1. - l * l - m * m));
    sl = l - a.A * (1. - n);
    sm = m - a.B * (1. - n);
  } else {
    const double xi = -a.A, eta = -a.B;
    const double X = l - xi, Y = m - eta;
    const double qa = xi * xi + eta * eta + 1., qb = X * xi + Y * eta, qc = X * X + Y * Y - 1.;
    const double n = (-qb + sqrt(fmax(0., qb * qb - qa * qc))) / qa;
    sl = l - xi * (1. - n);
    sm = m - eta * (1. - n);
  }
  const double px = sl / a.cellx + double(nx / 2), py = sm / a.celly + double(ny / 2);
  const double fx0 = floor(px), fy0 = floor(py);
  const int x0 = (int)fx0, y0 = (int)fy0;
  const double fx = px - fx0, fy = py - fy0;
  const int lo = a.cubic ? -1 : 0, hi = a.cubic ? 2 : 1;
  const bool ok = fx0 + lo >= 0. && fx0 + hi <= double(nx - 1) && fy0 + lo >= 0. && fy0 + hi <= double(ny - 1);
  // interpolation weights in FP32 (the coordinates above are FP64; the images are FP32)
  float wx[4] = {0.f, 0.f, 0.f, 0.f}, wy[4] = {0.f, 0.f, 0.f, 0.f};
  if (a.cubic) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      wx[i] = (float)keys(fx - double(i - 1));
      wy[i] = (float)keys(fy - double(i - 1));
    }
  } else {
    wx[0] = (float)(1. - fx);
    wx[1] = (float)fx;
    wy[0] = (float)(1. - fy);
    wy[1] = (float)fy;
  }
  const size_t s0 = ok ? (size_t)(y0 + lo) * nx + (size_t)(x0 + lo) : 0;
  for (int p = 0; p < nplanes; ++p) {
    float v = 0.f;
    if (ok) {
      const float* s = src + p * plane + s0;
      if (a.cubic) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float* row = s + (size_t)j * nx;
          v += wy[j] * (wx[0] * row[0] + wx[1] * row[1] + wx[2] * row[2] + wx[3] * row[3]);
        }
      } else {
        v = wy[0] * (wx[0] * s[0] + wx[1] * s[1]) + wy[1] * (wx[0] * s[nx] + wx[1] * s[nx + 1]);
      }
    }
    float* d = dst + p * plane + o;
    *d = a.add ? *d + v : v;
  }
}

}  // namespace

void launch_wplane_sums(const Ctx& c, const double* uvw, int nrow, double A, double B, double* sums6) {
  launch_fill_u32(c, sums6, 12, 0u);
  if (nrow <= 0) return;
  const int blocks = std::max(1, std::min(1024, (nrow + 255) / 256));
  hipLaunchKernelGGL(k_wplane_sums, dim3(blocks), dim3(256), 0, c.stream, uvw, nrow, A, B, sums6,
                     reinterpret_cast<unsigned long long*>(sums6 + 5));
  YG_HIP(hipGetLastError());
}

void launch_wplane_apply(const Ctx& c, const double* uvw, int nrow, double A, double B, double* out) {
  if (nrow <= 0) return;
  hipLaunchKernelGGL(k_wplane_apply, dim3((nrow + 255) / 256), dim3(256), 0, c.stream, uvw, nrow, A, B, out);
  YG_HIP(hipGetLastError());
}

void launch_snapshot_regrid(const Ctx& c, const float* src, float* dst, int nx, int ny, int nplanes,
                            const SnapshotRegridArgs& a) {
  if (nx <= 0 || ny <= 0 || nplanes <= 0) return;
  hipLaunchKernelGGL(k_snapshot_regrid, dim3((nx + 127) / 128, ny), dim3(128), 0, c.stream, src, dst, nx, ny, nplanes,
                     a);
  YG_HIP(hipGetLastError());
}

}  // namespace ygrid
