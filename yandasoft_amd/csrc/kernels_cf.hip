// Convolution-function generation on the device.
//
// CF tables live in HBM in oversample-major order: plane k = fracu + os*(fracv + os*z)
// (TableVisGridder.cc:720), each stored as (2S+1+2*kGuardV) rows x (2S+1+2*kGuardU) columns
// with the taps in the middle and zeros around them (see kGuardU/kGuardV).
#include <hip/hip_runtime.h>

#include "ygrid_internal.h"

namespace ygrid {

namespace {

__device__ __forceinline__ float cb_sign(int x, int y) { return ((x + y) & 1) ? -1.f : 1.f; }

// Taper x w-term phase screen, written pre-multiplied by the input checkerboard of the
// centred FFT (WProjectVisGridder::populateThisPlane, WProjectVisGridder.cc:632-660).
__global__ void k_wproj_screen(float2* __restrict__ buf, int nx, int ny, int qnx, int qny, double ccellx, double ccelly,
                               const float* __restrict__ ccfx, const float* __restrict__ ccfy,
                               const double* __restrict__ two_pi_w) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y;
  if (X >= nx) return;
  const int ix = X - nx / 2 + qnx / 2, iy = Y - ny / 2 + qny / 2;
  float2 out = make_float2(0.f, 0.f);
  if (ix >= 0 && ix < qnx && iy >= 0 && iy < qny) {
    double y2 = (double)(iy - qny / 2) * ccelly;
    y2 *= y2;
    double x2 = (double)(ix - qnx / 2) * ccellx;
    x2 *= x2;
    const double r2 = x2 + y2;
    if (r2 < 1.0) {
      const double phase = two_pi_w[blockIdx.z] * (1.0 - sqrt(1.0 - r2));
      const float wt = ccfx[ix] * ccfy[iy];
      const float s = cb_sign(X, Y);
      out = make_float2((float)(wt * cos(phase)) * s, (float)(-wt * sin(phase)) * s);
    }
  }
  buf[(size_t)blockIdx.z * nx * ny + (size_t)Y * nx + X] = out;
}

// argmax of |c|^2 (computed in double as SupportSearcher's ampFunction) with the lowest flat
// index among equal maxima (column-major first occurrence, SupportSearcher.tcc:98-124).
__device__ __forceinline__ unsigned long long pack_amp(double a, unsigned idx) {
  // a >= 0: the IEEE bit pattern of a float is monotone; keep the higher 32 bits for the
  // value and invert the index so that larger packed = larger value, then smaller index.
  const float af = (float)a;
  return ((unsigned long long)__float_as_uint(af) << 32) | (unsigned long long)(0xffffffffu - idx);
}

__global__ void k_argmax(const float2* __restrict__ plane, int n, unsigned long long* out) {
  unsigned long long best = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float2 v = plane[i];
    const double a = (double)v.x * v.x + (double)v.y * v.y;
    const unsigned long long pk = pack_amp(a, (unsigned)i);
    best = pk > best ? pk : best;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = __shfl_xor(best, o, 64);
    best = t > best ? t : best;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(out, best);
}

__device__ double block_sum(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = blockDim.x >> 6;
  for (int i = 0; i < nw; ++i) t += sh[i];
  return t;
}

// One workgroup per CF plane k in the batch (blockIdx.x = fracu + os*fracv, blockIdx.y = b).
// mode 0: normalise by the sum of real parts (WProjectVisGridder::normalise, :612-631);
// mode 1: multiply by `rescale` (AWProjectVisGridder.cc:482,526).
__global__ __launch_bounds__(256) void k_cf_cutout(const float2* __restrict__ buf, int nx, int ny, int os,
                                                   const int* __restrict__ supp, const int* __restrict__ offu,
                                                   const int* __restrict__ offv, const int32_t* __restrict__ cf_off_first,
                                                   int mode, double rescale, float2* __restrict__ table) {
  __shared__ double sh[8];
  const int frac = blockIdx.x, b = blockIdx.y;
  const int fracu = frac % os, fracv = frac / os;
  const int S = supp[b], ou = offu ? offu[b] : 0, ov = offv ? offv[b] : 0;
  const int W = 2 * S + 1;
  const float2* src = buf + (size_t)b * nx * ny;
  // WProjectVisGridder::populateItsConvFunc (:588-611): tap (ix+S, iy+S) <- buf(kx, ky)
  double local = 0.0;
  for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
    const int ix = t % W - S, iy = t / W - S;
    const int kx = (ix + ou) * os + fracu + nx / 2, ky = (iy + ov) * os + fracv + ny / 2;
    local += (double)src[(size_t)ky * nx + kx].x;
  }
  double scale = rescale;
  if (mode == 0) {
    const double norm = block_sum(local, sh);
    scale = norm > 0.0 ? 1.0 / norm : 1.0;
  }
  const float fs = (float)scale;
  const int P = W + 2 * kGuardU, R = W + 2 * kGuardV;
  // the os^2 planes of one z share S, so they are consecutive blocks of P*R elements
  float2* dst = table + cf_off_first[b] + (size_t)frac * (size_t)P * R;
  // mode 2: taps -S .. S-1 only, the last row and column of the plane zero (the cut-out loops of
  // AProjectWStackVisGridder.cc:342-350)
  const int Wc = mode == 2 ? W - 1 : W;
  for (int t = threadIdx.x; t < P * R; t += blockDim.x) {
    const int du = t % P - kGuardU, dv = t / P - kGuardV;
    float2 v = make_float2(0.f, 0.f);
    if (du >= 0 && du < Wc && dv >= 0 && dv < Wc) {
      const int kx = (du - S + ou) * os + fracu + nx / 2, ky = (dv - S + ov) * os + fracv + ny / 2;
      const float2 c = src[(size_t)ky * nx + kx];
      v = make_float2(c.x * fs, c.y * fs);
    }
    dst[t] = v;
  }
}

// SphFunc CF plane frac = fracu + os*fracv: C(ix,iy) = bx[fracu][ix]*by[fracv][iy] / sum|C|
// (SphFuncVisGridder.cc:137-167).
__global__ __launch_bounds__(256) void k_sph_cf(const double* __restrict__ bx, const double* __restrict__ by, int os,
                                                int S, const int32_t* __restrict__ cf_off, float2* __restrict__ table) {
  __shared__ double sh[8];
  const int frac = blockIdx.x;
  const int fracu = frac % os, fracv = frac / os;
  const int W = 2 * S + 1;
  double local = 0.0;
  for (int t = threadIdx.x; t < W * W; t += blockDim.x) {
    const int ix = t % W, iy = t / W;
    local += fabs(bx[fracu * W + ix] * by[fracv * W + iy]);
  }
  const double norm = block_sum(local, sh);
  const double inv = norm > 0.0 ? 1.0 / norm : 1.0;
  const int P = W + 2 * kGuardU, R = W + 2 * kGuardV;
  float2* dst = table + cf_off[frac];
  for (int t = threadIdx.x; t < P * R; t += blockDim.x) {
    const int du = t % P - kGuardU, dv = t / P - kGuardV;
    float v = 0.f;
    if (du >= 0 && du < W && dv >= 0 && dv < W) v = (float)(bx[fracu * W + du] * by[fracv * W + dv] * inv);
    dst[t] = make_float2(v, 0.f);
  }
}

__global__ void k_gather_cross(const float2* __restrict__ plane, int nx, int ny, int px, int py, float2* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nx) out[i] = plane[(size_t)py * nx + i];
  if (i < ny) out[nx + i] = plane[(size_t)i * nx + px];
}

}  // namespace

void launch_wproj_screen(const Ctx& c, float2* buf, int nx, int ny, int qnx, int qny, double ccellx, double ccelly,
                         const float* ccfx, const float* ccfy, const double* two_pi_w, int nb) {
  hipLaunchKernelGGL(k_wproj_screen, dim3((nx + 255) / 256, ny, nb), dim3(256), 0, c.stream, buf, nx, ny, qnx, qny,
                     ccellx, ccelly, ccfx, ccfy, two_pi_w);
  YG_HIP(hipGetLastError());
}

void launch_argmax(const Ctx& c, const float2* plane, int n, unsigned long long* out) {
  YG_HIP(hipMemsetAsync(out, 0, sizeof(unsigned long long), c.stream));
  int blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_argmax, dim3(blocks), dim3(256), 0, c.stream, plane, n, out);
  YG_HIP(hipGetLastError());
}

void launch_gather_cross(const Ctx& c, const float2* plane, int nx, int ny, int px, int py, float2* out) {
  const int n = nx > ny ? nx : ny;
  hipLaunchKernelGGL(k_gather_cross, dim3((n + 255) / 256), dim3(256), 0, c.stream, plane, nx, ny, px, py, out);
  YG_HIP(hipGetLastError());
}

void launch_cf_cutout(const Ctx& c, const float2* buf, int nx, int ny, int nb, int os, const int* supp, const int* offu,
                      const int* offv, const int32_t* cf_off_first, int mode, double rescale, float2* table) {
  hipLaunchKernelGGL(k_cf_cutout, dim3(os * os, nb), dim3(256), 0, c.stream, buf, nx, ny, os, supp, offu, offv,
                     cf_off_first, mode, rescale, table);
  YG_HIP(hipGetLastError());
}

void launch_sph_cf(const Ctx& c, const double* bx, const double* by, int os, int support, const int32_t* cf_off,
                   float2* table) {
  hipLaunchKernelGGL(k_sph_cf, dim3(os * os), dim3(256), 0, c.stream, bx, by, os, support, cf_off, table);
  YG_HIP(hipGetLastError());
}

}  // namespace ygrid
