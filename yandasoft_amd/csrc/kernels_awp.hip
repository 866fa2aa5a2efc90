// AWProject convolution functions and weights image on the device
// (AWProjectVisGridder.cc:259-611,618-770; DiskIllumination.cc:85-177, ATCAIllumination.cc:159-331).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "illumination.h"
#include "ygrid_internal.h"

namespace ygrid {

namespace {

__device__ __forceinline__ float cb_sign(int x, int y) { return ((x + y) & 1) ? -1.f : 1.f; }

// Aperture illumination in the uv plane (iU fastest), pre-multiplied by the input checkerboard of
// a centred transform.  One plane per channel of the batch.  As getPattern: the value (double)
// is rounded to complex float and multiplied by the pointing phase slope in complex float, then
// scaled by the normalisation.
__global__ void k_illum_pattern(float2* __restrict__ buf, int nu, int nv, const IllumParams* __restrict__ prm) {
#pragma clang fp contract(off)
  const int iU = blockIdx.x * blockDim.x + threadIdx.x;
  const int iV = blockIdx.y;
  if (iU >= nu) return;
  const IllumParams p = prm[blockIdx.z];
  const double ou = double(iU) - double(nu) / 2., ov = double(iV) - double(nv) / 2.;
  const double r2 = ou * ou + ov * ov;
  float2 v = make_float2(0.f, 0.f);
  if (r2 >= p.rmin2 && r2 <= p.rmax2) {
    double vr, vi;
    illum_value(p, ou, ov, r2, vr, vi);
    const double ph = p.lscaled * ou + p.mscaled * ov;
    const float a = (float)vr, b = (float)vi, c = (float)cos(ph), d = (float)(-sin(ph));
    const float s = p.scale * cb_sign(iU, iV);
    v = make_float2((a * c - b * d) * s, (a * d + b * c) * s);
  }
  buf[(size_t)blockIdx.z * nu * nv + (size_t)iV * nu + iU] = v;
}

// |A|^2 x taper x w-screen (AWProjectVisGridder.cc:397-418) for the w terms of the batch; A is
// the raw backward transform of the disk (its centring sign drops out of |A|^2, its 1/(nu nv)
// normalisation is `ascale`).  Written pre-multiplied by the forward transform's checkerboard.
__global__ void k_aw_screen(const float2* __restrict__ A, int nx, int ny, int qnx, int qny, double ccellx,
                            double ccelly, float ascale, const double* __restrict__ tx, const double* __restrict__ ty,
                            const double* __restrict__ two_pi_w, float2* __restrict__ out) {
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  const int iy = blockIdx.y;
  if (ix >= nx) return;
  const double y = (double(iy) - double(ny) / 2) * ccelly;
  const double x = (double(ix) - double(nx) / 2) * ccellx;
  const double r2 = x * x + y * y;
  const int qix = ix + qnx / 2 - nx / 2, qiy = iy + qny / 2 - ny / 2;
  float2 o = make_float2(0.f, 0.f);
  const bool taper = tx != nullptr;
  if (r2 < 1.0 && (!taper || (qix >= 0 && qix < qnx && qiy >= 0 && qiy < qny))) {
    const double phase = two_pi_w[blockIdx.z] * (1.0 - sqrt(1.0 - r2));
    const float2 a = A[(size_t)iy * nx + ix];
    const float ar = a.x * ascale, ai = a.y * ascale;
    const float t = taper ? (float)(tx[qix] * ty[qiy]) : 1.f;
    const float wt = (ar * ar + ai * ai) * t;
    const float s = cb_sign(ix, iy);
    o = make_float2((float)(wt * cos(phase)) * s, (float)(wt * -sin(phase)) * s);
  }
  out[(size_t)blockIdx.z * nx * ny + (size_t)iy * nx + ix] = o;
}

__global__ void k_scale_sign(float2* data, int nx, int ny, float gsign, float scale) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  float2* p = data + (size_t)blockIdx.z * nx * ny + (size_t)y * nx + x;
  const float s = cb_sign(x, y) * gsign * scale;
  float2 v = *p;
  v.x *= s;
  v.y *= s;
  *p = v;
}

// Weights image: place the frac-0 CF plane of row z at the centre of a cnx x cny plane
// (AWProjectVisGridder.cc:683-697), pre-multiplied by the centred backward transform's checkerboard.
// asym: taps -S .. S-1 only (AProjectWStackVisGridder.cc:432-441)
__global__ void k_place_cf(const float2* __restrict__ table, const int32_t* __restrict__ off, const int* __restrict__ supp,
                           const int* __restrict__ ou, const int* __restrict__ ov, int cnx, int cny, int asym,
                           double2* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, b = blockIdx.z;
  if (x >= cnx) return;
  const int S = supp[b];
  const int hi = asym ? S - 1 : S;
  const int ix = x - cnx / 2 - ou[b], iy = y - cny / 2 - ov[b];
  double2 v = make_double2(0.0, 0.0);
  if (ix >= -S && ix <= hi && iy >= -S && iy <= hi) {
    const int P = 2 * S + 1 + 2 * kGuardU;
    const float2 t = table[off[b] + (size_t)(iy + S + kGuardV) * P + ix + S + kGuardU];
    const double s = cb_sign(x, y);
    v = make_double2(t.x, t.y);
    v.x *= s;
    v.y *= s;
  }
  out[(size_t)b * cnx * cny + (size_t)y * cnx + x] = v;
}

// cOut[cp] += sum_b wt[b][cp] |val_b / taper|^2 where val_b is the raw transform of plane b
// (centring sign drops out of |.|^2).
// inv_max2 (or null): per plane b a factor on |val_b|^2 (1 / max |val_b|^2: AProjectWStack's
// transforms normalised to their peak)
__global__ void k_aw_weights_acc(const double2* __restrict__ planes, int nb, int cnx, int cny, const double* __restrict__ wt,
                                 int ncp, const double* __restrict__ tx, const double* __restrict__ ty,
                                 const unsigned long long* __restrict__ max2_bits, double* __restrict__ cout) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y, cp = blockIdx.z;
  if (x >= cnx) return;
  double t = 1.0;
  if (tx) t = tx[x] * ty[y];
  double acc = 0.0;
  for (int b = 0; b < nb; ++b) {
    const double w = wt[(size_t)b * ncp + cp];
    if (w == 0.0) continue;
    const double2 v = planes[(size_t)b * cnx * cny + (size_t)y * cnx + x];
    double vr = v.x, vi = v.y;
    if (tx) {
      if (t > 0) { vr /= t; vi /= t; } else { vr = vi = 0.0; }
    }
    double p = vr * vr + vi * vi;
    if (max2_bits) {
      const double mx = __longlong_as_double((long long)max2_bits[b]);
      p = mx > 0.0 ? p / mx : 0.0;
    }
    acc += w * p;
  }
  cout[(size_t)cp * cnx * cny + (size_t)y * cnx + x] += acc;
}

// per plane b: the largest |v|^2, as the bits of a non-negative double (they order as the values)
__global__ void k_plane_max2(const double2* __restrict__ planes, int n, unsigned long long* __restrict__ out) {
  const int b = blockIdx.y;
  double m = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double2 v = planes[(size_t)b * n + i];
    m = fmax(m, v.x * v.x + v.y * v.y);
  }
  for (int d = 32; d >= 1; d >>= 1) m = fmax(m, __shfl_xor(m, d));
  if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(&out[b], (unsigned long long)__double_as_longlong(m));
}

// AProjectWStack (AProjectWStackVisGridder.cc:296-317): the aperture power |A|^2 of the
// backward-transformed pattern (raw transform times ascale), in complex float, and its peak
// (float bits; AProjectWStack divides by it before the forward transform)
__global__ void k_aws_power(const float2* __restrict__ A, int n, float ascale, float2* __restrict__ out,
                            uint32_t* __restrict__ peak_bits) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float p = 0.f;
  if (i < n) {
    const float ar = A[i].x * ascale, ai = A[i].y * ascale;
    p = ar * ar + ai * ai;
    out[i] = make_float2(p, 0.f);
  }
  float m = p;
  for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d));
  if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(peak_bits, __float_as_uint(m));
}

// power / peak, pre-multiplied by the forward transform's checkerboard
__global__ void k_aws_norm(float2* __restrict__ buf, int nx, int ny, const uint32_t* __restrict__ peak_bits) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  const float pk = __uint_as_float(*peak_bits);
  const float inv = pk > 0.f ? 1.f / pk : 0.f;
  float2& v = buf[(size_t)y * nx + x];
  v.x = v.x * inv * cb_sign(x, y);
  v.y = v.y * inv * cb_sign(x, y);
}

// Zero-padding of centred spectra: out (nx x ny) <- in (cnx x cny) at the centre.
__global__ void k_pad_centre(const double2* __restrict__ in, int cnx, int cny, int nx, int ny, double2* __restrict__ out) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y;
  if (X >= nx) return;
  const int x = X - (nx - cnx) / 2, y = Y - (ny - cny) / 2;
  double2 v = make_double2(0.0, 0.0);
  if (x >= 0 && x < cnx && y >= 0 && y < cny) v = in[(size_t)blockIdx.z * cnx * cny + (size_t)y * cnx + x];
  out[(size_t)blockIdx.z * nx * ny + (size_t)Y * nx + X] = v;
}

__global__ void k_d2c(const double* __restrict__ in, size_t n, double scale, double2* __restrict__ out) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double2(in[i] * scale, 0.0);
}

// out[plane][y][x] = Re(in[plane][y+oy][x+ox])
__global__ void k_extract_real(const double2* __restrict__ in, int nxp, int nyp, int nx, int ny, float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  const int ox = (nxp - nx) / 2, oy = (nyp - ny) / 2;
  out[(size_t)blockIdx.z * nx * ny + (size_t)y * nx + x] = (float)in[(size_t)blockIdx.z * nxp * nyp + (size_t)(y + oy) * nxp + x + ox].x;
}

}  // namespace

static float gsign(int nx, int ny) { return (((nx / 2) + (ny / 2)) & 1) ? -1.f : 1.f; }

void launch_illum_pattern(const Ctx& c, float2* buf, int nu, int nv, const IllumParams* prm, int nb) {
  hipLaunchKernelGGL(k_illum_pattern, dim3((nu + 255) / 256, nv, nb), dim3(256), 0, c.stream, buf, nu, nv, prm);
  YG_HIP(hipGetLastError());
}

void launch_aw_screen(const Ctx& c, const float2* A, int nx, int ny, int qnx, int qny, double ccellx, double ccelly,
                      float ascale, const double* tx, const double* ty, const double* two_pi_w, int nb, float2* out) {
  hipLaunchKernelGGL(k_aw_screen, dim3((nx + 255) / 256, ny, nb), dim3(256), 0, c.stream, A, nx, ny, qnx, qny, ccellx,
                     ccelly, ascale, tx, ty, two_pi_w, out);
  YG_HIP(hipGetLastError());
}

void launch_scale_sign(const Ctx& c, float2* data, int nx, int ny, int nb, float scale) {
  hipLaunchKernelGGL(k_scale_sign, dim3((nx + 255) / 256, ny, nb), dim3(256), 0, c.stream, data, nx, ny, gsign(nx, ny),
                     scale);
  YG_HIP(hipGetLastError());
}

void launch_place_cf(const Ctx& c, const float2* table, const int32_t* off, const int* supp, const int* ou, const int* ov,
                     int cnx, int cny, int nb, double2* out, int asym) {
  hipLaunchKernelGGL(k_place_cf, dim3((cnx + 255) / 256, cny, nb), dim3(256), 0, c.stream, table, off, supp, ou, ov, cnx,
                     cny, asym, out);
  YG_HIP(hipGetLastError());
}

void launch_aw_weights_acc(const Ctx& c, const double2* planes, int nb, int cnx, int cny, const double* wt, int ncp,
                           const double* tx, const double* ty, double* cout, unsigned long long* max2_tmp) {
  if (max2_tmp) {
    launch_fill_u32(c, max2_tmp, 2 * (size_t)nb, 0u);
    const int n = cnx * cny;
    hipLaunchKernelGGL(k_plane_max2, dim3(std::max(1, std::min(64, (n + 255) / 256)), nb), dim3(256), 0, c.stream,
                       planes, n, max2_tmp);
  }
  hipLaunchKernelGGL(k_aw_weights_acc, dim3((cnx + 255) / 256, cny, ncp), dim3(256), 0, c.stream, planes, nb, cnx, cny,
                     wt, ncp, tx, ty, max2_tmp, cout);
  YG_HIP(hipGetLastError());
}

void launch_aws_cf_power(const Ctx& c, const float2* A, int nx, int ny, float ascale, float2* buf, uint32_t* tmp) {
  launch_fill_u32(c, tmp, 1, 0u);
  const int n = nx * ny;
  hipLaunchKernelGGL(k_aws_power, dim3((n + 255) / 256), dim3(256), 0, c.stream, A, n, ascale, buf, tmp);
  hipLaunchKernelGGL(k_aws_norm, dim3((nx + 255) / 256, ny), dim3(256), 0, c.stream, buf, nx, ny, tmp);
  YG_HIP(hipGetLastError());
}

void launch_pad_centre(const Ctx& c, const double2* in, int cnx, int cny, int nx, int ny, int nb, double2* out) {
  hipLaunchKernelGGL(k_pad_centre, dim3((nx + 255) / 256, ny, nb), dim3(256), 0, c.stream, in, cnx, cny, nx, ny, out);
  YG_HIP(hipGetLastError());
}

void launch_d2c(const Ctx& c, const double* in, size_t n, double scale, double2* out) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_d2c, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, in, n, scale, out);
  YG_HIP(hipGetLastError());
}

void launch_extract_real(const Ctx& c, const double2* in, int nxp, int nyp, int nx, int ny, int nb, float* out) {
  hipLaunchKernelGGL(k_extract_real, dim3((nx + 255) / 256, ny, nb), dim3(256), 0, c.stream, in, nxp, nyp, nx, ny, out);
  YG_HIP(hipGetLastError());
}

}  // namespace ygrid
