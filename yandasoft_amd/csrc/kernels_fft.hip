// Image <-> grid transforms on the device.
//
// The reference transforms with a centred FFT (the origin sits at element n/2 on each axis;
// forward exp(-i) unnormalised, inverse normalised by 1/n: scimath::fft2d as used at
// TableVisGridder.cc:1215,1312).  For even n a centred transform is the plain transform
// between two checkerboard multiplications by (-1)^(x+y), with an extra (-1)^(nx/2+ny/2) on
// the output; the checkerboards are fused into the kernels that read or write the buffers.
// finaliseGrid's 1/(nx*ny) of the inverse transform and its later *(nx*ny)
// (TableVisGridder.cc:1225-1226) cancel; both are dropped.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include "ygrid_internal.h"

namespace ygrid {

struct FFTPlan {
  hipfftHandle h = 0;
  int nx = 0, ny = 0, batch = 0;
  hipfftType type = HIPFFT_C2C;
};

void destroy_plan(FFTPlan* p) {
  if (p) {
    if (p->h) (void)hipfftDestroy(p->h);
    delete p;
  }
}

namespace {

__device__ __forceinline__ float cb_sign(int x, int y) { return ((x + y) & 1) ? -1.f : 1.f; }

__global__ void k_checkerboard(float2* data, int nx, int ny, float gsign) {
  const size_t plane = (size_t)nx * ny;
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  float2* p = data + blockIdx.z * plane + (size_t)y * nx + x;
  const float s = cb_sign(x, y) * gsign;
  float2 v = *p;
  v.x *= s;
  v.y *= s;
  *p = v;
}

// out[plane][y][x] = sign * Re(fftd[plane][y+oy][x+ox]) * ccfx[x+ox] * ccfy[y+oy]
__global__ void k_finalise_image(const float2* __restrict__ fftd, int nxp, int nyp, const double* __restrict__ ccfx,
                                 const double* __restrict__ ccfy, int nx, int ny, int ox, int oy, float gsign,
                                 float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  const int X = x + ox, Y = y + oy;
  const float2 v = fftd[(size_t)blockIdx.z * nxp * nyp + (size_t)Y * nxp + X];
  const double c = (ccfx ? ccfx[X] : 1.0) * (ccfy ? ccfy[Y] : 1.0);
  out[(size_t)blockIdx.z * nx * ny + (size_t)y * nx + x] = (float)((double)(cb_sign(X, Y) * gsign * v.x) * c);
}

// grid[plane][Y][X] = checkerboard * ccf * padded(model)
__global__ void k_model_to_grid(const float* __restrict__ model, int nx, int ny, int ox, int oy,
                                const double* __restrict__ ccfx, const double* __restrict__ ccfy, int nxp, int nyp,
                                float2* __restrict__ grid) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y;
  if (X >= nxp) return;
  const int x = X - ox, y = Y - oy;
  float val = 0.f;
  if (x >= 0 && x < nx && y >= 0 && y < ny) {
    const double c = (ccfx ? ccfx[X] : 1.0) * (ccfy ? ccfy[Y] : 1.0);
    val = (float)((double)model[(size_t)blockIdx.z * nx * ny + (size_t)y * nx + x] * c) * cb_sign(X, Y);
  }
  grid[(size_t)blockIdx.z * nxp * nyp + (size_t)Y * nxp + X] = make_float2(val, 0.f);
}

}  // namespace

static float global_sign(int nx, int ny) { return (((nx / 2) + (ny / 2)) & 1) ? -1.f : 1.f; }

void launch_checkerboard(const Ctx& c, float2* data, int nx, int ny, int batch) {
  hipLaunchKernelGGL(k_checkerboard, dim3((nx + 255) / 256, ny, batch), dim3(256), 0, c.stream, data, nx, ny, 1.f);
  YG_HIP(hipGetLastError());
}

static void ensure_plan(const Ctx& c, int nx, int ny, int batch, PlanPtr& cache, hipfftType type = HIPFFT_C2C) {
  if (cache && cache->nx == nx && cache->ny == ny && cache->batch == batch && cache->type == type) {
    YG_CHECK(hipfftSetStream(cache->h, c.stream) == HIPFFT_SUCCESS, 3, "hipfftSetStream failed");
    return;
  }
  if (cache) {
    cache.reset();
  }
  auto* p = new FFTPlan();
  int n[2] = {ny, nx};
  hipfftResult r = hipfftPlanMany(&p->h, 2, n, nullptr, 1, nx * ny, nullptr, 1, nx * ny, type, batch);
  if (r != HIPFFT_SUCCESS) {
    delete p;
    throw Error(3, "hipfftPlanMany failed (" + std::to_string((int)r) + ")");
  }
  p->nx = nx;
  p->ny = ny;
  p->batch = batch;
  p->type = type;
  cache.reset(p);
  YG_CHECK(hipfftSetStream(p->h, c.stream) == HIPFFT_SUCCESS, 3, "hipfftSetStream failed");
}

// In-place centred transform of `batch` planes (input and output both centred).
void fft_centred_c2c(const Ctx& c, float2* data, int nx, int ny, int batch, bool forward,
                     PlanPtr& cache) {
  YG_CHECK(nx % 2 == 0 && ny % 2 == 0, 1, "centred FFT needs even sizes");
  ensure_plan(c, nx, ny, batch, cache);
  hipLaunchKernelGGL(k_checkerboard, dim3((nx + 255) / 256, ny, batch), dim3(256), 0, c.stream, data, nx, ny, 1.f);
  hipfftResult r = hipfftExecC2C(cache->h, reinterpret_cast<hipfftComplex*>(data), reinterpret_cast<hipfftComplex*>(data),
                                 forward ? HIPFFT_FORWARD : HIPFFT_BACKWARD);
  YG_CHECK(r == HIPFFT_SUCCESS, 3, "hipfftExecC2C failed");
  hipLaunchKernelGGL(k_checkerboard, dim3((nx + 255) / 256, ny, batch), dim3(256), 0, c.stream, data, nx, ny,
                     global_sign(nx, ny));
  YG_HIP(hipGetLastError());
}

// fftd holds the un-centred result of a backward transform of a checkerboarded grid.
void launch_finalise_image(const Ctx& c, const float2* fftd, int nxp, int nyp, int nplanes, const double* ccfx,
                           const double* ccfy, int nx, int ny, float* out) {
  const int ox = (nxp - nx) / 2, oy = (nyp - ny) / 2;
  hipLaunchKernelGGL(k_finalise_image, dim3((nx + 255) / 256, ny, nplanes), dim3(256), 0, c.stream, fftd, nxp, nyp,
                     ccfx, ccfy, nx, ny, ox, oy, global_sign(nxp, nyp), out);
  YG_HIP(hipGetLastError());
}

void launch_model_to_grid(const Ctx& c, const float* model, int nx, int ny, int nplanes, const double* ccfx,
                          const double* ccfy, int nxp, int nyp, float2* grid) {
  const int ox = (nxp - nx) / 2, oy = (nyp - ny) / 2;
  hipLaunchKernelGGL(k_model_to_grid, dim3((nxp + 255) / 256, nyp, nplanes), dim3(256), 0, c.stream, model, nx, ny, ox,
                     oy, ccfx, ccfy, nxp, nyp, grid);
  YG_HIP(hipGetLastError());
}

// Plain (un-centred) transform used after the callers have applied the input checkerboard.
void fft_plain_c2c(const Ctx& c, float2* data, int nx, int ny, int batch, bool forward, PlanPtr& cache) {
  ensure_plan(c, nx, ny, batch, cache);
  hipfftResult r = hipfftExecC2C(cache->h, reinterpret_cast<hipfftComplex*>(data), reinterpret_cast<hipfftComplex*>(data),
                                 forward ? HIPFFT_FORWARD : HIPFFT_BACKWARD);
  YG_CHECK(r == HIPFFT_SUCCESS, 3, "hipfftExecC2C failed");
}

void launch_checkerboard_sign(const Ctx& c, float2* data, int nx, int ny, int batch) {
  hipLaunchKernelGGL(k_checkerboard, dim3((nx + 255) / 256, ny, batch), dim3(256), 0, c.stream, data, nx, ny,
                     global_sign(nx, ny));
  YG_HIP(hipGetLastError());
}

// ---- FP64 image <-> grid transforms (finaliseGrid / initialiseDegrid) ----
namespace {
__global__ void k_grid_to_z(const float2* __restrict__ g, int nx, int ny, double2* __restrict__ z) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  const size_t i = (size_t)blockIdx.z * nx * ny + (size_t)y * nx + x;
  const float2 v = g[i];
  const double s = cb_sign(x, y);
  z[i] = make_double2(v.x * s, v.y * s);
}

__global__ void k_finalise_image_z(const double2* __restrict__ fftd, int nxp, int nyp, const double* __restrict__ ccfx,
                                   const double* __restrict__ ccfy, int nx, int ny, int ox, int oy, double gsign,
                                   float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  const int X = x + ox, Y = y + oy;
  const double2 v = fftd[(size_t)blockIdx.z * nxp * nyp + (size_t)Y * nxp + X];
  const double c = (ccfx ? ccfx[X] : 1.0) * (ccfy ? ccfy[Y] : 1.0);
  out[(size_t)blockIdx.z * nx * ny + (size_t)y * nx + x] = (float)(cb_sign(X, Y) * gsign * v.x * c);
}

__global__ void k_model_to_z(const float* __restrict__ model, int nx, int ny, int ox, int oy,
                             const double* __restrict__ ccfx, const double* __restrict__ ccfy, int nxp, int nyp,
                             double2* __restrict__ z) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y;
  if (X >= nxp) return;
  const int x = X - ox, y = Y - oy;
  double val = 0.0;
  if (x >= 0 && x < nx && y >= 0 && y < ny) {
    const double c = (ccfx ? ccfx[X] : 1.0) * (ccfy ? ccfy[Y] : 1.0);
    val = (double)model[(size_t)blockIdx.z * nx * ny + (size_t)y * nx + x] * c * cb_sign(X, Y);
  }
  z[(size_t)blockIdx.z * nxp * nyp + (size_t)Y * nxp + X] = make_double2(val, 0.0);
}

__global__ void k_z_to_grid(const double2* __restrict__ z, int nx, int ny, double gsign, float2* __restrict__ g) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  const size_t i = (size_t)blockIdx.z * nx * ny + (size_t)y * nx + x;
  const double s = cb_sign(x, y) * gsign;
  const double2 v = z[i];
  g[i] = make_float2((float)(v.x * s), (float)(v.y * s));
}
}  // namespace

void fft_plain_z2z(const Ctx& c, double2* data, int nx, int ny, int batch, bool forward, PlanPtr& cache) {
  YG_CHECK(nx % 2 == 0 && ny % 2 == 0, 1, "centred FFT needs even sizes");
  ensure_plan(c, nx, ny, batch, cache, HIPFFT_Z2Z);
  hipfftResult r = hipfftExecZ2Z(cache->h, reinterpret_cast<hipfftDoubleComplex*>(data),
                                 reinterpret_cast<hipfftDoubleComplex*>(data), forward ? HIPFFT_FORWARD : HIPFFT_BACKWARD);
  YG_CHECK(r == HIPFFT_SUCCESS, 3, "hipfftExecZ2Z failed");
}
static void exec_z2z(const Ctx& c, double2* data, int nx, int ny, int batch, bool forward, PlanPtr& cache) {
  fft_plain_z2z(c, data, nx, ny, batch, forward, cache);
}

namespace {
__global__ void k_checkerboard_z(double2* data, int nx, int ny, double gsign) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  double2* p = data + (size_t)blockIdx.z * nx * ny + (size_t)y * nx + x;
  const double s = cb_sign(x, y) * gsign;
  double2 v = *p;
  v.x *= s;
  v.y *= s;
  *p = v;
}
}  // namespace

void fft_centred_z2z(const Ctx& c, double2* data, int nx, int ny, int batch, bool forward, PlanPtr& cache) {
  const dim3 g((nx + 255) / 256, ny, batch);
  hipLaunchKernelGGL(k_checkerboard_z, g, dim3(256), 0, c.stream, data, nx, ny, 1.0);
  fft_plain_z2z(c, data, nx, ny, batch, forward, cache);
  hipLaunchKernelGGL(k_checkerboard_z, g, dim3(256), 0, c.stream, data, nx, ny, (double)global_sign(nx, ny));
  YG_HIP(hipGetLastError());
}

// image[plane] = ccf * Re(centred unnormalised backward transform of grid[plane]), unpadded centre
void grid_to_image_z(const Ctx& c, const float2* grid, int nxp, int nyp, int nplanes, const double* ccfx,
                     const double* ccfy, int nx, int ny, double2* scratch, float* out, PlanPtr& cache) {
  const dim3 gp((nxp + 255) / 256, nyp, nplanes);
  hipLaunchKernelGGL(k_grid_to_z, gp, dim3(256), 0, c.stream, grid, nxp, nyp, scratch);
  exec_z2z(c, scratch, nxp, nyp, nplanes, false, cache);
  hipLaunchKernelGGL(k_finalise_image_z, dim3((nx + 255) / 256, ny, nplanes), dim3(256), 0, c.stream, scratch, nxp, nyp,
                     ccfx, ccfy, nx, ny, (nxp - nx) / 2, (nyp - ny) / 2, (double)global_sign(nxp, nyp), out);
  YG_HIP(hipGetLastError());
}

namespace {
// WStackVisGridder::multiply's screen at padded pixel (X, Y), in its float arithmetic
// (WStackVisGridder.cc:134-171); false outside r^2 < 1 (left unchanged there).
__device__ __forceinline__ bool wstack_screen(int X, int Y, int nxp, int nyp, float w2pi, float cellx, float celly,
                                              float& cr, float& ci) {
#pragma clang fp contract(off)
  float y2 = float(Y - nyp / 2) * celly;
  y2 *= y2;
  float x2 = float(X - nxp / 2) * cellx;
  x2 *= x2;
  const float r2 = x2 + y2;
  if (!((double)r2 < 1.0)) return false;
  const float phase = (float)((double)w2pi * (1.0 - sqrt(1.0 - (double)r2)));
  cr = cosf(phase);
  ci = -sinf(phase);
  return true;
}

__global__ void k_wstack_acc(const double2* __restrict__ z, int nxp, int nyp, float w2pi, float cellx, float celly,
                             double gsign, double* __restrict__ acc) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y;
  if (X >= nxp) return;
  const size_t i = (size_t)blockIdx.z * nxp * nyp + (size_t)Y * nxp + X;
  const double s = cb_sign(X, Y) * gsign;
  const double2 v = make_double2(z[i].x * s, z[i].y * s);
  float cr, ci;
  acc[i] += wstack_screen(X, Y, nxp, nyp, w2pi, cellx, celly, cr, ci) ? v.x * (double)cr - v.y * (double)ci : v.x;
}

__global__ void k_wstack_final(const double* __restrict__ acc, int nxp, int nyp, const double* __restrict__ ccfx,
                               const double* __restrict__ ccfy, int nx, int ny, int ox, int oy,
                               float* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y;
  if (x >= nx) return;
  const int X = x + ox, Y = y + oy;
  const double c = (ccfx ? ccfx[X] : 1.0) * (ccfy ? ccfy[Y] : 1.0);
  out[(size_t)blockIdx.z * nx * ny + (size_t)y * nx + x] = (float)(acc[(size_t)blockIdx.z * nxp * nyp + (size_t)Y * nxp + X] * c);
}

// conj(screen * m): m (cos p, sin p) inside r^2 < 1, (m, 0) outside; centring signs applied
__global__ void k_wstack_model_to_z(const float* __restrict__ model, int nx, int ny, int ox, int oy,
                                    const double* __restrict__ ccfx, const double* __restrict__ ccfy, int nxp, int nyp,
                                    float w2pi, float cellx, float celly, double2* __restrict__ z) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y;
  if (X >= nxp) return;
  const int x = X - ox, y = Y - oy;
  double2 val = make_double2(0.0, 0.0);
  if (x >= 0 && x < nx && y >= 0 && y < ny) {
    const double c = (ccfx ? ccfx[X] : 1.0) * (ccfy ? ccfy[Y] : 1.0);
    const double m = (double)model[(size_t)blockIdx.z * nx * ny + (size_t)y * nx + x] * c * cb_sign(X, Y);
    float cr, ci;
    if (m != 0.0 && wstack_screen(X, Y, nxp, nyp, w2pi, cellx, celly, cr, ci))
      val = make_double2(m * (double)cr, -m * (double)ci);
    else
      val = make_double2(m, 0.0);
  }
  z[(size_t)blockIdx.z * nxp * nyp + (size_t)Y * nxp + X] = val;
}
}  // namespace

void wstack_accumulate(const Ctx& c, const float2* grid, int nxp, int nyp, int nplanes, float w2pi, float cellx,
                       float celly, double2* scratch, double* acc, PlanPtr& cache) {
  const dim3 gp((nxp + 255) / 256, nyp, nplanes);
  hipLaunchKernelGGL(k_grid_to_z, gp, dim3(256), 0, c.stream, grid, nxp, nyp, scratch);
  exec_z2z(c, scratch, nxp, nyp, nplanes, false, cache);
  hipLaunchKernelGGL(k_wstack_acc, gp, dim3(256), 0, c.stream, scratch, nxp, nyp, w2pi, cellx, celly,
                     (double)global_sign(nxp, nyp), acc);
  YG_HIP(hipGetLastError());
}

namespace {
// stack += norm * cb(X, Y) * (Re g, im_scale * Im g): one plane of the w-stacking
// preconditioner's uv stack (WStackVisGridder::updatew, WStackVisGridder.cc:173-209, scales the
// imaginary part where |g| > 0; zero pixels are unchanged either way), with the centring signs
// and the inverse transform's normalisation applied here (linear)
__global__ void k_wstack_pcf_stack(const float2* __restrict__ g, int nxp, int nyp, double im_scale, double norm,
                                   double2* __restrict__ stack) {
  const int X = blockIdx.x * blockDim.x + threadIdx.x;
  const int Y = blockIdx.y;
  if (X >= nxp) return;
  const size_t i = (size_t)blockIdx.z * nxp * nyp + (size_t)Y * nxp + X;
  const float2 v = g[i];
  const double s = cb_sign(X, Y) * norm;
  double2 a = stack[i];
  a.x += s * (double)v.x;
  a.y += s * (double)v.y * im_scale;
  stack[i] = a;
}
}  // namespace

void wstack_pcf_stack(const Ctx& c, const float2* grid, int nxp, int nyp, int nplanes, double im_scale, double norm,
                      double2* stack) {
  hipLaunchKernelGGL(k_wstack_pcf_stack, dim3((nxp + 255) / 256, nyp, nplanes), dim3(256), 0, c.stream, grid, nxp, nyp,
                     im_scale, norm, stack);
  YG_HIP(hipGetLastError());
}

// image[plane] = Re(backward transform of a stack that carries its centring signs), unpadded
// centre, no correction
void stack_to_image_z(const Ctx& c, double2* stack, int nxp, int nyp, int nplanes, int nx, int ny, float* out,
                      PlanPtr& cache) {
  exec_z2z(c, stack, nxp, nyp, nplanes, false, cache);
  hipLaunchKernelGGL(k_finalise_image_z, dim3((nx + 255) / 256, ny, nplanes), dim3(256), 0, c.stream, stack, nxp, nyp,
                     nullptr, nullptr, nx, ny, (nxp - nx) / 2, (nyp - ny) / 2, (double)global_sign(nxp, nyp), out);
  YG_HIP(hipGetLastError());
}

void wstack_finalise(const Ctx& c, const double* acc, int nxp, int nyp, int nplanes, const double* ccfx,
                     const double* ccfy, int nx, int ny, float* out) {
  hipLaunchKernelGGL(k_wstack_final, dim3((nx + 255) / 256, ny, nplanes), dim3(256), 0, c.stream, acc, nxp, nyp, ccfx,
                     ccfy, nx, ny, (nxp - nx) / 2, (nyp - ny) / 2, out);
  YG_HIP(hipGetLastError());
}

void wstack_model_to_grid(const Ctx& c, const float* model, int nx, int ny, int nplanes, const double* ccfx,
                          const double* ccfy, int nxp, int nyp, float w2pi, float cellx, float celly, double2* scratch,
                          float2* grid, PlanPtr& cache) {
  const dim3 gp((nxp + 255) / 256, nyp, nplanes);
  hipLaunchKernelGGL(k_wstack_model_to_z, gp, dim3(256), 0, c.stream, model, nx, ny, (nxp - nx) / 2, (nyp - ny) / 2,
                     ccfx, ccfy, nxp, nyp, w2pi, cellx, celly, scratch);
  exec_z2z(c, scratch, nxp, nyp, nplanes, true, cache);
  hipLaunchKernelGGL(k_z_to_grid, gp, dim3(256), 0, c.stream, scratch, nxp, nyp, (double)global_sign(nxp, nyp), grid);
  YG_HIP(hipGetLastError());
}

// grid[plane] = centred unnormalised forward transform of ccf * padded(model[plane])
void model_to_grid_z(const Ctx& c, const float* model, int nx, int ny, int nplanes, const double* ccfx,
                     const double* ccfy, int nxp, int nyp, double2* scratch, float2* grid, PlanPtr& cache) {
  const dim3 gp((nxp + 255) / 256, nyp, nplanes);
  hipLaunchKernelGGL(k_model_to_z, gp, dim3(256), 0, c.stream, model, nx, ny, (nxp - nx) / 2, (nyp - ny) / 2, ccfx, ccfy,
                     nxp, nyp, scratch);
  exec_z2z(c, scratch, nxp, nyp, nplanes, true, cache);
  hipLaunchKernelGGL(k_z_to_grid, gp, dim3(256), 0, c.stream, scratch, nxp, nyp, (double)global_sign(nxp, nyp), grid);
  YG_HIP(hipGetLastError());
}

}  // namespace ygrid

namespace ygrid {
namespace {
__global__ void k_any_nonzero(const float* __restrict__ d, size_t n, int* flag) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    if (d[i] != 0.f) { *flag = 1; return; }
}
}  // namespace
// isModelEmpty test of TableVisGridder::initialiseDegrid (TableVisGridder.cc:1304)
void launch_any_nonzero(const Ctx& c, const float* data, size_t n, int* flag) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_any_nonzero, dim3(1024), dim3(256), 0, c.stream, data, n, flag);
  YG_HIP(hipGetLastError());
}
}  // namespace ygrid
