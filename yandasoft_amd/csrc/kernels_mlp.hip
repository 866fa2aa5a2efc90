// Learned convolution functions: a coordinate network (dense layers with SiLU, complex output)
// that maps (w, fractional oversampling offset u, v, tap dx, dy) to a CF tap, trained on the CF
// tables the W-projection generator produces and evaluated per sample at its exact w.
//
//   k_gemm        C = op(A) op(B) on 64x64 tiles on the matrix cores (FP32 MFMA 16x16x4, each
//                 wave a 32x32 quarter in 2x2 blocks) with the epilogues a dense layer needs:
//                 + bias, SiLU (keeping the pre-activation), times SiLU'(Z) for the backward
//                 pass; B may be FP16 (inference weights, widened to FP32 as they are staged).
//   k_colsum      bias gradients;  k_mse  loss and its gradient;  k_adam  the optimiser step.
//   k_lcf_inputs  per-sample network inputs;  k_lcf_scatter  network outputs into per-sample
//                 padded CF planes (the layout the tile kernels read).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "learned_cf.h"

namespace ygrid {
namespace {

constexpr int kT = 64;   // output tile
constexpr int kK = 16;   // k step
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float sigm(float z) { return 1.f / (1.f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigm(z);
  return s * (1.f + z * (1.f - s));
}

template <typename T>
__device__ __forceinline__ float ld(const T* p, size_t i) {
  return (float)p[i];
}

// C[M x N] (row stride ldc) = op(A) op(B) (+ epilogue), op(A) M x K, op(B) K x N.
// A: TA ? A[k*lda + m] : A[m*lda + k];  B: TB ? B[n*ldb + k] : B[k*ldb + n].
template <bool TA, bool TB, typename TBt, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const float* __restrict__ A, const TBt* __restrict__ B,
                                              float* __restrict__ C, int M, int N, int K, int lda, int ldb, int ldc,
                                              const float* __restrict__ bias, float* __restrict__ Zout,
                                              const float* __restrict__ Zin, int kslice) {
  __shared__ float sA[kK][kT + 4];
  __shared__ float sB[kK][kT + 4];
  const int tid = threadIdx.x;
  const int m0 = blockIdx.y * kT, n0 = blockIdx.x * kT;
  // wave (wm, wn) computes rows 32wm .. +31 and columns 32wn .. +31 of the tile in 2x2 blocks of
  // 16x16.  v_mfma_f32_16x16x4_f32: lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16] of
  // a block's 16x4 and 4x16 operands, and holds D[4 (l / 16) + r][l % 16] in register r.
  const int lane = tid & 63, wv = tid >> 6;
  const int wm = wv >> 1, wn = wv & 1;
  const int li = lane & 15, lk = lane >> 4;
  f32x4 acc[2][2];
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = f32x4{0.f, 0.f, 0.f, 0.f};
  // split K (gridDim.z > 1, no epilogue): this workgroup sums k in [kb, ke) and adds its partial
  // tile into C, which the caller has zeroed
  const int kb = blockIdx.z * kslice, ke = min(K, kb + kslice);
  for (int k0 = kb; k0 < ke; k0 += kK) {
    // 64 x 16 of A and 16 x 64 of B: 1024 values each, 4 per thread
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int e = tid + 256 * r;
      {
        int mm, kk;
        if (TA) { mm = e % kT; kk = e / kT; } else { kk = e % kK; mm = e / kK; }
        const int m = m0 + mm, k = k0 + kk;
        sA[kk][mm] = (m < M && k < ke) ? (TA ? A[(size_t)k * lda + m] : A[(size_t)m * lda + k]) : 0.f;
      }
      {
        int nn, kk;
        if (TB) { kk = e % kK; nn = e / kK; } else { nn = e % kT; kk = e / kT; }
        const int n = n0 + nn, k = k0 + kk;
        sB[kk][nn] = (n < N && k < ke) ? (TB ? ld(B, (size_t)n * ldb + k) : ld(B, (size_t)k * ldb + n)) : 0.f;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kK; kk += 4) {
      float a[2], b[2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) a[bi] = sA[kk + lk][32 * wm + 16 * bi + li];
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) b[bj] = sB[kk + lk][32 * wn + 16 * bj + li];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[bi], b[bj], acc[bi][bj], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 32 * wm + 16 * bi + 4 * lk + r;
        const int n = n0 + 32 * wn + 16 * bj + li;
        if (m >= M || n >= N) continue;
        const size_t o = (size_t)m * ldc + n;
        float v = acc[bi][bj][r];
        if (EPI == kEpiBiasSilu) {
          v += bias[n];
          if (Zout) Zout[o] = v;
          v = silu(v);
        } else if (EPI == kEpiBias) {
          v += bias[n];
        } else if (EPI == kEpiSiluGrad) {
          v *= silu_grad(Zin[o]);
        }
        if (gridDim.z > 1) atomicAdd(&C[o], v);
        else C[o] = v;
      }
}

// db[n] = sum over m of D[m * ldd + n]
__global__ void k_colsum(const float* __restrict__ D, int M, int N, int ldd, float* __restrict__ db) {
  __shared__ float s[256];
  const int n = blockIdx.x;
  float a = 0.f;
  for (int m = threadIdx.x; m < M; m += blockDim.x) a += D[(size_t)m * ldd + n];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[n] = s[0];
}

// loss = sum (y - t)^2 / (2M) over M rows of 2 outputs; dY = (y - t) / M
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, int M, float* __restrict__ dY,
                      double* __restrict__ loss) {
  __shared__ double s[256];
  double a = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < 2 * M; i += gridDim.x * blockDim.x) {
    const float d = Y[i] - T[i];
    dY[i] = d / (float)M;
    a += (double)d * d;
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, s[0] / (2.0 * M));
}

// Adam (Kingma & Ba 2015, algorithm 1) with bias correction; also refreshes the FP16 copy.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       _Float16* __restrict__ p16, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[i] = pi;
  p16[i] = (_Float16)pi;
}

__global__ void k_to_half(const float* __restrict__ p, _Float16* __restrict__ h, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) h[i] = (_Float16)p[i];
}

// rows idx[r] of X (cols 5) and T (cols 2) into Xb, Tb
__global__ void k_gather_rows(const float* __restrict__ X, const float* __restrict__ T, const int* __restrict__ idx,
                              int M, float* __restrict__ Xb, float* __restrict__ Tb) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= M) return;
  const int s = idx[r];
#pragma unroll
  for (int c = 0; c < kLcfIn; ++c) Xb[(size_t)r * kLcfIn + c] = X[(size_t)s * kLcfIn + c];
  Tb[2 * r] = T[2 * s];
  Tb[2 * r + 1] = T[2 * s + 1];
}

// Network inputs of the taps of samples [s0, s0 + ns): row (s - s0) * W*W + (dy+S)*W + (dx+S)
// = (w / wnorm, fracu / os, fracv / os, dx / S, dy / S); samples not gridded give zero rows.
__global__ void k_lcf_inputs(const float4* __restrict__ sin, int s0, int ns, int S, float inv_os, float inv_wnorm,
                             float* __restrict__ X) {
  const int W = 2 * S + 1;
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (long long)ns * W * W) return;
  const int s = s0 + (int)(r / (W * W));
  const int t = (int)(r % (W * W));
  const float4 q = sin[s];
  const float invS = S > 0 ? 1.f / (float)S : 0.f;
  float* x = X + (size_t)r * kLcfIn;
  x[0] = q.x * inv_wnorm;
  x[1] = q.y * inv_os;
  x[2] = q.z * inv_os;
  x[3] = (float)(t % W - S) * invS;
  x[4] = (float)(t / W - S) * invS;
}

// network outputs (scaled by 1 / tscale) into the padded CF planes of the samples, at the
// storage precision of the table (pitch P, the plane of sample s at s * plane_elems)
__global__ void k_lcf_scatter(const float4* __restrict__ sin, const float* __restrict__ Y, int s0, int ns, int S,
                              int P, long long plane_elems, float inv_tscale, void* __restrict__ table, int prec) {
  const int W = 2 * S + 1;
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= (long long)ns * W * W) return;
  const int s = s0 + (int)(r / (W * W));
  if (sin[s].w == 0.f) return;  // not gridded: its plane is never read
  const int t = (int)(r % (W * W));
  const int dx = t % W, dy = t / W;
  const size_t e = (size_t)s * plane_elems + (size_t)(dy + kGuardV) * P + (dx + kGuardU);
  const float re = Y[2 * r] * inv_tscale, im = Y[2 * r + 1] * inv_tscale;
  if (prec == kPrecF32) {
    reinterpret_cast<float2*>(table)[e] = make_float2(re, im);
  } else {
    uint16_t hr, hi;
    if (prec == kPrecF16) {
      hr = __builtin_bit_cast(uint16_t, (_Float16)re);
      hi = __builtin_bit_cast(uint16_t, (_Float16)im);
    } else {
      const uint32_t ur = __builtin_bit_cast(uint32_t, re), ui = __builtin_bit_cast(uint32_t, im);
      hr = (uint16_t)((ur + 0x7fffu + ((ur >> 16) & 1u)) >> 16);
      hi = (uint16_t)((ui + 0x7fffu + ((ui >> 16) & 1u)) >> 16);
    }
    reinterpret_cast<uint32_t*>(table)[e] = (uint32_t)hr | ((uint32_t)hi << 16);
  }
}

}  // namespace

void mlp_gemm(const Ctx& c, bool ta, bool tb, const float* A, const void* B, bool b_half, float* C, int M, int N, int K,
              int lda, int ldb, int ldc, int epi, const float* bias, float* Zout, const float* Zin) {
  if (M <= 0 || N <= 0) return;
  // Few output tiles over a long K (the weight gradients: 128 x 128 over a batch of rows):
  // split K over workgroups (slices of kSplitK), partial tiles added into a zeroed C.
  constexpr int kSplitK = 64;
  const int tiles = ((N + kT - 1) / kT) * ((M + kT - 1) / kT);
  int splits = 1;
  if (epi == kEpiNone && tiles < c.n_cu && K > 2 * kSplitK) splits = std::min((K + kSplitK - 1) / kSplitK, 64);
  const int kslice = splits > 1 ? kSplitK * ((K + kSplitK * splits - 1) / (kSplitK * splits)) : K;
  if (splits > 1) {
    splits = (K + kslice - 1) / kslice;
    YG_CHECK(ldc == N, 1, "split-K products need a dense C");
    YG_HIP(hipMemsetAsync(C, 0, (size_t)M * N * sizeof(float), c.stream));
  }
  const dim3 g((N + kT - 1) / kT, (M + kT - 1) / kT, splits), b(256);
#define YG_GEMM(TA, TB, TT, E)                                                                                    \
  hipLaunchKernelGGL((k_gemm<TA, TB, TT, E>), g, b, 0, c.stream, A, static_cast<const TT*>(B), C, M, N, K, lda, \
                     ldb, ldc, bias, Zout, Zin, kslice)
  if (b_half) {
    YG_CHECK(!ta && !tb, 1, "FP16 weights only in the forward product");
    if (epi == kEpiBiasSilu) YG_GEMM(false, false, _Float16, kEpiBiasSilu);
    else YG_GEMM(false, false, _Float16, kEpiBias);
  } else if (!ta && !tb) {
    if (epi == kEpiBiasSilu) YG_GEMM(false, false, float, kEpiBiasSilu);
    else if (epi == kEpiBias) YG_GEMM(false, false, float, kEpiBias);
    else YG_GEMM(false, false, float, kEpiNone);
  } else if (!ta && tb) {
    if (epi == kEpiSiluGrad) YG_GEMM(false, true, float, kEpiSiluGrad);
    else YG_GEMM(false, true, float, kEpiNone);
  } else if (ta && !tb) {
    YG_GEMM(true, false, float, kEpiNone);
  } else {
    throw Error(1, "unsupported product");
  }
#undef YG_GEMM
  YG_HIP(hipGetLastError());
}

void mlp_colsum(const Ctx& c, const float* D, int M, int N, int ldd, float* db) {
  hipLaunchKernelGGL(k_colsum, dim3(N), dim3(256), 0, c.stream, D, M, N, ldd, db);
  YG_HIP(hipGetLastError());
}

void mlp_mse(const Ctx& c, const float* Y, const float* T, int M, float* dY, double* loss) {
  const int blocks = std::min((2 * M + 255) / 256, 4 * c.n_cu);
  hipLaunchKernelGGL(k_mse, dim3(std::max(blocks, 1)), dim3(256), 0, c.stream, Y, T, M, dY, loss);
  YG_HIP(hipGetLastError());
}

void mlp_adam(const Ctx& c, float* p, const float* g, float* m, float* v, void* p16, int n, float lr, float b1, float b2,
              float eps, int step) {
  const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, c.stream, p, g, m, v, static_cast<_Float16*>(p16), n,
                     lr, b1, b2, eps, c1, c2);
  YG_HIP(hipGetLastError());
}

void mlp_to_half(const Ctx& c, const float* p, void* p16, int n) {
  hipLaunchKernelGGL(k_to_half, dim3((n + 255) / 256), dim3(256), 0, c.stream, p, static_cast<_Float16*>(p16), n);
  YG_HIP(hipGetLastError());
}

void mlp_gather(const Ctx& c, const float* X, const float* T, const int* idx, int M, float* Xb, float* Tb) {
  hipLaunchKernelGGL(k_gather_rows, dim3((M + 255) / 256), dim3(256), 0, c.stream, X, T, idx, M, Xb, Tb);
  YG_HIP(hipGetLastError());
}

void lcf_inputs(const Ctx& c, const float4* sin, int s0, int ns, int S, float inv_os, float inv_wnorm, float* X) {
  const long long n = (long long)ns * (2 * S + 1) * (2 * S + 1);
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lcf_inputs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, sin, s0, ns, S, inv_os,
                     inv_wnorm, X);
  YG_HIP(hipGetLastError());
}

void lcf_scatter(const Ctx& c, const float4* sin, const float* Y, int s0, int ns, int S, int P, long long plane_elems,
                 float inv_tscale, void* table, int prec) {
  const long long n = (long long)ns * (2 * S + 1) * (2 * S + 1);
  if (n <= 0) return;
  hipLaunchKernelGGL(k_lcf_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, sin, Y, s0, ns, S, P,
                     plane_elems, inv_tscale, table, prec);
  YG_HIP(hipGetLastError());
}

}  // namespace ygrid
