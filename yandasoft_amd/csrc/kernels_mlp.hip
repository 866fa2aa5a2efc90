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
//   k_lcf_fused   inference of a width-128 network in one kernel on the FP16 matrix cores
//                 (v_mfma_f32_32x32x16_f16, FP32 accumulation): the weights of every layer sit
//                 in LDS (packed by k_lcf_pack), a wave carries 64 rows through all layers in
//                 registers, and with per-sample inputs it forms the inputs and writes the
//                 outputs into the CF planes itself.
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

// ---- Fused FP16 inference (hidden width 128) ------------------------------------------------
//
// Layer products run transposed, out^T = W^T x^T: the MFMA's A operand is a 32 x 16 slice of W^T
// (32 output neurons, 16 inputs), the B operand 16 inputs of 32 rows, and D a 32 neuron x 32 row
// block: lane l holds row l % 32 and, in register r, neuron 8 (r / 4) + 4 (l / 32) + r % 4 of its
// block of 32.  A lane's D registers are therefore the next layer's B operand for the same row
// without any data movement, provided that layer's inputs are taken in that order: K step c,
// half j of lane half h is input neuron nrn(h, 8c + j) (below), and the packed weights follow it.
// The A and B operands of one instruction agree on which k each (lane half, half) holds, so the
// instruction's own k order does not matter.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
constexpr int kFH = 128;     // hidden width of the fused form
constexpr int kFRows = 64;   // rows per wave and round (two blocks of 32)

// input neuron held by lane half h at position rho = 8c + j of the K order
__host__ __device__ __forceinline__ int fused_nrn(int h, int rho) {
  return 32 * (rho >> 4) + 8 * ((rho & 15) >> 2) + 4 * h + (rho & 3);
}

struct FusedLayout {
  int L;                       // hidden layers (dims 5, 128 x L, 2)
  long long woff[17], boff[17];  // parameter offsets of each dense layer (0 .. L)
};

// LDS image, in uint4 (8 halfs): A0 [mb 4][lane 64] (K = the 5 inputs, input 8h + j), then per
// hidden layer l = 1 .. L-1 [mb 4][c 8][lane 64], then the output layer [c 8][lane 64] (outputs 0
// and 1 in rows 0, 1 of its 32-neuron block, zero rows otherwise); then FP32 biases [L][128] and
// the output layer's 2 (padded to 4).
__host__ __device__ inline long long fused_a0() { return 0; }
__host__ __device__ inline long long fused_ah(int l) { return 256 + (long long)(l - 1) * 2048; }
__host__ __device__ inline long long fused_al(int L) { return 256 + (long long)(L - 1) * 2048; }
__host__ __device__ inline long long fused_u4(int L) { return fused_al(L) + 512 + (L * kFH + 4) / 4; }

__global__ void k_lcf_pack(FusedLayout fl, const _Float16* __restrict__ p16, const float* __restrict__ p32,
                           _Float16* __restrict__ img) {
  const int L = fl.L;
  const long long nh = (fused_al(L) + 512) * 8;  // halfs of the weight part
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nh) {
    const int j = (int)(i & 7);
    const long long u = i >> 3;  // uint4 index
    const int lane = (int)(u & 63), h = lane >> 5, m = lane & 31;
    _Float16 v = (_Float16)0.f;
    if (u < 256) {
      const int mb = (int)(u >> 6), f = 8 * h + j;
      if (f < 5) v = p16[fl.woff[0] + (long long)f * kFH + 32 * mb + m];
    } else if (u < fused_al(L)) {
      const long long q = u - 256;
      const int l = 1 + (int)(q / 2048), r = (int)(q % 2048);
      const int mb = r / 512, c = (r / 64) & 7;
      v = p16[fl.woff[l] + (long long)fused_nrn(h, 8 * c + j) * kFH + 32 * mb + m];
    } else {
      const int c = (int)((u - fused_al(L)) >> 6);
      if (m < 2) v = p16[fl.woff[L] + (long long)fused_nrn(h, 8 * c + j) * 2 + m];
    }
    img[i] = v;
  } else {
    const long long b = i - nh;  // bias float index
    float* bias = reinterpret_cast<float*>(img + nh);
    if (b < (long long)L * kFH) bias[b] = p32[fl.boff[b / kFH] + b % kFH];
    else if (b < (long long)L * kFH + 4) bias[b] = b - (long long)L * kFH < 2 ? p32[fl.boff[L] + (b - (long long)L * kFH)] : 0.f;
  }
}

// *flag = 1 when the samples' network inputs differ from the cached ones (or n differs)
__global__ void k_lcf_differ(const float4* __restrict__ a, const float4* __restrict__ b, int n, int* flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 x = a[i], y = b[i];
  if (__float_as_uint(x.x) != __float_as_uint(y.x) || __float_as_uint(x.y) != __float_as_uint(y.y) ||
      __float_as_uint(x.z) != __float_as_uint(y.z) || __float_as_uint(x.w) != __float_as_uint(y.w))
    *flag = 1;
}

__global__ void k_lcf_keep(const float4* __restrict__ a, float4* __restrict__ b, int n, int* flag, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && flag) *flag = v;
  if (i < n) b[i] = a[i];
}

struct FusedArgs {
  const uint4* pack;
  int pack_u4, L;
  long long nrows;
  const float* x;  // rows mode: nrows x 5 inputs, nrows x 2 outputs
  float* y;
  const float4* sin;  // samples mode (k_lcf_inputs / k_lcf_scatter conventions)
  int S, P, prec;
  float inv_os, inv_wnorm, inv_tscale;
  long long plane_elems;
  void* table;
  // reuse: when *same is 0 the samples equal those of the last evaluation, whose planes are at
  // src (null: already in table); the kernel then copies them instead of evaluating
  const int* same;
  const void* src;
};

__device__ __forceinline__ half8 as_half8(uint4 u) { return __builtin_bit_cast(half8, u); }

// bias + SiLU of one 32-neuron block mb of a layer's D registers (both row blocks), packed as
// K steps 2mb, 2mb + 1 of the next layer's B operands
__device__ __forceinline__ void fused_epilogue(const f32x16& d0, const f32x16& d1, const float* __restrict__ bias,
                                               int mb, int h, half8 (&nxt)[2][8]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 b4 = *reinterpret_cast<const float4*>(bias + 32 * mb + 8 * q + 4 * h);
    const float bq[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int r = 4 * q + e;
      const float z0 = d0[r] + bq[e], z1 = d1[r] + bq[e];
      nxt[0][2 * mb + (r >> 3)][r & 7] = (_Float16)(z0 * (1.f / (1.f + __expf(-z0))));
      nxt[1][2 * mb + (r >> 3)][r & 7] = (_Float16)(z1 * (1.f / (1.f + __expf(-z1))));
    }
  }
}

template <bool SAMPLES>
__global__ __launch_bounds__(512, 1) void k_lcf_fused(FusedArgs a) {
  extern __shared__ uint4 s_img[];
  if (SAMPLES && a.same && *a.same == 0) {
    if (a.src) {
      const size_t n16 = (size_t)(a.nrows / ((2 * a.S + 1) * (2 * a.S + 1))) * (size_t)a.plane_elems *
                         (a.prec == kPrecF32 ? 8 : 4) / 16;
      const uint4* src = static_cast<const uint4*>(a.src);
      uint4* dst = static_cast<uint4*>(a.table);
      for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
    }
    return;
  }
  for (int i = threadIdx.x; i < a.pack_u4; i += blockDim.x) s_img[i] = a.pack[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
  const int h = lane >> 5, n = lane & 31;
  const int L = a.L;
  const uint4* A0 = s_img + fused_a0();
  const uint4* AL = s_img + fused_al(L);
  const float* bias = reinterpret_cast<const float*>(AL + 512);
  const int W = 2 * a.S + 1, W2 = W * W;
  const float invS = a.S > 0 ? 1.f / (float)a.S : 0.f;
  const f32x16 zero = {};
  for (long long base = ((long long)blockIdx.x * nwv + wv) * kFRows; base < a.nrows;
       base += (long long)gridDim.x * nwv * kFRows) {
    half8 bop[2][8];
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      const long long r = base + 32 * rb + n;
      float x[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
      if (r < a.nrows) {
        if constexpr (SAMPLES) {
          const long long s = r / W2;
          const int t = (int)(r - s * W2);
          const float4 q = a.sin[s];
          x[0] = q.x * a.inv_wnorm;
          x[1] = q.y * a.inv_os;
          x[2] = q.z * a.inv_os;
          x[3] = (float)(t % W - a.S) * invS;
          x[4] = (float)(t / W - a.S) * invS;
        } else {
#pragma unroll
          for (int f = 0; f < 5; ++f) x[f] = a.x[r * 5 + f];
        }
      }
      half8 v = {};
      if (h == 0) {
#pragma unroll
        for (int f = 0; f < 5; ++f) v[f] = (_Float16)x[f];
      }
      bop[rb][0] = v;
    }
    // layer 0 (K = the 5 inputs), then the hidden layers, one 32-neuron block at a time so that
    // only that block's sums are live beside the layer's inputs and outputs
    {
      half8 nxt[2][8];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const half8 w = as_half8(A0[mb * 64 + lane]);
        const f32x16 d0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(w, bop[0][0], zero, 0, 0, 0);
        const f32x16 d1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(w, bop[1][0], zero, 0, 0, 0);
        fused_epilogue(d0, d1, bias, mb, h, nxt);
      }
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int c = 0; c < 8; ++c) bop[rb][c] = nxt[rb][c];
    }
    for (int l = 1; l < L; ++l) {
      const uint4* Al = s_img + fused_ah(l);
      half8 nxt[2][8];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        f32x16 d0 = zero, d1 = zero;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
          const half8 w = as_half8(Al[(mb * 8 + c) * 64 + lane]);
          d0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(w, bop[0][c], d0, 0, 0, 0);
          d1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(w, bop[1][c], d1, 0, 0, 0);
        }
        fused_epilogue(d0, d1, bias + l * kFH, mb, h, nxt);
        // keep the scheduler from hoisting the next block's weight reads over this one (they
        // would all be live at once)
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int rb = 0; rb < 2; ++rb)
#pragma unroll
        for (int c = 0; c < 8; ++c) bop[rb][c] = nxt[rb][c];
    }
    f32x16 out[2] = {zero, zero};
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const half8 w = as_half8(AL[c * 64 + lane]);
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) out[rb] = __builtin_amdgcn_mfma_f32_32x32x16_f16(w, bop[rb][c], out[rb], 0, 0, 0);
    }
    // outputs 0 and 1 of row n sit in registers 0 and 1 of the lanes of half 0
    if (h == 0) {
      const float b0 = bias[L * kFH], b1 = bias[L * kFH + 1];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        const long long r = base + 32 * rb + n;
        if (r >= a.nrows) continue;
        const float y0 = out[rb][0] + b0, y1 = out[rb][1] + b1;
        if constexpr (SAMPLES) {
          const long long s = r / W2;
          if (a.sin[s].w == 0.f) continue;  // not gridded: its plane is never read
          const int t = (int)(r - s * W2);
          const size_t e = (size_t)s * a.plane_elems + (size_t)(t / W + kGuardV) * a.P + (t % W + kGuardU);
          const float re = y0 * a.inv_tscale, im = y1 * a.inv_tscale;
          if (a.prec == kPrecF32) {
            reinterpret_cast<float2*>(a.table)[e] = make_float2(re, im);
          } else {
            uint16_t hr, hi;
            if (a.prec == kPrecF16) {
              hr = __builtin_bit_cast(uint16_t, (_Float16)re);
              hi = __builtin_bit_cast(uint16_t, (_Float16)im);
            } else {
              const uint32_t ur = __builtin_bit_cast(uint32_t, re), ui = __builtin_bit_cast(uint32_t, im);
              hr = (uint16_t)((ur + 0x7fffu + ((ur >> 16) & 1u)) >> 16);
              hi = (uint16_t)((ui + 0x7fffu + ((ui >> 16) & 1u)) >> 16);
            }
            reinterpret_cast<uint32_t*>(a.table)[e] = (uint32_t)hr | ((uint32_t)hi << 16);
          }
        } else {
          a.y[2 * r] = y0;
          a.y[2 * r + 1] = y1;
        }
      }
    }
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

bool lcf_fused_supported(int hidden, int layers) { return hidden == kFH && layers >= 1 && fused_u4(layers) * 16 <= 160 * 1024; }

size_t lcf_fused_pack_bytes(int layers) { return (size_t)fused_u4(layers) * 16; }

void lcf_fused_pack(const Ctx& c, int layers, const long long* woff, const long long* boff, const void* p16,
                    const float* p32, void* img) {
  FusedLayout fl{};
  fl.L = layers;
  for (int l = 0; l <= layers; ++l) {
    fl.woff[l] = woff[l];
    fl.boff[l] = boff[l];
  }
  const long long n = (fused_al(layers) + 512) * 8 + (long long)layers * kFH + 4;
  hipLaunchKernelGGL(k_lcf_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, fl,
                     static_cast<const _Float16*>(p16), p32, static_cast<_Float16*>(img));
  YG_HIP(hipGetLastError());
}

static void launch_fused(const Ctx& c, const FusedArgs& a, bool samples) {
  if (a.nrows <= 0) return;
  const size_t lds = (size_t)a.pack_u4 * 16;
  static bool attr[2] = {false, false};
  const void* fn = samples ? reinterpret_cast<const void*>(&k_lcf_fused<true>)
                           : reinterpret_cast<const void*>(&k_lcf_fused<false>);
  if (!attr[samples]) {
    YG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr[samples] = true;
  }
  const long long waves = (a.nrows + kFRows - 1) / kFRows;
  const int blocks = (int)std::max<long long>(1, std::min<long long>(c.n_cu, (waves + 7) / 8));
  if (samples)
    hipLaunchKernelGGL(k_lcf_fused<true>, dim3(blocks), dim3(512), lds, c.stream, a);
  else
    hipLaunchKernelGGL(k_lcf_fused<false>, dim3(blocks), dim3(512), lds, c.stream, a);
  YG_HIP(hipGetLastError());
}

void lcf_fused_rows(const Ctx& c, const void* img, int layers, const float* x, float* y, long long n) {
  FusedArgs a{};
  a.pack = static_cast<const uint4*>(img);
  a.pack_u4 = (int)fused_u4(layers);
  a.L = layers;
  a.nrows = n;
  a.x = x;
  a.y = y;
  launch_fused(c, a, false);
}

void lcf_differ(const Ctx& c, const float4* a, const float4* b, int n, int* flag, bool compare) {
  // flag = !compare, then (compare) set where the inputs differ
  hipLaunchKernelGGL(k_lcf_keep, dim3(1), dim3(64), 0, c.stream, a, (float4*)nullptr, 0, flag, compare ? 0 : 1);
  if (compare && n > 0)
    hipLaunchKernelGGL(k_lcf_differ, dim3((n + 255) / 256), dim3(256), 0, c.stream, a, b, n, flag);
  YG_HIP(hipGetLastError());
}

void lcf_keep(const Ctx& c, const float4* a, float4* b, int n) {
  if (n > 0) hipLaunchKernelGGL(k_lcf_keep, dim3((n + 255) / 256), dim3(256), 0, c.stream, a, b, n, (int*)nullptr, 0);
  YG_HIP(hipGetLastError());
}

void lcf_fused_samples(const Ctx& c, const void* img, int layers, const float4* sin, int ns, int S, float inv_os,
                       float inv_wnorm, int P, long long plane_elems, float inv_tscale, void* table, int prec,
                       const int* same, const void* src) {
  FusedArgs a{};
  a.same = same;
  a.src = src;
  a.pack = static_cast<const uint4*>(img);
  a.pack_u4 = (int)fused_u4(layers);
  a.L = layers;
  a.nrows = (long long)ns * (2 * S + 1) * (2 * S + 1);
  a.sin = sin;
  a.S = S;
  a.P = P;
  a.prec = prec;
  a.inv_os = inv_os;
  a.inv_wnorm = inv_wnorm;
  a.inv_tscale = inv_tscale;
  a.plane_elems = plane_elems;
  a.table = table;
  launch_fused(c, a, true);
}

}  // namespace ygrid
