// Small-support kernels (CF support S <= kSmallSupportMax): the same passes as the 64x64 tile
// kernels (GridKernel::grid / ::degrid, GridKernel.cc:61-83,106-125, inside
// TableVisGridder::generic), shaped for footprints of at most (2S+1)^2 = 289 taps, where the tile
// kernels spend most of a pair's work on the empty parts of their 4x32-pixel blocks.
//
//   k_grid_small    persistent workgroups take work units: one 8-row band of the 64x64 tile of
//                   a work item of the static binning made with one item per key, so that a
//                   unit owns its band of the (image channel, grid, tile) region exclusively (the
//                   pass bins with one CF-plane group, or by w grid when stacking).  The band of
//                   every polarisation is accumulated in LDS (float LDS atomics, no per-tap HBM
//                   atomics) by groups of L lanes, one group per pair whose footprint reaches the
//                   band (the pair descriptor's block mask says which) and one lane per footprint
//                   column over the footprint's rows inside the band; the band is then added into
//                   the grid with plain loads and stores, only where the unit added something.
//   k_degrid_small  no binning: one group of L lanes per sample, in row-major (channel-fastest)
//                   order so that the channels of one row, whose footprints lie a few pixels
//                   apart, are read together; each lane sums its column over the footprint's
//                   rows and the group's sums meet by shuffles.  The per-sample result goes to the
//                   same accumulator as the tile kernels' (k_degrid_finish applies phasor and
//                   weight).
//
// L (lanes per sample) is the power of two >= 2S+1: 8 for S <= 3, 16 for S <= 7, 32 for S <= 15.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "device_util.h"
#include "ygrid_internal.h"

namespace ygrid {

namespace {

// complex tap e (element offset in the table) of an FP32 or 16-bit table
template <int CFP>
__device__ __forceinline__ float2 load_tap(const void* __restrict__ cf, long long e) {
  if constexpr (CFP == kPrecF32) {
    return reinterpret_cast<const float2*>(cf)[e];
  } else {
    const uint32_t w = reinterpret_cast<const uint32_t*>(cf)[e];
    return make_float2(widen16(w & 0xffffu, CFP), widen16(w >> 16, CFP));
  }
}

// the four MAC coefficients (a, b, c, d) of polarisation q of sample s: re += a*C.re + b*C.im,
// im += c*C.re + d*C.im (the layout the prep kernel writes)
// (i0 and i1 are even: two 8-byte, or two 4-byte, loads)
__device__ __forceinline__ float4 pol_coef(const Samples& smp, int s, int q) {
  const int ore = 2 * (q & 1), mre = 2 * (q >> 1);
  const int i0 = 4 * ore + mre, i1 = 4 * (ore + 1) + mre;
  if (smp.coef_prec == kPrecF32) {
    const float* c = smp.coef + (size_t)s * kCoef;
    const float2 a = *reinterpret_cast<const float2*>(c + i0), b = *reinterpret_cast<const float2*>(c + i1);
    return make_float4(a.x, a.y, b.x, b.y);
  }
  const uint16_t* c = smp.coef16 + (size_t)s * kCoef;
  const uint32_t a = *reinterpret_cast<const uint32_t*>(c + i0), b = *reinterpret_cast<const uint32_t*>(c + i1);
  return make_float4(widen16(a & 0xffffu, smp.coef_prec), widen16(a >> 16, smp.coef_prec),
                     widen16(b & 0xffffu, smp.coef_prec), widen16(b >> 16, smp.coef_prec));
}

// Rows of the band a workgroup of k_grid_small owns (BR: 4, 8 or 16, the context's
// small_band_rows, ygrid_ctx_set_small_band_rows; at least the rows one instruction covers):
// work unit u is band u % (64 / BR) of item u / (64 / BR), so that the bands of a tile are worked
// at the same time by several workgroups (a tile's pairs are the samples of ~64x64 pixels: one
// workgroup per tile would leave most of the chip idle where the samples fill few tiles).  A
// taller band stages fewer pairs twice but holds 4 x NP x BR x 512 bytes of LDS per workgroup.
constexpr int kGridSmallBS = 256;

// descriptors a workgroup of k_grid_small stages per round (two per thread)
constexpr int kGridSmallRound = 2 * kGridSmallBS;
constexpr int kGridSmallWaves = kGridSmallBS / 64;
// pairs a wave of k_grid_small loads before it adds them (taps and coefficients in flight;
// deeper batches were measured slower: the registers cost occupancy)
constexpr int kGridSmallBatch = 4;

// L: columns per lane row (the power of two >= 2S+1); a wave instruction covers 64/L band rows
// of one pair's footprint, lane = (row, column)
template <int NP, int CFP, int L, int BR>
__global__ __launch_bounds__(kGridSmallBS) void k_grid_small(const int4* __restrict__ items,
                                                            const int* __restrict__ nitems_p, int* __restrict__ counter,
                                                            SmallGridArgs a) {
  constexpr int kBands = kTile / BR;
  constexpr int BAND = BR * kTile;       // pixels of a band
  constexpr int RP = 64 / L;             // band rows per instruction
  constexpr int NSTEP = BR / RP;         // instructions per pair and band
  static_assert(NSTEP >= 1, "a band holds at least the rows of one instruction");
  constexpr int U = kGridSmallBatch;
  // each wave's own copy of the unit's band, every polarisation: [wave][pol][row][column]
  // (re, im).  A wave adds one pair at a time, its lanes on distinct pixels, so plain LDS reads
  // and writes suffice (a wave's LDS accesses execute in order; float LDS atomics are much
  // slower); the copies meet when the band is flushed.
  __shared__ float2 s_acc[kGridSmallWaves * NP * BAND];
  // a round's pairs that reach the band: sample index and header
  __shared__ int s_smp[kGridSmallRound];
  __shared__ SampleHdr s_hdr[kGridSmallRound];
  __shared__ int s_n;
  __shared__ int s_item;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6;
  const int lr = lane / L, lc = lane % L;  // the lane's row (of RP) and column
  float2* __restrict__ my = s_acc + w * NP * BAND;
  const int nunits = *nitems_p * kBands;
  const size_t plane = (size_t)a.nx * a.ny;
  int queue = blockIdx.x % kItemQueues, queues_left = kItemQueues;
  for (;;) {
    const int unit = next_item(counter, nunits, queue, queues_left, &s_item);
    if (unit < 0) break;
    const int band = unit % kBands;
    const int4 it = items[unit / kBands];
    const int tile = it.x % a.ntiles;
    const int kz = it.x / a.ntiles;
    const int ichan = kz / a.nzg;
    const int u0 = (tile % a.ntu) * kTile;
    const int vb = (tile / a.ntu) * kTile + band * BR;  // first row of the band
    // the pair descriptor's block bits 8w + 2k + j (blocks of 4 rows x 32 columns, four per
    // 16-row wave band w: bits 2 (row / 4) and 2 (row / 4) + 1) that cover this band's rows
    const uint32_t mshift = (uint32_t)(band * BR / 2);
    constexpr uint32_t kMask = BR >= 16 ? 0xffu : ((1u << (BR / 2)) - 1u);
    for (int i = tid; i < kGridSmallWaves * NP * BAND; i += kGridSmallBS) s_acc[i] = make_float2(0.f, 0.f);
    for (int base = it.y; base < it.z; base += kGridSmallRound) {
      // stage: the round's descriptors, kept where the footprint reaches the band (compacted
      // per wave: one LDS atomic reserves a wave's slots), with their sample headers
      if (tid == 0) s_n = 0;
      __syncthreads();
      // (the two descriptors per thread in plain registers: an array filled under conditions
      // would live in scratch)
      const int pa = base + tid, pb = base + kGridSmallBS + tid;
      const int4 da = pa < it.z ? a.pairs[pa] : make_int4(0, 0, 0, 0);
      const int4 db = pb < it.z ? a.pairs[pb] : make_int4(0, 0, 0, 0);
      const bool ka = pa < it.z && (((uint32_t)da.z >> mshift) & kMask) != 0u;
      const bool kb = pb < it.z && (((uint32_t)db.z >> mshift) & kMask) != 0u;
      const int4* hdr4 = reinterpret_cast<const int4*>(a.smp.hdr);
      const int4 ha = ka ? hdr4[da.x] : make_int4(0, 0, 0, 0);
      const int4 hb = kb ? hdr4[db.x] : make_int4(0, 0, 0, 0);
      const unsigned long long bala = __ballot(ka), balb = __ballot(kb);
      int wbase = 0;
      if (lane == 0 && (bala | balb)) wbase = atomicAdd(&s_n, __popcll(bala) + __popcll(balb));
      wbase = __shfl(wbase, 0, 64);
      const int ba = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bala >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bala, 0u));
      const int bb = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(balb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)balb, 0u));
      int4* shdr4 = reinterpret_cast<int4*>(s_hdr);
      if (ka) {
        s_smp[wbase + ba] = da.x;
        shdr4[wbase + ba] = ha;
      }
      if (kb) {
        s_smp[wbase + __popcll(bala) + bb] = db.x;
        shdr4[wbase + __popcll(bala) + bb] = hb;
      }
      __syncthreads();
      const int n = s_n;
      // wave w adds a contiguous block of the list (neighbouring entries are neighbouring
      // channels of one row), U pairs at a time: their taps and coefficients are loaded first
      const int per = (n + kGridSmallWaves - 1) / kGridSmallWaves;
      const int i0 = w * per, i1 = min(n, i0 + per);
      for (int i = i0; i < i1; i += U) {
        float2 t[U][NSTEP];
        float4 k[U][NP];
        int px[U][NSTEP];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const bool valid = i + u < i1;
          const int e = valid ? i + u : i0;
          const int s = s_smp[e];
          const SampleHdr h = s_hdr[e];
          const int S = h.spack & 0xffff, P = h.spack >> 16, W = 2 * S + 1;
          const int du = h.iu0 + lc - u0;
          const bool col = valid && lc < W && du >= 0 && du < kTile;
          const long long e0 = (long long)h.cfoff + (long long)kGuardV * P + kGuardU + lc;
#pragma unroll
          for (int st = 0; st < NSTEP; ++st) {
            const int br = st * RP + lr;          // band row
            const int r = vb + br - h.iv0;        // footprint row
            const bool on = col && r >= 0 && r < W;
            px[u][st] = on ? br * kTile + du : -1;
            t[u][st] = on ? load_tap<CFP>(a.cf, e0 + (long long)r * P) : make_float2(0.f, 0.f);
          }
#pragma unroll
          for (int q = 0; q < NP; ++q) k[u][q] = valid ? pol_coef(a.smp, s, q) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int st = 0; st < NSTEP; ++st)
            if (px[u][st] >= 0) {
#pragma unroll
              for (int q = 0; q < NP; ++q) {
                float2 g = my[q * BAND + px[u][st]];
                g.x += k[u][q].x * t[u][st].x + k[u][q].y * t[u][st].y;
                g.y += k[u][q].z * t[u][st].x + k[u][q].w * t[u][st].y;
                my[q * BAND + px[u][st]] = g;
              }
            }
      }
      __syncthreads();  // the round's slots are read before the next round stages
    }
    __syncthreads();  // every add of the unit is in (also when the item has no round)
    // the unit owns its band of the tile: add the waves' copies into the grid where they hold
    // something
    float2* __restrict__ g = reinterpret_cast<float2*>(a.grid) + a.wstride * (size_t)(kz % a.nzg);
    for (int i = tid; i < NP * BAND; i += kGridSmallBS) {
      float2 v = s_acc[i];
#pragma unroll
      for (int ww = 1; ww < kGridSmallWaves; ++ww) {
        const float2 x = s_acc[ww * NP * BAND + i];
        v.x += x.x;
        v.y += x.y;
      }
      const int q = i / BAND, px = i % BAND;
      const int u = u0 + px % kTile, vv = vb + px / kTile;
      if ((v.x != 0.f || v.y != 0.f) && u < a.nx && vv < a.ny) {
        float2* gp = g + (size_t)(ichan * NP + q) * plane + (size_t)vv * a.nx + u;
        float2 x = *gp;
        x.x += v.x;
        x.y += v.y;
        *gp = x;
      }
    }
    // next_item's barrier keeps the band from being cleared while another wave still reads it
  }
}

template <int NP, int CFP, int L>
__global__ __launch_bounds__(256) void k_degrid_small(SmallDegridArgs a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long gi = t / L;
  const int c = (int)(t % L);
  const long long n = (long long)a.nrow * a.nchan;
  const size_t plane = (size_t)a.nx * a.ny;
  float sr[NP], si[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) sr[q] = si[q] = 0.f;
  int s = -1;
  float gs = 1.f, igs = 1.f;
  if (CFP != kPrecF32 && a.round16 && a.gscale) {
    gs = *a.gscale;
    igs = 1.f / gs;  // a power of two: exact
  }
  if (gi < n) {
    const int row = (int)(gi / a.nchan), chan = (int)(gi % a.nchan);
    s = row + a.nrow * chan;
    const SampleHdr h = a.smp.hdr[s];
    if (h.iu0 == INT_MIN) {
      s = -1;
    } else {
      const int S = h.spack & 0xffff, P = h.spack >> 16, W = 2 * S + 1;
      if (c < W) {
        const int kz = a.smp.key[s];
        const float2* __restrict__ gp = reinterpret_cast<const float2*>(a.grid) + a.wstride * (size_t)(kz % a.nzg) +
                                        (size_t)(kz / a.nzg) * NP * plane + (size_t)h.iv0 * a.nx + h.iu0 + c;
        const long long e0 = (long long)h.cfoff + (long long)kGuardV * P + kGuardU + c;
        for (int r = 0; r < W; ++r) {
          const float2 C = load_tap<CFP>(a.cf, e0 + (long long)r * P);
#pragma unroll
          for (int q = 0; q < NP; ++q) {
            float2 G = gp[(size_t)q * plane + (size_t)r * a.nx];
            if (CFP != kPrecF32 && a.round16)
              G = make_float2(widen16(round16(G.x * gs, CFP), CFP) * igs, widen16(round16(G.y * gs, CFP), CFP) * igs);
            // C * conj(G) (GridKernel.cc:121-122)
            sr[q] += C.x * G.x + C.y * G.y;
            si[q] += C.y * G.x - C.x * G.y;
          }
        }
      }
    }
  }
  // the group's lanes are consecutive and L-aligned: butterfly sums within the group
#pragma unroll
  for (int q = 0; q < NP; ++q)
#pragma unroll
    for (int o = L / 2; o > 0; o >>= 1) {
      sr[q] += __shfl_xor(sr[q], o, 64);
      si[q] += __shfl_xor(si[q], o, 64);
    }
  // (the visibility cube is written by k_degrid_finish, one thread per sample in row order:
  // coalesced, where this kernel's channel-fastest order would scatter the writes)
  if (s >= 0 && c == 0) {
#pragma unroll
    for (int q = 0; q < NP; ++q)
      *reinterpret_cast<float2*>(a.acc + (size_t)s * 8 + 2 * q) = make_float2(sr[q], si[q]);
  }
}

template <int NP, int CFP, class F>
void with_lanes(int max_support, F&& f) {
  if (max_support <= 3) f(std::integral_constant<int, 8>{});
  else if (max_support <= 7) f(std::integral_constant<int, 16>{});
  else f(std::integral_constant<int, 32>{});
}

// band rows: the requested 4, 8 or 16 (0: 4 where one instruction covers 4 rows or fewer, as
// measured faster for supports 4 to 8, else 8), raised to the rows one instruction of L lanes
// covers
template <int L, class F>
void with_band_rows(int rows, F&& f) {
  const int r = std::max(rows > 0 ? rows : (64 / L <= 4 ? 4 : 8), 64 / L);
  if (r <= 4) f(std::integral_constant<int, (64 / L <= 4 ? 4 : 64 / L)>{});
  else if (r <= 8) f(std::integral_constant<int, 8>{});
  else f(std::integral_constant<int, 16>{});
}

template <int NP, class F>
void with_cf_prec(int prec, F&& f) {
  switch (prec) {
    case kPrecF32: f(std::integral_constant<int, kPrecF32>{}); break;
    case kPrecF16: f(std::integral_constant<int, kPrecF16>{}); break;
    case kPrecBF16: f(std::integral_constant<int, kPrecBF16>{}); break;
    default: throw Error(1, "unknown CF precision");
  }
}

template <class F>
void with_npol(int npol, F&& f) {
  switch (npol) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    default: throw Error(1, "1 to 4 polarisations per grid");
  }
}


}  // namespace

void launch_grid_small(const Ctx& c, const int4* items, const int* nitems, int* counter, int nitems_cap,
                       const SmallGridArgs& a, int npol, int prec, int max_support) {
  YG_CHECK(max_support <= kSmallSupportMax, 2, "support too large for the small-support kernels");
  with_npol(npol, [&](auto np) {
    constexpr int NP = decltype(np)::value;
    with_cf_prec<NP>(prec, [&](auto cp) {
      constexpr int CFP = decltype(cp)::value;
      with_lanes<NP, CFP>(max_support, [&](auto l) {
        constexpr int L = decltype(l)::value;
        with_band_rows<L>(c.small_band_rows, [&](auto br) {
          constexpr int BR = decltype(br)::value;
          // up to eight workgroups of 256 threads per CU (as the band's LDS allows)
          const long long units = (long long)nitems_cap * (kTile / BR);
          const int g = (int)std::max(1ll, std::min(units, 8ll * c.n_cu));
          hipLaunchKernelGGL((k_grid_small<NP, CFP, L, BR>), dim3(g), dim3(kGridSmallBS), 0, c.stream, items, nitems,
                             counter, a);
        });
      });
    });
  });
  YG_HIP(hipGetLastError());
}

void launch_degrid_small(const Ctx& c, const SmallDegridArgs& a, int npol, int prec, int max_support) {
  YG_CHECK(max_support <= kSmallSupportMax, 2, "support too large for the small-support kernels");
  const long long n = (long long)a.nrow * a.nchan;
  if (n == 0) return;
  with_npol(npol, [&](auto np) {
    constexpr int NP = decltype(np)::value;
    with_cf_prec<NP>(prec, [&](auto cp) {
      constexpr int CFP = decltype(cp)::value;
      with_lanes<NP, CFP>(max_support, [&](auto l) {
        constexpr int L = decltype(l)::value;
        const long long threads = n * L;
        YG_CHECK(threads / 256 < (long long)INT_MAX, 1, "too many samples for one launch");
        hipLaunchKernelGGL((k_degrid_small<NP, CFP, L>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                           c.stream, a);
      });
    });
  });
  YG_HIP(hipGetLastError());
}

}  // namespace ygrid
