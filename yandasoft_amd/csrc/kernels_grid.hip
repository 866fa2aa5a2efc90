// Grid / degrid kernels for gfx950 (CDNA4, wave64).
//
// Pass structure (one call of TableVisGridder::generic, TableVisGridder.cc:451-884):
//   k_prep        one thread per (row, chan): uvw -> (iu, fracu, iv, fracv) in FP64, CF plane,
//                 on-grid test, flags, noise weight, sum of weights; writes a SampleRec.
//   k_bin_*       counting sort of (sample, grid tile) pairs by (image chan, CF plane group, tile);
//                 the key order is plane-major so that concurrently running workgroups read the
//                 same few CF planes, which then stay resident in each XCD's L2.
//   k_grid        persistent workgroups take work items (<= kChunk pairs of one key) in key order.
//                 A workgroup owns a 64x64 tile of all polarisations in registers (16 pixels per
//                 lane), spreads every pair into it, and flushes once with float atomics: no per-tap
//                 HBM atomics (GridKernel::grid, GridKernel.cc:61-83).
//   k_degrid      same work items; the tile of the model grid is loaded into registers, each pair
//                 is a wave-level dot product over the CF taps (GridKernel::degrid, GridKernel.cc:106-125),
//                 partial sums of the tiles a footprint covers meet in per-sample float accumulators.
//   k_degrid_finish  phasor x weight and the add into the visibility cube (TableVisGridder.cc:767-773,861).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <climits>

#include "ygrid_internal.h"

namespace ygrid {

namespace {

constexpr int kChunk = 512;  // max pairs per work item

__device__ __forceinline__ int nint_d(double x) { return (x > 0.0) ? (int)(x + 0.5) : (int)(x - 0.5); }

// WDependentGridderBase::getWPlane (WDependentGridderBase.cc:118-150): -1 out of range,
// -2 when the gaussian sampling cannot be inverted.
__device__ int w_plane(const CIndexDesc& ci, double w) {
  int r;
  if (ci.sampling != 0 && ci.nw > 1) {
    const int half = (ci.nw - 1) / 2;
    const double x = w / ci.wmax;
    double idx = 0.0;
    if (ci.sampling == 1) {  // PowerWSampling::index (PowerWSampling.cc:76-84)
      if (x > 0) idx = pow(x, 1. / ci.exponent);
      else if (x < 0) idx = -pow(-x, 1. / ci.exponent);
    } else {  // GaussianWSampling::index (GaussianWSampling.cc:91-110)
      if (x == -1. || x == 1.) {
        idx = x;
      } else {
        const double e = 1. - fabs(x) / ci.g_amp;
        if (!(e > 0.)) return -2;
        const double ps = -log(e) * ci.g_ts;
        idx = (x < 0) ? -sqrt(ps) : ((x > 0) ? sqrt(ps) : 0.0);
      }
    }
    r = half + nint_d(idx * half);
  } else {
    r = ci.nw > 1 ? (ci.nw - 1) / 2 + nint_d(w / ci.wscale) : 0;
  }
  return (r < ci.nw && r > -1) ? r : -1;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ float wave_sum_f32(float v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// --------------------------------------------------------------------------------------
// prep
// --------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_prep(PrepParams p, PrepInputs in, PrepOutputs out) {
#pragma clang fp contract(off)
  const int n = p.nrow * p.nchan;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;  // s = row + nrow*chan
  const bool inb = s < n;
  const int row = inb ? s % p.nrow : 0;
  const int chan = inb ? s / p.nrow : 0;
  const bool grid_mode = p.mode != kModeDegrid;

  SampleRec rec;
  rec.iu0 = INT_MIN;
  rec.iv0 = 0;
  rec.cfoff = 0;
  rec.spack = 0;
  rec.key = 0;
  rec.conj = 0;
  rec.pad0 = rec.pad1 = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) rec.v[q] = 0.f;

  bool valid = inb;
  bool flagged = false, wflagged = false;
  int err = 0;
  int swrow = 0, ichan = 0;
  float wts[4] = {0.f, 0.f, 0.f, 0.f};
  unsigned long long taps = 0;

  if (valid) {
    const double f = in.freq[chan];
    const double os = (double)p.oversample;
    // TableVisGridder.cc:616-649
    const double uScaled = f * in.uvw[3 * row + 0] / (kC * p.uvcell_u);
    int iu = nint_d(uScaled);
    int fracu = nint_d(os * ((double)iu - uScaled));
    if (fracu < 0) { iu += 1; fracu += p.oversample; }
    else if (fracu >= p.oversample) { iu -= 1; fracu -= p.oversample; }
    iu += p.nx / 2;
    const double vScaled = f * in.uvw[3 * row + 1] / (kC * p.uvcell_v);
    int iv = nint_d(vScaled);
    int fracv = nint_d(os * ((double)iv - vScaled));
    if (fracv < 0) { iv += 1; fracv += p.oversample; }
    else if (fracv >= p.oversample) { iv -= 1; fracv -= p.oversample; }
    iv += p.ny / 2;
    // delay phasor (TableVisGridder.cc:652-655)
    const double phase = 2.0f * kPi * f * (in.delay ? in.delay[row] : 0.0) / kC;
    const float ph_r = (float)cos(phase), ph_i = (float)sin(phase);

    for (int pol = 0; pol < p.npol; ++pol)
      if (in.flag && in.flag[row + (size_t)p.nrow * (chan + (size_t)p.nchan * pol)]) { flagged = true; break; }
    ichan = in.chanmap ? in.chanmap[chan] : 0;
    if (ichan < 0) flagged = true;
    if (flagged) {
      valid = false;
    } else {
      // cIndex
      int plane = 0;
      if (p.ci.kind != 0) {
        const double w = in.uvw[3 * row + 2] / kC;
        const int iw = w_plane(p.ci, w * f);
        if (iw < 0) {
          if (!p.ci.wmaxclip || iw == -2) err |= 2;
          plane = -1;
        } else if (p.ci.kind == 2) {
          const int feed = in.feed1 ? in.feed1[row] : 0;
          const int c = p.ci.freqdep ? chan : 0;
          plane = iw + p.ci.nw * (c + p.ci.nchan_cf * (feed + p.ci.maxfeeds * p.ci.field));
        } else {
          plane = iw;
        }
      }
      if (plane < 0) {
        wflagged = true;
        valid = false;
      } else {
        const int cInd = fracu + p.oversample * (fracv + p.oversample * plane);
        if (cInd >= in.n_cf) { err |= 4; valid = false; }
        else {
          const int S = in.cf_supp[cInd];
          const int iuo = iu + (in.cf_ofs_u ? in.cf_ofs_u[plane] : 0);
          const int ivo = iv + (in.cf_ofs_v ? in.cf_ofs_v[plane] : 0);
          // strict on-grid test (TableVisGridder.cc:763-764)
          if (!((iuo - S) > 0 && (ivo - S) > 0 && (iuo + S) < p.nx && (ivo + S) < p.ny)) {
            valid = false;
          } else {
            rec.iu0 = iuo - S;
            rec.iv0 = ivo - S;
            rec.cfoff = in.cf_off[cInd];
            rec.spack = S | ((2 * S + 1 + 2 * kGuardU) << 16);
            rec.key = ichan * p.nzg + plane / p.zdiv;
            swrow = p.track_per_os ? cInd : plane;
            const float vw = in.visweight ? in.visweight[chan] : 1.0f;
            if (grid_mode) {
              if (p.mode == kModePSF && p.psf_feed >= 0) {
                if ((in.feed1 ? in.feed1[row] : 0) != p.psf_feed) valid = false;
              }
              if (p.mode == kModePCF) {
                rec.conj = (ivo < p.ny / 2 && iuo >= p.nx / 2) || (ivo <= p.ny / 2 && iuo < p.nx / 2);
              }
              for (int pol = 0; pol < p.npol && valid; ++pol) {
                const size_t ci = row + (size_t)p.nrow * (chan + (size_t)p.nchan * pol);
                const float nre = in.noise ? in.noise[ci].x : 1.0f;
                const float visNoise = nre * nre;
                const float wt = (visNoise > 0.f) ? (float)(1. / (double)visNoise) : 0.f;
                if (!(wt > 0.f)) { err |= 1; valid = false; break; }
                wts[pol] = wt;
                if (p.mode == kModeGrid) {
                  const float2 v = in.vis[ci];
                  // rVis = phasor * conj(vis) * wt (* visweight)  (TableVisGridder.cc:797-800)
                  float rr = (ph_r * v.x + ph_i * v.y) * wt;
                  float ri = (ph_i * v.x - ph_r * v.y) * wt;
                  rr *= vw;
                  ri *= vw;
                  rec.v[2 * pol] = rr;
                  rec.v[2 * pol + 1] = ri;
                } else if (p.mode == kModePSF) {
                  rec.v[2 * pol] = wt * vw;
                } else {
                  rec.v[2 * pol] = wt;
                }
              }
              if (!valid) rec.iu0 = INT_MIN;
            } else {
              rec.v[0] = ph_r * vw;
              rec.v[1] = ph_i * vw;
            }
            if (valid) taps = (unsigned long long)(2 * S + 1) * (2 * S + 1) * p.npol;
          }
        }
      }
    }
  }
  if (inb) out.recs[s] = rec;

  // counters (wave aggregated)
  const unsigned long long nvalid = __popcll(__ballot(valid));
  const unsigned long long nflag = __popcll(__ballot(flagged && grid_mode));
  const unsigned long long nwflag = __popcll(__ballot(wflagged));
  const unsigned long long tsum = wave_sum_u64(taps);
  const int lane = threadIdx.x & 63;
  if (lane == 0) {
    if (nvalid) atomicAdd(&out.counters[0], nvalid * (unsigned long long)p.npol);
    if (tsum) atomicAdd(&out.counters[1], tsum);
    if (nflag) atomicAdd(&out.counters[2], nflag);
    if (nwflag) atomicAdd(&out.counters[3], nwflag * (grid_mode ? 2ull : 1ull));
  }
  if (err) atomicOr(out.error, err);

  // sum of weights, aggregated per distinct (row of sumw, image chan) within the wave
  if (grid_mode && p.mode != kModePCF) {
    long long key = valid ? (long long)swrow + (long long)p.nsw_rows * (long long)p.npol * ichan : -1;
    bool done = !valid;
    while (true) {
      const unsigned long long active = __ballot(!done);
      if (active == 0) break;
      const int leader = __ffsll((long long)active) - 1;
      const long long lkey = __shfl(key, leader, 64);
      const bool match = !done && key == lkey;
      for (int pol = 0; pol < p.npol; ++pol) {
        const double t = wave_sum_f64(match ? (double)wts[pol] : 0.0);
        if (lane == leader) atomicAdd(&out.sumw[lkey + (long long)p.nsw_rows * pol], t);
      }
      if (match) done = true;
    }
  }
}

// --------------------------------------------------------------------------------------
// binning
// --------------------------------------------------------------------------------------
struct TileRange {
  int tu0, tu1, tv0, tv1;
};
__device__ __forceinline__ TileRange tiles_of(const SampleRec& r) {
  const int W = 2 * (r.spack & 0xffff) + 1;
  return {r.iu0 / kTile, (r.iu0 + W - 1) / kTile, r.iv0 / kTile, (r.iv0 + W - 1) / kTile};
}

__global__ void k_bin_count(const SampleRec* __restrict__ recs, int n, int ntu, int ntiles, int* counts) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const SampleRec r = recs[s];
  if (r.iu0 == INT_MIN) return;
  const TileRange t = tiles_of(r);
  const long long base = (long long)r.key * ntiles;
  for (int tv = t.tv0; tv <= t.tv1; ++tv)
    for (int tu = t.tu0; tu <= t.tu1; ++tu) atomicAdd(&counts[base + tv * ntu + tu], 1);
}

__global__ void k_bin_fill(const SampleRec* __restrict__ recs, int n, int ntu, int ntiles, int* cursor, int* pairs) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const SampleRec r = recs[s];
  if (r.iu0 == INT_MIN) return;
  const TileRange t = tiles_of(r);
  const long long base = (long long)r.key * ntiles;
  for (int tv = t.tv0; tv <= t.tv1; ++tv)
    for (int tu = t.tu0; tu <= t.tu1; ++tu) {
      const int pos = atomicAdd(&cursor[base + tv * ntu + tu], 1);
      pairs[pos] = s;
    }
}

__global__ void k_item_count(const int* __restrict__ counts, int nkeys, int* icount) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nkeys) return;
  icount[k] = (k < nkeys) ? (counts[k] + kChunk - 1) / kChunk : 0;
}

__global__ void k_item_emit(const int* __restrict__ counts, const int* __restrict__ offsets,
                            const int* __restrict__ ioffsets, int nkeys, int4* items) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  const int c = counts[k];
  if (c == 0) return;
  const int o = offsets[k];
  const int io = ioffsets[k];
  const int nch = (c + kChunk - 1) / kChunk;
  for (int j = 0; j < nch; ++j) {
    const int b = o + j * kChunk;
    items[io + j] = make_int4(k, b, min(b + kChunk, o + c), 0);
  }
}

// --------------------------------------------------------------------------------------
// grid: lane layout in a 64x64 tile.  Wave w owns rows [16w, 16w+16); lane (lu, lv) =
// (lane & 15, lane >> 4) owns pixels u = 16j + lu, v = 16w + 4k + lv for j, k in 0..3, i.e.
// 16 "blocks" (k, j) of 16x4 pixels per wave.  A (sample, tile) pair touches the blocks its
// footprint overlaps; all of a pair's CF loads are issued before its FMAs so that one L2
// latency is paid per pair rather than per block.
// --------------------------------------------------------------------------------------

// Complex MAC coefficients of one sample and polarisation: acc_re += a*c.x + b*c.y and
// acc_im += cc*c.x + d*c.y.  Plain: (v_r, -v_i, v_i, v_r) = v*C; PCF half plane: the same
// with conj(C) (TableVisGridder.cc:837-844).
struct MacCoef {
  float a, b, c, d;
};

__device__ __forceinline__ bool block_active(int dvb, int dub, int S2) {
  return !(dvb + 3 < 0 || dvb > S2 || dub + 15 < 0 || dub > S2);
}

template <int NP, int WPE>
__global__ __launch_bounds__(256, WPE) void k_grid(const int4* __restrict__ items, const int* __restrict__ nitems_p,
                                              int* __restrict__ counter, const int* __restrict__ pairs,
                                              const SampleRec* __restrict__ recs, const float2* __restrict__ cf,
                                              float* __restrict__ grid, int nx, int ny, int ntu, int ntiles, int nzg) {
  __shared__ int s_item;
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int lu = lane & 15, lv = lane >> 4;
  const int nitems = *nitems_p;
  for (;;) {
    if (tid == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= nitems) break;
    const int4 it = items[item];
    const int tile = it.x % ntiles;
    const int kz = it.x / ntiles;
    const int ichan = kz / nzg;
    const int u0 = (tile % ntu) * kTile, v0 = (tile / ntu) * kTile;

    float acc[4][4][NP][2];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int q = 0; q < NP; ++q) acc[k][j][q][0] = acc[k][j][q][1] = 0.f;

    for (int pi = it.y; pi < it.z; ++pi) {
      const int s = pairs[pi];
      const SampleRec& r = recs[s];
      const int S2 = 2 * (r.spack & 0xffff);
      const int P = r.spack >> 16;
      const int du0 = u0 - r.iu0;            // footprint column of tile column 0
      const int dvw = v0 - r.iv0 + 16 * w;   // footprint row of this wave's first row
      const float2* base = cf + r.cfoff + (long long)(dvw + kGuardV + lv) * P + (du0 + kGuardU + lu);
      float2 c[4][4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (block_active(dvw + 4 * k, du0 + 16 * j, S2)) c[k][j] = base[(long long)4 * k * P + 16 * j];
      MacCoef m[NP];
      const bool cj = r.conj != 0;
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        const float vr = r.v[2 * q], vi = r.v[2 * q + 1];
        m[q] = cj ? MacCoef{vr, vi, vi, -vr} : MacCoef{vr, -vi, vi, vr};
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (block_active(dvw + 4 * k, du0 + 16 * j, S2)) {
            const float2 cc = c[k][j];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
              acc[k][j][q][0] = fmaf(m[q].a, cc.x, acc[k][j][q][0]);
              acc[k][j][q][0] = fmaf(m[q].b, cc.y, acc[k][j][q][0]);
              acc[k][j][q][1] = fmaf(m[q].c, cc.x, acc[k][j][q][1]);
              acc[k][j][q][1] = fmaf(m[q].d, cc.y, acc[k][j][q][1]);
            }
          }
    }
    // flush
    const size_t plane = (size_t)nx * ny;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int v = v0 + 16 * w + 4 * k + lv;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int u = u0 + 16 * j + lu;
        if (u < nx && v < ny) {
#pragma unroll
          for (int q = 0; q < NP; ++q) {
            float* g = grid + 2 * ((size_t)(ichan * NP + q) * plane + (size_t)v * nx + u);
            if (acc[k][j][q][0] != 0.f) atomicAdd(g, acc[k][j][q][0]);
            if (acc[k][j][q][1] != 0.f) atomicAdd(g + 1, acc[k][j][q][1]);
          }
        }
      }
    }
  }
}

// --------------------------------------------------------------------------------------
// degrid
// --------------------------------------------------------------------------------------
constexpr int kDegridBatch = 32;

// Sum over the 64 lanes of a wave with DPP (row shifts, then row broadcasts); the total is
// returned wave-uniformly (read from lane 63).
__device__ __forceinline__ float wave_sum_dpp(float x) {
  x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x111, 0xf, 0xf, true));
  x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x112, 0xf, 0xf, true));
  x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x114, 0xf, 0xf, true));
  x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x118, 0xf, 0xf, true));
  x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x142, 0xa, 0xf, true));
  x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x143, 0xc, 0xf, true));
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), 63));
}

template <int NP, int WPE>
__global__ __launch_bounds__(256, WPE) void k_degrid(const int4* __restrict__ items, const int* __restrict__ nitems_p,
                                                int* __restrict__ counter, const int* __restrict__ pairs,
                                                const SampleRec* __restrict__ recs, const float2* __restrict__ cf,
                                                const float2* __restrict__ grid, int nx, int ny, int ntu, int ntiles,
                                                int nzg, float* __restrict__ acc_out) {
  __shared__ int s_item;
  __shared__ float s_part[kDegridBatch][4][2 * NP];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int lu = lane & 15, lv = lane >> 4;
  const int nitems = *nitems_p;
  for (;;) {
    if (tid == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int item = s_item;
    __syncthreads();
    if (item >= nitems) break;
    const int4 it = items[item];
    const int tile = it.x % ntiles;
    const int kz = it.x / ntiles;
    const int ichan = kz / nzg;
    const int u0 = (tile % ntu) * kTile, v0 = (tile / ntu) * kTile;
    const size_t plane = (size_t)nx * ny;

    float2 g[4][4][NP];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int v = v0 + 16 * w + 4 * k + lv;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int u = u0 + 16 * j + lu;
#pragma unroll
        for (int q = 0; q < NP; ++q)
          g[k][j][q] = (u < nx && v < ny) ? grid[(size_t)(ichan * NP + q) * plane + (size_t)v * nx + u]
                                          : make_float2(0.f, 0.f);
      }
    }

    for (int pb = it.y; pb < it.z; pb += kDegridBatch) {
      const int nb = min(kDegridBatch, it.z - pb);
      for (int b = 0; b < nb; ++b) {
        const int s = pairs[pb + b];
        const SampleRec& r = recs[s];
        const int S2 = 2 * (r.spack & 0xffff);
        const int P = r.spack >> 16;
        const int du0 = u0 - r.iu0;
        const int dvw = v0 - r.iv0 + 16 * w;
        const float2* base = cf + r.cfoff + (long long)(dvw + kGuardV + lv) * P + (du0 + kGuardU + lu);
        float2 c[4][4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (block_active(dvw + 4 * k, du0 + 16 * j, S2)) c[k][j] = base[(long long)4 * k * P + 16 * j];
        float a[NP][2];
#pragma unroll
        for (int q = 0; q < NP; ++q) a[q][0] = a[q][1] = 0.f;
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (block_active(dvw + 4 * k, du0 + 16 * j, S2)) {
              const float2 cc = c[k][j];
#pragma unroll
              for (int q = 0; q < NP; ++q) {
                // CF * conj(G)  (GridKernel.cc:121-122)
                a[q][0] = fmaf(cc.x, g[k][j][q].x, a[q][0]);
                a[q][0] = fmaf(cc.y, g[k][j][q].y, a[q][0]);
                a[q][1] = fmaf(cc.y, g[k][j][q].x, a[q][1]);
                a[q][1] = fmaf(-cc.x, g[k][j][q].y, a[q][1]);
              }
            }
#pragma unroll
        for (int q = 0; q < NP; ++q) {
          a[q][0] = wave_sum_dpp(a[q][0]);
          a[q][1] = wave_sum_dpp(a[q][1]);
        }
        if (lane < 2 * NP) {
          float v = a[0][0];
#pragma unroll
          for (int q = 0; q < 2 * NP; ++q)
            if (lane == q) v = a[q / 2][q % 2];
          s_part[b][w][lane] = v;
        }
      }
      __syncthreads();
      for (int t = tid; t < nb * 2 * NP; t += blockDim.x) {
        const int b = t / (2 * NP), q = t % (2 * NP);
        const float sum = s_part[b][0][q] + s_part[b][1][q] + s_part[b][2][q] + s_part[b][3][q];
        if (sum != 0.f) atomicAdd(&acc_out[(size_t)pairs[pb + b] * 8 + q], sum);
      }
      __syncthreads();
    }
  }
}

__global__ void k_degrid_finish(const SampleRec* __restrict__ recs, const float* __restrict__ acc, int nrow, int nchan,
                                int npol, float2* __restrict__ vis_out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nrow * nchan) return;
  const SampleRec r = recs[s];
  if (r.iu0 == INT_MIN) return;
  const int row = s % nrow, chan = s / nrow;
  const float fr = r.v[0], fi = r.v[1];
  for (int pol = 0; pol < npol; ++pol) {
    const float cr = acc[(size_t)s * 8 + 2 * pol], ci = acc[(size_t)s * 8 + 2 * pol + 1];
    float2* o = vis_out + row + (size_t)nrow * (chan + (size_t)nchan * pol);
    float2 v = *o;
    v.x += cr * fr - ci * fi;
    v.y += cr * fi + ci * fr;
    *o = v;
  }
}

}  // namespace

// ----------------------------------------------------------------------------------------
// launchers
// ----------------------------------------------------------------------------------------
void launch_prep(const Ctx& c, const PrepParams& p, const PrepInputs& in, const PrepOutputs& out) {
  const int n = p.nrow * p.nchan;
  if (n == 0) return;
  hipLaunchKernelGGL(k_prep, dim3((n + 255) / 256), dim3(256), 0, c.stream, p, in, out);
  YG_HIP(hipGetLastError());
}

void launch_binning(const Ctx& c, const SampleRec* recs, int n, int nx, int ny, int nkz, Binning& b) {
  const int ntu = (nx + kTile - 1) / kTile, ntv = (ny + kTile - 1) / kTile;
  const size_t ntiles = (size_t)ntu * ntv;
  const size_t nkeys = ntiles * (size_t)nkz;
  YG_CHECK(nkeys < (size_t)INT_MAX, 1, "too many binning keys");
  b.nkeys = nkeys;
  b.counts.resize((nkeys + 1) * sizeof(int));
  b.offsets.resize((nkeys + 1) * sizeof(int));
  b.cursor.resize((nkeys + 1) * sizeof(int));
  b.nitems.resize(sizeof(int) * 2);
  b.counter.resize(sizeof(int) * 2);
  int* counts = b.counts.as<int>();
  int* offsets = b.offsets.as<int>();
  YG_HIP(hipMemsetAsync(counts, 0, (nkeys + 1) * sizeof(int), c.stream));
  if (n > 0) hipLaunchKernelGGL(k_bin_count, dim3((n + 255) / 256), dim3(256), 0, c.stream, recs, n, ntu, (int)ntiles, counts);
  size_t tbytes = 0;
  YG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tbytes, counts, offsets, (int)(nkeys + 1), c.stream));
  b.temp.resize(tbytes);
  YG_HIP(hipcub::DeviceScan::ExclusiveSum(b.temp.get(), tbytes, counts, offsets, (int)(nkeys + 1), c.stream));
  // the total pair count bounds the pairs buffer: read it back (one small copy per pass)
  int npairs = 0;
  YG_HIP(hipMemcpyAsync(&npairs, offsets + nkeys, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  YG_HIP(hipStreamSynchronize(c.stream));
  YG_CHECK(npairs >= 0, 3, "pair count overflow");
  b.pairs.resize((size_t)(npairs > 0 ? npairs : 1) * sizeof(int));
  YG_HIP(hipMemcpyAsync(b.cursor.get(), offsets, (nkeys + 1) * sizeof(int), hipMemcpyDeviceToDevice, c.stream));
  if (n > 0)
    hipLaunchKernelGGL(k_bin_fill, dim3((n + 255) / 256), dim3(256), 0, c.stream, recs, n, ntu, (int)ntiles,
                       b.cursor.as<int>(), b.pairs.as<int>());
  // work items: reuse cursor as the per-key item count and its scan
  int* icount = b.cursor.as<int>();
  hipLaunchKernelGGL(k_item_count, dim3((nkeys + 1 + 255) / 256), dim3(256), 0, c.stream, counts, (int)nkeys, icount);
  DevBuf& ioffsets = b.ioffsets;
  ioffsets.resize((nkeys + 1) * sizeof(int));
  size_t tb2 = 0;
  YG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, icount, ioffsets.as<int>(), (int)(nkeys + 1), c.stream));
  b.temp.resize(tb2);
  YG_HIP(hipcub::DeviceScan::ExclusiveSum(b.temp.get(), tb2, icount, ioffsets.as<int>(), (int)(nkeys + 1), c.stream));
  const size_t max_items = (size_t)npairs / kChunk + nkeys + 1;
  b.items.resize(max_items * sizeof(int4));
  hipLaunchKernelGGL(k_item_emit, dim3((nkeys + 255) / 256), dim3(256), 0, c.stream, counts, offsets,
                     ioffsets.as<int>(), (int)nkeys, b.items.as<int4>());
  YG_HIP(hipMemcpyAsync(b.nitems.get(), ioffsets.as<int>() + nkeys, sizeof(int), hipMemcpyDeviceToDevice, c.stream));
  YG_HIP(hipGetLastError());
}

static int persistent_blocks(const Ctx& c) { return c.n_cu * (c.waves_per_simd >= 3 ? 3 : 2); }

void launch_grid(const Ctx& c, const Binning& b, const SampleRec* recs, const float2* cf, float2* grid, int nx, int ny,
                 int npol, int nzg) {
  const int ntu = (nx + kTile - 1) / kTile, ntv = (ny + kTile - 1) / kTile;
  const int ntiles = ntu * ntv;
  YG_HIP(hipMemsetAsync(b.counter.get(), 0, sizeof(int), c.stream));
  const dim3 g(persistent_blocks(c)), t(256);
  float* gf = reinterpret_cast<float*>(grid);
#define YG_GRID_CASE(NPV)                                                                                       \
  case NPV:                                                                                                     \
    if (c.waves_per_simd >= 3)                                                                                  \
      hipLaunchKernelGGL((k_grid<NPV, 3>), g, t, 0, c.stream, b.items.as<int4>(), b.nitems.as<int>(),         \
                         b.counter.as<int>(), b.pairs.as<int>(), recs, cf, gf, nx, ny, ntu, ntiles, nzg);       \
    else                                                                                                        \
      hipLaunchKernelGGL((k_grid<NPV, 2>), g, t, 0, c.stream, b.items.as<int4>(), b.nitems.as<int>(),         \
                         b.counter.as<int>(), b.pairs.as<int>(), recs, cf, gf, nx, ny, ntu, ntiles, nzg);       \
    break;
  switch (npol) {
    YG_GRID_CASE(1)
    YG_GRID_CASE(2)
    YG_GRID_CASE(3)
    YG_GRID_CASE(4)
    default:
      throw Error(1, "unsupported number of polarisations (1..4)");
  }
#undef YG_GRID_CASE
  YG_HIP(hipGetLastError());
}

void launch_degrid(const Ctx& c, const Binning& b, const SampleRec* recs, const float2* cf, const float2* grid, int nx,
                   int ny, int npol, int nzg, float* acc) {
  const int ntu = (nx + kTile - 1) / kTile, ntv = (ny + kTile - 1) / kTile;
  const int ntiles = ntu * ntv;
  YG_HIP(hipMemsetAsync(b.counter.get(), 0, sizeof(int), c.stream));
  const dim3 g(persistent_blocks(c)), t(256);
#define YG_DEGRID_CASE(NPV)                                                                                     \
  case NPV:                                                                                                     \
    if (c.waves_per_simd >= 3)                                                                                  \
      hipLaunchKernelGGL((k_degrid<NPV, 3>), g, t, 0, c.stream, b.items.as<int4>(), b.nitems.as<int>(),       \
                         b.counter.as<int>(), b.pairs.as<int>(), recs, cf, grid, nx, ny, ntu, ntiles, nzg, acc); \
    else                                                                                                        \
      hipLaunchKernelGGL((k_degrid<NPV, 2>), g, t, 0, c.stream, b.items.as<int4>(), b.nitems.as<int>(),       \
                         b.counter.as<int>(), b.pairs.as<int>(), recs, cf, grid, nx, ny, ntu, ntiles, nzg, acc); \
    break;
  switch (npol) {
    YG_DEGRID_CASE(1)
    YG_DEGRID_CASE(2)
    YG_DEGRID_CASE(3)
    YG_DEGRID_CASE(4)
    default:
      throw Error(1, "unsupported number of polarisations (1..4)");
  }
#undef YG_DEGRID_CASE
  YG_HIP(hipGetLastError());
}

void launch_degrid_finish(const Ctx& c, const SampleRec* recs, const float* acc, int nrow, int nchan, int npol,
                          float2* vis_out) {
  const int n = nrow * nchan;
  if (n == 0) return;
  hipLaunchKernelGGL(k_degrid_finish, dim3((n + 255) / 256), dim3(256), 0, c.stream, recs, acc, nrow, nchan, npol,
                     vis_out);
  YG_HIP(hipGetLastError());
}

}  // namespace ygrid
