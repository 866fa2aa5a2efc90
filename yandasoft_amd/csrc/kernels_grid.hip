// Grid / degrid kernels for gfx950 (CDNA4, wave64).
//
// Pass structure (one call of TableVisGridder::generic, TableVisGridder.cc:451-884):
//   k_prep        one thread per (row, chan): uvw -> (iu, fracu, iv, fracv) in FP64, CF plane,
//                 on-grid test, flags, noise weight, sum of weights; writes the sample header, key and coefficients.
//   k_bin_*       counting sort of (sample, grid tile) pairs by (image chan, CF plane group, tile);
//                 the key order is plane-major so that concurrently running workgroups read the
//                 same few CF planes, which then stay resident in each XCD's L2.
//   k_grid        persistent workgroups take work items (<= kDefaultItemPairs pairs of one key) in key order.
//                 A workgroup owns a 64x64 tile of all polarisations in registers (16 pixels per
//                 lane), spreads every pair into it, and flushes once with float atomics: no per-tap
//                 HBM atomics (GridKernel::grid, GridKernel.cc:61-83).
//   k_degrid_mm   same work items; the tile of the model grid is loaded into registers and four pairs
//                 at a time are dot products on the matrix cores (GridKernel::degrid, GridKernel.cc:106-125);
//                 partial sums of the tiles a footprint covers meet in per-sample float accumulators.
//   k_degrid      the one-polarisation form: each pair a wave-level dot product in FP32 FMAs.
//   k_degrid_finish  phasor x weight and the add into the visibility cube (TableVisGridder.cc:767-773,861).
// Around the passes: k_residual (the imaging loop's data - model), k_pol_to_image / k_pol_to_vis_add
// (polarisation frames, TableVisGridder.cc:682-693,860).  W stacking offsets the tile kernels' grid
// by the key's w plane (TileJob::wstride).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <type_traits>

#include "device_util.h"
#include "ygrid_internal.h"

namespace ygrid {

namespace {

constexpr int kChunk = 512;  // pairs per LDS staging round of a work item
// floats of one wave's flush area in k_grid: 4 rows x 64 columns x 4 outputs, one pad per 32
constexpr int kFlushWave = 4 * 64 * 4 + 4 * 64 * 4 / 32;

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

// MVDirection::separation: great-circle angle between two (longitude, latitude) directions
__device__ __forceinline__ double separation_d(double l1, double b1, double l2, double b2) {
  const double dl = l2 - l1;
  const double a = cos(b2) * sin(dl);
  const double b = cos(b1) * sin(b2) - sin(b1) * cos(b2) * cos(dl);
  const double c = sin(b1) * sin(b2) + cos(b1) * cos(b2) * cos(dl);
  return atan2(sqrt(a * a + b * b), c);
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
// Blocks loop over the samples in rounds of 256.  With lds_sumw the sum of weights of the block
// is first accumulated in shared memory (nsw_rows*npol*nchan_img doubles) and added to the
// global sums once at the end; otherwise per distinct row within each wave.
__global__ __launch_bounds__(256) void k_prep(PrepParams p, PrepInputs in, PrepOutputs out, int lds_sumw) {
#pragma clang fp contract(off)
  extern __shared__ double s_sumw[];
  // the block's counters, added to the pass's counters once per block (the counters share one
  // cache line: per-wave atomics from thousands of waves serialise there)
  __shared__ unsigned long long s_cnt[5];
  __shared__ int s_err;
  const int n = p.nrow * p.nchan;
  const bool grid_mode = p.mode != kModeDegrid;
  const bool sumw_on = grid_mode && p.mode != kModePCF;
  const bool use_lds = sumw_on && lds_sumw != 0;
  const int nsw = p.nsw_rows * p.npol * p.nchan_img;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  if (tid < 5) s_cnt[tid] = 0;
  if (tid == 0) s_err = 0;
  if (use_lds)
    for (int e = tid; e < nsw; e += blockDim.x) s_sumw[e] = 0.0;
  __syncthreads();
  unsigned long long c_valid = 0, c_taps = 0, c_flag = 0, c_wflag = 0, c_reject = 0;
  int err = 0;
  for (long long base = (long long)blockIdx.x * blockDim.x; base < n; base += (long long)gridDim.x * blockDim.x) {
  const int s = (int)base + tid;  // s = row + nrow*chan
  const bool inb = s < n;
  const int row = inb ? s % p.nrow : 0;
  const int chan = inb ? s / p.nrow : 0;

  SampleHdr rec;
  rec.iu0 = INT_MIN;
  rec.iv0 = 0;
  rec.cfoff = 0;
  rec.spack = 0;
  int key = 0;
  bool conj = false;
  float coef[kCoef];
#pragma unroll
  for (int q = 0; q < kCoef; ++q) coef[q] = 0.f;

  bool valid = inb;
  // rows whose pointing lies farther than MaxPointingSeparation from the image centre are left
  // out before anything else (TableVisGridder.cc:571-578)
  if (valid && p.max_sep > 0.0 && in.pointing1 &&
      separation_d(p.centre_lon, p.centre_lat, in.pointing1[2 * row], in.pointing1[2 * row + 1]) > p.max_sep) {
    valid = false;
    if (chan == 0) ++c_reject;
  }
  bool flagged = false, wflagged = false;
  const bool used = valid;
  int swrow = 0, ichan = 0;
  float wts[4] = {0.f, 0.f, 0.f, 0.f};
  float4 lcf = make_float4(0.f, 0.f, 0.f, 0.f);
  unsigned long long taps = 0;

  if (used) {
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
    // (no delays: the phase is 0 and the phasor exactly (1, 0), without the FP64 cos and sin)
    float ph_r = 1.f, ph_i = 0.f;
    if (in.delay) {
      const double phase = 2.0f * kPi * f * in.delay[row] / kC;
      ph_r = (float)cos(phase);
      ph_i = (float)sin(phase);
    }

    for (int pol = 0; pol < p.npol; ++pol)
      if (in.flag && in.flag[row + (size_t)p.nrow * (chan + (size_t)p.nchan * pol)]) { flagged = true; break; }
    ichan = in.chanmap ? in.chanmap[chan] : 0;
    if (ichan < 0) flagged = true;
    if (flagged) {
      valid = false;
    } else {
      // cIndex
      int plane = 0, gplane = 0;
      if (p.ci.kind != 0) {
        const double w = in.uvw[3 * row + 2] / kC;
        const int iw = w_plane(p.ci, w * f);
        if (iw < 0) {
          if (!p.ci.wmaxclip || iw == -2) err |= 2;
          plane = -1;
        } else if (p.ci.kind == 3) {
          gplane = iw;  // WStackVisGridder::gIndex (WStackVisGridder.cc:339-344); cIndex 0
          plane = 0;
        } else if (p.ci.kind == 4) {
          // AProjectWStackVisGridder::initIndices (AProjectWStackVisGridder.cc:125-172): the grid of
          // the w plane, the CF of (channel, feed, field)
          gplane = iw;
          const int feed = in.feed1 ? in.feed1[row] : 0;
          const int c = p.ci.freqdep ? chan : 0;
          plane = c + p.ci.nchan_cf * (feed + p.ci.maxfeeds * p.ci.field);
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
          const bool learned = p.lcf_plane_elems > 0;
          const int S = learned ? p.lcf_S : in.cf_supp[cInd];
          const int iuo = iu + (!learned && in.cf_ofs_u ? in.cf_ofs_u[plane] : 0);
          const int ivo = iv + (!learned && in.cf_ofs_v ? in.cf_ofs_v[plane] : 0);
          // strict on-grid test (TableVisGridder.cc:763-764)
          if (!((iuo - S) > 0 && (ivo - S) > 0 && (iuo + S) < p.nx && (ivo + S) < p.ny)) {
            valid = false;
          } else {
            rec.iu0 = iuo - S;
            rec.iv0 = ivo - S;
            rec.cfoff = learned ? (int32_t)((long long)s * p.lcf_plane_elems) : in.cf_off[cInd];
            if (learned) lcf = make_float4((float)(in.uvw[3 * row + 2] / kC * f), (float)fracu, (float)fracv, 1.f);
            rec.spack = S | ((2 * S + 1 + 2 * kGuardU) << 16);
            key = ichan * p.nzg + (p.ci.kind >= 3 ? gplane : plane / p.zdiv);
            swrow = p.track_per_os ? cInd : plane;
            const float vw = in.visweight ? in.visweight[chan] : 1.0f;
            if (grid_mode) {
              if (p.mode == kModePSF && p.psf_feed >= 0) {
                // the representative feed and field (TableVisGridder.cc:809-812)
                if ((in.feed1 ? in.feed1[row] : 0) != p.psf_feed) valid = false;
                const double* dish = in.dish1 ? in.dish1 : in.pointing1;
                if (valid && p.psf_field && dish &&
                    !(separation_d(p.psf_lon, p.psf_lat, dish[2 * row], dish[2 * row + 1]) < 1e-6))
                  valid = false;
              }
              if (p.mode == kModePCF) {
                conj = (ivo < p.ny / 2 && iuo >= p.nx / 2) || (ivo <= p.ny / 2 && iuo < p.nx / 2);
              }
#pragma unroll
              for (int pol = 0; pol < 4; ++pol) {
                if (pol >= p.npol || !valid) break;
                const size_t ci = row + (size_t)p.nrow * (chan + (size_t)p.nchan * pol);
                const float nre = in.noise ? in.noise[ci].x : 1.0f;
                const float visNoise = nre * nre;
                const float wt = (visNoise > 0.f) ? (float)(1. / (double)visNoise) : 0.f;
                if (!(wt > 0.f)) { err |= 1; valid = false; break; }
                wts[pol] = wt;
                float rr = 0.f, ri = 0.f;
                if (p.mode == kModeGrid) {
                  const float2 v = in.vis[ci];
                  // rVis = phasor * conj(vis) * wt (* visweight)  (TableVisGridder.cc:797-800)
                  rr = (ph_r * v.x + ph_i * v.y) * wt;
                  ri = (ph_i * v.x - ph_r * v.y) * wt;
                  rr *= vw;
                  ri *= vw;
                } else if (p.mode == kModePSF) {
                  rr = wt * vw;
                } else {
                  rr = wt;
                }
                // MAC coefficients of rVis*C, or rVis*conj(C) (TableVisGridder.cc:837-844):
                // re += a*C.re + b*C.im, im += c*C.re + d*C.im, stored as the gridding kernel's
                // MFMA B operands: coef[4*o + m], output o = 2*(pol&1) + (re 0 / im 1),
                // operand m = 2*(pol/2) + (C.re 0 / C.im 1)
                const int ore = 2 * (pol & 1), mre = 2 * (pol >> 1);
                coef[4 * ore + mre] = rr;
                coef[4 * ore + mre + 1] = conj ? ri : -ri;
                coef[4 * (ore + 1) + mre] = ri;
                coef[4 * (ore + 1) + mre + 1] = conj ? -rr : rr;
              }
              if (!valid) rec.iu0 = INT_MIN;
            } else {
              coef[0] = ph_r * vw;
              coef[1] = ph_i * vw;
            }
            if (valid) taps = (unsigned long long)(2 * S + 1) * (2 * S + 1) * p.npol;
          }
        }
      }
    }
  }
  if (inb) {
    out.smp.hdr[s] = rec;
    out.smp.key[s] = key;
    if (out.lcf_in) out.lcf_in[s] = valid ? lcf : make_float4(0.f, 0.f, 0.f, 0.f);
    if (p.vis_prec != kPrecF32 && p.mode != kModeDegrid) {
      // gridding coefficients stored at 16-bit precision: 32 bytes per sample
      uint4* c4 = reinterpret_cast<uint4*>(out.smp.coef16 + (size_t)s * kCoef);
      uint32_t wds[kCoef / 2];
#pragma unroll
      for (int q = 0; q < kCoef / 2; ++q)
        wds[q] = (uint32_t)round16(coef[2 * q], p.vis_prec) | ((uint32_t)round16(coef[2 * q + 1], p.vis_prec) << 16);
#pragma unroll
      for (int q = 0; q < kCoef / 8; ++q) c4[q] = make_uint4(wds[4 * q], wds[4 * q + 1], wds[4 * q + 2], wds[4 * q + 3]);
    } else {
      // (degridding reads only [0..1], the phasor times the weight)
      float4* c4 = reinterpret_cast<float4*>(out.smp.coef + (size_t)s * kCoef);
      const int nq = p.mode == kModeDegrid ? 1 : kCoef / 4;
#pragma unroll
      for (int q = 0; q < kCoef / 4; ++q)
        if (q < nq) c4[q] = make_float4(coef[4 * q], coef[4 * q + 1], coef[4 * q + 2], coef[4 * q + 3]);
    }
  }

  c_valid += valid ? 1 : 0;
  c_taps += taps;
  c_flag += (flagged && grid_mode) ? 1 : 0;
  c_wflag += wflagged ? 1 : 0;

  // sum of weights per (row of sumw, pol, image chan): TableVisGridder.cc:776-782
  if (use_lds) {
    // a wave whose valid lanes share one (row of sumw, image chan) -- always so for the gridders
    // with one CF row (Box, SphFunc, w stacking), whose lanes would otherwise all contend for the
    // same LDS words -- adds its wave sum once; other waves add per lane
    const int lkey = valid ? swrow + p.nsw_rows * p.npol * ichan : -1;
    const unsigned long long vm = __ballot(valid);
    const int leader = vm ? __ffsll((long long)vm) - 1 : 0;
    const int k0 = __shfl(lkey, leader, 64);
    if (vm != 0 && __all(!valid || lkey == k0)) {
      for (int pol = 0; pol < p.npol; ++pol) {
        const double t = wave_sum_f64(valid ? (double)wts[pol] : 0.0);
        if (lane == leader) atomicAdd(&s_sumw[k0 + p.nsw_rows * pol], t);
      }
    } else if (valid) {
      for (int pol = 0; pol < p.npol; ++pol)
        atomicAdd(&s_sumw[swrow + p.nsw_rows * (pol + p.npol * ichan)], (double)wts[pol]);
    }
  } else if (sumw_on) {
    // aggregated per distinct (row of sumw, image chan) within the wave
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
  }  // rounds

  // counters (wave aggregated)
  const unsigned long long nvalid = wave_sum_u64(c_valid);
  const unsigned long long nflag = wave_sum_u64(c_flag);
  const unsigned long long nwflag = wave_sum_u64(c_wflag);
  const unsigned long long tsum = wave_sum_u64(c_taps);
  const unsigned long long nrej = wave_sum_u64(c_reject);
  if (lane == 0) {
    if (nrej) atomicAdd(&s_cnt[4], nrej);
    if (nvalid) atomicAdd(&s_cnt[0], nvalid * (unsigned long long)p.npol);
    if (tsum) atomicAdd(&s_cnt[1], tsum);
    if (nflag) atomicAdd(&s_cnt[2], nflag);
    if (nwflag) atomicAdd(&s_cnt[3], nwflag * (grid_mode ? 2ull : 1ull));
  }
  if (err) atomicOr(&s_err, err);
  __syncthreads();
  if (tid < 5 && s_cnt[tid]) atomicAdd(&out.counters[tid], s_cnt[tid]);
  if (tid == 0 && s_err) atomicOr(out.error, s_err);
  if (use_lds)
    for (int e = tid; e < nsw; e += blockDim.x)
      if (s_sumw[e] != 0.0) atomicAdd(&out.sumw[e], s_sumw[e]);
}

// --------------------------------------------------------------------------------------
// binning
// --------------------------------------------------------------------------------------
struct TileRange {
  int tu0, tu1, tv0, tv1;
};
__device__ __forceinline__ TileRange tiles_of(const SampleHdr& r) {
  const int W = 2 * (r.spack & 0xffff) + 1;
  return {r.iu0 / kTile, (r.iu0 + W - 1) / kTile, r.iv0 / kTile, (r.iv0 + W - 1) / kTile};
}

// The binning kernels take the samples channel-fastest (thread t: row t / nchan, channel
// t % nchan), so that neighbouring lanes are neighbouring channels of one row: their footprints
// move by a few pixels per channel and mostly fall in the same (plane, tile) bins.  Lanes with
// equal bins in consecutive lanes ("runs") then share one atomic.
struct BinLane {
  int s;        // sample index (row + nrow*chan), -1 when the lane has none
  SampleHdr r;
  TileRange t;
  long long base;  // key * ntiles
};

__device__ __forceinline__ BinLane bin_lane(const Samples& smp, long long tglob, int nrow, int nchan, int ntiles) {
  BinLane b;
  b.s = -1;
  b.r.iu0 = INT_MIN;
  b.t = TileRange{0, -1, 0, -1};
  b.base = 0;
  const long long n = (long long)nrow * nchan;
  if (tglob < n) {
    const int row = (int)(tglob / nchan), chan = (int)(tglob % nchan);
    const int s = row + nrow * chan;
    const SampleHdr r = smp.hdr[s];
    if (r.iu0 != INT_MIN) {
      b.s = s;
      b.r = r;
      b.t = tiles_of(r);
      b.base = (long long)smp.key[s] * ntiles;
    }
  }
  return b;
}

// Runs of equal idx among the active lanes of a wave: head = lowest lane of the run, rank =
// lane - head, len = run length (meaningful at the head).
struct Run {
  bool head;
  int hlane, rank, len;
};
__device__ __forceinline__ Run wave_run(long long idx, bool active, int lane) {
  const long long prev = __shfl_up(idx, 1, 64);
  const int pa = __shfl_up(active ? 1 : 0, 1, 64);
  Run r;
  r.head = active && (lane == 0 || !pa || prev != idx);
  const unsigned long long heads = __ballot(r.head);
  const unsigned long long act = __ballot(active);
  const unsigned long long upto = (2ull << lane) - 1;  // lanes 0..lane
  const unsigned long long hb = heads & upto;
  r.hlane = hb ? 63 - __clzll(hb) : lane;
  r.rank = lane - r.hlane;
  const unsigned long long after = (heads | ~act) & ~upto;
  r.len = (after ? __ffsll((long long)after) - 1 : 64) - lane;
  return r;
}

// widest footprint in tiles over the wave (loop bounds shared by all lanes)
__device__ __forceinline__ int wave_max_int(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
  return v;
}

// counts (and cursor in k_bin_fill) of key k live at [k * kstride]: with few keys, kstride 32
// gives every key its own 128-byte line, so that atomics on different keys do not serialise on
// a shared line
__global__ void k_bin_count(Samples smp, int nrow, int nchan, int ntu, int ntiles, int* counts, int kstride) {
  const long long tglob = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const BinLane b = bin_lane(smp, tglob, nrow, nchan, ntiles);
  const int nv = wave_max_int(b.t.tv1 - b.t.tv0 + 1), nu = wave_max_int(b.t.tu1 - b.t.tu0 + 1);
  for (int dv = 0; dv < nv; ++dv)
    for (int du = 0; du < nu; ++du) {
      const int tv = b.t.tv0 + dv, tu = b.t.tu0 + du;
      const bool active = b.s >= 0 && tv <= b.t.tv1 && tu <= b.t.tu1;
      const long long idx = active ? b.base + tv * ntu + tu : -1;
      const Run run = wave_run(idx, active, lane);
      if (run.head) atomicAdd(&counts[idx * kstride], run.len);
    }
}

// bits lo..hi (clamped to 0..n-1) of an n-bit set
template <int N>
__device__ __forceinline__ uint32_t bit_range(int lo, int hi) {
  lo = max(lo, 0);
  hi = min(hi, N - 1);
  return lo > hi ? 0u : ((2u << hi) - (1u << lo));
}

// Descriptor of one (sample, tile) pair, read wave-uniformly by the tile kernels.  The tile is
// worked by four waves, wave w on rows [16w, 16w+16), each in 8 blocks (k, j) of 4 rows x 32
// columns (block rows 16w+4k .. +3, columns 32j .. +31 of the tile).
//   x: sample index
//   y: byte offset (in a CF buffer that starts pad bytes before the table, taps of esz bytes)
//      of the CF tap under the tile's pixel (0, 0), modulo 2^32; wave w adds 16*w rows
//   z: bits 8w+2k+j: block (k, j) of wave w overlaps the footprint
//   w: CF row pitch P (bits 0..15), and the signed bytes A (16..23), B (24..31) with: the two
//      taps of lane column pair x = 32j + 2lu (tile columns x, x+1) touch the footprint when
//      A <= x <= B (A = -1 - du0, B = 2S - du0, clamped to what x can take)
__device__ __forceinline__ int4 pair_descriptor(const SampleHdr& r, int s, int u0, int v0, uint32_t pad, uint32_t esz) {
  const int S2 = 2 * (r.spack & 0xffff), P = r.spack >> 16;
  const int du0 = u0 - r.iu0, dv0 = v0 - r.iv0;
  // block column j: du0+32j+31 >= 0 and du0+32j <= 2S; block row k of wave w, with
  // dvw = dv0+16w: dvw+4k+3 >= 0 and dvw+4k <= 2S  (>> is floor division by a power of two)
  const uint32_t cols = bit_range<2>(-((du0 + 31) >> 5), (S2 - du0) >> 5);
  uint32_t masks = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const int dvw = dv0 + 16 * w;
    const uint32_t rows = bit_range<4>(-((dvw + 3) >> 2), (S2 - dvw) >> 2);
    const uint32_t rows8 = (rows & 1u) | ((rows & 2u) << 1) | ((rows & 4u) << 2) | ((rows & 8u) << 3);
    masks |= (rows8 * cols) << (8 * w);
  }
  const uint32_t soff = pad + (uint32_t)(r.cfoff + (dv0 + kGuardV) * P + (du0 + kGuardU)) * esz;
  const int A = min(max(-1 - du0, -2), 64);
  const int B = min(max(S2 - du0, -1), 100);
  const uint32_t pw = (uint32_t)(P & 0xffff) | ((uint32_t)(A & 0xff) << 16) | ((uint32_t)(B & 0xff) << 24);
  return make_int4(s, (int)soff, (int)masks, (int)pw);
}

// zg_base (tables of 2 GiB or more): descriptor offsets relative to the sample's key group's
// first plane (the group: key % nzg)
__global__ void k_bin_fill(Samples smp, int nrow, int nchan, int ntu, int ntiles, uint32_t pad, uint32_t esz,
                           int* cursor, int kstride, int4* descs, const long long* __restrict__ zg_base, int nzg) {
  const long long tglob = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  BinLane b = bin_lane(smp, tglob, nrow, nchan, ntiles);
  if (zg_base && b.s >= 0) b.r.cfoff = (int32_t)((long long)b.r.cfoff - zg_base[smp.key[b.s] % nzg]);
  const int nv = wave_max_int(b.t.tv1 - b.t.tv0 + 1), nu = wave_max_int(b.t.tu1 - b.t.tu0 + 1);
  for (int dv = 0; dv < nv; ++dv)
    for (int du = 0; du < nu; ++du) {
      const int tv = b.t.tv0 + dv, tu = b.t.tu0 + du;
      const bool active = b.s >= 0 && tv <= b.t.tv1 && tu <= b.t.tu1;
      const long long idx = active ? b.base + tv * ntu + tu : -1;
      const Run run = wave_run(idx, active, lane);
      const int hpos = run.head ? atomicAdd(&cursor[idx * kstride], run.len) : 0;
      const int pos = __shfl(hpos, run.hlane, 64) + run.rank;
      if (active) descs[pos] = pair_descriptor(b.r, b.s, tu * kTile, tv * kTile, pad, esz);
    }
}

__global__ void k_item_count(const int* __restrict__ counts, int nkeys, int chunk, int* icount) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nkeys) return;
  icount[k] = (k < nkeys) ? (counts[k] + chunk - 1) / chunk : 0;
}

// work items of one job (its keys' pair ranges in chunks of <= chunk pairs) written from
// items[*base] on, tagged with the job index
__global__ void k_item_emit(const int* __restrict__ counts, const int* __restrict__ offsets,
                            const int* __restrict__ ioffsets, int nkeys, int chunk, const int* __restrict__ base,
                            int job, int4* items) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  const int c = counts[k];
  if (c == 0) return;
  const int o = offsets[k];
  const int io = *base + ioffsets[k];
  const int nch = (c + chunk - 1) / chunk;
  for (int j = 0; j < nch; ++j) {
    const int b = o + j * chunk;
    items[io + j] = make_int4(k, b, min(b + chunk, o + c), job);
  }
}

// *base_out = *base_in + (item count of the job: the last entry of its exclusive item scan)
__global__ void k_item_base(const int* __restrict__ base_in, const int* __restrict__ iscan_end, int* base_out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *base_out = *base_in + *iscan_end;
}

__global__ void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// dst[i * dstride] = src[i * sstride], i < n
__global__ void k_copy_strided_i32(const int* __restrict__ src, int sstride, int* __restrict__ dst, int dstride,
                                   size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i * dstride] = src[i * sstride];
}

// --------------------------------------------------------------------------------------
// grid: lane layout in a 64x64 tile.  Wave w owns rows [16w, 16w+16); lane (lu, lv) =
// (lane & 15, lane >> 4) owns pixels u = 32j + 2lu + e, v = 16w + 4k + lv for k in 0..3,
// j, e in 0..1, i.e. 8 "blocks" (k, j) of 32x4 pixels per wave, two adjacent pixels per lane
// and block, so that a block's CF taps are one 16-byte load per lane (the vector memory path
// moves four lanes per clock, whatever their width).  A (sample, tile) pair touches the blocks
// its footprint overlaps.  Everything that depends on the pair only (footprint geometry, which
// blocks are active, CF row offsets, MAC coefficients) is wave-uniform and lives in SGPRs; per
// lane there is one CF byte offset per column block, set out of the buffer's range for lanes
// whose taps lie beyond the footprint's guard band (such loads return zeros and touch no
// memory).  All of a pair's CF loads are issued before its FMAs; the pipelined form (PIPE)
// issues the loads of pair i+1 before the FMAs of pair i into a second register set.
// --------------------------------------------------------------------------------------

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Operands of one pair for one lane: the CF taps [k][j][e] and, for gridding, the four MFMA
// B operands of the sample (Samples::coef in the lane-major layout written by k_prep).
template <int CFP>
struct Taps {
  // FP32 tables: the taps as loaded; 16-bit tables: the packed (re, im) words, widened where the
  // taps are used (widening at the load would make the compiler wait for each load at once)
  typename std::conditional<CFP == kPrecF32, float2, uint32_t>::type t[4][2][2];
  f32x4 b;
  __device__ __forceinline__ float2 tap(int k, int j, int e) const {
    if constexpr (CFP == kPrecF32) {
      return t[k][j][e];
    } else {
      return make_float2(widen16(t[k][j][e] & 0xffffu, CFP), widen16(t[k][j][e] >> 16, CFP));
    }
  }
};

// s_waitcnt vmcnt(0) (expcnt and lgkmcnt unconstrained): every vector memory load issued so
// far has returned.  Placed before the loads of the next pair are issued, it tells the
// compiler's wait-count analysis that the current pair's CF registers are ready.
constexpr int kWaitVmcnt0 = 0x0F70;

// Buffer resource word 3 for raw 32-bit-element buffer accesses on gfx9-family parts
// (DATA_FORMAT = 32); bounds are checked against num_records in bytes.
constexpr int kBufferWord3 = 0x00020000;

// Staging of a work item: before its pairs are worked, the workgroup copies the item's pair
// descriptors (and, for gridding, the 16 MAC coefficients of each pair's sample) into LDS with
// all its threads at once, so that the latency of those scattered reads is paid once per item
// instead of once per pair on each wave's critical path.  A wave then reads a pair's descriptor
// from LDS (every lane the same 16 bytes) and makes it wave-uniform.
template <bool COEF>
__device__ __forceinline__ void stage_item(const int4* __restrict__ descs, const Samples& smp, int pb, int n,
                                           int4* s_desc, f32x4* s_coef, int tid, int nthreads) {
  for (int t = tid; t < n; t += nthreads) {
    const int4 d = descs[pb + t];
    s_desc[t] = d;
    if constexpr (COEF) {
      if (smp.coef16) {
        const uint4* src = reinterpret_cast<const uint4*>(smp.coef16 + (size_t)d.x * kCoef);
#pragma unroll
        for (int h = 0; h < kCoef / 8; ++h) {
          const uint4 u = src[h];
          const uint32_t wd[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
          for (int q = 0; q < 2; ++q)
            s_coef[(kCoef / 4) * t + 2 * h + q] =
                f32x4{widen16(wd[2 * q] & 0xffffu, smp.coef_prec), widen16(wd[2 * q] >> 16, smp.coef_prec),
                      widen16(wd[2 * q + 1] & 0xffffu, smp.coef_prec), widen16(wd[2 * q + 1] >> 16, smp.coef_prec)};
        }
      } else {
        const f32x4* src = reinterpret_cast<const f32x4*>(smp.coef + (size_t)d.x * kCoef);
#pragma unroll
        for (int q = 0; q < kCoef / 4; ++q) s_coef[(kCoef / 4) * t + q] = src[q];
      }
    }
  }
  __syncthreads();
}

struct PairU {
  int4 d;
};

// Geometry of a pair seen from wave w: bit 2k+j of mask = block (k, j) active, soff0 = byte
// offset of the CF tap under the wave's first pixel (lane part excluded), rowstep = bytes per 4
// CF rows, and the lane column-pair range [A, B] (see pair_descriptor).  The rows of an active
// block stay inside the padded CF plane (the guard band covers the 3 rows a block may overhang)
// and every offset of an active block is non-negative (the buffer starts pad bytes early).
struct PairGeom {
  int P, A, B;
  uint32_t mask, soff0, rowstep;
};

// The active blocks of a pair in one wave, from its 8-bit mask (bit 2k+j: block (k, j)), tested
// per block at run time.  (Straight-line code per block rectangle, chosen by one dispatch, was
// built and compiled to kernels that spill hundreds of registers; the tests stay.)
struct MaskShape {
  uint32_t mask;
};
template <int V>
using IC = std::integral_constant<int, V>;
// f(IC<k>, IC<j>) for the active blocks, rows first
template <int B = 0, class F>
__device__ __forceinline__ void for_blocks(MaskShape shp, F&& f) {
  if constexpr (B < 8) {
    if (shp.mask & (1u << B)) f(IC<(B >> 1)>{}, IC<(B & 1)>{});
    for_blocks<B + 1>(shp, f);
  }
}

template <int SH>
__device__ __forceinline__ PairGeom pair_geom(const PairU& p, int w) {
  PairGeom g;
  const uint32_t pw = (uint32_t)p.d.w;
  g.P = (int)(pw & 0xffffu);
  g.A = (int)(pw << 8) >> 24;
  g.B = (int)pw >> 24;
  g.mask = ((uint32_t)p.d.z >> (8 * w)) & 0xffu;
  g.soff0 = (uint32_t)p.d.y + (uint32_t)w * ((uint32_t)g.P << (4 + SH));  // 16 rows of 2^SH-byte taps
  g.rowstep = (uint32_t)g.P << (2 + SH);                                   // 4 rows
  return g;
}

// Per lane and column block j, the byte offset of the lane's two taps relative to soff0 (tile
// columns x, x+1 with x = 32j + 2lu).  A lane with no tap in the footprint gets an offset beyond
// the buffer's num_records (which stays below 2 GiB) so that its load returns zeros; the other
// lanes read inside the padded row (the partner of an edge tap lies in the guard band).
constexpr uint32_t kOutOfRange = 0x80000000u;
template <int SH>
__device__ __forceinline__ void lane_offsets(uint32_t (&vo)[2], const PairGeom& g, int lu, int lv) {
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int x = 32 * j + 2 * lu;
    const bool in = (x >= g.A) && (x <= g.B);
    vo[j] = in ? (uint32_t)(lv * g.P + x) << SH : kOutOfRange;
  }
}

// The CF taps of the active blocks: one 16-byte buffer load per lane and block, with the row in
// the scalar offset and the lane's position in the vector offset.  COEF: also the sample's MFMA
// B operands (the 16 bytes of the staged coefficients for the lane's output column o).
// CFP: storage precision of the table (FP32: 16 bytes per lane and block; FP16 / BF16: 8 bytes,
// widened to FP32 in registers).
template <int CFP>
constexpr int cf_shift() {
  return CFP == kPrecF32 ? 3 : 2;
}

template <bool COEF, int CFP, class SHP>
__device__ __forceinline__ void load_cf(Taps<CFP>& c, __amdgpu_buffer_rsrc_t cfr, const PairGeom& g, int lu, int lv,
                                        const f32x4* s_coef_pair, int o, SHP shp) {
  if constexpr (COEF) c.b = s_coef_pair[o];
  uint32_t vo[2];
  lane_offsets<cf_shift<CFP>()>(vo, g, lu, lv);
  for_blocks(shp, [&](auto kc, auto jc) {
        constexpr int k = decltype(kc)::value, j = decltype(jc)::value;
        // dword loads, merged into one dwordx4 (dwordx2) by the compiler (the vector-valued
        // buffer load builtins of this toolchain load only the first dword)
        const uint32_t so = g.soff0 + k * g.rowstep;
        if constexpr (CFP == kPrecF32) {
          uint32_t d[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) d[i] = __builtin_amdgcn_raw_buffer_load_b32(cfr, vo[j] + 4u * i, so, 0);
          c.t[k][j][0] = make_float2(__builtin_bit_cast(float, d[0]), __builtin_bit_cast(float, d[1]));
          c.t[k][j][1] = make_float2(__builtin_bit_cast(float, d[2]), __builtin_bit_cast(float, d[3]));
        } else {
#pragma unroll
          for (int e = 0; e < 2; ++e) c.t[k][j][e] = __builtin_amdgcn_raw_buffer_load_b32(cfr, vo[j] + 4u * e, so, 0);
        }
  });
}

// acc += v * C (or v * conj(C) in the PCF half plane, TableVisGridder.cc:837-844) for the
// active blocks, on the matrix cores: per block, tap e and output set m (polarisations 2m, 2m+1)
// two 4x4x1 MFMAs (16 independent 4x4 outer products per wave) add C.re (x) B[2m] and
// C.im (x) B[2m+1].  MFMA block b = lane/4 holds the taps of its lanes 4b..4b+3 (4 pixels) as A
// and the sample's coefficients for outputs (pol 2m re, pol 2m im, pol 2m+1 re, pol 2m+1 im) as B;
// the result for the pixel of lane 4b+i and output o lands in lane 4b+o, register i.
template <int NP, class T, class SHP>
__device__ __forceinline__ void grid_pair(f32x4 (&acc)[4][2][2][(NP + 1) / 2], const T& c, SHP shp) {
  for_blocks(shp, [&](auto kc, auto jc) {
        constexpr int k = decltype(kc)::value, j = decltype(jc)::value;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const float2 x = c.tap(k, j, e);
          acc[k][j][e][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(x.x, c.b[0], acc[k][j][e][0], 0, 0, 0);
          acc[k][j][e][0] = __builtin_amdgcn_mfma_f32_4x4x1f32(x.y, c.b[1], acc[k][j][e][0], 0, 0, 0);
          if constexpr (NP > 2) {
            acc[k][j][e][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(x.x, c.b[2], acc[k][j][e][1], 0, 0, 0);
            acc[k][j][e][1] = __builtin_amdgcn_mfma_f32_4x4x1f32(x.y, c.b[3], acc[k][j][e][1], 0, 0, 0);
          }
        }
  });
}


__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_cf_rsrc(const CFRef& t) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(reinterpret_cast<const char*>(t.ptr)) - t.pad, (short)0,
                                           (int)(t.bytes + t.pad), kBufferWord3);
}

// The CF resource of a work item of key group kz: the whole table, or (tables of 2 GiB or more)
// a resource based at the group's first plane that the descriptors' offsets are relative to.
// The group base is read through a generic pointer, a load the compiler takes as per-lane: the
// resource is rebuilt from readfirstlane'd words so that it stays in scalar registers (a
// per-lane resource makes every buffer load a waterfall loop and blocks their merging into
// 16-byte loads).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_item_rsrc(const TileJob& J, int kz) {
  const CFRef& t = J.cf;
  if (!t.zg_base) return make_cf_rsrc(t);
  const unsigned long long b = (unsigned long long)t.zg_base[kz % J.nzg] * t.esz;
  const unsigned long long rec = min(t.total - b + t.pad, kCFResourceMax);
  const unsigned long long base = reinterpret_cast<unsigned long long>(t.ptr) + b - t.pad;
  const unsigned long long ub =
      ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(base >> 32)) << 32) |
      (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)base);
  const int ur = __builtin_amdgcn_readfirstlane((int)rec);
  return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<char*>(ub), (short)0, ur, kBufferWord3);
}

// Work queues of the persistent tile kernels: the key-ordered items (w-plane-major) split into
// kItemQueues contiguous ranges, one per XCD, so that the workgroups of one XCD work through
// neighbouring keys and the CF planes they read stay in that XCD's L2.  A workgroup's XCD is
// taken as blockIdx.x % kItemQueues (the dispatcher places consecutive workgroups on the XCDs in
// turn; were it otherwise, only locality would suffer).  A workgroup takes items from its own
// range and, once that is exhausted, from the following ones; returns -1 when all are.

// The staged pairs whose footprint touches wave w's rows, as indices into the round's staging
// arrays: the wave then skips the pairs it has nothing to add for.  Returns their number.
__device__ __forceinline__ int build_wave_list(const int4* s_desc, int n, int w, int lane, uint16_t* list) {
  int cnt = 0;
  for (int b = 0; b < n; b += 64) {
    const int t = b + lane;
    const bool act = t < n && ((((uint32_t)s_desc[t].z) >> (8 * w)) & 0xffu) != 0u;
    const unsigned long long bal = __ballot(act);
    const int before = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
    if (act) list[cnt + before] = (uint16_t)t;
    cnt += __popcll(bal);
  }
  return __builtin_amdgcn_readfirstlane(cnt);
}

__device__ __forceinline__ PairU make_uniform(const int4& d) {
  PairU p;
  p.d = make_int4(__builtin_amdgcn_readfirstlane(d.x), __builtin_amdgcn_readfirstlane(d.y),
                  __builtin_amdgcn_readfirstlane(d.z), __builtin_amdgcn_readfirstlane(d.w));
  return p;
}

// Runs consume(c, geom, pair, i) for n staged pairs of a work item in wave w, with the CF taps
// (and, COEF, the MFMA B operands) of pair i in c.
// COMPACT: the pairs are list[0 .. n) (build_wave_list), else the staged pairs 0 .. n.  A pair's
// loads are issued for its active blocks.  The descriptor of pair i+1 is read
// from LDS before pair i is consumed and made wave-uniform after it, and the list entry one pair
// further ahead, so that no LDS latency stands between the waits and the products.  PIPE: two
// register sets; the loads of pair i+1 are issued before pair i is consumed.  consume must not
// issue vector memory loads whose results it waits for after returning (the wait before the
// next loads covers them).
template <bool COEF, bool PIPE, int CFP, bool COMPACT, class F>
__device__ __forceinline__ void for_each_pair(const int4* s_desc, const f32x4* s_coef, const uint16_t* list,
                                              __amdgpu_buffer_rsrc_t cfr, int n, int w, int lu, int lv, int o,
                                              F&& consume) {
  constexpr int SH = cf_shift<CFP>();
  if (n <= 0) return;
  constexpr int CW = kCoef / 4;  // f32x4 per staged pair
  auto entry = [&](int i) -> int {
    if constexpr (COMPACT) return list[i];
    else return i;
  };
  auto issue = [&](Taps<CFP>& c, const PairGeom& g, int ix) {
    load_cf<COEF, CFP>(c, cfr, g, lu, lv, COEF ? s_coef + CW * ix : s_coef, o, MaskShape{g.mask});
  };
  if constexpr (!PIPE) {
    int ixn = entry(0);
    int4 rn = s_desc[ixn];
    int ixq = n > 1 ? entry(1) : 0;
    for (int i = 0; i < n; ++i) {
      // Nothing is in flight here (the previous pair consumed its taps); saying so keeps the
      // wait-count analysis, which cannot correlate the loads of a shape with their uses, from
      // waiting for all loads before each load that reuses a register.
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
      const PairU p = make_uniform(rn);
      const int ix = ixn;
      if (i + 1 < n) {
        ixn = ixq;
        rn = s_desc[ixn];
      }
      if (i + 2 < n) ixq = entry(i + 2);
      const PairGeom g = pair_geom<SH>(p, w);
      Taps<CFP> c;
      issue(c, g, ix);
      consume(c, g, p, i);
    }
  } else {
    int i = 0;
    Taps<CFP> cA, cB;
    const int ix0 = entry(0);
    PairU pA = make_uniform(s_desc[ix0]);
    PairGeom gA = pair_geom<SH>(pA, w);
    issue(cA, gA, ix0);
    // (ixn, rn): list entry and raw descriptor of the next pair; ixq: entry of the one after
    int ixn = ix0, ixq = 0;
    int4 rn = make_int4(0, 0, 0, 0);
    if (n > 1) {
      ixn = entry(1);
      rn = s_desc[ixn];
    }
    if (n > 2) ixq = entry(2);
    PairU pB = pA;
    PairGeom gB = gA;
    for (;;) {
      // pair i: taps in flight into cA
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
      if (i + 1 < n) {
        pB = make_uniform(rn);
        gB = pair_geom<SH>(pB, w);
        issue(cB, gB, ixn);
      }
      if (i + 2 < n) {
        ixn = ixq;
        rn = s_desc[ixn];
      }
      if (i + 3 < n) ixq = entry(i + 3);
      __builtin_amdgcn_sched_barrier(0);
      consume(cA, gA, pA, i);
      __builtin_amdgcn_sched_barrier(0);
      if (++i >= n) break;
      // pair i: taps in flight into cB
      __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
      if (i + 1 < n) {
        pA = make_uniform(rn);
        gA = pair_geom<SH>(pA, w);
        issue(cA, gA, ixn);
      }
      if (i + 2 < n) {
        ixn = ixq;
        rn = s_desc[ixn];
      }
      if (i + 3 < n) ixq = entry(i + 3);
      __builtin_amdgcn_sched_barrier(0);
      consume(cB, gB, pB, i);
      __builtin_amdgcn_sched_barrier(0);
      if (++i >= n) break;
    }
  }
}

template <int NP, int WPE, bool PIPE, int CFP>
__global__ __launch_bounds__(256, WPE) void k_grid(const int4* __restrict__ items, const int* __restrict__ nitems_p,
                                              int* __restrict__ counter, const TileJob* __restrict__ jobs) {
  __shared__ int s_item;
  __shared__ int4 s_desc[kChunk];
  __shared__ f32x4 s_coef[kChunk * (kCoef / 4)];
  __shared__ uint16_t s_list[4][kChunk];  // per wave: the staged pairs that touch its rows
  static_assert(4 * kFlushWave <= kChunk * kCoef, "flush areas must fit in the staging space");
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  // Lane (lu, lv) = (lane & 15, lane >> 4) loads the taps of block row lv, columns 2lu, 2lu+1 (so
  // that the four lanes of a quad read 64 contiguous bytes: one L1 tag lookup).  The MFMA block
  // b = lane/4 then holds the four column pairs 4(b&3)..4(b&3)+3 of row b>>2; after the products
  // lane 4b+o holds output o = lane & 3 of those four column pairs (register i: column pair
  // 4(b&3)+i).
  const int lu = lane & 15, lv = lane >> 4;
  const int o = lane & 3;
  const int ob = (lane >> 2) & 3;
  constexpr int NS = (NP + 1) / 2;  // output sets of two polarisations
  const int nitems = *nitems_p;
  int queue = blockIdx.x % kItemQueues, queues_left = kItemQueues;
  for (;;) {
    const int item = next_item(counter, nitems, queue, queues_left, &s_item);
    if (item < 0) break;
    const int4 it = items[item];
    const TileJob& J = jobs[it.w];
    const int nx = J.nx, ny = J.ny;
    const int tile = it.x % J.ntiles;
    const int kz = it.x / J.ntiles;
    const __amdgpu_buffer_rsrc_t cfr = make_item_rsrc(J, kz);
    const int ichan = kz / J.nzg;
    const int u0 = (tile % J.ntu) * kTile, v0 = (tile / J.ntu) * kTile;
    float* __restrict__ grid = J.grid + 2 * J.wstride * (size_t)(kz % J.nzg);

    f32x4 acc[4][2][2][NS];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int e = 0; e < 2; ++e)
#pragma unroll
          for (int m = 0; m < NS; ++m) acc[k][j][e][m] = f32x4{0.f, 0.f, 0.f, 0.f};

    // the item's pairs in staging rounds of <= kChunk (the tile stays in registers throughout)
    const int n = it.z - it.y;
    for (int cb = 0; cb < n; cb += kChunk) {
      const int m = min(kChunk, n - cb);
      if (cb > 0) __syncthreads();  // every wave is done with the previous round's staging
      stage_item<true>(J.descs, J.smp, it.y + cb, m, s_desc, s_coef, tid, blockDim.x);
      const int mw = build_wave_list(s_desc, m, w, lane, s_list[w]);
      for_each_pair<true, PIPE, CFP, true>(s_desc, s_coef, s_list[w], cfr, mw, w, lu, lv, o,
                                           [&](const auto& c, const PairGeom& g, const PairU&, int) {
                                             grid_pair<NP>(acc, c, MaskShape{g.mask});
                                           });
    }
    // flush: lane 4b+o holds output o (pol 2m + o/2, re/im o&1) of row 4k+lv, column
    // 32j+8(b&3)+2i+e.  Per (m, k) the wave writes its 4 rows x 64 columns x 4 outputs to its own
    // LDS area (the coefficient staging space, free until the next item is staged) and reads them
    // back transposed, so that each float atomic instruction covers one row's 16 consecutive
    // pixels of two polarisations (two cache lines).  One float of padding per 32 spreads banks.
    // The barrier keeps a wave from overwriting coefficients the other waves still read.
    __syncthreads();
    float* s_fl = reinterpret_cast<float*>(s_coef) + w * kFlushWave;
    const size_t plane = (size_t)nx * ny;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int e = 0; e < 2; ++e)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              const int idx = ((lv * kTile) + 32 * j + 8 * ob + 2 * i + e) * 4 + o;
              s_fl[idx + (idx >> 5)] = acc[k][j][e][m][i];
            }
        __builtin_amdgcn_wave_barrier();
        const int q = 2 * m + (o >> 1);
        if (q < NP) {
          float* gq = grid + 2 * (size_t)(ichan * NP + q) * plane + (o & 1);
#pragma unroll
          for (int t = 0; t < 16; ++t) {
            const int idx = t * 64 + lane;
            const float val = s_fl[idx + (idx >> 5)];
            const int v = v0 + 16 * w + 4 * k + (t >> 2);
            const int u = u0 + 16 * (t & 3) + (lane >> 2);
            if (val != 0.f && u < nx && v < ny) atomicAdd(gq + 2 * ((size_t)v * nx + u), val);
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
}

// --------------------------------------------------------------------------------------
// degrid
// --------------------------------------------------------------------------------------
constexpr int kDegridBatch = 32;

template <int CTRL>
__device__ __forceinline__ float dpp(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL, 0xf, 0xf, true));
}

// Reduce-scatter of 8 values over the 64 lanes of a wave: on return, lane l holds the sum over
// all lanes of x[bitrev3(l & 7)] (bitrev3 reverses the three low bits).  Exchanges: lane l^1 and
// l^2 by DPP quad permutes (halving the values each time), l^4 by row shifts, l^8 by a row
// rotate, l^16 and l^32 by the gfx950 permlane swaps.
__device__ __forceinline__ float wave_reduce_scatter8(const float (&x)[8], int lane) {
  const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4;
  float y[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const float send = b0 ? x[i] : x[i + 4];
    const float keep = b0 ? x[i + 4] : x[i];
    y[i] = keep + dpp<0xB1>(send);  // quad_perm [1,0,3,2]: lane l^1
  }
  float z[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const float send = b1 ? y[i] : y[i + 2];
    const float keep = b1 ? y[i + 2] : y[i];
    z[i] = keep + dpp<0x4E>(send);  // quad_perm [2,3,0,1]: lane l^2
  }
  const float send = b2 ? z[0] : z[1];
  const float keep = b2 ? z[1] : z[0];
  const float from_up = dpp<0x104>(send);  // row_shl:4: lane l+4
  const float from_dn = dpp<0x114>(send);  // row_shr:4: lane l-4
  float r = keep + (b2 ? from_dn : from_up);
  r += dpp<0x128>(r);  // row_ror:8: lane l^8 within the row of 16
  const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(r), __float_as_uint(r), false, false);
  r = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);  // rows 2m and 2m+1 exchanged: lane l^16
  const auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(r), __float_as_uint(r), false, false);
  r = __uint_as_float(s32[0]) + __uint_as_float(s32[1]);  // halves exchanged: lane l^32
  return r;
}

// sum over the wave's active blocks of C * conj(G) per polarisation (GridKernel.cc:121-122);
// x[2q] / x[2q+1] = real / imaginary part for polarisation q, x[2NP..7] = 0
template <int NP, class T, class SHP>
__device__ __forceinline__ void degrid_pair(float (&x)[8], const T& c, const float2 (&gt)[4][2][2][NP], SHP shp) {
#pragma unroll
  for (int i = 0; i < 8; ++i) x[i] = 0.f;
  for_blocks(shp, [&](auto kc, auto jc) {
        constexpr int k = decltype(kc)::value, j = decltype(jc)::value;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const float2 cc = c.tap(k, j, e);
#pragma unroll
          for (int q = 0; q < NP; ++q) {
            x[2 * q] = fmaf(cc.x, gt[k][j][e][q].x, x[2 * q]);
            x[2 * q] = fmaf(cc.y, gt[k][j][e][q].y, x[2 * q]);
            x[2 * q + 1] = fmaf(cc.y, gt[k][j][e][q].x, x[2 * q + 1]);
            x[2 * q + 1] = fmaf(-cc.x, gt[k][j][e][q].y, x[2 * q + 1]);
          }
        }
  });
}

template <int NP, int WPE, bool PIPE, int CFP>
__global__ __launch_bounds__(256, WPE) void k_degrid(const int4* __restrict__ items, const int* __restrict__ nitems_p,
                                                int* __restrict__ counter, const TileJob* __restrict__ jobs) {
  __shared__ int s_item;
  __shared__ int4 s_desc[kChunk];
  __shared__ float s_part[kDegridBatch][4][8];
  __shared__ int s_sid[kDegridBatch];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lu = lane & 15, lv = lane >> 4;
  const int slot = ((lane & 1) << 2) | (lane & 2) | ((lane >> 2) & 1);  // value held after the reduce-scatter
  const int nitems = *nitems_p;
  int queue = blockIdx.x % kItemQueues, queues_left = kItemQueues;
  for (;;) {
    const int item = next_item(counter, nitems, queue, queues_left, &s_item);
    if (item < 0) break;
    const int4 it = items[item];
    const TileJob& J = jobs[it.w];
    const int nx = J.nx, ny = J.ny;
    const int tile = it.x % J.ntiles;
    const int kz = it.x / J.ntiles;
    const __amdgpu_buffer_rsrc_t cfr = make_item_rsrc(J, kz);
    const int ichan = kz / J.nzg;
    const int u0 = (tile % J.ntu) * kTile, v0 = (tile / J.ntu) * kTile;
    const size_t plane = (size_t)nx * ny;
    const float2* __restrict__ grid = reinterpret_cast<const float2*>(J.grid) + J.wstride * (size_t)(kz % J.nzg);
    float* __restrict__ acc_out = J.acc;

    float2 gt[4][2][2][NP];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int v = v0 + 16 * w + 4 * k + lv;
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int u = u0 + 32 * j + 2 * lu + e;
#pragma unroll
          for (int q = 0; q < NP; ++q)
            gt[k][j][e][q] = (u < nx && v < ny) ? grid[(size_t)(ichan * NP + q) * plane + (size_t)v * nx + u]
                                                : make_float2(0.f, 0.f);
        }
    }

    // per pair: wave dot product, reduce-scatter, one LDS slot per (pair, wave, value); every
    // kDegridBatch pairs (and at the end of each staging round of <= kChunk pairs) the four
    // waves' partial sums meet and go to the per-sample accumulators
    const int nall = it.z - it.y;
    for (int cb = 0; cb < nall; cb += kChunk) {
    const int n = min(kChunk, nall - cb);
    if (cb > 0) __syncthreads();  // every wave is done with the previous round's staging
    stage_item<false>(J.descs, Samples{}, it.y + cb, n, s_desc, nullptr, tid, blockDim.x);
    for_each_pair<false, PIPE, CFP, false>(s_desc, nullptr, nullptr, cfr, n, w, lu, lv, 0,
                           [&](const auto& c, const PairGeom& g, const PairU& p, int i) {
                             // a wave whose rows the footprint misses (mask 0) contributes zeros
                             // without the dot product and the reduction
                             float r = 0.f;
                             if (g.mask) {
                               float x[8];
                               degrid_pair<NP>(x, c, gt, MaskShape{g.mask});
                               r = wave_reduce_scatter8(x, lane);
                             }
                             const int b = i % kDegridBatch;
                             if (lane < 8) s_part[b][w][slot] = r;
                             if (tid == 0) s_sid[b] = p.d.x;
                             if (b == kDegridBatch - 1 || i == n - 1) {
                               __syncthreads();
                               const int nb = b + 1;
                               for (int t = tid; t < nb * 2 * NP; t += blockDim.x) {
                                 const int bb = t / (2 * NP), q = t % (2 * NP);
                                 const float sum =
                                     s_part[bb][0][q] + s_part[bb][1][q] + s_part[bb][2][q] + s_part[bb][3][q];
                                 if (sum != 0.f) atomicAdd(&acc_out[(size_t)s_sid[bb] * 8 + q], sum);
                               }
                               __syncthreads();
                             }
                           });
    }
  }
}

// ---- Degridding on the matrix cores: four pairs at once ------------------------------------
//
// Lane l = 4b + j of a wave: MFMA block b covers the tile columns 32h + 2b + e (e = 0, 1) of one
// row of the wave's 16, j selects one pair of the group of four (the A operand: that pair's CF
// tap at the pixel) and one output (the B operand: the model-grid value at the pixel; j = 0, 1:
// re, im of polarisation 2m; j = 2, 3: of 2m + 1).  One v_mfma_f32_4x4x1_16b_f32 then forms
// for 16 pixels every product of the four pairs' taps with the four outputs, all of them
// needed: per block, D1[m][i][j] accumulates C.re(pair i) * B[j] and D2[m][i][j] C.im(pair i)
// * B[j] over the wave's pixels, and re/im of C * conj(G) (GridKernel.cc:121-122) follow as
// D1[j] + D2[j^1] (j even) and D2[j^1] - D1[j] (j odd).  A lane loads its pair's two taps of
// one row and column pair (16 bytes for FP32 tables) where that pair's footprint covers the
// block, zeros elsewhere; a (row block, half) that none of the four pairs covers is skipped.
// (Loading coalesced instead, 16 lanes per pair, and moving the taps to the MFMA lanes with
// ds_bpermute was measured slower: the permutes cost more than the coalescing saves.)  The
// four pairs of a group are neighbours in binning order, which follows the sample order within
// a bin (consecutive rows of one channel), so their footprints mostly cover the same blocks;
// sorting each staging round by footprint in LDS was measured slower than this order.

// Geometry of group g0 (pairs g0 .. g0 + 3) for the loads of lane (c, s) = (lane >> 2,
// lane & 3): the lane's pair is g0 + s, its column pair 2c, 2c + 1 of each half.
struct QuadGeom {
  uint32_t base;  // byte offset of the lane's pair's tap under the wave's first pixel
  uint32_t rowb;  // bytes per CF row of that pair's plane
  uint32_t mask;  // that pair's blocks (k, h) of this wave, bit 2k + h
  uint32_t um;    // union of the four pairs' masks (wave-uniform)
  int sid;        // that pair's sample
  bool c0, c1;    // the lane's column pair (2c, 2c + 1 of half 0, 1) touches that footprint
};

template <int SH>
__device__ __forceinline__ QuadGeom quad_geom(const int4* s_desc, int g0, int n, int w, int s, int c) {
  QuadGeom q;
  const bool valid = g0 + s < n;
  const int4 d = s_desc[valid ? g0 + s : g0];
  const uint32_t pw = (uint32_t)d.w;
  const int P = (int)(pw & 0xffffu), A = (int)(pw << 8) >> 24, B = (int)pw >> 24;
  q.mask = valid ? ((uint32_t)d.z >> (8 * w)) & 0xffu : 0u;
  q.base = (uint32_t)d.y + (uint32_t)w * ((uint32_t)P << (4 + SH));  // 16 rows per wave
  q.rowb = (uint32_t)P << SH;
  q.sid = d.x;
  q.c0 = (2 * c >= A) && (2 * c <= B);
  q.c1 = (32 + 2 * c >= A) && (32 + 2 * c <= B);
  q.um = (uint32_t)__builtin_amdgcn_readlane((int)q.mask, 0) | (uint32_t)__builtin_amdgcn_readlane((int)q.mask, 1) |
         (uint32_t)__builtin_amdgcn_readlane((int)q.mask, 2) | (uint32_t)__builtin_amdgcn_readlane((int)q.mask, 3);
  return q;
}

template <int CFP>
constexpr int quad_dwords() {
  return CFP == kPrecF32 ? 4 : 2;
}

// The taps of row block k (rows 4k .. 4k + 3 of the wave, both halves) of the lane's pair and
// column pair c.
template <int CFP>
__device__ __forceinline__ void load_quad(uint32_t (&t)[4][2][quad_dwords<CFP>()], __amdgpu_buffer_rsrc_t cfr,
                                          const QuadGeom& q, int k, int c) {
  constexpr int SH = cf_shift<CFP>();
#pragma unroll
  for (int h = 0; h < 2; ++h)
    if (q.um & (1u << (2 * k + h))) {
      const bool on = ((q.mask >> (2 * k + h)) & 1u) && (h ? q.c1 : q.c0);
      const uint32_t col = (uint32_t)(32 * h + 2 * c) << SH;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t vo = on ? q.base + (uint32_t)(4 * k + r) * q.rowb + col : kOutOfRange;
#pragma unroll
        for (int i = 0; i < quad_dwords<CFP>(); ++i)
          t[r][h][i] = __builtin_amdgcn_raw_buffer_load_b32(cfr, vo + 4u * i, 0, 0);
      }
    }
}

template <int NP, int CFP>
__device__ __forceinline__ void mfma_quad(f32x4 (&D1)[(NP + 1) / 2], f32x4 (&D2)[(NP + 1) / 2],
                                          const uint32_t (&t)[4][2][quad_dwords<CFP>()],
                                          const float (&gt)[16][2][2][(NP + 1) / 2], uint32_t um, int k) {
#pragma unroll
  for (int h = 0; h < 2; ++h)
    if (um & (1u << (2 * k + h))) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          float cr, ci;
          if constexpr (CFP == kPrecF32) {
            cr = __builtin_bit_cast(float, t[r][h][2 * e]);
            ci = __builtin_bit_cast(float, t[r][h][2 * e + 1]);
          } else {
            cr = widen16(t[r][h][e] & 0xffffu, CFP);
            ci = widen16(t[r][h][e] >> 16, CFP);
          }
#pragma unroll
          for (int m = 0; m < (NP + 1) / 2; ++m) {
            D1[m] = __builtin_amdgcn_mfma_f32_4x4x1f32(cr, gt[4 * k + r][h][e][m], D1[m], 0, 0, 0);
            D2[m] = __builtin_amdgcn_mfma_f32_4x4x1f32(ci, gt[4 * k + r][h][e][m], D2[m], 0, 0, 0);
          }
        }
    }
}

template <int NP, int CFP>
__global__ __launch_bounds__(256, 2) void k_degrid_mm(const int4* __restrict__ items, const int* __restrict__ nitems_p,
                                                    int* __restrict__ counter, const TileJob* __restrict__ jobs) {
  constexpr int SH = cf_shift<CFP>();
  constexpr int M = (NP + 1) / 2;
  constexpr int ND = quad_dwords<CFP>();
  __shared__ int s_item;
  __shared__ int4 s_desc[kChunk];
  __shared__ float s_part[kDegridBatch][4][8];
  __shared__ int s_sid[kDegridBatch];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = lane >> 2, j = lane & 3;  // MFMA block (column pair of the loads), pair / output
  const float sgn = (j & 1) ? -1.f : 1.f;
  const int nitems = *nitems_p;
  int queue = blockIdx.x % kItemQueues, queues_left = kItemQueues;
  for (;;) {
    const int item = next_item(counter, nitems, queue, queues_left, &s_item);
    if (item < 0) break;
    const int4 it = items[item];
    const TileJob& J = jobs[it.w];
    const int nx = J.nx, ny = J.ny;
    const int tile = it.x % J.ntiles;
    const int kz = it.x / J.ntiles;
    const __amdgpu_buffer_rsrc_t cfr = make_item_rsrc(J, kz);
    const int ichan = kz / J.nzg;
    const int u0 = (tile % J.ntu) * kTile, v0 = (tile / J.ntu) * kTile;
    const size_t plane = (size_t)nx * ny;
    const float* __restrict__ gridf = J.grid + 2 * J.wstride * (size_t)(kz % J.nzg);
    float* __restrict__ acc_out = J.acc;

    // model-grid tile: gt[r][h][e][m] = output j of polarisation pair m at tile pixel
    // (16w + r, 32h + 2b + e)
    float gt[16][2][2][M];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int v = v0 + 16 * w + r;
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int u = u0 + 32 * h + 2 * b + e;
#pragma unroll
          for (int m = 0; m < M; ++m) {
            const int q = 2 * m + (j >> 1);
            gt[r][h][e][m] = (u < nx && v < ny && q < NP)
                                 ? gridf[2 * ((size_t)(ichan * NP + q) * plane + (size_t)v * nx + u) + (j & 1)]
                                 : 0.f;
          }
        }
    }

    const int nall = it.z - it.y;
    for (int cb = 0; cb < nall; cb += kChunk) {
      const int n = min(kChunk, nall - cb);
      if (cb > 0) __syncthreads();  // every wave is done with the previous round's staging
      stage_item<false>(J.descs, Samples{}, it.y + cb, n, s_desc, nullptr, tid, blockDim.x);
      uint32_t tA[4][2][ND], tB[4][2][ND];
      QuadGeom lq = quad_geom<SH>(s_desc, 0, n, w, j, b);
      load_quad<CFP>(tA, cfr, lq, 0, b);
      for (int g0 = 0; g0 < n; g0 += 4) {
        const QuadGeom cq = lq;  // this group; the taps of its row block 0 are in flight into tA
        f32x4 D1[M], D2[M];
#pragma unroll
        for (int m = 0; m < M; ++m) {
          D1[m] = f32x4{0.f, 0.f, 0.f, 0.f};
          D2[m] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        load_quad<CFP>(tB, cfr, cq, 1, b);
        mfma_quad<NP, CFP>(D1, D2, tA, gt, cq.um, 0);
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        load_quad<CFP>(tA, cfr, cq, 2, b);
        mfma_quad<NP, CFP>(D1, D2, tB, gt, cq.um, 1);
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        load_quad<CFP>(tB, cfr, cq, 3, b);
        mfma_quad<NP, CFP>(D1, D2, tA, gt, cq.um, 2);
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        if (g0 + 4 < n) {
          lq = quad_geom<SH>(s_desc, g0 + 4, n, w, j, b);
          load_quad<CFP>(tA, cfr, lq, 0, b);
        }
        mfma_quad<NP, CFP>(D1, D2, tB, gt, cq.um, 3);

        // lane (b, j), value 4m + i: re/im (j even/odd) of pair g0 + i, polarisation
        // 2m + (j >> 1), summed over block b; then summed over the 16 blocks
        float val[4 * M];
#pragma unroll
        for (int t = 0; t < 4 * M; ++t) val[t] = 0.f;
        if (cq.um) {
#pragma unroll
          for (int m = 0; m < M; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i) val[4 * m + i] = fmaf(sgn, D1[m][i], dpp<0xB1>(D2[m][i]));  // lane j^1
#pragma unroll
          for (int t = 0; t < 4 * M; ++t) {
            float r = val[t];
            r += dpp<0x124>(r);  // row_ror:4
            r += dpp<0x128>(r);  // row_ror:8: the four blocks of the row
            const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(r), __float_as_uint(r), false, false);
            r = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);
            const auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(r), __float_as_uint(r), false, false);
            val[t] = __uint_as_float(s32[0]) + __uint_as_float(s32[1]);
          }
        }
        const int gn = min(4, n - g0);
        if (lane < 4) {
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (i < gn) {
#pragma unroll
              for (int m = 0; m < M; ++m) s_part[(g0 + i) % kDegridBatch][w][4 * m + j] = val[4 * m + i];
            }
        }
        if (w == 0 && b == 0 && j < gn) s_sid[(g0 + j) % kDegridBatch] = cq.sid;  // lane j: pair j
        const int gend = g0 + gn;
        if (gend % kDegridBatch == 0 || gend == n) {
          __syncthreads();
          const int nb = gend - (g0 / kDegridBatch) * kDegridBatch;
          for (int t = tid; t < nb * 2 * NP; t += blockDim.x) {
            const int bb = t / (2 * NP), q = t % (2 * NP);
            const float sum = s_part[bb][0][q] + s_part[bb][1][q] + s_part[bb][2][q] + s_part[bb][3][q];
            if (sum != 0.f) atomicAdd(&acc_out[(size_t)s_sid[bb] * 8 + q], sum);
          }
          __syncthreads();
        }
      }
    }
  }
}

// ---- 16-bit tables on the 16-bit matrix cores --------------------------------------------
//
// v_mfma_f32_4x4x4_16b_f16 (bf16: _4x4x4bf16_1k) takes four 16-bit values per lane as A and as B
// and sums their products in FP32: C_b[i][o] += sum_k A_b[i][k] B_b[k][o] for the 16 blocks b of
// four lanes.  A 16-bit CF tap is one dword (re, im), so the two taps a lane loads (8 bytes) are
// an A operand as they arrive.
typedef _Float16 half4v __attribute__((ext_vector_type(4)));
typedef short short4v __attribute__((ext_vector_type(4)));

template <int CFP>
__device__ __forceinline__ f32x4 mma16(uint32_t a0, uint32_t a1, uint32_t b0, uint32_t b1, f32x4 c) {
  const uint2 a = make_uint2(a0, a1), b = make_uint2(b0, b1);
  if constexpr (CFP == kPrecF16)
    return __builtin_amdgcn_mfma_f32_4x4x4f16(__builtin_bit_cast(half4v, a), __builtin_bit_cast(half4v, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(__builtin_bit_cast(short4v, a), __builtin_bit_cast(short4v, b), c, 0,
                                                 0, 0);
}

// (lo, hi) rounded to nearest even in the 16-bit format prec, as one dword
__device__ __forceinline__ uint32_t pack16(float lo, float hi, int prec) {
  if (prec == kPrecBF16)
    return (uint32_t)__builtin_bit_cast(uint16_t, (__bf16)lo) | ((uint32_t)__builtin_bit_cast(uint16_t, (__bf16)hi) << 16);
  return (uint32_t)round16(lo, prec) | ((uint32_t)round16(hi, prec) << 16);
}

// Degridding with 16-bit taps (the visibilities' side also 16-bit): groups of four pairs as in
// k_degrid_mm, with lane l = 4b + j holding pair j's four taps at the tile columns 4b .. 4b + 3 of
// a row (one 16-byte load; a load instruction covers a whole row of the tile for four pairs).  K =
// (re, im) of two of those taps, so one MFMA forms the complete complex products of four pairs,
// two pixels and four outputs: the B operand of output j at pixels x, x + 1 is
// (G0.re, G0.im, G1.re, G1.im) for the real part of C * conj(G) and (-G0.im, G0.re, -G1.im, G1.re)
// for the imaginary part (GridKernel.cc:121-122).  The model-grid tile is rounded to the table's
// 16-bit format once per item; FP16 scales it first by the power of two the degridder computed
// for its grid (largest magnitude to [2^13, 2^14), exact), which the sums undo.
struct Quad16 {
  uint32_t base;  // byte offset of the lane's pair's tap under the wave's first pixel (column 4b)
  uint32_t rowb;  // bytes per CF row of that pair's plane
  uint32_t mask;  // that pair's row blocks of this wave (bit k), 0 where the lane's taps miss it
  uint32_t um;    // union of the four pairs' row blocks (wave-uniform)
  int sid;
};

__device__ __forceinline__ Quad16 quad16(const int4* s_desc, int g0, int n, int w, int j, int b) {
  Quad16 q;
  const bool valid = g0 + j < n;
  const int4 d = s_desc[valid ? g0 + j : g0];
  const uint32_t pw = (uint32_t)d.w;
  const int P = (int)(pw & 0xffffu), A = (int)(pw << 8) >> 24, B = (int)pw >> 24;
  const uint32_t m8 = valid ? ((uint32_t)d.z >> (8 * w)) & 0xffu : 0u;
  uint32_t m4 = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) m4 |= ((m8 >> (2 * k)) & 3u) ? (1u << k) : 0u;
  q.um = (uint32_t)__builtin_amdgcn_readlane((int)m4, 0) | (uint32_t)__builtin_amdgcn_readlane((int)m4, 1) |
         (uint32_t)__builtin_amdgcn_readlane((int)m4, 2) | (uint32_t)__builtin_amdgcn_readlane((int)m4, 3);
  const int x = 4 * b;
  q.mask = (x >= A - 2 && x <= B) ? m4 : 0u;
  q.base = (uint32_t)d.y + (uint32_t)w * ((uint32_t)P << 6) + ((uint32_t)x << 2);  // 16 rows of 4-byte taps
  q.rowb = (uint32_t)P << 2;
  q.sid = d.x;
  return q;
}

// the four taps of rows 4k .. 4k + 3 of the lane's pair and column quad (zeros where it misses)
__device__ __forceinline__ void load_q16(uint32_t (&t)[4][4], __amdgpu_buffer_rsrc_t cfr, const Quad16& q, int k) {
  if (q.um & (1u << k)) {
    const bool on = (q.mask >> k) & 1u;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const uint32_t vo = on ? q.base + (uint32_t)(4 * k + r) * q.rowb : kOutOfRange;
#pragma unroll
      for (int e = 0; e < 4; ++e) t[r][e] = __builtin_amdgcn_raw_buffer_load_b32(cfr, vo + 4u * e, 0, 0);
    }
  }
}

template <int NP, int CFP>
__global__ __launch_bounds__(256, 2) void k_degrid16(const int4* __restrict__ items, const int* __restrict__ nitems_p,
                                                   int* __restrict__ counter, const TileJob* __restrict__ jobs) {
  constexpr int M = (NP + 1) / 2;
  __shared__ int s_item;
  __shared__ int4 s_desc[kChunk];
  __shared__ float s_part[kDegridBatch][4][8];
  __shared__ int s_sid[kDegridBatch];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int b = lane >> 2, j = lane & 3;
  const int nitems = *nitems_p;
  int queue = blockIdx.x % kItemQueues, queues_left = kItemQueues;
  for (;;) {
    const int item = next_item(counter, nitems, queue, queues_left, &s_item);
    if (item < 0) break;
    const int4 it = items[item];
    const TileJob& J = jobs[it.w];
    const int nx = J.nx, ny = J.ny;
    const int tile = it.x % J.ntiles;
    const int kz = it.x / J.ntiles;
    const __amdgpu_buffer_rsrc_t cfr = make_item_rsrc(J, kz);
    const int ichan = kz / J.nzg;
    const int u0 = (tile % J.ntu) * kTile, v0 = (tile / J.ntu) * kTile;
    const size_t plane = (size_t)nx * ny;
    const float* __restrict__ gridf = J.grid + 2 * J.wstride * (size_t)(kz % J.nzg);
    float* __restrict__ acc_out = J.acc;

    // lane (b, j), B operands gt[r][e][m]: model grid of polarisation 2m + (j >> 1) at the tile
    // pixels (16w + r, 4b + 2e) and (16w + r, 4b + 2e + 1)
    const float scale = (CFP == kPrecF16 && J.gscale) ? *J.gscale : 1.f;
    const float inv_scale = 1.f / scale;  // a power of two: exact
    uint2 gt[16][2][M];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int v = v0 + 16 * w + r;
#pragma unroll
      for (int e = 0; e < 2; ++e)
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const int q = 2 * m + (j >> 1);
          float2 g2[2];
#pragma unroll
          for (int f = 0; f < 2; ++f) {
            const int u = u0 + 4 * b + 2 * e + f;
            g2[f] = (u < nx && v < ny && q < NP)
                        ? *reinterpret_cast<const float2*>(gridf +
                                                           2 * ((size_t)(ichan * NP + q) * plane + (size_t)v * nx + u))
                        : make_float2(0.f, 0.f);
            g2[f].x *= scale;
            g2[f].y *= scale;
          }
          if (j & 1)
            gt[r][e][m] = make_uint2(pack16(-g2[0].y, g2[0].x, CFP), pack16(-g2[1].y, g2[1].x, CFP));
          else
            gt[r][e][m] = make_uint2(pack16(g2[0].x, g2[0].y, CFP), pack16(g2[1].x, g2[1].y, CFP));
        }
    }

    const int nall = it.z - it.y;
    for (int cb = 0; cb < nall; cb += kChunk) {
      const int n = min(kChunk, nall - cb);
      if (cb > 0) __syncthreads();  // every wave is done with the previous round's staging
      stage_item<false>(J.descs, Samples{}, it.y + cb, n, s_desc, nullptr, tid, blockDim.x);
      uint32_t tA[4][4], tB[4][4];
      Quad16 lq = quad16(s_desc, 0, n, w, j, b);
      load_q16(tA, cfr, lq, 0);
      for (int g0 = 0; g0 < n; g0 += 4) {
        const Quad16 cq = lq;
        f32x4 D[M];
#pragma unroll
        for (int m = 0; m < M; ++m) D[m] = f32x4{0.f, 0.f, 0.f, 0.f};
        auto mfma_rows = [&](const uint32_t (&t)[4][4], int k) {
          if (cq.um & (1u << k)) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
              for (int e = 0; e < 2; ++e)
#pragma unroll
                for (int m = 0; m < M; ++m)
                  D[m] = mma16<CFP>(t[r][2 * e], t[r][2 * e + 1], gt[4 * k + r][e][m].x, gt[4 * k + r][e][m].y, D[m]);
          }
        };
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        load_q16(tB, cfr, cq, 1);
        mfma_rows(tA, 0);
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        load_q16(tA, cfr, cq, 2);
        mfma_rows(tB, 1);
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        load_q16(tB, cfr, cq, 3);
        mfma_rows(tA, 2);
        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
        if (g0 + 4 < n) {
          lq = quad16(s_desc, g0 + 4, n, w, j, b);
          load_q16(tA, cfr, lq, 0);
        }
        mfma_rows(tB, 3);

        // lane (b, j), value 4m + i: output j (re / im of polarisation 2m + (j >> 1)) of pair
        // g0 + i over block b; then summed over the 16 blocks
        float val[4 * M];
#pragma unroll
        for (int t = 0; t < 4 * M; ++t) val[t] = 0.f;
        if (cq.um) {
#pragma unroll
          for (int m = 0; m < M; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i) val[4 * m + i] = D[m][i] * inv_scale;
#pragma unroll
          for (int t = 0; t < 4 * M; ++t) {
            float r = val[t];
            r += dpp<0x124>(r);  // row_ror:4
            r += dpp<0x128>(r);  // row_ror:8: the four blocks of the row
            const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(r), __float_as_uint(r), false, false);
            r = __uint_as_float(s16[0]) + __uint_as_float(s16[1]);
            const auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(r), __float_as_uint(r), false, false);
            val[t] = __uint_as_float(s32[0]) + __uint_as_float(s32[1]);
          }
        }
        const int gn = min(4, n - g0);
        if (lane < 4) {
#pragma unroll
          for (int i = 0; i < 4; ++i)
            if (i < gn) {
#pragma unroll
              for (int m = 0; m < M; ++m) s_part[(g0 + i) % kDegridBatch][w][4 * m + j] = val[4 * m + i];
            }
        }
        if (w == 0 && b == 0 && j < gn) s_sid[(g0 + j) % kDegridBatch] = cq.sid;
        const int gend = g0 + gn;
        if (gend % kDegridBatch == 0 || gend == n) {
          __syncthreads();
          const int nb = gend - (g0 / kDegridBatch) * kDegridBatch;
          for (int t = tid; t < nb * 2 * NP; t += blockDim.x) {
            const int bb = t / (2 * NP), q = t % (2 * NP);
            const float sum = s_part[bb][0][q] + s_part[bb][1][q] + s_part[bb][2][q] + s_part[bb][3][q];
            if (sum != 0.f) atomicAdd(&acc_out[(size_t)s_sid[bb] * 8 + q], sum);
          }
          __syncthreads();
        }
      }
    }
  }
}

// Gridding with 16-bit taps and 16-bit gridding coefficients, two pairs per MFMA.  Lane layout:
// wave w owns rows 16w .. 16w + 15 of the 64x64 tile in 4 row blocks k of 4 rows; lane (lu, lv) =
// (lane & 15, lane >> 4) holds the four pixels (row 4k + lv, columns 4lu .. 4lu + 3) of every
// block, whose four taps are one 16-byte load (one load instruction covers a whole 4 x 64 block,
// half the load instructions of two-tap lanes).  For tap e of the lane: K = (re, im of pair p's
// tap, re, im of pair q's tap) and B = (the coefficients that multiply them for output o: two of
// pair p's, two of pair q's), which are a dword of each pair's 16-bit coefficients as staged, so
// one instruction adds both pairs' contributions to 4 pixels x 4 outputs and neither operand
// needs repacking.  Pairs are taken two at a time from the wave's list (the second of an odd
// count is empty); a block only one of the two covers gets zero taps for the other.
struct Geom16 {
  uint32_t mask;  // bit k: row block k of the wave overlaps the footprint
  uint32_t soff0, rowstep;
  uint32_t vo;    // lane offset of its four taps (out of range: none of them in the footprint)
};

__device__ __forceinline__ Geom16 geom16(const int4& raw, int w, int lu, int lv, bool valid) {
  Geom16 g;
  const PairU p = make_uniform(raw);
  const PairGeom pg = pair_geom<2>(p, w);
  uint32_t m4 = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) m4 |= ((pg.mask >> (2 * k)) & 3u) ? (1u << k) : 0u;
  g.mask = valid ? m4 : 0u;
  g.soff0 = pg.soff0;
  g.rowstep = pg.rowstep;
  // taps x .. x + 3 (x = 4 lu) meet the footprint when A - 2 <= x <= B (A, B as for column pairs);
  // the guard bands hold the up to 3 taps beyond either end
  const int x = 4 * lu;
  g.vo = (x >= pg.A - 2 && x <= pg.B) ? (uint32_t)(lv * pg.P + x) << 2 : kOutOfRange;
  return g;
}

// the four taps (dwords) of the lane in the blocks of um, zero where g does not cover the block
__device__ __forceinline__ void load16(uint32_t (&t)[4][4], __amdgpu_buffer_rsrc_t cfr, const Geom16& g, uint32_t um) {
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (um & (1u << k)) {
      if (g.mask & (1u << k)) {
        const uint32_t so = g.soff0 + k * g.rowstep;
#pragma unroll
        for (int e = 0; e < 4; ++e) t[k][e] = __builtin_amdgcn_raw_buffer_load_b32(cfr, g.vo + 4u * e, so, 0);
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) t[k][e] = 0u;
      }
    }
}

template <int NP, int CFP>
__global__ __launch_bounds__(256, 2) void k_grid16(const int4* __restrict__ items, const int* __restrict__ nitems_p,
                                                 int* __restrict__ counter, const TileJob* __restrict__ jobs) {
  __shared__ int s_item;
  __shared__ int4 s_desc[kChunk];
  __shared__ uint2 s_cf16[kChunk * 4];    // per staged pair and output o: its 16-bit coefficients
  __shared__ float s_flush[4 * kFlushWave];
  __shared__ uint16_t s_list[4][kChunk];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lu = lane & 15, lv = lane >> 4;
  const int o = lane & 3;
  const int ob = (lane >> 2) & 3;
  constexpr int NS = (NP + 1) / 2;
  const int nitems = *nitems_p;
  int queue = blockIdx.x % kItemQueues, queues_left = kItemQueues;
  for (;;) {
    const int item = next_item(counter, nitems, queue, queues_left, &s_item);
    if (item < 0) break;
    const int4 it = items[item];
    const TileJob& J = jobs[it.w];
    const int nx = J.nx, ny = J.ny;
    const int tile = it.x % J.ntiles;
    const int kz = it.x / J.ntiles;
    const __amdgpu_buffer_rsrc_t cfr = make_item_rsrc(J, kz);
    const int ichan = kz / J.nzg;
    const int u0 = (tile % J.ntu) * kTile, v0 = (tile / J.ntu) * kTile;
    float* __restrict__ grid = J.grid + 2 * J.wstride * (size_t)(kz % J.nzg);

    // acc[k][e][m]: lane 4b + o, register i: output o (polarisation 2m + o / 2, re / im o & 1) of
    // the pixel (row 4k + lv, column 4 (4 ob + i) + e)
    f32x4 acc[4][4][NS];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int m = 0; m < NS; ++m) acc[k][e][m] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int nall = it.z - it.y;
    for (int cb = 0; cb < nall; cb += kChunk) {
      const int n = min(kChunk, nall - cb);
      if (cb > 0) __syncthreads();
      // staging: descriptors and the samples' 16-bit coefficients (output o: dwords 2o, 2o + 1)
      for (int t = tid; t < n; t += blockDim.x) {
        const int4 d = J.descs[it.y + cb + t];
        s_desc[t] = d;
        const uint4* src = reinterpret_cast<const uint4*>(J.smp.coef16 + (size_t)d.x * kCoef);
        const uint4 c0 = src[0], c1 = src[1];
        s_cf16[4 * t + 0] = make_uint2(c0.x, c0.y);
        s_cf16[4 * t + 1] = make_uint2(c0.z, c0.w);
        s_cf16[4 * t + 2] = make_uint2(c1.x, c1.y);
        s_cf16[4 * t + 3] = make_uint2(c1.z, c1.w);
      }
      __syncthreads();
      const int mw = build_wave_list(s_desc, n, w, lane, s_list[w]);
      auto setup = [&](int i, Geom16& gp, Geom16& gq, uint2& bp, uint2& bq) {
        const int ip = s_list[w][i];
        const bool hq = i + 1 < mw;
        const int iq = hq ? s_list[w][i + 1] : ip;
        gp = geom16(s_desc[ip], w, lu, lv, true);
        gq = geom16(s_desc[iq], w, lu, lv, hq);
        bp = s_cf16[4 * ip + o];
        bq = hq ? s_cf16[4 * iq + o] : make_uint2(0u, 0u);
      };
      auto products = [&](const uint32_t (&ta)[4][4], const uint32_t (&tb)[4][4], uint32_t um, uint2 bp, uint2 bq) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (um & (1u << k)) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              acc[k][e][0] = mma16<CFP>(ta[k][e], tb[k][e], bp.x, bq.x, acc[k][e][0]);
              if constexpr (NP > 2) acc[k][e][1] = mma16<CFP>(ta[k][e], tb[k][e], bp.y, bq.y, acc[k][e][1]);
            }
          }
      };
      // two register sets: the taps of the next two pairs are loaded before these two's products
      if (mw > 0) {
        uint32_t a0[4][4], b0[4][4], a1[4][4], b1[4][4];
        uint32_t um0 = 0, um1 = 0;
        uint2 bp0, bq0, bp1, bq1;
        Geom16 gp, gq;
        setup(0, gp, gq, bp0, bq0);
        um0 = gp.mask | gq.mask;
        load16(a0, cfr, gp, um0);
        load16(b0, cfr, gq, um0);
        for (int i = 0; i < mw; i += 4) {
          __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
          if (i + 2 < mw) {
            setup(i + 2, gp, gq, bp1, bq1);
            um1 = gp.mask | gq.mask;
            load16(a1, cfr, gp, um1);
            load16(b1, cfr, gq, um1);
          }
          __builtin_amdgcn_sched_barrier(0);
          products(a0, b0, um0, bp0, bq0);
          __builtin_amdgcn_sched_barrier(0);
          if (i + 2 >= mw) break;
          __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
          if (i + 4 < mw) {
            setup(i + 4, gp, gq, bp0, bq0);
            um0 = gp.mask | gq.mask;
            load16(a0, cfr, gp, um0);
            load16(b0, cfr, gq, um0);
          }
          __builtin_amdgcn_sched_barrier(0);
          products(a1, b1, um1, bp1, bq1);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    // flush as k_grid: per (m, k) the wave writes its 4 rows x 64 columns x 4 outputs to its LDS
    // area and reads them back so that each float-atomic instruction covers one row's 16
    // consecutive pixels of two polarisations
    float* s_fl = s_flush + w * kFlushWave;
    const size_t plane = (size_t)nx * ny;
#pragma unroll
    for (int m = 0; m < NS; ++m) {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int idx = ((lv * kTile) + 16 * ob + 4 * i + e) * 4 + o;
            s_fl[idx + (idx >> 5)] = acc[k][e][m][i];
          }
        __builtin_amdgcn_wave_barrier();
        const int q = 2 * m + (o >> 1);
        if (q < NP) {
          float* gq = grid + 2 * (size_t)(ichan * NP + q) * plane + (o & 1);
#pragma unroll
          for (int t = 0; t < 16; ++t) {
            const int idx = t * 64 + lane;
            const float val = s_fl[idx + (idx >> 5)];
            const int v = v0 + 16 * w + 4 * k + (t >> 2);
            const int u = u0 + 16 * (t & 3) + (lane >> 2);
            if (val != 0.f && u < nx && v < ny) atomicAdd(gq + 2 * ((size_t)v * nx + u), val);
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
}

__global__ void k_degrid_finish(Samples smp, const float* __restrict__ acc, int nrow, int nchan, int npol,
                                float2* __restrict__ vis_out, int ld) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nrow * nchan) return;
  if (smp.hdr[s].iu0 == INT_MIN) return;
  const int row = s % nrow, chan = s / nrow;
  const float2 ph = *reinterpret_cast<const float2*>(smp.coef + (size_t)s * kCoef);
  const float fr = ph.x, fi = ph.y;
  for (int pol = 0; pol < npol; ++pol) {
    const float cr = acc[(size_t)s * 8 + 2 * pol], ci = acc[(size_t)s * 8 + 2 * pol + 1];
    float2* o = vis_out + row + (size_t)ld * (chan + (size_t)nchan * pol);
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
  const int blocks = (n + 255) / 256;
  // sums of weights in shared memory when they fit in 64 KB; then a few blocks per CU loop over
  // the samples, so that each block adds its sums to the global ones once
  const long long nsw = (long long)p.nsw_rows * p.npol * p.nchan_img;
  // Both forms run a bounded grid looping over the samples: every block ends with global atomics
  // on the pass's few counters (one cache line), which serialise there.
  const bool lds = p.mode != kModeDegrid && p.mode != kModePCF && nsw > 0 && nsw <= 8192;
  const int g = std::min(blocks, 2 * c.n_cu);
  if (lds) {
    hipLaunchKernelGGL(k_prep, dim3(g), dim3(256), (size_t)nsw * sizeof(double), c.stream, p, in, out, 1);
  } else {
    hipLaunchKernelGGL(k_prep, dim3(g), dim3(256), 0, c.stream, p, in, out, 0);
  }
  YG_HIP(hipGetLastError());
}

// Bytes by which the tile kernels' CF buffer resource starts before the table: 16 rows of the
// widest plane and a little (a wave with an active block starts at most 11 rows and 47 columns
// before its plane).
static uint32_t cf_pad_bytes(int max_support, int prec) {
  return (uint32_t)((16 * (2 * (size_t)max_support + 1 + 2 * kGuardU) + 64) * cf_elem_bytes(prec));
}

void launch_binning(const Ctx& c, const Samples& recs, int nrow, int nchan, int nx, int ny, int nkz, const CFView& cf,
                    Binning& b) {
  const int n = nrow * nchan;
  const int ntu = (nx + kTile - 1) / kTile, ntv = (ny + kTile - 1) / kTile;
  const size_t ntiles = (size_t)ntu * ntv;
  const size_t nkeys = ntiles * (size_t)nkz;
  YG_CHECK(nkeys < (size_t)INT_MAX, 1, "too many binning keys");
  b.nkeys = nkeys;
  b.counts.resize((nkeys + 1) * sizeof(int));
  b.offsets.resize((nkeys + 1) * sizeof(int));
  b.cursor.resize((nkeys + 1) * sizeof(int));
  b.nitems.resize(sizeof(int) * 2);
  b.counter.resize(sizeof(int) * kQueueInts);
  // nitems[1] = 0: the item base of this single-job binning
  YG_HIP(hipMemsetAsync(b.nitems.get(), 0, 2 * sizeof(int), c.stream));
  int* counts = b.counts.as<int>();
  int* offsets = b.offsets.as<int>();
  const int ks = bin_key_stride(nkeys);
  if (ks > 1) b.wide.resize((nkeys + 1) * ks * sizeof(int));
  int* cw = ks > 1 ? b.wide.as<int>() : counts;
  const int cgrid = (int)std::min<size_t>((nkeys + 256) / 256, 4 * (size_t)c.n_cu);
  YG_HIP(hipMemsetAsync(cw, 0, (nkeys + 1) * ks * sizeof(int), c.stream));
  if (n > 0) hipLaunchKernelGGL(k_bin_count, dim3((n + 255) / 256), dim3(256), 0, c.stream, recs, nrow, nchan, ntu,
                                (int)ntiles, cw, ks);
  if (ks > 1)
    hipLaunchKernelGGL(k_copy_strided_i32, dim3(cgrid), dim3(256), 0, c.stream, cw, ks, counts, 1, nkeys + 1);
  size_t tbytes = 0;
  YG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tbytes, counts, offsets, (int)(nkeys + 1), c.stream));
  b.temp.resize(tbytes);
  YG_HIP(hipcub::DeviceScan::ExclusiveSum(b.temp.get(), tbytes, counts, offsets, (int)(nkeys + 1), c.stream));
  // the total pair count bounds the pairs buffer: read it back (one small copy per pass)
  int npairs = 0;
  YG_HIP(hipMemcpyAsync(&npairs, offsets + nkeys, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  YG_HIP(hipStreamSynchronize(c.stream));
  YG_CHECK(npairs >= 0, 3, "pair count overflow");
  b.pairs.resize((size_t)(npairs > 0 ? npairs : 1) * sizeof(int4));
  int* cur = ks > 1 ? cw : b.cursor.as<int>();
  if (ks > 1)
    hipLaunchKernelGGL(k_copy_strided_i32, dim3(cgrid), dim3(256), 0, c.stream, offsets, 1, cw, ks, nkeys + 1);
  else
    YG_HIP(hipMemcpyAsync(cur, offsets, (nkeys + 1) * sizeof(int), hipMemcpyDeviceToDevice, c.stream));
  if (n > 0)
    hipLaunchKernelGGL(k_bin_fill, dim3((n + 255) / 256), dim3(256), 0, c.stream, recs, nrow, nchan, ntu, (int)ntiles,
                       cf_pad_bytes(cf.max_support, cf.prec), (uint32_t)cf_elem_bytes(cf.prec), cur, ks,
                       b.pairs.as<int4>(), cf.zg_base, cf.nzg);
  // work items: reuse cursor as the per-key item count and its scan
  int* icount = b.cursor.as<int>();
  const int chunk = c.item_pairs > 0 ? c.item_pairs : kDefaultItemPairs;
  hipLaunchKernelGGL(k_item_count, dim3((nkeys + 1 + 255) / 256), dim3(256), 0, c.stream, counts, (int)nkeys, chunk,
                     icount);
  DevBuf& ioffsets = b.ioffsets;
  ioffsets.resize((nkeys + 1) * sizeof(int));
  size_t tb2 = 0;
  YG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, icount, ioffsets.as<int>(), (int)(nkeys + 1), c.stream));
  b.temp.resize(tb2);
  YG_HIP(hipcub::DeviceScan::ExclusiveSum(b.temp.get(), tb2, icount, ioffsets.as<int>(), (int)(nkeys + 1), c.stream));
  const size_t max_items = (size_t)npairs / chunk + nkeys + 1;
  b.items.resize(max_items * sizeof(int4));
  hipLaunchKernelGGL(k_item_emit, dim3((nkeys + 255) / 256), dim3(256), 0, c.stream, counts, offsets,
                     ioffsets.as<int>(), (int)nkeys, chunk, b.nitems.as<int>() + 1, 0, b.items.as<int4>());
  YG_HIP(hipMemcpyAsync(b.nitems.get(), ioffsets.as<int>() + nkeys, sizeof(int), hipMemcpyDeviceToDevice, c.stream));
  YG_HIP(hipGetLastError());
}

void static_bins_shape(const Ctx& c, StaticBinsView& v, int nrow, int nchan, int nx, int ny, int nkz, const CFView& cf,
                       int chunk) {
  const int ntu = (nx + kTile - 1) / kTile, ntv = (ny + kTile - 1) / kTile;
  const size_t nkeys = (size_t)ntu * ntv * (size_t)nkz;
  YG_CHECK(nkeys < (size_t)INT_MAX, 1, "too many binning keys");
  // a footprint of W pixels starting anywhere in a tile spans at most ceil((63 + W) / 64) tiles
  const size_t W = 2 * (size_t)cf.max_support + 1;
  const size_t span = (63 + W + kTile - 1) / kTile;
  const size_t nsamp = (size_t)nrow * nchan;
  const size_t cap = nsamp * span * span;
  YG_CHECK(cap < (size_t)INT_MAX, 1, "too many (sample, tile) pairs for one pass");
  v.nkeys = nkeys;
  v.npairs_cap = cap;
  v.chunk = std::max(chunk, 16);
  v.nitems_cap = cap / v.chunk + std::min(nkeys, cap) + 1;
  v.kstride = bin_key_stride(nkeys);
  v.nrow = nrow;
  v.nchan = nchan;
  v.nx = nx;
  v.ny = ny;
  v.nkz = nkz;
  // the scan's temporary storage: a size query with null buffers
  size_t t1 = 0;
  YG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, (int*)nullptr, (int*)nullptr, (int)(nkeys + 1), c.stream));
  v.temp_bytes = std::max<size_t>(t1, 1);
}

void size_static_bins(const Ctx& c, StaticBins& sb, int nrow, int nchan, int nx, int ny, int nkz, const CFView& cf,
                      int chunk) {
  StaticBinsView& v = sb.v;
  static_bins_shape(c, v, nrow, nchan, nx, ny, nkz, cf, chunk);
  const size_t nk1 = v.nkeys + 1;
  sb.counts.resize(nk1 * sizeof(int));
  sb.offsets.resize(nk1 * sizeof(int));
  sb.cursor.resize(nk1 * sizeof(int));
  sb.icount.resize(nk1 * sizeof(int));
  sb.ioffsets.resize(nk1 * sizeof(int));
  sb.pairs.resize(std::max<size_t>(v.npairs_cap, 1) * sizeof(int4));
  sb.temp.resize(v.temp_bytes);
  if (v.kstride > 1) sb.wide.resize(nk1 * v.kstride * sizeof(int));
  v.wide = v.kstride > 1 ? sb.wide.as<int>() : nullptr;
  v.counts = sb.counts.as<int>();
  v.offsets = sb.offsets.as<int>();
  v.cursor = sb.cursor.as<int>();
  v.icount = sb.icount.as<int>();
  v.ioffsets = sb.ioffsets.as<int>();
  v.pairs = sb.pairs.as<int4>();
  v.temp = sb.temp.get();
}

void enqueue_static_binning(const Ctx& c, const StaticBinsView& sb, const Samples& smp, const CFView& cf,
                            int4* items_out, const int* base_in, int* base_out, int job) {
  const int n = sb.nrow * sb.nchan;
  const int ntu = (sb.nx + kTile - 1) / kTile, ntv = (sb.ny + kTile - 1) / kTile;
  const int ntiles = ntu * ntv;
  const int nk = (int)sb.nkeys;
  int* counts = sb.counts;
  int* offsets = sb.offsets;
  const int ks = sb.kstride;
  YG_CHECK(ks == 1 || sb.wide, 2, "padded binning counters not allocated");
  int* cw = ks > 1 ? sb.wide : counts;
  const int cgrid = std::min((nk + 256) / 256, 4 * c.n_cu);
  launch_fill_u32(c, reinterpret_cast<uint32_t*>(cw), ((size_t)nk + 1) * ks, 0u);
  if (n > 0)
    hipLaunchKernelGGL(k_bin_count, dim3((n + 255) / 256), dim3(256), 0, c.stream, smp, sb.nrow, sb.nchan, ntu, ntiles,
                       cw, ks);
  if (ks > 1)
    hipLaunchKernelGGL(k_copy_strided_i32, dim3(cgrid), dim3(256), 0, c.stream, cw, ks, counts, 1, (size_t)nk + 1);
  size_t tb = sb.temp_bytes;
  YG_HIP(hipcub::DeviceScan::ExclusiveSum(sb.temp, tb, counts, offsets, nk + 1, c.stream));
  int* cur = ks > 1 ? cw : sb.cursor;
  hipLaunchKernelGGL(k_copy_strided_i32, dim3(cgrid), dim3(256), 0, c.stream, offsets, 1, cur, ks, (size_t)nk + 1);
  if (n > 0)
    hipLaunchKernelGGL(k_bin_fill, dim3((n + 255) / 256), dim3(256), 0, c.stream, smp, sb.nrow, sb.nchan, ntu, ntiles,
                       cf_pad_bytes(cf.max_support, cf.prec), (uint32_t)cf_elem_bytes(cf.prec), cur, ks, sb.pairs,
                       cf.zg_base, cf.nzg);
  hipLaunchKernelGGL(k_item_count, dim3((nk + 1 + 255) / 256), dim3(256), 0, c.stream, counts, nk, sb.chunk,
                     sb.icount);
  tb = sb.temp_bytes;
  YG_HIP(hipcub::DeviceScan::ExclusiveSum(sb.temp, tb, sb.icount, sb.ioffsets, nk + 1, c.stream));
  hipLaunchKernelGGL(k_item_emit, dim3((nk + 255) / 256), dim3(256), 0, c.stream, counts, offsets, sb.ioffsets, nk,
                     sb.chunk, base_in, job, items_out);
  hipLaunchKernelGGL(k_item_base, dim3(1), dim3(64), 0, c.stream, base_in, sb.ioffsets + nk, base_out);
  YG_HIP(hipGetLastError());
}

namespace {
// flag = 1 when sample s differs between (ha, ka) and (hb, kb) in the fields binning uses
__global__ void k_samples_differ(const SampleHdr* __restrict__ ha, const int32_t* __restrict__ ka,
                                 const SampleHdr* __restrict__ hb, const int32_t* __restrict__ kb, int n, int* flag) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const SampleHdr a = ha[s], b = hb[s];
  const bool valid = a.iu0 != INT_MIN;
  const bool same = a.iu0 == b.iu0 &&
                    (!valid || (a.iv0 == b.iv0 && a.cfoff == b.cfoff && a.spack == b.spack && ka[s] == kb[s]));
  if (!same) *flag = 1;
}
}  // namespace

bool bin_samples(Ctx& c, const Samples& smp, int nrow, int nchan, int nx, int ny, int nkz, const CFView& cf) {
  BinCache& bc = c.bins;
  const int n = nrow * nchan;
  if (bc.valid && bc.nrow == nrow && bc.nchan == nchan && bc.nx == nx && bc.ny == ny && bc.nkz == nkz &&
      bc.smax == cf.max_support && bc.cf_prec == cf.prec && bc.zg_base == cf.zg_base) {
    bc.flag.resize(sizeof(int));
    YG_HIP(hipMemsetAsync(bc.flag.get(), 0, sizeof(int), c.stream));
    if (n > 0)
      hipLaunchKernelGGL(k_samples_differ, dim3((n + 255) / 256), dim3(256), 0, c.stream, smp.hdr, smp.key,
                         bc.hdr.as<SampleHdr>(), bc.key.as<int32_t>(), n, bc.flag.as<int>());
    int differ = 1;
    YG_HIP(hipMemcpyAsync(&differ, bc.flag.get(), sizeof(int), hipMemcpyDeviceToHost, c.stream));
    YG_HIP(hipStreamSynchronize(c.stream));
    if (!differ) return true;
  }
  bc.valid = false;
  launch_binning(c, smp, nrow, nchan, nx, ny, nkz, cf, bc.bin);
  bc.hdr.resize((size_t)(n > 0 ? n : 1) * sizeof(SampleHdr));
  bc.key.resize((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
  if (n > 0) {
    YG_HIP(hipMemcpyAsync(bc.hdr.get(), smp.hdr, (size_t)n * sizeof(SampleHdr), hipMemcpyDeviceToDevice, c.stream));
    YG_HIP(hipMemcpyAsync(bc.key.get(), smp.key, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToDevice, c.stream));
  }
  bc.nrow = nrow;
  bc.nchan = nchan;
  bc.nx = nx;
  bc.ny = ny;
  bc.nkz = nkz;
  bc.smax = cf.max_support;
  bc.cf_prec = cf.prec;
  bc.zg_base = cf.zg_base;
  bc.valid = true;
  return false;
}

static int persistent_blocks(const Ctx& c, const TileVariant& v) { return c.n_cu * (v.waves >= 3 ? 3 : 2); }

// kernel variant by polarisation count, register budget (waves per SIMD), pipelining and CF
// storage precision
template <int NP, int WPE, bool PIPE, int CFP>
struct TileKernels {
  static void run(hipStream_t s, dim3 g, bool forward, const int4* items, const int* nitems, int* counter,
                  const TileJob* jobs) {
    if (forward)
      hipLaunchKernelGGL((k_degrid<NP, WPE, PIPE, CFP>), g, dim3(256), 0, s, items, nitems, counter, jobs);
    else
      hipLaunchKernelGGL((k_grid<NP, WPE, PIPE, CFP>), g, dim3(256), 0, s, items, nitems, counter, jobs);
  }
};

// The tile kernels address the CF table through a buffer resource with 32-bit byte offsets
// that starts cf_pad_bytes before the table.
static CFRef cf_ref(const CFView& cf) {
  const size_t pad = cf_pad_bytes(cf.max_support, cf.prec);
  YG_CHECK(cf.bytes > 0, 1, "empty CF table");
  if (cf.zg_base) {
    // per-group resources (make_item_rsrc); the caller checked each group's span
    return CFRef{cf.ptr, 0u, (uint32_t)pad, cf.zg_base, (unsigned long long)cf.bytes, (uint32_t)cf_elem_bytes(cf.prec)};
  }
  // below 2 GiB, so that an out-of-range lane offset (2^31) plus any in-range scalar offset
  // neither wraps nor lands inside the buffer
  YG_CHECK(cf.bytes + pad <= kCFResourceMax, 1, "CF table of 2 GiB or more without key-group bases");
  return CFRef{cf.ptr, (uint32_t)cf.bytes, (uint32_t)pad, nullptr, (unsigned long long)cf.bytes,
               (uint32_t)cf_elem_bytes(cf.prec)};
}

size_t cf_resource_pad(int max_support, int prec) { return cf_pad_bytes(max_support, prec); }

// FP32 tables: every variant; 16-bit tables: the two-wave pipelined form only
template <int NP, int CFP, class F>
static void with_variant(const TileVariant& v, F&& f) {
  if constexpr (CFP != kPrecF32) {
    (void)v;
    f(TileKernels<NP, 2, true, CFP>{});
  } else if (v.waves >= 3) {
    if (v.pipeline) f(TileKernels<NP, 3, true, CFP>{});
    else f(TileKernels<NP, 3, false, CFP>{});
  } else {
    if (v.pipeline) f(TileKernels<NP, 2, true, CFP>{});
    else f(TileKernels<NP, 2, false, CFP>{});
  }
}

template <int NP, class F>
static void with_precision(const TileVariant& v, int prec, F&& f) {
  switch (prec) {
    case kPrecF32: with_variant<NP, kPrecF32>(v, f); break;
    case kPrecF16: with_variant<NP, kPrecF16>(v, f); break;
    case kPrecBF16: with_variant<NP, kPrecBF16>(v, f); break;
    default: throw Error(1, "unknown CF precision");
  }
}

template <class F>
static void with_tile_kernels(const TileVariant& v, int npol, int prec, F&& f) {
  switch (npol) {
    case 1: with_precision<1>(v, prec, f); break;
    case 2: with_precision<2>(v, prec, f); break;
    case 3: with_precision<3>(v, prec, f); break;
    case 4: with_precision<4>(v, prec, f); break;
    default: throw Error(1, "unsupported number of polarisations (1..4)");
  }
}

TileJob make_tile_job(const int4* descs, const Samples& smp, const CFView& cf, float* grid, float* acc, int nx, int ny,
                      int npol, int nzg, size_t wstride) {
  TileJob j{};
  j.descs = descs;
  j.smp = smp;
  j.cf = cf_ref(cf);
  j.grid = grid;
  j.acc = acc;
  j.nx = nx;
  j.ny = ny;
  j.ntu = (nx + kTile - 1) / kTile;
  j.ntiles = j.ntu * ((ny + kTile - 1) / kTile);
  j.nzg = nzg;
  j.wstride = wstride;
  j.npol = npol;
  j.gscale = cf.gscale;
  return j;
}

namespace {
__global__ void k_residual(const float2* __restrict__ data, float2* __restrict__ r, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float2 d = data[i], m = r[i];
    r[i] = make_float2(d.x - m.x, d.y - m.y);
  }
}
}  // namespace

namespace {
__global__ void k_pol_to_image(PolMatrix M, const float2* __restrict__ vis, const float2* __restrict__ noise,
                               const uint8_t* __restrict__ flag, size_t nsamp, float2* __restrict__ vis_o,
                               float2* __restrict__ noise_o, uint8_t* __restrict__ flag_o) {
  for (size_t s = blockIdx.x * (size_t)blockDim.x + threadIdx.x; s < nsamp; s += (size_t)gridDim.x * blockDim.x) {
    float2 v[4], n[4];
    bool flagged = false;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q < M.nin) {
        v[q] = vis ? vis[s + nsamp * q] : make_float2(0.f, 0.f);
        n[q] = noise ? noise[s + nsamp * q] : make_float2(1.f, 1.f);
        flagged |= flag ? flag[s + nsamp * q] != 0 : false;
      }
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      if (p >= M.nout) break;
      float ar = 0.f, ai = 0.f, vr = 0.f, vi = 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q >= M.nin) break;
        const float2 m = M.m[p][q];
        ar += m.x * v[q].x - m.y * v[q].y;
        ai += m.x * v[q].y + m.y * v[q].x;
        vr += m.x * m.x * n[q].x * n[q].x + m.y * m.y * n[q].y * n[q].y;
        vi += m.x * m.x * n[q].y * n[q].y + m.y * m.y * n[q].x * n[q].x;
      }
      if (vis_o) vis_o[s + nsamp * p] = make_float2(ar, ai);
      if (noise_o) noise_o[s + nsamp * p] = make_float2(sqrtf(vr), sqrtf(vi));
      if (flag_o) flag_o[s + nsamp * p] = flagged ? 1 : 0;
    }
  }
}

__global__ void k_pol_to_vis_add(PolMatrix M, const float2* __restrict__ img, size_t nsamp,
                                 float2* __restrict__ vis_out, int nrow, int ld) {
  for (size_t s = blockIdx.x * (size_t)blockDim.x + threadIdx.x; s < nsamp; s += (size_t)gridDim.x * blockDim.x) {
    float2 x[4];
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < M.nin) x[q] = img[s + nsamp * q];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      if (p >= M.nout) break;
      float ar = 0.f, ai = 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q >= M.nin) break;
        const float2 m = M.m[p][q];
        ar += m.x * x[q].x - m.y * x[q].y;
        ai += m.x * x[q].y + m.y * x[q].x;
      }
      // sample s = row + nrow * chan of polarisation p, in a cube of ld rows
      const size_t vo = (s % nrow) + (size_t)ld * (s / nrow + (nsamp / nrow) * p);
      float2 o = vis_out[vo];
      o.x += ar;
      o.y += ai;
      vis_out[vo] = o;
    }
  }
}
}  // namespace

void launch_pol_to_image(const Ctx& c, const PolMatrix& M, const float2* vis, const float2* noise,
                         const uint8_t* flag, size_t nsamp, float2* vis_o, float2* noise_o, uint8_t* flag_o) {
  YG_CHECK(M.nin >= 1 && M.nin <= 4 && M.nout >= 1 && M.nout <= 4, 1, "bad polarisation matrix");
  if (nsamp == 0) return;
  const unsigned blocks = (unsigned)std::min<size_t>((nsamp + 255) / 256, (size_t)8 * c.n_cu);
  hipLaunchKernelGGL(k_pol_to_image, dim3(blocks), dim3(256), 0, c.stream, M, vis, noise, flag, nsamp, vis_o,
                     noise_o, flag_o);
  YG_HIP(hipGetLastError());
}

void launch_pol_to_vis_add(const Ctx& c, const PolMatrix& M, const float2* img, size_t nsamp, float2* vis_out,
                           int nrow, int ld) {
  YG_CHECK(M.nin >= 1 && M.nin <= 4 && M.nout >= 1 && M.nout <= 4, 1, "bad polarisation matrix");
  if (nsamp == 0) return;
  YG_CHECK(nrow > 0 && ld >= nrow && nsamp % nrow == 0, 1, "bad visibility cube shape");
  const unsigned blocks = (unsigned)std::min<size_t>((nsamp + 255) / 256, (size_t)8 * c.n_cu);
  hipLaunchKernelGGL(k_pol_to_vis_add, dim3(blocks), dim3(256), 0, c.stream, M, img, nsamp, vis_out, nrow, ld);
  YG_HIP(hipGetLastError());
}

void launch_residual(const Ctx& c, const float2* data, float2* r, size_t n) {
  if (n == 0) return;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)8 * c.n_cu);
  hipLaunchKernelGGL(k_residual, dim3(blocks), dim3(256), 0, c.stream, data, r, n);
  YG_HIP(hipGetLastError());
}

void launch_fill_u32(const Ctx& c, void* p, size_t n, uint32_t value) {
  if (n == 0) return;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, (size_t)4 * c.n_cu);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, c.stream, (uint32_t*)p, n, value);
  YG_HIP(hipGetLastError());
}

template <int NP>
static void launch_degrid_mm(hipStream_t s, dim3 g, int prec, const int4* items, const int* nitems, int* counter,
                             const TileJob* jobs) {
  switch (prec) {
    case kPrecF32:
      hipLaunchKernelGGL((k_degrid_mm<NP, kPrecF32>), g, dim3(256), 0, s, items, nitems, counter, jobs);
      break;
    case kPrecF16:
      hipLaunchKernelGGL((k_degrid_mm<NP, kPrecF16>), g, dim3(256), 0, s, items, nitems, counter, jobs);
      break;
    case kPrecBF16:
      hipLaunchKernelGGL((k_degrid_mm<NP, kPrecBF16>), g, dim3(256), 0, s, items, nitems, counter, jobs);
      break;
    default: throw Error(1, "unknown CF precision");
  }
}

template <int NP>
static void launch_mma16(hipStream_t s, dim3 g, bool forward, int prec, const int4* items, const int* nitems,
                         int* counter, const TileJob* jobs) {
  if (prec == kPrecF16) {
    if (forward) hipLaunchKernelGGL((k_degrid16<NP, kPrecF16>), g, dim3(256), 0, s, items, nitems, counter, jobs);
    else hipLaunchKernelGGL((k_grid16<NP, kPrecF16>), g, dim3(256), 0, s, items, nitems, counter, jobs);
  } else if (prec == kPrecBF16) {
    if (forward) hipLaunchKernelGGL((k_degrid16<NP, kPrecBF16>), g, dim3(256), 0, s, items, nitems, counter, jobs);
    else hipLaunchKernelGGL((k_grid16<NP, kPrecBF16>), g, dim3(256), 0, s, items, nitems, counter, jobs);
  } else {
    throw Error(1, "the 16-bit matrix-core kernels need a 16-bit table");
  }
}

void launch_tile_jobs(const Ctx& c, const TileVariant& v0, bool forward, int npol, int prec_arg, const TileJob* jobs,
                      const int4* items, const int* nitems, int* counter) {
  launch_fill_u32(c, counter, kQueueInts, 0u);
  const int prec = prec_arg & ~kPrecMma16;
  if ((prec_arg & kPrecMma16) && prec != kPrecF32) {
    // 16-bit table and operands: the 16-bit matrix-core kernels (two workgroups per CU)
    const dim3 g(c.n_cu * 2);
    switch (npol) {
      case 1: launch_mma16<1>(c.stream, g, forward, prec, items, nitems, counter, jobs); break;
      case 2: launch_mma16<2>(c.stream, g, forward, prec, items, nitems, counter, jobs); break;
      case 3: launch_mma16<3>(c.stream, g, forward, prec, items, nitems, counter, jobs); break;
      case 4: launch_mma16<4>(c.stream, g, forward, prec, items, nitems, counter, jobs); break;
      default: throw Error(1, "unsupported number of polarisations (1..4)");
    }
    YG_HIP(hipGetLastError());
    return;
  }
  // one polarisation: half of the matrix form's outputs would be zeros; the VALU form is used
  if (forward && v0.matrix && npol >= 2) {
    const dim3 g(c.n_cu * 2);  // two workgroups (8 waves) per CU: <= 256 VGPRs per wave
    switch (npol) {
      case 1: launch_degrid_mm<1>(c.stream, g, prec, items, nitems, counter, jobs); break;
      case 2: launch_degrid_mm<2>(c.stream, g, prec, items, nitems, counter, jobs); break;
      case 3: launch_degrid_mm<3>(c.stream, g, prec, items, nitems, counter, jobs); break;
      case 4: launch_degrid_mm<4>(c.stream, g, prec, items, nitems, counter, jobs); break;
      default: throw Error(1, "unsupported number of polarisations (1..4)");
    }
    YG_HIP(hipGetLastError());
    return;
  }
  const TileVariant v = prec == kPrecF32 ? v0 : TileVariant{2, 1, 0};
  const dim3 g(persistent_blocks(c, v));
  with_tile_kernels(v, npol, prec, [&](auto k) { k.run(c.stream, g, forward, items, nitems, counter, jobs); });
  YG_HIP(hipGetLastError());
}

void launch_grid(const Ctx& c, Binning& b, const Samples& smp, const CFView& cf, float2* grid, int nx, int ny,
                 int npol, int nzg, size_t wstride) {
  b.job_host = make_tile_job(b.pairs.as<int4>(), smp, cf, reinterpret_cast<float*>(grid), nullptr, nx, ny, npol, nzg,
                             wstride);
  b.job.resize(sizeof(TileJob));
  YG_HIP(hipMemcpyAsync(b.job.get(), &b.job_host, sizeof(TileJob), hipMemcpyHostToDevice, c.stream));
  launch_tile_jobs(c, c.grid_variant, false, npol, tile_prec(cf), b.job.as<TileJob>(), b.items.as<int4>(),
                   b.nitems.as<int>(), b.counter.as<int>());
}

void launch_degrid(const Ctx& c, Binning& b, const CFView& cf, const float2* grid, int nx, int ny, int npol, int nzg,
                   float* acc, size_t wstride) {
  b.job_host = make_tile_job(b.pairs.as<int4>(), Samples{}, cf, const_cast<float*>(reinterpret_cast<const float*>(grid)),
                             acc, nx, ny, npol, nzg, wstride);
  b.job.resize(sizeof(TileJob));
  YG_HIP(hipMemcpyAsync(b.job.get(), &b.job_host, sizeof(TileJob), hipMemcpyHostToDevice, c.stream));
  launch_tile_jobs(c, c.degrid_variant, true, npol, tile_prec(cf), b.job.as<TileJob>(), b.items.as<int4>(),
                   b.nitems.as<int>(), b.counter.as<int>());
}

void launch_degrid_finish(const Ctx& c, const Samples& recs, const float* acc, int nrow, int nchan, int npol,
                          float2* vis_out, int ld) {
  const int n = nrow * nchan;
  if (n == 0) return;
  YG_CHECK(ld >= nrow, 1, "vis_out_rows must be 0 or at least nrow");
  hipLaunchKernelGGL(k_degrid_finish, dim3((n + 255) / 256), dim3(256), 0, c.stream, recs, acc, nrow, nchan, npol,
                     vis_out, ld);
  YG_HIP(hipGetLastError());
}

namespace {
__global__ void k_absmax_bits(const float* __restrict__ a, size_t n, uint32_t* out) {
  float m = 0.f;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(a[i]));
  for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_xor(m, d));
  // non-negative floats order as their bit patterns
  if ((threadIdx.x & 63) == 0 && m > 0.f) atomicMax(out, __float_as_uint(m));
}
__global__ void k_pow2_scale(const uint32_t* mx, float* scale) {
  const float m = __uint_as_float(*mx);
  *scale = m > 0.f ? ldexpf(1.f, 13 - ilogbf(m)) : 1.f;
}
}  // namespace

void launch_grid_scale16(const Ctx& c, const float* grid, size_t n, uint32_t* tmp, float* scale) {
  launch_fill_u32(c, tmp, 1, 0u);
  const int blocks = (int)std::min<size_t>((n + 255) / 256, (size_t)c.n_cu * 8);
  if (n > 0) hipLaunchKernelGGL(k_absmax_bits, dim3(std::max(blocks, 1)), dim3(256), 0, c.stream, grid, n, tmp);
  hipLaunchKernelGGL(k_pow2_scale, dim3(1), dim3(1), 0, c.stream, tmp, scale);
  YG_HIP(hipGetLastError());
}

namespace {
__global__ void k_cf_round(const float2* __restrict__ src, uint32_t* __restrict__ dst, size_t n, int prec) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float2 v = src[i];
  dst[i] = (uint32_t)round16(v.x, prec) | ((uint32_t)round16(v.y, prec) << 16);
}
}  // namespace

void launch_cf_round(const Ctx& c, const float2* src, uint32_t* dst, size_t n, int prec) {
  YG_CHECK(prec == kPrecF16 || prec == kPrecBF16, 1, "CF rounding needs a 16-bit precision");
  if (n == 0) return;
  hipLaunchKernelGGL(k_cf_round, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, src, dst, n, prec);
  YG_HIP(hipGetLastError());
}

}  // namespace ygrid
