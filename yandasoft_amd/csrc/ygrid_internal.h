// Internal declarations shared by the host code and the HIP kernels of libygrid.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ygrid {

constexpr double kC = 299792458.0;  // casacore::C::c
constexpr double kPi = 3.14159265358979323846;

// Grid tile owned by one workgroup in the grid/degrid kernels.
constexpr int kTile = 64;
// default pairs per tile-kernel work item (pairs of one grid tile and key; the tile is loaded or
// flushed once per item)
constexpr int kDefaultItemPairs = 2048;
// work-item queues of a tile-kernel launch (one per XCD; their counters live in one int array)
constexpr int kItemQueues = 8;
// ints between two queue counters (each its own 128-byte line: the workgroups of a launch hit
// them all at once) and the size of a launch's counter array
constexpr int kQueueStride = 32;
constexpr int kQueueInts = kItemQueues * kQueueStride;
// Guard band of zeros around every CF plane in device memory so that the 16x4-lane
// blocks of the grid/degrid kernels may read past the (2S+1)^2 footprint without a
// per-lane bounds test: columns -kGuardU..2S+kGuardU-1, rows -kGuardV..2S+kGuardV-1.
constexpr int kGuardU = 16;
constexpr int kGuardV = 4;
inline int cf_pitch(int support) { return 2 * support + 1 + 2 * kGuardU; }
inline int cf_rows(int support) { return 2 * support + 1 + 2 * kGuardV; }

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define YG_CHECK(cond, code, msg)                                 \
  do {                                                            \
    if (!(cond)) throw ::ygrid::Error((code), std::string(msg));  \
  } while (0)

#define YG_HIP(call)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      throw ::ygrid::Error(3, std::string(#call) + " failed: " + hipGetErrorString(e_));          \
  } while (0)

// Owning device allocation.
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { resize(bytes); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void resize(size_t bytes) {
    if (bytes <= n_) return;
    release();
    if (bytes == 0) return;
    YG_HIP(hipMalloc(&p_, bytes));
    n_ = bytes;
  }
  void release() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p_); }
  void* get() const { return p_; }
  size_t size() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// How the CF plane (before oversampling) of a sample is chosen: TableVisGridder::cIndex
// overrides (TableVisGridder.cc:1410-1412; WProjectVisGridder.cc:114-147,286-292;
// AWProjectVisGridder.cc:123-172,855-863).
struct CIndexDesc {
  int kind = 0;            // 0: constant 0; 1: w-plane; 2: w-plane + feed/chan (AW); 3: w plane
                           // selects the grid (w stacking), CF plane 0; 4: w plane selects the
                           // grid, CF plane of (chan, feed, field) (A-projection w stacking)
  int nw = 1;              // number of w planes
  double wscale = 1.0;     // linear sampling step (WDependentGridderBase.cc:60-65)
  double wmax = 1.0;
  int sampling = 0;        // 0 linear, 1 power law, 2 gaussian
  double exponent = 1.0;   // power law
  double g_ts = 1.0, g_amp = 1.0;  // gaussian: 2 sigma^2 and amplitude
  int wmaxclip = 0;
  int freqdep = 0;         // AW: channel-dependent CFs
  int nchan_cf = 1;        // AW: channels in the CF cache
  int maxfeeds = 1;        // AW
  int field = 0;           // AW: current field
};

// What kind of pass the prep kernel prepares.
enum PassMode : int { kModeGrid = 0, kModePSF = 1, kModePCF = 2, kModeDegrid = 3 };

// Per (row, chan) sample written by the prep kernel, as three arrays indexed by
// s = row + nrow*chan.  The grid/degrid kernels read a sample wave-uniformly (scalar loads).
// hdr.iu0 == INT32_MIN marks a sample that is not gridded.
struct alignas(16) SampleHdr {
  int32_t iu0, iv0;     // footprint origin: (iuOffset - S, ivOffset - S)
  int32_t cfoff;        // element offset of the padded CF plane in the CF table
  int32_t spack;        // support S (low 16 bits) | pitch << 16
};
static_assert(sizeof(SampleHdr) == 16, "SampleHdr must be 16 bytes");
constexpr int kCoef = 16;  // floats per sample in Samples::coef

// Storage precision of CF taps and of the gridding coefficients (the weighted visibilities):
// FP32, or FP16 / BF16 (rounded to nearest even when stored); arithmetic is FP32 throughout.
enum Prec : int { kPrecF32 = 0, kPrecF16 = 1, kPrecBF16 = 2 };
// bytes of one complex CF tap stored at precision p
inline int cf_elem_bytes(int p) { return p == kPrecF32 ? 8 : 4; }
struct Samples {
  SampleHdr* hdr;
  int32_t* key;         // binning key: (ichan, z-group) plane index
  // grid: the MAC coefficients (a, b, c, d) of each polarisation q, with
  //   grid_re += a*C.re + b*C.im, grid_im += c*C.re + d*C.im
  // (the value spread is v*C, or v*conj(C) in the PCF half plane), laid out as the gridding
  // kernel's MFMA operands: [4*o + m] with output o = 2*(q&1) + (0 re, 1 im) and operand
  // m = 2*(q/2) + (0 for C.re, 1 for C.im); degrid: [0..1] = the phasor times the weight.
  float* coef;
  // the gridding coefficients stored at 16-bit precision coef_prec (kCoef per sample, same
  // order) instead of coef, or null
  uint16_t* coef16;
  int coef_prec;
};

struct PrepParams {
  int nrow, nchan, npol;        // chunk shape (npol == image pols)
  int nx, ny;                   // padded grid
  int nchan_img;
  double uvcell_u, uvcell_v;
  int oversample;
  int mode;                     // PassMode
  int track_per_os;             // sum-of-weights per oversampled plane
  int nsw_rows;
  int psf_feed;                 // -1: all data
  CIndexDesc ci;
  int zdiv;                     // z-group divisor for binning keys
  int nzg;                      // number of z groups
  int vis_prec;                 // Prec of the stored gridding coefficients
  // learned CFs: every sample gets its own padded plane of support lcf_S at s * lcf_plane_elems
  // (0: the CF table's planes)
  int lcf_S;
  long long lcf_plane_elems;
  // MaxPointingSeparation (radians; <= 0: off) about the image centre (TableVisGridder.cc:571-578)
  double max_sep;
  double centre_lon, centre_lat;
  // PSF passes with a representative field (psf_feed >= 0): also its dish pointing (psf_field)
  int psf_field;
  double psf_lon, psf_lat;
};

struct PrepOutputs {
  Samples smp;                  // nrow*nchan
  double* sumw;                 // nsw_rows * npol * nchan_img
  unsigned long long* counters; // [0] samples, [1] taps, [2] flagged, [3] wflagged, [4] rows rejected by pointing
  int* error;                   // bit 1: zero noise weight; bit 2: w out of range
  float4* lcf_in;               // learned CFs: (w in wavelengths, fracu, fracv, gridded) per sample
};

struct PrepInputs {
  const double* uvw;
  const double* delay;
  const double* freq;
  const float2* vis;
  const float2* noise;
  const uint8_t* flag;
  const int32_t* feed1;
  const double* pointing1;     // nrow x 2 (lon, lat) or null (= the image centre)
  const double* dish1;         // nrow x 2 or null (= pointing1)
  const int32_t* chanmap;      // nchan: image channel or -1
  const float* visweight;      // nchan or null
  const int32_t* cf_off;       // per CF plane (incl. oversampling)
  const int32_t* cf_supp;
  const int32_t* cf_ofs_u;     // per plane before oversampling, or null
  const int32_t* cf_ofs_v;
  int n_cf;
};

// Device work for one grid/degrid pass, produced by the binning step.
// The CF buffer as a tile kernel addresses it: table address and size, and the pad (bytes, at
// least 16 rows of the widest plane) by which the kernel's buffer resource starts before it.
struct CFRef {
  const void* ptr;
  uint32_t bytes, pad;
  // tables of 2 GiB or more: the element offset of each key group's first plane (the pair
  // descriptors' offsets are relative to it, and each work item addresses its group through a
  // buffer resource based there); total bytes of the table.  Null: one resource for the table.
  const long long* zg_base;
  unsigned long long total;
  uint32_t esz;
};

// One grid of a tile-kernel launch: its pair descriptors, samples, CF buffer, grid (gridding:
// accumulated; degridding: read) and, degridding, the per-sample accumulators.  Work items name
// their job in int4::w.
struct TileJob {
  const int4* descs;
  Samples smp;
  CFRef cf;
  float* grid;
  float* acc;
  int nx, ny, ntu, ntiles, nzg, npol;
  size_t wstride;  // float2 elements between the grids of consecutive w planes (w stacking), else 0
  const float* gscale;  // see CFView::gscale (null: 1)
};

struct Binning {
  DevBuf counts, offsets, cursor, ioffsets, pairs, items, nitems, counter, temp, wide;
  size_t nkeys = 0, npairs_cap = 0;
  // the one-job table of a single-grid launch, and its host copy (kept until the pass has synced)
  DevBuf job;
  TileJob job_host{};
};

// hipStream + device id of a context.
// Register budget and pipelining of a tile kernel: waves = waves per SIMD the kernel is built
// for (2: <= 256 VGPRs, 3: <= 168), pipeline = CF loads of pair i+1 issued before pair i's
// arithmetic (a second register set).  matrix (degridding only): the matrix-core form that
// works four pairs at once (k_degrid_mm; waves and pipeline do not apply to it).
struct TileVariant {
  int waves;
  int pipeline;
  int matrix;
};

// The binning of the last pass run on a context, with the sample headers and keys it was built
// from.  A pass whose samples bin identically (same headers and keys, grid and support: e.g. the
// degridding and gridding of one chunk in a major cycle) reuses it.
struct BinCache {
  DevBuf hdr, key, flag;
  int nrow = -1, nchan = -1, nx = 0, ny = 0, nkz = 0, smax = 0, cf_prec = 0;
  const long long* zg_base = nullptr;
  bool valid = false;
  Binning bin;
};

// Spacing (in ints) of the per-key counters the binning kernels update atomically: with few
// keys each key gets its own 128-byte line (a padded array of at most 1 MB), else 1.
inline int bin_key_stride(size_t nkeys) { return nkeys + 1 <= 8192 ? 32 : 1; }

struct StaticBinsView {
  int *counts = nullptr, *offsets = nullptr, *cursor = nullptr, *icount = nullptr, *ioffsets = nullptr;
  // the padded counters when kstride > 1 (counts, then cursor), else null
  int* wide = nullptr;
  int kstride = 1;
  int4* pairs = nullptr;
  void* temp = nullptr;
  size_t nkeys = 0, npairs_cap = 0, nitems_cap = 0, temp_bytes = 0;
  int nrow = 0, nchan = 0, nx = 0, ny = 0, nkz = 0, chunk = 0;
};
struct StaticBins {
  DevBuf counts, offsets, cursor, icount, ioffsets, pairs, temp, wide;
  StaticBinsView v;
};

// Largest CF support the small-support kernels take (kernels_small.hip).
constexpr int kSmallSupportMax = 8;

// Buffers of the small-support gridding pass: its static binning (one work item per key) and
// the persistent launch's item list, item count and queue counters.
struct SmallGridBufs {
  StaticBins bins;
  DevBuf items, base, counter;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int n_cu = 256;
  TileVariant grid_variant{2, 1};
  TileVariant degrid_variant{2, 1, 1};
  int item_pairs = 0;  // pairs per work item of the tile kernels (0: the built-in default)
  // passes whose CF support is at most this use the small-support kernels (0: never;
  // at most kSmallSupportMax)
  int small_support_max = kSmallSupportMax;
  // rows of the LDS band a workgroup of the small-support gridding kernel owns (4, 8 or 16; 0:
  // chosen by the support)
  int small_band_rows = 0;
  BinCache bins;  // shared by the gridders of the context (their passes are ordered on stream)
  SmallGridBufs small;  // likewise
};

// Arguments of the small-support kernels (kernels_small.hip).  The grid holds nPol planes per
// image channel (float2 x nx x ny each); with w stacking, grid w of a key's (key % nzg) starts
// wstride float2 elements after grid w - 1, else wstride is 0.
struct SmallGridArgs {
  const int4* pairs;  // the binning's pair descriptors (x: the sample)
  Samples smp;
  const void* cf;     // the CF table (taps of cf_elem_bytes(prec) bytes; SampleHdr::cfoff indexes it)
  float* grid;
  int nx, ny, ntu, ntiles, nzg;
  size_t wstride;
};
struct SmallDegridArgs {
  Samples smp;
  const void* cf;
  const float* grid;
  float* acc;  // per sample 8 floats: (re, im) of C * conj(G) summed, per polarisation
  int nrow, nchan, nx, ny, nzg;
  size_t wstride;
  // 16-bit table with a 16-bit visibility side (CFView::mma16): the model grid's values are
  // rounded to the table's format, after the scale gscale (FP16; null: 1), as the tile kernels
  // round their operand
  int round16 = 0;
  const float* gscale = nullptr;
};
// Persistent small-support gridding over the items of a one-item-per-key static binning
// (nitems: device item count, nitems_cap its bound; counter: kItemQueues zeros).
void launch_grid_small(const Ctx&, const int4* items, const int* nitems, int* counter, int nitems_cap,
                       const SmallGridArgs& a, int npol, int prec, int max_support);
void launch_degrid_small(const Ctx&, const SmallDegridArgs& a, int npol, int prec, int max_support);

// Kernel launchers (kernels_grid.hip)
void launch_prep(const Ctx&, const PrepParams&, const PrepInputs&, const PrepOutputs&);
// The CF table as the tile kernels read it: taps stored at precision prec (cf_elem_bytes each),
// bytes in all (the kernels address it with 32-bit byte offsets), max_support the largest
// support of its planes.
struct CFView {
  const void* ptr;
  size_t bytes;
  int max_support;
  int prec;
  // 16-bit table whose passes' other operand (the gridding coefficients; the model grid when
  // degridding) is stored at the same precision: the tile kernels multiply on the 16-bit
  // matrix cores
  int mma16 = 0;
  // degridding with FP16 operands: a power-of-two scale of the model grid (device float) that
  // brings its largest magnitude to [2^13, 2^14)
  const float* gscale = nullptr;
  // tables of 2 GiB or more (see CFRef): per key group (nzg of them per image channel) the
  // element offset of the group's first plane, on the device; null when the table fits one
  // buffer resource
  const long long* zg_base = nullptr;
  int nzg = 0;
};
// bytes a CF table may span (with the resource's leading pad) for one buffer resource of the
// tile kernels; the largest pad the tile kernels put before a table of this support and precision
constexpr unsigned long long kCFResourceMax = 0x7F000000ull;
size_t cf_resource_pad(int max_support, int prec);
// the precision argument of launch_tile_jobs: the table's, with kPrecMma16 for mma16 views
constexpr int kPrecMma16 = 0x10;
inline int tile_prec(const CFView& v) { return v.prec | (v.mma16 ? kPrecMma16 : 0); }
// Bins valid records by (key, tile); returns nothing, work lands in b.  The samples are the
// nrow x nchan (row, channel) samples of one chunk, s = row + nrow*chan.  The pair descriptors
// address the CF table cf describes.
void launch_binning(const Ctx&, const Samples& smp, int nrow, int nchan, int nx, int ny, int nkeys_z, const CFView& cf,
                    Binning& b);
// Bins the samples into ctx.bins unless the cached binning was built from identical sample
// headers and keys with the same parameters; returns true when the cache was reused.
bool bin_samples(Ctx& c, const Samples& smp, int nrow, int nchan, int nx, int ny, int nkeys_z, const CFView& cf);
void launch_grid(const Ctx&, Binning& b, const Samples& smp, const CFView& cf, float2* grid, int nx, int ny,
                 int npol, int nzg, size_t wstride = 0);
void launch_degrid(const Ctx&, Binning& b, const CFView& cf, const float2* grid, int nx, int ny, int npol,
                   int nzg, float* acc, size_t wstride = 0);
// The job of one grid for the tile kernels (descs: its binning's pair descriptors).
// *scale = 2^(13 - e) with 2^e <= max |grid[i]| < 2^(e+1) over n floats (1 when all are zero);
// tmp: one device uint
void launch_grid_scale16(const Ctx&, const float* grid, size_t n, uint32_t* tmp, float* scale);
TileJob make_tile_job(const int4* descs, const Samples& smp, const CFView& cf, float* grid, float* acc, int nx, int ny,
                      int npol, int nzg, size_t wstride = 0);
// One persistent tile-kernel launch over the work items of several jobs (all with npol
// polarisations and CF precision prec): items[0 .. *nitems) in device memory, counter zeroed
// by the launch's own fill kernel.  forward: degridding.
void launch_tile_jobs(const Ctx&, const TileVariant& v, bool forward, int npol, int prec, const TileJob* jobs,
                      const int4* items, const int* nitems, int* counter);

// Sync-free binning of one pass into buffers sized by upper bounds (for batches, graph capture
// and the allocation-free entry point): no host readback, no allocation, no memset nodes.  The
// pass's work items go to items_out[*base_in ...) with job index `job`, and
// *base_out = *base_in + its item count.  The view holds the buffers and their sizes; they are
// owned by a StaticBins or carved from a caller's workspace.
// The sizes of the binning of nrow x nchan samples of support <= cf.max_support (no allocation:
// the scan's temporary size is a query).
void static_bins_shape(const Ctx&, StaticBinsView& v, int nrow, int nchan, int nx, int ny, int nkz, const CFView& cf,
                       int chunk);
// Sizes sb for nrow x nchan samples of support <= cf.max_support (allocates).
void size_static_bins(const Ctx&, StaticBins& sb, int nrow, int nchan, int nx, int ny, int nkz, const CFView& cf,
                      int chunk);
void enqueue_static_binning(const Ctx&, const StaticBinsView& sb, const Samples& smp, const CFView& cf, int4* items_out,
                            const int* base_in, int* base_out, int job);

// A fixed device workspace that a pass's buffers are carved from (256-byte aligned pieces);
// running out of it is an error, never an allocation.
struct PassArena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  void* take(size_t bytes) {
    const size_t a = (used + 255) & ~(size_t)255;
    YG_CHECK(a <= cap && bytes <= cap - a, 1,
             "the pass needs more than " + std::to_string((a + bytes + (1 << 20) - 1) >> 20) +
                 " MiB of device workspace; the allocation-free entry point has " + std::to_string(cap >> 20) +
                 " MiB: split the chunk");
    used = a + bytes;
    return base + a;
  }
  template <class T> T* take_n(size_t n) { return static_cast<T*>(take(n * sizeof(T))); }
};
// The allocation-free entry point's device workspace (noalloc.cpp): the library's file-scope
// device array `scratch`, resolved once per device when a context is created on it.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
void noalloc_setup(int device);
// the current device's copy of `scratch` (after noalloc_setup on that device)
void* noalloc_scratch(int device);
// fills n 32-bit words with value (a kernel: usable inside captured graphs)
void launch_fill_u32(const Ctx&, void* p, size_t n, uint32_t value);
// *flag |= (any of data[0 .. n) is non-zero)
void launch_any_nonzero(const Ctx& c, const float* data, size_t n, int* flag);
// Polarisation conversion of visibility cubes (casacore Cube(nRow, nChan, nPol) layout: the
// sample (row, chan) of polarisation p at index s + nsamp * p).  m: row-major nout x nin.
struct PolMatrix {
  float2 m[4][4];
  int nout, nin;
};
// To the image frame: vis_o = M vis; noise_o per output (re, im) = sqrt of the sums of
// Re(M)^2 and Im(M)^2 times the inputs' (re, im) / (im, re) variances (noise NULL: 1);
// flag_o = any input flagged.  Null outputs are skipped.
void launch_pol_to_image(const Ctx&, const PolMatrix& M, const float2* vis, const float2* noise,
                         const uint8_t* flag, size_t nsamp, float2* vis_o, float2* noise_o, uint8_t* flag_o);
// Back to the visibility frame: vis_out += M img, img the nsamp = nrow x nchan samples per
// polarisation, vis_out a cube of ld rows.
void launch_pol_to_vis_add(const Ctx&, const PolMatrix& M, const float2* img, size_t nsamp, float2* vis_out, int nrow,
                           int ld);
// r[i] = data[i] - r[i] over n complex values (the residual of a model cube r).
void launch_residual(const Ctx&, const float2* data, float2* r, size_t n);
// rotatedUVW / uvwRotationDelay on the device (kernels_uvw.hip): the tangent point's frame
// vectors (equatorial xyz) and the image centre's direction.
struct UvwTargets {
  double tu[3], tv[3], tw[3], cw[3];
};
UvwTargets uvw_targets(double tangent_lon, double tangent_lat, double centre_lon, double centre_lat);
// out (nrow x 3) = uvw rotated from each row's phase centre (pc: 2 doubles, or 2 per row) to the
// tangent point; delay (nrow, or null) = the path difference rephasing from the phase centre to
// the image centre.
void launch_uvw_rotate(const Ctx&, const double* uvw, const double* pc, int pc_per_row, const UvwTargets& T,
                       int nrow, double* out, double* delay);
// Snapshot imaging (kernels_snapshot.hip).  sums6: u.u, u.v, v.v, u.w, v.w and (as the bits of a
// non-negative double) the largest |w - A u - B v| over the rows.
void launch_wplane_sums(const Ctx&, const double* uvw, int nrow, double A, double B, double* sums6);
// out = uvw with w - A u - B v (nrow x 3)
void launch_wplane_apply(const Ctx&, const double* uvw, int nrow, double A, double B, double* out);
struct SnapshotRegridArgs {
  double A = 0, B = 0;        // the snapshot's plane w = A u + B v
  double cellx = 0, celly = 0;
  int to_target = 1;          // 1: snapshot frame -> target frame, 0: target -> snapshot
  int cubic = 1;              // 1: Keys cubic convolution, 0: bilinear
  int add = 0;                // 1: dst += regridded, 0: dst = regridded
  int clip_x0 = 0, clip_y0 = 0, clip_kx = 0, clip_ky = 0;  // pixels kept (the rest zeroed)
};
void launch_snapshot_regrid(const Ctx&, const float* src, float* dst, int nx, int ny, int nplanes,
                            const SnapshotRegridArgs& a);
// Rounds the FP32 CF table (n complex taps) to prec (FP16 or BF16), packed (re, im) in 32 bits.
void launch_cf_round(const Ctx&, const float2* src, uint32_t* dst, size_t n, int prec);
// vis_out: a cube of ld >= nrow rows
void launch_degrid_finish(const Ctx&, const Samples& smp, const float* acc, int nrow, int nchan, int npol,
                          float2* vis_out, int ld);

// FFT-domain helpers (kernels_fft.hip)
struct FFTPlan;
void destroy_plan(FFTPlan*);
struct FFTPlanDeleter { void operator()(FFTPlan* p) const { destroy_plan(p); } };
using PlanPtr = std::unique_ptr<FFTPlan, FFTPlanDeleter>;
// In-place centred 2-D transform of `batch` planes of ny rows x nx columns.
void fft_centred_c2c(const Ctx&, float2* data, int nx, int ny, int batch, bool forward, PlanPtr& cache);
// Plain transform (no centring) and the two checkerboard passes it is combined with.
void fft_plain_c2c(const Ctx&, float2* data, int nx, int ny, int batch, bool forward, PlanPtr& cache);
void launch_checkerboard(const Ctx&, float2* data, int nx, int ny, int batch);
void launch_checkerboard_sign(const Ctx&, float2* data, int nx, int ny, int batch);
// FP64 transforms between the (FP32) grid and the image, through an FP64 scratch of nplanes*nxp*nyp
void grid_to_image_z(const Ctx&, const float2* grid, int nxp, int nyp, int nplanes, const double* ccfx,
                     const double* ccfy, int nx, int ny, double2* scratch, float* out, PlanPtr& cache);
void model_to_grid_z(const Ctx&, const float* model, int nx, int ny, int nplanes, const double* ccfx,
                     const double* ccfy, int nxp, int nyp, double2* scratch, float2* grid, PlanPtr& cache);
// W stacking (WStackVisGridder.cc:134-171,255-289,314-327).  The w screen of plane w2pi = 2 pi w
// at padded pixel (X, Y): (cos p, -sin p), p = w2pi (1 - sqrt(1 - r^2)), r^2 = ((X - nxp/2) cellx)^2
// + ((Y - nyp/2) celly)^2, where r^2 < 1 (1 elsewhere), in the reference's float arithmetic.
// acc (nplanes x nyp x nxp) += Re(screen * centred unnormalised inverse transform of grid).
void wstack_accumulate(const Ctx&, const float2* grid, int nxp, int nyp, int nplanes, float w2pi, float cellx,
                       float celly, double2* scratch, double* acc, PlanPtr& cache);
// image = ccf * acc, the unpadded centre
void wstack_finalise(const Ctx&, const double* acc, int nxp, int nyp, int nplanes, const double* ccfx,
                     const double* ccfy, int nx, int ny, float* out);
// preconditioner stack (WStackVisGridder.cc:173-209,222-262): stack += norm * centring signs *
// (Re g, im_scale * Im g) for one w plane's grids
void wstack_pcf_stack(const Ctx&, const float2* grid, int nxp, int nyp, int nplanes, double im_scale, double norm,
                      double2* stack);
// image = Re(unnormalised backward transform of a stack carrying its centring signs), the
// unpadded centre, no correction
void stack_to_image_z(const Ctx&, double2* stack, int nxp, int nyp, int nplanes, int nx, int ny, float* out,
                      PlanPtr& cache);
// grid = centred unnormalised forward transform of conj(screen * ccf * padded(model))
void wstack_model_to_grid(const Ctx&, const float* model, int nx, int ny, int nplanes, const double* ccfx,
                          const double* ccfy, int nxp, int nyp, float w2pi, float cellx, float celly, double2* scratch,
                          float2* grid, PlanPtr& cache);
void launch_finalise_image(const Ctx&, const float2* fftd, int nxp, int nyp, int nplanes, const double* ccfx,
                           const double* ccfy, int nx, int ny, float* out);
void launch_model_to_grid(const Ctx&, const float* model, int nx, int ny, int nplanes, const double* ccfx,
                          const double* ccfy, int nxp, int nyp, float2* grid);

// CF generation kernels (kernels_cf.hip)
// Fill planes [0,nb) of buf (nx*ny complex each) with taper x phase screen for w terms wt[]
// (WProjectVisGridder::populateThisPlane, WProjectVisGridder.cc:632-660).
void launch_wproj_screen(const Ctx&, float2* buf, int nx, int ny, int qnx, int qny, double ccellx, double ccelly,
                         const float* ccfx, const float* ccfy, const double* two_pi_w, int nb);
// Argmax of |c|^2 (first in memory order among equals) of one plane; result (flat index) in *out.
void launch_argmax(const Ctx&, const float2* plane, int n, unsigned long long* out);
// Cut out os^2 CF planes per FFT plane, normalise by the sum of real parts (mode 0: WProject),
// or scale by `rescale` (mode 1: AWProject), and write them, zero-padded, into the table.
void launch_cf_cutout(const Ctx&, const float2* buf, int nx, int ny, int nb, int os, const int* supp_of_plane,
                      const int* offu_of_plane, const int* offv_of_plane, const int32_t* cf_off_first, int mode,
                      double rescale, float2* table);
// SphFunc CF: planes from 1-D factors bx[fracu][ix], by[fracv][iy], normalised by the sum
// of |C| (SphFuncVisGridder.cc:137-167).
void launch_sph_cf(const Ctx&, const double* bx, const double* by, int os, int support, const int32_t* cf_off,
                   float2* table);

// AWProject kernels (kernels_awp.hip)
struct IllumParams;  // illumination.h
void launch_illum_pattern(const Ctx&, float2* buf, int nu, int nv, const IllumParams* prm, int nb);
void launch_aw_screen(const Ctx&, const float2* A, int nx, int ny, int qnx, int qny, double ccellx, double ccelly,
                      float ascale, const double* tx, const double* ty, const double* two_pi_w, int nb, float2* out);
// multiply by the output checkerboard of a centred transform (incl. its global sign) and a scale
void launch_scale_sign(const Ctx&, float2* data, int nx, int ny, int nb, float scale);
// asym: taps -S .. S-1 only (AProjectWStack)
void launch_place_cf(const Ctx&, const float2* table, const int32_t* off, const int* supp, const int* ou,
                     const int* ov, int cnx, int cny, int nb, double2* out, int asym = 0);
// max2_tmp (2 * nb uints, or null): each plane's |.|^2 divided by its largest value (AProjectWStack)
void launch_aw_weights_acc(const Ctx&, const double2* planes, int nb, int cnx, int cny, const double* wt, int ncp,
                           const double* tx, const double* ty, double* cout, unsigned long long* max2_tmp = nullptr);
// AProjectWStack's CF before its forward transform: |A ascale|^2 / max, with the forward
// transform's checkerboard (tmp: one device uint)
void launch_aws_cf_power(const Ctx&, const float2* A, int nx, int ny, float ascale, float2* buf, uint32_t* tmp);
void launch_pad_centre(const Ctx&, const double2* in, int cnx, int cny, int nx, int ny, int nb, double2* out);
void launch_d2c(const Ctx&, const double* in, size_t n, double scale, double2* out);
void launch_extract_real(const Ctx&, const double2* in, int nxp, int nyp, int nx, int ny, int nb, float* out);
// plain / centred FP64 transforms
void fft_plain_z2z(const Ctx&, double2* data, int nx, int ny, int batch, bool forward, PlanPtr& cache);
void fft_centred_z2z(const Ctx&, double2* data, int nx, int ny, int batch, bool forward, PlanPtr& cache);

// Host helpers (spheroidal.cpp)
class Spheroidal {
 public:
  Spheroidal(double c, double alpha, int nterms = 24);
  double operator()(double nu) const;
 private:
  int m_;
  std::vector<double> d_;
  double norm0_;
  double series(double x) const;
};
// 1-D grid-correction factors of SphFuncVisGridder::correctConvolution (SphFuncVisGridder.cc:171-266).
void sph_grid_correction_xy(int nx, int ny, double alpha, int support, std::vector<double>& cx,
                            std::vector<double>& cy);

}  // namespace ygrid
