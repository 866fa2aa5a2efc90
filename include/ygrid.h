/*
 * ygrid.h -- C ABI of the MI355X-native convolutional gridder (yandasoft_amd/libygrid.so).
 *
 * This boundary replaces yandasoft's gridding path behind askap::synthesis::IVisGridder
 * (askap/gridding/IVisGridder.h:64-140) as implemented by TableVisGridder and its
 * Box/SphFunc/WProject/AWProject subclasses.  A plugin adapter deriving from
 * TableVisGridder forwards each virtual call to the function cited beside it.
 *
 * Conventions
 *  - every function returns YGRID_OK (0) or a non-zero status; the message of the last
 *    failure on the calling thread is returned by ygrid_last_error().  No C++ exception
 *    crosses this boundary (the adapter turns a non-zero status into ASKAPTHROW).
 *  - arrays follow casacore's column-major layouts so that an adapter can pass the
 *    accessor's storage without reordering:
 *      visibility / noise / flag cubes (nRow, nChan, nPol): row + nRow*(chan + nChan*pol)
 *      images (nx, ny, nPol, nChan): x + nx*(y + ny*(pol + nPol*chan))
 *    complex values are interleaved float pairs (casacore::Complex).
 *  - chunk arrays may live in host memory or in device (HBM) memory (ygrid_chunk.mem).
 *  - gridders are not thread-safe; clones are independent (IVisGridder.h:75).
 */
#ifndef YGRID_H
#define YGRID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YGRID_VERSION 1

enum ygrid_status {
  YGRID_OK = 0,
  YGRID_ERR_ARG = 1,              /* invalid argument */
  YGRID_ERR_STATE = 2,            /* call out of order (e.g. grid before initialise) */
  YGRID_ERR_DEVICE = 3,           /* HIP runtime / hipFFT failure */
  YGRID_ERR_DATA = 4,             /* data check failed (zero noise weight, w out of range, ...) */
  YGRID_ERR_UNKNOWN_GRIDDER = 5   /* factory name not registered */
};

enum ygrid_mem { YGRID_MEM_HOST = 0, YGRID_MEM_DEVICE = 1 };

typedef struct ygrid_ctx ygrid_ctx;
typedef struct ygrid_gridder ygrid_gridder;

/* Image geometry, the part of scimath::Axes + shape the gridder uses
 * (TableVisGridder::initialiseCellSize, TableVisGridder.cc:1038-1049;
 *  FrequencyMapper::setupImage, FrequencyMapper.cc:46-53). */
typedef struct {
  int nx, ny;           /* unpadded image size */
  int npol, nchan;      /* image polarisations and channels */
  double cellx, celly;  /* direction-axis increments (radians; sign is ignored) */
  int has_freq_axis;    /* FREQUENCY axis present: map accessor channels to image channels */
  double freq_start, freq_end;
  /* image centre direction (TableVisGridder::getImageCentre, TableVisGridder.cc:919-934),
   * longitude/latitude in radians; used by the A-projection beam offsets */
  double centre_lon, centre_lat;
  /* casacore Stokes codes of the image polarisations (npol; I=1 Q=2 U=3 V=4 RR=5 RL=6 LR=7 LL=8
   * XX=9 XY=10 YX=11 YY=12; host memory), the grid's stokes axis (TableVisGridder::initStokes,
   * TableVisGridder.cc:982-994), or NULL: the visibilities' own frame (no conversion) */
  const int32_t *stokes;
} ygrid_image_desc;

/* One accessor chunk (IConstDataAccessor / IDataAccessor getters used by
 * TableVisGridder::generic, TableVisGridder.cc:487-567). */
typedef struct {
  int nrow, nchan, npol;
  int mem;                   /* enum ygrid_mem, applies to every pointer below */
  const double *uvw;         /* rotatedUVW: nrow x 3 (metres), row-major (u,v,w) */
  const double *delay;       /* uvwRotationDelay: nrow (metres) or NULL (= 0) */
  const double *freq;        /* frequency: nchan (Hz) */
  const float *vis;          /* visibility cube (complex) */
  float *vis_out;            /* degrid: rwVisibility cube, model is ADDED in place */
  const float *noise;        /* noise cube (complex) or NULL (= 1) */
  const uint8_t *flag;       /* flag cube or NULL */
  const int32_t *feed1;      /* feed1: nrow or NULL (= 0) */
  const double *pointing1;   /* pointingDir1: nrow x 2 (longitude, latitude; rad) or NULL (= image centre) */
  /* stokes(): casacore Stokes codes of the visibility polarisations (npol, host memory), or NULL:
   * the image's frame.  When they differ from the image's, visibilities, noise and flags are
   * converted to the image frame before gridding and the model back to this frame after
   * degridding (PolConverter as used at TableVisGridder.cc:517-524,682-693,860; no parallactic
   * angle rotation) */
  const int32_t *stokes;
  /* dishPointing1: nrow x 2 (longitude, latitude; rad) or NULL (= pointing1): the representative
   * field of a PSF pass is the first used row's feed and dish pointing (TableVisGridder.cc:581-593,
   * 809-812) */
  const double *dish_pointing1;
  /* rows of the cube vis_out points into (its leading dimension), 0: nrow.  Direction-dependent
   * calibration degrids source s into rows s*nrow .. (s+1)*nrow-1 of a cube of nsource*nrow rows
   * (TableVisGridder.cc:567,861, itsSourceIndex): pass vis_out at row s*nrow and the cube's rows */
  int vis_out_rows;
  /* feed1PA: nrow float32 parallactic angle of each row's first feed (rad) or NULL (= 0).  Read
   * by A-projection with an illumination that rotates with it (ATCA with feed legs): the cache of
   * convolution functions is rebuilt when a row's angle differs from the cached one by more than
   * the gridder's patolerance (AProjectGridderBase.cc:202-253, AWProjectVisGridder.cc:369) */
  const float *feed1_pa;
} ygrid_chunk;

/* Counters reported by TableVisGridder's destructor (TableVisGridder.cc:182-317). */
typedef struct {
  double samples_gridded, taps_gridded;
  double samples_degridded, taps_degridded;
  double vectors_flagged, vectors_wflagged;
  double time_grid_s, time_degrid_s, time_cf_s;
  int support;        /* itsSupport: largest CF support */
  int oversample;     /* itsOverSample */
  int n_cf;           /* number of CF planes (incl. oversampling) */
  /* device time (HIP events on the gridder's stream) when profiling is enabled */
  double ms_prep, ms_bin, ms_grid_kernel, ms_degrid_kernel;
  double n_grid_launches, n_degrid_launches;
  double n_bin_reused; /* passes that reused the context's cached sample binning */
  /* tile-kernel variants of the context: waves per SIMD the kernel is built for, pipelined */
  int grid_waves, grid_pipeline, degrid_waves, degrid_pipeline;
  /* storage precision of the CF taps and of the gridding coefficients (parset keys cfprecision,
   * visprecision): 0 FP32, 1 FP16, 2 BF16; accumulation is FP32 in every case */
  int cf_precision, vis_precision;
  /* 1 when the gridder's CFs come from an attached learned network (ygrid_gridder_set_learned_cf) */
  int learned_cf_active;
  /* 1 when degridding runs on the matrix cores, four samples at once (2 or more polarisations;
   * environment YGRID_DEGRID_VARIANT="w,p,0" selects the VALU form) */
  int degrid_matrix;
  /* rows left out because their pointing lies farther than MaxPointingSeparation from the image
   * centre (itsRowsRejectedDueToMaxPointingSeparation, TableVisGridder.cc:571-578) */
  double rows_rejected_pointing;
  /* passes that ran the small-support kernels (CF support <= the context's small-support limit,
   * ygrid_ctx_set_small_support) instead of the 64x64 tile kernels */
  double small_support_passes;
  /* A-projection CF cache (AProjectGridderBase::updateStats, AProjectGridderBase.cc:117-140,
   * 272-282): calls of initConvolutionFunction, feed/field CF sets generated, and those of them
   * generated after the parallactic angle or the frequencies invalidated the cache */
  double cf_iterations, cf_generations, cf_generations_pa, cf_generations_freq;
  /* snapshot imaging (gridder.snapshotimaging): w-plane fits made, image-plane regrids done, and
   * the current plane w = snapshot_a u + snapshot_b v; zero for other gridders */
  double snapshot_fits, snapshot_regrids, snapshot_a, snapshot_b;
} ygrid_stats;

/* ---- library / context ---- */
const char *ygrid_last_error(void);
int ygrid_version(void);
/* Per-process device context; gridders created from it share its stream. */
int ygrid_ctx_create(int device, ygrid_ctx **out);
int ygrid_ctx_destroy(ygrid_ctx *ctx);
/* Use an existing hipStream_t (may be NULL for the default stream). */
int ygrid_ctx_set_stream(ygrid_ctx *ctx, void *hip_stream);
int ygrid_ctx_synchronize(ygrid_ctx *ctx);
/* Drops the context's cached sample binning (the next grid/degrid pass bins again).  A pass
 * reuses the binning of the previous pass on the context when its samples map to the same
 * pixels, CF planes and supports, as the degridding and gridding of one chunk do; no
 * counterpart in the reference (TableVisGridder recomputes the mapping on every pass). */
int ygrid_ctx_clear_cache(ygrid_ctx *ctx);
/* Passes whose CF support is at most max_support (0..8; default 8, environment
 * YGRID_SMALL_SUPPORT at context creation) run the small-support kernels: gridding accumulates
 * each 64x64 tile in LDS with one work item per tile, degridding reads every sample's
 * footprint directly without binning.  0 keeps every pass on the 64x64 tile kernels.  Same
 * results either way (within FP32 accumulation order); performance tuning with no counterpart
 * in the reference. */
int ygrid_ctx_set_small_support(ygrid_ctx *ctx, int max_support);
/* Rows (4, 8 or 16; 0, the default: 4 for supports 4 to 8, 8 below) of the LDS band a workgroup
 * of the small-support gridding kernel owns in its 64-pixel-wide tile; raised to the rows one
 * instruction covers (8 for supports up to 3).  Same results to FP32 accumulation order.  No
 * counterpart in the reference. */
int ygrid_ctx_set_small_band_rows(ygrid_ctx *ctx, int rows);
/* Autotuning of that LDS tile shape for one gridder and chunk: a clone of the gridder, set up for
 * img, grids the chunk `trials` times at each band height; the fastest (host clock around the
 * synchronised pass) becomes the context's setting and is returned with its time.  The gridder
 * itself is not changed.  A pass whose support is above the small-support limit runs the tile
 * kernels, on which the setting has no effect (*best_rows is then the unchanged setting).  No
 * counterpart in the reference. */
int ygrid_gridder_autotune_small(ygrid_gridder *g, const ygrid_image_desc *img, const ygrid_chunk *acc, int trials,
                                 int *best_rows, double *best_ms);

/* ---- factory: VisGridderFactory::make / createGridder (VisGridderFactory.cc:112-286) ----
 * name: "GPUBox", "GPUSphFunc", "GPUWProject", "GPUWStack", "GPUAWProject" or "GPUAProjectWStack"
 * (AProjectWStackVisGridder); the last two with illumination "disk" or "ATCA" and the
 * illumination.* keys of AProjectGridderBase::makeIllumination (AProjectGridderBase.cc:290-348).
 * parset: the gridder.<name>. subset as "key=value" lines (';' also separates); the global keys
 * of ygrid_gridder_configure may be given here too. */
int ygrid_gridder_create(ygrid_ctx *ctx, const char *name, const char *parset, ygrid_gridder **out);
/* The options VisGridderFactory::make applies to a created gridder (VisGridderFactory.cc:137-232),
 * as "key=value" text: padding (VisGridderWithPadding::setPaddingFactor), alldatapsf
 * (TableVisGridder::useAllDataForPSF), oversampleweight (trackWeightPerPlane), cleargrids
 * (doClearGrid), MaxPointingSeparation (maxPointingSeparation: radians, or a quantity such as
 * "1.5deg"; negative: off), parotation (doPARotation: only false is supported).  Call before
 * initialise_grid / initialise_degrid.
 * snapshotimaging=true wraps the gridder in SnapShotImagingGridderAdapter's behaviour
 * (VisGridderFactory.cc:248-270, SnapShotImagingGridderAdapter.cc): snapshotimaging.wtolerance
 * (metres, required), .clipping, .weightsclipping, .reprojectpsf, .interpmethod (cubic | linear),
 * .coorddecimation (accepted; coordinates are computed for every pixel); .longtrack is refused.
 * Every call on the handle then goes through the wrapper; ygrid_stats reports its fits and
 * regrids.  Not for ygrid_imaging_chunk, batches or ygrid_degrid_noalloc. */
int ygrid_gridder_configure(ygrid_gridder *g, const char *options);
/* IVisGridder::clone (IVisGridder.h:75; TableVisGridder.cc:126-168) */
int ygrid_gridder_clone(const ygrid_gridder *g, ygrid_gridder **out);
int ygrid_gridder_destroy(ygrid_gridder *g);

/* IVisGridder::initialiseGrid (IVisGridder.h:81-84; TableVisGridder.cc:997-1030) */
int ygrid_initialise_grid(ygrid_gridder *g, const ygrid_image_desc *img, int dopsf, int dopcf);
/* IVisGridder::grid (IVisGridder.h:88; TableVisGridder.cc:908-913 -> generic(acc,false)) */
int ygrid_grid(ygrid_gridder *g, const ygrid_chunk *acc);
/* IVisGridder::finaliseGrid (IVisGridder.h:92; TableVisGridder.cc:1109-1236).
 * out: nx*ny*npol*nchan floats (unpadded image), host or device per mem. */
int ygrid_finalise_grid(ygrid_gridder *g, float *out, int mem);
/* IVisGridder::finaliseWeights (IVisGridder.h:97; TableVisGridder.cc:1255-1284) */
int ygrid_finalise_weights(ygrid_gridder *g, float *out, int mem);
/* IVisGridder::initialiseDegrid (IVisGridder.h:102-103; TableVisGridder.cc:1286-1319).
 * model: nx*ny*npol*nchan floats (unpadded image), host or device per mem. */
int ygrid_initialise_degrid(ygrid_gridder *g, const ygrid_image_desc *img, const float *model, int mem);
/* IVisGridder::degrid (IVisGridder.h:115; TableVisGridder.cc:898-902 -> generic(acc,true)) */
int ygrid_degrid(ygrid_gridder *g, const ygrid_chunk *acc);
/* Allocation-free form of ygrid_degrid: no counterpart in the reference.  For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same result as ygrid_degrid on the same gridder and chunk, but every device buffer
 * of the pass (sample records, binning, work items, per-sample sums) comes from one fixed
 * workspace of YGRID_NOALLOC_WORKSPACE_BYTES that the library holds on each GPU until the process
 * ends; no hipMalloc/hipFree and no allocating library routine is called.  Limits: the chunk must
 * be in device memory (mem = YGRID_MEM_DEVICE) and in the image's polarisation frame; the
 * gridder's CF cache must exist (one ordinary ygrid_degrid pass on it or on the gridder it was
 * cloned from), and A-projection and learned-CF gridders are refused; a chunk whose pass needs
 * more workspace fails with YGRID_ERR_ARG and a message giving the size (split the chunk).  All
 * calls on one GPU share its workspace: the library serialises them with a per-GPU lock, so calls
 * that would overlap on one GPU wait for each other. */
#define YGRID_NOALLOC_WORKSPACE_BYTES (64u << 20)
int ygrid_degrid_noalloc(ygrid_gridder *g, const ygrid_chunk *acc);
/* IVisGridder::finaliseDegrid (IVisGridder.h:118; TableVisGridder.cc:1364-1372) */
int ygrid_finalise_degrid(ygrid_gridder *g);
/* IVisGridder::isModelEmpty (IVisGridder.h:139; TableVisGridder.cc:1467-1470) */
int ygrid_is_model_empty(const ygrid_gridder *g, int *out);
/* VisWeightsMultiFrequency via initVisWeights/customiseForContext (IVisGridder.h:107-111;
 * VisWeightsMultiFrequency.cc:59-75): order < 0 disables visibility weighting. */
int ygrid_set_mfs_weight(ygrid_gridder *g, double ref_freq, int order);

/* Gridder statistics (TableVisGridder.cc:182-317). */
int ygrid_get_stats(const ygrid_gridder *g, ygrid_stats *out);
/* Record HIP events around the prep, binning and grid/degrid kernels of every pass
 * (the reference's own timers, TableVisGridder.cc:465-500,879-883). */
int ygrid_set_profiling(ygrid_gridder *g, int enable);
/* Copy CF plane k (k = fracu + os*(fracv + os*plane), TableVisGridder.cc:720) to host:
 * (2S+1)^2 complex values, ix fastest.  *support receives S; out may be NULL to query S. */
int ygrid_get_cf(const ygrid_gridder *g, int k, float *out, int *support);
/* Copy the padded grid (complex, casacore layout) to host; *n receives the float count. */
int ygrid_get_grid(const ygrid_gridder *g, float *out, size_t *n);

/* ---- imaging-equation data loop ----
 * One chunk of ImageFFTEquation::calcImagingEquations' loop over the data
 * (ImageFFTEquation.cc:437-480), device-resident: the model visibilities are the sum of the
 * predictions of the `degridders` (initialised with initialise_degrid; those whose model is
 * empty add nothing, `:445-455`), the residual is data - model (`:464-465`), and every one of
 * the `gridders` (initialised with initialise_grid: residual, PSF and preconditioner gridders,
 * `:467-476`) grids that residual.  chunk->vis holds the data; chunk->vis_out, when not NULL,
 * receives the residual cube (the model is not added to it); chunk->mem applies to every array.
 * All gridders must belong to one context; either list may be empty. */
int ygrid_imaging_chunk(ygrid_gridder *const *degridders, int ndegrid, ygrid_gridder *const *gridders, int ngrid,
                        const ygrid_chunk *chunk);

/* ---- the accessor's coordinate step on the device ----
 * Replaces IConstDataAccessor::rotatedUVW(tangentPoint) and
 * ::uvwRotationDelay(tangentPoint, imageCentre) as TableVisGridder::generic calls them
 * (TableVisGridder.cc:481-495; the accessor is base-accessors', outside the reference).  Each
 * row's uvw (metres, frame of its phase centre: u east, v north, w toward the centre) is
 * rotated to the tangent point's frame (uvw_out, nrow x 3) and delay_out (nrow, or NULL) receives
 * the path difference b.s_centre - w that rephases the row from its phase centre to the image
 * centre, in the sense ygrid_chunk.delay takes it.  The outputs are ygrid_chunk's uvw and delay.
 * FP64; mem applies to every pointer; with YGRID_MEM_DEVICE the call is asynchronous on the
 * context's stream. */
typedef struct {
  int nrow;
  int mem;                     /* enum ygrid_mem */
  const double *uvw;           /* nrow x 3, row-major (u, v, w) */
  const double *phase_centre;  /* (lon, lat) radians: 2 values, or nrow x 2 when per_row */
  int phase_centre_per_row;
  double tangent_lon, tangent_lat;  /* the image's tangent point (reference direction) */
  double centre_lon, centre_lat;    /* the image centre */
  double *uvw_out;             /* nrow x 3 */
  double *delay_out;           /* nrow, or NULL */
} ygrid_uvw_rotation;
int ygrid_uvw_rotate(ygrid_ctx *ctx, const ygrid_uvw_rotation *r);

/* ---- batched passes: no counterpart in the reference, whose imaging equation runs one gridder
 * after the other (ImageFFTEquation.cc:362-387).  For callers that grid or degrid many images per
 * major cycle (beams x channels x polarisations of uneven sizes).
 * A batch holds passes (gridder, chunk in device memory); ygrid_batch_run gives every gridder the
 * result of ygrid_grid / ygrid_degrid on its chunk (by how the gridder was initialised), running
 * per pass the prep and binning kernels and then ONE persistent tile-kernel launch per (grid or
 * degrid, polarisations, CF precision) over the work items of all passes.  A gridder takes part
 * in one pass per batch; chunks must keep their addresses and shapes while the batch is used
 * (their contents may change between runs).  Not thread-safe; gridders of a batch share its
 * context. */
typedef struct ygrid_batch ygrid_batch;
typedef struct {
  int grid_waves, grid_pipeline, degrid_waves, degrid_pipeline; /* tile-kernel variant chosen */
  int item_pairs;   /* pairs per work item chosen */
  double ms;        /* best run time of the batch, milliseconds */
  int trials;       /* configurations timed */
} ygrid_tuning;
int ygrid_batch_create(ygrid_ctx *ctx, ygrid_batch **out);
int ygrid_batch_destroy(ygrid_batch *b);
int ygrid_batch_add(ygrid_batch *b, ygrid_gridder *g, const ygrid_chunk *acc);
/* enable: the next run captures the whole device sequence in a HIP graph, later runs replay it */
int ygrid_batch_set_graph(ygrid_batch *b, int enable);
int ygrid_batch_run(ygrid_batch *b);
/* Times the tile-kernel variants and work-item sizes on this batch (min over `trials` runs each)
 * and keeps the fastest on the context.  Each trial is a full run: grids and degridded
 * visibilities accumulate it, so initialise them again before the runs that count. */
int ygrid_batch_autotune(ygrid_batch *b, int trials, ygrid_tuning *best);
/* wall time of the last run (ms) and number of graph captures so far */
int ygrid_batch_info(const ygrid_batch *b, double *last_ms, int *captures);

/* ---- learned convolution functions: no counterpart in the reference (its CFs are FFT tables
 * on a fixed set of w planes, WProjectVisGridder.cc:435-522).  For callers that want CFs at each
 * sample's exact w instead of the nearest plane.
 * A coordinate network (dense layers of `hidden` units with SiLU, `layers` of them, complex output)
 * maps (w, fracu, fracv, dx, dy) to a CF tap.  It is trained on the GPU on a W-projection
 * gridder's own CF table; every holdout_every-th w plane is left out of training and the table
 * judges the network's error there.  A W-projection gridder with an attached network gives every
 * sample its own CF plane from the network at the sample's w (FP16 weights, FP32 accumulation),
 * as long as that held-out error is <= max_err; otherwise, and when detached, its table. */
typedef struct ygrid_lcf ygrid_lcf;
typedef struct {
  int epochs;          /* passes over the training taps */
  int batch;           /* examples per step */
  float lr;            /* peak Adam learning rate (cosine decay to lr/20) */
  unsigned long long seed;
  int holdout_every;   /* every n-th w plane is held out */
} ygrid_lcf_train_opts;
typedef struct {
  int epochs, steps, examples, cfs, holdout_cfs;
  double first_loss, last_loss;                /* mean over the epoch, targets scaled by 1/peak */
  double holdout_max_err, holdout_rms_err;     /* |network - table| / max|table| on held-out planes */
  double train_s;
  int support;                                 /* support of the network's planes */
} ygrid_lcf_report;
int ygrid_lcf_create(ygrid_ctx *ctx, int hidden, int layers, unsigned long long seed, ygrid_lcf **out);
int ygrid_lcf_destroy(ygrid_lcf *l);
int ygrid_lcf_num_params(const ygrid_lcf *l, int *n);
/* parameters (host, FP32): per layer the weights (inputs x outputs, row-major) then the biases */
int ygrid_lcf_get_params(const ygrid_lcf *l, float *out);
int ygrid_lcf_set_params(ygrid_lcf *l, const float *in);
/* forget the network's last evaluation: the next pass evaluates its samples' CFs even when they
 * equal the last ones (a pass over the samples of the previous pass otherwise reuses its planes).
 * No counterpart in the reference: its CFs are FFT tables. */
int ygrid_lcf_clear_cache(ygrid_lcf *l);
/* one forward and backward pass over n examples x (n x 5) with targets t (n x 2), host arrays:
 * loss = sum (y - t)^2 / (2n) and its gradient (num_params floats) */
int ygrid_lcf_loss_grad(ygrid_lcf *l, const float *x, const float *t, int n, double *loss, float *grad);
/* inference of n rows (host arrays); half: 0 FP32, 1 FP16 weights with FP32 activations, 2 the
 * fused FP16 matrix-core form (FP16 weights and activations, FP32 sums: the form gridding uses when
 * the network has hidden width 128 and at most 5 layers) */
int ygrid_lcf_forward(ygrid_lcf *l, const float *x, float *y, int n, int half);
/* trains on the CF table of a W-projection gridder (which must have run a pass) */
int ygrid_lcf_train_wproject(ygrid_lcf *l, const ygrid_gridder *wproject, const ygrid_lcf_train_opts *opts,
                             ygrid_lcf_report *report);
/* attach (lcf non-null) or detach a network; max_err bounds its held-out error */
int ygrid_gridder_set_learned_cf(ygrid_gridder *g, ygrid_lcf *lcf, double max_err);

#ifdef __cplusplus
}
#endif
#endif /* YGRID_H */
