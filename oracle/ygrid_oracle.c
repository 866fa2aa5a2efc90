/*
 * ygrid_oracle.c -- CPU ORACLE FOR THE TEST SUITE ONLY.
 *
 * This file is test infrastructure.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or as the timed
 * CPU baseline).  The product (yandasoft_amd/libygrid.so) never links or calls it.
 *
 * It restates, in plain C, the arithmetic of yandasoft's convolutional gridding path:
 *
 *   orc_nint ............ askap::nint (round half away from zero; the base-askap
 *                         header is not vendored, tie behaviour is "parity unpinned"
 *                         and fixtures avoid exact .5 ties)
 *   orc_sph_* ........... prolate spheroidal psi_alpha(c, nu) by the Legendre
 *                         eigen-expansion of apps/gridderExperiments.cc:218-234,409-494
 *                         (scimath::SpheroidalFunction is an absent dependency,
 *                         base-scimath tags/1.9.0; pinned by the Schwab m=6 rational
 *                         approximation quoted in askap/gridding/SphFuncVisGridder.cc:268-333
 *                         and by scipy.special.pro_ang1 in tests/)
 *   orc_generic_f32/f64 . TableVisGridder::generic (askap/gridding/TableVisGridder.cc:451-884)
 *                         with GridKernel::grid/degrid, pointer variant
 *                         (askap/gridding/GridKernel.cc:61-83, 106-125)
 *   orc_support_search .. SupportSearcher::findPeak/doSupportSearch
 *                         (askap/gridding/SupportSearcher.tcc:87-257) and
 *                         symmetricalSupport (askap/gridding/SupportSearcher.cc:108-117)
 *
 * The f32 variant keeps the reference's single-precision order of operations
 * (compile with -ffp-contract=off); the f64 variant is the FP64 oracle gridder.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_C 299792458.0
#define ORC_PI 3.14159265358979323846

/* askap::nint: nearest integer, halves away from zero. */
int orc_nint(double x) { return (x > 0.0) ? (int)(x + 0.5) : (int)(x - 0.5); }

/* ------------------------------------------------------------------------- */
/* Prolate spheroidal function (Legendre eigen-expansion).                     */
/* ------------------------------------------------------------------------- */

/* Helper matrix of the prolate spheroidal eigenproblem in the basis of normalised
 * associated Legendre functions P_{m+r}^m, r = 0,2,4,... (gridderExperiments.cc:409-427). */
static void orc_sph_matrix(double *B, int n, double c, int m) {
  const double c2 = c * c;
  memset(B, 0, sizeof(double) * (size_t)n * (size_t)n);
  for (int row = 0; row < n; ++row) {
    const int r = 2 * row;
    const int l = r + m;
    B[row * n + row] = (double)(l * (l + 1)) +
                       c2 * ((double)(2 * l + 3) * (l + m) * (l - m) + (double)(2 * l - 1) * (l + m + 1) * (l - m + 1)) /
                           ((double)(2 * l + 1) * (2 * l - 1) * (2 * l + 3));
    if (row >= 1) {
      B[row * n + row - 1] =
          c2 / (double)(2 * l - 1) * sqrt((double)(l + m) * (l + m - 1) * (l - m) * (l - m - 1) / ((double)(2 * l + 1) * (2 * l - 3)));
    }
    if (row + 1 < n) {
      B[row * n + row + 1] = c2 / (double)(2 * l + 3) *
                             sqrt((double)(l + m + 1) * (l + m + 2) * (l - m + 1) * (l - m + 2) / ((double)(2 * l + 1) * (2 * l + 5)));
    }
  }
}

/* Cyclic Jacobi rotations: eigenvector of the smallest eigenvalue of the symmetric B. */
static void orc_jacobi_min_evec(double *A, int n, double *evec) {
  double *V = (double *)calloc((size_t)n * n, sizeof(double));
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[p * n + q] * A[p * n + q];
    if (off < 1e-30) break;
    for (int p = 0; p < n; ++p) {
      for (int q = p + 1; q < n; ++q) {
        const double apq = A[p * n + q];
        if (fabs(apq) < 1e-300) continue;
        const double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
        for (int k = 0; k < n; ++k) {
          const double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = cs * akp - sn * akq;
          A[k * n + q] = sn * akp + cs * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = cs * apk - sn * aqk;
          A[q * n + k] = sn * apk + cs * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = cs * vkp - sn * vkq;
          V[k * n + q] = sn * vkp + cs * vkq;
        }
      }
    }
  }
  int imin = 0;
  for (int i = 1; i < n; ++i)
    if (A[i * n + i] < A[imin * n + imin]) imin = i;
  for (int i = 0; i < n; ++i) evec[i] = V[i * n + imin];
  free(V);
}

/* Coefficients of psi for bandwidth c and integer order m = alpha (nterms <= 64). */
int orc_sph_coeffs(double c, int m, int nterms, double *coeffs) {
  if (nterms < 2 || nterms > 64) return -1;
  double *B = (double *)malloc(sizeof(double) * (size_t)nterms * nterms);
  orc_sph_matrix(B, nterms, c, m);
  orc_jacobi_min_evec(B, nterms, coeffs);
  free(B);
  return 0;
}

/* Sum_r coeffs[r] * Pbar_{m+2r}^m(x), with Pbar normalised to unit L2 norm on [-1,1]. */
static double orc_legendre_series(const double *coeffs, int nterms, int m, double x) {
  /* Pbar_m^m = sqrt((2m+1)/2 * (2m-1)!!^2/(2m)!) (1-x^2)^(m/2), Condon-Shortley phase dropped
     (a constant sign cancels in the normalisation at x = 0). */
  double pmm = sqrt(0.5);
  const double s = sqrt(fmax(0.0, 1.0 - x * x));
  for (int k = 1; k <= m; ++k) pmm *= s * sqrt((2.0 * k + 1.0) / (2.0 * k));
  double pl_2 = 0.0, pl_1 = pmm; /* Pbar_{l-1}, Pbar_l for l = m */
  double sum = coeffs[0] * pmm;
  int l = m;
  const int lmax = m + 2 * (nterms - 1);
  /* normalised three-term recurrence in l at fixed m */
  double prev = 0.0, cur = pmm;
  (void)pl_2; (void)pl_1;
  while (l < lmax) {
    const int lp = l + 1;
    const double a = sqrt((4.0 * lp * lp - 1.0) / ((double)lp * lp - (double)m * m));
    const double b = (l == m) ? 0.0 : sqrt(((double)l * l - (double)m * m) / (4.0 * l * l - 1.0));
    const double next = a * (x * cur - b * prev);
    prev = cur;
    cur = next;
    l = lp;
    if (((l - m) & 1) == 0) sum += coeffs[(l - m) / 2] * cur;
  }
  return sum;
}

/* psi_alpha(c, nu) normalised to 1 at nu = 0; zero for |nu| >= 1. */
double orc_sph_value(const double *coeffs, int nterms, int m, double nu) {
  if (fabs(nu) >= 1.0) return 0.0;
  const double s0 = orc_legendre_series(coeffs, nterms, m, 0.0);
  const double s = orc_legendre_series(coeffs, nterms, m, nu);
  return (s / s0) * pow(1.0 - nu * nu, -(double)m / 2.0);
}

/* Schwab's m=6, alpha=1 rational approximation (SphFuncVisGridder.cc:276-332, kept there
 * as a comment); used only to pin orc_sph_value. */
double orc_schwab_grdsf(double nu) {
  static const double p[2][5] = {{8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1},
                                 {4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2}};
  static const double q[2][3] = {{1.0000000, 8.212018e-1, 2.078043e-1}, {1.0000000, 9.599102e-1, 2.918724e-1}};
  int part;
  double nuend;
  if (nu >= 0.0 && nu < 0.75) { part = 0; nuend = 0.75; }
  else if (nu >= 0.75 && nu <= 1.0) { part = 1; nuend = 1.0; }
  else return 0.0;
  double top = p[part][0], bot = q[part][0];
  const double d = nu * nu - nuend * nuend;
  for (int k = 1; k <= 4; ++k) top += p[part][k] * pow(d, k);
  for (int k = 1; k <= 2; ++k) bot += q[part][k] * pow(d, k);
  double v = (bot != 0.0) ? top / bot : 0.0;
  return v < 0.0 ? 0.0 : v;
}

/* ------------------------------------------------------------------------- */
/* Support search.                                                            */
/* ------------------------------------------------------------------------- */

/* in: complex double matrix, casacore column-major (ix fastest), nrow x ncol.
 * value > 0 replaces the peak amplitude (absolute cutoff form).
 * out[0..1] = peak pos, out[2..3] = blc, out[4..5] = trc; *peak = peak amplitude.
 * returns 0, or -1 (no peak), -2 (peak at edge), -3 (support not found). */
int orc_support_search(const double *in, int nrow, int ncol, double cutoff, double value, int *out, double *peak) {
  int px = 0, py = 0;
  double pval = -1.0;
  for (int iy = 0; iy < ncol; ++iy) {
    double tn = -1.0;
    int tx = 0, ty = 0;
    for (int ix = 0; ix < nrow; ++ix) {
      const double re = in[2 * ((size_t)iy * nrow + ix)], im = in[2 * ((size_t)iy * nrow + ix) + 1];
      const double n = re * re + im * im;
      if (tn < n) { tx = ix; ty = iy; tn = n; }
    }
    if (pval < 0 || pval * pval < tn) { px = tx; py = ty; pval = sqrt(tn); }
  }
  if (!(pval > 0.0)) return -1;
  if (value > 0) pval = value;
  *peak = pval;
  out[0] = px; out[1] = py;
  if (!(px > 0 && py > 0 && px + 1 < nrow && py + 1 < ncol)) return -2;
  const double thr = fabs(cutoff * pval) * fabs(cutoff * pval);
#define AMP(ix, iy) (in[2 * ((size_t)(iy) * nrow + (ix))] * in[2 * ((size_t)(iy) * nrow + (ix))] + \
                     in[2 * ((size_t)(iy) * nrow + (ix)) + 1] * in[2 * ((size_t)(iy) * nrow + (ix)) + 1])
  int blc0 = -1, blc1 = -1, trc0 = -1, trc1 = -1;
  for (int ix = 0; ix <= px; ++ix) if (AMP(ix, py) > thr) { blc0 = ix; break; }
  for (int iy = 0; iy <= py; ++iy) if (AMP(px, iy) > thr) { blc1 = iy; break; }
  for (int ix = nrow - 1; ix >= px; --ix) if (AMP(ix, py) > thr) { trc0 = ix; break; }
  for (int iy = ncol - 1; iy >= py; --iy) if (AMP(px, iy) > thr) { trc1 = iy; break; }
#undef AMP
  out[2] = blc0; out[3] = blc1; out[4] = trc0; out[5] = trc1;
  if (blc0 < 0 || blc1 < 0 || trc0 < 0 || trc1 < 0) return -3;
  return 0;
}

/* SupportSearcher::symmetricalSupport with centre = shape/2. */
int orc_symmetrical_support(const int *res, int nrow, int ncol) {
  const int c0 = nrow / 2, c1 = ncol / 2;
  const int l10 = res[4] - c0, l11 = res[5] - c1, l20 = c0 - res[2], l21 = c1 - res[3];
  const int xm = abs(l10) > abs(l20) ? abs(l10) : abs(l20);
  const int ym = abs(l11) > abs(l21) ? abs(l11) : abs(l21);
  return 2 * (xm > ym ? xm : ym);
}

/* ------------------------------------------------------------------------- */
/* TableVisGridder::generic                                                   */
/* ------------------------------------------------------------------------- */

enum { ORC_GRID = 0, ORC_PSF = 1, ORC_PCF = 2, ORC_DEGRID = 3 };

typedef struct {
  int nx, ny;          /* padded grid shape (itsShape(0), itsShape(1)) */
  int npol_img;        /* image polarisations (itsShape(2)) */
  int nchan_img;       /* image channels (itsShape(3)) */
  double uvcell_u, uvcell_v; /* itsUVCellSize (wavelengths) */
  int oversample;
  int mode;            /* ORC_GRID/PSF/PCF/DEGRID */
  int track_per_os;    /* itsTrackWeightPerOversamplePlane */
  int nsw_rows;        /* first dimension of itsSumWeights */
  int use_all_psf;     /* itsUseAllDataForPSF */
  int stale_offgrid;   /* degrid: add the previous on-grid value for off-grid samples (reference) */
  double max_sep;      /* itsMaxPointingSeparation (radians; <= 0: off) */
  double centre_lon, centre_lat; /* image centre (getImageCentre) */
} orc_params;

/* Per-row pointings and the PSF state that TableVisGridder keeps across calls
 * (itsFirstGriddedVis, itsFeedUsedForPSF, itsPointingUsedForPSF, TableVisGridder.cc:581-593)
 * and its count of rows rejected by the pointing test (:571-578). */
typedef struct {
  const double *pointing1; /* pointingDir1: nrow x 2 (lon, lat), or NULL (= the image centre) */
  const double *dish1;     /* dishPointing1: nrow x 2, or NULL (= pointing1) */
  int psf_set;             /* the representative feed and field are chosen */
  int psf_feed, psf_field; /* psf_field: psf_lon/psf_lat hold the representative dish pointing */
  double psf_lon, psf_lat;
  double rows_rejected;
} orc_rows;

/* MVDirection::separation: the great-circle angle between two directions */
static double orc_separation(double l1, double b1, double l2, double b2) {
  const double dl = l2 - l1;
  const double a = cos(b2) * sin(dl);
  const double b = cos(b1) * sin(b2) - sin(b1) * cos(b2) * cos(dl);
  const double c = sin(b1) * sin(b2) + cos(b1) * cos(b2) * cos(dl);
  return atan2(sqrt(a * a + b * b), c);
}

typedef struct {
  double samples_gridded, taps_gridded, samples_degridded, taps_degridded;
  double vectors_flagged, vectors_wflagged;
  int error; /* 0 ok; 1 zero noise weight; 2 bad frequency; 3 cf index overflow */
} orc_counters;

/* Shared inputs.  Cube layouts follow casacore column-major order:
 *   vis/noise/flag (nrow, nchan, npol)  -> row + nrow*(chan + nchan*pol)
 *   cmap (nrow, npol, nchan)            -> row + nrow*(pol + npol*chan)
 *   grid (nx, ny, npol_img, nchan_img)  -> x + nx*(y + ny*(pol + npol*chan))
 *   sumw (nsw_rows, npol_img, nchan_img)
 * CF plane k (k = fracu + os*(fracv + os*plane)) starts at cf_off[k], is (2S+1)^2 with
 * S = cf_supp[k], ix fastest.  cf_ofs_u/v (per plane before oversampling) may be NULL. */
#define ORC_GENERIC_BODY(REAL)                                                                                     \
  const int os = p->oversample;                                                                                    \
  const int isPSF = p->mode == ORC_PSF, isPCF = p->mode == ORC_PCF, forward = p->mode == ORC_DEGRID;               \
  REAL stale_r[4] = {0, 0, 0, 0}, stale_i[4] = {0, 0, 0, 0};                                                       \
  orc_rows local_rows = {NULL, NULL, 0, -1, 0, 0.0, 0.0, 0.0};                                                      \
  orc_rows *rw = rows ? rows : &local_rows;                                                                        \
  const double *dish = rw->dish1 ? rw->dish1 : rw->pointing1;                                                      \
  for (int i = 0; i < nrow; ++i) {                                                                                 \
    if (p->max_sep > 0. && rw->pointing1 &&                                                                        \
        orc_separation(p->centre_lon, p->centre_lat, rw->pointing1[2 * i], rw->pointing1[2 * i + 1]) > p->max_sep) { \
      rw->rows_rejected += 1;                                                                                      \
      continue;                                                                                                    \
    }                                                                                                              \
    if (!rw->psf_set && isPSF) {                                                                                   \
      if (!p->use_all_psf) {                                                                                       \
        rw->psf_feed = feed1 ? feed1[i] : 0;                                                                       \
        rw->psf_field = dish != NULL;                                                                              \
        if (dish) { rw->psf_lon = dish[2 * i]; rw->psf_lat = dish[2 * i + 1]; }                                    \
      }                                                                                                            \
      rw->psf_set = 1;                                                                                             \
    }                                                                                                              \
    for (int chan = 0; chan < nchan; ++chan) {                                                                     \
      const double f = freq[chan];                                                                                 \
      if (chan == 0) {                                                                                             \
        const double r = f / ORC_C;                                                                                \
        if (!(r > 0.1 && r < 30000)) { cnt->error = 2; return; }                                                   \
      }                                                                                                            \
      const double uScaled = f * uvw[3 * i + 0] / (ORC_C * p->uvcell_u);                                           \
      int iu = orc_nint(uScaled);                                                                                  \
      int fracu = orc_nint(os * ((double)iu - uScaled));                                                           \
      if (fracu < 0) { iu += 1; fracu += os; } else if (fracu >= os) { iu -= 1; fracu -= os; }                     \
      iu += p->nx / 2;                                                                                             \
      const double vScaled = f * uvw[3 * i + 1] / (ORC_C * p->uvcell_v);                                           \
      int iv = orc_nint(vScaled);                                                                                  \
      int fracv = orc_nint(os * ((double)iv - vScaled));                                                           \
      if (fracv < 0) { iv += 1; fracv += os; } else if (fracv >= os) { iv -= 1; fracv -= os; }                     \
      iv += p->ny / 2;                                                                                             \
      const double phase = 2.0f * ORC_PI * f * (delay ? delay[i] : 0.0) / ORC_C;                                   \
      const float ph_r = (float)cos(phase), ph_i = (float)sin(phase);                                              \
      int allPolGood = 1;                                                                                          \
      for (int pol = 0; pol < npol; ++pol)                                                                         \
        if (flag && flag[i + (size_t)nrow * (chan + (size_t)nchan * pol)]) { allPolGood = 0; break; }               \
      const int imageChan = chanmap ? chanmap[chan] : 0;                                                           \
      if (!(allPolGood && imageChan >= 0)) {                                                                       \
        if (!forward) cnt->vectors_flagged += 1;                                                                   \
        continue;                                                                                                  \
      }                                                                                                            \
      int wGood = 1;                                                                                               \
      REAL out_r[4] = {0, 0, 0, 0}, out_i[4] = {0, 0, 0, 0};                                                       \
      for (int pol = 0; pol < p->npol_img; ++pol) {                                                                \
        const int plane = cmap ? cmap[i + (size_t)nrow * (pol + (size_t)npol * chan)] : 0;                         \
        if (plane < 0) { wGood = 0; cnt->vectors_wflagged += 1; break; }                                           \
        const int cInd = fracu + os * (fracv + os * plane);                                                        \
        if (cInd >= ncf) { cnt->error = 3; return; }                                                               \
        const int support = cf_supp[cInd];                                                                         \
        const int ld = 2 * support + 1;                                                                            \
        const int iuo = iu + (cf_ofs_u ? cf_ofs_u[plane] : 0);                                                     \
        const int ivo = iv + (cf_ofs_v ? cf_ofs_v[plane] : 0);                                                     \
        GRID_T *g2 = grid + 2 * (size_t)p->nx * p->ny * (pol + (size_t)p->npol_img * imageChan);                   \
        const CF_T *cf = cf_data + 2 * cf_off[cInd];                                                               \
        if (!((iuo - support) > 0 && (ivo - support) > 0 && (iuo + support) < p->nx && (ivo + support) < p->ny)) { \
          if (forward && p->stale_offgrid) { out_r[pol] = stale_r[pol]; out_i[pol] = stale_i[pol]; }               \
          continue;                                                                                                \
        }                                                                                                          \
        if (forward) {                                                                                             \
          REAL cr = 0, ci = 0;                                                                                     \
          for (int sv = -support; sv <= support; ++sv) {                                                           \
            const CF_T *w = cf + 2 * (size_t)(sv + support) * ld;                                                  \
            const GRID_T *gp = g2 + 2 * ((size_t)(ivo + sv) * p->nx + (iuo - support));                           \
            for (int su = -support; su <= support; ++su, w += 2, gp += 2) {                                        \
              const REAL wr = w[0], wi = w[1], gr = gp[0], gi = gp[1];                                             \
              const REAL tr = wr * gr + wi * gi;                                                                   \
              const REAL ti = -wr * gi + wi * gr;                                                                  \
              cr += tr;                                                                                            \
              ci += ti;                                                                                            \
            }                                                                                                      \
          }                                                                                                        \
          cnt->samples_degridded += 1;                                                                             \
          cnt->taps_degridded += (double)ld * ld;                                                                  \
          if (visweight) { cr *= (REAL)visweight[chan]; ci *= (REAL)visweight[chan]; }                             \
          out_r[pol] = cr * ph_r - ci * ph_i;                                                                      \
          out_i[pol] = cr * ph_i + ci * ph_r;                                                                      \
          stale_r[pol] = out_r[pol]; stale_i[pol] = out_i[pol];                                                    \
        } else {                                                                                                   \
          const float nre = noise ? noise[2 * (i + (size_t)nrow * (chan + (size_t)nchan * pol))] : 1.0f;             \
          const float visNoise = nre * nre;                                                                        \
          const float wt = (visNoise > 0.) ? (float)(1. / visNoise) : 0.f;                                        \
          if (!(wt > 0.f)) { cnt->error = 1; return; }                                                             \
          const int swRow = p->track_per_os ? cInd : plane;                                                        \
          REAL rr, ri;                                                                                             \
          if (!isPSF && !isPCF) {                                                                                  \
            const size_t vi = 2 * (i + (size_t)nrow * (chan + (size_t)nchan * pol));                               \
            const REAL vr = vis[vi], vim = -(REAL)vis[vi + 1];                                                     \
            rr = ((REAL)ph_r * vr - (REAL)ph_i * vim) * wt;                                                        \
            ri = ((REAL)ph_r * vim + (REAL)ph_i * vr) * wt;                                                        \
            if (visweight) { rr *= (REAL)visweight[chan]; ri *= (REAL)visweight[chan]; }                           \
          } else {                                                                                                 \
            rr = wt; ri = 0;                                                                                       \
            if (isPSF && visweight) rr *= (REAL)visweight[chan];                                                   \
          }                                                                                                        \
          int doIt = 1;                                                                                            \
          if (isPSF && !p->use_all_psf)                                                                            \
            doIt = (feed1 ? feed1[i] : 0) == rw->psf_feed &&                                                       \
                   (!rw->psf_field ||                                                                              \
                    orc_separation(rw->psf_lon, rw->psf_lat, dish[2 * i], dish[2 * i + 1]) < 1e-6);                 \
          if (!doIt) continue;                                                                                     \
          int conjCF = 0;                                                                                          \
          if (isPCF) conjCF = (ivo < p->ny / 2 && iuo >= p->nx / 2) || (ivo <= p->ny / 2 && iuo < p->nx / 2);       \
          for (int sv = -support; sv <= support; ++sv) {                                                           \
            const CF_T *w = cf + 2 * (size_t)(sv + support) * ld;                                                  \
            GRID_T *gp = g2 + 2 * ((size_t)(ivo + sv) * p->nx + (iuo - support));                                  \
            for (int su = -support; su <= support; ++su, w += 2, gp += 2) {                                        \
              const REAL wr = w[0], wi = conjCF ? -(REAL)w[1] : (REAL)w[1];                                        \
              const REAL a = rr * wr - ri * wi;                                                                    \
              const REAL b = rr * wi + ri * wr;                                                                    \
              gp[0] += a;                                                                                          \
              gp[1] += b;                                                                                          \
            }                                                                                                      \
          }                                                                                                        \
          cnt->samples_gridded += 1;                                                                               \
          cnt->taps_gridded += (double)ld * ld;                                                                    \
          if (!isPCF) sumw[swRow + (size_t)p->nsw_rows * (pol + (size_t)p->npol_img * imageChan)] += wt;           \
        }                                                                                                          \
      }                                                                                                            \
      if (wGood) {                                                                                                 \
        if (forward) {                                                                                             \
          for (int pol = 0; pol < npol; ++pol) {                                                                   \
            VIS_T *o = vis_out + 2 * (i + (size_t)nrow * (chan + (size_t)nchan * pol));                             \
            o[0] += out_r[pol];                                                                                    \
            o[1] += out_i[pol];                                                                                    \
          }                                                                                                        \
        }                                                                                                          \
      } else if (!forward) {                                                                                       \
        cnt->vectors_wflagged += 1;                                                                                \
      }                                                                                                            \
    }                                                                                                              \
  }

/* FP32 restatement (grid/vis/CF single precision). Polarisation conversion is the
 * identity (image pols == vis pols), the in-scope configuration. */
void orc_generic_f32(const orc_params *p, int nrow, int nchan, int npol, const double *uvw, const double *delay,
                     const double *freq, const float *vis, const float *noise, const uint8_t *flag, const int *feed1,
                     const int *cmap, const int *chanmap, const float *visweight, const float *cf_data,
                     const int64_t *cf_off, const int *cf_supp, int ncf, const int *cf_ofs_u, const int *cf_ofs_v,
                     float *grid, double *sumw, float *vis_out, orc_counters *cnt, orc_rows *rows) {
#define GRID_T float
#define CF_T float
#define VIS_T float
  ORC_GENERIC_BODY(float)
#undef GRID_T
#undef CF_T
#undef VIS_T
}

/* FP64 oracle gridder: same loop with double CF, grid and accumulation. */
void orc_generic_f64(const orc_params *p, int nrow, int nchan, int npol, const double *uvw, const double *delay,
                     const double *freq, const float *vis, const float *noise, const uint8_t *flag, const int *feed1,
                     const int *cmap, const int *chanmap, const float *visweight, const double *cf_data,
                     const int64_t *cf_off, const int *cf_supp, int ncf, const int *cf_ofs_u, const int *cf_ofs_v,
                     double *grid, double *sumw, double *vis_out, orc_counters *cnt, orc_rows *rows) {
#define GRID_T double
#define CF_T double
#define VIS_T double
  ORC_GENERIC_BODY(double)
#undef GRID_T
#undef CF_T
#undef VIS_T
}

/* Vectorised psi evaluation for the Python wrapper. */
void orc_sph_values(const double *coeffs, int nterms, int m, const double *nu, int n, double *out) {
  for (int i = 0; i < n; ++i) out[i] = orc_sph_value(coeffs, nterms, m, nu[i]);
}
