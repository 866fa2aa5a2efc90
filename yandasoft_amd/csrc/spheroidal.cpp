// Prolate spheroidal function and the 1-D grid correction, on the host (FP64).
//
// grdsf = scimath::SpheroidalFunction(c = pi*support, alpha) (SphFuncVisGridder.h:101-106,
// SphFuncVisGridder.cc:46): psi_alpha(c, nu) = S_{alpha,alpha}(c, nu) (1 - nu^2)^(-alpha/2),
// normalised to 1 at nu = 0.  S is expanded in associated Legendre functions P_{m+2r}^m;
// the expansion coefficients are the eigenvector of the lowest eigenvalue of the symmetric
// tridiagonal matrix of the prolate operator in the L2-normalised Legendre basis (the method
// of apps/gridderExperiments.cc:218-234,409-494).  Here the eigenvalue is isolated by Sturm
// bisection and the eigenvector found by inverse iteration (Thomas algorithm).
#include <algorithm>
#include <cmath>
#include <complex>

#include "ygrid_internal.h"

namespace ygrid {

namespace {

// Number of eigenvalues of the tridiagonal (d, e) below x (Sturm count).
int sturm_count(const std::vector<double>& d, const std::vector<double>& e, double x) {
  int cnt = 0;
  double q = 1.0;
  for (size_t i = 0; i < d.size(); ++i) {
    const double off = (i == 0) ? 0.0 : e[i - 1] * e[i - 1];
    q = d[i] - x - (i == 0 ? 0.0 : off / q);
    if (q == 0.0) q = -1e-300;
    if (q < 0) ++cnt;
  }
  return cnt;
}

}  // namespace

Spheroidal::Spheroidal(double c, double alpha, int nterms) : m_((int)alpha), d_(nterms, 0.0) {
  const int n = nterms;
  const double c2 = c * c;
  const int m = m_;
  std::vector<double> diag(n), off(n > 1 ? n - 1 : 1, 0.0);
  for (int k = 0; k < n; ++k) {
    const double l = m + 2.0 * k;
    diag[k] = l * (l + 1) + c2 * (2 * l * l + 2 * l - 1 - 2.0 * m * m) / ((2 * l - 1) * (2 * l + 3));
    if (k + 1 < n) {
      // <P_l^m | x^2 | P_{l+2}^m> in the normalised basis
      off[k] = c2 / (2 * l + 3) * std::sqrt((l + m + 1) * (l + m + 2) * (l - m + 1) * (l - m + 2) / ((2 * l + 1) * (2 * l + 5)));
    }
  }
  // Gershgorin interval, then bisection for the smallest eigenvalue.
  double lo = 1e300, hi = -1e300;
  for (int k = 0; k < n; ++k) {
    const double r = (k > 0 ? std::fabs(off[k - 1]) : 0.0) + (k + 1 < n ? std::fabs(off[k]) : 0.0);
    lo = std::min(lo, diag[k] - r);
    hi = std::max(hi, diag[k] + r);
  }
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (sturm_count(diag, off, mid) >= 1) hi = mid; else lo = mid;
    if (hi - lo <= 1e-14 * std::max(1.0, std::fabs(hi))) break;
  }
  const double lambda = 0.5 * (lo + hi) - 1e-10 * std::max(1.0, std::fabs(lo));
  // inverse iteration: solve (T - lambda) y = x with the Thomas algorithm
  std::vector<double> x(n, 1.0), cp(n), dp(n);
  for (int it = 0; it < 8; ++it) {
    double b0 = diag[0] - lambda;
    cp[0] = (n > 1 ? off[0] : 0.0) / b0;
    dp[0] = x[0] / b0;
    for (int k = 1; k < n; ++k) {
      const double den = (diag[k] - lambda) - off[k - 1] * cp[k - 1];
      cp[k] = (k + 1 < n ? off[k] : 0.0) / den;
      dp[k] = (x[k] - off[k - 1] * dp[k - 1]) / den;
    }
    x[n - 1] = dp[n - 1];
    for (int k = n - 2; k >= 0; --k) x[k] = dp[k] - cp[k] * x[k + 1];
    double nrm = 0.0;
    for (double v : x) nrm += v * v;
    nrm = std::sqrt(nrm);
    for (double& v : x) v /= nrm;
  }
  d_ = x;
  norm0_ = series(0.0);
}

// sum_k d_k * Pbar_{m+2k}^m(x) with Pbar_l^m = sqrt((2l+1)/2 (l-m)!/(l+m)!) P_l^m,
// P_l^m built by the upward recurrence (l-m+1) P_{l+1} = (2l+1) x P_l - (l+m) P_{l-1}.
double Spheroidal::series(double x) const {
  const int m = m_;
  const int n = (int)d_.size();
  double pmm = 1.0;  // (2m-1)!! (1-x^2)^(m/2)
  const double s = std::sqrt(std::max(0.0, 1.0 - x * x));
  for (int k = 1; k <= m; ++k) pmm *= (2.0 * k - 1.0) * s;
  double pprev = 0.0, p = pmm, sum = 0.0;
  const int lmax = m + 2 * (n - 1);
  for (int l = m; l <= lmax; ++l) {
    if (((l - m) & 1) == 0) {
      const double lnorm = 0.5 * (std::log(2.0 * l + 1.0) - std::log(2.0)) +
                           0.5 * (std::lgamma(l - m + 1.0) - std::lgamma(l + m + 1.0));
      sum += d_[(l - m) / 2] * p * std::exp(lnorm);
    }
    const double pnext = ((2.0 * l + 1.0) * x * p - (l + m) * pprev) / (l - m + 1.0);
    pprev = p;
    p = pnext;
  }
  return sum;
}

double Spheroidal::operator()(double nu) const {
  if (std::fabs(nu) >= 1.0) return 0.0;
  return series(nu) / norm0_ * std::pow(1.0 - nu * nu, -(double)m_ / 2.0);
}

// SphFuncVisGridder::correctConvolution 1-D factors (SphFuncVisGridder.cc:171-266): grdsf on
// each image axis, edge value interpolated (SphFuncVisGridder.tcc:37-51), low-pass filtered by
// keeping the 7 central coefficients of its centred transform, and inverted where
// |value| > 1e-10.  The 7 coefficients are computed directly (O(7n)).  Both axes are scaled by
// the one factor maxBefore/bufx(nx/2) taken from the x axis, including the 1/n of each axis'
// inverse transform (SphFuncVisGridder.cc:216-238), so the y factor of a non-square grid
// carries nx/ny exactly as the reference's does.
static std::vector<double> filtered_axis(const Spheroidal& sph, int n, double* max_before) {
  YG_CHECK(n > 3 && n % 2 == 0, 1, "grid correction needs an even axis length > 3");
  const int half = n / 2;
  std::vector<double> buf(n);
  for (int i = 0; i < n; ++i) buf[i] = sph(std::fabs(double(i - half)) / double(half));
  buf[0] = 3.0 * (buf[1] - buf[2]) + buf[3];
  if (max_before) *max_before = buf[half];
  const int filt = 3;
  std::complex<double> F[2 * filt + 1];
  for (int k = -filt; k <= filt; ++k) {
    std::complex<double> acc(0.0, 0.0);
    for (int i = 0; i < n; ++i) {
      const double ph = -2.0 * kPi * double(k) * double(i - half) / double(n);
      acc += buf[i] * std::complex<double>(std::cos(ph), std::sin(ph));
    }
    F[k + filt] = acc;
  }
  std::vector<double> val(n);
  for (int i = 0; i < n; ++i) {
    std::complex<double> acc(0.0, 0.0);
    for (int k = -filt; k <= filt; ++k) {
      const double ph = 2.0 * kPi * double(k) * double(i - half) / double(n);
      acc += F[k + filt] * std::complex<double>(std::cos(ph), std::sin(ph));
    }
    val[i] = acc.real() / double(n);
  }
  return val;
}

void sph_grid_correction_xy(int nx, int ny, double alpha, int support, std::vector<double>& cx,
                            std::vector<double>& cy) {
  const Spheroidal sph(kPi * support, alpha);
  double max_before = 1.0;
  const std::vector<double> vx = filtered_axis(sph, nx, &max_before);
  const std::vector<double> vy = filtered_axis(sph, ny, nullptr);
  const double scale = max_before / vx[nx / 2];
  cx.resize(nx);
  cy.resize(ny);
  for (int i = 0; i < nx; ++i) {
    const double v = vx[i] * scale;
    cx[i] = std::fabs(v) > 1e-10 ? 1.0 / v : 0.0;
  }
  for (int i = 0; i < ny; ++i) {
    const double v = vy[i] * scale;
    cy[i] = std::fabs(v) > 1e-10 ? 1.0 / v : 0.0;
  }
}

}  // namespace ygrid
