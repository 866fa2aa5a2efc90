// Aperture illumination patterns of the A-projection gridder on the uv grid of the pattern buffer
// (IBasicIllumination::getPattern): the uniform disk (DiskIllumination.cc:85-177) and the ATCA
// model (ATCAIllumination.cc:159-331: a Jamesian taper with a defocusing phase, feed-leg shadows
// rotated with the parallactic angle and the wedges behind them).  One function serves the device
// kernel that fills the buffer and the host loop that sums the normalisation, so that both see the
// same values.  Products and sums are not fused, as in the reference's double arithmetic.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace ygrid {

enum IllumKind : int { kIllumDisk = 0, kIllumATCA = 1 };

struct IllumParams {
  double lscaled, mscaled;  // 2 pi cell l, 2 pi cell m: the pointing-offset phase slope
  double rmin2, rmax2;      // annulus (blockage, dish) in pattern cells^2
  float scale;              // nU nV / sqrt(sum |value|^2): the pattern's normalisation
  int kind;                 // IllumKind
  // ATCA: which effects are simulated
  int taper, legs, wedges;
  double defocus;           // maximum defocusing phase (rad), at the dish edge
  double leg_half;          // feed-leg half width (cells)
  double leg_angle;         // feed-leg rotation plus the parallactic angle (rad)
  double leg_shadow;        // feed-leg shadowing factor
  double wedge_tan;         // tan of the wedge opening angle
  double wedge_r0;          // wedge starting radius (cells)
  double wedge_s1, wedge_s2;  // wedge shadowing factors (legs along dU, along dV)
};

// ATCAIllumination::jamesian and its inner and outer forms (ATCAIllumination.cc:119-157; the
// constants are the reference's)
__host__ __device__ inline double atca_inner_jamesian(double fr) {
#pragma clang fp contract(off)
  const double dsub = 1.8 / 11.0;
  return 0.58 + 0.27 * atan(130 * (fr - 1.1 * dsub));
}
__host__ __device__ inline double atca_outer_jamesian(double fr) {
#pragma clang fp contract(off)
  const double a2 = 0.95, a1 = 8.;
  return 1. - a2 * exp(a1 * (fr * fr - 1.));
}
__host__ __device__ inline double atca_jamesian(double fr) {
#pragma clang fp contract(off)
  if (fr > 1.) return 0.;
  const double cd = 0.3, transition = 0.4;
  const double v = fr < transition ? atca_inner_jamesian(fr) / atca_inner_jamesian(transition)
                                   : atca_outer_jamesian(fr) / atca_outer_jamesian(transition);
  return v * (1. - cd * (1. - fr));
}

// The illumination at pattern offset (ou, ov) (cells from the buffer centre) inside the annulus,
// before the pointing-offset phase slope: (1, 0) for the disk.
__host__ __device__ inline void illum_value(const IllumParams& p, double ou, double ov, double r2, double& vr,
                                            double& vi) {
#pragma clang fp contract(off)
  vr = 1.0;
  vi = 0.0;
  if (p.kind != kIllumATCA) return;
  if (p.taper) {
    const double frs = r2 / p.rmax2;
    const double extra = p.defocus * frs;
    const double j = atca_jamesian(sqrt(frs));
    vr = j * cos(extra);
    vi = j * sin(extra);
  }
  if (p.legs) {
    // offsets in the frame of the legs; each shadow that covers the cell multiplies the value in
    // turn, as getPattern does
    const double ca = cos(p.leg_angle), sa = sin(p.leg_angle);
    const double dU = ca * ou + sa * ov;
    const double dV = -sa * ou + ca * ov;
    auto shade = [&](double f) {
      vr *= f;
      vi *= f;
    };
    if (fabs(dU) < p.leg_half || fabs(dV) < p.leg_half) shade(p.leg_shadow);
    if (p.wedges) {
      const double t = p.wedge_tan, r0 = p.wedge_r0;
      if (dU > r0 && fabs(dV) < t * fabs(dU - r0)) shade(p.leg_shadow * p.wedge_s1);
      if (dU < -r0 && fabs(dV) < t * fabs(dU + r0)) shade(p.leg_shadow * p.wedge_s1);
      if (dV > r0 && fabs(dU) < t * fabs(dV - r0)) shade(p.leg_shadow * p.wedge_s2);
      if (dV < -r0 && fabs(dU) < t * fabs(dV + r0)) shade(p.leg_shadow * p.wedge_s2);
    }
  }
}

}  // namespace ygrid
