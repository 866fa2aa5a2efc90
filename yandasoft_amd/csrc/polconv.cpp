// Polarisation conversion around the gridding passes.
//
// TableVisGridder::generic converts each sample's visibility vector from the accessor's
// polarisation frame to the image's before gridding (PolConverter::convert and ::noise,
// TableVisGridder.cc:517-524,682-693) and the degridded image-frame vector back to the
// accessor's frame before adding it to the visibilities (:860); a sample is used only when none
// of its visibility polarisations is flagged (:657-664).  PolConverter is part of base-scimath,
// which the reference does not include; the conversion here follows the casacore / IAU
// definitions of the polarisation products in terms of the Stokes parameters
//   XX = I + Q, XY = U + iV, YX = U - iV, YY = I - Q,
//   RR = I + V, RL = Q + iU, LR = Q - iU, LL = I - V,
// with out = S_out S_in^+ in (S: products x Stokes; S^+ = S^H (S S^H)^-1, the minimum-norm
// inverse: Stokes parameters that the inputs do not determine are taken as zero when predicting,
// and an image polarisation that the visibilities do not determine is an error when imaging).
// The noise of each output's real and imaginary parts combines the inputs' as independent
// errors.  No parallactic-angle rotation (itsPARotation) is applied.
//
// On the device the conversion is an elementwise pass over the chunk's cubes before the
// gridding pass (visibilities, noise, flags to the image frame) and after the degridding pass
// (image-frame model to the visibility frame, added to vis_out); the passes themselves run on
// image-frame cubes.
#include <algorithm>
#include <complex>
#include <vector>

#include "gridders.h"

namespace ygrid {

namespace {

using cd = std::complex<double>;

// Polarisation product in terms of (I, Q, U, V) for a casacore Stokes code.
void stokes_row(int code, cd (&r)[4]) {
  const cd i1(0., 1.);
  for (cd& x : r) x = 0.;
  switch (code) {
    case 1: r[0] = 1.; break;  // I
    case 2: r[1] = 1.; break;  // Q
    case 3: r[2] = 1.; break;  // U
    case 4: r[3] = 1.; break;  // V
    case 5: r[0] = 1.; r[3] = 1.; break;    // RR = I + V
    case 6: r[1] = 1.; r[2] = i1; break;    // RL = Q + iU
    case 7: r[1] = 1.; r[2] = -i1; break;   // LR = Q - iU
    case 8: r[0] = 1.; r[3] = -1.; break;   // LL = I - V
    case 9: r[0] = 1.; r[1] = 1.; break;    // XX = I + Q
    case 10: r[2] = 1.; r[3] = i1; break;   // XY = U + iV
    case 11: r[2] = 1.; r[3] = -i1; break;  // YX = U - iV
    case 12: r[0] = 1.; r[1] = -1.; break;  // YY = I - Q
    default: throw Error(1, "unsupported Stokes code " + std::to_string(code) + " (supported: I Q U V RR RL LR LL XX XY YX YY)");
  }
}

// Inverse of an n x n complex matrix (n <= 4) by Gauss-Jordan with partial pivoting.
std::vector<cd> invert(std::vector<cd> a, int n) {
  std::vector<cd> inv((size_t)n * n, 0.);
  for (int i = 0; i < n; ++i) inv[(size_t)i * n + i] = 1.;
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (std::abs(a[(size_t)r * n + c]) > std::abs(a[(size_t)piv * n + c])) piv = r;
    YG_CHECK(std::abs(a[(size_t)piv * n + c]) > 1e-12, 1,
             "the polarisation list is degenerate (repeated or dependent products)");
    if (piv != c)
      for (int k = 0; k < n; ++k) {
        std::swap(a[(size_t)piv * n + k], a[(size_t)c * n + k]);
        std::swap(inv[(size_t)piv * n + k], inv[(size_t)c * n + k]);
      }
    const cd d = a[(size_t)c * n + c];
    for (int k = 0; k < n; ++k) {
      a[(size_t)c * n + k] /= d;
      inv[(size_t)c * n + k] /= d;
    }
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      const cd f = a[(size_t)r * n + c];
      if (f == 0.) continue;
      for (int k = 0; k < n; ++k) {
        a[(size_t)r * n + k] -= f * a[(size_t)c * n + k];
        inv[(size_t)r * n + k] -= f * inv[(size_t)c * n + k];
      }
    }
  }
  return inv;
}

const char* stokes_name(int code) {
  static const char* names[] = {"?", "I", "Q", "U", "V", "RR", "RL", "LR", "LL", "XX", "XY", "YX", "YY"};
  return code >= 1 && code <= 12 ? names[code] : "?";
}

}  // namespace

std::vector<cd> pol_matrix(const std::vector<int>& from, const std::vector<int>& to, bool determined) {
  const int ni = (int)from.size(), no = (int)to.size();
  YG_CHECK(ni >= 1 && ni <= 4 && no >= 1 && no <= 4, 1, "1 to 4 polarisations per frame");
  std::vector<cd> Si((size_t)ni * 4), So((size_t)no * 4);
  for (int i = 0; i < ni; ++i) stokes_row(from[i], *reinterpret_cast<cd(*)[4]>(&Si[(size_t)i * 4]));
  for (int o = 0; o < no; ++o) stokes_row(to[o], *reinterpret_cast<cd(*)[4]>(&So[(size_t)o * 4]));
  // pinv(Si) = Si^H (Si Si^H)^-1  (4 x ni)
  std::vector<cd> G((size_t)ni * ni, 0.);
  for (int a = 0; a < ni; ++a)
    for (int b = 0; b < ni; ++b)
      for (int k = 0; k < 4; ++k) G[(size_t)a * ni + b] += Si[(size_t)a * 4 + k] * std::conj(Si[(size_t)b * 4 + k]);
  const std::vector<cd> Gi = invert(G, ni);
  std::vector<cd> P((size_t)4 * ni, 0.);
  for (int k = 0; k < 4; ++k)
    for (int b = 0; b < ni; ++b)
      for (int a = 0; a < ni; ++a) P[(size_t)k * ni + b] += std::conj(Si[(size_t)a * 4 + k]) * Gi[(size_t)a * ni + b];
  std::vector<cd> M((size_t)no * ni, 0.);
  for (int o = 0; o < no; ++o)
    for (int b = 0; b < ni; ++b)
      for (int k = 0; k < 4; ++k) M[(size_t)o * ni + b] += So[(size_t)o * 4 + k] * P[(size_t)k * ni + b];
  // imaging: every output must be determined by the inputs, M Si == So (predicting: Stokes
  // parameters the inputs do not hold are zero)
  for (int o = 0; o < no && determined; ++o)
    for (int k = 0; k < 4; ++k) {
      cd x = 0.;
      for (int b = 0; b < ni; ++b) x += M[(size_t)o * ni + b] * Si[(size_t)b * 4 + k];
      if (std::abs(x - So[(size_t)o * 4 + k]) > 1e-9) {
        std::string f;
        for (int c : from) f += std::string(f.empty() ? "" : ",") + stokes_name(c);
        throw Error(1, std::string("polarisation ") + stokes_name(to[o]) + " cannot be formed from " + f);
      }
    }
  for (cd& x : M) {  // exact small fractions (1/2, i/2) stay exact
    if (std::abs(x.real()) < 1e-15) x.real(0.);
    if (std::abs(x.imag()) < 1e-15) x.imag(0.);
  }
  return M;
}

namespace {

PolMatrix to_device_matrix(const std::vector<cd>& M, int no, int ni) {
  PolMatrix pm{};
  pm.nout = no;
  pm.nin = ni;
  for (int o = 0; o < no; ++o)
    for (int i = 0; i < ni; ++i)
      pm.m[o][i] = make_float2((float)M[(size_t)o * ni + i].real(), (float)M[(size_t)o * ni + i].imag());
  return pm;
}

}  // namespace

bool Gridder::needsPolConversion(const ygrid_chunk& a) const {
  if (!a.stokes || img_stokes_.empty()) return false;
  if (a.npol != (int)img_stokes_.size()) return true;
  for (int p = 0; p < a.npol; ++p)
    if (a.stokes[p] != img_stokes_[p]) return true;
  return false;
}

void Gridder::convertedPass(const ygrid_chunk& a, bool forward) {
  YG_CHECK(a.nrow >= 0 && a.nchan > 0 && a.npol > 0 && a.npol <= 4, 1, "bad chunk shape");
  const Ctx& c = *ctx_;
  const std::vector<int> vis_pols(a.stokes, a.stokes + a.npol);
  const int nimg = (int)img_stokes_.size();
  const size_t nsamp = (size_t)a.nrow * a.nchan;
  DeviceChunk dc(c, a);
  ygrid_chunk ic = dc.c;
  ic.npol = nimg;
  ic.stokes = nullptr;
  if (!forward) {
    // visibilities, noise and flags in the image frame
    const bool want_vis = a.vis && !(dopsf_ || dopcf_);
    const PolMatrix M = to_device_matrix(pol_matrix(vis_pols, img_stokes_), nimg, a.npol);
    const size_t cube = nsamp * nimg;
    DevBuf v(want_vis ? cube * sizeof(float2) : 0), n(cube * sizeof(float2)), f(dc.c.flag ? cube : 0);
    launch_pol_to_image(c, M, want_vis ? reinterpret_cast<const float2*>(dc.c.vis) : nullptr,
                        reinterpret_cast<const float2*>(dc.c.noise), dc.c.flag, nsamp,
                        want_vis ? v.as<float2>() : nullptr, n.as<float2>(), dc.c.flag ? f.as<uint8_t>() : nullptr);
    ic.vis = want_vis ? v.as<float>() : nullptr;
    ic.noise = n.as<float>();
    ic.flag = dc.c.flag ? f.as<uint8_t>() : nullptr;
    generic(ic, false);
    YG_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  // degridding: the image-frame model of every sample, then back to the visibility frame
  YG_CHECK(a.vis_out != nullptr, 1, "degridding needs vis_out");
  const PolMatrix Mb = to_device_matrix(pol_matrix(img_stokes_, vis_pols, false), a.npol, nimg);
  const size_t cube = nsamp * nimg, vcube = nsamp * a.npol;
  DevBuf model(std::max<size_t>(cube, 1) * sizeof(float2)), f(dc.c.flag ? cube : 0);
  launch_fill_u32(c, model.get(), 2 * cube, 0u);
  if (dc.c.flag) {
    // flags to the image frame (any visibility polarisation flagged: the sample is skipped)
    PolMatrix M1{};
    M1.nin = a.npol;
    M1.nout = nimg;
    launch_pol_to_image(c, M1, nullptr, nullptr, dc.c.flag, nsamp, nullptr, nullptr, f.as<uint8_t>());
  }
  ic.vis = nullptr;
  ic.noise = nullptr;
  ic.flag = dc.c.flag ? f.as<uint8_t>() : nullptr;
  ic.vis_out = model.as<float>();
  ic.vis_out_rows = 0;
  generic(ic, true);
  const int ld = a.vis_out_rows > 0 ? a.vis_out_rows : a.nrow;
  YG_CHECK(ld >= a.nrow, 1, "vis_out_rows must be 0 or at least nrow");
  if (a.mem == YGRID_MEM_DEVICE) {
    launch_pol_to_vis_add(c, Mb, model.as<float2>(), nsamp, reinterpret_cast<float2*>(a.vis_out), a.nrow, ld);
  } else if (vcube) {
    // the span of the (ld-row) cube this chunk's rows occupy
    const size_t span = (size_t)a.nrow + (size_t)ld * ((size_t)a.nchan * a.npol - 1);
    DevBuf out(span * sizeof(float2));
    YG_HIP(hipMemcpyAsync(out.get(), a.vis_out, span * sizeof(float2), hipMemcpyHostToDevice, c.stream));
    launch_pol_to_vis_add(c, Mb, model.as<float2>(), nsamp, out.as<float2>(), a.nrow, ld);
    YG_HIP(hipMemcpyAsync(a.vis_out, out.get(), span * sizeof(float2), hipMemcpyDeviceToHost, c.stream));
    YG_HIP(hipStreamSynchronize(c.stream));
  }
  YG_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace ygrid
