"""16-bit storage modes (parset cfprecision / visprecision = fp16 | bf16) against the oracle.

The CF taps and the gridding coefficients (the weighted visibilities) are rounded to FP16 or
BF16, round to nearest even, when stored; products and sums stay FP32.

* Rounding pin: one visibility on an exact pixel grids the CF plane itself (frac 0, weight 1,
  phasor 1), so the grid patch must equal round(CF) * round(V) bit for bit, with numpy's
  float16 conversion (round to nearest even) and the BF16 carry rule as references.
* Error against the FP64 oracle gridder on the same inputs, per precision: the grid, the dirty
  image where the grid correction's gain is <= 10 (max |image - image64| / max |image64|), and
  the degridded visibilities.  The rounding of the CF taps is a white error floor in the image
  plane, which the correction (1/grdsf, up to ~1e3 at the image edge) lifts toward O(1) of the
  peak at the edge, so the edge is not a measure of the storage precision.  The bound BOUND is
  what the precision's rounding allows (FP16 keeps 11 significant bits: a relative step of
  2^-11 = 4.9e-4; BF16 8 bits: 3.9e-3); the measured values are printed.
* The target of 1e-4 of the peak for the 16-bit form is asserted for FP16 in
  test_fp16_reaches_1e4_of_peak, an expected failure: FP16 taps give 2-5e-4 on the grid, the
  image and the degridded visibilities at every support measured (3 to 14; the errors of one
  plane's taps are shared by all its samples), FP32 coefficients included, and a power-of-two
  scale of the table changes nothing (its taps lie far above FP16's subnormal range).  BF16 is
  reported only.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.test_gpu_parity import ARC, degrid_both, gain_of, make_data, ochunk, pchunk, random_model, rel

pytestmark = pytest.mark.gpu

PRECS = ("fp16", "bf16")
# grid, image (gain <= 10) and degridded-visibility error vs FP64, relative to the peak: four
# relative rounding steps of the precision (2^-11 FP16, 2^-8 BF16)
BOUND = {"fp16": 4 * 2.0 ** -11, "bf16": 4 * 2.0 ** -8}
TARGET = 1e-4  # the 16-bit form's image-error target against the FP64 oracle


def inner_err(img, img64, gain):
    peak = float(np.max(np.abs(img64)))
    err = np.abs(np.asarray(img, np.float64) - img64)
    return float(np.max(err[..., gain <= 10.0])) / peak


def round_np(x, prec):
    x = np.asarray(x, np.float32)
    if prec == "fp32":
        return x
    if prec == "fp16":
        return x.astype(np.float16).astype(np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def round_c(z, prec):
    z = np.asarray(z, np.complex64)
    return (round_np(z.real, prec) + 1j * round_np(z.imag, prec)).astype(np.complex64)


@pytest.mark.parametrize("tiles", [False, True])
@pytest.mark.parametrize("cfp,visp", [("fp32", "fp32"), ("fp16", "fp32"), ("bf16", "fp32"), ("fp16", "fp16"),
                                      ("bf16", "bf16")])
def test_single_sample_grids_rounded_cf(gpu_ctx, cfp, visp, tiles):
    """tiles: the 64x64 tile kernels instead of the small-support ones (support 3); with both
    storage precisions 16-bit they multiply on the 16-bit matrix cores, whose FP32 sums of
    16-bit products are exact here too."""
    announce(f"single sample, cfprecision={cfp} visprecision={visp}, tile kernels {tiles}")
    import yandasoft_amd as ya
    if tiles:
        gpu_ctx = ya.Context(0)
        gpu_ctx.set_small_support(0)
    N, cell, f = 256, 30 * ARC, 1.4e9
    du = 1.0 / (cell * N)
    lam = O.C_LIGHT / f
    m = (10, -7)
    uvw = np.array([[m[0] * du * lam, m[1] * du * lam, 0.0]])
    V = np.float32(1.0 + 3.0 * 2.0 ** -13)  # rounds to 1.0 in FP16 and BF16
    vis = np.full((1, 1, 1), V, np.complex64)
    params = {"alpha": 1.0, "support": 3, "oversample": 128, "cfprecision": cfp, "visprecision": visp}
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", params)
    g.initialise_grid(ya.ImageDesc(N, N, 1, 1, cell, cell))
    g.grid(ya.Chunk(uvw=uvw, freq=np.array([f]), vis=vis, noise=np.ones_like(vis)))
    st = g.stats()
    assert (st.cf_precision, st.vis_precision) == ({"fp32": 0, "fp16": 1, "bf16": 2}[cfp],
                                                   {"fp32": 0, "fp16": 1, "bf16": 2}[visp])
    grid = g.grid_array().reshape(N, N)
    cf = g.cf_plane(0)
    S = (cf.shape[0] - 1) // 2
    iu, iv = m[0] + N // 2, m[1] + N // 2
    patch = grid[iv - S:iv + S + 1, iu - S:iu + S + 1]
    want = round_c(cf, cfp) * round_np(V, visp)
    assert np.array_equal(patch, want.astype(np.complex64))
    # nothing outside the footprint
    assert np.count_nonzero(grid) == np.count_nonzero(want)


WP = dict(wmax=3000.0, nwplanes=9, oversample=4, maxsupport=256, cutoff=1e-2)


def _wproject_errors(gpu_ctx, prec, visprec=None):
    """(grid, image (gain <= 10), degrid) errors of W-projection at storage precision prec vs FP64,
    and the FP32 mode's grid and image errors."""
    import yandasoft_amd as ya
    N, cell = 256, 30 * ARC
    d = make_data(1500, 2, 4, N, cell, f0=1.0e9, frac=0.35, wmax_m=2500.0 * O.C_LIGHT / 1.005e9)
    prm = O.WProjectParams(**WP)
    o = O.OracleGridder("WProject", "f64", wproject=prm)
    o.initialise_grid((N, N, 4, 1), (cell, cell))
    o.grid(ochunk(d))
    grid64 = o.grid_.copy()
    img64 = o.finalise_grid()
    gain = gain_of(o)
    eg, eh = {}, {}
    for p in ("fp32", prec):
        params = dict(WP, cfprecision=p, visprecision=visprec or p)
        g = ya.Gridder(gpu_ctx, "GPUWProject", params)
        g.initialise_grid(ya.ImageDesc(N, N, 4, 1, cell, cell))
        g.grid(pchunk(d))
        eg[p] = rel(g.grid_array().reshape(grid64.shape), grid64)
        eh[p] = inner_err(g.finalise_grid(), img64, gain)
    # degrid with the same storage precision against the FP64 oracle
    model = random_model(N, 4)
    gv, _, _, _ = degrid_both(gpu_ctx, "GPUWProject", dict(WP, cfprecision=prec, visprecision=visprec or prec),
                              dict(kind="WProject", wproject=prm), d, N, cell, 4, model)
    od = O.OracleGridder("WProject", "f64", wproject=prm)
    od.initialise_degrid((N, N, 4, 1), (cell, cell), model)
    ov64 = od.degrid(ochunk(d))
    ev = rel(gv, ov64)
    print(f"[precision] wproject cf {prec} vis {visprec or prec}: grid err {eg[prec]:.3e} (fp32 {eg['fp32']:.3e}), "
          f"image err (gain<=10) {eh[prec]:.3e} (fp32 {eh['fp32']:.3e}), degrid err {ev:.3e}", flush=True)
    return eg, eh, ev


@pytest.mark.parametrize("prec", PRECS)
def test_wproject_mixed_precision_error_vs_fp64(gpu_ctx, prec):
    announce(f"wproject cf/vis precision {prec} vs fp64 oracle")
    eg, eh, ev = _wproject_errors(gpu_ctx, prec)
    assert eg[prec] <= BOUND[prec] and eh[prec] <= BOUND[prec] and ev <= BOUND[prec]
    assert eg[prec] > 10 * eg["fp32"]  # the 16-bit mode is in effect


@pytest.mark.parametrize("prec", PRECS)
def test_wproject_16bit_matrix_cores_and_widened_form_agree(gpu_ctx, prec):
    """visprecision equal to cfprecision runs the 16-bit matrix-core tile kernels (coefficients,
    and when degridding the model grid, rounded to the same format); visprecision fp32 runs the
    FP32 ones on widened taps.  Both stay within the precision's bound against FP64."""
    announce(f"wproject {prec}: 16-bit matrix-core form vs widened form")
    eg, eh, ev = _wproject_errors(gpu_ctx, prec, visprec="fp32")
    assert eg[prec] <= BOUND[prec] and eh[prec] <= BOUND[prec] and ev <= BOUND[prec]
    eg2, eh2, ev2 = _wproject_errors(gpu_ctx, prec, visprec=prec)
    assert eg2[prec] <= BOUND[prec] and eh2[prec] <= BOUND[prec] and ev2 <= BOUND[prec]
    assert eg2[prec] != eg[prec]  # the coefficients' rounding is in effect


@pytest.mark.xfail(strict=True, reason="FP16 CF taps (11-bit mantissa) give 2-5e-4 of the peak against the FP64 "
                                        "oracle, above the 1e-4 target, with FP32 coefficients too")
def test_fp16_reaches_1e4_of_peak(gpu_ctx):
    announce("wproject fp16 taps, fp32 coefficients: the 1e-4 target")
    eg, eh, ev = _wproject_errors(gpu_ctx, "fp16", visprec="fp32")
    assert eg["fp16"] <= TARGET and eh["fp16"] <= TARGET and ev <= TARGET


@pytest.mark.parametrize("prec", PRECS)
def test_awproject_mixed_precision_error_vs_fp64(gpu_ctx, prec):
    announce(f"awproject disk, 3 feeds, cf/vis precision {prec} vs fp64 oracle")
    import yandasoft_amd as ya
    N, cell = 128, 60 * ARC
    centre = (0.5, -0.7)
    nrow = 900
    d = make_data(nrow, 2, 1, N, cell, f0=1.4e9, df=1e7, frac=0.3, wmax_m=0.0)
    feed = (np.arange(nrow) % 3).astype(np.int32)
    off = math.radians(1.0)
    point = np.zeros((nrow, 2))
    point[:, 0] = centre[0] + (feed - 1) * off
    point[:, 1] = centre[1]
    params = {"wmax": 100.0, "nwplanes": 1, "oversample": 4, "maxsupport": 128, "maxfeeds": 3, "cutoff": 0.01,
              "diameter": "12m", "blockage": "0m", "cfprecision": prec, "visprecision": prec}
    prm = O.AWProjectParams(wmax=100.0, nwplanes=1, oversample=4, maxsupport=128, maxfeeds=3, cutoff=0.01,
                            diameter=12.0, blockage=0.0)
    g = ya.Gridder(gpu_ctx, "GPUAWProject", params)
    img = ya.ImageDesc(N, N, 1, 1, cell, cell, centre=centre)
    g.initialise_grid(img)
    ch = ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], noise=d["noise"], feed1=feed,
                  pointing1=np.ascontiguousarray(point))
    g.grid(ch)
    o = O.OracleGridder("AWProject", "f64", wproject=prm)
    o.aw_lm = {k: O.direction_lm(tuple(point[k]), centre) for k in range(3)}
    o.aw_nchan = 2
    o.initialise_grid((N, N, 1, 1), (cell, cell))
    och = ochunk(d)
    och.feed1 = feed
    o.grid(och)
    eg = rel(g.grid_array().reshape(o.grid_.shape), o.grid_)
    eh = inner_err(g.finalise_grid(), o.finalise_grid(), gain_of(o))
    model = random_model(N, 1)
    dg = ya.Gridder(gpu_ctx, "GPUAWProject", params)
    dg.initialise_degrid(img, model.astype(np.float32))
    out = np.zeros_like(d["vis"])
    dg.degrid(ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], vis_out=out, noise=d["noise"], feed1=feed,
                       pointing1=np.ascontiguousarray(point)))
    od = O.OracleGridder("AWProject", "f64", wproject=prm)
    od.aw_lm, od.aw_nchan = o.aw_lm, 2
    od.initialise_degrid((N, N, 1, 1), (cell, cell), model)
    ev = rel(np.transpose(out, (2, 1, 0)), od.degrid(och))
    print(f"[precision] awproject {prec}: grid err {eg:.3e}, image err (gain<=10) {eh:.3e}, degrid err {ev:.3e}",
          flush=True)
    assert eg <= BOUND[prec] and eh <= BOUND[prec] and ev <= BOUND[prec]


def test_bad_precision_is_rejected(gpu_ctx):
    import yandasoft_amd as ya
    with pytest.raises(ya.YgridError) as e:
        ya.Gridder(gpu_ctx, "GPUSphFunc", {"cfprecision": "fp8"})
    assert e.value.code == 1 and "cfprecision" in str(e.value)
