"""The small-support kernels (CF support <= 8: gridding into a 64x64 LDS tile per work item with
one item per key, degridding straight from each sample's footprint) against the 64x64 tile
kernels and against the FP32 oracle.

The context's small-support limit selects the kernels (ygrid_ctx_set_small_support: 0 = the tile
kernels only, 8 = the default); every case runs both and checks

* the small-support pass was taken (stats.small_support_passes) or not;
* grid and degridded visibilities of the two paths agree to 1e-6 of the peak (FP32 accumulation
  order is all that differs);
* both agree with the FP32 oracle to 1e-5 of the peak.

Cases: support 0 (Box), 3 (SphFunc, 8 lanes per sample), 5 and 8 (W-projection with
limitsupport, 16 and 32 lanes), 1..4 polarisations, several channels with flags and delays, FP16
and BF16 storage, PSF and preconditioner passes, a grid whose side is not a multiple of the
tile, and w stacking (one grid per w plane).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.test_gpu_parity import ARC, TOL, make_data, ochunk, pchunk, random_model, rel

pytestmark = pytest.mark.gpu

WP8 = dict(wmax=3000.0, nwplanes=9, oversample=4, maxsupport=256, cutoff=1e-2, limitsupport=8)


def _run(ctx, small, name, params, d, N, cell, npol, model, dopsf=False, dopcf=False):
    """grid (grid array) and degrid (vis cube) with the given small-support limit"""
    import yandasoft_amd as ya
    ctx.set_small_support(small)
    try:
        img = ya.ImageDesc(N, N, npol, 1, cell, cell)
        g = ya.Gridder(ctx, name, params)
        g.initialise_grid(img, dopsf=dopsf, dopcf=dopcf)
        g.grid(pchunk(d))
        out = np.zeros_like(d["vis"])
        dg = ya.Gridder(ctx, name, params)
        dg.initialise_degrid(img, model.astype(np.float32))
        dg.degrid(pchunk(d, vis_out=out))
        return g.grid_array(), out, g.stats(), dg.stats()
    finally:
        ctx.set_small_support(8)


def _oracle(kind_kw, d, N, cell, npol, model, dopsf=False, dopcf=False):
    o = O.OracleGridder(precision="f32", **kind_kw)
    o.initialise_grid((N, N, npol, 1), (cell, cell), dopsf=dopsf, dopcf=dopcf)
    o.grid(ochunk(d))
    od = O.OracleGridder(precision="f32", **kind_kw)
    od.initialise_degrid((N, N, npol, 1), (cell, cell), model)
    return o.grid_, np.transpose(od.degrid(ochunk(d)), (2, 1, 0)), o


CASES = [
    ("box S=0, 2 pols", "GPUBox", {}, dict(kind="Box"), 128, 20, 2, 2, {}),
    ("sphfunc S=3 os 128, 1 pol", "GPUSphFunc", {"oversample": 128}, dict(kind="SphFunc"), 256, 10, 1, 1, {}),
    ("sphfunc S=3 os 8, 4 pols, 3 ch, flags, delays, noise", "GPUSphFunc", {"oversample": 8},
     dict(kind="SphFunc", sph_oversample=8), 256, 10, 4, 3, dict(flags=True, delay=True, noise_var=True)),
    ("sphfunc S=3, 3 pols, grid 200 (edge tiles)", "GPUSphFunc", {"oversample": 16},
     dict(kind="SphFunc", sph_oversample=16), 200, 12, 3, 2, {}),
    ("wproject S<=5, 2 pols", "GPUWProject", dict(WP8, limitsupport=5),
     dict(kind="WProject", wproject=O.WProjectParams(**dict(WP8, limitsupport=5))), 256, 30, 2, 2, dict(w=True)),
    ("wproject S<=8, 4 pols", "GPUWProject", dict(WP8), dict(kind="WProject", wproject=O.WProjectParams(**WP8)), 256,
     30, 4, 2, dict(w=True)),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_small_support_matches_tile_kernels_and_oracle(gpu_ctx, case):
    label, name, params, okw, N, cell_as, npol, nchan, opt = case
    announce(f"small-support kernels: {label}")
    cell = cell_as * ARC
    kw = dict(opt)
    w = kw.pop("w", False)
    if w:
        kw.update(f0=1.0e9, frac=0.35, wmax_m=2500.0 * O.C_LIGHT / 1.005e9)
    d = make_data(2500, nchan, npol, N, cell, seed=60 + npol, **kw)
    model = random_model(N, npol, seed=3)
    gs, vs, sgs, sds = _run(gpu_ctx, 8, name, params, d, N, cell, npol, model)
    gt, vt, sgt, sdt = _run(gpu_ctx, 0, name, params, d, N, cell, npol, model)
    assert sgs.small_support_passes == 1 and sds.small_support_passes == 1
    assert sgt.small_support_passes == 0 and sdt.small_support_passes == 0
    assert sgs.support <= 8
    og, ov, o = _oracle(okw, d, N, cell, npol, model)
    assert sgs.samples_gridded == o.counters.samples_gridded and sgs.taps_gridded == o.counters.taps_gridded
    assert rel(gs, gt) < 1e-6 and rel(vs, vt) < 1e-6
    assert rel(gs.reshape(og.shape), og) < TOL
    assert rel(vs, ov) < TOL


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_small_support_16bit_storage_matches_tile_kernels(gpu_ctx, prec):
    """Both paths read the same rounded taps and coefficients and, degridding, round the model grid to
    the same 16-bit format (the tile kernels multiply on the 16-bit matrix cores, the small ones
    widen to FP32: the products are equal, their sums differ in order only)."""
    announce(f"small-support kernels: {prec} CF taps and coefficients vs the tile kernels")
    N, cell, npol = 256, 10 * ARC, 4
    d = make_data(2500, 2, npol, N, cell, seed=71)
    params = {"oversample": 32, "cfprecision": prec, "visprecision": prec}
    model = random_model(N, npol, seed=5)
    gs, vs, sgs, _ = _run(gpu_ctx, 8, "GPUSphFunc", params, d, N, cell, npol, model)
    gt, vt, _, _ = _run(gpu_ctx, 0, "GPUSphFunc", params, d, N, cell, npol, model)
    assert sgs.small_support_passes == 1 and sgs.cf_precision == {"fp16": 1, "bf16": 2}[prec]
    assert rel(gs, gt) < 1e-6 and rel(vs, vt) < 1e-6


@pytest.mark.parametrize("mode", ["psf", "pcf"])
def test_small_support_psf_and_pcf_passes(gpu_ctx, mode):
    announce(f"small-support kernels: {mode} pass vs the tile kernels and the oracle")
    N, cell = 128, 20 * ARC
    d = make_data(1500, 1, 1, N, cell, noise_var=True, seed=72)
    params = {"oversample": 32, "alldatapsf": True}
    model = random_model(N, 1)
    flags = dict(dopsf=mode == "psf", dopcf=mode == "pcf")
    gs, _, sgs, _ = _run(gpu_ctx, 8, "GPUSphFunc", params, d, N, cell, 1, model, **flags)
    gt, _, _, _ = _run(gpu_ctx, 0, "GPUSphFunc", params, d, N, cell, 1, model, **flags)
    og, _, _ = _oracle(dict(kind="SphFunc", sph_oversample=32), d, N, cell, 1, model, **flags)
    assert sgs.small_support_passes == 1
    assert rel(gs, gt) < 1e-6
    assert rel(gs.reshape(og.shape), og) < TOL


def test_small_support_w_stacking_matches_tile_kernels(gpu_ctx):
    """W stacking: degridding reads the sample's own w grid with the small-support kernel;
    gridding stays on the tile kernels (its (w grid, tile) keys hold few samples each)."""
    announce("small-support kernels: w stacking, 17 w grids, 2 pols, vs the tile kernels")
    N, cell, npol = 128, 30 * ARC, 2
    d = make_data(2000, 2, npol, N, cell, f0=1.0e9, frac=0.35, wmax_m=2500.0 * O.C_LIGHT / 1.005e9, seed=73)
    params = {"wmax": 3000.0, "nwplanes": 17}
    model = random_model(N, npol, seed=6)
    gs, vs, sgs, sds = _run(gpu_ctx, 8, "GPUWStack", params, d, N, cell, npol, model)
    gt, vt, _, _ = _run(gpu_ctx, 0, "GPUWStack", params, d, N, cell, npol, model)
    assert sgs.small_support_passes == 0 and sds.small_support_passes == 1
    assert rel(gs, gt) < 1e-6 and rel(vs, vt) < 1e-6


def test_small_support_limit_is_checked(gpu_ctx):
    import yandasoft_amd as ya
    with pytest.raises(ya.YgridError):
        gpu_ctx.set_small_support(9)
    with pytest.raises(ya.YgridError):
        gpu_ctx.set_small_support(-1)


@pytest.mark.parametrize("rows", [4, 8, 16])
@pytest.mark.parametrize("case", [CASES[2], CASES[4], CASES[5]], ids=["sphfunc S=3 4 pols", "wproject S<=5", "wproject S<=8"])
def test_small_support_band_rows_match_oracle(gpu_ctx, case, rows):
    """Every LDS band height of the small-support gridding kernel (ygrid_ctx_set_small_band_rows;
    4 is raised to 8 for supports up to 3) grids as the oracle does (1e-5) and as the default
    choice by support (1e-6: FP32 accumulation order)."""
    label, name, params, okw, N, cell_as, npol, nchan, opt = case
    announce(f"small-support band rows {rows}: {label}")
    cell = cell_as * ARC
    kw = dict(opt)
    if kw.pop("w", False):
        kw.update(f0=1.0e9, frac=0.35, wmax_m=2500.0 * O.C_LIGHT / 1.005e9)
    d = make_data(2500, nchan, npol, N, cell, seed=90 + npol, **kw)
    model = random_model(N, npol, seed=3)
    try:
        gpu_ctx.set_small_band_rows(rows)
        g, _, st, _ = _run(gpu_ctx, 8, name, params, d, N, cell, npol, model)
    finally:
        gpu_ctx.set_small_band_rows(0)
    g8, _, _, _ = _run(gpu_ctx, 8, name, params, d, N, cell, npol, model)
    assert st.small_support_passes == 1
    og, _, _ = _oracle(okw, d, N, cell, npol, model)
    assert rel(g.reshape(og.shape), og) < TOL
    assert rel(g, g8) < 1e-6


def test_small_support_band_rows_autotune(gpu_ctx):
    """ygrid_gridder_autotune_small times the three band heights on a clone and keeps the fastest
    in the context; the gridder's own grid is untouched and a later pass still matches the oracle."""
    import yandasoft_amd as ya
    label, name, params, okw, N, cell_as, npol, nchan, opt = CASES[2]
    announce("small-support band rows autotune")
    cell = cell_as * ARC
    d = make_data(2500, nchan, npol, N, cell, seed=97, **{k: v for k, v in opt.items() if k != "w"})
    img = ya.ImageDesc(N, N, npol, 1, cell, cell)
    g = ya.Gridder(gpu_ctx, name, params)
    g.initialise_grid(img)
    try:
        rows, ms = g.autotune_small(img, pchunk(d), trials=2)
        assert rows in (4, 8, 16) and ms > 0
        assert not np.any(g.grid_array())
        g.grid(pchunk(d))
        o = O.OracleGridder(precision="f32", **okw)
        o.initialise_grid((N, N, npol, 1), (cell, cell))
        o.grid(ochunk(d))
        assert rel(g.grid_array().reshape(o.grid_.shape), o.grid_) < TOL
        with pytest.raises(ya.YgridError):
            gpu_ctx.set_small_band_rows(12)
    finally:
        gpu_ctx.set_small_band_rows(0)
