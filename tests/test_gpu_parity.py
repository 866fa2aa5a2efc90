"""HIP gridder (through the C ABI) against the CPU oracle on identical seeded inputs.

Bar: max|gpu - oracle| / max|oracle| <= 1e-5 (FP32) for the grid, the sum of weights and the
degridded visibilities; CF planes to 1e-5 of their peak.

Dirty images (image_ok): the grid correction multiplies the image edges by 1/grdsf (up to ~1e4
at 128^2), which lifts the rounding of any FP32 grid accumulation there -- the reference's
included: the oracle's own FP32 restatement differs from its FP64 form by ~4e-5 of the peak at
the edges.  An image is therefore pinned in three parts, none statistical: against the FP64
oracle image, <= 1e-5 of the peak where the correction gain is <= 10 and, everywhere, <= 1e-5 of
the peak once divided by the gain (the gridder's error before the correction amplifies it);
and, everywhere including the edges, equal to the oracle's FP64 finalise (transform,
correction, crop) of the GPU's own grid to 1e-6 of its peak (the FP32 output rounding).  With
the grid parity above this fixes the whole image.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce

TOL = 1e-5
ARC = math.pi / 180 / 3600

pytestmark = pytest.mark.gpu


def rel(a, b):
    b = np.asarray(b)
    return float(np.max(np.abs(np.asarray(a) - b)) / max(np.max(np.abs(b)), 1e-300))


def image_ok(gi, oi64, gain, refinalise=None):
    """gain: (ny, nx) grid-correction factor of the oracle gridder; refinalise: callable returning
    the oracle's FP64 finalise of the GPU's own grid (None: not checked)."""
    peak = max(float(np.max(np.abs(oi64))), 1e-300)
    err = np.abs(np.asarray(gi, np.float64) - oi64)
    inner = float(np.max(err[..., gain <= 10.0])) / peak
    normed = float(np.max(err / gain)) / peak
    fin = 0.0
    if refinalise is not None:
        f = refinalise()
        fin = float(np.max(np.abs(np.asarray(gi, np.float64) - f))) / max(float(np.max(np.abs(f))), 1e-300)
    return inner <= TOL and normed <= TOL and fin <= 1e-6, (inner, normed, fin)


def finalise_with(o, grid):
    """The oracle's finaliseGrid (FP64) applied to another grid of the same shape."""
    saved = o.grid_
    o.grid_ = np.asarray(grid, np.complex128).reshape(saved.shape)
    try:
        return o.finalise_grid()
    finally:
        o.grid_ = saved


def gain_of(o):
    return o.correct(np.ones((1, 1, o.ny, o.nx)))[0, 0]


def make_data(nrow, nchan, npol, N, cell, f0=1.4e9, df=5e6, frac=0.45, wmax_m=0.0, seed=7, flags=False,
              delay=False, noise_var=False):
    rng = np.random.default_rng(seed)
    freq = f0 + df * np.arange(nchan)
    lam = O.C_LIGHT / freq.max()
    uvmax = frac * (N / 2) * (1 / (cell * N)) * lam
    r = uvmax * np.sqrt(rng.random(nrow))
    th = 2 * np.pi * rng.random(nrow)
    w = (rng.random(nrow) * 2 - 1) * wmax_m
    uvw = np.ascontiguousarray(np.stack([r * np.cos(th), r * np.sin(th), w], 1))
    vis = (rng.normal(size=(npol, nchan, nrow)) + 1j * rng.normal(size=(npol, nchan, nrow))).astype(np.complex64)
    noise = np.ones_like(vis)
    if noise_var:
        noise = (0.5 + rng.random((npol, nchan, nrow))).astype(np.float32).astype(np.complex64)
    flag = None
    if flags:
        flag = (rng.random((npol, nchan, nrow)) < 0.05).astype(np.uint8)
    dl = rng.normal(size=nrow) * 0.3 if delay else None
    return dict(uvw=uvw, freq=freq, vis=vis, noise=noise, flag=flag, delay=dl)


def ochunk(d):
    t = lambda a: None if a is None else np.transpose(a, (2, 1, 0))
    return O.Chunk(d["uvw"], d["freq"], t(d["vis"]), noise=t(d["noise"]),
                   flag=None if d["flag"] is None else t(d["flag"]).astype(bool), delay=d["delay"])


def pchunk(d, vis_out=None):
    import yandasoft_amd as ya
    return ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], vis_out=vis_out, noise=d["noise"], flag=d["flag"],
                    delay=d["delay"])


def grid_both(ctx, name, params, ogrid_kwargs, d, N, cell, npol, nchan_img=1, dopsf=False, dopcf=False,
              freq_axis=None, mfs=None, extra_oracle=None):
    """Grids d with the GPU gridder and the oracle (FP32 and FP64); checks the dirty image
    (image_ok) and returns (gpu gridder, f32 oracle, gpu grid, oracle grid, gpu image,
    f32 oracle image, gpu weights, oracle weights)."""
    import yandasoft_amd as ya
    g = ya.Gridder(ctx, name, params)
    if mfs:
        g.set_mfs_weight(*mfs)
    img = ya.ImageDesc(N, N, npol, nchan_img, cell, cell, freq_axis)
    g.initialise_grid(img, dopsf=dopsf, dopcf=dopcf)
    g.grid(pchunk(d))
    ggrid = g.grid_array().reshape(nchan_img, npol, N, N)
    gimg = g.finalise_grid()
    gw = g.finalise_weights()
    imgs = {}
    for prec in ("f64", "f32"):
        o = O.OracleGridder(precision=prec, **ogrid_kwargs)
        for k, v in (extra_oracle or {}).items():
            setattr(o, k, v)
        if mfs:
            o.visweight = lambda f, r=mfs[0], k=mfs[1]: ((f - r) / r) ** k
        o.initialise_grid((N, N, npol, nchan_img), (cell, cell), dopsf=dopsf, dopcf=dopcf, freq_axis=freq_axis)
        o.grid(ochunk(d))
        imgs[prec] = o.finalise_grid()
    ow = o.finalise_weights()
    ok, info = image_ok(gimg, imgs["f64"], gain_of(o), lambda: finalise_with(o, ggrid))
    assert ok, f"image parity (inner, gain-normalised, own-grid finalise) = {info}"
    return g, o, ggrid, o.grid_, gimg, imgs["f32"], gw, ow


def degrid_both(ctx, name, params, ogrid_kwargs, d, N, cell, npol, model, nchan_img=1, freq_axis=None):
    import yandasoft_amd as ya
    g = ya.Gridder(ctx, name, params)
    img = ya.ImageDesc(N, N, npol, nchan_img, cell, cell, freq_axis)
    g.initialise_degrid(img, model.astype(np.float32))
    out = np.zeros_like(d["vis"])
    g.degrid(pchunk(d, vis_out=out))
    o = O.OracleGridder(precision="f32", **ogrid_kwargs)
    o.initialise_degrid((N, N, npol, nchan_img), (cell, cell), model, freq_axis=freq_axis)
    ov = o.degrid(ochunk(d))
    return np.transpose(out, (2, 1, 0)), ov, g, o


def random_model(N, npol, nchan=1, seed=11, npts=20):
    rng = np.random.default_rng(seed)
    m = np.zeros((nchan, npol, N, N))
    for _ in range(npts):
        m[:, :, rng.integers(N // 4, 3 * N // 4), rng.integers(N // 4, 3 * N // 4)] += rng.normal(size=(nchan, npol))
    return m


# ---------------------------------------------------------------------------------------
def test_sphfunc_cf_planes_match_oracle(gpu_ctx):
    announce("sphfunc cf planes S=3 os=128")
    import yandasoft_amd as ya
    N, cell = 256, 10 * ARC
    d = make_data(50, 1, 1, N, cell)
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {"support": 3, "oversample": 128})
    g.initialise_grid(ya.ImageDesc(N, N, 1, 1, cell, cell))
    g.grid(pchunk(d))
    ocf = O.sphfunc_cf(1.0, 3, 128)
    for k in (0, 1, 127, 128, 5000, 16383):
        assert rel(g.cf_plane(k), ocf.planes[k]) < TOL, k


@pytest.mark.parametrize("npol,nchan,flags,delay,noise_var", [(1, 1, False, False, False), (4, 3, True, True, True)])
def test_sphfunc_grid_matches_oracle(gpu_ctx, npol, nchan, flags, delay, noise_var):
    announce(f"sphfunc grid npol={npol} nchan={nchan} flags={flags}")
    N, cell = 256, 10 * ARC
    d = make_data(3000, nchan, npol, N, cell, flags=flags, delay=delay, noise_var=noise_var)
    g, o, gg, og, gi, oi, gw, ow = grid_both(gpu_ctx, "GPUSphFunc", {"oversample": 128},
                                            dict(kind="SphFunc"), d, N, cell, npol)
    assert rel(gg, og) < TOL
    # image checked inside grid_both
    assert rel(gw, ow) < 1e-6
    s = g.stats()
    assert s.samples_gridded == o.counters.samples_gridded
    assert s.taps_gridded == o.counters.taps_gridded
    assert s.vectors_flagged == o.counters.vectors_flagged


@pytest.mark.parametrize("npol", [1, 4])
def test_sphfunc_degrid_matches_oracle(gpu_ctx, npol):
    announce(f"sphfunc degrid npol={npol}")
    N, cell = 256, 10 * ARC
    d = make_data(3000, 2, npol, N, cell, flags=True, delay=True)
    gv, ov, g, o = degrid_both(gpu_ctx, "GPUSphFunc", {"oversample": 128}, dict(kind="SphFunc"), d, N, cell, npol,
                               random_model(N, npol))
    assert rel(gv, ov) < TOL
    assert g.stats().taps_degridded == o.counters.taps_degridded


def test_box_grid_and_degrid(gpu_ctx):
    announce("box grid/degrid")
    N, cell = 128, 20 * ARC
    d = make_data(2000, 2, 2, N, cell)
    _, _, gg, og, gi, oi, gw, ow = grid_both(gpu_ctx, "GPUBox", {}, dict(kind="Box"), d, N, cell, 2)
    assert rel(gg, og) < TOL and rel(gw, ow) < 1e-6
    gv, ov, _, _ = degrid_both(gpu_ctx, "GPUBox", {}, dict(kind="Box"), d, N, cell, 2, random_model(N, 2))
    assert rel(gv, ov) < TOL


def test_psf_and_pcf_modes(gpu_ctx):
    announce("sphfunc psf/pcf")
    N, cell = 128, 20 * ARC
    d = make_data(1500, 1, 1, N, cell, noise_var=True)
    _, _, gg, og, gi, oi, gw, ow = grid_both(gpu_ctx, "GPUSphFunc", {"oversample": 32, "alldatapsf": True},
                                            dict(kind="SphFunc", sph_oversample=32), d, N, cell, 1, dopsf=True)
    assert rel(gg, og) < TOL and rel(gw, ow) < 1e-6
    _, _, gg, og, gi, oi, _, _ = grid_both(gpu_ctx, "GPUSphFunc", {"oversample": 32},
                                          dict(kind="SphFunc", sph_oversample=32), d, N, cell, 1, dopcf=True)
    assert rel(gg, og) < TOL


def test_mfs_weights_and_frequency_axis(gpu_ctx):
    announce("mfs weights + spectral image")
    N, cell = 128, 20 * ARC
    d = make_data(1500, 6, 1, N, cell, f0=1.3e9, df=2e7)
    fa = (1.3e9, 1.4e9)
    _, _, gg, og, gi, oi, gw, ow = grid_both(gpu_ctx, "GPUSphFunc", {"oversample": 16},
                                            dict(kind="SphFunc", sph_oversample=16), d, N, cell, 1, nchan_img=3,
                                            freq_axis=fa, mfs=(1.35e9, 1))
    assert rel(gg, og) < TOL and rel(gw, ow) < 1e-6


WP = dict(wmax=3000.0, nwplanes=9, oversample=4, maxsupport=256, cutoff=1e-2)


def _wp_oracle_params(**kw):
    p = dict(WP)
    p.update(kw)
    return O.WProjectParams(**{k: v for k, v in p.items()})


@pytest.mark.parametrize("extra", [{}, {"variablesupport": True}, {"variablesupport": True, "offsetsupport": True},
                                   {"wsampling": "powerlaw", "wsampling.exponent": 0.5},
                                   {"wsampling": "gaussian", "wsampling.nwplanes50": 4.0}])
def test_wproject_cf_grid_degrid_match_oracle(gpu_ctx, extra):
    announce(f"wproject {extra}")
    N, cell = 256, 30 * ARC
    d = make_data(1500, 2, 4, N, cell, f0=1.0e9, frac=0.35, wmax_m=2500.0 * O.C_LIGHT / 1.005e9)
    params = dict(WP)
    params.update(extra)
    okw = dict(WP)
    # parset "wsampling.<param>" keys are the oracle's <param> fields
    okw.update({k.split(".", 1)[1] if k.startswith("wsampling.") else k: v for k, v in extra.items()})
    prm = O.WProjectParams(**okw)
    g, o, gg, og, gi, oi, gw, ow = grid_both(gpu_ctx, "GPUWProject", params, dict(kind="WProject", wproject=prm), d,
                                            N, cell, 4)
    for k in (0, 7, 16 * 4 + 5, 16 * 8 + 15):
        ref = o.cf.planes[k]
        got = g.cf_plane(k)
        assert got.shape == ref.shape, k
        assert rel(got, ref) < TOL, k
    assert rel(gg, og) < TOL
    # image checked inside grid_both
    assert rel(gw, ow) < 1e-6
    gv, ov, _, _ = degrid_both(gpu_ctx, "GPUWProject", params, dict(kind="WProject", wproject=prm), d, N, cell, 4,
                               random_model(N, 4))
    assert rel(gv, ov) < TOL


@pytest.mark.parametrize("cutoff", [1e-2, 1e-3])
def test_wproject_pcf_matches_oracle(gpu_ctx, cutoff):
    """Preconditioner form (WProjectVisGridder::doPCFGridder): 3x3 kernels carrying the fitted
    W-kernel width, gridded with conj(C) in half the plane (TableVisGridder.cc:837-844)."""
    announce(f"wproject pcf, cutoff {cutoff}")
    N, cell = 256, 30 * ARC
    d = make_data(1500, 2, 2, N, cell, f0=1.0e9, frac=0.35, wmax_m=2500.0 * O.C_LIGHT / 1.005e9, noise_var=True)
    params = dict(WP, cutoff=cutoff)
    prm = O.WProjectParams(**params)
    g, o, gg, og, gi, oi, gw, ow = grid_both(gpu_ctx, "GPUWProject", params, dict(kind="WProject", wproject=prm), d,
                                            N, cell, 2, dopcf=True)
    assert g.stats().support == 1
    for k in (0, 5, 16 * 4 + 3, 16 * 8 + 15):
        assert np.array_equal(g.cf_plane(k), o.cf.planes[k].astype(np.complex64)), k
    # the centre plane (w = 0) carries the nominal 3 pixels, the outer ones more
    assert g.cf_plane(16 * 4)[1, 1].imag == 3.0 and g.cf_plane(0)[1, 1].imag > 3.0
    assert rel(gg, og) < TOL
    assert rel(gw, ow) < 1e-6


def test_awproject_pcf_matches_oracle(gpu_ctx):
    """Preconditioner form of AWProject (AWProjectVisGridder.cc:436-466,511-516): the support found
    on the first (beam-centred) plane, carried by 3x3 kernels."""
    announce("awproject pcf, 2 feeds, 2 chans")
    import yandasoft_amd as ya
    N, cell = 128, 60 * ARC
    centre = (0.5, -0.7)
    d = make_data(600, 2, 2, N, cell, f0=1.4e9, df=1e7, frac=0.3, wmax_m=400.0 * O.C_LIGHT / 1.41e9, seed=9)
    feed = (np.arange(600) % 2).astype(np.int32)
    point = np.zeros((600, 2))
    point[:, 0] = centre[0] + np.where(feed == 1, math.radians(0.5), 0.0)
    point[:, 1] = centre[1]
    params = {"wmax": 500.0, "nwplanes": 3, "oversample": 4, "maxsupport": 128, "maxfeeds": 2, "cutoff": 0.01,
              "diameter": "12m", "blockage": "1m"}
    prm = O.AWProjectParams(wmax=500.0, nwplanes=3, oversample=4, maxsupport=128, maxfeeds=2, cutoff=0.01,
                            diameter=12.0, blockage=1.0)
    g = ya.Gridder(gpu_ctx, "GPUAWProject", params)
    g.initialise_grid(ya.ImageDesc(N, N, 2, 1, cell, cell, centre=centre), dopcf=True)
    g.grid(ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], noise=d["noise"], feed1=feed,
                    pointing1=np.ascontiguousarray(point)))
    o = O.OracleGridder("AWProject", "f32", wproject=prm)
    o.aw_lm = {0: O.direction_lm(tuple(point[0]), centre), 1: O.direction_lm(tuple(point[1]), centre)}
    o.aw_nchan = 2
    o.initialise_grid((N, N, 2, 1), (cell, cell), dopcf=True)
    och = ochunk(d)
    och.feed1 = feed
    o.grid(och)
    assert o.support > 1
    for k in (0, 7, 16 * 5 + 9, 16 * 11 + 15):
        assert np.array_equal(g.cf_plane(k), o.cf.planes[k].astype(np.complex64)), k
    assert g.cf_plane(0)[1, 1] == complex(1.0, o.support)
    assert rel(g.grid_array().reshape(1, 2, N, N), o.grid_) < TOL


@pytest.mark.parametrize("taper", [False, True])
def test_awproject_disk_matches_oracle(gpu_ctx, taper):
    announce(f"awproject disk 12m, 2 feeds, 2 chans, spheroidaltaper={taper}")
    import yandasoft_amd as ya
    N, cell = 128, 60 * ARC
    centre = (0.5, -0.7)
    d = make_data(600, 2, 2, N, cell, f0=1.4e9, df=1e7, frac=0.3, wmax_m=400.0 * O.C_LIGHT / 1.41e9)
    feed = (np.arange(600) % 2).astype(np.int32)
    off = math.radians(0.5)
    point = np.zeros((600, 2))
    point[:, 0] = centre[0] + np.where(feed == 1, off, 0.0)
    point[:, 1] = centre[1]
    params = {"wmax": 500.0, "nwplanes": 3, "oversample": 4, "maxsupport": 128, "maxfeeds": 2, "cutoff": 0.01,
              "diameter": "12m", "blockage": "1m", "spheroidaltaper": taper}
    prm = O.AWProjectParams(wmax=500.0, nwplanes=3, oversample=4, maxsupport=128, maxfeeds=2, cutoff=0.01,
                            diameter=12.0, blockage=1.0, spheroidaltaper=taper)
    g = ya.Gridder(gpu_ctx, "GPUAWProject", params)
    img = ya.ImageDesc(N, N, 2, 1, cell, cell, centre=centre)
    g.initialise_grid(img)
    ch = ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], noise=d["noise"], feed1=feed,
                  pointing1=np.ascontiguousarray(point))
    g.grid(ch)
    o = O.OracleGridder("AWProject", "f32", wproject=prm)
    o.aw_lm = {0: O.direction_lm(tuple(point[0]), centre), 1: O.direction_lm(tuple(point[1]), centre)}
    o.aw_nchan = 2
    o.initialise_grid((N, N, 2, 1), (cell, cell))
    och = ochunk(d)
    och.feed1 = feed
    o.grid(och)
    assert g.stats().support == o.support
    os_ = 4
    for feedn in (0, 1):
        for chan in (0, 1):
            for iw in (0, 2):
                z = iw + 3 * (chan + 2 * feedn)
                for frac in (0, 9):
                    k = frac + os_ * os_ * z
                    assert rel(g.cf_plane(k), o.cf.planes[k]) < TOL, (feedn, chan, iw, frac)
    gg = g.grid_array().reshape(1, 2, N, N)
    assert rel(gg, o.grid_) < TOL
    o64 = O.OracleGridder("AWProject", "f64", wproject=prm)
    o64.aw_lm, o64.aw_nchan = o.aw_lm, 2
    o64.initialise_grid((N, N, 2, 1), (cell, cell))
    o64.grid(och)
    ok, info = image_ok(g.finalise_grid(), o64.finalise_grid(), gain_of(o), lambda: finalise_with(o, gg))
    assert ok, info
    gw, ow = g.finalise_weights(), o.finalise_weights()
    if taper:
        # the weights divide |FT(CF)|^2 by the taper squared (AWProjectVisGridder.cc:724-742), ~1e-6
        # at the corners, which blows the FP32 rounding of the CF table up there: compare the
        # central half, where the taper is >= ~0.3
        q = slice(N // 4, 3 * N // 4)
        gw, ow = gw[..., q, q], ow[..., q, q]
    assert rel(gw, ow) < TOL
    # degrid
    model = random_model(N, 2)
    dg = ya.Gridder(gpu_ctx, "GPUAWProject", params)
    dg.initialise_degrid(img, model.astype(np.float32))
    out = np.zeros_like(d["vis"])
    dg.degrid(ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], vis_out=out, noise=d["noise"], feed1=feed,
                       pointing1=np.ascontiguousarray(point)))
    od = O.OracleGridder("AWProject", "f32", wproject=prm)
    od.aw_lm = o.aw_lm
    od.aw_nchan = 2
    od.initialise_degrid((N, N, 2, 1), (cell, cell), model)
    ov = od.degrid(och)
    assert rel(np.transpose(out, (2, 1, 0)), ov) < TOL


@pytest.mark.parametrize("sampling", ["linear", "powerlaw"])
def test_awproject_variable_and_offset_support_match_oracle(gpu_ctx, sampling):
    """variablesupport + offsetsupport (AWProjectVisGridder.cc:444-503) with linear or power-law w
    planes: per-plane supports and offsets (a 1.5 deg beam offset moves the w != 0 kernels' peaks by
    more than a grid pixel), CF planes, grid, image, weights (offsets applied,
    AWProjectVisGridder.cc:706-712) and degridding against the oracle; the preconditioner form
    carries every plane's own support at its centre."""
    announce(f"awproject variable + offset support, {sampling} w sampling")
    import yandasoft_amd as ya
    N, cell = 128, 120 * ARC
    centre = (0.5, -0.7)
    d = make_data(600, 2, 2, N, cell, f0=1.4e9, df=1e7, frac=0.3, wmax_m=2700.0 * O.C_LIGHT / 1.41e9, seed=11)
    feed = (np.arange(600) % 2).astype(np.int32)
    point = np.zeros((600, 2))
    point[:, 0] = centre[0] + np.where(feed == 1, math.radians(1.5), 0.0)
    point[:, 1] = centre[1]
    params = {"wmax": 3000.0, "nwplanes": 5, "oversample": 4, "maxsupport": 128, "maxfeeds": 2, "cutoff": 0.01,
              "diameter": "12m", "blockage": "1m", "variablesupport": True, "offsetsupport": True,
              "wsampling": sampling, "wsampling.exponent": 2.0}
    prm = O.AWProjectParams(wmax=3000.0, nwplanes=5, oversample=4, maxsupport=128, maxfeeds=2, cutoff=0.01,
                            diameter=12.0, blockage=1.0, variablesupport=True, offsetsupport=True,
                            wsampling=sampling, exponent=2.0)
    lm = {0: O.direction_lm(tuple(point[0]), centre), 1: O.direction_lm(tuple(point[1]), centre)}
    img = ya.ImageDesc(N, N, 2, 1, cell, cell, centre=centre)
    ch = ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], noise=d["noise"], feed1=feed,
                  pointing1=np.ascontiguousarray(point))
    och = ochunk(d)
    och.feed1 = feed

    def oracle(prec, **kw):
        o = O.OracleGridder("AWProject", prec, wproject=prm)
        o.aw_lm, o.aw_nchan = lm, 2
        o.initialise_grid((N, N, 2, 1), (cell, cell), **kw)
        o.grid(och)
        return o

    g = ya.Gridder(gpu_ctx, "GPUAWProject", params)
    g.initialise_grid(img)
    g.grid(ch)
    o = oracle("f32")
    offs = o.cf.offsets
    assert any(v != (0, 0) for v in offs), offs
    os_ = 4
    sizes = set()
    for z in range(5 * 2 * 2):
        for frac in (0, 6, 15):
            k = frac + os_ * os_ * z
            gp, op = g.cf_plane(k), o.cf.planes[k]
            assert gp.shape == op.shape, (z, gp.shape, op.shape)
            assert rel(gp, op) < TOL, (z, frac)
        sizes.add(o.cf.planes[os_ * os_ * z].shape[0])
    assert len(sizes) > 1
    assert g.stats().support == max(sizes) // 2
    gg = g.grid_array().reshape(1, 2, N, N)
    assert rel(gg, o.grid_) < TOL
    o64 = oracle("f64")
    ok, info = image_ok(g.finalise_grid(), o64.finalise_grid(), gain_of(o), lambda: finalise_with(o, gg))
    assert ok, info
    assert rel(g.finalise_weights(), o.finalise_weights()) < TOL
    # degrid
    model = random_model(N, 2)
    dg = ya.Gridder(gpu_ctx, "GPUAWProject", params)
    dg.initialise_degrid(img, model.astype(np.float32))
    out = np.zeros_like(d["vis"])
    dg.degrid(ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], vis_out=out, noise=d["noise"], feed1=feed,
                       pointing1=np.ascontiguousarray(point)))
    od = O.OracleGridder("AWProject", "f32", wproject=prm)
    od.aw_lm, od.aw_nchan = lm, 2
    od.initialise_degrid((N, N, 2, 1), (cell, cell), model)
    assert rel(np.transpose(out, (2, 1, 0)), od.degrid(och)) < TOL
    # preconditioner form: 3x3 planes whose centre holds each plane's own support, offsets applied
    gp = ya.Gridder(gpu_ctx, "GPUAWProject", params)
    gp.initialise_grid(img, dopcf=True)
    gp.grid(ch)
    op = oracle("f32", dopcf=True)
    centres = set()
    for z in range(5 * 2 * 2):
        k = os_ * os_ * z + 5
        assert np.array_equal(gp.cf_plane(k), op.cf.planes[k].astype(np.complex64)), z
        centres.add(gp.cf_plane(k)[1, 1].imag)
    assert len(centres) > 1, centres
    assert rel(gp.grid_array().reshape(1, 2, N, N), op.grid_) < TOL


def test_wproject_wmax_handling(gpu_ctx):
    announce("wproject wmaxclip / w out of range")
    import yandasoft_amd as ya
    N, cell = 128, 30 * ARC
    d = make_data(500, 1, 1, N, cell, f0=1.0e9, frac=0.3, wmax_m=5000.0 * O.C_LIGHT / 1.0e9)
    p = dict(WP, wmax=2000.0)
    g = ya.Gridder(gpu_ctx, "GPUWProject", p)
    g.initialise_grid(ya.ImageDesc(N, N, 1, 1, cell, cell))
    with pytest.raises(ya.YgridError):
        g.grid(pchunk(d))
    p["wmaxclip"] = True
    _, o, gg, og, gi, oi, _, _ = grid_both(gpu_ctx, "GPUWProject", p,
                                          dict(kind="WProject", wproject=_wp_oracle_params(wmax=2000.0, wmaxclip=True)),
                                          d, N, cell, 1)
    assert rel(gg, og) < TOL
    assert o.counters.vectors_wflagged > 0


def test_edge_cases(gpu_ctx):
    announce("edge cases: empty chunk, all flagged, zero noise, off-grid, empty model")
    import yandasoft_amd as ya
    N, cell = 128, 20 * ARC
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 8})
    g.initialise_grid(ya.ImageDesc(N, N, 1, 1, cell, cell))
    d = make_data(0, 1, 1, N, cell)
    g.grid(pchunk(d))
    d = make_data(100, 1, 1, N, cell)
    d["flag"] = np.ones((1, 1, 100), np.uint8)
    g.grid(pchunk(d))
    assert np.all(g.grid_array() == 0)
    assert g.stats().vectors_flagged == 100
    d = make_data(100, 1, 1, N, cell)
    d["noise"][0, 0, 5] = 0
    with pytest.raises(ya.YgridError):
        g.grid(pchunk(d))
    # off-grid samples are dropped by grid and leave visibilities untouched in degrid
    d = make_data(50, 1, 1, N, cell, frac=2.0)
    g2 = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 8})
    g2.initialise_degrid(ya.ImageDesc(N, N, 1, 1, cell, cell), random_model(N, 1).astype(np.float32))
    out = np.full_like(d["vis"], 7 + 7j)
    g2.degrid(pchunk(d, vis_out=out))
    o = O.OracleGridder("SphFunc", "f32", sph_oversample=8)
    o.initialise_degrid((N, N, 1, 1), (cell, cell), random_model(N, 1))
    ochk = ochunk(d)
    ov = o.degrid(ochk)
    lam = O.C_LIGHT / d["freq"][0]
    r = np.hypot(d["uvw"][:, 0], d["uvw"][:, 1]) / lam * cell * N
    ongrid = (np.abs(d["uvw"][:, 0]) / lam * cell * N < N / 2 - 5) & (np.abs(d["uvw"][:, 1]) / lam * cell * N < N / 2 - 5)
    offgrid = (np.abs(d["uvw"][:, 0]) / lam * cell * N > N / 2 + 5) | (np.abs(d["uvw"][:, 1]) / lam * cell * N > N / 2 + 5)
    assert np.all(out[0, 0, offgrid] == 7 + 7j)
    got = np.transpose(out, (2, 1, 0))[ongrid] - (7 + 7j)
    assert rel(got, ov[ongrid]) < TOL
    # an empty model makes degridding a no-op
    g3 = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 8})
    g3.initialise_degrid(ya.ImageDesc(N, N, 1, 1, cell, cell), np.zeros((1, 1, N, N), np.float32))
    assert g3.is_model_empty()
    out = np.ones_like(d["vis"])
    g3.degrid(pchunk(d, vis_out=out))
    assert np.all(out == 1)


def test_clone_is_independent(gpu_ctx):
    announce("clone")
    import yandasoft_amd as ya
    N, cell = 128, 20 * ARC
    d = make_data(800, 1, 1, N, cell)
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 16})
    g.initialise_grid(ya.ImageDesc(N, N, 1, 1, cell, cell))
    g.grid(pchunk(d))
    c = g.clone()
    c.initialise_grid(ya.ImageDesc(N, N, 1, 1, cell, cell))
    c.grid(pchunk(d))
    c.grid(pchunk(d))
    assert rel(c.grid_array(), 2 * g.grid_array()) < TOL


def test_unknown_gridder_is_rejected(gpu_ctx):
    import yandasoft_amd as ya
    with pytest.raises(ya.YgridError) as e:
        ya.Gridder(gpu_ctx, "NoSuchGridder", {})
    assert e.value.code == 5


def test_binning_cache_reuse_and_invalidation(gpu_ctx):
    announce("binning cache: degrid after grid of one chunk reuses it; changed samples bin again")
    import yandasoft_amd as ya
    N, cell, npol = 128, 20 * ARC, 2
    img = ya.ImageDesc(N, N, npol, 1, cell, cell)
    d = make_data(3000, 3, npol, N, cell, seed=11)
    rng = np.random.default_rng(5)
    model = np.zeros((1, npol, N, N), np.float32)
    model[0, :, rng.integers(20, 108, 30), rng.integers(20, 108, 30)] = rng.normal(size=(30, npol))

    # the tile kernels' binning (the small-support kernels do not bin when degridding)
    gpu_ctx.set_small_support(0)
    try:
        _binning_cache_checks(gpu_ctx, img, d, model, N, cell, npol)
    finally:
        gpu_ctx.set_small_support(8)


def _binning_cache_checks(gpu_ctx, img, d, model, N, cell, npol):
    import yandasoft_amd as ya
    gpu_ctx.clear_cache()
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 8})
    g.initialise_grid(img)
    g.grid(pchunk(d))
    dg = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 8})
    dg.initialise_degrid(img, model)
    out1 = np.zeros_like(d["vis"])
    dg.degrid(pchunk(d, vis_out=out1))
    assert dg.stats().n_bin_reused == 1 and g.stats().n_bin_reused == 0

    gpu_ctx.clear_cache()
    out2 = np.zeros_like(d["vis"])
    dg.degrid(pchunk(d, vis_out=out2))
    assert dg.stats().n_bin_reused == 1
    assert rel(out1, out2) < 1e-6

    # different samples (and, separately, only different flags) must not reuse the binning
    for d2 in (make_data(3000, 3, npol, N, cell, seed=12), dict(d, flag=make_data(3000, 3, npol, N, cell, flags=True)["flag"])):
        g.initialise_grid(img)
        g.grid(pchunk(d2))
        assert g.stats().n_bin_reused == 0
        gpu_ctx.clear_cache()
        ref = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 8})
        ref.initialise_grid(img)
        ref.grid(pchunk(d2))
        assert rel(g.grid_array(), ref.grid_array()) < 1e-6


@pytest.mark.parametrize("small", [0, 8])
def test_dense_tiles_span_staging_rounds_and_items(gpu_ctx, small):
    """Every sample within ~3 pixels of one tile corner: ~4000 pairs per (tile, key), so that a
    work item of the tile kernels is staged in several rounds of 512 pairs and a key is split
    over several items (taken from several XCD queues); the small-support kernels (small = 8)
    take the same corner as four heavily loaded tiles of one item each."""
    announce(f"dense tile corner, SphFunc os 2, 4 pols, grid + degrid vs oracle, small-support limit {small}")
    gpu_ctx.set_small_support(small)
    try:
        _dense_corner(gpu_ctx)
    finally:
        gpu_ctx.set_small_support(8)


def _dense_corner(gpu_ctx):
    import yandasoft_amd as ya
    N, cell, npol = 256, 30 * ARC, 4
    d = make_data(2500, 2, npol, N, cell, frac=0.02, seed=32)
    params = {"oversample": 2}
    okw = dict(kind="SphFunc", sph_oversample=2)
    img = ya.ImageDesc(N, N, npol, 1, cell, cell)
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", params)
    g.initialise_grid(img)
    g.grid(pchunk(d))
    o = O.OracleGridder(precision="f32", **okw)
    o.initialise_grid((N, N, npol, 1), (cell, cell))
    o.grid(ochunk(d))
    assert g.stats().samples_gridded == o.counters.samples_gridded
    assert rel(g.grid_array().reshape(1, npol, N, N), o.grid_) < TOL
    model = random_model(N, npol)
    dg = ya.Gridder(gpu_ctx, "GPUSphFunc", params)
    dg.initialise_degrid(img, model.astype(np.float32))
    out = np.zeros_like(d["vis"])
    dg.degrid(pchunk(d, vis_out=out))
    od = O.OracleGridder(precision="f32", **okw)
    od.initialise_degrid((N, N, npol, 1), (cell, cell), model)
    assert rel(np.transpose(out, (2, 1, 0)), od.degrid(ochunk(d))) < TOL


@pytest.mark.parametrize("npol", [2, 3, 4])
@pytest.mark.parametrize("matrix", [1, 0])
def test_degrid_forms_match_oracle(monkeypatch, matrix, npol):
    """Both degridding forms against the oracle on W-projection data (supports up to ~20, so
    that footprints cover tiles partially and several tiles): the matrix-core form (four
    samples per MFMA group, the default for 2 or more polarisations) and the VALU form
    (one sample per wave), selected through the context's environment."""
    announce(f"degrid form matrix={matrix} npol={npol}")
    import yandasoft_amd as ya
    monkeypatch.setenv("YGRID_DEGRID_VARIANT", f"2,1,{matrix}")
    ctx = ya.Context(0)
    try:
        N, cell = 256, 30 * ARC
        d = make_data(900, 3, npol, N, cell, f0=1.0e9, frac=0.35, wmax_m=2500.0 * O.C_LIGHT / 1.005e9, seed=40 + npol)
        gv, ov, g, _ = degrid_both(ctx, "GPUWProject", dict(WP), dict(kind="WProject", wproject=_wp_oracle_params()),
                                   d, N, cell, npol, random_model(N, npol, seed=npol))
        assert g.stats().degrid_matrix == matrix
        assert rel(gv, ov) < TOL
    finally:
        ctx.close()
