"""The A-projection workload at its full size (160k rows x 32 channels around 1.4 GHz, 9 feeds on a
3 x 3 grid of 1 degree spacing, 12 m uniform disk, 2048^2, oversample 8):

* a subset of rows (every feed, every channel) against the FP32 and FP64 oracle: CF planes, grid,
  image, weights and degridded visibilities, bar 1e-5 of the peak (images as image_ok);
* the same subset with 16-bit CF storage (FP16, BF16) against the FP64 oracle, within the
  storage precision's rounding bound (test_gpu_precision's BOUND), errors printed;
* all rows through size-independent properties: the adjoint identity of grid and degrid
  (sum_s D_s(M) conj(V_s) = sum_p G(V)_p conj(M_p)) and linearity.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.test_gpu_parity import finalise_with, image_ok, rel
from tests.test_gpu_precision import BOUND
from yandasoft_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def awf():
    return W.awproject_feeds()


def _prm(wl):
    p = wl.params
    return O.AWProjectParams(wmax=p["wmax"], nwplanes=p["nwplanes"], oversample=p["oversample"],
                             maxsupport=p["maxsupport"], maxfeeds=p["maxfeeds"], cutoff=p["cutoff"], diameter=12.0,
                             blockage=0.0, freqdep=True)


def _subset(wl, step=400):
    rows = np.arange(0, wl.nrow, step)
    return (rows, np.ascontiguousarray(wl.uvw[rows]), np.ascontiguousarray(wl.vis[:, :, rows]),
            np.ascontiguousarray(wl.noise[:, :, rows]), np.ascontiguousarray(wl.feed1[rows]),
            np.ascontiguousarray(wl.pointing1[rows]))


def _chunk(wl, uvw, vis, noise, feed, point, vis_out=None):
    import yandasoft_amd as ya
    return ya.Chunk(uvw=uvw, freq=wl.freq, vis=vis, vis_out=vis_out, noise=noise, feed1=feed, pointing1=point)


def _oracle(wl, prec, feed_lm):
    o = O.OracleGridder("AWProject", prec, wproject=_prm(wl))
    o.aw_lm = dict(feed_lm)
    o.aw_nchan = wl.nchan
    return o


def _feed_lm(wl):
    lm = {}
    for f in range(9):
        r = int(np.nonzero(wl.feed1 == f)[0][0])
        lm[f] = O.direction_lm(tuple(wl.pointing1[r]), wl.centre)
    return lm


def _model(wl, seed=4):
    rng = np.random.default_rng(seed)
    m = np.zeros((1, 1, wl.ny, wl.nx), np.float32)
    for _ in range(25):
        m[0, 0, rng.integers(wl.ny // 4, 3 * wl.ny // 4), rng.integers(wl.nx // 4, 3 * wl.nx // 4)] = rng.normal()
    return m


def test_awproject_feeds_subset_vs_oracle(gpu_ctx, awf):
    announce("awproject feeds (9 x 32 ch, 2048^2): row subset vs fp32/fp64 oracle")
    import yandasoft_amd as ya
    wl = awf
    rows, uvw, vis, noise, feed, point = _subset(wl)
    img = ya.ImageDesc(wl.nx, wl.ny, 1, 1, wl.cell, wl.cell, centre=wl.centre)
    g = ya.Gridder(gpu_ctx, wl.gridder, wl.params)
    g.initialise_grid(img)
    g.grid(_chunk(wl, uvw, vis, noise, feed, point))
    lm = _feed_lm(wl)
    och = O.Chunk(uvw, wl.freq, np.transpose(vis, (2, 1, 0)), noise=np.transpose(noise, (2, 1, 0)))
    och.feed1 = feed
    o, o64 = _oracle(wl, "f32", lm), _oracle(wl, "f64", lm)
    for x in (o, o64):
        x.initialise_grid((wl.nx, wl.ny, 1, 1), (wl.cell, wl.cell))
        x.grid(och)
    assert g.stats().support == o.support
    nw, os_ = 1, wl.params["oversample"]
    for f in (0, 4, 8):
        for c in (0, 31):
            for frac in (0, 27):
                k = frac + os_ * os_ * (0 + nw * (c + wl.nchan * f))
                assert rel(g.cf_plane(k), o.cf.planes[k]) < TOL, (f, c, frac)
    gg = g.grid_array().reshape(o.grid_.shape)
    assert rel(gg, o.grid_) < TOL
    assert rel(gg, o64.grid_) < TOL
    ok, info = image_ok(g.finalise_grid(), o64.finalise_grid(), np.ones((wl.ny, wl.nx)), lambda: finalise_with(o, gg))
    assert ok, info
    assert rel(g.finalise_weights(), o64.finalise_weights()) < 1e-5
    model = _model(wl)
    d = ya.Gridder(gpu_ctx, wl.gridder, wl.params)
    d.initialise_degrid(img, model)
    out = np.zeros_like(vis)
    d.degrid(_chunk(wl, uvw, vis, noise, feed, point, vis_out=out))
    od = _oracle(wl, "f64", lm)
    od.initialise_degrid((wl.nx, wl.ny, 1, 1), (wl.cell, wl.cell), model.astype(np.float64))
    assert rel(np.transpose(out, (2, 1, 0)), od.degrid(och)) < TOL


@pytest.mark.parametrize("prec", ["fp16", "bf16"])
def test_awproject_feeds_16bit_storage_vs_fp64(gpu_ctx, awf, prec):
    announce(f"awproject feeds: {prec} CF taps and coefficients vs fp64 oracle (row subset)")
    import yandasoft_amd as ya
    wl = awf
    rows, uvw, vis, noise, feed, point = _subset(wl)
    params = dict(wl.params, cfprecision=prec, visprecision=prec)
    img = ya.ImageDesc(wl.nx, wl.ny, 1, 1, wl.cell, wl.cell, centre=wl.centre)
    g = ya.Gridder(gpu_ctx, wl.gridder, params)
    g.initialise_grid(img)
    g.grid(_chunk(wl, uvw, vis, noise, feed, point))
    lm = _feed_lm(wl)
    och = O.Chunk(uvw, wl.freq, np.transpose(vis, (2, 1, 0)), noise=np.transpose(noise, (2, 1, 0)))
    och.feed1 = feed
    o64 = _oracle(wl, "f64", lm)
    o64.initialise_grid((wl.nx, wl.ny, 1, 1), (wl.cell, wl.cell))
    o64.grid(och)
    eg = rel(g.grid_array().reshape(o64.grid_.shape), o64.grid_)
    eh = rel(g.finalise_grid(), o64.finalise_grid())
    model = _model(wl)
    d = ya.Gridder(gpu_ctx, wl.gridder, params)
    d.initialise_degrid(img, model)
    out = np.zeros_like(vis)
    d.degrid(_chunk(wl, uvw, vis, noise, feed, point, vis_out=out))
    od = _oracle(wl, "f64", lm)
    od.initialise_degrid((wl.nx, wl.ny, 1, 1), (wl.cell, wl.cell), model.astype(np.float64))
    ev = rel(np.transpose(out, (2, 1, 0)), od.degrid(och))
    print(f"[precision] awproject feeds {prec}: grid err {eg:.3e}, image err {eh:.3e}, degrid err {ev:.3e}",
          flush=True)
    assert eg <= BOUND[prec] and eh <= BOUND[prec] and ev <= BOUND[prec]


def test_awproject_feeds_all_rows_adjoint_and_linearity(gpu_ctx, awf):
    announce("awproject feeds all 160k rows x 32 ch: adjoint identity + linearity")
    import yandasoft_amd as ya
    wl = awf
    img = ya.ImageDesc(wl.nx, wl.ny, 1, 1, wl.cell, wl.cell, centre=wl.centre)
    g = ya.Gridder(gpu_ctx, wl.gridder, wl.params)
    g.initialise_grid(img)
    g.grid(_chunk(wl, wl.uvw, wl.vis, wl.noise, wl.feed1, wl.pointing1))
    G = g.grid_array().astype(np.complex128)
    s = g.stats()
    assert s.samples_gridded == wl.nrow * wl.nchan  # every sample lies on the grid
    model = _model(wl, seed=9)
    d = ya.Gridder(gpu_ctx, wl.gridder, wl.params)
    d.initialise_degrid(img, model)
    M = d.grid_array().astype(np.complex128)
    out = np.zeros_like(wl.vis)
    d.degrid(_chunk(wl, wl.uvw, wl.vis, wl.noise, wl.feed1, wl.pointing1, vis_out=out))
    lhs = np.sum(out.astype(np.complex128) * np.conj(wl.vis.astype(np.complex128)))
    rhs = np.sum(G * np.conj(M))
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs))
    g2 = ya.Gridder(gpu_ctx, wl.gridder, wl.params)
    g2.initialise_grid(img)
    g2.grid(_chunk(wl, wl.uvw, np.ascontiguousarray(2 * wl.vis), wl.noise, wl.feed1, wl.pointing1))
    assert rel(g2.grid_array(), 2 * G) < TOL
