"""A-projection with w stacking (GPUAProjectWStack; AProjectWStackVisGridder.cc) against the
oracle's composition: OracleWStack with an 'AProjectWStack' OracleGridder per w plane.  The CFs
hold the A term only, one per (channel, feed); every sample grids into the grid of its w plane
with its feed's CF; the image sums the w-screened inverse transforms without a correction; the
weights place each used CF over taps -S..S-1, normalise its transform to a peak of 1 and sum
the squared moduli; degridding predicts every sample from the conjugate-screened model of its
plane."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.test_gpu_parity import ARC, TOL, make_data, ochunk, random_model, rel

pytestmark = pytest.mark.gpu

N, CELL, CENTRE = 128, 60 * ARC, (0.5, -0.7)
PARAMS = {"wmax": 2000.0, "nwplanes": 7, "oversample": 4, "maxsupport": 128, "maxfeeds": 2,
          "diameter": "12m", "blockage": "1m"}


def _prm():
    return O.AWProjectParams(wmax=2000.0, nwplanes=7, oversample=4, maxsupport=128, maxfeeds=2, diameter=12.0,
                             blockage=1.0)


def _case(seed):
    d = make_data(700, 2, 2, N, CELL, f0=1.4e9, df=1.2e8, frac=0.3, wmax_m=1800.0 * O.C_LIGHT / 1.52e9, seed=seed)
    feed = (np.arange(700) % 2).astype(np.int32)
    point = np.zeros((700, 2))
    point[:, 0] = CENTRE[0] + np.where(feed == 1, math.radians(0.7), 0.0)
    point[:, 1] = CENTRE[1]
    lm = {0: O.direction_lm(tuple(point[0]), CENTRE), 1: O.direction_lm(tuple(point[1]), CENTRE)}
    point = np.ascontiguousarray(point)
    och = ochunk(d)
    och.feed1, och.pointing1 = feed, point
    return d, feed, point, lm, och


def _oracle(prec, lm):
    prm = _prm()

    def make():
        # alldatapsf false (the default): the PSF grids the first row's feed and pointing only
        o = O.OracleGridder("AProjectWStack", prec, wproject=prm, use_all_psf=False)
        o.aw_lm, o.aw_nchan = lm, 2
        return o
    return O.OracleWStack(prm.wmax, prm.nwplanes, precision=prec, make=make)


@pytest.mark.parametrize("dopsf", [False, True])
def test_aprojectwstack_grid_image_weights_match_oracle(gpu_ctx, dopsf):
    announce(f"aprojectwstack grid dopsf={dopsf}: 12m disk, 2 feeds (0.7 deg offset), 2 chans, 7 w planes")
    import yandasoft_amd as ya
    d, feed, point, lm, och = _case(31)
    g = ya.Gridder(gpu_ctx, "GPUAProjectWStack", dict(PARAMS))
    img = ya.ImageDesc(N, N, 2, 1, CELL, CELL, centre=CENTRE)
    g.initialise_grid(img, dopsf=dopsf)
    g.grid(ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], noise=d["noise"], feed1=feed, pointing1=point))
    o = _oracle("f32", lm)
    grids = o.grid((N, N, 2, 1), (CELL, CELL), och, dopsf=dopsf)
    # the CF planes: (chan, feed) blocks of os^2 oversampling planes, support 1 + 2 int(R) / os
    # (R = 4.17 pattern cells of 1.44 m at 1.4 GHz)
    assert g.stats().support == 3
    cfo = _oracle("f32", lm)
    probe = cfo.make()
    probe.initialise_grid((N, N, 2, 1), (CELL, CELL), dopsf=dopsf)
    probe.grid(och)
    for z in range(4):
        for frac in (0, 5, 15):
            k = frac + 16 * z
            assert rel(g.cf_plane(k), probe.cf.planes[k]) < TOL, (z, frac)
    gg = g.grid_array().reshape(7, 1, 2, N, N)
    used = [iw for iw in range(7) if np.abs(grids[iw]).max() > 0]
    assert len(used) >= 4
    peak = max(np.abs(x).max() for x in grids)
    for iw in range(7):
        assert np.abs(gg[iw] - grids[iw]).max() <= TOL * peak, iw
    gi = g.finalise_grid()
    # the finalisation alone: the oracle's sum of screened transforms of the GPU's own grids
    o.grids = [np.asarray(gg[iw], np.complex128) for iw in range(7)]
    assert rel(gi, o.finalise_grid()) < 1e-5
    # and the whole chain against the oracle in double precision
    o64 = _oracle("f64", lm)
    o64.grid((N, N, 2, 1), (CELL, CELL), och, dopsf=dopsf)
    assert rel(gi, o64.finalise_grid()) < TOL
    assert rel(g.finalise_weights(), o.sumw) < TOL


def test_aprojectwstack_degrid_matches_oracle(gpu_ctx):
    announce("aprojectwstack degrid: 12m disk, 2 feeds, 2 chans, 7 w planes")
    import yandasoft_amd as ya
    d, feed, point, lm, och = _case(32)
    model = random_model(N, 2, seed=4)
    g = ya.Gridder(gpu_ctx, "GPUAProjectWStack", dict(PARAMS))
    g.initialise_degrid(ya.ImageDesc(N, N, 2, 1, CELL, CELL, centre=CENTRE), model.astype(np.float32))
    out = np.zeros_like(d["vis"])
    g.degrid(ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], vis_out=out, noise=d["noise"], feed1=feed,
                      pointing1=point))
    ov = _oracle("f32", lm).degrid((N, N, 2, 1), (CELL, CELL), model, och)
    assert rel(np.transpose(out, (2, 1, 0)), ov) < TOL


def test_aprojectwstack_preconditioner_pass_matches_oracle(gpu_ctx):
    """The preconditioner pass: AProjectWStackVisGridder has no form of its own, so its A-term
    CFs (unshifted, as for the PSF) go through the generic preconditioner branch (1 for every
    sample, conjugate CF in half the uv plane) into the grid of each sample's w plane, and
    WStackVisGridder::finaliseGrid's preconditioner form (:222-262) stacks the used planes with
    their imaginary parts scaled (updatew) into one normalised inverse transform."""
    announce("aprojectwstack preconditioner pass")
    import yandasoft_amd as ya
    d, feed, point, lm, och = _case(33)
    g = ya.Gridder(gpu_ctx, "GPUAProjectWStack", dict(PARAMS))
    g.initialise_grid(ya.ImageDesc(N, N, 2, 1, CELL, CELL, centre=CENTRE), dopcf=True)
    g.grid(ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], noise=d["noise"], feed1=feed, pointing1=point))
    o = _oracle("f32", lm)
    grids = o.grid((N, N, 2, 1), (CELL, CELL), och, dopcf=True)
    gg = g.grid_array().reshape(7, 1, 2, N, N)
    peak = max(np.abs(x).max() for x in grids)
    for iw in range(7):
        assert np.abs(gg[iw] - grids[iw]).max() <= TOL * peak, iw
    gi = g.finalise_grid()
    o.grids = [np.asarray(gg[iw], np.complex128) for iw in range(7)]
    assert rel(gi, o.finalise_grid()) < 1e-5
    o64 = _oracle("f64", lm)
    o64.grid((N, N, 2, 1), (CELL, CELL), och, dopcf=True)
    assert rel(gi, o64.finalise_grid()) < TOL


def test_aprojectwstack_rejects_the_taper(gpu_ctx):
    import yandasoft_amd as ya
    with pytest.raises(Exception, match="[Tt]aper"):
        ya.Gridder(gpu_ctx, "GPUAProjectWStack", dict(PARAMS, spheroidaltaper=True))
