"""GPUAWProject with the ATCA illumination (ATCAIllumination.cc:183-331) and the cache policy of
AProjectGridderBase::validateCFCache (AProjectGridderBase.cc:202-253) against the CPU oracle:
CF planes, grids and degridded visibilities to 1e-5 of the peak (FP32), and the CF-generation
counters of updateStats (AProjectGridderBase.cc:272-282) equal to the oracle's, over chunks whose
parallactic angles and frequencies move."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.test_gpu_parity import ARC, make_data, ochunk, random_model, rel

pytestmark = pytest.mark.gpu
TOL = 1e-5
N, CELL = 128, 60 * ARC
CENTRE = (0.5, -0.7)

BASE = {"wmax": 500.0, "nwplanes": 3, "oversample": 4, "maxsupport": 128, "maxfeeds": 2, "cutoff": 0.01,
        "diameter": "22m", "blockage": "2m", "illumination": "ATCA", "patolerance": "0.1rad"}


def _prm(**kw):
    p = dict(wmax=500.0, nwplanes=3, oversample=4, maxsupport=128, maxfeeds=2, cutoff=0.01, diameter=22.0,
             blockage=2.0, illumination="ATCA", patolerance=0.1)
    p.update(kw)
    return O.AWProjectParams(**p)


def _chunks(pas, seed=5, nrow=400, f0=1.4e9):
    """Chunks of nrow rows (2 feeds, 2 channels, 2 pols), row r of chunk k at angle pas[k][r % len]."""
    out = []
    for k, pa in enumerate(pas):
        d = make_data(nrow, 2, 2, N, CELL, f0=f0, df=1e7, frac=0.3, wmax_m=400.0 * O.C_LIGHT / 1.41e9,
                      seed=seed + k)
        feed = (np.arange(nrow) % 2).astype(np.int32)
        point = np.zeros((nrow, 2))
        point[:, 0] = CENTRE[0] + np.where(feed == 1, math.radians(0.5), 0.0)
        point[:, 1] = CENTRE[1]
        d["feed1"] = feed
        d["point"] = np.ascontiguousarray(point)
        d["pa"] = np.ascontiguousarray(np.resize(np.asarray(pa, np.float32), nrow))
        out.append(d)
    return out


def _gchunk(d, vis_out=None):
    import yandasoft_amd as ya
    return ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], vis_out=vis_out, noise=d["noise"], feed1=d["feed1"],
                    pointing1=d["point"], feed1_pa=d["pa"])


def _ochunk(d):
    c = ochunk(d)
    c.feed1 = d["feed1"]
    c.feed1_pa = d["pa"]
    return c


def _oracle(prm, precision="f32"):
    o = O.OracleGridder("AWProject", precision, wproject=prm)
    lm = O.direction_lm((CENTRE[0], CENTRE[1]), CENTRE)
    lm1 = O.direction_lm((CENTRE[0] + math.radians(0.5), CENTRE[1]), CENTRE)
    o.aw_lm = {0: lm, 1: lm1}
    o.aw_nchan = 2
    return o


def _check_cfs(g, o, os_=4, nw=3):
    for feed in (0, 1):
        for chan in (0, 1):
            for iw in (0, 2):
                z = iw + nw * (chan + 2 * feed)
                for frac in (0, 9):
                    k = frac + os_ * os_ * z
                    assert rel(g.cf_plane(k), o.cf.planes[k]) < TOL, (feed, chan, iw, frac)


def _stats_equal(g, o):
    s = g.stats()
    got = (s.cf_iterations, s.cf_generations, s.cf_generations_pa, s.cf_generations_freq)
    assert got == tuple(float(x) for x in o.aw_stats), (got, o.aw_stats)


@pytest.mark.parametrize("legs", [True, False])
def test_atca_grid_follows_the_parallactic_angle(gpu_ctx, legs):
    """Three chunks: angle 0.2 (drops the empty cache at the first row), 0.25 then 0.6 within the
    second chunk (0.6 is 0.4 off the cached 0.2: the cache is rebuilt with the first rows of each
    feed, at 0.25), 0.6 again (no rebuild).  Without feed legs the pattern is symmetric and no
    angle is checked."""
    announce(f"atca grid, feed legs {legs}, 3 chunks of moving parallactic angle")
    import yandasoft_amd as ya
    params = dict(BASE, **{"illumination.feedlegs": legs, "illumination.tapering.defocusing": "0.5rad"})
    prm = _prm(feedlegs=legs, defocusing=0.5)
    chunks = _chunks([[0.2], [0.25, 0.25, 0.6], [0.6]])
    g = ya.Gridder(gpu_ctx, "GPUAWProject", params)
    g.initialise_grid(ya.ImageDesc(N, N, 2, 1, CELL, CELL, centre=CENTRE))
    o = _oracle(prm)
    o.initialise_grid((N, N, 2, 1), (CELL, CELL))
    for d in chunks:
        g.grid(_gchunk(d))
        o.grid(_ochunk(d))
        _check_cfs(g, o)
        _stats_equal(g, o)
    assert rel(g.grid_array().reshape(1, 2, N, N), o.grid_) < TOL
    s = g.stats()
    # with legs: chunks 1 and 2 rebuild both feeds for the angle (the first chunk's frequency check
    # is skipped); without: only the first chunk's frequencies build the cache
    expect = (4, 4, 0) if legs else (2, 0, 2)
    assert (s.cf_generations, s.cf_generations_pa, s.cf_generations_freq) == expect


def test_atca_degrid_and_frequency_invalidation(gpu_ctx):
    """Degridding with asymmetric ATCA patterns at two angles, then a chunk whose frequencies moved
    by 1e-4 (beyond freqtolerance 1e-6) at the same angle: the cache is rebuilt for the new
    frequencies and counted as a frequency rebuild."""
    announce("atca degrid, angle change then frequency change")
    import yandasoft_amd as ya
    params = dict(BASE, **{"freqtolerance": "1e-6", "illumination.feedlegs.wedges.shadowing": "[0.9, 0.3]"})
    prm = _prm(freqtolerance=1e-6, wedges_shadowing=(0.9, 0.3))
    chunks = _chunks([[0.0], [1.0]], seed=11) + _chunks([[1.0]], seed=13, f0=1.4e9 * (1 + 1e-4))
    model = random_model(N, 2)
    g = ya.Gridder(gpu_ctx, "GPUAWProject", params)
    g.initialise_degrid(ya.ImageDesc(N, N, 2, 1, CELL, CELL, centre=CENTRE), model.astype(np.float32))
    o = _oracle(prm)
    o.initialise_degrid((N, N, 2, 1), (CELL, CELL), model)
    for d in chunks:
        out = np.zeros_like(d["vis"])
        g.degrid(_gchunk(d, vis_out=out))
        ov = o.degrid(_ochunk(d))
        _check_cfs(g, o)
        _stats_equal(g, o)
        assert rel(np.transpose(out, (2, 1, 0)), ov) < TOL
    s = g.stats()
    assert (s.cf_generations, s.cf_generations_pa, s.cf_generations_freq) == (6, 2, 4)


def test_atca_parset_checks(gpu_ctx):
    import yandasoft_amd as ya
    for bad in ({"illumination": "SKA_MID"}, {"illumination.feedlegs.wedges.shadowing": "[0.1, 0.2, 0.3]"},
                {"illumination.feedlegs.shadowing": "1.5"}, {"illumination.feedlegs.wedges.angle": "95deg"}):
        with pytest.raises(ya.YgridError):
            ya.Gridder(gpu_ctx, "GPUAWProject", dict(BASE, **bad))
