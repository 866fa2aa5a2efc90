"""The factory's global gridder options and the accessor fields they read, against the oracle:
MaxPointingSeparation row rejection (TableVisGridder.cc:571-578), the representative PSF feed and
field (:581-593,809-812), the direction-dependent row offset of degridded visibilities
(:567,861, itsSourceIndex) and ygrid_gridder_configure (VisGridderFactory.cc:137-232)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.devmem import DevArray
from tests.test_gpu_parity import ARC, TOL, make_data, ochunk, pchunk, random_model, rel

pytestmark = pytest.mark.gpu

CENTRE = (1.0, -0.5)


def _fields(nrow, seed):
    rng = np.random.default_rng(seed)
    far = rng.random(nrow) < 0.4
    point = np.tile(np.array(CENTRE), (nrow, 1))
    point[far, 0] += math.radians(2.0) / math.cos(CENTRE[1])
    feed = (rng.random(nrow) < 0.5).astype(np.int32)
    return far, np.ascontiguousarray(point), feed


def test_max_pointing_separation_matches_oracle(gpu_ctx):
    announce("MaxPointingSeparation grid/weights/degrid")
    import yandasoft_amd as ya
    N, cell = 128, 20 * ARC
    d = make_data(1200, 3, 2, N, cell, seed=21)
    far, point, _ = _fields(1200, 22)
    img = ya.ImageDesc(N, N, 2, 1, cell, cell, centre=CENTRE)
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 32})
    g.configure({"MaxPointingSeparation": "1deg"})
    g.initialise_grid(img)
    c = pchunk(d)
    c.pointing1 = point
    g.grid(c)
    o = O.OracleGridder("SphFunc", "f32", sph_oversample=32)
    o.max_sep, o.centre = math.radians(1.0), CENTRE
    o.initialise_grid((N, N, 2, 1), (cell, cell))
    och = ochunk(d)
    och.pointing1 = point
    o.grid(och)
    assert rel(g.grid_array().reshape(o.grid_.shape), o.grid_) < TOL
    assert rel(g.finalise_weights(), o.finalise_weights()) < 1e-6
    st = g.stats()
    assert st.rows_rejected_pointing == np.count_nonzero(far) == o.rows.rows_rejected
    assert st.samples_gridded == o.counters.samples_gridded
    # degridding leaves the rejected rows unchanged
    model = random_model(N, 2)
    dg = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 32, "MaxPointingSeparation": math.radians(1.0)})
    dg.initialise_degrid(img, model.astype(np.float32))
    out = np.zeros_like(d["vis"])
    c2 = pchunk(d, vis_out=out)
    c2.pointing1 = point
    dg.degrid(c2)
    assert not np.any(out[:, :, far])
    od = O.OracleGridder("SphFunc", "f32", sph_oversample=32)
    od.max_sep, od.centre = math.radians(1.0), CENTRE
    od.initialise_degrid((N, N, 2, 1), (cell, cell), model)
    ov = od.degrid(och)
    assert rel(np.transpose(out, (2, 1, 0)), ov) < TOL


@pytest.mark.parametrize("dish", [False, True])
def test_psf_representative_feed_and_field_match_oracle(gpu_ctx, dish):
    announce(f"PSF representative feed/field over two chunks, dish pointing given={dish}")
    import yandasoft_amd as ya
    N, cell = 128, 20 * ARC
    d = make_data(1000, 2, 1, N, cell, seed=31)
    far, point, feed = _fields(1000, 32)
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 32})
    g.initialise_grid(ya.ImageDesc(N, N, 1, 1, cell, cell, centre=CENTRE), dopsf=True)
    o = O.OracleGridder("SphFunc", "f32", sph_oversample=32, use_all_psf=False)
    o.initialise_grid((N, N, 1, 1), (cell, cell), dopsf=True)
    for sl in (slice(0, 500), slice(500, None)):
        part = {k: (None if v is None else np.ascontiguousarray(v[sl] if k in ("uvw", "delay") else
                                                               (v[..., sl] if k in ("vis", "noise", "flag") else v)))
                for k, v in d.items()}
        c = pchunk(part)
        c.feed1 = np.ascontiguousarray(feed[sl])
        if dish:
            c.dish_pointing1 = np.ascontiguousarray(point[sl])
        else:
            c.pointing1 = np.ascontiguousarray(point[sl])
        g.grid(c)
        och = ochunk(part)
        och.feed1 = np.ascontiguousarray(feed[sl])
        och.pointing1 = np.ascontiguousarray(point[sl])
        o.grid(och)
    sel = (feed == feed[0]) & (far == far[0])
    assert 0 < np.count_nonzero(sel) < sel.size
    assert rel(g.grid_array().reshape(o.grid_.shape), o.grid_) < TOL
    assert g.stats().samples_gridded == o.counters.samples_gridded


class _At:
    """A device address where a chunk array starts (data_ptr, as the binding reads it)."""

    def __init__(self, p):
        self.p = p

    def data_ptr(self):
        return self.p


@pytest.mark.parametrize("device", [False, True])
def test_vis_out_rows_offsets_the_prediction(gpu_ctx, device):
    """A direction-dependent calibration chunk degrids source s into rows s*nrow.. of a cube of
    nsource*nrow rows (vis_out at row s*nrow, vis_out_rows = nsource*nrow): the other rows are
    untouched and the source's rows equal an ordinary prediction."""
    announce(f"vis_out_rows: source 1 of 3, device={device}")
    import yandasoft_amd as ya
    N, cell, nrow, nchan, npol = 128, 20 * ARC, 400, 3, 2
    d = make_data(nrow, nchan, npol, N, cell, seed=41)
    model = random_model(N, npol)
    img = ya.ImageDesc(N, N, npol, 1, cell, cell)
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 32})
    g.initialise_degrid(img, model.astype(np.float32))
    ref = np.zeros_like(d["vis"])
    g.degrid(pchunk(d, vis_out=ref))
    big = np.full((npol, nchan, 3 * nrow), 7.0 + 1j, np.complex64)  # casacore Cube(3 nrow, nchan, npol)
    if device:
        dev = DevArray(big)
        c = ya.Chunk(uvw=DevArray(d["uvw"]), freq=DevArray(d["freq"]), vis=DevArray(d["vis"]),
                     noise=DevArray(d["noise"]), vis_out=_At(dev.data_ptr() + 8 * nrow), device=True,
                     vis_out_rows=3 * nrow)
        g.degrid(c)
        got = dev.numpy()
    else:
        flat = big.reshape(-1)
        view = flat[nrow:]  # the cube from row nrow on, same leading dimension
        c = ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], noise=d["noise"], vis_out=view,
                     vis_out_rows=3 * nrow)
        g.degrid(c)
        got = big
    assert np.all(got[:, :, :nrow] == 7.0 + 1j) and np.all(got[:, :, 2 * nrow:] == 7.0 + 1j)
    assert rel(got[:, :, nrow:2 * nrow] - (7.0 + 1j), ref) < 1e-6


def test_configure_options(gpu_ctx):
    announce("ygrid_gridder_configure: padding, MaxPointingSeparation, refusals")
    import yandasoft_amd as ya
    N, cell = 128, 20 * ARC
    d = make_data(600, 2, 1, N, cell, seed=51)
    img = ya.ImageDesc(N, N, 1, 1, cell, cell)
    a = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 32, "padding": 1.5})
    b = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 32})
    b.configure({"padding": 1.5, "alldatapsf": "true", "oversampleweight": "false", "cleargrids": "false"})
    for g in (a, b):
        g.initialise_grid(img)
        g.grid(pchunk(d))
    # the same up to FP32 accumulation order (the order in which a tile's pairs are added is
    # fixed by atomics at run time, so two passes agree to rounding, not bit for bit)
    ia, ib = a.finalise_grid(), b.finalise_grid()
    assert ia.shape == ib.shape and rel(ia, ib) < 1e-6
    with pytest.raises(ya.YgridError):
        b.configure({"parotation": "true"})
    with pytest.raises(ya.YgridError):
        b.configure({"swappols": "true"})
    with pytest.raises(ya.YgridError):
        b.configure({"nosuchoption": 1})
