"""The imaging equation's data loop on the device (ygrid_imaging_chunk; ImageFFTEquation.cc:437-480)
against the same steps composed from the oracle: model visibilities = sum of the model
degridders' predictions (an empty model adds nothing), residual = data - model, and the
residual and PSF gridders grid that residual.  Host and device chunks; the residual cube the
call returns is checked too, and so is the composition of the single-pass entry points."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.devmem import DevArray
from tests.test_gpu_parity import ARC, TOL, WP, _wp_oracle_params, make_data, ochunk, pchunk, random_model, rel

pytestmark = pytest.mark.gpu


def _gridders(ctx, img, models):
    import yandasoft_amd as ya
    dgs = []
    for m in models:
        g = ya.Gridder(ctx, "GPUWProject", dict(WP))
        g.initialise_degrid(img, m.astype(np.float32))
        dgs.append(g)
    res = ya.Gridder(ctx, "GPUWProject", dict(WP))
    res.initialise_grid(img)
    psf = ya.Gridder(ctx, "GPUWProject", dict(WP))
    psf.initialise_grid(img, dopsf=True)
    return dgs, res, psf


@pytest.mark.parametrize("device", [False, True])
def test_imaging_chunk_matches_oracle(gpu_ctx, device):
    announce(f"imaging chunk: 2 models + 1 empty, residual + PSF gridders, device={device}")
    import yandasoft_amd as ya
    N, cell, npol = 128, 60 * ARC, 2
    d = make_data(700, 3, npol, N, cell, f0=1.0e9, frac=0.35, wmax_m=1500.0 * O.C_LIGHT / 1.005e9, seed=50)
    img = ya.ImageDesc(N, N, npol, 1, cell, cell)
    m1, m2 = random_model(N, npol, seed=1), random_model(N, npol, seed=2)
    dgs, res, psf = _gridders(gpu_ctx, img, (m1, m2, np.zeros_like(m1)))
    assert dgs[2].is_model_empty()
    if device:
        out = DevArray(np.zeros_like(d["vis"]))
        chunk = ya.Chunk(uvw=DevArray(d["uvw"]), freq=DevArray(d["freq"]), vis=DevArray(d["vis"]),
                         noise=DevArray(d["noise"]), vis_out=out, device=True)
    else:
        out = np.zeros_like(d["vis"])
        chunk = pchunk(d, vis_out=out)
    ya.imaging_chunk(dgs, [res, psf], chunk)
    resid = out.numpy() if device else out

    # the oracle: model = sum of the predictions, residual = data - model, then grid it
    okw = dict(kind="WProject", wproject=_wp_oracle_params())
    model = 0
    for m in (m1, m2):
        o = O.OracleGridder(precision="f32", **okw)
        o.initialise_degrid((N, N, npol, 1), (cell, cell), m)
        model = model + o.degrid(ochunk(d))
    oresid = np.transpose(d["vis"], (2, 1, 0)) - model
    assert rel(np.transpose(resid, (2, 1, 0)), oresid) < TOL
    dr = dict(d)
    dr["vis"] = np.ascontiguousarray(np.transpose(oresid, (2, 1, 0))).astype(np.complex64)
    og = O.OracleGridder(precision="f32", **okw)
    og.initialise_grid((N, N, npol, 1), (cell, cell))
    og.grid(ochunk(dr))
    assert rel(res.grid_array().reshape(1, npol, N, N), og.grid_) < TOL
    op = O.OracleGridder(precision="f32", **okw)
    op.initialise_grid((N, N, npol, 1), (cell, cell), dopsf=True)
    op.grid(ochunk(dr))
    assert rel(psf.grid_array().reshape(1, npol, N, N), op.grid_) < TOL

    # the same steps through the single-pass entry points
    dgs2, res2, _ = _gridders(gpu_ctx, img, (m1, m2))
    pred = np.zeros_like(d["vis"])
    for g in dgs2:
        g.degrid(pchunk(d, vis_out=pred))
    d2 = dict(d)
    d2["vis"] = (d["vis"] - pred).astype(np.complex64)
    res2.grid(pchunk(d2))
    assert rel(resid, d2["vis"]) < TOL
    assert rel(res.grid_array(), res2.grid_array()) < TOL


def test_imaging_chunk_rejects_bad_calls(gpu_ctx):
    announce("imaging chunk: argument errors")
    import yandasoft_amd as ya
    N, cell, npol = 64, 60 * ARC, 1
    d = make_data(50, 1, npol, N, cell, seed=51)
    img = ya.ImageDesc(N, N, npol, 1, cell, cell)
    dgs, res, _ = _gridders(gpu_ctx, img, (random_model(N, npol, seed=3),))
    with pytest.raises(ya.YgridError):
        ya.imaging_chunk([], [], pchunk(d))  # no gridders
    nodata = pchunk(d)
    nodata.vis = None
    with pytest.raises(ya.YgridError):
        ya.imaging_chunk(dgs, [res], nodata)  # the data visibilities are required
    other = ya.Context(0)
    try:
        g = ya.Gridder(other, "GPUSphFunc", {})
        g.initialise_grid(img)
        with pytest.raises(ya.YgridError):
            ya.imaging_chunk(dgs, [g], pchunk(d))  # gridders of two contexts
        with pytest.raises(ya.YgridError):
            ya.imaging_chunk([res], [], pchunk(d))  # a gridding gridder in the degridder list
    finally:
        other.close()
