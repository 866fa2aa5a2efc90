"""W stacking (GPUWStack; WStackVisGridder.cc) against the oracle's composition (OracleWStack):
every sample grids into the grid of its w plane with the SphFunc kernel; the image is the sum
over the planes of the real part of the w-screened inverse transforms, corrected; degridding
predicts every sample from the grid of its plane, filled from the conjugate-screened corrected
model."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.test_gpu_parity import ARC, TOL, gain_of, image_ok, make_data, ochunk, pchunk, random_model, rel

pytestmark = pytest.mark.gpu

WS = dict(wmax=2000.0, nwplanes=9)


def _data(npol, seed):
    N, cell = 128, 60 * ARC
    return N, cell, make_data(900, 3, npol, N, cell, f0=1.0e9, frac=0.35, wmax_m=1800.0 * O.C_LIGHT / 1.01e9,
                              seed=seed, flags=True)


@pytest.mark.parametrize("npol", [1, 2])
def test_wstack_grid_and_image_match_oracle(gpu_ctx, npol):
    announce(f"wstack grid npol={npol}")
    import yandasoft_amd as ya
    N, cell, d = _data(npol, 70 + npol)
    g = ya.Gridder(gpu_ctx, "GPUWStack", dict(WS))
    g.initialise_grid(ya.ImageDesc(N, N, npol, 1, cell, cell))
    g.grid(pchunk(d))
    o = O.OracleWStack(WS["wmax"], WS["nwplanes"])
    grids = o.grid((N, N, npol, 1), (cell, cell), ochunk(d))
    gg = g.grid_array().reshape(WS["nwplanes"], 1, npol, N, N)
    used = [iw for iw in range(WS["nwplanes"]) if np.abs(grids[iw]).max() > 0]
    assert len(used) >= 3  # the data spread over several planes
    peak = max(np.abs(x).max() for x in grids)
    for iw in range(WS["nwplanes"]):
        assert np.abs(gg[iw] - grids[iw]).max() <= TOL * peak, iw
    gi = g.finalise_grid()
    o64 = O.OracleWStack(WS["wmax"], WS["nwplanes"], precision="f64")
    o64.grid((N, N, npol, 1), (cell, cell), ochunk(d))

    def own_grids():
        saved = o64.grids
        o64.grids = [np.asarray(gg[iw], np.complex128) for iw in range(WS["nwplanes"])]
        try:
            return o64.finalise_grid()
        finally:
            o64.grids = saved

    # as for the single-grid images (test_gpu_parity's image_ok)
    ok, info = image_ok(gi, o64.finalise_grid(), gain_of(o.o), own_grids)
    assert ok, f"image parity (inner, gain-normalised, own-grid finalise) = {info}"
    assert rel(g.finalise_weights(), o.sumw) < 1e-6


@pytest.mark.parametrize("npol", [1, 4])
def test_wstack_degrid_matches_oracle(gpu_ctx, npol):
    announce(f"wstack degrid npol={npol}")
    import yandasoft_amd as ya
    N, cell, d = _data(npol, 80 + npol)
    model = random_model(N, npol, seed=9)
    g = ya.Gridder(gpu_ctx, "GPUWStack", dict(WS))
    g.initialise_degrid(ya.ImageDesc(N, N, npol, 1, cell, cell), model.astype(np.float32))
    out = np.zeros_like(d["vis"])
    g.degrid(pchunk(d, vis_out=out))
    ov = O.OracleWStack(WS["wmax"], WS["nwplanes"]).degrid((N, N, npol, 1), (cell, cell), model, ochunk(d))
    assert rel(np.transpose(out, (2, 1, 0)), ov) < TOL


def test_wstack_rejects_an_even_plane_count(gpu_ctx):
    announce("wstack: an even number of w planes is an error")
    import yandasoft_amd as ya
    with pytest.raises(ya.YgridError):
        ya.Gridder(gpu_ctx, "GPUWStack", {"nwplanes": 8})  # odd number of planes


@pytest.mark.parametrize("npol", [1, 2])
def test_wstack_pcf_matches_oracle(gpu_ctx, npol):
    """The preconditioner pass (WStackVisGridder.cc:173-262): every plane's grid of the 1 + 3i
    nearest-neighbour kernel (conjugated in half the plane) and the stacked, w-scaled,
    once-transformed image against the oracle's composition."""
    announce(f"wstack pcf npol={npol}")
    import yandasoft_amd as ya
    N, cell, d = _data(npol, 100 + npol)
    g = ya.Gridder(gpu_ctx, "GPUWStack", dict(WS))
    g.initialise_grid(ya.ImageDesc(N, N, npol, 1, cell, cell), dopcf=True)
    g.grid(pchunk(d))
    o = O.OracleWStack(WS["wmax"], WS["nwplanes"], precision="f64")
    grids = o.grid((N, N, npol, 1), (cell, cell), ochunk(d), dopcf=True)
    gg = g.grid_array().reshape(WS["nwplanes"], 1, npol, N, N)
    used = [iw for iw in range(WS["nwplanes"]) if np.abs(grids[iw]).max() > 0]
    assert len(used) >= 3 and any(o.pcf_scale(iw) != 1.0 for iw in used)
    peak = max(np.abs(x).max() for x in grids)
    for iw in range(WS["nwplanes"]):
        assert np.abs(gg[iw] - grids[iw]).max() <= TOL * peak, iw
    assert rel(g.finalise_grid(), o.finalise_grid()) < TOL


SAMPLINGS = [
    ("powerlaw", {"wsampling": "powerlaw", "wsampling.exponent": 0.5}, dict(kind="powerlaw", exponent=0.5)),
    ("gaussian", {"wsampling": "gaussian", "wsampling.nwplanes50": 4.0}, dict(kind="gaussian", nwplanes50=4.0)),
    ("wmaxclip", {"wmaxclip": True}, dict(wmaxclip=True)),
]


@pytest.mark.parametrize("case", SAMPLINGS, ids=[c[0] for c in SAMPLINGS])
def test_wstack_sampling_kinds_and_wmaxclip(gpu_ctx, case):
    """Non-linear w sampling (power law, gaussian: the planes' w from WDependentGridderBase's
    samplings, PowerWSampling.cc / GaussianWSampling.cc) and wmaxclip (samples beyond wmax left
    out, TableVisGridder.cc:698-705): every plane's grid and the degridded visibilities against
    the oracle's composition with the same sampling."""
    name, params, okw = case
    announce(f"wstack w sampling {name}")
    import yandasoft_amd as ya
    npol = 2
    N, cell = 128, 60 * ARC
    # wmaxclip: w up to 1.3 wmax, so that some samples lie beyond the last plane
    wspan = 1.3 * WS["wmax"] if name == "wmaxclip" else 0.9 * WS["wmax"]
    d = make_data(900, 3, npol, N, cell, f0=1.0e9, frac=0.35, wmax_m=wspan * O.C_LIGHT / 1.01e9, seed=95, flags=True)
    g = ya.Gridder(gpu_ctx, "GPUWStack", dict(WS, **params))
    g.initialise_grid(ya.ImageDesc(N, N, npol, 1, cell, cell))
    g.grid(pchunk(d))
    o = O.OracleWStack(WS["wmax"], WS["nwplanes"], **okw)
    grids = o.grid((N, N, npol, 1), (cell, cell), ochunk(d))
    gg = g.grid_array().reshape(WS["nwplanes"], 1, npol, N, N)
    assert len([iw for iw in range(WS["nwplanes"]) if np.abs(grids[iw]).max() > 0]) >= 3
    peak = max(np.abs(x).max() for x in grids)
    for iw in range(WS["nwplanes"]):
        assert np.abs(gg[iw] - grids[iw]).max() <= TOL * peak, iw
    if name == "wmaxclip":
        assert g.stats().vectors_wflagged > 0  # some samples were beyond wmax
    model = random_model(N, npol, seed=19)
    dg = ya.Gridder(gpu_ctx, "GPUWStack", dict(WS, **params))
    dg.initialise_degrid(ya.ImageDesc(N, N, npol, 1, cell, cell), model.astype(np.float32))
    out = np.zeros_like(d["vis"])
    dg.degrid(pchunk(d, vis_out=out))
    ov = O.OracleWStack(WS["wmax"], WS["nwplanes"], **okw).degrid((N, N, npol, 1), (cell, cell), model, ochunk(d))
    assert rel(np.transpose(out, (2, 1, 0)), ov) < TOL
