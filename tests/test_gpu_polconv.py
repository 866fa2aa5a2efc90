"""Polarisation conversion around the gridding passes (TableVisGridder.cc:517-524,682-693,860)
against the oracle: the GPU gridder given the visibilities' Stokes codes and the image's grids
(or predicts) in the image frame; the oracle converts with oracle.pol_to_image / pol_to_vis and
grids (degrids) in the image frame.  Linear and circular feeds, Stokes I and IQUV images,
noise and flags; undetermined image polarisations and unknown codes are errors."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.devmem import DevArray
from tests.test_gpu_parity import ARC, TOL, make_data, ochunk, random_model, rel

pytestmark = pytest.mark.gpu

LIN, CIRC, IQUV = (9, 10, 11, 12), (5, 6, 7, 8), (1, 2, 3, 4)


def _chunk(d, stokes, vis_out=None, device=False):
    import yandasoft_amd as ya
    if device:
        return ya.Chunk(uvw=DevArray(d["uvw"]), freq=DevArray(d["freq"]), vis=DevArray(d["vis"]),
                        noise=DevArray(d["noise"]), flag=None if d["flag"] is None else DevArray(d["flag"]),
                        vis_out=vis_out, stokes=stokes, device=True)
    return ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], vis_out=vis_out, noise=d["noise"], flag=d["flag"],
                    stokes=stokes)


@pytest.mark.parametrize("feeds,image", [(LIN, (1,)), (LIN, IQUV), (CIRC, IQUV), ((9, 12), (1, 2))])
def test_grid_converts_to_the_image_frame(gpu_ctx, feeds, image):
    announce(f"polconv grid {feeds} -> {image}")
    import yandasoft_amd as ya
    N, cell = 128, 60 * ARC
    d = make_data(800, 3, len(feeds), N, cell, seed=60, flags=True, noise_var=True)
    d["noise"] = (np.real(d["noise"]) * (1 + 1j)).astype(np.complex64)  # sigma on both parts, as accessors
    img = ya.ImageDesc(N, N, len(image), 1, cell, cell, stokes=image)
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 16})
    g.initialise_grid(img)
    g.grid(_chunk(d, feeds))
    v, n, f = O.pol_to_image(d["vis"], d["noise"], d["flag"].astype(bool), feeds, image)
    di = dict(d, vis=v.astype(np.complex64), noise=n.astype(np.complex64), flag=f.astype(np.uint8))
    o = O.OracleGridder(precision="f32", kind="SphFunc", sph_oversample=16)
    o.initialise_grid((N, N, len(image), 1), (cell, cell))
    o.grid(ochunk(di))
    assert rel(g.grid_array().reshape(1, len(image), N, N), o.grid_) < TOL
    assert rel(g.finalise_weights(), o.finalise_weights()) < 1e-5
    # the PSF gridder converts the noise (weights) but grids no data
    p = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 16})
    p.initialise_grid(img, dopsf=True)
    p.grid(_chunk(d, feeds))
    op = O.OracleGridder(precision="f32", kind="SphFunc", sph_oversample=16)
    op.initialise_grid((N, N, len(image), 1), (cell, cell), dopsf=True)
    op.grid(ochunk(di))
    assert rel(p.grid_array().reshape(1, len(image), N, N), op.grid_) < TOL


@pytest.mark.parametrize("device", [False, True])
@pytest.mark.parametrize("feeds,image", [(LIN, IQUV), (LIN, (1,)), (CIRC, (1, 4))])
def test_degrid_converts_to_the_visibility_frame(gpu_ctx, feeds, image, device):
    announce(f"polconv degrid {image} -> {feeds} device={device}")
    import yandasoft_amd as ya
    N, cell = 128, 60 * ARC
    d = make_data(700, 2, len(feeds), N, cell, seed=61, flags=True)
    img = ya.ImageDesc(N, N, len(image), 1, cell, cell, stokes=image)
    model = random_model(N, len(image), seed=7)
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 16})
    g.initialise_degrid(img, model.astype(np.float32))
    start = (0.25 * d["vis"]).astype(np.complex64)  # the model is added to what vis_out holds
    out = DevArray(start) if device else start.copy()
    g.degrid(_chunk(d, feeds, vis_out=out, device=device))
    got = out.numpy() if device else out
    o = O.OracleGridder(precision="f32", kind="SphFunc", sph_oversample=16)
    o.initialise_degrid((N, N, len(image), 1), (cell, cell), model)
    _, _, fi = O.pol_to_image(None, None, d["flag"].astype(bool), feeds, image)
    di = dict(d, vis=np.zeros((len(image),) + d["vis"].shape[1:], np.complex64), noise=None,
              flag=fi.astype(np.uint8))
    di["noise"] = np.ones_like(di["vis"])
    mi = np.transpose(o.degrid(ochunk(di)), (2, 1, 0))  # image-frame prediction (npol_img, nchan, nrow)
    want = start + O.pol_to_vis(mi, image, feeds)
    assert rel(got, want) < TOL


def test_conversion_errors(gpu_ctx):
    announce("polconv: undetermined and unknown polarisations")
    import yandasoft_amd as ya
    N, cell = 64, 60 * ARC
    d = make_data(50, 1, 2, N, cell, seed=62)
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 8})
    g.initialise_grid(ya.ImageDesc(N, N, 1, 1, cell, cell, stokes=(4,)))
    with pytest.raises(ya.YgridError):
        g.grid(_chunk(d, (9, 12)))  # V needs the cross hands
    with pytest.raises(ya.YgridError):
        ya.Gridder(gpu_ctx, "GPUSphFunc", {}).initialise_grid(ya.ImageDesc(N, N, 1, 1, cell, cell, stokes=(13,)))
    # identical frames: no conversion, the plain pass
    g2 = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 8})
    g2.initialise_grid(ya.ImageDesc(N, N, 2, 1, cell, cell, stokes=(9, 12)))
    g2.grid(_chunk(d, (9, 12)))
    g3 = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 8})
    g3.initialise_grid(ya.ImageDesc(N, N, 2, 1, cell, cell))
    g3.grid(_chunk(d, None))
    assert rel(g2.grid_array(), g3.grid_array()) < 1e-6
