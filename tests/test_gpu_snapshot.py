"""Snapshot imaging (gridder.snapshotimaging; SnapShotImagingGridderAdapter.cc) on the device
against the oracle's composition (oracle.OracleSnapshot: the least-squares w-plane fit, the
plane-relative w, the wrapped gridder's snapshot images regridded between the slanted and the
target SIN frames, pinned in test_oracle_pins): images, weights, fits and regrids, degridding,
the PSF pass, host and device chunks, and the reference's own acceptance bar for a wide-field
source (wtermtest.py: position within one cell, flux within 10%)."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.devmem import DevArray
from tests.test_gpu_parity import ARC, TOL, random_model, rel

pytestmark = pytest.mark.gpu

N, CELL = 512, 40 * ARC
F = np.array([1.2e9, 1.21e9])
KX, KY = 200, -150                          # the source, in pixels from the tangent point
PLANES = [(1.5, 0.5), (-1.0, 1.2), (0.3, -2.0)]


def _chunks(nrow=700, seed=5):
    """Coplanar snapshots (w = A u + B v) of a 1 Jy source at (KX, KY) pixels, 2 channels, 1 pol."""
    rng = np.random.default_rng(seed)
    lam = O.C_LIGHT / F.max()
    uvmax = 0.4 * (N / 2) / (N * CELL) * lam
    l0, m0 = KX * CELL, KY * CELL
    n0 = math.sqrt(1 - l0 * l0 - m0 * m0)
    out = []
    for A, B in PLANES:
        r = uvmax * np.sqrt(rng.random(nrow))
        th = 2 * np.pi * rng.random(nrow)
        u, v = r * np.cos(th), r * np.sin(th)
        uvw = np.ascontiguousarray(np.column_stack([u, v, A * u + B * v]))
        ph = (u * l0 + v * m0 + uvw[:, 2] * (n0 - 1))[None, :] * F[:, None] / O.C_LIGHT
        vis = np.ascontiguousarray(np.exp(2j * np.pi * ph)[None].astype(np.complex64))  # (1, nchan, nrow)
        out.append((uvw, vis))
    return out


def _oracle(prec="f32", **kw):
    return O.OracleSnapshot(lambda: O.OracleGridder("SphFunc", prec), wtolerance=1.0, **kw)


def _ya_chunk(uvw, vis, device=False, vis_out=None):
    import yandasoft_amd as ya
    noise = np.ones_like(vis)
    if device:
        return ya.Chunk(uvw=DevArray(uvw), freq=DevArray(F), vis=DevArray(vis), noise=DevArray(noise),
                        vis_out=vis_out, device=True)
    return ya.Chunk(uvw=uvw, freq=F, vis=vis, noise=noise, vis_out=vis_out)


def _och(uvw, vis):
    return O.Chunk(uvw, F, np.transpose(vis, (2, 1, 0)), noise=np.ones((uvw.shape[0], len(F), 1), np.complex64))


OPTS = {"snapshotimaging": True, "snapshotimaging.wtolerance": 1.0}


@pytest.mark.parametrize("case", ["cubic-host", "cubic-device", "linear-clipped"])
def test_snapshot_grid_image_weights_match_oracle(gpu_ctx, case):
    announce(f"snapshot imaging grid {case}: 3 coplanar snapshots, SphFunc, {N}^2")
    import yandasoft_amd as ya
    interp = "linear" if case.startswith("linear") else "cubic"
    clip = 0.1 if case.endswith("clipped") else 0.0
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {})
    g.configure(dict(OPTS, **{"snapshotimaging.interpmethod": interp, "snapshotimaging.clipping": clip}))
    g.initialise_grid(ya.ImageDesc(N, N, 1, 1, CELL, CELL))
    o = _oracle(interp=interp, clipping=clip)
    o.initialise_grid((N, N, 1, 1), (CELL, CELL))
    dev = case.endswith("device")
    for uvw, vis in _chunks():
        g.grid(_ya_chunk(uvw, vis, dev))
        o.grid(_och(uvw, vis))
    gi, oi = g.finalise_grid(), o.finalise_grid()
    st = g.stats()
    assert (st.snapshot_fits, st.snapshot_regrids) == (3, 6) and (o.fits, o.regrids) == (3, 6)
    assert (st.snapshot_a, st.snapshot_b) == pytest.approx(PLANES[-1], abs=1e-9)
    q = slice(N // 4, 3 * N // 4)
    assert rel(gi[..., q, q], oi[..., q, q]) < TOL, rel(gi[..., q, q], oi[..., q, q])
    assert rel(gi, oi) < 1e-4
    gw, ow = g.finalise_weights(), o.finalise_weights()
    assert rel(gw, ow) < 1e-6
    # the reference's bar (wtermtest.py): the source at its true pixel, flux within 10%
    im = gi[0, 0]
    assert np.unravel_index(np.argmax(im), im.shape) == (N // 2 + KY, N // 2 + KX)
    assert abs(im.max() / (3 * 700 * 2) - 1.0) < 0.1


def test_snapshot_degrid_and_psf_match_oracle(gpu_ctx):
    announce("snapshot imaging degrid (model regridded per plane) and PSF pass-through")
    import yandasoft_amd as ya
    model = random_model(N, 1, seed=3, npts=12)
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {})
    g.configure(OPTS)
    g.initialise_degrid(ya.ImageDesc(N, N, 1, 1, CELL, CELL), model.astype(np.float32))
    o = _oracle()
    o.initialise_degrid((N, N, 1, 1), (CELL, CELL), model.astype(np.float32).astype(np.float64))
    for uvw, vis in _chunks(seed=6):
        out = np.zeros_like(vis)
        g.degrid(_ya_chunk(uvw, vis, vis_out=out))
        ov = o.degrid(_och(uvw, vis))
        assert rel(np.transpose(out, (2, 1, 0)), ov) < TOL
    g.finalise_degrid()
    assert g.stats().snapshot_regrids == 3
    # PSF: the wrapped gridder directly, on the plane-relative w (no regrid by default)
    p = ya.Gridder(gpu_ctx, "GPUSphFunc", {})
    p.configure(OPTS)
    p.initialise_grid(ya.ImageDesc(N, N, 1, 1, CELL, CELL), dopsf=True)
    op = _oracle()
    op.initialise_grid((N, N, 1, 1), (CELL, CELL), dopsf=True)
    for uvw, vis in _chunks(seed=7):
        p.grid(_ya_chunk(uvw, vis))
        op.grid(_och(uvw, vis))
    gp, opi = p.finalise_grid(), op.finalise_grid()
    q = slice(N // 4, 3 * N // 4)
    assert rel(gp[..., q, q], opi[..., q, q]) < TOL
    assert p.stats().snapshot_regrids == 0 and p.stats().snapshot_fits == 3
    # an empty model degrids nothing
    e = ya.Gridder(gpu_ctx, "GPUSphFunc", {})
    e.configure(OPTS)
    e.initialise_degrid(ya.ImageDesc(N, N, 1, 1, CELL, CELL), np.zeros((1, 1, N, N), np.float32))
    assert e.is_model_empty()


def test_snapshot_options_are_checked(gpu_ctx):
    import yandasoft_amd as ya
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {})
    with pytest.raises(Exception, match="wtolerance"):
        g.configure({"snapshotimaging": True})
    with pytest.raises(Exception, match="longtrack"):
        g.configure(dict(OPTS, **{"snapshotimaging.longtrack": True}))
    with pytest.raises(Exception, match="interpmethod"):
        g.configure(dict(OPTS, **{"snapshotimaging.interpmethod": "lanczos"}))
    with pytest.raises(Exception, match="Clipping"):
        g.configure(dict(OPTS, **{"snapshotimaging.clipping": 1.0}))
    # a snapshot gridder is refused where its wrapper would be bypassed
    s = ya.Gridder(gpu_ctx, "GPUSphFunc", {})
    s.configure(OPTS)
    s.initialise_grid(ya.ImageDesc(N, N, 1, 1, CELL, CELL))
    uvw, vis = _chunks(nrow=50)[0]
    with pytest.raises(ya.YgridError, match="snapshot"):
        ya.imaging_chunk([], [s], _ya_chunk(uvw, vis))
    b = ya.Batch(gpu_ctx)
    with pytest.raises(ya.YgridError, match="snapshot"):
        b.add(s, _ya_chunk(uvw, vis, device=True))


def test_snapshot_around_wprojection_matches_oracle(gpu_ctx):
    """The reference's own use (wtermtest.py: gridder = WProject, snapshotimaging.wtolerance):
    rows off their snapshot's plane by up to 40 m (160 wavelengths; tolerance 300 wavelengths, so
    each snapshot is one fit) and the W-projection kernel (wmax 250 wavelengths) takes the
    plane-relative w; images, weights and degridding against the oracle's composition, and a
    clone grids as the original."""
    announce("snapshot imaging around W-projection: non-coplanar snapshots, 256^2, 2 pols")
    import yandasoft_amd as ya
    N2, cell = 256, 60 * ARC
    wp = dict(wmax=250.0, nwplanes=5, oversample=4, maxsupport=128, cutoff=1e-2)
    rng = np.random.default_rng(31)
    lam = O.C_LIGHT / F.max()
    uvmax = 0.35 * (N2 / 2) / (N2 * cell) * lam
    chunks = []
    for A, B in ((0.8, -0.3), (-0.6, 0.9)):
        r = uvmax * np.sqrt(rng.random(600))
        th = 2 * np.pi * rng.random(600)
        u, v = r * np.cos(th), r * np.sin(th)
        uvw = np.ascontiguousarray(np.column_stack([u, v, A * u + B * v + rng.uniform(-40, 40, 600)]))
        vis = np.ascontiguousarray((rng.normal(size=(2, len(F), 600)) + 1j * rng.normal(size=(2, len(F), 600)))
                                   .astype(np.complex64))
        chunks.append((uvw, vis))
    opts = {"snapshotimaging": True, "snapshotimaging.wtolerance": 300.0}
    g = ya.Gridder(gpu_ctx, "GPUWProject", dict(wp))
    g.configure(opts)
    img = ya.ImageDesc(N2, N2, 2, 1, cell, cell)
    g.initialise_grid(img)
    c = g.clone()
    c.initialise_grid(img)
    o = O.OracleSnapshot(lambda: O.OracleGridder("WProject", "f32", wproject=O.WProjectParams(**wp)), wtolerance=300.0)
    o.initialise_grid((N2, N2, 2, 1), (cell, cell))
    for uvw, vis in chunks:
        noise = np.ones_like(vis)
        g.grid(ya.Chunk(uvw=uvw, freq=F, vis=vis, noise=noise))
        c.grid(ya.Chunk(uvw=uvw, freq=F, vis=vis, noise=noise))
        o.grid(O.Chunk(uvw, F, np.transpose(vis, (2, 1, 0)), noise=np.transpose(noise, (2, 1, 0))))
    gi, oi, ci = g.finalise_grid(), o.finalise_grid(), c.finalise_grid()
    q = slice(N2 // 4, 3 * N2 // 4)
    assert g.stats().snapshot_fits == 2 and o.fits == 2
    assert rel(gi[..., q, q], oi[..., q, q]) < TOL
    # the clone: the same to FP32 accumulation order (the small-support kernel's per-band pair
    # lists are compacted by LDS atomics, so their order varies between passes; the image edge,
    # where the correction gain reaches ~1e3, amplifies that to ~1e-4 of the peak)
    assert rel(ci[..., q, q], gi[..., q, q]) < 1e-6
    assert rel(g.finalise_weights(), o.finalise_weights()) < 1e-6
    model = random_model(N2, 2, seed=8, npts=10)
    d = ya.Gridder(gpu_ctx, "GPUWProject", dict(wp))
    d.configure(opts)
    d.initialise_degrid(img, model.astype(np.float32))
    od = O.OracleSnapshot(lambda: O.OracleGridder("WProject", "f32", wproject=O.WProjectParams(**wp)), wtolerance=300.0)
    od.initialise_degrid((N2, N2, 2, 1), (cell, cell), model.astype(np.float32).astype(np.float64))
    for uvw, vis in chunks:
        out = np.zeros_like(vis)
        d.degrid(ya.Chunk(uvw=uvw, freq=F, vis=vis, noise=np.ones_like(vis), vis_out=out))
        ov = od.degrid(O.Chunk(uvw, F, np.transpose(vis, (2, 1, 0))))
        assert rel(np.transpose(out, (2, 1, 0)), ov) < TOL


def test_snapshot_psf_with_reprojection_matches_oracle(gpu_ctx):
    """reprojectpsf (SnapShotImagingGridderAdapter.cc:195-211, setPSFReprojection): the PSF pass
    is finalised and regridded per snapshot like the dirty image."""
    announce("snapshot imaging PSF with reprojectpsf")
    import yandasoft_amd as ya
    g = ya.Gridder(gpu_ctx, "GPUSphFunc", {})
    g.configure(dict(OPTS, **{"snapshotimaging.reprojectpsf": True}))
    g.initialise_grid(ya.ImageDesc(N, N, 1, 1, CELL, CELL), dopsf=True)
    o = _oracle(reprojectpsf=True)
    o.initialise_grid((N, N, 1, 1), (CELL, CELL), dopsf=True)
    for uvw, vis in _chunks(seed=8):
        g.grid(_ya_chunk(uvw, vis))
        o.grid(_och(uvw, vis))
    gi, oi = g.finalise_grid(), o.finalise_grid()
    q = slice(N // 4, 3 * N // 4)
    assert g.stats().snapshot_regrids == 6 and o.regrids == 6
    assert rel(gi[..., q, q], oi[..., q, q]) < TOL
    assert rel(g.finalise_weights(), o.finalise_weights()) < 1e-6
