"""Benchmark-size workloads on the GPU.

* the SphFunc disc workload at full size (160k visibilities, 2048^2, S=3, os=128) against the
  oracle (grid, image, weights, degridded visibilities), bar 1e-5 of peak;
* the W-projection array geometry (4096^2, 129 w planes, os=4, S=64): a subset of rows against
  the oracle (whose CF planes are generated only for the w planes that subset uses), and all
  rows through size-independent properties: the adjoint identity of grid and degrid
  (sum_s D_s(M) conj(W_s) = sum_p G(W)_p conj(M_p)) and linearity.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.test_gpu_parity import finalise_with, gain_of, image_ok
from yandasoft_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-5


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


def _pchunk(uvw, freq, vis, noise, vis_out=None):
    import yandasoft_amd as ya
    return ya.Chunk(uvw=np.ascontiguousarray(uvw), freq=np.ascontiguousarray(freq),
                    vis=None if vis is None else np.ascontiguousarray(vis),
                    vis_out=vis_out, noise=np.ascontiguousarray(noise))


def test_sphfunc_disc_full_size(gpu_ctx):
    announce("sphfunc disc 160k vis 2048^2 S=3 os=128")
    import yandasoft_amd as ya
    wl = W.sphfunc_disc()
    img = ya.ImageDesc(wl.nx, wl.ny, 1, 1, wl.cell, wl.cell)
    g = ya.Gridder(gpu_ctx, wl.gridder, wl.params)
    g.initialise_grid(img)
    g.grid(_pchunk(wl.uvw, wl.freq, wl.vis, wl.noise))
    ggrid = g.grid_array().reshape(1, 1, wl.ny, wl.nx)
    gimg = g.finalise_grid()
    o = O.OracleGridder("SphFunc", "f32", sph_oversample=128)
    o.initialise_grid((wl.nx, wl.ny, 1, 1), (wl.cell, wl.cell))
    ch = O.Chunk(wl.uvw, wl.freq, np.transpose(wl.vis, (2, 1, 0)), noise=np.transpose(wl.noise, (2, 1, 0)))
    o.grid(ch)
    o64 = O.OracleGridder("SphFunc", "f64", sph_oversample=128)
    o64.initialise_grid((wl.nx, wl.ny, 1, 1), (wl.cell, wl.cell))
    o64.grid(O.Chunk(wl.uvw, wl.freq, np.transpose(wl.vis, (2, 1, 0)), noise=np.transpose(wl.noise, (2, 1, 0))))
    # 160k FP32 accumulations: judged against the FP64 oracle grid (the FP32 restatement's own
    # accumulation order is no closer to it than ours)
    assert rel(ggrid, o64.grid_) < TOL
    assert rel(ggrid, o.grid_) < 2 * TOL
    ok, info = image_ok(gimg, o64.finalise_grid(), gain_of(o), lambda: finalise_with(o, ggrid))
    assert ok, info
    assert rel(g.finalise_weights(), o.finalise_weights()) < 1e-6
    # degrid a point-source model
    rng = np.random.default_rng(8)
    model = np.zeros((1, 1, wl.ny, wl.nx), np.float32)
    for _ in range(40):
        model[0, 0, rng.integers(wl.ny // 4, 3 * wl.ny // 4), rng.integers(wl.nx // 4, 3 * wl.nx // 4)] = rng.normal()
    d = ya.Gridder(gpu_ctx, wl.gridder, wl.params)
    d.initialise_degrid(img, model)
    out = np.zeros_like(wl.vis)
    d.degrid(_pchunk(wl.uvw, wl.freq, wl.vis, wl.noise, vis_out=out))
    od = O.OracleGridder("SphFunc", "f32", sph_oversample=128)
    od.initialise_degrid((wl.nx, wl.ny, 1, 1), (wl.cell, wl.cell), model)
    ov = od.degrid(ch)
    assert rel(np.transpose(out, (2, 1, 0)), ov) < TOL


@pytest.fixture(scope="module")
def wproj():
    return W.wproject_array()


def test_wproject_array_subset_vs_oracle(gpu_ctx, wproj):
    announce("wproject array geometry 4096^2 S=64: row subset vs oracle")
    import yandasoft_amd as ya
    wl = wproj
    rows = np.arange(0, wl.nrow, 97)
    chans = np.array([0, 63])
    uvw = wl.uvw[rows]
    freq = wl.freq[chans]
    vis = np.ascontiguousarray(wl.vis[:, chans][:, :, rows])
    noise = np.ones_like(vis)
    prm = O.WProjectParams(wmax=wl.params["wmax"], nwplanes=wl.params["nwplanes"], cutoff=wl.params["cutoff"],
                           oversample=wl.params["oversample"], maxsupport=wl.params["maxsupport"],
                           limitsupport=wl.params["limitsupport"])
    ws = prm.sampling()
    planes = np.array([[ws.plane(uvw[i, 2] / O.C_LIGHT * f) for f in freq] for i in range(len(rows))])
    # keep the rows whose w planes (at both channels) are among the 8 most used ones, so that the
    # oracle transforms few planes; plane 0 is added for the common support
    vals, cnt = np.unique(planes, return_counts=True)
    keep_planes = set(vals[np.argsort(-cnt)][:8].tolist())
    sel = [i for i in range(len(rows)) if set(planes[i].tolist()) <= keep_planes]
    assert len(sel) >= 10, (len(sel), keep_planes)
    assert len({p for i in sel for p in planes[i]}) >= 3
    uvw, vis, noise = uvw[sel], np.ascontiguousarray(vis[:, :, sel]), np.ascontiguousarray(noise[:, :, sel])
    img = ya.ImageDesc(wl.nx, wl.ny, 4, 1, wl.cell, wl.cell)
    g = ya.Gridder(gpu_ctx, "GPUWProject", wl.params)
    g.initialise_grid(img)
    g.grid(_pchunk(uvw, freq, vis, noise))
    o = O.OracleGridder("WProject", "f32", wproject=prm)
    o.only_planes = set(keep_planes)
    o.initialise_grid((wl.nx, wl.ny, 4, 1), (wl.cell, wl.cell))
    ch = O.Chunk(uvw, freq, np.transpose(vis, (2, 1, 0)), noise=np.transpose(noise, (2, 1, 0)))
    o.grid(ch)
    assert o.support == 64 and g.stats().support == 64
    for z in keep_planes:
        for frac in (0, 5, 15):
            k = frac + 16 * z
            assert rel(g.cf_plane(k), o.cf.planes[k]) < TOL, k
    ggrid = g.grid_array().reshape(1, 4, wl.ny, wl.nx)
    assert rel(ggrid, o.grid_) < TOL
    # degrid from a model with a few sources
    rng = np.random.default_rng(5)
    model = np.zeros((1, 4, wl.ny, wl.nx), np.float32)
    for _ in range(10):
        model[0, :, rng.integers(1000, 3000), rng.integers(1000, 3000)] = rng.normal(size=4)
    d = ya.Gridder(gpu_ctx, "GPUWProject", wl.params)
    d.initialise_degrid(img, model)
    out = np.zeros_like(vis)
    d.degrid(_pchunk(uvw, freq, vis, noise, vis_out=out))
    od = O.OracleGridder("WProject", "f32", wproject=prm)
    od.only_planes = set(keep_planes)
    od.initialise_degrid((wl.nx, wl.ny, 4, 1), (wl.cell, wl.cell), model.astype(np.float64))
    ov = od.degrid(ch)
    assert rel(np.transpose(out, (2, 1, 0)), ov) < TOL


def test_wproject_array_all_rows_adjoint_and_linearity(gpu_ctx, wproj):
    announce("wproject array all rows: adjoint identity + linearity")
    import yandasoft_amd as ya
    wl = wproj
    # a quarter of the channels keeps host memory modest; geometry and CFs are the full ones
    chans = np.arange(0, wl.nchan, 4)
    freq = wl.freq[chans]
    V = np.ascontiguousarray(wl.vis[:, chans])
    noise = np.ones_like(V)
    img = ya.ImageDesc(wl.nx, wl.ny, 4, 1, wl.cell, wl.cell)
    g = ya.Gridder(gpu_ctx, "GPUWProject", wl.params)
    g.initialise_grid(img)
    g.grid(_pchunk(wl.uvw, freq, V, noise))
    G = g.grid_array().astype(np.complex128)
    s = g.stats()
    assert s.taps_gridded == s.samples_gridded * (2 * 64 + 1) ** 2
    rng = np.random.default_rng(9)
    model = np.zeros((1, 4, wl.ny, wl.nx), np.float32)
    for _ in range(30):
        model[0, :, rng.integers(500, 3500), rng.integers(500, 3500)] = rng.normal(size=4)
    d = ya.Gridder(gpu_ctx, "GPUWProject", wl.params)
    d.initialise_degrid(img, model)
    M = d.grid_array().astype(np.complex128)
    out = np.zeros_like(V)
    d.degrid(_pchunk(wl.uvw, freq, V, noise, vis_out=out))
    lhs = np.sum(out.astype(np.complex128) * np.conj(V.astype(np.complex128)))
    rhs = np.sum(G * np.conj(M))
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), abs(rhs))
    # linearity: grid(2V) - 2 grid(V) = 0
    g2 = ya.Gridder(gpu_ctx, "GPUWProject", wl.params)
    g2.initialise_grid(img)
    g2.grid(_pchunk(wl.uvw, freq, np.ascontiguousarray(2 * V), noise))
    assert rel(g2.grid_array(), 2 * G) < TOL


def test_cf_table_over_2gib_matches_oracle(gpu_ctx):
    """A W-projection table of 2.78 GiB (129 w planes x 64 oversampling planes, support 96: the
    reference's table has no size limit, TableVisGridder.h:356).  The tile kernels address it
    through one 32-bit buffer resource per w-plane group (pair descriptors relative to the group's
    first plane); samples on planes at the start, middle and end of the table (offsets beyond
    2 GiB) grid and degrid as the oracle does."""
    announce("CF table of 2.78 GiB (129 w planes, oversample 8, support 96): grid and degrid vs oracle")
    import yandasoft_amd as ya
    from tests.test_gpu_parity import random_model
    N, cell, f = 512, 30 * math.pi / 180 / 3600, 1.0e9
    params = dict(wmax=32000.0, nwplanes=129, oversample=8, maxsupport=2048, cutoff=1e-3, limitsupport=96)
    prm = O.WProjectParams(**params)
    rng = np.random.default_rng(31)
    planes = [0, 1, 64, 127, 128]
    nrow = 250
    lam = O.C_LIGHT / f
    pl = rng.choice(planes, nrow)
    w = ((pl - 64) * 500.0 + rng.uniform(-200.0, 200.0, nrow)) * lam      # metres
    uvmax = 0.8 * (N / 2 - 97) / (N * cell) * lam
    r = uvmax * np.sqrt(rng.random(nrow))
    th = 2 * np.pi * rng.random(nrow)
    uvw = np.ascontiguousarray(np.stack([r * np.cos(th), r * np.sin(th), w], 1))
    vis = (rng.normal(size=(1, 1, nrow)) + 1j * rng.normal(size=(1, 1, nrow))).astype(np.complex64)
    noise = np.ones_like(vis)
    freq = np.array([f])
    img = ya.ImageDesc(N, N, 1, 1, cell, cell)
    g = ya.Gridder(gpu_ctx, "GPUWProject", params)
    g.initialise_grid(img)
    g.grid(ya.Chunk(uvw=uvw, freq=freq, vis=vis, noise=noise))
    st = g.stats()
    assert st.support == 96
    table_bytes = 129 * 64 * (2 * 96 + 9) * (2 * 96 + 33) * 8
    assert table_bytes >= 2 ** 31
    print(f"[big cf] table {table_bytes / 2 ** 30:.2f} GiB, samples gridded {st.samples_gridded:.0f}", flush=True)
    o = O.OracleGridder("WProject", "f32", wproject=prm)
    o.only_planes = set(planes)
    o.initialise_grid((N, N, 1, 1), (cell, cell))
    ch = O.Chunk(uvw, freq, np.transpose(vis, (2, 1, 0)), noise=np.transpose(noise, (2, 1, 0)))
    o.grid(ch)
    assert rel(g.grid_array().reshape(1, 1, N, N), o.grid_) < 1e-5
    for k in (0, 64 * 64 + 9, 64 * 128 + 63):  # first plane, a middle one, the table's last
        assert rel(g.cf_plane(k), o.cf.planes[k]) < 1e-5, k
    model = random_model(N, 1)
    dg = ya.Gridder(gpu_ctx, "GPUWProject", params)
    dg.initialise_degrid(img, model.astype(np.float32))
    out = np.zeros_like(vis)
    dg.degrid(ya.Chunk(uvw=uvw, freq=freq, vis=vis, vis_out=out, noise=noise))
    od = O.OracleGridder("WProject", "f32", wproject=prm)
    od.only_planes = set(planes)
    od.initialise_degrid((N, N, 1, 1), (cell, cell), model)
    assert rel(np.transpose(out, (2, 1, 0)), od.degrid(ch)) < 1e-5
