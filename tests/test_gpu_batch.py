"""Batched passes (ygrid_batch_*): one persistent tile-kernel launch over several gridders, HIP
graph capture and replay, and the autotuner.

A batch must give every gridder what its own grid() / degrid() on its chunk gives (the same
kernels; only float-atomic summation order may differ: 1e-6 of the peak), on every graph replay,
with eager work and synchronisations between the replays.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.devmem import DevArray
from tests.test_gpu_parity import ARC, make_data, random_model, rel

pytestmark = pytest.mark.gpu
TOL = 1e-6

WP = dict(wmax=3000.0, nwplanes=9, oversample=4, maxsupport=256, cutoff=1e-2)


def _cases():
    """(gridder name, params, N, cell, npol, data) of uneven sizes and kinds."""
    out = []
    for i, (N, nrow, nchan) in enumerate([(256, 1500, 2), (128, 300, 5), (256, 40, 1), (128, 900, 3)]):
        d = make_data(nrow, nchan, 4, N, 30 * ARC * 256 / N, f0=1.0e9, frac=0.35,
                      wmax_m=2500.0 * O.C_LIGHT / 1.005e9, seed=100 + i)
        out.append(("GPUWProject", dict(WP), N, 30 * ARC * 256 / N, 4, d))
    d = make_data(700, 2, 4, 256, 30 * ARC, f0=1.4e9, frac=0.4, seed=200)
    out.append(("GPUSphFunc", {}, 256, 30 * ARC, 4, d))
    return out


def _dev_chunk(d, vis_out=None):
    import yandasoft_amd as ya
    return ya.Chunk(uvw=DevArray(d["uvw"]), freq=DevArray(d["freq"]), vis=DevArray(d["vis"]),
                    noise=DevArray(d["noise"]), vis_out=vis_out, device=True)


def _grid_single(ctx, case):
    import yandasoft_amd as ya
    name, params, N, cell, npol, d = case
    g = ya.Gridder(ctx, name, params)
    g.initialise_grid(ya.ImageDesc(N, N, npol, 1, cell, cell))
    g.grid(_dev_chunk(d))
    return g.grid_array()


def test_batched_grid_matches_single_passes_and_graph_replays(gpu_ctx):
    announce("batch: 5 gridders (WProject 256/128, SphFunc), eager, graph capture + 3 replays")
    import yandasoft_amd as ya
    cases = _cases()
    ref = [_grid_single(gpu_ctx, c) for c in cases]
    gs, chunks = [], []
    b = ya.Batch(gpu_ctx)
    for name, params, N, cell, npol, d in cases:
        g = ya.Gridder(gpu_ctx, name, params)
        g.initialise_grid(ya.ImageDesc(N, N, npol, 1, cell, cell))
        ch = _dev_chunk(d)
        b.add(g, ch)
        gs.append(g)
        chunks.append(ch)
    b.run()
    for g, r, c in zip(gs, ref, cases):
        assert rel(g.grid_array(), r) < TOL, c[0]
    taps_eager = [g.stats().taps_gridded for g in gs]
    assert all(t > 0 for t in taps_eager)
    b.set_graph(True)
    for rep in range(4):  # capture, then replays
        for g, c in zip(gs, cases):
            N, cell, npol = c[2], c[3], c[4]
            g.initialise_grid(ya.ImageDesc(N, N, npol, 1, cell, cell))
        # eager work of the program (a single-gridder pass on the same context) between replays
        assert rel(_grid_single(gpu_ctx, cases[2]), ref[2]) < TOL
        b.run()
        for g, r, c in zip(gs, ref, cases):
            assert rel(g.grid_array(), r) < TOL, (rep, c[0])
    assert b.info()[1] == 1  # captured once, replayed after
    for g, t in zip(gs, taps_eager):
        assert g.stats().taps_gridded == pytest.approx(5 * t)


def test_batched_degrid_matches_single_passes(gpu_ctx):
    announce("batch: degrid passes of 5 gridders, eager and graph")
    import yandasoft_amd as ya
    cases = _cases()
    refs, gs, outs = [], [], []
    b = ya.Batch(gpu_ctx)
    for name, params, N, cell, npol, d in cases:
        model = random_model(N, npol).astype(np.float32)
        g = ya.Gridder(gpu_ctx, name, params)
        g.initialise_degrid(ya.ImageDesc(N, N, npol, 1, cell, cell), model)
        o = DevArray(np.zeros_like(d["vis"]))
        g.degrid(_dev_chunk(d, vis_out=o))
        refs.append(o.numpy())
        g2 = ya.Gridder(gpu_ctx, name, params)
        g2.initialise_degrid(ya.ImageDesc(N, N, npol, 1, cell, cell), model)
        o2 = DevArray(np.zeros_like(d["vis"]))
        b.add(g2, _dev_chunk(d, vis_out=o2))
        gs.append(g2)
        outs.append(o2)
    b.run()
    for o, r in zip(outs, refs):
        assert rel(o.numpy(), r) < TOL
    b.set_graph(True)
    for rep in range(3):
        for o in outs:
            o.zero_()
        b.run()
        for o, r in zip(outs, refs):
            assert rel(o.numpy(), r) < TOL, rep


def test_autotune_picks_a_variant_and_results_hold(gpu_ctx):
    announce("batch autotune")
    import yandasoft_amd as ya
    cases = _cases()[:3]
    ref = [_grid_single(gpu_ctx, c) for c in cases]
    ctx = ya.Context(0)
    gs = []
    b = ya.Batch(ctx)
    for name, params, N, cell, npol, d in cases:
        g = ya.Gridder(ctx, name, params)
        g.initialise_grid(ya.ImageDesc(N, N, npol, 1, cell, cell))
        b.add(g, _dev_chunk(d))
        gs.append(g)
    t = b.autotune(1)
    assert t.trials == 12 and t.ms > 0
    assert t.grid_waves in (2, 3) and t.item_pairs in (256, 512, 2048)
    st = gs[0].stats()
    assert (st.grid_waves, st.grid_pipeline) == (t.grid_waves, t.grid_pipeline)
    for g, c in zip(gs, cases):
        g.initialise_grid(ya.ImageDesc(c[2], c[2], c[4], 1, c[3], c[3]))
    b.run()
    for g, r in zip(gs, ref):
        assert rel(g.grid_array(), r) < TOL
    b.close()
    ctx.close()


def test_batch_rejects_host_chunks_and_repeated_gridders(gpu_ctx):
    import yandasoft_amd as ya
    name, params, N, cell, npol, d = _cases()[2]
    g = ya.Gridder(gpu_ctx, name, params)
    g.initialise_grid(ya.ImageDesc(N, N, npol, 1, cell, cell))
    b = ya.Batch(gpu_ctx)
    host = ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], noise=d["noise"])
    c = host.c()
    import ctypes
    rc = ya.lib().ygrid_batch_add(b.h, g.h, ctypes.byref(c))
    assert rc == 1 and b"device memory" in ya.lib().ygrid_last_error()
    b.add(g, _dev_chunk(d))
    with pytest.raises(ya.YgridError):
        b.add(g, _dev_chunk(d))
