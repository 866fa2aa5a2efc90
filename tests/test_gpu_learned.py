"""Learned CFs (ygrid_lcf_*): the GPU network against the FP64 oracle network, training on a
W-projection table, and what the network gains over the nearest-plane table.

* forward (FP32 and FP16 weights) and loss/gradient of the GPU kernels vs oracle/mlp.py;
* training: the loss falls and the table judges the held-out planes;
* gain: visibilities at continuous w gridded with the network at each sample's w against
  (a) the same gridder with its coarse w-plane table (nearest plane) and (b) an FP64 oracle
  gridder with 1025 w planes standing in for continuous w: the network must be closer to (b).
* fallback: with an error bound the network does not meet, the gridder uses its table.
"""
import numpy as np
import pytest

from oracle import mlp as M
from oracle import oracle as O
from tests.conftest import announce
from tests.test_gpu_parity import ARC, make_data, ochunk, pchunk, rel

pytestmark = pytest.mark.gpu


def test_forward_and_gradients_match_oracle(gpu_ctx):
    announce("learned CF: forward + gradients vs oracle")
    import yandasoft_amd as ya
    lcf = ya.LearnedCF(gpu_ctx, hidden=128, layers=4, seed=5)
    p = lcf.get_params()
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (300, 5)).astype(np.float32)
    t = rng.normal(0, 0.5, (300, 2)).astype(np.float32)
    y32 = lcf.forward(x, half=False)
    ref = M.forward(p, x, 128, 4)
    assert rel(y32, ref) < 1e-5
    # FP16 weights (biases stay FP32), FP32 products and sums: the oracle with the same rounding
    y16 = lcf.forward(x, half=True)
    p16 = p.astype(np.float64)
    o = 0
    for a, b in zip(M.dims(128, 4)[:-1], M.dims(128, 4)[1:]):
        p16[o:o + a * b] = p[o:o + a * b].astype(np.float16)
        o += a * b + b
    ref16 = M.forward(p16, x, 128, 4)
    assert rel(y16, ref16) < 1e-5
    # the fused FP16 matrix-core form (the gridders' inference): FP16 weights, inputs and hidden
    # activations, FP32 sums -- against the oracle with the same roundings (FP64 sums; an
    # activation whose FP32 and FP64 pre-activations round to neighbouring FP16 values differs by
    # one step of 2^-11), and within the propagated rounding bound of the exact network
    y2 = lcf.forward(x, half=2)
    assert rel(y2, M.forward_f16(p, x, 128, 4)) < 1e-3
    assert rel(y2, ref) < 4 * 5 * 2.0 ** -11
    loss, g = lcf.loss_grad(x, t)
    oloss, og = M.loss_grad(p, x, t, 128, 4)
    assert loss == pytest.approx(oloss, rel=1e-5)
    assert np.max(np.abs(g - og)) <= 1e-5 * np.max(np.abs(og))


WP = dict(wmax=3000.0, oversample=4, maxsupport=256, cutoff=1e-2)


def _grid(ctx, params, d, N, cell, lcf=None, max_err=1.0):
    import yandasoft_amd as ya
    g = ya.Gridder(ctx, "GPUWProject", params)
    if lcf is not None:
        g.set_learned_cf(lcf, max_err)
    g.initialise_grid(ya.ImageDesc(N, N, 1, 1, cell, cell))
    g.grid(pchunk(d))
    return g


def test_learned_cf_trains_and_beats_nearest_plane(gpu_ctx):
    announce("learned CF: train on a 129-plane table, grid continuous w, vs 9-plane table and 1025-plane oracle")
    import yandasoft_amd as ya
    N, cell = 256, 30 * ARC
    d = make_data(1500, 2, 1, N, cell, f0=1.0e9, frac=0.35, wmax_m=2500.0 * O.C_LIGHT / 1.005e9, seed=21)
    # training table: 129 w planes (2064 CFs), generated by a first pass
    tg = _grid(gpu_ctx, dict(WP, nwplanes=129), d, N, cell)
    lcf = ya.LearnedCF(gpu_ctx, hidden=128, layers=4, seed=2)
    rep = lcf.train_wproject(tg, epochs=300, batch=1024, lr=3e-3, seed=3, holdout_every=8)
    print(f"[learned] {rep.cfs} CFs ({rep.examples} taps) + {rep.holdout_cfs} held out, support {rep.support}, "
          f"loss {rep.first_loss:.3e} -> {rep.last_loss:.3e}, held-out max err {rep.holdout_max_err:.3e} "
          f"rms {rep.holdout_rms_err:.3e}, {rep.train_s:.1f} s", flush=True)
    assert rep.cfs + rep.holdout_cfs == 129 * 16 and rep.holdout_cfs == 16 * 16
    assert rep.last_loss < 0.05 * rep.first_loss
    assert rep.holdout_max_err < 0.05
    # continuous-w stand-in: FP64 oracle with 1025 planes
    o = O.OracleGridder("WProject", "f64", wproject=O.WProjectParams(**dict(WP, nwplanes=1025)))
    o.initialise_grid((N, N, 1, 1), (cell, cell))
    o.grid(ochunk(d))
    truth = o.grid_[0, 0]
    coarse = _grid(gpu_ctx, dict(WP, nwplanes=9), d, N, cell).grid_array().reshape(N, N)
    lg = _grid(gpu_ctx, dict(WP, nwplanes=129), d, N, cell, lcf=lcf, max_err=1.0)
    assert lg.stats().learned_cf_active == 1
    learned = lg.grid_array().reshape(N, N)
    fine = tg.grid_array().reshape(N, N)
    e_coarse, e_learned, e_fine = rel(coarse, truth), rel(learned, truth), rel(fine, truth)
    print(f"[learned] grid error vs 1025-plane FP64 oracle: 9-plane table {e_coarse:.3e}, "
          f"129-plane table {e_fine:.3e}, network at each sample's w {e_learned:.3e}", flush=True)
    assert e_learned < e_coarse


def test_learned_planes_are_evaluated_once_per_chunk(gpu_ctx):
    """A pass over the samples of the network's last evaluation copies its planes instead of
    evaluating again (a degridding after the gridding of the same chunk); other samples are
    evaluated.  Results equal those of a fresh evaluation (to the float-atomic summation order)."""
    announce("learned CF: one network evaluation per chunk shared by the grid and degrid passes")
    import yandasoft_amd as ya
    N, cell = 128, 60 * ARC
    dA = make_data(400, 2, 1, N, cell, f0=1.0e9, frac=0.35, wmax_m=1000.0 * O.C_LIGHT / 1.005e9, seed=5)
    dB = make_data(400, 2, 1, N, cell, f0=1.0e9, frac=0.35, wmax_m=1000.0 * O.C_LIGHT / 1.005e9, seed=6)
    params = dict(WP, nwplanes=33)
    tg = _grid(gpu_ctx, params, dA, N, cell)
    lcf = ya.LearnedCF(gpu_ctx, hidden=128, layers=2, seed=2)
    lcf.train_wproject(tg, epochs=3, batch=1024, lr=3e-3, seed=3, holdout_every=8)
    rng = np.random.default_rng(4)
    model = np.zeros((1, 1, N, N), np.float32)
    model[0, 0, rng.integers(32, 96, 15), rng.integers(32, 96, 15)] = 1.0
    img = ya.ImageDesc(N, N, 1, 1, cell, cell)

    def degrid(d, clear):
        if clear:
            lcf.clear_cache()
        g = ya.Gridder(gpu_ctx, "GPUWProject", params)
        g.set_learned_cf(lcf, 1.0)
        g.initialise_degrid(img, model)
        out = np.zeros_like(d["vis"])
        g.degrid(pchunk(d, vis_out=out))
        assert g.stats().learned_cf_active == 1
        return out

    gA = _grid(gpu_ctx, params, dA, N, cell, lcf=lcf)   # evaluates A
    reuse_A = degrid(dA, clear=False)                    # copies A's planes from the gridder's buffer
    fresh_A = degrid(dA, clear=True)
    assert rel(reuse_A, fresh_A) < 1e-6
    after_A = degrid(dB, clear=False)                    # other samples: evaluated
    fresh_B = degrid(dB, clear=True)
    assert rel(after_A, fresh_B) < 1e-6
    assert rel(fresh_A, fresh_B) > 1e-2                  # the two chunks' planes do differ
    # the gridder's own buffer: a second pass over A reuses what it holds
    gA.initialise_grid(img)
    gA.grid(pchunk(dA))
    g1 = gA.grid_array().copy()
    gA.initialise_grid(img)
    gA.grid(pchunk(dA))
    assert rel(gA.grid_array(), g1) < 1e-6


def test_learned_cf_falls_back_to_the_table(gpu_ctx):
    announce("learned CF: fallback when the held-out error exceeds the bound")
    import yandasoft_amd as ya
    N, cell = 128, 60 * ARC
    d = make_data(300, 1, 1, N, cell, f0=1.0e9, frac=0.35, wmax_m=1000.0 * O.C_LIGHT / 1.005e9, seed=4)
    params = dict(WP, nwplanes=17)
    tg = _grid(gpu_ctx, params, d, N, cell)
    lcf = ya.LearnedCF(gpu_ctx, hidden=16, layers=2, seed=2)
    rep = lcf.train_wproject(tg, epochs=1, batch=1024, lr=1e-3, seed=3, holdout_every=4)
    assert rep.holdout_max_err > 1e-9
    g = _grid(gpu_ctx, params, d, N, cell, lcf=lcf, max_err=1e-9)
    assert g.stats().learned_cf_active == 0
    assert rel(g.grid_array(), tg.grid_array()) < 1e-6
    with pytest.raises(ya.YgridError):
        sph = ya.Gridder(gpu_ctx, "GPUSphFunc", {})
        sph.set_learned_cf(lcf, 1.0)


def test_learned_cf_full_size_beats_the_129_plane_table(gpu_ctx):
    """The continuous-w workload at its full size (160k rows, uv in a 2 km disc, w uniform in
    +-300 m, 2048^2, cutoff 1e-2): a network trained on a 1441-plane table (20,176 CFs, every 8th
    plane held out) grids every row at its own w; against an FP64 oracle with 1025 w planes it
    must be closer than the 129-plane table the plain gridder uses."""
    announce("learned CF full size: 160k rows 2048^2, trained on ~20k CFs, vs 129-plane table and 1025-plane oracle")
    import yandasoft_amd as ya
    from yandasoft_amd import workloads as W
    wl = W.learned_wdisc()
    img = ya.ImageDesc(wl.nx, wl.ny, 1, 1, wl.cell, wl.cell)
    ch = ya.Chunk(uvw=wl.uvw, freq=wl.freq, vis=wl.vis, noise=wl.noise)

    def grid(params, lcf=None):
        g = ya.Gridder(gpu_ctx, "GPUWProject", params)
        if lcf is not None:
            g.set_learned_cf(lcf, 1.0)
        g.initialise_grid(img)
        g.grid(ch)
        return g

    tg = grid(dict(wl.params, nwplanes=1441))
    lcf = ya.LearnedCF(gpu_ctx, hidden=128, layers=4, seed=2)
    rep = lcf.train_wproject(tg, epochs=40, batch=1024, lr=3e-3, seed=3, holdout_every=8)
    print(f"[learned] {rep.cfs} CFs ({rep.examples} taps) + {rep.holdout_cfs} held out, support {rep.support}, "
          f"loss {rep.first_loss:.3e} -> {rep.last_loss:.3e}, held-out max err {rep.holdout_max_err:.3e} "
          f"rms {rep.holdout_rms_err:.3e}, {rep.steps} steps in {rep.train_s:.1f} s", flush=True)
    assert rep.cfs == (1441 - 180) * 16 and rep.holdout_cfs == 180 * 16
    assert rep.last_loss < 0.05 * rep.first_loss
    p = wl.params
    o = O.OracleGridder("WProject", "f64", wproject=O.WProjectParams(
        wmax=p["wmax"], nwplanes=1025, cutoff=p["cutoff"], oversample=p["oversample"], maxsupport=p["maxsupport"]))
    o.initialise_grid((wl.nx, wl.ny, 1, 1), (wl.cell, wl.cell))
    o.grid(O.Chunk(wl.uvw, wl.freq, np.transpose(wl.vis, (2, 1, 0)), noise=np.transpose(wl.noise, (2, 1, 0))))
    truth = o.grid_[0, 0]
    assert tg.stats().support == o.support
    table = grid(dict(wl.params, nwplanes=129)).grid_array().reshape(wl.ny, wl.nx)
    lg = grid(dict(wl.params, nwplanes=129), lcf)
    assert lg.stats().learned_cf_active == 1
    learned = lg.grid_array().reshape(wl.ny, wl.nx)
    fine = tg.grid_array().reshape(wl.ny, wl.nx)
    e_table, e_learned, e_fine = rel(table, truth), rel(learned, truth), rel(fine, truth)
    print(f"[learned] grid error vs 1025-plane FP64 oracle: 129-plane table {e_table:.3e}, "
          f"network at each sample's w {e_learned:.3e} (1441-plane table {e_fine:.3e})", flush=True)
    assert e_learned < e_table
