"""The allocation-free degridding entry point (ygrid_degrid_noalloc) against ygrid_degrid: the
same gridder state and chunk give the same visibilities (to the float-atomic summation order,
within the 1e-5 bar), on two calls in a row, the second with a smaller chunk; the fixed
workspace is refused, with a message, for a chunk that does not fit it; and the oracle agrees."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.conftest import announce
from tests.devmem import DevArray
from tests.test_gpu_parity import ARC, TOL, WP, _wp_oracle_params, make_data, ochunk, random_model, rel

pytestmark = pytest.mark.gpu


def _dev_chunk(d, nrow=None, nchan=None):
    import yandasoft_amd as ya
    nrow = nrow or d["uvw"].shape[0]
    nchan = nchan or d["freq"].shape[0]
    vis = np.ascontiguousarray(d["vis"][:, :nchan, :nrow])
    noise = np.ascontiguousarray(d["noise"][:, :nchan, :nrow])
    out = DevArray(np.zeros_like(vis))
    c = ya.Chunk(uvw=DevArray(np.ascontiguousarray(d["uvw"][:nrow])), freq=DevArray(d["freq"][:nchan]),
                 vis=DevArray(vis), noise=DevArray(noise), vis_out=out, device=True)
    sub = dict(d, uvw=np.ascontiguousarray(d["uvw"][:nrow]), freq=d["freq"][:nchan], vis=vis, noise=noise,
               flag=None, delay=None)
    return c, out, sub


CASES = [
    ("GPUSphFunc", {"oversample": 32}, dict(kind="SphFunc", sph_oversample=32), 1, 0.0),
    ("GPUWProject", dict(WP), None, 4, 2500.0),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_noalloc_degrid_matches_degrid_on_two_calls(gpu_ctx, case):
    name, params, okw, npol, wmax = CASES[case]
    announce(f"noalloc degrid {name} npol={npol}")
    import yandasoft_amd as ya
    if okw is None:
        okw = dict(kind="WProject", wproject=_wp_oracle_params())
    N, cell = 256, 30 * ARC
    d = make_data(1200, 3, npol, N, cell, f0=1.0e9, frac=0.35, wmax_m=wmax * O.C_LIGHT / 1.005e9, seed=60 + case)
    model = random_model(N, npol, seed=5 + case)
    img = ya.ImageDesc(N, N, npol, 1, cell, cell)
    g = ya.Gridder(gpu_ctx, name, params)
    g.initialise_degrid(img, model.astype(np.float32))
    # references from the ordinary entry point (the first builds the CF cache)
    refs = []
    for nrow, nchan in ((1200, 3), (500, 2)):
        c, out, _ = _dev_chunk(d, nrow, nchan)
        g.degrid(c)
        refs.append(out.numpy())
    # two allocation-free calls in a row, the second smaller
    for (nrow, nchan), ref in zip(((1200, 3), (500, 2)), refs):
        c, out, sub = _dev_chunk(d, nrow, nchan)
        g.degrid_noalloc(c)
        got = out.numpy()
        assert np.count_nonzero(ref) > 0
        assert rel(got, ref) < TOL, (nrow, nchan)
        o = O.OracleGridder(precision="f32", **okw)
        o.initialise_degrid((N, N, npol, 1), (cell, cell), model)
        assert rel(np.transpose(got, (2, 1, 0)), o.degrid(ochunk(sub))) < TOL, (nrow, nchan)


def test_noalloc_degrid_clone_and_refusals(gpu_ctx):
    announce("noalloc degrid: clone, missing CF, host chunk, workspace limit")
    import yandasoft_amd as ya
    N, cell = 256, 30 * ARC
    d = make_data(300, 2, 1, N, cell, f0=1.0e9, frac=0.35, seed=71)
    model = random_model(N, 1, seed=3)
    img = ya.ImageDesc(N, N, 1, 1, cell, cell)
    fresh = ya.Gridder(gpu_ctx, "GPUSphFunc", {"oversample": 32})
    fresh.initialise_degrid(img, model.astype(np.float32))
    c, out, _ = _dev_chunk(d)
    with pytest.raises(ya.YgridError) as e:
        fresh.degrid_noalloc(c)  # no CF cache yet
    assert e.value.code == 2
    fresh.degrid(c)
    ref = out.numpy()
    # a clone shares the CF cache: the allocation-free pass runs at once
    cl = fresh.clone()
    cl.initialise_degrid(img, model.astype(np.float32))
    c2, out2, _ = _dev_chunk(d)
    cl.degrid_noalloc(c2)
    assert rel(out2.numpy(), ref) < TOL
    # a host chunk is refused
    h = ya.Chunk(uvw=d["uvw"], freq=d["freq"], vis=d["vis"], vis_out=np.zeros_like(d["vis"]), noise=d["noise"])
    with pytest.raises(ya.YgridError) as e:
        cl.degrid_noalloc(h)
    assert e.value.code == 1
    # a chunk whose pass needs more than the workspace: refused with the size, nothing written
    big = make_data(400000, 1, 1, N, cell, f0=1.0e9, frac=0.35, seed=72)
    cb, outb, _ = _dev_chunk(big)
    with pytest.raises(ya.YgridError) as e:
        cl.degrid_noalloc(cb)
    assert e.value.code == 1 and "MiB" in str(e.value)
    assert not np.any(outb.numpy())
    # and the workspace still serves the next call
    c3, out3, _ = _dev_chunk(d)
    cl.degrid_noalloc(c3)
    assert rel(out3.numpy(), ref) < TOL
